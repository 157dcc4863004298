// Shared host-side plumbing of libb3dhip: error reporting and the limb-count dispatch.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include <initializer_list>

namespace b3d {
void set_error(const std::string& msg);
int check_launch(const char* what);

// Limb planes are addressed with 32-bit byte offsets (buffer loads, see mp.hpp): every plane stride (entries
// per limb plane, which bounds every index into it) must stay below 2^30.  Returns 0, or sets the error and
// returns 2.
inline int check_planes(const char* what, std::initializer_list<int64_t> strides) {
  for (int64_t ps : strides)
    if (ps < 0 || ps >= (int64_t(1) << 30)) {
      set_error(std::string(what) + ": a limb plane of " + std::to_string(ps) +
                " entries exceeds the 2^30 entries that 32-bit plane offsets address (split the batch)");
      return 2;
    }
  return 0;
}

// The pole step's dot products on int8 MFMA (b3d_pole.hip): X entry (xr, u, j, k) at
// Xm[xr S SB + u SB + j B + k] (xr = xrow[task], or task when xrow is null), output (task, s0, n = u B + k) at
// Om[orow S SB + s0 SB + n], constant row ctab[task] S + s0 with the digit and correction tables of
// b3d_pole_tables and row exponents Ce.
int pole_mma(int L, const int32_t* Xm, const int64_t* Xe, int64_t xps, const int64_t* xrow, const int32_t* ctab,
             const int8_t* d8, const int32_t* corr, const int64_t* Ce, const int64_t* orow, int64_t T, int S, int64_t B,
             int32_t* Om, int64_t* Oe, int64_t ops, hipStream_t s);
// largest structure count S the MFMA pole step takes at L limbs (its operands and digit rows fit in LDS)
int pole_mma_max_s(int L);
}  // namespace b3d

// Workgroups are dealt to the 8 XCDs (each with its own L2) round robin by dispatch index.  This maps
// dispatch index b of a G-workgroup grid to a logical block so that each XCD runs one contiguous range of
// logical blocks: neighbouring blocks, which read the same data (the rows of one Cauchy group), share an L2.
// A bijection of [0, G).
__device__ __forceinline__ unsigned b3d_xcd_block(unsigned b, unsigned G) {
  const unsigned xcd = b & 7u, idx = b >> 3, per = G >> 3, rem = G & 7u;
  return xcd < rem ? xcd * (per + 1u) + idx : rem * (per + 1u) + (xcd - rem) * per + idx;
}

// Limb counts with compiled kernels.  L = 2 ceil(precision / 64) + 2 (precision rounded up to whole
// 64-bit limbs, plus >= 60 guard bits): every precision that is a multiple of 128 bits up to 1536 bits,
// plus the multiples of 64 bits the configured workloads use.
#define B3D_FOR_EACH_L(X) X(4) X(6) X(10) X(12) X(14) X(18) X(22) X(26) X(30) X(34) X(38) X(42) X(46) X(50) X(54)

#define B3D_CASE_L(Lv, ...) \
  case Lv: {                \
    constexpr int LL = Lv;  \
    __VA_ARGS__;            \
  } break;

// Expands to a switch over the compiled limb counts; the statement may use the constexpr LL.
#define B3D_DISPATCH_L(L, ...)                                                     \
  switch (L) {                                                                    \
    B3D_CASE_L(4, __VA_ARGS__)                                                           \
    B3D_CASE_L(6, __VA_ARGS__)                                                           \
    B3D_CASE_L(10, __VA_ARGS__)                                                          \
    B3D_CASE_L(12, __VA_ARGS__)                                                          \
    B3D_CASE_L(14, __VA_ARGS__)                                                          \
    B3D_CASE_L(18, __VA_ARGS__)                                                          \
    B3D_CASE_L(22, __VA_ARGS__)                                                          \
    B3D_CASE_L(26, __VA_ARGS__)                                                          \
    B3D_CASE_L(30, __VA_ARGS__)                                                          \
    B3D_CASE_L(34, __VA_ARGS__)                                                          \
    B3D_CASE_L(38, __VA_ARGS__)                                                          \
    B3D_CASE_L(42, __VA_ARGS__)                                                          \
    B3D_CASE_L(46, __VA_ARGS__)                                                          \
    B3D_CASE_L(50, __VA_ARGS__)                                                          \
    B3D_CASE_L(54, __VA_ARGS__)                                                          \
    default:                                                                      \
      b3d::set_error("unsupported limb count L=" + std::to_string(L) +            \
                     " (compiled: 4 6 10 12 14 18 22 26 30 34 38 42 46 50 54)");     \
      return 2;                                                                   \
  }

static inline unsigned b3d_blocks(int64_t n, int th) { return (unsigned)((n + th - 1) / th); }
