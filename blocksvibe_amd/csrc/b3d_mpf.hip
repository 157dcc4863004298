// libb3dhip: elementwise multiprecision kernels (normalise, add, multiply, row alignment) and the C-ABI
// error plumbing.  Arrays are limb planes: limb k of entry i at m[k * ps + i] (ps = plane stride),
// exponents e[i] int64.
#include "common.hpp"
#include "mp.hpp"
#include "../../include/b3dhip.h"

#include <mutex>
#include <string>

namespace b3d {
static thread_local std::string g_err;
void set_error(const std::string& msg) { g_err = msg; }
int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error(std::string(what) + ": " + hipGetErrorString(e));
    return 1;
  }
  return 0;
}
}  // namespace b3d

static inline unsigned grid_for(int64_t n, int th) {
  int64_t b = (n + th - 1) / th;
  if (b > (1 << 22)) b = 1 << 22;
  if (b < 1) b = 1;
  return (unsigned)b;
}

// ---------------------------------------------------------------------------------------------- kernels

// wide two's complement (LW <= 2L + 4 limbs, sign-extended) times 2^e_in -> normalised L limbs
template <int L>
__global__ __launch_bounds__(256) void k_norm_wide(const int32_t* __restrict__ W, i64 wps, int LW,
                                                   const i64* __restrict__ ein, int32_t* __restrict__ om,
                                                   i64* __restrict__ oe, i64 ops, i64 n) {
  constexpr int NW = 2 * L + 4;
  for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
    u32 acc[NW];
    u32 top = 0;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
      if (k < LW) {
        acc[k] = (u32)W[k * wps + i];
        top = acc[k];
      } else {
        acc[k] = (top >> 31) ? 0xFFFFFFFFu : 0u;
      }
    }
    u32 m[L];
    i64 e = mp::norm<L, NW>(acc, ein[i], m);
    mp::store(m, om, ops, i);
    oe[i] = e;
  }
}

template <int L>
__global__ __launch_bounds__(256) void k_add(const int32_t* __restrict__ am, const i64* __restrict__ ae, i64 aps,
                                             i64 sa, const int32_t* __restrict__ bm, const i64* __restrict__ be,
                                             i64 bps, i64 sb, int32_t* __restrict__ cm, i64* __restrict__ ce,
                                             i64 cps, i64 n) {
  for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
    u32 a[L], b[L], r[L];
    mp::load(a, am, aps, i * sa);
    mp::load(b, bm, bps, i * sb);
    i64 e = mp::add<L>(a, ae[i * sa], b, be[i * sb], r);
    mp::store(r, cm, cps, i);
    ce[i] = e;
  }
}

template <int L>
__global__ __launch_bounds__(256) void k_mul(const int32_t* __restrict__ am, const i64* __restrict__ ae, i64 aps,
                                             i64 sa, const int32_t* __restrict__ bm, const i64* __restrict__ be,
                                             i64 bps, i64 sb, int32_t* __restrict__ cm, i64* __restrict__ ce,
                                             i64 cps, i64 n) {
  for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
    u32 a[L], b[L];
    mp::load(a, am, aps, i * sa);
    mp::load(b, bm, bps, i * sb);
    u32 r[L];
    i64 e;
    if (a[L - 1] == 0u || b[L - 1] == 0u) {
      mp::zero(r);
      e = B3D_EZ;
    } else {
      e = mp::mul_norm<L>(a, b, ae[i * sa] + be[i * sb], r);
    }
    mp::store(r, cm, cps, i);
    ce[i] = e;
  }
}

// Row alignment: entry (r, k) of a (rows, K) array at index r * rs + k * ks.  Output: aligned mantissas
// (two's complement, limb planes, contiguous (rows, K)) and one exponent per row (EZ if the row is zero).
template <int L>
__global__ __launch_bounds__(256) void k_align_rows(const int32_t* __restrict__ m, const i64* __restrict__ e,
                                                    i64 ps, i64 rows, i64 K, i64 rs, i64 ks,
                                                    int32_t* __restrict__ om, i64 ops, i64* __restrict__ oe) {
  for (i64 t = blockIdx.x * (i64)blockDim.x + threadIdx.x; t < rows * K; t += (i64)gridDim.x * blockDim.x) {
    const i64 r = t / K, k = t % K;
    i64 E = B3D_EZ;
    for (i64 j = 0; j < K; ++j) {
      const i64 idx = r * rs + j * ks;
      if (mp::load_limb(m, ps, L - 1, idx) != 0u) {
        const i64 ej = e[idx];
        E = ej > E ? ej : E;
      }
    }
    u32 a[L];
    const i64 idx = r * rs + k * ks;
    mp::load(a, m, ps, idx);
    if (a[L - 1] != 0u) {
      const u32 na = mp::to_mag(a);
      const i64 sh = E - e[idx];
      mp::shr(a, sh > 32 * L ? 32 * L : (int)sh);
      mp::cneg(a, na);
    }
    mp::store(a, om, ops, t);
    if (k == 0) oe[r] = E;
  }
}

// ---------------------------------------------------------------------------------------------- C-ABI

// 28-bit magnitude limbs + sign mask of each entry (the pole step's constant operand form)
template <int L>
__global__ __launch_bounds__(256) void k_to28(const int32_t* __restrict__ m, i64 ps, i64 n, int32_t* __restrict__ o28,
                                              i64 ops, uint32_t* __restrict__ osg) {
  constexpr int NL = mp::R28<L>::NL;
  for (i64 i = blockIdx.x * (i64)blockDim.x + threadIdx.x; i < n; i += (i64)gridDim.x * blockDim.x) {
    u32 a[L], a28[NL];
    mp::load(a, m, ps, i);
    osg[i] = mp::split28<L>(a, a28);
    mp::store(a28, o28, ops, i);
  }
}

extern "C" {

const char* b3d_last_error(void) { return b3d::g_err.c_str(); }

int b3d_abi_version(void) { return B3D_ABI_VERSION; }

int b3d_clear_hip_error(void) { return (int)hipGetLastError(); }

int b3d_end_capture(void* stream) {
  hipStream_t s = (hipStream_t)stream;
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s, &st) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  if (st == hipStreamCaptureStatusNone) return 0;
  hipGraph_t g = nullptr;
  (void)hipStreamEndCapture(s, &g);
  if (g) (void)hipGraphDestroy(g);
  (void)hipGetLastError();
  return 1;
}

int b3d_thread_capture_mode(int mode) {
  if (mode < 0 || mode > 2) {
    b3d::set_error("b3d_thread_capture_mode: mode must be 0 (global), 1 (thread-local) or 2 (relaxed)");
    return -1;
  }
  hipStreamCaptureMode m = (hipStreamCaptureMode)mode;
  const hipError_t st = hipThreadExchangeStreamCaptureMode(&m);
  if (st != hipSuccess) {
    b3d::set_error(std::string("b3d_thread_capture_mode: ") + hipGetErrorString(st));
    return -1;
  }
  return (int)m;
}

int b3d_supported_limbs(int L) {
  switch (L) {
#define B3D_SUP(x) case x:
    B3D_FOR_EACH_L(B3D_SUP)
    return 1;
#undef B3D_SUP
    default:
      return 0;
  }
}

int b3d_mpf_norm_wide(int L, int LW, const int32_t* W, int64_t wps, const int64_t* ein, int32_t* om,
                      int64_t* oe, int64_t ops, int64_t n, void* stream) {
  if (n <= 0) return 0;
  if (LW < 1 || LW > 2 * L + 4) {
    b3d::set_error("b3d_mpf_norm_wide: need 1 <= LW <= 2L+4");
    return 2;
  }
  if (b3d::check_planes("b3d_mpf_norm_wide", {ops})) return 2;
  hipStream_t s = (hipStream_t)stream;
  B3D_DISPATCH_L(L, k_norm_wide<LL><<<dim3(grid_for(n, 256)), dim3(256), 0, s>>>(W, wps, LW, ein, om, oe, ops, n));
  return b3d::check_launch("b3d_mpf_norm_wide");
}

int b3d_limbs28(int L) { return b3d_supported_limbs(L) ? (32 * L - 2 + 27) / 28 : -1; }

int b3d_mpf_to28(int L, const int32_t* m, int64_t ps, int64_t n, int32_t* o28, int64_t ops, uint32_t* osg,
                 void* stream) {
  if (n <= 0) return 0;
  if (b3d::check_planes("b3d_mpf_to28", {ps, ops})) return 2;
  hipStream_t s = (hipStream_t)stream;
  B3D_DISPATCH_L(L, k_to28<LL><<<dim3(grid_for(n, 256)), dim3(256), 0, s>>>(m, ps, n, o28, ops, osg));
  return b3d::check_launch("b3d_mpf_to28");
}

int b3d_mpf_add(int L, const int32_t* am, const int64_t* ae, int64_t aps, int64_t sa, const int32_t* bm,
                const int64_t* be, int64_t bps, int64_t sb, int32_t* cm, int64_t* ce, int64_t cps, int64_t n,
                void* stream) {
  if (n <= 0) return 0;
  if (b3d::check_planes("b3d_mpf_add", {aps, bps, cps})) return 2;
  hipStream_t s = (hipStream_t)stream;
  B3D_DISPATCH_L(L, k_add<LL><<<dim3(grid_for(n, 256)), dim3(256), 0, s>>>(am, ae, aps, sa, bm, be, bps, sb, cm, ce, cps, n));
  return b3d::check_launch("b3d_mpf_add");
}

int b3d_mpf_mul(int L, const int32_t* am, const int64_t* ae, int64_t aps, int64_t sa, const int32_t* bm,
                const int64_t* be, int64_t bps, int64_t sb, int32_t* cm, int64_t* ce, int64_t cps, int64_t n,
                void* stream) {
  if (n <= 0) return 0;
  if (b3d::check_planes("b3d_mpf_mul", {aps, bps, cps})) return 2;
  hipStream_t s = (hipStream_t)stream;
  B3D_DISPATCH_L(L, k_mul<LL><<<dim3(grid_for(n, 256)), dim3(256), 0, s>>>(am, ae, aps, sa, bm, be, bps, sb, cm, ce, cps, n));
  return b3d::check_launch("b3d_mpf_mul");
}

int b3d_mpf_align_rows(int L, const int32_t* m, const int64_t* e, int64_t ps, int64_t rows, int64_t K,
                       int64_t rs, int64_t ks, int32_t* om, int64_t ops, int64_t* oe, void* stream) {
  if (rows <= 0 || K <= 0) return 0;
  if (b3d::check_planes("b3d_mpf_align_rows", {ps, ops})) return 2;
  hipStream_t s = (hipStream_t)stream;
  B3D_DISPATCH_L(L, k_align_rows<LL><<<dim3(grid_for(rows * K, 256)), dim3(256), 0, s>>>(m, e, ps, rows, K, rs, ks, om, ops, oe));
  return b3d::check_launch("b3d_mpf_align_rows");
}

}  // extern "C"
