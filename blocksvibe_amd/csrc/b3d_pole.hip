// libb3dhip: the pole step's small dot products on int8 MFMA (reference blocksvibe/smallgemm.py:14-88, the
// two products of recursion_gpu.py:481-503).
//
// One output is out = norm(sum_j a_j X_j) over S terms (the pole step's rounding, as mp::small_dot):
// a_j the constant row cr's pre-aligned mantissas (two's complement, common exponent ea), X_j the entries
// x_j aligned over j (E = largest exponent, |x_j| >> (E - e_j) truncated, the sign of x_j put back).  The
// exact integer sum is formed in radix 2^8:
//   X_j as 4L bytes of two's complement (bytes 0 .. 4L-2 unsigned, the top byte signed),
//   a_j as 4L + 1 balanced digits d in [-128, 127] (b3d_pole_tables),
//   column c = sum_j sum_i x_j[i] d_j[c - i]  (c < 8L), an int8 GEMM against the Toeplitz matrix of d_j,
// with the unsigned bytes fed to the MFMA as u ^ 0x80 = u - 128 and the difference added back per column
// (corr[c] = 128 sum_j sum_{i <= 4L-2} d_j[c - i], a property of the constant row alone).  Every column
// sum is an exact int32 (|c| < S 4L 2^14 + |corr|), and carrying the columns into 32-bit limbs gives the
// same 2L + 2 limb integer as the VALU path, so the rounding is identical.  The accumulators start from the
// corrections.
//
// Work item = (constant group (task, s0), block of 64 rows): the rows share the constant row, so one
// workgroup stages the rows' aligned bytes in LDS (A operand, 16-byte chunks; up to six operands at a
// time, in the space the column sums take afterwards), the constant digits
// reversed in four byte-shifted copies (so that every B fragment d_j[c - i0 - 0..15] is four aligned LDS
// dwords), and runs v_mfma_i32_16x16x64_i8 over 4 row tiles x (8L / 16) column tiles (column tiles dealt
// round robin to the four waves).  Column sums go through LDS to a carry pass (four segments per row in
// parallel) and a per-row assembly + norm.
#include "common.hpp"
#include "mp.hpp"
#include "../../include/b3dhip.h"

typedef int v4i __attribute__((ext_vector_type(4)));

template <int L>
struct PoleMma {
  static constexpr int CPJ = (4 * L + 15) / 16;              // 16-byte chunks per operand (4L bytes)
  static constexpr int NCOL = 8 * L;                         // product columns
  static constexpr int NCT = NCOL / 16;                      // column tiles (L even)
  static constexpr int MT = (NCT + 3) / 4;                   // column tiles per wave
  static constexpr int P0 = NCOL - 1;                        // reversed digits: R[p] = d[P0 - p]
  static constexpr int RL4 = (NCOL + 16 * CPJ - 1 + 3) / 4 * 4;  // bytes per reversed copy
  static constexpr int DS = 4 * L + 4;                       // table bytes per constant entry
  static constexpr int CS = NCOL + 1;                        // LDS dwords per row of column sums (odd)
  static constexpr int LC = 64 * CS * 4;                     // column sums of 64 rows
  // operands staged per batch: as many as fit in the column-sum area (6 for every compiled L), so that
  // with S <= JB the staging is done before any accumulator is live
  static constexpr int JB0 = (LC - 64 * 16) / (64 * 16 * CPJ + 4 * RL4);
  static constexpr int JB = JB0 < 1 ? 1 : (JB0 > 16 ? 16 : JB0);
  static constexpr int XROW = 16 * CPJ * JB + 16;            // LDS bytes per row: JB operands + pad
  static constexpr int LX = 64 * XROW + 4 * JB * RL4 + JB * DS;  // staging: rows, 4 copies and the digits of
                                                                  // JB operands
  static constexpr int LDS = LX > LC ? LX : LC;
  static constexpr int SEGL = L / 2;                         // limbs per carry segment (4 segments)
};

// balanced base-256 digits of a two's complement constant (4L + 1 digits, then zero padding) and the
// per-row column corrections
template <int L>
__global__ __launch_bounds__(256) void k_sdigits(const int32_t* __restrict__ m, i64 ps, i64 n,
                                                 int8_t* __restrict__ d8) {
  using P = PoleMma<L>;
  for (i64 t = blockIdx.x * (i64)blockDim.x + threadIdx.x; t < n; t += (i64)gridDim.x * blockDim.x) {
    u32 x[L];
    mp::load(x, m, ps, t);
    int8_t* o = d8 + t * P::DS;
    int carry = 0;
#pragma unroll
    for (int i = 0; i < 4 * L - 1; ++i) {
      const int u = (int)((x[i >> 2] >> (8 * (i & 3))) & 0xFFu) + carry;
      carry = u >= 128 ? 1 : 0;
      o[i] = (int8_t)(u - 256 * carry);
    }
    const int top = (int)(int8_t)(x[L - 1] >> 24) + carry;   // in [-128, 128]
    o[4 * L - 1] = (int8_t)(top == 128 ? -128 : top);
    o[4 * L] = (int8_t)(top == 128 ? 1 : 0);
    for (int i = 4 * L + 1; i < P::DS; ++i) o[i] = 0;
  }
}

template <int L>
__global__ __launch_bounds__(256) void k_scorr(const int8_t* __restrict__ d8, int S, i64 nrows,
                                               int32_t* __restrict__ corr) {
  using P = PoleMma<L>;
  for (i64 t = blockIdx.x * (i64)blockDim.x + threadIdx.x; t < nrows * P::NCOL; t += (i64)gridDim.x * blockDim.x) {
    const i64 cr = t / P::NCOL;
    const int c = (int)(t % P::NCOL);
    int s = 0;
    for (int j = 0; j < S; ++j) {
      const int8_t* d = d8 + (cr * S + j) * P::DS;
      const int lo = c - 4 * L > 0 ? c - 4 * L : 0, hi = c < 4 * L - 2 ? c : 4 * L - 2;
      for (int i = lo; i <= hi; ++i) s += d[c - i];
    }
    corr[t] = 128 * s;
  }
}

// X entry (xr, u, j, k) = Xm[xr S SB + u SB + j B + k]; output (orow, s0, n) = Om[orow S SB + s0 SB + n],
// n = u B + k < SB = S B; constant row cr = ctab[task] S + s0, digits d8[(cr S + j) DS ..], corr[cr NCOL ..],
// exponent Ce[cr].  xrow / orow null: the task index itself.  ONE: S <= JB, the operands staged once.
// Two waves per SIMD (the LDS of two workgroups per CU): the compiler keeps the accumulators in VGPRs
// within that budget rather than in AGPRs next to the epilogue's registers.
template <int L, bool ONE>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) void k_pole_mma(const int32_t* __restrict__ Xm, const i64* __restrict__ Xe, i64 xps,
                                                  const int64_t* __restrict__ xrow, const int32_t* __restrict__ ctab,
                                                  const int8_t* __restrict__ d8, const int32_t* __restrict__ corr,
                                                  const i64* __restrict__ Ce, const int64_t* __restrict__ orow,
                                                  i64 T, int S, i64 B, int32_t* __restrict__ Om,
                                                  i64* __restrict__ Oe, i64 ops) {
  using P = PoleMma<L>;
  __shared__ __align__(16) unsigned char lds[P::LDS];
  __shared__ i64 cseg[4][64];
  unsigned char* const X = lds;                     // [64][XROW]
  unsigned char* const Dc = lds + 64 * P::XROW;     // [JB operands][4 copies][RL4]
  int8_t* const Dr = (int8_t*)(Dc + 4 * P::JB * P::RL4);  // [JB operands][DS] digits as in the table
  int32_t* const C = (int32_t*)lds;                 // [64][CS], after the products
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int r = lane & 15, g = lane >> 4;
  const i64 SB = (i64)S * B;
  const i64 NRB = (SB + 63) / 64;
  const i64 items = T * S * NRB;
  for (i64 it = blockIdx.x; it < items; it += gridDim.x) {
    const i64 grp = it / NRB, rb = it - grp * NRB;
    const i64 task = grp / S;
    const int s0 = (int)(grp - task * S);
    const i64 cr = (i64)ctab[task] * S + s0;
    const i64 xb = (xrow ? xrow[task] : task) * S * SB;
    // this thread's row (each wave covers the 64 rows and stages every fourth operand)
    const i64 n = rb * 64 + lane;
    const bool valid = n < SB;
    const i64 u = valid ? n / B : 0, k = valid ? n - (n / B) * B : 0;
    const i64 xrowbase = xb + u * SB + k;
    i64 E = B3D_EZ;   // zero entries carry EZ
#pragma unroll
    for (int j = 0; j < 16; ++j)   // S <= 16: every exponent load in flight at once
      if (j < S) {
        const i64 ej = valid ? Xe[xrowbase + j * B] : B3D_EZ;
        E = ej > E ? ej : E;
      }
    // column corrections: the accumulators start from them (the same for every row)
    int32_t cinit[P::MT];
#pragma unroll
    for (int mt = 0; mt < P::MT; ++mt) {
      const int ct = w + 4 * mt < P::NCT ? w + 4 * mt : P::NCT - 1;
      cinit[mt] = corr[cr * P::NCOL + 16 * ct + r];
    }
    const i64 ea = Ce[cr];
    __syncthreads();   // the previous item's epilogue is done with LDS
    v4i acc[4][P::MT];
    for (int jb = 0; ONE ? jb < 1 : jb * P::JB < S; ++jb) {
      const int nb = S - P::JB * jb < P::JB ? S - P::JB * jb : P::JB;
      if (jb > 0) __syncthreads();   // the previous batch's fragments are read
      // A: row `lane`, operands jl = w, w + 4, .. of the batch: bytes u ^ 0x80 except the signed top byte,
      // zero padding
      for (int jl = w; jl < nb; jl += 4) {
        const int j = P::JB * jb + jl;
        u32* row = (u32*)(X + lane * P::XROW + jl * 16 * P::CPJ);
        {
          u32 x[L];
          mp::zero(x);
          if (valid && E != B3D_EZ) {
            const i64 idx = xrowbase + (i64)j * B;
            mp::load(x, Xm, xps, idx);
            const i64 sh = E - Xe[idx];
            if (x[L - 1] != 0u && sh < 32 * L) {
              const u32 sm = mp::to_mag(x);
              mp::shr(x, (int)sh);
              mp::cneg(x, sm);
            } else {
              mp::zero(x);
            }
          }
#pragma unroll
          for (int q = 0; q < 4 * P::CPJ; ++q)
            row[q] = q < L - 1 ? (x[q] ^ 0x80808080u) : (q == L - 1 ? (x[q] ^ 0x00808080u) : 0u);
        }
      }
      // the batch's digit rows (contiguous in the table) into LDS, one dword per thread
      {
        const u32* src = (const u32*)(d8 + (cr * S + P::JB * jb) * P::DS);
        for (int q = tid; q < nb * (P::DS / 4); q += 256) ((u32*)Dr)[q] = src[q];
      }
      __syncthreads();
      // B source: operand jl's digits reversed, four copies shifted by 0..3 bytes:
      // copy s, byte p' = d[P0 - p' - s] (zero outside 0 .. 4L)
      for (int q = tid; q < nb * 4 * (P::RL4 / 4); q += 256) {
        const int jl = q / (4 * (P::RL4 / 4));
        const int rem = q - jl * 4 * (P::RL4 / 4);
        const int s = rem / (P::RL4 / 4), wd = rem - s * (P::RL4 / 4);
        const int8_t* d = Dr + jl * P::DS;
        u32 v = 0u;
#pragma unroll
        for (int b = 0; b < 4; ++b) {
          const int di = P::P0 - 4 * wd - s - b;
          const u32 byte = (di >= 0 && di <= 4 * L) ? (u32)(uint8_t)d[di] : 0u;
          v |= byte << (8 * b);
        }
        ((u32*)(Dc + (jl * 4 + s) * P::RL4))[wd] = v;
      }
      __syncthreads();
      if (jb == 0) {   // (ONE: no accumulator is live while the operands are staged)
#pragma unroll
        for (int rt = 0; rt < 4; ++rt)
#pragma unroll
          for (int mt = 0; mt < P::MT; ++mt) acc[rt][mt] = v4i{cinit[mt], cinit[mt], cinit[mt], cinit[mt]};
      }
      const int nch = nb * P::CPJ;
      const int nks = (nch + 3) / 4;
      for (int ks = 0; ks < nks; ++ks) {
        const int q = 4 * ks + g;   // this lane's chunk: operand q / CPJ, bytes 16 (q % CPJ) ..
        v4i A[4];
#pragma unroll
        for (int rt = 0; rt < 4; ++rt) A[rt] = *(const v4i*)(X + (16 * rt + r) * P::XROW + 16 * q);
        const int jl = q / P::CPJ;
        const int i0 = 16 * (q - jl * P::CPJ);
        const bool live = jl < nb;                  // padding chunks past the batch: B fragment 0
        const int jc = live ? jl : nb - 1;
#pragma unroll
        for (int mt = 0; mt < P::MT; ++mt) {
          // column tile (the waves with one tile less repeat their last one; never written back)
          const int ct = w + 4 * mt < P::NCT ? w + 4 * mt : P::NCT - 1;
          const int o = P::P0 - (16 * ct + r) + i0;   // R[o + b] = d[c - i0 - b]
          const int sft = o & 3;
          const u32* src = (const u32*)(Dc + (jc * 4 + sft) * P::RL4 + (o - sft));
          const v4i Bf = v4i{live ? (int)src[0] : 0, live ? (int)src[1] : 0, live ? (int)src[2] : 0,
                             live ? (int)src[3] : 0};
#pragma unroll
          for (int rt = 0; rt < 4; ++rt) acc[rt][mt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[rt], Bf, acc[rt][mt], 0, 0, 0);
        }
      }
    }
    __syncthreads();   // products done: the column sums overwrite the staging area
#pragma unroll
    for (int rt = 0; rt < 4; ++rt)
#pragma unroll
      for (int mt = 0; mt < P::MT; ++mt) {
        const int ct = w + 4 * mt;
        if (ct < P::NCT)
#pragma unroll
          for (int i = 0; i < 4; ++i) C[(16 * rt + 4 * g + i) * P::CS + 16 * ct + r] = acc[rt][mt][i];
        __builtin_amdgcn_sched_barrier(0);   // tile by tile: the accumulators are not all read out at once
      }
    __syncthreads();
    // carry pass: row `lane`, segment w = columns [2L w, 2L w + 2L) -> L/2 limbs written in place + a carry
    if (valid) {
      int32_t* cr_row = C + lane * P::CS + 2 * L * w;
      i64 carry = 0;
#pragma unroll 2
      for (int q = 0; q < P::SEGL; ++q) {
        i64 v = carry;
#pragma unroll
        for (int b = 0; b < 4; ++b) v += (i64)cr_row[4 * q + b] * ((i64)1 << (8 * b));
        cr_row[q] = (int32_t)(u32)v;
        carry = v >> 32;
      }
      cseg[w][lane] = carry;
    }
    __syncthreads();
    // assembly and rounding: one thread per row
    if (w == 0 && valid) {
      u32 W[2 * L + 2];
#pragma unroll
      for (int s = 0; s < 4; ++s)
#pragma unroll
        for (int q = 0; q < P::SEGL; ++q) W[s * P::SEGL + q] = (u32)C[lane * P::CS + 2 * L * s + q];
      const i64 c3 = cseg[3][lane];
      W[2 * L] = (u32)c3;
      W[2 * L + 1] = (u32)((u64)c3 >> 32);
#pragma unroll
      for (int s = 0; s < 3; ++s) {
        const i64 cs = cseg[s][lane];
        const u32 ext = cs < 0 ? 0xFFFFFFFFu : 0u;
        u32 cy = 0u;
#pragma unroll
        for (int q = (s + 1) * P::SEGL; q < 2 * L + 2; ++q)
          W[q] = mp::addc(W[q], q == (s + 1) * P::SEGL ? (u32)cs : (q == (s + 1) * P::SEGL + 1 ? (u32)((u64)cs >> 32) : ext), cy);
      }
      u32 out[L];
      i64 e;
      if (E == B3D_EZ || ea == B3D_EZ) {
        mp::zero(out);
        e = B3D_EZ;
      } else {
        e = mp::norm<L, 2 * L + 2>(W, ea + E, out);
      }
      const i64 o = (orow ? orow[task] : task) * S * SB + s0 * SB + n;
      mp::store(out, Om, ops, o);
      Oe[o] = e;
    }
  }
}

int b3d::pole_mma(int L, const int32_t* Xm, const int64_t* Xe, int64_t xps, const int64_t* xrow, const int32_t* ctab,
                  const int8_t* d8, const int32_t* corr, const int64_t* Ce, const int64_t* orow, int64_t T, int S,
                  int64_t B, int32_t* Om, int64_t* Oe, int64_t ops, hipStream_t s) {
  const int64_t items = T * S * ((S * B + 63) / 64);
  const unsigned grid = (unsigned)(items < (1 << 22) ? items : (1 << 22));
  B3D_DISPATCH_L(L, if (S <= PoleMma<LL>::JB) k_pole_mma<LL, true><<<dim3(grid), dim3(256), 0, s>>>(Xm, Xe, xps, xrow, ctab, d8, corr, Ce, orow, T, S, B, Om, Oe, ops);
                    else k_pole_mma<LL, false><<<dim3(grid), dim3(256), 0, s>>>(Xm, Xe, xps, xrow, ctab, d8, corr, Ce, orow, T, S, B, Om, Oe, ops));
  return b3d::check_launch("b3d_pole_level (MFMA)");
}

extern "C" {

int b3d_pole_digits_bytes(int L) { return b3d_supported_limbs(L) ? 4 * L + 4 : -1; }

int b3d_pole_tables(int L, int S, const int32_t* m, int64_t ps, int64_t nrows, int8_t* d8, int32_t* corr,
                    void* stream) {
  if (nrows <= 0) return 0;
  if (S < 1 || S > 16) {
    b3d::set_error("b3d_pole_tables: structure count S must be 1..16");
    return 2;
  }
  if (b3d::check_planes("b3d_pole_tables", {ps})) return 2;
  hipStream_t s = (hipStream_t)stream;
  const int64_t n = nrows * S;
  const unsigned gn = (unsigned)((n + 255) / 256 < (1 << 22) ? (n + 255) / 256 : (1 << 22));
  B3D_DISPATCH_L(L, k_sdigits<LL><<<dim3(gn), dim3(256), 0, s>>>(m, ps, n, d8));
  if (b3d::check_launch("b3d_pole_tables (digits)")) return 1;
  const int64_t nc = nrows * 8 * L;
  const unsigned gc = (unsigned)((nc + 255) / 256 < (1 << 22) ? (nc + 255) / 256 : (1 << 22));
  B3D_DISPATCH_L(L, k_scorr<LL><<<dim3(gc), dim3(256), 0, s>>>(d8, S, nrows, corr));
  return b3d::check_launch("b3d_pole_tables (corrections)");
}

}  // extern "C"
