// libb3dhip: one level of the batched radial recursion (reference blocksvibe/recursion_gpu.py:438-510,
// cauchy.py:15-87, smallgemm.py:14-88), as three gfx950 kernels:
//
//   k_cauchy   H[h, n] = h_inf[g(h), n] + sum_i x_i * 2 / (D'_h - D_i), x_i = res[col[g, i], n]:
//              gathers the prime-pole residues straight from the slot buffer (no staging copy), aligns
//              them per column, divides by the one-limb integer D' - D_i, sums exactly, rounds once and
//              adds the h_inf entry.  k_cauchy_groups with a wave table (a group's rows over consecutive
//              lanes, the group's data in scalar registers) or k_cauchy_pairs (the same with the rows in pairs,
//              one lane both rows of a pair); without one, k_cauchy_rows when the rows,
//              padded to whole waves, keep at least 3/4 of the lanes busy, k_cauchy otherwise;
//   k_pole_r   Y[t, b0, a, k] = sum_{b < SR} Ralign[R_t][b0][b] * H[hrow_t][a, b][k]
//   k_pole_l   res[slot_t][a0, b0][k] = sum_{a < SL} QLalign[qL_t][a0][a] * Y[t, b0, a, k]
//
// Layouts (limb planes, entries innermost = right-hand side k), SL x SR structure entries per residue (left
// a < SL, right b < SR), S >= SL, SR the row stride of the constant tables:
//   slot buffer res: (L, nslots, SL*SR, B), H: (L, nH, SL*SR, B), Y: (L, T, SR, SL, B),
//   R / qL constants: aligned per row, as 28-bit magnitude limb planes (NL, ntab, S, S) + sign masks
//   (ntab, S, S) (b3d_mpf_to28), (ntab, S) row exponents.
#include "common.hpp"
#include "mp.hpp"
#include "../../include/b3dhip.h"

static inline unsigned grid_for(int64_t n, int th) {
  int64_t b = (n + th - 1) / th;
  if (b > (1 << 22)) b = 1 << 22;
  if (b < 1) b = 1;
  return (unsigned)b;
}

// Row-uniform form: each row's N columns are laid over W >= N lanes, W a multiple of 64 (lanes n >= N idle),
// so every wave lies in one row h and takes the group's column list, divisors and plane bases from scalar
// registers.  Up to L = 34 that leaves at most 96 VGPRs, five waves per SIMD (the compiler is asked for that
// occupancy; the term loop waits on its limb loads, so more waves hide more of that latency).
template <int L>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(L <= 34 ? 5 : 1))) void k_cauchy_rows(
    const int32_t* __restrict__ rm, const i64* __restrict__ re, i64 rps, i64 N, const int32_t* __restrict__ hm,
    const i64* __restrict__ he, i64 hps, const int64_t* __restrict__ ghrow, const int32_t* __restrict__ rg,
    const int32_t* __restrict__ rd, i64 nH, const int64_t* __restrict__ col, const int32_t* __restrict__ cold,
    const int32_t* __restrict__ ncol, int Kc, int32_t* __restrict__ om, i64* __restrict__ oe, i64 ops, i64 W) {
  const unsigned lb = b3d_xcd_block(blockIdx.x, gridDim.x);   // a group's rows stay on one XCD's L2
  for (i64 u = lb * (i64)blockDim.x + threadIdx.x; u < nH * W; u += (i64)gridDim.x * blockDim.x) {
    const i64 n = u % W;
    const int h = __builtin_amdgcn_readfirstlane((int)(u / W));
    if (n >= N) continue;
    const int g = rg[h];
    u32 r[L];
    const i64 er = mp::cauchy_entry<L, true>(rm, re, rps, N, n, col + (i64)g * Kc, cold + (i64)g * Kc, ncol[g],
                                             rd[h], hm, he, hps, ghrow[g], r);
    const i64 t = (i64)h * N + n;
    mp::store(r, om, ops, t);
    oe[t] = er;
  }
}

// Group-uniform form: the rows of one group (which gather the same residue columns from the same h_inf row)
// are laid out column by column over consecutive lanes, a wave at a time, padded only at the end of the group:
// a wave may span several rows of its group, so only the row's position D' (one VGPR) varies across its
// lanes, while the group's column list, divisor positions and plane bases are scalars.  Wave w covers
// entries woff[w] + lane of the group whose rows start at wrow[w] (wcnt[w] rows, N columns each).
template <int L>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(L <= 34 ? 5 : 1))) void k_cauchy_groups(
    const int32_t* __restrict__ rm, const i64* __restrict__ re, i64 rps, i64 N, const int32_t* __restrict__ hm,
    const i64* __restrict__ he, i64 hps, const int64_t* __restrict__ ghrow, const int32_t* __restrict__ rg,
    const int32_t* __restrict__ rd, const int64_t* __restrict__ col, const int32_t* __restrict__ cold,
    const int32_t* __restrict__ ncol, int Kc, int32_t* __restrict__ om, i64* __restrict__ oe, i64 ops,
    const int32_t* __restrict__ wrow, const int32_t* __restrict__ wcnt, const int32_t* __restrict__ woff, i64 nw) {
  const unsigned lb = b3d_xcd_block(blockIdx.x, gridDim.x);   // a group's waves stay on one XCD's L2
  const i64 w = (i64)lb * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (w >= nw) return;
  const int h0 = __builtin_amdgcn_readfirstlane(wrow[w]);
  const int cnt = __builtin_amdgcn_readfirstlane(wcnt[w]);
  const i64 o = (i64)__builtin_amdgcn_readfirstlane(woff[w]) + (threadIdx.x & 63);
  if (o >= (i64)cnt * N) return;
  const int g = __builtin_amdgcn_readfirstlane(rg[h0]);
  const i64 h = h0 + o / N, n = o % N;
  u32 r[L];
  const i64 er = mp::cauchy_entry<L, true>(rm, re, rps, N, n, col + (i64)g * Kc, cold + (i64)g * Kc, ncol[g], rd[h],
                                           hm, he, hps, ghrow[g], r);
  const i64 t = h * N + n;
  mp::store(r, om, ops, t);
  oe[t] = er;
}

// Group-uniform form, two rows per lane: the rows of one group are taken in pairs, and the pairs laid out column by column over consecutive lanes, a wave at a time, padded
// only at the end of the group: a wave may span several pairs of its group, so only the rows' positions D'
// (VGPRs) vary across its lanes, while the group's column list, divisor positions and plane bases are
// scalars.  A lane computes both rows of its pair at its column (mp::cauchy_entry2): each term is loaded and
// put on the grid once for the two, only the division and the sum are per row.  Wave w covers pair entries
// woff[w] + lane of the group whose rows start at wrow[w] (wcnt[w] rows, ceil(wcnt[w] / 2) pairs, N columns
// each).  The two accumulators take the kernel to 168 VGPRs up to L = 34, three waves per SIMD (asked for:
// at four the compiler spills): it does a third less work per entry than k_cauchy_groups (a term's load and
// shift are about a third of its cost, the division by D - D_i the rest) with fewer waves in flight, so it
// pays on levels with many waves to run (15% on the heaviest nmax = 6 group) and not on small ones.
template <int L>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(L <= 34 ? 3 : 1))) void k_cauchy_pairs(
    const int32_t* __restrict__ rm, const i64* __restrict__ re, i64 rps, i64 N, const int32_t* __restrict__ hm,
    const i64* __restrict__ he, i64 hps, const int64_t* __restrict__ ghrow, const int32_t* __restrict__ rg,
    const int32_t* __restrict__ rd, const int64_t* __restrict__ col, const int32_t* __restrict__ cold,
    const int32_t* __restrict__ ncol, int Kc, int32_t* __restrict__ om, i64* __restrict__ oe, i64 ops,
    const int32_t* __restrict__ wrow, const int32_t* __restrict__ wcnt, const int32_t* __restrict__ woff, i64 nw) {
  const unsigned lb = b3d_xcd_block(blockIdx.x, gridDim.x);   // a group's waves stay on one XCD's L2
  const i64 w = (i64)lb * (blockDim.x >> 6) + (threadIdx.x >> 6);
  if (w >= nw) return;
  const int h0 = __builtin_amdgcn_readfirstlane(wrow[w]);
  const int cnt = __builtin_amdgcn_readfirstlane(wcnt[w]);
  const i64 o = (i64)__builtin_amdgcn_readfirstlane(woff[w]) + (threadIdx.x & 63);
  if (o >= (i64)((cnt + 1) / 2) * N) return;
  const int g = __builtin_amdgcn_readfirstlane(rg[h0]);
  const i64 p = o / N, n = o % N;
  const i64 ha = h0 + 2 * p;
  const bool has2 = 2 * p + 1 < cnt;
  u32 r[L];
  i64 e1, e2;
  mp::cauchy_entry2<L>(rm, re, rps, N, n, col + (i64)g * Kc, cold + (i64)g * Kc, ncol[g], rd[ha],
                       has2 ? rd[ha + 1] : 0, has2, hm, he, hps, ghrow[g], r, e1, e2, om, ops, ha * N + n,
                       (ha + 1) * N + n);
  oe[ha * N + n] = e1;
  if (has2) oe[(ha + 1) * N + n] = e2;
}

// General form (rows that would leave many lanes idle if padded to whole waves): every index per lane.
template <int L>
__global__ __launch_bounds__(256) void k_cauchy(const int32_t* __restrict__ rm, const i64* __restrict__ re, i64 rps,
                                                i64 N, const int32_t* __restrict__ hm, const i64* __restrict__ he,
                                                i64 hps, const int64_t* __restrict__ ghrow,
                                                const int32_t* __restrict__ rg, const int32_t* __restrict__ rd, i64 nH,
                                                const int64_t* __restrict__ col, const int32_t* __restrict__ cold,
                                                const int32_t* __restrict__ ncol, int Kc, int32_t* __restrict__ om,
                                                i64* __restrict__ oe, i64 ops) {
  const unsigned lb = b3d_xcd_block(blockIdx.x, gridDim.x);   // a group's rows stay on one XCD's L2
  for (i64 t = lb * (i64)blockDim.x + threadIdx.x; t < nH * N; t += (i64)gridDim.x * blockDim.x) {
    const i64 n = t % N, h = t / N;
    const int g = rg[h];
    u32 r[L];
    const i64 er = mp::cauchy_entry<L, false>(rm, re, rps, N, n, col + (i64)g * Kc, cold + (i64)g * Kc, ncol[g],
                                              rd[h], hm, he, hps, ghrow[g], r);
    mp::store(r, om, ops, t);
    oe[t] = er;
  }
}

// The pole step's two products, a workgroup per block of columns: a column is the SR (R product) or SL (q L
// product) X entries one output row combines; every output of the column shares them, so the workgroup
// aligns each column once (exponent from its nonzero terms), converts each term once into 28-bit limbs in
// LDS (thread (column, s) stages term s), and every thread then forms its output's dot product from LDS
// (mp::small_dot_staged, the same bits as mp::small_dot).  Threads: s fastest within a column.  kofs (may be
// null): per right-hand side k an offset added to the task's constant row, so that right-hand sides of
// different operators (their tables stacked into one) share the launch.
template <int L>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(L <= 34 ? 2 : 1))) void k_pole_r(const int32_t* __restrict__ Hm, const i64* __restrict__ He, i64 hps,
                                                const int64_t* __restrict__ thrw, const int32_t* __restrict__ tR,
                                                i64 T, int SL, int SR, int S, i64 B, const int32_t* __restrict__ R28,
                                                const uint32_t* __restrict__ Rsg, const i64* __restrict__ Re,
                                                i64 rps, const int32_t* __restrict__ kofs, int32_t* __restrict__ Ym,
                                                i64* __restrict__ Ye, i64 yps) {
  constexpr int NL = mp::R28<L>::NL;
  extern __shared__ int32_t st[];
  const int CPB = 256 / SR;                   // columns (task, a, k) per block, SR outputs b0 each
  const i64 ncol = T * SL * B, NE = (i64)SR * SL * B;
  const int lc = threadIdx.x / SR, b0 = threadIdx.x % SR;
  const unsigned lb = b3d_xcd_block(blockIdx.x, gridDim.x);   // a task's outputs stay on one XCD's L2
  for (i64 c0 = (i64)lb * CPB; c0 < ncol; c0 += (i64)gridDim.x * CPB) {
    const i64 col = c0 + lc;
    const bool act = lc < CPB && col < ncol;
    i64 task = 0, a = 0, k = 0, xbase = 0, E = B3D_EZ;
    if (act) {
      k = col % B;
      a = (col / B) % SL;
      task = col / ((i64)SL * B);
      xbase = thrw[task] * NE + a * SR * B + k;   // H[hrow][a*SR + b][k], b stride B
      for (int j = 0; j < SR; ++j) {
        const i64 idx = xbase + j * B;
        if (mp::load_limb(Hm, hps, L - 1, idx) != 0u) {
          const i64 ej = He[idx];
          E = ej > E ? ej : E;
        }
      }
      mp::stage_term<L>(Hm, He, hps, xbase + b0 * B, E, st + threadIdx.x * (NL + 1));
    }
    __syncthreads();
    if (act) {
      const i64 r = tR[task] + (kofs ? kofs[k] : 0);
      u32 out[L];
      const i64 e = mp::small_dot_staged<L>(R28, Rsg, rps, (r * S + b0) * S, Re[r * S + b0],
                                            st + lc * SR * (NL + 1), SR, E, out);
      const i64 t = ((task * SR + b0) * SL + a) * B + k;   // Y (task, b0, a, k)
      mp::store(out, Ym, yps, t);
      Ye[t] = e;
    }
    __syncthreads();
  }
}

template <int L>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(L <= 34 ? 2 : 1))) void k_pole_l(const int32_t* __restrict__ Ym, const i64* __restrict__ Ye, i64 yps,
                                                const int32_t* __restrict__ tQL, const int64_t* __restrict__ tslot,
                                                i64 T, int SL, int SR, int S, i64 B, const int32_t* __restrict__ Q28,
                                                const uint32_t* __restrict__ Qsg, const i64* __restrict__ Qe, i64 qps,
                                                const int32_t* __restrict__ kofs, int32_t* __restrict__ rm,
                                                i64* __restrict__ re, i64 rps) {
  constexpr int NL = mp::R28<L>::NL;
  extern __shared__ int32_t st[];
  const int CPB = 256 / SL;                   // columns (task, b0, k) per block, SL outputs a0 each
  const i64 ncol = T * SR * B, NE = (i64)SL * SR * B;
  const int lc = threadIdx.x / SL, a0 = threadIdx.x % SL;
  const unsigned lb = b3d_xcd_block(blockIdx.x, gridDim.x);
  for (i64 c0 = (i64)lb * CPB; c0 < ncol; c0 += (i64)gridDim.x * CPB) {
    const i64 col = c0 + lc;
    const bool act = lc < CPB && col < ncol;
    i64 task = 0, b0 = 0, k = 0, xbase = 0, E = B3D_EZ;
    if (act) {
      k = col % B;
      b0 = (col / B) % SR;
      task = col / ((i64)SR * B);
      xbase = (task * SR + b0) * SL * B + k;   // Y[task][b0][a][k], a stride B
      for (int j = 0; j < SL; ++j) {
        const i64 idx = xbase + j * B;
        if (mp::load_limb(Ym, yps, L - 1, idx) != 0u) {
          const i64 ej = Ye[idx];
          E = ej > E ? ej : E;
        }
      }
      mp::stage_term<L>(Ym, Ye, yps, xbase + a0 * B, E, st + threadIdx.x * (NL + 1));
    }
    __syncthreads();
    if (act) {
      const i64 q = tQL[task] + (kofs ? kofs[k] : 0);
      u32 out[L];
      const i64 e = mp::small_dot_staged<L>(Q28, Qsg, qps, (q * S + a0) * S, Qe[q * S + a0],
                                            st + lc * SL * (NL + 1), SL, E, out);
      const i64 o = tslot[task] * NE + (a0 * SR + b0) * B + k;
      mp::store(out, rm, rps, o);
      re[o] = e;
    }
    __syncthreads();
  }
}

extern "C" {

int b3d_cauchy_level(int L, const int32_t* rm, const int64_t* re, int64_t rps, int64_t N, const int32_t* hm,
                     const int64_t* he, int64_t hps, const int64_t* ghrow, const int32_t* rg, const int32_t* rd,
                     int64_t nH, const int64_t* col, const int32_t* cold, const int32_t* ncol, int Kc, int32_t* om,
                     int64_t* oe, int64_t ops, const int32_t* wrow, const int32_t* wcnt, const int32_t* woff,
                     int64_t nw, int rows_per_lane, void* stream) {
  if (nH <= 0 || N <= 0) return 0;
  if (b3d::check_planes("b3d_cauchy_level", {rps, hps, ops})) return 2;
  hipStream_t s = (hipStream_t)stream;
  if (wrow && wcnt && woff && nw > 0) {
    if (nw > (int64_t(1) << 33)) {
      b3d::set_error("b3d_cauchy_level: too many waves");
      return 2;
    }
    if (rows_per_lane == 2) {
      B3D_DISPATCH_L(L, k_cauchy_pairs<LL><<<dim3((unsigned)((nw + 3) / 4)), dim3(256), 0, s>>>(rm, re, rps, N, hm, he, hps, ghrow, rg, rd, col, cold, ncol, Kc, om, oe, ops, wrow, wcnt, woff, nw));
    } else if (rows_per_lane == 1) {
      B3D_DISPATCH_L(L, k_cauchy_groups<LL><<<dim3((unsigned)((nw + 3) / 4)), dim3(256), 0, s>>>(rm, re, rps, N, hm, he, hps, ghrow, rg, rd, col, cold, ncol, Kc, om, oe, ops, wrow, wcnt, woff, nw));
    } else {
      b3d::set_error("b3d_cauchy_level: rows_per_lane must be 1 or 2");
      return 2;
    }
    return b3d::check_launch("b3d_cauchy_level");
  }
  // row-uniform kernel unless padding the rows to whole waves would idle more than a quarter of the lanes
  const int64_t W = (N + 63) / 64 * 64;
  if (4 * (W - N) <= W) {
    B3D_DISPATCH_L(L, k_cauchy_rows<LL><<<dim3(grid_for(nH * W, 256)), dim3(256), 0, s>>>(rm, re, rps, N, hm, he, hps, ghrow, rg, rd, nH, col, cold, ncol, Kc, om, oe, ops, W));
  } else {
    B3D_DISPATCH_L(L, k_cauchy<LL><<<dim3(grid_for(nH * N, 256)), dim3(256), 0, s>>>(rm, re, rps, N, hm, he, hps, ghrow, rg, rd, nH, col, cold, ncol, Kc, om, oe, ops));
  }
  return b3d::check_launch("b3d_cauchy_level");
}

int b3d_pole_level(int L, int SL, int SR, int S, int64_t B, int64_t T, const int32_t* Hm, const int64_t* He,
                   int64_t hps, const int64_t* thrw, const int32_t* tR, const int32_t* tQL, const int64_t* tslot,
                   const int32_t* R28, const uint32_t* Rsg, const int64_t* Re, int64_t rtps, const int32_t* Q28,
                   const uint32_t* Qsg, const int64_t* Qe, int64_t qps, const int8_t* Rd8, const int32_t* Rcorr,
                   const int8_t* Qd8, const int32_t* Qcorr, const int32_t* kR, const int32_t* kQ, int32_t* Ym,
                   int64_t* Ye, int64_t yps, int32_t* rm, int64_t* re, int64_t rps, void* stream) {
  if (T <= 0 || B <= 0) return 0;
  if ((kR == nullptr) != (kQ == nullptr)) {
    b3d::set_error("b3d_pole_level: give both per-column offsets kR and kQ or neither");
    return 2;
  }
  if (S < 1 || S > 16 || SL < 1 || SR < 1 || SL > S || SR > S) {
    b3d::set_error("b3d_pole_level: need 1 <= SL, SR <= S <= 16");
    return 2;
  }
  if (b3d::check_planes("b3d_pole_level", {hps, rtps, qps, yps, rps})) return 2;
  hipStream_t s = (hipStream_t)stream;
  if (Rd8 && Rcorr && Qd8 && Qcorr && !kR && SL == S && SR == S && (int64_t)S * B >= 32 &&
      S <= b3d::pole_mma_max_s(L)) {
    // square structure blocks, at least 32 rows sharing each constant row: int8 MFMA products
    if (b3d::pole_mma(L, Hm, He, hps, thrw, tR, Rd8, Rcorr, Re, nullptr, T, S, B, Ym, Ye, yps, s)) return 1;
    return b3d::pole_mma(L, Ym, Ye, yps, nullptr, tQL, Qd8, Qcorr, Qe, tslot, T, S, B, rm, re, rps, s);
  }
  // one block per 256 / SR (R product) or 256 / SL (q L product) columns, each staging its terms in LDS
  const int nl = b3d_limbs28(L);
  const size_t lds = 256 * (size_t)(nl + 1) * sizeof(int32_t);
  const int64_t colR = T * SL * B, colL = T * SR * B;
  B3D_DISPATCH_L(L, k_pole_r<LL><<<dim3(grid_for(colR * SR, 256 / SR * SR)), dim3(256), lds, s>>>(Hm, He, hps, thrw, tR, T, SL, SR, S, B, R28, Rsg, Re, rtps, kR, Ym, Ye, yps));
  if (b3d::check_launch("b3d_pole_level (R)")) return 1;
  B3D_DISPATCH_L(L, k_pole_l<LL><<<dim3(grid_for(colL * SL, 256 / SL * SL)), dim3(256), lds, s>>>(Ym, Ye, yps, tQL, tslot, T, SL, SR, S, B, Q28, Qsg, Qe, qps, kQ, rm, re, rps));
  return b3d::check_launch("b3d_pole_level (L)");
}

}  // extern "C"
