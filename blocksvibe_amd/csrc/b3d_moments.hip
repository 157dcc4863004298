// libb3dhip: derivative moments accumulated during the recursion (reference blocksvibe/derivatives.py:46-189,
// 448-510): v_i[k] = sum_m n^k r_c^n res[m][i], n = m + shift_i, for k <= R, per moment row
// (parity, output spin, pole) and column (structure, right-hand side).
//
//   k_mom_stage  after level m: y = res[slot_t] * r_c^(n_t), written to the staging window at w = m - m0
//   k_mom_flush  every W levels: v[row, col, k] += sum_w y_w n_w^k, with n_w^k exact integers (up to KLMAX
//                limbs); the W products are put on one fixed-point grid (T = E - 32 (L + 1), E an upper
//                bound of the largest product's top bit), truncated there, summed exactly in L + 3 limbs,
//                rounded once, then added to v.
// Layouts: staging (L, W, nrows, N), moments (L, nrows, N, R + 1) (k innermost: the R + 1 threads of a
// (row, col) share their loads of y).
#include "common.hpp"
#include "mp.hpp"
#include "../../include/b3dhip.h"

static inline unsigned grid_for(int64_t n, int th) {
  int64_t b = (n + th - 1) / th;
  if (b > (1 << 22)) b = 1 << 22;
  if (b < 1) b = 1;
  return (unsigned)b;
}

template <int L>
__global__ __launch_bounds__(256) void k_mom_stage(const int32_t* __restrict__ rm, const i64* __restrict__ re, i64 rps,
                                                   const int64_t* __restrict__ slot, const int64_t* __restrict__ row,
                                                   const int32_t* __restrict__ nn, i64 T, i64 N,
                                                   const int32_t* __restrict__ pm, const i64* __restrict__ pe, i64 pps,
                                                   int32_t* __restrict__ sm, i64* __restrict__ se, i64 sps, i64 woff) {
  for (i64 t = blockIdx.x * (i64)blockDim.x + threadIdx.x; t < T * N; t += (i64)gridDim.x * blockDim.x) {
    const i64 c = t % N, task = t / N;
    u32 a[L], b[L], r[L];
    mp::load(a, rm, rps, slot[task] * N + c);
    const i64 n = nn[task];
    mp::load(b, pm, pps, n);
    i64 e = B3D_EZ;
    if (a[L - 1] == 0u || b[L - 1] == 0u) {
      mp::zero(r);
    } else {
      e = mp::mul_norm<L>(a, b, re[slot[task] * N + c] + pe[n], r);
    }
    const i64 o = woff + row[task] * N + c;
    mp::store(r, sm, sps, o);
    se[o] = e;
  }
}

template <int L, int KLM>
__global__ __launch_bounds__(256) void k_mom_flush(const int32_t* __restrict__ sm, const i64* __restrict__ se, i64 sps,
                                                   int W, int m0, const int32_t* __restrict__ srow,
                                                   const uint32_t* __restrict__ NK, i64 nkps, int KL,
                                                   const int32_t* __restrict__ NKb, int32_t* __restrict__ vm,
                                                   i64* __restrict__ ve, i64 vps, i64 nrows, i64 N, int R1, i64 B,
                                                   const int32_t* __restrict__ jk, int Kc) {
  // thread (row, structure entry s2, j): moment k of right-hand side b, b R1 + k = jk[j] (or j without a list)
  const i64 S2 = N / B;
  for (i64 t = blockIdx.x * (i64)blockDim.x + threadIdx.x; t < nrows * S2 * Kc; t += (i64)gridDim.x * blockDim.x) {
    const int j = (int)(t % Kc);
    const i64 rs = t / Kc;
    const int code = jk ? jk[j] : j;
    const int k = code % R1;
    const i64 c = (rs % S2) * B + code / R1, row = rs / S2;
    u32 r[L];
    const i64 er = mp::mom_window<L, KLM>(sm, se, sps, W, nrows, N, row, c, m0 + srow[row], k, R1, NK, nkps, KL, NKb, r);
    if (r[L - 1] == 0u) continue;
    const i64 vo = (row * N + c) * R1 + k;
    u32 v[L], out[L];
    mp::load(v, vm, vps, vo);
    const i64 eo = mp::add<L>(v, ve[vo], r, er, out);
    mp::store(out, vm, vps, vo);
    ve[vo] = eo;
  }
}

extern "C" {

int b3d_moments_stage(int L, const int32_t* rm, const int64_t* re, int64_t rps, const int64_t* slot,
                      const int64_t* row, const int32_t* nn, int64_t T, int64_t N, const int32_t* pm,
                      const int64_t* pe, int64_t pps, int32_t* sm, int64_t* se, int64_t sps, int64_t woff,
                      void* stream) {
  if (T <= 0 || N <= 0) return 0;
  if (b3d::check_planes("b3d_moments_stage", {rps, pps, sps})) return 2;
  hipStream_t s = (hipStream_t)stream;
  B3D_DISPATCH_L(L, k_mom_stage<LL><<<grid_for(T * N, 256), 256, 0, s>>>(rm, re, rps, slot, row, nn, T, N, pm, pe,
                                                                          pps, sm, se, sps, woff));
  return b3d::check_launch("b3d_moments_stage");
}

int b3d_moments_flush(int L, int KL, const int32_t* sm, const int64_t* se, int64_t sps, int W, int m0,
                      const int32_t* srow, const uint32_t* NK, int64_t nkps, const int32_t* NKb, int32_t* vm,
                      int64_t* ve, int64_t vps, int64_t nrows, int64_t N, int R1, int64_t B, const int32_t* jk,
                      int Kc, void* stream) {
  if (nrows <= 0 || N <= 0 || R1 <= 0) return 0;
  if (B <= 0 || N % B != 0) {
    b3d::set_error("b3d_moments_flush: N must be a multiple of B > 0");
    return 2;
  }
  if (!jk) Kc = (int)(B * R1);
  if (Kc <= 0) return 0;
  if (b3d::check_planes("b3d_moments_flush", {sps, vps})) return 2;
  hipStream_t s = (hipStream_t)stream;
  const int64_t n = nrows * (N / B) * Kc;
  // the products y n^k run over exactly the limbs n^k has in the table (KL): no product by a zero limb
  if (KL == 1) {
    B3D_DISPATCH_L(L, k_mom_flush<LL, 1><<<grid_for(n, 256), 256, 0, s>>>(sm, se, sps, W, m0, srow, NK, nkps, KL,
                                                                           NKb, vm, ve, vps, nrows, N, R1, B, jk,
                                                                           Kc));
  } else if (KL == 2) {
    B3D_DISPATCH_L(L, k_mom_flush<LL, 2><<<grid_for(n, 256), 256, 0, s>>>(sm, se, sps, W, m0, srow, NK, nkps, KL,
                                                                           NKb, vm, ve, vps, nrows, N, R1, B, jk,
                                                                           Kc));
  } else if (KL == 3) {
    B3D_DISPATCH_L(L, k_mom_flush<LL, 3><<<grid_for(n, 256), 256, 0, s>>>(sm, se, sps, W, m0, srow, NK, nkps, KL,
                                                                           NKb, vm, ve, vps, nrows, N, R1, B, jk,
                                                                           Kc));
  } else if (KL == 4) {
    B3D_DISPATCH_L(L, k_mom_flush<LL, 4><<<grid_for(n, 256), 256, 0, s>>>(sm, se, sps, W, m0, srow, NK, nkps, KL,
                                                                           NKb, vm, ve, vps, nrows, N, R1, B, jk,
                                                                           Kc));
  } else if (KL <= 8) {
    B3D_DISPATCH_L(L, k_mom_flush<LL, 8><<<grid_for(n, 256), 256, 0, s>>>(sm, se, sps, W, m0, srow, NK, nkps, KL,
                                                                           NKb, vm, ve, vps, nrows, N, R1, B, jk,
                                                                           Kc));
  } else {
    b3d::set_error("b3d_moments_flush: n^k needs more than 8 limbs");
    return 2;
  }
  return b3d::check_launch("b3d_moments_flush");
}

}  // extern "C"
