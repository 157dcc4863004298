// Multiprecision helpers for gfx950 kernels: fixed-size limb arrays held in VGPRs.
//
// Every array index below is a compile-time constant after unrolling (shifts by a run-time amount are
// barrel shifters built from selects), so the arrays stay in registers instead of falling back to scratch
// memory.  Number format: value = m * 2^e, m two's complement over L 32-bit limbs (limb 0 least
// significant), normalised bitlen(|m|) = NB = 32 L - 2, zero = (0, EZ).
#pragma once
#ifndef B3D_HOST_CHECK
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

typedef uint32_t u32;
typedef uint64_t u64;
typedef int64_t i64;

#define B3D_EZ (-(1LL << 62))

namespace mp {

template <int N>
__device__ __forceinline__ void zero(u32 (&a)[N]) {
#pragma unroll
  for (int k = 0; k < N; ++k) a[k] = 0u;
}

// two's complement negation in place
template <int N>
__device__ __forceinline__ void neg(u32 (&a)[N]) {
  u64 c = 1;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    u64 v = (u64)(~a[k]) + c;
    a[k] = (u32)v;
    c = v >> 32;
  }
}

template <int N>
__device__ __forceinline__ bool is_zero(const u32 (&a)[N]) {
  u32 o = 0;
#pragma unroll
  for (int k = 0; k < N; ++k) o |= a[k];
  return o == 0u;
}

template <int N>
__device__ __forceinline__ int bitlen(const u32 (&a)[N]) {
  int bl = 0;
#pragma unroll
  for (int k = 0; k < N; ++k)
    if (a[k]) bl = 32 * k + 32 - __clz(a[k]);
  return bl;
}

// a >>= 32*q limbs (q >= 0), zero fill
template <int N>
__device__ __forceinline__ void shr_limbs(u32 (&a)[N], int q) {
  if (q >= N) {
    zero(a);
    return;
  }
#pragma unroll
  for (int b = 0; (1 << b) < N; ++b) {
    const int st = 1 << b;
    const bool on = (q >> b) & 1;
#pragma unroll
    for (int k = 0; k < N; ++k) a[k] = on ? (k + st < N ? a[k + st] : 0u) : a[k];
  }
}

// a <<= 32*q limbs (q >= 0), zero fill, bits shifted out are lost
template <int N>
__device__ __forceinline__ void shl_limbs(u32 (&a)[N], int q) {
  if (q >= N) {
    zero(a);
    return;
  }
#pragma unroll
  for (int b = 0; (1 << b) < N; ++b) {
    const int st = 1 << b;
    const bool on = (q >> b) & 1;
#pragma unroll
    for (int k = N - 1; k >= 0; --k) a[k] = on ? (k - st >= 0 ? a[k - st] : 0u) : a[k];
  }
}

// logical right shift by s >= 0 bits (unsigned magnitude), truncating
template <int N>
__device__ __forceinline__ void shr(u32 (&a)[N], int s) {
  shr_limbs(a, s >> 5);
  const u32 r = (u32)(s & 31);
  if (r) {
#pragma unroll
    for (int k = 0; k < N; ++k) {
      u32 hi = (k + 1 < N) ? a[k + 1] : 0u;
      a[k] = (a[k] >> r) | (hi << (32u - r));
    }
  }
}

// left shift by s >= 0 bits
template <int N>
__device__ __forceinline__ void shl(u32 (&a)[N], int s) {
  shl_limbs(a, s >> 5);
  const u32 r = (u32)(s & 31);
  if (r) {
#pragma unroll
    for (int k = N - 1; k >= 0; --k) {
      u32 lo = (k > 0) ? a[k - 1] : 0u;
      a[k] = (a[k] << r) | (lo >> (32u - r));
    }
  }
}

// acc += x (N >= M limbs, x zero-extended), two's complement wrap
template <int N, int M>
__device__ __forceinline__ void add_into(u32 (&acc)[N], const u32 (&x)[M]) {
  u64 c = 0;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    u64 v = (u64)acc[k] + (k < M ? x[k] : 0u) + c;
    acc[k] = (u32)v;
    c = v >> 32;
  }
}

// acc -= x (N >= M limbs)
template <int N, int M>
__device__ __forceinline__ void sub_from(u32 (&acc)[N], const u32 (&x)[M]) {
  u32 b = 0;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    u64 v = (u64)acc[k] - (u64)(k < M ? x[k] : 0u) - (u64)b;
    acc[k] = (u32)v;
    b = (u32)(v >> 63);
  }
}

// p = a * b (magnitudes), p has NA + NB limbs
template <int NA, int NBL>
__device__ __forceinline__ void mul_full(const u32 (&a)[NA], const u32 (&b)[NBL], u32 (&p)[NA + NBL]) {
#pragma unroll
  for (int k = 0; k < NA + NBL; ++k) p[k] = 0u;
#pragma unroll
  for (int x = 0; x < NA; ++x) {
    const u32 ax = a[x];
    u64 c = 0;
#pragma unroll
    for (int y = 0; y < NBL; ++y) {
      u64 v = (u64)ax * b[y] + (u64)p[x + y] + c;
      p[x + y] = (u32)v;
      c = v >> 32;
    }
    p[x + NBL] = (u32)c;
  }
}

// Normalise an exact two's complement integer acc (NW limbs) times 2^e to L limbs: NB = 32L-2
// significant bits, round half away from zero.  acc is destroyed.  Returns the exponent; m receives the
// two's complement mantissa (zero: m = 0, e = EZ).
template <int L, int NW>
__device__ __forceinline__ i64 norm(u32 (&acc)[NW], i64 e, u32 (&m)[L]) {
  const int NBITS = 32 * L - 2;
  const bool neg_ = (acc[NW - 1] >> 31) != 0u;
  if (neg_) neg(acc);
  const int bl = bitlen(acc);
  if (bl == 0) {
    zero(m);
    return B3D_EZ;
  }
  int s = bl - NBITS;
  if (s > 0) {
    // add 2^(s-1), then shift right by s
    const int pos = s - 1;
    const int pq = pos >> 5;
    const u32 pb = 1u << (pos & 31);
    u64 c = 0;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
      u64 v = (u64)acc[k] + (k == pq ? pb : 0u) + c;
      acc[k] = (u32)v;
      c = v >> 32;
    }
    shr(acc, s);
    // a carry out of the rounding gives exactly 2^NB: one more bit down
    const bool over = (acc[L - 1] >> 30) != 0u || (NW > L && acc[L] != 0u);
    if (over) {
      shr(acc, 1);
      s += 1;
    }
#pragma unroll
    for (int k = 0; k < L; ++k) m[k] = acc[k];
  } else {
#pragma unroll
    for (int k = 0; k < L; ++k) m[k] = k < NW ? acc[k] : 0u;
    if (s < 0) shl(m, -s);
  }
  if (neg_) neg(m);
  return e + s;
}

// load / store an entry of a limb-plane array (plane stride ps)
template <int N>
__device__ __forceinline__ void load(u32 (&a)[N], const int32_t* __restrict__ p, i64 ps, i64 idx) {
#pragma unroll
  for (int k = 0; k < N; ++k) a[k] = (u32)p[k * ps + idx];
}

template <int N>
__device__ __forceinline__ void store(const u32 (&a)[N], int32_t* __restrict__ p, i64 ps, i64 idx) {
#pragma unroll
  for (int k = 0; k < N; ++k) p[k * ps + idx] = (int32_t)a[k];
}

// sign-magnitude from two's complement (in place); returns true if negative
template <int N>
__device__ __forceinline__ bool to_mag(u32 (&a)[N]) {
  const bool n = (a[N - 1] >> 31) != 0u;
  if (n) neg(a);
  return n;
}

// a + b for normalised operands: both put on the grid 2^(max(ea,eb) - 32), the smaller truncated, exact
// sum in L + 3 limbs, one rounding.  a, b are two's complement (not modified); result in r.
template <int L>
__device__ __forceinline__ i64 add(const u32 (&a)[L], i64 ea, const u32 (&b)[L], i64 eb, u32 (&r)[L]) {
  const bool za = (a[L - 1] == 0u), zb = (b[L - 1] == 0u);
  if (za && zb) {
    zero(r);
    return B3D_EZ;
  }
  if (za) {
#pragma unroll
    for (int k = 0; k < L; ++k) r[k] = b[k];
    return eb;
  }
  if (zb) {
#pragma unroll
    for (int k = 0; k < L; ++k) r[k] = a[k];
    return ea;
  }
  const i64 tt = (ea > eb ? ea : eb) - 32;
  // magnitudes (negated in L limbs, then zero-extended)
  u32 ma[L], mb[L];
#pragma unroll
  for (int k = 0; k < L; ++k) {
    ma[k] = a[k];
    mb[k] = b[k];
  }
  const bool na = to_mag(ma), nb = to_mag(mb);
  u32 wa[L + 3], wb[L + 3];
#pragma unroll
  for (int k = 0; k < L + 3; ++k) {
    wa[k] = k < L ? ma[k] : 0u;
    wb[k] = k < L ? mb[k] : 0u;
  }
  {
    const i64 da = ea - tt;  // <= 32
    if (da >= 0) shl(wa, (int)da);
    else shr(wa, (-da) > 32 * (L + 3) ? 32 * (L + 3) : (int)(-da));
    const i64 db = eb - tt;
    if (db >= 0) shl(wb, (int)db);
    else shr(wb, (-db) > 32 * (L + 3) ? 32 * (L + 3) : (int)(-db));
  }
  if (na) neg(wa);
  if (nb) neg(wb);
  add_into(wa, wb);
  return norm<L, L + 3>(wa, tt, r);
}

// q = floor(|x| * 2^32 / d) for a magnitude x of L limbs and 0 < d < 2^15: L + 1 quotient limbs by long
// division one 32-bit limb at a time in FP64 (full rate on CDNA4): the partial dividend rem 2^32 + limb
// < 2^47 is exact in a double, the estimate floor(cur * (1/d)) is within +-1 of the true digit (relative
// error < 2^-52 on a digit < 2^32), and one correction makes it exact; the remainder cur - q d is exact.
template <int L>
__device__ __forceinline__ void div_small(const u32 (&x)[L], u32 d, u32 (&q)[L + 1]) {
  const double dd = (double)d;
  const double rd = 1.0 / dd;
  double rem = 0.0;
#pragma unroll
  for (int k = L; k >= 0; --k) {
    const u32 limb = k > 0 ? x[k - 1] : 0u;
    const double cur = rem * 4294967296.0 + (double)limb;
    double qd = floor(cur * rd);
    double r = fma(-qd, dd, cur);
    if (r < 0.0) {
      qd -= 1.0;
      r += dd;
    } else if (r >= dd) {
      qd += 1.0;
      r -= dd;
    }
    q[k] = (u32)qd;
    rem = r;
  }
}


// acc += a * x, a an L-limb two's complement value, x an L-limb magnitude, acc NW >= 2L + 1 limbs
// (wrapping).  Product scanning: the column sums of the unsigned product bits(a) * x are accumulated in
// 96 bits and folded into acc one limb at a time (no product array, no carry propagation per row); a
// negative a contributes bits(a) = a + 2^(32L), corrected by subtracting x at limb offset L.
template <int L, int NW>
__device__ __forceinline__ void mac_signed(u32 (&acc)[NW], const u32 (&a)[L], const u32 (&x)[L]) {
  u64 t01 = 0;
  u32 t2 = 0;
#pragma unroll
  for (int s = 0; s < 2 * L - 1; ++s) {
#pragma unroll
    for (int i = (s - L + 1 > 0 ? s - L + 1 : 0); i <= (s < L - 1 ? s : L - 1); ++i) {
      const u64 prod = (u64)a[i] * x[s - i];
      const u64 n = t01 + prod;
      t2 += (n < prod) ? 1u : 0u;
      t01 = n;
    }
    const u64 n = t01 + (u64)acc[s];
    t2 += (n < t01) ? 1u : 0u;
    acc[s] = (u32)n;
    t01 = (n >> 32) | ((u64)t2 << 32);
    t2 = 0;
  }
  u64 c = t01;
#pragma unroll
  for (int s = 2 * L - 1; s < NW; ++s) {
    const u64 n = (u64)acc[s] + (c & 0xFFFFFFFFull);
    acc[s] = (u32)n;
    c = (c >> 32) + (n >> 32);
  }
  if (a[L - 1] >> 31) {                      // acc -= x * 2^(32 L)
    u32 b = 0;
#pragma unroll
    for (int s = L; s < NW; ++s) {
      const u64 v = (u64)acc[s] - (u64)(s - L < L ? x[s - L] : 0u) - (u64)b;
      acc[s] = (u32)v;
      b = (u32)(v >> 63);
    }
  }
}

// out = norm(sum_j A[j] * X[j]) over K = S terms: A pre-aligned (mantissas a_j, exponent ea), X entries
// aligned here (over j).  Shared by both pole-step products.
template <int L>
__device__ __forceinline__ i64 small_dot(const int32_t* __restrict__ am, i64 aps, i64 abase, i64 ea,
                                         const int32_t* __restrict__ xm, const i64* __restrict__ xe, i64 xps,
                                         i64 xbase, i64 xstride, int S, u32 (&out)[L]) {
  i64 E = B3D_EZ;
  for (int j = 0; j < S; ++j) {
    const i64 idx = xbase + j * xstride;
    if (xm[(L - 1) * xps + idx] != 0) {
      const i64 ej = xe[idx];
      E = ej > E ? ej : E;
    }
  }
  if (E == B3D_EZ || ea == B3D_EZ) {
    mp::zero(out);
    return B3D_EZ;
  }
  u32 acc[2 * L + 2];
  mp::zero(acc);
  for (int j = 0; j < S; ++j) {
    const i64 idx = xbase + j * xstride;
    u32 x[L];
    mp::load(x, xm, xps, idx);
    if (x[L - 1] == 0u) continue;
    const bool nx = mp::to_mag(x);
    const i64 sh = E - xe[idx];
    if (sh >= 32 * L) continue;
    mp::shr(x, (int)sh);
    u32 a[L];
    mp::load(a, am, aps, abase + j);
    if (nx) mp::neg(a);                     // sign of x folded into a (|a| < 2^(32L-2): no overflow)
    mac_signed(acc, a, x);
  }
  return mp::norm<L, 2 * L + 2>(acc, ea + E, out);
}

// One Cauchy-step entry: h + sum_i x_i * 2 / (D - D_i) for column n of a group whose prime poles are the
// slots colg[i] (i < nc) at doubled positions coldg[i].  The x_i (slot buffer rm/re, plane stride rps,
// N columns per slot) are aligned to their largest exponent E; term i = sign * floor(|x_i| 2^32 / |d|)
// >> (E - e_i) on the grid 2^(E - 31); exact sum in L + 3 limbs, one rounding, then add(h) with the
// h_inf entry hidx of hm/he.  Returns the exponent, mantissa in r.
template <int L>
__device__ __forceinline__ i64 cauchy_entry(const int32_t* __restrict__ rm, const i64* __restrict__ re, i64 rps,
                                            i64 N, i64 n, const int64_t* __restrict__ colg,
                                            const int32_t* __restrict__ coldg, int nc, int D,
                                            const int32_t* __restrict__ hm, const i64* __restrict__ he, i64 hps,
                                            i64 hidx, u32 (&r)[L]) {
  i64 E = B3D_EZ;
  for (int i = 0; i < nc; ++i) {
    const i64 idx = colg[i] * N + n;
    if (rm[(L - 1) * rps + idx] != 0) {
      const i64 ei = re[idx];
      E = ei > E ? ei : E;
    }
  }
  u32 acc[L + 3];
  zero(acc);
  if (E != B3D_EZ) {
    for (int i = 0; i < nc; ++i) {
      int d = D - coldg[i];
      if (d == 0) continue;
      const i64 idx = colg[i] * N + n;
      u32 x[L];
      load(x, rm, rps, idx);
      if (x[L - 1] == 0u) continue;
      bool ng = to_mag(x);
      if (d < 0) {
        ng = !ng;
        d = -d;
      }
      const i64 sh = E - re[idx];
      if (sh >= 32 * (L + 1)) continue;
      u32 q[L + 1];
      div_small<L>(x, (u32)d, q);
      shr(q, (int)sh);
      if (ng) sub_from(acc, q);
      else add_into(acc, q);
    }
  }
  u32 c[L];
  const i64 ec = norm<L, L + 3>(acc, E - 31, c);
  u32 hv[L];
  load(hv, hm, hps, hidx);
  return add<L>(c, ec, hv, he[hidx], r);
}

// One windowed moment: sum_w y_w n_w^k for the staged entries y_w = stage[w, row, c] (w < W), n_w = n0 + w,
// exact n^k from NK (KL limb planes, entry n * R1 + k, bit lengths NKb).  E = max_w (e_w + NB + bitlen n_w^k)
// bounds the top bit; every product is put on the grid 2^T, T = E - 32 (L + 1) (truncated), summed exactly
// in L + 3 limbs and rounded once.  Returns the exponent (EZ and r = 0 if every y_w is zero).
template <int L, int KLM>
__device__ __forceinline__ i64 mom_window(const int32_t* __restrict__ sm, const i64* __restrict__ se, i64 sps, int W,
                                          i64 nrows, i64 N, i64 row, i64 c, int n0, int k, int R1,
                                          const uint32_t* __restrict__ NK, i64 nkps, int KL,
                                          const int32_t* __restrict__ NKb, u32 (&r)[L]) {
  const int NBITS = 32 * L - 2;
  i64 E = B3D_EZ;
  for (int w = 0; w < W; ++w) {
    const i64 o = ((i64)w * nrows + row) * N + c;
    if (sm[(L - 1) * sps + o] == 0) continue;
    const i64 x = se[o] + NBITS + NKb[(i64)(n0 + w) * R1 + k];
    E = x > E ? x : E;
  }
  if (E == B3D_EZ) {
    zero(r);
    return B3D_EZ;
  }
  const i64 T = E - 32 * (L + 1);
  u32 acc[L + 3];
  zero(acc);
  for (int w = 0; w < W; ++w) {
    const i64 o = ((i64)w * nrows + row) * N + c;
    u32 y[L];
    load(y, sm, sps, o);
    if (y[L - 1] == 0u) continue;
    const bool ny = to_mag(y);
    const i64 ni = (i64)(n0 + w) * R1 + k;
    u32 nk[KLM];
#pragma unroll
    for (int q = 0; q < KLM; ++q) nk[q] = q < KL ? NK[q * nkps + ni] : 0u;  // the table has KL limb planes
    u32 p[L + KLM];
    mul_full(y, nk, p);
    const i64 sh = T - se[o];
    if (sh >= 0) {
      if (sh >= 32 * (L + KLM)) continue;
      shr(p, (int)sh);
    } else {
      shl(p, (int)(-sh));
    }
    u32 q[L + 2];
#pragma unroll
    for (int z = 0; z < L + 2; ++z) q[z] = z < L + KLM ? p[z] : 0u;
    if (ny) sub_from(acc, q);
    else add_into(acc, q);
  }
  return norm<L, L + 3>(acc, T, r);
}

}  // namespace mp
