// CPU build of the gfx950 kernels' per-thread multiprecision helpers (blocksvibe_amd/csrc/mp.hpp), so that
// tests can check them bit for bit against the oracle on inputs of any size without a GPU.  The device
// intrinsics have host stand-ins (a ballot is the lane's own predicate, alignbit is a funnel shift).
#include <math.h>
#include <stdint.h>
#define B3D_HOST_CHECK 1
#define __device__
#define __forceinline__ inline
#include "../../blocksvibe_amd/csrc/mp.hpp"

#define DISP(L, CALL)                                                                    \
  switch (L) {                                                                           \
    case 2: { constexpr int LL = 2; return CALL; }                                       \
    case 4: { constexpr int LL = 4; return CALL; }                                       \
    case 6: { constexpr int LL = 6; return CALL; }                                       \
    case 10: { constexpr int LL = 10; return CALL; }                                     \
    case 34: { constexpr int LL = 34; return CALL; }                                     \
    case 50: { constexpr int LL = 50; return CALL; }                                     \
  }                                                                                      \
  return -1;

template <int L> static void copy_in(u32 (&d)[L], const uint32_t* s) { for (int k = 0; k < L; ++k) d[k] = s[k]; }
template <int L> static void copy_out(uint32_t* d, const u32 (&s)[L]) { for (int k = 0; k < L; ++k) d[k] = s[k]; }

template <int L> static int64_t add_t(const uint32_t* a, int64_t ea, const uint32_t* b, int64_t eb, uint32_t* r) {
  u32 A[L], B[L], R[L];
  copy_in(A, a);
  copy_in(B, b);
  const int64_t e = mp::add<L>(A, ea, B, eb, R);
  copy_out(r, R);
  return e;
}

template <int L> static int64_t mul_t(const uint32_t* a, int64_t ea, const uint32_t* b, int64_t eb, uint32_t* r) {
  u32 A[L], B[L], R[L];
  copy_in(A, a);
  copy_in(B, b);
  if (A[L - 1] == 0u || B[L - 1] == 0u) {
    for (int k = 0; k < L; ++k) r[k] = 0;
    return B3D_EZ;
  }
  const int64_t e = mp::mul_norm<L>(A, B, ea + eb, R);
  copy_out(r, R);
  return e;
}

// norm of a two's complement integer given in LW limbs, sign-extended to the kernels' widths
template <int L, int NW> static int64_t norm_w(const uint32_t* w, int LW, int64_t e, uint32_t* r) {
  u32 acc[NW];
  const u32 ext = (w[LW - 1] >> 31) ? 0xFFFFFFFFu : 0u;
  for (int k = 0; k < NW; ++k) acc[k] = k < LW ? w[k] : ext;
  u32 R[L];
  const int64_t eo = mp::norm<L, NW>(acc, e, R);
  copy_out(r, R);
  return eo;
}
template <int L> static int64_t norm_t(const uint32_t* w, int LW, int64_t e, uint32_t* r) {
  if (LW <= L + 3) return norm_w<L, L + 3>(w, LW, e, r);
  if (LW <= 2 * L + 2) return norm_w<L, 2 * L + 2>(w, LW, e, r);
  return norm_w<L, 2 * L + 4>(w, LW, e, r);
}

template <int L> static int64_t divs_t(const uint32_t* x, uint32_t d, uint32_t* q) {
  u32 X[L + 1];   // m 2^32
  X[0] = 0u;
  for (int k = 0; k < L; ++k) X[k + 1] = x[k];
  mp::div_floor(X, d);
  for (int k = 0; k < L + 1; ++k) q[k] = X[k];
  return 0;
}

template <int L> static int64_t shr_t(const uint32_t* x, int s, int left, uint32_t* out) {
  u32 X[L + 3];
  for (int k = 0; k < L + 3; ++k) X[k] = x[k];
  if (left == 1) mp::shl(X, s);
  else if (left == 2) mp::sar(X, s, 0u - (X[L + 2] >> 31));
  else mp::shr(X, s);
  for (int k = 0; k < L + 3; ++k) out[k] = X[k];
  return 0;
}

template <int L>
static int64_t cauchy_t(const int32_t* rm, const int64_t* re, int64_t rps, int64_t N, int64_t n, const int64_t* colg,
                        const int32_t* coldg, int nc, int D, const int32_t* hm, const int64_t* he, int64_t hps,
                        int64_t hrow, int uni, uint32_t* r) {
  u32 R[L];
  const int64_t e = uni ? mp::cauchy_entry<L, true>(rm, re, rps, N, n, colg, coldg, nc, D, hm, he, hps, hrow, R)
                        : mp::cauchy_entry<L, false>(rm, re, rps, N, n, colg, coldg, nc, D, hm, he, hps, hrow, R);
  copy_out(r, R);
  return e;
}

// cauchy_entry2 (two rows of a group per lane): both rows' mantissas go to planes om (L x 2, entries 0 and 1),
// exponents to oe[0..1] (row 2 only when has2)
template <int L>
static int64_t cauchy2_t(const int32_t* rm, const int64_t* re, int64_t rps, int64_t N, int64_t n, const int64_t* colg,
                         const int32_t* coldg, int nc, int D1, int D2, int has2, const int32_t* hm, const int64_t* he,
                         int64_t hps, int64_t hrow, int32_t* om, int64_t* oe) {
  u32 R[L];
  int64_t e1 = 0, e2 = 0;
  mp::cauchy_entry2<L>(rm, re, rps, N, n, colg, coldg, nc, D1, D2, has2 != 0, hm, he, hps, hrow, R, e1, e2, om, 2, 0,
                       1);
  oe[0] = e1;
  if (has2) oe[1] = e2;
  return 0;
}

template <int L>
static int64_t dot_t(const int32_t* am, int64_t aps, int64_t abase, int64_t ea, const int32_t* xm, const int64_t* xe,
                     int64_t xps, int64_t xbase, int64_t xstride, int S, uint32_t* r) {
  // the constant operand in the pole step's form: 28-bit magnitude limb planes and sign masks (as b3d_mpf_to28)
  constexpr int NL = mp::R28<L>::NL;
  static int32_t a28[NL * 64];
  static uint32_t sg[64];
  for (int j = 0; j < aps && j < 64; ++j) {
    u32 a[L], t[NL];
    for (int k = 0; k < L; ++k) a[k] = (u32)am[k * aps + j];
    sg[j] = mp::split28<L>(a, t);
    for (int k = 0; k < NL; ++k) a28[k * aps + j] = (int32_t)t[k];
  }
  u32 R[L];
  const int64_t e = mp::small_dot<L>(a28, sg, aps, abase, ea, xm, xe, xps, xbase, xstride, S, R);
  copy_out(r, R);
  return e;
}

template <int L>
static int64_t mom_t(const int32_t* sm, const int64_t* se, int64_t sps, int W, int64_t nrows, int64_t N, int64_t row,
                     int64_t c, int n0, int k, int R1, const uint32_t* NK, int64_t nkps, int KL, const int32_t* NKb,
                     uint32_t* r) {
  u32 R[L];
  const int64_t e = mp::mom_window<L, 4>(sm, se, sps, W, nrows, N, row, c, n0, k, R1, NK, nkps, KL, NKb, R);
  copy_out(r, R);
  return e;
}

// tri_entry on host arrays: v (L planes of n entries, exponents ve), table (TL planes, signs), one output
template <int L, int TL> static int64_t tri_t(const int32_t* vm, const int64_t* ve, int64_t vps, int64_t vbase,
                                              int64_t vstride, int K, int e, const int32_t* tm, const uint32_t* tsg,
                                              int64_t tps, int64_t tb, int64_t texp, uint32_t* r) {
  u32 R[L];
  const int64_t eo = mp::tri_entry<L, TL>(vm, ve, vps, vbase, vstride, K, e + 1, tm, tsg, tps, tb, texp, R);
  copy_out(r, R);
  return eo;
}
template <int L> static int64_t tri_tl(int TL, const int32_t* vm, const int64_t* ve, int64_t vps, int64_t vbase,
                                       int64_t vstride, int K, int e, const int32_t* tm, const uint32_t* tsg,
                                       int64_t tps, int64_t tb, int64_t texp, uint32_t* r) {
  if (TL == 2) return tri_t<L, 2>(vm, ve, vps, vbase, vstride, K, e, tm, tsg, tps, tb, texp, r);
  if (TL == 4) return tri_t<L, 4>(vm, ve, vps, vbase, vstride, K, e, tm, tsg, tps, tb, texp, r);
  return tri_t<L, 8>(vm, ve, vps, vbase, vstride, K, e, tm, tsg, tps, tb, texp, r);
}

// small_dot_staged on host arrays: the S terms staged by stage_term into a local buffer first
template <int L> static int64_t dot_staged_t(const int32_t* am, int64_t aps, int64_t abase, int64_t ea,
                                             const int32_t* xm, const int64_t* xe, int64_t xps, int64_t xbase,
                                             int64_t xstride, int S, uint32_t* r) {
  constexpr int NL = mp::R28<L>::NL;
  // the constant row as 28-bit magnitude limbs + sign masks, as b3d_mpf_to28 stores it
  int32_t a28[NL * 16];
  uint32_t sg[16];
  for (int j = 0; j < S; ++j) {
    u32 a[L];
    for (int k = 0; k < L; ++k) a[k] = (u32)am[k * aps + abase + j];
    u32 t[NL];
    sg[j] = mp::split28<L>(a, t);
    for (int k = 0; k < NL; ++k) a28[k * S + j] = (int32_t)t[k];
  }
  int64_t E = B3D_EZ;
  for (int j = 0; j < S; ++j) {
    const int64_t idx = xbase + j * xstride;
    if (xm[(L - 1) * xps + idx] != 0) E = xe[idx] > E ? xe[idx] : E;
  }
  int32_t st[16 * (NL + 1)];
  for (int j = 0; j < S; ++j) mp::stage_term<L>(xm, xe, xps, xbase + j * xstride, E, st + j * (NL + 1));
  u32 R[L];
  const int64_t e = mp::small_dot_staged<L>(a28, sg, S, 0, ea, st, S, E, R);
  copy_out(r, R);
  return e;
}

extern "C" {
int64_t hc_add(int L, const uint32_t* a, int64_t ea, const uint32_t* b, int64_t eb, uint32_t* r) {
  DISP(L, add_t<LL>(a, ea, b, eb, r))
}
int64_t hc_mul(int L, const uint32_t* a, int64_t ea, const uint32_t* b, int64_t eb, uint32_t* r) {
  DISP(L, mul_t<LL>(a, ea, b, eb, r))
}
int64_t hc_norm(int L, const uint32_t* w, int LW, int64_t e, uint32_t* r) { DISP(L, norm_t<LL>(w, LW, e, r)) }
int64_t hc_div(int L, const uint32_t* x, uint32_t d, uint32_t* q) { DISP(L, divs_t<LL>(x, d, q)) }
int64_t hc_shift(int L, const uint32_t* x, int s, int left, uint32_t* out) { DISP(L, shr_t<LL>(x, s, left, out)) }
int64_t hc_cauchy(int L, const int32_t* rm, const int64_t* re, int64_t rps, int64_t N, int64_t n, const int64_t* colg,
                  const int32_t* coldg, int nc, int D, const int32_t* hm, const int64_t* he, int64_t hps, int64_t hrow,
                  int uni, uint32_t* r) {
  DISP(L, cauchy_t<LL>(rm, re, rps, N, n, colg, coldg, nc, D, hm, he, hps, hrow, uni, r))
}
int64_t hc_cauchy2(int L, const int32_t* rm, const int64_t* re, int64_t rps, int64_t N, int64_t n,
                   const int64_t* colg, const int32_t* coldg, int nc, int D1, int D2, int has2, const int32_t* hm,
                   const int64_t* he, int64_t hps, int64_t hrow, int32_t* om, int64_t* oe) {
  DISP(L, cauchy2_t<LL>(rm, re, rps, N, n, colg, coldg, nc, D1, D2, has2, hm, he, hps, hrow, om, oe))
}
int64_t hc_dot(int L, const int32_t* am, int64_t aps, int64_t abase, int64_t ea, const int32_t* xm, const int64_t* xe,
               int64_t xps, int64_t xbase, int64_t xstride, int S, uint32_t* r) {
  DISP(L, dot_t<LL>(am, aps, abase, ea, xm, xe, xps, xbase, xstride, S, r))
}
int64_t hc_mom(int L, const int32_t* sm, const int64_t* se, int64_t sps, int W, int64_t nrows, int64_t N, int64_t row,
               int64_t c, int n0, int k, int R1, const uint32_t* NK, int64_t nkps, int KL, const int32_t* NKb,
               uint32_t* r) {
  DISP(L, mom_t<LL>(sm, se, sps, W, nrows, N, row, c, n0, k, R1, NK, nkps, KL, NKb, r))
}
int64_t hc_tri(int L, int TL, const int32_t* vm, const int64_t* ve, int64_t vps, int64_t vbase, int64_t vstride, int K,
               int e, const int32_t* tm, const uint32_t* tsg, int64_t tps, int64_t tb, int64_t texp, uint32_t* r) {
  DISP(L, tri_tl<LL>(TL, vm, ve, vps, vbase, vstride, K, e, tm, tsg, tps, tb, texp, r))
}
int64_t hc_dot_staged(int L, const int32_t* am, int64_t aps, int64_t abase, int64_t ea, const int32_t* xm,
                      const int64_t* xe, int64_t xps, int64_t xbase, int64_t xstride, int S, uint32_t* r) {
  DISP(L, dot_staged_t<LL>(am, aps, abase, ea, xm, xe, xps, xbase, xstride, S, r))
}
}
