// Multiprecision helpers for gfx950 kernels: fixed-size limb arrays held in VGPRs.
//
// Every array index below is a compile-time constant after unrolling (shifts by a run-time amount are
// barrel shifters built from selects), so the arrays stay in registers instead of falling back to scratch
// memory.  Carry chains use the hardware carry flag (__builtin_addc / __builtin_subc -> v_add_co /
// v_addc_co, one instruction per limb); a barrel stage that no lane of the wave needs is skipped with a
// wave-uniform branch (ballot), and sign handling is branch-free (xor with a sign mask, carry in).
// Number format: value = m * 2^e, m two's complement over L 32-bit limbs (limb 0 least significant),
// normalised bitlen(|m|) = NB = 32 L - 2, zero = (0, EZ).
#pragma once
#ifndef B3D_HOST_CHECK
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

typedef uint32_t u32;
typedef uint64_t u64;
typedef int64_t i64;

#define B3D_EZ (-(1LL << 62))

#ifdef B3D_HOST_CHECK
#define B3D_ANY(p) (p)
static inline u32 b3d_alignbit(u32 hi, u32 lo, u32 r) {
  r &= 31u;
  return r ? (lo >> r) | (hi << (32u - r)) : lo;
}
#else
#define B3D_ANY(p) (__builtin_amdgcn_ballot_w64(p) != 0)
#define b3d_alignbit(hi, lo, r) __builtin_amdgcn_alignbit((hi), (lo), (r))
#endif

namespace mp {

__device__ __forceinline__ u32 addc(u32 a, u32 b, u32& c) {
  unsigned co;
  const u32 r = __builtin_addc(a, b, c, &co);
  c = co;
  return r;
}

__device__ __forceinline__ u32 subb(u32 a, u32 b, u32& br) {
  unsigned bo;
  const u32 r = __builtin_subc(a, b, br, &bo);
  br = bo;
  return r;
}

template <int N>
__device__ __forceinline__ void zero(u32 (&a)[N]) {
#pragma unroll
  for (int k = 0; k < N; ++k) a[k] = 0u;
}

// all-ones if a (two's complement) is negative, else 0
template <int N>
__device__ __forceinline__ u32 sign_mask(const u32 (&a)[N]) {
  return 0u - (a[N - 1] >> 31);
}

// a = (a ^ m) + (m & 1): two's complement negation for m = ~0, identity for m = 0
template <int N>
__device__ __forceinline__ void cneg(u32 (&a)[N], u32 m) {
  u32 c = m & 1u;
#pragma unroll
  for (int k = 0; k < N; ++k) a[k] = addc(a[k] ^ m, 0u, c);
}

template <int N>
__device__ __forceinline__ void neg(u32 (&a)[N]) {
  cneg(a, 0xFFFFFFFFu);
}

// sign-magnitude from two's complement (in place); returns the sign mask (~0 if it was negative)
template <int N>
__device__ __forceinline__ u32 to_mag(u32 (&a)[N]) {
  const u32 m = sign_mask(a);
  cneg(a, m);
  return m;
}

template <int N>
__device__ __forceinline__ bool is_zero(const u32 (&a)[N]) {
  u32 o = 0;
#pragma unroll
  for (int k = 0; k < N; ++k) o |= a[k];
  return o == 0u;
}

// bit length.  The values normalised here have their top bit within the highest few limbs, so the top four
// limbs are searched with selects and the rest only when some lane of the wave has a nonzero value below
// them (one OR per limb decides that).
template <int N>
__device__ __forceinline__ int bitlen(const u32 (&a)[N]) {
  constexpr int H = N < 4 ? N : 4;
  int bl = 0;
#pragma unroll
  for (int k = N - H; k < N; ++k)
    if (a[k]) bl = 32 * k + 32 - __builtin_clz(a[k]);
  if (N > H) {
    u32 low = 0;
#pragma unroll
    for (int k = 0; k < N - H; ++k) low |= a[k];
    if (B3D_ANY(bl == 0 && low != 0u)) {
      int b2 = 0;
#pragma unroll
      for (int k = 0; k < N - H; ++k)
        if (a[k]) b2 = 32 * k + 32 - __builtin_clz(a[k]);
      bl = bl ? bl : b2;
    }
  }
  return bl;
}

// out[0..NO) = in >> s (NI input limbs, s >= 0), the vacated bits filled from `fill` (0: logical shift;
// the sign mask of a two's complement input: arithmetic shift = floor division by 2^s).  Barrel stages from
// the largest: after the stage of 2^b limbs, only the NO + 2^b lowest limbs can still reach the output.
// Stages no lane needs are skipped; the bit part is one funnel shift per output limb.
template <int NO, int NI>
__device__ __forceinline__ void shr_to(u32 (&out)[NO], u32 (&a)[NI], int s, u32 fill = 0u) {
  int q = s >> 5;
  if (q > NI) q = NI;
  constexpr int TOP = (NI >= 64) ? 6 : (NI >= 32) ? 5 : (NI >= 16) ? 4 : (NI >= 8) ? 3 : (NI >= 4) ? 2 : (NI >= 2) ? 1 : 0;
#pragma unroll
  for (int b = TOP; b >= 0; --b) {
    const int st = 1 << b;
    const bool on = (q >> b) & 1;
    if (B3D_ANY(on)) {
#pragma unroll
      for (int k = 0; k < NI; ++k)
        if (k < NO + st) a[k] = on ? (k + st < NI ? a[k + st] : fill) : a[k];
    }
  }
  // q == NI (the clamp) has already moved fill limbs into every limb the output reads
  const u32 r = (u32)(s & 31);
#pragma unroll
  for (int k = 0; k < NO; ++k) {
    const u32 lo = k < NI ? a[k] : fill;
    const u32 hi = k + 1 < NI ? a[k + 1] : fill;
    out[k] = b3d_alignbit(hi, lo, r);
  }
}

// logical right shift in place by s >= 0 bits (truncating)
template <int N>
__device__ __forceinline__ void shr(u32 (&a)[N], int s) {
  u32 t[N];
  shr_to(t, a, s);
#pragma unroll
  for (int k = 0; k < N; ++k) a[k] = t[k];
}

// right shift in place by s >= 0 bits, the vacated bits filled from `fill` (the sign mask of a two's
// complement value: floor division by 2^s).  Barrel stages as in shr_to; the funnel pass runs upward, so
// each limb reads its upper neighbour before that one is rewritten.
template <int N>
__device__ __forceinline__ void sar(u32 (&a)[N], int s, u32 fill) {
  int q = s >> 5;
  if (q > N) q = N;
  constexpr int TOP = (N >= 64) ? 6 : (N >= 32) ? 5 : (N >= 16) ? 4 : (N >= 8) ? 3 : (N >= 4) ? 2 : (N >= 2) ? 1 : 0;
#pragma unroll
  for (int b = TOP; b >= 0; --b) {
    const int st = 1 << b;
    const bool on = (q >> b) & 1;
    if (B3D_ANY(on)) {
#pragma unroll
      for (int k = 0; k < N; ++k) a[k] = on ? (k + st < N ? a[k + st] : fill) : a[k];
    }
  }
  const u32 r = (u32)(s & 31);
#pragma unroll
  for (int k = 0; k < N; ++k) a[k] = b3d_alignbit(k + 1 < N ? a[k + 1] : fill, a[k], r);
}

// left shift in place by s >= 0 bits (bits shifted out are lost)
template <int N>
__device__ __forceinline__ void shl(u32 (&a)[N], int s) {
  int q = s >> 5;
  if (q > N) q = N;
  constexpr int TOP = (N >= 64) ? 6 : (N >= 32) ? 5 : (N >= 16) ? 4 : (N >= 8) ? 3 : (N >= 4) ? 2 : (N >= 2) ? 1 : 0;
#pragma unroll
  for (int b = TOP; b >= 0; --b) {
    const int st = 1 << b;
    const bool on = (q >> b) & 1;
    if (B3D_ANY(on)) {
#pragma unroll
      for (int k = N - 1; k >= 0; --k) a[k] = on ? (k - st >= 0 ? a[k - st] : 0u) : a[k];
    }
  }
  const u32 r = (u32)(s & 31);
  if (B3D_ANY(r != 0u)) {
#pragma unroll
    for (int k = N - 1; k >= 0; --k) {
      const u32 lo = k > 0 ? a[k - 1] : 0u;
      a[k] = r ? b3d_alignbit(a[k], lo, 32u - r) : a[k];
    }
  }
}

// acc += x (N >= M limbs, x zero-extended), two's complement wrap
template <int N, int M>
__device__ __forceinline__ void add_into(u32 (&acc)[N], const u32 (&x)[M]) {
  u32 c = 0;
#pragma unroll
  for (int k = 0; k < N; ++k) acc[k] = addc(acc[k], k < M ? x[k] : 0u, c);
}

// acc -= x (N >= M limbs)
template <int N, int M>
__device__ __forceinline__ void sub_from(u32 (&acc)[N], const u32 (&x)[M]) {
  u32 b = 0;
#pragma unroll
  for (int k = 0; k < N; ++k) acc[k] = subb(acc[k], k < M ? x[k] : 0u, b);
}

// acc += x if m == 0, acc -= x if m == ~0 (x of M <= N limbs, extended with `fill`: 0 for a magnitude, the
// sign mask for a two's complement value)
template <int N, int M>
__device__ __forceinline__ void cadd(u32 (&acc)[N], const u32 (&x)[M], u32 m, u32 fill = 0u) {
  u32 c = m & 1u;
#pragma unroll
  for (int k = 0; k < N; ++k) acc[k] = addc(acc[k], (k < M ? x[k] : fill) ^ m, c);
}

// p = a * b (magnitudes), p has NA + NBL limbs
template <int NA, int NBL>
__device__ __forceinline__ void mul_full(const u32 (&a)[NA], const u32 (&b)[NBL], u32 (&p)[NA + NBL]) {
#pragma unroll
  for (int k = 0; k < NA + NBL; ++k) p[k] = 0u;
#pragma unroll
  for (int x = 0; x < NA; ++x) {
    u32 c = 0;
#pragma unroll
    for (int y = 0; y < NBL; ++y) {
      const u64 v = (u64)a[x] * b[y] + (u64)p[x + y];   // < 2^64: no overflow
      u32 cc = 0;
      p[x + y] = addc((u32)v, c, cc);
      c = (u32)(v >> 32) + cc;                         // <= 2^32 - 1
    }
    p[x + NBL] = c;
  }
}

// Normalise an exact two's complement integer acc (NW limbs) times 2^e to L limbs: NB = 32L-2
// significant bits, round half away from zero.  acc is destroyed.  Returns the exponent; m receives the
// two's complement mantissa (zero: m = 0, e = EZ).
template <int L, int NW>
__device__ __forceinline__ i64 norm(u32 (&acc)[NW], i64 e, u32 (&m)[L]) {
  const int NBITS = 32 * L - 2;
  const u32 sm = to_mag(acc);
  const int bl = bitlen(acc);
  if (bl == 0) {
    zero(m);
    return B3D_EZ;
  }
  int s = bl - NBITS;
  if (s > 0) {
    // floor(|W| / 2^(s-1)) has NB + 1 bits; + 1 then >> 1 rounds half away from zero
    u32 t[L];
    shr_to(t, acc, s - 1);
    u32 c = 1u;
#pragma unroll
    for (int k = 0; k < L; ++k) t[k] = addc(t[k], 0u, c);
#pragma unroll
    for (int k = 0; k < L; ++k) m[k] = b3d_alignbit(k + 1 < L ? t[k + 1] : 0u, t[k], 1u);
    // a carry out of the rounding gives exactly 2^NB: 2^(NB-1) one exponent up
    const bool over = (m[L - 1] >> 30) != 0u;
    if (B3D_ANY(over)) {
#pragma unroll
      for (int k = 0; k < L; ++k) m[k] = over ? (k == L - 1 ? (1u << 29) : 0u) : m[k];
    }
    s += over ? 1 : 0;
  } else {
#pragma unroll
    for (int k = 0; k < L; ++k) m[k] = k < NW ? acc[k] : 0u;
    shl(m, -s);
  }
  cneg(m, sm);
  return e + s;
}

// load / store an entry of a limb-plane array (plane stride ps entries, base p uniform over the wave).
// On the GPU each limb is one buffer instruction: the plane base walks in SGPRs (the empty asm keeps the
// compiler from hoisting all N plane bases into scalar registers, which would spill) and the per-lane
// offset idx * 4 is a 32-bit VGPR, so no VALU address arithmetic is spent per limb.  The entry points check
// that every plane stride is below 2^30 entries.
#ifdef B3D_HOST_CHECK
template <int N>
__device__ __forceinline__ void load(u32 (&a)[N], const int32_t* __restrict__ p, i64 ps, i64 idx) {
  for (int k = 0; k < N; ++k) a[k] = (u32)p[k * ps + idx];
}

template <int N>
__device__ __forceinline__ void store(const u32 (&a)[N], int32_t* __restrict__ p, i64 ps, i64 idx) {
  for (int k = 0; k < N; ++k) p[k * ps + idx] = (int32_t)a[k];
}

__device__ __forceinline__ u32 load_limb(const int32_t* __restrict__ p, i64 ps, int k, i64 idx) {
  return (u32)p[k * ps + idx];
}

template <int NL = -1, int N>
__device__ __forceinline__ void load_desc(u32 (&a)[N], const int32_t* __restrict__ p, i64 ps, i64 idx) {
  constexpr int C = NL < 0 ? N : NL;
  for (int k = C - 1; k >= 0; --k) a[N - C + k] = (u32)p[k * ps + idx];
}
#else
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const char* base) {
  // raw buffer (stride 0), every 32-bit offset in range, 32-bit data format (dword 3 = 0x00020000)
  return __builtin_amdgcn_make_buffer_rsrc((void*)base, (short)0, (int)0xFFFFFFFF, 0x00020000);
}

template <int N>
__device__ __forceinline__ void load(u32 (&a)[N], const int32_t* __restrict__ p, i64 ps, i64 idx) {
  const u32 off = (u32)idx * 4u;
  const char* base = (const char*)p;
  const i64 step = ps * 4;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    a[k] = __builtin_amdgcn_raw_buffer_load_b32(plane_rsrc(base), off, 0, 0);
    base += step;
    asm volatile("" : "+s"(base));
  }
}

template <int N>
__device__ __forceinline__ void store(const u32 (&a)[N], int32_t* __restrict__ p, i64 ps, i64 idx) {
  const u32 off = (u32)idx * 4u;
  const char* base = (const char*)p;
  const i64 step = ps * 4;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    __builtin_amdgcn_raw_buffer_store_b32(a[k], plane_rsrc(base), off, 0, 0);
    base += step;
    asm volatile("" : "+s"(base));
  }
}

__device__ __forceinline__ u32 load_limb(const int32_t* __restrict__ p, i64 ps, int k, i64 idx) {
  return __builtin_amdgcn_raw_buffer_load_b32(plane_rsrc((const char*)p + k * ps * 4), (u32)idx * 4u, 0, 0);
}

// the same load issued from the top limb down, for consumers that start with the top limb (the long division):
// the first limbs they need return first.  With NL given, the entry's NL limbs go to the top NL of a[N].
template <int NL = -1, int N>
__device__ __forceinline__ void load_desc(u32 (&a)[N], const int32_t* __restrict__ p, i64 ps, i64 idx) {
  constexpr int C = NL < 0 ? N : NL;
  const u32 off = (u32)idx * 4u;
  const i64 step = ps * 4;
  const char* base = (const char*)p + (C - 1) * step;
#pragma unroll
  for (int k = C - 1; k >= 0; --k) {
    a[N - C + k] = __builtin_amdgcn_raw_buffer_load_b32(plane_rsrc(base), off, 0, 0);
    base -= step;
    asm volatile("" : "+s"(base));
  }
}
#endif

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
  u32 ma[L], mb[L];
#pragma unroll
  for (int k = 0; k < L; ++k) {
    ma[k] = a[k];
    mb[k] = b[k];
  }
  const u32 sa = to_mag(ma), sb = to_mag(mb);
  u32 wa[L + 1], wb[L + 1];
#pragma unroll
  for (int k = 0; k < L + 1; ++k) {
    wa[k] = k < L ? ma[k] : 0u;
    wb[k] = k < L ? mb[k] : 0u;
  }
  // the larger exponent is shifted up by exactly 32 (one limb), the other down by its distance
  const i64 da = ea - tt, db = eb - tt;   // <= 32, one of them == 32
  if (da >= 0) shl(wa, (int)da);
  else shr(wa, (-da) > 32 * (L + 1) ? 32 * (L + 1) : (int)(-da));
  if (db >= 0) shl(wb, (int)db);
  else shr(wb, (-db) > 32 * (L + 1) ? 32 * (L + 1) : (int)(-db));
  u32 acc[L + 3];
  zero(acc);
  cadd(acc, wa, sa);
  cadd(acc, wb, sb);
  return norm<L, L + 3>(acc, tt, r);
}

// x = floor(x / d) in place for a signed N-limb value x (two's complement) and 0 < d < 2^15: the quotient in
// two's complement (the top digit signed, the others in [0, 2^32)), by long division one 32-bit limb at a
// time in FP64 (full rate on CDNA4) with remainders in [0, d).  The partial dividend cur = rem 2^32 + limb
// (|cur| < 2^47; the top one is the signed top limb) is exact; fma(cur, 1/d, 2^-19) is within 2^-20 of
// cur/d + 2^-19 (rounding of 1/d and of the fma, each < 2^-21 on a value below 2^32 in magnitude), and the
// fractional part of cur/d is at most 1 - 1/d < 1 - 2^-15, so its floor is the exact digit; the remainder
// cur - q d is exact.
template <int N>
__device__ __forceinline__ void div_floor(u32 (&x)[N], u32 d) {
  const double dd = (double)d;
  const double rd = 1.0 / dd;
  const double bias = 1.0 / 524288.0;   // 2^-19
  double cur = (double)(int32_t)x[N - 1];
  double qd = floor(fma(cur, rd, bias));
  double rem = fma(-qd, dd, cur);
  x[N - 1] = (u32)(int32_t)qd;
#pragma unroll
  for (int k = N - 2; k >= 0; --k) {
    cur = fma(rem, 4294967296.0, (double)x[k]);
    qd = floor(fma(cur, rd, bias));
    rem = fma(-qd, dd, cur);
    x[k] = (u32)qd;
  }
}

// ---- 28-bit limbs for sums of products: a_i y_k < 2^56 in magnitude, so a column of up to 58 products
// plus a normalised limb and a carry stays inside int64 and every product is one v_mad_i64_i32 (no carry
// flags).  Signs are folded into the limbs of one operand (each limb negated).

template <int L>
struct R28 {
  static constexpr int NL = (32 * L - 2 + 27) / 28;  // limbs of a magnitude below 2^(32L - 2)
  static constexpr int NC = 2 * NL - 1;              // product columns
};

// y = sign * x in NL 28-bit limbs (x a magnitude of L limbs below 2^(32L-2); m = 0 or ~0 for the sign)
template <int L, int NL>
__device__ __forceinline__ void to28(const u32 (&x)[L], u32 m, int32_t (&y)[NL]) {
#pragma unroll
  for (int k = 0; k < NL; ++k) {
    const int bit = 28 * k, w = bit >> 5, off = bit & 31;
    const u32 lo = w < L ? x[w] : 0u;
    const u32 hi = w + 1 < L ? x[w + 1] : 0u;
    const u32 v = b3d_alignbit(hi, lo, (u32)off) & 0x0FFFFFFFu;
    y[k] = (int32_t)((v ^ m) - m);
  }
}

// acc += a * y with acc = NC normalised limbs in [0, 2^28) and a signed top limb acc[NC] (weight 2^(28 NC));
// a: nonnegative 28-bit limbs, y: signed limbs.  Product scanning, one int64 column at a time.
template <int NL>
__device__ __forceinline__ void mac28(int32_t (&acc)[2 * NL], const int32_t (&a)[NL], const int32_t (&y)[NL]) {
  constexpr int NC = 2 * NL - 1;
  i64 carry = 0;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    // two independent partial sums (alternate terms): two multiply-add chains in flight per column
    i64 s0 = (i64)acc[c] + carry, s1 = 0;
    const int lo = c - NL + 1 > 0 ? c - NL + 1 : 0, hi = c < NL - 1 ? c : NL - 1;
#pragma unroll
    for (int i = lo; i <= hi; ++i) {
      if ((i - lo) & 1) s1 += (i64)a[i] * (i64)y[c - i];
      else s0 += (i64)a[i] * (i64)y[c - i];
    }
    const i64 s = s0 + s1;
    acc[c] = (int32_t)(s & 0x0FFFFFFF);
    carry = s >> 28;
  }
  acc[NC] += (int32_t)carry;
}

// The columns c >= C0 of mac28 only (their partial products i + k = c), with no carry into column C0: acc
// then holds the upper part of the product, short of the carry out of the columns below C0.
template <int C0, int NL>
__device__ __forceinline__ void mac28_from(int32_t (&acc)[2 * NL], const int32_t (&a)[NL], const int32_t (&y)[NL]) {
  constexpr int NC = 2 * NL - 1;
  i64 carry = 0;
#pragma unroll
  for (int c = C0; c < NC; ++c) {
    i64 s0 = (i64)acc[c] + carry, s1 = 0;
    const int lo = c - NL + 1 > 0 ? c - NL + 1 : 0, hi = c < NL - 1 ? c : NL - 1;
#pragma unroll
    for (int i = lo; i <= hi; ++i) {
      if ((i - lo) & 1) s1 += (i64)a[i] * (i64)y[c - i];
      else s0 += (i64)a[i] * (i64)y[c - i];
    }
    const i64 s = s0 + s1;
    acc[c] = (int32_t)(s & 0x0FFFFFFF);
    carry = s >> 28;
  }
  acc[NC] += (int32_t)carry;
}

// the integer held by acc (NC limbs of 28 bits + signed top) as NW two's complement 32-bit limbs
template <int NA, int NW>
__device__ __forceinline__ void pack28(const int32_t (&acc)[NA], u32 (&W)[NW]) {
  constexpr int NC = NA - 1;
  constexpr int TB = 28 * NC, W0 = TB >> 5, OFF = TB & 31;
  static_assert(W0 + 1 < NW, "top limb must fit the 32-bit accumulator");
#pragma unroll
  for (int w = 0; w < NW; ++w) W[w] = 0u;
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    const int bit = 28 * k, w = bit >> 5, off = bit & 31;
    const u32 v = (u32)acc[k];
    W[w] |= v << off;
    if (off > 4 && w + 1 < NW) W[w + 1] |= v >> (32 - off);
  }
  const i64 T = acc[NC];
  const u64 t = (u64)(T * ((i64)1 << OFF));
  W[W0] |= (u32)t;
  W[W0 + 1] |= (u32)(t >> 32);
  const u32 ext = T < 0 ? 0xFFFFFFFFu : 0u;
#pragma unroll
  for (int w = W0 + 2; w < NW; ++w) W[w] = ext;
}

// First product column of the truncated products (small_dot, mul_norm): the dropped columns stay at
// least 60 bits below the rounding position of a sum whose top bit is within 60 bits of the largest
// possible one.
template <int L>
struct ProdCut {
  static constexpr int NB = 32 * L - 2;
  static constexpr int NL = R28<L>::NL;
  static constexpr int C0 = (NB - 132) / 28 > 0 ? (NB - 132) / 28 : 0;
  static constexpr int NH = 2 * NL - C0;                      // 28-bit columns of U (top one signed)
  static constexpr int NWH = (28 * (NH - 1) + 32) / 32 + 2;   // two's complement 32-bit limbs holding U
};

// U = the columns >= C0 of acc (units of 2^(28 C0)) as two's complement limbs; returns whether norm(U) is
// certainly norm(W) for every W = U 2^(28 C0) + D with |D| < 2^(28 C0 + 38): bits [39, 71) of |U| hold both
// a 0 and a 1 and bitlen(|U|) >= NB + 72, so adding D changes neither the bit length of |W| nor its bits
// from the half-unit position of the rounding up.
template <int L>
__device__ __forceinline__ bool upper_sure(const int32_t (&acc)[2 * R28<L>::NL], u32 (&U)[ProdCut<L>::NWH]) {
  constexpr int C0 = ProdCut<L>::C0, NH = ProdCut<L>::NH, NWH = ProdCut<L>::NWH;
  int32_t hi[NH];
#pragma unroll
  for (int c = 0; c < NH; ++c) hi[c] = acc[C0 + c];
  pack28(hi, U);
  if constexpr (C0 == 0) {
    return true;
  } else {
    u32 A[NWH];
#pragma unroll
    for (int k = 0; k < NWH; ++k) A[k] = U[k];
    to_mag(A);
    const u32 w = b3d_alignbit(A[2], A[1], 7u);
    return bitlen(A) >= ProdCut<L>::NB + 72 && w != 0u && w != 0xFFFFFFFFu;
  }
}

// out = norm(sum_j A[j] * X[j]) over K = S <= 16 terms: A pre-aligned and stored as 28-bit magnitude limbs (NL
// limb planes a28, plane stride aps) with sign masks asg (exponent ea), X entries aligned here (over j).
// Shared by both pole-step products.  The exact sum W is accumulated in 28-bit limbs, first without the
// partial products below column C0 (about 40 % of them at L = 34): the upper part U = W - D, where
// |D| < S (C0 + 1) 2^(28 C0 + 28) <= 2^(28 C0 + 38) for the dropped D (S <= 16, C0 < 64).  When upper_sure
// holds, norm(U) = norm(W) exactly; otherwise (a sum that cancels far below its terms, or a rare bit
// pattern) the lane recomputes W in full: the result is always the exactly rounded sum.
template <int L>
__device__ __forceinline__ i64 small_dot(const int32_t* __restrict__ a28, const uint32_t* __restrict__ asg, i64 aps,
                                         i64 abase, i64 ea, const int32_t* __restrict__ xm,
                                         const i64* __restrict__ xe, i64 xps, i64 xbase, i64 xstride, int S,
                                         u32 (&out)[L]) {
  constexpr int NL = R28<L>::NL;
  constexpr int C0 = ProdCut<L>::C0, NWH = ProdCut<L>::NWH;
  i64 E = B3D_EZ;
  for (int j = 0; j < S; ++j) {
    const i64 idx = xbase + j * xstride;
    if (load_limb(xm, xps, L - 1, idx) != 0u) {
      const i64 ej = xe[idx];
      E = ej > E ? ej : E;
    }
  }
  if (E == B3D_EZ || ea == B3D_EZ) {
    mp::zero(out);
    return B3D_EZ;
  }
  // one term's operands: the constant's 28-bit limbs and the aligned X entry in 28-bit limbs with the sign
  // of the product; false if the term is zero
  auto term = [&](int j, int32_t (&a)[NL], int32_t (&y28)[NL]) -> bool {
    const i64 idx = xbase + j * xstride;
    u32 x[L];
    mp::load(x, xm, xps, idx);
    if (x[L - 1] == 0u) return false;
    const u32 sx = mp::to_mag(x);
    const i64 sh = E - xe[idx];
    if (sh >= 32 * L) return false;
    mp::shr(x, (int)sh);
    u32 au[NL];
    mp::load(au, a28, aps, abase + j);
#pragma unroll
    for (int k = 0; k < NL; ++k) a[k] = (int32_t)au[k];
    to28(x, sx ^ asg[abase + j], y28);
    return true;
  };
  int32_t acc[2 * NL];
#pragma unroll
  for (int c = 0; c < 2 * NL; ++c) acc[c] = 0;
  for (int j = 0; j < S; ++j) {
    int32_t a[NL], y28[NL];
    if (term(j, a, y28)) mac28_from<C0>(acc, a, y28);
  }
  u32 U[NWH];
  const bool sure = upper_sure<L>(acc, U);
  if (B3D_ANY(!sure)) {
    if (!sure) {
      int32_t full[2 * NL];
#pragma unroll
      for (int c = 0; c < 2 * NL; ++c) full[c] = 0;
      for (int j = 0; j < S; ++j) {
        int32_t a[NL], y28[NL];
        if (term(j, a, y28)) mac28(full, a, y28);
      }
      u32 W[2 * L + 2];
      pack28(full, W);
      return mp::norm<L, 2 * L + 2>(W, ea + E, out);
    }
  }
  return mp::norm<L, NWH>(U, ea + E + 28 * C0, out);
}

// small_dot with its X terms already aligned and converted (by the workgroup, into LDS): term j at
// st + j (NL + 1): NL signed 28-bit limbs (the sign of X folded in) and a flag (0: a zero term), E the terms'
// common exponent.  The constant's sign is folded into its limbs here.  Same rounding as small_dot (truncated
// product, certainty test, full product where it fails): the same bits.
template <int L>
__device__ __forceinline__ i64 small_dot_staged(const int32_t* __restrict__ a28, const uint32_t* __restrict__ asg,
                                                i64 aps, i64 abase, i64 ea, const int32_t* st, int S, i64 E,
                                                u32 (&out)[L]) {
  constexpr int NL = R28<L>::NL;
  constexpr int C0 = ProdCut<L>::C0, NWH = ProdCut<L>::NWH;
  if (E == B3D_EZ || ea == B3D_EZ) {
    mp::zero(out);
    return B3D_EZ;
  }
  auto term = [&](int j, int32_t (&a)[NL], int32_t (&y)[NL]) -> bool {
    const int32_t* t = st + j * (NL + 1);
    if (t[NL] == 0) return false;
    u32 au[NL];
    mp::load(au, a28, aps, abase + j);
    const u32 m = asg[abase + j];
#pragma unroll
    for (int k = 0; k < NL; ++k) a[k] = (int32_t)((au[k] ^ m) - m);
#pragma unroll
    for (int k = 0; k < NL; ++k) y[k] = t[k];
    return true;
  };
  int32_t acc[2 * NL];
#pragma unroll
  for (int c = 0; c < 2 * NL; ++c) acc[c] = 0;
  for (int j = 0; j < S; ++j) {
    int32_t a[NL], y[NL];
    if (term(j, a, y)) mac28_from<C0>(acc, a, y);
  }
  u32 U[NWH];
  const bool sure = upper_sure<L>(acc, U);
  if (B3D_ANY(!sure)) {
    if (!sure) {
      int32_t full[2 * NL];
#pragma unroll
      for (int c = 0; c < 2 * NL; ++c) full[c] = 0;
      for (int j = 0; j < S; ++j) {
        int32_t a[NL], y[NL];
        if (term(j, a, y)) mac28(full, a, y);
      }
      u32 W[2 * L + 2];
      pack28(full, W);
      return mp::norm<L, 2 * L + 2>(W, ea + E, out);
    }
  }
  return mp::norm<L, NWH>(U, ea + E + 28 * C0, out);
}

// Stage one term for small_dot_staged: X entry idx aligned to the exponent E (magnitude truncated), in 28-bit
// limbs with its sign, and the flag; a zero entry or one shifted out entirely is a zero term.
template <int L>
__device__ __forceinline__ void stage_term(const int32_t* __restrict__ xm, const i64* __restrict__ xe, i64 xps, i64 idx,
                                           i64 E, int32_t* t) {
  constexpr int NL = R28<L>::NL;
  u32 x[L];
  mp::load(x, xm, xps, idx);
  const i64 sh = E - xe[idx];
  if (x[L - 1] == 0u || E == B3D_EZ || sh >= 32 * L) {
    t[NL] = 0;
    return;
  }
  const u32 sx = mp::to_mag(x);
  mp::shr(x, (int)sh);
  int32_t y[NL];
  to28(x, sx, y);
#pragma unroll
  for (int k = 0; k < NL; ++k) t[k] = y[k];
  t[NL] = 1;
}

// a two's complement value of L limbs as 28-bit magnitude limbs and a sign mask (the pole step's constant form)
template <int L>
__device__ __forceinline__ u32 split28(u32 (&a)[L], u32 (&a28)[R28<L>::NL]) {
  constexpr int NL = R28<L>::NL;
  const u32 sa = to_mag(a);
  int32_t t[NL];
  to28(a, 0u, t);
#pragma unroll
  for (int k = 0; k < NL; ++k) a28[k] = (u32)t[k];
  return sa;
}

// out = norm(a * b) for two's complement a, b of L limbs (both destroyed), exponent e of the product
// (the columns below ProdCut<L>::C0 first left out, as in small_dot: the dropped part of one product is below
// (C0 + 1) 2^(28 C0 + 28); the product of two normalised values has no cancellation, so upper_sure fails
// only on its rare bit pattern, and then the lane completes the product)
template <int L>
__device__ __forceinline__ i64 mul_norm(u32 (&a)[L], u32 (&b)[L], i64 e, u32 (&out)[L]) {
  constexpr int NL = R28<L>::NL;
  constexpr int C0 = ProdCut<L>::C0, NWH = ProdCut<L>::NWH;
  const u32 sa = to_mag(a), sb = to_mag(b);
  int32_t a28[NL], b28[NL], acc[2 * NL];
#pragma unroll
  for (int c = 0; c < 2 * NL; ++c) acc[c] = 0;
  to28(a, 0u, a28);
  to28(b, sa ^ sb, b28);
  mac28_from<C0>(acc, a28, b28);
  u32 U[NWH];
  const bool sure = upper_sure<L>(acc, U);
  if (B3D_ANY(!sure)) {
    if (!sure) {
#pragma unroll
      for (int c = 0; c < 2 * NL; ++c) acc[c] = 0;
      mac28(acc, a28, b28);
      u32 W[2 * L + 2];
      pack28(acc, W);
      return norm<L, 2 * L + 2>(W, e, out);
    }
  }
  return norm<L, NWH>(U, e + 28 * C0, out);
}

// One Cauchy term on the grid 2^G: acc += sign(d) floor(m 2^32 / |d| / 2^sh), sh = G + 31 - e >= 0, for the
// signed mantissa m (limbs x) with exponent e; nothing for d == 0, m == 0 or sh >= 32 (L + 1).  The value is
// put on the grid first, X = floor(m 2^(32 - sh)) (L + 1 limbs, arithmetic shift), then divided:
// floor(floor(a / 2^sh) / |d|) = floor(a / (2^sh |d|)) for integer a, so the result is the same single floor,
// and one (L + 1)-limb buffer carries the term from the load to the sum.
template <int L>
__device__ __forceinline__ void cauchy_term(u32 (&acc)[L + 2], u32 (&X)[L + 1], i64 e, int d, i64 G) {
  if (d == 0 || X[L] == 0u) return;
  const i64 sh = G + 31 - e;
  if (sh >= 32 * (L + 1)) return;
  const u32 fill = 0u - (X[L] >> 31);
  sar(X, (int)sh, fill);                 // floor(m 2^32 / 2^sh)
  div_floor(X, (u32)(d < 0 ? -d : d));   // floor(. / |d|), same sign as m
  cadd(acc, X, 0u - (u32)(d < 0), fill);  // + or - by the sign of d
}

// One Cauchy-step entry: h + sum_i x_i * 2 / (D - D_i) for column n of a group whose prime poles are the
// slots colg[i] (i < nc) at doubled positions coldg[i], h the h_inf entry (hrow, n) of hm/he, with one
// rounding.  E = largest exponent of the nonzero x_i (slot buffer rm/re, plane stride rps, N columns per slot);
// grid G = max(E - 31, e_h - 32) (either side alone when the other is absent); term i =
// sign(d) floor(m_i 2^32 / |d| / 2^(G + 31 - e_i)) (signed mantissa, floor toward -infinity; none when the
// shift reaches 32 (L + 1)) and h truncated onto 2^G; exact sum in L + 2 limbs (every term and h are
// below 2^(32 L + 30) in magnitude, so fewer than 2^32 of them cannot overflow), one rounding.  Returns the
// exponent, mantissa in r.
// UNI: the group (colg, coldg, nc, D, hrow) is the same for every lane of the wave.  Then the column list
// is read through the scalar cache and each slot's plane base colg[i] * N is a scalar, so a term costs the
// limb loads alone (lane offset n * 4); otherwise every index is a per-lane vector load.
template <int L, bool UNI = false>
__device__ __forceinline__ i64 cauchy_entry(const int32_t* __restrict__ rm, const i64* __restrict__ re, i64 rps,
                                            i64 N, i64 n, const int64_t* __restrict__ colg,
                                            const int32_t* __restrict__ coldg, int nc, int D,
                                            const int32_t* __restrict__ hm, const i64* __restrict__ he, i64 hps,
                                            i64 hrow, u32 (&r)[L]) {
  // slot i's base (UNI: a scalar offset, lane index n) or its entry (lane index 0)
  auto at = [&](i64 slot, i64& base, i64& lane) {
    if constexpr (UNI) {
      base = slot * N;
      lane = n;
    } else {
      base = 0;
      lane = slot * N + n;
    }
  };
  // zero entries carry the exponent EZ (the canonical zero every kernel writes), below that of any nonzero
  // entry, so E is the plain maximum of the exponents: independent loads, no limb read
  i64 E = B3D_EZ;
#pragma unroll 4
  for (int i = 0; i < nc; ++i) {
    i64 b, l;
    at(colg[i], b, l);
    const i64 ei = re[b + l];
    E = ei > E ? ei : E;
  }
  // only the top limb of h here: its other limbs are loaded after the terms, so they hold no registers
  // through the loop
  i64 hb, hl;
  at(hrow, hb, hl);
  const bool hz = load_limb(hm + hb, hps, L - 1, hl) == 0u;
  const i64 eh = he[hb + hl];
  if (E == B3D_EZ && hz) {
    zero(r);
    return B3D_EZ;
  }
  const i64 G = E == B3D_EZ ? eh - 32 : (hz ? E - 31 : ((E - 31) > (eh - 32) ? (E - 31) : (eh - 32)));
  u32 acc[L + 2];
  zero(acc);
  if (E != B3D_EZ) {
    // UNI: the next column's slot and position are read one term ahead (scalar registers)
    i64 cn = colg[0];
    int dn = coldg[0];
    for (int i = 0; i < nc; ++i) {
      i64 c;
      int d;
      if constexpr (UNI) {
        c = cn;
        d = D - dn;
        if (i + 1 < nc) {
          cn = colg[i + 1];
          dn = coldg[i + 1];
        }
      } else {
        c = colg[i];
        d = D - coldg[i];
      }
      if (d == 0) continue;
      i64 b, l;
      at(c, b, l);
      u32 X[L + 1];   // m 2^32
      X[0] = 0u;
      load_desc<L>(X, rm + b, rps, l);
      cauchy_term<L>(acc, X, re[b + l], d, G);
    }
  }
  if (!hz) {
    // |m_h| 2^(e_h - G) truncated = floor(|m_h| 2^32 / 2^(32 - e_h + G)), e_h - G <= 32
    u32 y[L + 1];
    y[0] = 0u;
    load_desc<L>(y, hm + hb, hps, hl);
    const u32 sh_ = to_mag(y);
    const i64 s = 32 - (eh - G);
    sar(y, s > 32 * (L + 1) ? 32 * (L + 1) : (int)s, 0u);
    cadd(acc, y, sh_);
  }
  return norm<L, L + 2>(acc, G, r);
}

// Two rows of one group at column n (positions D1, D2; has2 false: only the first): cauchy_entry<L, true> for
// each, sharing what depends on the column alone (the grid G, every term's load and shift, the h_inf part);
// only the division by D - D_i and the sum are per row.  Same bits as two cauchy_entry calls.
template <int L>
__device__ __forceinline__ void cauchy_entry2(const int32_t* __restrict__ rm, const i64* __restrict__ re, i64 rps,
                                              i64 N, i64 n, const int64_t* __restrict__ colg,
                                              const int32_t* __restrict__ coldg, int nc, int D1, int D2, bool has2,
                                              const int32_t* __restrict__ hm, const i64* __restrict__ he, i64 hps,
                                              i64 hrow, u32 (&r)[L], i64& e1, i64& e2, int32_t* __restrict__ om,
                                              i64 ops, i64 t1, i64 t2) {
  i64 E = B3D_EZ;
#pragma unroll 4
  for (int i = 0; i < nc; ++i) {
    const i64 ei = re[colg[i] * N + n];
    E = ei > E ? ei : E;
  }
  const i64 hb = hrow * N;
  const bool hz = load_limb(hm + hb, hps, L - 1, n) == 0u;
  const i64 eh = he[hb + n];
  if (E == B3D_EZ && hz) {
    zero(r);
    store(r, om, ops, t1);
    if (has2) store(r, om, ops, t2);
    e1 = e2 = B3D_EZ;
    return;
  }
  const i64 G = E == B3D_EZ ? eh - 32 : (hz ? E - 31 : ((E - 31) > (eh - 32) ? (E - 31) : (eh - 32)));
  u32 acc1[L + 2], acc2[L + 2];
  zero(acc1);
  zero(acc2);
  if (E != B3D_EZ) {
    i64 cn = colg[0];
    int dn = coldg[0];
    for (int i = 0; i < nc; ++i) {
      const i64 c = cn;
      const int d1 = D1 - dn, d2 = has2 ? D2 - dn : 0;
      if (i + 1 < nc) {
        cn = colg[i + 1];
        dn = coldg[i + 1];
      }
      if (d1 == 0 && d2 == 0) continue;
      u32 X[L + 1];
      X[0] = 0u;
      load_desc<L>(X, rm + c * N, rps, n);
      if (X[L] == 0u) continue;
      const i64 sh = G + 31 - re[c * N + n];
      if (sh >= 32 * (L + 1)) continue;
      const u32 fill = 0u - (X[L] >> 31);
      sar(X, (int)sh, fill);
      if (d1 != 0) {
        u32 Q[L + 1];
#pragma unroll
        for (int k = 0; k <= L; ++k) Q[k] = X[k];
        div_floor(Q, (u32)(d1 < 0 ? -d1 : d1));
        cadd(acc1, Q, 0u - (u32)(d1 < 0), fill);
      }
      if (d2 != 0) {
        div_floor(X, (u32)(d2 < 0 ? -d2 : d2));
        cadd(acc2, X, 0u - (u32)(d2 < 0), fill);
      }
    }
  }
  if (!hz) {
    u32 y[L + 1];
    y[0] = 0u;
    load_desc<L>(y, hm + hb, hps, n);
    const u32 sh_ = to_mag(y);
    const i64 s = 32 - (eh - G);
    sar(y, s > 32 * (L + 1) ? 32 * (L + 1) : (int)s, 0u);
    cadd(acc1, y, sh_);
    cadd(acc2, y, sh_);
  }
  e1 = norm<L, L + 2>(acc1, G, r);
  store(r, om, ops, t1);
  if (has2) {
    e2 = norm<L, L + 2>(acc2, G, r);
    store(r, om, ops, t2);
  }
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
    if (load_limb(sm, sps, L - 1, o) == 0u) continue;
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
    const u32 sy = to_mag(y);
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
    cadd(acc, q, sy);
  }
  return norm<L, L + 3>(acc, T, r);
}

// One entry a[e] = sum_{k < kend} T[e][k] v[k] of a product by a matrix of exact integers (kend = e + 1: a
// lower-triangular one; kend = K: a full row), with the multiprecision GEMM's semantics: the column v[0..K)
// aligned to its largest exponent E (magnitudes truncated), the sum of the aligned products exact, one
// rounding.  T[e][k] = +-tm[k] 2^texp (magnitude of TL limbs at tm planes, entry tb + k, plane stride tps;
// sign mask tsg[tb + k]): a constant row of exact integers aligned per row is the same integers scaled by a
// power of two, so the rounded result is identical to the GEMM's.  v[k]: entry vbase + k vstride of the planes
// vm / ve (plane stride vps).
template <int L, int TL>
__device__ __forceinline__ i64 tri_entry(const int32_t* __restrict__ vm, const i64* __restrict__ ve, i64 vps,
                                         i64 vbase, i64 vstride, int K, int kend, const int32_t* __restrict__ tm,
                                         const uint32_t* __restrict__ tsg, i64 tps, i64 tb, i64 texp,
                                         u32 (&out)[L]) {
  i64 E = B3D_EZ;   // zeros carry EZ, below every exponent
  for (int k = 0; k < K; ++k) {
    const i64 x = ve[vbase + k * vstride];
    E = x > E ? x : E;
  }
  if (E == B3D_EZ) {
    zero(out);
    return B3D_EZ;
  }
  u32 acc[L + TL + 1];
  zero(acc);
  for (int k = 0; k < kend; ++k) {
    const i64 ek = ve[vbase + k * vstride];
    const u32 st = tsg[tb + k];
    if (ek == B3D_EZ) continue;
    u32 t[TL];
    load(t, tm, tps, tb + k);
    u32 x[L];
    load(x, vm, vps, vbase + k * vstride);
    const u32 sv = to_mag(x);
    const i64 d = E - ek;
    shr(x, d > 32 * L ? 32 * L : (int)d);
    u32 p[L + TL];
    mul_full(t, x, p);
    cadd(acc, p, st ^ sv);
  }
  return norm<L, L + TL + 1>(acc, texp + E, out);
}

}  // namespace mp
