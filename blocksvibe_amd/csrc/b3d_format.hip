// libb3dhip host side: blocks_3d's decimal output of MPF values (reference blocksvibe/b3djson.py:30-68
// fmt_big / fmt_gmpy, :150-177 derivs_text; blocks_3d prints mpf_float at stream precision p in the general
// format).
//
// b3d_format_numbers: the exact value m 2^e of each entry (two's complement limb planes, as the kernels store
// them) as p significant decimal digits, rounded to nearest with ties to even, trailing zeros removed, fixed
// notation for decimal exponents -4 <= X < p, else d.ddde+XX.  Digits come from exact multiprecision
// arithmetic on 64-bit limbs: the integer part by repeated division by 10^19, the fraction (shifted so that
// the binary point falls between limbs) by repeated multiplication by 10^19, each step carrying the next 19
// digits out of the top limb, until p + 1 significant digits are known; a sticky flag records whether
// anything nonzero remains below them.  Entries are independent, so the work is split over host threads.
// b3d_poly_sizes: derivs_text's truncation of negligible top coefficients of each polynomial (keep up to the
// last coefficient c with max|later| < 2^eps |c|), by exact comparisons; 0 for an all-zero polynomial.
// b3d_format_polys: both at once, each polynomial's kept coefficients as the JSON list body blocks_3d
// writes ("c0",\n"c1",... or "0"), split over host threads by polynomial.
#include <stdint.h>
#include <stdio.h>

#include <algorithm>
#include <string>
#include <thread>
#include <vector>

#include "common.hpp"
#include "../../include/b3dhip.h"

namespace {

typedef uint32_t u32;
typedef uint64_t u64;
typedef unsigned __int128 u128;

const u64 TEN19 = 10000000000000000000ull;   // the largest power of ten below 2^64

// little-endian base 2^64 magnitude with its used length (buffers are reused across entries)
struct Nat {
  std::vector<u64> d;
  size_t n = 0;
  void trim() {
    while (n && d[n - 1] == 0u) --n;
  }
  void fit(size_t k) {
    if (d.size() < k) d.resize(k);
  }
};

// two decimal digits of each value 0..99
struct Pairs {
  char d[200];
  Pairs() {
    for (int i = 0; i < 100; ++i) {
      d[2 * i] = (char)('0' + i / 10);
      d[2 * i + 1] = (char)('0' + i % 10);
    }
  }
};
const Pairs PAIRS;

// the 2k decimal digits of x < 10^(2k), zero padded, two at a time
template <int K>
inline void digits_pairs(u32 x, char* o) {
  for (int k = K - 1; k >= 0; --k) {
    const u32 q = x / 100u, r = x - 100u * q;
    o[2 * k] = PAIRS.d[2 * r];
    o[2 * k + 1] = PAIRS.d[2 * r + 1];
    x = q;
  }
}

// the 19 decimal digits of x < 10^19, zero padded: 10^19 = 10^11 * 10^8 splits x into a high part below 10^8
// (then 10^8 = 10^8) and a low part below 10^11, each converted two digits at a time
void digits19(u64 x, char* o) {
  const u64 hi = x / 100000000000ull;            // < 10^8
  const u64 lo = x - hi * 100000000000ull;        // < 10^11
  digits_pairs<4>((u32)hi, o);                    // 8 digits
  const u64 lh = lo / 1000000ull;                 // < 10^5
  const u32 ll = (u32)(lo - lh * 1000000ull);     // < 10^6
  o[8] = (char)('0' + (u32)lh / 10000u);          // 5 digits of lh
  digits_pairs<2>((u32)lh % 10000u, o + 9);
  digits_pairs<3>(ll, o + 13);                    // 6 digits
}

// n /= 10^19, returning the remainder
u64 div19(Nat& n) {
  u128 r = 0;
  for (size_t k = n.n; k-- > 0;) {
    const u128 v = (r << 64) | n.d[k];
    n.d[k] = (u64)(v / TEN19);
    r = v % TEN19;
  }
  n.trim();
  return (u64)r;
}

// decimal digits of an integer (nonzero) appended to sig; n is consumed
void integer_digits(Nat& n, std::string& sig, std::vector<u64>& chunks) {
  chunks.clear();
  while (n.n) chunks.push_back(div19(n));
  char buf[20];
  digits19(chunks.back(), buf);
  int k = 0;
  while (k < 18 && buf[k] == '0') ++k;
  sig.append(buf + k, 19 - k);
  for (size_t c = chunks.size() - 1; c-- > 0;) {
    digits19(chunks[c], buf);
    sig.append(buf, 19);
  }
}

// dst = src << s (bits), base 2^64
void shl_into(const Nat& src, int64_t s, Nat& dst) {
  const int64_t w = s / 64, b = s % 64;
  dst.fit(src.n + w + 1);
  std::fill(dst.d.begin(), dst.d.begin() + src.n + w + 1, 0ull);
  for (size_t k = 0; k < src.n; ++k) {
    dst.d[k + w] |= src.d[k] << b;
    if (b) dst.d[k + w + 1] |= src.d[k] >> (64 - b);
  }
  dst.n = src.n + w + 1;
  dst.trim();
}

struct Scratch {
  Nat a, t;
  std::string sig;
  std::vector<u64> chunks;
};

// the %.{p}g text of a 2^e (a > 0, a.n limbs), with a leading '-' if neg, appended to out
void format_one(Scratch& S, int64_t e, int p, bool neg, std::string& out) {
  Nat &a = S.a, &t = S.t;
  std::string& sig = S.sig;   // significant digits from the first nonzero one
  sig.clear();
  int64_t X;                  // decimal exponent of the first significant digit
  bool sticky = false;        // a nonzero remainder below the digits in sig
  if (e >= 0) {
    shl_into(a, e, t);
    integer_digits(t, sig, S.chunks);
    X = (int64_t)sig.size() - 1;
  } else {
    const int64_t fb = -e;                    // fraction bits
    const int64_t bl = 64 * (int64_t)(a.n - 1) + (64 - __builtin_clzll(a.d[a.n - 1]));
    X = -1;                                   // lowered by each leading zero of the fraction
    if (bl > fb) {                            // integer part a >> fb
      const int64_t w = fb / 64, b = fb % 64;
      t.fit(a.n - w);
      t.n = a.n - w;
      for (size_t k = 0; k < t.n; ++k)
        t.d[k] = (a.d[k + w] >> b) | (b && k + w + 1 < a.n ? a.d[k + w + 1] << (64 - b) : 0ull);
      t.trim();
      integer_digits(t, sig, S.chunks);
      X = (int64_t)sig.size() - 1;
    }
    // fraction a mod 2^fb, shifted up so that the binary point falls on a limb boundary: W limbs below it;
    // each multiplication by 10^19 then carries the next 19 digits out of the top limb
    const int64_t W = (fb + 63) / 64, up = 64 * W - fb;
    t.fit(W + 1);
    std::fill(t.d.begin(), t.d.begin() + W + 1, 0ull);
    for (size_t k = 0; k < a.n; ++k) {
      const int64_t pos = 64 * (int64_t)k + up;   // bit position of limb k after the shift
      if (pos >= 64 * W) break;
      const int64_t w = pos / 64, b = pos % 64;
      t.d[w] |= a.d[k] << b;
      if (b && w + 1 < W) t.d[w + 1] |= a.d[k] >> (64 - b);
    }
    // limbs [lo, hi) hold the nonzero part; a carry out of limb hi - 1 below the point grows it upwards
    // (zero digits: the fraction is still below 10^-19 of the current scale), out of limb W - 1 it is the
    // next 19 digits
    size_t lo = 0, hi = (size_t)W;
    while (hi > 0 && t.d[hi - 1] == 0u) --hi;
    while (lo < hi && t.d[lo] == 0u) ++lo;
    char buf[20];
    while ((int64_t)sig.size() < p + 1 && lo < hi) {
      u64 carry = 0;
      for (size_t k = lo; k < hi; ++k) {
        const u128 v = (u128)t.d[k] * TEN19 + carry;
        t.d[k] = (u64)v;
        carry = (u64)(v >> 64);
      }
      u64 chunk = 0;
      if (hi < (size_t)W) {
        if (carry) t.d[hi++] = carry;
      } else {
        chunk = carry;
      }
      while (lo < hi && t.d[lo] == 0u) ++lo;
      digits19(chunk, buf);
      int k = 0;
      if (sig.empty()) {
        while (k < 19 && buf[k] == '0') ++k;
        X -= k;
      }
      sig.append(buf + k, 19 - k);
    }
    sticky = lo < hi;
  }
  // round to p digits, ties to even
  if ((int64_t)sig.size() > p) {
    const char r = sig[p];
    bool below = sticky;
    for (size_t k = p + 1; k < sig.size() && !below; ++k) below = sig[k] != '0';
    sig.resize(p);
    if (r > '5' || (r == '5' && (below || ((sig[p - 1] - '0') & 1)))) {
      int k = p - 1;
      while (k >= 0 && sig[k] == '9') sig[k--] = '0';
      if (k >= 0) {
        ++sig[k];
      } else {                                // 99..9 -> 100..0: one more decimal place
        sig[0] = '1';
        ++X;
      }
    }
  }
  size_t nd = sig.size();
  while (nd > 1 && sig[nd - 1] == '0') --nd;
  if (neg) out += '-';
  if (X < -4 || X >= p) {
    out += sig[0];
    if (nd > 1) {
      out += '.';
      out.append(sig, 1, nd - 1);
    }
    char eb[32];
    const int len = snprintf(eb, sizeof eb, "e%c%02lld", X < 0 ? '-' : '+', (long long)(X < 0 ? -X : X));
    out.append(eb, len);
  } else if (X >= 0) {
    if ((int64_t)nd > X + 1) {
      out.append(sig, 0, X + 1);
      out += '.';
      out.append(sig, X + 1, nd - (X + 1));
    } else {
      out.append(sig, 0, nd);
      out.append(X + 1 - nd, '0');
    }
  } else {
    out += "0.";
    out.append(-1 - X, '0');
    out.append(sig, 0, nd);
  }
}

// magnitude (base 2^64) and sign of entry i of L two's complement 32-bit limb planes
bool entry(const int32_t* m, int64_t ps, int L, int64_t i, Nat& a) {
  u32 v[64];
  for (int k = 0; k < L; ++k) v[k] = (u32)m[k * ps + i];
  const bool neg = (v[L - 1] >> 31) != 0u;
  if (neg) {
    u64 c = 1;
    for (int k = 0; k < L; ++k) {
      const u64 s = (u64)(u32)~v[k] + c;
      v[k] = (u32)s;
      c = s >> 32;
    }
  }
  const int n64 = (L + 1) / 2;
  a.fit(n64);
  for (int k = 0; k < n64; ++k) a.d[k] = (u64)v[2 * k] | (2 * k + 1 < L ? (u64)v[2 * k + 1] << 32 : 0ull);
  a.n = n64;
  a.trim();
  return neg;
}

int64_t bitlen(const Nat& n) {
  return n.n ? 64 * (int64_t)(n.n - 1) + (64 - __builtin_clzll(n.d[n.n - 1])) : 0;
}

// x 2^ex < y 2^ey for nonzero magnitudes x, y
bool less(const Nat& x, int64_t ex, const Nat& y, int64_t ey, Nat& tx, Nat& ty) {
  const int64_t bx = bitlen(x) + ex, by = bitlen(y) + ey;
  if (bx != by) return bx < by;
  // equal top bits: the shift below is under the longer one's length
  shl_into(x, ex > ey ? ex - ey : 0, tx);
  shl_into(y, ey > ex ? ey - ex : 0, ty);
  if (tx.n != ty.n) return tx.n < ty.n;
  for (size_t k = tx.n; k-- > 0;)
    if (tx.d[k] != ty.d[k]) return tx.d[k] < ty.d[k];
  return false;
}

// kept length of the polynomial with coefficients at entries base .. base + deg - 1 (0: all zero)
int poly_size(const int32_t* m, const int64_t* e, int64_t ps, int L, int64_t base, int deg, int64_t eps_exp,
              Nat& mx, Nat& c, Nat& t1, Nat& t2) {
  int size = 0;
  int64_t emx = 0;
  mx.n = 0;   // running maximum of |later coefficients| (n = 0: none yet)
  for (int d = deg - 1; d >= 0; --d) {
    const int64_t i = base + d;
    entry(m, ps, L, i, c);
    if (!c.n) continue;   // 0 < 2^eps 0 is false: zeros never set the size
    if (!mx.n || less(mx, emx, c, e[i] + eps_exp, t1, t2)) size = d + 1;
    if (!mx.n || less(mx, emx, c, e[i], t1, t2)) {
      std::swap(mx, c);
      emx = e[i];
    }
  }
  return size;
}

// split [0, n) over threads; work(t, i0, i1) fills per-thread text and ends; then the texts are concatenated
// into out (cap bytes) and end[] made absolute
template <class F>
int run_split(const char* what, int64_t n, int T, char* out, int64_t cap, int64_t* end, F work) {
  std::vector<std::string> text(T);
  std::vector<std::vector<int64_t>> ends(T);
  auto job = [&](int t) { work(n * t / T, n * (t + 1) / T, text[t], ends[t]); };
  std::vector<std::thread> pool;
  for (int t = 1; t < T; ++t) pool.emplace_back(job, t);
  job(0);
  for (auto& th : pool) th.join();
  int64_t pos = 0, i = 0;
  for (int t = 0; t < T; ++t) {
    if (pos + (int64_t)text[t].size() > cap) {
      b3d::set_error(std::string(what) + ": output buffer too small");
      return 2;
    }
    std::copy(text[t].begin(), text[t].end(), out + pos);
    for (int64_t x : ends[t]) end[i++] = pos + x;
    pos += (int64_t)text[t].size();
  }
  return 0;
}

int nthreads(int threads, int64_t n, int64_t per) {
  int t = threads > 0 ? threads : (int)std::max(1u, std::thread::hardware_concurrency());
  t = (int)std::min<int64_t>(t, std::max<int64_t>(1, n / per));
  return std::max(1, std::min(t, 256));
}

}  // namespace

extern "C" {

int b3d_format_numbers(int L, const int32_t* m, const int64_t* e, int64_t ps, int64_t n, int p, char* out,
                       int64_t cap, int64_t* end, int threads) {
  if (n <= 0) return 0;
  if (L < 1 || L > 64 || p < 1 || p > 2000) {
    b3d::set_error("b3d_format_numbers: need 1 <= L <= 64 and 1 <= p <= 2000");
    return 2;
  }
  return run_split("b3d_format_numbers", n, nthreads(threads, n, 64), out, cap, end,
                   [&](int64_t i0, int64_t i1, std::string& text, std::vector<int64_t>& ends) {
                     text.reserve((size_t)(i1 - i0) * (p + 8));
                     ends.reserve(i1 - i0);
                     Scratch S;
                     for (int64_t i = i0; i < i1; ++i) {
                       const bool neg = entry(m, ps, L, i, S.a);
                       if (S.a.n)
                         format_one(S, e[i], p, neg, text);
                       else
                         text += '0';
                       ends.push_back((int64_t)text.size());
                     }
                   });
}

int b3d_poly_sizes(int L, const int32_t* m, const int64_t* e, int64_t ps, int64_t npoly, int deg, int64_t eps_exp,
                   int32_t* sizes) {
  if (npoly <= 0) return 0;
  if (L < 1 || L > 64 || deg < 1) {
    b3d::set_error("b3d_poly_sizes: need 1 <= L <= 64 and deg >= 1");
    return 2;
  }
  Nat mx, c, t1, t2;
  for (int64_t q = 0; q < npoly; ++q) sizes[q] = poly_size(m, e, ps, L, q * deg, deg, eps_exp, mx, c, t1, t2);
  return 0;
}

int b3d_format_polys(int L, const int32_t* m, const int64_t* e, int64_t ps, int64_t npoly, int deg, int p,
                     int64_t eps_exp, char* out, int64_t cap, int64_t* end, int threads) {
  if (npoly <= 0) return 0;
  if (L < 1 || L > 64 || deg < 1 || p < 1 || p > 2000) {
    b3d::set_error("b3d_format_polys: need 1 <= L <= 64, deg >= 1 and 1 <= p <= 2000");
    return 2;
  }
  return run_split("b3d_format_polys", npoly, nthreads(threads, npoly * deg, 64), out, cap, end,
                   [&](int64_t q0, int64_t q1, std::string& text, std::vector<int64_t>& ends) {
                     text.reserve((size_t)(q1 - q0) * deg * (p + 12));
                     ends.reserve(q1 - q0);
                     Scratch S;
                     Nat mx, c, t1, t2;
                     for (int64_t q = q0; q < q1; ++q) {
                       const int size = poly_size(m, e, ps, L, q * deg, deg, eps_exp, mx, c, t1, t2);
                       if (!size) text += "\"0\"";
                       for (int d = 0; d < size; ++d) {
                         const int64_t i = q * deg + d;
                         text += d ? ",\n\"" : "\"";
                         const bool neg = entry(m, ps, L, i, S.a);
                         if (S.a.n)
                           format_one(S, e[i], p, neg, text);
                         else
                           text += '0';
                         text += '"';
                       }
                       ends.push_back((int64_t)text.size());
                     }
                   });
}

}  // extern "C"
