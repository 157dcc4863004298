// libb3dhip: the JSON writer's decimal text on the GPU (the same text as b3d_format_polys in b3d_format.hip,
// reference blocksvibe/b3djson.py:30-68 fmt_big / fmt_gmpy and :150-177 derivs_text).  Three kernels:
//
//   k_fmt_sizes   per polynomial: its kept length (the truncation of negligible top coefficients, exact
//                 comparisons), 0 for an all-zero polynomial
//   k_fmt_text    per element (a kept coefficient, or the "0" of an all-zero polynomial): the %.{p}g text of
//                 the exact value.  The fraction is held as FW 32-bit limbs with the binary point above the top
//                 limb (the integer part shifted out of the window); each
//                 multiplication by 10^9 carries the next nine digits out of the top limb; the integer part
//                 (the mantissa shifted right) is converted by repeated division by 10^9.  Digits go to a
//                 per-element staging row, are rounded half to even with a sticky flag, and the text is laid
//                 out from them.  Values outside that range (more than 32 FW fraction bits, exponent >= 0)
//                 get length -1 and are formatted by the host formatter.
//   k_fmt_pack    per element: '"' text '"' and ",\n" unless it ends its polynomial, at its offset in the
//                 packed bodies (offsets: an exclusive prefix sum of the piece lengths)
//
// Staging rows and texts are stored byte-transposed (byte c of element k at c * n + k), so a wave's lanes
// write consecutive bytes.
#include "common.hpp"
#include "mp.hpp"
#include "../../include/b3dhip.h"

namespace {

template <int L>
struct Fmt {
  static constexpr int FW = L + 16;   // fraction limbs: values down to 2^(32 L - 2 - 32 FW)
};

// |x| 2^ex < |y| 2^ey for nonzero two's complement entries i, j of the planes (ey += eps)
template <int L>
__device__ __forceinline__ bool lt_abs(const int32_t* m, i64 ps, i64 i, i64 ei, i64 j, i64 ej) {
  u32 a[L], b[L];
  mp::load(a, m, ps, i);
  mp::load(b, m, ps, j);
  mp::to_mag(a);
  mp::to_mag(b);
  const int ba = mp::bitlen(a), bb = mp::bitlen(b);
  const i64 ta = ba + ei, tb = bb + ej;
  if (ta != tb) return ta < tb;
  // same top bit: put both on the larger bit length (the shift is below 32 L)
  if (ba < bb) mp::shl(a, bb - ba);
  if (bb < ba) mp::shl(b, ba - bb);
  for (int k = L - 1; k >= 0; --k)
    if (a[k] != b[k]) return a[k] < b[k];
  return false;
}

// The top limb t of a normalised value (magnitude bit length 32 L - 2) whose bit length it alone certifies:
// t in [2^29, 2^30) (positive) or t in (-2^30, -2^29) (negative; the end points also occur for values of other
// bit lengths).
__device__ __forceinline__ bool top_certifies(u32 t) {
  const int32_t v = (int32_t)t;
  return (v >= (1 << 29) && v < (1 << 30)) || (v > -(1 << 30) && v < -(1 << 29));
}

// lt_abs with the exponents first: when both top limbs certify the bit length 32 L - 2, |x| 2^ex < |y| 2^ey is
// decided by ex < ey unless ex == ey; only then (or for an uncertified value) are the limbs compared.  ti: the
// top limb of entry i, already loaded.
template <int L>
__device__ __forceinline__ bool lt_abs_fast(const int32_t* m, i64 ps, i64 i, u32 ti, i64 ei, i64 j, u32 tj, i64 ej) {
  if (top_certifies(ti) && top_certifies(tj) && ei != ej) return ei < ej;
  return lt_abs<L>(m, ps, i, ei, j, ej);
}

}  // namespace

template <int L>
__global__ __launch_bounds__(256) void k_fmt_sizes(const int32_t* __restrict__ m, const i64* __restrict__ e, i64 ps,
                                                   const int64_t* __restrict__ pbase, i64 npoly, int deg, i64 eps,
                                                   int32_t* __restrict__ sizes) {
  for (i64 q = blockIdx.x * (i64)blockDim.x + threadIdx.x; q < npoly; q += (i64)gridDim.x * blockDim.x) {
    const i64 base = pbase[q];
    int size = 0;
    i64 mx = -1, emx = 0;
    u32 tmx = 0u;
    // the exponents are loaded eight coefficients ahead of the comparisons (they do not depend on them), so a
    // thread waits for one exponent load per group instead of one per coefficient
    constexpr int G = 8;
    i64 ev[G];
#pragma unroll
    for (int u = 0; u < G; ++u) {
      ev[u] = B3D_EZ;
      if (deg - 1 - u >= 0) ev[u] = e[base + deg - 1 - u];
    }
    for (int d = deg - 1; d >= 0; --d) {
      const i64 i = base + d;
      const i64 ei = ev[0];
#pragma unroll
      for (int u = 0; u + 1 < G; ++u) ev[u] = ev[u + 1];
      ev[G - 1] = B3D_EZ;
      if (d - G >= 0) ev[G - 1] = e[base + d - G];
      if (ei == B3D_EZ) continue;   // canonical zero: 0 < 2^eps 0 is false
      const u32 ti = (u32)m[(i64)(L - 1) * ps + i];
      if (mx < 0 || lt_abs_fast<L>(m, ps, mx, tmx, emx, i, ti, ei + eps)) size = d + 1;
      if (mx < 0 || lt_abs_fast<L>(m, ps, mx, tmx, emx, i, ti, ei)) {
        mx = i;
        emx = ei;
        tmx = ti;
      }
    }
    sizes[q] = size;
  }
}

template <int L>
__global__ __launch_bounds__(256) void k_fmt_text(const int32_t* __restrict__ m, const i64* __restrict__ e, i64 ps,
                                                  const int64_t* __restrict__ ent, i64 n, int p, uint8_t* __restrict__ dg,
                                                  uint8_t* __restrict__ tx, int32_t* __restrict__ len) {
  constexpr int FW = Fmt<L>::FW;
  for (i64 k = blockIdx.x * (i64)blockDim.x + threadIdx.x; k < n; k += (i64)gridDim.x * blockDim.x) {
    const i64 i = ent[k];
    // text and digit staging in groups of four bytes: byte c of element k at word (c / 4) * n + k, byte c % 4,
    // so a wave's accesses to one group cover 256 contiguous bytes.  The text is collected a word at a time
    // (its bytes are produced in order) and stored whole.
    u32* const txw = (u32*)tx;
    u32 tw = 0u;
    auto put = [&](int c, char v) {
      tw |= (u32)(uint8_t)v << (8 * (c & 3));
      if ((c & 3) == 3) {
        txw[(i64)(c >> 2) * n + k] = tw;
        tw = 0u;
      }
    };
    auto dat = [&](int c) -> i64 { return (((i64)(c >> 2) * n + k) << 2) + (c & 3); };
    auto dput = [&](int c, int v) { dg[dat(c)] = (uint8_t)v; };
    auto dget = [&](int c) -> int { return dg[dat(c)]; };
    auto dword = [&](int g) -> u32 { return ((const u32*)dg)[(i64)g * n + k]; };
    if (i < 0 || e[i] == B3D_EZ) {   // the "0" of an all-zero polynomial, or a zero coefficient
      txw[k] = (u32)'0';
      len[k] = 1;
      continue;
    }
    u32 x[L];
    mp::load(x, m, ps, i);
    const bool neg = mp::to_mag(x) != 0u;
    const i64 ex = e[i];
    const int bl = mp::bitlen(x);
    const i64 fb = -ex;   // fraction bits
    if (ex >= 0 || fb > 32 * FW) {
      len[k] = -1;          // left to the host formatter
      continue;
    }
    // digits: nd significant digits staged, X the decimal exponent of the first
    int nd = 0, X = -1;
    if (bl > fb) {
      // integer part (the mantissa shifted right by fb, at most L limbs) by repeated division by 10^9; the
      // digits come least significant first (the top chunk without its leading zeros) and are then reversed
      u32 t[L];
#pragma unroll
      for (int q = 0; q < L; ++q) t[q] = x[q];
      mp::shr(t, (int)fb);
      bool more = true;
      while (more) {
        u64 rem = 0;
        u32 any = 0u;
#pragma unroll
        for (int q = L - 1; q >= 0; --q) {
          const u64 cur = (rem << 32) | t[q];
          const u64 qq = cur / 1000000000ull;
          t[q] = (u32)qq;
          rem = cur - qq * 1000000000ull;
          any |= t[q];
        }
        more = any != 0u;
        u32 chunk = (u32)rem;
        for (int r = 0; r < 9; ++r) {
          if (!more && chunk == 0u) break;
          dput(nd++, (int)(chunk % 10u));
          chunk /= 10u;
        }
      }
      for (int a = 0, b = nd - 1; a < b; ++a, --b) {
        const int va = dget(a), vb = dget(b);
        dput(a, vb);
        dput(b, va);
      }
      X = nd - 1;
    }
    // the fraction shifted to the top of FW limbs (the integer part's bits are shifted out of the window)
    u32 F[FW];
#pragma unroll
    for (int q = 0; q < FW; ++q) F[q] = q < L ? x[q] : 0u;
    mp::shl(F, (int)(32 * FW - fb));
    bool nz = false;
#pragma unroll
    for (int q = 0; q < FW; ++q) nz |= F[q] != 0u;
    while (nd < p + 1 && nz) {
      u32 carry = 0;
#pragma unroll
      for (int q = 0; q < FW; ++q) {
        const u64 v = (u64)F[q] * 1000000000u + carry;
        F[q] = (u32)v;
        carry = (u32)(v >> 32);
      }
      nz = false;
#pragma unroll
      for (int q = 0; q < FW; ++q) nz |= F[q] != 0u;
      u32 chunk = carry, dig[9];   // nine digits (constant divisors: multiply-high sequences)
#pragma unroll
      for (int q = 8; q >= 0; --q) {
        dig[q] = chunk % 10u;
        chunk /= 10u;
      }
#pragma unroll
      for (int q = 0; q < 9; ++q) {
        if (nd == 0 && dig[q] == 0u) {
          --X;
          continue;
        }
        dput(nd++, (int)dig[q]);
      }
    }
    const bool sticky = nz;
    // round to p digits, half to even
    if (nd > p) {
      const int r = dget(p);
      bool below = sticky;
      for (int q = p + 1; q < nd && !below; ++q) below = dget(q) != 0;
      nd = p;
      if (r > 5 || (r == 5 && (below || (dget(p - 1) & 1)))) {
        int q = p - 1;
        while (q >= 0 && dget(q) == 9) dput(q--, 0);
        if (q >= 0) {
          dput(q, dget(q) + 1);
        } else {   // 99..9 -> 100..0: one more decimal place
          dput(0, 1);
          ++X;
        }
      }
    }
    while (nd > 1 && dget(nd - 1) == 0) --nd;
    // layout (C's %g): d.ddde+XX outside -4 <= X < p, fixed notation inside
    int c = 0;
    // digits q0 .. q1 - 1 as text, read a staged word (four digits) at a time
    auto digits = [&](int q0, int q1) {
      int q = q0;
      while (q < q1) {
        const u32 w = dword(q >> 2);
        const int qe = q1 < ((q | 3) + 1) ? q1 : ((q | 3) + 1);
        for (; q < qe; ++q) put(c++, (char)('0' + ((w >> (8 * (q & 3))) & 0xFFu)));
      }
    };
    if (neg) put(c++, '-');
    if (X < -4 || X >= p) {
      put(c++, (char)('0' + dget(0)));
      if (nd > 1) {
        put(c++, '.');
        digits(1, nd);
      }
      put(c++, 'e');
      put(c++, X < 0 ? '-' : '+');
      int ax = X < 0 ? -X : X;
      char tmp[12];
      int w = 0;
      do {
        tmp[w++] = (char)('0' + ax % 10);
        ax /= 10;
      } while (ax);
      if (w < 2) tmp[w++] = '0';
      while (w) put(c++, tmp[--w]);
    } else if (X >= 0) {
      digits(0, nd < X + 1 ? nd : X + 1);
      for (int q = nd; q <= X; ++q) put(c++, '0');
      if (nd > X + 1) {
        put(c++, '.');
        digits(X + 1, nd);
      }
    } else {
      put(c++, '0');
      put(c++, '.');
      for (int q = 0; q < -1 - X; ++q) put(c++, '0');
      digits(0, nd);
    }
    if (c & 3) txw[(i64)(c >> 2) * n + k] = tw;   // the last, partial word
    len[k] = c;
  }
}

// the text of element k (l bytes, in groups of four: word (c / 4) * n + k) to o: four words (16 bytes) loaded
// per step, then stored byte by byte (o has any alignment); returns o + l
__device__ __forceinline__ uint8_t* copy_text(const uint8_t* __restrict__ tx, i64 n, i64 k, int l, uint8_t* o) {
  const u32* src = (const u32*)tx + k;
  for (int g0 = 0; 4 * g0 < l; g0 += 4) {
    u32 w[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) w[j] = 4 * (g0 + j) < l ? src[(i64)(g0 + j) * n] : 0u;
#pragma unroll
    for (int j = 0; j < 16; ++j)
      if (4 * g0 + j < l) o[4 * g0 + j] = (uint8_t)(w[j >> 2] >> (8 * (j & 3)));
  }
  return o + l;
}

__global__ __launch_bounds__(256) void k_fmt_pack(const uint8_t* __restrict__ tx, const int32_t* __restrict__ len,
                                                  const uint8_t* __restrict__ last, const int32_t* __restrict__ pre,
                                                  const int32_t* __restrict__ post, const uint8_t* __restrict__ strtab,
                                                  const int64_t* __restrict__ stroff, const int64_t* __restrict__ off,
                                                  i64 n, uint8_t* __restrict__ out) {
  for (i64 k = blockIdx.x * (i64)blockDim.x + threadIdx.x; k < n; k += (i64)gridDim.x * blockDim.x) {
    if (last[k] > 1) continue;   // an unused slot: no text
    uint8_t* o = out + off[k];
    if (pre && pre[k] >= 0)
      for (i64 c = stroff[pre[k]]; c < stroff[pre[k] + 1]; ++c) *o++ = strtab[c];
    const int l = len[k];
    *o++ = '"';
    o = copy_text(tx, n, k, l, o);
    *o++ = '"';
    if (!last[k]) {
      *o++ = ',';
      *o++ = '\n';
    }
    if (post && post[k] >= 0)
      for (i64 c = stroff[post[k]]; c < stroff[post[k] + 1]; ++c) *o++ = strtab[c];
  }
}

// Packed files with coalesced stores: a block takes blockDim.x consecutive elements, whose pieces are one
// contiguous range of out; each thread lays its piece out in LDS (at its offset in the range, shifted so that LDS
// and global addresses agree mod 4), then the block stores the range as aligned words (bytes at its two ends,
// which it shares with its neighbours).  A block whose range exceeds the LDS buffer (cap bytes; long header
// strings) writes each piece straight to global memory as k_fmt_pack does.
__device__ __forceinline__ i64 piece_bytes(const int32_t* len, const uint8_t* last, const int32_t* pre, const int32_t* post,
                                           const int64_t* stroff, i64 k) {
  if (last[k] > 1) return 0;
  i64 b = (len[k] > 0 ? len[k] : 0) + 2 + (last[k] ? 0 : 2);   // len < 0 (left to the host) writes no digits
  if (pre[k] >= 0) b += stroff[pre[k] + 1] - stroff[pre[k]];
  if (post[k] >= 0) b += stroff[post[k] + 1] - stroff[post[k]];
  return b;
}

__device__ __forceinline__ void write_piece(const uint8_t* __restrict__ tx, const int32_t* __restrict__ len,
                                            const uint8_t* __restrict__ last, const int32_t* __restrict__ pre,
                                            const int32_t* __restrict__ post, const uint8_t* __restrict__ strtab,
                                            const int64_t* __restrict__ stroff, i64 n, i64 k, uint8_t* o) {
  if (pre[k] >= 0)
    for (i64 c = stroff[pre[k]]; c < stroff[pre[k] + 1]; ++c) *o++ = strtab[c];
  const int l = len[k];
  *o++ = '"';
  o = copy_text(tx, n, k, l, o);
  *o++ = '"';
  if (!last[k]) {
    *o++ = ',';
    *o++ = '\n';
  }
  if (post[k] >= 0)
    for (i64 c2 = stroff[post[k]]; c2 < stroff[post[k] + 1]; ++c2) *o++ = strtab[c2];
}

__global__ __launch_bounds__(128) void k_fmt_pack_lds(const uint8_t* __restrict__ tx, const int32_t* __restrict__ len,
                                                      const uint8_t* __restrict__ last, const int32_t* __restrict__ pre,
                                                      const int32_t* __restrict__ post,
                                                      const uint8_t* __restrict__ strtab,
                                                      const int64_t* __restrict__ stroff, const int64_t* __restrict__ off,
                                                      i64 n, uint8_t* __restrict__ out, int cap) {
  extern __shared__ uint32_t sw[];
  uint8_t* sb = (uint8_t*)sw;
  const i64 k0 = (i64)blockIdx.x * blockDim.x;
  const i64 k1 = (k0 + blockDim.x < n) ? k0 + blockDim.x : n;
  const i64 k = k0 + threadIdx.x;
  const i64 base = off[k0];
  const i64 endb = off[k1 - 1] + piece_bytes(len, last, pre, post, stroff, k1 - 1);
  const int sh = (int)(base & 3);
  const i64 total = endb - base + sh;             // bytes of the LDS image, from the aligned word below base
  if (total > cap) {
    if (k < k1 && last[k] <= 1) write_piece(tx, len, last, pre, post, strtab, stroff, n, k, out + off[k]);
    return;
  }
  if (k < k1 && last[k] <= 1) write_piece(tx, len, last, pre, post, strtab, stroff, n, k, sb + (off[k] - base + sh));
  __syncthreads();
  uint32_t* ow = (uint32_t*)(out + (base - sh));  // aligned: base - sh is a multiple of 4
  const i64 nw = (total + 3) >> 2;
  for (i64 w = threadIdx.x; w < nw; w += blockDim.x) {
    const i64 lo = 4 * w, hi = lo + 4;
    if (lo >= sh && hi <= total) {
      ow[w] = sw[w];                              // a whole word of this block
    } else {
      for (i64 b = lo; b < hi; ++b)               // the partial words at the two ends
        if (b >= sh && b < total) out[base - sh + b] = sb[b];
    }
  }
}

static inline unsigned grid_for(int64_t n, int th) {
  int64_t b = (n + th - 1) / th;
  if (b > (1 << 22)) b = 1 << 22;
  if (b < 1) b = 1;
  return (unsigned)b;
}

extern "C" {

int b3d_fmt_sizes(int L, const int32_t* m, const int64_t* e, int64_t ps, const int64_t* pbase, int64_t npoly, int deg,
                  int64_t eps_exp, int32_t* sizes, void* stream) {
  if (npoly <= 0) return 0;
  if (deg < 1) {
    b3d::set_error("b3d_fmt_sizes: need deg >= 1");
    return 2;
  }
  if (b3d::check_planes("b3d_fmt_sizes", {ps})) return 2;
  hipStream_t s = (hipStream_t)stream;
  B3D_DISPATCH_L(L, k_fmt_sizes<LL><<<dim3(grid_for(npoly, 256)), dim3(256), 0, s>>>(m, e, ps, pbase, npoly, deg, eps_exp, sizes));
  return b3d::check_launch("b3d_fmt_sizes");
}

int b3d_fmt_text(int L, const int32_t* m, const int64_t* e, int64_t ps, const int64_t* ent, int64_t n, int p,
                 uint8_t* digits, uint8_t* text, int32_t* len, void* stream) {
  if (n <= 0) return 0;
  if (p < 1 || p > 2000) {
    b3d::set_error("b3d_fmt_text: need 1 <= p <= 2000");
    return 2;
  }
  if (b3d::check_planes("b3d_fmt_text", {ps})) return 2;
  hipStream_t s = (hipStream_t)stream;
  B3D_DISPATCH_L(L, k_fmt_text<LL><<<dim3(grid_for(n, 256)), dim3(256), 0, s>>>(m, e, ps, ent, n, p, digits, text, len));
  return b3d::check_launch("b3d_fmt_text");
}

int b3d_fmt_pack(const uint8_t* text, const int32_t* len, const uint8_t* last, const int64_t* off, int64_t n,
                 uint8_t* out, void* stream) {
  if (n <= 0) return 0;
  hipStream_t s = (hipStream_t)stream;
  k_fmt_pack<<<dim3(grid_for(n, 256)), dim3(256), 0, s>>>(text, len, last, nullptr, nullptr, nullptr, nullptr, off,
                                                          n, out);
  return b3d::check_launch("b3d_fmt_pack");
}

int b3d_fmt_pack_files(const uint8_t* text, const int32_t* len, const uint8_t* last, const int32_t* pre,
                       const int32_t* post, const uint8_t* strtab, const int64_t* stroff, int64_t nstr,
                       const int64_t* off, int64_t n, uint8_t* out, void* stream) {
  if (n <= 0) return 0;
  if (nstr < 0 || (nstr > 0 && (!strtab || !stroff))) {
    b3d::set_error("b3d_fmt_pack_files: need a string table (strtab, stroff[nstr + 1]) for nstr > 0");
    return 2;
  }
  hipStream_t s = (hipStream_t)stream;
  // LDS image per block: 128 pieces of up to ~470 bytes (p + 20 at p = 309 is 329), within the 64 KiB a
  // workgroup may allocate
  constexpr int kThreads = 128, kCap = 60 * 1024;
  const int64_t blocks = (n + kThreads - 1) / kThreads;
  if (blocks > (int64_t(1) << 31) - 1) {
    b3d::set_error("b3d_fmt_pack_files: too many elements");
    return 2;
  }
  k_fmt_pack_lds<<<dim3((unsigned)blocks), dim3(kThreads), kCap, s>>>(text, len, last, pre, post, strtab, stroff, off,
                                                                       n, out, kCap);
  return b3d::check_launch("b3d_fmt_pack_files");
}

}  // extern "C"
