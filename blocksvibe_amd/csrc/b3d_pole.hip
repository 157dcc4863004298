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
// Work item = (task, block of 64 rows), one workgroup of 8 waves: the rows' aligned bytes (A operand,
// 16-byte chunks) are staged in LDS once for the task's S constant rows, with every constant row's digits
// and corrections; for each constant row, the math waves build its digits reversed in four byte-shifted
// copies (so that every B fragment d_j[c - i0 - 0..15] is four aligned LDS dwords) and run
// v_mfma_i32_16x16x64_i8 over 4 row tiles x (8L / 16) column tiles (dealt round robin to waves 0-3), while
// waves 4-7 carry the previous constant row's column sums (four segments per row) and round them.
#include "common.hpp"
#include "mp.hpp"
#include "../../include/b3dhip.h"

typedef int v4i __attribute__((ext_vector_type(4)));

template <int L>
struct PoleMma {
  static constexpr int CPJ = (4 * L + 15) / 16;              // 16-byte chunks per operand (4L bytes)
  static constexpr int NCOL = 8 * L;                         // product columns
  static constexpr int NCT = NCOL / 16;                      // column tiles (L even)
  static constexpr int MT = (NCT + 3) / 4;                   // column tiles per math wave
  static constexpr int P0 = NCOL - 1;                        // reversed digits: R[p] = d[P0 - p]
  static constexpr int RL4 = (NCOL + 16 * CPJ - 1 + 3) / 4 * 4;  // bytes per reversed copy
  static constexpr int DS = 4 * L + 4;                       // table bytes per constant entry
  static constexpr int CS = NCOL + 1;                        // LDS dwords per row of column sums (odd)
  static constexpr int LC = 64 * CS * 4;                     // column sums of 64 rows
  static constexpr int SEGL = L / 2;                         // limbs per carry segment (4 segments)
  // K order of the products: the 16-byte chunk ci of operand j sits in K-step ks, lane group g (a K-step is
  // 4 chunks).  Whole groups of 4 operands share a K-step per ci (ks = (j / 4) CPJ + ci, g = j % 4), so the
  // K-step's Toeplitz tiles are zero outside column tiles ci .. ci + CPJ; the remaining operands follow with
  // their chunks packed in order (slot = (j - 4 FG) CPJ + ci, ks = FG CPJ + slot / 4, g = slot % 4).
  static constexpr int nks(int J) { return (J / 4) * CPJ + ((J % 4) * CPJ + 3) / 4; }
  // LDS of a task with J operands: column sums, the 64 rows' operands, 4 reversed copies of one constant
  // row's digits, the digits and corrections of all J constant rows; then (own arrays) the constant rows'
  // exponents and the row exponents
  static constexpr int bytes(int J) {
    return LC + 64 * (64 * nks(J) + 16) + 4 * J * RL4 + J * J * DS + J * NCOL * 4 + 16 * 8 + 64 * 8;
  }
  static constexpr int jmax() {
    int j = 1;
    while (j < 16 && bytes(j + 1) <= 160 * 1024) ++j;
    return j;
  }
  static constexpr int JB = jmax();                          // largest S the MFMA path takes (6 at L = 34)
  static constexpr int XROW = 64 * nks(JB) + 16;             // LDS bytes per staged row
  static constexpr int LDS = bytes(JB) - 16 * 8 - 64 * 8;
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
// exponent Ce[cr].  xrow / orow null: the task index itself.  Requires S <= JB.
// Work item = (task, block of 64 rows): the operands X are the same for the S constant rows s0, so they are
// staged once, with the digits and corrections of all S rows; then waves 0-3 run the products of row s0
// (reversed digit copies, MFMA, column sums into LDS) while waves 4-7 carry and round the column sums of
// row s0 - 1.
template <int L>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(2))) void k_pole_mma(
    const int32_t* __restrict__ Xm, const i64* __restrict__ Xe, i64 xps, const int64_t* __restrict__ xrow,
    const int32_t* __restrict__ ctab, const int8_t* __restrict__ d8, const int32_t* __restrict__ corr,
    const i64* __restrict__ Ce, const int64_t* __restrict__ orow, i64 T, int S, i64 B, int32_t* __restrict__ Om,
    i64* __restrict__ Oe, i64 ops) {
  using P = PoleMma<L>;
  __shared__ __align__(16) unsigned char lds[P::LDS];
  __shared__ i64 Cex[16];                                             // [S s0] constant row exponents
  __shared__ i64 Erow[64];
  int32_t* const C = (int32_t*)lds;                                   // [64][CS]
  unsigned char* const X = lds + P::LC;                               // [64][XROW]
  unsigned char* const Dc = X + 64 * P::XROW;                         // [JB operands][4 copies][RL4]
  int8_t* const Dr = (int8_t*)(Dc + 4 * P::JB * P::RL4);              // [S s0][S j][DS]
  int32_t* const Cr = (int32_t*)(Dr + P::JB * P::JB * P::DS);         // [S s0][NCOL] (DS: whole dwords)
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const bool math = w < 4;
  const int r = lane & 15, g = lane >> 4;
  const i64 SB = (i64)S * B;
  const i64 NRB = (SB + 63) / 64;
  const i64 items = T * NRB;
  const int nks = P::nks(S);
  for (i64 it = blockIdx.x; it < items; it += gridDim.x) {
    const i64 task = it / NRB, rb = it - task * NRB;
    const i64 cr0 = (i64)ctab[task] * S;
    const i64 xb = (xrow ? xrow[task] : task) * S * SB;
    const i64 ob = (orow ? orow[task] : task) * S * SB;
    const i64 n = rb * 64 + lane;                  // this thread's row
    const bool valid = n < SB;
    const i64 u = valid ? n / B : 0, k = valid ? n - (n / B) * B : 0;
    const i64 xrowbase = xb + u * SB + k;
    i64 E = B3D_EZ;                                // zero entries carry EZ
#pragma unroll
    for (int j = 0; j < 16; ++j)                   // S <= 16: every exponent load in flight at once
      if (j < S) {
        const i64 ej = valid ? Xe[xrowbase + j * B] : B3D_EZ;
        E = ej > E ? ej : E;
      }
    __syncthreads();                               // the previous item is done with LDS
    if (w == 0) Erow[lane] = E;
    // operands: row `lane`, operand jl = w, w + 8, ..: bytes u ^ 0x80 except the signed top byte, zero padding
    const int FG = S / 4;
    for (int jl = w; jl < S; jl += 8) {
      unsigned char* row = X + lane * P::XROW;
      u32 x[L];
      mp::zero(x);
      if (valid && E != B3D_EZ) {
        const i64 idx = xrowbase + (i64)jl * B;
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
      for (int ci = 0; ci < P::CPJ; ++ci) {
        int ks, gg;
        if (jl < 4 * FG) {
          ks = (jl >> 2) * P::CPJ + ci;
          gg = jl & 3;
        } else {
          const int slot = (jl - 4 * FG) * P::CPJ + ci;
          ks = FG * P::CPJ + (slot >> 2);
          gg = slot & 3;
        }
        u32* dst = (u32*)(row + (ks * 4 + gg) * 16);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const int qd = 4 * ci + t;
          dst[t] = qd < L - 1 ? (x[qd] ^ 0x80808080u) : (qd == L - 1 ? (x[qd] ^ 0x00808080u) : 0u);
        }
      }
    }
    // the task's S x S digit rows and S correction rows (contiguous in the tables), S row exponents
    {
      const u32* src = (const u32*)(d8 + cr0 * S * P::DS);
      for (int q = tid; q < S * S * (P::DS / 4); q += 512) ((u32*)Dr)[q] = src[q];
      const int32_t* cs = corr + cr0 * P::NCOL;
      for (int q = tid; q < S * P::NCOL; q += 512) Cr[q] = cs[q];
      if (tid < S) Cex[tid] = Ce[cr0 + tid];
    }
    // reversed digit copies of constant row s0 (math waves): copy sft, byte p' = d[P0 - p' - sft], zero outside
    // 0 .. 4L.  P0 = 8L - 1 = 3 mod 4, so the reversed string's dword m is the byte swap of the digits' dword
    // 2L - 1 - m, and copy sft's dword wd is the byte funnel of reversed dwords wd + 1 : wd by sft.
    auto build_copies = [&](int s0) {
      for (int q = tid; q < S * (P::RL4 / 4); q += 256) {
        const int jl = q / (P::RL4 / 4), wd = q - jl * (P::RL4 / 4);
        const u32* dw = (const u32*)(Dr + (s0 * S + jl) * P::DS);    // DS / 4 = L + 1 dwords, digit 4L in the last
        const int i0 = 2 * L - 1 - wd, i1 = 2 * L - 2 - wd;           // digit dwords of reversed dwords wd, wd + 1
        const u32 r0 = (i0 >= 0 && i0 <= L) ? __builtin_bswap32(dw[i0]) : 0u;
        const u32 r1 = (i1 >= 0 && i1 <= L) ? __builtin_bswap32(dw[i1]) : 0u;
        u32* dst = (u32*)(Dc + jl * 4 * P::RL4) + wd;
        dst[0] = r0;
#pragma unroll
        for (int sft = 1; sft < 4; ++sft) dst[sft * (P::RL4 / 4)] = __builtin_amdgcn_alignbyte(r1, r0, (u32)sft);
      }
    };
    // rounding of constant row s0's column sums (epilogue waves): thread = (row, segment), the 4 segments of a
    // row in 4 neighbouring lanes of one wave.  Segment sg carries columns [2L sg, 2L sg + 2L) into L/2 limbs
    // and a carry; the carries ripple up through the segments (lane shuffles), then the magnitude, its bit
    // length and the rounding window (read from LDS at the row's bit offset) give norm(W) exactly as
    // mp::norm<L, 2L + 2> does.
    // value of lane k of this lane's quad (the 4 segments of a row), by DPP quad permutation
    auto quad32 = [](u32 x, int k) -> u32 {
      switch (k) {
        case 0: return (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x00, 0xF, 0xF, false);
        case 1: return (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0x55, 0xF, 0xF, false);
        case 2: return (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0xAA, 0xF, 0xF, false);
        default: return (u32)__builtin_amdgcn_update_dpp(0, (int)x, 0xFF, 0xF, 0xF, false);
      }
    };
    auto quad64 = [&](i64 x, int k) -> i64 {
      return (i64)(((u64)quad32((u32)((u64)x >> 32), k) << 32) | quad32((u32)x, k));
    };
    auto round_row = [&](int s0) {
      const int et = tid - 256, er = et >> 2, sg = et & 3;
      const i64 en = rb * 64 + er;
      int32_t* crow = C + er * P::CS;
      u32 v[P::SEGL];
      i64 carry = 0;
      {
        i64 part[P::SEGL];
#pragma unroll
        for (int q = 0; q < P::SEGL; ++q) {
          i64 x = 0;
#pragma unroll
          for (int b = 0; b < 4; ++b) x += (i64)crow[2 * L * sg + 4 * q + b] * ((i64)1 << (8 * b));
          part[q] = x;
        }
#pragma unroll
        for (int q = 0; q < P::SEGL; ++q) {
          const i64 x = part[q] + carry;
          v[q] = (u32)x;
          carry = x >> 32;
        }
      }
      // carries up through the segments: segment k adds the final carry-out of segment k - 1
      i64 cout = carry;   // this segment's carry-out once its carry-in is added
#pragma unroll
      for (int k = 1; k < 4; ++k) {
        const i64 cin = quad64(cout, k - 1);
        if (sg == k) {
          const u32 ext = cin < 0 ? 0xFFFFFFFFu : 0u;
          u32 cy = 0u;
#pragma unroll
          for (int q = 0; q < P::SEGL; ++q)
            v[q] = mp::addc(v[q], q == 0 ? (u32)cin : (q == 1 ? (u32)((u64)cin >> 32) : ext), cy);
          cout = carry + (i64)cy - (cin < 0 ? 1 : 0);
        }
      }
      const i64 top = quad64(cout, 3);   // W's limbs 2L, 2L + 1
      const bool neg = top < 0;
      // magnitude: ~W + 1; the +1 reaches segment k when every lower segment of W is zero
      u32 anyv = 0u;
#pragma unroll
      for (int q = 0; q < P::SEGL; ++q) anyv |= v[q];
      const unsigned zmask = (anyv == 0u ? 1u : 0u) << sg;
      const unsigned zall = quad32(zmask, 0) | quad32(zmask, 1) | quad32(zmask, 2) | quad32(zmask, 3);
      const unsigned below = (1u << sg) - 1u;
      if (neg) {
        u32 cy = ((zall & below) == below) ? 1u : 0u;
#pragma unroll
        for (int q = 0; q < P::SEGL; ++q) v[q] = mp::addc(~v[q], 0u, cy);
      }
      u64 tmag = (u64)top;
      if (neg) tmag = ~tmag + ((zall & 15u) == 15u ? 1u : 0u);
      // bit length of |W|
      int hb = 0;
#pragma unroll
      for (int q = 0; q < P::SEGL; ++q)
        if (v[q]) hb = 32 * (sg * P::SEGL + q) + 32 - __builtin_clz(v[q]);
      hb = max(max((int)quad32((u32)hb, 0), (int)quad32((u32)hb, 1)), max((int)quad32((u32)hb, 2), (int)quad32((u32)hb, 3)));
      const int bl = tmag ? 64 * L + 64 - __builtin_clzll(tmag) : hb;
      // |W| limbs into the row's LDS (its column sums are consumed), then one lane per row takes the window
      u32* mrow = (u32*)crow;
#pragma unroll
      for (int q = 0; q < P::SEGL; ++q) mrow[sg * P::SEGL + q] = v[q];
      if (sg == 3) {
        mrow[2 * L] = (u32)tmag;
        mrow[2 * L + 1] = (u32)(tmag >> 32);
      }
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      if (sg == 0 && en < SB) {
        const i64 ea = Cex[s0], Er = Erow[er];
        u32 out[L];
        i64 e = B3D_EZ;
        mp::zero(out);
        if (bl != 0 && Er != B3D_EZ && ea != B3D_EZ) {
          // floor(|W| / 2^(s - 1)), s = bl - NB (negative s: |W| 2^(1 - s), exact), NB + 1 bits in L limbs
          int s = bl - (32 * L - 2);
          const int sb = s - 1;
          const int w0 = sb >> 5;                    // floor, also for negative sb
          const u32 bs = (u32)(sb & 31);
          u32 t[L];
          u32 lo = (w0 >= 0 && w0 < 2 * L + 2) ? mrow[w0] : 0u;
#pragma unroll
          for (int k = 0; k < L; ++k) {
            const int i = w0 + k + 1;
            const u32 hi = (i >= 0 && i < 2 * L + 2) ? mrow[i] : 0u;
            t[k] = b3d_alignbit(hi, lo, bs);
            lo = hi;
          }
          u32 cy = 1u;
#pragma unroll
          for (int k = 0; k < L; ++k) t[k] = mp::addc(t[k], 0u, cy);
#pragma unroll
          for (int k = 0; k < L; ++k) out[k] = b3d_alignbit(k + 1 < L ? t[k + 1] : 0u, t[k], 1u);
          const bool over = (out[L - 1] >> 30) != 0u;
          if (over) {
#pragma unroll
            for (int k = 0; k < L; ++k) out[k] = k == L - 1 ? (1u << 29) : 0u;
          }
          s += over ? 1 : 0;
          mp::cneg(out, neg ? 0xFFFFFFFFu : 0u);
          e = ea + Er + s;
        }
        const i64 o = ob + s0 * SB + en;
        mp::store(out, Om, ops, o);
        Oe[o] = e;
      }
    };
    __syncthreads();
    if (math) build_copies(0);
    __syncthreads();
    for (int step = 0; step <= S; ++step) {
      v4i acc[4][P::MT];
      // phase 1: math waves multiply constant row `step`, epilogue waves round row step - 1
      if (math) {
        if (step < S) {
          const int32_t* cc = Cr + step * P::NCOL;
#pragma unroll
          for (int mt = 0; mt < P::MT; ++mt) {
            const int ct = w + 4 * mt < P::NCT ? w + 4 * mt : P::NCT - 1;
            const int c0 = cc[16 * ct + r];          // the accumulators start from the corrections
#pragma unroll
            for (int rt = 0; rt < 4; ++rt) acc[rt][mt] = v4i{c0, c0, c0, c0};
          }
          const int FGk = (S / 4) * P::CPJ;         // K-steps of whole operand groups
          for (int ks = 0; ks < nks; ++ks) {
            // this lane group's chunk (operand j, chunk ci), and the column tiles the K-step can reach
            int j, ci, cmin, cmax;
            if (ks < FGk) {
              ci = ks % P::CPJ;
              j = (ks / P::CPJ) * 4 + g;
              cmin = ci;
              cmax = ci + P::CPJ;
            } else {
              const int s0l = (ks - FGk) * 4, nleft = (S % 4) * P::CPJ;
              const int slot = s0l + g;
              j = (S / 4) * 4 + slot / P::CPJ;
              ci = slot % P::CPJ;
              // the K-step's slots run over ci = s0l % CPJ .. last % CPJ, wrapping to the next operand (ci 0)
              // at most once
              const int last = (s0l + 3 < nleft ? s0l + 3 : nleft - 1);
              const bool wrap = (s0l / P::CPJ) != (last / P::CPJ);
              cmin = wrap ? 0 : s0l % P::CPJ;
              cmax = (wrap ? P::CPJ - 1 : last % P::CPJ) + P::CPJ;
            }
            v4i A[4];
#pragma unroll
            for (int rt = 0; rt < 4; ++rt) A[rt] = *(const v4i*)(X + (16 * rt + r) * P::XROW + (ks * 4 + g) * 16);
            const int i0 = 16 * ci;
            // lane groups past the operands: B fragment 0 (loads from a valid operand, masked)
            const int lmask = j < S ? -1 : 0;
            const int jc = j < S ? j : S - 1;
            // every fragment of the K-step is read first (one LDS wait per K-step); column tiles the K-step
            // cannot reach hold zero fragments and skip their products (a wave-uniform test)
            v4i Bf[P::MT];
#pragma unroll
            for (int mt = 0; mt < P::MT; ++mt) {
              // column tile (the waves with one tile less repeat their last one; never written back)
              const int ct = w + 4 * mt < P::NCT ? w + 4 * mt : P::NCT - 1;
              const int o = P::P0 - (16 * ct + r) + i0;   // R[o + b] = d[c - i0 - b]
              const int sft = o & 3;
              const u32* src = (const u32*)(Dc + (jc * 4 + sft) * P::RL4 + (o - sft));
              Bf[mt] = v4i{(int)src[0] & lmask, (int)src[1] & lmask, (int)src[2] & lmask, (int)src[3] & lmask};
            }
#pragma unroll
            for (int mt = 0; mt < P::MT; ++mt) {
              const int ct = w + 4 * mt;
              if (ct < P::NCT && ct >= cmin && ct <= cmax) {
#pragma unroll
                for (int rt = 0; rt < 4; ++rt)
                  acc[rt][mt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[rt], Bf[mt], acc[rt][mt], 0, 0, 0);
              }
            }
          }
        }
      } else if (step >= 1) {
        round_row(step - 1);
      }
      __syncthreads();
      // phase 2: column sums of row `step` into LDS (the epilogue has read row step - 1's), then the
      // reversed digit copies of row step + 1 (the products of row `step` are done)
      if (math && step < S) {
#pragma unroll
        for (int rt = 0; rt < 4; ++rt)
#pragma unroll
          for (int mt = 0; mt < P::MT; ++mt) {
            const int ct = w + 4 * mt;
            if (ct < P::NCT)
#pragma unroll
              for (int i = 0; i < 4; ++i) C[(16 * rt + 4 * g + i) * P::CS + 16 * ct + r] = acc[rt][mt][i];
          }
        if (step + 1 < S) build_copies(step + 1);
      }
      __syncthreads();
    }
  }
}

int b3d::pole_mma(int L, const int32_t* Xm, const int64_t* Xe, int64_t xps, const int64_t* xrow, const int32_t* ctab,
                  const int8_t* d8, const int32_t* corr, const int64_t* Ce, const int64_t* orow, int64_t T, int S,
                  int64_t B, int32_t* Om, int64_t* Oe, int64_t ops, hipStream_t s) {
  const int64_t items = T * ((S * B + 63) / 64);
  const unsigned grid = (unsigned)(items < (1 << 22) ? items : (1 << 22));
  B3D_DISPATCH_L(L, k_pole_mma<LL><<<dim3(grid), dim3(512), 0, s>>>(Xm, Xe, xps, xrow, ctab, d8, corr, Ce, orow, T, S, B, Om, Oe, ops));
  return b3d::check_launch("b3d_pole_level (MFMA)");
}

int b3d::pole_mma_max_s(int L) {
  B3D_DISPATCH_L(L, return PoleMma<LL>::JB);
  return 0;
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
