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
  // LDS of a task with J operands: column sums, the 64 rows' operands, 4 reversed copies of one constant
  // row's digits, the digits and corrections of all J constant rows; then (own arrays) the constant rows'
  // exponents, the row exponents and the segment carries
  static constexpr int bytes(int J) {
    return LC + 64 * (16 * CPJ * J + 16) + 4 * J * RL4 + J * J * DS + J * NCOL * 4 + 16 * 8 + 64 * 8 + 4 * 64 * 8;
  }
  static constexpr int jmax() {
    int j = 1;
    while (j < 16 && bytes(j + 1) <= 160 * 1024) ++j;
    return j;
  }
  static constexpr int JB = jmax();                          // largest S the MFMA path takes (6 at L = 34)
  static constexpr int XROW = 16 * CPJ * JB + 16;            // LDS bytes per staged row
  static constexpr int LDS = bytes(JB) - 16 * 8 - 64 * 8 - 4 * 64 * 8;
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
  __shared__ i64 cseg[4][64];
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
  const int nch = S * P::CPJ;
  const int nks = (nch + 3) / 4;
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
    for (int jl = w; jl < S; jl += 8) {
      u32* row = (u32*)(X + lane * P::XROW + jl * 16 * P::CPJ);
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
      for (int q = 0; q < 4 * P::CPJ; ++q)
        row[q] = q < L - 1 ? (x[q] ^ 0x80808080u) : (q == L - 1 ? (x[q] ^ 0x00808080u) : 0u);
    }
    // the task's S x S digit rows and S correction rows (contiguous in the tables), S row exponents
    {
      const u32* src = (const u32*)(d8 + cr0 * S * P::DS);
      for (int q = tid; q < S * S * (P::DS / 4); q += 512) ((u32*)Dr)[q] = src[q];
      const int32_t* cs = corr + cr0 * P::NCOL;
      for (int q = tid; q < S * P::NCOL; q += 512) Cr[q] = cs[q];
      if (tid < S) Cex[tid] = Ce[cr0 + tid];
    }
    __syncthreads();
    for (int step = 0; step <= S; ++step) {
      v4i acc[4][P::MT];
      // phase A: math waves build the reversed copies of row `step`; epilogue waves carry row step - 1
      if (math) {
        if (step < S)
#pragma unroll 4
          for (int q = tid; q < S * 4 * (P::RL4 / 4); q += 256) {
            const int jl = q / (4 * (P::RL4 / 4));
            const int rem = q - jl * 4 * (P::RL4 / 4);
            const int sft = rem / (P::RL4 / 4), wd = rem - sft * (P::RL4 / 4);
            const int8_t* d = Dr + (step * S + jl) * P::DS;
            u32 v = 0u;
#pragma unroll
            for (int b = 0; b < 4; ++b) {
              const int di = P::P0 - 4 * wd - sft - b;
              const u32 byte = (di >= 0 && di <= 4 * L) ? (u32)(uint8_t)d[di] : 0u;
              v |= byte << (8 * b);
            }
            ((u32*)(Dc + (jl * 4 + sft) * P::RL4))[wd] = v;
          }
      } else if (step >= 1 && valid) {
        // row `lane`, segment w - 4 = columns [2L (w-4), 2L (w-4) + 2L) -> L/2 limbs in place + a carry
        const int sg = w - 4;
        int32_t* crow = C + lane * P::CS + 2 * L * sg;
        // all of the segment's column sums are read first (one LDS wait), then carried
        i64 part[P::SEGL];
#pragma unroll
        for (int q = 0; q < P::SEGL; ++q) {
          i64 v = 0;
#pragma unroll
          for (int b = 0; b < 4; ++b) v += (i64)crow[4 * q + b] * ((i64)1 << (8 * b));
          part[q] = v;
        }
        i64 carry = 0;
#pragma unroll
        for (int q = 0; q < P::SEGL; ++q) {
          const i64 v = part[q] + carry;
          crow[q] = (int32_t)(u32)v;
          carry = v >> 32;
        }
        cseg[sg][lane] = carry;
      }
      __syncthreads();
      // phase B: math waves multiply row `step`; epilogue wave 4 rounds row step - 1
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
          for (int ks = 0; ks < nks; ++ks) {
            const int q = 4 * ks + g;   // this lane's chunk: operand q / CPJ, bytes 16 (q % CPJ) ..
            v4i A[4];
#pragma unroll
            for (int rt = 0; rt < 4; ++rt) A[rt] = *(const v4i*)(X + (16 * rt + r) * P::XROW + 16 * q);
            const int jl = q / P::CPJ;
            const int i0 = 16 * (q - jl * P::CPJ);
            // padding chunks past the operands: B fragment 0 (loads from a valid operand, masked)
            const int lmask = jl < S ? -1 : 0;
            const int jc = jl < S ? jl : S - 1;
            // every fragment of the K-step is read before the products, one LDS wait per K-step
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
            for (int mt = 0; mt < P::MT; ++mt)
#pragma unroll
              for (int rt = 0; rt < 4; ++rt)
                acc[rt][mt] = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[rt], Bf[mt], acc[rt][mt], 0, 0, 0);
          }
        }
      } else if (w == 4 && step >= 1 && valid) {
        const int s0 = step - 1;
        u32 W[2 * L + 2];
#pragma unroll
        for (int sg = 0; sg < 4; ++sg)
#pragma unroll
          for (int q = 0; q < P::SEGL; ++q) W[sg * P::SEGL + q] = (u32)C[lane * P::CS + 2 * L * sg + q];
        const i64 c3 = cseg[3][lane];
        W[2 * L] = (u32)c3;
        W[2 * L + 1] = (u32)((u64)c3 >> 32);
#pragma unroll
        for (int sg = 0; sg < 3; ++sg) {
          const i64 cs = cseg[sg][lane];
          const u32 ext = cs < 0 ? 0xFFFFFFFFu : 0u;
          u32 cy = 0u;
#pragma unroll
          for (int q = (sg + 1) * P::SEGL; q < 2 * L + 2; ++q)
            W[q] = mp::addc(W[q], q == (sg + 1) * P::SEGL ? (u32)cs : (q == (sg + 1) * P::SEGL + 1 ? (u32)((u64)cs >> 32) : ext), cy);
        }
        const i64 ea = Cex[s0], Er = Erow[lane];
        u32 out[L];
        i64 e;
        if (Er == B3D_EZ || ea == B3D_EZ) {
          mp::zero(out);
          e = B3D_EZ;
        } else {
          e = mp::norm<L, 2 * L + 2>(W, ea + Er, out);
        }
        const i64 o = ob + s0 * SB + n;
        mp::store(out, Om, ops, o);
        Oe[o] = e;
      }
      __syncthreads();
      // phase C: column sums of row `step` into LDS (the epilogue has read row step - 1's)
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
