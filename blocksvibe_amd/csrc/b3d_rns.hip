// libb3dhip: exact multiprecision GEMM through a residue number system on int8 MFMA.
//
// C[b] = A[b] B[b] with A aligned per row and B per column (one common exponent per line), the aligned
// mantissa products summed exactly, one rounding per entry.  This replaces the unvendored sdpvibe.mpf.gemm
// of the reference (called at blocksvibe/recursion_gpu.py:53-79, derivatives.py:540-580,
// output.py:189,211).
//
// Residues: nmod primes p < 2^15 whose product M exceeds 2^(2 NB + log2 K + 4), so that the signed result
// is recovered exactly from its residues.  A residue r (balanced, |r| < 2^14) is split into two signed
// int8 digits r = 256 h + l (h in [-64, 64], l in [-128, 127]); per prime, four int8 GEMMs on
// v_mfma_i32_16x16x64_i8 (int32 accumulation, exact for K < 2^17) give
//   r_A r_B summed = 2^16 (h h) + 2^8 (h l + l h) + (l l).
// Stages: (1) per-line exponent maxima, (2) aligned limbs -> residue digits (int8 planes, K contiguous,
// padded to 16), (3) MFMA products reduced mod p -> uint16 residues, (4) CRT lift (y_t = c_t (M/p_t)^-1,
// C = sum_t y_t M/p_t - round(sum_t y_t / p_t) M, exact in 2L + 2 limbs) and normalisation.
#include "common.hpp"
#include "mp.hpp"
#include "../../include/b3dhip.h"

#include <mutex>

typedef int v4i __attribute__((ext_vector_type(4)));

static inline unsigned grid_for(int64_t n, int th) {
  int64_t b = (n + th - 1) / th;
  if (b > (1 << 22)) b = 1 << 22;
  if (b < 1) b = 1;
  return (unsigned)b;
}

// y = x mod p in [0, p) for an integral double |x| < 2^52 (q within one of x/p, one correction each way)
__device__ __forceinline__ double fmod_p(double x, double pd, double rp) {
  double y = fma(-floor(x * rp), pd, x);
  y = y < 0.0 ? y + pd : y;
  return y >= pd ? y - pd : y;
}

// the balanced residue |y| <= p/2 of an integral y in [0, p) as two int8 digits y = 256 h + l, packed h | l << 8
__device__ __forceinline__ uint16_t digit_pair(double y, double pd) {
  const int yb = y > 0.5 * pd ? (int)(y - pd) : (int)y;
  const int h = (yb + 128) >> 8;
  return (uint16_t)(((u32)h & 0xFFu) | (((u32)(yb - 256 * h) & 0xFFu) << 8));
}

// (1) exponent maximum per line: entry (b, line, k) at b*sb + line*sl + k*sk; Eline[b * lines + line]
__global__ __launch_bounds__(256) void k_line_max(const int32_t* __restrict__ m, const i64* __restrict__ e, i64 top,
                                                  i64 nb, i64 lines, i64 K, i64 sb, i64 sl, i64 sk,
                                                  i64* __restrict__ Eline) {
  for (i64 t = blockIdx.x * (i64)blockDim.x + threadIdx.x; t < nb * lines; t += (i64)gridDim.x * blockDim.x) {
    const i64 b = t / lines, line = t % lines;
    i64 E = B3D_EZ;
    const i64 base = b * sb + line * sl;
    for (i64 k = 0; k < K; ++k) {
      const i64 idx = base + k * sk;
      if (m[top + idx] != 0) {
        const i64 x = e[idx];
        E = x > E ? x : E;
      }
    }
    Eline[t] = E;
  }
}

// (2) residue digits on MFMA.  x mod p_t = sum_j byte_j(x) c_tj mod p_t with c_tj = 2^(8 j) mod p_t
// (balanced, split into int8 digits c = 256 ch + cl): an int8 GEMM of the entries' bytes (stored as
// byte - 128 = byte ^ 0x80, signed) against the constant weights w8 = [ch | cl][prime][KB] (KB = 4L bytes
// padded to a multiple of 64), corrected by wcorr_t = 128 sum_j c_tj.  One wave converts 64 consecutive
// entries t (t enumerates (b, line, k), k fastest, zero in the padding line >= lines or k >= K):
//   phase 1  lane = entry: load, align to the line exponent, magnitude bytes (^ 0x80) into LDS
//   phase 2  per 16 primes: 4 entry tiles x KB/64 steps x 2 digit halves of v_mfma_i32_16x16x64_i8
//   phase 3  Y = 256 Yh + Yl + wcorr mod p, sign, balanced residue -> int8 digits (h, l), staged in LDS
//            and written as 64 contiguous bytes per output plane.
// Output planes out[(t*2 + d) * plane + entry], d = 0 high digit, 1 low digit; the K dimension is padded to
// Kp = K rounded up to 16 (zero digits), so the number of entries is a multiple of 16.
template <int L>
__global__ __launch_bounds__(64) void k_to_residues(const int32_t* __restrict__ m, const i64* __restrict__ e, i64 ps,
                                                    i64 nb, i64 lines, i64 K, i64 sb, i64 sl, i64 sk,
                                                    const i64* __restrict__ Eline, i64 estride, int Lp, int Kp,
                                                    int nmod, const uint16_t* __restrict__ primes,
                                                    const int8_t* __restrict__ w8, const int32_t* __restrict__ wcorr,
                                                    const double* __restrict__ rp, int8_t* __restrict__ out,
                                                    i64 plane) {
  constexpr int KB = (4 * L + 63) / 64 * 64;  // bytes per entry, padded
  constexpr int NKS = KB / 64;                // MFMA K steps
  constexpr int ROW = KB + 16;                // LDS row stride (bytes)
  const int np16 = (nmod + 15) / 16 * 16;
  __shared__ __align__(16) int8_t xb[64 * ROW];
  __shared__ __align__(16) int8_t stage[32 * 64];
  __shared__ u32 sgn[64];
  const int lane = threadIdx.x;
  const int r = lane & 15, g = lane >> 4;
  const i64 total = nb * (i64)Lp * Kp;  // a multiple of 16 (Kp is)
  for (i64 t0 = (i64)blockIdx.x * 64; t0 < total; t0 += (i64)gridDim.x * 64) {
    // phase 1
    {
      const i64 t = t0 + lane;
      const i64 k = t % Kp, line = (t / Kp) % Lp, b = t / ((i64)Kp * Lp);
      u32 x[L];
      mp::zero(x);
      u32 sm = 0;
      if (t < total && line < lines && k < K) {
        const i64 idx = b * sb + line * sl + k * sk;
        mp::load(x, m, ps, idx);
        if (x[L - 1] != 0u) {
          sm = mp::to_mag(x);
          const i64 sh = Eline[b * estride + line] - e[idx];
          mp::shr(x, sh > 32 * L ? 32 * L : (int)sh);
        }
      }
      u32* row = (u32*)(xb + lane * ROW);
#pragma unroll
      for (int q = 0; q < KB / 4; ++q) row[q] = q < L ? (x[q] ^ 0x80808080u) : 0u;
      sgn[lane] = sm;
    }
    __syncthreads();
    v4i A[4][NKS];
#pragma unroll
    for (int et = 0; et < 4; ++et)
#pragma unroll
      for (int ks = 0; ks < NKS; ++ks) A[et][ks] = *(const v4i*)(xb + (et * 16 + r) * ROW + ks * 64 + g * 16);
    u32 negs = 0;   // sign flags of this lane's 16 output entries et*16 + 4g + i, bit et*4 + i
#pragma unroll
    for (int et = 0; et < 4; ++et)
#pragma unroll
      for (int i = 0; i < 4; ++i) negs |= (sgn[et * 16 + 4 * g + i] ? 1u : 0u) << (et * 4 + i);
    for (int pt = 0; pt < np16; pt += 16) {
      // phase 2
      v4i Bh[NKS], Bl[NKS];
#pragma unroll
      for (int ks = 0; ks < NKS; ++ks) {
        Bh[ks] = *(const v4i*)(w8 + ((i64)0 * np16 + pt + r) * KB + ks * 64 + g * 16);
        Bl[ks] = *(const v4i*)(w8 + ((i64)1 * np16 + pt + r) * KB + ks * 64 + g * 16);
      }
      const int tm = pt + r;
      const bool live = tm < nmod;
      const double pd = live ? (double)primes[tm] : 1.0;
      const double rpt = live ? rp[tm] : 1.0;
      const double corrd = live ? (double)wcorr[tm] : 0.0;
#pragma unroll
      for (int et = 0; et < 4; ++et) {
        v4i ah = {0, 0, 0, 0}, al = {0, 0, 0, 0};
#pragma unroll
        for (int ks = 0; ks < NKS; ++ks) {
          ah = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[et][ks], Bh[ks], ah, 0, 0, 0);
          al = __builtin_amdgcn_mfma_i32_16x16x64_i8(A[et][ks], Bl[ks], al, 0, 0, 0);
        }
        // phase 3: entries et*16 + 4g + i, prime tm, all in FP64 (|256 Yh + Yl + wcorr| < 2^31: exact)
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int ent = et * 16 + 4 * g + i;
          double c = fmod_p(fma((double)ah[i], 256.0, (double)al[i] + corrd), pd, rpt);
          c = ((negs >> (et * 4 + i)) & 1u) && c != 0.0 ? pd - c : c;    // residue of the signed entry
          const uint16_t hl = digit_pair(c, pd);
          stage[(r * 2 + 0) * 64 + ent] = (int8_t)(hl & 0xFFu);
          stage[(r * 2 + 1) * 64 + ent] = (int8_t)(hl >> 8);
        }
      }
      __syncthreads();
      {
        const int pd = lane >> 1, half = lane & 1;   // plane (prime pt + pd/2, digit pd%2), 32 bytes
        const int tq = pt + (pd >> 1);
        if (tq < nmod) {
          const v4i* src = (const v4i*)(stage + pd * 64 + half * 32);
          v4i* dst = (v4i*)(out + (i64)(tq * 2 + (pd & 1)) * plane + t0 + half * 32);
          if (t0 + half * 32 < total) dst[0] = src[0];          // the tile's tail: 16-byte groups
          if (t0 + half * 32 + 16 < total) dst[1] = src[1];
        }
      }
      __syncthreads();
    }
  }
}

// (3) per-prime products on MFMA: each wave one 16 x 16 output tile; the 4 waves of a block cover
// (16 wm) x (64 / wm) outputs (wm = 1, 2 or 4, chosen by the caller's tuner).  The epilogue reduces each sum
// mod p and multiplies by (M/p)^-1 mod p in FP64, so it writes the CRT coefficient y_t directly, balanced,
// as an int8 digit pair; staged in LDS and written by rows: Cr[(t * nb + b) * M * Nc + i * Nc + j].
__global__ __launch_bounds__(256) void k_rns_mma(const int8_t* __restrict__ A8, i64 aplane, int a_shared,
                                                 const int8_t* __restrict__ B8, i64 bplane, int Mp, int Np, int Kp,
                                                 int M, int N, const uint16_t* __restrict__ primes,
                                                 const uint16_t* __restrict__ inv, const double* __restrict__ rp,
                                                 uint16_t* __restrict__ Cr, i64 nb, int wm) {
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wn = 4 / wm, bw = 16 * wn;
  const int tm = blockIdx.y;
  const i64 b = blockIdx.z;
  const int ntn = (Np + bw - 1) / bw;
  const int tile_m = blockIdx.x / ntn, tile_n = blockIdx.x % ntn;
  const int wi = wave / wn, wj = wave % wn;
  const int row0 = tile_m * 16 * wm + 16 * wi, col0 = tile_n * bw + 16 * wj;
  __shared__ uint16_t st[64][64];   // the block's residues, stored by rows
  const i64 ab = a_shared ? 0 : b;
  const int8_t* Ah = A8 + (i64)(tm * 2 + 0) * aplane + (ab * Mp) * Kp;
  const int8_t* Al = A8 + (i64)(tm * 2 + 1) * aplane + (ab * Mp) * Kp;
  const int8_t* Bh = B8 + (i64)(tm * 2 + 0) * bplane + (b * Np) * Kp;
  const int8_t* Bl = B8 + (i64)(tm * 2 + 1) * bplane + (b * Np) * Kp;
  const int r = lane & 15, g = lane >> 4;
  v4i hh = {0, 0, 0, 0}, mid = {0, 0, 0, 0}, ll = {0, 0, 0, 0};
  for (int k0 = 0; k0 < (col0 < Np && row0 < Mp ? Kp : 0); k0 += 64) {
    // K is padded to 16 only: the 16-byte groups beyond Kp of the last step are zero
    const bool kin = k0 + g * 16 < Kp;
    const i64 ao = (i64)(row0 + r) * Kp + k0 + g * 16;
    const i64 bo = (i64)(col0 + r) * Kp + k0 + g * 16;
    const v4i z = {0, 0, 0, 0};
    const v4i ah = kin ? *(const v4i*)(Ah + ao) : z;
    const v4i al = kin ? *(const v4i*)(Al + ao) : z;
    const v4i bh = kin ? *(const v4i*)(Bh + bo) : z;
    const v4i bl = kin ? *(const v4i*)(Bl + bo) : z;
    hh = __builtin_amdgcn_mfma_i32_16x16x64_i8(ah, bh, hh, 0, 0, 0);
    mid = __builtin_amdgcn_mfma_i32_16x16x64_i8(ah, bl, mid, 0, 0, 0);
    mid = __builtin_amdgcn_mfma_i32_16x16x64_i8(al, bh, mid, 0, 0, 0);
    ll = __builtin_amdgcn_mfma_i32_16x16x64_i8(al, bl, ll, 0, 0, 0);
  }
  const double pd = (double)primes[tm], rpt = rp[tm], ivd = (double)inv[tm];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    // c = sum mod p (|sum| < 2^31, exact in FP64), then the CRT coefficient y = c (M/p)^-1 mod p (< 2^46)
    const double c = fmod_p(fma((double)hh[i], 65536.0, fma((double)mid[i], 256.0, (double)ll[i])), pd, rpt);
    st[16 * wi + 4 * g + i][16 * wj + r] = digit_pair(fmod_p(c * ivd, pd, rpt), pd);
  }
  __syncthreads();
  {
    const int tpr = 4 * wn;   // threads per row of the block tile
    const int rl = threadIdx.x / tpr, cl = (threadIdx.x % tpr) * 4;
    const int row = tile_m * 16 * wm + rl, cb = tile_n * bw;
    if (row < M) {
      uint16_t* dst = Cr + ((i64)tm * nb + b) * M * N + (i64)row * N + cb;
#pragma unroll
      for (int c = 0; c < 4; ++c)
        if (cb + cl + c < N) dst[cl + c] = st[rl][cl + c];
    }
  }
}

// (4) CRT lift on MFMA and normalisation.  With y_t = c_t (M/p_t)^-1 mod p_t taken balanced (|y_t| < 2^14,
// split y_t = 256 h_t + l_t into int8 digits by the MFMA kernel) and q = round(sum_t y_t / p_t)
// (|q| <= nmod / 2 < 128),
//   C = sum_t y_t M/p_t - q M
// is one int8 GEMM against the constant signed radix-256 digits of M/p_t (row l_t), of 256 M/p_t (row h_t)
// and of M (row -q): c8[byte column][k], k < CK.  Byte columns run along the MFMA rows, so each lane holds
// the sums of four consecutive byte columns = one 32-bit limb's worth (|sum| < 2^48 in int64).  One wave
// lifts 64 outputs (4 MFMA column tiles of 16) per pass:
//   phase 1  lane = output: the y digit pairs (one coalesced row per prime, written by the MFMA kernel)
//            and -q into an LDS row, then each lane loads its MFMA operand fragments (output r of tile ot,
//            bytes 64 ks + 16 g ..)
//   phase 2  per 16 byte columns: CK/64 steps of v_mfma_i32_16x16x64_i8 per tile -> limb sums into LDS
//   phase 3  lane = output: carry propagation into 2L + 2 limbs, one rounding, store.
// Output entry (b, i, j) at b*csb + i*csi + j*csj (+ column offset j0).
template <int L>
constexpr int crt_ckmax() {
  return (2 * ((2 * (32 * L - 2) + 20) / 14 + 2) + 1 + 63) / 64 * 64;
}

template <int L>
__global__ __launch_bounds__(64) void k_crt(const uint16_t* __restrict__ Cr, i64 nb, int M, int N, int nmod,
                                            const double* __restrict__ rp, const int8_t* __restrict__ c8, int CK,
                                            const i64* __restrict__ EA, int a_shared, const i64* __restrict__ EB,
                                            i64 j0, i64 Nfull, int32_t* __restrict__ cm, i64* __restrict__ ce,
                                            i64 cps, i64 csb, i64 csi, i64 csj) {
  constexpr int LW = 2 * L + 2;
  constexpr int CN = (4 * LW + 2 + 15) / 16 * 16;
  constexpr int NRT = CN / 16;
  constexpr int VW = LW + 1;
  constexpr int NKS = crt_ckmax<L>() / 64;
  constexpr int YROW = crt_ckmax<L>() + 16;
  static_assert(64 * YROW <= 64 * VW * 8, "digit rows must fit the limb-sum buffer they share");
  __shared__ i64 V[64 * VW];
  int8_t* yv = (int8_t*)V;
  const int lane = threadIdx.x, r = lane & 15, g = lane >> 4;
  const i64 total = nb * (i64)M * N;
  const int nks = CK / 64;
  for (i64 o0 = (i64)blockIdx.x * 64; o0 < total; o0 += (i64)gridDim.x * 64) {
    // phase 1: lane = output (coalesced residue loads), digits into LDS rows, then each lane's operand
    // fragments (output r of tile ot, bytes 64 ks + 16 g ..) into registers
    {
      const i64 o = o0 + lane;
      int8_t* yrow = yv + lane * YROW;
      double qs = 0.0;
      if (o < total) {
        // 64 loads in flight per lane: the kernel runs one wave per SIMD (its LDS buffer), so the wave's loads
        // are the only memory parallelism there is, and its register budget has room for them
        constexpr int NF = 64;
        for (int t0 = 0; t0 < nmod; t0 += NF) {
          uint32_t c[NF];
#pragma unroll
          for (int u = 0; u < NF; ++u) c[u] = t0 + u < nmod ? (uint32_t)Cr[(i64)(t0 + u) * total + o] : 0u;
#pragma unroll
          for (int u = 0; u < NF; ++u) {
            const int t = t0 + u;
            if (t < nmod) {
              // digit pair (h, l) of the balanced y_t: y_t = 256 h + l
              const int yb = 256 * (int)(int8_t)(c[u] & 0xFFu) + (int)(int8_t)(c[u] >> 8);
              qs = fma((double)yb, rp[t], qs);
              *(uint16_t*)(yrow + 2 * t) = (uint16_t)c[u];
            }
          }
        }
      } else {
        for (int t = 0; t < nmod; ++t) *(uint16_t*)(yrow + 2 * t) = 0;
      }
      yrow[2 * nmod] = (int8_t)(-(int)rint(qs));
      for (int k = 2 * nmod + 1; k < CK; ++k) yrow[k] = 0;
    }
    __syncthreads();
    v4i Bf[4][NKS];
#pragma unroll
    for (int ot = 0; ot < 4; ++ot)
#pragma unroll
      for (int ks = 0; ks < NKS; ++ks)
        Bf[ot][ks] = ks < nks ? *(const v4i*)(yv + (16 * ot + r) * YROW + ks * 64 + g * 16) : v4i{0, 0, 0, 0};
    __syncthreads();
    // phase 2
    for (int rt = 0; rt < NRT; ++rt) {
      v4i acc[4];
#pragma unroll
      for (int ot = 0; ot < 4; ++ot) acc[ot] = v4i{0, 0, 0, 0};
#pragma unroll
      for (int ks = 0; ks < NKS; ++ks) {
        if (ks < nks) {
          const v4i a = *(const v4i*)(c8 + (i64)(rt * 16 + r) * CK + ks * 64 + g * 16);
#pragma unroll
          for (int ot = 0; ot < 4; ++ot) acc[ot] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a, Bf[ot][ks], acc[ot], 0, 0, 0);
        }
      }
      const int k = 4 * rt + g;
      if (k < LW) {
#pragma unroll
        for (int ot = 0; ot < 4; ++ot)
          V[(16 * ot + r) * VW + k] = (i64)acc[ot][0] + (i64)acc[ot][1] * 256 + (i64)acc[ot][2] * 65536 +
                                      (i64)acc[ot][3] * 16777216;
      }
    }
    __syncthreads();
    // phase 3
    const i64 o = o0 + lane;
    if (o < total) {
      const i64 j = o % N, i = (o / N) % M, b = o / ((i64)N * M);
      const i64 ea = EA[(a_shared ? 0 : b) * M + i];
      const i64 eb = EB[b * Nfull + j0 + j];
      u32 out[L];
      i64 eo = B3D_EZ;
      if (ea == B3D_EZ || eb == B3D_EZ) {
        mp::zero(out);
      } else {
        u32 acc[LW];
        i64 carry = 0;
#pragma unroll
        for (int k = 0; k < LW; ++k) {
          const i64 v = V[lane * VW + k] + carry;
          acc[k] = (u32)v;
          carry = v >> 32;   // arithmetic: floor division by 2^32
        }
        eo = mp::norm<L, LW>(acc, ea + eb, out);
      }
      const i64 oo = b * csb + i * csi + (j0 + j) * csj;
      mp::store(out, cm, cps, oo);
      ce[oo] = eo;
    }
    __syncthreads();
  }
}

extern "C" {

int64_t b3d_rns_workspace_bytes(int nmod, int64_t nbA, int64_t nb, int64_t M, int64_t Nc, int64_t K) {
  const int64_t Mp = (M + 15) / 16 * 16, Np = (Nc + 63) / 64 * 64, Kp = (K + 15) / 16 * 16;
  int64_t bytes = 0;
  bytes += 2 * (int64_t)nmod * nbA * Mp * Kp;  // A digits
  bytes += 2 * (int64_t)nmod * nb * Np * Kp;   // B digits (one chunk)
  bytes += 2 * (int64_t)nmod * nb * M * Nc;    // residues of C (uint16)
  bytes += 8 * (nbA * M + nb * (int64_t)0);    // row exponents (column exponents are separate)
  return bytes + 1024;
}

int b3d_rns_gemm(int L, int nmod, const uint16_t* primes, const int8_t* w8, const int32_t* wcorr, const uint16_t* inv,
                 const double* rp, const int8_t* c8, int CK, int64_t nb, int64_t M, int64_t N, int64_t K, int a_shared,
                 const int32_t* Am, const int64_t* Ae, int64_t aps, int64_t asb, int64_t asi, int64_t ask,
                 const int32_t* Bm, const int64_t* Be, int64_t bps, int64_t bsb, int64_t bsk, int64_t bsj, int32_t* Cm,
                 int64_t* Ce, int64_t cps, int64_t csb, int64_t csi, int64_t csj, int64_t Nc, int tile_wm,
                 int64_t* EB, void* workspace, int64_t ws_bytes, void* stream) {
  if (nb <= 0 || M <= 0 || N <= 0) return 0;
  if (K <= 0 || K > (1 << 16)) {
    b3d::set_error("b3d_rns_gemm: need 0 < K <= 2^16");
    return 2;
  }
  if (nb > 65535) {
    b3d::set_error("b3d_rns_gemm: batch must be < 65536");
    return 2;
  }
  const int ckmax = (2 * ((2 * (32 * L - 2) + 20) / 14 + 2) + 1 + 63) / 64 * 64;   // crt_ckmax<L>()
  if (nmod <= 0 || nmod > 255 || CK % 64 != 0 || CK < 2 * nmod + 1 || CK > ckmax) {
    b3d::set_error("b3d_rns_gemm: CRT digit table does not fit (need 2 nmod + 1 <= CK <= bound, CK % 64 == 0)");
    return 2;
  }
  if (tile_wm != 1 && tile_wm != 2 && tile_wm != 4) {
    b3d::set_error("b3d_rns_gemm: tile_wm must be 1, 2 or 4");
    return 2;
  }
  if (Nc <= 0 || Nc > N) Nc = N;
  const int64_t nbA = a_shared ? 1 : nb;
  if (b3d_rns_workspace_bytes(nmod, nbA, nb, M, Nc, K) > ws_bytes) {
    b3d::set_error("b3d_rns_gemm: workspace too small for the column chunk");
    return 2;
  }
  if (b3d::check_planes("b3d_rns_gemm", {aps, bps, cps})) return 2;
  hipStream_t s = (hipStream_t)stream;
  const int Mpad = (int)((M + 15) / 16 * 16), Npad = (int)((Nc + 63) / 64 * 64), Kpad = (int)((K + 15) / 16 * 16);
  char* w = (char*)workspace;
  int8_t* A8 = (int8_t*)w;
  w += 2 * (int64_t)nmod * nbA * Mpad * Kpad;
  int8_t* B8 = (int8_t*)w;
  w += 2 * (int64_t)nmod * nb * Npad * Kpad;
  uint16_t* Cr = (uint16_t*)w;
  w += 2 * (int64_t)nmod * nb * M * Nc;
  w = (char*)(((uintptr_t)w + 255) & ~(uintptr_t)255);
  int64_t* EA = (int64_t*)w;
  const int64_t aplane = nbA * Mpad * Kpad, bplane = nb * (int64_t)Npad * Kpad;
  // A: rows
  k_line_max<<<grid_for(nbA * M, 256), 256, 0, s>>>(Am, Ae, (int64_t)(L - 1) * aps, nbA, M, K, asb, asi, ask, EA);
  B3D_DISPATCH_L(L, k_to_residues<LL><<<grid_for(nbA * (int64_t)Mpad * Kpad, 64), 64, 0, s>>>(
                        Am, Ae, aps, nbA, M, K, asb, asi, ask, EA, M, Mpad, Kpad, nmod, primes, w8, wcorr, rp, A8,
                        aplane));
  // B: columns (all of them once: exponents over k per column)
  k_line_max<<<grid_for(nb * N, 256), 256, 0, s>>>(Bm, Be, (int64_t)(L - 1) * bps, nb, N, K, bsb, bsj, bsk, EB);
  if (b3d::check_launch("b3d_rns_gemm (prep)")) return 1;
  for (int64_t j0 = 0; j0 < N; j0 += Nc) {
    const int64_t nc = (N - j0) < Nc ? (N - j0) : Nc;
    const int Np = (int)((nc + 63) / 64 * 64);
    // column chunk [j0, j0 + nc): shift the B base pointer, column exponents indexed with the full N
    B3D_DISPATCH_L(L, k_to_residues<LL><<<grid_for(nb * (int64_t)Np * Kpad, 64), 64, 0, s>>>(
                          Bm + j0 * bsj, Be + j0 * bsj, bps, nb, nc, K, bsb, bsj, bsk, EB + j0, N, Np, Kpad, nmod,
                          primes, w8, wcorr, rp, B8, nb * (int64_t)Np * Kpad));
    const int bh = 16 * tile_wm, bw = 64 / tile_wm;
    dim3 grid((unsigned)(((Mpad + bh - 1) / bh) * ((Np + bw - 1) / bw)), (unsigned)nmod, (unsigned)nb);
    k_rns_mma<<<grid, 256, 0, s>>>(A8, aplane, a_shared, B8, nb * (int64_t)Np * Kpad, Mpad, Np, Kpad, (int)M,
                                   (int)nc, primes, inv, rp, Cr, nb, tile_wm);
    B3D_DISPATCH_L(L, k_crt<LL><<<grid_for(nb * M * nc, 64), 64, 0, s>>>(
                          Cr, nb, (int)M, (int)nc, nmod, rp, c8, CK, EA, a_shared, EB, j0, N, Cm, Ce,
                          cps, csb, csi, csj));
    if (b3d::check_launch("b3d_rns_gemm (chunk)")) return 1;
  }
  (void)bplane;
  return 0;
}

}  // extern "C"

// ---- the allocation-free form (b3d_rns_gemm_noalloc) ----
// Its scratch: the column exponents EB, then the workspace of one column chunk.  A device global, so each
// GPU the library runs on holds its own copy (at its own address) from the module's load until the process
// ends.  Host code never uses this name as an address: the per-device address comes from
// hipGetSymbolAddress.
static __device__ __align__(256) char scratch[B3D_RNS_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch[kMaxDevices] = {};

// the scratch's device address on the current device (looked up once per device), or null with the error set
char* device_scratch() {
  int dev = -1;
  hipError_t st = hipGetDevice(&dev);
  if (st != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    b3d::set_error(std::string("b3d_rns_gemm_noalloc: no usable current device (") + hipGetErrorString(st) + ")");
    return nullptr;
  }
  std::lock_guard<std::mutex> g(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    st = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (st != hipSuccess || !p) {
      b3d::set_error(std::string("b3d_rns_gemm_noalloc: hipGetSymbolAddress(scratch) failed: ") +
                     hipGetErrorString(st));
      return nullptr;
    }
    g_scratch[dev] = (char*)p;
  }
  return g_scratch[dev];
}
}  // namespace

extern "C" {

int b3d_rns_gemm_noalloc(int L, int nmod, const uint16_t* primes, const int8_t* w8, const int32_t* wcorr,
                         const uint16_t* inv, const double* rp, const int8_t* c8, int CK, int64_t nb, int64_t M,
                         int64_t N, int64_t K, int a_shared, const int32_t* Am, const int64_t* Ae, int64_t aps,
                         int64_t asb, int64_t asi, int64_t ask, const int32_t* Bm, const int64_t* Be, int64_t bps,
                         int64_t bsb, int64_t bsk, int64_t bsj, int32_t* Cm, int64_t* Ce, int64_t cps, int64_t csb,
                         int64_t csi, int64_t csj, int tile_wm, void* stream) {
  if (nb <= 0 || M <= 0 || N <= 0) return 0;
  if (K <= 0 || K > (1 << 16) || nmod <= 0 || nmod > 255 || nb > 65535 || M > (int64_t(1) << 30) ||
      N > (int64_t(1) << 30)) {
    b3d::set_error("b3d_rns_gemm_noalloc: need 0 < K <= 2^16, 0 < nmod <= 255, nb < 65536, M, N <= 2^30");
    return 2;
  }
  const int64_t cap = B3D_RNS_NOALLOC_SCRATCH_BYTES;
  const int64_t nbA = a_shared ? 1 : nb;
  const int64_t eb_bytes = (nb * N * 8 + 255) / 256 * 256;
  const int64_t room = cap - eb_bytes;
  // the widest column chunk whose workspace fits beside EB (the workspace grows by `per` for every 64 columns)
  const int64_t fixed = b3d_rns_workspace_bytes(nmod, nbA, nb, M, 0, K);
  const int64_t per = b3d_rns_workspace_bytes(nmod, nbA, nb, M, 64, K) - fixed;
  int64_t Nc = N;
  if (room <= 0 || b3d_rns_workspace_bytes(nmod, nbA, nb, M, N, K) > room) {
    Nc = room > fixed ? (room - fixed) / per * 64 : 0;
    if (Nc > N) Nc = N;
  }
  if (Nc < 1 || b3d_rns_workspace_bytes(nmod, nbA, nb, M, Nc, K) > room) {
    const int64_t need = eb_bytes + b3d_rns_workspace_bytes(nmod, nbA, nb, M, N < 64 ? N : 64, K);
    b3d::set_error("b3d_rns_gemm_noalloc: this product needs at least " + std::to_string(need) +
                   " bytes of scratch (column exponents and one chunk of " + std::to_string(N < 64 ? N : 64) +
                   " columns) but the fixed scratch holds " + std::to_string(cap) +
                   " bytes; use b3d_rns_gemm with a workspace");
    return 2;
  }
  char* base = device_scratch();
  if (!base) return 1;
  return b3d_rns_gemm(L, nmod, primes, w8, wcorr, inv, rp, c8, CK, nb, M, N, K, a_shared, Am, Ae, aps, asb, asi, ask,
                      Bm, Be, bps, bsb, bsk, bsj, Cm, Ce, cps, csb, csi, csj, Nc, tile_wm, (int64_t*)base,
                      base + eb_bytes, room, stream);
}

}  // extern "C"
