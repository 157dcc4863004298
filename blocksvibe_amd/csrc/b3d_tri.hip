// libb3dhip: products of multiprecision columns by small matrices of exact integers, in the multiprecision
// GEMM's semantics.  First use: the derivative stage's residues a_i[e] = sum_{k <= e} C(e, k) D_i^(e - k)
// v_i[k] (reference blocksvibe/derivatives.py:13 and the batched GEMM at :553-554), a batch of small
// lower-triangular products.  As a GEMM through residues it costs a residue conversion of every v entry and a
// CRT lift of every output for K <= Lambda + 1 products; here each output is one thread: the column aligned as
// the GEMM aligns it, K multiplications of an L-limb magnitude by a TL-limb integer (the matrix entries
// C(e, k) d^(e - k) 2^k with d = 2 D_i are exact integers of at most TL limbs, with the row's unit 2^-e), an
// exact sum and one rounding -- the same bits as the GEMM.  The same kernel takes full rows of M x K integer
// matrices (b3d_int_product): the stage's G = (D_i^w) v (reference derivatives.py:560), whose B operand (the
// moments, K = poles per spin) is the largest one the stage would convert to residues, and the binomial
// product C(e, d) G (:567).
//
// Layouts: v (L, nb, K, Nc) (entry (b, k, c) at (b K + k) Nc + c); tables tm (TL, ntab, M, K) magnitude limb
// planes, tsg (ntab, M, K) sign masks, texp (ntab, M) unit exponents; tsel[b] the table of batch b; output
// entry (b, e, c) at (b / nper) obatch + (b % nper) orow + e Nc + c.
#include <string>

#include "common.hpp"
#include "mp.hpp"
#include "../../include/b3dhip.h"

template <int L, int TL>
__global__ __launch_bounds__(256) void k_int_product(const int32_t* __restrict__ vm, const i64* __restrict__ ve,
                                                     i64 vps, i64 nb, int M, int K, int lower, i64 Nc,
                                                     const int64_t* __restrict__ tsel, const int32_t* __restrict__ tm,
                                                     const uint32_t* __restrict__ tsg, i64 tps,
                                                     const i64* __restrict__ texp, int32_t* __restrict__ om,
                                                     i64* __restrict__ oe, i64 ops, i64 nper, i64 obatch, i64 orow) {
  const i64 n = nb * M * Nc;
  for (i64 t = blockIdx.x * (i64)blockDim.x + threadIdx.x; t < n; t += (i64)gridDim.x * blockDim.x) {
    const i64 c = t % Nc;
    const int e = (int)((t / Nc) % M);
    const i64 b = t / ((i64)M * Nc);
    const i64 ts = tsel[b];
    u32 r[L];
    const i64 er = mp::tri_entry<L, TL>(vm, ve, vps, b * K * Nc + c, Nc, K, lower ? e + 1 : K, tm, tsg, tps,
                                        (ts * M + e) * K, texp[ts * M + e], r);
    const i64 o = (b / nper) * obatch + (b % nper) * orow + (i64)e * Nc + c;
    mp::store(r, om, ops, o);
    oe[o] = er;
  }
}

static inline unsigned grid_for(int64_t n, int th) {
  int64_t b = (n + th - 1) / th;
  if (b > (1 << 22)) b = 1 << 22;
  if (b < 1) b = 1;
  return (unsigned)b;
}

static int int_product(const char* what, int L, int TL, const int32_t* vm, const int64_t* ve, int64_t vps,
                       int64_t nb, int M, int K, int lower, int64_t Nc, const int64_t* tsel, const int32_t* tm,
                       const uint32_t* tsg, int64_t tps, const int64_t* texp, int32_t* om, int64_t* oe, int64_t ops,
                       int64_t nper, int64_t obatch, int64_t orow, hipStream_t s) {
  if (nb <= 0 || M <= 0 || K <= 0 || Nc <= 0) return 0;
  if (M > 4096 || K > 4096 || nper <= 0) {
    b3d::set_error(std::string(what) + ": need 1 <= M, K <= 4096 and nper >= 1");
    return 2;
  }
  if (b3d::check_planes(what, {vps, tps, ops})) return 2;
  const unsigned g = grid_for(nb * M * Nc, 256);
  switch (TL) {
    case 2:
      B3D_DISPATCH_L(L, k_int_product<LL, 2><<<dim3(g), dim3(256), 0, s>>>(vm, ve, vps, nb, M, K, lower, Nc, tsel, tm, tsg, tps, texp, om, oe, ops, nper, obatch, orow));
      break;
    case 4:
      B3D_DISPATCH_L(L, k_int_product<LL, 4><<<dim3(g), dim3(256), 0, s>>>(vm, ve, vps, nb, M, K, lower, Nc, tsel, tm, tsg, tps, texp, om, oe, ops, nper, obatch, orow));
      break;
    case 8:
      B3D_DISPATCH_L(L, k_int_product<LL, 8><<<dim3(g), dim3(256), 0, s>>>(vm, ve, vps, nb, M, K, lower, Nc, tsel, tm, tsg, tps, texp, om, oe, ops, nper, obatch, orow));
      break;
    default:
      b3d::set_error(std::string(what) + ": TL must be 2, 4 or 8");
      return 2;
  }
  return b3d::check_launch(what);
}

extern "C" {

int b3d_tri_product(int L, int TL, const int32_t* vm, const int64_t* ve, int64_t vps, int64_t nb, int K, int64_t Nc,
                    const int64_t* tsel, const int32_t* tm, const uint32_t* tsg, int64_t tps, const int64_t* texp,
                    int32_t* om, int64_t* oe, int64_t ops, int64_t nper, int64_t obatch, int64_t orow, void* stream) {
  if (K > 64) {
    b3d::set_error("b3d_tri_product: need 1 <= K <= 64 and nper >= 1");
    return 2;
  }
  return int_product("b3d_tri_product", L, TL, vm, ve, vps, nb, K, K, 1, Nc, tsel, tm, tsg, tps, texp, om, oe, ops,
                     nper, obatch, orow, (hipStream_t)stream);
}

int b3d_int_product(int L, int TL, const int32_t* vm, const int64_t* ve, int64_t vps, int64_t nb, int M, int K,
                    int lower, int64_t Nc, const int64_t* tsel, const int32_t* tm, const uint32_t* tsg, int64_t tps,
                    const int64_t* texp, int32_t* om, int64_t* oe, int64_t ops, int64_t nper, int64_t obatch,
                    int64_t orow, void* stream) {
  if (lower && M > K) {
    b3d::set_error("b3d_int_product: a lower-triangular product needs M <= K");
    return 2;
  }
  return int_product("b3d_int_product", L, TL, vm, ve, vps, nb, M, K, lower, Nc, tsel, tm, tsg, tps, texp, om, oe,
                     ops, nper, obatch, orow, (hipStream_t)stream);
}

}  // extern "C"
