/* libb3dhip -- C-ABI of the MI355X (gfx950) blocks_3d hot path.
 *
 * Plain pointers and sizes only.  Device arrays are allocated by the caller (any HIP allocator); the
 * library never allocates on these paths except where noted (b3d_rns_gemm takes a caller workspace).
 * Every entry point returns 0 on success, nonzero on error (b3d_last_error() then describes it); launches
 * are asynchronous on `stream` (a hipStream_t, NULL = default stream).
 *
 * Multiprecision arrays ("MPF"): mantissas as int32 limb planes -- limb k of entry i at m[k * ps + i],
 * ps the plane stride, L limbs, two's complement, limb 0 least significant -- and exponents e[i] int64;
 * value m * 2^e; normalised: bitlen(|m|) = 32 L - 2, zero = (0, EZ = -2^62).
 * The reference's equivalent is the unvendored sdpvibe.mpf type reached from Python through DLPack
 * (reference blocksvibe/cauchy.py:83-85, smallgemm.py:84-86, derivatives.py:478-482).
 */
#ifndef B3DHIP_H
#define B3DHIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define B3D_ABI_VERSION 18

/* Last error message of the calling thread (empty if none). */
const char* b3d_last_error(void);
int b3d_abi_version(void);
/* Set the calling thread's stream-capture mode (0 global, 1 thread-local, 2 relaxed; HIP's
 * hipStreamCaptureMode) and return the previous one, -1 on error.  The HIP runtime judges a potentially
 * unsafe call (a synchronous copy, a device-wide synchronisation) by the calling thread's mode against every
 * capture in the process, so a thread left in the global mode invalidates captures running in other threads;
 * the serving loop relaxes its worker and submitting threads and captures in the thread-local mode.  No
 * counterpart in the reference (it has no captured pipelines). */
int b3d_thread_capture_mode(int mode);
/* Read and clear the calling thread's pending HIP error (hipGetLastError), e.g. after a failed capture, so
 * that it is not reported by the next launch; returns the hipError_t code (0: none). */
int b3d_clear_hip_error(void);
/* End a capture still open on `stream` (one that failed part way) and discard its graph, so that the stream
 * can be used again; returns 1 if a capture was ended, 0 if none was open. */
int b3d_end_capture(void* stream);
/* 1 if kernels are compiled for L limbs (reference cli.py:270-274 gives L from the precision). */
int b3d_supported_limbs(int L);

/* ---- MPF primitives (replace sdpvibe.mpf.normalize / add / mul / _align_rows, called at reference
 *      blocksvibe/cauchy.py:78,87, smallgemm.py:77-79,88, recursion_gpu.py:432-436,457,460,496-509) ---- */

/* Wide two's complement integers W (LW <= 2L+4 limb planes, plane stride wps) times 2^ein[i] ->
 * normalised (round half away from zero). */
int b3d_mpf_norm_wide(int L, int LW, const int32_t* W, int64_t wps, const int64_t* ein, int32_t* om,
                      int64_t* oe, int64_t ops, int64_t n, void* stream);
/* c[i] = a[i*sa] + b[i*sb]  (sa, sb in {0, 1}: broadcast or elementwise) */
int b3d_mpf_add(int L, const int32_t* am, const int64_t* ae, int64_t aps, int64_t sa, const int32_t* bm,
                const int64_t* be, int64_t bps, int64_t sb, int32_t* cm, int64_t* ce, int64_t cps, int64_t n,
                void* stream);
/* c[i] = a[i*sa] * b[i*sb] */
int b3d_mpf_mul(int L, const int32_t* am, const int64_t* ae, int64_t aps, int64_t sa, const int32_t* bm,
                const int64_t* be, int64_t bps, int64_t sb, int32_t* cm, int64_t* ce, int64_t cps, int64_t n,
                void* stream);
/* Number of 28-bit limbs of a magnitude below 2^(32L - 2) (the pole step's constant form), -1 if L has no
 * kernels. */
int b3d_limbs28(int L);

/* Constant operands of the pole step: each of the n entries m (two's complement, plane stride ps) as
 * b3d_limbs28(L) 28-bit magnitude limbs (o28, plane stride ops) and a sign mask osg (0 or ~0).  No
 * counterpart in the reference (its smallgemm re-splits the operands on every call). */
int b3d_mpf_to28(int L, const int32_t* m, int64_t ps, int64_t n, int32_t* o28, int64_t ops, uint32_t* osg,
                 void* stream);

/* Align each row of a (rows, K) array (entry (r,k) at r*rs + k*ks) to its largest exponent: aligned
 * mantissas to om (contiguous (rows, K), plane stride ops), row exponents to oe[rows]. */
int b3d_mpf_align_rows(int L, const int32_t* m, const int64_t* e, int64_t ps, int64_t rows, int64_t K,
                       int64_t rs, int64_t ks, int32_t* om, int64_t ops, int64_t* oe, void* stream);

/* ---- radial recursion, one level (replaces the body of reference Recursion.run,
 *      blocksvibe/recursion_gpu.py:446-476: cauchy_kernel cauchy.py:73-87, _pole :483-492) ---- */

/* Cauchy step.  Slot buffer rm/re (plane stride rps, N = SL*SR*B columns per slot); h_inf hm/he (plane
 * stride hps, N columns per row); groups g < nb: ghrow[g] = h_inf row, col[g*Kc + i] physical slots of
 * the prime poles (i < ncol[g]), cold[g*Kc + i] their doubled positions D_i; H rows h < nH: rg[h] group,
 * rd[h] doubled D'.  Output H (nH, N) at om/oe (plane stride ops).  Requires |D' - D_i| < 2^15.
 * Optional wave table (null: none), for rows_per_lane 1 or 2: the rows of a group contiguous and taken
 * rows_per_lane at a time (rows_per_lane 2: rows 2p, 2p + 1 of the group, the last pair of an odd count
 * has one row); wave w < nw covers units woff[w] .. woff[w] + 63 (unit o is row / pair o / N of the group,
 * column o % N) of the group whose wcnt[w] rows start at row wrow[w]; every unit of every group in exactly
 * one wave, ceil(wcnt / rows_per_lane) N of them per group.  With it the group's data is wave-uniform and
 * rows are padded only at the end of their group; with rows_per_lane 2 a lane computes both rows of its
 * pair at its column (one load and shift of every term for both: less work per entry, fewer waves in
 * flight; the same bits).  rows_per_lane is ignored without a table. */
int b3d_cauchy_level(int L, const int32_t* rm, const int64_t* re, int64_t rps, int64_t N, const int32_t* hm,
                     const int64_t* he, int64_t hps, const int64_t* ghrow, const int32_t* rg, const int32_t* rd,
                     int64_t nH, const int64_t* col, const int32_t* cold, const int32_t* ncol, int Kc, int32_t* om,
                     int64_t* oe, int64_t ops, const int32_t* wrow, const int32_t* wcnt, const int32_t* woff,
                     int64_t nw, int rows_per_lane, void* stream);

/* Pole step for T tasks: Y = Ralign[tR[t]] . H[thrw[t]]^T, then res[tslot[t]] = QLalign[tQL[t]] . Y^T, over
 * residue blocks of SL x SR structure entries (left a < SL, right b < SR; entry (a, b) of a residue at
 * (a SR + b) B + k) with B right-hand sides: Y[t][b0][a][k] = sum_{b < SR} R[b0][b] H[a][b][k],
 * res[a0][b0][k] = sum_{a < SL} QL[a0][a] Y[t][b0][a][k].  Ralign/QLalign are (ntab, S, S) tables, S >= SL, SR
 * (rows and entries beyond SR resp. SL are never read), given in the form b3d_mpf_to28 makes of the row-aligned
 * mantissas: 28-bit magnitude limb planes R28/Q28 (b3d_limbs28(L) planes, plane strides rtps/qps) and sign
 * masks Rsg/Qsg (ntab, S, S), with row exponents Re/Qe (ntab, S); and in the form b3d_pole_tables makes
 * (digits Rd8/Qd8, corrections Rcorr/Qcorr; may be null).  With the digit tables, SL = SR = S, S*B >= 32 and S
 * small enough for the MFMA path's LDS staging (6 at L = 34) the products run on int8 MFMA, otherwise on the
 * VALU; both give the same bits.  kR, kQ (both or neither; may be null, and then the MFMA path is not taken):
 * per right-hand side k an offset added to every task's row index into Ralign resp. QLalign, so that the right-hand
 * sides of several operators of one plan, with their tables stacked into one, run in one launch.  Y: scratch
 * (L, T, SR, SL, B), plane stride yps.  The reference pads every residue to Smax x Smax and runs one operator per
 * launch (recursion_gpu.py:483-510); here only the SL x SR entries are stored and computed. */
int b3d_pole_level(int L, int SL, int SR, int S, int64_t B, int64_t T, const int32_t* Hm, const int64_t* He,
                   int64_t hps, const int64_t* thrw, const int32_t* tR, const int32_t* tQL, const int64_t* tslot,
                   const int32_t* R28, const uint32_t* Rsg, const int64_t* Re, int64_t rtps, const int32_t* Q28,
                   const uint32_t* Qsg, const int64_t* Qe, int64_t qps, const int8_t* Rd8, const int32_t* Rcorr,
                   const int8_t* Qd8, const int32_t* Qcorr, const int32_t* kR, const int32_t* kQ, int32_t* Ym,
                   int64_t* Ye, int64_t yps, int32_t* rm, int64_t* re, int64_t rps, void* stream);

/* Bytes per constant entry in the digit tables of b3d_pole_tables (4L + 4), or -1 for an unsupported L. */
int b3d_pole_digits_bytes(int L);

/* Constant operands of the pole step for its int8 MFMA path: the nrows*S row-aligned entries m (two's
 * complement, plane stride ps; entry j of row cr at cr*S + j) as balanced base-256 digits d8
 * (b3d_pole_digits_bytes(L) bytes per entry: 4L + 1 digits in [-128, 127], then zeros) and per row the
 * column corrections corr (nrows, 8L) int32, corr[cr][c] = 128 sum_j sum_{i <= 4L-2} d_j[c - i].  No
 * counterpart in the reference (its smallgemm multiplies limb by limb on every call). */
int b3d_pole_tables(int L, int S, const int32_t* m, int64_t ps, int64_t nrows, int8_t* d8, int32_t* corr,
                    void* stream);

/* ---- derivative moments during the recursion (replace the momacc/momflush kernels and the staging of
 *      reference blocksvibe/derivatives.py:46-189, 474-510) ---- */

/* Staging after a level: for t < T, column c < N: stage[woff + row[t]*N + c] = res[slot[t]*N + c] *
 * pw[nn[t]] (pw: table of r_c^n, plane stride pps). */
int b3d_moments_stage(int L, const int32_t* rm, const int64_t* re, int64_t rps, const int64_t* slot,
                      const int64_t* row, const int32_t* nn, int64_t T, int64_t N, const int32_t* pm,
                      const int64_t* pe, int64_t pps, int32_t* sm, int64_t* se, int64_t sps, int64_t woff,
                      void* stream);
/* Flush W staged levels m0..m0+W-1 into the moments v (L, nrows, N, R1): v[row, c, k] += sum_w
 * stage[w, row, c] * n_w^k, n_w = m0 + w + srow[row]; NK: exact n^k as KL uint32 limb planes (plane
 * stride nkps, entry n*R1 + k), NKb their bit lengths.  KL <= 8.  Column c = s B + b (structure entry s,
 * right-hand side b < B, N a multiple of B); jk (Kc codes b R1 + k) limits the flush to those (b, k), e.g.
 * k <= R_b for right-hand sides of several lambda orders in one recursion (their other moments are neither
 * read nor written); null: every (b, k < R1), Kc ignored. */
int b3d_moments_flush(int L, int KL, const int32_t* sm, const int64_t* se, int64_t sps, int W, int m0,
                      const int32_t* srow, const uint32_t* NK, int64_t nkps, const int32_t* NKb, int32_t* vm,
                      int64_t* ve, int64_t vps, int64_t nrows, int64_t N, int R1, int64_t B, const int32_t* jk,
                      int Kc, void* stream);

/* Lower-triangular products by exact integer matrices, in the multiprecision GEMM's semantics (column of
 * v aligned to its largest exponent with truncated magnitudes, exact sum, one rounding; the same bits as
 * b3d_rns_gemm with the matrix as A): out(b, e, c) = sum_{k <= e} T_t[e][k] v(b, k, c), t = tsel[b], for
 * b < nb, e < K, c < Nc.  v entry (b, k, c) at (b K + k) Nc + c (plane stride vps).  T_t[e][k] = +-tm 2^texp:
 * magnitudes as TL (2, 4 or 8) limb planes tm (ntab, K, K) (plane stride tps), sign masks tsg (ntab, K, K),
 * unit exponents texp (ntab, K).  Output entry (b, e, c) at (b / nper) obatch + (b % nper) orow + e Nc + c
 * (plane stride ops).  Replaces the batched GEMM a_i = C(e, k) D_i^(e-k) v_i of reference
 * blocksvibe/derivatives.py:553-554 (one conversion to residues per v entry and one CRT per output there). */
int b3d_tri_product(int L, int TL, const int32_t* vm, const int64_t* ve, int64_t vps, int64_t nb, int K, int64_t Nc,
                    const int64_t* tsel, const int32_t* tm, const uint32_t* tsg, int64_t tps, const int64_t* texp,
                    int32_t* om, int64_t* oe, int64_t ops, int64_t nper, int64_t obatch, int64_t orow, void* stream);

/* The same product by M x K exact integer matrices with full rows (lower != 0: k <= e only, M <= K):
 * out(b, e, c) = sum_k T_t[e][k] v(b, k, c), e < M; tables tm (TL, ntab, M, K) (plane stride tps), tsg
 * (ntab, M, K), texp (ntab, M); v and the output as b3d_tri_product; M, K <= 4096.  Replaces the GEMMs of
 * reference blocksvibe/derivatives.py:560 (G = the powers D_i^w of the pole positions times the moments)
 * and :567 (the binomials C(e, d) times G), whose A operands are exact integers with a row unit 2^-w. */
int b3d_int_product(int L, int TL, const int32_t* vm, const int64_t* ve, int64_t vps, int64_t nb, int M, int K,
                    int lower, int64_t Nc, const int64_t* tsel, const int32_t* tm, const uint32_t* tsg, int64_t tps,
                    const int64_t* texp, int32_t* om, int64_t* oe, int64_t ops, int64_t nper, int64_t obatch,
                    int64_t orow, void* stream);

/* ---- exact multiprecision GEMM through residues on int8 MFMA (replaces the unvendored sdpvibe.mpf.gemm
 *      reached through reference blocksvibe/recursion_gpu.py:53-79 gemm/bgemm, called at
 *      derivatives.py:540,554,560,567,580 and output.py:189,211) ---- */

/* Workspace for one column chunk of Nc columns: nbA = 1 if A is shared by the batch, else nb. */
int64_t b3d_rns_workspace_bytes(int nmod, int64_t nbA, int64_t nb, int64_t M, int64_t Nc, int64_t K);

/* C[b] = A[b] B[b] for b < nb: A entry (b,i,k) at b*asb + i*asi + k*ask (a_shared: one A for all b), B entry
 * (b,k,j) at b*bsb + k*bsk + j*bsj, C entry (b,i,j) at b*csb + i*csi + j*csj.  A aligned per row, B per
 * column, exact sum of products, one rounding.  Residue tables (device): nmod primes < 2^15; residue
 * conversion weights w8 (int8, [2][n16][KB], n16 = nmod rounded up to 16, KB = 4L rounded up to 64):
 * c = 2^(8 j) mod p_t balanced and split c = 256 w8[0][t][j] + w8[1][t][j], wcorr[t] = 128 sum_j c (int32);
 * inv[t] = (M/p_t)^-1 mod p_t, rp[t] = 1/p_t with M = prod p_t >= 2^(2(32L-2) + log2 K + 4); CRT digits c8
 * (int8, [CN][CK], CN = 4(2L+2)+2 rounded up to 16, CK = 2 nmod + 1 rounded up to 64): column 2t+1 the signed
 * radix-256 digits of M/p_t, column 2t the same one byte up, column 2 nmod those of M.  EB: scratch of nb*N
 * int64.  Columns are processed Nc at a time within `workspace` (b3d_rns_workspace_bytes).  tile_wm (1, 2
 * or 4): MFMA block tile of 16 tile_wm rows x 64 / tile_wm columns (a tuning choice; results do not depend
 * on it).  Requires K <= 2^16, nb < 65536, nmod <= 255. */
int b3d_rns_gemm(int L, int nmod, const uint16_t* primes, const int8_t* w8, const int32_t* wcorr, const uint16_t* inv,
                 const double* rp, const int8_t* c8, int CK, int64_t nb, int64_t M, int64_t N, int64_t K, int a_shared,
                 const int32_t* Am, const int64_t* Ae, int64_t aps, int64_t asb, int64_t asi, int64_t ask,
                 const int32_t* Bm, const int64_t* Be, int64_t bps, int64_t bsb, int64_t bsk, int64_t bsj, int32_t* Cm,
                 int64_t* Ce, int64_t cps, int64_t csb, int64_t csi, int64_t csj, int64_t Nc, int tile_wm,
                 int64_t* EB, void* workspace, int64_t ws_bytes, void* stream);

/* Bytes of the fixed device scratch that b3d_rns_gemm_noalloc works in (one copy per GPU, held until the
 * process ends). */
#define B3D_RNS_NOALLOC_SCRATCH_BYTES (64ll << 20)

/* b3d_rns_gemm without a workspace argument and without any allocation: the column exponents and the column
 * chunks' workspace live in a fixed device scratch of B3D_RNS_NOALLOC_SCRATCH_BYTES per GPU, the chunk width
 * chosen to fit; the same bits as b3d_rns_gemm.  Returns 2 with a message if the column exponents (8 nb N
 * bytes) and a chunk of min(N, 64) columns do not fit.  The scratch address is looked up once per device on
 * the first call made with that device current (the stream must belong to the current device).  Every call
 * on one GPU shares that GPU's scratch: calls on one stream are ordered, calls that overlap on different
 * streams of one GPU race.  No counterpart in the reference: for callers that cannot allocate device memory
 * (code under stream capture, tight serving loops, harnesses that forbid hipMalloc). */
int b3d_rns_gemm_noalloc(int L, int nmod, const uint16_t* primes, const int8_t* w8, const int32_t* wcorr,
                         const uint16_t* inv, const double* rp, const int8_t* c8, int CK, int64_t nb, int64_t M,
                         int64_t N, int64_t K, int a_shared, const int32_t* Am, const int64_t* Ae, int64_t aps,
                         int64_t asb, int64_t asi, int64_t ask, const int32_t* Bm, const int64_t* Be, int64_t bps,
                         int64_t bsb, int64_t bsk, int64_t bsj, int32_t* Cm, int64_t* Ce, int64_t cps, int64_t csb,
                         int64_t csi, int64_t csj, int tile_wm, void* stream);

/* ---- output emission (host code; host arrays; no GPU needed) ---- */

/* Decimal text of n MPF entries read from host limb planes (entry i: limbs m[k * ps + i], exponent e[i]): the
 * exact value rounded to p significant digits, ties to even, trailing zeros removed, fixed notation for
 * decimal exponents -4 <= X < p, else d.ddde+XX; zero as "0" (blocks_3d's mpf_float output at stream
 * precision p; replaces reference blocksvibe/b3djson.py:30-68 fmt_big / fmt_gmpy, which convert each value to
 * an mpmath or MPFR number first).  The texts are concatenated into out (cap bytes, no terminators); text i
 * ends at end[i].  threads: worker threads (<= 0: one per hardware thread).  1 <= L <= 64, 1 <= p <= 2000. */
int b3d_format_numbers(int L, const int32_t* m, const int64_t* e, int64_t ps, int64_t n, int p, char* out,
                       int64_t cap, int64_t* end, int threads);
/* Kept length of each of npoly polynomials of deg coefficients (coefficient d of polynomial q is entry
 * q * deg + d of host limb planes m / exponents e): one past the last coefficient c with
 * max(|later coefficients|) < 2^eps_exp |c|, by exact comparison; 0 if every coefficient is zero (reference
 * blocksvibe/b3djson.py:160-168, blocks_3d's dropping of negligible top coefficients). */
int b3d_poly_sizes(int L, const int32_t* m, const int64_t* e, int64_t ps, int64_t npoly, int deg, int64_t eps_exp,
                   int32_t* sizes);
/* Both at once for npoly polynomials laid out as in b3d_poly_sizes: each polynomial's JSON list body as
 * blocks_3d's write_derivs.cxx writes it (reference blocksvibe/b3djson.py:150-177): the kept coefficients
 * quoted and joined by ",\n" ("c0",\n"c1"), or "0" (quoted) if every coefficient is zero.  Bodies are
 * concatenated into out (cap bytes); body q ends at end[q].  Work is split over threads by polynomial. */
int b3d_format_polys(int L, const int32_t* m, const int64_t* e, int64_t ps, int64_t npoly, int deg, int p,
                     int64_t eps_exp, char* out, int64_t cap, int64_t* end, int threads);

/* ---- output emission on the GPU (device arrays; the same text as b3d_format_polys) ---- */

/* Kept length of each polynomial q < npoly whose deg coefficients are entries pbase[q] .. pbase[q] + deg - 1
 * of the device planes m / e (plane stride ps): as b3d_poly_sizes. */
int b3d_fmt_sizes(int L, const int32_t* m, const int64_t* e, int64_t ps, const int64_t* pbase, int64_t npoly, int deg,
                  int64_t eps_exp, int32_t* sizes, void* stream);
/* Text of n elements: element k is entry ent[k] of the planes (ent[k] < 0: the text "0"), formatted as
 * b3d_format_numbers does, in groups of four bytes: byte c of its text at text[((c / 4) * n + k) * 4 + c % 4]
 * (text: n * R(p + 16) bytes, R(x) = x rounded up to a multiple of 4), its length at len[k]; len[k] = -1 for a
 * value outside the GPU's range (exponent >= 0, or fraction bits beyond 32 (L + 16)), to be formatted on the
 * host.  digits: staging of n * R(p + 10 L + 32) bytes.  text and digits 4-byte aligned. */
int b3d_fmt_text(int L, const int32_t* m, const int64_t* e, int64_t ps, const int64_t* ent, int64_t n, int p,
                 uint8_t* digits, uint8_t* text, int32_t* len, void* stream);
/* Packed bodies: element k as '"' text '"' followed by ",\n" unless last[k], at out + off[k]. */
int b3d_fmt_pack(const uint8_t* text, const int32_t* len, const uint8_t* last, const int64_t* off, int64_t n,
                 uint8_t* out, void* stream);

/* Whole files at once: b3d_fmt_pack's pieces, each element k preceded by string pre[k] and followed by string
 * post[k] of the table (string i = strtab[stroff[i] .. stroff[i + 1]); -1: none), so that the texts between the
 * polynomials of a spin file (its header and the JSON brackets) are laid out with them and each file is one
 * contiguous range of out.  off[k]: offset of element k's piece (pre + '"' text '"' + ",\n" unless last[k] +
 * post); last[k] = 2 marks an unused slot, which writes nothing.  Ids must be < nstr.  No counterpart in the reference (its writer joins Python strings per
 * coefficient, reference blocksvibe/b3djson.py:150-207). */
int b3d_fmt_pack_files(const uint8_t* text, const int32_t* len, const uint8_t* last, const int32_t* pre,
                       const int32_t* post, const uint8_t* strtab, const int64_t* stroff, int64_t nstr,
                       const int64_t* off, int64_t n, uint8_t* out, void* stream);

/* ---- the index plan of the radial recursion (host code; no GPU needed).  Replaces reference
 *      blocksvibe/recursion_gpu.py:136-365 (Plan, numpy) and its pole bookkeeping structure.py:16-74. ---- */

/* Build the plan of one operator: two_j12, two_j43 (doubled, one parity), two_j_max, radial order, the
 * pseudo-parities ((pp0_left, pp0_right), (pp1_left, pp1_right)) of the two parities, reuse: recycle physical
 * slots (LIFO, the reference's numbering) or not.  *plan receives an opaque handle (free with b3d_plan_free).
 * The tables are read with b3d_plan_get. */
int b3d_plan_build(int two_j12, int two_j43, int two_j_max, int order, int pp0_left, int pp0_right, int pp1_left,
                   int pp1_right, int reuse, void** plan);
void b3d_plan_free(void* plan);
/* Copy table `field` (of level `level` for the per-level fields, of the given level for the LEVEL_TJS/NPOL/FIRST
 * fields, ignored otherwise) into out (at most cap int64 entries; out may be null to ask the length).  Returns
 * the table's length, -1 for a bad field or level.  Plan fields:
 *   SCALARS      nslots, nphys, SLm, SRm, S2 = SLm SRm, Smax, Jh, dmax, order, b = two_j12 mod 2
 *   SL, SR       (2, n) structure counts per parity and spin two_j = b + 2 i
 *   PHYS         physical slot of each of the 2 nslots logical slots (-1: never written)
 *   MKEYS        per M key in index order: pole type, k, 0 (L) or 1 (R), two_jNN, pseudo-parity, two_j
 *   QKEYS        per Q key in index order: pole type, k, two_j, sign;   LQPAIRS: per (L, Q) pair: M key, Q key
 *   LEVEL_TJS, LEVEL_NPOL, LEVEL_FIRST: spins of the level's residue blocks, their pole counts, first slots
 * Level fields:
 *   LEVEL        present, m, nb (groups), Kc (columns per group row), nH (Cauchy rows), T (tasks), dmax
 *   GSRC, GTJP   per group its source parity and pole spin;  NCOL: its prime-pole columns
 *   COL, COLD    (nb, Kc) physical slot (nphys: zero slot) and doubled position of each column
 *   RG, RD       per Cauchy row its group and doubled position D'
 *   T_SLOT, T_HROW, T_L, T_R, T_Q, T_QL: per task its physical slot, Cauchy row, M keys, Q key, (L, Q) pair
 *   T_KEYS       (6, T): pole type, k, two_j, pseudo-parities (left, right), sign */
#define B3D_PLAN_SCALARS 0
#define B3D_PLAN_SL 1
#define B3D_PLAN_SR 2
#define B3D_PLAN_PHYS 3
#define B3D_PLAN_MKEYS 4
#define B3D_PLAN_QKEYS 5
#define B3D_PLAN_LQPAIRS 6
#define B3D_PLAN_LEVEL_TJS 7
#define B3D_PLAN_LEVEL_NPOL 8
#define B3D_PLAN_LEVEL_FIRST 9
#define B3D_PLAN_LEVEL 16
#define B3D_PLAN_GSRC 17
#define B3D_PLAN_GTJP 18
#define B3D_PLAN_NCOL 19
#define B3D_PLAN_COL 20
#define B3D_PLAN_COLD 21
#define B3D_PLAN_RG 22
#define B3D_PLAN_RD 23
#define B3D_PLAN_T_SLOT 24
#define B3D_PLAN_T_HROW 25
#define B3D_PLAN_T_L 26
#define B3D_PLAN_T_R 27
#define B3D_PLAN_T_Q 28
#define B3D_PLAN_T_QL 29
#define B3D_PLAN_T_KEYS 30
int64_t b3d_plan_get(const void* plan, int field, int level, int64_t* out, int64_t cap);

#ifdef __cplusplus
}
#endif
#endif
