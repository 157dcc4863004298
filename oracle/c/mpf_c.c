/* C restatement of the oracle's Cauchy-term and pole-step dot-product arithmetic (oracle/mpf_ref.py:
 * norm, cauchy, gemm_aligned), bit for bit.  TEST INFRASTRUCTURE ONLY: used by tests/ (pinned against
 * mpf_ref on random and edge-case inputs) and by the CPU baseline leg of bench.py, which times it on all
 * host threads.  Nothing in the product links it.
 *
 * Number format: value m * 2^e, m two's complement over L 32-bit limbs (limb 0 least significant),
 * normalised bitlen(|m|) = 32 L - 2, zero = (0, OC_EZ).  Arrays here are entry-major: entry i's limbs at
 * m[i * L .. i * L + L - 1].
 */
#include <pthread.h>
#include <stdlib.h>
#include <stdint.h>
#include <string.h>
#include <time.h>

#define OC_EZ (-(INT64_C(1) << 62))
#define OC_MAXL 64
#define OC_MAXW (2 * OC_MAXL + 8)

typedef uint32_t u32;
typedef uint64_t u64;

static int bitlen(const u32 *a, int n) {
  for (int k = n - 1; k >= 0; --k)
    if (a[k]) return 32 * k + 32 - __builtin_clz(a[k]);
  return 0;
}

static void negate(u32 *a, int n) {
  u64 c = 1;
  for (int k = 0; k < n; ++k) {
    u64 t = (u64)(u32)~a[k] + c;
    a[k] = (u32)t;
    c = t >> 32;
  }
}

static int is_neg(const u32 *a, int n) { return (int)(a[n - 1] >> 31); }

/* out[0..no) = a >> s (logical, a has n limbs) */
static void shr(const u32 *a, int n, int s, u32 *out, int no) {
  int q = s >> 5, r = s & 31;
  for (int k = 0; k < no; ++k) {
    u64 lo = (k + q < n) ? a[k + q] : 0u;
    u64 hi = (k + q + 1 < n) ? a[k + q + 1] : 0u;
    out[k] = r ? (u32)((lo >> r) | (hi << (32 - r))) : (u32)lo;
  }
}

/* out[0..no) = a << s (a has n limbs) */
static void shl(const u32 *a, int n, int s, u32 *out, int no) {
  int q = s >> 5, r = s & 31;
  for (int k = 0; k < no; ++k) {
    int j = k - q;
    u64 lo = (j >= 0 && j < n) ? a[j] : 0u;
    u64 lo1 = (j - 1 >= 0 && j - 1 < n) ? a[j - 1] : 0u;
    out[k] = r ? (u32)((lo << r) | (lo1 >> (32 - r))) : (u32)lo;
  }
}

/* W (two's complement, n limbs) * 2^e -> normalised L limbs, round half away from zero
 * (mpf_ref.norm) */
int64_t oc_norm(const u32 *W, int n, int64_t e, int L, u32 *out) {
  u32 a[OC_MAXW + 1], r[OC_MAXW + 1];
  memcpy(a, W, (size_t)n * 4);
  a[n] = 0;
  int s = is_neg(a, n);
  if (s) negate(a, n);
  int bl = bitlen(a, n);
  memset(out, 0, (size_t)L * 4);
  if (!bl) return OC_EZ;
  int NB = 32 * L - 2, sh = bl - NB;
  if (sh > 0) {
    /* add 2^(sh - 1), then shift */
    int b = sh - 1;
    u64 c = (u64)1 << (b & 31);
    for (int k = b >> 5; k <= n && c; ++k) {
      u64 t = (u64)a[k] + c;
      a[k] = (u32)t;
      c = t >> 32;
    }
    shr(a, n + 1, sh, r, L + 1);
    if (bitlen(r, L + 1) > NB) {
      shr(r, L + 1, 1, out, L);
      sh += 1;
    } else {
      memcpy(out, r, (size_t)L * 4);
    }
  } else {
    shl(a, n, -sh, out, L);
  }
  if (s) negate(out, L);
  return e + sh;
}

/* acc (n limbs, two's complement) += sign * x (x magnitude, nx limbs) */
static void acc_add(u32 *acc, int n, const u32 *x, int nx, int sub) {
  u64 c = sub ? 1 : 0;
  for (int k = 0; k < n; ++k) {
    u32 xv = k < nx ? x[k] : 0u;
    if (sub) xv = ~xv;
    u64 t = (u64)acc[k] + xv + c;
    acc[k] = (u32)t;
    c = t >> 32;
  }
}

/* h + sum_i x_i * 2 / d_i with one rounding (mpf_ref.cauchy): x entry-major (K x L); d_i == 0 or x_i == 0
 * contribute nothing; hm == NULL or a zero h means no h */
int64_t oc_cauchy(const u32 *xm, const int64_t *xe, const int32_t *d, int K, int L, const u32 *hm, int64_t he,
                  u32 *out) {
  int64_t E = OC_EZ;
  int any = 0;
  for (int i = 0; i < K; ++i) {
    if (bitlen(xm + (size_t)i * L, L) == 0) continue;
    if (!any || xe[i] > E) E = xe[i];
    any = 1;
  }
  const int hnz = hm && bitlen(hm, L) > 0;
  memset(out, 0, (size_t)L * 4);
  if (!any && !hnz) return OC_EZ;
  int64_t G;
  if (!any) G = he - 32;
  else if (!hnz) G = E - 31;
  else G = (E - 31 > he - 32) ? E - 31 : he - 32;
  u32 acc[OC_MAXL + 3], a[OC_MAXL + 2], q[OC_MAXL + 1], t[OC_MAXL + 1];
  memset(acc, 0, sizeof(acc));
  for (int i = 0; i < K; ++i) {
    const u32 *x = xm + (size_t)i * L;
    if (d[i] == 0 || bitlen(x, L) == 0) continue;
    int64_t sh = G + 31 - xe[i];
    if (sh >= 32 * (L + 1)) continue;
    /* floor(m * 2^32 / |d|) of the signed mantissa: long division with a signed top limb, remainders >= 0 */
    int64_t dd = d[i] < 0 ? -(int64_t)d[i] : (int64_t)d[i];
    int64_t cur = (int32_t)x[L - 1];
    int64_t qt = cur >= 0 ? cur / dd : -((-cur + dd - 1) / dd);
    int64_t rem = cur - qt * dd;
    q[L] = (u32)(int32_t)qt;
    for (int k = L - 1; k >= 0; --k) {
      u64 c = ((u64)rem << 32) | (k > 0 ? x[k - 1] : 0u);
      q[k] = (u32)(c / (u64)dd);
      rem = (int64_t)(c % (u64)dd);
    }
    /* arithmetic shift right by sh (floor), sign-extended into L + 3 limbs, added with the sign of d */
    u32 ext[OC_MAXL + 3];
    const u32 fill = (q[L] >> 31) ? 0xFFFFFFFFu : 0u;
    for (int k = 0; k < L + 3; ++k) {
      int64_t bit = (int64_t)k * 32 + sh;
      int w = (int)(bit >> 5), r = (int)(bit & 31);
      u64 lo = w < L + 1 ? q[w] : fill, hi = w + 1 < L + 1 ? q[w + 1] : fill;
      ext[k] = r ? (u32)((lo >> r) | (hi << (32 - r))) : (u32)lo;
    }
    acc_add(acc, L + 3, ext, L + 3, d[i] < 0);
  }
  if (hnz) {
    memcpy(a, hm, (size_t)L * 4);
    a[L] = 0;
    int sh_ = is_neg(a, L);
    if (sh_) negate(a, L);
    int64_t tt = he - G; /* <= 32 */
    if (tt >= 0) shl(a, L + 1, (int)tt, t, L + 1);
    else if (-tt < 32 * (L + 1)) shr(a, L + 1, (int)(-tt), t, L + 1);
    else memset(t, 0, (size_t)(L + 1) * 4);
    acc_add(acc, L + 3, t, L + 1, sh_);
  }
  return oc_norm(acc, L + 3, G, L, out);
}

/* one pole-step entry from pre-aligned operands (mpf_ref.gemm_aligned): exact sum of K products of
 * L-limb signed mantissas, one rounding */
int64_t oc_dot_aligned(const u32 *am, int64_t ae, const u32 *bm, int64_t be, int K, int L, u32 *out) {
  memset(out, 0, (size_t)L * 4);
  if (ae == OC_EZ || be == OC_EZ) return OC_EZ;
  u32 acc[OC_MAXW], x[OC_MAXL], y[OC_MAXL], p[2 * OC_MAXL];
  memset(acc, 0, sizeof(acc));
  for (int i = 0; i < K; ++i) {
    memcpy(x, am + (size_t)i * L, (size_t)L * 4);
    memcpy(y, bm + (size_t)i * L, (size_t)L * 4);
    int sx = is_neg(x, L), sy = is_neg(y, L);
    if (sx) negate(x, L);
    if (sy) negate(y, L);
    memset(p, 0, (size_t)2 * L * 4);
    for (int j = 0; j < L; ++j) {
      if (!x[j]) continue;
      u64 c = 0;
      for (int k = 0; k < L; ++k) {
        u64 t = (u64)x[j] * y[k] + p[j + k] + c;
        p[j + k] = (u32)t;
        c = t >> 32;
      }
      p[j + L] = (u32)c;
    }
    acc_add(acc, 2 * L + 2, p, 2 * L, sx != sy);
  }
  return oc_norm(acc, 2 * L + 2, ae + be, L, out);
}

/* ------------------------------------------------------------------ multithreaded timing (CPU baseline) */

typedef struct {
  int L, K, kind;
  double seconds;
  u64 seed;
  double units, elapsed;
} job_t;

static u64 xs64(u64 *s) {
  u64 x = *s;
  x ^= x << 13;
  x ^= x >> 7;
  x ^= x << 17;
  return *s = x;
}

static double now(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* random normalised mantissa with exponent spread 0..3 below the top */
static void rnd_mant(u32 *m, int L, u64 *s) {
  for (int k = 0; k < L; ++k) m[k] = (u32)xs64(s);
  m[L - 1] &= 0x1FFFFFFFu;
  m[L - 1] |= 0x20000000u;
  if (xs64(s) & 1) negate(m, L);
}

static void *worker(void *arg) {
  job_t *j = (job_t *)arg;
  int L = j->L, K = j->K;
  u32 xm[64 * OC_MAXL], bm[64 * OC_MAXL], out[OC_MAXL];
  int64_t xe[64];
  int32_t d[64];
  u64 s = j->seed | 1;
  for (int i = 0; i < K; ++i) {
    rnd_mant(xm + (size_t)i * L, L, &s);
    rnd_mant(bm + (size_t)i * L, L, &s);
    xe[i] = -(int64_t)(32 * L) + (int64_t)(xs64(&s) % 5);
    d[i] = (int32_t)(xs64(&s) % 600) - 300;
    if (d[i] == 0) d[i] = 1;
  }
  double t0 = now(), t = t0, units = 0;
  volatile u32 sink = 0;
  while (t - t0 < j->seconds) {
    for (int r = 0; r < 64; ++r) {
      if (j->kind == 0) {
        oc_cauchy(xm, xe, d, K, L, bm, -(int64_t)(32 * L), out);
        units += K;
      } else {
        oc_dot_aligned(xm, -100, bm, -100, K, L, out);
        units += 1;
      }
      sink ^= out[0];
      xm[r % K * L] ^= out[1] | 1u;
    }
    t = now();
  }
  j->units = units;
  j->elapsed = t - t0;
  return 0;
}

/* per-unit wall time on nthreads threads running concurrently: kind 0 = one Cauchy term (groups of K
 * terms), kind 1 = one K-term pole-step dot product.  Returns seconds per unit for the whole machine
 * (total elapsed / total units across threads, i.e. the aggregate throughput's inverse). */
double oc_time(int kind, int L, int K, int nthreads, double seconds) {
  if (nthreads < 1) nthreads = 1;
  if (nthreads > 256) nthreads = 256;
  if (K > 64) K = 64;
  pthread_t th[256];
  job_t jobs[256];
  for (int i = 0; i < nthreads; ++i) {
    jobs[i].L = L;
    jobs[i].K = K;
    jobs[i].kind = kind;
    jobs[i].seconds = seconds;
    jobs[i].seed = 0x9E3779B97F4A7C15ull * (u64)(i + 1);
    pthread_create(&th[i], 0, worker, &jobs[i]);
  }
  double units = 0, el = 0;
  for (int i = 0; i < nthreads; ++i) {
    pthread_join(th[i], 0);
    units += jobs[i].units;
    if (jobs[i].elapsed > el) el = jobs[i].elapsed;
  }
  return units > 0 ? el / units : 0.0;
}

/* ------------------------------------------------------------------ the recursion, one level at a time */
/* The same level loop as the GPU library's b3d_cauchy_level / b3d_pole_level, restated on the host over the
 * same index tables (mpf_ref semantics: recursion_ref.recursion_bits), split over nthreads threads.  Arrays are
 * entry-major; a residue slot holds N = S2 B entries, entry (a SR + b) B + k. */

/* mpf_ref.align of K entries m[i * stride] (entry units): magnitudes shifted right onto the largest exponent,
 * truncated, sign kept; returns that exponent (OC_EZ if every entry is zero) */
static int64_t align_vec(const u32 *m, const int64_t *e, int64_t stride, int K, int L, u32 *out) {
  int64_t E = OC_EZ;
  int any = 0;
  for (int i = 0; i < K; ++i) {
    const u32 *x = m + (size_t)(i * stride) * L;
    if (bitlen(x, L) == 0) continue;
    if (!any || e[i * stride] > E) E = e[i * stride];
    any = 1;
  }
  memset(out, 0, (size_t)K * L * 4);
  if (!any) return OC_EZ;
  u32 a[OC_MAXL];
  for (int i = 0; i < K; ++i) {
    const u32 *x = m + (size_t)(i * stride) * L;
    if (bitlen(x, L) == 0) continue;
    memcpy(a, x, (size_t)L * 4);
    int neg = is_neg(a, L);
    if (neg) negate(a, L);
    int64_t s = E - e[i * stride];
    if (s >= 32 * L) continue;
    shr(a, L, (int)s, out + (size_t)i * L, L);
    if (neg) negate(out + (size_t)i * L, L);
  }
  return E;
}

typedef struct {
  int L, Kc, S, SL, SR, nthreads, tid;
  int64_t N, B, units;
  const u32 *rm, *hm, *Hm, *Rm, *Qm;
  const int64_t *re, *he, *He, *Re, *Qe, *ghrow, *col, *thrw, *tslot;
  const int32_t *rg, *rd, *cold, *ncol, *tR, *tQL;
  u32 *om, *resm;
  int64_t *oe, *rese;
} lev_t;

static void *cauchy_worker(void *arg) {
  lev_t *j = (lev_t *)arg;
  const int L = j->L, Kc = j->Kc;
  const int64_t N = j->N;
  u32 *xs = (u32 *)malloc((size_t)(Kc > 0 ? Kc : 1) * L * 4);
  int64_t *xe = (int64_t *)malloc((size_t)(Kc > 0 ? Kc : 1) * 8);
  int32_t *ds = (int32_t *)malloc((size_t)(Kc > 0 ? Kc : 1) * 4);
  for (int64_t u = j->tid; u < j->units; u += j->nthreads) {
    const int64_t h = u / N, c = u % N;
    const int g = j->rg[h];
    const int K = j->ncol[g];
    for (int i = 0; i < K; ++i) {
      const int64_t src = j->col[(int64_t)g * Kc + i] * N + c;
      memcpy(xs + (size_t)i * L, j->rm + (size_t)src * L, (size_t)L * 4);
      xe[i] = j->re[src];
      ds[i] = j->rd[h] - j->cold[(int64_t)g * Kc + i];
    }
    const int64_t hi = j->ghrow[g] * N + c;
    j->oe[h * N + c] = oc_cauchy(xs, xe, ds, K, L, j->hm + (size_t)hi * L, j->he[hi], j->om + (size_t)(h * N + c) * L);
  }
  free(xs);
  free(xe);
  free(ds);
  return 0;
}

static void run_threads(lev_t *base, void *(*fn)(void *)) {
  int n = base->nthreads < 1 ? 1 : (base->nthreads > 256 ? 256 : base->nthreads);
  pthread_t th[256];
  lev_t jobs[256];
  for (int i = 0; i < n; ++i) {
    jobs[i] = *base;
    jobs[i].nthreads = n;
    jobs[i].tid = i;
    pthread_create(&th[i], 0, fn, &jobs[i]);
  }
  for (int i = 0; i < n; ++i) pthread_join(th[i], 0);
}

/* H (nH, N) of one level: H[h][c] = cauchy(res[col[g][i]][c] for i < ncol[g], rd[h] - cold[g][i], hinf[ghrow[g]][c]),
 * g = rg[h] (b3d_cauchy_level's contract) */
int oc_cauchy_level(int L, int64_t N, const u32 *rm, const int64_t *re, const u32 *hm, const int64_t *he,
                    const int64_t *ghrow, const int32_t *rg, const int32_t *rd, int64_t nH, const int64_t *col,
                    const int32_t *cold, const int32_t *ncol, int Kc, u32 *Hm, int64_t *He, int nthreads) {
  if (L < 1 || L > OC_MAXL) return 2;
  lev_t j;
  memset(&j, 0, sizeof(j));
  j.L = L; j.N = N; j.rm = rm; j.re = re; j.hm = hm; j.he = he; j.ghrow = ghrow; j.rg = rg; j.rd = rd;
  j.col = col; j.cold = cold; j.ncol = ncol; j.Kc = Kc; j.om = Hm; j.oe = He; j.nthreads = nthreads;
  j.units = nH * N;
  run_threads(&j, cauchy_worker);
  return 0;
}

static void *pole_worker(void *arg) {
  lev_t *j = (lev_t *)arg;
  const int L = j->L, S = j->S, SL = j->SL, SR = j->SR;
  const int64_t B = j->B, N = (int64_t)SL * SR * B;
  u32 *hal = (u32 *)malloc((size_t)SL * SR * L * 4), *y = (u32 *)malloc((size_t)SR * SL * L * 4);
  u32 *yal = (u32 *)malloc((size_t)SL * L * 4);
  int64_t *hae = (int64_t *)malloc((size_t)SL * 8), *ye = (int64_t *)malloc((size_t)SR * SL * 8);
  for (int64_t u = j->tid; u < j->units; u += j->nthreads) {
    const int64_t t = u / B, k = u % B;
    const int64_t hrow = j->thrw[t] * N;
    /* the H operand: for each a, the entries b < SR at column k, aligned */
    for (int a = 0; a < SL; ++a) {
      const int64_t base = hrow + ((int64_t)a * SR) * B + k;
      hae[a] = align_vec(j->Hm + (size_t)base * L, j->He + base, B, SR, L, hal + (size_t)a * SR * L);
    }
    /* Y[b0][a] = R[b0] . H[a] */
    const int64_t rt = (int64_t)j->tR[t] * S;
    for (int b0 = 0; b0 < SR; ++b0)
      for (int a = 0; a < SL; ++a)
        ye[b0 * SL + a] = oc_dot_aligned(j->Rm + (size_t)((rt + b0) * S) * L, j->Re[rt + b0], hal + (size_t)a * SR * L,
                                         hae[a], SR, L, y + (size_t)(b0 * SL + a) * L);
    /* res[a0][b0] = qL[a0] . Y[b0] */
    const int64_t qt = (int64_t)j->tQL[t] * S;
    const int64_t out = j->tslot[t] * N;
    for (int b0 = 0; b0 < SR; ++b0) {
      const int64_t E = align_vec(y + (size_t)b0 * SL * L, ye + b0 * SL, 1, SL, L, yal);
      for (int a0 = 0; a0 < SL; ++a0) {
        const int64_t o = out + ((int64_t)a0 * SR + b0) * B + k;
        j->rese[o] = oc_dot_aligned(j->Qm + (size_t)((qt + a0) * S) * L, j->Qe[qt + a0], yal, E, SL, L,
                                    j->resm + (size_t)o * L);
      }
    }
  }
  free(hal);
  free(y);
  free(yal);
  free(hae);
  free(ye);
  return 0;
}

/* the pole step of one level (b3d_pole_level's contract): for task t, Y = R[tR[t]] . H[thrw[t]]^T, then
 * res[tslot[t]] = qL[tQL[t]] . Y^T; R and qL tables (ntab, S, S) entry-major with row exponents (ntab, S),
 * already aligned per row */
int oc_pole_level(int L, int S, int SL, int SR, int64_t B, int64_t T, const u32 *Hm, const int64_t *He,
                  const int64_t *thrw, const int32_t *tR, const int32_t *tQL, const int64_t *tslot, const u32 *Rm,
                  const int64_t *Re, const u32 *Qm, const int64_t *Qe, u32 *rm, int64_t *re, int nthreads) {
  if (L < 1 || L > OC_MAXL || SL < 1 || SR < 1 || SL > S || SR > S) return 2;
  lev_t j;
  memset(&j, 0, sizeof(j));
  j.L = L; j.S = S; j.SL = SL; j.SR = SR; j.B = B; j.Hm = Hm; j.He = He; j.thrw = thrw; j.tR = tR; j.tQL = tQL;
  j.tslot = tslot; j.Rm = Rm; j.Re = Re; j.Qm = Qm; j.Qe = Qe; j.resm = rm; j.rese = re; j.nthreads = nthreads;
  j.units = T * B;
  run_threads(&j, pole_worker);
  return 0;
}
