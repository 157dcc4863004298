"""The gfx950 kernels' per-thread multiprecision helpers (blocksvibe_amd/csrc/mp.hpp), compiled for the
host CPU (tests/host/mp_host.cpp), against the oracle bit for bit: add, mul (28-bit limb products), norm at
every accumulator width the kernels use, the FP64 one-limb division (every digit exact, divisors up to
2^15 - 1), barrel shifts (logical and arithmetic), Cauchy entries with h_inf, pole-step dot products, moment
windows and the derivative stage's triangular products by exact integer matrices.  CPU only;
the same source is what the GPU runs."""
import ctypes
import os
import random
import subprocess

import numpy as np
import pytest

from oracle import mpf_ref as R

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "host", "mp_host.so")
CLANG = "/opt/rocm/llvm/bin/clang++"
P, I64 = ctypes.c_void_p, ctypes.c_int64


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(CLANG):
        pytest.skip("clang++ from ROCm is needed for the host build of mp.hpp")
    tmp = f"{SO}.{os.getpid()}.tmp"
    subprocess.run([CLANG, "-O1", "-std=c++17", "-fPIC", "-shared", "-o", tmp,
                    os.path.join(HERE, "host", "mp_host.cpp")], check=True)
    os.replace(tmp, SO)
    lb = ctypes.CDLL(SO)
    for f in (lb.hc_add, lb.hc_mul, lb.hc_norm, lb.hc_div, lb.hc_shift, lb.hc_cauchy, lb.hc_dot, lb.hc_mom,
              lb.hc_tri, lb.hc_dot_staged, lb.hc_cauchy2):
        f.restype = I64
    return lb


def limbs(m, L):
    return np.array(R.to_limbs(m, L), dtype=np.uint32)


def ptr(a):
    return a.ctypes.data_as(P)


def rp(rng, L, spread=40):
    if rng.random() < 0.08:
        return 0, R.EZ
    nb = R.NB(L)
    m = rng.getrandbits(nb - 1) | (1 << (nb - 1))
    if rng.random() < 0.1:
        m = (1 << (nb - 1)) if rng.random() < 0.5 else (1 << nb) - 1
    return (-m if rng.random() < 0.5 else m), rng.randint(-spread, spread) - nb


def got(r, e):
    m = R.from_limbs(list(r))
    return (m, e) if m else (0, R.EZ)


LS = [2, 4, 10, 34, 50]


@pytest.mark.parametrize("L", LS)
def test_add_mul(lib, L):
    rng = random.Random(L)
    for i in range(600):
        a, b = rp(rng, L), rp(rng, L)
        if i % 7 == 0 and a[0]:
            b = (-a[0], a[1])
        if i % 11 == 3 and a[0]:
            b = (-a[0] + 1, a[1])
        r = np.zeros(L, np.uint32)
        e = lib.hc_add(L, ptr(limbs(a[0], L)), I64(a[1]), ptr(limbs(b[0], L)), I64(b[1]), ptr(r))
        assert got(r, e) == R.add(a, b, L), (i, a, b)
        e = lib.hc_mul(L, ptr(limbs(a[0], L)), I64(a[1]), ptr(limbs(b[0], L)), I64(b[1]), ptr(r))
        assert got(r, e) == R.mul(a, b, L), (i, a, b)


@pytest.mark.parametrize("L", LS)
def test_norm_widths(lib, L):
    rng = random.Random(100 + L)
    for LW in (L + 3, 2 * L + 2, 2 * L + 4):
        for i in range(400):
            bits = rng.randint(1, 32 * LW - 2)
            w = rng.getrandbits(bits)
            if i % 5 == 0 and bits > 32 * L:                   # exact ties
                s = bits - R.NB(L)
                w = (w >> s << s) | (1 << (s - 1))
            if i % 9 == 0 and bits > 32 * L:                   # all ones below the top: rounding carries out
                w = (1 << bits) - 1
            if i % 13 == 0:
                w = 0
            w = -w if rng.random() < 0.5 else w
            r = np.zeros(L, np.uint32)
            e = lib.hc_norm(L, ptr(limbs(w, LW)), LW, I64(12345), ptr(r))
            assert (R.from_limbs(list(r)), e) == R.norm(w, 12345, L), (LW, i)


@pytest.mark.parametrize("L", LS)
def test_div_small_exact(lib, L):
    """floor(m 2^32 / d) of a signed L-limb mantissa, every quotient digit exact (floor toward -infinity)"""
    rng = random.Random(200 + L)
    ds = [1, 2, 3, 7, 255, 256, 1000, (1 << 15) - 1, (1 << 15) - 2, (1 << 14) + 1]
    for i in range(600):
        d = ds[i % len(ds)] if i < 3 * len(ds) else rng.randint(1, (1 << 15) - 1)
        nb = R.NB(L)
        x = rng.choice([rng.getrandbits(nb), (1 << nb) - 1, (1 << (32 * L - 1)) - 1, d * rng.getrandbits(nb - 16),
                        d * ((1 << (32 * L - 16)) - 1), 1, 0])
        if i % 2:
            x = -x
        if i % 97 == 5:
            x = -(1 << (32 * L - 1))                                   # the most negative L-limb value
        q = np.zeros(L + 1, np.uint32)
        lib.hc_div(L, ptr(limbs(x, L)), d, ptr(q))
        assert R.from_limbs(list(q)) == (x << 32) // d, (d, i, x)


def test_div_every_divisor(lib):
    """every divisor 0 < d < 2^15 with dividends whose digits meet the extremes: exact multiples of d (each
    remainder 0, where a digit estimate below the exact quotient gives q - 1), remainders d - 1 at every step
    (the largest fractional part), all-ones and the most negative value"""
    L = 4
    nb = 32 * L
    for d in range(1, 1 << 15):
        for x in (d * ((1 << (nb - 16)) - 1), ((1 << (nb - 1)) - 1), -(1 << (nb - 1)), -d * 12345,
                  (((1 << (nb - 16)) // d) * d) - 1):
            q = np.zeros(L + 1, np.uint32)
            lib.hc_div(L, ptr(limbs(x, L)), d, ptr(q))
            assert R.from_limbs(list(q)) == (x << 32) // d, (d, x)


@pytest.mark.parametrize("L", [4, 34])
def test_shifts(lib, L):
    rng = random.Random(300 + L)
    W = L + 3
    for i in range(300):
        x = rng.getrandbits(32 * W)
        s = rng.choice([0, 1, 31, 32, 33, 64, rng.randint(0, 32 * W + 40)])
        out = np.zeros(W, np.uint32)
        lib.hc_shift(L, ptr(limbs(x, W)), s, 0, ptr(out))
        assert R.from_limbs(list(out) + [0]) == x >> s
        lib.hc_shift(L, ptr(limbs(x, W)), s, 1, ptr(out))
        assert R.from_limbs(list(out) + [0]) == (x << s) & ((1 << (32 * W)) - 1)
        lib.hc_shift(L, ptr(limbs(x, W)), s, 2, ptr(out))                # arithmetic: floor(x / 2^s)
        xs = x - (1 << (32 * W)) if x >> (32 * W - 1) else x
        assert R.from_limbs(list(out)) == xs >> s, (i, s)


@pytest.mark.parametrize("L", [4, 10, 34])
def test_cauchy_entry(lib, L):
    rng = random.Random(400 + L)
    for trial in range(300):
        nc = rng.randint(1, 12)
        N, nslot = 3, nc + 2
        vals = [rp(rng, L, spread=rng.choice([3, 200])) for _ in range(nslot * N)]
        if trial % 17 == 0:
            vals = [(0, R.EZ)] * len(vals)
        rm = np.stack([limbs(v[0], L) for v in vals], 1).view(np.int32).copy()
        re_ = np.array([v[1] for v in vals], np.int64)
        colg = np.array(rng.sample(range(nslot), nc), np.int64)
        big = trial % 2 == 1
        coldg = np.array([rng.randint(-16000, 16000) if big else rng.randint(-300, 300) for _ in range(nc)], np.int32)
        D = rng.randint(-16000, 16000) if big else (int(coldg[rng.randrange(nc)]) if rng.random() < 0.5
                                                   else rng.randint(-300, 300))
        hv = [rp(rng, L, spread=rng.choice([3, 300])) for _ in range(2 * N)]     # h_inf rows 0 and 1
        n = rng.randrange(N)
        if trial % 5 == 0:
            hv[N + n] = (0, R.EZ)
        hm = np.stack([limbs(v[0], L) for v in hv], 1).view(np.int32).copy()
        he = np.array([v[1] for v in hv], np.int64)
        xs = [vals[int(c) * N + n] for c in colg]
        ds = [D - int(c) for c in coldg]
        want = R.cauchy(xs, ds, L, hv[N + n])
        for uni in (0, 1):                                  # per-lane indices and wave-uniform group
            r = np.zeros(L, np.uint32)
            e = lib.hc_cauchy(L, ptr(rm), ptr(re_), I64(nslot * N), I64(N), I64(n), ptr(colg), ptr(coldg), nc, D,
                              ptr(hm), ptr(he), I64(2 * N), I64(1), uni, ptr(r))
            assert got(r, e) == want, (trial, uni)


@pytest.mark.parametrize("L", [4, 10, 34, 50])
def test_cauchy_entry2(lib, L):
    """two rows of a group per lane (the kernel of large levels) against the oracle for each row: a second row
    or none, either row at a divisor 0 (the term left out for that row only), h_inf zero, every column zero"""
    rng = random.Random(450 + L)
    for trial in range(200):
        nc = rng.randint(1, 12)
        N, nslot = 3, nc + 2
        vals = [rp(rng, L, spread=rng.choice([3, 200])) for _ in range(nslot * N)]
        if trial % 13 == 0:
            vals = [(0, R.EZ)] * len(vals)
        rm = np.stack([limbs(v[0], L) for v in vals], 1).view(np.int32).copy()
        re_ = np.array([v[1] for v in vals], np.int64)
        colg = np.array(rng.sample(range(nslot), nc), np.int64)
        big = trial % 2 == 1
        coldg = np.array([rng.randint(-16000, 16000) if big else rng.randint(-300, 300) for _ in range(nc)], np.int32)
        pick = lambda: int(coldg[rng.randrange(nc)])
        D1 = pick() if trial % 4 == 1 else (rng.randint(-16000, 16000) if big else rng.randint(-300, 300))
        D2 = pick() if trial % 4 == 2 else (rng.randint(-16000, 16000) if big else rng.randint(-300, 300))
        has2 = trial % 7 != 3
        hv = [rp(rng, L, spread=rng.choice([3, 300])) for _ in range(2 * N)]
        n = rng.randrange(N)
        if trial % 5 == 0:
            hv[N + n] = (0, R.EZ)
        hm = np.stack([limbs(v[0], L) for v in hv], 1).view(np.int32).copy()
        he = np.array([v[1] for v in hv], np.int64)
        xs = [vals[int(c) * N + n] for c in colg]
        om = np.zeros((L, 2), np.int32)
        oe = np.zeros(2, np.int64)
        lib.hc_cauchy2(L, ptr(rm), ptr(re_), I64(nslot * N), I64(N), I64(n), ptr(colg), ptr(coldg), nc, D1, D2,
                       int(has2), ptr(hm), ptr(he), I64(2 * N), I64(1), ptr(om), ptr(oe))
        for row, D in enumerate((D1, D2) if has2 else (D1,)):
            want = R.cauchy(xs, [D - int(c) for c in coldg], L, hv[N + n])
            assert got(om[:, row].view(np.uint32), int(oe[row])) == want, (trial, row)


@pytest.mark.parametrize("L", [2, 4, 10, 34, 50])
def test_small_dot(lib, L):
    rng = random.Random(500 + L)
    for trial in range(300):
        S = rng.randint(1, 5)
        arow = [rp(rng, L, spread=rng.choice([3, 100])) for _ in range(S)]
        am_, ea = R.align(arow)
        xs = [rp(rng, L, spread=rng.choice([3, 100])) for _ in range(S)]
        A = np.stack([limbs(v, L) for v in am_], 1).view(np.int32).copy()
        X = np.stack([limbs(v[0], L) for v in xs], 1).view(np.int32).copy()
        XE = np.array([v[1] for v in xs], np.int64)
        r = np.zeros(L, np.uint32)
        e = lib.hc_dot(L, ptr(A), I64(S), I64(0), I64(ea), ptr(X), ptr(XE), I64(S), I64(0), I64(1), S, ptr(r))
        xm, xe = R.align(xs)
        assert got(r, e) == R.gemm_aligned(am_, ea, xm, xe, L), trial


@pytest.mark.parametrize("L", [4, 6, 10, 34, 50])
def test_small_dot_cancellation(lib, L):
    """sums far below their terms (near and exact cancellation), sums whose top bit is tens of bits below the
    largest term product, power-of-two patterns: the truncated product's certainty test must send these
    to the full product, so the result stays the exactly rounded sum"""
    rng = random.Random(550 + L)
    nb = R.NB(L)
    for trial in range(300):
        S = rng.randint(2, 6)
        kind = trial % 4
        arow = [rp(rng, L, spread=3) for _ in range(S)]
        xs = [rp(rng, L, spread=3) for _ in range(S)]
        if kind == 0:                                     # x_1 = -x_0 a_0 / a_1 up to a few low bits
            arow[1] = (arow[0][0], arow[0][1])
            xs[1] = (-xs[0][0] + rng.randint(-(1 << rng.randint(0, nb)), 1 << rng.randint(0, nb)), xs[0][1])
            if xs[1][0] == 0 or abs(xs[1][0]).bit_length() != nb:
                xs[1] = (-xs[0][0], xs[0][1])
        elif kind == 1:                                   # the largest a meets a small x and vice versa
            d = rng.randint(20, 200)
            arow = [((1 << (nb - 1)) + rng.getrandbits(nb - 2), -nb), ((1 << (nb - 1)) + rng.getrandbits(nb - 2), -nb - d)]
            xs = [((1 << (nb - 1)) + rng.getrandbits(nb - 2), -nb - d), ((1 << (nb - 1)) + rng.getrandbits(nb - 2), -nb)]
            S = 2
        elif kind == 2:                                   # powers of two and all-ones mantissas
            arow = [(rng.choice([1 << (nb - 1), (1 << nb) - 1, -(1 << (nb - 1))]), rng.randint(-2, 2) - nb)
                    for _ in range(S)]
            xs = [(rng.choice([1 << (nb - 1), (1 << (nb - 1)) + 1, (1 << nb) - 1]), rng.randint(-2, 2) - nb)
                  for _ in range(S)]
        am_, ea = R.align(arow)
        A = np.stack([limbs(v, L) for v in am_], 1).view(np.int32).copy()
        X = np.stack([limbs(v[0], L) for v in xs], 1).view(np.int32).copy()
        XE = np.array([v[1] for v in xs], np.int64)
        r = np.zeros(L, np.uint32)
        e = lib.hc_dot(L, ptr(A), I64(S), I64(0), I64(ea), ptr(X), ptr(XE), I64(S), I64(0), I64(1), S, ptr(r))
        xm, xe = R.align(xs)
        want = R.gemm_aligned(am_, ea, xm, xe, L)
        assert got(r, e) == want, (trial, kind)
        # the LDS-staged form of the kernels (terms aligned and converted first, constant's sign in its limbs)
        r2 = np.zeros(L, np.uint32)
        e2 = lib.hc_dot_staged(L, ptr(A), I64(S), I64(0), I64(ea), ptr(X), ptr(XE), I64(S), I64(0), I64(1), S,
                               ptr(r2))
        assert got(r2, e2) == want, ("staged", trial, kind)


@pytest.mark.parametrize("L", [4, 10, 34])
def test_mom_window(lib, L):
    rng = random.Random(600 + L)
    for trial in range(200):
        W, nrows, N, R1 = rng.randint(1, 8), 3, 2, rng.randint(1, 12)
        nmax = 80
        tab = [n ** k for n in range(nmax) for k in range(R1)]
        KL = max(1, (max(tab).bit_length() + 31) // 32)
        if KL > 4:
            continue
        NK = np.array([[(x >> (32 * l)) & 0xFFFFFFFF for x in tab] for l in range(KL)], dtype=np.uint32)
        NKb = np.array([x.bit_length() for x in tab], dtype=np.int32)
        vals = [rp(rng, L, spread=rng.choice([3, 60])) if rng.random() < 0.8 else (0, R.EZ)
                for _ in range(W * nrows * N)]
        sm = np.stack([limbs(v[0], L) for v in vals], 1).view(np.int32).copy()
        se = np.array([v[1] for v in vals], np.int64)
        row, c, k, n0 = rng.randrange(nrows), rng.randrange(N), rng.randrange(R1), rng.randint(0, nmax - W - 1)
        r = np.zeros(L, np.uint32)
        e = lib.hc_mom(L, ptr(sm), ptr(se), I64(len(vals)), W, I64(nrows), I64(N), I64(row), I64(c), n0, k, R1,
                       ptr(NK), I64(NK.shape[1]), KL, ptr(NKb), ptr(r))
        ys = [vals[(w * nrows + row) * N + c] for w in range(W)]
        nks = [(n0 + w) ** k for w in range(W)]
        assert got(r, e) == R.mom_flush(ys, nks, (0, R.EZ), L), trial


def tri_tables(ds, K):
    """the derivative stage's matrices C(e, k) D^(e-k), D = d / 2, as exact integers C(e, k) d^(e-k) 2^k with
    the row unit 2^-e: table 0 all zero (padding), then one per d"""
    from math import comb
    ints = [[0] * (K * K)] + [[comb(e, k) * d ** (e - k) * 2 ** k if k <= e else 0 for e in range(K) for k in range(K)]
                              for d in ds]
    return ints


@pytest.mark.parametrize("L", [4, 10, 34, 50])
def test_tri_entry(lib, L):
    """mp::tri_entry against the oracle's GEMM (row of C(e, k) D^(e-k) as normalised numbers, column of v)"""
    from fractions import Fraction
    rng = random.Random(700 + L)
    for trial in range(150):
        K = rng.randint(1, 20 if L >= 10 else 12)
        ds = [rng.choice([0, 1, -1, rng.randint(-320, 320)]) for _ in range(3)]
        ints = tri_tables(ds, K)
        bits = max(abs(x).bit_length() for row in ints for x in row)
        TL = next(t for t in (2, 4, 8) if bits <= 32 * t)
        ntab = len(ints)
        mags = [abs(x) for row in ints for x in row]
        tm = np.array([[(x >> (32 * l)) & 0xFFFFFFFF for x in mags] for l in range(TL)], dtype=np.uint32).view(np.int32)
        tsg = np.array([0xFFFFFFFF if x < 0 else 0 for row in ints for x in row], dtype=np.uint32)
        spread = rng.choice([2, 40, 2000])
        ncol = 3
        vals = [rp(rng, L, spread=spread) if rng.random() < 0.85 else (0, R.EZ) for _ in range(K * ncol)]
        if trial % 10 == 0:
            vals = [(0, R.EZ)] * (K * ncol)                      # all-zero columns
        vm = np.stack([limbs(v[0], L) for v in vals], 1).view(np.int32).copy()
        ve = np.array([v[1] for v in vals], np.int64)
        t, e, c = rng.randrange(ntab), rng.randrange(K), rng.randrange(ncol)
        r = np.zeros(L, np.uint32)
        ex = lib.hc_tri(L, TL, ptr(vm), ptr(ve), I64(len(vals)), I64(c), I64(ncol), K, e, ptr(tm), ptr(tsg),
                        I64(len(mags)), I64((t * K + e) * K), I64(-e), ptr(r))
        arow = [R.from_fraction(Fraction(ints[t][e * K + k], 2 ** e), L) for k in range(K)]
        col = [[vals[k * ncol + c]] for k in range(K)]
        assert got(r, ex) == R.gemm([arow], col, L)[0][0], (trial, t, e, c)
