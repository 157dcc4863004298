"""The oracle's C restatement (oracle/c/mpf_c.c) against mpf_ref, bit for bit: norm, Cauchy sums and
pole-step dot products on random and edge-case inputs.  CPU only."""
import random

import pytest

from oracle import c_ref as C
from oracle import mpf_ref as R


@pytest.fixture(scope="module", autouse=True)
def built():
    from oracle.c import build as B  # noqa: F401
    B.build()


def rnum(rng, L, spread=3, zero=0.0):
    if rng.random() < zero:
        return 0, R.EZ
    nb = R.NB(L)
    m = rng.getrandbits(nb - 1) | (1 << (nb - 1))
    return (-m if rng.random() < 0.5 else m), rng.randint(-spread, spread) - nb


@pytest.mark.parametrize("L", [2, 3, 4, 10, 34])
def test_norm(L):
    rng = random.Random(L)
    n = 2 * L + 2
    for _ in range(300):
        W = rng.getrandbits(rng.randint(1, 32 * n - 2)) * rng.choice((-1, 1))
        e = rng.randint(-500, 500)
        assert C.norm(W, n, e, L) == R.norm(W, e, L)


def test_norm_ties_small():
    L = 2
    NB = R.NB(L)
    for k in range(1, 40):
        for low in range(0, 1 << min(k, 6)):
            W = ((1 << NB) - 1 - low) << k | (1 << (k - 1))     # exact ties and carries into a new bit
            for s in (1, -1):
                assert C.norm(s * W, 2 * L + 2, 0, L) == R.norm(s * W, 0, L)


@pytest.mark.parametrize("L", [2, 4, 10, 34])
def test_cauchy(L):
    rng = random.Random(100 + L)
    for trial in range(200):
        K = rng.randint(1, 64)
        spread = rng.choice([0, 3, 40, 32 * L + 100])
        xs = [rnum(rng, L, spread, zero=0.2) for _ in range(K)]
        ds = [rng.choice([0, 1, -1, 2, (1 << 15) - 1, -((1 << 15) - 1)] + [rng.randint(-300, 300)])
              for _ in range(K)]
        assert C.cauchy(xs, ds, L) == R.cauchy(xs, ds, L), (trial, K, spread)
        h = rnum(rng, L, rng.choice([0, 5, 40, 32 * L + 100]), zero=0.1)
        h = (h[0], h[1] + rng.choice([0, 0, -60, 60]))
        assert C.cauchy(xs, ds, L, h) == R.cauchy(xs, ds, L, h), (trial, K, spread, "h")
    assert C.cauchy([(0, R.EZ)] * 3, [1, 2, 3], L) == (0, R.EZ)
    h = rnum(rng, L)
    assert C.cauchy([(0, R.EZ)] * 3, [1, 2, 3], L, h) == h == R.cauchy([(0, R.EZ)] * 3, [1, 2, 3], L, h)


@pytest.mark.parametrize("L", [2, 4, 10, 34])
def test_dot_aligned(L):
    rng = random.Random(200 + L)
    for trial in range(200):
        K = rng.randint(1, 8)
        a = [rnum(rng, L, rng.choice([0, 5, 70]), zero=0.2) for _ in range(K)]
        b = [rnum(rng, L, rng.choice([0, 5, 70]), zero=0.2) for _ in range(K)]
        am, ae = R.align(a)
        bm, be = R.align(b)
        assert C.dot_aligned(am, ae, bm, be, L) == R.gemm_aligned(am, ae, bm, be, L), trial


def test_timer_runs():
    assert C.time_unit(0, 10, 16, 2, 0.05) > 0
    assert C.time_unit(1, 10, 5, 2, 0.05) > 0


@pytest.mark.parametrize("tj12,tj43,tjmax,order,pp,L,B,reuse", [
    (4, 2, 4, 5, ((0, 1), (1, 0)), 10, 2, False),
    (8, 8, 2, 4, ((0, 0), (1, 1)), 6, 1, True),
    (3, 1, 5, 5, ((0, 1), (1, 0)), 4, 3, False)])
def test_c_recursion_equals_recursion_bits(tj12, tj43, tjmax, order, pp, L, B, reuse):
    """oracle/c_chain.py (the C level loop over the product plan's index tables) bit for bit against
    recursion_ref.recursion_bits, every residue of every right-hand side (with slot reuse: the last residue
    written to each slot)"""
    import random
    from fractions import Fraction as Fr
    import numpy as np
    from blocksvibe_amd.plan import Plan
    from oracle import c_chain as CC
    from oracle import recursion_ref as RR
    from oracle import structure_ref as OS
    rng = random.Random(tj12 * 7 + order)
    rnd = lambda: Fr(rng.uniform(-2, 2)) + Fr(rng.random()) * Fr(2) ** -90
    plan = Plan(tj12, tj43, tjmax, order, pp, reuse=reuse)
    S, SL, SR, S2, O1 = plan.Smax, plan.SLm, plan.SRm, plan.S2, order + 1
    M = {}
    for (P, d, tjNN, ppp, tj) in plan.Mkeys:
        rows = len(OS.so3((ppp + OS.parity(P)) % 2, tjNN, tj))
        cols = len(OS.so3(ppp, tjNN, OS.pole_two_j(P, tj)))
        M[(P, d, tjNN, ppp, tj)] = [[R.from_fraction(rnd(), L) for _ in range(cols)] for _ in range(rows)]
    Mtab = [(0, R.EZ)] * (max(len(plan.Mkeys), 1) * S * S)
    for key, i in plan.Mkeys.items():
        for a, row in enumerate(M[key]):
            for b, x in enumerate(row):
                Mtab[(i * S + a) * S + b] = x
    H = []
    flat = [(0, R.EZ)] * (2 * O1 * plan.Jh * S2 * B)
    for k in range(B):
        h = {}
        for s in range(2):
            for m in range(O1):
                for j in range(plan.Jh):
                    tj = tj12 % 2 + 2 * j
                    for a in range(len(OS.so3(pp[s][0], tj12, tj))):
                        for b in range(len(OS.so3(pp[s][1], tj43, tj))):
                            h[(s, m, j, a, b)] = R.from_fraction(rnd(), L)
                            flat[((((s * O1 + m) * plan.Jh + j) * SL + a) * SR + b) * B + k] = h[(s, m, j, a, b)]
        H.append(h)
    hm, he = CC.planes(flat, L)
    rm, re = CC.recursion(plan, Mtab, hm, he, L, B, threads=3)
    got = {}
    for k in range(B):
        bits = RR.recursion_bits(H[k], M, tj12, tj43, tjmax, order, pp, L, S)
        final = {}
        for (p, m, j, q, a, b), v in bits.items():
            if a >= SL or b >= SR:
                continue
            slot = plan.phys_slot(p * plan.nslots + plan.offset(m, j) + q)
            final[(slot, a, b)] = v                     # with reuse the last write wins (loop order = level order)
        for (slot, a, b), v in final.items():
            i = (slot * S2 + a * SR + b) * B + k
            g = (R.from_limbs(rm[i].tolist()), int(re[i]))
            g = g if g[0] else (0, R.EZ)
            assert g == v, (k, slot, a, b)
        got[k] = len(final)
    assert all(got.values())
