"""Pins of the recursion oracles: the exact recursion (oracle.recursion_ref.recursion) against a direct
evaluation of the fill_residues formulas for the first two levels of a tiny operator, and the bit-level
restatement of the GPU semantics (recursion_bits) against the exact recursion to 2^-(32L - 64)."""
import math
import random
from fractions import Fraction as F

from oracle import mpf_ref as R
from oracle import recursion_ref as RR
from oracle import structure_ref as OS


def tiny(tj12, tj43, tjmax, order, pp, seed):
    from blocksvibe_amd.plan import Plan
    rng = random.Random(seed)
    plan = Plan(tj12, tj43, tjmax, order, pp, reuse=False)
    M = {}
    for (P, d, tjNN, ppp, tj) in plan.Mkeys:
        rows = len(OS.so3((ppp + OS.parity(P)) % 2, tjNN, tj))
        cols = len(OS.so3(ppp, tjNN, OS.pole_two_j(P, tj)))
        M[(P, d, tjNN, ppp, tj)] = [[F(rng.randint(-9, 9), rng.randint(1, 5)) for _ in range(cols)] for _ in range(rows)]
    H = {}
    for s in range(2):
        for m in range(order + 1):
            for j in range(plan.Jh):
                tj = tj12 % 2 + 2 * j
                for a in range(len(OS.so3(pp[s][0], tj12, tj))):
                    for b in range(len(OS.so3(pp[s][1], tj43, tj))):
                        H[(s, m, j, a, b)] = F(rng.randint(-9, 9), rng.randint(1, 7))
    return M, H


def test_first_levels_closed_form():
    tj12, tj43, tjmax, order, pp = 2, 2, 2, 2, ((0, 1), (1, 0))
    M, H = tiny(tj12, tj43, tjmax, order, pp, 5)
    res = RR.recursion(H, M, tj12, tj43, tjmax, order, pp)
    checked = 0
    for (p, m, jidx, pidx, a0, b0), v in res.items():
        tj = OS.two_j_values(m, order, tj12, tjmax)[jidx]
        P = OS.poles(tj, tj12, tj43, order - m - OS.j_excess(tj, tjmax))[pidx]
        tjp = OS.pole_two_j(P, tj)
        src = (p + OS.parity(P)) % 2
        LM, RM = M[(P, 0, tj12, pp[src][0], tj)], M[(P, 1, tj43, pp[src][1], tj)]
        sign = -1 if (tjp - tj) % 4 else 1
        Dp = OS.delta_2(P, tj) + 2 * OS.shift(P)

        def hfull(a, b):
            x = H.get((src, m, tjp // 2, a, b), F(0))
            for i in OS.poles(tjp, tj12, tj43, m):
                mm = m - OS.shift(i)
                q = OS.position(tjp, tj12, tj43, order - m + OS.shift(i) - OS.j_excess(tjp, tjmax), i)
                d = Dp - OS.delta_2(i, tjp)
                if d:
                    x += res.get((src, mm, tjp // 2, q, a, b), F(0)) * F(2, d)
            return x

        want = OS.q_value(P, tj) * sign * sum(LM[a0][a] * RM[b0][b] * hfull(a, b)
                                              for a in range(len(LM[0])) for b in range(len(RM[0])))
        assert v == want
        checked += 1
    assert checked > 20


def test_bits_within_bound():
    tj12, tj43, tjmax, order, pp, L = 4, 2, 4, 5, ((0, 1), (1, 0)), 4
    M, H = tiny(tj12, tj43, tjmax, order, pp, 6)
    ex = RR.recursion(H, M, tj12, tj43, tjmax, order, pp)
    Hp = {k: R.from_fraction(v, L) for k, v in H.items()}
    Mp = {k: [[R.from_fraction(x, L) for x in row] for row in v] for k, v in M.items()}
    from blocksvibe_amd.plan import Plan
    Sm = Plan(tj12, tj43, tjmax, order, pp).Smax
    bits = RR.recursion_bits(Hp, Mp, tj12, tj43, tjmax, order, pp, L, Sm)
    blk = {}
    for k, v in ex.items():
        blk[k[:4]] = max(blk.get(k[:4], 0), abs(v))
    worst = max(abs(R.value(*bits[k]) - v) / blk[k[:4]] for k, v in ex.items() if blk[k[:4]])
    assert -math.log2(worst) > 32 * L - 64


def test_linear_in_h_inf():
    tj12, tj43, tjmax, order, pp = 2, 2, 4, 4, ((0, 0), (1, 1))
    M, H1 = tiny(tj12, tj43, tjmax, order, pp, 8)
    _, H2 = tiny(tj12, tj43, tjmax, order, pp, 9)
    r1 = RR.recursion(H1, M, tj12, tj43, tjmax, order, pp)
    r2 = RR.recursion(H2, M, tj12, tj43, tjmax, order, pp)
    r3 = RR.recursion({k: H1[k] - 3 * H2[k] for k in H1}, M, tj12, tj43, tjmax, order, pp)
    assert r1.keys() == r3.keys()
    assert all(r3[k] == r1[k] - 3 * r2[k] for k in r1)
