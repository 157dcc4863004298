"""Pins of the derivative oracle and of the product's exact numerator matrix:
  * blocks_3d's pole shifting matrix S = A^-1 V (exact LU in the oracle) equals the Lagrange closed form
    S_ki = (x'_k / x_i)^h l_k(x_i), computed here independently;
  * kept_pole_order >= order leaves the numerators unchanged;
  * the product's numerator columns (closed form, exact rationals) reproduce the oracle's pole-shifted
    together_with_factors on exact inputs."""
from fractions import Fraction as F

import mpmath as mp

from oracle import derivs_ref as DR
from oracle import structure_ref as OS


def lagrange_S(xs, keep):
    h = len(keep) // 2
    S = []
    for k, yk in enumerate(keep):
        row = []
        for x in xs:
            l = F(1)
            for m, ym in enumerate(keep):
                if m != k:
                    l *= F(x - ym) / (yk - ym)
            row.append((F(yk) / x) ** h * l)
        S.append(row)
    return S


def test_shift_matrix_closed_form():
    for keep, xs in (([F(-1, 2), F(-3, 2), F(-5, 2)], [F(-7, 2), F(-9, 2), F(-1, 2)]),
                     ([F(-1), F(-2), F(-3), F(-4)], [F(-5), F(-11, 2), F(-13, 2)])):
        assert DR.shift_matrix(xs, keep) == lagrange_S(xs, keep)


def _toy_inputs(order, R):
    hv = lambda p, m, j, a, b: mp.mpf(m + 1) / (j + 2) * (1 if p == 0 else -1)
    rv = lambda p, m, j, q, a, b: mp.mpf(q + 1) / (m + 3) / (j + 1)
    return hv, rv


def test_kept_at_least_order_is_identity():
    mp.mp.prec = 300
    hv, rv = _toy_inputs(6, 2)
    a = DR.numerators(hv, rv, 2, 0, 0, 0, 6, 2, 2, 2, kept=None)
    b = DR.numerators(hv, rv, 2, 0, 0, 0, 6, 2, 2, 2, kept=6)
    assert a == b


def test_product_numerator_columns_match_oracle():
    """the product's exact numerator matrix applied to random (polynomial, residue) data equals the
    oracle's shift_poles + together_with_factors"""
    from blocksvibe_amd.derivatives import numerator_columns
    mp.mp.prec = 400
    tj12, tj43, order, kept, R = 2, 2, 9, 4, 2
    for tj in (0, 2, 4):
        cols, npj = numerator_columns(tj, tj12, tj43, order, kept, R)
        poles = OS.poles(tj, tj12, tj43, order)
        ub = F(tj, 2) + (F(1, 2) if tj <= 1 else 1)
        xs = [F(OS.delta_2(P, tj), 2) - ub for P in poles]
        keep = [F(OS.delta_2(P, tj), 2) - ub for P in OS.poles(tj, tj12, tj43, kept) if F(OS.delta_2(P, tj), 2) != ub]
        prot = [x for x in xs if x == 0]
        unp = [i for i, x in enumerate(xs) if x != 0]
        S = DR.shift_matrix([xs[i] for i in unp], keep)
        new = prot + keep
        # random residues a_i and polynomial coefficients P[t] (exact rationals)
        a = [F(3 * i + 1, 7) for i in range(len(xs))]
        P = [F(t + 2, 5) for t in range(R + 1)]
        # oracle: P(x + ub) Pi_new(x) + sum_new a'_k Pi_new / (x - x_k)
        rprot = [a[i] for i, x in enumerate(xs) if x == 0]
        rkeep = [sum(S[k][j] * a[i] for j, i in enumerate(unp)) for k in range(len(keep))]
        rnew = rprot + rkeep

        def pmul(p, q):
            out = [F(0)] * (len(p) + len(q) - 1)
            for i, x in enumerate(p):
                for j, y in enumerate(q):
                    out[i + j] += x * y
            return out
        Pi = [F(1)]
        for x in new:
            Pi = pmul(Pi, [-x, F(1)])
        shifted = [F(0)]
        pw = [F(1)]
        for t in range(R + 1):
            shifted = [u + v for u, v in zip(shifted + [F(0)] * (len(pw) - len(shifted)), [P[t] * y for y in pw])]
            pw = pmul(pw, [ub, F(1)])
        want = pmul(shifted, Pi)
        for r, x in zip(rnew, new):
            q = [F(1)]
            for y in new:
                if y != x:
                    q = pmul(q, [-y, F(1)])
            want = [u + v for u, v in zip(want + [F(0)] * max(0, len(q) - len(want)),
                                          [r * c for c in q] + [F(0)] * max(0, len(want) - len(q)))]
        got = [F(0)] * max(len(want), npj + R + 1)
        for key, poly in cols.items():
            coef = P[key[1]] if isinstance(key, tuple) else a[key]
            for d, c in enumerate(poly):
                got[d] += coef * c
        want = want + [F(0)] * (len(got) - len(want))
        assert got == want, tj
