"""Residues -> numerator polynomials, restated from blocks_3d's algorithm.  TEST INFRASTRUCTURE
(oracle/__init__.py).

numerators() restates reference tests/test_blocks_derivatives.py:13-80 (an mpmath port of blocks_3d's
residues_to_derivatives, restore_r_delta, shift_fraction, together_with_factors) and adds blocks_3d's
shift_poles as described at reference blocksvibe/derivatives.py:27-33: the unprotected poles x_i != 0 are
replaced by the kept ones x'_k with residues a' = A^-1 V a, A_rk = x'_k^(r-h), V_ri = x_i^(r-h), r < size,
h = size // 2 -- solved here by exact Gaussian elimination over the rationals (independent of the closed
form the product uses).  Pinned by tests/test_oracle_pins.py: numerator_e(x) / prod (x - x_i) equals the closed
form of the truncated radial series' e-th log r derivative at rational x off the poles (every step: the moment
weights, restore_r_delta's division, the shift to x, together_with_factors), and, with pole shifting, the same
identity over the kept poles when the other residues vanish; by tests/test_oracle_derivs.py: the LU solve
against the Lagrange closed form.
"""
from fractions import Fraction
from math import comb

import mpmath as mp

from . import structure_ref as S


def pmul(p, q):
    out = [mp.mpf(0)] * (len(p) + len(q) - 1)
    for i, x in enumerate(p):
        for j, y in enumerate(q):
            out[i + j] += x * y
    return out


def padd(p, q):
    n = max(len(p), len(q))
    return [(p[i] if i < len(p) else 0) + (q[i] if i < len(q) else 0) for i in range(n)]


def pdiv_lin(p, r):
    """p(x) = (x - r) q(x) + rem"""
    q = [mp.mpf(0)] * max(len(p) - 1, 0)
    acc = mp.mpf(0)
    for k in range(len(p) - 1, 0, -1):
        acc = p[k] + r * acc
        q[k - 1] = acc
    return q, (p[0] + r * acc if p else mp.mpf(0))


def solve_exact(A, V):
    """A^-1 V over the rationals (A square, nonsingular)"""
    n = len(A)
    M = [list(A[r]) + list(V[r]) for r in range(n)]
    w = len(M[0])
    for c in range(n):
        piv = next(r for r in range(c, n) if M[r][c] != 0)
        M[c], M[piv] = M[piv], M[c]
        inv = 1 / M[c][c]
        M[c] = [x * inv for x in M[c]]
        for r in range(n):
            if r != c and M[r][c] != 0:
                f = M[r][c]
                M[r] = [x - f * y for x, y in zip(M[r], M[c])]
    return [row[n:w] for row in M]


def shift_matrix(xs, keep):
    """S (size x len(xs)) with a' = S a for the unprotected old poles xs (exact Fractions)"""
    size = len(keep)
    h = size // 2
    A = [[Fraction(k) ** (r - h) for k in keep] for r in range(size)]
    V = [[Fraction(x) ** (r - h) for x in xs] for r in range(size)]
    return solve_exact(A, V)


def numerators(h, res, tj, p, a, b, order, R, tj12, tj43, kept=None):
    """h(p, m, j, a, b), res(p, m, j, pole_index, a, b) -> [numerator polynomial (x^k coefficients) for
    e = 0..R] of one (spin, parity, structure)"""
    rc = 3 - 2 * mp.sqrt(2)
    W = [[(mp.mpf(n) ** k if (n or k) else mp.mpf(1)) * rc ** n for n in range(order + 1)] for k in range(R + 1)]
    ji = (tj - tj12 % 2) // 2
    c = [sum(W[k][m] * h(p, m, ji, a, b) for m in range(order + 1)) for k in range(R + 1)]
    poles = []
    for P in S.poles(tj, tj12, tj43, order):
        s = S.shift(P)
        v = [sum(W[k][ii + s] * res(p, ii, ji, S.position(tj, tj12, tj43, order - ii, P), a, b)
                 for ii in range(order - s + 1)) for k in range(R + 1)]
        poles.append((Fraction(S.delta_2(P, tj), 2), v))
    rdelta = lambda e, d: [mp.mpf(0)] * (e - d) + [mp.mpf(comb(e, d))] if d <= e else [mp.mpf(0)]
    ub = Fraction(tj, 2) + (Fraction(1, 2) if tj <= 1 else 1)
    xs = [D - ub for D, _ in poles]
    if kept is not None and kept < order:
        prot = [x for x in xs if x == 0]
        keep = [Fraction(S.delta_2(P, tj), 2) - ub for P in S.poles(tj, tj12, tj43, kept)
                if Fraction(S.delta_2(P, tj), 2) != ub]
        unp = [i for i, x in enumerate(xs) if x != 0]
        Sm = shift_matrix([xs[i] for i in unp], keep) if keep and unp else None
        new = prot + keep
    else:
        new = xs
    num = lambda f: mp.mpf(f.numerator) / f.denominator
    prod = [mp.mpf(1)]
    for x in new:
        prod = pmul(prod, [-num(x), mp.mpf(1)])
    comp = []
    for i in range(len(new)):
        q = [mp.mpf(1)]
        for k, x in enumerate(new):
            if k != i:
                q = pmul(q, [-num(x), mp.mpf(1)])
        comp.append(q)
    out = []
    for e in range(R + 1):
        poly, resid = [mp.mpf(0)], []
        for d in range(R + 1):
            poly = padd(poly, [t * c[d] for t in rdelta(e, d)])
        for D, v in poles:
            rdp = [mp.mpf(0)]
            for d in range(R + 1):
                rdp = padd(rdp, [t * v[d] for t in rdelta(e, d)])
            q, rem = pdiv_lin(rdp, num(D))
            poly = padd(poly, q)
            resid.append(rem)
        shifted, power = [mp.mpf(0)], [mp.mpf(1)]
        for t in poly:
            shifted = padd(shifted, [t * y for y in power])
            power = pmul(power, [num(ub), mp.mpf(1)])
        if kept is not None and kept < order:
            rprot = [r for r, x in zip(resid, xs) if x == 0]
            runp = [resid[i] for i in unp]
            rkeep = [sum((num(Sm[k][i]) * runp[i] for i in range(len(unp))), mp.mpf(0)) for k in range(len(keep))] \
                if Sm is not None else [mp.mpf(0)] * len(keep)
            rnew = rprot + rkeep
        else:
            rnew = resid
        tot = pmul(shifted, prod)
        for r, q in zip(rnew, comp):
            tot = padd(tot, [r * y for y in q])
        out.append(tot)
    return out
