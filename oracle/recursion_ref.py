"""The radial recursion of blocks_3d in exact arithmetic.  TEST INFRASTRUCTURE (oracle/__init__.py).

recursion()      -- restates reference blocksvibe/recursion_ref.py:11-44 (blocks_3d's fill_residues /
                    set_residue): any number type (Fraction for exact results).
recursion_bits() -- the GPU kernels' bit-level semantics for the same loop, on top of oracle/mpf_ref.py:
                    Cauchy step = mpf_ref.cauchy over all prime poles of the (src, j') block with
                    h_inf in the same exact sum (one rounding); pole step = two rounded S x S products with the constant operand
                    pre-aligned per row (R, and q L with q L rounded once), the residue operand aligned
                    per column.  Structures are padded to Smax (zero entries: no effect on alignment).

Inputs: h_inf[(pp, m, jidx, a, b)] and M[((t,k), delta, two_jNN, ppp, two_j)] -> matrix.
Output: res[(p, m, jidx, pole_index, a, b)].
Pinned by tests/test_golden.py: recursion() equals, residue for residue, the reference's own
recursion_ref.recursion on eight seeded operators (tests/golden/recursion.json, written by
tests/golden/make_fixtures.py from the reference's module); by tests/test_oracle_recursion.py: recursion_bits
against recursion() to its error bound.
"""
from fractions import Fraction

from . import mpf_ref as R
from . import structure_ref as S


def _tasks(two_j12, two_j43, two_j_max, order):
    """per task: (m, jidx, pidx, P, tj, tjp, dprime, terms, sign) in the loop order of fill_residues"""
    for m in range(order):
        for jidx, tj in enumerate(S.two_j_values(m, order, two_j12, two_j_max)):
            pl = S.poles(tj, two_j12, two_j43, order - m - S.j_excess(tj, two_j_max))
            for pidx, P in enumerate(pl):
                tjp = S.pole_two_j(P, tj)
                dprime = S.delta_2(P, tj) + 2 * S.shift(P)
                prime = S.poles(tjp, two_j12, two_j43, m)
                jpex = S.j_excess(tjp, two_j_max)
                terms = [(m - S.shift(i), S.position(tjp, two_j12, two_j43, order - m + S.shift(i) - jpex, i),
                          dprime - S.delta_2(i, tjp)) for i in prime]
                sign = -1 if (tjp - tj) % 4 else 1
                yield m, jidx, pidx, P, tj, tjp, dprime, terms, sign


def recursion(h_inf, M, two_j12, two_j43, two_j_max, order, pseudo_parities, num=Fraction):
    res = {}
    zero = num(0)
    for m, jidx, pidx, P, tj, tjp, dprime, terms, sign in _tasks(two_j12, two_j43, two_j_max, order):
        Q = num(S.q_value(P, tj))
        for p in range(2):
            src = (p + S.parity(P)) % 2
            LM = M.get((P, 0, two_j12, pseudo_parities[src][0], tj))
            RM = M.get((P, 1, two_j43, pseudo_parities[src][1], tj))
            if not LM or not LM[0] or not RM or not RM[0]:
                continue
            nl, nr = len(LM[0]), len(RM[0])
            coef = [(mm, q, num(Fraction(0) if d == 0 else Fraction(2, d))) for mm, q, d in terms]
            h = [[h_inf.get((src, m, tjp // 2, a, b), zero)
                  + sum((res.get((src, mm, tjp // 2, q, a, b), zero) * c for mm, q, c in coef), zero)
                  for b in range(nr)] for a in range(nl)]
            inter = [[sum((h[a][b] * RM[b0][b] for b in range(nr)), zero) * sign
                      for b0 in range(len(RM))] for a in range(nl)]
            for a0 in range(len(LM)):
                for b0 in range(len(RM)):
                    res[(p, m, jidx, pidx, a0, b0)] = Q * sum((LM[a0][a] * inter[a][b0] for a in range(nl)), zero)
    return res


def recursion_bits(h_inf, M, two_j12, two_j43, two_j_max, order, pseudo_parities, L, Smax):
    """h_inf and M values given as exact normalised (m, e) pairs at L limbs; returns (m, e) pairs"""
    Z = (0, R.EZ)
    res = {}
    for m, jidx, pidx, P, tj, tjp, dprime, terms, sign in _tasks(two_j12, two_j43, two_j_max, order):
        Qp = R.from_fraction(S.q_value(P, tj) * sign, L)
        for p in range(2):
            src = (p + S.parity(P)) % 2
            LM = M.get((P, 0, two_j12, pseudo_parities[src][0], tj))
            RM = M.get((P, 1, two_j43, pseudo_parities[src][1], tj))
            if not LM or not LM[0] or not RM or not RM[0]:
                continue
            Lp = [[LM[a][b] if a < len(LM) and b < len(LM[0]) else Z for b in range(Smax)] for a in range(Smax)]
            Rp = [[RM[a][b] if a < len(RM) and b < len(RM[0]) else Z for b in range(Smax)] for a in range(Smax)]
            # Cauchy step, per structure entry (a, b)
            H = [[None] * Smax for _ in range(Smax)]
            for a in range(Smax):
                for b in range(Smax):
                    xs = [res.get((src, mm, tjp // 2, q, a, b), Z) for mm, q, d in terms]
                    ds = [d for mm, q, d in terms]
                    H[a][b] = R.cauchy(xs, ds, L, h_inf.get((src, m, tjp // 2, a, b), Z))
            # pole step: Y[b0][a] = sum_b R[b0][b] H[a][b];  out[a0][b0] = sum_a qL[a0][a] Y[b0][a]
            Ral = [R.align(Rp[b0]) for b0 in range(Smax)]
            qL = [[R.mul(Lp[a0][a], Qp, L) for a in range(Smax)] for a0 in range(Smax)]
            qLal = [R.align(qL[a0]) for a0 in range(Smax)]
            Y = [[None] * Smax for _ in range(Smax)]
            for b0 in range(Smax):
                for a in range(Smax):
                    hm, he = R.align(H[a])
                    Y[b0][a] = R.gemm_aligned(Ral[b0][0], Ral[b0][1], hm, he, L)
            for a0 in range(Smax):
                for b0 in range(Smax):
                    ym, ye = R.align(Y[b0])
                    res[(p, m, jidx, pidx, a0, b0)] = R.gemm_aligned(qLal[a0][0], qLal[a0][1], ym, ye, L)
    return res
