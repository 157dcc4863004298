"""GPU post-recursion stages (blocksvibe_amd.derivatives: moments staged and flushed during a recursion
with recycled slots, then the c / a / G / H / numerator GEMMs) against oracle.derivs_ref on the exact
residues of the same random operator (reference tests/test_blocks_derivatives.py:83-118), with and
without pole shifting.  Bound: 2^-(32L - 40) relative to the largest coefficient of each polynomial;
coefficients beyond the polynomial's degree are exactly zero."""
from fractions import Fraction

import mpmath as mp
import pytest

from oracle import derivs_ref as DR
from oracle import mpf_ref as R
from oracle import recursion_ref as RR
from oracle import structure_ref as OS
from tests.test_gpu_recursion import h_tensor, random_operator

pytestmark = pytest.mark.gpu


def f2m(x):
    return mp.mpf(x.numerator) / x.denominator

CASES = [dict(tj12=4, tj43=2, tjmax=6, order=7, R=3, L=10, B=2, seed=1, kept=None),
         dict(tj12=8, tj43=8, tjmax=4, order=6, R=5, L=10, B=1, seed=2, kept=None),
         dict(tj12=3, tj43=1, tjmax=5, order=5, R=0, L=10, B=2, seed=3, kept=None),
         dict(tj12=2, tj43=2, tjmax=6, order=9, R=4, L=10, B=2, seed=4, kept=4),
         dict(tj12=4, tj43=4, tjmax=4, order=8, R=3, L=18, B=1, seed=5, kept=6)]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c['tj12']}-{c['tj43']}-{c['tjmax']}-o{c['order']}-R{c['R']}-k{c['kept']}")
def test_derivs(case):
    from blocksvibe_amd.derivatives import Derivs
    from blocksvibe_amd.plan import Plan
    from blocksvibe_amd.recursion import Recursion
    c = case
    L, B, pp, Rr = c["L"], c["B"], ((0, 1), (1, 0)), c["R"]
    plan = Plan(c["tj12"], c["tj43"], c["tjmax"], c["order"], pp, reuse=True)
    M, H = random_operator(plan, c["tj12"], c["tj43"], pp, c["order"], B, c["seed"])
    h = h_tensor(H, plan, L)
    rec = Recursion(plan, M, L)
    der = Derivs(plan, Rr, L, kept_pole_order=c["kept"])
    der.start(B, window=3)
    rec.run(h, hook=der.accumulate)
    num = der.finish(h, chunk=2)
    got = num.to_host()
    SR, S2 = plan.SRm, plan.S2
    Dout = der.Dout
    mp.mp.prec = 32 * L + 64
    worst, cnt = 0, 0
    for k in range(B):
        exact = RR.recursion(H[k], M, c["tj12"], c["tj43"], c["tjmax"], c["order"], pp)
        hk = H[k]
        hf = lambda p, m, j, a, b: f2m(hk.get((p, m, j, a, b), Fraction(0)))
        rf = lambda p, m, j, q, a, b: f2m(exact.get((p, m, j, q, a, b), Fraction(0)))
        for jo, tj in enumerate(der.tjs):
            for p in range(2):
                for a in range(len(OS.so3(pp[p][0], c["tj12"], tj))):
                    for b in range(len(OS.so3(pp[p][1], c["tj43"], tj))):
                        ref = DR.numerators(hf, rf, tj, p, a, b, c["order"], Rr, c["tj12"], c["tj43"], kept=c["kept"])
                        for e, poly in enumerate(ref):
                            big = max(abs(x) for x in poly)
                            g = [R.value(*got[((((jo * Dout + i) * (Rr + 1) + e) * 2 + p) * S2 + a * SR + b) * B + k])
                                 for i in range(Dout)]
                            assert all(x == 0 for x in g[len(poly):]), (jo, p, a, b, e)
                            if big == 0:
                                continue
                            worst = max(worst, max(abs(f2m(x) - y)
                                                   for x, y in zip(g, poly)) / big)
                            cnt += 1
    bits = -float(mp.log(worst, 2)) if worst else float("inf")
    print(f"{cnt} numerators (R={Rr}, B={B}, kept={c['kept']}): max error relative to the largest coefficient 2^-{bits:.1f}")
    assert bits > 32 * L - 40, bits


@pytest.mark.parametrize("L,tj12,tj43,order,Rmax,Bs,window", [(10, 4, 2, 7, 5, [2, 1, 3], 3),
                                                               (34, 8, 8, 6, 4, [2, 2], 8),
                                                               (10, 2, 2, 9, 3, [1, 2, 1, 2], 4)])
def test_orders_in_one_recursion(L, tj12, tj43, order, Rmax, Bs, window):
    """the right-hand sides of several lambda orders (R = Rmax, Rmax - 2, ...) in ONE recursion with the moment
    flush limited per column to its order's k <= R (derivatives.run_orders) against one recursion per order:
    every order's numerators bit for bit, at uneven right-hand-side counts per order"""
    import torch
    from blocksvibe_amd import derivatives as DV
    from blocksvibe_amd import mpf
    from blocksvibe_amd.derivatives import Derivs
    from blocksvibe_amd.plan import Plan
    from blocksvibe_amd.recursion import Recursion
    pp = ((0, 1), (1, 0))
    plan = Plan(tj12, tj43, 2 if tj12 == 8 else 6, order, pp, reuse=True)
    M, H = random_operator(plan, tj12, tj43, pp, order, sum(Bs), 11)
    h_all = h_tensor(H, plan, L)
    hs, b0 = [], 0
    for B in Bs:
        hs.append(mpf.MPF(h_all.m[..., b0:b0 + B].contiguous(), h_all.e[..., b0:b0 + B].contiguous()))
        b0 += B
    Rs = [max(Rmax - 2 * i, 0) for i in range(len(Bs))]
    rec = Recursion(plan, M, L)
    der = Derivs(plan, Rmax, L)
    print(f"orders R={Rs}, right-hand sides {Bs}, L={L}", flush=True)
    want = []
    for h, R in zip(hs, Rs):
        dn = der.for_order(R)
        dn.start(h.shape[-1], window=window)
        rec.run(h, hook=dn.accumulate)
        want.append(dn.finish(h))
    got = DV.run_orders(rec, der, hs, Rs, window=window)
    got2 = DV.run_orders(rec, der, hs, Rs, h_all=h_all, window=window)
    for w, g, g2, R in zip(want, got, got2, Rs):
        assert w.shape == g.shape == g2.shape, (R, w.shape, g.shape)
        assert torch.equal(w.m, g.m) and torch.equal(w.e, g.e), ("order", R)
        assert torch.equal(w.m, g2.m) and torch.equal(w.e, g2.e), ("order, resident h_all", R)
