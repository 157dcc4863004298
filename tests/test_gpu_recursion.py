"""Batched GPU radial recursion (blocksvibe_amd.recursion) against the oracle, on random operators and
initial data with several right-hand sides (reference tests/test_blocks_recursion.py:12-52):
  * bit for bit against oracle.recursion_ref.recursion_bits (the kernels' exact rounding semantics), with the
    MFMA pole step where it applies and again with the VALU pole step and two Cauchy rows per lane on every level;
  * against the reference's own recursion vectors (tests/golden/recursion.json) within 2^-(32L - 64);
  * within 2^-(32L - 64) of the exact rational recursion relative to each (pole, spin) block;
  * with slot recycling (reuse=True) every residue, captured right after its level, is identical.
Residues hold SL x SR structure entries (left and right structure counts; asymmetric cases 3 x 2, 2 x 1,
4 x 2).  Rows whose SL SR B columns fill whole waves to at least 3/4 take the row-uniform Cauchy kernel (cases
with 63 and 64 columns), the others the general one; with SL = SR = S and S B >= 32 the pole step runs on int8
MFMA (S 5, B 7).
"""
import random
from fractions import Fraction as F

import pytest
import torch

from oracle import mpf_ref as R
from oracle import recursion_ref as RR
from oracle import structure_ref as OS

pytestmark = pytest.mark.gpu

CASES = [dict(tj12=4, tj43=2, tjmax=6, order=7, L=10, B=3, seed=1, pp=((0, 1), (1, 0))),
         # 3 x 3 structures, B 7: rows of 63 columns on 64 lanes (one idle lane per wave), row-uniform Cauchy
         dict(tj12=4, tj43=4, tjmax=6, order=7, L=10, B=7, seed=6, pp=((0, 1), (1, 0))),
         # 5 x 5, B 3: 75 columns per row, waves inside one row and waves across two in the general kernel
         dict(tj12=8, tj43=8, tjmax=4, order=6, L=10, B=3, seed=2, pp=((0, 1), (1, 0))),
         dict(tj12=3, tj43=1, tjmax=5, order=6, L=6, B=2, seed=3, pp=((0, 1), (1, 0))),
         dict(tj12=2, tj43=2, tjmax=6, order=8, L=34, B=2, seed=4, pp=((0, 0), (1, 1))),
         # 5 x 5, B 7: 35 rows per constant row, the pole step on int8 MFMA
         dict(tj12=8, tj43=8, tjmax=4, order=6, L=10, B=7, seed=7, pp=((0, 1), (1, 0))),
         # 4 x 2, B 8: 64 columns per row, the row-uniform Cauchy kernel
         dict(tj12=6, tj43=2, tjmax=4, order=5, L=34, B=8, seed=5, pp=((0, 1), (1, 0))),
         # 5 x 5 at L = 34 (the heaviest operator group's structures and limbs), B 2: the VALU pole kernels and the
         # MFMA ones against the oracle directly
         dict(tj12=8, tj43=8, tjmax=2, order=5, L=34, B=2, seed=8, pp=((0, 1), (1, 0)))]


def random_operator(plan, tj12, tj43, pp, order, B, seed):
    rng = random.Random(seed)
    rnd = lambda: F(rng.uniform(-2, 2)) + F(rng.random()) * F(2) ** -150
    M = {}
    for (P, d, tjNN, ppp, tj) in plan.Mkeys:
        rows = len(OS.so3((ppp + OS.parity(P)) % 2, tjNN, tj))
        cols = len(OS.so3(ppp, tjNN, OS.pole_two_j(P, tj)))
        M[(P, d, tjNN, ppp, tj)] = [[rnd() for _ in range(cols)] for _ in range(rows)]
    H = []
    for k in range(B):
        h = {}
        for s in range(2):
            for m in range(order + 1):
                for j in range(plan.Jh):
                    tj = tj12 % 2 + 2 * j
                    for a in range(len(OS.so3(pp[s][0], tj12, tj))):
                        for b in range(len(OS.so3(pp[s][1], tj43, tj))):
                            h[(s, m, j, a, b)] = rnd()
        H.append(h)
    return M, H


def h_tensor(H, plan, L):
    """h_inf (2, order + 1, Jh, SL SR, B), structure entry (a, b) at a SR + b"""
    from blocksvibe_amd import mpf
    SL, SR, O1 = plan.SLm, plan.SRm, plan.order + 1
    B = len(H)
    pairs = [(0, R.EZ)] * (2 * O1 * plan.Jh * SL * SR * B)
    for k, h in enumerate(H):
        for (s, m, j, a, b), v in h.items():
            pairs[((((s * O1 + m) * plan.Jh + j) * SL + a) * SR + b) * B + k] = R.from_fraction(v, L)
    return mpf.from_pairs(pairs, L, (2, O1, plan.Jh, SL * SR, B))


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c['tj12']}-{c['tj43']}-{c['tjmax']}-{c['order']}-L{c['L']}")
def test_recursion_bits_and_exact(case):
    from blocksvibe_amd.plan import Plan
    from blocksvibe_amd.recursion import Recursion
    c = case
    L, B, pp = c["L"], c["B"], c["pp"]
    plan = Plan(c["tj12"], c["tj43"], c["tjmax"], c["order"], pp, reuse=False)
    plan_r = Plan(c["tj12"], c["tj43"], c["tjmax"], c["order"], pp, reuse=True)
    M, H = random_operator(plan, c["tj12"], c["tj43"], pp, c["order"], B, c["seed"])
    print(f"case {c}: {sum(l.T for l in plan.levels if l)} tasks, {plan.SLm} x {plan.SRm} structures", flush=True)
    h = h_tensor(H, plan, L)
    rec = Recursion(plan, M, L)
    rec.mma = True                       # the int8 MFMA pole step where SL = SR = S, S B >= 32, the VALU one elsewhere
    res = rec.run(h)
    got = res.to_host()
    # the same recursion with the VALU pole step everywhere and two Cauchy rows per lane on every level
    # (k_cauchy_pairs, which the default threshold only selects on levels that fill the chip)
    rec_v = Recursion(plan, M, L)
    rec_v.mma, rec_v.pair_min_waves = False, 0
    got_v = rec_v.run(h).to_host()
    # slot reuse: capture the slots written at each level right after it
    rec_r = Recursion(plan_r, M, L)
    cap = {}

    def hook(m, r):
        lev = plan_r.levels[m]
        idx = torch.from_numpy(lev.t_slot).cuda()
        cap[m] = (lev.t_slot.copy(), r.take(idx).to_host())

    rec_r.run(h, hook=hook)
    Sm, SR, S2 = plan.Smax, plan.SRm, plan.S2
    Mp = {k: [[R.from_fraction(x, L) for x in row] for row in v] for k, v in M.items()}
    worst = 0
    for k in range(B):
        hp = {key: R.from_fraction(v, L) for key, v in H[k].items()}
        bits = RR.recursion_bits(hp, Mp, c["tj12"], c["tj43"], c["tjmax"], c["order"], pp, L, Sm)
        exact = RR.recursion(H[k], M, c["tj12"], c["tj43"], c["tjmax"], c["order"], pp)
        blk = {}
        for key, v in exact.items():
            blk[key[:4]] = max(blk.get(key[:4], 0), abs(v))
        for (p, m, j, q, a, b), v in bits.items():
            if a >= plan.SLm or b >= SR:                 # the oracle pads to Smax x Smax: zeros there
                assert v == (0, R.EZ), ("padding", (p, m, j, q, a, b), v)
                continue
            slot = rec.slot(p, m, j, q)
            g = got[(slot * S2 + a * SR + b) * B + k]
            assert g == v, ("bits", k, (p, m, j, q, a, b), g, v)
            assert got_v[(slot * S2 + a * SR + b) * B + k] == v, ("bits, VALU pole step, paired Cauchy rows", k)
            ex = exact.get((p, m, j, q, a, b))
            if ex is not None and blk[(p, m, j, q)]:
                worst = max(worst, abs(R.value(*g) - ex) / blk[(p, m, j, q)])
    bits_err = (worst.denominator.bit_length() - worst.numerator.bit_length()) if worst else float("inf")
    print(f"max error relative to the (pole, spin) block: 2^-{bits_err:.1f} (bound 2^-{32 * L - 64})")
    assert bits_err > 32 * L - 64
    # reuse: every residue captured after its level equals the non-reuse result
    for m, (slots, vals) in cap.items():
        lev_n = plan.levels[m]
        # the same tasks in the same order in both plans
        for ti in range(len(slots)):
            for e in range(S2 * B):
                assert vals[ti * S2 * B + e] == got[int(lev_n.t_slot[ti]) * S2 * B + e], ("reuse", m, ti, e)


@pytest.mark.parametrize("case", [CASES[1], CASES[2], CASES[6], CASES[3]],
                         ids=lambda c: f"{c['tj12']}-{c['tj43']}-B{c['B']}")
def test_cauchy_kernels_agree(case):
    """the group-uniform Cauchy kernels (wave tables; one row per lane and two) against the kernels used
    without them (rows padded to whole waves at 63 and 64 columns, every index per lane at 75 and 4
    columns), bit for bit"""
    from blocksvibe_amd.plan import Plan
    from blocksvibe_amd.recursion import Recursion
    c = case
    L, B, pp = c["L"], c["B"], c["pp"]
    plan = Plan(c["tj12"], c["tj43"], c["tjmax"], c["order"], pp, reuse=True)
    M, H = random_operator(plan, c["tj12"], c["tj43"], pp, c["order"], B, c["seed"])
    h = h_tensor(H, plan, L)
    print(f"cauchy kernels: {plan.S2 * B} columns per row", flush=True)
    outs = []
    for waves, pmin in ((True, 0), (True, 1 << 62), (False, 0)):
        rec = Recursion(plan, M, L)
        rec.group_waves, rec.pair_min_waves = waves, pmin
        outs.append(rec.run(h))
    a = outs[0]
    for b in outs[1:]:
        assert torch.equal(a.e, b.e) and torch.equal(a.m, b.m)


@pytest.mark.parametrize("i,L", [(0, 10), (1, 10), (2, 10), (3, 10), (4, 34), (5, 10), (6, 10), (7, 18)])
def test_recursion_golden_vectors(i, L):
    """the GPU recursion on the inputs of the reference's recursion vectors (written by reference
    blocksvibe/recursion_ref.py, tests/golden/make_fixtures.py): every residue within 2^-(32L - 64) of the
    vector, relative to its (pole, spin) block, as reference tests/test_blocks_recursion.py:45-52 bounds it"""
    import json
    import os
    from blocksvibe_amd.plan import Plan
    from blocksvibe_amd.recursion import Recursion
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "recursion.json")) as f:
        case = json.load(f)["data"][i]
    tj12, tj43, tjmax, order, pp = case["case"]
    pp = tuple(tuple(x) for x in pp)
    M = {((k[0][0], k[0][1]), k[1], k[2], k[3], k[4]): [[F(x) for x in row] for row in k[5]] for k in case["M"]}
    hd = {tuple(x[:5]): F(x[5]) for x in case["h_inf"]}
    want = {tuple(x[:6]): F(x[6]) for x in case["res"]}
    plan = Plan(tj12, tj43, tjmax, order, pp, reuse=False)
    print(f"golden recursion {case['case']} L={L}: {len(want)} residues", flush=True)
    rec = Recursion(plan, M, L)
    got = rec.run(h_tensor([hd], plan, L)).to_host()
    S2, SR = plan.S2, plan.SRm
    blk = {}
    for key, v in want.items():
        blk[key[:4]] = max(blk.get(key[:4], 0), abs(v))
    worst = F(0)
    for (p, m, j, q, a, b), v in want.items():
        g = got[rec.slot(p, m, j, q) * S2 + a * SR + b]
        if blk[(p, m, j, q)]:
            worst = max(worst, abs(R.value(*g) - v) / blk[(p, m, j, q)])
    bits = (worst.denominator.bit_length() - worst.numerator.bit_length()) if worst else float("inf")
    print(f"max error relative to the (pole, spin) block: 2^-{bits} (bound 2^-{32 * L - 64})")
    assert bits > 32 * L - 64


@pytest.mark.parametrize("L,tj12,tj43", [(10, 4, 2), (34, 8, 8)])
def test_operators_of_one_plan_in_one_recursion(L, tj12, tj43):
    """right-hand sides of two operators of one plan (other M matrices) in one recursion (recursion.Operators,
    the pole step's per-column table offsets): every column bit-identical to its own operator's recursion run
    alone, and the operators' results differ (so the columns really read their own tables)"""
    from blocksvibe_amd import mpf
    from blocksvibe_amd.plan import Plan
    from blocksvibe_amd.recursion import Operators, Recursion
    plan = Plan(tj12, tj43, 4, 6, ((0, 1), (1, 0)))
    ra = Recursion(plan, None, L, seed=21)
    rb = ra.with_matrices(None, seed=22)
    h = mpf.random_device((2, plan.order + 1, plan.Jh, plan.S2, 4), L, seed=23)
    cols = [0, 1, 1, 0]
    mixed = ra.run(h, operators=Operators([ra, rb], cols))
    mixed = (mixed.m.cpu(), mixed.e.cpu())
    alone = []
    for r in (ra, rb):
        x = r.run(h)
        alone.append((x.m.cpu(), x.e.cpu()))
    for k, i in enumerate(cols):
        assert torch.equal(mixed[0][..., k], alone[i][0][..., k]) and torch.equal(mixed[1][..., k], alone[i][1][..., k]), k
    assert not torch.equal(alone[0][0], alone[1][0])
