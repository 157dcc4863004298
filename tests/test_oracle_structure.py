"""Pins of the pole bookkeeping and plan oracles, and the product Plan against them:
  * published counts: 10,134 logical / 4,062 physical slots at (2j12, 2j43) = (4, 4), 2j_max 20, order 16,
    and 699,214 / 294,860 at (8, 8), 2j_max 120, order 60 (reference docs/BLOCKS_PLAN.md:209-210);
  * the plan invariants: every slot read as a Cauchy column was written at an earlier level, and no
    physical slot is reused while a residue in it is still to be read;
  * the vectorised product Plan equals the per-task oracle PlanRef (reference tests/test_blocks_plan.py:11-13
    shapes, with and without reuse)."""
import numpy as np
import pytest

from oracle import plan_ref as PR
from oracle import structure_ref as OS

PP = ((0, 0), (1, 1))
CASES = [(8, 8, 30, 30, ((0, 0), (1, 1))), (4, 2, 20, 24, ((0, 1), (1, 0))), (3, 1, 9, 12, ((0, 1), (1, 1))),
         (5, 5, 11, 10, ((1, 0), (0, 1))), (4, 4, 20, 16, ((0, 0), (1, 1))), (0, 0, 6, 8, ((0, 0), (1, 1))),
         (8, 4, 4, 7, ((0, 1), (1, 0)))]


def test_known_counts_oracle():
    p = PR.PlanRef(4, 4, 20, 16, PP, reuse=True)
    assert 2 * p.nslots == 10134 and p.nphys == 4062


def test_known_counts_product():
    from blocksvibe_amd.plan import Plan
    p = Plan(4, 4, 20, 16, PP, reuse=True)
    assert 2 * p.nslots == 10134 and p.nphys == 4062
    p = Plan(8, 8, 120, 60, PP, reuse=True)
    assert 2 * p.nslots == 699214 and p.nphys == 294860


def test_pole_closed_forms():
    # positions of all poles of a spin are 0..n-1; shifts and spins follow the type rules
    for tj in range(0, 14):
        for order in range(0, 9):
            pl = OS.poles(tj, 4, 2, order)
            assert [OS.position(tj, 4, 2, order, P) for P in pl] == list(range(len(pl)))
            for t, k in pl:
                assert OS.shift((t, k)) == (k if t < 2 else (2 * k if t == 2 else 2 * k - 1))
    assert OS.q_value((0, 1), 4) == -4 and OS.q_value((2, 1), 4) == OS.Fraction(-48, 7)


def test_plan_invariants_oracle():
    p = PR.PlanRef(4, 2, 10, 10, ((0, 1), (1, 0)), reuse=True)
    occupants = {}
    for t, ph in p.phys.items():
        occupants.setdefault(ph, []).append(t)
    for ph, ts in occupants.items():
        ts.sort(key=lambda t: p.born[t])
        for a, b in zip(ts, ts[1:]):
            assert p.born[b] >= max(p.last.get(a, -1), p.born[a] + 1)
    for lev in p.levels:
        if lev is None:
            continue
        for g in lev["groups"].values():
            for c, _ in g["cols"]:
                if c in p.born:
                    assert p.born[c] < lev["m"]


@pytest.mark.parametrize("case", CASES)
@pytest.mark.parametrize("reuse", [False, True])
def test_product_plan_equals_oracle(case, reuse):
    from blocksvibe_amd.plan import Plan
    tj12, tj43, tjmax, order, pp = case
    ref = PR.PlanRef(tj12, tj43, tjmax, order, pp, reuse=reuse)
    new = Plan(tj12, tj43, tjmax, order, pp, reuse=reuse)
    assert ref.nslots == new.nslots and ref.nphys == new.nphys
    assert ref.Mkeys == new.Mkeys and ref.Qkeys == new.Qkeys
    for t in range(2 * ref.nslots):
        assert ref.phys_slot(t) == new.phys_slot(t)
    for a, b in zip(ref.levels, new.levels):
        assert (a is None) == (b is None)
        if a is None:
            continue
        groups = sorted(a["groups"].items(), key=lambda kv: kv[1]["gi"])
        assert [(s, tjp) for (s, tjp), _ in groups] == list(zip(b.gsrc.tolist(), b.gtjp.tolist()))
        for gi, (_, g) in enumerate(groups):
            nc = b.ncol[gi]
            assert nc == len(g["cols"])
            assert [ref.phys_slot(c) for c, _ in g["cols"]] == b.col[gi, :nc].tolist()
            assert [d for _, d in g["cols"]] == b.cold[gi, :nc].tolist()
        rows = {}
        for gi, (_, g) in enumerate(groups):
            for d, r in g["rows"].items():
                rows[(gi, r)] = d
        assert len(a["tasks"]) == b.T
        for ti, t in enumerate(a["tasks"]):
            h = b.t_hrow[ti]
            assert (b.rg[h], b.rd[h]) == (t["group"], rows[(t["group"], t["row"])])
            assert ref.phys_slot(t["slot"]) == b.t_slot[ti]
            assert (t["L"], t["R"], t["Q"]) == (b.t_L[ti], b.t_R[ti], b.t_q[ti])
