"""The oracle and the product's host bookkeeping against golden vectors written from the reference's own modules
(tests/golden/make_fixtures.py: reference blocksvibe/structure.py, plan_ref.py, recursion_ref.py), exactly:

  oracle.structure_ref   every function over a grid of spins and orders (negative orders included, where pole
                         counts go below zero)
  oracle.plan_ref        slot offsets, M and Q keys in first-appearance order, per level the groups (source
                         parity, spin), their Cauchy columns (physical slots) and differences D' - D_i, the tasks'
                         slots, H rows and M / Q keys, and the physical slot map with and without reuse
  blocksvibe_amd.plan    the same quantities from the product's vectorised plan
  oracle.recursion_ref   every residue of the exact recursion on seeded operators (Fraction arithmetic)
  blocksvibe_amd.structure  the vectorised pole counts and positions, where the recursion reads them
"""
import json
import os
from fractions import Fraction

import numpy as np
import pytest

from oracle import plan_ref as OP
from oracle import recursion_ref as RR
from oracle import structure_ref as OS

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(GOLDEN, name)) as f:
        return json.load(f)["data"]


def F(s):
    return Fraction(s)


def test_structure_vectors():
    d = load("structure.json")
    for tj, tj12, tj43, order, counts in d["range"]:
        assert [OS.count(t, tj, tj12, tj43, order) for t in range(4)] == counts, (tj, tj12, tj43, order)
    for tj, tj12, tj43, order, poles, first_pos in d["poles"]:
        assert [list(p) for p in OS.poles(tj, tj12, tj43, order)] == [p[:2] for p in poles]
        assert [OS.position(tj, tj12, tj43, order, (t, 1)) for t in range(4)] == first_pos, (tj, tj12, tj43, order)
        for t, k, pos, sh, ptj, d2, par, qv, jex in poles:
            P = (t, k)
            assert OS.position(tj, tj12, tj43, order, P) == pos
            assert (OS.shift(P), OS.pole_two_j(P, tj), OS.delta_2(P, tj), OS.parity(P)) == (sh, ptj, d2, par)
            assert OS.q_value(P, tj) == F(qv)
            assert OS.j_excess(OS.pole_two_j(P, tj), 6) == jex
    for m, order, tj12, tjmax, vals in d["two_j_values"]:
        assert OS.two_j_values(m, order, tj12, tjmax) == vals
    for pp, a, b, vals in d["so3"]:
        assert list(OS.so3(pp, a, b)) == vals
    for dp, dd, c in d["cauchy"]:
        assert OS.cauchy_coeff(dp, dd) == F(c)


def test_product_structure_vectors():
    """the product's vectorised positions equal the reference's at every nonnegative order: the recursion asks for
    a position at order - m + shift(i) - j_excess(tj') >= 1 only (a prime pole i of spin tj' at level m - shift(i)
    has tj' <= tjmax + 2 (order - m + shift(i) - 1)), where no pole count is negative"""
    from blocksvibe_amd import structure as PS
    d = load("structure.json")
    for tj, tj12, tj43, order, poles, _ in d["poles"]:
        if order < 0:
            continue
        for t, k, pos, *_ in poles:
            assert int(PS.index_of_pole(tj, tj12, tj43, order, t, k)) == pos


def _ref_level_arrays(p, lev, reuse):
    """the reference plan_ref level arrays (gsrc, gj, col, cdiff, t_slot, t_hrow, t_L, t_R, t_q) from the
    oracle's PlanRef level (groups by first appearance, rows by first appearance of D')"""
    groups = sorted(lev["groups"].items(), key=lambda kv: kv[1]["gi"])
    nb = len(groups)
    Mr = max(len(g["rows"]) for _, g in groups)
    Kc = max(1, max(len(g["cols"]) for _, g in groups))
    z = p.nphys
    col = [[z] * Kc for _ in range(nb)]
    cdiff = [[[0] * Kc for _ in range(Mr)] for _ in range(nb)]
    for gi, ((src, tjp), g) in enumerate(groups):
        for c, (slot, di) in enumerate(g["cols"]):
            col[gi][c] = p.phys_slot(slot)
            for dd, r in g["rows"].items():
                cdiff[gi][r][c] = dd - di
    ts = lev["tasks"]
    return {"m": lev["m"], "nb": nb, "Mr": Mr, "Kc": Kc, "gsrc": [k[0] for k, _ in groups],
            "gj": [k[1] // 2 for k, _ in groups], "col": col, "cdiff": cdiff,
            "t_slot": [p.phys_slot(t["slot"]) for t in ts], "t_hrow": [t["group"] * Mr + t["row"] for t in ts],
            "t_L": [t["L"] for t in ts], "t_R": [t["R"] for t in ts], "t_q": [t["Q"] for t in ts]}


def _cases():
    return [(c, i) for i, c in enumerate(load("plan.json"))]


@pytest.mark.parametrize("case,i", _cases(), ids=lambda x: str(x) if isinstance(x, int) else "")
def test_oracle_plan_vectors(case, i):
    tj12, tj43, tjmax, order, pp, reuse = case["case"]
    pp = tuple(tuple(x) for x in pp)
    p = OP.PlanRef(tj12, tj43, tjmax, order, pp, reuse=reuse)
    assert p.nslots == case["nslots"] and p.nphys == case["nphys"]
    assert [[list(k[0]), k[1], k[2], k[3], k[4]] for k, _ in sorted(p.Mkeys.items(), key=lambda kv: kv[1])] == case["Mkeys"]
    assert [[list(k[0]), k[1], k[2]] for k, _ in sorted(p.Qkeys.items(), key=lambda kv: kv[1])] == case["Qkeys"]
    assert [[m, j, v] for (m, j), v in sorted(p.offset.items())] == case["offset"]
    assert sorted([t, v] for t, v in p.phys.items()) == case["phys"]
    assert len(p.levels) == len(case["levels"])
    dmax = 1
    for lev, want in zip(p.levels, case["levels"]):
        assert (lev is None) == (want is None)
        if lev is None:
            continue
        got = _ref_level_arrays(p, lev, reuse)
        for k, v in want.items():
            assert got[k] == v, (lev["m"], k)
        dmax = max([dmax] + [abs(x) for g in want["cdiff"] for r in g for x in r])
    assert dmax == case["dmax"]


@pytest.mark.parametrize("case,i", _cases(), ids=lambda x: str(x) if isinstance(x, int) else "")
def test_product_plan_vectors(case, i):
    """the product Plan: the same keys, slot counts and physical map (logical -> physical), and per level the
    same tasks (slots, M and Q keys) and Cauchy columns"""
    from blocksvibe_amd.plan import Plan
    tj12, tj43, tjmax, order, pp, reuse = case["case"]
    pp = tuple(tuple(x) for x in pp)
    p = Plan(tj12, tj43, tjmax, order, pp, reuse=reuse)
    assert [[list(k[0]), k[1], k[2], k[3], k[4]] for k, _ in sorted(p.Mkeys.items(), key=lambda kv: kv[1])] == case["Mkeys"]
    assert p.nslots == case["nslots"] and p.nphys == case["nphys"]
    phys = dict((t, v) for t, v in case["phys"])
    for t in range(2 * case["nslots"]):
        assert p.phys_slot(t) == phys.get(t, case["nphys"]), t
    live = [lv for lv in case["levels"] if lv is not None]
    assert len([lv for lv in p.levels if lv]) == len(live)
    for lev, want in zip([lv for lv in p.levels if lv], live):
        assert lev.m == want["m"]
        assert sorted(np.asarray(lev.t_slot).tolist()) == sorted(want["t_slot"])
        # every task's (slot, L key, R key) triple as in the reference (order within a level may differ)
        ref = sorted(zip(want["t_slot"], want["t_L"], want["t_R"]))
        qL = np.asarray(p.LQpairs)[np.asarray(lev.t_qL)] if len(p.LQpairs) else np.zeros((0, 2), np.int64)
        got = sorted(zip(np.asarray(lev.t_slot).tolist(), qL[:, 0].tolist(), np.asarray(lev.t_R).tolist()))
        assert got == ref, lev.m
        # the Cauchy columns of every group (unborn prime poles read the zero slot)
        z = case["nphys"]
        cols = sorted(c for g in want["col"] for c in g if c != z)
        gc = np.asarray(lev.col).reshape(len(lev.ncol), -1)
        gcols = sorted(int(c) for g in range(len(lev.ncol)) for c in gc[g, :int(lev.ncol[g])] if c != z)
        assert gcols == cols, lev.m


def _recursion_cases():
    return load("recursion.json")


@pytest.mark.parametrize("i", range(8))
def test_oracle_recursion_vectors(i):
    case = _recursion_cases()[i]
    tj12, tj43, tjmax, order, pp = case["case"]
    pp = tuple(tuple(x) for x in pp)
    M = {((k[0][0], k[0][1]), k[1], k[2], k[3], k[4]): [[F(x) for x in row] for row in k[5]] for k in case["M"]}
    h = {tuple(x[:5]): F(x[5]) for x in case["h_inf"]}
    got = RR.recursion(h, M, tj12, tj43, tjmax, order, pp)
    want = {tuple(x[:6]): F(x[6]) for x in case["res"]}
    assert got.keys() == want.keys()
    for k, v in want.items():
        assert got[k] == v, k
