"""The radial recursion on the host through the C restatement (oracle/c/mpf_c.c: oc_cauchy_level, oc_pole_level),
level by level over a plan's index tables -- the bit-level semantics of oracle.recursion_ref.recursion_bits at
sizes the Python oracle cannot reach (a whole operator group at order 60, 1024 bits).  TEST INFRASTRUCTURE (see
oracle/__init__.py): used by the GPU tests as the checker and by bench.py's CPU baseline as the timed CPU path.

The index tables are those of the product's Plan (blocksvibe_amd.plan), itself pinned against the reference's
plan_ref vectors (tests/test_golden.py); the constant tables are built here from the M matrices with mpf_ref
(each row of R aligned; q L entries rounded once, then aligned per row), not taken from the GPU.  Pinned by
tests/test_oracle_c.py against recursion_bits on small operators.
"""
import ctypes
import time

import numpy as np

from . import c_ref as C
from . import mpf_ref as R
from . import structure_ref as OS

_P, _I64, _I = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int


def _lib():
    lb = C.lib()
    if not getattr(lb, "_chain_ready", False):
        lb.oc_cauchy_level.argtypes = [_I, _I64, _P, _P, _P, _P, _P, _P, _P, _I64, _P, _P, _P, _I, _P, _P, _I]
        lb.oc_cauchy_level.restype = _I
        lb.oc_pole_level.argtypes = [_I, _I, _I, _I, _I64, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _I]
        lb.oc_pole_level.restype = _I
        lb._chain_ready = True
    return lb


def planes(pairs, L):
    """list of (m, e) -> entry-major limbs (n, L) uint32 and exponents (n,) int64"""
    m = np.array([R.to_limbs(x, L) for x, _ in pairs], dtype=np.uint32).reshape(-1, L)
    e = np.array([ex if x else R.EZ for x, ex in pairs], dtype=np.int64)
    return np.ascontiguousarray(m), e


def tables(plan, Mtab, L):
    """aligned constant tables: R rows (nM, S, S) and q L rows (nLQ, S, S), each with its row exponents.
    Mtab: (m, e) pairs of the (nM, S, S) M matrices (zero padded to S x S)"""
    S = plan.Smax
    nM = max(len(plan.Mkeys), 1)
    Rp, Re = [], []
    for i in range(nM):
        for r in range(S):
            am, E = R.align(Mtab[(i * S + r) * S:(i * S + r + 1) * S])
            Rp += [(x, E) for x in am]
            Re.append(E)
    q = [(0, R.EZ)] * max(len(plan.Qkeys), 1)
    for ((t, k), tj, sign), i in plan.Qkeys.items():
        q[i] = R.from_fraction(OS.q_value((t, k), tj) * sign, L)
    Qp, Qe = [], []
    lq = plan.LQpairs if len(plan.LQpairs) else np.zeros((1, 2), np.int64)
    for lk, qk in lq:
        for r in range(S):
            row = [R.mul(Mtab[(int(lk) * S + r) * S + c], q[int(qk)], L) for c in range(S)]
            am, E = R.align(row)
            Qp += [(x, E) for x in am]
            Qe.append(E)
    Rm, _ = planes([(x, 0) for x, _ in Rp], L)
    Qm, _ = planes([(x, 0) for x, _ in Qp], L)
    return Rm, np.array(Re, dtype=np.int64), Qm, np.array(Qe, dtype=np.int64)


def recursion(plan, Mtab, hm, he, L, B, threads=1, timing=None):
    """Run the recursion of `plan` for B right-hand sides: Mtab as for tables(); h_inf as entry-major planes
    hm (rows * N, L) uint32 / he (rows * N,) in (2, order + 1, Jh, SL SR, B) order.  Returns the slot buffer
    (nphys + 1, SL SR, B) as planes (m (n, L) uint32, e (n,) int64).  timing: dict that receives the seconds of
    the level loop."""
    lb = _lib()
    N = plan.S2 * B
    Rm, Re, Qm, Qe = tables(plan, Mtab, L)
    nslot = plan.nphys + 1
    rm = np.zeros((nslot * N, L), dtype=np.uint32)
    re = np.full(nslot * N, R.EZ, dtype=np.int64)
    O1, Jh = plan.order + 1, plan.Jh
    maxH = max((lv.nH for lv in plan.levels if lv), default=1)
    Hm = np.zeros((maxH * N, L), dtype=np.uint32)
    He = np.full(maxH * N, R.EZ, dtype=np.int64)
    hm = np.ascontiguousarray(hm, dtype=np.uint32)
    he = np.ascontiguousarray(he, dtype=np.int64)
    P = lambda a: a.ctypes.data
    t0 = time.time()
    for lev in plan.levels:
        if lev is None:
            continue
        col = np.ascontiguousarray(lev.col.reshape(-1), dtype=np.int64)
        cold = np.ascontiguousarray(lev.cold.reshape(-1), dtype=np.int32)
        ncol = np.ascontiguousarray(lev.ncol, dtype=np.int32)
        rg = np.ascontiguousarray(lev.rg, dtype=np.int32)
        rd = np.ascontiguousarray(lev.rd, dtype=np.int32)
        ghrow = np.ascontiguousarray((lev.gsrc * O1 + lev.m) * Jh + lev.gtjp // 2, dtype=np.int64)
        st = lb.oc_cauchy_level(L, N, P(rm), P(re), P(hm), P(he), P(ghrow), P(rg), P(rd), lev.nH, P(col), P(cold),
                                P(ncol), lev.Kc, P(Hm), P(He), threads)
        assert st == 0, st
        thrw = np.ascontiguousarray(lev.t_hrow, dtype=np.int64)
        tR = np.ascontiguousarray(lev.t_R, dtype=np.int32)
        tQL = np.ascontiguousarray(lev.t_qL, dtype=np.int32)
        tslot = np.ascontiguousarray(lev.t_slot, dtype=np.int64)
        st = lb.oc_pole_level(L, plan.Smax, plan.SLm, plan.SRm, B, lev.T, P(Hm), P(He), P(thrw), P(tR), P(tQL),
                              P(tslot), P(Rm), P(Re), P(Qm), P(Qe), P(rm), P(re), threads)
        assert st == 0, st
    if timing is not None:
        timing["seconds"] = time.time() - t0
    return rm, re
