"""Batched radial recursion of blocks_3d on the GPU (reference blocksvibe/recursion_gpu.py:379-514).

For one operator (a Plan) the level loop runs, per level m, two libb3dhip kernels' worth of work:

  Cauchy step  H(src, m, j', D') = h_inf[src][m][j'] + sum_i 2 / (D' - D_i) res[src][m - shift_i][j'][i]
  pole step    res[p][m][j][P]   = Q(P, j) L_P H R_P^T        (SL x SL left, SR x SR right structure matrices)

over all right-hand sides (four-point structures x lambda orders) at once: the batch B is the innermost
dimension of every array.  A residue holds SL x SR structure entries (the operator's left and right
structure counts), not a padded Smax x Smax square.  Constants (M matrices, Q values, their pre-aligned
forms) are built once per operator and reused by every run().
"""
import os
import threading
import time
from fractions import Fraction

import numpy as np
import torch

from . import _lib, mpf
from . import structure as S

_WAVES_LOCK = threading.Lock()


def number_pair(x, L):
    """a number of any supported kind (decimal string, int, Fraction, mpmath mpf, (m, e) pair) ->
    normalised (m, e), rounded half away from zero from its exact value"""
    if isinstance(x, tuple):
        return mpf.norm_int(int(x[0]), int(x[1]), L)
    if isinstance(x, str):
        x = Fraction(x)
    if isinstance(x, int):
        return mpf.norm_int(x, 0, L)
    if isinstance(x, Fraction):
        return mpf.norm_fraction(x.numerator, x.denominator, L)
    return mpf.from_mpmath(x, L)


def _split28(m, L):
    """row-aligned mantissas (L limb planes) -> (28-bit magnitude limb planes, sign masks), on the GPU"""
    n = m[0].numel()
    nl = _lib.lib().b3d_limbs28(L)
    o28 = torch.empty((nl,) + tuple(m.shape[1:]), dtype=torch.int32, device=m.device)
    sg = torch.empty(tuple(m.shape[1:]), dtype=torch.int32, device=m.device)
    _lib.call("b3d_mpf_to28", L, m.data_ptr(), n, n, o28.data_ptr(), n, sg.data_ptr(), mpf.stream_ptr())
    return o28, sg


def _pole_tables(m, L, S):
    """row-aligned mantissas (L limb planes of (rows, S, S)) -> (balanced base-256 digits, column
    corrections) for the pole step's int8 MFMA products, on the GPU"""
    n = m[0].numel()
    ds = _lib.lib().b3d_pole_digits_bytes(L)
    d8 = torch.empty((n, ds), dtype=torch.int8, device=m.device)
    corr = torch.empty((n // S, 8 * L), dtype=torch.int32, device=m.device)
    _lib.call("b3d_pole_tables", L, S, m.data_ptr(), n, n // S, d8.data_ptr(), corr.data_ptr(), mpf.stream_ptr())
    return d8, corr


class Recursion:
    """Device constants and level tables of one Plan at L limbs.
    M: dict {Mkey: matrix (rows of numbers)} for every key of plan.Mkeys (absent keys: zero matrices), an
    MTable-like object with .pairs(key) -> rows of (m, e), or None for random O(1) matrices (synthetic
    operators; the cost does not depend on the values)."""

    def __init__(self, plan, M, L, seed=0):
        if not _lib.lib().b3d_supported_limbs(L):
            raise _lib.B3DError(f"no kernels compiled for L={L}")
        if plan.dmax >= 1 << 15:
            raise _lib.B3DError(f"|D' - D_i| up to {plan.dmax}: the Cauchy kernel needs < 2^15")
        self.plan, self.L = plan, L
        Sm = plan.Smax
        nM = len(plan.Mkeys)
        if M is None:
            self.Mtab = mpf.random((max(nM, 1), Sm, Sm), L, spread=1, generator=np.random.default_rng(seed))
        else:
            vals = [(0, mpf.EZ)] * (max(nM, 1) * Sm * Sm)
            for key, i in plan.Mkeys.items():
                mat = M.get(key) if isinstance(M, dict) else M.pairs(key)
                for a, row in enumerate(mat or []):
                    for b, x in enumerate(row):
                        vals[(i * Sm + a) * Sm + b] = number_pair(x, L)
            self.Mtab = mpf.from_pairs(vals, L, (max(nM, 1), Sm, Sm))
        qv = [Fraction(0)] * max(len(plan.Qkeys), 1)
        for ((t, k), tj, sign), i in plan.Qkeys.items():
            qv[i] = S.q_value(t, k, tj) * sign
        self.Qtab = mpf.from_fractions(qv, L, (len(qv),))
        self._constants()
        self._upload_levels()

    def with_matrices(self, M, seed=0):
        """a Recursion of the same plan and limb count for other M matrices (another Delta_12, Delta_43 or
        external spins), sharing the uploaded level tables and Q values: only the M-dependent tables
        (R rows aligned, q L products aligned) are rebuilt"""
        r = object.__new__(Recursion)
        r.__dict__.update(self.__dict__)
        Sm, L, nM = self.plan.Smax, self.L, len(self.plan.Mkeys)
        if M is None:
            r.Mtab = mpf.random((max(nM, 1), Sm, Sm), L, spread=1, generator=np.random.default_rng(seed))
        else:
            vals = [(0, mpf.EZ)] * (max(nM, 1) * Sm * Sm)
            for key, i in self.plan.Mkeys.items():
                mat = M.get(key) if isinstance(M, dict) else M.pairs(key)
                for a, row in enumerate(mat or []):
                    for b, x in enumerate(row):
                        vals[(i * Sm + a) * Sm + b] = number_pair(x, L)
            r.Mtab = mpf.from_pairs(vals, L, (max(nM, 1), Sm, Sm))
        r._constants()
        return r

    # pole-step products on int8 MFMA where at least 32 rows share a constant row (bit-identical to the VALU
    # path); BLOCKSVIBE_POLE_MMA=1 / 0 selects the default
    mma = os.environ.get("BLOCKSVIBE_POLE_MMA", "0") == "1"
    # the Cauchy step's group-uniform kernel (per-level wave tables); False: the library's kernels without them
    group_waves = True
    # two rows per lane (k_cauchy_pairs) on levels with at least this many waves of pairs (one wave for each
    # of the 1024 SIMDs at the kernel's three waves per SIMD), one row per lane below it, where the level
    # is too small to fill the chip and the one-row kernel's five waves per SIMD win; measured on the nmax = 6
    # 1024-bit workload: 1536 .. 6144 give the same time, 0 (pairs everywhere) and never are 5% slower
    pair_min_waves = 3072

    def _constants(self):
        plan, Sm, L = self.plan, self.plan.Smax, self.L
        self.Ralign_m, self.Ralign_e = mpf.align_rows(self.Mtab)            # (L, nM, S, S), (nM, S)
        lq = torch.from_numpy(plan.LQpairs if len(plan.LQpairs) else np.zeros((1, 2), np.int64)).to(mpf.dev())
        Lm = self.Mtab.take(lq[:, 0])                                       # (nLQ, S, S)
        q = self.Qtab.take(lq[:, 1])                                        # (nLQ,)
        qx = mpf.MPF(q.m[:, :, None, None].expand(L, len(lq), Sm, Sm).contiguous(),
                     q.e[:, None, None].expand(len(lq), Sm, Sm).contiguous())
        self.QL = mpf.mul(Lm, qx)
        self.QLalign_m, self.QLalign_e = mpf.align_rows(self.QL)
        # the pole step reads both constant operands as 28-bit magnitude limbs + sign masks
        self.R28, self.Rsg = _split28(self.Ralign_m, L)
        self.QL28, self.QLsg = _split28(self.QLalign_m, L)
        self.Rd8, self.Rcorr = _pole_tables(self.Ralign_m, L, Sm)
        self.QLd8, self.QLcorr = _pole_tables(self.QLalign_m, L, Sm)

    def _upload_levels(self):
        plan, d = self.plan, mpf.dev()
        T = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a).astype(dt)).to(d)
        O1, Jh = plan.order + 1, plan.Jh
        self.levels = []
        for lev in plan.levels:
            if lev is None:
                self.levels.append(None)
                continue
            self.levels.append({
                "m": lev.m, "nH": lev.nH, "T": lev.T, "Kc": lev.Kc,
                "col": T(lev.col.reshape(-1), np.int64), "cold": T(lev.cold.reshape(-1), np.int32),
                "ncol": T(lev.ncol, np.int32), "rg": T(lev.rg, np.int32), "rd": T(lev.rd, np.int32),
                "ghrow": T((lev.gsrc * O1 + lev.m) * Jh + lev.gtjp // 2, np.int64),
                "t_hrow": T(lev.t_hrow, np.int64), "t_R": T(lev.t_R, np.int32), "t_qL": T(lev.t_qL, np.int32),
                "t_slot": T(lev.t_slot, np.int64)})
        # algorithmic bytes of the level's Cauchy launch per column (structure entry x right-hand side):
        # every gathered residue read once, one h_inf entry per (source, spin) group, every H entry written once
        nb = 4 * self.L + 8
        for lev, lv in zip(plan.levels, self.levels):
            if lv is not None:
                g = np.unique(lev.rg)
                lv["cauchy_bytes"] = nb * (int(lev.ncol[g].sum()) + len(g) + lev.nH)
        live = [lv for lv in self.levels if lv is not None]
        self.maxH = max((lv["nH"] for lv in live), default=1)
        self.maxT = max((lv["T"] for lv in live), default=1)
        self._waves = {}

    def wave_tables(self, N):
        """per level, the Cauchy step's wave table for N columns per row (cached per N and pair_min_waves):
        each group's rows (contiguous in the level's row order) taken one or two at a time (pairs 2p, 2p + 1;
        an odd count ends with one row), laid out column-major over consecutive lanes, 64 lanes a wave,
        padded at the end of each group -> (first row of the group, rows of the group, first unit, waves,
        rows per lane): two when the level has at least pair_min_waves waves of pairs"""
        key = (N, self.pair_min_waves)
        tabs = self._waves.get(key)
        if tabs is not None:
            return tabs
        with _WAVES_LOCK:         # several streams may ask at once: built once, its stream synchronised
            if key not in self._waves:
                tabs = self._wave_tables(N)
                torch.cuda.current_stream().synchronize()
                self._waves[key] = tabs
            return self._waves[key]

    def _wave_tables(self, N):
        d, tabs = mpf.dev(), []
        for lev in self.plan.levels:
            if lev is None:
                tabs.append(None)
                continue
            rg = lev.rg
            starts = np.flatnonzero(np.r_[True, rg[1:] != rg[:-1]])
            assert len(starts) == len(np.unique(rg)), "a level's rows must be grouped"
            cnts = np.diff(np.r_[starts, len(rg)])
            rpl = 2 if int((((cnts + 1) // 2 * N + 63) // 64).sum()) >= self.pair_min_waves else 1
            nwg = ((cnts + rpl - 1) // rpl * N + 63) // 64
            wrow = np.repeat(starts, nwg).astype(np.int32)
            wcnt = np.repeat(cnts, nwg).astype(np.int32)
            woff = ((np.arange(int(nwg.sum())) - np.repeat(np.cumsum(nwg) - nwg, nwg)) * 64).astype(np.int32)
            assert cnts.max() * N < 1 << 31
            tabs.append((torch.from_numpy(wrow).to(d), torch.from_numpy(wcnt).to(d), torch.from_numpy(woff).to(d),
                         len(wrow), rpl))
        return tabs

    def run(self, h_inf, hook=None, prof=None, res=None, timer=None, operators=None):
        """h_inf: MPF (2, order + 1, Jh, SL SR, B) (structure entry (a, b) at a SR + b).  Returns the residue
        slot buffer MPF (nphys + 1, SL SR, B) (physical slots; the last one is the zero slot).  hook(m, res)
        runs after each level (e.g. the derivative moments).  timer: a list that receives (start event, end
        event, algorithmic bytes) of every Cauchy launch, events recorded on the launch stream.  operators: an
        Operators of this plan and L whose right-hand sides are the columns of h_inf (each column with its own
        operator's M matrices); None: every column with this Recursion's."""
        plan, L, Sm = self.plan, self.L, self.plan.Smax
        SL, SR, S2 = plan.SLm, plan.SRm, plan.S2
        assert h_inf.shape[:4] == (2, plan.order + 1, plan.Jh, S2), (h_inf.shape, plan.order, plan.Jh, S2)
        B = h_inf.shape[-1]
        N = S2 * B
        h = h_inf.contiguous()
        # With slot recycling every physical slot is written in full (by the pole step of its level) before
        # anything reads it, so only the zero slot needs clearing; without it, logical slots that are never
        # written keep their place in the buffer and are cleared with the rest.  H and Y are written in full
        # for the rows a level reads.
        if res is None:
            res = (mpf.empty if plan.reuse else mpf.zeros)((plan.nphys + 1, S2, B), L)
        elif not plan.reuse:
            res.m.zero_()
            res.e.fill_(mpf.EZ)
        if plan.reuse:
            res.m[:, plan.nphys].zero_()
            res.e[plan.nphys].fill_(mpf.EZ)
        H = mpf.empty((self.maxH, S2, B), L)
        Y = mpf.empty((self.maxT, SR, SL, B), L)
        st = mpf.stream_ptr()
        P = lambda t: t.data_ptr()
        waves = self.wave_tables(N) if self.group_waves else [(None, None, None, 0, 1)] * len(self.levels)
        if operators is not None:
            if operators.kR.numel() != B or operators.plan is not plan or operators.L != L:
                raise _lib.B3DError("operators: one operator per right-hand side, of this plan and limb count")
            C, kR, kQ, mma = operators, P(operators.kR), P(operators.kQ), False
        else:
            C, kR, kQ, mma = self, 0, 0, self.mma
        for lv, wt in zip(self.levels, waves):
            if lv is None:
                continue
            t0 = time.time() if prof is not None else None
            if timer is not None:
                ev0 = torch.cuda.Event(enable_timing=True)
                ev0.record()
            _lib.call("b3d_cauchy_level", L, P(res.m), P(res.e), res.numel(), N, P(h.m), P(h.e), h.numel(),
                      P(lv["ghrow"]), P(lv["rg"]), P(lv["rd"]), lv["nH"], P(lv["col"]), P(lv["cold"]),
                      P(lv["ncol"]), lv["Kc"], P(H.m), P(H.e), H.numel(), *(P(x) if x is not None else 0 for x in wt[:3]),
                      wt[3], wt[4], st)
            if timer is not None:
                ev1 = torch.cuda.Event(enable_timing=True)
                ev1.record()
                timer.append((ev0, ev1, lv["cauchy_bytes"] * N))
            if prof is not None:
                torch.cuda.synchronize()
                t1 = time.time()
                prof["cauchy"] = prof.get("cauchy", 0.0) + t1 - t0
            _lib.call("b3d_pole_level", L, SL, SR, Sm, B, lv["T"], P(H.m), P(H.e), H.numel(), P(lv["t_hrow"]),
                      P(lv["t_R"]), P(lv["t_qL"]), P(lv["t_slot"]), P(C.R28), P(C.Rsg), P(C.Ralign_e),
                      C.R28[0].numel(), P(C.QL28), P(C.QLsg), P(C.QLalign_e), C.QL28[0].numel(),
                      P(self.Rd8) if mma else 0, P(self.Rcorr) if mma else 0,
                      P(self.QLd8) if mma else 0, P(self.QLcorr) if mma else 0, kR, kQ,
                      P(Y.m), P(Y.e), Y.numel(), P(res.m), P(res.e), res.numel(), st)
            if prof is not None:
                torch.cuda.synchronize()
                t2 = time.time()
                prof["pole"] = prof.get("pole", 0.0) + t2 - t1
            if hook is not None:
                hook(lv["m"], res)
                if prof is not None:
                    torch.cuda.synchronize()
                    prof["moments"] = prof.get("moments", 0.0) + time.time() - t2
        return res

    def slot(self, p, m, jidx, pidx):
        """physical slot of residue (p, m, jidx, pidx) (meaningful after run() only without slot reuse)"""
        return self.plan.phys_slot(p * self.plan.nslots + self.plan.offset(m, jidx) + pidx)


class Operators:
    """Right-hand sides of several operators of one plan and limb count in one recursion: the pole step's
    constant tables (R rows, q L rows: the only part of the recursion that depends on the M matrices) of every
    operator stacked into one, operator by operator, and per right-hand side the row offsets of its operator
    (b3d_pole_level's kR, kQ).  recs: Recursion objects of one plan and L (Recursion.with_matrices);
    col_op: per right-hand side (column of h_inf) the index of its operator in recs.  The Cauchy step, the
    moments and everything after the recursion do not depend on M, so the columns of a batch may mix
    operators freely and each keeps the bits it has alone."""

    def __init__(self, recs, col_op):
        r0 = recs[0]
        if any(r.plan is not r0.plan or r.L != r0.L for r in recs):
            raise _lib.B3DError("Operators: every operator must share one plan and limb count")
        if not col_op or min(col_op) < 0 or max(col_op) >= len(recs):
            raise _lib.B3DError("Operators: col_op indexes recs")
        self.plan, self.L = r0.plan, r0.L
        self.R28 = torch.cat([r.R28 for r in recs], 1)
        self.Rsg = torch.cat([r.Rsg for r in recs], 0)
        self.Ralign_e = torch.cat([r.Ralign_e for r in recs], 0)
        self.QL28 = torch.cat([r.QL28 for r in recs], 1)
        self.QLsg = torch.cat([r.QLsg for r in recs], 0)
        self.QLalign_e = torch.cat([r.QLalign_e for r in recs], 0)
        nR, nQ = r0.R28.shape[1], r0.QL28.shape[1]
        dev = r0.R28.device
        self.kR = torch.tensor([i * nR for i in col_op], dtype=torch.int32, device=dev)
        self.kQ = torch.tensor([i * nQ for i in col_op], dtype=torch.int32, device=dev)
