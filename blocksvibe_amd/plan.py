"""Index plan of the batched radial recursion for one operator (two_j12, two_j43, two_j_max, order,
pseudo-parities): which residue every task writes, which H row it reads, which prime-pole residues every
Cauchy row sums over, and which M / Q constants it multiplies by.  Pure integers, built on the host once per
operator and uploaded once (Recursion); depends on no number.

Semantics follow reference blocksvibe/plan_ref.py:13-121 (and its vectorised twin recursion_gpu.py:136-365):
  * logical slot of residue (p, m, jidx, pole) = p * nslots + first[m][jidx] + pole index;
  * per level m, groups (src, two_j') in order of first appearance, rows D' per group in order of first
    appearance, columns = the prime poles of two_j' at order m (residues of level m - shift);
  * M keys (pole, 0|1, two_jNN, ppp, two_j) and Q keys (pole, two_j, sign) by first appearance over all
    levels, L before R per task; (L, Q) pairs likewise;
  * physical slots: a residue's slot is recycled once the last level reading it as a Cauchy column has
    run (or right after its own level if nobody reads it), LIFO free list.
Differences by design: H rows are stored compactly (no padding to the largest group), and the Cauchy
coefficients are never tabulated densely -- the kernel computes D' - D_i from the row/column positions.
"""
import numpy as np

from . import structure as S


def first_rank(codes):
    """rank of every code among the distinct codes in order of first appearance; and those codes"""
    codes = np.asarray(codes, dtype=np.int64)
    if codes.size == 0:
        return np.zeros(0, np.int64), np.zeros(0, np.int64)
    uniq, first, inv = np.unique(codes, return_index=True, return_inverse=True)
    by_first = np.argsort(first, kind="stable")
    rank = np.empty(len(uniq), dtype=np.int64)
    rank[by_first] = np.arange(len(uniq))
    return rank[inv.reshape(-1)], uniq[by_first]


class Level:
    """index arrays of one radial level (host numpy)"""
    __slots__ = ("m", "nb", "Kc", "gsrc", "gtjp", "ncol", "col", "cold", "nH", "rg", "rd", "T", "t_slot",
                 "t_hrow", "t_L", "t_R", "t_q", "t_qL", "t_keys", "dmax")


class Plan:
    def __init__(self, two_j12, two_j43, two_j_max, order, pseudo_parities, reuse=True):
        self.tj12, self.tj43, self.tjmax, self.order = two_j12, two_j43, two_j_max, order
        self.pp = tuple(tuple(int(x) for x in p) for p in pseudo_parities)
        b = two_j12 % 2
        self.b = b
        tj_all = np.arange(b, two_j_max + 2 * order + 3, 2)
        self.SL = np.array([[len(S.so3(self.pp[p][0], two_j12, int(tj))) for tj in tj_all] for p in range(2)])
        self.SR = np.array([[len(S.so3(self.pp[p][1], two_j43, int(tj))) for tj in tj_all] for p in range(2)])
        nsp = (two_j_max + 2 * order - b) // 2 + 1                 # spins up to two_j_max + 2 order
        # structure entries per residue: SLm left x SRm right (a residue block (a, b) sits at a * SRm + b);
        # Smax = max of the two is the row stride of the M-matrix tables
        self.SLm = int(self.SL[:, :nsp].max())
        self.SRm = int(self.SR[:, :nsp].max())
        self.S2 = self.SLm * self.SRm
        self.Smax = max(self.SLm, self.SRm)
        self.Jh = len(range(b, two_j_max + 2 * order + 1, 2))      # spins of h_inf
        # logical slot layout
        self.level_tjs, self.level_npol, self.level_first = [], [], []
        n = 0
        for m in range(order):
            tjs = S.spins_at_level(m, order, two_j12, two_j_max)
            ords = order - m - S.spin_excess(tjs, two_j_max)
            npol = S.n_poles(tjs, two_j12, two_j43, ords).sum(0)
            first = n + np.concatenate([[0], np.cumsum(npol)[:-1]]).astype(np.int64) if len(npol) else \
                np.zeros(0, np.int64)
            n += int(npol.sum())
            self.level_tjs.append(tjs)
            self.level_npol.append(npol)
            self.level_first.append(first)
        self.nslots = n
        self._lstart = np.concatenate([[0], np.cumsum([len(x) for x in self.level_tjs])[:-1]]).astype(np.int64)
        self._npol_flat = np.concatenate(self.level_npol).astype(np.int64) if order else np.zeros(0, np.int64)
        self._first_flat = np.concatenate(self.level_first).astype(np.int64) if order else np.zeros(0, np.int64)
        self.levels = [self._level(m) for m in range(order)]
        self._keys()
        self.reuse = reuse
        self._physical(reuse)

    # ------------------------------------------------------------------------------------------------
    def ji(self, tj):
        return (np.asarray(tj) - self.b) // 2

    def offset(self, m, jidx):
        return int(self.level_first[m][jidx])

    def _level(self, m):
        tj12, tj43, tjmax, order, ns = self.tj12, self.tj43, self.tjmax, self.order, self.nslots
        tjs, npol = self.level_tjs[m], self.level_npol[m]
        tot = int(npol.sum())
        if tot == 0:
            return None
        jidx = np.repeat(np.arange(len(tjs)), npol)
        pidx = np.arange(tot) - np.repeat(np.cumsum(npol) - npol, npol)
        tj = tjs[jidx]
        ords = order - m - S.spin_excess(tj, tjmax)
        t, k = S.pole_of_index(tj, tj12, tj43, ords, pidx)
        # tasks: parity p innermost
        jidx, pidx, tj, t, k = (np.repeat(a, 2) for a in (jidx, pidx, tj, t, k))
        p = np.tile(np.array([0, 1], dtype=np.int64), tot)
        tjp = S.pole_spin(t, k, tj)
        dprime = S.pole_delta2(t, k, tj) + 2 * S.shift(t, k)
        src = (p + (t == 3)) % 2
        keep = ((self.SL[p, self.ji(tj)] > 0) & (self.SR[p, self.ji(tj)] > 0) &
                (self.SL[src, self.ji(tjp)] > 0) & (self.SR[src, self.ji(tjp)] > 0))
        if not keep.any():
            return None
        jidx, pidx, tj, t, k, p, tjp, dprime, src = (a[keep] for a in (jidx, pidx, tj, t, k, p, tjp, dprime, src))
        lev = Level()
        lev.m = m
        # groups (src, two_j') and rows D'
        gi, gcodes = first_rank(src * (1 << 32) + tjp)
        nb = len(gcodes)
        gsrc, gtjp = gcodes >> 32, gcodes & 0xFFFFFFFF
        dmin = int(dprime.min())
        ri, rcodes = first_rank(gi * (1 << 32) + (dprime - dmin))
        rgrp = rcodes >> 32
        by_g = np.argsort(rgrp, kind="stable")                     # rows grouped, first appearance within
        cnt = np.bincount(rgrp, minlength=nb)
        gstart = np.concatenate([[0], np.cumsum(cnt)[:-1]])
        hrow_of_code = np.empty(len(rcodes), dtype=np.int64)
        hrow_of_code[by_g] = np.arange(len(rcodes))
        lev.nH = len(rcodes)
        lev.rg = rgrp[by_g].astype(np.int32)
        lev.rd = ((rcodes[by_g] & 0xFFFFFFFF) + dmin).astype(np.int32)
        # columns: prime poles of two_j' at order m, residues of level m - shift
        ncol = S.n_poles(gtjp, tj12, tj43, m).sum(0) if nb else np.zeros(0, np.int64)
        Kc = max(1, int(ncol.max()) if nb else 1)
        cg = np.repeat(np.arange(nb), ncol)
        cc = np.arange(int(ncol.sum())) - np.repeat(np.cumsum(ncol) - ncol, ncol)
        ctj = gtjp[cg]
        ti, ki = S.pole_of_index(ctj, tj12, tj43, m, cc)
        si = S.shift(ti, ki)
        lvl = m - si
        cord = order - m + si - S.spin_excess(ctj, tjmax)
        cpos = S.index_of_pole(ctj, tj12, tj43, cord, ti, ki)
        col = np.full((nb, Kc), -1, dtype=np.int64)
        cold = np.zeros((nb, Kc), dtype=np.int32)
        if len(cg):
            cji = self.ji(ctj)
            # the read residue must exist at its level (same spin block, index within its pole list)
            fi = self._lstart[lvl] + cji
            assert np.all(cji < np.array([len(x) for x in self.level_tjs])[lvl]), "prime-pole spin outside its level"
            assert np.all(cpos < self._npol_flat[fi]), "prime-pole position outside its residue block"
            fst = self._first_flat[fi]
            col[cg, cc] = gsrc[cg] * ns + fst + cpos
            cold[cg, cc] = S.pole_delta2(ti, ki, ctj)
        lev.nb, lev.Kc, lev.gsrc, lev.gtjp, lev.ncol, lev.col, lev.cold = nb, Kc, gsrc, gtjp, ncol.astype(np.int32), \
            col, cold
        # largest |D' - D_i| of the level (the Cauchy kernel divides by it with one-limb arithmetic)
        dm = 0
        if len(cg):
            rows_of = [lev.rd[lev.rg == g] for g in range(nb)]
            for g in range(nb):
                if ncol[g] and len(rows_of[g]):
                    c = cold[g, :ncol[g]]
                    dm = max(dm, int(np.abs(rows_of[g][:, None] - c[None, :]).max()))
        lev.dmax = dm
        # tasks
        lev.T = len(p)
        lev.t_slot = p * ns + self.level_first[m][jidx] + pidx
        lev.t_hrow = hrow_of_code[ri]
        sign = np.where((tjp - tj) % 4 != 0, -1, 1)
        ppa = np.array(self.pp)
        lev.t_keys = (t, k, tj, ppa[src, 0], ppa[src, 1], sign)
        return lev

    def _keys(self):
        """M keys (L before R per task), Q keys and (L, Q) pairs, by first appearance over all levels"""
        lv = [lev for lev in self.levels if lev is not None]
        if not lv:
            self.Mkeys, self.Qkeys, self.LQpairs = {}, {}, np.zeros((0, 2), np.int64)
            return
        t, k, tj, pL, pR, sign = (np.concatenate([lev.t_keys[i] for lev in lv]) for i in range(6))
        KM = self.order + 2
        TJM = self.tjmax + 2 * self.order + 4
        PPM = 2

        def mcode(d, pq):
            return (((t * KM + k) * 2 + d) * PPM + pq) * TJM + tj

        midx, mcodes = first_rank(np.stack([mcode(0, pL), mcode(1, pR)], 1).reshape(-1))
        tju, u = mcodes % TJM, mcodes // TJM
        ppu, u = u % PPM, u // PPM
        du, u = u % 2, u // 2
        ku, tu = u % KM, u // KM
        self.Mkeys = {((int(a), int(b_)), int(d), self.tj12 if d == 0 else self.tj43, int(c), int(e)): i
                      for i, (a, b_, d, c, e) in enumerate(zip(tu, ku, du, ppu, tju))}
        qidx, qcodes = first_rank(((t * KM + k) * TJM + tj) * 2 + (sign > 0))
        su, u = qcodes % 2, qcodes // 2
        tju, u = u % TJM, u // TJM
        self.Qkeys = {((int(x // KM), int(x % KM)), int(a), 1 if s else -1): i
                      for i, (x, a, s) in enumerate(zip(u, tju, su))}
        nQ = len(self.Qkeys)
        lq, lqcodes = first_rank(midx[0::2] * nQ + qidx)
        self.LQpairs = np.stack([lqcodes // nQ, lqcodes % nQ], 1) if len(lqcodes) else np.zeros((0, 2), np.int64)
        o = 0
        for lev in lv:
            n = lev.T
            lev.t_L = midx[2 * o:2 * (o + n):2].copy()
            lev.t_R = midx[2 * o + 1:2 * (o + n):2].copy()
            lev.t_q = qidx[o:o + n].copy()
            lev.t_qL = lq[o:o + n].copy()
            o += n
        self.dmax = max(lev.dmax for lev in lv)

    def _physical(self, reuse):
        """physical slots: without reuse one per written logical slot; with reuse a logical slot's physical slot
        goes back to a LIFO stack once its last reader's level has run (the level after its own if nothing
        reads it), and the slots born at a level take the most recently returned physical slots first, in
        logical order, before new ones are numbered.  Returns wait in per-level lists that are handed back in
        the order their return level was first scheduled (a skipped level hands back several at once), so the
        numbering, and with it the published slot counts, is reproduced exactly."""
        N = 2 * self.nslots
        written = np.full(N, -1, dtype=np.int64)           # level that writes the slot
        read_to = np.full(N, -1, dtype=np.int64)           # last level that reads it
        for lev in self.levels:
            if lev is None:
                continue
            written[lev.t_slot] = lev.m
            cols = lev.col.reshape(-1)
            np.maximum.at(read_to, cols[cols >= 0], lev.m)
        phys = np.full(N, -1, dtype=np.int64)
        if not reuse:
            live = np.flatnonzero(written >= 0)
            phys[live] = live
            self.nphys = N
        else:
            stack = np.zeros(0, dtype=np.int64)              # free physical slots, top at the end
            returns = {}                                     # return level -> [physical slot arrays]
            made = 0
            for lev in self.levels:
                if lev is None:
                    continue
                here = lev.m
                due = [r for r in returns if r <= here]      # scheduling order, not numeric order
                if due:
                    stack = np.concatenate([stack] + [a for r in due for a in returns.pop(r)])
                born = np.sort(lev.t_slot)
                k = min(len(born), len(stack))
                got = np.empty(len(born), dtype=np.int64)
                got[:k] = stack[len(stack) - k:][::-1]
                got[k:] = np.arange(made, made + len(born) - k)
                made += len(born) - k
                stack = stack[:len(stack) - k]
                phys[born] = got
                back_at = np.maximum(read_to[born], here + 1)
                levels_in_order = back_at[np.sort(np.unique(back_at, return_index=True)[1])]
                for r in levels_in_order:
                    returns.setdefault(int(r), []).append(got[back_at == r])
            self.nphys = made
        self.phys = phys
        self.zero_slot = self.nphys
        for lev in self.levels:
            if lev is None:
                continue
            lev.t_slot = phys[lev.t_slot]
            c = lev.col
            pc = np.full(c.shape, self.nphys, dtype=np.int64)
            ok = c >= 0
            pc[ok] = phys[c[ok]]
            lev.col = np.where(pc >= 0, pc, self.nphys)

    # ------------------------------------------------------------------------------------------------
    def phys_slot(self, logical):
        """physical slot of a logical one (the zero slot if it is never written)"""
        if logical >= len(self.phys) or self.phys[logical] < 0:
            return self.zero_slot
        return int(self.phys[logical])

    def counts(self):
        """(Cauchy output entries, pole-step output entries) per right-hand side (SLm x SRm structure entries
        per residue) -- what one run computes"""
        h = sum(lev.nH for lev in self.levels if lev) * self.S2
        r = sum(lev.T for lev in self.levels if lev) * self.S2
        return h, r

    def cauchy_terms(self):
        """sum over levels of (H rows x prime-pole columns of their group): Cauchy terms per structure
        entry and right-hand side"""
        tot = 0
        for lev in self.levels:
            if lev is None:
                continue
            tot += int(np.sum(lev.ncol[lev.rg]))
        return tot
