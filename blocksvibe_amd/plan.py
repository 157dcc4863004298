"""Index plan of the batched radial recursion for one operator (two_j12, two_j43, two_j_max, order,
pseudo-parities): which residue every task writes, which H row it reads, which prime-pole residues every
Cauchy row sums over, and which M / Q constants it multiplies by.  Pure integers, built on the host once per
operator by the library (b3d_plan_build, C++, libb3dhip) and uploaded once (Recursion); depends on no number.

Semantics follow reference blocksvibe/plan_ref.py:13-121 (and its vectorised twin recursion_gpu.py:136-365):
  * logical slot of residue (p, m, jidx, pole) = p * nslots + first[m][jidx] + pole index;
  * per level m, groups (src, two_j') in order of first appearance, rows D' per group in order of first
    appearance, columns = the prime poles of two_j' at order m (residues of level m - shift);
  * M keys (pole, 0|1, two_jNN, ppp, two_j) and Q keys (pole, two_j, sign) by first appearance over all
    levels, L before R per task; (L, Q) pairs likewise;
  * physical slots: a residue's slot is recycled once the last level reading it as a Cauchy column has
    run (or right after its own level if nobody reads it), LIFO free list, returns handed back in the order
    their return level was first scheduled (so the numbering, and the published slot counts, are the
    reference's).
Differences by design: H rows are stored compactly (no padding to the largest group), and the Cauchy
coefficients are never tabulated densely -- the kernel computes D' - D_i from the row/column positions.
"""
import numpy as np

from . import _lib


class Level:
    """index arrays of one radial level (host numpy)"""
    __slots__ = ("m", "nb", "Kc", "gsrc", "gtjp", "ncol", "col", "cold", "nH", "rg", "rd", "T", "t_slot",
                 "t_hrow", "t_L", "t_R", "t_q", "t_qL", "t_keys", "dmax")


class Plan:
    def __init__(self, two_j12, two_j43, two_j_max, order, pseudo_parities, reuse=True):
        self.tj12, self.tj43, self.tjmax, self.order = two_j12, two_j43, two_j_max, order
        self.pp = tuple(tuple(int(x) for x in p) for p in pseudo_parities)
        self.reuse = reuse
        lib, K = _lib.lib(), _lib.constants()
        h = _lib.ctypes.c_void_p()
        _lib.call("b3d_plan_build", two_j12, two_j43, two_j_max, order, self.pp[0][0], self.pp[0][1], self.pp[1][0],
                  self.pp[1][1], 1 if reuse else 0, _lib.ctypes.byref(h))
        try:
            def get(field, level=0):
                n = lib.b3d_plan_get(h, K["B3D_PLAN_" + field], level, None, 0)
                if n < 0:
                    raise _lib.B3DError(f"b3d_plan_get: no table {field} at level {level}")
                out = np.empty(n, dtype=np.int64)
                lib.b3d_plan_get(h, K["B3D_PLAN_" + field], level, out.ctypes.data, n)
                return out

            (self.nslots, self.nphys, self.SLm, self.SRm, self.S2, self.Smax, self.Jh, self.dmax, _,
             self.b) = (int(x) for x in get("SCALARS"))
            self.SL = get("SL").reshape(2, -1)
            self.SR = get("SR").reshape(2, -1)
            self.level_tjs = [get("LEVEL_TJS", m) for m in range(order)]
            self.level_npol = [get("LEVEL_NPOL", m) for m in range(order)]
            self.level_first = [get("LEVEL_FIRST", m) for m in range(order)]
            self.phys = get("PHYS")
            mk = get("MKEYS").reshape(-1, 6)
            self.Mkeys = {((int(t), int(k)), int(d), int(nn), int(ppp), int(tj)): i
                          for i, (t, k, d, nn, ppp, tj) in enumerate(mk.tolist())}
            qk = get("QKEYS").reshape(-1, 4)
            self.Qkeys = {((int(t), int(k)), int(tj), int(sg)): i for i, (t, k, tj, sg) in enumerate(qk.tolist())}
            self.LQpairs = get("LQPAIRS").reshape(-1, 2)
            self.levels = [self._level(get, m) for m in range(order)]
        finally:
            lib.b3d_plan_free(h)
        self.zero_slot = self.nphys

    @staticmethod
    def _level(get, m):
        present, mm, nb, Kc, nH, T, dmax = (int(x) for x in get("LEVEL", m))
        if not present:
            return None
        lev = Level()
        lev.m, lev.nb, lev.Kc, lev.nH, lev.T, lev.dmax = mm, nb, Kc, nH, T, dmax
        lev.gsrc, lev.gtjp = get("GSRC", m), get("GTJP", m)
        lev.ncol = get("NCOL", m).astype(np.int32)
        lev.col = get("COL", m).reshape(nb, Kc)
        lev.cold = get("COLD", m).astype(np.int32).reshape(nb, Kc)
        lev.rg, lev.rd = get("RG", m).astype(np.int32), get("RD", m).astype(np.int32)
        lev.t_slot, lev.t_hrow = get("T_SLOT", m), get("T_HROW", m)
        lev.t_L, lev.t_R, lev.t_q, lev.t_qL = get("T_L", m), get("T_R", m), get("T_Q", m), get("T_QL", m)
        lev.t_keys = tuple(get("T_KEYS", m).reshape(6, T))
        return lev

    # ------------------------------------------------------------------------------------------------
    def ji(self, tj):
        return (np.asarray(tj) - self.b) // 2

    def offset(self, m, jidx):
        return int(self.level_first[m][jidx])

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
