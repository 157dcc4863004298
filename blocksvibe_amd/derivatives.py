"""blocks_3d's post-recursion stages on the GPU: residues -> derivatives at r_c -> r^Delta restored ->
Delta -> x shift -> numerator polynomials, for every output spin, parity, structure and right-hand side
(reference blocksvibe/derivatives.py:1-34, 313-582; blocks_3d's conformal_block_derivatives,
restore_r_delta, shift_block, together_with_factors, shift_poles).

Per (spin j, parity, structure), with w[k][n] = n^k r_c^n, r_c = 3 - 2 sqrt 2, R = Lambda - n_lambda,
poles i of spin j at full order (positions D_i, shifts s_i), x = Delta - ub_j:
  c[k]    = sum_{m <= order} w[k][m] h_inf[m]                       (one GEMM, K = order + 1)
  v_i[k]  = sum_m w[k][m + s_i] res[m][i]                           (accumulated level by level)
  a_i[e]  = sum_{d <= e} C(e, d) D_i^(e - d) v_i[d]                 (triangular products, exact integers)
  G[w][d] = sum_i D_i^w v_i[d];  H[e][s] = sum_d C(e, d) G[s-1-d][d] (GEMMs)
  P_e[t]  = C(e, t) c[e - t] + H[e][e - t]
  numerator_e(x) = P_e(x + ub) Pi(x) + sum_i a_i[e] Pi_i(x)          (one GEMM with an exact matrix)
Pole shifting (kept_pole_order < order) replaces the unprotected poles by the kept ones with residues
S a, S_ki = (y'_k / y_i)^h l_k(y_i) (closed form of blocks_3d's A^-1 V, Lagrange basis l_k on the kept poles,
h = size // 2), folded into the numerator matrix as exact rationals, rounded once.
"""
import copy
import threading
from fractions import Fraction
from math import comb, lcm

import mpmath as mp
import numpy as np
import torch

from . import _lib, mpf, rns
from . import structure as S

_ORDERS_LOCK = threading.Lock()


def _polymul_lin(p, a):
    """p(y) (y + a), integer coefficients, lowest first"""
    out = [0] * (len(p) + 1)
    for k, c in enumerate(p):
        out[k] += a * c
        out[k + 1] += c
    return out


def _divide_root(p, r):
    """p(y) / (y - r) for a root r of p (exact synthetic division)"""
    n = len(p) - 1
    q = [0] * n
    acc = 0
    for k in range(n, 0, -1):
        acc = p[k] + r * acc
        q[k - 1] = acc
    assert p[0] + r * acc == 0
    return q


def numerator_columns(tj, tj12, tj43, order, kept, R):
    """exact numerator matrix of one spin: dict column -> list of x^k coefficients (Fractions).
    Columns: i (< number of full-order poles) for the pole terms, ('t', t) for Pi(x) (x + ub)^t."""
    two_ub = S.upper_bound_2(tj)
    full = S.poles_list(tj, tj12, tj43, order)
    roots = [int(S.pole_delta2(t, k, tj)) - two_ub for t, k in full]          # y_i = 2 x_i (integers)
    if kept is None:
        new = list(roots)
    else:
        new = [y for y in roots if y == 0] + \
              [int(S.pole_delta2(t, k, tj)) - two_ub for t, k in S.poles_list(tj, tj12, tj43, kept)
               if int(S.pole_delta2(t, k, tj)) != two_ub]
    npj = len(new)
    Pi = [1]
    for y0 in new:
        Pi = _polymul_lin(Pi, -y0)
    cols = {}
    q = Pi
    for t in range(R + 1):
        cols[("t", t)] = [Fraction(c) * Fraction(2) ** (k - (npj + t)) for k, c in enumerate(q)]
        q = _polymul_lin(q, two_ub)
    comp = [_divide_root(Pi, y0) for y0 in new]
    if kept is None:
        for i, qi in enumerate(comp):
            cols[i] = [Fraction(c) * Fraction(2) ** (k - (npj - 1)) for k, c in enumerate(qi)]
        return cols, npj
    nprot = sum(1 for y in new if y == 0)
    keep = new[nprot:]
    size = len(keep)
    h = size // 2
    # S_ki = (y'_k / y)^h prod_{m != k} (y - y'_m) / (y'_k - y'_m): integer weights over the common
    # denominator y^h D, D = lcm |prod_{m != k} (y'_k - y'_m)|
    Dk = [1] * size
    for k in range(size):
        for m in range(size):
            if m != k:
                Dk[k] *= keep[k] - keep[m]
    D = 1
    for d in Dk:
        D = lcm(D, abs(d))
    fac = [keep[k] ** h * (D // Dk[k]) for k in range(size)]           # exact: Dk divides D up to sign
    Q = np.array([comp[nprot + k] + [0] * (npj - len(comp[nprot + k])) for k in range(size)], dtype=object) \
        if size else None
    for i, y in enumerate(roots):
        if y == 0:
            cols[i] = [Fraction(c) * Fraction(2) ** (k - (npj - 1)) for k, c in enumerate(comp[new.index(0)])]
            continue
        if size == 0:
            cols[i] = [Fraction(0)]
            continue
        diff = [y - km for km in keep]
        pre, suf = [1] * (size + 1), [1] * (size + 1)
        for k in range(size):
            pre[k + 1] = pre[k] * diff[k]
        for k in range(size - 1, -1, -1):
            suf[k] = suf[k + 1] * diff[k]
        w = np.array([fac[k] * pre[k] * suf[k + 1] for k in range(size)], dtype=object)
        poly = w.dot(Q)
        den = y ** h * D
        cols[i] = [Fraction(int(c), den) * Fraction(2) ** (d - (npj - 1)) for d, c in enumerate(poly)]
    return cols, npj


class Derivs:
    """Post-recursion stages for one Plan and a batch of right-hand sides whose lambda order has
    R = Lambda - n_lambda.  two_js: output spins (default: every two_j <= two_j_max)."""

    def __init__(self, plan, R, L, two_js=None, kept_pole_order=None):
        self.plan, self.R, self.L = plan, R, L
        self.kept = kept_pole_order if kept_pole_order is not None and kept_pole_order < plan.order else None
        order, tj12, tj43 = plan.order, plan.tj12, plan.tj43
        self.tjs = list(two_js) if two_js is not None else list(range(tj12 % 2, plan.tjmax + 1, 2))
        self.Jo = len(self.tjs)
        self.poles = [S.poles_list(tj, tj12, tj43, order) for tj in self.tjs]
        self.NP = max(1, max(len(p) for p in self.poles))
        self._levels()
        self._constants()
        self._finish_consts()
        self._orders = {R: self}
        self._jk = {}                 # moment-flush column lists per r1 pattern (shared by every order's copy)

    # --------------------------------------------------------------------------------------- bookkeeping
    def _levels(self):
        pl, order, Jo, NP = self.plan, self.plan.order, self.Jo, self.NP
        tj12, tj43 = pl.tj12, pl.tj43
        jo_of = {tj: jo for jo, tj in enumerate(self.tjs)}
        srow = np.zeros((2, Jo, NP), dtype=np.int32)
        for jo, ps in enumerate(self.poles):
            srow[:, jo, :len(ps)] = [int(S.shift(np.int64(t), np.int64(k))) for t, k in ps]
        self.srow_host = srow.reshape(-1)
        self.levels = []
        for m in range(order):
            tjs = pl.level_tjs[m]
            jsel = np.array([j for j, tj in enumerate(tjs) if int(tj) in jo_of], dtype=np.int64)
            if len(jsel) == 0:
                self.levels.append(None)
                continue
            npol = pl.level_npol[m][jsel]
            tot = int(npol.sum())
            if tot == 0:
                self.levels.append(None)
                continue
            rep = np.repeat(np.arange(len(jsel)), npol)
            pidx = np.arange(tot) - np.repeat(np.cumsum(npol) - npol, npol)
            jidx = jsel[rep]
            tj = tjs[jidx]
            jo = np.array([jo_of[int(x)] for x in tj], dtype=np.int64)
            t, k = S.pole_of_index(tj, tj12, tj43, order - m - S.spin_excess(tj, pl.tjmax), pidx)
            i = S.index_of_pole(tj, tj12, tj43, order, t, k)
            nn = m + S.shift(t, k)
            slots, rows, ns = [], [], []
            for p in range(2):
                ph = pl.phys[p * pl.nslots + pl.level_first[m][jidx] + pidx]
                ok = ph >= 0
                slots.append(ph[ok])
                rows.append(((p * Jo + jo) * NP + i)[ok])
                ns.append(nn[ok])
            # interleave as (pole, parity) like the plan's task order
            sl, rw, nn_ = (np.concatenate(x) for x in (slots, rows, ns))
            if len(sl) == 0:
                self.levels.append(None)
                continue
            self.levels.append((sl.astype(np.int64), rw.astype(np.int64), nn_.astype(np.int32)))

    def _constants(self):
        L, R, order, Jo, NP = self.L, self.R, self.plan.order, self.Jo, self.NP
        pl = self.plan
        mp.mp.prec = 32 * L + 64
        rc = 3 - 2 * mp.sqrt(2)
        self.rcpow = mpf.from_mp([rc ** n for n in range(order + 1)], L, (order + 1,))
        self.Wh = mpf.from_mp([(mp.mpf(n) ** k if (n or k) else mp.mpf(1)) * rc ** n
                               for k in range(R + 1) for n in range(order + 1)], L, (R + 1, order + 1))
        # C(e, k) D^(e - k), D = d / 2, per distinct doubled pole position d (index 0: zero padding)
        d2 = sorted({int(S.pole_delta2(t, k, tj)) for tj, ps in zip(self.tjs, self.poles) for t, k in ps})
        self.d2 = d2
        di = {d: n + 1 for n, d in enumerate(d2)}
        pairs = [(0, mpf.EZ)] * ((R + 1) * (R + 1))
        for d in d2:
            for e in range(R + 1):
                for k in range(R + 1):
                    pairs.append(mpf.norm_int(comb(e, k) * d ** (e - k), k - e, L) if k <= e else (0, mpf.EZ))
        self.Ttab = mpf.from_pairs(pairs, L, (len(d2) + 1, R + 1, R + 1))
        self.tri = self._tri_tables(d2, R + 1)
        didx = np.zeros((Jo, NP), dtype=np.int64)
        for jo, (tj, ps) in enumerate(zip(self.tjs, self.poles)):
            didx[jo, :len(ps)] = [di[int(S.pole_delta2(t, k, tj))] for t, k in ps]
        self.vand = self._vand_tables([[int(S.pole_delta2(t, k, tj)) for t, k in ps]
                                       for tj, ps in zip(self.tjs, self.poles)], NP, max(R, 1))
        self.didx = torch.from_numpy(didx).to(mpf.dev())
        # numerator matrix (Jo, Dout, NP + R + 1): columns Pi_i (pole terms), then Pi (x + ub)^t
        mats, npjs = [], []
        for tj in self.tjs:
            cols, npj = numerator_columns(tj, pl.tj12, pl.tj43, order, self.kept, R)
            mats.append(cols)
            npjs.append(npj)
        self.Dout = max(npjs) + R + 1
        W = NP + R + 1
        vals = [(0, mpf.EZ)] * (Jo * self.Dout * W)
        for jo, cols in enumerate(mats):
            for key, poly in cols.items():
                c = NP + key[1] if isinstance(key, tuple) else key
                for kdeg, v in enumerate(poly):
                    if v:
                        vals[(jo * self.Dout + kdeg) * W + c] = mpf.norm_fraction(v.numerator, v.denominator, L)
        self.npjs = npjs
        self.Anum = mpf.from_pairs(vals, L, (Jo, self.Dout, W))
        # exact n^k for the moment flush, n up to order plus the largest shift
        nmax = order + int(self.srow_host.max(initial=0)) + 1
        v = [n ** k for n in range(nmax + 1) for k in range(R + 1)]
        KL = max(1, (max(v).bit_length() + 31) // 32)
        if KL > 8:
            raise _lib.B3DError("n^k exceeds 8 limbs")
        self.KL = KL
        self.NK = torch.from_numpy(np.array([[(x >> (32 * l)) & 0xFFFFFFFF for x in v] for l in range(KL)],
                                            dtype=np.uint32).view(np.int32)).to(mpf.dev())
        self.NKb = torch.tensor([x.bit_length() for x in v], dtype=torch.int32, device=mpf.dev())
        self.srow = torch.from_numpy(self.srow_host).to(mpf.dev())
        self.dev_levels = [None if lv is None else tuple(torch.from_numpy(a).to(mpf.dev()) for a in lv)
                           for lv in self.levels]

    # --------------------------------------------------------------------------------------- moments
    def start(self, B, window=8, r1=None):
        """zero the moment accumulators for B right-hand sides.  r1: per right-hand side the number of moments
        k it needs (at most R + 1; right-hand sides of several lambda orders in one recursion, each with its own
        R): the flushes compute only those (split then hands each order its columns); None: all R + 1"""
        self.B, self.N = B, self.plan.S2 * B
        self.nrows = 2 * self.Jo * self.NP
        self.mom = mpf.zeros((self.nrows, self.N, self.R + 1), self.L)
        self.jk, self.Kc = None, B * (self.R + 1)
        if r1 is not None:
            # built on the first use of a pattern (eagerly: a graph capture of the same shape comes later) and
            # kept, so that a captured run reads the same list
            key = (self.R, tuple(r1))
            jk = self._jk.get(key)
            if jk is None:
                assert len(r1) == B and all(0 < x <= self.R + 1 for x in r1)
                codes = [b * (self.R + 1) + k for b in range(B) for k in range(r1[b])]
                jk = self._jk.setdefault(key, torch.tensor(codes, dtype=torch.int32, device=self.mom.e.device))
            self.jk, self.Kc = jk, int(jk.numel())
        self.W = window
        # the flush reads a staged entry only when its top limb is nonzero (zero: nothing staged there), so only
        # that limb plane is cleared, here and after each flush; the other limbs and the exponents of an
        # entry are written by the stage kernel before they are read
        self.stage = mpf.empty((window, self.nrows, self.N), self.L)
        self.stage.m[self.L - 1].zero_()
        self.m0, self.pending = 0, False

    def _flush(self):
        P = lambda x: x.data_ptr()
        _lib.call("b3d_moments_flush", self.L, self.KL, P(self.stage.m), P(self.stage.e), self.stage.numel(), self.W,
                  self.m0, P(self.srow), P(self.NK), self.NK.shape[1], P(self.NKb), P(self.mom.m), P(self.mom.e),
                  self.mom.numel(), self.nrows, self.N, self.R + 1, self.B, P(self.jk) if self.jk is not None else 0,
                  self.Kc, mpf.stream_ptr())
        self.stage.m[self.L - 1].zero_()
        self.pending = False

    def accumulate(self, m, res):
        """recursion hook: stage level m's residues of output spins (times r_c^n); flush every W levels"""
        if self.pending and m - self.m0 >= self.W:
            self._flush()
        lv = self.dev_levels[m]
        if lv is None:
            return
        sl, rw, nn = lv
        if not self.pending:
            self.m0, self.pending = m, True
        w = m - self.m0
        P = lambda x: x.data_ptr()
        _lib.call("b3d_moments_stage", self.L, P(res.m), P(res.e), res.numel(), P(sl), P(rw), P(nn), sl.numel(),
                  self.N, P(self.rcpow.m), P(self.rcpow.e), self.rcpow.numel(), P(self.stage.m), P(self.stage.e),
                  self.stage.numel(), w * self.nrows * self.N, mpf.stream_ptr())
        if w == self.W - 1:
            self._flush()

    def split(self, parts):
        """after one recursion over the right-hand sides of several lambda orders (start(B, r1=...)): parts =
        [(the order's Derivs (for_order(R_i)), first right-hand side b0, count Bi)] -> those Derivs, each holding
        the moments k <= R_i of its own columns b0 .. b0 + Bi, ready for finish (h_inf: the same columns).  The
        moments are exactly the ones the order's own run would have accumulated (every kernel treats columns
        independently)."""
        if self.pending:
            self._flush()
        self.stage = None
        S2, R1 = self.plan.S2, self.R + 1
        m4 = self.mom.reshape(self.nrows, S2, self.B, R1)
        out = []
        for d, b0, Bi in parts:
            assert d.R <= self.R and d.plan is self.plan and 0 <= b0 and b0 + Bi <= self.B
            sub = m4[:, :, b0:b0 + Bi, :d.R + 1].contiguous()
            d.B, d.N, d.nrows = Bi, S2 * Bi, self.nrows
            d.mom = sub.reshape(self.nrows, S2 * Bi, d.R + 1)
            d.stage, d.pending, d.jk, d.Kc = None, False, None, Bi * (d.R + 1)
            out.append(d)
        self.mom = None
        return out

    # --------------------------------------------------------------------------------------- final maps
    def finish(self, h_inf, chunk=None):
        """h_inf: MPF (2, order + 1, Jh, SL SR, B).  Returns numerators MPF (Jo, Dout, R+1, 2, SL SR, B):
        [spin, x^k coefficient, log r derivative e, parity, structure a*SR+b, right-hand side]"""
        if self.pending:
            self._flush()
        self.stage = None
        L, R, Jo, NP, order = self.L, self.R, self.Jo, self.NP, self.plan.order
        R1, Rw = R + 1, max(R, 1)
        S2, B, N = self.plan.S2, self.B, self.N
        Nc = 2 * N
        jidx, cidx, hidx, gidx, cb = self._fc
        H = mpf.MPF(h_inf.m[:, :, :order + 1][:, :, :, jidx], h_inf.e[:, :order + 1][:, :, jidx])  # (2,O1,Jo,S2,B)
        H = H.permute(1, 2, 0, 3, 4).reshape(order + 1, Jo * Nc)
        c = rns.gemm(self.Wh, H).reshape(R1, Jo, Nc)                                  # c[d]
        mom = self.mom.reshape(2, Jo, NP, N, R1).permute(1, 2, 4, 0, 3)                # (Jo, NP, R1, 2, N)
        chunk = chunk or Jo
        outs = []
        for j0 in range(0, Jo, chunk):
            j1 = min(Jo, j0 + chunk)
            J = j1 - j0
            v = mom[j0:j1].reshape(J * NP, R1, Nc)
            # the GEMM operands of the pole tables, only where a product kernel by exact integers is not taken
            T = (self.Ttab.take(self.didx[j0:j1].reshape(-1))                           # (J NP, e, d)
                 if self.tri is None or self.vand is None else None)
            # numerator GEMM operand (J, NP + R1, e Nc): rows < NP the residues a_i[e], then P below
            Bn = mpf.MPF(torch.empty((L, J, NP + R1, R1 * Nc), dtype=torch.int32, device=v.e.device),
                         torch.empty((J, NP + R1, R1 * Nc), dtype=torch.int64, device=v.e.device))
            if self.tri is not None:
                TL, tm, tsg, texp = self.tri
                v = v.contiguous()
                _lib.call("b3d_tri_product", L, TL, v.m.data_ptr(), v.e.data_ptr(), v.numel(), J * NP, R1, Nc,
                          self.didx[j0:j1].contiguous().data_ptr(), tm.data_ptr(), tsg.data_ptr(), tsg.numel(),
                          texp.data_ptr(), Bn.m.data_ptr(), Bn.e.data_ptr(), Bn.numel(), NP, (NP + R1) * R1 * Nc,
                          R1 * Nc, mpf.stream_ptr())
            else:
                a = rns.gemm(T, v)                                                      # (J NP, e, Nc)
                Bn.m[:, :, :NP] = a.m.reshape(L, J, NP, R1 * Nc)
                Bn.e[:, :NP] = a.e.reshape(J, NP, R1 * Nc)
            if self.vand is not None:                                                   # (J, w, d Nc)
                G = self._vand_product(v.reshape(J, NP, R1 * Nc), j0, j1, Rw)
            else:
                Vand = mpf.MPF(T.m[:, :, :Rw, 0].reshape(L, J, NP, Rw).permute(0, 1, 3, 2).contiguous(),
                               T.e[:, :Rw, 0].reshape(J, NP, Rw).permute(0, 2, 1).contiguous())  # (J, w, NP)
                G = rns.gemm(Vand, v.reshape(J, NP, R1 * Nc))
            Gz = mpf.cat([G.reshape(J, Rw * R1, Nc), mpf.zeros((J, 1, Nc), L)], 1)
            Gs = Gz.permute(1, 0, 2).take(gidx).reshape(R1, R1, J, Nc)                  # (s, d, J, Nc)
            Gs = Gs.permute(1, 2, 0, 3).reshape(R1, J * R1 * Nc)                        # (d, J s Nc)
            Hm = self._binomial_product(Gs).reshape(R1, J, R1, Nc)                      # (e, J, s, Nc)
            Hz = mpf.cat([Hm.permute(0, 2, 1, 3).reshape(R1 * R1, J, Nc), mpf.zeros((1, J, Nc), L)], 0)
            cz = mpf.cat([c[:, j0:j1], mpf.zeros((1, J, Nc), L)], 0)                    # (d | 0, J, Nc)
            cbx = mpf.MPF(cb.m.expand(L, R1 * R1, J, Nc).contiguous(), cb.e.expand(R1 * R1, J, Nc).contiguous())
            P = mpf.add(mpf.mul(cz.take(cidx), cbx), Hz.take(hidx))                     # ((t, e), J, Nc)
            P = P.reshape(R1, R1, J, Nc).permute(2, 0, 1, 3).reshape(J, R1, R1 * Nc)    # (J, t, e Nc)
            Bn.m[:, :, NP:] = P.m
            Bn.e[:, NP:] = P.e
            A = self.Anum[j0:j1]
            num = rns.gemm(A.contiguous(), Bn)                                          # (J, Dout, e Nc)
            outs.append(num.reshape(J, self.Dout, R1, 2, S2, B))
        return outs[0] if len(outs) == 1 else mpf.cat(outs, 0)

    @staticmethod
    def _vand_tables(d2s, NP, Rw):
        """the powers D_i^w (w < Rw) of each spin's pole positions D_i = d_i / 2 as exact integers d_i^w with the
        row unit 2^-w, zero for the padding poles, for b3d_int_product: (TL, magnitude limb planes
        (TL, Jo, Rw, NP), sign masks (Jo, Rw, NP), unit exponents (Jo, Rw)); None if an entry needs more than 8
        limbs"""
        Jo = len(d2s)
        ints = [[(d2[i] ** w if i < len(d2) else 0) for w in range(Rw) for i in range(NP)] for d2 in d2s]
        bits = max(abs(x).bit_length() for row in ints for x in row)
        TL = next((t for t in (2, 4, 8) if bits <= 32 * t), None)
        if TL is None:
            return None
        mags = [abs(x) for row in ints for x in row]
        tm = np.array([[(x >> (32 * l)) & 0xFFFFFFFF for x in mags] for l in range(TL)], dtype=np.uint64)
        tm = torch.from_numpy(tm.astype(np.uint32).view(np.int32).reshape(TL, Jo, Rw, NP)).to(mpf.dev())
        tsg = torch.tensor([-1 if x < 0 else 0 for row in ints for x in row], dtype=torch.int32,
                           device=mpf.dev()).reshape(Jo, Rw, NP)
        texp = torch.tensor([-w for _ in range(Jo) for w in range(Rw)], dtype=torch.int64,
                            device=mpf.dev()).reshape(Jo, Rw)
        return TL, tm, tsg, texp

    def _vand_product(self, v, j0, j1, Rw):
        """G[j] = (D_i^w)_{w, i} v[j] for the spins j0 .. j1 (v: (J, NP, N)): the GEMM's bits (each column
        aligned to its largest exponent, exact sum, one rounding) by b3d_int_product with the exact powers"""
        TL, tm, tsg, texp = self.vand
        J, NP, N = v.shape
        v = v.contiguous()
        out = mpf.empty((J, Rw, N), self.L, v.e.device)
        _lib.call("b3d_int_product", self.L, TL, v.m.data_ptr(), v.e.data_ptr(), v.numel(), J, Rw, NP, 0, N,
                  self._jsel[j0:j1].data_ptr(), tm.data_ptr(), tsg.data_ptr(), tsg.numel(), texp.data_ptr(),
                  out.m.data_ptr(), out.e.data_ptr(), out.numel(), 1, Rw * N, 0, mpf.stream_ptr())
        return out

    def _binomial_product(self, Gs):
        """C Gs for Gs (R + 1, N), C[e][d] = binomial(e, d): the multiprecision GEMM's bits (each column aligned to its
        largest exponent, exact sum, one rounding) by the triangular product with the exact binomials"""
        TL, tm, tsg, texp, tsel = self._cbtri
        R1, N = Gs.shape
        Gs = Gs.contiguous()
        out = mpf.empty((R1, N), self.L, Gs.e.device)
        _lib.call("b3d_tri_product", self.L, TL, Gs.m.data_ptr(), Gs.e.data_ptr(), Gs.numel(), 1, R1, N,
                  tsel.data_ptr(), tm.data_ptr(), tsg.data_ptr(), tsg.numel(), texp.data_ptr(), out.m.data_ptr(),
                  out.e.data_ptr(), out.numel(), 1, 0, 0, mpf.stream_ptr())
        return out

    @staticmethod
    def _tri_tables(d2, R1):
        """the same matrices C(e, k) D^(e - k) as exact integers C(e, k) d^(e - k) 2^k (d = 2 D) with the row
        unit 2^-e, for b3d_tri_product: (TL, magnitude limb planes (TL, ntab, R1, R1), sign masks
        (ntab, R1, R1), unit exponents (ntab, R1)); None if an entry needs more than 8 limbs"""
        ints = [[0] * (R1 * R1)] + [[comb(e, k) * d ** (e - k) * 2 ** k if k <= e else 0
                                     for e in range(R1) for k in range(R1)] for d in d2]
        bits = max(abs(x).bit_length() for row in ints for x in row)
        TL = next((t for t in (2, 4, 8) if bits <= 32 * t), None)
        if TL is None:
            return None
        mags = np.array([abs(x) for row in ints for x in row], dtype=object)
        tm = np.array([[(int(x) >> (32 * l)) & 0xFFFFFFFF for x in mags] for l in range(TL)], dtype=np.uint64)
        tm = torch.from_numpy(tm.astype(np.uint32).view(np.int32).reshape(TL, len(ints), R1, R1)).to(mpf.dev())
        tsg = torch.tensor([-1 if x < 0 else 0 for row in ints for x in row], dtype=torch.int32,
                           device=mpf.dev()).reshape(len(ints), R1, R1)
        texp = torch.tensor([-e for _ in ints for e in range(R1)], dtype=torch.int64,
                            device=mpf.dev()).reshape(len(ints), R1)
        return TL, tm, tsg, texp

    def _finish_consts(self):
        """index tables and binomials of finish() for this R (device constants, built once so that finish()
        itself only launches device work and can be captured in a graph)"""
        R1, Rw, L, dev = self.R + 1, max(self.R, 1), self.L, mpf.dev()
        jidx = torch.tensor([(tj - self.plan.tj12 % 2) // 2 for tj in self.tjs], device=dev)
        te = [(t, e) for t in range(R1) for e in range(R1)]
        cidx = torch.tensor([e - t if t <= e else R1 for t, e in te], device=dev)
        hidx = torch.tensor([e * R1 + (e - t) if t <= e else R1 * R1 for t, e in te], device=dev)
        cbv = [mpf.norm_int(comb(e, t), 0, L) if t <= e else (0, mpf.EZ) for t, e in te]
        gidx = torch.tensor([(s - 1 - d) * R1 + d if 1 <= s and d <= s - 1 else Rw * R1
                             for s in range(R1) for d in range(R1)], device=dev)
        cb = mpf.from_pairs(cbv, L, (R1 * R1, 1, 1))
        self._fc = (jidx, cidx, hidx, gidx, cb)
        self._jsel = torch.arange(len(self.tjs), dtype=torch.int64, device=dev)      # spin j's table is j
        # Hm = C Gs by b3d_tri_product: C[e][d] = binomial(e, d) is lower triangular with small exact integer
        # entries (row unit 2^0), so the triangular product gives the GEMM's bits without the residue conversions
        ints = [comb(e, d) if d <= e else 0 for e in range(R1) for d in range(R1)]
        TL = next(t for t in (2, 4, 8) if max(ints).bit_length() <= 32 * t)
        tm = np.array([[(x >> (32 * l)) & 0xFFFFFFFF for x in ints] for l in range(TL)], dtype=np.uint64)
        self._cbtri = (TL, torch.from_numpy(tm.astype(np.uint32).view(np.int32).reshape(TL, 1, R1, R1)).to(dev),
                       torch.zeros((1, R1, R1), dtype=torch.int32, device=dev),
                       torch.zeros((1, R1), dtype=torch.int64, device=dev),
                       torch.zeros(1, dtype=torch.int64, device=dev))

    def for_order(self, R):
        """the same stages for a smaller R (a higher lambda order of the task): the constants are built once
        per R and cached; each call returns its own object for the run state (start / accumulate / finish)"""
        assert R <= self.R
        d = self._orders.get(R)
        if d is None:
            # several streams may ask at once: one builds, and its stream is synchronised before the constants
            # are shared
            with _ORDERS_LOCK:
                d = self._orders.get(R)
                if d is None:
                    d = self._build_order(R)
                    torch.cuda.current_stream().synchronize()
                    self._orders[R] = d
        return copy.copy(d)                          # shared constants, own run state (start/accumulate)

    def _build_order(self, R):
        d = object.__new__(Derivs)
        d.__dict__.update(self.__dict__)
        d.R = R
        R1 = R + 1
        d.Dout = max(self.npjs) + R + 1
        W = self.NP + R + 1
        # numerator matrix: the pole columns, then t <= R; rows up to the smaller degree
        A = self.Anum
        d.Anum = mpf.MPF(A.m[:, :, :d.Dout, :W].contiguous(), A.e[:, :d.Dout, :W].contiguous())
        d.Ttab = mpf.MPF(self.Ttab.m[:, :, :R1, :R1].contiguous(), self.Ttab.e[:, :R1, :R1].contiguous())
        if self.tri is not None:
            TL, tm, tsg, texp = self.tri
            d.tri = (TL, tm[:, :, :R1, :R1].contiguous(), tsg[:, :R1, :R1].contiguous(), texp[:, :R1].contiguous())
        if self.vand is not None:
            TL, tm, tsg, texp = self.vand
            Rw = max(R, 1)
            d.vand = (TL, tm[:, :, :Rw].contiguous(), tsg[:, :Rw].contiguous(), texp[:, :Rw].contiguous())
        d.Wh = mpf.MPF(self.Wh.m[:, :R1].contiguous(), self.Wh.e[:R1].contiguous())
        v = [n ** k for n in range(self.NK.shape[1] // (self.R + 1)) for k in range(R1)]
        d.KL = max(1, (max(v).bit_length() + 31) // 32)
        d.NK = torch.from_numpy(np.array([[(x >> (32 * l)) & 0xFFFFFFFF for x in v] for l in range(d.KL)],
                                         dtype=np.uint32).view(np.int32)).to(mpf.dev())
        d.NKb = torch.tensor([x.bit_length() for x in v], dtype=torch.int32, device=mpf.dev())
        d._finish_consts()
        return d


def run_orders(rec, der, hs, Rs, operators=None, timer=None, h_all=None, window=8):
    """the recursion and the derivative stages for the right-hand sides of several lambda orders of one operator
    and sign in ONE recursion: hs[i] is order i's h_inf MPF (2, order + 1, Jh, S2, B_i), Rs[i] its number of
    log r derivatives R_i (der.for_order(R_i) holds its constants).  The columns of all orders run through
    each level together (one launch per level for all of them), the moment flushes compute each column's
    moments k <= R_i only, and each order's numerators come from its own columns (Derivs.split, finish): the
    same bits as one recursion per order.  h_all: the hs concatenated along the right-hand sides, if the caller
    keeps them so; window: levels per moment flush (as Derivs.start).  Returns the numerators per order (as
    Derivs.finish)."""
    Bs = [h.shape[-1] for h in hs]
    if h_all is None:
        h_all = hs[0] if len(hs) == 1 else mpf.cat(hs, len(hs[0].shape) - 1)
    dm = der.for_order(max(Rs))
    dm.start(sum(Bs), window=window, r1=[R + 1 for R, B in zip(Rs, Bs) for _ in range(B)])
    rec.run(h_all, hook=dm.accumulate, operators=operators, timer=timer)
    parts, b0 = [], 0
    for R, B in zip(Rs, Bs):
        parts.append((der.for_order(R), b0, B))
        b0 += B
    return [d.finish(h_all[..., b0:b0 + B]) for d, (_, b0, B) in zip(dm.split(parts), parts)]

