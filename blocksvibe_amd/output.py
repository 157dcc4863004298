"""blocks_3d's write_output on the GPU: numerator polynomials of all lambda orders of a task -> derivatives
in the xt or xt_radial coordinates (reference blocksvibe/output.py:1-212; blocks_3d's
lambda_lr_to_lrho_lrhob, congruence_transform, zzb_to_xt, lrho_to_xt_radial).

Every step is a scalar linear map acting on each x^k coefficient separately, so the whole chain is one
matrix W (outputs x inputs) per (coordinate, Lambda, parity, Delta_1+2), built once and applied to every
spin, structure, coefficient and task as one RNS GEMM:
  xt_radial: W = pre R_z S  (order O = Lambda - parity; pre = power_prefactor(O, D) for parity 0,
             one_minus_power_prefactor(O) power_prefactor(O, D + 1) for parity 1);
  xt:        F_u = the (lambda, log r) -> (lambda_rho, lambda_rhobar) image of input u = (n_lambda index, e)
             (one antidiagonal, sigma-symmetric), Z_u = M F_u M^T with M = P(D/2) R_z S (P only if
             Delta_1+2 != 0), W[(n, b), u] = ((m - p)/2)! / m! sum_{k, l} (-1)^l C(n, k) C(m, l)
             Z_u(n + m - k - l, k + l), m = p + 2 b.
S is the Stirling-number matrix scaled by r_c^-n, R_z the inverse of blocks_3d's rho -> z series matrix.

Construction (build_map_device): the (O + 1)^2 matrices S, R_z, P and their product M on the host in mpmath
(O^3 operations), then W on the GPU at LG >= L + 4 limbs (128 guard bits, as the host construction's
32 L + 128): for each antidiagonal s, T_s[(a, b), k] = M[a, k] M[b, s - k] (elementwise products over the
pairs a + b <= O the map reads) and Z[(a, b), u] = T_s G_s with G_s[k, u] = F_u(k, s - k) (exact); then
W = C Z with C[(n, b), (a, b')] = ((m - p)/2)! / m! sum_l (-1)^l C(n, b' - l) C(m, l) on a + b' = n + m
(exact), both RNS GEMMs; W is rounded once more to L limbs.  build_map is the same map entirely in mpmath
(O^5 operations at O = Lambda: minutes at Lambda = 43), kept to check the device construction.
"""
import math

import mpmath as mp
import numpy as np
import torch

from . import mpf, rns


def _binom_gen(k, j):
    """binomial coefficient generalised to negative upper arguments (blocks_3d's binomial_coefficient)"""
    def rising(a, n):
        r = 1
        for q in range(n):
            r *= a + q
        return r
    if k < j:
        return mp.mpf(0) if j < 0 else mp.mpf(rising(k - j + 1, j)) / math.factorial(j)
    return mp.mpf(rising(j + 1, k - j)) / math.factorial(k - j)


def _falling(x, n):
    """x (x - 1) ... (x - n + 1), and 1 / ((x + 1) ... (x - n)) for n < 0"""
    if n < 0:
        r = mp.mpf(1)
        for q in range(-n):
            r *= x + 1 + q
        return 1 / r
    r = mp.mpf(1)
    for q in range(n):
        r *= x - q
    return r


def stirling_matrix(O):
    """S[n][k] = r_c^-n s(n, k), s the signed Stirling numbers of the first kind"""
    s = [[0] * (O + 1) for _ in range(O + 1)]
    s[0][0] = 1
    for n in range(O):
        for k in range(n + 1):
            s[n + 1][k + 1] += s[n][k]
            s[n + 1][k] += -n * s[n][k]
    rc = 3 - 2 * mp.sqrt(2)
    return mp.matrix([[rc ** -n * s[n][k] for k in range(O + 1)] for n in range(O + 1)])


def rho_to_z(O):
    """inverse of the matrix of z-derivatives of rho-derivatives at the crossing point (blocks_3d rho_z)"""
    z = mp.matrix(O + 1, O + 1)
    s2 = mp.sqrt(2)
    for p in range(O + 1):
        for n in range(p + 1):
            acc = mp.mpf(0)
            for k in range(n + 1):
                acc += (_binom_gen(n, k) * mp.mpf(2) ** -n * (-1) ** k * (4 * s2) ** (n - k)
                        * _binom_gen(-2 * n, p - k - n) * (4 - 2 * s2) ** (-n - p + k))
            z[p, n] = mp.factorial(p) / mp.factorial(n) * acc
    return mp.inverse(z)


def power_prefactor(O, D):
    P = mp.matrix(O + 1, O + 1)
    for n in range(O + 1):
        for k in range(n + 1):
            P[n, k] = _falling(-D, n - k) * mp.mpf(2) ** (n - k + D) * _binom_gen(n, k)
    return P


def one_minus_power_prefactor(O):
    P = mp.matrix(O + 1, O + 1)
    h = mp.mpf(1) / 2
    for n in range(O + 1):
        for k in range(n + 1):
            P[n, k] = (-1) ** (n - k) * _falling(-h, n - k) * mp.mpf(2) ** (n - k + h) * _binom_gen(n, k)
    return P


def build_map(coordinate, Lam, parity, delta_1_plus_2, L):
    """(W as a list of rows of mpf, inputs [(idx, e)], outputs [(n, b)]) at 32 L + 128 bits"""
    mp.mp.prec = 32 * L + 128
    D = mp.mpf(delta_1_plus_2)
    p = parity
    if coordinate == "xt_radial":
        O = Lam - p
        pre = power_prefactor(O, D) if p == 0 else one_minus_power_prefactor(O) * power_prefactor(O, D + 1)
        T = pre * rho_to_z(O) * stirling_matrix(O)
        inputs = [(0, e) for e in range(O + 1)]
        outputs = [(i, 0) for i in range(O + 1)]
        W = [[T[i, e] for e in range(O + 1)] for i in range(O + 1)]
        return W, inputs, outputs
    if coordinate != "xt":
        raise ValueError(coordinate)
    O = Lam
    sig = -1 if p % 2 else 1
    M = rho_to_z(O) * stirling_matrix(O)
    if D != 0:
        M = power_prefactor(O, D / 2) * M
    inputs = [(idx, e) for idx in range((O - p) // 2 + 1) for e in range(O - (p + 2 * idx) + 1)]
    outputs = [(n, b) for n in range(O - p + 1) for b in range((O - n - p) // 2 + 1)]
    W = [[mp.mpf(0)] * len(inputs) for _ in outputs]
    for u, (idx, e) in enumerate(inputs):
        s = p + 2 * idx + e
        F = mp.matrix(O + 1, O + 1)
        for i in range(s + 1):
            j = s - i
            hi, lo = (i, j) if i >= j else (j, i)
            if lo > min(hi, O - hi):
                continue
            acc = 0
            for a in range(max(0, e - lo), min(hi, e) + 1):
                b = e - a
                acc += (-1) ** (lo - b) * math.comb(hi, a) * math.comb(lo, b)
            v = mp.mpf(acc) * mp.mpf(2) ** -s
            F[i, j] = v if i >= j else sig * v
        Z = M * F * M.T
        for o, (n, b) in enumerate(outputs):
            m = p + 2 * b
            fac = mp.factorial((m - p) // 2) / mp.factorial(m)
            tot = mp.mpf(0)
            for jj in range(n + m + 1):
                c = sum((-1) ** l * math.comb(n, jj - l) * math.comb(m, l)
                        for l in range(max(0, jj - n), min(m, jj) + 1))
                if c:
                    tot += c * Z[n + m - jj, jj]
            W[o][u] = fac * tot
    return W, inputs, outputs


def guard_limbs(L):
    """the limb count of the map construction: the smallest compiled one >= L + 4"""
    from . import _lib
    for LG in range(L + 4, L + 17):
        if _lib.lib().b3d_supported_limbs(LG):
            return LG
    raise ValueError(f"no compiled limb count in [{L + 4}, {L + 16}] for the map construction at L = {L}")


def _xt_factors(O, p):
    """host exact parts of the xt map: inputs, outputs, the per-input antidiagonal images F_u(k, s - k) as
    Fractions, and the output rows' coefficients on the Z pairs (a, b) (a + b = n + m) as Fractions"""
    from fractions import Fraction
    sig = -1 if p % 2 else 1
    inputs = [(idx, e) for idx in range((O - p) // 2 + 1) for e in range(O - (p + 2 * idx) + 1)]
    outputs = [(n, b) for n in range(O - p + 1) for b in range((O - n - p) // 2 + 1)]
    G = []
    for idx, e in inputs:
        s = p + 2 * idx + e
        col = []
        for i in range(s + 1):
            j = s - i
            hi, lo = (i, j) if i >= j else (j, i)
            if lo > min(hi, O - hi):
                col.append(Fraction(0))
                continue
            acc = 0
            for a in range(max(0, e - lo), min(hi, e) + 1):
                acc += (-1) ** (lo - (e - a)) * math.comb(hi, a) * math.comb(lo, e - a)
            v = Fraction(acc, 1 << s)
            col.append(v if i >= j else sig * v)
        G.append((s, col))
    C = []
    for n, b in outputs:
        m = p + 2 * b
        fac = Fraction(math.factorial((m - p) // 2), math.factorial(m))
        row = {}
        for jj in range(n + m + 1):
            c = sum((-1) ** l * math.comb(n, jj - l) * math.comb(m, l) for l in range(max(0, jj - n), min(m, jj) + 1))
            if c:
                row[(n + m - jj, jj)] = fac * c
        C.append(row)
    return inputs, outputs, G, C


def build_map_device(coordinate, Lam, parity, delta_1_plus_2, L):
    """(W as a device MPF (outputs, inputs) at L limbs, inputs [(idx, e)], outputs [(n, b)]): the map of
    build_map, its O^5 part on the GPU (module docstring)"""
    p = parity
    if coordinate == "xt_radial":
        W, inputs, outputs = build_map(coordinate, Lam, parity, delta_1_plus_2, L)   # O^3 on the host
        return mpf.from_mp([x for row in W for x in row], L, (len(outputs), len(inputs))), inputs, outputs
    if coordinate != "xt":
        raise ValueError(coordinate)
    LG = guard_limbs(L)
    O = Lam
    mp.mp.prec = 32 * LG + 64
    D = mp.mpf(delta_1_plus_2)
    M = rho_to_z(O) * stirling_matrix(O)
    if D != 0:
        M = power_prefactor(O, D / 2) * M
    inputs, outputs, G, C = _xt_factors(O, p)
    O1 = O + 1
    Md = mpf.from_mp([M[i, j] for i in range(O1) for j in range(O1)], LG, (O1 * O1,))
    pairs = [(a, b) for a in range(O1) for b in range(O1 - a)]
    pidx = {ab: i for i, ab in enumerate(pairs)}
    dev = Md.e.device
    pa = torch.tensor([a for a, _ in pairs], dtype=torch.int64, device=dev)
    pb = torch.tensor([b for _, b in pairs], dtype=torch.int64, device=dev)
    Z = mpf.zeros((len(pairs), len(inputs)), LG)
    by_s = {}
    for u, (s, _) in enumerate(G):
        by_s.setdefault(s, []).append(u)
    for s, us in sorted(by_s.items()):
        k = torch.arange(s + 1, dtype=torch.int64, device=dev)
        Ta = Md.take((pa[:, None] * O1 + k[None, :]).reshape(-1))
        Tb = Md.take((pb[:, None] * O1 + (s - k)[None, :]).reshape(-1))
        T = mpf.mul(Ta, Tb).reshape(len(pairs), s + 1)
        Gs = mpf.from_fractions([G[u][1][kk] for kk in range(s + 1) for u in us], LG, (s + 1, len(us)))
        Zs = rns.gemm(T, Gs)
        cols = torch.tensor(us, dtype=torch.int64, device=dev)
        Z.m[:, :, cols] = Zs.m
        Z.e[:, cols] = Zs.e
    Cd = mpf.from_fractions([row.get(ab, 0) for row in C for ab in pairs], LG, (len(outputs), len(pairs)))
    Wg = rns.gemm(Cd, Z)
    W = mpf.norm_wide(Wg.m.reshape(LG, -1).contiguous(), Wg.e.reshape(-1).contiguous(), L)
    return W.reshape(len(outputs), len(inputs)), inputs, outputs


class Output:
    """write_output for one coordinate ("xt" | "xt_radial"), Lambda, parity (= (1 - four_pt_sign) / 2) and
    Delta_1+2 (decimal string) at L limbs."""

    def __init__(self, coordinate, Lam, parity, delta_1_plus_2, L, W=None, inputs=None, outputs=None):
        self.coord, self.Lam, self.parity, self.d12, self.L = coordinate, Lam, parity, delta_1_plus_2, L
        if W is None:
            W, inputs, outputs = build_map_device(coordinate, Lam, parity, delta_1_plus_2, L)
        self.W, self.inputs, self.outputs = W, list(inputs), list(outputs)
        self.nL = (Lam - parity) // 2 + 1
        # rows of each lambda order's derivatives that the map reads (device index tables, built once)
        self._rows = {}
        for idx in sorted({i for i, _ in self.inputs}):
            self._rows[idx] = torch.tensor([e for i, e in self.inputs if i == idx], device=mpf.dev())

    def apply_per_order(self, nums):
        """nums[idx]: MPF (Jo, Dout_idx, R_idx + 1, 2, S2, nq) for the task's lambda orders idx (all of them for
        xt, the first for xt_radial).  Returns MPF (Jo, 2, S2, nq, outputs, Dmax) (a permuted view of
        apply_raw's result)."""
        return self.apply_raw(nums).permute(1, 2, 3, 4, 0, 5)

    def apply_raw(self, nums):
        """as apply_per_order, in the GEMM's own contiguous layout (outputs, Jo, 2, S2, nq, Dmax)"""
        Dmax = max(x.shape[1] for x in nums)
        rows = []
        for idx, x in enumerate(nums):
            if idx not in self._rows:
                continue
            Jo, D, R1, _, S2, nq = x.shape
            X = x.permute(2, 0, 3, 4, 5, 1)                                      # (e, jo, p, s, q, d)
            if D < Dmax:
                X = mpf.cat([X, mpf.zeros((R1, Jo, 2, S2, nq, Dmax - D), self.L)], 5)
            rows.append(X.reshape(R1, -1).take(self._rows[idx]))
        X = mpf.cat(rows, 0)
        Jo, _, _, _, S2, nq = nums[0].shape
        return rns.gemm(self.W, X).reshape(len(self.outputs), Jo, 2, S2, nq, Dmax)


def cached(coordinate, Lam, parity, delta_1_plus_2, L, cache_dir=None):
    """Output whose map is kept on disk under cache_dir (its host construction takes seconds to minutes at
    large Lambda), keyed by everything the map depends on (reference blocksvibe/cli.py:281-309 keeps the same
    cache); without cache_dir, or if the file cannot be read or written, the map is built in memory"""
    import hashlib
    import os
    if not cache_dir:
        return Output(coordinate, Lam, parity, delta_1_plus_2, L)
    tag = hashlib.sha1(f"{coordinate}|{Lam}|{parity}|{delta_1_plus_2}|{L}|map3".encode()).hexdigest()[:20]
    path = os.path.join(cache_dir, f"wmap_{coordinate}_L{Lam}_p{parity}_{tag}.pt")
    try:
        st = torch.load(path, weights_only=True)
        W = mpf.MPF(st["m"].to(mpf.dev()), st["e"].to(mpf.dev()))
        return Output(coordinate, Lam, parity, delta_1_plus_2, L, W=W,
                      inputs=[tuple(x) for x in st["inputs"].tolist()], outputs=[tuple(x) for x in st["outputs"].tolist()])
    except (OSError, KeyError, RuntimeError):
        pass
    o = Output(coordinate, Lam, parity, delta_1_plus_2, L)
    try:
        os.makedirs(cache_dir, exist_ok=True)
        tmp = f"{path}.{os.getpid()}"
        torch.save({"m": o.W.m.cpu(), "e": o.W.e.cpu(), "inputs": torch.tensor(o.inputs),
                    "outputs": torch.tensor(o.outputs)}, tmp)
        os.replace(tmp, path)
    except OSError:
        pass
    return o
