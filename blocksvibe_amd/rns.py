"""Exact multiprecision GEMM through a residue number system on int8 MFMA (libb3dhip b3d_rns_gemm).

Semantics (the build's own definition of the reference's unvendored sdpvibe.mpf.gemm, reference
blocksvibe/smallgemm.py:1-6 "A aligned per row, B per column, exact products, one rounding per entry"):
C[i, j] = round(sum_k Abar[i, k] Bbar[k, j]) with Abar the row-aligned mantissas of A (exponent eA[i]) and
Bbar the column-aligned mantissas of B (exponent eB[j]).

Tables: the nmod largest primes below 2^15 with product M > 2^(2 NB + ceil(log2 K) + 4); int8 weights of
the byte values 2^(8 j) mod p for the MFMA residue conversion; the CRT constants (M/p)^-1 mod p and the
signed radix-256 digits of M/p and M for the MFMA CRT lift.
"""
import math

import numpy as np
import torch

from . import _lib, mpf

_PRIMES = None
_TABLES = {}
WORKSPACE_BYTES = 1 << 31


def primes_desc():
    """primes below 2^15, largest first"""
    global _PRIMES
    if _PRIMES is None:
        n = 1 << 15
        sieve = np.ones(n, dtype=bool)
        sieve[:2] = False
        for i in range(2, int(n ** 0.5) + 1):
            if sieve[i]:
                sieve[i * i::i] = False
        _PRIMES = [int(p) for p in np.flatnonzero(sieve)[::-1]]
    return _PRIMES


def n_moduli(L, K):
    need = 2 * (32 * L - 2) + max(1, math.ceil(math.log2(max(K, 2)))) + 4
    bits, n = 0.0, 0
    for p in primes_desc():
        bits += math.log2(p)
        n += 1
        if bits > need:
            return n
    raise ValueError("not enough primes")


def conversion_weights(ps, L):
    """int8 weights of the MFMA residue conversion: for byte j < 4L of a magnitude and prime p_t,
    c = 2^(8 j) mod p_t taken balanced (|c| <= p_t / 2 < 2^14) and split c = 256 ch + cl (cl in [-128, 127],
    |ch| <= 64): w8[0, t, j] = ch, w8[1, t, j] = cl, bytes padded to a multiple of 64 and primes to a
    multiple of 16 with zeros.  The kernel feeds byte - 128; wcorr[t] = 128 sum_j c restores sum_j byte_j c."""
    KB = (4 * L + 63) // 64 * 64
    n16 = (len(ps) + 15) // 16 * 16
    w8 = np.zeros((2, n16, KB), dtype=np.int8)
    wcorr = np.zeros(n16, dtype=np.int32)
    for t, p in enumerate(ps):
        c = np.array([pow(2, 8 * j, p) for j in range(4 * L)], dtype=np.int64)
        c = np.where(c > p // 2, c - p, c)
        ch = (c + 128) >> 8
        cl = c - 256 * ch
        w8[0, t, :4 * L] = ch
        w8[1, t, :4 * L] = cl
        wcorr[t] = 128 * int(c.sum())
    return w8, wcorr


def signed_digits(X, n):
    """X >= 0 as n signed radix-256 digits in [-128, 127] (X = sum d_j 256^j)"""
    d = []
    for _ in range(n):
        v = X & 255
        if v >= 128:
            v -= 256
        d.append(v)
        X = (X - v) >> 8
    assert X == 0, "not enough digits"
    return d


def crt_digits(ps, L):
    """int8 matrix of the MFMA CRT lift, [byte column][k]: k = 2t + 1 (low digit l_t of y_t): the signed
    radix-256 digits of M/p_t; k = 2t (high digit h_t, weight 256): the same digits one column up;
    k = 2 nmod (the coefficient -q): the digits of M.  Byte columns cover 2L + 2 limbs plus two, rounded up
    to 16; k is padded with zeros to a multiple of 64."""
    n = len(ps)
    LW = 2 * L + 2
    CN = (4 * LW + 2 + 15) // 16 * 16
    CK = (2 * n + 1 + 63) // 64 * 64
    M = 1
    for p in ps:
        M *= p
    c8 = np.zeros((CN, CK), dtype=np.int8)
    for t, p in enumerate(ps):
        d = signed_digits(M // p, CN - 1)
        c8[:CN - 1, 2 * t + 1] = d
        c8[1:, 2 * t] = d
    c8[:, 2 * n] = signed_digits(M, CN)
    return c8


def tables(L, nmod):
    key = (L, nmod)
    if key in _TABLES:
        return _TABLES[key]
    ps = primes_desc()[:nmod]
    M = 1
    for p in ps:
        M *= p
    LW = 2 * L + 2
    assert M.bit_length() < 32 * LW - 1, "residue product does not fit the CRT accumulator"
    w8, wcorr = conversion_weights(ps, L)
    inv = np.array([pow((M // p) % p, -1, p) for p in ps], dtype=np.uint16)
    rp = np.array([1.0 / p for p in ps], dtype=np.float64)
    c8 = crt_digits(ps, L)
    d = mpf.dev()
    t = {"nmod": nmod,
         "primes": torch.from_numpy(np.array(ps, dtype=np.uint16).view(np.int16)).to(d),
         "w8": torch.from_numpy(w8).to(d),
         "wcorr": torch.from_numpy(wcorr).to(d),
         "inv": torch.from_numpy(inv.view(np.int16)).to(d),
         "rp": torch.from_numpy(rp).to(d),
         "c8": torch.from_numpy(c8).to(d), "CK": int(c8.shape[1])}
    _TABLES[key] = t
    return t


def _chunk_cols(nmod, nbA, nb, M, N, K, budget):
    lib = _lib.lib()
    per = lib.b3d_rns_workspace_bytes(nmod, nbA, nb, M, 64, K) - lib.b3d_rns_workspace_bytes(nmod, nbA, nb, M, 0, K)
    fixed = lib.b3d_rns_workspace_bytes(nmod, nbA, nb, M, 0, K)
    nc = max(64, int((budget - fixed) // max(per / 64.0, 1)) // 64 * 64)
    return min(nc, N)


def gemm_strided(L, nb, M, N, K, A, a_strides, a_shared, B, b_strides, C, c_strides, budget=None):
    """low-level call: A, B, C are MPF (any shapes) addressed through (batch, row, col) strides"""
    if nb == 0 or M == 0 or N == 0:
        return C
    nmod = n_moduli(L, K)
    t = tables(L, nmod)
    nbA = 1 if a_shared else nb
    budget = budget or WORKSPACE_BYTES
    Nc = _chunk_cols(nmod, nbA, nb, M, N, K, budget)
    ws_bytes = _lib.lib().b3d_rns_workspace_bytes(nmod, nbA, nb, M, Nc, K)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=A.e.device)
    EB = torch.empty(nb * N, dtype=torch.int64, device=A.e.device)
    P = lambda x: x.data_ptr()
    _lib.call("b3d_rns_gemm", L, nmod, P(t["primes"]), P(t["w8"]), P(t["wcorr"]), P(t["inv"]), P(t["rp"]),
              P(t["c8"]), t["CK"],
              nb, M, N, K, 1 if a_shared else 0,
              P(A.m), P(A.e), A.m[0].numel(), *a_strides,
              P(B.m), P(B.e), B.m[0].numel(), *b_strides,
              P(C.m), P(C.e), C.m[0].numel(), *c_strides, Nc, P(EB), P(ws), ws_bytes, mpf.stream_ptr())
    return C


def gemm(A, B, out=None, budget=None):
    """A: MPF (M, K) (shared) or (nb, M, K); B: MPF (K, N) or (nb, K, N) -> MPF (M, N) or (nb, M, N)"""
    A, B = A.contiguous(), B.contiguous()
    L = A.L
    assert B.L == L
    if len(B.shape) == 2:
        assert len(A.shape) == 2
        Mr, K = A.shape
        K2, N = B.shape
        assert K == K2
        C = out or mpf.zeros((Mr, N), L, A.e.device)
        return gemm_strided(L, 1, Mr, N, K, A, (0, K, 1), True, B, (0, N, 1), C, (0, N, 1), budget)
    nb, K, N = B.shape
    if len(A.shape) == 2:
        Mr, K2 = A.shape
        shared, asb = True, 0
    else:
        nbA, Mr, K2 = A.shape
        assert nbA == nb
        shared, asb = False, Mr * K2
    assert K == K2, (A.shape, B.shape)
    C = out or mpf.zeros((nb, Mr, N), L, A.e.device)
    return gemm_strided(L, nb, Mr, N, K, A, (asb, K, 1), shared, B, (K * N, N, 1), C, (Mr * N, N, 1), budget)
