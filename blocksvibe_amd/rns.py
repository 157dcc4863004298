"""Exact multiprecision GEMM through a residue number system on int8 MFMA (libb3dhip b3d_rns_gemm).

Semantics (the build's own definition of the reference's unvendored sdpvibe.mpf.gemm, reference
blocksvibe/smallgemm.py:1-6 "A aligned per row, B per column, exact products, one rounding per entry"):
C[i, j] = round(sum_k Abar[i, k] Bbar[k, j]) with Abar the row-aligned mantissas of A (exponent eA[i]) and
Bbar the column-aligned mantissas of B (exponent eB[j]).

Tables: the nmod largest primes below 2^15 with product M > 2^(2 NB + ceil(log2 K) + 4); residues of the
limb weights 2^(32 l) mod p; the CRT constants (M/p)^-1 mod p and the limbs of M/p and M.
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
    pw = np.array([[pow(2, 32 * l, p) for l in range(L)] for p in ps], dtype=np.uint32)
    inv = np.array([pow((M // p) % p, -1, p) for p in ps], dtype=np.uint16)
    rp = np.array([1.0 / p for p in ps], dtype=np.float64)
    mask = (1 << 32) - 1
    Mp = np.array([[((M // p) >> (32 * q)) & mask for q in range(LW)] for p in ps], dtype=np.uint32)
    Mf = np.array([(M >> (32 * q)) & mask for q in range(LW)], dtype=np.uint32)
    d = mpf.dev()
    t = {"nmod": nmod,
         "primes": torch.from_numpy(np.array(ps, dtype=np.uint16).view(np.int16)).to(d),
         "pw": torch.from_numpy(pw.view(np.int32)).to(d),
         "inv": torch.from_numpy(inv.view(np.int16)).to(d),
         "rp": torch.from_numpy(rp).to(d),
         "Mp": torch.from_numpy(Mp.view(np.int32)).to(d),
         "Mf": torch.from_numpy(Mf.view(np.int32)).to(d)}
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
    _lib.call("b3d_rns_gemm", L, nmod, P(t["primes"]), P(t["pw"]), P(t["inv"]), P(t["rp"]), P(t["Mp"]), P(t["Mf"]),
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
