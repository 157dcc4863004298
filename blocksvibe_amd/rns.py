"""Exact multiprecision GEMM through a residue number system on int8 MFMA (libb3dhip b3d_rns_gemm).

Semantics (the build's own definition of the reference's unvendored sdpvibe.mpf.gemm, reference
blocksvibe/smallgemm.py:1-6 "A aligned per row, B per column, exact products, one rounding per entry"):
C[i, j] = round(sum_k Abar[i, k] Bbar[k, j]) with Abar the row-aligned mantissas of A (exponent eA[i]) and
Bbar the column-aligned mantissas of B (exponent eB[j]).

Tables: the nmod largest primes below 2^15 with product M > 2^(2 NB + ceil(log2 K) + 4); int8 weights of
the byte values 2^(8 j) mod p for the MFMA residue conversion; the CRT constants (M/p)^-1 mod p and the
signed radix-256 digits of M/p and M for the MFMA CRT lift.
"""
import json
import math
import os
import threading

import numpy as np
import torch

from . import _lib, mpf

_PRIMES = None
_TABLES = {}
_TABLE_LOCK = threading.Lock()
WORKSPACE_BYTES = 1 << 31

# MFMA block tiles of b3d_rns_gemm (16 wm rows x 64 / wm columns).  By default (BLOCKSVIBE_RNS_TUNE unset or
# "on") the first GEMM of a shape class times every candidate (they compute the same bits) and keeps the fastest
# for the rest of the process; BLOCKSVIBE_RNS_TUNE=<file> also keeps the choices in that JSON file across
# processes (and only tunes classes it does not hold).  BLOCKSVIBE_RNS_TUNE=off: the tiles of the file
# BLOCKSVIBE_RNS_TILES names if any, else wm = 1.
TILES = (1, 2, 4)
_TUNED = {}
_TUNE_LOCK = threading.Lock()
_TUNE_LOADED = [False]


def _tune_mode():
    return os.environ.get("BLOCKSVIBE_RNS_TUNE", "") or "on"


def _tune_file():
    f = _tune_mode()
    return f if f and f not in ("on", "off") else (os.environ.get("BLOCKSVIBE_RNS_TILES") or None)


def _shape_class(L, nb, M, N, K, shared):
    up = lambda x: 1 << max(0, (int(x) - 1).bit_length())
    return (L, up(nb), up(M), up(N), up(K), int(bool(shared)))


def tuned():
    """the tile chosen per shape class so far"""
    with _TUNE_LOCK:
        return dict(_TUNED)


def _load_tuning():
    f = _tune_file()
    if _TUNE_LOADED[0] or not f:
        return
    _TUNE_LOADED[0] = True
    try:
        for k, v in json.load(open(f)).items():
            _TUNED[tuple(int(x) for x in k.split(","))] = int(v)
    except (OSError, ValueError):
        pass


def _save_tuning():
    f = _tune_file()
    if not f or _tune_mode() == "off":
        return
    tmp = f + ".tmp"
    with open(tmp, "w") as fh:
        json.dump({",".join(str(x) for x in k): v for k, v in _TUNED.items()}, fh)
    os.replace(tmp, f)


def _run_tuned(key, run):
    """run(wm) with the tuned tile of this shape class, tuning it first if it is new (not while a graph is being
    captured: the default tile is used then)"""
    with _TUNE_LOCK:
        _load_tuning()
        wm = _TUNED.get(key)
    if wm is not None or _tune_mode() == "off" or torch.cuda.is_current_stream_capturing():
        run(wm or 1)
        return
    times = {}
    run(TILES[0])                                  # warm-up (code objects, caches); its result is the same
    for t in TILES:
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        run(t)
        e1.record()
        e1.synchronize()
        times[t] = e0.elapsed_time(e1)
    with _TUNE_LOCK:
        _TUNED[key] = min(times, key=times.get)
        _save_tuning()


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
    """the device tables of (L, nmod), built once per process.  Callers on several streams share them: they are
    built under a lock and the building stream is synchronised before they are published"""
    key = (L, nmod)
    t = _TABLES.get(key)
    if t is not None:
        return t
    with _TABLE_LOCK:
        if key not in _TABLES:
            t = _build_tables(L, nmod)
            torch.cuda.current_stream().synchronize()
            _TABLES[key] = t
        return _TABLES[key]


def _build_tables(L, nmod):
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
    return t


def _chunk_cols(nmod, nbA, nb, M, N, K, budget):
    lib = _lib.lib()
    per = lib.b3d_rns_workspace_bytes(nmod, nbA, nb, M, 64, K) - lib.b3d_rns_workspace_bytes(nmod, nbA, nb, M, 0, K)
    fixed = lib.b3d_rns_workspace_bytes(nmod, nbA, nb, M, 0, K)
    nc = max(64, int((budget - fixed) // max(per / 64.0, 1)) // 64 * 64)
    return min(nc, N)


def gemm_strided(L, nb, M, N, K, A, a_strides, a_shared, B, b_strides, C, c_strides, budget=None, tile=None):
    """low-level call: A, B, C are MPF (any shapes) addressed through (batch, row, col) strides; tile forces an
    MFMA block tile (else the tuned one)"""
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

    def run(wm):
        _lib.call("b3d_rns_gemm", L, nmod, P(t["primes"]), P(t["w8"]), P(t["wcorr"]), P(t["inv"]), P(t["rp"]),
                  P(t["c8"]), t["CK"],
                  nb, M, N, K, 1 if a_shared else 0,
                  P(A.m), P(A.e), A.m[0].numel(), *a_strides,
                  P(B.m), P(B.e), B.m[0].numel(), *b_strides,
                  P(C.m), P(C.e), C.m[0].numel(), *c_strides, Nc, wm, P(EB), P(ws), ws_bytes, mpf.stream_ptr())

    if tile is not None:
        run(tile)
    else:
        _run_tuned(_shape_class(L, nb, M, N, K, a_shared), run)
    return C


def gemm_strided_noalloc(L, nb, M, N, K, A, a_strides, a_shared, B, b_strides, C, c_strides, tile=1):
    """gemm_strided through b3d_rns_gemm_noalloc: no workspace and no allocation by the library (its fixed
    per-GPU scratch; B3DError if the product does not fit in it).  The residue tables are built on the first
    use of an (L, nmod) pair, so a caller that must not allocate at all calls once beforehand."""
    if nb == 0 or M == 0 or N == 0:
        return C
    nmod = n_moduli(L, K)
    t = tables(L, nmod)
    P = lambda x: x.data_ptr()
    _lib.call("b3d_rns_gemm_noalloc", L, nmod, P(t["primes"]), P(t["w8"]), P(t["wcorr"]), P(t["inv"]), P(t["rp"]),
              P(t["c8"]), t["CK"], nb, M, N, K, 1 if a_shared else 0,
              P(A.m), P(A.e), A.m[0].numel(), *a_strides,
              P(B.m), P(B.e), B.m[0].numel(), *b_strides,
              P(C.m), P(C.e), C.m[0].numel(), *c_strides, tile, mpf.stream_ptr())
    return C


def gemm(A, B, out=None, budget=None, tile=None):
    """A: MPF (M, K) (shared) or (nb, M, K); B: MPF (K, N) or (nb, K, N) -> MPF (M, N) or (nb, M, N)"""
    A, B = A.contiguous(), B.contiguous()
    L = A.L
    assert B.L == L
    if len(B.shape) == 2:
        assert len(A.shape) == 2
        Mr, K = A.shape
        K2, N = B.shape
        assert K == K2
        C = out or (mpf.empty if K else mpf.zeros)((Mr, N), L, A.e.device)   # written in full when K > 0
        return gemm_strided(L, 1, Mr, N, K, A, (0, K, 1), True, B, (0, N, 1), C, (0, N, 1), budget, tile)
    nb, K, N = B.shape
    if len(A.shape) == 2:
        Mr, K2 = A.shape
        shared, asb = True, 0
    else:
        nbA, Mr, K2 = A.shape
        assert nbA == nb
        shared, asb = False, Mr * K2
    assert K == K2, (A.shape, B.shape)
    C = out or (mpf.empty if K else mpf.zeros)((nb, Mr, N), L, A.e.device)
    return gemm_strided(L, nb, Mr, N, K, A, (asb, K, 1), shared, B, (K * N, N, 1), C, (Mr * N, N, 1), budget, tile)
