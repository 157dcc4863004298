"""The allocation-free residue GEMM (b3d_rns_gemm_noalloc, fixed per-GPU scratch) against b3d_rns_gemm with a
caller workspace, bit for bit: two calls in a row on one stream, the first wide enough to need several column
chunks inside the scratch, the second smaller (the scratch keeps the first call's contents); a spot check of
entries against the oracle's exact big-integer product; and the host-side rejection of a product whose column
exponents alone exceed the scratch (no launch: the entry point returns before touching any pointer)."""
import ctypes
import random

import pytest

from oracle import mpf_ref as R

pytestmark = pytest.mark.gpu


def rand_mat(rng, L, nb, M, N):
    from blocksvibe_amd import mpf
    nbits = 32 * L - 2
    vals = []
    for _ in range(nb * M * N):
        if rng.random() < 0.03:
            vals.append((0, R.EZ))
            continue
        m = rng.getrandbits(nbits - 1) | (1 << (nbits - 1))
        vals.append((-m if rng.random() < 0.5 else m, rng.randint(-50, 50) - nbits))
    return vals, mpf.from_pairs(vals, L, (nb, M, N))


def run_pair(L, nb, M, K, N, shared, seed):
    """(noalloc result, workspace result, inputs) of one product"""
    from blocksvibe_amd import mpf, rns
    rng = random.Random(seed)
    a_vals, A = rand_mat(rng, L, 1 if shared else nb, M, K)
    b_vals, B = rand_mat(rng, L, nb, K, N)
    asb = 0 if shared else M * K
    args = (L, nb, M, N, K, A, (asb, K, 1), shared, B, (K * N, N, 1))
    C1 = mpf.zeros((nb, M, N), L, A.e.device)
    C2 = mpf.zeros((nb, M, N), L, A.e.device)
    rns.gemm_strided_noalloc(*args, C1, (M * N, N, 1), tile=1)
    rns.gemm_strided(*args, C2, (M * N, N, 1), tile=1)
    return C1.to_host(), C2.to_host(), a_vals, b_vals


def test_noalloc_matches_workspace_gemm_twice():
    from blocksvibe_amd import rns
    # the first call: 4 x 12 x 2000 at K = 61, L = 34 -- its column exponents and workspace exceed the
    # 64 MiB scratch in one chunk, so the entry point runs several column chunks
    L, nb, M, K, N = 34, 4, 12, 61, 2000
    nmod = rns.n_moduli(L, K)
    lib = rns._lib.lib()
    one_chunk = 8 * nb * N + lib.b3d_rns_workspace_bytes(nmod, nb, nb, M, N, K)
    assert one_chunk > (64 << 20), one_chunk
    print(f"noalloc call 1: L={L} nb={nb} M={M} K={K} N={N} nmod={nmod}", flush=True)
    got, want, a_vals, b_vals = run_pair(L, nb, M, K, N, False, seed=11)
    assert got == want
    rng = random.Random(5)
    for _ in range(12):                                        # entries against the exact product
        b, i, j = rng.randrange(nb), rng.randrange(M), rng.randrange(N)
        row = [a_vals[(b * M + i) * K + k] for k in range(K)]
        col = [[b_vals[(b * K + k) * N + j]] for k in range(K)]
        assert got[(b * M + i) * N + j] == R.gemm([row], col, L)[0][0], (b, i, j)
    # the second call, smaller, in the same scratch (shared A, another L and K)
    print("noalloc call 2: L=18 nb=2 M=5 K=24 N=100 shared", flush=True)
    got2, want2, _, _ = run_pair(18, 2, 5, 24, 100, True, seed=12)
    assert got2 == want2


def test_noalloc_rejects_oversized_product():
    from blocksvibe_amd import _lib, rns
    L, K = 10, 12
    nmod = rns.n_moduli(L, K)
    t = rns.tables(L, nmod)
    null = ctypes.c_void_p(0)
    # nb * N * 8 bytes of column exponents alone exceed the scratch: rejected before any pointer is used
    st = _lib.lib().b3d_rns_gemm_noalloc(L, nmod, t["primes"].data_ptr(), t["w8"].data_ptr(), t["wcorr"].data_ptr(),
                                        t["inv"].data_ptr(), t["rp"].data_ptr(), t["c8"].data_ptr(), t["CK"],
                                        4, 8, 1 << 22, K, 1, null, null, 1, 0, K, 1, null, null, 1, 1, 1, 1,
                                        null, null, 1, 1, 1, 1, 1, null)
    assert st == 2
    assert "fixed scratch" in _lib.lib().b3d_last_error().decode()
