"""Residue-number-system GEMM on int8 MFMA (blocksvibe_amd.rns) bit for bit against the CPU big-integer
product of the aligned mantissas (oracle.mpf_ref.gemm): batched and shared-A forms, K not a multiple of
the MFMA depth, zero rows/columns, wide exponent spreads, several column chunks, L from 4 to 50."""
import random

import pytest

from oracle import mpf_ref as R

pytestmark = pytest.mark.gpu


def rand_pair(rng, L, spread):
    if rng.random() < 0.05:
        return 0, R.EZ
    nb = 32 * L - 2
    m = rng.getrandbits(nb - 1) | (1 << (nb - 1))
    if rng.random() < 0.05:
        m = (1 << nb) - 1
    return (-m if rng.random() < 0.5 else m), rng.randint(-spread, spread) - nb


def mat(rng, L, r, c, spread):
    return [[rand_pair(rng, L, spread) for _ in range(c)] for _ in range(r)]


@pytest.mark.parametrize("L,nb,M,K,N,shared,budget", [
    (4, 3, 5, 70, 100, False, None),
    (10, 2, 17, 1, 33, True, None),
    (10, 4, 3, 130, 200, False, 1 << 20),       # several column chunks
    (34, 2, 12, 61, 150, True, None),
    (34, 3, 6, 136, 80, False, None),
    (50, 1, 4, 40, 70, False, None),
    (34, 5, 12, 12, 300, False, None),           # the derivative stages' K = R + 1
    (18, 2, 9, 24, 64, True, None),
    (26, 1, 16, 17, 20, True, None),
])
def test_rns_gemm_bit_exact(L, nb, M, K, N, shared, budget):
    from blocksvibe_amd import mpf, rns
    rng = random.Random(L * 1000 + K)
    spread = 60
    As = [mat(rng, L, M, K, spread)] if shared else [mat(rng, L, M, K, spread) for _ in range(nb)]
    Bs = [mat(rng, L, K, N, spread) for _ in range(nb)]
    Bs[0][min(2, K - 1)] = [(0, R.EZ)] * N                     # a zero row of B
    for k in range(K):                                          # a zero column of B
        Bs[-1][k][N // 2] = (0, R.EZ)
    As[0][M - 1] = [(0, R.EZ)] * K                              # a zero row of A
    flatA = [x for a in As for row in a for x in row]
    flatB = [x for b in Bs for row in b for x in row]
    A = mpf.from_pairs(flatA, L, (M, K) if shared else (nb, M, K))
    B = mpf.from_pairs(flatB, L, (nb, K, N))
    print(f"L={L} nb={nb} M={M} K={K} N={N} shared={shared} nmod={rns.n_moduli(L, K)}", flush=True)
    C = rns.gemm(A, B, budget=budget).to_host()                 # tuned tile (every candidate ran once)
    for b in range(nb):
        ref = R.gemm(As[0] if shared else As[b], Bs[b], L)
        for i in range(M):
            for j in range(N):
                assert C[(b * M + i) * N + j] == ref[i][j], (b, i, j)
    for tile in rns.TILES:                                      # every MFMA block tile gives the same bits
        assert rns.gemm(A, B, budget=budget, tile=tile).to_host() == C, tile


def test_tile_tuner(monkeypatch, tmp_path):
    """BLOCKSVIBE_RNS_TUNE=<file>: the first GEMM of a shape class times every tile, keeps the fastest, writes it to
    the file; the result is the same bits as the default tile"""
    from blocksvibe_amd import mpf, rns
    L, nb, M, K, N = 10, 3, 20, 24, 300
    rng = random.Random(5)
    A = mpf.from_pairs([x for _ in range(nb) for row in mat(rng, L, M, K, 20) for x in row], L, (nb, M, K))
    B = mpf.from_pairs([x for _ in range(nb) for row in mat(rng, L, K, N, 20) for x in row], L, (nb, K, N))
    want = rns.gemm(A, B, tile=1).to_host()
    f = tmp_path / "tiles.json"
    monkeypatch.setenv("BLOCKSVIBE_RNS_TUNE", str(f))
    got = rns.gemm(A, B).to_host()
    key = rns._shape_class(L, nb, M, N, K, False)
    assert got == want and key in rns.tuned() and rns.tuned()[key] in rns.TILES
    assert f.exists() and ",".join(str(x) for x in key) in f.read_text()
