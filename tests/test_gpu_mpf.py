"""MPF primitives on the GPU (libb3dhip) bit for bit against the oracle's specification (oracle/mpf_ref.py):
normalisation of wide integers (ties included), add, multiply, row alignment."""
import random

import numpy as np
import pytest
import torch

from oracle import mpf_ref as R

pytestmark = pytest.mark.gpu


def rand_pair(rng, L, spread=40):
    if rng.random() < 0.08:
        return 0, R.EZ
    nb = 32 * L - 2
    m = rng.getrandbits(nb - 1) | (1 << (nb - 1))
    if rng.random() < 0.1:                       # powers of two and all-ones mantissas
        m = (1 << (nb - 1)) if rng.random() < 0.5 else (1 << nb) - 1
    return (-m if rng.random() < 0.5 else m), rng.randint(-spread, spread) - nb


@pytest.mark.parametrize("L", [4, 10, 34])
def test_add_mul(L):
    from blocksvibe_amd import mpf
    rng = random.Random(L)
    n = 3000
    a = [rand_pair(rng, L) for _ in range(n)]
    b = [rand_pair(rng, L) for _ in range(n)]
    for i in range(0, n, 7):                      # exact cancellations and near-cancellations
        b[i] = (-a[i][0], a[i][1]) if a[i][0] else b[i]
    for i in range(3, n, 11):
        b[i] = (-a[i][0] + 1, a[i][1]) if a[i][0] else b[i]
    A, B = mpf.from_pairs(a, L, (n,)), mpf.from_pairs(b, L, (n,))
    got_add = mpf.add(A, B).to_host()
    got_mul = mpf.mul(A, B).to_host()
    for i in range(n):
        assert got_add[i] == R.add(a[i], b[i], L), ("add", i, a[i], b[i])
        assert got_mul[i] == R.mul(a[i], b[i], L), ("mul", i)


@pytest.mark.parametrize("L", [4, 10, 34])
def test_norm_wide(L):
    from blocksvibe_amd import mpf
    rng = random.Random(100 + L)
    for LW in (L + 3, 2 * L + 2):
        n = 2000
        ws = []
        for i in range(n):
            bits = rng.randint(1, 32 * LW - 1)
            w = rng.getrandbits(bits)
            if i % 5 == 0 and bits > 32 * L:     # exact ties: ...x100...0 below the kept bits
                s = bits - (32 * L - 2)
                w = (w >> s << s) | (1 << (s - 1))
            if i % 13 == 0:
                w = 0
            ws.append(-w if rng.random() < 0.5 else w)
        es = [rng.randint(-10 ** 6, 10 ** 6) for _ in range(n)]
        W = torch.from_numpy(mpf.to_limb_array(ws, LW).view(np.int32)).cuda()
        E = torch.tensor(es, dtype=torch.int64, device="cuda")
        got = mpf.norm_wide(W, E, L).to_host()
        for i in range(n):
            assert got[i] == R.norm(ws[i], es[i], L), (LW, i)


@pytest.mark.parametrize("L", [4, 10, 34])
def test_align_rows(L):
    from blocksvibe_amd import mpf
    rng = random.Random(200 + L)
    rows, K = 300, 7
    a = [[rand_pair(rng, L, spread=3 * 32 * L if r % 3 == 0 else 40) for _ in range(K)] for r in range(rows)]
    a[5] = [(0, R.EZ)] * K
    A = mpf.from_pairs([x for row in a for x in row], L, (rows, K))
    om, oe = mpf.align_rows(A)
    gm = mpf.to_pairs(om.cpu().numpy(), np.zeros((rows * K,), np.int64))
    ge = oe.cpu().numpy()
    for r in range(rows):
        ms, E = R.align(a[r])
        assert int(ge[r]) == E, r
        for k in range(K):
            assert gm[r * K + k][0] == ms[k], (r, k)
