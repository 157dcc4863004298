"""The oracle's C restatement (oracle/c/mpf_c.c) against mpf_ref, bit for bit: norm, Cauchy sums and
pole-step dot products on random and edge-case inputs.  CPU only."""
import random

import pytest

from oracle import c_ref as C
from oracle import mpf_ref as R


@pytest.fixture(scope="module", autouse=True)
def built():
    from oracle.c import build as B  # noqa: F401
    B.build()


def rnum(rng, L, spread=3, zero=0.0):
    if rng.random() < zero:
        return 0, R.EZ
    nb = R.NB(L)
    m = rng.getrandbits(nb - 1) | (1 << (nb - 1))
    return (-m if rng.random() < 0.5 else m), rng.randint(-spread, spread) - nb


@pytest.mark.parametrize("L", [2, 3, 4, 10, 34])
def test_norm(L):
    rng = random.Random(L)
    n = 2 * L + 2
    for _ in range(300):
        W = rng.getrandbits(rng.randint(1, 32 * n - 2)) * rng.choice((-1, 1))
        e = rng.randint(-500, 500)
        assert C.norm(W, n, e, L) == R.norm(W, e, L)


def test_norm_ties_small():
    L = 2
    NB = R.NB(L)
    for k in range(1, 40):
        for low in range(0, 1 << min(k, 6)):
            W = ((1 << NB) - 1 - low) << k | (1 << (k - 1))     # exact ties and carries into a new bit
            for s in (1, -1):
                assert C.norm(s * W, 2 * L + 2, 0, L) == R.norm(s * W, 0, L)


@pytest.mark.parametrize("L", [2, 4, 10, 34])
def test_cauchy(L):
    rng = random.Random(100 + L)
    for trial in range(200):
        K = rng.randint(1, 64)
        spread = rng.choice([0, 3, 40, 32 * L + 100])
        xs = [rnum(rng, L, spread, zero=0.2) for _ in range(K)]
        ds = [rng.choice([0, 1, -1, 2, (1 << 15) - 1, -((1 << 15) - 1)] + [rng.randint(-300, 300)])
              for _ in range(K)]
        assert C.cauchy(xs, ds, L) == R.cauchy(xs, ds, L), (trial, K, spread)
        h = rnum(rng, L, rng.choice([0, 5, 40, 32 * L + 100]), zero=0.1)
        h = (h[0], h[1] + rng.choice([0, 0, -60, 60]))
        assert C.cauchy(xs, ds, L, h) == R.cauchy(xs, ds, L, h), (trial, K, spread, "h")
    assert C.cauchy([(0, R.EZ)] * 3, [1, 2, 3], L) == (0, R.EZ)
    h = rnum(rng, L)
    assert C.cauchy([(0, R.EZ)] * 3, [1, 2, 3], L, h) == h == R.cauchy([(0, R.EZ)] * 3, [1, 2, 3], L, h)


@pytest.mark.parametrize("L", [2, 4, 10, 34])
def test_dot_aligned(L):
    rng = random.Random(200 + L)
    for trial in range(200):
        K = rng.randint(1, 8)
        a = [rnum(rng, L, rng.choice([0, 5, 70]), zero=0.2) for _ in range(K)]
        b = [rnum(rng, L, rng.choice([0, 5, 70]), zero=0.2) for _ in range(K)]
        am, ae = R.align(a)
        bm, be = R.align(b)
        assert C.dot_aligned(am, ae, bm, be, L) == R.gemm_aligned(am, ae, bm, be, L), trial


def test_timer_runs():
    assert C.time_unit(0, 10, 16, 2, 0.05) > 0
    assert C.time_unit(1, 10, 5, 2, 0.05) > 0
