"""Pins of the bit-level number specification (oracle.mpf_ref) against exact rational arithmetic: each
operation's result is the exact value up to its stated error bound, norm() is round-to-nearest with ties
away from zero (checked against both neighbours), and conversions round-trip."""
import random
from fractions import Fraction as F

from oracle import mpf_ref as R


def ulp(m, e):
    return F(2) ** e


def test_norm_nearest_ties_away():
    rng = random.Random(1)
    for L in (1, 2, 4):
        nb = R.NB(L)
        for _ in range(3000):
            W = rng.getrandbits(rng.randint(1, 3 * 32 * L)) * rng.choice((-1, 1))
            if rng.random() < 0.2 and abs(W).bit_length() > nb + 1:     # exact tie
                s = abs(W).bit_length() - nb
                W = (abs(W) >> s << s | (1 << (s - 1))) * (1 if W > 0 else -1)
            e0 = rng.randint(-100, 100)
            m, e = R.norm(W, e0, L)
            if W == 0:
                assert (m, e) == (0, R.EZ)
                continue
            assert abs(m).bit_length() == nb
            x, y = F(W) * F(2) ** e0, R.value(m, e)
            err = abs(x - y)
            assert err <= ulp(m, e) / 2
            if err == ulp(m, e) / 2:                                  # ties: away from zero
                assert abs(y) > abs(x)


def test_from_fraction_roundtrip():
    rng = random.Random(2)
    for L in (2, 10):
        for _ in range(500):
            x = F(rng.randint(-10 ** 30, 10 ** 30), rng.randint(1, 10 ** 20))
            m, e = R.from_fraction(x, L)
            if x == 0:
                continue
            assert abs(x - R.value(m, e)) <= F(2) ** e / 2
            assert R.from_fraction(R.value(m, e), L) == (m, e)
            assert R.from_limbs(R.to_limbs(m, L)) == m


def rp(rng, L, spread):
    nb = R.NB(L)
    m = rng.getrandbits(nb - 1) | (1 << (nb - 1))
    return m * rng.choice((-1, 1)), rng.randint(-spread, spread) - nb


def test_add_mul_error_bounds():
    rng = random.Random(3)
    for L in (1, 2, 10):
        for _ in range(2000):
            a, b = rp(rng, L, 80), rp(rng, L, 80)
            x = R.value(*a) + R.value(*b)
            m, e = R.add(a, b, L)
            tt = max(a[1], b[1]) - 32
            assert abs(R.value(m, e) - x) <= (F(2) ** e / 2 if m else 0) + F(2) ** tt
            m, e = R.mul(a, b, L)
            assert abs(R.value(m, e) - R.value(*a) * R.value(*b)) <= F(2) ** e / 2


def test_gemm_and_cauchy_bounds():
    rng = random.Random(4)
    L, K = 2, 7
    for _ in range(200):
        A = [[rp(rng, L, 30) for _ in range(K)] for _ in range(3)]
        B = [[rp(rng, L, 30) for _ in range(2)] for _ in range(K)]
        C = R.gemm(A, B, L)
        for i in range(3):
            _, ea = R.align(A[i])
            for j in range(2):
                _, eb = R.align([B[k][j] for k in range(K)])
                exact = sum(R.value(*A[i][k]) * R.value(*B[k][j]) for k in range(K))
                amax = max(abs(R.value(*A[i][k])) for k in range(K))
                bmax = max(abs(R.value(*B[k][j])) for k in range(K))
                bound = K * (F(2) ** ea * bmax + F(2) ** eb * amax) + F(2) ** C[i][j][1]
                assert abs(R.value(*C[i][j]) - exact) <= bound
        xs = [rp(rng, L, 30) for _ in range(K)]
        ds = [rng.choice((-1, 1)) * rng.randint(0, 50) for _ in range(K)]
        m, e = R.cauchy(xs, ds, L)
        exact = sum(R.value(*x) * F(2, d) for x, d in zip(xs, ds) if d)
        _, E = R.align(xs)
        assert abs(R.value(m, e) - exact) <= 3 * K * F(2) ** E + (F(2) ** e if m else 0)
        # with h in the same sum: every term and h truncated on the grid G, one rounding
        h = rp(rng, L, 60)
        m, e = R.cauchy(xs, ds, L, h)
        G = max(E - 31, h[1] - 32)
        assert abs(R.value(m, e) - (exact + R.value(*h))) <= (K + 1) * F(2) ** G + (F(2) ** e / 2 if m else 0)


def test_mom_flush_exact_when_no_truncation():
    # products that land exactly on the grid are summed exactly
    L = 2
    ys = [(1 << (R.NB(L) - 1), -R.NB(L)), (-(1 << (R.NB(L) - 1)), -R.NB(L) - 3), (0, R.EZ)]
    nks = [5, 9, 7]
    m, e = R.mom_flush(ys, nks, (0, R.EZ), L)
    assert R.value(m, e) == F(1, 2) * 5 - F(1, 16) * 9
