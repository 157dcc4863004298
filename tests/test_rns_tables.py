"""The RNS tables' arithmetic, emulated in numpy with the kernels' integer widths (CPU): the MFMA residue
conversion (bytes ^ 0x80 as int8 against the int8 weight digits, int32 sums, + wcorr) gives x mod p for
every prime, and the digit bounds that keep every int32 MFMA sum exact hold."""
import random

import numpy as np
import pytest

from blocksvibe_amd import rns


@pytest.mark.parametrize("L", [4, 10, 34, 50])
def test_conversion_weights(L):
    K = 300
    nmod = rns.n_moduli(L, K)
    ps = rns.primes_desc()[:nmod]
    w8, wcorr = rns.conversion_weights(ps, L)
    KB = w8.shape[2]
    assert KB % 64 == 0 and w8.shape[1] % 16 == 0
    assert np.abs(w8[0]).max() <= 64 and w8[1].min() >= -128 and w8[1].max() <= 127
    rng = random.Random(L)
    xs = [0, 1, (1 << (32 * L - 2)) - 1, (1 << (32 * L)) - 1] + [rng.getrandbits(32 * L - 2) for _ in range(20)]
    for x in xs:
        b = np.zeros(KB, dtype=np.int64)
        b[:4 * L] = [((x >> (8 * j)) & 255) ^ 0x80 for j in range(4 * L)]
        b[:4 * L] = np.where(b[:4 * L] > 127, b[:4 * L] - 256, b[:4 * L])     # the kernel's signed bytes
        yh = (w8[0].astype(np.int64) @ b)[:nmod]
        yl = (w8[1].astype(np.int64) @ b)[:nmod]
        assert np.abs(yh).max() < 2 ** 31 and np.abs(yl).max() < 2 ** 31
        y = 256 * yh + yl + wcorr[:nmod]
        assert np.abs(y).max() < 2 ** 31
        assert [int(v) % p for v, p in zip(y, ps)] == [x % p for p in ps]


@pytest.mark.parametrize("L", [4, 10, 34, 50])
def test_crt_digits(L):
    """the MFMA CRT lift, emulated: balanced y digits and -q against c8, int32 column sums, four columns per
    limb, carry propagation -> the signed product, for |C| up to M / 16"""
    K = 1 << 16
    nmod = rns.n_moduli(L, K)
    ps = rns.primes_desc()[:nmod]
    c8 = rns.crt_digits(ps, L)
    CN, CK = c8.shape
    LW = 2 * L + 2
    M = 1
    for p in ps:
        M *= p
    inv = [pow((M // p) % p, -1, p) for p in ps]
    rng = random.Random(L)
    cs = [0, 1, -1, M // 16 - 1, -(M // 16 - 1)] + [rng.randrange(-(M // 16), M // 16) for _ in range(30)]
    for C in cs:
        y = [(C % p) * iv % p for p, iv in zip(ps, inv)]
        yb = [v - p if v > p // 2 else v for v, p in zip(y, ps)]
        q = round(sum(v / p for v, p in zip(yb, ps)))
        assert abs(q) <= 127
        col = np.zeros(CK, dtype=np.int64)
        for t, v in enumerate(yb):
            h = (v + 128) >> 8
            col[2 * t], col[2 * t + 1] = h, v - 256 * h
            assert -128 <= col[2 * t + 1] <= 127 and abs(h) <= 64
        col[2 * nmod] = -q
        S = c8.astype(np.int64) @ col                                   # byte-column sums
        assert np.abs(S).max() < 2 ** 31
        V = [int(S[4 * k]) + int(S[4 * k + 1]) * 256 + int(S[4 * k + 2]) * 65536 + int(S[4 * k + 3]) * 16777216
             for k in range(CN // 4)]
        assert max(abs(v) for v in V) < 2 ** 62
        limbs, carry = [], 0
        for k in range(LW):
            v = V[k] + carry
            limbs.append(v & 0xFFFFFFFF)
            carry = v >> 32
        val = sum(x << (32 * k) for k, x in enumerate(limbs))
        if val >> (32 * LW - 1):
            val -= 1 << (32 * LW)
        assert val == C
