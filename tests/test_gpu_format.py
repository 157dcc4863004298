"""The JSON writer's text made on the GPU (b3djson.poly_bodies_device: b3d_fmt_sizes, b3d_fmt_text, b3d_fmt_pack)
against the host formatter (b3djson.poly_bodies: b3d_format_polys, itself checked text for text against the
Python formatter in tests/test_format_native.py), byte for byte, on device planes: random magnitudes over a
wide exponent range, integer parts of up to the whole mantissa, values the GPU leaves to the host (exponent
>= 0, fraction beyond its window), exact ties and all-nines carries at the digit count, zero coefficients, all-zero
polynomials and negligible top coefficients, several chunks, L 10..50."""
import random

import numpy as np
import pytest
import torch

from blocksvibe_amd import b3djson as BJ

pytestmark = pytest.mark.gpu

EZ = -(1 << 62)


def planes(vals, L):
    from blocksvibe_amd import mpf
    m = mpf.to_limb_array([v for v, _ in vals], L).view(np.int32).reshape(L, -1)
    e = np.array([x if v else EZ for v, x in vals], dtype=np.int64)
    return m, e


def values(L, n, p, rng):
    nb = 32 * L - 2
    out = []
    for _ in range(n):
        r = rng.random()
        if r < 0.08:
            out.append((0, EZ))
            continue
        mag = rng.getrandbits(nb - 1) | (1 << (nb - 1))
        if r < 0.5:
            ex = -nb + rng.randint(-150, 150)                 # the common range: |x| ~ 2^+-150
        elif r < 0.6:
            ex = -nb + rng.randint(60, nb)                    # integer parts of 60 bits up to the mantissa
        elif r < 0.65:
            ex = rng.randint(0, 300)                          # exponent >= 0 (host)
        elif r < 0.75:
            ex = -nb - rng.randint(400, 700)                  # tiny: near and beyond the fraction window
        elif r < 0.85:                                        # ties and nines at the digit count
            # j >= p - 18: the tie's integer part stays below 2^64 (the GPU's range)
            j = rng.randint(max(1, p - 18), p + 40) if rng.random() < 0.7 else rng.randint(1, 30)
            if rng.random() < 0.3:
                K, j = 2 * 10 ** p - 1, 1                     # 99..9.5
            else:                                             # K 2^-j with digits K 5^j: p + 1 ending in 5
                lo, hi = -(-10 ** p // 5 ** j), (10 ** (p + 1) - 1) // 5 ** j
                K = rng.randint(lo, hi) | 1 if lo <= hi else 1
            if K.bit_length() > nb:
                K, j = 3, 1
            out.append((K * rng.choice((-1, 1)), -j))
            continue
        else:
            ex = -nb + rng.randint(-40, 40)
        out.append((mag * rng.choice((-1, 1)), ex))
    return out


@pytest.mark.parametrize("L,p", [(34, 308), (10, 77), (50, 481), (34, 17), (18, 5), (10, 1)])
def test_device_bodies_match_host(L, p):
    rng = random.Random(L * 1000 + p)
    npoly, D = 700, 13
    vals = values(L, npoly * D, p, rng)
    for q in range(0, npoly, 37):                             # all-zero polynomials
        vals[q * D:(q + 1) * D] = [(0, EZ)] * D
    for q in range(5, npoly, 23):                             # negligible top coefficients
        m0, e0 = vals[q * D]
        if m0:
            vals[q * D + D - 1] = (m0, e0 - 1000 * L)
    m, e = planes(vals, L)
    eps = BJ.trim_exponent(p)
    host = BJ.poly_bodies(m.reshape(L, npoly, D), e.reshape(npoly, D), p, eps)
    dm = torch.from_numpy(m.copy()).cuda().reshape(L, npoly, D)
    de = torch.from_numpy(e.copy()).cuda().reshape(npoly, D)
    print(f"L={L} p={p}: {npoly} polynomials x {D}", flush=True)
    dev = BJ.poly_bodies_device(dm, de, p, eps, chunk=2000)    # several chunks
    assert len(dev) == len(host)
    bad = [i for i, (a, b) in enumerate(zip(dev, host)) if a != b]
    assert not bad, (bad[:3], bytes(dev[bad[0]])[:200], bytes(host[bad[0]])[:200])
    sel = np.array([5, 3, 600, 0], dtype=np.int64)
    got = BJ.poly_bodies_device(dm, de, p, eps, select=torch.from_numpy(sel).cuda())
    assert got == [host[i] for i in sel]
    print(f"identical: {sum(len(x) for x in dev)} bytes", flush=True)
