"""The arithmetic of the pole step's int8 MFMA path (blocksvibe_amd/csrc/b3d_pole.hip), emulated exactly on the
CPU with numpy: balanced base-256 digits of the constants, the Toeplitz column sums of the two's complement
bytes fed as u ^ 0x80 with the per-column correction, the carry pass in four segments and their assembly.
Checked against the exact integer sum of products and, after rounding, against the oracle's
gemm_aligned (the VALU path's rounding); the kernel's segmented rounding (segment carries rippled up, the
magnitude from per-segment zero flags, the bit length, the window at the row's bit offset) against the
oracle's norm.  CPU only."""
import random

import numpy as np
import pytest

from oracle import mpf_ref as R


def sdigits(a, L):
    """4L + 1 balanced digits of a two's complement L-limb value (k_sdigits)"""
    u = a & ((1 << (32 * L)) - 1)
    by = [(u >> (8 * i)) & 0xFF for i in range(4 * L)]
    d, carry = [], 0
    for i in range(4 * L - 1):
        t = by[i] + carry
        carry = 1 if t >= 128 else 0
        d.append(t - 256 * carry)
    top = (by[4 * L - 1] - 256 if by[4 * L - 1] >= 128 else by[4 * L - 1]) + carry
    d += [-128, 1] if top == 128 else [top, 0]
    assert all(-128 <= x <= 127 for x in d)
    assert sum(x << (8 * i) for i, x in enumerate(d)) == a
    return d


def corrections(ds, L):
    """corr[c] = 128 sum_j sum_{i <= 4L-2} d_j[c - i] (k_scorr)"""
    out = np.zeros(8 * L, dtype=np.int64)
    for d in ds:
        for c in range(8 * L):
            out[c] += 128 * sum(d[c - i] for i in range(max(0, c - 4 * L), min(4 * L - 2, c) + 1))
    return out


def fed_bytes(x, L):
    """the A operand bytes of one aligned entry: two's complement, u ^ 0x80 as int8 except the signed top"""
    u = x & ((1 << (32 * L)) - 1)
    by = [(u >> (8 * i)) & 0xFF for i in range(4 * L)]
    s = [(b ^ 0x80) - 256 if (b ^ 0x80) >= 128 else (b ^ 0x80) for b in by[:-1]]
    s.append(by[-1] - 256 if by[-1] >= 128 else by[-1])
    return s


def column_sums(xs, ds, L):
    C = np.zeros(8 * L, dtype=np.int64)
    for x, d in zip(xs, ds):
        C += np.convolve(np.array(fed_bytes(x, L), dtype=np.int64), np.array(d, dtype=np.int64))[:8 * L]
    return C


def carry_segments(C, corr, L):
    """four segments of 2L columns -> L/2 limbs + a signed carry each, then the assembly of the kernel"""
    segl = L // 2
    limbs, carries = [], []
    for s in range(4):
        carry = 0
        for q in range(segl):
            v = carry
            for b in range(4):
                c = 2 * L * s + 4 * q + b
                v += (int(C[c]) + int(corr[c])) << (8 * b)
            limbs.append(v & 0xFFFFFFFF)
            carry = v >> 32
        carries.append(carry)
    W = sum(x << (32 * k) for k, x in enumerate(limbs)) + (carries[3] << (32 * 2 * L))
    for s in range(3):
        W += carries[s] << (32 * segl * (s + 1))
    return W


def segmented_round(C, corr, L, e):
    """the kernel's rounding of one row's column sums (round_row in b3d_pole.hip), restated"""
    segl, M32 = L // 2, (1 << 32) - 1
    v, carry = [], []
    for s in range(4):
        c, limbs = 0, []
        for q in range(segl):
            x = c + sum((int(C[2 * L * s + 4 * q + b]) + int(corr[2 * L * s + 4 * q + b])) << (8 * b) for b in range(4))
            limbs.append(x & M32)
            c = x >> 32
        v.append(limbs)
        carry.append(c)
    cout = [carry[0]]
    for k in range(1, 4):
        cin = cout[k - 1]
        U = sum(x << (32 * q) for q, x in enumerate(v[k])) + (cin & ((1 << (32 * segl)) - 1) if cin >= 0 else
                                                              (1 << (32 * segl)) + cin)
        cy = U >> (32 * segl)
        v[k] = [(U >> (32 * q)) & M32 for q in range(segl)]
        cout.append(carry[k] + cy - (1 if cin < 0 else 0))
    top = cout[3]
    neg = top < 0
    zero = [all(x == 0 for x in seg) for seg in v]
    if neg:
        for k in range(4):
            cy = 1 if all(zero[:k]) else 0
            U = sum((~x & M32) << (32 * q) for q, x in enumerate(v[k])) + cy
            v[k] = [(U >> (32 * q)) & M32 for q in range(segl)]
        tmag = (~top & ((1 << 64) - 1)) + (1 if all(zero) else 0)
    else:
        tmag = top
    mag = [x for seg in v for x in seg] + [tmag & M32, (tmag >> 32) & M32]
    A = sum(x << (32 * k) for k, x in enumerate(mag))
    bl = A.bit_length()
    if bl == 0:
        return 0, R.EZ
    s = bl - R.NB(L)
    sb = s - 1
    w0, bs = sb >> 5, sb & 31
    get = lambda i: mag[i] if 0 <= i < 2 * L + 2 else 0
    t = [((get(w0 + k + 1) << 32 | get(w0 + k)) >> bs) & M32 for k in range(L)]
    T = sum(x << (32 * k) for k, x in enumerate(t)) + 1
    m = (T >> 1) & ((1 << (32 * L)) - 1)
    if m >> (32 * L - 2):
        m, s = 1 << (32 * L - 3), s + 1
    return (-m if neg else m), e + s


@pytest.mark.parametrize("L", [4, 6, 10, 34])
def test_pole_digits_exact(L):
    rng = random.Random(900 + L)
    nb = R.NB(L)
    for trial in range(40 if L == 34 else 120):
        S = rng.randint(1, 6)
        rnd = lambda: (0, R.EZ) if rng.random() < 0.1 else (
            rng.choice((-1, 1)) * (rng.getrandbits(nb - 1) | (1 << (nb - 1))), rng.randint(-60, 60) - nb)
        arow = [rnd() for _ in range(S)]
        if trial % 7 == 0:                                  # extreme constants: +-(2^NB - 1), -2^(NB-1)
            arow[0] = (rng.choice(((1 << nb) - 1, -((1 << nb) - 1), -(1 << (nb - 1)))), -nb)
        am, ea = R.align(arow)
        xs = [rnd() for _ in range(S)]
        xm, xe = R.align(xs)                                # x aligned over j (truncated magnitudes)
        ds = [sdigits(a, L) for a in am]
        C = column_sums(xm, ds, L)
        assert np.abs(C).max() < 2 ** 31                   # exact int32 MFMA accumulation
        corr = corrections(ds, L)
        assert np.abs(C + corr).max() < 2 ** 31
        W = carry_segments(C, corr, L)
        exact = sum(a * x for a, x in zip(am, xm))
        assert W % (1 << (32 * (2 * L + 2))) == exact % (1 << (32 * (2 * L + 2))), trial
        assert abs(exact) < 1 << (32 * (2 * L + 2) - 1)
        if ea != R.EZ and xe != R.EZ:
            assert R.norm(exact, ea + xe, L) == R.gemm_aligned(am, ea, xm, xe, L), trial
            assert segmented_round(C, corr, L, ea + xe) == R.norm(exact, ea + xe, L), trial


@pytest.mark.parametrize("L", [4, 10, 34])
def test_segmented_round_edges(L):
    """column sums of chosen integers W (base-256 digits, signed top): cancellation to tiny magnitudes,
    negative values whose low segments are zero, powers of two and values just below them"""
    rng = random.Random(950 + L)
    nb = R.NB(L)
    segbits = 32 * (L // 2)
    cases = [1, -1, 2, -2, 3, (1 << nb) - 1, -((1 << nb) - 1), 1 << nb, -(1 << nb), (1 << (2 * nb)) + 1,
             -(1 << (2 * nb)), -(1 << segbits), -(1 << (2 * segbits)), -(3 << (3 * segbits)), 1 << (3 * segbits),
             (1 << (3 * segbits)) - 1, -((1 << (2 * nb + 3)) - 1)]
    cases += [rng.choice((-1, 1)) * rng.getrandbits(rng.randint(1, 2 * nb + 3)) for _ in range(200)]
    for W in cases:
        assert abs(W) < 1 << (64 * L + 60)
        u = W % (1 << (64 * L))                        # low 8L digits; the rest goes to the top column
        C = np.array([(u >> (8 * c)) & 0xFF for c in range(8 * L)], dtype=np.int64)
        C[8 * L - 1] += (W - u) >> (8 * (8 * L - 1))   # top column carries the signed remainder
        got = segmented_round(C, np.zeros(8 * L, dtype=np.int64), L, 77)
        assert got == R.norm(W, 77, L), W
