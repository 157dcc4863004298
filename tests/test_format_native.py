"""The C-ABI library's host formatter of the JSON derivative tables (b3d_format_numbers, b3d_poly_sizes,
b3d_format_polys in blocksvibe_amd/csrc/b3d_format.hip; the last reached through b3djson.poly_bodies), text
for text against the Python
formatter fmt_exact (itself checked against the decimal module in tests/test_json_cpu.py) and the truncation
rule trimmed_length (reference blocksvibe/b3djson.py:160-168), on values read from two's complement limb
planes as the kernels leave them.  Edge cases: ties at every rounding position (exact binary values ending
in 5 right after the last kept digit), carries into a new leading digit (...9995 -> 1000...), the
fixed/scientific switch at X = -4 and X = p, huge and tiny exponents, the extreme mantissas of L limbs,
unnormalised and zero entries, threads.  Host code only: runs without a GPU."""
import random

import numpy as np
import pytest

from blocksvibe_amd import _lib
from blocksvibe_amd import b3djson as BJ
from blocksvibe_amd import mpf

EZ = mpf.EZ


def planes(vals, L):
    """exact (m, e) pairs -> (L, n) int32 limb planes and (n,) exponents, zeros canonical"""
    m = mpf.to_limb_array([v for v, _ in vals], L).view(np.int32).reshape(L, -1)
    e = np.array([x if v else EZ for v, x in vals], dtype=np.int64)
    return m, e


def native(vals, L, p, threads=0):
    """b3d_format_numbers on the values"""
    m, e = planes(vals, L)
    n = len(vals)
    cap = n * (p + 32)
    buf = np.empty(cap, dtype=np.uint8)
    end = np.empty(n, dtype=np.int64)
    _lib.call("b3d_format_numbers", L, m.ctypes.data, e.ctypes.data, n, n, p, buf.ctypes.data, cap, end.ctypes.data,
              threads)
    text = buf[:int(end[-1])].tobytes().decode()
    b = [0] + end.tolist()
    return [text[x:y] for x, y in zip(b, b[1:])]


def body(poly, p, eps):
    """the Python writer's list body of one polynomial of exact (m, e) pairs"""
    if all(v == 0 for v, _ in poly):
        return "\"0\""
    return ",\n".join(f"\"{BJ.fmt_exact(v, x, p) if v else '0'}\"" for v, x in poly[:BJ.trimmed_length(poly, eps)])


def ties(p, rng):
    """exact binary values K 2^-j (K odd) whose decimal digits N = K 5^j are p + 1 significant digits ending
    in 5: a tie at p digits; 10^(p+1) - 5 (all nines, so the rounding carries into a new digit); and the same
    values nudged up by 2^-(j+40) (just above the tie); both signs"""
    out = [(2 * 10 ** p - 1, -1)]                       # (2 10^p - 1) / 2 = 99..9.5
    for _ in range(40):
        j = rng.randint(1, 40)
        lo, hi = -(-10 ** p // 5 ** j), (10 ** (p + 1) - 1) // 5 ** j
        if lo > hi:
            continue
        K = rng.randint(lo, hi) | 1
        if K > hi:
            continue
        out.append((K, -j))
    out += [((K << 40) + 1, j - 40) for K, j in out]
    return [(K * rng.choice((-1, 1)), j) for K, j in out]


@pytest.mark.parametrize("L", [4, 10, 18, 34, 50])
def test_format_matches_python(L):
    rng = random.Random(40 + L)
    nb = 32 * L - 2
    ps = sorted({1, 2, 3, 5, 17, BJ.digits10(32 * L - 64) if L > 2 else 9, BJ.digits10(32 * L)})
    vals = [(0, EZ), (1, 0), (-1, 0), (1, -1), (-(1 << (32 * L - 1)), 0), ((1 << (32 * L - 1)) - 1, -(32 * L)),
            (1, 3000), (-3, -5000), (5, -1), (25, -2), (125, -3), (1, -14), (1, -13)]   # 2^-14 = 6.1e-05
    for _ in range(300):
        bits = rng.choice((nb, nb, rng.randint(1, 32 * L - 1)))
        m = rng.getrandbits(bits) | (1 << (bits - 1))
        vals.append((m * rng.choice((-1, 1)), rng.choice((-nb, -nb - 30, -nb + 30, rng.randint(-3000, 2000)))))
    for p in ps:
        cur = vals + [v for v in ties(p, rng) if abs(v[0]) < 1 << (32 * L - 1)]
        got = native(cur, L, p)
        want = [BJ.fmt_exact(v, x, p) if v else "0" for v, x in cur]
        bad = [(v, g, w) for v, g, w in zip(cur, got, want) if g != w]
        assert not bad, (L, p, bad[:3])


def test_switch_points_and_layout():
    L = 4
    cases = [(10 ** 4 - 1, 0, 5, "9999"), (10 ** 5, 0, 5, "1e+05"), (99995, -1, 4, "5e+04"),
             (19999, -1, 4, "1e+04"),                   # 9999.5: tie, 9999 odd -> 10000, X = p -> scientific
             (3, -1, 1, "2"), (5, -1, 1, "2"), (7, -1, 1, "4"),          # 1.5, 2.5, 3.5 to one digit
             (1, -13, 5, "0.00012207"), (1, -14, 5, "6.1035e-05"),       # X = -4 fixed, X = -5 scientific
             (1234, -23, 5, "0.0001471"), (1, -17, 5, "7.6294e-06"), (1234, 0, 3, "1.23e+03"),
             (100, 0, 5, "100"), (5, -1, 5, "2.5"), (-3, 200, 4, "-4.821e+60"), (1, -400, 3, "3.87e-121")]
    for m, e, p, want in cases:
        assert native([(m, e)], L, p) == [want] == [BJ.fmt_exact(m, e, p)], (m, e, p)


def test_threads_and_order():
    L = 10
    rng = random.Random(3)
    vals = [(rng.getrandbits(310) * rng.choice((-1, 1)), rng.randint(-400, 100)) for _ in range(3000)]
    one = native(vals, L, 96, threads=1)
    assert native(vals, L, 96, threads=7) == one
    assert one == [BJ.fmt_exact(v, x, 96) if v else "0" for v, x in vals]


@pytest.mark.parametrize("L", [4, 34])
def test_poly_sizes_match_trimmed_length(L):
    rng = random.Random(70 + L)
    nb = 32 * L - 2
    p = BJ.digits10(32 * L)
    eps = BJ.trim_exponent(p)
    polys = []
    for _ in range(400):
        D = rng.randint(1, 9)
        base = rng.randint(-200, 200)
        poly = []
        for _ in range(D):
            r = rng.random()
            if r < 0.25:
                poly.append((0, EZ))
            else:
                m = rng.getrandbits(nb) | (1 << (nb - 1))
                # exponents around the threshold: a later coefficient 2^eps times a lower one, exactly or not
                x = base + rng.choice((0, eps, eps + 1, eps - 1, rng.randint(2 * eps, 10)))
                poly.append((m * rng.choice((-1, 1)), x - nb))
        if rng.random() < 0.2 and len(poly) > 1:      # exact 2^eps ratio: max|later| == 2^eps |c| keeps c out
            m0 = 1 << (nb - 1)
            poly[0] = (m0, base - nb)
            poly[-1] = (-m0, base + eps - nb)
        polys.append(poly)
    D = max(len(q) for q in polys)
    padded = [q + [(0, EZ)] * (D - len(q)) for q in polys]
    m, e = planes([c for q in padded for c in q], L)
    sizes = np.empty(len(polys), dtype=np.int32)
    _lib.call("b3d_poly_sizes", L, m.ctypes.data, e.ctypes.data, e.size, len(polys), D, eps, sizes.ctypes.data)
    want = [0 if all(v == 0 for v, _ in q) else BJ.trimmed_length(q, eps) for q in padded]
    assert sizes.tolist() == want
    for threads in (1, 5):
        got = BJ.poly_bodies(m.reshape(L, len(polys), D), e.reshape(len(polys), D), p, eps, threads=threads)
        assert got == [body(q, p, eps).encode() for q in padded]


def test_derivs_text_same_from_bodies():
    L, p = 10, 96
    eps = BJ.trim_exponent(p)
    rng = random.Random(11)
    rows = [[[(rng.getrandbits(300) * rng.choice((-1, 1)), rng.randint(-330, -290)) if rng.random() < 0.8
              else (0, EZ) for _ in range(4)] for _ in range(3)] for _ in range(2)]
    rows[1][2] = [(0, EZ)] * 4
    flat = [c for r in rows for q in r for c in q]
    m, e = planes(flat, L)
    bodies = BJ.poly_bodies(m.reshape(L, 6, 4), e.reshape(6, 4), p, eps)
    assert BJ.derivs_text([bodies[0:3], bodies[3:6]], p, eps) == BJ.derivs_text(rows, p, eps)


def test_select_and_errors():
    L, p = 4, 20
    vals = [(3, -1), (5, -2), (0, EZ), (7, 0)]
    m, e = planes(vals, L)
    got = BJ.poly_bodies(m.reshape(L, 4, 1), e.reshape(4, 1), p, -900, select=np.array([3, 0, 2]))
    assert got == [b'"7"', b'"1.5"', b'"0"']
    buf = np.empty(4, dtype=np.uint8)
    end = np.empty(1, dtype=np.int64)
    with pytest.raises(_lib.B3DError, match="too small"):
        _lib.call("b3d_format_numbers", L, m.ctypes.data, e.ctypes.data, 4, 1, 30, buf.ctypes.data, 2,
                  end.ctypes.data, 1)
    with pytest.raises(_lib.B3DError, match="too small"):
        _lib.call("b3d_format_polys", L, m.ctypes.data, e.ctypes.data, 4, 2, 2, 30, -900, buf.ctypes.data, 3,
                  end.ctypes.data, 1)
