"""The exact decimal formatter of the JSON writer (blocksvibe_amd.b3djson.fmt_exact): C's %.{p}g of the exact
binary value (round half to even, trailing zeros removed, exponent rules of the general format), checked
against Python's decimal module at high precision and against float formatting; blocks_3d's trailing
coefficient truncation (reference blocksvibe/b3djson.py:159-168)."""
import decimal
import random

from blocksvibe_amd import b3djson as BJ


def dec_digits(m, e, p):
    """(sign, significant digits without trailing zeros, decimal exponent X) of m 2^e rounded to p digits,
    half to even, by the decimal module"""
    with decimal.localcontext(decimal.Context(prec=4000, Emax=10 ** 6, Emin=-10 ** 6)):
        exact = decimal.Decimal(m) * (decimal.Decimal(2) ** e)
    r = decimal.Context(prec=p, rounding=decimal.ROUND_HALF_EVEN, Emax=10 ** 6, Emin=-10 ** 6).plus(exact)
    sign, digs, exp = r.as_tuple()
    d = "".join(map(str, digs))
    return sign, d.rstrip("0") or "0", exp + len(d) - 1


def parse_g(s):
    sign = s.startswith("-")
    s = s.lstrip("-")
    if "e" in s:
        mant, ex = s.split("e")
        X = int(ex)
        d = mant.replace(".", "")
    else:
        ip, _, fp = s.partition(".")
        if ip != "0":
            X = len(ip) - 1
            d = ip + fp
        else:
            z = len(fp) - len(fp.lstrip("0"))
            X = -z - 1
            d = fp.lstrip("0")
    return int(sign), d.rstrip("0") or "0", X


def test_against_decimal():
    rng = random.Random(5)
    for p in (5, 17, 77, 308):
        for _ in range(300):
            m = rng.getrandbits(rng.randint(1, 1100)) * rng.choice((-1, 1))
            e = rng.randint(-1400, 300)
            if m == 0:
                continue
            out = BJ.fmt_exact(m, e, p)
            sign, d, X = parse_g(out)
            assert (sign, d, X) == dec_digits(m, e, p), (m, e, p)
            assert ("e" in out) == (X < -4 or X >= p)


def test_c_layout():
    # printf("%.5g") / ("%.3g") layouts of C
    assert BJ.fmt_exact(1234, -23, 5) == "0.0001471"                      # 1234 / 2^23 = 0.000147104...
    assert BJ.fmt_exact(1, -17, 5) == "7.6294e-06"
    assert BJ.fmt_exact(1234, 0, 3) == "1.23e+03"
    assert BJ.fmt_exact(100, 0, 5) == "100"
    assert BJ.fmt_exact(5, -1, 5) == "2.5"
    assert BJ.fmt_exact(-3, 200, 4) == "-4.821e+60"


def test_against_float_repr():
    for x in (1.5, -0.000123456789, 123456.0, 1e-5, 2.5e17, 0.1, 3.0):
        m, e = float(x).as_integer_ratio()
        # value m / 2^k
        k = e.bit_length() - 1
        assert BJ.fmt_exact(m, -k, 12) == BJ.fmt_double(x, 12), x


def test_digits10_and_truncation():
    assert BJ.digits10(256) == 77 and BJ.digits10(1024) == 308 and BJ.digits10(512) == 154
    # top coefficients negligible against a lower one are dropped
    poly = [(1 << 20, 0), (3, -2000), (5, -2100)]
    assert BJ.trimmed_length(poly, -900) == 1
    poly = [(1, 0), (1, -10)]
    assert BJ.trimmed_length(poly, -900) == 2
