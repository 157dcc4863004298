"""Pins of the oracle's numerator and coordinate chains by the mathematics they implement (no restated formula):

  oracle.derivs_ref.numerators  (residues_to_derivatives, restore_r_delta, shift_fraction, together_with_factors,
      reference tests/test_blocks_derivatives.py:36-80): with h(Delta, r) = r^Delta [sum_m h_m r^m +
      sum_P (Delta - D_P)^-1 sum_i res_(i,P) r^(i + s_P)] (the truncated radial series with its poles), the e-th
      log r derivative at r_c, divided by r_c^Delta, is in closed form
          G_e(Delta) = sum_m h_m (Delta + m)^e r_c^m + sum_P (Delta - D_P)^-1 sum_i res_(i,P) (Delta + i + s_P)^e r_c^(i+s_P),
      and numerator_e(x) / prod_P (x - x_P) must equal G_e(x + ub) at any rational x off the poles.  With pole
      shifting, residues of the poles beyond the kept ones set to zero: the kept poles map to themselves, so the
      same identity holds with the kept poles alone.
  oracle.output_ref.reference_xt / reference_radial  (lambda_lr_to_lrho_lrhob, the congruence transforms,
      zzb_to_xt, reference tests/test_blocks_output.py:13-47): for a known function h(rho, rhobar) the inputs are
      i^s d_x^e d_theta^s h(e^(x + i theta), e^(x - i theta)) at (log r_c, 0) and the outputs must be the x, t
      derivatives of (z zbar)^(-D/2) h(rho(z), rho(zbar)), z = x + y, zbar = x - y, t = y^2 (divided by y for
      the odd part, whose orientation is y = (zbar - z) / 2), at x = 1/2, t = 0 -- all by mpmath numerical
      differentiation.
"""
import random
from fractions import Fraction

import mpmath as mp
import pytest

from oracle import derivs_ref as DR
from oracle import output_ref as OR
from oracle import structure_ref as OS


def _series_data(tj, tj12, tj43, order, p, a, b, seed, zero_beyond=None):
    """random h_m and residues res_(i,P) for one (spin, parity, structure); returns the oracle's callables and the
    data keyed by pole"""
    rng = random.Random(seed)
    val = lambda: mp.mpf(rng.randint(-999, 999)) / 256
    ji = (tj - tj12 % 2) // 2
    hm = {m: val() for m in range(order + 1)}
    data = {}
    for P in OS.poles(tj, tj12, tj43, order):
        keep = zero_beyond is None or P in OS.poles(tj, tj12, tj43, zero_beyond)
        for i in range(order - OS.shift(P) + 1):
            data[(i, P)] = val() if keep else mp.mpf(0)
    by_index = {(i, OS.position(tj, tj12, tj43, order - i, P)): v for (i, P), v in data.items()}
    hf = lambda pp, m, j, aa, bb: hm[m] if (pp, j, aa, bb) == (p, ji, a, b) else mp.mpf(0)
    rf = lambda pp, i, j, q, aa, bb: by_index.get((i, q), mp.mpf(0)) if (pp, j, aa, bb) == (p, ji, a, b) else mp.mpf(0)
    return hf, rf, hm, data


def _series_derivative(e, Delta, hm, data, tj):
    rc = 3 - 2 * mp.sqrt(2)
    tot = sum(v * (Delta + m) ** e * rc ** m for m, v in hm.items())
    for (i, P), v in data.items():
        s = OS.shift(P)
        DP = mp.mpf(OS.delta_2(P, tj)) / 2
        tot += v * (Delta + i + s) ** e * rc ** (i + s) / (Delta - DP)
    return tot


def _ub(tj):
    return Fraction(tj, 2) + (Fraction(1, 2) if tj <= 1 else 1)


@pytest.mark.parametrize("tj,tj12,tj43,order,R,p", [(0, 2, 2, 6, 3, 0), (2, 2, 2, 6, 3, 1), (4, 4, 2, 7, 2, 0),
                                                    (1, 1, 3, 6, 3, 0), (3, 3, 1, 5, 4, 1), (6, 2, 2, 5, 2, 0)])
def test_numerators_equal_the_truncated_series(tj, tj12, tj43, order, R, p):
    mp.mp.prec = 400
    hf, rf, hm, data = _series_data(tj, tj12, tj43, order, p, 0, 0, seed=tj * 100 + order)
    nums = DR.numerators(hf, rf, tj, p, 0, 0, order, R, tj12, tj43)
    ub = _ub(tj)
    xs = [Fraction(OS.delta_2(P, tj), 2) - ub for P in OS.poles(tj, tj12, tj43, order)]
    for x0 in (Fraction(1, 3), Fraction(-7, 5), Fraction(11, 4)):
        assert x0 not in xs
        X = mp.mpf(x0.numerator) / x0.denominator
        den = mp.mpf(1)
        for x in xs:
            den *= X - mp.mpf(x.numerator) / x.denominator
        for e, poly in enumerate(nums):
            got = sum(c * X ** k for k, c in enumerate(poly)) / den
            want = _series_derivative(e, X + mp.mpf(ub.numerator) / ub.denominator, hm, data, tj)
            assert abs(got - want) <= mp.mpf(2) ** -360 * (1 + abs(want)), (e, x0)


@pytest.mark.parametrize("tj,tj12,tj43,order,kept,R,p", [(2, 2, 2, 8, 4, 3, 0), (4, 4, 2, 9, 5, 2, 1),
                                                         (0, 2, 2, 7, 3, 3, 0)])
def test_shifted_numerators_exact_when_far_residues_vanish(tj, tj12, tj43, order, kept, R, p):
    mp.mp.prec = 400
    hf, rf, hm, data = _series_data(tj, tj12, tj43, order, p, 0, 0, seed=7 + tj, zero_beyond=kept)
    nums = DR.numerators(hf, rf, tj, p, 0, 0, order, R, tj12, tj43, kept=kept)
    ub = _ub(tj)
    all_x = [Fraction(OS.delta_2(P, tj), 2) - ub for P in OS.poles(tj, tj12, tj43, order)]
    kept_x = [Fraction(OS.delta_2(P, tj), 2) - ub for P in OS.poles(tj, tj12, tj43, kept)]
    new = [x for x in all_x if x == 0] + [x for x in kept_x if x != 0]
    assert len(new) < len(all_x)                                       # the shifting path is taken
    for x0 in (Fraction(2, 7), Fraction(-13, 3)):
        X = mp.mpf(x0.numerator) / x0.denominator
        den = mp.mpf(1)
        for x in new:
            den *= X - mp.mpf(x.numerator) / x.denominator
        for e, poly in enumerate(nums):
            got = sum(c * X ** k for k, c in enumerate(poly)) / den
            want = _series_derivative(e, X + mp.mpf(ub.numerator) / ub.denominator, hm, data, tj)
            assert abs(got - want) <= mp.mpf(2) ** -330 * (1 + abs(want)), (e, x0)


def _rho(z):
    return z / (1 + mp.sqrt(1 - z)) ** 2


def _h_even(r1, r2):
    return mp.exp(r1 + r2) * (1 + r1 * r2) + (r1 * r2) ** 2


def _h_odd(r1, r2):
    return (r1 - r2) * mp.cosh(r1 * r2) + r1 ** 3 - r2 ** 3


@pytest.mark.parametrize("parity,D", [(0, "0"), (0, "2.3"), (1, "0"), (1, "-1.25")])
def test_xt_chain_numerical(parity, D):
    mp.mp.dps = 60
    Lam, D = 4, mp.mpf(D)
    h = _h_odd if parity else _h_even
    x0 = mp.log(3 - 2 * mp.sqrt(2))
    O = Lam
    num = [[mp.mpf(0)] * (O + 1) for _ in range(O // 2 + 1)]
    for idx in range((O - parity) // 2 + 1):
        s = 2 * idx + parity
        for e in range(O - s + 1):
            d = mp.diff(lambda x, th: h(mp.exp(x + 1j * th), mp.exp(x - 1j * th)), (x0, 0), (e, s))
            num[idx][e] = mp.re(mp.mpc(0, 1) ** s * d)
    got = OR.reference_xt(num, Lam, parity, D)

    # the odd part's orientation: the chain pairs the first argument of h (rho = r e^(i theta)) with zbar, i.e.
    # y = (zbar - z) / 2 (blocks_3d's convention, which the reference's port carries; the mathematics fixes the
    # chain up to this sign of the parity-1 tables)
    sgn = -1 if parity else 1

    def K(x, y):
        z, zb = x + sgn * y, x - sgn * y
        return (z * zb) ** (-D / 2) * h(_rho(z), _rho(zb))

    for n in range(O - parity + 1):
        for bi, m in enumerate(range(parity, O - n + 1, 2)):
            b = (m - parity) // 2
            want = mp.diff(K, (mp.mpf(0.5), 0), (n, m)) * mp.factorial(b) / mp.factorial(m)
            assert abs(got[n][bi] - want) < mp.mpf(10) ** -25 * (1 + abs(want)), (n, m)


@pytest.mark.parametrize("parity,D", [(0, "0"), (0, "1.7"), (1, "0"), (1, "0.6")])
def test_radial_chain_numerical(parity, D):
    """the radial derivatives: parity 0, d_z^i [z^-D h(rho(z), rho(z))] at z = 1/2 from the log r derivatives of
    h(r, r) at r_c; parity 1, d_x^i d_y K(x, y) at (1/2, 0) (K as in the xt chain, y oriented as there) from the
    log r derivatives of i d_theta h"""
    mp.mp.dps = 60
    Lam, D = 5, mp.mpf(D)
    x0 = mp.log(3 - 2 * mp.sqrt(2))
    h = _h_odd if parity else _h_even
    num0 = [mp.re(mp.mpc(0, 1) ** parity * mp.diff(lambda x, th: h(mp.exp(x + 1j * th), mp.exp(x - 1j * th)),
                                                      (x0, 0), (e, parity)))
            for e in range(Lam - parity + 1)]
    got = OR.reference_radial(num0, Lam, parity, D)
    sgn = -1 if parity else 1

    def K(x, y):
        z, zb = x + sgn * y, x - sgn * y
        return (z * zb) ** (-D / 2) * h(_rho(z), _rho(zb))

    for i in range(Lam - parity + 1):
        want = mp.diff(K, (mp.mpf(0.5), 0), (i, parity))
        assert abs(got[i] - want) < mp.mpf(10) ** -25 * (1 + abs(want)), i
