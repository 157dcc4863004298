"""Scalar pole/spin bookkeeping of blocks_3d's radial recursion.  TEST INFRASTRUCTURE (oracle/__init__.py).

Restates reference blocksvibe/structure.py:16-83 (a port of blocks_3d's Pole.hxx, Pole_Types.cxx,
two_j_values.cxx, so3_tensor_parity.cxx, fill_residues.cxx).  Spins doubled; pole = (type 0..3, k >= 1).
Pinned by tests/test_golden.py (every function against vectors written from the reference's structure.py
itself, tests/golden/structure.json) and tests/test_oracle_structure.py: the published task / slot counts
(reference docs/BLOCKS_PLAN.md:116-123,209-210), and closed forms of the pole positions.
"""
from fractions import Fraction
from functools import lru_cache


def count(t, tj, tj12, tj43, order):
    """number of poles of type t (reference structure.py:16-24)"""
    even = tj % 2 == 0
    if t == 0:
        return order
    if t == 1:
        return min(tj // 2, order)
    if t == 2:
        return order // 2 if even else min(tj // 2, tj12 // 2, tj43 // 2, order // 2)
    return min(tj // 2, tj12 // 2, tj43 // 2, (order + 1) // 2) if even else (order + 1) // 2


@lru_cache(maxsize=None)
def poles(tj, tj12, tj43, order):
    return tuple((t, k) for t in range(4) for k in range(1, count(t, tj, tj12, tj43, order) + 1))


def position(tj, tj12, tj43, order, pole):
    """index of the pole in poles(tj, tj12, tj43, order); the counts of the lower types are summed as they are,
    negative ones included (reference structure.py:31-33; the recursion asks for positions at orders where a
    count can be negative, reference recursion_ref.py:27)"""
    t, k = pole
    return k - 1 + sum(count(u, tj, tj12, tj43, order) for u in range(t))


def shift(pole):
    t, k = pole
    return k if t < 2 else (2 * k if t == 2 else 2 * k - 1)


def pole_two_j(pole, tj):
    t, k = pole
    return tj + 2 * k if t == 0 else (tj - 2 * k if t == 1 else tj)


def delta_2(pole, tj):
    """twice the pole position in Delta"""
    t, k = pole
    return (2 - tj - 2 * k, 4 + tj - 2 * k, 3 - 2 * k, 4 - 2 * k)[t]


def parity(pole):
    return 1 if pole[0] == 3 else 0


def q_value(pole, tj):
    t, k = pole
    d = Fraction(4) ** shift(pole)
    if t in (0, 1):
        return -d
    if t == 2:
        return -d * Fraction(k * (1 + tj - 2 * k), 1 + tj + 2 * k)
    return d * Fraction((2 + tj - 2 * k) * (2 * k - 1), 2 * tj + 4 * k)


def two_j_values(m, order, tj12, tjmax):
    return list(range(tj12 % 2, tjmax + 2 * (order - m - 1) + 1, 2))


@lru_cache(maxsize=None)
def so3(pp, tj1, tj2):
    return tuple(c for c in range(abs(tj1 - tj2), tj1 + tj2 + 1, 2) if ((tj1 + tj2 - c) - 2 * pp) % 4 == 0)


def j_excess(tj, tjmax):
    return (tj - tjmax) // 2 if tj > tjmax else 0


def cauchy_coeff(dprime_2, pole_delta_2):
    diff = dprime_2 - pole_delta_2
    return Fraction(0) if diff == 0 else Fraction(2, diff)
