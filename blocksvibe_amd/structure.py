"""Pole and spin bookkeeping of blocks_3d's radial recursion (Erramilli, Iliesiu, Kravchuk, Landry, Poland,
Simmons-Duffin, arXiv:2011.01959), numpy-vectorised.  Same definitions as reference
blocksvibe/structure.py:16-127 (blocks_3d's Pole.hxx, Pole_Types.cxx, two_j_values.cxx,
so3_tensor_parity.cxx, fill_residues.cxx).

Spins are doubled (two_j).  A pole is (type t, k >= 1), t = 0..3 for blocks_3d's types I..IV.
"""
from fractions import Fraction
from functools import lru_cache

import numpy as np


def n_poles(tj, tj12, tj43, order):
    """pole counts per type, int64 (4, *shape) (reference structure.py:16-24), clipped at 0 as the reference's
    vectorised counts_v (structure.py:91-99); the recursion asks only at orders >= 1, where no count is negative"""
    tj = np.asarray(tj, dtype=np.int64)
    order = np.asarray(order, dtype=np.int64)
    tj, order = np.broadcast_arrays(tj, order)
    half = tj // 2
    even = (tj % 2) == 0
    small = np.minimum(half, min(tj12 // 2, tj43 // 2))
    c = np.stack([order,
                  np.minimum(half, order),
                  np.where(even, order // 2, np.minimum(small, order // 2)),
                  np.where(even, np.minimum(small, (order + 1) // 2), (order + 1) // 2)])
    return np.maximum(c, 0)


def pole_of_index(tj, tj12, tj43, order, idx):
    """(type, k) of the idx-th pole (type-major order) of spin tj at `order`"""
    c = np.cumsum(n_poles(tj, tj12, tj43, order), axis=0)
    idx = np.asarray(idx, dtype=np.int64)
    t = (idx >= c[0]).astype(np.int64) + (idx >= c[1]) + (idx >= c[2])
    before = np.where(t == 0, 0, np.take_along_axis(c, np.maximum(t - 1, 0)[None], 0)[0])
    return t, idx - before + 1


def index_of_pole(tj, tj12, tj43, order, t, k):
    c = np.cumsum(n_poles(tj, tj12, tj43, order), axis=0)
    t = np.asarray(t, dtype=np.int64)
    before = np.where(t == 0, 0, np.take_along_axis(c, np.maximum(t - 1, 0)[None], 0)[0])
    return before + np.asarray(k, dtype=np.int64) - 1


def shift(t, k):
    """radial order at which the pole appears (reference structure.py:37-39)"""
    return np.where(t < 2, k, np.where(t == 2, 2 * k, 2 * k - 1))


def pole_spin(t, k, tj):
    """doubled spin of the pole's residue block (reference structure.py:42-44)"""
    return np.where(t == 0, tj + 2 * k, np.where(t == 1, tj - 2 * k, tj))


def pole_delta2(t, k, tj):
    """twice the pole position in Delta (reference structure.py:47-50)"""
    t = np.asarray(t)
    return np.choose(t, [2 - tj - 2 * k, 4 + tj - 2 * k, 3 - 2 * k, 4 - 2 * k])


def spin_excess(tj, tjmax):
    return np.where(tj > tjmax, (tj - tjmax) // 2, 0)


def q_value(t, k, tj):
    """Q(pole, two_j) of blocks_3d's fill_residues, exact (reference structure.py:56-64)"""
    s = int(shift(np.int64(t), np.int64(k)))
    d = Fraction(4) ** s
    if t in (0, 1):
        return -d
    if t == 2:
        return -d * Fraction(k * (1 + tj - 2 * k), 1 + tj + 2 * k)
    return d * Fraction((2 + tj - 2 * k) * (2 * k - 1), 2 * tj + 4 * k)


def spins_at_level(m, order, tj12, tjmax):
    """doubled spins of the residue blocks at radial level m (reference structure.py:67-68)"""
    return np.arange(tj12 % 2, tjmax + 2 * (order - m - 1) + 1, 2, dtype=np.int64)


@lru_cache(maxsize=None)
def so3(pseudo_parity, tj1, tj2):
    """SO(3) structures (doubled spins) of the given pseudo-parity (reference structure.py:71-74)"""
    return tuple(c for c in range(abs(tj1 - tj2), tj1 + tj2 + 1, 2)
                 if ((tj1 + tj2 - c) - 2 * pseudo_parity) % 4 == 0)


def poles_list(tj, tj12, tj43, order):
    """list of (type, k) in type-major order"""
    c = n_poles(tj, tj12, tj43, order)
    return [(t, k) for t in range(4) for k in range(1, int(c[t]) + 1)]


def upper_bound_2(tj):
    """twice the unitarity bound offset: Delta = x + ub (reference derivatives.py:387)"""
    return tj + 1 if tj <= 1 else tj + 2
