"""Bit-level specification of the multiprecision float ("MPF") used by the HIP kernels, restated with
Python integers.  TEST INFRASTRUCTURE: see oracle/__init__.py.

Number format (the build's own; the reference's unvendored sdpvibe.mpf fixes only its outline,
reference blocksvibe/cauchy.py:31-33, smallgemm.py:30-32, derivatives.py:66-93, cli.py:270-274):
  * a value is m * 2^e with m a signed integer of L 32-bit limbs (two's complement) and e an int64;
  * normalised: bitlen(|m|) == NB = 32 L - 2, or m == 0 and e == EZ.

Every operation below is "exact integer intermediate, then one rounding by norm()", with the integer
intermediate defined exactly, so a GPU result can be compared bit for bit.

Pinned by tests/test_oracle_mpf.py: norm against exact rationals (error <= 1/2 ulp, ties away from zero),
align/gemm/cauchy/add against Fraction arithmetic with their stated error bounds, and brute force over all
small cases at L = 2.
"""
from fractions import Fraction

EZ = -(1 << 62)


def NB(L):
    return 32 * L - 2


def bitlen(x):
    return abs(x).bit_length()


def norm(W, e, L):
    """exact W * 2^e -> normalised (m, e'): NB significant bits, round half away from zero"""
    if W == 0:
        return 0, EZ
    a = abs(W)
    s = a.bit_length() - NB(L)
    if s > 0:
        a = (a + (1 << (s - 1))) >> s
        if a.bit_length() > NB(L):
            a >>= 1
            s += 1
    else:
        a <<= -s
    return (-a if W < 0 else a), e + s


def from_fraction(x, L):
    """exact rational -> normalised, round half away from zero (magnitude rounding)"""
    x = Fraction(x)
    if x == 0:
        return 0, EZ
    a = abs(x)
    # e such that 2^(NB-1) <= a / 2^e < 2^NB
    e = a.numerator.bit_length() - a.denominator.bit_length() - NB(L)
    while a / Fraction(2) ** e >= (1 << NB(L)):
        e += 1
    while a / Fraction(2) ** e < (1 << (NB(L) - 1)):
        e -= 1
    y = a / Fraction(2) ** e
    m = int(y)
    if y - m >= Fraction(1, 2):
        m += 1
    if m.bit_length() > NB(L):
        m >>= 1
        e += 1
    return (-m if x < 0 else m), e


def value(m, e):
    if m == 0:
        return Fraction(0)
    return Fraction(m) * Fraction(2) ** e


def align(xs):
    """list of normalised (m, e) -> (aligned mantissas, E): E = max exponent of the nonzero entries;
    each magnitude shifted right by E - e_k with truncation (sign kept).  All zero -> ([0...], EZ)."""
    nz = [e for m, e in xs if m != 0]
    if not nz:
        return [0] * len(xs), EZ
    E = max(nz)
    out = []
    for m, e in xs:
        if m == 0:
            out.append(0)
            continue
        a = abs(m) >> (E - e)
        out.append(-a if m < 0 else a)
    return out, E


def gemm(A, B, L):
    """A: rows x K, B: K x cols, lists of (m, e).  Rows of A aligned (per row over k), columns of B
    aligned (per column over k); exact sum of the aligned products; one rounding."""
    rows, K, cols = len(A), len(B), len(B[0]) if B else 0
    Aal = [align(A[i]) for i in range(rows)]
    Bal = [align([B[k][j] for k in range(K)]) for j in range(cols)]
    C = []
    for i in range(rows):
        am, ae = Aal[i]
        row = []
        for j in range(cols):
            bm, be = Bal[j]
            if ae == EZ or be == EZ:
                row.append((0, EZ))
                continue
            row.append(norm(sum(am[k] * bm[k] for k in range(K)), ae + be, L))
        C.append(row)
    return C


def gemm_aligned(am, ae, bm, be, L):
    """one entry from pre-aligned operands (am: K mantissas with exponent ae, bm likewise)"""
    if ae == EZ or be == EZ:
        return 0, EZ
    return norm(sum(x * y for x, y in zip(am, bm)), ae + be, L)


def add(a, b, L):
    """a + b: both operands put on the grid 2^(max(ea, eb) - 32) (the smaller one truncated toward zero
    on its magnitude), summed exactly, rounded once."""
    (ma, ea), (mb, eb) = a, b
    if ma == 0:
        return (mb, eb) if mb != 0 else (0, EZ)
    if mb == 0:
        return ma, ea
    tt = max(ea, eb) - 32

    def grid(m, e):
        s = e - tt
        x = abs(m) << s if s >= 0 else abs(m) >> (-s)
        return -x if m < 0 else x

    return norm(grid(ma, ea) + grid(mb, eb), tt, L)


def mul(a, b, L):
    (ma, ea), (mb, eb) = a, b
    if ma == 0 or mb == 0:
        return 0, EZ
    return norm(ma * mb, ea + eb, L)


def cauchy(xs, ds, L, h=(0, EZ)):
    """h + sum_i x_i * 2 / d_i for small nonzero integers d_i (d_i == 0: no term), one rounding.
    E = the largest exponent of the nonzero x_i (all columns of the group, whatever d); the grid is
    G = max(E - 31, e_h - 32) (E - 31 if h is zero, e_h - 32 if every x_i is zero); term i is
    sign(d_i) * floor(m_i * 2^(32 - sh_i) / |d_i|), sh_i = G + 31 - e_i >= 0, with the signed mantissa m_i
    and floor toward -infinity (no term when sh_i >= 32 (L + 1)); h is truncated toward zero onto 2^G
    (|m_h| 2^(e_h - G), e_h - G <= 32); the terms and h are summed exactly and rounded once by norm."""
    nz = [e for m, e in xs if m != 0]
    mh, eh = h
    if not nz and mh == 0:
        return 0, EZ
    if not nz:
        G = eh - 32
    elif mh == 0:
        G = max(nz) - 31
    else:
        G = max(max(nz) - 31, eh - 32)
    acc = 0
    for (m, e), d in zip(xs, ds):
        if d == 0 or m == 0:
            continue
        sh = G + 31 - e
        if sh >= 32 * (L + 1):
            continue
        q = (m << 32) // (abs(d) << sh)
        acc += q if d > 0 else -q
    if mh:
        t = eh - G
        a = abs(mh) << t if t >= 0 else abs(mh) >> (-t)
        acc += -a if mh < 0 else a
    return norm(acc, G, L)


def mom_flush(ys, nks, acc_prev, L):
    """windowed moment update: v + sum_w y_w * n_w^k (exact integers n_w^k).
    ys: list of (m, e) (zero allowed), nks: matching exact integers n_w^k >= 0.
    E = max over nonzero y_w of e_w + NB + bitlen(n_w^k) (an upper bound of the top bit); every product
    |y_w| n_w^k is put on the grid 2^T, T = E - 32 (L + 1), truncated toward zero, summed exactly with its
    sign, rounded once; the rounded window sum is then added to v by add()."""
    cand = [e + NB(L) + nk.bit_length() for (m, e), nk in zip(ys, nks) if m != 0]
    if not cand:
        return acc_prev
    E = max(cand)
    T = E - 32 * (L + 1)
    s = 0
    for (m, e), nk in zip(ys, nks):
        if m == 0 or nk == 0:
            continue
        p = abs(m) * nk
        sh = T - e
        q = p >> sh if sh >= 0 else p << (-sh)
        s += -q if m < 0 else q
    if s == 0:
        return acc_prev
    return add(acc_prev, norm(s, T, L), L)


# ------------------------------------------------------------------ limb-plane conversion (host arrays)

def to_limbs(m, L):
    """signed mantissa -> list of L uint32 limbs (two's complement, limb 0 least significant)"""
    u = m & ((1 << (32 * L)) - 1)
    return [(u >> (32 * k)) & 0xFFFFFFFF for k in range(L)]


def from_limbs(limbs):
    L = len(limbs)
    u = 0
    for k, x in enumerate(limbs):
        u |= (int(x) & 0xFFFFFFFF) << (32 * k)
    if u >> (32 * L - 1):
        u -= 1 << (32 * L)
    return u
