"""blocks_3d's write_output maps, line by line in mpmath.  TEST INFRASTRUCTURE (oracle/__init__.py).

Restates reference blocksvibe/output.py:35-95 (stirling, rho_z, power_prefactor,
one_minus_power_prefactor) and reference tests/test_blocks_output.py:13-47 (congruence, the xt chain
lambda_lr_to_lrho_lrhob -> congruence transforms -> zzb_to_xt) plus the xt_radial product.
Pinned by tests/test_oracle_output.py: Stirling numbers against their textbook values, rho_z against
numerical z-derivatives of rho(z)^a at z = 1/2 (rho(1/2) = 3 - 2 sqrt 2), power_prefactor against
numerical derivatives of the prefactor function; and by tests/test_oracle_pins.py: the whole xt and xt_radial
chains against numerical x, t derivatives of (z zbar)^(-D/2) h(rho(z), rho(zbar)) for known even and odd
functions h, fed their numerical (log r, angle) derivatives (the odd part's orientation y = (zbar - z) / 2 is
the one convention the mathematics leaves open).
"""
import math

import mpmath as mp


def binom_gen(k, j):
    def poch(a, n):
        r = 1
        for q in range(n):
            r *= a + q
        return r
    if k < j:
        return 0 if j < 0 else mp.mpf(poch(k - j + 1, j)) / math.factorial(j)
    return mp.mpf(poch(j + 1, k - j)) / math.factorial(k - j)


def factorial_power(x, n):
    def poch(a, n):
        r = mp.mpf(1)
        for k in range(n):
            r *= a + k
        return r
    return 1 / poch(x + 1, -n) if n < 0 else poch(x + 1 - n, n)


def stirling_numbers(O):
    """signed Stirling numbers of the first kind s(n, k), n, k <= O"""
    S1 = [[0] * (O + 1) for _ in range(O + 1)]
    S1[0][0] = 1
    for n in range(O):
        for k in range(n + 1):
            S1[n + 1][k + 1] = -n * S1[n][k + 1] + S1[n][k]
    return S1


def stirling(O):
    S1 = stirling_numbers(O)
    rc = 3 - 2 * mp.sqrt(2)
    return mp.matrix([[rc ** -n * S1[n][k] for k in range(O + 1)] for n in range(O + 1)])


def rho_z_forward(O):
    """matrix z[p][n] before inversion (rho-derivative index n -> z-derivative index p)"""
    z = mp.matrix(O + 1, O + 1)
    s2 = mp.sqrt(2)
    for p in range(O + 1):
        for n in range(p + 1):
            acc = mp.mpf(0)
            for k in range(n + 1):
                acc += (binom_gen(n, k) * mp.mpf(0.5) ** n * (-1) ** k * (4 * s2) ** (n - k) * binom_gen(-2 * n, p - k - n)
                        * (4 - 2 * s2) ** (-n - p + k))
            z[p, n] = mp.factorial(p) / mp.factorial(n) * acc
    return z


def rho_z(O):
    return mp.inverse(rho_z_forward(O))


def power_prefactor(O, D):
    P = mp.matrix(O + 1, O + 1)
    for n in range(O + 1):
        for k in range(n + 1):
            P[n, k] = factorial_power(-D, n - k) * mp.mpf(2) ** (n - k + D) * binom_gen(n, k)
    return P


def one_minus_power_prefactor(O):
    P = mp.matrix(O + 1, O + 1)
    h = mp.mpf(0.5)
    for n in range(O + 1):
        for k in range(n + 1):
            P[n, k] = (-1) ** (n - k) * factorial_power(-h, n - k) * mp.mpf(2) ** (n - k + h) * binom_gen(n, k)
    return P


def congruence(parity, T, inp):
    O = len(inp) - 1
    inter = [[mp.mpf(0)] * (O + 1) for _ in range(O + 1)]
    for i in range(O + 1):
        for k in range(O + 1):
            for l in range(O + 1 - k):
                v = inp[k][l] if l <= k else (inp[l][k] if parity % 2 == 0 else -inp[l][k])
                inter[i][l] += T[i, k] * v
    return [[sum((T[j, l] * inter[i][l] for l in range(O)), mp.mpf(0)) for j in range(min(i, O - i) + 1)]
            for i in range(O + 1)]


def reference_xt(num, Lam, parity, D):
    """num[idx][e] (scalars) -> xt[n][b]"""
    O = Lam
    F = [[sum(((-1) ** (j - l) * mp.mpf(2) ** (-i - j) * mp.binomial(i, k) * mp.binomial(j, l)
               * num[(i + j - k - l - parity) // 2][k + l]
               for k in range(i + 1) for l in range((i + j - k - parity + 2) % 2, j + 1, 2)), mp.mpf(0))
          for j in range(min(i, O - i) + 1)] for i in range(O + 1)]
    Z = congruence(parity, rho_z(O), congruence(parity, stirling(O), F))
    if D != 0:
        Z = congruence(parity, power_prefactor(O, D / 2), Z)
    out = []
    for n in range(O - parity + 1):
        row = []
        for m in range(parity, O - n + 1, 2):
            r = mp.mpf(0)
            for k in range(n + 1):
                for l in range(m + 1):
                    i, j = n + m - k - l, k + l
                    z = Z[i][j] if i >= j else (1 if parity % 2 == 0 else -1) * Z[j][i]
                    r += (-1) ** l * mp.factorial((m - parity) // 2) / mp.factorial(m) * mp.binomial(n, k) * mp.binomial(m, l) * z
            row.append(r)
        out.append(row)
    return out


def reference_radial(num0, Lam, parity, D):
    """num0[e] (lambda order n = parity) -> xt_radial[i]"""
    O = Lam - parity
    pre = power_prefactor(O, D) if parity == 0 else one_minus_power_prefactor(O) * power_prefactor(O, D + 1)
    T = pre * rho_z(O) * stirling(O)
    return [sum((T[i, j] * num0[j] for j in range(O + 1)), mp.mpf(0)) for i in range(O + 1)]
