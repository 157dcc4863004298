"""Pins of the write_output oracle (oracle.output_ref) by mathematics that fixes each matrix:
Stirling numbers of the first kind (textbook values and the log-derivative identity), rho_z against
numerical z-derivatives of f(rho(z)) at z = 1/2 (rho(1/2) = 3 - 2 sqrt 2), the prefactor matrices
against numerical derivatives of z^-D g(z) and (1 - z)^-1/2 g(z)."""
import mpmath as mp

from oracle import output_ref as OR


def test_stirling_textbook():
    s = OR.stirling_numbers(6)
    assert s[5][:6] == [0, 24, -50, 35, -10, 1]
    assert s[6][:7] == [0, -120, 274, -225, 85, -15, 1]


def test_stirling_log_derivatives():
    mp.mp.dps = 50
    O, a = 6, mp.mpf("0.731")
    S = OR.stirling(O)
    rc = 3 - 2 * mp.sqrt(2)
    for n in range(O + 1):
        lhs = sum(S[n, k] * a ** k * rc ** a for k in range(O + 1))       # (d/dlog r)^k r^a = a^k r^a
        rhs = mp.ff(a, n) * rc ** (a - n)                                  # d^n/dr^n r^a
        assert abs(lhs - rhs) < mp.mpf(10) ** -40 * (1 + abs(rhs))


def test_rho_z_numerical():
    mp.mp.dps = 60
    O = 5
    Zi = OR.rho_z(O)
    rho = lambda z: z / (1 + mp.sqrt(1 - z)) ** 2
    rc = rho(mp.mpf(0.5))
    assert abs(rc - (3 - 2 * mp.sqrt(2))) < mp.mpf(10) ** -55
    for a in (mp.mpf("0.37"), mp.mpf(3)):
        f = [mp.diff(lambda r: r ** a, rc, n) for n in range(O + 1)]
        g = [mp.diff(lambda z: rho(z) ** a, mp.mpf(0.5), p) for p in range(O + 1)]
        for p in range(O + 1):
            got = sum(Zi[p, n] * f[n] for n in range(O + 1))
            assert abs(got - g[p]) < mp.mpf(10) ** -30 * (1 + abs(g[p]))


def test_prefactors_numerical():
    mp.mp.dps = 60
    O, D = 5, mp.mpf("2.3")
    gfun = lambda z: mp.exp(z) * (1 + z ** 2)
    g = [mp.diff(gfun, mp.mpf(0.5), k) for k in range(O + 1)]
    P = OR.power_prefactor(O, D)
    Q = OR.one_minus_power_prefactor(O)
    for n in range(O + 1):
        want = mp.diff(lambda z: z ** -D * gfun(z), mp.mpf(0.5), n)
        assert abs(sum(P[n, k] * g[k] for k in range(O + 1)) - want) < mp.mpf(10) ** -30 * (1 + abs(want))
        want = mp.diff(lambda z: (1 - z) ** mp.mpf(-0.5) * gfun(z), mp.mpf(0.5), n)
        assert abs(sum(Q[n, k] * g[k] for k in range(O + 1)) - want) < mp.mpf(10) ** -30 * (1 + abs(want))
