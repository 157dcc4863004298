"""The write_output map built on the GPU (blocksvibe_amd.output.build_map_device: host S, R_z, P, M in mpmath,
then elementwise products and two RNS GEMMs at L + 4 or more limbs, one rounding to L limbs) against the same
map built entirely on the host in mpmath at 32 L + 128 bits (build_map, itself checked against the oracle's
restatement of the reference's chain in tests/test_output_map.py).

Both are roundings of the same exact map: every entry must agree to within 2 units in the last place of L
limbs, taken relative to the largest entry of its row (the map is applied as a row-aligned GEMM, so that is
the precision it is used at).  A large Lambda is timed against the host construction's known cost."""
import time

import mpmath as mp
import pytest

pytestmark = pytest.mark.gpu


def _check(coord, Lam, p, D, L):
    from blocksvibe_amd import mpf
    from blocksvibe_amd.output import build_map, build_map_device
    Wd, ins, outs = build_map_device(coord, Lam, p, D, L)
    Wh, ins_h, outs_h = build_map(coord, Lam, p, D, L)
    assert ins == ins_h and outs == outs_h
    got = Wd.to_host()
    K = len(ins)
    mp.mp.prec = 32 * L + 128
    worst = -10 ** 9
    for o in range(len(outs)):
        row = [mp.mpf(x) for x in Wh[o]]
        mx = max(abs(x) for x in row)
        for u in range(K):
            m, e = got[o * K + u]
            v = mp.mpf(m) * mp.mpf(2) ** e if m else mp.mpf(0)
            d = abs(v - row[u])
            if d:
                worst = max(worst, int(mp.floor(mp.log(d / mx, 2))))
            assert d <= mp.mpf(2) ** -(mpf.NB(L) - 2) * mx, (o, u, v, row[u])
    return worst


@pytest.mark.parametrize("coord,Lam,p,D,L", [("xt", 7, 0, "0", 10), ("xt", 9, 1, "2.5", 10), ("xt", 11, 0, "6", 34),
                                             ("xt", 6, 1, "1.25", 50), ("xt_radial", 7, 1, "0.6", 10),
                                             ("xt", 5, 0, "-0.75", 4)])
def test_device_map_equals_host_map(coord, Lam, p, D, L):
    from blocksvibe_amd import mpf
    worst = _check(coord, Lam, p, D, L)
    print(f"{coord} Lambda={Lam} parity={p} D={D} L={L}: worst difference 2^{worst} of the row maximum "
          f"(NB = {mpf.NB(L)})")


def test_large_lambda_is_fast():
    """Lambda = 27 at 1536 bits: the host construction takes about 27 s (O^5 mpmath operations); the device
    construction's O^5 part is two GEMMs"""
    from blocksvibe_amd.output import build_map_device
    t = time.time()
    W, ins, outs = build_map_device("xt", 27, 0, "6", 50)
    dt = time.time() - t
    print(f"Lambda=27, L=50: {len(outs)} x {len(ins)} map in {dt:.2f} s")
    assert W.shape == (len(outs), len(ins))
    assert dt < 20
