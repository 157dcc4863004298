"""The write_output composite maps (blocksvibe_amd.output.build_map, host mpmath) against the oracle's
line-by-line port of blocks_3d's coordinate transforms (oracle.output_ref), on random numerators: the
map construction itself, without the GPU."""
import random

import mpmath as mp
import pytest

from oracle import output_ref as OR


@pytest.mark.parametrize("Lam,parity,D", [(7, 0, "2.3"), (7, 1, "2.3"), (6, 0, "0"), (9, 1, "6"), (5, 1, "0")])
def test_map_matches_oracle(Lam, parity, D):
    from blocksvibe_amd.output import build_map
    L = 10
    rng = random.Random(Lam * 10 + parity)
    nL, Rr = (Lam - parity) // 2 + 1, Lam - parity
    for coord in ("xt", "xt_radial"):
        W, inputs, outputs = build_map(coord, Lam, parity, D, L)
        mp.mp.prec = 32 * L + 128
        num = [[mp.mpf(rng.uniform(-1, 1)) if e <= Rr - 2 * idx else mp.mpf(0) for e in range(Rr + 1)]
               for idx in range(nL)]
        got = [sum((W[o][u] * num[idx][e] for u, (idx, e) in enumerate(inputs)), mp.mpf(0))
               for o in range(len(outputs))]
        if coord == "xt":
            ref = [x for row in OR.reference_xt(num, Lam, parity, mp.mpf(D)) for x in row]
        else:
            ref = OR.reference_radial(num[0], Lam, parity, mp.mpf(D))
        assert len(got) == len(ref)
        big = max(abs(x) for x in ref)
        err = max(abs(a - b) for a, b in zip(got, ref)) / big
        assert err < mp.mpf(2) ** -(32 * L + 100), (coord, float(mp.log(err, 2)) if err else None)
