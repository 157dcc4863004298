"""write_output on the GPU (blocksvibe_amd.output: one composite matrix per (coordinate, Lambda, parity,
Delta_1+2), applied as one RNS GEMM) against oracle.output_ref (line-by-line port of blocks_3d's
lambda_lr_to_lrho_lrhob, congruence_transform, zzb_to_xt, lrho_to_xt_radial) on random numerator
coefficients (reference tests/test_blocks_output.py:50-79).  Bound 2^-(32L - 40) relative to the largest
output."""
import random

import mpmath as mp
import pytest

from oracle import mpf_ref as R
from oracle import output_ref as OR

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("Lam,parity,D,L,seed", [(7, 0, "2.3", 10, 0), (7, 1, "2.3", 10, 0), (6, 0, "0", 10, 0),
                                                   (9, 1, "6", 10, 2), (11, 0, "6", 34, 3)])
def test_output(Lam, parity, D, L, seed):
    from blocksvibe_amd import mpf
    from blocksvibe_amd.output import Output
    rng = random.Random(seed)
    mp.mp.prec = 32 * L + 128
    nL, Rr = (Lam - parity) // 2 + 1, Lam - parity
    num = [[mp.mpf(rng.uniform(-1, 1)) * mp.mpf(2) ** rng.randint(-20, 20) if e <= Lam - parity - 2 * idx else mp.mpf(0)
            for e in range(Rr + 1)] for idx in range(nL)]
    worst = 0
    for coord in ("xt", "xt_radial"):
        o = Output(coord, Lam, parity, D, L)
        # numerators per lambda order idx as Derivs output (Jo=1, Dout=1, R_idx + 1, 2, S2=1, nq=1), parity 0
        nums = []
        for idx in range(nL if coord == "xt" else 1):
            R1 = Rr - 2 * idx + 1
            vals = [R.norm(*_pair(num[idx][e]), L) if p == 0 else (0, R.EZ) for e in range(R1) for p in range(2)]
            nums.append(mpf.from_pairs(vals, L, (1, 1, R1, 2, 1, 1)))
        g = o.apply_per_order(nums).to_host()                # (Jo, 2, S2, nq, outputs, Dmax)
        got = [R.value(*g[k]) for k in range(len(o.outputs))]
        mp.mp.prec = 32 * L + 128
        if coord == "xt":
            ref = [x for row in OR.reference_xt(num, Lam, parity, mp.mpf(D)) for x in row]
        else:
            ref = OR.reference_radial(num[0], Lam, parity, mp.mpf(D))
        big = max(abs(x) for x in ref)
        err = max(abs(mp.mpf(a.numerator) / a.denominator - b) for a, b in zip(got, ref)) / big
        worst = max(worst, err)
    bits = -float(mp.log(worst, 2))
    print(f"Lambda {Lam}, parity {parity}, Delta_1+2 {D}: max error relative to the largest output 2^-{bits:.1f}")
    assert bits > 32 * L - 40, bits


def _pair(x):
    sign, man, exp, bc = x._mpf_
    return (-int(man) if sign else int(man)), int(exp)
