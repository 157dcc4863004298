"""The moment flush (b3d_moments_flush, reference blocksvibe/derivatives.py:474-510) through the C-ABI against the
oracle's window update (oracle.mpf_ref.mom_flush), bit for bit: every width of the exact n^k table the kernel is
instantiated for (1, 2, 3, 4 and 8 limb planes: the products run over exactly the table's limbs), the same bits
when the table carries extra zero limb planes, zero staged entries and zero table values, and the per-column
(right-hand side, k) list that limits a flush to each column's own moments (several lambda orders in one
recursion) while every other moment keeps its value."""
import random

import numpy as np
import pytest
import torch

from oracle import mpf_ref as R

pytestmark = pytest.mark.gpu


def _tables(vals, planes):
    """exact integers -> (KL limb planes as int32 (planes, n), bit lengths)"""
    from blocksvibe_amd import mpf
    nk = np.array([[(v >> (32 * l)) & 0xFFFFFFFF for v in vals] for l in range(planes)], dtype=np.uint64)
    NK = torch.from_numpy(nk.astype(np.uint32).view(np.int32)).to(mpf.dev())
    NKb = torch.tensor([v.bit_length() for v in vals], dtype=torch.int32, device=mpf.dev())
    return NK, NKb


def _flush(L, stage, m0, srow, NK, NKb, v, R1, B, jk=None):
    from blocksvibe_amd import _lib, mpf
    W, nrows, N = stage.shape
    out = v.clone()
    P = lambda t: t.data_ptr()
    _lib.call("b3d_moments_flush", L, NK.shape[0], P(stage.m), P(stage.e), stage.numel(), W, m0, P(srow), P(NK),
              NK.shape[1], P(NKb), P(out.m), P(out.e), out.numel(), nrows, N, R1, B,
              P(jk) if jk is not None else 0, int(jk.numel()) if jk is not None else 0, mpf.stream_ptr())
    torch.cuda.synchronize()
    return out


@pytest.mark.parametrize("L,limbs", [(10, 1), (10, 2), (10, 3), (34, 3), (10, 4), (18, 7)])
def test_flush_matches_oracle(L, limbs):
    from blocksvibe_amd import mpf
    rng = np.random.default_rng(1000 + 10 * L + limbs)
    W, nrows, S2, B, R1, m0 = 3, 4, 3, 2, 4, 2
    N = S2 * B
    print(f"moment flush L={L}, n^k of {limbs} limbs", flush=True)
    stage = mpf.random((W, nrows, N), L, spread=60, generator=rng)
    z = torch.from_numpy(rng.random((W, nrows, N)) < 0.2).to(stage.m.device)
    stage.m[:, z] = 0
    stage.e[z] = mpf.EZ
    v0 = mpf.random((nrows, N, R1), L, spread=80, generator=rng)
    zv = torch.from_numpy(rng.random((nrows, N, R1)) < 0.2).to(v0.m.device)
    v0.m[:, zv] = 0
    v0.e[zv] = mpf.EZ
    srow_h = rng.integers(0, 3, nrows)
    srow = torch.from_numpy(srow_h.astype(np.int32)).to(mpf.dev())
    nmax = m0 + W - 1 + int(srow_h.max())
    prng = random.Random(7 * L + limbs)
    vals = [prng.getrandbits(prng.randint(1, 32 * limbs)) for _ in range((nmax + 1) * R1)]
    vals[-1] = (1 << (32 * limbs)) - 1                       # the widest value the table holds
    vals = [0 if prng.random() < 0.1 else x for x in vals[:-1]] + vals[-1:]
    assert max(vals).bit_length() <= 32 * limbs
    sp, vp = stage.to_host(), v0.to_host()
    want = []
    for row in range(nrows):
        for c in range(N):
            for k in range(R1):
                ys = [sp[(w * nrows + row) * N + c] for w in range(W)]
                nks = [vals[(m0 + w + int(srow_h[row])) * R1 + k] for w in range(W)]
                want.append(R.mom_flush(ys, nks, vp[(row * N + c) * R1 + k], L))
    outs = []
    for planes in sorted({limbs, limbs + 1 if limbs < 4 else 8, 8}):
        NK, NKb = _tables(vals, planes)
        got = _flush(L, stage, m0, srow, NK, NKb, v0, R1, B).to_host()
        assert got == want, ("planes", planes)
        outs.append(planes)
    # the column list: right-hand side b keeps only its first r1[b] moments; all others stay as they were
    r1 = [R1, 2]
    codes = [b * R1 + k for b in range(B) for k in range(r1[b])]
    jk = torch.tensor(codes, dtype=torch.int32, device=mpf.dev())
    NK, NKb = _tables(vals, limbs)
    got = _flush(L, stage, m0, srow, NK, NKb, v0, R1, B, jk).to_host()
    for row in range(nrows):
        for c in range(N):
            for k in range(R1):
                i = (row * N + c) * R1 + k
                assert got[i] == (want[i] if k < r1[c % B] else vp[i]), (row, c, k)
    print(f"identical to the oracle with {outs} limb planes and with the column list", flush=True)
