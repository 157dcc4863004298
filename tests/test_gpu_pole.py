"""The pole step's two product paths through the C-ABI (b3d_pole_level): the int8 MFMA path (digit tables,
S*B >= 32 rows per constant row) against the VALU path (28-bit limbs, no digit tables) on the same random
inputs, bit for bit, for both products (Y = R H^T and res = QL Y^T).  Covers several row blocks per
task, S up to the most the MFMA path stages in LDS (6 at L = 34, 3 at L = 50, 16 at L = 4), S = 8 at
L = 34 (taken by the VALU path), zero entries and exponent spreads wide enough that aligned entries
vanish.  Residue blocks of SL x SR < S x S structure entries against the same products on the S x S blocks
padded with zeros (the entries the reference's padded layout holds), bit for bit.  The VALU path itself is
checked against the oracle bit for bit in tests/test_gpu_recursion.py; the digit arithmetic against exact
integers in tests/test_pole_digits.py."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _rand(shape, L, seed, zero_frac=0.1, spread=40):
    from blocksvibe_amd import mpf
    x = mpf.random(shape, L, spread=spread, generator=np.random.default_rng(seed))
    m, e = x.m.clone(), x.e.clone()
    z = torch.from_numpy(np.random.default_rng(seed + 1).random(tuple(shape)) < zero_frac).to(m.device)
    m[:, z] = 0
    e[z] = mpf.EZ
    return mpf.MPF(m, e)


@pytest.mark.parametrize("L,S,B", [(34, 5, 11), (34, 6, 20), (34, 8, 5), (10, 6, 7), (50, 3, 12), (4, 5, 13),
                                   (4, 16, 3)])
def test_pole_mma_matches_valu(L, S, B):
    from blocksvibe_amd import _lib, mpf
    from blocksvibe_amd.recursion import _pole_tables, _split28
    print(f"pole level L={L} S={S} B={B}", flush=True)
    ntab, nH, T = 3, 4, 7
    rng = np.random.default_rng(L * 100 + S)
    Rm, Re = mpf.align_rows(_rand((ntab, S, S), L, 10 + L, spread=3))
    Qm, Qe = mpf.align_rows(_rand((ntab, S, S), L, 20 + L, spread=3))
    R28, Rsg = _split28(Rm, L)
    Q28, Qsg = _split28(Qm, L)
    Rd8, Rc = _pole_tables(Rm, L, S)
    Qd8, Qc = _pole_tables(Qm, L, S)
    H = _rand((nH, S * S, B), L, 30 + L, spread=int(rng.integers(20, 400)))
    dev = H.m.device
    thrw = torch.from_numpy(rng.integers(0, nH, T)).to(dev)
    tR = torch.from_numpy(rng.integers(0, ntab, T).astype(np.int32)).to(dev)
    tQL = torch.from_numpy(rng.integers(0, ntab, T).astype(np.int32)).to(dev)
    tslot = torch.from_numpy(rng.permutation(T + 2)[:T]).to(dev)
    P = lambda t: t.data_ptr()
    outs = []
    for tables in (True, False):
        Y = mpf.zeros((T, S, S, B), L)
        res = mpf.zeros((T + 2, S * S, B), L)
        tb = [P(Rd8), P(Rc), P(Qd8), P(Qc)] if tables else [0, 0, 0, 0]
        _lib.call("b3d_pole_level", L, S, S, S, B, T, P(H.m), P(H.e), H.numel(), P(thrw), P(tR), P(tQL), P(tslot),
                  P(R28), P(Rsg), P(Re), R28[0].numel(), P(Q28), P(Qsg), P(Qe), Q28[0].numel(), *tb, 0, 0,
                  P(Y.m), P(Y.e), Y.numel(), P(res.m), P(res.e), res.numel(), mpf.stream_ptr())
        torch.cuda.synchronize()
        outs.append((Y.m.cpu(), Y.e.cpu(), res.m.cpu(), res.e.cpu()))
    (ym1, ye1, rm1, re1), (ym0, ye0, rm0, re0) = outs
    assert torch.equal(ye1, ye0) and torch.equal(ym1, ym0), "Y = R H^T"
    assert torch.equal(re1, re0) and torch.equal(rm1, rm0), "res = QL Y^T"
    nz = int((re1 != mpf.EZ).sum())
    assert nz > 0
    print(f"identical; {nz} nonzero outputs", flush=True)


@pytest.mark.parametrize("L,SL,SR,S,B", [(34, 5, 1, 5, 11), (34, 2, 5, 5, 11), (10, 3, 2, 4, 7), (50, 4, 3, 4, 5)])
def test_pole_rectangular_matches_padded(L, SL, SR, S, B):
    from blocksvibe_amd import _lib, mpf
    from blocksvibe_amd.recursion import _split28
    print(f"pole level L={L} SL={SL} SR={SR} S={S} B={B}", flush=True)
    ntab, nH, T = 3, 4, 6
    rng = np.random.default_rng(L + 10 * SL + SR)
    Rm, Re = mpf.align_rows(_rand((ntab, S, S), L, 40 + L, spread=3))
    Qm, Qe = mpf.align_rows(_rand((ntab, S, S), L, 50 + L, spread=3))
    R28, Rsg = _split28(Rm, L)
    Q28, Qsg = _split28(Qm, L)
    Hs = _rand((nH, SL, SR, B), L, 60 + L, spread=int(rng.integers(20, 200)))
    Hp = mpf.zeros((nH, S, S, B), L)                         # the same entries in S x S blocks, zeros elsewhere
    Hp.m[:, :, :SL, :SR] = Hs.m
    Hp.e[:, :SL, :SR] = Hs.e
    dev = Hs.m.device
    thrw = torch.from_numpy(rng.integers(0, nH, T)).to(dev)
    tR = torch.from_numpy(rng.integers(0, ntab, T).astype(np.int32)).to(dev)
    tQL = torch.from_numpy(rng.integers(0, ntab, T).astype(np.int32)).to(dev)
    tslot = torch.from_numpy(rng.permutation(T + 2)[:T]).to(dev)
    P = lambda t: t.data_ptr()
    outs = []
    for sl, sr, H in ((SL, SR, Hs), (S, S, Hp)):
        Y = mpf.zeros((T, sr, sl, B), L)
        res = mpf.zeros((T + 2, sl, sr, B), L)
        _lib.call("b3d_pole_level", L, sl, sr, S, B, T, P(H.m), P(H.e), H.numel(), P(thrw), P(tR), P(tQL),
                  P(tslot), P(R28), P(Rsg), P(Re), R28[0].numel(), P(Q28), P(Qsg), P(Qe), Q28[0].numel(), 0, 0, 0, 0,
                  0, 0, P(Y.m), P(Y.e), Y.numel(), P(res.m), P(res.e), res.numel(), mpf.stream_ptr())
        torch.cuda.synchronize()
        outs.append(res)
    got, pad = outs
    assert torch.equal(got.e, pad.e[:, :SL, :SR]) and torch.equal(got.m, pad.m[:, :, :SL, :SR])
    nz = int((got.e != mpf.EZ).sum())
    assert nz > 0
    print(f"identical; {nz} nonzero outputs", flush=True)
