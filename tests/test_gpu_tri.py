"""The derivative stage's products by exact integer matrices (b3d_tri_product, reference
blocksvibe/derivatives.py:553-554; b3d_int_product, :560 and :567) against the residue-number GEMM
(b3d_rns_gemm) that computes the same a_i[e] = sum_{k <= e} C(e, k) D_i^(e-k) v_i[k] and G = (D_i^w) v, bit
for bit, through the C-ABI: every table width TL
(2, 4, 8 limbs), L 4..50, zero-padding tables, all-zero and mixed columns, exponent spreads that truncate
aligned entries to zero, the strided output layout finish() uses; and the whole Derivs.finish() with the
kernel against the same finish() through the GEMM.  The kernel's per-thread arithmetic is checked against
the oracle's GEMM on the CPU in tests/test_mp_host.py (test_tri_entry); the GEMM against the big-integer
product in tests/test_gpu_rns.py."""
from math import comb

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


@pytest.mark.parametrize("L,K,dmax", [(34, 12, 250), (34, 6, 40), (10, 12, 250), (4, 5, 9), (50, 20, 300),
                                      (18, 1, 5)])
def test_tri_matches_rns_gemm(L, K, dmax):
    from blocksvibe_amd import _lib, mpf, rns
    from blocksvibe_amd.derivatives import Derivs
    print(f"tri product L={L} K={K} |d|<={dmax}", flush=True)
    rng = np.random.default_rng(L * 10 + K)
    d2 = sorted({int(x) for x in rng.integers(-dmax, dmax + 1, 6)} | {0})
    TL, tm, tsg, texp = Derivs._tri_tables(d2, K)
    pairs = [(0, mpf.EZ)] * (K * K)
    for d in d2:
        pairs += [mpf.norm_int(comb(e, k) * d ** (e - k), k - e, L) if k <= e else (0, mpf.EZ)
                  for e in range(K) for k in range(K)]
    Ttab = mpf.from_pairs(pairs, L, (len(d2) + 1, K, K))
    nper, nj, Nc = 5, 3, 70
    nb = nper * nj
    tsel = torch.from_numpy(rng.integers(0, len(d2) + 1, nb)).to(mpf.dev())
    v = _rand((nb, K, Nc), L, L + K, spread=int(rng.integers(3, 300)))
    v.m[:, 2] = 0                                            # one all-zero batch
    v.e[2] = mpf.EZ
    want = rns.gemm(Ttab.take(tsel), v)                       # (nb, K, Nc)
    extra = 2                                                 # rows after the nper products, as in finish()
    out = mpf.zeros((nj, nper + extra, K * Nc), L)
    P = lambda t: t.data_ptr()
    _lib.call("b3d_tri_product", L, TL, P(v.m), P(v.e), v.numel(), nb, K, Nc, P(tsel), P(tm), P(tsg), tsg.numel(),
              P(texp), P(out.m), P(out.e), out.numel(), nper, (nper + extra) * K * Nc, K * Nc, mpf.stream_ptr())
    torch.cuda.synchronize()
    got_m = out.m[:, :, :nper].reshape(L, nb, K, Nc)
    got_e = out.e[:, :nper].reshape(nb, K, Nc)
    assert torch.equal(got_e, want.e) and torch.equal(got_m, want.m)
    assert bool((out.e[:, nper:] == mpf.EZ).all())          # rows outside the products untouched (zeros)
    nz = int((want.e != mpf.EZ).sum())
    assert nz > 0
    print(f"identical; TL={TL}, {nz} nonzero outputs", flush=True)


@pytest.mark.parametrize("kept,lower", [(None, False), (4, False), (4, True)])
def test_finish_tri_equals_gemm(kept, lower):
    """Derivs.finish() with the exact-integer product kernels (a_i triangular, G by the pole positions' powers)
    against finish() with the residue GEMM for both, bit for bit (lower: the constants sliced for a smaller R,
    as a higher lambda order uses them)"""
    import copy
    from blocksvibe_amd.derivatives import Derivs
    from blocksvibe_amd.plan import Plan
    from blocksvibe_amd.recursion import Recursion
    from tests.test_gpu_recursion import h_tensor, random_operator
    L, B, pp, R = 10, 2, ((0, 1), (1, 0)), 4
    print(f"finish kept={kept} lower={lower}", flush=True)
    plan = Plan(2, 2, 6, 9, pp, reuse=True)
    M, H = random_operator(plan, 2, 2, pp, 9, B, 4)
    h = h_tensor(H, plan, L)
    rec = Recursion(plan, M, L)
    base = Derivs(plan, R, L, kept_pole_order=kept)
    outs = []
    for tri in (True, False):
        der = base.for_order(R - 2) if lower else copy.copy(base)
        assert der.tri is not None and der.vand is not None
        if not tri:
            der.tri = der.vand = None
        der.start(B, window=3)
        rec.run(h, hook=der.accumulate)
        outs.append(der.finish(h, chunk=2))
    a, b = outs
    assert torch.equal(a.e, b.e) and torch.equal(a.m, b.m)
    print(f"finish identical: {int((a.e != -(1 << 62)).sum())} nonzero numerator coefficients", flush=True)


@pytest.mark.parametrize("L,M,K,dmax,lower", [(34, 11, 124, 240, False), (34, 12, 12, 250, True), (10, 3, 40, 9, False),
                                              (50, 6, 17, 300, False), (4, 1, 5, 5, False), (18, 5, 9, 60, True)])
def test_int_product_matches_rns_gemm(L, M, K, dmax, lower):
    """b3d_int_product (full rows of M x K exact integer matrices d^w 2^-w, and the lower-triangular form)
    against b3d_rns_gemm with the same matrices as A, bit for bit: per-batch tables (the derivative stage's
    Vandermonde rows of pole positions, zero padding columns included), mixed and all-zero columns, wide
    exponent spreads"""
    from blocksvibe_amd import _lib, mpf, rns
    print(f"int product L={L} M={M} K={K} |d|<={dmax} lower={lower}", flush=True)
    rng = np.random.default_rng(L * 100 + M * 10 + K)
    nb, Nc = 4, 96
    ints = []
    for b in range(nb):
        npole = K - int(rng.integers(0, 3))                   # a few padding columns at the end
        d = [int(x) for x in rng.integers(-dmax, dmax + 1, npole)] + [0] * (K - npole)
        row = [[(d[i] ** w if i < npole else 0) if (not lower or i <= w) else 0 for i in range(K)] for w in range(M)]
        ints.append(row)
    bits = max(abs(x).bit_length() for t in ints for r in t for x in r)
    TL = next(t for t in (2, 4, 8) if bits <= 32 * t)
    mags = [abs(x) for t in ints for r in t for x in r]
    tm = np.array([[(x >> (32 * l)) & 0xFFFFFFFF for x in mags] for l in range(TL)], dtype=np.uint64)
    tm = torch.from_numpy(tm.astype(np.uint32).view(np.int32).reshape(TL, nb, M, K)).to(mpf.dev())
    tsg = torch.tensor([-1 if x < 0 else 0 for t in ints for r in t for x in r], dtype=torch.int32,
                       device=mpf.dev()).reshape(nb, M, K)
    texp = torch.tensor([-w for _ in range(nb) for w in range(M)], dtype=torch.int64, device=mpf.dev()).reshape(nb, M)
    A = mpf.from_pairs([mpf.norm_int(x, -w, L) for t in ints for w, r in enumerate(t) for x in r], L, (nb, M, K))
    v = _rand((nb, K, Nc), L, L + K, spread=int(rng.integers(3, 300)))
    v.m[:, 1] = 0                                            # one all-zero batch
    v.e[1] = mpf.EZ
    want = rns.gemm(A, v)                                    # (nb, M, Nc)
    out = mpf.zeros((nb, M, Nc), L)
    tsel = torch.arange(nb, dtype=torch.int64, device=mpf.dev())
    P = lambda t: t.data_ptr()
    _lib.call("b3d_int_product", L, TL, P(v.m), P(v.e), v.numel(), nb, M, K, int(lower), Nc, P(tsel), P(tm), P(tsg),
              tsg.numel(), P(texp), P(out.m), P(out.e), out.numel(), 1, M * Nc, 0, mpf.stream_ptr())
    torch.cuda.synchronize()
    assert torch.equal(out.e, want.e) and torch.equal(out.m, want.m)
    nz = int((want.e != mpf.EZ).sum())
    assert nz > 0
    print(f"identical; TL={TL}, {nz} nonzero outputs", flush=True)
