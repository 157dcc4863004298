"""The five workload configurations of the hot path, each run on the GPU and checked (reference
tests/test_blocks_recursion.py:12-52, tests/test_dropin.py:20-59):

  cfg1  four scalar externals, Delta_12 = Delta_43 = 0, even spins 0..10, xt, order 20, nmax 2, kept 6, 256 bits,
        one task through bin/blocks_3d: every JSON number within 2^-(256 - 32) of the oracle chain (exact setup
        values -> recursion (the C restatement at 50 limbs, bit-identical to recursion_ref.recursion_bits) ->
        numerators with pole shifting (derivs_ref) -> xt (output_ref)) relative to its spin file's largest number
  cfg2  four spin-1/2 externals, Delta_12 = 0.3, Delta_43 = -0.2, spins 0..20, xt_radial, order 40, nmax 4, kept
        10, 512 bits, structures of both four-point parities and both signs in one batched run through bin/blocks_3d:
        a structure of sign +1 at every spin and one of sign -1 at three spins against the oracle chain within
        2^-(512 - 32)
  cfg3  the heaviest operator group of the nmax = 6 workload ((2j12, 2j43) = (8, 8), spins 0..30, order 60,
        1024 bits), two right-hand sides: every residue slot bit for bit against the C restatement of the
        recursion (oracle/c_chain.py; 25 of the 60 levels take k_cauchy_pairs, the others k_cauchy_groups, and
        the pole step is the VALU path on 5 x 5 blocks), and sampled numerators (kept 14) within 2^-(1024 - 32) of oracle.derivs_ref
        relative to each polynomial's largest coefficient
  cfg4  a serving mix of limb counts 10..42 and nmax 2..8 with both signs: each request's result identical alone
        and in the mix
  cfg5  four spin-2 externals at (8, 8), spins 0..64, order 90, nmax 10, kept 20, 1536 bits (L = 50), xt_radial:
        one right-hand side through the whole chain, identical with and without slot recycling and alone or in a
        batch of two; and the whole chain of a batch of two captured as one HIP graph, every replay identical
        to the eager run of its input
Operators and h_inf are seeded synthetic data of the real shapes (blocks_3d's own M and h_inf need its sources).
"""
import json
import os
import subprocess
import sys
from fractions import Fraction

import mpmath as mp
import numpy as np
import pytest
import torch

from oracle import derivs_ref as DR
from oracle import mpf_ref as R
from oracle import output_ref as OR
from oracle import structure_ref as OS
from tests.test_gpu_dropin import read_setup

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def run_cli(tmp_path, lines):
    helper = tmp_path / "b3d_setup"
    helper.write_text(f"#!/bin/bash\nexec {sys.executable} {ROOT}/tests/synthetic_b3d_setup.py \"$@\"\n")
    helper.chmod(0o755)
    f = tmp_path / "lines.txt"
    f.write_text("".join(l + "\n" for l in lines))
    env = dict(os.environ, BLOCKSVIBE_B3D_SETUP=str(helper), BLOCKSVIBE_CACHE=str(tmp_path / "cache"),
               BLOCKSVIBE_DAEMON="0", PYTHONPATH=ROOT)
    env.pop("BLOCKSVIBE_FALLBACK", None)
    r = subprocess.run([os.path.join(ROOT, "bin", "blocks_3d"), "batch", "--lines", str(f)], env=env,
                       capture_output=True, text=True, timeout=900)
    print(r.stderr[-2000:])
    assert r.returncode == 0


def c_residues(M, h, tj12, tj43, tjmax, order, pp, L=50):
    """the recursion of one right-hand side through the C restatement (oracle/c_chain.py, bit-identical to
    recursion_ref.recursion_bits) at L = 50 limbs, i.e. within 2^-1530 of the exact recursion relative to each
    (pole, spin) block -- far below any tolerance checked here; exact setup values in (Fractions), residues out as
    a function (p, m, j, pole index, a, b) -> mpmath number"""
    from blocksvibe_amd.plan import Plan
    from oracle import c_chain as CC
    plan = Plan(tj12, tj43, tjmax, order, pp, reuse=False)
    S, SL, SR, S2, O1 = plan.Smax, plan.SLm, plan.SRm, plan.S2, order + 1
    Mtab = [(0, R.EZ)] * (max(len(plan.Mkeys), 1) * S * S)
    for key, i in plan.Mkeys.items():
        for a, row in enumerate(M.get(key) or []):
            for b, x in enumerate(row):
                Mtab[(i * S + a) * S + b] = R.from_fraction(x, L)
    flat = [(0, R.EZ)] * (2 * O1 * plan.Jh * S2)
    for (s, m, j, a, b), v in h.items():
        if m <= order and j < plan.Jh and a < SL and b < SR:
            flat[(((s * O1 + m) * plan.Jh + j) * SL + a) * SR + b] = R.from_fraction(v, L)
    hm, he = CC.planes(flat, L)
    rm, re = CC.recursion(plan, Mtab, hm, he, L, 1, threads=max(1, min(16, len(os.sched_getaffinity(0)))))
    cache = {}

    def value(p, m, j, q, a, b):
        key = (p, m, j, q, a, b)
        if key not in cache:
            i = plan.phys_slot(p * plan.nslots + plan.offset(m, j) + q) * S2 + a * SR + b
            v = R.from_limbs(rm[i].tolist())
            cache[key] = mp.mpf(v) * mp.mpf(2) ** int(re[i]) if v else mp.mpf(0)
        return cache[key]
    return value


def oracle_files(tmp_path, precision, two_js, tj12, tj43, order, Lam, tjmax, kept, radial, structures, D,
                 check, coord):
    """compare the spin files of structures[i] for i, spins in check (out_<i>/b_<j>.json; spins None: every
    spin) with the oracle chain on the exact setup values; returns the worst error in bits relative to each file's
    largest number"""
    from blocksvibe_amd.setup import write_synthetic
    sfile = tmp_path / "oracle_setup.bin"
    write_synthetic(str(sfile), precision, list(two_js), tj12, tj43, order, Lam, tjmax, radial, structures, seed=1)
    M, stasks = read_setup(str(sfile))
    mp.mp.prec = precision + 96                         # before any mpmath value of the oracle is made
    D = mp.mpf(D)
    num = lambda x: mp.mpf(x.numerator) / x.denominator if isinstance(x, Fraction) else mp.mpf(x)
    worst_bits = float("inf")
    for i, spins in check:
        q4, sign = structures[i]
        pp, hs = stasks[(tuple(q4), sign)]
        p0 = (1 - sign) // 2
        ns = [p0] if radial else list(range(p0, Lam + 1, 2))
        res = {n: c_residues(M, hs[n], tj12, tj43, tjmax, order, pp) for n in ns}
        for tj in (range(tj12 % 2, tjmax + 1, 2) if spins is None else spins):
            doc = json.load(open(tmp_path / f"out_{i}" / f"b_{tj / 2:g}.json"))
            assert doc["j_internal"] == tj / 2 and doc["four_pt_sign"] == sign and doc["coordinates"] == [coord]
            derivs = doc["derivs"]
            cmp, big = [], mp.mpf(0)
            for p in range(2):
                nl, nr = len(OS.so3(pp[p][0], tj12, tj)), len(OS.so3(pp[p][1], tj43, tj))
                assert len(derivs[p]) == nl and all(len(x) == nr for x in derivs[p])
                for a in range(nl):
                    for b in range(nr):
                        nums = {}
                        for n in ns:
                            hf = lambda pp_, m, j, aa, bb, n=n: num(hs[n].get((pp_, m, j, aa, bb), 0))
                            rf = lambda pp_, m, j, q, aa, bb, n=n: res[n](pp_, m, j, q, aa, bb)
                            nums[n] = DR.numerators(hf, rf, tj, p, a, b, order, Lam - n, tj12, tj43, kept=kept)
                        deg = max(len(poly) for n in ns for poly in nums[n])
                        for k in range(deg):
                            v = [[nums[n][e][k] if e < len(nums[n]) and k < len(nums[n][e]) else mp.mpf(0)
                                  for e in range(Lam - p0 + 1)] for n in ns]
                            if coord == "xt":
                                table = OR.reference_xt(v, Lam, p0, D)
                            else:
                                table = [[x] for x in OR.reference_radial(v[0], Lam, p0, D)]
                            for d0, row in enumerate(table):
                                for d1, want in enumerate(row):
                                    printed = derivs[p][a][b][0][d0][d1]
                                    g = mp.mpf(printed[k]) if k < len(printed) and printed != ["0"] else mp.mpf(0)
                                    cmp.append((g, want))
                                    big = max(big, abs(want))
            worst = max(abs(g - w) for g, w in cmp) / big
            worst_bits = min(worst_bits, -float(mp.log(worst, 2)) if worst else float("inf"))
    return worst_bits


def test_cfg1_scalar_cli(tmp_path):
    base = ("--j-external 0,0,0,0 --j-12 0 --j-43 0 --delta-12 0 --delta-43 0 --j-internal 0-10 --delta-1-plus-2 0 "
            "--order 20 --lambda 3 --coordinates xt --kept-pole-order 6 --precision 256 --num-threads 4")
    run_cli(tmp_path, [f"{base} --four-pt-struct 0,0,0,0 --four-pt-sign 1 -o {tmp_path}/out_0/b_{{}}.json"])
    bits = oracle_files(tmp_path, 256, (0, 0, 0, 0), 0, 0, 20, 3, 20, 6, False, [((0, 0, 0, 0), 1)], "0", [(0, None)], "xt")
    print(f"cfg1: worst 2^-{bits:.1f} relative to each spin file")
    assert bits > 256 - 32


def test_cfg2_spin_half_radial_cli(tmp_path):
    base = ("--j-external 0.5,0.5,0.5,0.5 --j-12 1 --j-43 1 --delta-12 0.3 --delta-43 -0.2 --j-internal 0-20 "
            "--delta-1-plus-2 0 --order 40 --lambda 7 --coordinates xt_radial --kept-pole-order 10 --precision 512 "
            "--num-threads 4")
    structures = [((1, -1, 1, -1), 1), ((1, 1, 1, 1), 1), ((1, -1, -1, 1), -1), ((-1, 1, 1, 1), -1)]
    lines = [f"{base} --four-pt-struct {','.join(f'{x / 2:g}' for x in q)} --four-pt-sign {s} "
             f"-o {tmp_path}/out_{i}/b_{{}}.json" for i, (q, s) in enumerate(structures)]
    run_cli(tmp_path, lines)
    for i in range(len(structures)):
        assert len(os.listdir(tmp_path / f"out_{i}")) == 21 + 1            # 21 spin files and the .profile
    bits = oracle_files(tmp_path, 512, (1, 1, 1, 1), 2, 2, 40, 7, 40, 10, True, structures, "0",
                        [(0, None), (2, [0, 14, 40])], "xt_radial")
    print(f"cfg2: worst 2^-{bits:.1f} relative to each spin file")
    assert bits > 512 - 32


def test_cfg3_heavy_group_bit_exact():
    from blocksvibe_amd import mpf
    from blocksvibe_amd.derivatives import Derivs
    from blocksvibe_amd.plan import Plan
    from blocksvibe_amd.recursion import Recursion
    from oracle import c_chain as CC
    L, B, order, tjmax, pp = 34, 2, 60, 60, ((0, 0), (1, 1))
    plan = Plan(8, 8, tjmax, order, pp, reuse=True)
    rec = Recursion(plan, None, L, seed=3)
    h = mpf.random_device((2, order + 1, plan.Jh, plan.S2, B), L, seed=5)
    der = Derivs(plan, 11, L, kept_pole_order=14)
    der.start(B)
    res = rec.run(h, hook=der.accumulate)
    nums = der.finish(h)
    torch.cuda.synchronize()
    gm = np.ascontiguousarray(res.m.reshape(L, -1).cpu().numpy().T).view(np.uint32)
    ge = res.e.reshape(-1).cpu().numpy()
    Mtab = R_pairs(rec.Mtab)
    hm = np.ascontiguousarray(h.m.reshape(L, -1).cpu().numpy().T).view(np.uint32)
    he = h.e.reshape(-1).cpu().numpy()
    threads = max(1, min(16, len(os.sched_getaffinity(0))))
    print(f"cfg3 (8,8): {sum(l.T for l in plan.levels if l)} tasks x {plan.S2} entries x {B} on {threads} threads",
          flush=True)
    timing = {}
    cm, ce = CC.recursion(plan, Mtab, hm, he, L, B, threads=threads, timing=timing)
    print(f"C recursion: {timing['seconds']:.1f} s", flush=True)
    # both Cauchy kernels met the oracle: two rows per lane on the levels with enough waves, one on the others
    rpl = [w[4] for w in rec.wave_tables(plan.S2 * B) if w is not None]
    assert 1 in rpl and 2 in rpl, rpl
    zero = ge == R.EZ
    assert np.array_equal(zero, ce == R.EZ)
    assert np.array_equal(ge[~zero], ce[~zero])
    assert np.array_equal(gm[~zero], cm[~zero])
    # sampled numerators against the oracle, from the exact residues of a run without slot recycling
    plan_n = Plan(8, 8, tjmax, order, pp, reuse=False)
    rec_n = rec.__class__(plan_n, None, L, seed=3)
    res_n = rec_n.run(h)
    S2, SR = plan.S2, plan.SRm
    got = nums
    mp.mp.prec = 32 * L + 64
    for (tj, p, a, b, k) in ((20, 0, 1, 2, 0), (7 * 2, 1, 0, 0, 1)):
        jo = der.tjs.index(tj)
        ji = tj // 2
        hv = lambda pp_, m, j, aa, bb: mp_of(h, ((((pp_ * (order + 1) + m) * plan.Jh + j) * S2 + aa * SR + bb) * B + k))
        cache = {}

        def rv(pp_, m, j, q, aa, bb):
            key = (pp_, m, j, q, aa, bb)
            if key not in cache:
                slot = rec_n.slot(pp_, m, j, q)
                cache[key] = mp_of(res_n, (slot * S2 + aa * SR + bb) * B + k)
            return cache[key]
        ref = DR.numerators(hv, rv, tj, p, a, b, order, 11, 8, 8, kept=14)
        worst = mp.mpf(0)
        for e, poly in enumerate(ref):
            big = max(abs(x) for x in poly)
            g = [mp_of(got, (((jo * der.Dout + i) * 12 + e) * 2 + p) * S2 * B + (a * SR + b) * B + k)
                 for i in range(der.Dout)]
            assert all(x == 0 for x in g[len(poly):])
            err = max(abs(x - y) for x, y in zip(g, poly)) / big
            worst = max(worst, err)
        bits = float(-mp.log(worst, 2)) if worst else float("inf")
        print(f"cfg3 numerators spin {tj / 2:g} parity {p}: within 2^-{bits:.1f} of the largest coefficient", flush=True)
        # the required agreement for 1024-bit blocks: 2^-(precision - 32)
        assert bits > 1024 - 32, (tj, p, bits)


def R_pairs(x):
    """device MPF -> list of (m, e) pairs"""
    return x.to_host()


def mp_of(x, i):
    """entry i of a device MPF as an mpmath number"""
    L = x.m.shape[0]
    m = x.m.reshape(L, -1)[:, i].cpu().numpy().astype(np.uint32)
    e = int(x.e.reshape(-1)[i])
    v = R.from_limbs(m.tolist())
    return mp.mpf(v) * mp.mpf(2) ** e if v else mp.mpf(0)


def test_cfg4_serving_mix_alone_equals_mix():
    from blocksvibe_amd import mpf
    from blocksvibe_amd.plan import Plan
    from blocksvibe_amd.serve import Request, Server
    order_of = {2: 20, 4: 40, 6: 60, 8: 60}
    shapes = [(0, 0, 8, 2, 10), (1, 1, 5, 4, 18), (2, 2, 6, 6, 26), (4, 2, 4, 8, 34), (2, 0, 4, 2, 42)]
    makers = []
    for op_id, (tj12, tj43, tjmax, nmax, L) in enumerate(shapes):
        pp = ((0, 0), (1, 1))
        order, Lam = order_of[nmax], 2 * nmax - 1
        plan = Plan(tj12, tj43, tjmax, order, pp)
        key = (tj12, tj43, tjmax, order, Lam, 2 * nmax + 2, L, pp, op_id)
        for i, sign in enumerate((1, -1, 1)):
            p0 = (1 - sign) // 2
            hs = [mpf.random_device((2, order + 1, plan.Jh, plan.S2, 1), L, seed=100 * op_id + 10 * i + n)
                  for n in range(p0, Lam + 1, 2)]
            makers.append(lambda key=key, sign=sign, hs=hs: Request(key, sign, hs, coords=("xt", "xt_radial"), d12="1.5"))
    torch.cuda.synchronize()
    srv = Server(streams=2)
    alone = [{k: (v.m.cpu(), v.e.cpu()) for k, v in srv.submit(mk()).wait(600).items()} for mk in makers]
    rs = [srv.submit(mk()) for mk in makers]
    mix = [{k: (v.m.cpu(), v.e.cpu()) for k, v in r.wait(600).items()} for r in rs]
    srv.close()
    print(f"cfg4 mix: batches {srv.batch_sizes}")
    assert max(srv.batch_sizes) > 1
    for a, b in zip(alone, mix):
        assert a.keys() == b.keys()
        for k in a:
            assert torch.equal(a[k][0], b[k][0]) and torch.equal(a[k][1], b[k][1]), k


def test_cfg5_large_rhs_properties():
    from blocksvibe_amd import mpf
    from blocksvibe_amd.derivatives import Derivs
    from blocksvibe_amd.output import Output
    from blocksvibe_amd.plan import Plan
    from blocksvibe_amd.recursion import Recursion
    L, order, tjmax, Lam, kept, pp = 50, 90, 128, 19, 20, ((0, 0), (1, 1))
    outs = {}
    for reuse in (True, False):
        plan = Plan(8, 8, tjmax, order, pp, reuse=reuse)
        print(f"cfg5 reuse={reuse}: {sum(l.T for l in plan.levels if l)} tasks, {plan.nphys} physical slots",
              flush=True)
        rec = Recursion(plan, None, L, seed=11)
        der = Derivs(plan, Lam, L, kept_pole_order=kept)
        out = Output("xt_radial", Lam, 0, "0", L)
        h2 = mpf.random_device((2, order + 1, plan.Jh, plan.S2, 2), L, seed=12)
        for B in ((2,) if not reuse else (1, 2)):
            h = h2 if B == 2 else mpf.MPF(h2.m[..., :1].contiguous(), h2.e[..., :1].contiguous())
            dn = der.for_order(Lam)
            dn.start(B)
            rec.run(h, hook=dn.accumulate)
            r = out.apply_raw([dn.finish(h)])                       # (outputs, Jo, 2, S2, B, D)
            outs[(reuse, B)] = (r.m[..., 0, :].cpu(), r.e[..., 0, :].cpu())
            torch.cuda.synchronize()
        del rec, der
        torch.cuda.empty_cache()
    a = outs[(True, 1)]
    for k in ((True, 2), (False, 2)):
        b = outs[k]
        assert torch.equal(a[0], b[0]) and torch.equal(a[1], b[1]), k
    assert (a[1] != R.EZ).any()


def test_cfg5_chain_captured_in_one_graph():
    """cfg5's whole chain for a batch of two right-hand sides (recursion with the moment hook, numerators,
    write_output) captured as one HIP graph (the serving loop's capture: thread-local mode, a side stream) and
    replayed with its static input refilled: every replay bit-identical to the eager run of the same input,
    with another input in between"""
    from blocksvibe_amd import mpf
    from blocksvibe_amd.derivatives import Derivs
    from blocksvibe_amd.output import Output
    from blocksvibe_amd.plan import Plan
    from blocksvibe_amd.recursion import Recursion
    from blocksvibe_amd.serve import _Graph
    L, order, tjmax, Lam, kept, pp = 50, 90, 128, 19, 20, ((0, 0), (1, 1))
    plan = Plan(8, 8, tjmax, order, pp, reuse=True)
    rec = Recursion(plan, None, L, seed=11)
    der = Derivs(plan, Lam, L, kept_pole_order=kept)
    out = Output("xt_radial", Lam, 0, "0", L)

    def chain(hs):
        h = hs[0]
        dn = der.for_order(Lam)
        dn.start(h.shape[-1])
        rec.run(h, hook=dn.accumulate)
        return out.apply_raw([dn.finish(h)])

    host = lambda r: (r.m.cpu(), r.e.cpu())
    ha = mpf.random_device((2, order + 1, plan.Jh, plan.S2, 2), L, seed=21)
    hb = mpf.random_device((2, order + 1, plan.Jh, plan.S2, 2), L, seed=22)
    want_a, want_b = host(chain([ha])), host(chain([hb]))     # eager: builds every lazily made table
    torch.cuda.synchronize()
    g = _Graph(chain, [ha])
    print(f"cfg5 graph: {plan.nphys} physical slots, pool {g.bytes / 2**30:.2f} GiB", flush=True)
    for h, want in ((ha, want_a), (hb, want_b), (ha, want_a)):
        got = host(g.replay([h]))
        assert torch.equal(got[0], want[0]) and torch.equal(got[1], want[1])
    assert not torch.equal(want_a[0], want_b[0])
