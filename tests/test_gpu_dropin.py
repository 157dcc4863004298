"""The drop-in command line end to end on the GPU: `blocks_3d batch --lines` with a synthetic b3d_setup
helper (same binary format, seeded values), every per-spin JSON file compared with the oracle chain on the
exact setup values: exact recursion (oracle.recursion_ref) -> numerators (oracle.derivs_ref, pole shifting by
exact LU) -> xt coordinates (oracle.output_ref).  Bound: 2^-(precision - 32) relative to each spin file's
largest number (reference tests/test_dropin.py:53-59 uses -(prec - 40)); the text skeleton must parse as JSON
with blocks_3d's nesting [parity][a][b][coordinate][deriv_0][deriv_1][coefficient]."""
import json
import os
import subprocess
import sys
from fractions import Fraction

import mpmath as mp
import numpy as np
import pytest

from oracle import derivs_ref as DR
from oracle import output_ref as OR
from oracle import recursion_ref as RR
from oracle import structure_ref as OS

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

GROUP = ("--j-external 1,1,1,1 --j-12 1 --j-43 1 --delta-12 0 --delta-43 0 --j-internal 0-3 "
         "--delta-1-plus-2 2.3 --order 6 --lambda 3 --coordinates xt --kept-pole-order 4 --precision 256 "
         "--num-threads 2")
TASKS = [("1,0,1,0", 1), ("1,0,1,0", -1), ("0,1,0,-1", 1)]


def read_setup(path):
    raw = np.fromfile(path, dtype=np.int64)
    K = int(raw[1])
    pos = [2]

    def words(k):
        out = raw[pos[0]:pos[0] + k]
        pos[0] += k
        return out

    def num(rec):
        size, ex = int(rec[0]), int(rec[1])
        if size == 0:
            return Fraction(0)
        mag = int.from_bytes(rec[2:].astype("<i8").tobytes(), "little")
        return (-1 if size < 0 else 1) * Fraction(mag) * Fraction(2) ** (64 * (ex - K))
    nM = int(words(1)[0])
    hdr = words(8 * nM).reshape(nM, 8)
    M = {}
    for h in hdr:
        r, c = int(h[6]), int(h[7])
        vals = [num(words(2 + K)) for _ in range(r * c)]
        M[((int(h[0]), int(h[1])), int(h[2]), int(h[3]), int(h[4]), int(h[5]))] = \
            [vals[a * c:(a + 1) * c] for a in range(r)]
    tasks = {}
    for _ in range(int(words(1)[0])):
        t = words(10)
        q4, sign = tuple(int(x) for x in t[:4]), int(t[4])
        pp = ((int(t[5]), int(t[6])), (int(t[7]), int(t[8])))
        hs = {}
        for _ in range(int(t[9])):
            n, cnt = (int(x) for x in words(2))
            idx = words(5 * cnt).reshape(cnt, 5)
            hs[n] = {tuple(int(x) for x in i): num(words(2 + K)) for i in idx}
        tasks[(q4, sign)] = (pp, hs)
    return M, tasks


def cli_env(tmp_path, **extra):
    helper = tmp_path / "b3d_setup"
    helper.write_text(f"#!/bin/bash\nexec {sys.executable} {ROOT}/tests/synthetic_b3d_setup.py \"$@\"\n")
    helper.chmod(0o755)
    env = dict(os.environ, BLOCKSVIBE_B3D_SETUP=str(helper), BLOCKSVIBE_VERBOSE="1",
               BLOCKSVIBE_CACHE=str(tmp_path / "cache"), BLOCKSVIBE_SOCKET=f"/tmp/b3d-{os.getpid()}-{tmp_path.name}.sock",
               BLOCKSVIBE_DAEMON_IDLE="20", PYTHONPATH=ROOT, **extra)
    env.pop("BLOCKSVIBE_FALLBACK", None)
    return env


def stop(env):
    from blocksvibe_amd import cli
    cli.stop_daemon(env["BLOCKSVIBE_SOCKET"])


def test_dropin_batch_json(tmp_path):
    lines = tmp_path / "lines.txt"
    lines.write_text("".join(f"{GROUP} --four-pt-struct {q} --four-pt-sign {s} -o {tmp_path}/out_{i}/b_{{}}.json\n"
                             for i, (q, s) in enumerate(TASKS)))
    env = cli_env(tmp_path)
    try:
        r = subprocess.run([os.path.join(ROOT, "bin", "blocks_3d"), "batch", "--lines", str(lines)], env=env,
                           capture_output=True, text=True, timeout=600)
    finally:
        stop(env)
    print(r.stderr[-3000:])
    assert r.returncode == 0
    # the same synthetic setup file, regenerated, gives the exact inputs
    from blocksvibe_amd.setup import write_synthetic
    q4s = [(tuple(2 * int(x) for x in q.split(",")), s) for q, s in TASKS]
    sfile = tmp_path / "setup.bin"
    write_synthetic(str(sfile), 256, [2, 2, 2, 2], 2, 2, 6, 3, 6, False, q4s, seed=1)
    M, stasks = read_setup(str(sfile))
    mp.mp.prec = 400                          # before any mpmath value of the oracle is made
    order, Lam, kept, tj12, tj43, tjmax, D = 6, 3, 4, 2, 2, 6, mp.mpf("2.3")
    f2m = lambda x: mp.mpf(x.numerator) / x.denominator
    worst_all = mp.mpf(0)
    for i, ((q4, sign), (qs, _)) in enumerate(zip(q4s, TASKS)):
        pp, hs = stasks[(q4, sign)]
        p0 = (1 - sign) // 2
        ns = list(range(p0, Lam + 1, 2))
        res = {n: RR.recursion(hs[n], M, tj12, tj43, tjmax, order, pp) for n in ns}
        for tj in range(0, tjmax + 1, 2):
            path = tmp_path / f"out_{i}" / f"b_{tj // 2}.json"
            doc = json.load(open(path))
            assert doc["j_internal"] == tj / 2 and doc["four_pt_sign"] == sign and doc["coordinates"] == ["xt"]
            derivs = doc["derivs"]
            table_max, cmp = mp.mpf(0), []
            for p in range(2):
                nl, nr = len(OS.so3(pp[p][0], tj12, tj)), len(OS.so3(pp[p][1], tj43, tj))
                assert len(derivs[p]) == nl and all(len(x) == nr for x in derivs[p])
                for a in range(nl):
                    for b in range(nr):
                        nums = {}
                        for n in ns:
                            hf = lambda pp_, m, j, aa, bb, n=n: f2m(hs[n].get((pp_, m, j, aa, bb), Fraction(0)))
                            rf = lambda pp_, m, j, q, aa, bb, n=n: f2m(res[n].get((pp_, m, j, q, aa, bb), Fraction(0)))
                            nums[n] = DR.numerators(hf, rf, tj, p, a, b, order, Lam - n, tj12, tj43, kept=kept)
                        deg = max(len(poly) for n in ns for poly in nums[n])
                        for k in range(deg):
                            num = [[nums[n][e][k] if e < len(nums[n]) and k < len(nums[n][e]) else mp.mpf(0)
                                    for e in range(Lam - p0 + 1)] for n in ns]
                            xt = OR.reference_xt(num, Lam, p0, D)
                            for d0, row in enumerate(xt):
                                for d1, v in enumerate(row):
                                    printed = derivs[p][a][b][0][d0][d1]
                                    g = mp.mpf(printed[k]) if k < len(printed) and printed != ["0"] else None
                                    cmp.append((g, v))
                                    table_max = max(table_max, abs(v))
            worst = max((abs((g if g is not None else 0) - v) for g, v in cmp), default=mp.mpf(0)) / table_max
            worst_all = max(worst_all, worst)
    bits = -float(mp.log(worst_all, 2)) if worst_all else float("inf")
    print(f"drop-in JSON vs oracle chain: worst 2^-{bits:.1f} relative to each spin table")
    assert bits > 256 - 32


def test_concurrent_invocations_share_a_batch(tmp_path):
    """two blocks_3d processes started together with tasks of one operator: the daemon runs them as one device
    batch, and their files equal those of the same tasks run alone in their own process (BLOCKSVIBE_DAEMON=0)"""
    from blocksvibe_amd import cli
    env = cli_env(tmp_path, BLOCKSVIBE_WINDOW="3")
    args = [f"{GROUP} --four-pt-struct {q} --four-pt-sign {s}".split() for q, s in (TASKS[0], TASKS[2])]
    try:
        procs = [subprocess.Popen([os.path.join(ROOT, "bin", "blocks_3d")] + a + ["-o", f"{tmp_path}/c{i}/b_{{}}.json"],
                                  env=env, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True)
                 for i, a in enumerate(args)]
        outs = [p.communicate(timeout=600) for p in procs]
        stats = cli.daemon_stats(env["BLOCKSVIBE_SOCKET"])
    finally:
        stop(env)
    for p, (_, err) in zip(procs, outs):
        print(err[-2000:])
        assert p.returncode == 0
    print("daemon batches:", stats["batch_sizes"])
    assert 2 in stats["batch_sizes"]
    alone = dict(env, BLOCKSVIBE_DAEMON="0")
    for i, a in enumerate(args):
        r = subprocess.run([os.path.join(ROOT, "bin", "blocks_3d")] + a + ["-o", f"{tmp_path}/a{i}/b_{{}}.json"],
                           env=alone, capture_output=True, text=True, timeout=600)
        assert r.returncode == 0, r.stderr[-2000:]
        for tj in range(0, 7, 2):
            got = open(tmp_path / f"c{i}" / f"b_{tj // 2}.json").read().splitlines()
            want = open(tmp_path / f"a{i}" / f"b_{tj // 2}.json").read()
            want = want.replace(f"{tmp_path}/a{i}/", f"{tmp_path}/c{i}/").splitlines()
            first = next((k for k, (x, y) in enumerate(zip(got, want)) if x != y), None)
            assert first is None and len(got) == len(want), (i, tj, first, got[first:first + 2], want[first:first + 2])
