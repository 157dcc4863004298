"""The drop-in command line on the CPU (no GPU): blocks_3d option parsing (unique prefixes, -p param
files, precedence of the command line), spin ranges, output templates, task grouping, the fallback of
unsupported coordinates, and batch --lines exit status when the GPU path cannot run (no setup helper and
no fallback binary -> status 1), as reference blocksvibe/cli.py:36-207,503-528,640-676 behave; the daemon's
socket protocol (ping, stats, stop) without a GPU."""
import os
import subprocess
import sys

import pytest

from blocksvibe_amd import cli
from blocksvibe_amd import taskspec as TS

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BASE = ("--j-external 2,2,2,2 --j-12 2 --j-43 2 --delta-12 0 --delta-43 0 --j-internal 0-10 "
        "--delta-1-plus-2 2.3 --four-pt-struct 1,2,1,0 --four-pt-sign 1 --order 16 --lambda 7 --coordinates xt "
        "--kept-pole-order 16 --precision 256 --num-threads 2 -o out/b_{}.json").split()


def test_parse_prefix_and_param_file(tmp_path):
    p = tmp_path / "t.param"
    p.write_text("# comment\n[section]\norder = 20\nlambda=9\n")
    v = TS.read_options(["--j-ext", "2,2,2,2", "--lambda=7", "-p", str(p)])
    assert v["j-external"] == "2,2,2,2" and v["lambda"] == "7" and v["order"] == "20"
    with pytest.raises(TS.UsageError):
        TS.read_options(["--j", "1"])                      # ambiguous prefix
    with pytest.raises(TS.UsageError):
        TS.read_options(["--order", "1", "--order", "2"])


def test_spins_and_template():
    assert TS.spin_set("0-10", 2) == [0, 2, 4, 6, 8, 10, 12, 14, 16, 18, 20]
    assert TS.spin_set("0.5-3.5,5.5", 1) == [1, 3, 5, 7, 11]
    assert TS.spin_set("2,4-13", 2) == [4, 8, 10, 12, 14, 16, 18, 20, 22, 24, 26]
    with pytest.raises(TS.UsageError):
        TS.spin_set("0.5", 2)
    assert TS.fill_template("x/b_{}.json", 2.0) == "x/b_2.json"
    assert TS.fill_template("x/b_{}.json", 2.5) == "x/b_2.5.json"
    assert TS.fill_template("x/b_{:.1f}.json", 2.0) == "x/b_2.0.json"


def test_task_and_groups():
    t = TS.TaskSpec(BASE)
    assert t.two_js == [4, 4, 4, 4] and t.q4 == [2, 4, 2, 0] and t.p0 == 0
    assert t.lambda_orders() == [0, 2, 4, 6] and t.tjmax == 20 and t.unsupported() is None
    u = TS.TaskSpec(BASE[:-4] + ["--num-threads", "1", "-o", "o_{}.json"])
    assert t.group_key() == u.group_key()
    z = TS.TaskSpec([a if a != "xt" else "zzb" for a in BASE])
    assert z.unsupported()
    with pytest.raises(TS.UsageError):
        TS.TaskSpec([a if a != "1,2,1,0" else "1,2,1" for a in BASE])


def test_batch_without_helper_reports_failure(tmp_path):
    lines = tmp_path / "lines.txt"
    lines.write_text(" ".join(BASE) + "\n# a comment\n")
    env = dict(os.environ, BLOCKSVIBE_B3D_SETUP=str(tmp_path / "missing"), BLOCKSVIBE_DAEMON="0")
    env.pop("BLOCKSVIBE_FALLBACK", None)
    r = subprocess.run([sys.executable, "-m", "blocksvibe_amd", "batch", "--lines", str(lines)], cwd=tmp_path,
                       env=dict(env, PYTHONPATH=ROOT), capture_output=True, text=True, timeout=300)
    assert r.returncode == 1, r.stderr
    assert "no fallback" in r.stderr


def test_fallback_is_used_for_unsupported(tmp_path):
    stub = tmp_path / "stub.sh"
    stub.write_text("#!/bin/bash\necho \"$@\" > " + str(tmp_path / "called.txt") + "\nexit 0\n")
    stub.chmod(0o755)
    argv = [a if a != "xt" else "zzb" for a in BASE]
    env = dict(os.environ, BLOCKSVIBE_FALLBACK=str(stub), PYTHONPATH=ROOT)
    r = subprocess.run([sys.executable, "-m", "blocksvibe_amd"] + argv, cwd=tmp_path, env=env,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert "--coordinates zzb" in (tmp_path / "called.txt").read_text()


def test_param_file_errors(tmp_path):
    p = tmp_path / "bad.param"
    p.write_text("order = 2\nnot an option\n")
    with pytest.raises(TS.UsageError, match="unrecognised line in param file"):
        TS.read_options(["-p", str(p)])
    with pytest.raises(TS.UsageError, match="Could not open"):
        TS.read_options(["-p", str(tmp_path / "none")])
    with pytest.raises(TS.UsageError, match="requires an argument"):
        TS.read_options(["--order"])
    with pytest.raises(TS.UsageError, match="unexpected argument"):
        TS.read_options(["order"])
    ok = tmp_path / "ok.param"
    ok.write_text("lambda=3\n")
    assert TS.read_options([f"-p{ok}"])["lambda"] == "3"                # value glued to the short option
    assert TS.read_options(["-h"]) == {"help": "1"}


def test_batch_argvs(tmp_path):
    lines = tmp_path / "l.txt"
    lines.write_text("--order 3 -o 'a b/{}.json'\n\n  # skipped\n--order 4\n")
    assert cli.batch_argvs(["--lines", str(lines), "x.param"]) == [["--order", "3", "-o", "a b/{}.json"],
                                                                   ["--order", "4"], ["-p", "x.param"]]


def test_daemon_protocol_without_gpu(tmp_path):
    """a daemon process answers ping and stop on its socket and removes the socket when it exits"""
    import time
    path = f"/tmp/b3d-test-{os.getpid()}.sock"
    env = dict(os.environ, PYTHONPATH=ROOT)
    proc = subprocess.Popen([sys.executable, "-m", "blocksvibe_amd.daemon", "--socket", path, "--idle", "60"],
                            env=env, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    try:
        for _ in range(300):
            if os.path.exists(path) and cli._ask(path, {"op": "ping"}, 10):
                break
            time.sleep(0.1)
        assert cli._ask(path, {"op": "ping"}, 10)["ok"]
        assert cli._ask(path, {"op": "nonsense"}, 10) == {"ok": False, "error": "unknown op 'nonsense'"}
        assert cli.stop_daemon(path)
        assert proc.wait(60) == 0
        assert not os.path.exists(path)
    finally:
        if proc.poll() is None:
            proc.kill()
