"""The drop-in command line on the CPU (no GPU): blocks_3d option parsing (unique prefixes, -p param
files, precedence of the command line), spin ranges, output templates, task grouping, the fallback of
unsupported coordinates, and batch --lines exit status when the GPU path cannot run (no setup helper and
no fallback binary -> status 1), as reference blocksvibe/cli.py:36-207,503-528,640-676 behave."""
import os
import subprocess
import sys

import pytest

from blocksvibe_amd import cli

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BASE = ("--j-external 2,2,2,2 --j-12 2 --j-43 2 --delta-12 0 --delta-43 0 --j-internal 0-10 "
        "--delta-1-plus-2 2.3 --four-pt-struct 1,2,1,0 --four-pt-sign 1 --order 16 --lambda 7 --coordinates xt "
        "--kept-pole-order 16 --precision 256 --num-threads 2 -o out/b_{}.json").split()


def test_parse_prefix_and_param_file(tmp_path):
    p = tmp_path / "t.param"
    p.write_text("# comment\n[section]\norder = 20\nlambda=9\n")
    v = cli.parse_args(["--j-ext", "2,2,2,2", "--lambda=7", "-p", str(p)])
    assert v["j-external"] == "2,2,2,2" and v["lambda"] == "7" and v["order"] == "20"
    with pytest.raises(cli.UsageError):
        cli.parse_args(["--j", "1"])                      # ambiguous prefix
    with pytest.raises(cli.UsageError):
        cli.parse_args(["--order", "1", "--order", "2"])


def test_spins_and_template():
    assert cli.spins("0-10", 2) == [0, 2, 4, 6, 8, 10, 12, 14, 16, 18, 20]
    assert cli.spins("0.5-3.5,5.5", 1) == [1, 3, 5, 7, 11]
    assert cli.spins("2,4-13", 2) == [4, 8, 10, 12, 14, 16, 18, 20, 22, 24, 26]
    with pytest.raises(cli.UsageError):
        cli.spins("0.5", 2)
    assert cli.fmt_template("x/b_{}.json", 2.0) == "x/b_2.json"
    assert cli.fmt_template("x/b_{}.json", 2.5) == "x/b_2.5.json"
    assert cli.fmt_template("x/b_{:.1f}.json", 2.0) == "x/b_2.0.json"


def test_task_and_groups():
    t = cli.Task(BASE)
    assert t.two_js == [4, 4, 4, 4] and t.q4 == [2, 4, 2, 0] and t.p0 == 0
    assert t.lambda_orders() == [0, 2, 4, 6] and t.tjmax == 20 and t.unsupported() is None
    u = cli.Task(BASE[:-4] + ["--num-threads", "1", "-o", "o_{}.json"])
    assert t.group_key() == u.group_key()
    z = cli.Task([a if a != "xt" else "zzb" for a in BASE])
    assert z.unsupported()
    with pytest.raises(cli.UsageError):
        cli.Task([a if a != "1,2,1,0" else "1,2,1" for a in BASE])


def test_batch_without_helper_reports_failure(tmp_path):
    lines = tmp_path / "lines.txt"
    lines.write_text(" ".join(BASE) + "\n# a comment\n")
    env = dict(os.environ, BLOCKSVIBE_B3D_SETUP=str(tmp_path / "missing"), BLOCKSVIBE_SPOOL_OFF="1")
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
