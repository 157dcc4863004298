"""bench.py's host-side pieces (CPU only): the committed PMC summary of the default workload is found and
parsed into the roofline's traffic figure, and only for that workload."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_pmc_traffic_for_the_default_workload():
    t = bench.pmc_traffic(bench.WORKLOAD)
    assert t is not None and t > 0
    kv = dict(l.split("=", 1) for l in open(bench.PMC_SUMMARY).read().splitlines() if "=" in l and not l.startswith("#"))
    assert abs(t - float(kv["traffic_bytes_per_launch "])) < 1


def test_pmc_traffic_only_for_its_workload():
    assert bench.pmc_traffic(bench.WORKLOAD + "; groups heavy only") is None

