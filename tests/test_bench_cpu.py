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


class _B:
    def __init__(self, g, c):
        self.g, self.c = g, c

    def cost(self):
        return self.c


def test_lanes_balance():
    g = [object() for _ in range(5)]
    bs = [_B(g[0], 9), _B(g[0], 8), _B(g[1], 5), _B(g[1], 4), _B(g[2], 7), _B(g[3], 3), _B(g[4], 2), _B(g[4], 1)]
    ls = bench.lanes(bs, 2)
    assert sorted(id(b) for l in ls for b in l) == sorted(id(b) for b in bs)
    loads = sorted(sum(b.c for b in l) for l in ls)
    assert loads == [19, 20]
    assert bench.lanes(bs[:1], 3) == [bs[:1]]
