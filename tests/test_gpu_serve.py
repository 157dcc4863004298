"""The serving loop (blocksvibe_amd.serve): requests of several operators and both signs, first one at a
time, then all submitted at once so that they are batched and overlap on two streams.  Every result must be
bit-identical alone and in the mix (each right-hand-side column is computed independently), batches of
more than one request must have formed, requests of two operators of one shape (other M matrices) must have
shared a batch, and latencies are recorded for every request."""
import pytest
import torch

pytestmark = pytest.mark.gpu


def _requests(L):
    from blocksvibe_amd import mpf
    from blocksvibe_amd.plan import Plan
    from blocksvibe_amd.serve import Request
    specs = []
    for op_id, (tj12, tj43, pp, kept) in enumerate([(2, 2, ((0, 0), (1, 1)), 6), (4, 2, ((0, 1), (1, 0)), 8),
                                                     (2, 2, ((0, 0), (1, 1)), 6)]):
        order, Lam, tjmax = 8, 5, 6
        plan = Plan(tj12, tj43, tjmax, order, pp)
        key = (tj12, tj43, tjmax, order, Lam, kept, L, pp, op_id)
        for i, sign in enumerate((1, 1, -1, 1)):
            p0 = (1 - sign) // 2
            ns = list(range(p0, Lam + 1, 2))
            seed = 1000 * op_id + 10 * i
            specs.append((key, sign, [mpf.random_device((2, order + 1, plan.Jh, plan.S2, 1), L, seed=seed + n)
                                      for n in ns]))
    return [lambda s=s: Request(s[0], s[1], s[2], coords=("xt", "xt_radial"), d12="2.5") for s in specs]


def _host(res):
    return {k: (v.m.cpu(), v.e.cpu()) for k, v in res.items()}


def test_alone_and_in_mix_identical():
    from blocksvibe_amd.serve import Server
    L = 10
    makers = _requests(L)
    srv = Server(streams=2)
    alone = []
    for mk in makers:
        r = srv.submit(mk())
        alone.append(_host(r.wait(300)))
    rs = [srv.submit(mk()) for mk in makers]
    mix = [_host(r.wait(300)) for r in rs]
    srv.close()
    assert max(srv.batch_sizes) > 1, srv.batch_sizes
    assert srv.stacked, "operators 0 and 2 (one shape, other M matrices) never shared a batch"
    for a, b in zip(alone, mix):
        assert a.keys() == b.keys()
        for k in a:
            assert torch.equal(a[k][0], b[k][0]) and torch.equal(a[k][1], b[k][1]), k
    lat = srv.latencies()
    assert len(lat) == 2 * len(makers) and all(x > 0 for x in lat)
    print(f"batches {srv.batch_sizes}; latency p50 {lat[len(lat) // 2] * 1e3:.1f} ms, max {lat[-1] * 1e3:.1f} ms")
