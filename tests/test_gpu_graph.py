"""Captured request pipelines in the serving loop (blocksvibe_amd.serve): a batch shape that has run once is
captured as a HIP graph the next time and replayed afterwards.  Every replay must be bit-identical to the
eager result of the same inputs, including replays with new inputs, with eager work of another operator and
device synchronisations between the replays."""
import pytest
import torch

pytestmark = pytest.mark.gpu


def _req(key, sign, plan, L, seed):
    from blocksvibe_amd import mpf
    from blocksvibe_amd.serve import Request
    order, Lam = key[3], key[4]
    p0 = (1 - sign) // 2
    hs = [mpf.random_device((2, order + 1, plan.Jh, plan.S2, 1), L, seed=seed + n)
          for n in range(p0, Lam + 1, 2)]
    return lambda: Request(key, sign, hs, coords=("xt",), d12="1.5")


def _host(res):
    return {k: (v.m.cpu(), v.e.cpu()) for k, v in res.items()}


def _same(a, b):
    return a.keys() == b.keys() and all(torch.equal(a[k][0], b[k][0]) and torch.equal(a[k][1], b[k][1]) for k in a)


def test_graph_replays_match_eager():
    from blocksvibe_amd.plan import Plan
    from blocksvibe_amd.serve import Server
    L, pp = 10, ((0, 0), (1, 1))
    plan = Plan(4, 2, 6, 10, pp)
    key = (4, 2, 6, 10, 5, 8, L, pp, 0)
    other_plan = Plan(2, 2, 6, 8, pp)
    other = (2, 2, 6, 8, 5, 6, L, pp, 1)
    first, second = _req(key, 1, plan, L, 100), _req(key, 1, plan, L, 200)
    noise = _req(other, -1, other_plan, L, 300)

    eager = Server(streams=1, graphs=False)
    want1 = _host(eager.submit(first()).wait(300))
    want2 = _host(eager.submit(second()).wait(300))
    eager.close()
    assert not _same(want1, want2)

    srv = Server(streams=1, graphs=True)
    got = [_host(srv.submit(first()).wait(300))]         # eager, shape seen
    got.append(_host(srv.submit(first()).wait(300)))    # captured, first replay
    for i in range(3):
        _host(srv.submit(noise()).wait(300))            # eager work of another operator in between
        torch.cuda.synchronize()
        mk = second if i % 2 == 0 else first
        got.append((mk, _host(srv.submit(mk()).wait(300))))
    srv.close()
    assert srv.replays >= 4, srv.replays
    assert _same(got[0], want1) and _same(got[1], want1)
    for mk, res in got[2:]:
        assert _same(res, want2 if mk is second else want1)
    print(f"replays {srv.replays}, graph pools {srv.graph_bytes / 2**20:.1f} MiB")


def test_capture_beside_synchronous_copies():
    """Captures in a worker thread while the submitting thread makes synchronous host-device copies and
    another worker runs eagerly: no capture is invalidated, and the replays equal the eager results."""
    import threading
    from blocksvibe_amd.plan import Plan
    from blocksvibe_amd.serve import Server
    L, pp = 10, ((0, 0), (1, 1))
    plan = Plan(4, 2, 6, 10, pp)
    key = (4, 2, 6, 10, 5, 8, L, pp, 0)
    other_plan = Plan(2, 2, 6, 8, pp)
    other = (2, 2, 6, 8, 5, 6, L, pp, 1)
    first, noise = _req(key, 1, plan, L, 400), _req(other, -1, other_plan, L, 500)

    eager = Server(streams=1, graphs=False)
    want = _host(eager.submit(first()).wait(300))
    eager.close()

    srv = Server(streams=2, graphs=True)
    _host(srv.submit(first()).wait(300))                 # eager: the shape is seen
    got = []
    for _ in range(3):
        r = srv.submit(first())                          # captured the first time, then replayed
        n = srv.submit(noise())
        while not r.done.is_set():                       # synchronous copies in this thread meanwhile
            torch.tensor([1, 2, 3], device="cuda").cpu()
        got.append(_host(r.wait(300)))
        n.wait(300)
    srv.close()
    assert srv.capture_failures == 0, srv.capture_failures
    assert srv.replays >= 2, srv.replays
    assert all(_same(g, want) for g in got)
