"""Captured request pipelines in the serving loop (blocksvibe_amd.serve): a batch shape that has run once is
captured as a HIP graph the next time and replayed afterwards.  Every replay must be bit-identical to the
eager result of the same inputs, including replays with new inputs, with eager work of another operator and
device synchronisations between the replays; and the chains of tasks of different limb counts, orders, nmax and
spin ranges captured together as one graph."""
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


def test_uneven_tasks_in_one_graph():
    """tasks of different limb counts, orders, nmax and spin ranges captured together as one HIP graph (one
    launch replays the chains of all of them): every replay identical to the eager runs of its inputs"""
    from blocksvibe_amd import mpf
    from blocksvibe_amd.derivatives import Derivs
    from blocksvibe_amd.output import Output
    from blocksvibe_amd.plan import Plan
    from blocksvibe_amd.recursion import Recursion
    from blocksvibe_amd.serve import _Graph
    pp = ((0, 0), (1, 1))
    tasks = []
    for i, (tj12, tj43, tjmax, order, L, Lam, kept, B) in enumerate(
            [(2, 2, 6, 10, 10, 5, 6, 2), (4, 2, 12, 20, 18, 7, 8, 1), (0, 0, 16, 14, 26, 3, 4, 3)]):
        plan = Plan(tj12, tj43, tjmax, order, pp)
        tasks.append((plan, Recursion(plan, None, L, seed=40 + i), Derivs(plan, Lam, L, kept_pole_order=kept),
                      Output("xt_radial", Lam, 0, "1.5", L), Lam, L, B))

    def chains(hs):
        outs = []
        for (plan, rec, der, out, Lam, L, B), h in zip(tasks, hs):
            dn = der.for_order(Lam)
            dn.start(B)
            rec.run(h, hook=dn.accumulate)
            outs.append(out.apply_raw([dn.finish(h)]))
        return outs

    def inputs(seed):
        return [mpf.random_device((2, plan.order + 1, plan.Jh, plan.S2, B), L, seed=seed + i)
                for i, (plan, _, _, _, _, L, B) in enumerate(tasks)]

    host = lambda rs: [(r.m.cpu(), r.e.cpu()) for r in rs]
    same = lambda a, b: all(torch.equal(x[0], y[0]) and torch.equal(x[1], y[1]) for x, y in zip(a, b))
    ha, hb = inputs(500), inputs(600)
    want_a, want_b = host(chains(ha)), host(chains(hb))
    torch.cuda.synchronize()
    g = _Graph(chains, ha)
    for h, want in ((ha, want_a), (hb, want_b), (ha, want_a)):
        assert same(host(g.replay(h)), want)
    assert not same(want_a, want_b)
