"""The chain on several HIP streams at once, each driven by its own thread (as bench.py's lanes do): two
operator groups, each with both signs, dealt so that one group's two sign batches run concurrently on different
streams and the constants built on first use (per-order derivative tables, wave tables, RNS tables) are first
asked for by two streams at the same time.  Every result must be bit-identical to the same batches run one
after the other on one stream."""
import threading

import pytest
import torch

pytestmark = pytest.mark.gpu

L, ORDER, LAM, TJMAX, KEPT, B = 10, 10, 5, 8, 6, 3
PP = ((0, 0), (1, 1))


def _groups():
    from blocksvibe_amd.derivatives import Derivs
    from blocksvibe_amd.output import Output
    from blocksvibe_amd.plan import Plan
    from blocksvibe_amd.recursion import Recursion
    from blocksvibe_amd import mpf
    out = []
    for gi, (tj12, tj43) in enumerate([(4, 4), (2, 0)]):
        plan = Plan(tj12, tj43, TJMAX, ORDER, PP, reuse=True)
        rec = Recursion(plan, None, L, seed=gi + 1)
        der = Derivs(plan, LAM, L, kept_pole_order=KEPT)
        for sign in (1, -1):
            p0 = (1 - sign) // 2
            ns = list(range(p0, LAM + 1, 2))
            hs = [mpf.random_device((2, ORDER + 1, plan.Jh, plan.S2, B), L, seed=100 * gi + 10 * p0 + i)
                  for i in range(len(ns))]
            out.append((rec, der, Output("xt", LAM, p0, "6", L), ns, hs))
    torch.cuda.synchronize()
    return out


def _run(batch):
    rec, der, outmap, ns, hs = batch
    nums = []
    for n, h in zip(ns, hs):
        dn = der.for_order(LAM - n)
        dn.start(h.shape[-1])
        rec.run(h, hook=dn.accumulate)
        nums.append(dn.finish(h))
    r = outmap.apply_raw(nums)
    return r.m.clone(), r.e.clone()


def test_lanes_on_streams_equal_one_stream():
    lanes = [[0, 3], [1, 2]]              # group 0's two signs on different streams, group 1's likewise
    got = {}
    errors = []
    streams = [torch.cuda.Stream() for _ in lanes]

    def worker(k):
        try:
            with torch.cuda.stream(streams[k]):
                for i in lanes[k]:
                    got[i] = _run(batches[i])
                torch.cuda.current_stream().synchronize()
        except BaseException as exc:
            errors.append(exc)

    batches = _groups()                   # fresh objects: every lazily built table is first built concurrently
    th = [threading.Thread(target=worker, args=(k,)) for k in range(len(lanes))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    assert not errors, errors
    want = [_run(b) for b in _groups()]
    torch.cuda.synchronize()
    for i, (m, e) in enumerate(want):
        assert torch.equal(got[i][0], m) and torch.equal(got[i][1], e), i
