"""Device-resident serving loop for blocks_3d requests (continuous batching).

Requests (one blocks_3d task each: operator, four-point sign, lambda orders, h_inf) arrive at any time.
Worker threads, one per HIP stream, repeatedly take the oldest waiting request and every waiting request
compatible with it (same operator shape -- plan, limb count, Lambda, kept pole order -- sign and lambda
orders, any M matrices), up to what the HBM budget still admits, and run them as one batch: one recursion over
all their structures and lambda orders (the pole step reads each right-hand side's own operator tables,
recursion.Operators; each order's moments up to its own derivative count, derivatives.run_orders), the
numerator GEMMs per order and the write_output GEMM.  Requests that arrive while batches are on the GPU join the
next batch as soon as a worker and memory are free, and batches on different streams overlap on the
device.  Plans, uploaded level tables, operator constants (M, qL, R tables) and write_output maps are cached
on the GPU across requests.  A batch shape (operator, sign, batch size, coordinates) that has run once is
captured as a HIP graph the next time and replayed from then on (static inputs refilled by device copies),
within a memory budget for the graphs' private pools.  The HIP runtime judges a potentially unsafe call (a
synchronous copy, a device-wide synchronisation) by the calling thread's capture mode against every capture in
the process, so the worker threads and the threads that submit requests are switched to the relaxed mode, a
capture runs alone (the other workers' device work waits for it) on a fresh stream in the thread-local mode,
and a capture that fails anyway is ended and leaves its batch shape to eager runs.

Every kernel of the chain treats each right-hand-side column independently (per-column alignment, one
rounding per entry, A operands that do not depend on the batch), so a request's result is bit-identical
whether it runs alone or in any mix (tests/test_gpu_serve.py).

Analogue of the reference's spool batching of concurrent invocations and per-GPU lock (reference
blocksvibe/cli.py:238-265, 312-321, 569-637), in one process.
"""
import collections
import contextlib
import sys
import threading
import time

import torch

from . import _lib, mpf
from .derivatives import Derivs, run_orders
from .output import cached as output_cached
from .plan import Plan
from .recursion import Operators, Recursion


class Request:
    """one task: operator (tj12, tj43, tjmax, order, Lambda, kept, L, pp, op_id), sign, h_inf per lambda order
    (device MPF (2, order + 1, Jh, S^2, 1)), output coordinates and Delta_1+2.  op_id distinguishes operators
    with the same shape but different M matrices (Delta_12, Delta_43, external spins); M: the operator's
    matrices as setup.read gives them (None: seeded synthetic ones), used when the operator is first built.

    sink(raw, q, done): optional consumer of the results on the device.  raw maps (coordinate, Delta_1+2) to the
    batch's write_output result (outputs, Jo, 2, S^2, nq, Dmax) (output.Output.apply_raw), q is this request's
    column.  It is called on the worker thread with the batch's stream current and must enqueue everything that
    reads raw on that stream before it returns (the stream is synchronised before raw may be reused); the request
    completes when it calls done() (from any thread), e.g. once its files are written.  Without a sink, result
    holds the request's own copy {(coordinate, Delta_1+2): MPF (Jo, 2, S^2, outputs, Dmax)}."""
    __slots__ = ("key", "sign", "hs", "coords", "d12", "spins", "t_arrival", "t_start", "t_done", "result", "done",
                 "error", "blocks", "M", "sink")

    def __init__(self, key, sign, hs, coords=("xt",), d12="0", spins=None, M=None, sink=None):
        self.key, self.sign, self.hs, self.coords, self.d12, self.spins = key, sign, hs, tuple(coords), d12, spins
        self.M, self.sink = M, sink
        self.t_arrival = self.t_start = self.t_done = None
        self.result, self.error = None, None
        self.done = threading.Event()
        self.blocks = len(spins) if spins else len(range(key[0] % 2, key[2] + 1, 2))

    def wait(self, timeout=None):
        self.done.wait(timeout)
        if self.error:
            raise self.error
        return self.result


class _Operator:
    """cached device state of one operator: the shape-dependent part (plan, level tables, Q values,
    derivative constants) is shared by every operator of the same shape; only the M-dependent tables are
    built per operator"""

    def __init__(self, key, shapes, seed, M=None):
        tj12, tj43, tjmax, order, Lam, kept, L, pp, op_id = key
        sk = (tj12, tj43, tjmax, order, pp, L)
        if sk not in shapes:
            plan = Plan(tj12, tj43, tjmax, order, pp, reuse=True)
            shapes[sk] = {"plan": plan, "rec": Recursion(plan, None, L, seed=0), "derivs": {}}
        self.shape = shapes[sk]
        self.plan = self.shape["plan"]
        self.rec = self.shape["rec"].with_matrices(M, seed=seed)
        self.key = key

    def derivs_for(self, p0):
        tj12, tj43, tjmax, order, Lam, kept, L, pp, op_id = self.key
        dk = (Lam - p0, kept)
        if dk not in self.shape["derivs"]:
            self.shape["derivs"][dk] = Derivs(self.plan, Lam - p0, L, kept_pole_order=kept)
        return self.shape["derivs"][dk]

    def per_rhs_bytes(self, p0):
        d = self.derivs_for(p0)
        pl, L = self.plan, self.key[6]
        S2, nb = pl.S2, 4 * L + 8
        return S2 * nb * ((pl.nphys + 1) + (d.R + 1 + 8) * 2 * d.Jo * d.NP)

    def h_bytes(self):
        pl, L = self.plan, self.key[6]
        return pl.S2 * (4 * L + 8) * 2 * (pl.order + 1) * pl.Jh


_MODE_THREAD_LOCAL, _MODE_RELAXED = 1, 2     # hipStreamCaptureMode
_tls = threading.local()


def _capture_mode(mode):
    """set the calling thread's stream-capture mode, return the previous one"""
    prev = _lib.lib().b3d_thread_capture_mode(mode)
    if prev < 0:
        raise _lib.B3DError(_lib.lib().b3d_last_error().decode())
    return prev


def _relax_thread():
    """relaxed capture checking for the calling thread (once per thread): its synchronous copies and
    synchronisations then leave the captures of the worker threads alone"""
    if not getattr(_tls, "relaxed", False):
        _capture_mode(_MODE_RELAXED)
        _tls.relaxed = True


class _Gate:
    """many eager sections or one capture at a time (a capture waits for the running eager sections and
    holds new ones back)"""

    def __init__(self):
        self.c = threading.Condition()
        self.active, self.capturing, self.waiting = 0, False, 0

    @contextlib.contextmanager
    def eager(self):
        with self.c:
            while self.capturing or self.waiting:
                self.c.wait()
            self.active += 1
        try:
            yield
        finally:
            with self.c:
                self.active -= 1
                self.c.notify_all()

    @contextlib.contextmanager
    def device(self):
        """device work of a helper thread (the file emitter's copies): waits only while a capture runs, not for
        captures that wait (a worker inside an eager section may be waiting on this very helper)"""
        with self.c:
            while self.capturing:
                self.c.wait()
            self.active += 1
        try:
            yield
        finally:
            with self.c:
                self.active -= 1
                self.c.notify_all()

    @contextlib.contextmanager
    def capture(self):
        with self.c:
            self.waiting += 1
            while self.capturing or self.active:
                self.c.wait()
            self.waiting -= 1
            self.capturing = True
        try:
            yield
        finally:
            with self.c:
                self.capturing = False
                self.c.notify_all()


class _Graph:
    """one captured request pipeline (operator, sign, batch size, coordinates): static h_inf inputs, static
    outputs; replays are serialised by a lock, and results are copied out before the next replay"""

    def __init__(self, fn, hs):
        self.hin = [mpf.MPF(h.m.clone(), h.e.clone()) for h in hs]
        work = torch.cuda.current_stream()
        work.synchronize()
        cap = torch.cuda.Stream()
        self.graph = torch.cuda.CUDAGraph()
        before = torch.cuda.memory_allocated()
        # the thread-local mode makes an unsafe call of this thread during the capture an error rather than
        # an eagerly executed step missing from the graph
        prev = _capture_mode(_MODE_THREAD_LOCAL)
        try:
            with torch.cuda.graph(self.graph, stream=cap, capture_error_mode="thread_local"):
                self.out = fn(self.hin)
        except BaseException:
            # a failed capture can leave the capture open and the capture stream current: end it, go back
            _lib.lib().b3d_end_capture(cap.cuda_stream)
            _lib.lib().b3d_clear_hip_error()
            torch.cuda.set_stream(work)
            raise
        finally:
            _capture_mode(prev)
        cap.synchronize()
        self.bytes = max(0, torch.cuda.memory_allocated() - before)
        self.lock = threading.Lock()

    def replay(self, hs):
        for a, b in zip(self.hin, hs):
            a.m.copy_(b.m)
            a.e.copy_(b.e)
        self.graph.replay()
        return self.out


class Server:
    def __init__(self, hbm_budget=None, streams=2, max_batch=64, graphs=True, graph_budget=None, window=0.0,
                 map_cache=None):
        """window: seconds a request that could be batched may wait for companions (requests of the same
        operator from other callers, e.g. concurrent blocks_3d invocations) before its batch starts;
        map_cache: directory keeping write_output maps across processes (output.cached)"""
        self.window, self.map_cache = window, map_cache
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info()
            hbm_budget = int(0.5 * free)
        self.budget = hbm_budget
        self.in_use = 0
        self.max_batch = max_batch
        self.lock = threading.Condition()
        self.queue = collections.deque()
        self.ops, self.plans, self.outs = {}, {}, {}
        self.stacked = {}           # recursion.Operators of mixed-operator batches, by the columns' operators
        self.mixed_batches = 0      # batches whose requests belong to more than one operator
        self.setup_lock = threading.Lock()
        self.stop = False
        self.completed = []
        self.batch_sizes = []
        # captured pipelines: a batch shape seen once runs eagerly (building every constant it needs), the
        # next time it is captured and from then on replayed
        self.use_graphs = graphs
        self.graph_budget = graph_budget if graph_budget is not None else self.budget // 4
        self.graphs, self.seen, self.graph_bytes, self.replays = {}, set(), 0, 0
        self.capture_failures = 0
        self.gate = _Gate()
        self.graph_lock = threading.Lock()
        self.streams = [torch.cuda.Stream() for _ in range(streams)]
        self.workers = [threading.Thread(target=self._work, args=(s,), daemon=True) for s in self.streams]
        for w in self.workers:
            w.start()

    # ------------------------------------------------------------------------------------------ API
    def submit(self, req):
        _relax_thread()
        req.t_arrival = time.time()
        with self.lock:
            self.queue.append(req)
            self.lock.notify_all()
        return req

    def close(self):
        with self.lock:
            self.stop = True
            self.lock.notify_all()
        for w in self.workers:
            w.join()

    def latencies(self):
        return sorted(r.t_done - r.t_arrival for r in self.completed)

    # ------------------------------------------------------------------------------------------ internals
    def _operator(self, key, M=None):
        with self.setup_lock:
            if key not in self.ops:
                self.ops[key] = _Operator(key, self.plans, seed=len(self.ops) + 1, M=M)
            return self.ops[key]

    def output(self, coord, Lam, p0, d12, L):
        """the cached write_output map of (coordinate, Lambda, parity, Delta_1+2, L)"""
        return self._output(coord, Lam, p0, d12, L)

    def _output(self, coord, Lam, p0, d12, L):
        k = (coord, Lam, p0, d12, L)
        with self.setup_lock:
            if k not in self.outs:
                self.outs[k] = output_cached(coord, Lam, p0, d12, L, self.map_cache)
            return self.outs[k]

    def _take_batch(self):
        """oldest request plus compatible ones within the HBM budget (caller holds the lock)"""
        while True:
            if self.stop and not self.queue:
                return None, 0
            if self.queue:
                head = self.queue[0]
                op = self.ops.get(head.key)
                if op is None:                              # first request of an operator: build it unlocked
                    self.lock.release()
                    try:
                        with self.gate.eager():
                            self._operator(head.key, head.M)
                            torch.cuda.current_stream().synchronize()
                    finally:
                        self.lock.acquire()
                    continue
                p0 = (1 - head.sign) // 2
                with self.gate.eager(), self.setup_lock:    # may build derivative constants
                    # the recursion takes each request's right-hand sides once per lambda order
                    per = (op.per_rhs_bytes(p0) + op.h_bytes()) * len(head.hs)
                room = self.budget - self.in_use
                if per <= room or self.in_use == 0:
                    batch = [head]
                    nmax = max(1, min(self.max_batch, room // per if per else self.max_batch))
                    for r in list(self.queue)[1:]:
                        if len(batch) >= nmax:
                            break
                        # the same shape (every key field but the operator id), sign and lambda orders; the
                        # operator of another M must already be built (otherwise it heads a batch later)
                        if (r.key[:-1] == head.key[:-1] and r.sign == head.sign and len(r.hs) == len(head.hs)
                                and (r.key == head.key or r.key in self.ops)):
                            batch.append(r)
                    wait = self.window - (time.time() - head.t_arrival)
                    if wait > 0 and len(batch) < nmax and not self.stop:
                        self.lock.wait(min(wait, 0.05))           # room for companions to arrive
                        continue
                    for r in batch:
                        self.queue.remove(r)
                    cost = per * len(batch)
                    self.in_use += cost
                    self.batch_sizes.append(len(batch))
                    return batch, cost
            self.lock.wait(0.05)

    def _work(self, stream):
        _relax_thread()
        while True:
            with self.lock:
                batch, cost = self._take_batch()
            if batch is None:
                return
            deferred = ()
            try:
                with torch.cuda.stream(stream):
                    deferred = self._run(batch)     # returns with the stream synchronised
                t = time.time()
                for r in batch:
                    if r not in deferred:
                        r.t_done = t
            except Exception as e:          # surface to the callers
                deferred = ()
                for r in batch:
                    r.error = e
                    r.t_done = time.time()
            now = [r for r in batch if r not in deferred]
            with self.lock:
                self.in_use -= cost
                self.completed.extend(now)
                self.lock.notify_all()
            for r in now:
                r.done.set()

    def _complete(self, r, error=None):
        """a sink's done(): the request's results are consumed"""
        r.t_done = time.time()
        if error is not None:
            r.error = error
        with self.lock:
            self.completed.append(r)
        r.done.set()

    def _operators(self, batch):
        """recursion.Operators of a batch whose requests belong to more than one operator (None otherwise),
        cached per assignment of operators to columns (a captured batch keeps reading the same tables).  The
        recursion takes the batch's right-hand sides once per lambda order (derivatives.run_orders), so the
        assignment repeats per order."""
        keys = [r.key for r in batch]
        if all(k == keys[0] for k in keys):
            return None
        nh = len(batch[0].hs)
        with self.setup_lock:
            ops = self.stacked.get((tuple(keys), nh))
            if ops is None:
                order = list(dict.fromkeys(keys))
                ops = Operators([self.ops[k].rec for k in order],
                                [order.index(k) for r in batch for k in [r.key] * r.hs[0].shape[-1]] * nh)
                self.stacked[(tuple(keys), nh)] = ops
                while len(self.stacked) > 64:
                    self.stacked.pop(next(iter(self.stacked)))
            return ops

    def _pipeline(self, op, key, sign, coords, hs, operators=None):
        """device work of one batch: one recursion over the right-hand sides of all its lambda orders with the
        moment hook (derivatives.run_orders; operators: the columns' own operators when they differ), each
        order's numerators, then the write_output map per (coordinate, Delta_1+2)"""
        tj12, tj43, tjmax, order, Lam, kept, L, pp, op_id = key
        p0 = (1 - sign) // 2
        der = op.derivs_for(p0)
        nums = run_orders(op.rec, der, hs, [Lam - (p0 + 2 * i) for i in range(len(hs))], operators=operators)
        return {(c, d): self._output(c, Lam, p0, d, L).apply_raw(nums) for c, d in coords}

    def _run(self, batch):
        t0 = time.time()
        head = batch[0]
        op = self._operator(head.key, head.M)
        with self.gate.eager():                     # device work: never beside a capture
            operators = self._operators(batch)
        if operators is not None:
            self.mixed_batches += 1
        for r in batch:
            r.t_start = t0
        coords = tuple(sorted({(c, r.d12) for r in batch for c in r.coords}))
        nh = len(head.hs)
        inputs = lambda: [mpf.cat([r.hs[i] for r in batch], 4) if len(batch) > 1 else head.hs[i] for i in range(nh)]
        gkey = (tuple(r.key for r in batch), head.sign, nh, coords)
        fn = lambda x: self._pipeline(op, head.key, head.sign, coords, x, operators)
        graph, capture = None, False
        if self.use_graphs:
            with self.graph_lock:
                graph = self.graphs.get(gkey) or None          # None: claimed or failed, run eagerly
                capture = gkey not in self.graphs and gkey in self.seen and self.graph_bytes < self.graph_budget
                if capture:
                    self.graphs[gkey] = None            # claimed: other workers run this shape eagerly meanwhile
                self.seen.add(gkey)
        if capture:
            # no lock is held while capturing: the pipeline takes the setup lock for cached constants
            try:
                with self.gate.capture():
                    graph = _Graph(fn, inputs())
            except Exception as e:                          # the shape stays eager
                graph = None
                msg = str(e).splitlines()[0] if str(e) else type(e).__name__
                print(f"serve: capture of {gkey[:2]} x{len(batch)} failed, running it eagerly ({msg})",
                      file=sys.stderr, flush=True)
            with self.graph_lock:
                self.graphs[gkey] = graph if graph is not None else False
                if graph is not None:
                    self.graph_bytes += graph.bytes
                else:
                    self.capture_failures += 1
        with self.gate.eager():
            hs = inputs()
            if graph is not None:
                with graph.lock:
                    outs = graph.replay(hs)
                    self.replays += 1
                    deferred = self._deliver(batch, outs)
                    torch.cuda.current_stream().synchronize()
            else:
                deferred = self._deliver(batch, fn(hs))
                torch.cuda.current_stream().synchronize()
        return deferred

    def _deliver(self, batch, outs):
        """each request's results: handed to its sink (on the device), or copied out of the batch's outputs
        (which a graph replay reuses).  Returns the requests whose sinks complete them later."""
        deferred = []
        for q, r in enumerate(batch):
            if r.sink is not None:
                r.sink(outs, q, lambda err=None, r=r: self._complete(r, err))
                deferred.append(r)
                continue
            res = {}
            for k, v in outs.items():
                if k[0] in r.coords and k[1] == r.d12:
                    m = v.m[:, :, :, :, :, q].permute(0, 2, 3, 4, 1, 5).contiguous()   # (L, Jo, 2, S2, outputs, D)
                    e = v.e[:, :, :, :, q].permute(1, 2, 3, 0, 4).contiguous()
                    res[k] = mpf.MPF(m, e)
            r.result = res
        return deferred
