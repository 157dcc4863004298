"""The GPU side of blocks_3d invocations: a serving loop (serve.Server) and a spin-file emitter (emit.Emitter)
that every invocation of a process shares -- the serving daemon's (daemon.py), or the invoking process's own.

An invocation hands over groups of tasks with their b3d_setup file (M matrices and h_inf): each task becomes one
request of the serving loop, whose operator (plan, M-dependent tables, derivative constants) is keyed by the
group and pseudo-parities, so requests of one operator from concurrent invocations run as one device batch.  A
request's sink makes its spin files on the GPU as soon as its batch's write_output result exists; the task is
done when its last file is closed.  Reference: blocksvibe/cli.py:312-498 (one group at a time per process,
behind a GPU lock, files formatted on the CPU afterwards).
"""
import hashlib
import os
import threading
import time

from . import b3djson as BJ
from . import emit as EM
from . import structure as S


def limbs(precision, override=None):
    """L of the GPU numbers: override (BLOCKSVIBE_LIMBS) or 2 ceil(precision / 64) + 2"""
    if override:
        return int(override)
    from .mpf import limbs_for_precision
    return limbs_for_precision(precision)


def operator_id(task, pp):
    """the operator's identity: tasks of one group and pseudo-parities share M and every constant"""
    return hashlib.sha1(repr((task.group_key(), pp)).encode()).hexdigest()[:20]


class TaskRun:
    """one task in flight: its request, and the outcome (None: ok, else the error text)"""
    __slots__ = ("task", "req", "error", "t_submit", "t_done")


class Runner:
    def __init__(self, streams=2, window=0.0, map_cache=None, writers=16):
        from .emit import Emitter
        from .serve import Server
        self.srv = Server(streams=streams, window=window, map_cache=map_cache)
        self.emitter = Emitter(writers=writers, buffers=4, gate=self.srv.gate.device)
        self.lock = threading.Lock()
        self.plans = {}

    def close(self):
        self.srv.close()
        self.emitter.close()

    def stats(self):
        return {"batch_sizes": list(self.srv.batch_sizes), "operators": len(self.srv.ops),
                "files": self.emitter.files, "bytes": self.emitter.bytes}

    def run_group(self, tasks, setup_path, L, cwd, m_path=None):
        """Run one group's tasks (TaskSpec, one group_key) from their setup file; paths relative to cwd.  Blocks
        until every file is written; returns [(error text or None, seconds)] per task."""
        from . import setup as SU
        from .serve import _relax_thread
        _relax_thread()                 # this thread uploads inputs beside the server's captures
        t0 = time.time()
        with self.srv.gate.device():
            M, stasks = SU.read(setup_path, L, m_path=m_path)
        by_q4 = {(st.q4, st.sign): st for st in stasks}
        runs = []
        for t in tasks:
            run = TaskRun()
            run.task, run.req, run.error, run.t_submit, run.t_done = t, None, None, None, None
            runs.append(run)
            st = by_q4.get((tuple(t.q4), t.sign))
            if st is None:
                run.error = f"setup file has no structure q4={t.q4} sign={t.sign}"
                continue
            try:
                with self.srv.gate.device():                          # uploads h_inf: never beside a capture
                    run.req = self._request(t, st, M, L, cwd)
            except Exception as err:                                  # noqa: BLE001 (reported per task)
                run.error = f"{type(err).__name__}: {err}"
        for run in runs:
            if run.req is not None:
                run.t_submit = time.time()
                self.srv.submit(run.req)
        out = []
        for run in runs:
            if run.req is not None:
                try:
                    run.req.wait()
                except Exception as err:                              # noqa: BLE001
                    run.error = f"{type(err).__name__}: {err}"
            out.append((run.error, time.time() - t0))
        return out

    def _request(self, t, st, M, L, cwd):
        from .serve import Request
        pp = st.pp
        order, Lam = t.order, t.Lambda
        kept = min(t.kept_pole_order, order)
        key = (t.two_j12, t.two_j43, t.tjmax, order, Lam, kept, L, pp, operator_id(t, pp))
        plan = self._plan(t.two_j12, t.two_j43, t.tjmax, order, pp)
        hs = [_h_inf(st, n, order, plan, L) for n in t.lambda_orders()]
        coords = [c for c in t.coords_in_order() if c in ("xt", "xt_radial")]
        p = BJ.digits10(t.precision)
        eps = BJ.trim_exponent(p)
        srv, emitter = self.srv, self.emitter
        p0 = t.p0

        def sink(raw, q, done):
            import torch
            parts = [raw[(c, t.delta_1_plus_2)] for c in coords]
            perms, rows, bases, at = [], [], [], 0
            for c, r in zip(coords, parts):
                perm, rr = EM.output_rows(srv.output(c, Lam, p0, t.delta_1_plus_2, L).outputs)
                perms.append(perm)
                rows.append(rr)
                bases.append(at)
                at += r.e.numel() // r.e.shape[-1]
            _, Jo, _, S2, nq, D = parts[0].shape
            if any(r.e.shape[-1] != D for r in parts):
                raise ValueError("coordinates with different polynomial degrees")
            if len(parts) == 1:
                mm, ee = parts[0].m.reshape(L, -1), parts[0].e.reshape(-1)
            else:
                mm = torch.cat([r.m.reshape(L, -1) for r in parts], 1)
                ee = torch.cat([r.e.reshape(-1) for r in parts])
            tjs = list(range(t.two_j12 % 2, t.tjmax + 1, 2))
            jobs = []
            for tj in t.two_js_internal:
                structs = [(len(S.so3(pp[k][0], t.two_j12, tj)), len(S.so3(pp[k][1], t.two_j43, tj))) for k in range(2)]
                polys = EM.raw_polys(structs, plan.SRm, tjs.index(tj), Jo, S2, nq, q, perms, bases)
                jobs.append(EM.FileJob(os.path.join(cwd, t.path(tj / 2.0)), BJ.header(t, tj, pp, p),
                                       EM.skeleton(p, structs, rows), polys))
            emitter.emit(mm, ee, D, p, eps, jobs, on_done=done)

        return Request(key, t.sign, hs, coords=tuple(coords), d12=t.delta_1_plus_2, spins=t.two_js_internal,
                       M=M, sink=sink)

    def _plan(self, tj12, tj43, tjmax, order, pp):
        from .plan import Plan
        k = (tj12, tj43, tjmax, order, pp)
        with self.lock:
            if k not in self.plans:
                self.plans[k] = Plan(tj12, tj43, tjmax, order, pp, reuse=True)
            return self.plans[k]


def _h_inf(st, n, order, plan, L):
    """the task's h_inf at lambda order n as the serving loop takes it: (2, order + 1, Jh, SL SR, 1)"""
    from . import setup as SU
    return SU.h_inf_tensor(st, [n], order, plan.Jh, plan.SLm, plan.SRm, L)


def write_profile(task, cwd, timers):
    """blocks_3d's .profile beside the task's outputs: timers [(name, seconds, runs)], ".main" first"""
    path = os.path.join(cwd, task.path(0.0) + ".profile")
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    BJ.write_profile(path, task, task.num_threads, timers)
