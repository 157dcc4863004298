"""Throughput of the batched blocks_3d GPU chain on one MI355X.

Workload (the nmax=6 spin-2 external batch): four spin-2 externals, every operator group
(2j12, 2j43) in {0,2,4,6,8}^2 (25 groups), 22 four-point structures per group (11 with sign +1, 11 with
sign -1), internal spins 0..30, order 60, Lambda 11 (6 lambda orders per task), kept pole order 14,
1024 bits (34 limbs), xt coordinates, Delta_1+2 = 6.  550 tasks x 31 spins = 17,050 blocks (one block =
one (task, internal spin) table = one blocks_3d JSON file).  Operators (M matrices) and h_inf are seeded
synthetic data of the real shapes (the cost of every stage does not depend on the values).

One step = the whole chain for every task of the workload, inputs resident in HBM: radial recursion
(Cauchy + pole steps), derivative moments, numerator GEMMs, write_output GEMM.  Per-operator constants
(plan, M tables, numerator and output maps) are built before the timed region.  The JSON text formatting
(host CPU) is not part of a step.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--groups all|heavy] [--plain]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

ORDER, LAMBDA, KEPT, PREC, TJMAX, D12 = 60, 11, 14, 1024, 60, "6"
WORKLOAD = ("nmax=6 spin-2 externals: 25 operator groups x 22 structures x 6 lambda orders, order 60, kept 14, "
            "1024 bits, xt")
TWO_JS = (4, 4, 4, 4)
NQ_PER_SIGN = 11


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--groups", default="all", help="all | heavy (2j12 = 2j43 = 8 only) | comma list a:b")
    ap.add_argument("--plain-tasks", type=int, default=4,
                    help="tasks of the heaviest group run by the plain per-task path (0: skip)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline sample")
    ap.add_argument("--serve", type=float, default=30.0, help="seconds of the mixed serving stream (0: skip)")
    ap.add_argument("--no-json", action="store_true", help="skip the JSON emission sample")
    ap.add_argument("--json-out", default=None)
    return ap.parse_args()


def group_list(spec):
    if spec == "all":
        return [(a, b) for a in (0, 2, 4, 6, 8) for b in (0, 2, 4, 6, 8)]
    if spec == "heavy":
        return [(8, 8)]
    return [tuple(int(x) for x in g.split(":")) for g in spec.split(",")]


_OUTPUTS = {}


def output_map(p0, L):
    from blocksvibe_amd.output import Output
    if (p0, L) not in _OUTPUTS:
        _OUTPUTS[(p0, L)] = Output("xt", LAMBDA, p0, D12, L)
    return _OUTPUTS[(p0, L)]


class GroupRun:
    """one operator group: constants and resident inputs for both signs"""

    def __init__(self, tj12, tj43, L, seed):
        import torch
        from blocksvibe_amd import mpf
        from blocksvibe_amd.derivatives import Derivs
        from blocksvibe_amd.plan import Plan
        from blocksvibe_amd.recursion import Recursion
        self.tj12, self.tj43, self.L = tj12, tj43, L
        self.plan = Plan(tj12, tj43, TJMAX, ORDER, ((0, 0), (1, 1)), reuse=True)
        self.rec = Recursion(self.plan, None, L, seed=seed)
        self.subs = []
        S2, Jh = self.plan.S2, self.plan.Jh
        der = Derivs(self.plan, LAMBDA, L, kept_pole_order=KEPT)     # R = Lambda covers every lambda order
        for sign in (1, -1):
            p0 = (1 - sign) // 2
            out = output_map(p0, L)
            ns = list(range(p0, LAMBDA + 1, 2))
            hs = [mpf.random_device((2, ORDER + 1, Jh, S2, NQ_PER_SIGN), L, seed=seed * 100 + 10 * p0 + i)
                  for i, n in enumerate(ns)]
            self.subs.append((sign, p0, der, out, ns, hs))
        torch.cuda.synchronize()

    def blocks(self):
        return 2 * NQ_PER_SIGN * (TJMAX // 2 + 1)

    def run(self, prof=None, keep=False, timer=None):
        """batched path: per sign, per lambda order one recursion over all structures"""
        results = []
        for sign, p0, der, out, ns, hs in self.subs:
            nums = []
            for n, h in zip(ns, hs):
                dn = der.for_order(LAMBDA - n)
                dn.start(h.shape[-1])
                self.rec.run(h, hook=dn.accumulate, prof=prof, timer=timer)
                nums.append(dn.finish(h))
            r = out.apply_per_order(nums)
            if keep:
                results.append(r)
        return results

    def run_plain(self, task_limit=None):
        """plain per-task path: one structure at a time (B = 1), no batching"""
        count = 0
        for sign, p0, der, out, ns, hs in self.subs:
            for q in range(NQ_PER_SIGN):
                if task_limit is not None and count >= task_limit:
                    return count
                nums = []
                for n, h in zip(ns, hs):
                    dn = der.for_order(LAMBDA - n)
                    hq = h[:, :, :, :, q:q + 1].contiguous()
                    dn.start(1)
                    self.rec.run(hq, hook=dn.accumulate)
                    nums.append(dn.finish(hq))
                out.apply_per_order(nums)
                count += 1
        return count


def serving_stream(seconds, seed=7):
    """Mixed request stream (the serving configuration): single-task requests drawn from external spins
    {0, 1/2, 1, 2}, nmax 2..8 (Lambda = 2 nmax - 1, order 20..60, kept 2 nmax + 2), precision
    {256, 512, 768, 1024, 1280} bits, internal spins up to 20, several Delta_12/Delta_43 operators per shape
    (distinct M matrices), both signs; Poisson arrivals at ~0.8 of the capacity measured by a first burst.
    Operators and maps are built when first seen (a warm-up pass over the pool precedes the stream)."""
    import random
    import numpy as np
    import torch
    from blocksvibe_amd import mpf
    from blocksvibe_amd.plan import Plan
    from blocksvibe_amd.serve import Request, Server
    rng = random.Random(seed)
    order_of = {2: 20, 3: 30, 4: 40, 5: 50, 6: 60, 7: 60, 8: 60}
    pool = []
    for i in range(10):
        ext = rng.choice([0, 1, 2, 4])
        tj12 = rng.choice(range(ext % 2, 2 * ext + 1, 2))
        tj43 = rng.choice(range(ext % 2, 2 * ext + 1, 2))
        nmax = rng.randint(2, 8)
        prec = rng.choice([256, 512, 768, 1024, 1280])
        L = mpf.limbs_for_precision(prec)
        tjmax = rng.choice([8, 12, 16, 20]) + tj12 % 2
        Lam, order = 2 * nmax - 1, order_of[nmax]
        pp = ((0, 0), (1, 1))
        plan = Plan(tj12, tj43, tjmax, order, pp)
        for variant in range(3):                                   # distinct Delta_12, Delta_43
            pool.append(((tj12, tj43, tjmax, order, Lam, 2 * nmax + 2, L, pp, 10 * i + variant), plan))
    hcache = {}

    def make(entry, sign, k):
        key, plan = entry
        p0 = (1 - sign) // 2
        ns = list(range(p0, key[4] + 1, 2))
        hk = (key[:8], sign, k % 2)
        if hk not in hcache:
            hcache[hk] = [mpf.random_device((2, key[3] + 1, plan.Jh, plan.S2, 1), key[6], seed=k * 31 + n)
                          for n in ns]
        return Request(key, sign, hcache[hk], coords=("xt",), d12="0")

    srv = Server(streams=2)
    t0 = time.time()
    warm = [srv.submit(make(e, s, 0)) for e in pool for s in (1, -1)]
    for i, r in enumerate(warm):
        r.wait(600)
        print(f"# serving warm-up {i + 1}/{len(warm)} at {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
    warm_s = time.time() - t0
    for e in pool:                                  # every input of the stream exists before it starts
        for s in (1, -1):
            make(e, s, 1)
    torch.cuda.synchronize()
    burst = [srv.submit(make(rng.choice(pool), rng.choice((1, -1)), k)) for k in range(40)]
    t0 = time.time()
    for r in burst:
        r.wait(600)
    cap = len(burst) / (time.time() - t0)
    srv.completed.clear()
    print(f"# serving capacity {cap:.2f} requests/s; streaming for {seconds:.0f} s", file=sys.stderr, flush=True)
    rate = 0.8 * cap
    reqs, t0, nxt, k = [], time.time(), 0.0, 0
    while time.time() - t0 < seconds:
        now = time.time() - t0
        if now >= nxt:
            reqs.append(srv.submit(make(rng.choice(pool), rng.choice((1, -1)), k)))
            k += 1
            nxt += rng.expovariate(rate)
        else:
            time.sleep(min(0.005, nxt - now))
    for r in reqs:
        r.wait(600)
    wall = time.time() - t0
    srv.close()
    lat = sorted(r.t_done - r.t_arrival for r in reqs)
    q = lambda f: lat[min(len(lat) - 1, int(f * len(lat)))]
    return {"requests": len(reqs), "rate_per_s": rate, "capacity_per_s": cap, "p50_ms": q(0.5) * 1e3,
            "p99_ms": q(0.99) * 1e3, "blocks_per_s": sum(r.blocks for r in reqs) / wall,
            "mean_batch": float(np.mean(srv.batch_sizes)) if srv.batch_sizes else 0.0, "warmup_s": warm_s,
            "operators": len(pool), "graph_replays": srv.replays,
            "graphs": sum(1 for g in srv.graphs.values() if g), "capture_failures": srv.capture_failures,
            "graph_pool_mib": srv.graph_bytes / 2 ** 20}


PMC_SUMMARY = os.path.join(ROOT, "profiles", "k_cauchy_nmax6_1024bit_pmc.txt")


def pmc_traffic(workload):
    """HBM bytes per Cauchy-step launch from the committed rocprofv3 --pmc summary of this workload
    (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled by its reading for a plain copy of known size),
    or None when there is no summary for this workload"""
    try:
        kv = {k.strip(): v.strip() for k, v in (line.split("=", 1) for line in open(PMC_SUMMARY).read().splitlines()
                                                if "=" in line and not line.startswith("#"))}
    except OSError:
        return None
    if kv.get("workload", "") != workload or "traffic_bytes_per_launch" not in kv:
        return None
    return float(kv["traffic_bytes_per_launch"])


def roofline(timer, workload):
    """The Cauchy step (one launch per level: k_cauchy_pairs on large levels, k_cauchy_groups on small
    ones), the kernel with the largest share of device time: algorithmic bytes per launch (gathered
    residues, h_inf entries and H entries, each counted once) over its average launch time, both from the
    timed region (HIP events recorded around every launch on the launch stream)"""
    import torch
    torch.cuda.synchronize()
    ms = [a.elapsed_time(b) for a, b, _ in timer]
    nbytes = sum(x for _, _, x in timer)
    avg_s = sum(ms) / len(ms) * 1e-3
    per_launch = nbytes / len(timer)
    achieved = per_launch / avg_s / 1e9
    return {"kernel": "Cauchy step: k_cauchy_pairs<34> / k_cauchy_groups<34>", "bound": "hbm", "achieved": achieved, "peak": 8000.0, "unit": "GB/s",
            "frac": achieved / 8000.0, "traffic": pmc_traffic(workload), "launches": len(timer),
            "avg_launch_ms": avg_s * 1e3, "algorithmic_bytes_per_launch": per_launch}


def cpu_baseline(runs, seconds=12.0):
    """The oracle's C restatement of the recursion arithmetic (oracle/c/mpf_c.c: Cauchy entries and
    pole-step dot products at 34 limbs, bit-identical to the GPU kernels' spec), timed on all host threads
    available to this process (at most 16) on a bounded sample, scaled by the workload's exact entry and
    term counts.  Derivative and output stages are not counted, so the figure is an upper bound on the
    CPU rate."""
    from oracle import c_ref as C
    L = 34
    th = max(1, min(16, len(os.sched_getaffinity(0))))
    part = seconds / 8
    # Cauchy: cost = entries * a + terms * b, from entries of 1 and of 4 terms
    t1 = C.time_unit(0, L, 1, th, part) * 1
    t4 = C.time_unit(0, L, 4, th, part) * 4
    b = (t4 - t1) / 3
    a = t1 - b
    # pole step per entry of a task's SL x SR result: an SR-term dot (R) and an SL-term dot (q L)
    tdot = {k: C.time_unit(1, L, k, th, part) for k in range(1, 6)}
    tot = 0.0
    for g in runs:
        pl = g.plan
        S2 = pl.S2
        rhs = 2 * NQ_PER_SIGN * len(g.subs[0][4])
        entries = sum(lv.nH for lv in pl.levels if lv) * S2
        terms = pl.cauchy_terms() * S2
        pole = sum(lv.T for lv in pl.levels if lv) * S2 * (tdot[min(5, pl.SRm)] + tdot[min(5, pl.SLm)])
        tot += rhs * (entries * a + terms * b + pole)
    blocks = sum(g.blocks() for g in runs)
    return {"value": blocks / tot, "unit": "blocks/s", "cores": th, "kind": "port",
            "sample": f"oracle/c/mpf_c.c on {th} threads for {seconds:.0f} s: Cauchy entry {a * 1e9:.0f} ns + "
                      f"{b * 1e9:.0f} ns per term, 5-term pole dot {tdot[5] * 1e9:.0f} ns (aggregate per unit), "
                      f"scaled by the workload's recursion entry/term counts; derivative and output stages "
                      f"not counted"}


def json_emission(runs, dt_step, blocks_step):
    """Text emission of the spin files, timed on a sample: the spin files of the heaviest group's first task
    of sign +1 (every kept coefficient of every structure, output and spin), by the library's host
    formatter (b3d_format_polys) on up to 16 host threads and by its GPU formatter (b3d_fmt_*, with the
    transfer of the text), each timed on its second run (allocations cached, as in a run over many tasks).
    The whole step's emission is estimated from the sample's rate and the exact count of coefficients the
    step's spin files hold (before the truncation of negligible top coefficients, so an upper bound)."""
    import torch
    from blocksvibe_amd import b3djson as BJ
    from blocksvibe_amd import structure as S
    threads = max(1, min(16, os.cpu_count() or 1))
    p = BJ.digits10(PREC)
    eps = BJ.trim_exponent(p)
    pp = ((0, 0), (1, 1))

    def layout(g, sign_index, qs):
        """indices of the polynomials that the spin files of tasks qs hold, in the result's
        (Jo, 2, S2, nq, outputs) order"""
        sign, p0, der, out, ns, hs = g.subs[sign_index]
        SR, S2 = g.plan.SRm, g.plan.S2
        no, nq = len(out.outputs), NQ_PER_SIGN
        sel = []
        for jo, tj in enumerate(der.tjs):
            for k in range(2):
                nl = len(S.so3(pp[k][0], g.tj12, tj))
                nr = len(S.so3(pp[k][1], g.tj43, tj))
                for a in range(nl):
                    for b in range(nr):
                        base = ((jo * 2 + k) * S2 + a * SR + b) * nq
                        sel += [(base + q) * no + o for q in qs for o in range(no)]
        return sel

    heavy = max(runs, key=lambda g: g.plan.S2)
    r = heavy.run(keep=True)[0]                                 # (Jo, 2, S2, nq, outputs, Dmax), sign +1
    L, D = r.m.shape[0], r.e.shape[-1]
    sel = torch.tensor(layout(heavy, 0, [0]), dtype=torch.int64, device=r.e.device)   # the first task's files
    m = r.m.reshape(L, -1, D).index_select(1, sel).cpu().numpy()
    e = r.e.reshape(-1, D).index_select(0, sel).cpu().numpy()
    for _ in range(2):                                          # the second run: buffers allocated
        t = time.time()
        bodies = BJ.poly_bodies(m, e, p, eps, threads=threads)
        dt = time.time() - t
    nbytes = sum(len(x) for x in bodies)
    numbers = sum(x.numbers() for x in bodies)
    rate = numbers / dt
    # the same bodies made on the GPU from the device result, copied to the host
    for _ in range(2):                                          # the second run: pinned buffers cached
        torch.cuda.synchronize()
        t = time.time()
        dbodies = BJ.poly_bodies_device(r.m.reshape(L, -1, D), r.e.reshape(-1, D), p, eps, select=sel)
        torch.cuda.synchronize()
        dtd = time.time() - t
    rated = numbers / dtd
    total = 0
    for g in runs:
        for si, (sign, p0, der, out, ns, hs) in enumerate(g.subs):
            total += len(layout(g, si, range(NQ_PER_SIGN))) * der.for_order(LAMBDA - p0).Dout
    est, estd = total / rate, total / rated
    return {"sample": f"group {heavy.tj12},{heavy.tj43}, sign +1, first task's {TJMAX // 2 + 1} spin files: "
                      f"{len(sel)} polynomials x {D} coefficients, {numbers} kept numbers, "
                      f"{nbytes / 1e6:.1f} MB of list bodies",
            "host": {"threads": threads, "seconds": dt, "numbers_per_s": rate, "mb_per_s": nbytes / dt / 1e6,
                     "step_emission_s_estimate": est,
                     "blocks_per_s_with_json_overlapped": blocks_step / max(dt_step, est)},
            "gpu": {"seconds": dtd, "numbers_per_s": rated, "mb_per_s": nbytes / dtd / 1e6,
                    "identical_to_host": dbodies == bodies, "step_emission_s_estimate": estd,
                    "blocks_per_s_with_json_serial": blocks_step / (dt_step + estd)},
            "step_numbers_upper_bound": total,
            "note": "host: formatting on host threads while the GPU runs the next groups; gpu: text made on the "
                    "GPU after each group and copied to the host (serial with the chain); file writes not counted"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    import torch
    from blocksvibe_amd import mpf
    L = mpf.limbs_for_precision(PREC)
    groups = group_list(args.groups)
    t0 = time.time()
    runs = []
    for i, (a, b) in enumerate(groups):
        runs.append(GroupRun(a, b, L, seed=i + 1))
        print(f"# set up group {a},{b} ({i + 1}/{len(groups)}) at {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    setup_s = time.time() - t0
    print(f"# setup of {len(runs)} operator groups: {setup_s:.1f} s", file=sys.stderr, flush=True)
    for w in range(args.warmup):
        for g in runs:
            g.run()
        torch.cuda.synchronize()
        print(f"# warmup step {w} done", file=sys.stderr, flush=True)
    prof = {}
    group_s = {}
    timer = []
    torch.cuda.synchronize()
    t0 = time.time()
    for s in range(args.steps):
        for gi, g in enumerate(runs):
            tg = time.time()
            g.run(timer=timer)
            torch.cuda.synchronize()
            group_s[(g.tj12, g.tj43)] = group_s.get((g.tj12, g.tj43), 0.0) + time.time() - tg
            print(f"# step {s} group {g.tj12},{g.tj43} done at {time.time() - t0:.1f} s", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    dt = time.time() - t0
    blocks = sum(g.blocks() for g in runs) * args.steps
    res = {"metric": "blocks/s at nmax=6, 1024-bit vs plain GPU & CPU blocks_3d; p50/p99 latency",
           "value": blocks / dt, "unit": "blocks/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "u32 limbs (1086-bit mpf)",
           "data": "synthetic (seeded random M matrices and h_inf of the real shapes)",
           "config": {"workload": WORKLOAD if args.groups == "all" else f"{WORKLOAD}; groups {args.groups} only",
                      "groups": len(runs), "tasks": 22 * len(runs), "blocks_per_step": blocks // args.steps}}
    res["roofline"] = roofline(timer, res["config"]["workload"])
    print(f"# roofline: {res['roofline']}", file=sys.stderr, flush=True)
    if args.plain_tasks > 0:
        heavy = max(runs, key=lambda g: g.plan.S2)
        torch.cuda.synchronize()
        tp = time.time()
        n = heavy.run_plain(task_limit=args.plain_tasks)
        torch.cuda.synchronize()
        tp = time.time() - tp
        nb = n * (TJMAX // 2 + 1)
        batched = heavy.blocks() * args.steps / group_s[(heavy.tj12, heavy.tj43)]
        res["plain_path"] = {"group": f"{heavy.tj12},{heavy.tj43}", "tasks": n, "blocks_per_s": nb / tp,
                             "batched_blocks_per_s_same_group": batched, "batched_over_plain": batched / (nb / tp),
                             "note": "plain: one structure at a time (B = 1 per recursion), same kernels"}
        print(f"# plain path: {res['plain_path']}", file=sys.stderr, flush=True)
    if not args.no_json:
        res["json_emission"] = json_emission(runs, dt / args.steps, blocks // args.steps)
        print(f"# json emission: {res['json_emission']}", file=sys.stderr, flush=True)
    if args.serve > 0:
        res["serving"] = serving_stream(args.serve)
        print(f"# serving: {res['serving']}", file=sys.stderr, flush=True)
    from blocksvibe_amd import rns
    tiles = rns.tuned()
    res["rns_tiles"] = {f"wm{t}": sum(1 for v in tiles.values() if v == t) for t in rns.TILES}
    if not args.no_cpu:
        res["cpu_baseline"] = cpu_baseline(runs)
    print(json.dumps(res), flush=True)
    if args.json_out:
        with open(args.json_out, "w") as f:
            json.dump(res, f)


if __name__ == "__main__":
    main()
