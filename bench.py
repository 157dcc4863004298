"""End-to-end throughput of blocks_3d's GPU chain on one MI355X: blocks written per second.

Workload (the nmax=6 spin-2 external batch): four spin-2 externals, every operator group (2j12, 2j43) in
{0,2,4,6,8}^2 (25 groups), 22 four-point structures per group (11 with sign +1, 11 with sign -1), internal
spins 0..30, order 60, Lambda 11 (6 lambda orders per task), kept pole order 14, 1024 bits (34 limbs), xt,
Delta_1+2 = 6: 550 tasks x 31 spins = 17,050 blocks; one block = one (task, internal spin) table = one
blocks_3d JSON spin file.  M matrices and h_inf are seeded synthetic data of the real shapes (the cost of
every stage does not depend on the values).

The 50 (operator group, sign) batches -- 11 structures x 6 lambda orders each -- are dealt into 10 steps of
about equal cost; 10 consecutive steps are one pass over the whole workload.  A step is, for each of its
batches: the radial recursion (Cauchy + pole steps), derivative moments, numerator GEMMs and the write_output
GEMM with inputs resident in HBM, then every spin file of the batch made on the GPU (exact decimal text, JSON
layout), copied to the host and written to local disk by host threads, overlapped with the next batch's
device work.  The batches of the K timed steps go to the lanes (two HIP streams) through one queue, so a lane
starts the next step's work without waiting for the other at a step boundary (--step-sync: a barrier after
every step instead); the timed region ends when the last file of the K steps is closed.  `value` = blocks
(files closed) / wall time of the timed steps.  Files go to a temporary directory and are removed right after they are closed, so a run
does not fill the disk (the write into the page cache is counted; like blocks_3d, nothing is fsync'ed).

Beside it: the device-only rate (one pass without the files), the per-operator constant setup time, the
plain per-task path (B = 1), a mixed serving stream (p50/p99 to files written) and a CPU baseline.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--serve SECONDS] [--no-cpu] [--step-sync]
                    [--keep-files DIR]
"""
import argparse
import json
import os
import shutil
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

ORDER, LAMBDA, KEPT, PREC, TJMAX, D12 = 60, 11, 14, 1024, 60, "6"
TWO_JS = (4, 4, 4, 4)
NQ_PER_SIGN = 11
ORDERS_PER_RUN = 1       # lambda orders per recursion (0: all of a sign's orders in one)
BINS = 10
PP = ((0, 0), (1, 1))
WORKLOAD = ("nmax=6 spin-2 externals: 25 operator groups x 22 structures x 6 lambda orders, order 60, kept 14, "
            "1024 bits, xt")
METRIC = "blocks/s at nmax=6, 1024-bit vs plain GPU & CPU blocks_3d; p50/p99 latency"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10, help="timed steps (10 = one pass over the workload)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--groups", default="all", help="all | heavy (2j12 = 2j43 = 8 only) | comma list a:b")
    ap.add_argument("--plain-tasks", type=int, default=2, help="tasks of the heaviest group on the plain path")
    ap.add_argument("--no-device-pass", action="store_true", help="skip the device-only pass")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline sample")
    ap.add_argument("--serve", type=float, default=20.0, help="seconds of the mixed serving stream (0: skip)")
    ap.add_argument("--writers", type=int, default=16)
    ap.add_argument("--streams", type=int, default=2, help="lanes of operator groups per step, one HIP stream each")
    ap.add_argument("--step-sync", action="store_true",
                    help="synchronise the lanes at every step boundary (default: the K steps' batches in one queue)")
    ap.add_argument("--keep-files", default=None, help="write the files under this directory and keep them")
    ap.add_argument("--json-out", default=None)
    ap.add_argument("--orders-per-run", type=int, default=1,
                    help="lambda orders whose right-hand sides share one recursion (0: all of a sign's orders)")
    return ap.parse_args()


def log(msg):
    print(f"# {msg}", file=sys.stderr, flush=True)


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


def structures(sign):
    """NQ_PER_SIGN distinct four-point structures (doubled q4) of the spin-2 externals for one sign (their labels
    only reach the files' headers; the inputs are synthetic)"""
    from blocksvibe_amd import setup as SU
    p0 = (1 - sign) // 2
    vals = range(-TWO_JS[0], TWO_JS[0] + 1, 2)
    out = [q for q in ((a, b, c, d) for a in vals for b in vals for c in vals for d in vals)
           if SU.four_point_parity(TWO_JS, q) == p0]
    return out[:NQ_PER_SIGN]


class Batch:
    """one (operator group, sign): 11 structures x the sign's lambda orders, resident inputs, and the CLI tasks
    whose spin files it makes"""

    def __init__(self, group, sign, der, outdir, seed):
        from blocksvibe_amd import b3djson as BJ
        from blocksvibe_amd import emit as EM
        from blocksvibe_amd import mpf
        from blocksvibe_amd import structure as S
        from blocksvibe_amd.taskspec import TaskSpec
        self.g, self.sign, self.der = group, sign, der
        self.p0 = (1 - sign) // 2
        pl = group.plan
        self.out = output_map(self.p0, group.L)
        self.ns = list(range(self.p0, LAMBDA + 1, 2))
        self.hs = [mpf.random_device((2, ORDER + 1, pl.Jh, pl.S2, NQ_PER_SIGN), group.L,
                                     seed=seed * 100 + 10 * self.p0 + i) for i in range(len(self.ns))]
        self.tasks = []
        for qi, q4 in enumerate(structures(sign)):
            argv = ["--j-external", ",".join(f"{x / 2:g}" for x in TWO_JS), "--j-internal", f"0-{TJMAX // 2}",
                    "--j-12", f"{group.tj12 / 2:g}", "--j-43", f"{group.tj43 / 2:g}", "--delta-12", "0",
                    "--delta-43", "0", "--delta-1-plus-2", D12, "--four-pt-struct", ",".join(f"{x / 2:g}" for x in q4),
                    "--four-pt-sign", str(sign), "--order", str(ORDER), "--lambda", str(LAMBDA), "--coordinates",
                    "xt", "--kept-pole-order", str(KEPT), "--num-threads", "16", "--precision", str(PREC), "-o",
                    os.path.join(outdir, f"g{group.tj12}_{group.tj43}", f"s{sign}_q{qi}", "block_{}.json")]
            self.tasks.append(TaskSpec(argv, "blocks_3d"))
        # per internal spin: the file's shape (skeleton) and its polynomials for structure 0
        self.p = BJ.digits10(PREC)
        self.eps = BJ.trim_exponent(self.p)
        perm, rows = EM.output_rows(self.out.outputs)
        self.spins = []
        for tj in range(0, TJMAX + 1, 2):
            jo = der.tjs.index(tj)
            structs = [(len(S.so3(PP[k][0], group.tj12, tj)), len(S.so3(PP[k][1], group.tj43, tj))) for k in range(2)]
            sk = EM.skeleton(self.p, structs, [rows])
            base = EM.raw_polys(structs, pl.SRm, jo, der.Jo, pl.S2, NQ_PER_SIGN, 0, [perm], [0])
            self.spins.append((tj, sk, base))

    def blocks(self):
        return len(self.tasks) * len(self.spins)

    def cost(self):
        """relative device cost (recursion terms and products, times the structure entries)"""
        pl = self.g.plan
        pole = sum(lv.T for lv in pl.levels if lv) * (pl.SLm + pl.SRm)
        return pl.S2 * (pl.cauchy_terms() + pole) * len(self.ns)

    def run(self, timer=None):
        """the chain -> write_output result (outputs, Jo, 2, S2, nq, Dmax), on the device: one recursion per
        ORDERS_PER_RUN lambda orders (one by default: with two lanes the end-to-end step measured the same as with
        several orders per recursion, at a smaller HBM footprint; 0: all of a sign's orders in one recursion, each
        order's moments and numerators from its own columns)"""
        from blocksvibe_amd import derivatives as DV
        k = ORDERS_PER_RUN or len(self.ns)
        nums = []
        for i0 in range(0, len(self.ns), k):
            ns, hs = self.ns[i0:i0 + k], self.hs[i0:i0 + k]
            nums += DV.run_orders(self.g.rec, self.der, hs, [LAMBDA - n for n in ns], timer=timer)
        return self.out.apply_raw(nums)

    def jobs(self):
        from blocksvibe_amd import b3djson as BJ
        from blocksvibe_amd.emit import FileJob
        out = []
        for q, t in enumerate(self.tasks):
            for tj, sk, base in self.spins:
                out.append(FileJob(t.path(tj / 2.0), BJ.header(t, tj, PP, self.p), sk, base + q))
        return out

    def emit(self, emitter, r):
        L, D = r.m.shape[0], r.e.shape[-1]
        emitter.emit(r.m.reshape(L, -1), r.e.reshape(-1), D, self.p, self.eps, self.jobs())


class GroupRun:
    """one operator group: plan, recursion constants, derivative constants, and its two sign batches"""

    def __init__(self, tj12, tj43, L, seed, outdir):
        from blocksvibe_amd.derivatives import Derivs
        from blocksvibe_amd.plan import Plan
        from blocksvibe_amd.recursion import Recursion
        self.tj12, self.tj43, self.L = tj12, tj43, L
        self.plan = Plan(tj12, tj43, TJMAX, ORDER, PP, reuse=True)
        self.rec = Recursion(self.plan, None, L, seed=seed)
        der = Derivs(self.plan, LAMBDA, L, kept_pole_order=KEPT)       # R = Lambda covers every lambda order
        self.batches = [Batch(self, s, der, outdir, seed) for s in (1, -1)]

    def run_plain(self, task_limit):
        """plain per-task path: one structure at a time (B = 1), no batching; device only"""
        count = 0
        for b in self.batches:
            for q in range(NQ_PER_SIGN):
                if count >= task_limit:
                    return count
                nums = []
                for n, h in zip(b.ns, b.hs):
                    dn = b.der.for_order(LAMBDA - n)
                    hq = h[:, :, :, :, q:q + 1].contiguous()
                    dn.start(1)
                    self.rec.run(hq, hook=dn.accumulate)
                    nums.append(dn.finish(hq))
                b.out.apply_raw(nums)
                count += 1
        return count


def make_bins(batches, nbins):
    """batches dealt into nbins steps of about equal cost (largest first, each to the cheapest step so far)"""
    bins = [[] for _ in range(nbins)]
    load = [0.0] * nbins
    for b in sorted(batches, key=lambda x: -x.cost()):
        i = min(range(nbins), key=lambda k: load[k])
        bins[i].append(b)
        load[i] += b.cost()
    return [x for x in bins if x]


_STREAMS = []


def step(bins, i, emitter, timer=None, streams=1, count=1, progress=None):
    """count steps from step i: the chain and the spin files of bins i % len(bins) .. (i + count - 1) %
    len(bins); returns their blocks.  streams > 1: the batches run in that many lanes, each lane on its own HIP
    stream driven by its own thread, so that one lane's small levels and file emission share the GPU with
    another lane's chain; the lanes take the batches of all count steps from one queue (step by step, largest
    first within a step), so no lane waits for the other at a step boundary.  progress(done batches, total) is
    called as batches finish."""
    import queue
    import threading
    import torch
    batches = [b for k in range(i, i + count) for b in sorted(bins[k % len(bins)], key=lambda x: -x.cost())]
    nl = min(streams, len(batches))
    done = [0]
    done_lock = threading.Lock()

    def run_one(b):
        r = b.run(timer=timer)
        if emitter is not None:
            b.emit(emitter, r)
        del r
        if progress is not None:
            with done_lock:
                done[0] += 1
                progress(done[0], len(batches))

    if nl <= 1:
        for b in batches:
            run_one(b)
    else:
        # the lanes take the batches from one queue, so that both streams stay busy until the last batch has
        # started (a fixed split leaves one idle when the costs are uneven)
        while len(_STREAMS) < nl:
            _STREAMS.append(torch.cuda.Stream())
        work = queue.Queue()
        for b in batches:
            work.put(b)
        main = torch.cuda.current_stream()
        errors = []

        def worker(k):
            try:
                s = _STREAMS[k]
                s.wait_stream(main)
                with torch.cuda.stream(s):
                    while not errors:
                        try:
                            b = work.get_nowait()
                        except queue.Empty:
                            break
                        run_one(b)
            except BaseException as exc:            # re-raised on the calling thread
                errors.append(exc)

        th = [threading.Thread(target=worker, args=(k,)) for k in range(nl)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        if errors:
            raise errors[0]
    if emitter is not None:
        emitter.drain()
    torch.cuda.synchronize()
    return sum(b.blocks() for b in batches)


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


def roofline(timer):
    """The Cauchy step (one launch per level: k_cauchy_pairs on large levels, k_cauchy_groups on small ones),
    the kernel with the largest share of device time: algorithmic bytes per launch (gathered residues, h_inf
    entries and H entries, each counted once) over its average launch time, both from the device-only pass
    (HIP events recorded around every launch on the launch stream)"""
    import torch
    torch.cuda.synchronize()
    ms = [a.elapsed_time(b) for a, b, _ in timer]
    nbytes = sum(x for _, _, x in timer)
    avg_s = sum(ms) / len(ms) * 1e-3
    per_launch = nbytes / len(timer)
    achieved = per_launch / avg_s / 1e9
    return {"kernel": "Cauchy step: k_cauchy_pairs<34> / k_cauchy_groups<34>", "bound": "hbm", "achieved": achieved,
            "peak": 8000.0, "unit": "GB/s", "frac": achieved / 8000.0, "traffic": pmc_traffic(WORKLOAD),
            "launches": len(timer), "avg_launch_ms": avg_s * 1e3, "algorithmic_bytes_per_launch": per_launch,
            "note": "VALU-bound (one-limb long divisions), not HBM-bound: valu_issue gives the limit that applies",
            "valu_issue": valu_issue()}


COUNTERS_SUMMARY = os.path.join(ROOT, "profiles", "heavy_group_counters.txt")


def valu_issue():
    """the Cauchy kernels' VALU issue utilisation (SQ_INSTS_VALU per launch over the 614.4 G wave-instructions/s
    peak) from the committed rocprofv3 counters of the heaviest operator group, or None without them"""
    out = {}
    try:
        for line in open(COUNTERS_SUMMARY):
            f = line.split()
            if f and f[0] in ("k_cauchy_pairs<34>", "k_cauchy_groups<34>") and f[4].endswith("%"):
                out[f[0]] = float(f[4].rstrip("%")) / 100.0
    except (OSError, IndexError, ValueError):
        return None
    return {"per_kernel": out, "source": "profiles/heavy_group_counters.txt"} if out else None


def _random_pairs(rng, n, L, spread):
    """n random normalised (mantissa, exponent) pairs at L limbs, exponents within +-spread of 2^0"""
    nb = 32 * L - 2
    out = []
    for _ in range(n):
        m = int(rng.integers(1 << 62)) << (nb - 63) | (1 << (nb - 1))
        out.append((m if rng.random() < .5 else -m, int(rng.integers(-spread, spread + 1)) - nb))
    return out


def _host_threads():
    return max(1, len(os.sched_getaffinity(0)))


def chain_products(batch):
    """multiprecision multiply-adds of the stages after the recursion for one batch (all its lambda orders), in
    units of one L x L limb product: the moments (each staged residue times r_c^n, then R + 1 products by the
    exact n^k of KL limbs), the derivative GEMMs of Derivs.finish (c = Wh H, the triangular a_i of TL-limb
    integers, G = Vand v, Hm = Cb G, P, the numerators Anum [a | P]) and the write_output GEMM"""
    pl, B, L = batch.g.plan, NQ_PER_SIGN, batch.g.L
    S2 = pl.S2
    Nc = 2 * S2 * B
    tot, dmax, jo = 0.0, 0, 1
    for n in batch.ns:
        dn = batch.der.for_order(LAMBDA - n)
        R1, Rw, Jo, NP, Dout = dn.R + 1, max(dn.R, 1), dn.Jo, dn.NP, dn.Dout
        tl = dn.tri[0] / L if dn.tri is not None else 1.0
        staged = sum(len(lv[0]) for lv in dn.levels if lv is not None) * S2 * B
        tot += staged * (1.0 + R1 * dn.KL / L)                                     # moments
        tot += R1 * (ORDER + 1) * Jo * Nc                                           # c
        tot += Jo * NP * Nc * R1 * (R1 + 1) / 2 * tl                                # a (triangular)
        tot += Jo * Rw * NP * R1 * Nc + R1 ** 3 * Jo * Nc + R1 ** 2 * Jo * Nc       # G, Hm, P
        tot += Jo * Dout * (NP + R1) * R1 * Nc                                      # numerators
        dmax, jo = max(dmax, Dout), Jo
    W = batch.out.W
    tot += W.shape[0] * W.shape[1] * jo * 2 * S2 * B * dmax                          # write_output
    return tot


def cpu_baseline(groups, seconds=12.0, sample_s=10.0, json_bytes=None, json_rate=None):
    """The oracle's C restatement of the recursion (oracle/c/mpf_c.c through oracle/c_chain.py: the level loop
    of Cauchy and pole steps, bit-identical to the GPU kernels' spec) on every host thread of this process.

    Measured: one whole recursion of a real operator group of the workload at its real shape (order 60,
    2 j_max 60, 1024 bits, the 11 structures of one sign as right-hand sides), the largest group whose
    predicted time fits sample_s.  The other groups are priced by a per-unit model (Cauchy entry, Cauchy term,
    pole dot product, each timed on the same threads) calibrated by that run.  The stages after the recursion
    (moments, derivative and write_output GEMMs: chain_products) are priced at the C dot product's time per L x
    L limb product on the same threads, and the JSON text at the library's multithreaded host formatter's
    measured rate (json_rate bytes/s over json_bytes per pass), when given.  File writes are not counted."""
    import numpy as np
    from oracle import c_chain as CC
    from oracle import c_ref as C
    L = 34
    th = _host_threads()
    part = seconds / 8
    t1 = C.time_unit(0, L, 1, th, part) * 1
    t4 = C.time_unit(0, L, 4, th, part) * 4
    b = (t4 - t1) / 3
    a = t1 - b
    tdot = {k: C.time_unit(1, L, k, th, part) for k in range(1, 6)}

    def per_rhs(pl):
        entries = sum(lv.nH for lv in pl.levels if lv) * pl.S2
        terms = pl.cauchy_terms() * pl.S2
        pole = sum(lv.T for lv in pl.levels if lv) * pl.S2 * (tdot[min(5, pl.SRm)] + tdot[min(5, pl.SLm)])
        return entries * a + terms * b + pole

    fits = [g for g in groups if per_rhs(g.plan) * NQ_PER_SIGN <= sample_s]
    g = max(fits, key=lambda x: per_rhs(x.plan)) if fits else min(groups, key=lambda x: per_rhs(x.plan))
    pl = g.plan
    rng = np.random.default_rng(5)
    Mtab = _random_pairs(rng, max(len(pl.Mkeys), 1) * pl.Smax * pl.Smax, L, 1)
    hm, he = CC.planes(_random_pairs(rng, 2 * (ORDER + 1) * pl.Jh * pl.S2 * NQ_PER_SIGN, L, 2), L)
    tm = {}
    CC.recursion(pl, Mtab, hm, he, L, NQ_PER_SIGN, threads=th, timing=tm)
    measured = tm["seconds"]
    ratio = measured / (per_rhs(pl) * NQ_PER_SIGN)
    rec_s = ratio * sum(per_rhs(x.plan) * NQ_PER_SIGN * len(bt.ns) for x in groups for bt in x.batches)
    tmac = tdot[5] / 5
    prods = sum(chain_products(bt) for x in groups for bt in x.batches)
    post_s = prods * tmac
    json_s = json_bytes / json_rate if json_bytes and json_rate else 0.0
    tot = rec_s + post_s + json_s
    blocks = sum(bt.blocks() for x in groups for bt in x.batches)
    sample_blocks = g.batches[0].blocks() / (measured * len(g.batches[0].ns))
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"value": blocks / tot, "unit": "blocks/s", "cores": th, "kind": "port", "cpu_model": model,
            "host_cpus": os.cpu_count(),
            "pass_s": tot, "recursion_s": rec_s, "post_recursion_s": post_s, "json_text_s": json_s,
            "post_recursion_products": prods, "s_per_limb_product": tmac,
            "measured_group": f"{g.tj12},{g.tj43}", "measured_recursion_s": measured,
            "measured_group_blocks_per_s": sample_blocks, "model_calibration": ratio,
            "sample": f"oracle/c/mpf_c.c level loop (oracle/c_chain.py) on {th} threads: one whole recursion of "
                      f"operator group ({g.tj12},{g.tj43}) at order {ORDER}, 1024 bits, {NQ_PER_SIGN} right-hand "
                      f"sides timed ({measured:.1f} s); the other groups priced by per-unit costs (Cauchy entry "
                      f"{a * 1e9:.0f} ns + {b * 1e9:.0f} ns per term, 5-term pole dot {tdot[5] * 1e9:.0f} ns, "
                      f"{seconds:.0f} s of timing) scaled by that run ({ratio:.2f}x); moments, derivative and "
                      f"write_output GEMMs priced at the C dot's {tmac * 1e9:.0f} ns per 34 x 34 limb product "
                      f"({prods:.3g} per pass); JSON text at the library's host formatter's measured rate; file "
                      f"writes not counted (blocks_3d itself is not on this machine)"}


def leg(res, key, fn):
    """res[key] = fn(), or the error it raised: the measurements beside the headline never cost the JSON line"""
    try:
        res[key] = fn()
    except Exception as e:                    # noqa: BLE001 -- reported in the JSON, the run goes on
        res[key] = {"error": f"{type(e).__name__}: {str(e).splitlines()[0] if str(e) else ''}"}
    log(f"{key}: {res[key]}")


def emission_check(batch, emitter_cls, outdir):
    """one spin file of a batch written by the emitter against the library's host formatter (same bytes)"""
    import numpy as np
    import torch
    from blocksvibe_amd import emit as EM
    r = batch.run()
    L, D = r.m.shape[0], r.e.shape[-1]
    job = batch.jobs()[len(batch.spins) // 2]
    job = EM.FileJob(os.path.join(outdir, "check.json"), job.head, job.skel, job.polys)
    em = emitter_cls(writers=2, buffers=1)
    em.emit(r.m.reshape(L, -1), r.e.reshape(-1), D, batch.p, batch.eps, [job])
    em.close()
    sel = torch.from_numpy(job.polys).to(r.e.device)
    m = r.m.reshape(L, -1, D).index_select(1, sel).cpu().numpy().reshape(L, -1)
    e = r.e.reshape(-1, D).index_select(0, sel).cpu().numpy().reshape(-1)
    local = EM.FileJob(job.path, job.head, job.skel, np.arange(len(job.polys)))
    EM.assemble_host(m, e, D, batch.p, batch.eps, [local])                # warm
    th = time.time()
    want = EM.assemble_host(m, e, D, batch.p, batch.eps, [local])[0]
    th = time.time() - th
    got = open(job.path, "rb").read()
    os.unlink(job.path)
    return {"file_bytes": len(got), "identical_to_host_formatter": got == want, "host_formatter_s": th,
            "host_formatter_gb_per_s": len(want) / 1e9 / max(th, 1e-9), "host_formatter_threads": _host_threads()}


def serving_stream(seconds, outdir, seed=7):
    """The mixed serving stream: single-task requests drawn from external spins {0, 1/2, 1, 2}, nmax 2..8
    (Lambda = 2 nmax - 1, order from nmax, kept 2 nmax + 2), precision {256, 512, 768, 1024, 1280} bits, internal
    spins up to 10..40, two Delta_12/Delta_43 operators per shape, both signs; Poisson arrivals at 0.8 of the
    capacity a first burst measures.  Each request's spin files are made and written by the emitter; its latency
    runs from submission to its last file closed.  Operators, plans and maps are built in a warm-up pass over the
    pool before the stream."""
    import random
    import numpy as np
    import torch
    from blocksvibe_amd import b3djson as BJ
    from blocksvibe_amd import emit as EM
    from blocksvibe_amd import mpf
    from blocksvibe_amd import setup as SU
    from blocksvibe_amd import structure as S
    from blocksvibe_amd.emit import Emitter
    from blocksvibe_amd.plan import Plan
    from blocksvibe_amd.serve import Request, Server
    from blocksvibe_amd.taskspec import TaskSpec
    rng = random.Random(seed)
    order_of = {2: 20, 3: 30, 4: 40, 5: 50, 6: 60, 7: 60, 8: 60}
    pool = []
    for i in range(6):
        ext = rng.choice([0, 1, 2, 4])
        tj12 = rng.choice(range(ext % 2, 2 * ext + 1, 2))
        tj43 = rng.choice(range(ext % 2, 2 * ext + 1, 2))
        nmax = rng.randint(2, 8)
        prec = rng.choice([256, 512, 768, 1024, 1280])
        L = mpf.limbs_for_precision(prec)
        tjmax = rng.choice([20, 40, 60, 80]) + tj12 % 2
        Lam, order = 2 * nmax - 1, order_of[nmax]
        plan = Plan(tj12, tj43, tjmax, order, PP)
        for variant in range(2):
            d12s = (f"{rng.uniform(-0.5, 0.5):.4f}", f"{rng.uniform(-0.5, 0.5):.4f}")
            pool.append({"key": (tj12, tj43, tjmax, order, Lam, 2 * nmax + 2, L, PP, 10 * i + variant), "plan": plan,
                         "ext": ext, "prec": prec, "deltas": d12s})
    srv = Server(streams=2)
    emitter = Emitter(writers=16, buffers=4, unlink=True, gate=srv.gate.device)
    hcache, counter = {}, [0]

    def make(entry, sign, k):
        tj12, tj43, tjmax, order, Lam, kept, L, pp, op_id = entry["key"]
        plan, ext = entry["plan"], entry["ext"]
        p0 = (1 - sign) // 2
        ns = list(range(p0, Lam + 1, 2))
        hk = (op_id, sign, k % 2)
        if hk not in hcache:
            hcache[hk] = [mpf.random_device((2, order + 1, plan.Jh, plan.S2, 1), L, seed=k * 31 + n) for n in ns]
        vals = range(-ext, ext + 1, 2)
        q4 = next((q for q in ((a, b, c, d) for a in vals for b in vals for c in vals for d in vals)
                   if SU.four_point_parity((ext,) * 4, q) == p0), (ext,) * 4)
        counter[0] += 1
        argv = ["--j-external", ",".join([f"{ext / 2:g}"] * 4), "--j-internal", f"{tj12 % 2 / 2:g}-{tjmax / 2:g}",
                "--j-12", f"{tj12 / 2:g}", "--j-43", f"{tj43 / 2:g}", "--delta-12", entry["deltas"][0],
                "--delta-43", entry["deltas"][1], "--four-pt-struct", ",".join(f"{x / 2:g}" for x in q4),
                "--four-pt-sign", str(sign), "--order", str(order), "--lambda", str(Lam), "--coordinates", "xt",
                "--kept-pole-order", str(kept), "--num-threads", "16", "--precision", str(entry["prec"]), "-o",
                os.path.join(outdir, "serve", f"r{counter[0]}", "block_{}.json")]
        task = TaskSpec(argv, "blocks_3d")
        p = BJ.digits10(task.precision)
        eps = BJ.trim_exponent(p)

        def sink(raw, q, done, task=task, p=p, eps=eps):
            out = srv.output("xt", Lam, p0, "0", L)
            res = raw[("xt", "0")]
            _, Jo, _, S2, nq, D = res.shape
            perm, rows = EM.output_rows(out.outputs)
            jobs = []
            for jo, tj in enumerate(task.two_js_internal):
                structs = [(len(S.so3(pp[kk][0], tj12, tj)), len(S.so3(pp[kk][1], tj43, tj))) for kk in range(2)]
                polys = EM.raw_polys(structs, plan.SRm, jo, Jo, S2, nq, q, [perm], [0])
                jobs.append(EM.FileJob(task.path(tj / 2.0), BJ.header(task, tj, pp, p), EM.skeleton(p, structs, [rows]),
                                       polys))
            emitter.emit(res.m.reshape(res.m.shape[0], -1), res.e.reshape(-1), D, p, eps, jobs, on_done=done)

        return Request(entry["key"], sign, hcache[hk], coords=("xt",), d12="0", sink=sink)

    t0 = time.time()
    failed = []

    def wait(r):
        """a request's end; a failed one is counted (its error kept) and the stream goes on"""
        try:
            r.wait(600)
        except Exception as e:                # noqa: BLE001 -- reported in the JSON
            failed.append(f"{type(e).__name__}: {str(e).splitlines()[0] if str(e) else ''}")

    try:
        return _serve(srv, emitter, rng, pool, make, seconds, wait, failed)
    finally:
        srv.close()                               # a second close is a no-op
        emitter.close()


def _serve(srv, emitter, rng, pool, make, seconds, wait, failed):
    import numpy as np
    import torch
    t0 = time.time()
    warm = [srv.submit(make(e, s, 0)) for e in pool for s in (1, -1)]
    for i, r in enumerate(warm):
        wait(r)
        log(f"serving warm-up {i + 1}/{len(warm)} at {time.time() - t0:.1f} s")
    warm_s = time.time() - t0
    for e in pool:                                  # every input of the stream exists before it starts
        for s in (1, -1):
            make(e, s, 1)
    torch.cuda.synchronize()
    srv.batch_sizes.clear()
    srv.mixed_batches = 0
    burst = [srv.submit(make(rng.choice(pool), rng.choice((1, -1)), k)) for k in range(24)]
    t0 = time.time()
    for r in burst:
        wait(r)
    cap = len(burst) / (time.time() - t0)
    burst_batches, burst_mixed = list(srv.batch_sizes), srv.mixed_batches
    srv.mixed_batches = 0
    srv.completed.clear()
    srv.batch_sizes.clear()
    log(f"serving capacity {cap:.2f} requests/s; streaming for {seconds:.0f} s")
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
        wait(r)
    wall = time.time() - t0
    srv.close()
    emitter.drain()
    lat = sorted(r.t_done - r.t_arrival for r in reqs if r.error is None)
    q = lambda f: lat[min(len(lat) - 1, int(f * len(lat)))]
    return {"requests": len(reqs), "rate_per_s": rate, "capacity_per_s": cap, "p50_ms": q(0.5) * 1e3,
            "p99_ms": q(0.99) * 1e3, "blocks_per_s": sum(len(range(r.key[0] % 2, r.key[2] + 1, 2)) for r in reqs) / wall,
            "mean_batch": float(np.mean(srv.batch_sizes)) if srv.batch_sizes else 0.0,
            "burst_mean_batch": float(np.mean(burst_batches)) if burst_batches else 0.0,
            "burst_mixed_operator_batches": burst_mixed, "mixed_operator_batches": srv.mixed_batches,
            "warmup_s": warm_s,
            "operators": len(pool), "graph_replays": srv.replays,
            "graphs": sum(1 for g in srv.graphs.values() if g), "capture_failures": srv.capture_failures,
            "latency_to": "last spin file closed", "files_written": emitter.files, "failed_requests": len(failed),
            "first_failure": failed[0] if failed else None}


def fs_type(path):
    """the file system type holding path (from /proc/mounts: the longest mount point that prefixes it)"""
    best, kind = "", "unknown"
    try:
        for line in open("/proc/mounts"):
            f = line.split()
            if len(f) > 2 and os.path.realpath(path).startswith(f[1].rstrip("/") + "/") and len(f[1]) > len(best):
                best, kind = f[1], f[2]
    except OSError:
        pass
    return kind


def main():
    args = parse()
    global ORDERS_PER_RUN
    ORDERS_PER_RUN = args.orders_per_run
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    import torch
    from blocksvibe_amd import mpf
    from blocksvibe_amd.emit import Emitter
    L = mpf.limbs_for_precision(PREC)
    outdir = args.keep_files or tempfile.mkdtemp(prefix="b3d_bench_")
    t0 = time.time()
    groups = []
    for i, (a, b) in enumerate(group_list(args.groups)):
        groups.append(GroupRun(a, b, L, seed=i + 1, outdir=outdir))
        log(f"set up group {a},{b} ({len(groups)}) at {time.time() - t0:.1f} s")
    torch.cuda.synchronize()
    setup_s = time.time() - t0
    batches = [bt for g in groups for bt in g.batches]
    bins = make_bins(batches, BINS)
    log(f"constants of {len(groups)} operator groups: {setup_s:.1f} s; {len(bins)} steps per pass")
    emitter = Emitter(writers=args.writers, buffers=4, unlink=args.keep_files is None)
    for w in range(args.warmup):
        step(bins, w, emitter, streams=args.streams)
        log(f"warmup step {w} done")
    torch.cuda.synchronize()
    blocks = 0
    per_step = []
    files0, bytes0, write0, wait0 = emitter.files, emitter.bytes, emitter.write_s, emitter.wait_s
    t0 = time.time()
    if args.step_sync:
        for s in range(args.steps):
            ts = time.time()
            blocks += step(bins, args.warmup + s, emitter, streams=args.streams)
            per_step.append(time.time() - ts)
            log(f"step {s} done at {time.time() - t0:.1f} s ({per_step[-1]:.2f} s)")
    else:
        # the K steps' batches in one queue: the lanes go from one step's work to the next without waiting
        # for each other at the step boundaries; the region still holds exactly K steps of work
        nb_total = sum(len(bins[(args.warmup + s) % len(bins)]) for s in range(args.steps))
        report = max(1, nb_total // 10)
        blocks = step(bins, args.warmup, emitter, streams=args.streams, count=args.steps,
                      progress=lambda d, n: d % report == 0 and log(f"{d}/{n} batches done at "
                                                                     f"{time.time() - t0:.1f} s"))
    torch.cuda.synchronize()
    dt = time.time() - t0
    res = {"metric": METRIC, "value": blocks / dt, "unit": "blocks/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "u32 limbs (1086-bit mpf)",
           "data": "synthetic (seeded random M matrices and h_inf of the real shapes)",
           "config": {"workload": WORKLOAD if args.groups == "all" else f"{WORKLOAD}; groups {args.groups} only",
                      "groups": len(groups), "tasks": 2 * NQ_PER_SIGN * len(groups),
                      "blocks_per_pass": sum(b.blocks() for b in batches), "steps_per_pass": len(bins),
                      "blocks_timed": blocks,
                      "step": "chain + GPU JSON text + copy + spin files written and closed"},
           "end_to_end": {"files": emitter.files - files0, "bytes_written": emitter.bytes - bytes0,
                          "gb_per_s_written": (emitter.bytes - bytes0) / 1e9 / max(dt, 1e-9),
                          "host_fallback_units": emitter.fallback_units,
                          "step_s_min": min(per_step) if per_step else None,
                          "step_s_max": max(per_step) if per_step else None,
                          "steps_pipelined": not args.step_sync, "constants_setup_s": setup_s,
                          "writer_busy_s": emitter.write_s - write0, "writers": args.writers,
                          "streams": args.streams,
                          "chain_waited_for_writers_s": emitter.wait_s - wait0, "out_dir": outdir,
                          "out_fs": fs_type(outdir)}}
    emitter.close()
    log(f"end to end: {res['value']:.1f} blocks/s")
    if not args.no_device_pass:
        timer = []
        torch.cuda.synchronize()
        td = time.time()
        nb = sum(step(bins, i, None, timer=timer) for i in range(len(bins)))
        td = time.time() - td
        res["device_only"] = {"blocks_per_s": nb / td, "pass_s": td,
                              "note": "the same chain without the spin files (text, copy, writes)"}
        res["roofline"] = roofline(timer)
        log(f"device only: {nb / td:.1f} blocks/s; roofline {res['roofline']}")
    heavy = max(groups, key=lambda g: g.plan.S2)
    leg(res, "emission_check", lambda: emission_check(heavy.batches[0], Emitter, outdir))

    def plain():
        torch.cuda.synchronize()
        tb = time.time()
        step([heavy.batches], 0, None)
        tb = time.time() - tb
        tp = time.time()
        n = heavy.run_plain(args.plain_tasks)
        torch.cuda.synchronize()
        tp = time.time() - tp
        nbp = n * (TJMAX // 2 + 1)
        batched = sum(b.blocks() for b in heavy.batches) / tb
        return {"group": f"{heavy.tj12},{heavy.tj43}", "tasks": n, "blocks_per_s": nbp / tp,
                "batched_blocks_per_s_same_group": batched, "batched_over_plain": batched / (nbp / tp),
                "note": "device only; plain: one structure at a time (B = 1 per recursion), same kernels"}

    if args.plain_tasks > 0:
        leg(res, "plain_path", plain)
    if args.serve > 0:
        leg(res, "serving", lambda: serving_stream(args.serve, outdir))
    from blocksvibe_amd import rns
    tiles = rns.tuned()
    res["rns_tiles"] = {f"wm{t}": sum(1 for v in tiles.values() if v == t) for t in rns.TILES}
    ms = torch.cuda.memory_stats()
    res["device_memory"] = {"peak_reserved_gb": ms.get("reserved_bytes.all.peak", 0) / 1e9,
                            "peak_allocated_gb": ms.get("allocated_bytes.all.peak", 0) / 1e9,
                            "alloc_retries": ms.get("num_alloc_retries", 0),
                            "device_allocs": ms.get("num_device_alloc", 0), "device_frees": ms.get("num_device_free", 0),
                            "orders_per_run": ORDERS_PER_RUN}
    if not args.no_cpu:
        ec = res.get("emission_check", {})
        per_pass = res["end_to_end"]["bytes_written"] * res["config"]["blocks_per_pass"] / max(blocks, 1)
        rate = ec.get("host_formatter_gb_per_s")
        leg(res, "cpu_baseline", lambda: cpu_baseline(groups, json_bytes=per_pass,
                                                     json_rate=rate * 1e9 if rate else None))
    if not args.keep_files:
        shutil.rmtree(outdir, ignore_errors=True)
    print(json.dumps(res), flush=True)
    if args.json_out:
        with open(args.json_out, "w") as f:
            json.dump(res, f)


if __name__ == "__main__":
    main()
