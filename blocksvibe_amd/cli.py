"""Drop-in blocks_3d command line on MI355X (reference blocksvibe/cli.py:1-680).

    blocks_3d [blocks_3d options]             same options and param file (-p) as blocks_3d, same per-spin
                                              JSON files and .profile
    blocks_3d batch [--lines FILE] [PARAM_FILE ...]
                                              many tasks; tasks of one (j_external, j12, j43, Delta_12,
                                              Delta_43, order, Lambda, j_max, precision, kept) group share
                                              one setup run and one GPU operator, their right-hand sides
                                              (q4, lambda order) run batched

Supported on the GPU: coordinates xt and/or xt_radial, with or without pole shifting.  Anything else, or a
failure of the GPU path (including a missing b3d_setup helper), runs the official blocks_3d instead:
$BLOCKSVIBE_FALLBACK with $BLOCKSVIBE_JEMALLOC preloaded if set.

Environment: BLOCKSVIBE_B3D_SETUP (helper), BLOCKSVIBE_LIMBS, BLOCKSVIBE_MAX_RHS, BLOCKSVIBE_CACHE (write_output
maps and M matrices, default ~/.cache/blocksvibe), BLOCKSVIBE_LOCKDIR, BLOCKSVIBE_SPOOL / _SPOOL_QUIET /
_SPOOL_MAX / _SPOOL_OFF=1, BLOCKSVIBE_VERBOSE=1.
"""
import fcntl
import hashlib
import os
import re
import shlex
import subprocess
import sys
import tempfile
import time

OPTS = {  # blocks_3d option -> (attribute, default); None = required
    "j-external": ("j_external", None), "j-internal": ("j_internal", None), "j-12": ("j12", None),
    "j-43": ("j43", None), "delta-12": ("delta_12", None), "delta-43": ("delta_43", None),
    "delta-1-plus-2": ("delta_1_plus_2", "0"), "four-pt-struct": ("four_pt_struct", None),
    "four-pt-sign": ("four_pt_sign", None), "order": ("order", None), "lambda": ("lambda", None),
    "coordinates": ("coordinates", "zzb"), "kept-pole-order": ("kept_pole_order", None),
    "num-threads": ("num_threads", None), "precision": ("precision", None),
    "out-template": ("out_template", None), "debug": ("debug", "false"), "profile": ("profile", "true"),
    "param-file": ("param_file", ""), "help": ("help", None),
}
SHORT = {"o": "out-template", "p": "param-file", "h": "help"}
GPU_COORDS = {"xt", "xt_radial"}


class UsageError(Exception):
    pass


def log(*a):
    if os.environ.get("BLOCKSVIBE_VERBOSE") == "1":
        print("[blocksvibe]", *a, file=sys.stderr, flush=True)


def _long(name):
    """boost::program_options long-name matching: exact or unique prefix"""
    if name in OPTS:
        return name
    cand = [k for k in OPTS if k.startswith(name)]
    if len(cand) != 1:
        raise UsageError(f"unrecognised option '--{name}'")
    return cand[0]


def parse_args(argv):
    """command line first, then the param file (key=value) for options not given"""
    vals, i = {}, 0
    while i < len(argv):
        a = argv[i]
        if a.startswith("--"):
            name, eq, v = a[2:].partition("=")
            name = _long(name)
            has = bool(eq)
        elif a.startswith("-") and len(a) >= 2 and a[1] in SHORT:
            name, v = SHORT[a[1]], a[2:]
            has = v != ""
        else:
            raise UsageError(f"unexpected argument '{a}'")
        if name == "help":
            vals["help"] = "1"
            i += 1
            continue
        if not has:
            if i + 1 >= len(argv):
                raise UsageError(f"option '--{name}' requires an argument")
            v = argv[i + 1]
            i += 1
        if name in vals:
            raise UsageError(f"option '--{name}' cannot be specified more than once")
        vals[name] = v
        i += 1
    if vals.get("param-file"):
        try:
            lines = open(vals["param-file"]).read().splitlines()
        except OSError:
            raise UsageError(f"Could not open '{vals['param-file']}'")
        for line in lines:
            line = line.split("#", 1)[0].strip()
            if not line or line.startswith("["):
                continue
            k, eq, v = line.partition("=")
            k = k.strip()
            if not eq or k not in OPTS:
                raise UsageError(f"unrecognised line in param file: '{line}'")
            vals.setdefault(k, v.strip())
    return vals


def _two(s):
    x = 2 * float(s)
    if x != int(x):
        raise UsageError(f"not a whole or half integer: {s}")
    return int(x)


def _bool(s):
    s = s.strip().lower()
    if s in ("1", "true", "yes", "on"):
        return True
    if s in ("0", "false", "no", "off"):
        return False
    raise UsageError(f"not a boolean: {s}")


def spins(spec, two_j12):
    """j-internal ranges '0.5-3.5,5.5,8.5-16.5-2' -> sorted doubled spins"""
    out = set()
    for s in spec.split(","):
        parts = s.split("-")
        if not s or len(parts) > 3:
            raise UsageError(f"Found invalid element in j-internal: '{s}'")
        first = _two(parts[0])
        last = _two(parts[1]) if len(parts) > 1 else first
        if first % 2 != two_j12 % 2:
            raise UsageError("j_12 and j_internal must all be either whole or half integers")
        step = 2 * int(parts[2]) if len(parts) > 2 else 2
        if (last - first) % step:
            raise UsageError(f"The step of a spin range in j_internal must exactly divide the difference: {s}")
        out.update(range(first, last + 1, step))
    return sorted(out)


class Task:
    """one blocks_3d invocation (normalised options)"""

    def __init__(self, argv, prog="blocks_3d"):
        self.argv = [prog] + list(argv)
        v = parse_args(argv)
        self.help = "help" in v
        if self.help:
            return
        missing = [k for k, (_, d) in OPTS.items() if d is None and k != "help" and k not in v]
        if missing:
            raise UsageError(f"the option '--{missing[0]}' is required but missing")
        g = lambda k: v.get(k, OPTS[k][1])
        self.two_j12, self.two_j43 = _two(g("j-12")), _two(g("j-43"))
        if self.two_j12 % 2 != self.two_j43 % 2:
            raise UsageError("j_12 and j_43 must both be either integers or half integers")
        self.two_js = [_two(x) for x in g("j-external").split(",")]
        self.q4 = [_two(x) for x in g("four-pt-struct").split(",")]
        if len(self.two_js) != 4 or len(self.q4) != 4:
            raise UsageError("j-external and four-pt-struct take 4 values")
        self.sign = int(g("four-pt-sign"))
        if self.sign not in (1, -1):
            raise UsageError(f"Four-pt-sign must be 1 or -1, but got: {self.sign}")
        for i in range(4):
            if (self.q4[i] - self.two_js[i]) % 2 or not -self.two_js[i] <= self.q4[i] <= self.two_js[i]:
                raise UsageError(f"invalid four-pt-struct entry {i}")
        self.two_js_internal = spins(g("j-internal"), self.two_j12)
        self.delta_12, self.delta_43, self.delta_1_plus_2 = g("delta-12"), g("delta-43"), g("delta-1-plus-2")
        self.order, self.Lambda = int(g("order")), int(g("lambda"))
        self.kept_pole_order, self.num_threads = int(g("kept-pole-order")), int(g("num-threads"))
        self.precision = int(g("precision"))
        self.out_template = g("out-template")
        self.coordinates = set(g("coordinates").split(","))
        bad = self.coordinates - {"zzb", "yyb", "xt", "ws", "xt_radial", "ws_radial"}
        if bad:
            raise UsageError(f"Invalid entry for coordinates: '{bad.pop()}'")
        self.radial = all(c.endswith("_radial") for c in self.coordinates)
        self.debug, self.profile = _bool(g("debug")), _bool(g("profile"))
        self.p0 = (1 - self.sign) // 2
        self.tjmax = self.two_js_internal[-1]

    def unsupported(self):
        if not self.coordinates <= GPU_COORDS:
            return f"coordinates {sorted(self.coordinates - GPU_COORDS)} (GPU: xt, xt_radial)"
        return None

    def group_key(self):
        """tasks sharing the recursion operator and the M matrices"""
        return (tuple(self.two_js), self.two_j12, self.two_j43, self.delta_12, self.delta_43, self.order,
                self.Lambda, self.tjmax, self.precision, min(self.kept_pole_order, self.order))

    def lambda_orders(self):
        return [self.p0] if self.radial else list(range(self.p0, self.Lambda + 1, 2))

    def path(self, x):
        return fmt_template(self.out_template, x)


def fmt_template(tmpl, x):
    """fmt::format(template, double) for the '{}' / '{:spec}' fields blocks_3d's templates use"""
    def rep(m):
        spec = m.group(1)
        if spec:
            return format(x, spec)
        s = repr(float(x))
        return s[:-2] if s.endswith(".0") else s
    return re.sub(r"\{(?::([^}]*))?\}", rep, tmpl)


# -------------------------------------------------------------------------------------------- fallback

def fallback(argv, reason):
    exe = os.environ.get("BLOCKSVIBE_FALLBACK")
    if not exe:
        print(f"blocksvibe: {reason}, and no fallback (set BLOCKSVIBE_FALLBACK to the official blocks_3d)",
              file=sys.stderr)
        sys.exit(1)
    env = dict(os.environ)
    je = os.environ.get("BLOCKSVIBE_JEMALLOC")
    if je:
        env["LD_PRELOAD"] = (je + " " + env.get("LD_PRELOAD", "")).strip()
    log(f"fallback to {exe}: {reason}")
    sys.stdout.flush()
    os.execve(exe, [exe] + list(argv), env)


def _run_fallback(argv, reason):
    exe = os.environ.get("BLOCKSVIBE_FALLBACK")
    if not exe:
        print(f"blocksvibe: {reason}, and no fallback (set BLOCKSVIBE_FALLBACK)", file=sys.stderr)
        return 1
    env = dict(os.environ)
    if os.environ.get("BLOCKSVIBE_JEMALLOC"):
        env["LD_PRELOAD"] = (os.environ["BLOCKSVIBE_JEMALLOC"] + " " + env.get("LD_PRELOAD", "")).strip()
    log(f"fallback to {exe}: {reason}")
    return subprocess.run([exe] + list(argv), env=env).returncode


def gpu_ids():
    for var in ("HIP_VISIBLE_DEVICES", "ROCR_VISIBLE_DEVICES", "CUDA_VISIBLE_DEVICES"):
        if os.environ.get(var) is not None:
            return [x for x in os.environ[var].split(",") if x.strip() and x.strip() != "-1"]
    import torch
    return [str(i) for i in range(torch.cuda.device_count())]


class GpuLock:
    """exclusive use of one GPU through a lock file per GPU id"""

    def __init__(self, ids):
        d = os.environ.get("BLOCKSVIBE_LOCKDIR", tempfile.gettempdir())
        self.f = None
        for gid in ids:
            f = open(os.path.join(d, f"blocksvibe-gpu{gid}.lock"), "w")
            try:
                fcntl.flock(f, fcntl.LOCK_EX | fcntl.LOCK_NB)
                self.f, self.gid = f, gid
                break
            except OSError:
                f.close()
        if self.f is None:
            gid = ids[os.getpid() % len(ids)]
            self.f = open(os.path.join(d, f"blocksvibe-gpu{gid}.lock"), "w")
            t = time.time()
            fcntl.flock(self.f, fcntl.LOCK_EX)
            self.gid = gid
            log(f"waited {time.time() - t:.1f} s for GPU {gid}")

    def release(self):
        if self.f:
            fcntl.flock(self.f, fcntl.LOCK_UN)
            self.f.close()
            self.f = None


# -------------------------------------------------------------------------------------------- GPU run

def limbs(precision):
    if os.environ.get("BLOCKSVIBE_LIMBS"):
        return int(os.environ["BLOCKSVIBE_LIMBS"])
    from .mpf import limbs_for_precision
    return limbs_for_precision(precision)


def cache_dir():
    return os.environ.get("BLOCKSVIBE_CACHE", os.path.expanduser("~/.cache/blocksvibe"))


def cached_output(coord, Lam, p0, d12, L):
    """Output with its write_output map cached on disk (host construction: seconds to minutes)"""
    import torch
    from . import mpf
    from .output import Output
    d = cache_dir()
    key = hashlib.sha1(f"{coord}|{Lam}|{p0}|{d12}|{L}|amd1".encode()).hexdigest()[:20]
    path = os.path.join(d, f"wmap_{coord}_L{Lam}_p{p0}_{key}.pt")
    if os.path.exists(path):
        try:
            st = torch.load(path, weights_only=True)
            W = mpf.MPF(st["m"].to(mpf.dev()), st["e"].to(mpf.dev()))
            return Output(coord, Lam, p0, d12, L, W=W, inputs=[tuple(x) for x in st["inputs"].tolist()],
                          outputs=[tuple(x) for x in st["outputs"].tolist()])
        except Exception as e:
            log(f"cache {path}: {e}")
    o = Output(coord, Lam, p0, d12, L)
    try:
        os.makedirs(d, exist_ok=True)
        tmp = path + f".{os.getpid()}"
        torch.save({"m": o.W.m.cpu(), "e": o.W.e.cpu(), "inputs": torch.tensor(o.inputs),
                    "outputs": torch.tensor(o.outputs)}, tmp)
        os.replace(tmp, path)
    except OSError as e:
        log(f"cannot write cache {path}: {e}")
    return o


def max_rhs(plan, der, L, free=None):
    """right-hand sides per batch from the free HBM: residue slots + moments + their staging + h_inf,
    x2 for temporaries; capped by BLOCKSVIBE_MAX_RHS"""
    import torch
    if free is None:
        free, _ = torch.cuda.mem_get_info()
    S2, nb = plan.S2, 4 * L + 8
    per = S2 * nb * ((plan.nphys + 1) + 2 * der.Jo * der.NP * (der.R + 1) + 8 * 2 * der.Jo * der.NP
                     + 2 * (plan.order + 1) * plan.Jh)
    n = max(1, int(0.6 * free / (2 * per)))
    # every limb plane of the chain stays below 2^30 entries (32-bit plane offsets in the kernels): residue
    # slots, moments, numerators (Jo, Dout, R+1, 2, S2, B) and the write_output result (Jo, 2, S2, B, outputs,
    # Dmax), with Dout, outputs <= 4 (R + 1) + NP as generous bounds
    wide = max(plan.nphys + 1, 2 * der.Jo * der.NP * (der.R + 1),
               der.Jo * (4 * (der.R + 1) + der.NP) ** 2 * 2) * S2
    n = max(1, min(n, ((1 << 30) - 1) // wide))
    cap = os.environ.get("BLOCKSVIBE_MAX_RHS")
    return min(n, int(cap)) if cap else n


def run_group(tasks, timers):
    """tasks: Tasks with one group_key.  Writes all their outputs."""
    from . import setup as SU
    t0 = time.time()
    T = tasks[0]
    q4s, seen = [], set()
    for t in tasks:
        k = (tuple(t.q4), t.sign)
        if k not in seen:
            seen.add(k)
            q4s.append(k)
    radial = all(t.radial for t in tasks)
    threads = max(t.num_threads for t in tasks)
    parities = sorted({SU.four_point_parity(T.two_js, q) for q, _ in q4s})
    mkey = hashlib.sha1(repr((T.group_key(), parities, "m1")).encode()).hexdigest()[:24]
    mcache = os.path.join(cache_dir(), f"M_{mkey}.bin")
    have_m = os.path.exists(mcache)
    with tempfile.TemporaryDirectory(prefix="blocksvibe_") as tmp:
        sfile = os.path.join(tmp, "setup.bin")
        SU.run_helper(sfile, T.two_js, T.two_j12, T.two_j43, T.delta_12, T.delta_43, T.order, T.Lambda,
                      T.tjmax, radial, q4s, T.precision, threads, skip_m=have_m)
        if not have_m:
            try:
                os.makedirs(cache_dir(), exist_ok=True)
                SU.save_m(sfile, mcache)
            except OSError as e:
                log(f"cannot cache M matrices: {e}")
        t1 = time.time()
        timers.append((".main.setup", t1 - t0, 1))
        ids = gpu_ids()
        if not ids:
            raise RuntimeError("no GPU")
        lock = GpuLock(ids)
        try:
            t2 = time.time()
            if t2 - t1 > 0.5:
                timers.append((".main.gpu_wait", t2 - t1, 1))
            results = gpu_group(tasks, q4s, sfile, mcache if have_m else None, timers)
        finally:
            lock.release()
    t3 = time.time()
    write_outputs(tasks, results)
    timers.append((".main.write_output", time.time() - t3, 1))


def gpu_group(tasks, q4s, sfile, mfile, timers):
    """the GPU chain for one group, writing every task's spin files (write_task_files) -> results for
    write_outputs: {(q4, sign): (spins, pp, {(coord, d12): (m, e) host arrays (L, Jo, 2, S2, outputs, Dmax) /
    (Jo, 2, S2, outputs, Dmax)}, {(coord, d12): outputs list}, SR)} with S2 = SL SR structure entries, entry
    (a, b) at a SR + b; empty, since the files are already written"""
    import torch
    from . import mpf
    from . import setup as SU
    from .derivatives import Derivs
    from .plan import Plan
    from .recursion import Recursion
    T = tasks[0]
    L = limbs(T.precision)
    t0 = time.time()
    M, stasks = SU.read(sfile, L, m_path=mfile)
    by_q4 = {(st.q4, st.sign): st for st in stasks}
    sub = {}
    for k in q4s:
        st = by_q4[(tuple(k[0]), k[1])]
        sub.setdefault((st.pp, st.sign), []).append(st)
    torch.cuda.synchronize()
    timers.append((".main.gpu.read_setup", time.time() - t0, 1))
    results, write_s = {}, 0.0
    for (pp, sign), sts in sub.items():
        keys = {(s.q4, s.sign) for s in sts}
        ts = [t for t in tasks if t.sign == sign and (tuple(t.q4), t.sign) in keys]
        p0 = (1 - sign) // 2
        radial = all(t.radial for t in ts)
        ns = [p0] if radial else list(range(p0, T.Lambda + 1, 2))
        tjs = sorted(set().union(*[t.two_js_internal for t in ts]))
        coords = [c for c in ("xt", "xt_radial") if any(c in t.coordinates for t in ts)]
        d12s = sorted({t.delta_1_plus_2 for t in ts})
        t1 = time.time()
        plan = Plan(T.two_j12, T.two_j43, T.tjmax, T.order, pp, reuse=True)
        rec = Recursion(plan, M, L)
        der = Derivs(plan, T.Lambda - p0, L, two_js=tjs, kept_pole_order=T.kept_pole_order)
        outs = {(c, d): cached_output(c, T.Lambda, p0, d, L) for c in coords for d in d12s}
        torch.cuda.synchronize()
        t2 = time.time()
        timers.append((".main.gpu.constants", t2 - t1, 1))
        SL, SR, Jh = plan.SLm, plan.SRm, plan.Jh
        bmax = max_rhs(plan, der, L)
        chunks = [sts[i:i + bmax] for i in range(0, len(sts), bmax)]
        for ch in chunks:
            nums = []
            for n in ns:
                dn = der.for_order(T.Lambda - n)
                hs = [SU.h_inf_tensor(st, [n], T.order, Jh, SL, SR, L) for st in ch]
                h = mpf.cat(hs, 4) if len(hs) > 1 else hs[0]
                dn.start(len(ch))
                rec.run(h, hook=dn.accumulate)
                nums.append(dn.finish(h))
                del h, hs
            rs = {cd: o.apply_per_order(nums) for cd, o in outs.items()}   # (Jo, 2, S2, nq, outputs, Dmax)
            del nums
            tw = time.time()
            for q, st in enumerate(ch):
                # every spin file of the tasks of this structure, their numbers formatted on the GPU
                mine = [t for t in ts if (tuple(t.q4), t.sign) == (st.q4, st.sign)]
                write_task_files(mine, tjs, pp, plan, rs, {cd: o.outputs for cd, o in outs.items()}, q)
            write_s += time.time() - tw
            del rs
        torch.cuda.synchronize()
        t3 = time.time()
        timers.append((".main.gpu.chain", t3 - t2 - write_s, len(ns) * len(chunks)))
        log(f"pp {pp} sign {sign}: {len(sts)} structure(s) x {len(ns)} lambda order(s), batch <= {bmax}: "
            f"constants {t2 - t1:.2f} s, chain {t3 - t2:.2f} s")
    timers.append((".main.gpu.write_output", write_s, 1))
    return results


def write_task_files(mine, tjs, pp, plan, rs, outputs, q):
    """the spin files of the CLI tasks `mine` (one four-point structure and sign: column q of the write_output
    results rs[(coord, d12)] on the device, (Jo, 2, SL SR, nq, outputs, Dmax)): the polynomials each file
    holds, as JSON list bodies made on the GPU (b3djson.poly_bodies_device), then the files"""
    import torch
    from . import b3djson as BJ
    from . import structure as S
    if not mine:
        return
    ps = {BJ.digits10(t.precision) for t in mine}
    SR, S2 = plan.SRm, plan.S2
    for p in ps:
        eps = BJ.trim_exponent(p)
        need = sorted({tj for t in mine if BJ.digits10(t.precision) == p for tj in t.two_js_internal})
        bodies = {}
        for cd, r in rs.items():
            if not any(cd[0] in t.coordinates and cd[1] == t.delta_1_plus_2 for t in mine):
                continue
            Jo, _, _, nq, no, D = r.e.shape
            sel, spans = [], {}
            for tj in need:
                jo = tjs.index(tj)
                structs = [(len(S.so3(pp[k][0], plan.tj12, tj)), len(S.so3(pp[k][1], plan.tj43, tj))) for k in range(2)]
                a0 = len(sel)
                for k in range(2):
                    for a in range(structs[k][0]):
                        for b in range(structs[k][1]):
                            base = ((jo * 2 + k) * S2 + a * SR + b) * nq + q
                            sel += [base * no + o for o in range(no)]
                spans[tj] = (a0, len(sel), structs)
            got = BJ.poly_bodies_device(r.m.reshape(r.m.shape[0], -1, D), r.e.reshape(-1, D), p, eps,
                                        select=torch.tensor(sel, dtype=torch.int64, device=r.e.device))
            bodies[cd] = {tj: (got[a0:a1], structs) for tj, (a0, a1, structs) in spans.items()}
        for t in mine:
            if BJ.digits10(t.precision) != p:
                continue
            for tj in t.two_js_internal:
                coord_rows, structs = [], None
                for c in [c for c in BJ.COORD_ORDER if c in t.coordinates]:
                    cd = (c, t.delta_1_plus_2)
                    polys, structs = bodies[cd][tj]
                    no = len(outputs[cd])
                    rows_k, i = [], 0
                    for k in range(2):
                        rows_a = []
                        for a in range(structs[k][0]):
                            rows_b = []
                            for b in range(structs[k][1]):
                                rows = {}
                                for o, (n, _) in enumerate(outputs[cd]):
                                    rows.setdefault(n, []).append(polys[i])
                                    i += 1
                                rows_b.append([rows[n] for n in sorted(rows)])
                            rows_a.append(rows_b)
                        rows_k.append(rows_a)
                    coord_rows.append(rows_k)
                path = t.path(tj / 2.0)
                os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
                BJ.write_spin(path, t, tj, pp, structs, coord_rows)


def write_outputs(tasks, results):
    """the spin files of the tasks that have host results (the GPU path writes its files itself, from bodies
    made on the GPU): numbers formatted by the library's host formatter on up to num_threads threads, one
    spin file after another"""
    jobs = [(i, tj) for i, t in enumerate(tasks) if (tuple(t.q4), t.sign) in results for tj in t.two_js_internal]
    if not jobs:
        return
    threads = max(1, min(max(t.num_threads for t in tasks), os.cpu_count() or 1))
    for j in jobs:
        _write_one(tasks, results, j, threads)


def _write_one(tasks, results, job, threads):
    import numpy as np

    from . import b3djson as BJ
    from . import structure as S
    i, tj = job
    t = tasks[i]
    tjs, pp, vals, outs, SR = results[(tuple(t.q4), t.sign)]
    jo = tjs.index(tj)
    p = BJ.digits10(t.precision)
    structs = [(len(S.so3(pp[k][0], t.two_j12, tj)), len(S.so3(pp[k][1], t.two_j43, tj))) for k in range(2)]
    coord_rows = []
    for c in [c for c in BJ.COORD_ORDER if c in t.coordinates]:
        m, e = vals[(c, t.delta_1_plus_2)]                          # (L, Jo, 2, S2, outputs, Dmax)
        outputs = outs[(c, t.delta_1_plus_2)]
        mj, ej = m[:, jo], e[jo]
        _, S2, no, _ = ej.shape                                     # (2, S2, outputs, Dmax)
        sel, where = [], {}
        for k in range(2):
            for a in range(structs[k][0]):
                for b in range(structs[k][1]):
                    for o in range(no):
                        where[(k, a, b, o)] = len(sel)
                        sel.append((k * S2 + a * SR + b) * no + o)
        bodies = BJ.poly_bodies(mj, ej, p, BJ.trim_exponent(p), select=np.array(sel, dtype=np.int64),
                                threads=threads)
        rows_k = []
        for k in range(2):
            rows_a = []
            for a in range(structs[k][0]):
                rows_b = []
                for b in range(structs[k][1]):
                    rows = {}
                    for o, (n, _) in enumerate(outputs):
                        rows.setdefault(n, []).append(bodies[where[(k, a, b, o)]])
                    rows_b.append([rows[n] for n in sorted(rows)])
                rows_a.append(rows_b)
            rows_k.append(rows_a)
        coord_rows.append(rows_k)
    path = t.path(tj / 2.0)
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    BJ.write_spin(path, t, tj, pp, structs, coord_rows)


# -------------------------------------------------------------------------------------------- entry points

def run_tasks(tasks, argvs):
    """GPU for the supported tasks (grouped), the official blocks_3d for the others -> exit status per task"""
    groups, status = {}, [0] * len(tasks)
    for i, t in enumerate(tasks):
        why = t.unsupported()
        if why:
            status[i] = _run_fallback(argvs[i], why)
        else:
            groups.setdefault(t.group_key(), []).append(i)
    for key, idx in groups.items():
        ts = [tasks[i] for i in idx]
        timers = []
        t0 = time.time()
        try:
            run_group(ts, timers)
        except Exception as e:
            import traceback
            log(traceback.format_exc())
            for i in idx:
                status[i] = _run_fallback(argvs[i], f"GPU run failed: {type(e).__name__}: {e}")
            continue
        total = time.time() - t0
        for t in ts:
            if t.profile:
                write_profile(t, total, timers)
    return status


def write_profile(t, total, timers):
    from . import b3djson as BJ
    path = fmt_template(t.out_template, 0.0) + ".profile"
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    BJ.write_profile(path, t, t.num_threads, [(".main", total, 1)] + timers)


def _alive(pid):
    try:
        os.kill(pid, 0)
        return True
    except ProcessLookupError:
        return False
    except PermissionError:
        return True


def _atomic_write(path, text):
    tmp = f"{path}.tmp{os.getpid()}"
    with open(tmp, "w") as f:
        f.write(text)
    os.replace(tmp, path)


def spool_run(task, argv):
    """join concurrent invocations of one group into one GPU batch: every invocation drops its arguments in
    a spool directory per group; the holder of the group's leader lock waits until no new task arrived
    for BLOCKSVIBE_SPOOL_QUIET s (at most BLOCKSVIBE_SPOOL_MAX s), claims the pending tasks, runs them as one
    batch and writes each task's exit status; a task claimed by a dead leader is put back"""
    import json
    import uuid
    root = os.environ.get("BLOCKSVIBE_SPOOL", os.path.join(tempfile.gettempdir(), "blocksvibe-spool"))
    d = os.path.join(root, hashlib.sha1(repr(task.group_key()).encode()).hexdigest()[:20])
    os.makedirs(d, exist_ok=True)
    quiet = float(os.environ.get("BLOCKSVIBE_SPOOL_QUIET", "3"))
    tmax = float(os.environ.get("BLOCKSVIBE_SPOOL_MAX", "30"))
    me = f"{os.getpid()}-{uuid.uuid4().hex[:8]}"
    _atomic_write(os.path.join(d, me + ".task"), json.dumps({"argv0": task.argv[0], "argv": argv}))
    done = os.path.join(d, me + ".done")
    lockf = open(os.path.join(d, "leader.lock"), "w")
    while True:
        if os.path.exists(done):
            st = int(open(done).read().strip() or 1)
            os.remove(done)
            return st
        try:
            fcntl.flock(lockf, fcntl.LOCK_EX | fcntl.LOCK_NB)
        except OSError:
            time.sleep(0.5)
            if not os.path.exists(os.path.join(d, me + ".task")):
                for f in [f for f in os.listdir(d) if f.startswith(me + ".claimed.")]:
                    pid = int(f.rsplit(".", 1)[1])
                    if not _alive(pid) and not os.path.exists(done):
                        try:
                            os.rename(os.path.join(d, f), os.path.join(d, me + ".task"))
                        except OSError:
                            pass
            continue
        try:
            t0, last, n = time.time(), time.time(), -1
            while True:
                pending = sorted(f for f in os.listdir(d) if f.endswith(".task"))
                if len(pending) != n:
                    n, last = len(pending), time.time()
                if time.time() - last >= quiet or time.time() - t0 >= tmax:
                    break
                time.sleep(0.25)
            mine = []
            for f in pending:
                tid = f[:-5]
                try:
                    os.rename(os.path.join(d, f), os.path.join(d, f"{tid}.claimed.{os.getpid()}"))
                    mine.append(tid)
                except OSError:
                    pass
            if mine:
                specs = [json.load(open(os.path.join(d, f"{tid}.claimed.{os.getpid()}"))) for tid in mine]
                argvs = [sp["argv"] for sp in specs]
                log(f"spool leader: batch of {len(mine)} task(s) after {time.time() - t0:.1f} s")
                try:
                    tasks = [Task(sp["argv"], sp["argv0"]) for sp in specs]
                    status = run_tasks(tasks, argvs)
                except Exception as e:
                    log(f"spool batch failed: {e}")
                    status = [_run_fallback(a, f"batch failed: {e}") for a in argvs]
                for tid, st in zip(mine, status):
                    _atomic_write(os.path.join(d, tid + ".done"), str(st))
                    os.remove(os.path.join(d, f"{tid}.claimed.{os.getpid()}"))
        finally:
            fcntl.flock(lockf, fcntl.LOCK_UN)


def main(argv=None):
    argv = list(sys.argv[1:] if argv is None else argv)
    prog = os.path.basename(sys.argv[0]) if sys.argv and sys.argv[0] else "blocks_3d"
    if argv[:1] == ["batch"]:
        rest, argvs = argv[1:], []
        while rest:
            if rest[0] == "--lines":
                for line in open(rest[1]):
                    if line.strip() and not line.lstrip().startswith("#"):
                        argvs.append(shlex.split(line))
                rest = rest[2:]
            else:
                argvs.append(["-p", rest[0]])
                rest = rest[1:]
        try:
            tasks = [Task(a, prog) for a in argvs]
        except UsageError as e:
            print(f"Error: {e}", file=sys.stderr)
            return 1
        return max(run_tasks(tasks, argvs), default=0)
    try:
        task = Task(argv, sys.argv[0] if sys.argv and sys.argv[0] else prog)
    except UsageError as e:
        if os.environ.get("BLOCKSVIBE_FALLBACK"):
            fallback(argv, f"options not understood ({e})")
        print(f"Error: {e}", file=sys.stderr)
        return 1
    if task.help:
        if os.environ.get("BLOCKSVIBE_FALLBACK"):
            fallback(argv, "--help")
        print(__doc__)
        return 0
    why = task.unsupported()
    if why:
        fallback(argv, why)
    if os.environ.get("BLOCKSVIBE_SPOOL_OFF") != "1":
        return spool_run(task, argv)
    return run_tasks([task], [argv])[0]


if __name__ == "__main__":
    sys.exit(main())
