"""`blocks_3d` on MI355X: a drop-in for blocks_3d's command line (reference blocksvibe/cli.py).

    blocks_3d [blocks_3d options]            one task: same options, -p param files, per-spin JSON files and .profile
    blocks_3d batch [--lines FILE] [PARAM_FILE ...]
                                             many tasks: one blocks_3d command line per line of FILE (blank and '#'
                                             lines skipped), or one param file per argument

Coordinates xt and xt_radial (with or without pole shifting) run on the GPU; anything else, and any task the GPU
path fails, runs the official blocks_3d ($BLOCKSVIBE_FALLBACK, with $BLOCKSVIBE_JEMALLOC preloaded if set).

How an invocation runs: its tasks are grouped by operator (external spins, j12, j43, Delta_12, Delta_43, order,
Lambda, maximal spin, precision, kept pole order); for each group this process runs the b3d_setup helper on the CPU
(M matrices cached on disk) and then hands the group to the GPU's serving daemon (daemon.py) over a Unix socket,
starting the daemon if none is running.  The daemon owns the GPU, batches the tasks of concurrent invocations that
share an operator, writes every spin file and answers with a status per task.  BLOCKSVIBE_DAEMON=0 runs the same
GPU code inside this process instead.

Environment: BLOCKSVIBE_B3D_SETUP (helper), BLOCKSVIBE_LIMBS, BLOCKSVIBE_CACHE (M matrices and write_output maps,
default ~/.cache/blocksvibe), BLOCKSVIBE_SOCKET (daemon socket), BLOCKSVIBE_DAEMON (0: in process),
BLOCKSVIBE_DAEMON_IDLE (seconds before an idle daemon exits), BLOCKSVIBE_WINDOW (seconds a task may wait for
companions), BLOCKSVIBE_FALLBACK, BLOCKSVIBE_JEMALLOC, BLOCKSVIBE_VERBOSE=1.
"""
import fcntl
import hashlib
import os
import shlex
import socket
import subprocess
import sys
import tempfile
import time

from .taskspec import TaskSpec, UsageError, fill_template, spin_set  # noqa: F401  (re-exported interface)

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def log(*parts):
    if os.environ.get("BLOCKSVIBE_VERBOSE") == "1":
        print("[blocksvibe]", *parts, file=sys.stderr, flush=True)


# -------------------------------------------------------------------------- the official binary

def _fallback_env():
    env = dict(os.environ)
    pre = os.environ.get("BLOCKSVIBE_JEMALLOC")
    if pre:
        env["LD_PRELOAD"] = f"{pre} {env.get('LD_PRELOAD', '')}".strip()
    return env


def fallback_exec(argv, reason):
    """replace this process by the official blocks_3d (exit 1 with a message if there is none)"""
    exe = os.environ.get("BLOCKSVIBE_FALLBACK")
    if not exe:
        print(f"blocksvibe: {reason}, and no fallback (set BLOCKSVIBE_FALLBACK to the official blocks_3d)",
              file=sys.stderr)
        sys.exit(1)
    log(f"running {exe}: {reason}")
    sys.stdout.flush()
    os.execve(exe, [exe] + list(argv), _fallback_env())


def fallback_run(argv, reason):
    """the official blocks_3d on one task as a child process -> its exit status (1 if there is none)"""
    exe = os.environ.get("BLOCKSVIBE_FALLBACK")
    if not exe:
        print(f"blocksvibe: {reason}, and no fallback (set BLOCKSVIBE_FALLBACK)", file=sys.stderr)
        return 1
    log(f"running {exe}: {reason}")
    return subprocess.run([exe] + list(argv), env=_fallback_env()).returncode


# -------------------------------------------------------------------------- the CPU setup step

def cache_dir():
    return os.environ.get("BLOCKSVIBE_CACHE", os.path.expanduser("~/.cache/blocksvibe"))


def setup_group(tasks, workdir):
    """run b3d_setup for one group's tasks -> dict(setup=file, m_file=cached M file or None, seconds)"""
    from . import setup as SU
    t0 = time.time()
    head = tasks[0]
    structures = list(dict.fromkeys((tuple(t.q4), t.sign) for t in tasks))
    parities = sorted({SU.four_point_parity(head.two_js, q) for q, _ in structures})
    tag = hashlib.sha1(repr((head.group_key(), parities, "m1")).encode()).hexdigest()[:24]
    m_file = os.path.join(cache_dir(), f"M_{tag}.bin")
    cached = os.path.exists(m_file)
    out = os.path.join(workdir, f"setup_{tag}.bin")
    SU.run_helper(out, head.two_js, head.two_j12, head.two_j43, head.delta_12, head.delta_43, head.order,
                  head.Lambda, head.tjmax, all(t.radial for t in tasks), structures, head.precision,
                  max(t.num_threads for t in tasks), skip_m=cached)
    if not cached:                             # this setup file holds the M matrices: keep them for next time
        try:
            os.makedirs(cache_dir(), exist_ok=True)
            SU.save_m(out, m_file)
        except OSError as err:
            log(f"M matrices not cached: {err}")
    return {"setup": out, "m_file": m_file if cached else None, "seconds": time.time() - t0}


# -------------------------------------------------------------------------- the GPU service

def socket_path():
    if os.environ.get("BLOCKSVIBE_SOCKET"):
        return os.environ["BLOCKSVIBE_SOCKET"]
    gid = "0"
    for var in ("HIP_VISIBLE_DEVICES", "ROCR_VISIBLE_DEVICES", "CUDA_VISIBLE_DEVICES"):
        if os.environ.get(var):
            gid = os.environ[var].split(",")[0].strip() or "0"
            break
    return os.path.join(tempfile.gettempdir(), f"blocksvibe-{os.getuid()}", f"gpu{gid}.sock")


def _connect(path):
    s = socket.socket(socket.AF_UNIX, socket.SOCK_STREAM)
    try:
        s.connect(path)
    except OSError:
        s.close()
        return None
    return s


def _ask(path, msg, timeout=None):
    from .daemon import recv, send
    s = _connect(path)
    if s is None:
        return None
    with s:
        s.settimeout(timeout)
        send(s, msg)
        return recv(s)


def ensure_daemon(path):
    """the daemon's socket path once a daemon answers there (started if needed), or None"""
    if _ask(path, {"op": "ping"}, 30) is not None:
        return path
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    with open(path + ".lock", "w") as lockf:
        fcntl.flock(lockf, fcntl.LOCK_EX)          # one starter at a time
        if _ask(path, {"op": "ping"}, 30) is not None:
            return path
        if os.path.exists(path):
            os.unlink(path)                        # a socket nobody serves
        env = dict(os.environ, PYTHONPATH=PKG_ROOT + os.pathsep + os.environ.get("PYTHONPATH", ""))
        with open(path + ".log", "ab") as out:
            proc = subprocess.Popen([sys.executable, "-m", "blocksvibe_amd.daemon", "--socket", path], env=env,
                                    stdin=subprocess.DEVNULL, stdout=out, stderr=out, cwd="/",
                                    start_new_session=True)
        limit = time.time() + float(os.environ.get("BLOCKSVIBE_DAEMON_START", "300"))
        while time.time() < limit:
            if proc.poll() is not None:
                log(f"daemon exited with {proc.returncode} (see {path}.log)")
                return None
            if os.path.exists(path) and _ask(path, {"op": "ping"}, 30) is not None:
                log(f"daemon {proc.pid} serving {path}")
                return path
            time.sleep(0.2)
    return None


def run_on_gpu(groups, cwd):
    """groups: [{"tasks": [TaskSpec], "setup", "m_file", "limbs"}] -> per group [(error or None, seconds)]"""
    payload = [{"argvs": [t.argv for t in g["tasks"]], "setup": g["setup"], "m_file": g["m_file"],
                "limbs": g["limbs"]} for g in groups]
    if os.environ.get("BLOCKSVIBE_DAEMON", "1") != "0":
        path = ensure_daemon(socket_path())
        if path is not None:
            reply = _ask(path, {"op": "run", "cwd": cwd, "groups": payload})
            if reply is not None and reply.get("ok"):
                return reply["results"]
            why = "no reply" if reply is None else reply.get("error")
            return [[(f"daemon: {why}", 0.0)] * len(g["tasks"]) for g in groups]
        log("no daemon; running on the GPU in this process")
    from .runner import Runner
    runner = Runner(streams=2, map_cache=cache_dir())
    try:
        return [runner.run_group(g["tasks"], g["setup"], g["limbs"], cwd, g["m_file"]) for g in groups]
    finally:
        runner.close()


# -------------------------------------------------------------------------- invocations

def run_tasks(tasks, argvs):
    """every task of an invocation -> exit status per task"""
    from .runner import limbs, write_profile
    status = [0] * len(tasks)
    groups = {}
    for i, t in enumerate(tasks):
        why = t.unsupported()
        if why:
            status[i] = fallback_run(argvs[i], why)
        else:
            groups.setdefault(t.group_key(), []).append(i)
    if not groups:
        return status
    cwd = os.getcwd()
    t0 = time.time()
    with tempfile.TemporaryDirectory(prefix="blocksvibe_") as work:
        ready, failed = [], {}
        for idx in groups.values():
            ts = [tasks[i] for i in idx]
            try:
                g = setup_group(ts, work)
            except Exception as err:                                  # noqa: BLE001 (the fallback runs them)
                for i in idx:
                    failed[i] = f"setup failed: {type(err).__name__}: {err}"
                continue
            g.update(tasks=ts, index=idx, limbs=limbs(ts[0].precision, os.environ.get("BLOCKSVIBE_LIMBS")))
            ready.append(g)
        results = run_on_gpu(ready, cwd) if ready else []
    for g, res in zip(ready, results):
        for i, t, (err, gpu_s) in zip(g["index"], g["tasks"], res):
            if err:
                failed[i] = f"GPU run failed: {err}"
            elif t.profile:
                write_profile(t, cwd, [(".main", time.time() - t0, 1), (".main.setup", g["seconds"], 1),
                                       (".main.gpu", gpu_s, 1)])
    for i, why in sorted(failed.items()):
        log(f"task {i}: {why}")
        status[i] = fallback_run(argvs[i], why)
    return status


def batch_argvs(args):
    """the task command lines of `batch [--lines FILE] [PARAM_FILE ...]`"""
    out, rest = [], list(args)
    while rest:
        word = rest.pop(0)
        if word == "--lines":
            if not rest:
                raise UsageError("batch: --lines needs a file")
            with open(rest.pop(0)) as f:
                out += [shlex.split(line) for line in f if line.strip() and not line.lstrip().startswith("#")]
        else:
            out.append(["-p", word])
    return out


def main(argv=None):
    argv = list(sys.argv[1:] if argv is None else argv)
    prog = os.path.basename(sys.argv[0]) if sys.argv and sys.argv[0] else "blocks_3d"
    if argv[:1] == ["batch"]:
        try:
            argvs = batch_argvs(argv[1:])
            tasks = [TaskSpec(a, prog) for a in argvs]
        except (UsageError, OSError) as err:
            print(f"Error: {err}", file=sys.stderr)
            return 1
        return max(run_tasks(tasks, argvs), default=0)
    try:
        task = TaskSpec(argv, sys.argv[0] if sys.argv and sys.argv[0] else prog)
    except UsageError as err:
        if os.environ.get("BLOCKSVIBE_FALLBACK"):
            fallback_exec(argv, f"options not understood ({err})")
        print(f"Error: {err}", file=sys.stderr)
        return 1
    if task.help:
        if os.environ.get("BLOCKSVIBE_FALLBACK"):
            fallback_exec(argv, "--help")
        print(__doc__)
        return 0
    why = task.unsupported()
    if why:
        fallback_exec(argv, why)
    return run_tasks([task], [argv])[0]


def stop_daemon(path=None):
    """ask the daemon at path (default: this GPU's) to exit; True if one answered"""
    reply = _ask(path or socket_path(), {"op": "stop"}, 30)
    return bool(reply and reply.get("ok"))


def daemon_stats(path=None):
    return _ask(path or socket_path(), {"op": "stats"}, 30)


if __name__ == "__main__":
    sys.exit(main())
