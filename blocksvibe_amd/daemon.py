"""The serving daemon: the one process that owns a GPU for blocks_3d invocations.

    python -m blocksvibe_amd.daemon --socket PATH [--idle SECONDS] [--window SECONDS]

It listens on a Unix socket; each `blocks_3d` invocation (cli.py) is a client that runs its own b3d_setup step
on the CPU, then sends its task groups here and waits for their files.  All clients share one runner.Runner
(serving loop + emitter): operators, plans, derivative constants and write_output maps stay cached on the GPU
between invocations, and requests of one operator from concurrent invocations run as one device batch (they may
wait `window` seconds for companions).  The daemon exits after `idle` seconds without a client.  This replaces the
reference's per-GPU lock files and its spool directory of waiting invocations (reference
blocksvibe/cli.py:238-265, 569-637).

Protocol: one JSON document per line.  Client -> daemon:
  {"op": "run", "cwd": DIR, "groups": [{"argvs": [[prog, arg...], ...], "setup": FILE, "m_file": FILE|null,
                                        "limbs": L}, ...]}
  {"op": "stats"} | {"op": "stop"} | {"op": "ping"}
Daemon -> client: {"ok": true, "results": [[[error|null, seconds], ...] per group]} (run), the runner's stats,
or {"ok": false, "error": TEXT}.
"""
import argparse
import json
import os
import socket
import sys
import threading
import time


def send(conn, obj):
    conn.sendall((json.dumps(obj) + "\n").encode())


def recv(conn):
    """one JSON line from conn (None at end of stream)"""
    buf = bytearray()
    while True:
        chunk = conn.recv(1 << 16)
        if not chunk:
            return json.loads(buf) if buf.strip() else None
        buf += chunk
        if buf.endswith(b"\n"):
            return json.loads(buf)


class Daemon:
    def __init__(self, path, idle=120.0, window=0.2):
        self.path, self.idle = path, idle
        self.runner = None
        self.window = window
        self.active = 0
        self.last = time.time()
        self.lock = threading.Lock()
        self.stopping = False
        self.sock = socket.socket(socket.AF_UNIX, socket.SOCK_STREAM)
        self.sock.bind(path)
        self.sock.listen(64)
        self.sock.settimeout(1.0)

    def _runner(self):
        with self.lock:
            if self.runner is None:
                from .runner import Runner
                self.runner = Runner(streams=2, window=self.window, map_cache=os.environ.get("BLOCKSVIBE_CACHE"))
            return self.runner

    def serve(self):
        try:
            while not self.stopping:
                try:
                    conn, _ = self.sock.accept()
                except socket.timeout:
                    with self.lock:
                        if self.active == 0 and time.time() - self.last > self.idle:
                            break
                    continue
                with self.lock:
                    self.active += 1
                threading.Thread(target=self._client, args=(conn,), daemon=True).start()
        finally:
            self.sock.close()
            try:
                os.unlink(self.path)
            except OSError:
                pass
            if self.runner is not None:
                self.runner.close()

    def _client(self, conn):
        try:
            msg = recv(conn)
            if msg is None:
                return
            op = msg.get("op")
            if op == "ping":
                send(conn, {"ok": True, "pid": os.getpid()})
            elif op == "stats":
                send(conn, dict(self._runner().stats(), ok=True))
            elif op == "stop":
                self.stopping = True
                send(conn, {"ok": True})
            elif op == "run":
                send(conn, {"ok": True, "results": self.run(msg)})
            else:
                send(conn, {"ok": False, "error": f"unknown op {op!r}"})
        except Exception as err:                                      # noqa: BLE001 (reported to the client)
            try:
                send(conn, {"ok": False, "error": f"{type(err).__name__}: {err}"})
            except OSError:
                pass
        finally:
            conn.close()
            with self.lock:
                self.active -= 1
                self.last = time.time()

    def run(self, msg):
        from .taskspec import TaskSpec
        runner = self._runner()
        results = [None] * len(msg["groups"])

        def one(i, g):
            try:
                tasks = [TaskSpec(a[1:], a[0]) for a in g["argvs"]]
                results[i] = runner.run_group(tasks, g["setup"], g["limbs"], msg["cwd"], g.get("m_file"))
            except Exception as err:                                  # noqa: BLE001
                results[i] = [(f"{type(err).__name__}: {err}", 0.0)] * len(g["argvs"])

        threads = [threading.Thread(target=one, args=(i, g)) for i, g in enumerate(msg["groups"])]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        return results


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m blocksvibe_amd.daemon")
    ap.add_argument("--socket", required=True)
    ap.add_argument("--idle", type=float, default=float(os.environ.get("BLOCKSVIBE_DAEMON_IDLE", "120")))
    ap.add_argument("--window", type=float, default=float(os.environ.get("BLOCKSVIBE_WINDOW", "0.2")))
    a = ap.parse_args(argv)
    Daemon(a.socket, a.idle, a.window).serve()
    return 0


if __name__ == "__main__":
    sys.exit(main())
