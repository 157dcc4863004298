"""Spin-file emission: the JSON files of many blocks (one (task, internal spin) table each) made on the GPU and
written by host threads while the device goes on with the next work (reference blocksvibe/cli.py:453-498 and
b3djson.py:97-207 write one file at a time, formatting every coefficient through MPFR in a fork pool).

A spin file is its header followed by the derivative tables: the write_output result's polynomials in a fixed
nesting ([parity][a][b][coordinate][deriv_0][deriv_1]) with fixed text between them.  That text depends only
on the file's shape (structure counts, coordinates, output rows), so it is worked out once per shape from the
host writer itself (`skeleton`: b3djson.derivs_section with numbered placeholders) and kept as a table of
strings.  Files are then made in units of at most `unit_slots` coefficient slots:

  device, on the caller's stream: kept lengths (b3d_fmt_sizes); the digits of every coefficient slot
    (b3d_fmt_text); piece lengths and their prefix sum; the files' bytes in one buffer (b3d_fmt_pack_files:
    each polynomial preceded by its text from the string table, the header before a file's first one, the
    closing text after its last), so every file is one contiguous byte range;
  copy stream: the bytes and the file offsets into a page-locked host buffer;
  host: a dispatcher thread waits for the copy and hands each file's range to writer threads (one write call
    per file).

The caller's thread never waits for the device: units are enqueued and the dispatcher synchronises on their
copies.  Values the GPU formatter leaves to the host (exponent >= 0, or fractions beyond its window) make the
dispatcher redo that unit with the library's host formatter (`assemble_host`, byte-identical).
"""
import concurrent.futures as cf
import contextlib
import os
import queue
import threading
import time

import numpy as np

from . import _lib
from . import b3djson as BJ

UNIT_SLOTS = 1 << 22

_SKELETONS = {}
_SKEL_LOCK = threading.Lock()


class Skeleton:
    """the fixed text of one spin-file shape: texts[i] precedes polynomial i of the derivs section (texts[0]
    opens it), tail closes the file.  As a string table: strings (unique texts), ids[i] the id of texts[i],
    tail_id."""
    __slots__ = ("texts", "tail", "strings", "ids", "tail_id", "npoly")

    def __init__(self, texts, tail):
        self.texts, self.tail = texts, tail
        self.npoly = len(texts)
        index = {}
        self.ids = np.array([index.setdefault(t, len(index)) for t in texts], dtype=np.int32)
        self.tail_id = index.setdefault(tail, len(index))
        self.strings = [None] * len(index)
        for t, i in index.items():
            self.strings[i] = t.encode("ascii")


def skeleton(p, structs, rows_per_coord):
    """Skeleton of a spin file with structs[k] = (n_left, n_right) structures of parity k and, per coordinate
    (in the file's order), rows_per_coord[c] = the number of deriv_1 entries of each deriv_0 row"""
    key = (p, tuple(tuple(x) for x in structs), tuple(tuple(r) for r in rows_per_coord))
    with _SKEL_LOCK:
        sk = _SKELETONS.get(key)
    if sk is not None:
        return sk
    coord_rows = [[[[None] * nr for _ in range(nl)] for nl, nr in structs] for _ in rows_per_coord]
    i = 0
    for k, (nl, nr) in enumerate(structs):
        for a in range(nl):
            for b in range(nr):
                for c, rows in enumerate(rows_per_coord):
                    rr = []
                    for n1 in rows:
                        rr.append([BJ.Body(i + t) for t in range(n1)])
                        i += n1
                    coord_rows[c][k][a][b] = rr
    texts, cur = [], []
    for x in BJ.derivs_section(p, structs, coord_rows):
        if isinstance(x, str):
            cur.append(x)
        else:
            assert x == len(texts), "placeholders out of order"
            texts.append("".join(cur))
            cur = []
    assert len(texts) == i
    sk = Skeleton(texts, "".join(cur))
    with _SKEL_LOCK:
        _SKELETONS[key] = sk
    return sk


def output_rows(outputs):
    """(perm, rows): the write_output outputs (deriv_0, deriv_1) in file order -- grouped by deriv_0 ascending,
    in their order within a group -- and the size of each group"""
    groups = {}
    for o, (n, _) in enumerate(outputs):
        groups.setdefault(n, []).append(o)
    perm = [o for n in sorted(groups) for o in groups[n]]
    return np.array(perm, dtype=np.int64), [len(groups[n]) for n in sorted(groups)]


class FileJob:
    """one spin file: its path, header text (b3djson.header), skeleton, and the plane polynomial index of each
    of its polynomials in file order"""
    __slots__ = ("path", "head", "skel", "polys")

    def __init__(self, path, head, skel, polys):
        assert len(polys) == skel.npoly, (len(polys), skel.npoly)
        self.path, self.head, self.skel, self.polys = path, head, skel, np.asarray(polys, dtype=np.int64)


def raw_polys(structs, SR, jo, Jo, S2, nq, q, perms, bases):
    """plane polynomial indices of one spin file in write_output results laid out (outputs, Jo, 2, S2, nq, D)
    (output.Output.apply_raw), one result per coordinate at plane polynomial offset bases[c], outputs taken in
    the order perms[c]"""
    out = []
    for k, (nl, nr) in enumerate(structs):
        for a in range(nl):
            for b in range(nr):
                s = a * SR + b
                for perm, base in zip(perms, bases):
                    out.append(base + (((perm * Jo + jo) * 2 + k) * S2 + s) * nq + q)
    return np.concatenate(out) if out else np.zeros(0, dtype=np.int64)


def _file_bytes(job, bodies):
    """a file's bytes from its polynomials' list bodies (host assembly)"""
    sk = job.skel
    parts = [job.head.encode("ascii")]
    for t, body in zip(sk.texts, bodies):
        parts.append(t.encode("ascii"))
        parts.append(bytes(body))
    parts.append(sk.tail.encode("ascii"))
    return b"".join(parts)


def assemble_host(m, e, D, p, eps, jobs, threads=0):
    """every file of jobs as bytes, from host planes m (L, npoly * D) / e (npoly * D) through the library's host
    formatter (b3djson.poly_bodies): the reference the GPU path is checked against, and its fallback"""
    L = m.shape[0]
    allp = np.concatenate([j.polys for j in jobs]) if jobs else np.zeros(0, dtype=np.int64)
    bodies = BJ.poly_bodies(m.reshape(L, -1, D), e.reshape(-1, D), p, eps, select=allp, threads=threads)
    out, i = [], 0
    for j in jobs:
        out.append(_file_bytes(j, bodies[i:i + len(j.polys)]))
        i += len(j.polys)
    return out


def _write_file(path, data, unlink):
    fd = os.open(path, os.O_WRONLY | os.O_CREAT | os.O_TRUNC, 0o644)
    try:
        view = memoryview(data)
        while len(view):
            view = view[os.write(fd, view):]
    finally:
        os.close(fd)
    if unlink:
        os.unlink(path)
    return len(data)


class _Unit:
    __slots__ = ("jobs", "ready", "packed", "dmeta", "buf", "meta", "m2", "e1", "D", "p", "eps", "ub", "empty", "call")


class _Call:
    """the units of one emit() call and its completion callback"""
    __slots__ = ("left", "on_done", "error", "lock")

    def __init__(self, on_done):
        self.left, self.on_done, self.error, self.lock = 0, on_done, None, threading.Lock()

    def unit_done(self, error=None):
        with self.lock:
            self.left -= 1
            if error is not None and self.error is None:
                self.error = error
            last = self.left == 0
        if last and self.on_done is not None:
            self.on_done(self.error)


class Emitter:
    """Writes spin files from device results, overlapped with the caller's device work.

    emit(m, e, D, p, eps, jobs, on_done=None): m (L, npoly * D) int32 / e (npoly * D) int64 device planes
    (contiguous; polynomial q = entries q D .. q D + D - 1), jobs: FileJob list.  Returns once the work is
    enqueued on the current stream (which may then reuse m, e); on_done(error) is called (from a writer thread)
    when the call's files are written; drain() returns when every file so far is written (closed) and raises
    the first error.  Thread-safe: several threads may emit, each on its own stream.  unlink: remove each file
    after writing it (throughput runs that must not fill the disk)."""

    def __init__(self, writers=16, buffers=8, unit_slots=UNIT_SLOTS, unlink=False, pending=16, gate=None):
        """gate: a context manager factory held around the dispatcher's device work (serve.Server's gate.device,
        so that no copy or synchronisation of this thread runs beside one of the server's stream captures)"""
        import torch
        self.gate = gate or contextlib.nullcontext
        self.torch = torch
        self.unit_slots = unit_slots
        self.unlink = unlink
        self.copy_stream = torch.cuda.Stream()
        self.pool = cf.ThreadPoolExecutor(max(1, writers))
        self.nbuf = max(1, buffers)
        self.free = queue.Queue()
        self.made = 0
        self.buf_lock = threading.Lock()
        # units whose bytes are still on the device (made, not yet copied): emit() waits beyond this many
        self.pending = threading.BoundedSemaphore(max(1, pending))
        self.q = queue.Queue()
        self.errors = []
        self.lock = threading.Condition()
        self.inflight = 0
        self.files = 0
        self.bytes = 0
        self.fallback_units = 0
        self.wait_s = 0.0                 # time emit() callers waited for the copies and writes to catch up
        self.write_s = 0.0                # summed time of the file writes over the writer threads
        self.buffer_wait_s = 0.0          # time the copies waited for a free host buffer
        self.dirs = set()
        self.thread = threading.Thread(target=self._dispatch, daemon=True)
        self.thread.start()

    # ------------------------------------------------------------------------------------------ API
    def emit(self, m, e, D, p, eps, jobs, on_done=None):
        for j in jobs:
            d = os.path.dirname(j.path)
            if d and d not in self.dirs:
                os.makedirs(d, exist_ok=True)
                self.dirs.add(d)
        units = [[j for j in jobs if not len(j.polys)]]
        slots = 0
        units.append([])
        for j in jobs:
            if not len(j.polys):
                continue
            n = len(j.polys) * D
            if units[-1] and slots + n > self.unit_slots:
                units.append([])
                slots = 0
            units[-1].append(j)
            slots += n
        units = [u for u in units if u]
        call = _Call(on_done)
        call.left = len(units)
        if not units:
            if on_done is not None:
                on_done(None)
            return
        for u in units:
            self._enqueue(m, e, D, p, eps, u, call)

    def drain(self):
        with self.lock:
            while self.inflight:
                self.lock.wait()
            if self.errors:
                err = self.errors[0]
                self.errors.clear()
                raise err

    def close(self):
        try:
            self.drain()
        finally:
            self.q.put(None)
            self.thread.join()
            self.pool.shutdown(wait=True)

    # ------------------------------------------------------------------------------------------ internals
    def _buffer(self, nbytes):
        """a page-locked host buffer of at least nbytes (waits for a free one once `buffers` exist)"""
        torch = self.torch
        with self.buf_lock:
            make = self.made < self.nbuf and self.free.empty()
            if make:
                self.made += 1
        if make:
            with self.gate():
                return torch.empty(max(nbytes, 1 << 20), dtype=torch.uint8, pin_memory=True)
        t0 = time.time()
        buf = self.free.get()
        with self.lock:
            self.buffer_wait_s += time.time() - t0
        if buf.numel() < nbytes:
            del buf
            with self.gate():
                buf = torch.empty(nbytes, dtype=torch.uint8, pin_memory=True)
        return buf

    def _enqueue(self, m2, e1, D, p, eps, jobs, call):
        torch = self.torch
        from . import _lib
        u = _Unit()
        u.jobs, u.m2, u.e1, u.D, u.p, u.eps, u.call = jobs, m2, e1, D, p, eps, call
        u.buf = None
        with self.lock:
            self.inflight += 1
        npoly = sum(len(j.polys) for j in jobs)
        u.empty = npoly == 0
        if u.empty:                                   # files without polynomials: header and closing text
            self.q.put(u)
            return
        t0 = time.time()
        self.pending.acquire()
        with self.lock:
            self.wait_s += time.time() - t0
        # host tables: strings (per file: header + opening text; per skeleton: its texts), polynomial indices,
        # the string before each polynomial and after each file's last one
        strings, skel_base, pre, post, polys, first = [], {}, [], [], [], []
        at = 0
        for j in jobs:
            sk = j.skel
            if id(sk) not in skel_base:
                skel_base[id(sk)] = len(strings)
                strings += sk.strings
            base = skel_base[id(sk)]
            pr = sk.ids + base
            pr[0] = len(strings)
            strings.append((j.head + sk.texts[0]).encode("ascii"))
            po = np.full(len(j.polys), -1, dtype=np.int32)
            po[-1] = base + sk.tail_id
            pre.append(pr)
            post.append(po)
            polys.append(j.polys)
            first.append(at * D)
            at += len(j.polys)
        pre, post, polys = np.concatenate(pre), np.concatenate(post), np.concatenate(polys)
        slen = np.array([0] + [len(s) for s in strings], dtype=np.int64)     # slen[id + 1]; slen[0] = none
        stroff = np.concatenate([[0], np.cumsum(slen[1:])]).astype(np.int64)
        strtab = np.frombuffer(b"".join(strings), dtype=np.uint8).copy()
        n = npoly * D
        ub = n * (p + 20) + int(slen[pre + 1].sum() + slen[post + 1].sum())
        u.ub = ub
        dev = e1.device
        L = m2.shape[0]
        ps = e1.numel()
        st = torch.cuda.current_stream()
        sptr = st.cuda_stream
        P = lambda t: t.data_ptr()
        # small tables through page-locked staging: asynchronous, the caller's thread does not wait for the stream
        up = lambda a: torch.from_numpy(np.ascontiguousarray(a)).pin_memory().to(dev, non_blocking=True)
        tpoly, tpre, tpost, tslen = up(polys), up(pre), up(post), up(slen)
        tstr = up(strtab) if strtab.size else torch.zeros(1, dtype=torch.uint8, device=dev)
        tsoff, tfirst = up(stroff), up(np.array(first, dtype=np.int64))
        pbase = tpoly * D
        sizes = torch.empty(npoly, dtype=torch.int32, device=dev)
        _lib.call("b3d_fmt_sizes", L, P(m2), P(e1), ps, P(pbase), npoly, D, eps, P(sizes), sptr)
        cnt = sizes.clamp(min=1)
        d = torch.arange(D, device=dev, dtype=torch.int32)
        active = d[None, :] < cnt[:, None]
        lastf = d[None, :] == (cnt - 1)[:, None]
        ent = torch.where((sizes > 0)[:, None], pbase[:, None] + d[None, :].to(torch.int64),
                          torch.full((1, 1), -1, dtype=torch.int64, device=dev)).reshape(-1).contiguous()
        last = torch.where(active, lastf.to(torch.uint8), torch.full((1, 1), 2, dtype=torch.uint8, device=dev))
        last = last.reshape(-1).contiguous()
        pre_el = torch.full((npoly, D), -1, dtype=torch.int32, device=dev)
        pre_el[:, 0] = tpre
        pre_el = pre_el.reshape(-1)
        post_el = torch.where(lastf, tpost[:, None], torch.full((1, 1), -1, dtype=torch.int32, device=dev))
        post_el = post_el.reshape(-1).contiguous()
        dg = torch.empty(n * ((p + 10 * L + 32 + 3) // 4 * 4), dtype=torch.uint8, device=dev)
        tx = torch.empty(n * ((p + 16 + 3) // 4 * 4), dtype=torch.uint8, device=dev)   # groups of four bytes
        ln = torch.empty(n, dtype=torch.int32, device=dev)
        _lib.call("b3d_fmt_text", L, P(m2), P(e1), ps, P(ent), n, p, P(dg), P(tx), P(ln), sptr)
        del dg
        bad = (ln < 0).sum().reshape(1).to(torch.int64)
        piece = torch.where(active.reshape(-1), ln.clamp(min=0).to(torch.int64) + 2 + 2 * (last == 0).to(torch.int64),
                            torch.zeros((), dtype=torch.int64, device=dev))
        piece = piece + tslen[(pre_el + 1).to(torch.int64)] + tslen[(post_el + 1).to(torch.int64)]
        end = torch.cumsum(piece, 0)
        off = (end - piece).contiguous()
        packed = torch.empty(ub, dtype=torch.uint8, device=dev)
        _lib.call("b3d_fmt_pack_files", P(tx), P(ln), P(last), P(pre_el), P(post_el), P(tstr), P(tsoff),
                  len(strings), P(off), n, P(packed), sptr)
        u.packed, u.dmeta = packed, torch.cat([off[tfirst], end[-1:], bad])
        u.ready = torch.cuda.Event()
        u.ready.record(st)
        self.q.put(u)

    def _dispatch(self):
        # this thread's HIP calls (page-locked allocations, event waits and synchronisations) run beside stream
        # captures of other threads (the serving loop's): in the relaxed capture mode they leave those alone
        if _lib.lib().b3d_thread_capture_mode(2) < 0:
            with self.lock:
                self.errors.append(_lib.B3DError(_lib.lib().b3d_last_error().decode()))
        while True:
            u = self.q.get()
            if u is None:
                return
            try:
                self._finish(u)
            except BaseException as err:                  # surfaced by drain() and the call's on_done
                self._release(u)
                if getattr(u, "packed", None) is not None:
                    u.packed = u.dmeta = None
                    self.pending.release()
                with self.lock:
                    self.errors.append(err)
                    self.inflight -= 1
                    self.lock.notify_all()
                u.call.unit_done(err)

    def _finish(self, u):
        if u.empty:
            datas = [_file_bytes(j, []) for j in u.jobs]
            self._write_all(u, [(j.path, x) for j, x in zip(u.jobs, datas)])
            return
        # the copy: a host buffer (waiting here, not in emit(), when the writers are behind), then the bytes
        torch = self.torch
        u.buf = self._buffer(u.ub)
        with self.gate():
            u.meta = torch.empty(u.dmeta.numel(), dtype=torch.int64, pin_memory=True)
            cs = self.copy_stream
            cs.wait_event(u.ready)
            with torch.cuda.stream(cs):
                u.buf[:u.ub].copy_(u.packed, non_blocking=True)
                u.meta.copy_(u.dmeta, non_blocking=True)
            done = torch.cuda.Event()
            done.record(cs)
            done.synchronize()
            u.packed = u.dmeta = None
        self.pending.release()
        meta = u.meta.numpy()
        F = len(u.jobs)
        starts, total, bad = meta[:F], int(meta[F]), int(meta[F + 1])
        if bad:
            # values outside the GPU formatter's range: this unit's files from the host formatter
            self.fallback_units += 1
            torch = self.torch
            allp = np.concatenate([j.polys for j in u.jobs])
            L, D = u.m2.shape[0], u.D
            with self.gate():
                sel = torch.from_numpy(allp).to(u.e1.device)
                m = u.m2.reshape(L, -1, D).index_select(1, sel).cpu().numpy().reshape(L, -1)
                e = u.e1.reshape(-1, D).index_select(0, sel).cpu().numpy().reshape(-1)
            local, at = [], 0
            for j in u.jobs:
                local.append(FileJob(j.path, j.head, j.skel, np.arange(at, at + len(j.polys))))
                at += len(j.polys)
            datas = assemble_host(m, e, D, u.p, u.eps, local)
            self._release(u)
            self._write_all(u, [(j.path, x) for j, x in zip(u.jobs, datas)])
            return
        ends = np.append(starts[1:], total)
        mv = memoryview(u.buf.numpy())
        self._write_all(u, [(j.path, mv[int(a):int(b)]) for j, a, b in zip(u.jobs, starts, ends)], release=True)

    def _release(self, u):
        if u.buf is not None:
            self.free.put(u.buf)
            u.buf = None

    def _write_all(self, u, items, release=False):
        left = [len(items)]
        lk = threading.Lock()
        u.m2 = u.e1 = None

        first_error = []

        def one(path, data):
            try:
                t0 = time.time()
                nb = _write_file(path, data, self.unlink)
                dt = time.time() - t0
                with self.lock:
                    self.write_s += dt
            except BaseException as err:
                with self.lock:
                    self.errors.append(err)
                first_error.append(err)
                nb = 0
            with lk:
                left[0] -= 1
                last = left[0] == 0
            with self.lock:
                self.files += 1
                self.bytes += nb
            if last:
                if release:
                    self._release(u)
                with self.lock:
                    self.inflight -= 1
                    self.lock.notify_all()
                u.call.unit_done(first_error[0] if first_error else None)

        if not items:
            with self.lock:
                self.inflight -= 1
                self.lock.notify_all()
            u.call.unit_done()
            return
        for path, data in items:
            self.pool.submit(one, path, data)
