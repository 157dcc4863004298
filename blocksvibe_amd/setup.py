"""blocks_3d's setup stages for the GPU chain: M matrices and h_infinity, produced by the C++ helper
b3d_setup (linked against blocks_3d's own sources) into one binary file, read here into device MPF arrays
(reference blocksvibe/setup.py:1-180, csrc/b3d_setup.cxx:15-21 for the format).

Binary layout (little-endian int64 words), K = 64-bit limbs per number:
  "B3DSETUP" K
  nM; nM x [ptype pvalue delta(0=12,1=43) two_jNN prime_pseudo_parity two_j rows cols]; all matrix entries
  ntasks; per task: q4[4] sign pp00 pp01 pp10 pp11 nn; per n: n count; count x [pp m j a b]; count numbers
  number = [size exp limb_0 .. limb_{K-1}] (GMP mpf: sign(size) sum_t limb_t 2^(64 t) 2^(64 (exp - K))).

write_synthetic() writes a file of the same format with seeded random values, for machines without the
blocks_3d sources (the cost of every GPU stage does not depend on the values).
"""
import os
import shutil
import subprocess

import numpy as np
import torch

from . import mpf
from . import structure as S

HERE = os.path.dirname(os.path.abspath(__file__))


def helper_path():
    """b3d_setup: $BLOCKSVIBE_B3D_SETUP, <repo>/build/b3d_setup, or on PATH"""
    for p in (os.environ.get("BLOCKSVIBE_B3D_SETUP"), os.path.join(HERE, "..", "build", "b3d_setup"),
              shutil.which("b3d_setup")):
        if p and os.path.isfile(p) and os.access(p, os.X_OK):
            return os.path.abspath(p)
    return None


def four_point_parity(two_js, q4):
    """blocks_3d's four_point_parity (M depends on q4 only through it)"""
    return (sum(a - b for a, b in zip(two_js, q4)) // 2) % 2


def run_helper(out, two_js, two_j12, two_j43, delta_12, delta_43, order, Lam, two_j_max, radial, q4s,
               precision, threads, skip_m=False):
    exe = helper_path()
    if exe is None:
        raise FileNotFoundError("b3d_setup not found (set BLOCKSVIBE_B3D_SETUP)")
    args = [exe, f"out={out}", f"precision={precision}", f"threads={threads}",
            "two_js=" + ",".join(map(str, two_js)), f"two_j12={two_j12}", f"two_j43={two_j43}",
            f"delta_12={delta_12}", f"delta_43={delta_43}", f"order={order}", f"lambda={Lam}",
            f"two_j_max={two_j_max}", f"radial={int(radial)}"]
    args += ["q4=" + ",".join(map(str, q)) + f",{s}" for q, s in q4s]
    if skip_m:
        args.append("skip_m=1")
    r = subprocess.run(args, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"b3d_setup failed ({r.returncode}): {r.stderr.strip()[-2000:]}")


def numbers_to_pairs_host(rec, K, L):
    """records (n, 2 + K) int64 -> list of normalised (m, e) (host, exact)"""
    out = []
    for row in rec:
        size, ex = int(row[0]), int(row[1])
        if size == 0:
            out.append((0, mpf.EZ))
            continue
        mag = int.from_bytes(row[2:].astype("<i8").tobytes(), "little", signed=False)
        out.append(mpf.norm_int(-mag if size < 0 else mag, 64 * (ex - K), L))
    return out


def numbers_to_mpf(rec, K, L):
    """records (n, 2 + K) int64 -> device MPF (n,), normalised on the GPU (b3d_mpf_norm_wide)"""
    n = rec.shape[0]
    if n == 0:
        return mpf.zeros((0,), L)
    size, ex = rec[:, 0], rec[:, 1]
    w = np.zeros((n, 2 * K + 1), dtype=np.uint32)                    # + one sign limb
    w[:, :2 * K] = np.ascontiguousarray(rec[:, 2:]).view(np.uint32).reshape(n, 2 * K)
    neg = size < 0
    if neg.any():                                                    # two's complement of the magnitudes
        x = w[neg].astype(np.uint64)
        x = (~x) & 0xFFFFFFFF
        carry = np.ones(x.shape[0], dtype=np.uint64)
        for k in range(x.shape[1]):
            v = x[:, k] + carry
            x[:, k] = v & 0xFFFFFFFF
            carry = v >> 32
        w[neg] = x.astype(np.uint32)
    ein = np.where(size == 0, 0, 64 * (ex - K)).astype(np.int64)
    W = torch.from_numpy(np.ascontiguousarray(w.T).view(np.int32)).to(mpf.dev())
    if 2 * K + 1 > 2 * L + 4:
        raise ValueError(f"setup numbers have {K} limbs, too wide for L={L}")
    return mpf.norm_wide(W, torch.from_numpy(ein).to(mpf.dev()), L)


class MTable:
    """M matrices of a setup file: index[key] -> record, key ((type, value), delta, two_jNN, ppp, two_j)"""

    def __init__(self, hdr, rec, K, L):
        self.hdr, self.rec, self.K, self.L = hdr, rec, K, L
        self.index = {((int(h[0]), int(h[1])), int(h[2]), int(h[3]), int(h[4]), int(h[5])): i
                      for i, h in enumerate(hdr)}
        cnt = hdr[:, 6] * hdr[:, 7] if len(hdr) else np.zeros(0, np.int64)
        self.start = np.concatenate([[0], np.cumsum(cnt)[:-1]]) if len(hdr) else np.zeros(0, np.int64)
        self._pairs = None

    def pairs(self, key):
        """rows of normalised (m, e) of the matrix of `key` (None if absent)"""
        i = self.index.get(key)
        if i is None:
            return None
        if self._pairs is None:
            self._pairs = numbers_to_pairs_host(self.rec, self.K, self.L)
        r, c = int(self.hdr[i, 6]), int(self.hdr[i, 7])
        s = int(self.start[i])
        return [self._pairs[s + a * c:s + (a + 1) * c] for a in range(r)]


class SetupTask:
    """one four-point structure: q4 (doubled), sign, pseudo-parities pp, h[n] = (index array (count, 5):
    pp m j a b, device MPF (count,))"""

    def __init__(self, q4, sign, pp, h):
        self.q4, self.sign, self.pp, self.h = q4, sign, pp, h


def read(path, L, m_path=None):
    """-> (MTable, [SetupTask]); M from m_path if given (an M-only cache file)"""
    if m_path is not None:
        M, _ = read(m_path, L)
        _, tasks = read(path, L)
        return M, tasks
    raw = np.fromfile(path, dtype=np.int64)
    if raw[:1].view("S8")[0] != b"B3DSETUP":
        raise ValueError(f"{path}: not a b3d_setup file")
    K = int(raw[1])
    pos = 2

    def words(k):
        nonlocal pos
        out = raw[pos:pos + k]
        pos += k
        return out

    nM = int(words(1)[0])
    hdr = words(8 * nM).reshape(nM, 8).copy()
    nvals = int((hdr[:, 6] * hdr[:, 7]).sum()) if nM else 0
    M = MTable(hdr, words((2 + K) * nvals).reshape(nvals, 2 + K).copy(), K, L)
    tasks = []
    for _ in range(int(words(1)[0])):
        t = words(10)
        q4, sign = tuple(int(x) for x in t[:4]), int(t[4])
        pp = ((int(t[5]), int(t[6])), (int(t[7]), int(t[8])))
        h = {}
        for _ in range(int(t[9])):
            n, cnt = (int(x) for x in words(2))
            idx = words(5 * cnt).reshape(cnt, 5).copy()
            h[n] = (idx, numbers_to_mpf(words((2 + K) * cnt).reshape(cnt, 2 + K), K, L))
        tasks.append(SetupTask(q4, sign, pp, h))
    if pos != len(raw):
        raise ValueError(f"{path}: trailing data")
    return M, tasks


def save_m(path, cache):
    """copy the M section of a setup file into `cache` (a setup file with no tasks), atomically"""
    raw = np.fromfile(path, dtype=np.int64)
    K, nM = int(raw[1]), int(raw[2])
    hdr = raw[3:3 + 8 * nM].reshape(nM, 8)
    end = 3 + 8 * nM + (2 + K) * int((hdr[:, 6] * hdr[:, 7]).sum())
    tmp = f"{cache}.{os.getpid()}"
    np.concatenate([raw[:end], np.zeros(1, dtype=np.int64)]).tofile(tmp)
    os.replace(tmp, cache)


def h_inf_tensor(task, ns, order, Jh, SL, SR, L):
    """h_inf MPF (2, order + 1, Jh, SL SR, len(ns)) for the lambda orders ns of a task (structure entry (a, b)
    at a SR + b)"""
    B = len(ns)
    out = mpf.zeros((2 * (order + 1) * Jh * SL * SR * B,), L)
    for k, n in enumerate(ns):
        idx, v = task.h[n]
        s, m, j, a, b = idx.T
        ok = (m <= order) & (j < Jh) & (a < SL) & (b < SR)
        flat = ((((s * (order + 1) + m) * Jh + j) * SL + a) * SR + b) * B + k
        fi = torch.from_numpy(flat[ok]).to(mpf.dev())
        okt = torch.from_numpy(ok).to(mpf.dev())
        out.m[:, fi] = v.m[:, okt]
        out.e[fi] = v.e[okt]
    return out.reshape(2, order + 1, Jh, SL * SR, B)


# ------------------------------------------------------------------------------------ synthetic files

def _number_words(rng, K, spread=2):
    """a random GMP-style record [size, exp, K limbs] with |x| in ~[2^-spread, 2^spread]"""
    limbs = rng.integers(0, 1 << 63, size=K, dtype=np.int64) * 2 + rng.integers(0, 2, size=K)
    limbs = limbs.astype(np.int64)
    limbs[-1] |= np.int64(1) << 62
    size = K if rng.integers(0, 2) else -K
    exp = int(rng.integers(-spread // 64, spread // 64 + 1)) if spread >= 64 else 0
    return np.concatenate([[size, exp], limbs])


def write_synthetic(path, precision, two_js, two_j12, two_j43, order, Lam, two_j_max, radial, q4s, seed=0,
                    pseudo_parities=None, skip_m=False):
    """a setup file with seeded random M matrices (every key the recursion needs) and h_inf entries
    (every (pp, m <= order, j, a, b) the chain reads) for the tasks q4s = [(q4, sign)].  Each task's h_inf
    comes from a generator seeded by (seed, sign, q4) alone, so it does not depend on skip_m or on the other
    tasks of the file."""
    from .plan import Plan
    rng = np.random.default_rng(seed)
    rounded = ((precision - 1) // 64 + 1) * 64
    K = rounded // 64 + 1
    pps = {}
    for q4, sign in q4s:
        par = four_point_parity(two_js, q4)
        pps[(tuple(q4), sign)] = pseudo_parities or (((0, 0), (1, 1)) if par == 0 else ((0, 1), (1, 0)))
    words = [np.frombuffer(b"B3DSETUP", dtype=np.int64), np.array([K], np.int64)]
    keys = {}
    if not skip_m:
        for pp in set(pps.values()):
            plan = Plan(two_j12, two_j43, two_j_max, order, pp, reuse=False)
            for key in plan.Mkeys:
                (t, k), d, tjNN, ppp, tj = key
                rows = len(S.so3((ppp + (1 if t == 3 else 0)) % 2, tjNN, tj))
                cols = len(S.so3(ppp, tjNN, int(S.pole_spin(np.int64(t), np.int64(k), np.int64(tj)))))
                keys[key] = (rows, cols)
    words.append(np.array([len(keys)], np.int64))
    for ((t, k), d, tjNN, ppp, tj), (r, c) in keys.items():
        words.append(np.array([t, k, d, tjNN, ppp, tj, r, c], np.int64))
    for (r, c) in keys.values():
        for _ in range(r * c):
            words.append(_number_words(rng, K))
    words.append(np.array([len(q4s)], np.int64))
    Jh = len(range(two_j12 % 2, two_j_max + 2 * order + 1, 2))
    for q4, sign in q4s:
        pp = pps[(tuple(q4), sign)]
        p0 = (1 - sign) // 2
        hrng = np.random.default_rng([seed, sign + 1] + [int(q) + 64 for q in q4])
        ns = [p0] if radial else list(range(p0, Lam + 1, 2))
        words.append(np.array(list(q4) + [sign, pp[0][0], pp[0][1], pp[1][0], pp[1][1], len(ns)], np.int64))
        idx = []
        for s in range(2):
            for m in range(order + 1):
                for j in range(Jh):
                    tj = two_j12 % 2 + 2 * j
                    for a in range(len(S.so3(pp[s][0], two_j12, tj))):
                        for b in range(len(S.so3(pp[s][1], two_j43, tj))):
                            idx.append((s, m, j, a, b))
        idx = np.array(idx, dtype=np.int64).reshape(-1, 5)
        for n in ns:
            words.append(np.array([n, len(idx)], np.int64))
            words.append(idx.reshape(-1))
            vals = hrng.integers(0, 1 << 62, size=(len(idx), K), dtype=np.int64)
            vals[:, -1] |= np.int64(1) << 62
            sizes = np.where(hrng.integers(0, 2, size=len(idx)) == 1, K, -K)
            exps = np.zeros(len(idx), dtype=np.int64)
            words.append(np.concatenate([sizes[:, None], exps[:, None], vals], 1).reshape(-1))
    tmp = f"{path}.{os.getpid()}"
    np.concatenate(words).astype(np.int64).tofile(tmp)
    os.replace(tmp, path)
