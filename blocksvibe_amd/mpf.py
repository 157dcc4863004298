"""Device multiprecision floats (MPF) for the HIP kernels, and exact host-side conversions.

An MPF holds `m`, a torch int32 tensor of shape (L, *shape) (limb planes, two's complement, limb 0 least
significant), and `e`, a torch int64 tensor of `shape`; value = m * 2^e.  Normalised: bitlen(|m|) = NB =
32 L - 2, zero = (0, EZ).  The format replaces the reference's unvendored sdpvibe.mpf.MPF (reference
blocksvibe/setup.py:72, cauchy.py:86); the limb-count rule is reference cli.py:270-274.

torch is used for device memory and streams only; every arithmetic operation is a libb3dhip kernel.
"""
import numpy as np
import torch

from . import _lib

EZ = -(1 << 62)


def NB(L):
    return 32 * L - 2


def limbs_for_precision(precision):
    """limbs for a blocks_3d precision: rounded up to 64-bit words, plus >= 60 guard bits
    (reference cli.py:270-274)"""
    rounded = ((precision - 1) // 64 + 1) * 64
    return -(-(rounded + 62) // 32)


def stream_ptr():
    return torch.cuda.current_stream().cuda_stream


def dev():
    return torch.device("cuda")


class MPF:
    __slots__ = ("m", "e")

    def __init__(self, m, e):
        self.m, self.e = m, e

    @property
    def L(self):
        return self.m.shape[0]

    @property
    def shape(self):
        return tuple(self.e.shape)

    def numel(self):
        return self.e.numel()

    def reshape(self, *shape):
        return MPF(self.m.reshape(self.L, *shape), self.e.reshape(*shape))

    def contiguous(self):
        return MPF(self.m.contiguous(), self.e.contiguous())

    def __getitem__(self, idx):
        """index the entry dimensions (torch semantics)"""
        if not isinstance(idx, tuple):
            idx = (idx,)
        return MPF(self.m[(slice(None),) + idx], self.e[idx])

    def take(self, idx):
        """entries along the first entry dimension (idx: long tensor)"""
        return MPF(self.m[:, idx], self.e[idx])

    def permute(self, *perm):
        return MPF(self.m.permute(0, *[p + 1 for p in perm]).contiguous(), self.e.permute(*perm).contiguous())

    def clone(self):
        return MPF(self.m.clone(), self.e.clone())

    def to_host(self):
        """-> list of (mantissa int, exponent int) in row-major entry order"""
        return to_pairs(self.m.cpu().numpy(), self.e.cpu().numpy())


def zeros(shape, L, device=None):
    device = device or dev()
    if isinstance(shape, int):
        shape = (shape,)
    return MPF(torch.zeros((L,) + tuple(shape), dtype=torch.int32, device=device),
               torch.full(tuple(shape), EZ, dtype=torch.int64, device=device))


def empty(shape, L, device=None):
    """uninitialised MPF, for outputs that a kernel writes in full before anything reads them"""
    device = device or dev()
    if isinstance(shape, int):
        shape = (shape,)
    return MPF(torch.empty((L,) + tuple(shape), dtype=torch.int32, device=device),
               torch.empty(tuple(shape), dtype=torch.int64, device=device))


def cat(parts, dim):
    return MPF(torch.cat([p.m for p in parts], dim + 1), torch.cat([p.e for p in parts], dim))


# ----------------------------------------------------------------------------- exact host conversions

def norm_int(W, e, L):
    """exact W * 2^e -> normalised (m, e'), NB bits, round half away from zero"""
    if W == 0:
        return 0, EZ
    a = -W if W < 0 else W
    nb = 32 * L - 2
    s = a.bit_length() - nb
    if s > 0:
        a = (a + (1 << (s - 1))) >> s
        if a.bit_length() > nb:
            a >>= 1
            s += 1
    else:
        a <<= -s
    return (-a if W < 0 else a), e + s


def norm_fraction(num, den, L):
    """exact num/den (den > 0) -> normalised, round half away from zero"""
    if num == 0:
        return 0, EZ
    a = -num if num < 0 else num
    nb = 32 * L - 2
    e = a.bit_length() - den.bit_length() - nb

    def quo(e):
        n_, d_ = (a << -e, den) if e <= 0 else (a, den << e)
        q, r = divmod(n_, d_)
        return q, r, d_

    q, r, d = quo(e)
    while q.bit_length() > nb:
        e += 1
        q, r, d = quo(e)
    while q.bit_length() < nb:
        e -= 1
        q, r, d = quo(e)
    if 2 * r >= d:
        q += 1
        if q.bit_length() > nb:
            q >>= 1
            e += 1
    return (-q if num < 0 else q), e


def from_mpmath(x, L):
    """mpmath mpf (exact binary) -> normalised (m, e)"""
    import mpmath as mp
    sign, man, exp, bc = mp.mpf(x)._mpf_
    if not man:
        return 0, EZ
    return norm_int(-int(man) if sign else int(man), int(exp), L)


def to_limb_array(mants, L):
    """signed mantissas -> uint32 numpy array (L, n), two's complement"""
    n = len(mants)
    mask = (1 << (32 * L)) - 1
    buf = b"".join((int(v) & mask).to_bytes(4 * L, "little") for v in mants)
    return np.frombuffer(buf, dtype="<u4").reshape(n, L).T.copy() if n else np.zeros((L, 0), np.uint32)


def to_pairs(m, e):
    """host limb planes (L, ...) int32 and exponents (...) -> list of (mantissa, exponent) ints"""
    L = m.shape[0]
    m = np.ascontiguousarray(m.reshape(L, -1).T).view(np.uint8)
    e = e.reshape(-1)
    out = []
    fb = int.from_bytes
    for i in range(e.shape[0]):
        v = fb(m[i].tobytes(), "little", signed=True)
        out.append((v, int(e[i]) if v else EZ))
    return out


def from_pairs(pairs, L, shape, device=None):
    """list of normalised (m, e) -> device MPF"""
    device = device or dev()
    m = to_limb_array([p[0] for p in pairs], L).view(np.int32)
    e = np.array([p[1] if p[0] else EZ for p in pairs], dtype=np.int64)
    return MPF(torch.from_numpy(m).reshape((L,) + tuple(shape)).to(device),
               torch.from_numpy(e).reshape(tuple(shape)).to(device))


def from_ints(ints, exps, L, shape, device=None):
    """exact values ints[i] * 2^exps[i] -> normalised device MPF"""
    return from_pairs([norm_int(int(v), int(x), L) for v, x in zip(ints, exps)], L, shape, device)


def from_mp(values, L, shape, device=None):
    return from_pairs([from_mpmath(v, L) for v in values], L, shape, device)


def from_fractions(values, L, shape, device=None):
    return from_pairs([norm_fraction(v.numerator, v.denominator, L) for v in values], L, shape, device)


# ----------------------------------------------------------------------------- device operations

def _ptr(t):
    return t.data_ptr()


def add(a, b, out=None):
    """elementwise a + b (b may have a single entry: broadcast)"""
    L, n = a.L, a.numel()
    out = out or empty(a.shape, L, a.e.device)
    a, b = a.contiguous(), b.contiguous()
    sb = 0 if b.numel() == 1 else 1
    _lib.call("b3d_mpf_add", L, _ptr(a.m), _ptr(a.e), a.numel(), 1, _ptr(b.m), _ptr(b.e), b.numel(), sb,
              _ptr(out.m), _ptr(out.e), out.numel(), n, stream_ptr())
    return out


def mul(a, b, out=None):
    """elementwise a * b (b may have a single entry: broadcast)"""
    L, n = a.L, a.numel()
    out = out or empty(a.shape, L, a.e.device)
    a, b = a.contiguous(), b.contiguous()
    sb = 0 if b.numel() == 1 else 1
    _lib.call("b3d_mpf_mul", L, _ptr(a.m), _ptr(a.e), a.numel(), 1, _ptr(b.m), _ptr(b.e), b.numel(), sb,
              _ptr(out.m), _ptr(out.e), out.numel(), n, stream_ptr())
    return out


def align_rows(a):
    """a: MPF (..., K) -> (aligned mantissas (L, ..., K) int32, row exponents (...))"""
    a = a.contiguous()
    K = a.shape[-1]
    rows = a.numel() // K if K else 0
    om = torch.empty_like(a.m)
    oe = torch.empty(a.shape[:-1], dtype=torch.int64, device=a.e.device)
    _lib.call("b3d_mpf_align_rows", a.L, _ptr(a.m), _ptr(a.e), a.numel(), rows, K, K, 1, _ptr(om), om[0].numel(),
              _ptr(oe), stream_ptr())
    return om, oe


def norm_wide(W, ein, L):
    """W: int32 tensor (LW, n) of wide two's complement integers, ein: int64 (n) -> normalised MPF (n)"""
    LW, n = W.shape[0], W.shape[1]
    out = empty((n,), L, W.device)
    _lib.call("b3d_mpf_norm_wide", L, LW, _ptr(W), n, _ptr(ein), _ptr(out.m), _ptr(out.e), n, n, stream_ptr())
    return out


def random(shape, L, spread=2, generator=None, device=None):
    """random normalised MPF with |x| in ~[2^-spread, 2^spread] (synthetic operators: the cost of every
    kernel is independent of the values), built on the host from a seeded numpy generator"""
    n = int(np.prod(shape)) if shape else 1
    rng = generator if generator is not None else np.random.default_rng(0)
    words = rng.integers(0, 1 << 32, size=(n, L), dtype=np.uint64).astype(np.uint32)
    nb = 32 * L - 2
    words[:, L - 1] = (words[:, L - 1] & 0x3FFFFFFF) | 0x20000000          # bitlen exactly NB
    neg = rng.integers(0, 2, size=n).astype(bool)
    m = words.T.copy()                                                      # (L, n) magnitudes
    if neg.any():                                                           # two's complement of negatives
        mm = m[:, neg].astype(np.uint64)
        mm = (~mm) & 0xFFFFFFFF
        carry = np.ones(mm.shape[1], dtype=np.uint64)
        for k in range(L):
            v = mm[k] + carry
            mm[k] = v & 0xFFFFFFFF
            carry = v >> 32
        m[:, neg] = mm.astype(np.uint32)
    e = rng.integers(-spread, spread + 1, size=n).astype(np.int64) - nb
    device = device or dev()
    return MPF(torch.from_numpy(m.view(np.int32)).reshape((L,) + tuple(shape)).to(device),
               torch.from_numpy(e).reshape(tuple(shape)).to(device))


def random_device(shape, L, spread=2, seed=0, device=None):
    """random normalised MPF generated on the device (uniform mantissa bits with bitlen exactly NB, random
    sign, exponent spread +-spread around 2^0): synthetic inputs without a host round trip"""
    device = device or dev()
    n = int(np.prod(shape)) if shape else 1
    g = torch.Generator(device=device)
    g.manual_seed(seed)
    m = torch.randint(-(1 << 31), (1 << 31) - 1, (L, n), dtype=torch.int32, device=device, generator=g)
    m[L - 1] = (m[L - 1] & 0x3FFFFFFF) | 0x20000000                       # magnitude bitlen exactly NB
    neg = torch.randint(0, 2, (n,), device=device, generator=g).bool()
    W = torch.cat([m, torch.zeros((1, n), dtype=torch.int32, device=device)], 0)
    e = torch.randint(-spread, spread + 1, (n,), dtype=torch.int64, device=device, generator=g) - (32 * L - 2)
    out = norm_wide(W, e, L)                                              # already normalised: exact copy
    sgn = torch.zeros((L, n), dtype=torch.int32, device=device)           # +-2^(NB-1) * 2^-(NB-1)
    sgn[L - 1] = torch.where(neg, torch.tensor(-(1 << 29), dtype=torch.int32, device=device),
                             torch.tensor(1 << 29, dtype=torch.int32, device=device))
    one = MPF(sgn, torch.full((n,), -(32 * L - 3), dtype=torch.int64, device=device))
    return mul(out, one).reshape(*shape)
