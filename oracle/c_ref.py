"""ctypes access to oracle/c/liboracle_c.so, the C restatement of mpf_ref's Cauchy-term and pole-step
arithmetic.  TEST INFRASTRUCTURE: see oracle/__init__.py.  Build with `python oracle/c/build.py` (done by
__graft_entry__.build())."""
import ctypes
import os

from . import mpf_ref as R

_LIB = None
_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c", "liboracle_c.so")


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(_PATH):
            from .c import build as B
            B.build()
        lb = ctypes.CDLL(_PATH)
        P, I64, I = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
        lb.oc_cauchy.argtypes = [P, P, P, I, I, P, I64, P]
        lb.oc_cauchy.restype = I64
        lb.oc_dot_aligned.argtypes = [P, I64, P, I64, I, I, P]
        lb.oc_dot_aligned.restype = I64
        lb.oc_norm.argtypes = [P, I, I64, I, P]
        lb.oc_norm.restype = I64
        lb.oc_time.argtypes = [I, I, I, I, ctypes.c_double]
        lb.oc_time.restype = ctypes.c_double
        _LIB = lb
    return _LIB


def _limbs(ms, L):
    arr = (ctypes.c_uint32 * (max(1, len(ms)) * L))()
    for i, m in enumerate(ms):
        for k, v in enumerate(R.to_limbs(m, L)):
            arr[i * L + k] = v
    return arr


def _result(out, e, L):
    m = R.from_limbs(list(out)[:L])
    return (m, e) if m else (0, R.EZ)


def norm(W, n, e, L):
    """mpf_ref.norm of a two's complement integer held in n limbs"""
    w = _limbs([W], n)
    out = (ctypes.c_uint32 * L)()
    e2 = lib().oc_norm(w, n, e, L, out)
    return _result(out, e2, L)


def cauchy(xs, ds, L, h=(0, R.EZ)):
    K = len(xs)
    xm = _limbs([m for m, _ in xs], L)
    xe = (ctypes.c_int64 * max(1, K))(*[e for _, e in xs])
    d = (ctypes.c_int32 * max(1, K))(*ds)
    out = (ctypes.c_uint32 * L)()
    e = lib().oc_cauchy(xm, xe, d, K, L, _limbs([h[0]], L), h[1], out)
    return _result(out, e, L)


def dot_aligned(am, ae, bm, be, L):
    K = len(am)
    out = (ctypes.c_uint32 * L)()
    e = lib().oc_dot_aligned(_limbs(am, L), ae, _limbs(bm, L), be, K, L, out)
    return _result(out, e, L)


def time_unit(kind, L, K, threads, seconds):
    """aggregate seconds per unit on `threads` concurrent threads: kind 0 one Cauchy term, 1 one K-term
    dot product"""
    return lib().oc_time(kind, L, K, threads, seconds)
