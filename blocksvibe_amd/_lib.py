"""ctypes binding of libb3dhip.so (include/b3dhip.h).  The HIP library is the only compute path: if it is
missing or fails to load, every numeric entry point raises -- there is no CPU fallback."""
import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libb3dhip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "b3dhip.h")

_lib = None

P, I64, I32, SZ = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int, ctypes.c_size_t
_C = {"const int32_t*": P, "int32_t*": P, "const int64_t*": P, "int64_t*": P, "void*": P, "const void*": P,
      "const int8_t*": P, "int8_t*": P, "const uint32_t*": P, "uint32_t*": P, "const uint16_t*": P,
      "uint16_t*": P, "const double*": P, "double*": P, "int64_t": I64, "int": I32, "size_t": SZ,
      "int*": P, "float*": P, "const float*": P, "char*": P,
      "uint8_t*": P, "const uint8_t*": P, "void**": P}


class B3DError(RuntimeError):
    pass


def declarations():
    """(name, return type, [argument types]) of every function declared in include/b3dhip.h"""
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    out = []
    for m in re.finditer(r"^\s*(const char\*|int64_t|int|void)\s+(b3d_\w+)\s*\(([^)]*)\)\s*;", txt, flags=re.M):
        ret, name, args = m.group(1), m.group(2), m.group(3).strip()
        types = []
        if args and args != "void":
            for a in args.split(","):
                a = " ".join(a.split())
                t = a.rsplit(" ", 1)[0] if not a.endswith("*") else a
                t = t.replace(" *", "*")
                if t not in _C:
                    raise ValueError(f"unmapped C type {t!r} in {name}")
                types.append(_C[t])
        out.append((name, ret, types))
    return out


def constants():
    """the integer #define constants of include/b3dhip.h (B3D_...)"""
    return {m.group(1): int(m.group(2)) for m in re.finditer(r"^#define\s+(B3D_\w+)\s+(-?\d+)\s*$", open(HEADER).read(),
                                                               flags=re.M)}


def lib():
    """the loaded library (raises B3DError if it is not built or does not load)"""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise B3DError(f"{LIBPATH} is not built (python -m blocksvibe_amd.build)")
        try:
            L = ctypes.CDLL(LIBPATH)
        except OSError as e:
            raise B3DError(f"cannot load {LIBPATH}: {e}")
        for name, ret, types in declarations():
            f = getattr(L, name)
            f.restype = {"const char*": ctypes.c_char_p, "void": None, "int64_t": I64}.get(ret, I32)
            f.argtypes = types
        _lib = L
    return _lib


def call(name, *args):
    """call a status-returning entry point; nonzero status -> B3DError with the library's message"""
    st = getattr(lib(), name)(*args)
    if st != 0:
        raise B3DError(f"{name} failed ({st}): {lib().b3d_last_error().decode()}")
