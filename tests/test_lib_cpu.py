"""The C-ABI library (no GPU needed): libb3dhip.so loads, exports every entry point declared in
include/b3dhip.h with the header's ABI version, and reports its compiled limb counts."""
import ctypes

import pytest


def test_exports_every_declared_symbol():
    from blocksvibe_amd import _lib
    lib = ctypes.CDLL(_lib.LIBPATH)
    decl = _lib.declarations()
    assert len(decl) >= 10
    for name, _, _ in decl:
        assert hasattr(lib, name), name


def test_abi_and_limbs():
    import re
    from blocksvibe_amd import _lib
    v = int(re.search(r"#define B3D_ABI_VERSION (\d+)", open(_lib.HEADER).read()).group(1))
    assert _lib.lib().b3d_abi_version() == v
    for L in (10, 18, 26, 34, 42, 50):
        assert _lib.lib().b3d_supported_limbs(L) == 1
    assert _lib.lib().b3d_supported_limbs(7) == 0


def test_limb_rule():
    from blocksvibe_amd.mpf import limbs_for_precision
    # reference cli.py:270-274: 256 -> 10, 512 -> 18, 768 -> 26, 1024 -> 34, 1280 -> 42, 1536 -> 50
    assert [limbs_for_precision(p) for p in (256, 512, 768, 1024, 1280, 1536)] == [10, 18, 26, 34, 42, 50]
    assert limbs_for_precision(384) == 14 and limbs_for_precision(200) == 10


def test_cauchy_rows_per_lane_checked():
    """b3d_cauchy_level rejects a wave table for other than one or two rows per lane before it launches
    anything (the pointers below are never dereferenced)"""
    import numpy as np
    from blocksvibe_amd import _lib
    buf = np.zeros(64, np.int64)
    p = buf.ctypes.data
    with pytest.raises(_lib.B3DError, match="rows_per_lane"):
        _lib.call("b3d_cauchy_level", 34, p, p, 1, 1, p, p, 1, p, p, p, 1, p, p, p, 1, p, p, 1, p, p, p, 1, 3, 0)


def test_plan_builder_arguments():
    """b3d_plan_build (host code) rejects arguments outside its domain with a message, and b3d_plan_get answers
    -1 for an unknown table or a level out of range; the plan object frees cleanly"""
    from blocksvibe_amd import _lib
    h = ctypes.c_void_p()
    for args in ((2, 1, 10, 5, 0, 0, 1, 1), (2, 2, 10, -1, 0, 0, 1, 1), (2, 2, 10, 5, 0, 2, 1, 1)):
        with pytest.raises(_lib.B3DError, match="b3d_plan_build"):
            _lib.call("b3d_plan_build", *args, 1, ctypes.byref(h))
    _lib.call("b3d_plan_build", 2, 2, 6, 4, 0, 0, 1, 1, 1, ctypes.byref(h))
    K = _lib.constants()
    lib = _lib.lib()
    assert lib.b3d_plan_get(h, K["B3D_PLAN_SCALARS"], 0, None, 0) == 10
    assert lib.b3d_plan_get(h, 99, 0, None, 0) == -1
    assert lib.b3d_plan_get(h, K["B3D_PLAN_COL"], 4, None, 0) == -1
    assert lib.b3d_plan_get(h, K["B3D_PLAN_LEVEL_TJS"], -1, None, 0) == -1
    lib.b3d_plan_free(h)
