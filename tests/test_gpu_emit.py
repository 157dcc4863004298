"""The emitter's GPU path (b3d_fmt_sizes / b3d_fmt_text / b3d_fmt_pack_files, copy, writer threads) against the
host assembly from the library's host formatter (emit.assemble_host, itself checked against the host writer on
the CPU), file for file, byte for byte: several units, units from two threads on two streams, an all-zero
polynomial, a shape without polynomials, and a unit with a value outside the GPU formatter's range (its host
fallback)."""
import os
import threading

import numpy as np
import pytest
import torch

from oracle.mpf_ref import EZ

pytestmark = pytest.mark.gpu

ARGS = ("--j-external 2,2,2,2 --j-12 2 --j-43 1 --delta-12 0.25 --delta-43 -0.5 --j-internal 0-5 "
        "--delta-1-plus-2 2.3 --four-pt-struct 1,2,1,0 --four-pt-sign -1 --order 8 --lambda 3 "
        "--coordinates xt --kept-pole-order 8 --precision 512 --num-threads 3 -o {}/b{}_{{}}.json")


def planes(rng, L, npoly, D, big=False):
    m = rng.integers(-2 ** 31, 2 ** 31, size=(L, npoly, D), dtype=np.int64).astype(np.int32)
    m[L - 1] = (m[L - 1] & 0x3FFFFFFF) | 0x20000000
    m[L - 1] *= np.where(rng.random((npoly, D)) < 0.5, -1, 1).astype(np.int32)
    e = rng.integers(-600, -560, size=(npoly, D)).astype(np.int64)
    e[:, 2:] -= rng.choice([0, 300], size=(npoly, D - 2))                 # negligible top coefficients
    zero = rng.random((npoly, D)) < 0.1
    zero[5] = True
    m[:, zero] = 0
    e[zero] = EZ
    m, e = m.reshape(L, -1), e.reshape(-1)
    if big is not False:                                                   # outside the GPU formatter's range
        i = big * D + 1
        m[:, i] = rng.integers(-2 ** 31, 2 ** 31, size=L, dtype=np.int64).astype(np.int32)
        m[L - 1, i] = 0x20000000
        e[i] = 4
    return m, e


def jobs_for(tmp, tag, L, D, p, nq, q, with_empty=False):
    from blocksvibe_amd import b3djson as BJ
    from blocksvibe_amd import emit as EM
    from blocksvibe_amd import structure as S
    from blocksvibe_amd.taskspec import TaskSpec
    t = TaskSpec(ARGS.format(tmp, tag).split())
    pp = ((0, 1), (1, 0))
    tjs = t.two_js_internal
    Jo = len(tjs)
    SR = max(len(S.so3(pp[k][1], t.two_j43, tj)) for k in range(2) for tj in tjs)
    SL = max(len(S.so3(pp[k][0], t.two_j12, tj)) for k in range(2) for tj in tjs)
    outs = [(n, b) for n in range(4) for b in range((3 - n) // 2 + 1)]
    perm, rows = EM.output_rows(outs)
    jobs = []
    for jo, tj in enumerate(tjs):
        structs = [(len(S.so3(pp[k][0], t.two_j12, tj)), len(S.so3(pp[k][1], t.two_j43, tj))) for k in range(2)]
        polys = EM.raw_polys(structs, SR, jo, Jo, SL * SR, nq, q, [perm], [0])
        jobs.append(EM.FileJob(t.path(tj / 2.0), BJ.header(t, tj, pp, p), EM.skeleton(p, structs, [rows]), polys))
    if with_empty:
        sk = EM.skeleton(p, [(0, 1), (0, 0)], [rows])
        jobs.append(EM.FileJob(os.path.join(str(tmp), f"empty{tag}.json"), "{\n", sk, []))
    return jobs, len(outs) * Jo * 2 * SL * SR * nq


@pytest.mark.parametrize("big", [False, True])
def test_gpu_files_equal_host_assembly(tmp_path, big):
    from blocksvibe_amd import b3djson as BJ
    from blocksvibe_amd import emit as EM
    L, D, nq = 18, 6, 3
    p = BJ.digits10(512)
    eps = BJ.trim_exponent(p)
    rng = np.random.default_rng(11 if big else 12)
    jobs0, npoly = jobs_for(tmp_path, 0, L, D, p, nq, 0, with_empty=True)
    jobs1, _ = jobs_for(tmp_path, 1, L, D, p, nq, 2)
    m, e = planes(rng, L, npoly, D, big=int(jobs0[1].polys[2]) if big else False)
    dm, de = torch.from_numpy(m).cuda(), torch.from_numpy(e).cuda()
    print(f"emit L={L} D={D} polys={npoly} files={len(jobs0) + len(jobs1)} big={big}", flush=True)
    em = EM.Emitter(writers=4, buffers=2, unit_slots=3 * D * 40)          # several units per call
    done = []
    em.emit(dm, de, D, p, eps, jobs0, on_done=lambda err: done.append(err))
    s = torch.cuda.Stream()
    th = threading.Thread(target=lambda: (torch.cuda.set_stream(s), em.emit(dm, de, D, p, eps, jobs1)))
    th.start()
    th.join()
    em.close()
    assert done == [None]
    assert (em.fallback_units > 0) == big
    want = EM.assemble_host(m, e, D, p, eps, jobs0 + jobs1)
    for j, w in zip(jobs0 + jobs1, want):
        assert open(j.path, "rb").read() == w, j.path
