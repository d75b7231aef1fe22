"""tdm_solve_noalloc, the entry point that allocates nothing, against the
vtable's tdm.solve (which allocates its factor table per object): same
kernels, so the answers agree bit for bit.  Two calls in a row on the same
GPU, the second with a smaller system (the fixed table keeps the first
call's contents and must be rewritten), then a size beyond the table's
maximum, which must fail with an error and leave the data untouched."""
import ctypes as C

import numpy as np
import pytest

import simplevofsolver_amd as sva
from simplevofsolver_amd.hip import DeviceBuffer

pytestmark = pytest.mark.gpu


def _system(n, rng, periodic):
    l = rng.uniform(0.2, 1.5, n)
    u = rng.uniform(0.2, 1.5, n)
    c = -(l + u) - rng.uniform(0.05, 1.0, n)
    return l, c, u


def _rhs(n, nrhs, cplx, rng):
    q = rng.standard_normal((nrhs, n))
    return q + 1j * rng.standard_normal((nrhs, n)) if cplx else q


def _vtable(l, c, u, q, periodic):
    nrhs, n = q.shape
    cplx = np.iscomplexobj(q)
    t = sva.Tdm(n, nrhs, periodic, cplx, ndims=2)
    try:
        t.l[:], t.c[:], t.u[:] = l, c, u
        buf = DeviceBuffer.from_array(q)
        t.solve(buf.ptr)
        out = buf.download(q.dtype, q.shape)
        buf.free()
        return out
    finally:
        t.destruct()


def _noalloc(lib, l, c, u, q, periodic):
    nrhs, n = q.shape
    dl, dc, du = (DeviceBuffer.from_array(np.ascontiguousarray(a, dtype=np.float64)) for a in (l, c, u))
    dq = DeviceBuffer.from_array(q)
    try:
        rc = lib.tdm_solve_noalloc(n, nrhs, periodic, np.iscomplexobj(q), dl.ptr, dc.ptr, du.ptr, dq.ptr, None)
        assert rc == 0, lib.vof_last_error()
        # the copy on the null stream waits for the solve enqueued there
        return dq.download(q.dtype, q.shape)
    finally:
        for b in (dl, dc, du, dq):
            b.free()


def _lib():
    lib = sva.load(2)
    vp = C.c_void_p
    lib.tdm_solve_noalloc.argtypes = [C.c_int, C.c_int, C.c_bool, C.c_bool, vp, vp, vp, vp, vp]
    return lib


def test_noalloc_solve_matches_vtable_twice():
    lib = _lib()
    rng = np.random.default_rng(2024)
    cases = [(3000, 700, True, True), (40, 17, False, False)]   # the second smaller
    for n, nrhs, periodic, cplx in cases:
        print(f"noalloc n={n} nrhs={nrhs} periodic={periodic} complex={cplx}", flush=True)
        l, c, u = _system(n, rng, periodic)
        q = _rhs(n, nrhs, cplx, rng)
        want = _vtable(l, c, u, q, periodic)
        got = _noalloc(lib, l, c, u, q, periodic)
        assert np.array_equal(got, want)


def test_noalloc_rejects_sizes_beyond_its_table():
    lib = _lib()
    n = 16385
    q = np.ones((1, n))
    dq = DeviceBuffer.from_array(q)
    dl = DeviceBuffer.from_array(np.ones(n))
    try:
        rc = lib.tdm_solve_noalloc(n, 1, False, False, dl.ptr, dl.ptr, dl.ptr, dq.ptr, None)
        assert rc != 0
        assert b"exceeds" in lib.vof_last_error()
        assert np.array_equal(dq.download(np.float64, q.shape), q)
    finally:
        dq.free()
        dl.free()
