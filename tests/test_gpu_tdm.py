"""The exported tri-diagonal solver (the reference's ``tdm`` vtable,
src/tdm.c:66-106,147-324) and ``linear_system_init`` (src/linear_system.c:28-85)
on the GPU, through the C ABI, against the oracle's restatement of tdm.c.

Bar: bit for bit.  The GPU computes the factors of a system once per call
with the reference's operations in the reference's order, and every line then
runs the same recurrences as the reference's gtsv, so each value is the same
sequence of IEEE operations.  Cases: real and complex right-hand sides,
periodic (Sherman-Morrison) and not, sizes from the smallest the reference's
loops allow to lines longer than the LDS tiles hold, up to 10^4 right-hand
sides, and both singular guards (tdm.c:39-44 and :94)."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as OL
import simplevofsolver_amd as sva
from simplevofsolver_amd.hip import DeviceBuffer

pytestmark = pytest.mark.gpu


def _system(n, rng, periodic, singular=False):
    if singular:
        # the discrete Laplacian: Neumann ends (last pivot 0) or periodic (den 0)
        l, u = np.ones(n), np.ones(n)
        c = -2.0 * np.ones(n)
        if not periodic:
            c[0] = c[-1] = -1.0
        return l, c, u
    l = rng.uniform(0.2, 1.5, n)
    u = rng.uniform(0.2, 1.5, n)
    c = -(l + u) - rng.uniform(0.05, 1.0, n)
    return l, c, u


def _rhs(n, nrhs, cplx, rng, singular):
    q = rng.standard_normal((nrhs, n))
    if cplx:
        q = q + 1j * rng.standard_normal((nrhs, n))
    if singular:
        q -= q.mean(axis=1, keepdims=True)
    return q


def _gpu_solve(l, c, u, q, periodic, ndims=2):
    nrhs, n = q.shape
    cplx = np.iscomplexobj(q)
    t = sva.Tdm(n, nrhs, periodic, cplx, ndims=ndims)
    try:
        t.l[:], t.c[:], t.u[:] = l, c, u
        host = np.ascontiguousarray(q, dtype=np.complex128 if cplx else np.float64)
        buf = DeviceBuffer.from_array(host, ndims)
        t.solve(buf.ptr)
        out = buf.download(host.dtype, host.shape)
        buf.free()
        return out
    finally:
        t.destruct()


SIZES = [(2, 5), (3, 1), (3, 7), (5, 64), (16, 300), (64, 10000), (255, 33), (256, 10000), (257, 65),
         (1000, 9), (4099, 3), (9000, 3)]


@pytest.mark.parametrize("n,nrhs", SIZES, ids=[f"n{n}-rhs{r}" for n, r in SIZES])
@pytest.mark.parametrize("periodic", [False, True], ids=["thomas", "periodic"])
@pytest.mark.parametrize("cplx", [False, True], ids=["real", "complex"])
def test_tdm_matches_oracle(n, nrhs, periodic, cplx):
    if periodic and n < 3:
        pytest.skip("a periodic system needs n >= 3 (gtsv on n - 1 rows reads v[n - 3])")
    print(f"tdm n={n} nrhs={nrhs} periodic={periodic} complex={cplx}", flush=True)
    rng = np.random.default_rng(1000 * n + nrhs + 2 * periodic + cplx)
    l, c, u = _system(n, rng, periodic)
    q = _rhs(n, nrhs, cplx, rng, False)
    ref = OL.tdm_solve(l, c, u, q.ravel(), periodic=periodic).reshape(q.shape)
    got = _gpu_solve(l, c, u, q, periodic)
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("n", [8, 256, 512])
@pytest.mark.parametrize("periodic", [False, True], ids=["neumann", "periodic"])
@pytest.mark.parametrize("cplx", [False, True], ids=["real", "complex"])
def test_tdm_singular_guards(n, periodic, cplx):
    """The Laplacian's zero mode: the last pivot (Neumann ends, tdm.c:39-44)
    or the Sherman-Morrison denominator (periodic, tdm.c:94) falls below
    DBL_EPSILON, or stays just above it by roundoff; the GPU takes the
    reference's branch in every case."""
    rng = np.random.default_rng(n + 7 * periodic + cplx)
    l, c, u = _system(n, rng, periodic, singular=True)
    q = _rhs(n, 40, cplx, rng, True)
    ref = OL.tdm_solve(l, c, u, q.ravel(), periodic=periodic).reshape(q.shape)
    got = _gpu_solve(l, c, u, q, periodic)
    assert np.array_equal(got, ref)


def test_tdm_reuse_and_3d_library():
    """One object solved twice with new coefficients (the reference's
    predictor refills l, c, u between solves) and the 3D library's copy."""
    rng = np.random.default_rng(3)
    n, nrhs = 40, 100
    t = sva.Tdm(n, nrhs, False, False, ndims=3)
    try:
        for _ in range(2):
            l, c, u = _system(n, rng, False)
            t.l[:], t.c[:], t.u[:] = l, c, u
            q = rng.standard_normal((nrhs, n))
            buf = DeviceBuffer.from_array(q, 3)
            t.solve(buf.ptr)
            got = buf.download(np.float64, q.shape)
            ref = OL.tdm_solve(l, c, u, q.ravel()).reshape(q.shape)
            assert np.array_equal(got, ref)
        sz, nr = C.c_int(), C.c_int()
        assert t.vt.get_size(t.info, C.byref(sz)) == 0 and sz.value == n
        assert t.vt.get_nrhs(t.info, C.byref(nr)) == 0 and nr.value == nrhs
    finally:
        t.destruct()


@pytest.mark.parametrize("ndims", [2, 3])
def test_linear_system_init_builds_the_x_solver(ndims):
    """linear_system_init (linear_system.c:28-85): x1pncl holds the whole
    grid, tdm_x is a non-periodic real system of size glsizes[0] with
    glsizes[1] (x glsizes[2]) right-hand sides; solving x1pncl through it
    matches the oracle; a second init of the same object fails."""
    lib = sva.load(ndims)
    LS = sva._linear_system_struct(ndims)
    gl = (24, 16, 6)[:ndims]
    ls = LS()
    implicit = (C.c_bool * ndims)(True, *([False] * (ndims - 1)))
    sizes = (C.c_size_t * ndims)(*gl)
    assert lib.linear_system_init(None, implicit, sizes, C.byref(ls)) == 0
    try:
        assert ls.is_initialised and list(ls.x1pncl_mysizes) == list(gl)
        vt = sva.tdm_vtable(ndims)
        sz, nr = C.c_int(), C.c_int()
        assert vt.get_size(ls.tdm_x, C.byref(sz)) == 0 and sz.value == gl[0]
        assert vt.get_nrhs(ls.tdm_x, C.byref(nr)) == 0 and nr.value == int(np.prod(gl[1:]))
        # fill the host coefficient arrays as the predictor does, the pencil on the device
        rng = np.random.default_rng(ndims)
        l, c, u = _system(gl[0], rng, False)
        for name, arr in (("get_l", l), ("get_c", c), ("get_u", u)):
            p = C.POINTER(C.c_double)()
            assert getattr(vt, name)(ls.tdm_x, C.byref(p)) == 0
            np.ctypeslib.as_array(p, shape=(gl[0],))[:] = arr
        q = rng.standard_normal((nr.value, gl[0]))
        sva.hip.copy(ls.x1pncl, q.ctypes.data_as(C.c_void_p), q.nbytes, 1, ndims)
        assert vt.solve(ls.tdm_x, ls.x1pncl) == 0
        got = np.empty_like(q)
        sva.hip.copy(got.ctypes.data_as(C.c_void_p), ls.x1pncl, q.nbytes, 2, ndims)
        assert np.array_equal(got, OL.tdm_solve(l, c, u, q.ravel()).reshape(q.shape))
        assert lib.linear_system_init(None, implicit, sizes, C.byref(ls)) != 0
        assert b"already initialised" in lib.vof_last_error()
    finally:
        assert lib.linear_system_finalise(C.byref(ls)) == 0
    assert lib.linear_system_finalise(C.byref(ls)) != 0
