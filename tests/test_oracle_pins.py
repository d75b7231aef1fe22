"""Pins of the CPU oracle (no GPU).

The reference ships no tests, fixtures or golden vectors, and it cannot be
compiled here (empty SimpleDecomp / SimpleNpyIO submodules, no FFTW), so each
oracle function is pinned by what the mathematics fixes: dense solves,
library transforms, closed forms and invariants.  Each check is chosen so that
a plausible slip (dropped term, wrong sign, index or weight) fails it.
"""
import numpy as np
import pytest
import scipy.fft
import scipy.optimize

import oracle_lib as OL
from oracle_lib import Oracle
from simplevofsolver_amd.ic import InitialCondition, make_xgrid, random_fields, vof_discs, _pad_cells

RNG = np.random.default_rng(1234)


# --- tdm.c: Thomas and Sherman-Morrison --------------------------------------

def _dense(l, c, u, periodic):
    n = len(c)
    A = np.diag(c) + np.diag(u[:-1], 1) + np.diag(l[1:], -1)
    if periodic:
        A[0, -1] = l[0]
        A[-1, 0] = u[-1]
    return A


@pytest.mark.parametrize("periodic", [False, True])
@pytest.mark.parametrize("cplx", [False, True])
@pytest.mark.parametrize("n", [3, 5, 16])
def test_tdm_matches_dense_solve(periodic, cplx, n):
    l = RNG.uniform(-1, 1, n)
    u = RNG.uniform(-1, 1, n)
    c = 3.0 + RNG.uniform(0, 1, n)
    nrhs = 4
    q = RNG.standard_normal((nrhs, n))
    if cplx:
        q = q + 1j * RNG.standard_normal((nrhs, n))
    x = OL.tdm_solve(l, c, u, q, periodic=periodic)
    A = _dense(l, c, u, periodic)
    for r in range(nrhs):
        ref = np.linalg.solve(A, q[r])
        assert np.max(np.abs(x[r] - ref)) < 1e-13 * np.max(np.abs(ref))


def test_tdm_singular_neumann_gauge():
    """Neumann Laplacian (zero row sums) with a consistent right-hand side:
    the last pivot vanishes and the guard fixes x[n-1] = 0 (tdm.c:38-44)."""
    n = 8
    l = np.ones(n)
    u = np.ones(n)
    c = -l - u
    c[0] += l[0]
    c[-1] += u[-1]
    b = RNG.standard_normal(n)
    b -= b.mean()
    x = OL.tdm_solve(l, c, u, b)
    assert x[-1] == 0.0
    A = _dense(l, c, u, False)
    assert np.max(np.abs(A @ x - b)) < 1e-12


def test_tdm_periodic_singular_gauge():
    """Periodic Laplacian: den < DBL_EPSILON sets x[n-1] = 0 (tdm.c:94)."""
    n = 8
    l = np.ones(n)
    u = np.ones(n)
    c = -2.0 * np.ones(n)
    b = RNG.standard_normal(n)
    b -= b.mean()
    x = OL.tdm_solve(l, c, u, b, periodic=True)
    A = _dense(l, c, u, True)
    assert np.max(np.abs(A @ x - b)) < 1e-10


# --- FFTW conventions used by compute_potential/{dct,dft}.c ------------------

@pytest.mark.parametrize("n", [8, 12, 30, 64, 96, 192, 256])
def test_fft_matches_numpy(n):
    z = RNG.standard_normal(n) + 1j * RNG.standard_normal(n)
    assert np.max(np.abs(OL.fft(z, -1) - np.fft.fft(z))) < 1e-13 * n
    assert np.max(np.abs(OL.fft(z, +1) - n * np.fft.ifft(z))) < 1e-13 * n


@pytest.mark.parametrize("n", [8, 12, 64, 96, 256])
def test_real_transforms_match_fftw_definition(n):
    x = RNG.standard_normal(n)
    assert np.max(np.abs(OL.rfft(x) - np.fft.rfft(x))) < 1e-13 * n
    z = np.fft.rfft(x)
    # c2r is unnormalised and ignores the imaginary parts of the DC / Nyquist bins
    z_bad = z.copy()
    z_bad[0] += 0.7j
    z_bad[-1] += 0.3j
    assert np.max(np.abs(OL.irfft(z_bad, n) - n * x)) < 1e-12 * n


@pytest.mark.parametrize("n", [4, 16, 128])
def test_redft_pair_matches_scipy(n):
    x = RNG.standard_normal(n)
    assert np.max(np.abs(OL.redft10(x) - scipy.fft.dct(x, type=2))) < 1e-12 * n
    assert np.max(np.abs(OL.redft01(x) - scipy.fft.dct(x, type=3))) < 1e-12 * n
    # REDFT01(REDFT10(x)) = 2n x
    assert np.max(np.abs(OL.redft01(OL.redft10(x)) - 2 * n * x)) < 1e-11 * n


# --- THINC intercept / indicator (curvature_tensor.c:50-92, update/main.c) ---

def _vof_of(d, n, ndims):
    gp = np.array([-0.2886751345948129, 0.2886751345948129])
    s = 0.0
    if ndims == 2:
        for a in gp:
            for b in gp:
                s += 0.25 / (1.0 + np.exp(-2.0 * (n[0] * a + n[1] * b + d)))
    else:
        for a in gp:
            for b in gp:
                for cc in gp:
                    s += 0.125 / (1.0 + np.exp(-2.0 * (n[0] * a + n[1] * b + n[2] * cc + d)))
    return s


@pytest.mark.parametrize("ndims", [2, 3])
def test_intercept_solves_the_quadrature_equation(ndims):
    for _ in range(50):
        n = RNG.standard_normal(ndims)
        n /= np.linalg.norm(n)
        vof = RNG.uniform(1e-6, 1 - 1e-6)
        d = OL.intercept(vof, n, ndims)
        # independent root of sum_g w_g H(n . x_g + d) = vof
        ref = scipy.optimize.brentq(lambda dd: _vof_of(dd, n, ndims) - vof, -50, 50, xtol=1e-15)
        assert abs(d - ref) < 1e-9 * max(1.0, abs(ref))
        assert abs(_vof_of(d, n, ndims) - vof) < 1e-11


def test_intercept_symmetries():
    for _ in range(20):
        n = RNG.standard_normal(2)
        n /= np.linalg.norm(n)
        assert abs(OL.intercept(0.5, n)) < 1e-14
        v = RNG.uniform(0.01, 0.99)
        assert abs(OL.intercept(v, n) + OL.intercept(1 - v, n)) < 1e-10


def test_indicator_closed_form():
    n = np.array([0.6, 0.8, 0.3])
    x = np.array([0.1, -0.2])
    assert abs(OL.indicator(n, x) - 1 / (1 + np.exp(-2 * (0.06 - 0.16 + 0.3)))) < 1e-15


# --- grids and cases ----------------------------------------------------------

def _case(nx, ny, fields=None, lx=1.0, ly=1.0, uniformx=True, params=None):
    xf, xc = make_xgrid(lx, nx, uniformx)
    f = {"ux": np.zeros((ny, nx + 1)), "uy": np.zeros((ny, nx + 2)), "p": np.zeros((ny, nx + 2)),
         "t": np.zeros((ny, nx + 2)), "vof": np.zeros((ny, nx + 2))}
    if fields:
        f.update(fields)
    prm = dict(Ra=1e4, Pr=1.0, We=10.0, coef_dt_adv=0.4, coef_dt_dif=0.95, add_buoyancy=0,
               t_xm=0.0, t_xp=0.0)
    if params:
        prm.update(params)
    return InitialCondition((lx, ly), (nx, ny), xf, xc, f, prm)


def _field2d(o, name, nx, ny):
    shapes = {"ux": (ny + 2, nx + 1), "uy": (ny + 2, nx + 2), "p": (ny + 2, nx + 2),
              "psi": (ny + 2, nx + 2), "t": (ny + 2, nx + 2), "curv": (ny + 2, nx + 2),
              "vof": (ny + 4, nx + 2)}
    return o.get(name).reshape(shapes[name])


# --- interface geometry ---------------------------------------------------------

def test_curvature_of_a_circle():
    """kappa = -div(n) of the corner normals ~ 1/R at the interface (the band
    mean; this corner-normal scheme is noisy cell by cell, so a missing
    factor 1/2 or a sign slip is what the check catches)."""
    nx = ny = 128
    R = 0.25
    xf, _ = make_xgrid(1.0, nx)
    vof = _pad_cells(vof_discs(xf, ny, 1.0, [(0.5, 0.5, R)]))
    o = Oracle(_case(nx, ny, {"vof": vof}))
    o.curvature_tensor()
    curv = _field2d(o, "curv", nx, ny)[1:-1, 1:-1]
    v = vof[:, 1:-1]
    band = (v > 0.2) & (v < 0.8)
    k = curv[band]
    assert band.sum() > 50
    assert abs(np.mean(k) - 1 / R) < 0.06 / R


def test_flat_interface_normal_and_curvature():
    """A y-uniform interface: the cell normal is exactly (-1, 0), the
    curvature is y-uniform and zero away from the interface; as the corner
    normals drop to zero in single-phase regions (0 * 1 / fmax(|g|, eps)),
    the two cells beside a jump carry -1/2 d(n_x)/dx = +-1/dx."""
    nx, ny = 16, 8
    dx = 1.0 / nx
    v = np.zeros((ny, nx + 2))
    v[:, 1:9] = 1.0
    v[:, 9] = 0.25     # exactly representable: the corner sums cancel exactly
    o = Oracle(_case(nx, ny, {"vof": v}))
    o.curvature_tensor()
    curv = _field2d(o, "curv", nx, ny)
    assert np.all(curv[1:-1] == curv[1])
    row = curv[1, 1:-1]
    nz = np.nonzero(row)[0]
    # the wall (vof = 0 against 1 in cell 1) and both sides of the interface
    assert list(nz) == [0, 7, 9]
    assert row[0] == 1.0 / dx and row[7] == 1.0 / dx and row[9] == -1.0 / dx
    nrm = o.get("normal").reshape(ny + 2, nx, 3)
    # cell i = 9 (0-based column 8 of the normal array): n = (-1, 0)
    assert np.all(nrm[1:-1, 8, 0] == -1.0)
    assert np.all(nrm[1:-1, 8, 1] == 0.0)


# --- VOF transport invariants ---------------------------------------------------

@pytest.mark.parametrize("uniformx", [True, False])
def test_vof_volume_conserved_and_zero_velocity_is_identity(uniformx):
    ic = random_fields((32, 24), seed=3, uniformx=uniformx)
    o = Oracle(ic)
    _, _, vol0 = o.vof_stats()
    for _ in range(3):
        o.integrate()
    _, _, vol1 = o.vof_stats()
    assert abs(vol1 - vol0) < 1e-14 * vol0
    ic2 = random_fields((32, 24), seed=3, uniformx=uniformx)
    ic2.fields["ux"][:] = 0.0
    ic2.fields["uy"][:] = 0.0
    o2 = Oracle(ic2)
    v0 = o2.get("vof")
    for rk in range(3):
        o2.curvature_tensor()
        o2.update_vof(rk, 1e-3)
    assert np.array_equal(o2.get("vof"), v0)


def test_vof_mirror_symmetry():
    """A droplet centred in x, with a velocity field mirror-antisymmetric in
    x (ux(lx - x) = -ux(x), uy(lx - x) = uy(x)), stays mirror symmetric."""
    nx, ny = 32, 32
    xf, _ = make_xgrid(1.0, nx)
    vof = _pad_cells(vof_discs(xf, ny, 1.0, [(0.5, 0.4, 0.2)]))
    rng = np.random.default_rng(7)
    uy = np.zeros((ny, nx + 2))
    half = rng.uniform(-0.1, 0.1, size=(ny, nx // 2))
    uy[:, 1:nx // 2 + 1] = half
    uy[:, nx // 2 + 1:nx + 1] = half[:, ::-1]
    ux = np.zeros((ny, nx + 1))
    hx = rng.uniform(-0.1, 0.1, size=(ny, nx // 2 - 1))
    ux[:, 1:nx // 2] = hx
    ux[:, nx // 2 + 1:nx] = -hx[:, ::-1]
    o = Oracle(_case(nx, ny, {"vof": vof, "ux": ux, "uy": uy}))
    for rk in range(3):
        o.curvature_tensor()
        o.update_vof(rk, 2e-3)
    v = _field2d(o, "vof", nx, ny)[2:-2, 1:-1]
    assert np.max(np.abs(v - v[:, ::-1])) < 1e-13


# --- Poisson: discrete eigenfunctions and dense solves ---------------------------

def _poisson_case(nx, ny, uniformx, seed):
    """A velocity field whose divergence is a chosen right-hand side."""
    rng = np.random.default_rng(seed)
    ic = _case(nx, ny, uniformx=uniformx, ly=2.0)
    ic.fields["ux"][:, 1:-1] = rng.uniform(-1, 1, (ny, nx - 1))
    ic.fields["uy"][:, 1:-1] = rng.uniform(-1, 1, (ny, nx))
    return ic


def _dense_laplacian(xf, ny, ly):
    """Neumann-in-x (walls), periodic-in-y cell-centred Laplacian as the
    reference discretises it (dft.c:160-190 in x, 1/dy^2 stencil in y)."""
    nx = len(xf) - 1
    xc = np.concatenate([[0.0], 0.5 * (xf[1:] + xf[:-1]), [xf[-1]]])
    dxf = np.diff(xf)
    dxc = np.diff(xc)
    dy = ly / ny
    N = nx * ny
    A = np.zeros((N, N))
    idx = lambda i, j: (j % ny) * nx + i  # noqa: E731
    for j in range(ny):
        for i in range(nx):
            r = idx(i, j)
            if i > 0:
                w = 1.0 / dxc[i] / dxf[i]
                A[r, idx(i - 1, j)] += w
                A[r, r] -= w
            if i < nx - 1:
                w = 1.0 / dxc[i + 1] / dxf[i]
                A[r, idx(i + 1, j)] += w
                A[r, r] -= w
            A[r, idx(i, j - 1)] += 1.0 / dy ** 2
            A[r, idx(i, j + 1)] += 1.0 / dy ** 2
            A[r, r] -= 2.0 / dy ** 2
    return A


@pytest.mark.parametrize("uniformx,use_dct", [(True, True), (True, False), (False, False)])
def test_poisson_matches_dense_solve(uniformx, use_dct):
    nx, ny = 8, 8
    ic = _poisson_case(nx, ny, uniformx, 5)
    o = Oracle(ic)
    dt = 0.1
    rk = 1
    o.potential(rk, dt, use_dct=use_dct)
    psi = _field2d(o, "psi", nx, ny)[1:-1, 1:-1]
    # right-hand side: div(u*) / (gamma dt)
    ux = _field2d(o, "ux", nx, ny)[1:-1]
    uy = _field2d(o, "uy", nx, ny)
    dxf = np.diff(ic.xf)
    dy = 2.0 / ny
    div = (ux[:, 1:] - ux[:, :-1]) / dxf[None, :] + (uy[2:, 1:-1] - uy[1:-1, 1:-1]) / dy
    gamma = (25.0 - 17.0) / 60.0
    rhs = (div / (gamma * dt)).ravel()
    A = _dense_laplacian(ic.xf, ny, 2.0)
    # the divergence already satisfies the dxf-weighted compatibility condition
    ref = np.linalg.lstsq(A, rhs, rcond=None)[0].reshape(ny, nx)
    w = np.broadcast_to(dxf, psi.shape)
    a = psi - np.sum(psi * w) / np.sum(w)
    b = ref - np.sum(ref * w) / np.sum(w)
    assert np.max(np.abs(a - b)) < 1e-10 * np.max(np.abs(b))


def test_poisson_dct_and_dft_paths_agree_modulo_constant():
    nx, ny = 32, 48
    ic = _poisson_case(nx, ny, True, 9)
    o1, o2 = Oracle(ic), Oracle(ic)
    o1.potential(0, 0.05, use_dct=True)
    o2.potential(0, 0.05, use_dct=False)
    a = _field2d(o1, "psi", nx, ny)[1:-1, 1:-1]
    b = _field2d(o2, "psi", nx, ny)[1:-1, 1:-1]
    assert np.max(np.abs((a - a.mean()) - (b - b.mean()))) < 1e-11 * np.max(np.abs(b))


def test_projection_makes_velocity_divergence_free():
    ic = random_fields((32, 40), seed=21)
    o = Oracle(ic)
    dt = o.decide_dt()
    o.potential(0, dt)
    o.correct(0, dt)
    umax = max(np.max(np.abs(o.get("ux"))), np.max(np.abs(o.get("uy"))))
    assert o.max_divergence() < 1e-12 * umax * 32


# --- predictor operators -----------------------------------------------------------

def test_advection_conserves_kinetic_energy():
    """With zero diffusivity, pressure and surface force, the explicit sources
    of ux, uy are pure advection; for a discretely divergence-free velocity the
    skew-symmetric form gives sum u . adv(u) dV = 0 (predict/ux.c:71-122)."""
    from simplevofsolver_amd.ic import velocity_from_streamfunction_2d
    nx, ny = 24, 32
    xf, _ = make_xgrid(1.0, nx, False)
    fn = lambda X, Y: np.sin(np.pi * X) ** 2 * np.cos(2 * np.pi * Y + 0.3) + 0.3 * np.sin(2 * np.pi * X) * np.sin(4 * np.pi * Y)  # noqa: E731
    ux, uy = velocity_from_streamfunction_2d(xf, ny, 1.0, fn)
    ic = _case(nx, ny, {"ux": ux, "uy": uy}, uniformx=False, params=dict(Ra=np.inf))
    o = Oracle(ic)
    o.predict(0, 1e-3)
    sa = o.get("srcux_a").reshape(ny, nx - 1)
    sb = o.get("srcuy_a").reshape(ny, nx)
    dxf = np.diff(xf)
    xc = np.concatenate([[0.0], 0.5 * (xf[1:] + xf[:-1]), [1.0]])
    dxc = np.diff(xc)
    dy = 1.0 / ny
    e = np.sum(ux[:, 1:-1] * sa * dxc[None, 1:-1] * dy) + np.sum(uy[:, 1:-1] * sb * dxf[None, :] * dy)
    scale = np.sum(np.abs(ux[:, 1:-1] * sa * dxc[None, 1:-1] * dy)) + np.sum(np.abs(uy[:, 1:-1] * sb * dxf[None, :] * dy))
    assert abs(e) < 1e-13 * scale


def test_diffusion_operators_are_exact_on_quadratics_and_modes():
    """srcg of T holds t_dif * Lx T (implicit x); srca adds t_dif * Lyy T."""
    nx, ny = 16, 16
    xf, xc = make_xgrid(1.0, nx)
    X = xc[None, :]
    t = np.broadcast_to(X ** 2, (ny, nx + 2)).copy()
    o = Oracle(_case(nx, ny, {"t": t}, params=dict(Ra=1.0, Pr=1.0, t_xm=0.0, t_xp=1.0)))
    o.predict(0, 1e-3)
    sg = o.get("srct_g").reshape(ny, nx)
    assert np.max(np.abs(sg[:, 1:-1] - 2.0)) < 1e-10
    # y mode: cos(2 pi m y / ly) is an exact eigenvector of the y stencil
    ly, m = 1.0, 3
    dy = ly / ny
    yc = (np.arange(ny) + 0.5) * dy
    t2 = np.broadcast_to(np.cos(2 * np.pi * m * yc)[:, None], (ny, nx + 2)).copy()
    o2 = Oracle(_case(nx, ny, {"t": t2}, params=dict(Ra=1.0, Pr=1.0)))
    o2.predict(0, 1e-3)
    sa = o2.get("srct_a").reshape(ny, nx)
    lam = -4.0 / dy ** 2 * np.sin(np.pi * m * dy / ly) ** 2
    assert np.max(np.abs(sa - lam * np.cos(2 * np.pi * m * yc)[:, None])) < 1e-9 * abs(lam)


def test_rk3_amplification_of_an_advected_mode():
    """Uniform uy = V (walls uy = V), ux = 0, zero diffusivity: every field
    except T is steady, and each Fourier mode of T(y) evolves through one
    step by the RK3 stability polynomial 1 + z + z^2/2 + z^3/6 with
    z = -i V sin(k dy) dt / dy (central advection, t.c:95-118).  Pins the
    RK coefficients, the source carry and the scalar advection operator."""
    nx, ny = 8, 64
    V = 0.7
    dy = 1.0 / ny
    yc = (np.arange(ny) + 0.5) * dy
    m = 5
    k = 2 * np.pi * m
    t = np.broadcast_to(np.cos(k * yc)[:, None], (ny, nx + 2)).copy()
    uy = np.full((ny, nx + 2), V)
    ic = _case(nx, ny, {"t": t, "uy": uy}, params=dict(Ra=np.inf, uy_xm=V, uy_xp=V))
    o = Oracle(ic)
    dt = 0.3 * dy / V
    for rk in range(3):
        o.curvature_tensor()
        o.force()
        o.update_vof(rk, dt)
        o.predict(rk, dt)
        o.potential(rk, dt)
        o.correct(rk, dt)
        o.update_pressure(rk, dt)
    z = -1j * V * np.sin(k * dy) / dy * dt
    amp = 1 + z + z ** 2 / 2 + z ** 3 / 6
    expect = np.real(amp * np.exp(1j * k * yc))
    got = _field2d(o, "t", nx, ny)[1:-1, 1:-1]
    assert np.max(np.abs(got - expect[:, None])) < 1e-13


# --- time step --------------------------------------------------------------------

def test_decide_dt_closed_form():
    nx, ny = 16, 32
    ic = _case(nx, ny, ly=2.0)
    ic.fields["ux"][3, 5] = 2.0
    ic.fields["uy"][7, 9] = -3.0
    o = Oracle(ic)
    dx, dy = 1.0 / nx, 2.0 / ny
    adv = 0.4 * min(dx / (2.0 + 1e-8), dy / (3.0 + 1e-8))
    dif = 0.95 / 0.01 * 0.5 / 2 * dy * dy     # m_dif = t_dif = 0.01 (Ra=1e4, Pr=1)
    assert abs(o.decide_dt() - min(adv, dif)) < 1e-15


# --- 3D extrapolation reduces to the 2D reference step ----------------------------

def test_3d_z_uniform_reproduces_2d_step():
    ic2 = random_fields((16, 20), seed=31, uniformx=False)
    nz = 4
    f3 = {}
    for name, a in ic2.fields.items():
        f3[name] = np.broadcast_to(a[None], (nz,) + a.shape).copy()
    f3["uz"] = np.zeros((nz,) + ic2.fields["uy"].shape)
    ic3 = InitialCondition(ic2.lengths + (1.0,), ic2.glsizes + (nz,), ic2.xf, ic2.xc, f3, ic2.params)
    o2, o3 = Oracle(ic2), Oracle(ic3)
    # dt differs by construction (0.5 / NDIMS in decide_dt.c:92), so both
    # cases advance with the same fixed dt
    dt = 0.5 * min(o2.decide_dt(), o3.decide_dt())
    for _ in range(2):
        for o in (o2, o3):
            for rk in range(3):
                o.curvature_tensor()
                o.force()
                o.update_vof(rk, dt)
                o.predict(rk, dt)
                o.potential(rk, dt, use_dct=False)
                o.correct(rk, dt)
                o.update_pressure(rk, dt)
    nx, ny = ic2.glsizes
    for name, pad in (("ux", nx + 1), ("uy", nx + 2), ("t", nx + 2), ("vof", nx + 2)):
        a2 = o2.get(name)
        hy = 2 if name == "vof" else 1
        a3 = o3.get(name).reshape(nz + 2 * hy, ny + 2 * hy, pad)
        for kk in range(hy, nz + hy):
            assert np.max(np.abs(a3[kk].ravel() - a2)) < 1e-13 * max(1e-300, np.max(np.abs(a2))), name
    assert np.max(np.abs(o3.get("uz"))) < 1e-13


def yz_symmetric_case(nx, n, seed, uniformx):
    """A 3D case (ny = nz, ly = lz) invariant under the swap y <-> z: scalar
    fields and ux symmetric in (j, k), uz the transpose of uy."""
    base = random_fields((nx, n, n), seed=seed, uniformx=uniformx)
    swap = lambda a: a.transpose(1, 0, 2)
    f = {k: 0.5 * (base.fields[k] + swap(base.fields[k])) for k in ("ux", "p", "t")}
    # the droplets and their mirror images, with interiors and exteriors
    # exactly 1 and 0: where the vof varies only at rounding level the
    # reference's corner normals (gradient / max(|gradient|, DBL_EPSILON))
    # are normalised rounding noise, whose curvature then enters the CSF
    # force of the neighbouring interface faces, so such a field is not
    # symmetric beyond its noise
    v = np.maximum(base.fields["vof"], swap(base.fields["vof"]))
    v[v < 1e-12] = 0.0
    v[v > 1.0 - 1e-12] = 1.0
    f["vof"] = v
    f["uy"] = 0.5 * (base.fields["uy"] + swap(base.fields["uz"]))
    f["uz"] = swap(f["uy"]).copy()
    return InitialCondition(base.lengths, base.glsizes, base.xf, base.xc, f, dict(base.params))


def yz_asymmetry(get, nx, n):
    """Largest relative violation of the y <-> z symmetry by the padded
    fields that `get` returns."""
    worst = 0.0
    for a, b, pad, h in (("ux", "ux", nx + 1, 1), ("p", "p", nx + 2, 1), ("t", "t", nx + 2, 1),
                         ("vof", "vof", nx + 2, 2), ("uy", "uz", nx + 2, 1)):
        fa = get(a).reshape(n + 2 * h, n + 2 * h, pad)
        fb = get(b).reshape(n + 2 * h, n + 2 * h, pad).transpose(1, 0, 2)
        worst = max(worst, float(np.max(np.abs(fa - fb)) / max(1e-300, np.max(np.abs(fa)))))
    return worst


@pytest.mark.parametrize("uniformx", [True, False])
def test_3d_step_is_symmetric_in_y_and_z(uniformx):
    """Two full 3D steps keep a y <-> z symmetric state symmetric: every
    z-direction term of the 3D step (advection, diffusion, CSF force, THINC
    fluxes, corner normals, z transform of the Poisson solve, correction)
    equals its y counterpart, which the z-uniform test ties to the 2D step.
    The y and z terms enter sums in different orders, so the state stays
    symmetric to rounding (measured: <= 5e-12 relative over three steps)."""
    nx, n = 10, 12
    ic = yz_symmetric_case(nx, n, seed=41, uniformx=uniformx)
    assert np.max(np.abs(ic.fields["uz"])) > 0.05
    o = Oracle(ic)
    assert yz_asymmetry(o.get, nx, n) == 0.0
    for _ in range(3):
        o.integrate()
        assert yz_asymmetry(o.get, nx, n) < 1e-10


# --- diagnostics: momentum and interface measure ---------------------------------

@pytest.mark.parametrize("uniformx", [True, False])
def test_momentum_of_uniform_velocities(uniformx):
    """x momentum lives on the faces i = 2..nx (width DXC), y momentum on the
    cell centres (width DXF): a stretched grid tells the two apart."""
    nx, ny, lx, ly = 16, 8, 1.0, 2.0
    ux = np.zeros((ny, nx + 1))
    ux[:, 1:-1] = 0.7
    uy = np.full((ny, nx + 2), -0.3)
    o = Oracle(_case(nx, ny, {"ux": ux, "uy": uy}, lx=lx, ly=ly, uniformx=uniformx))
    _, xc = make_xgrid(lx, nx, uniformx)
    m = o.momentum()
    assert abs(m[0] - 0.7 * (xc[nx] - xc[1]) * ly) < 1e-14
    assert abs(m[1] - (-0.3) * lx * ly) < 1e-14


def test_interface_measure_of_planar_and_circular_interfaces():
    """Central differences telescope across a monotone jump, so a slab with
    two planar interfaces measures exactly 2 ly; a disc of radius R measures
    its perimeter 2 pi R to discretisation accuracy."""
    nx, ny, ly = 16, 8, 1.5
    v = np.zeros((ny, nx + 2))
    v[:, 3] = 0.6
    v[:, 4:9] = 1.0
    v[:, 9] = 0.3
    o = Oracle(_case(nx, ny, {"vof": v}, ly=ly))
    assert abs(o.interface_area() - 2 * ly) < 1e-13
    n, R = 128, 0.25
    xf, _ = make_xgrid(1.0, n)
    o2 = Oracle(_case(n, n, {"vof": _pad_cells(vof_discs(xf, n, 1.0, [(0.5, 0.5, R)]))}))
    assert abs(o2.interface_area() - 2 * np.pi * R) < 0.03 * 2 * np.pi * R


# --- logging/nusselt and statistics ----------------------------------------------

def _linear_t_case(nx, ny, uniformx, lx=1.0, ly=1.0, params=None, fields=None):
    xf, xc = make_xgrid(lx, nx, uniformx)
    t = np.tile(0.5 - xc / lx, (ny, 1))
    f = {"t": t}
    if fields:
        f.update(fields)
    prm = dict(t_xm=0.5, t_xp=-0.5)
    if params:
        prm.update(params)
    return _case(nx, ny, f, lx=lx, ly=ly, uniformx=uniformx, params=prm)


@pytest.mark.parametrize("uniformx", [True, False])
def test_nusselt_of_the_conductive_state(uniformx):
    """Linear temperature between the walls and no flow: every definition
    gives Nu = 1 (the wall fluxes and the dissipation both telescope, the
    doubled half-cell weights at the walls included)."""
    o = Oracle(_linear_t_case(16, 8, uniformx))
    nu = o.nusselt()
    assert np.max(np.abs(nu - 1.0)) < 1e-13, nu


@pytest.mark.parametrize("uniformx", [True, False])
def test_nusselt_of_a_linear_shear_and_a_uniform_injection(uniformx):
    """uy = c x (walls uy_xm = 0, uy_xp = c lx): the kinetic dissipation is
    m_dif c^2 lx ly exactly; a uniform ux against uniform T injects
    u0 T0 (xc[nx] - xc[1]) ly."""
    nx, ny, lx, ly, c = 16, 8, 1.0, 2.0, 0.3
    xf, xc = make_xgrid(lx, nx, uniformx)
    uy = np.tile(c * xc, (ny, 1))
    o = Oracle(_case(nx, ny, {"uy": uy}, lx=lx, ly=ly, uniformx=uniformx,
                     params=dict(uy_xm=0.0, uy_xp=c * lx)))
    md, td = o.lib.orc_m_dif(o.h), o.lib.orc_t_dif(o.h)
    nu = o.nusselt()
    assert abs(nu[2] - (md * c * c * lx * ly / (td * ly) + 1.0)) < 1e-12
    u0, t0 = 0.4, 0.25
    ux = np.zeros((ny, nx + 1))
    ux[:, 1:-1] = u0
    o2 = Oracle(_case(nx, ny, {"ux": ux, "t": np.full((ny, nx + 2), t0)}, lx=lx, ly=ly, uniformx=uniformx,
                      params=dict(t_xm=t0, t_xp=t0)))
    nu2 = o2.nusselt()
    assert abs(nu2[1] - (u0 * t0 * (xc[nx] - xc[1]) * ly / (td * ly) + 1.0)) < 1e-12
    assert abs(nu2[0]) < 1e-13 and abs(nu2[3]) < 1e-13


def test_nusselt_3d_reduces_to_2d_and_is_symmetric_in_y_and_z():
    """A z-uniform 3D field gives the 2D numbers (the reference flux gains
    the factor lz); a field varying in y only equals its y<->z transpose."""
    nx, ny = 12, 10
    ic2 = random_fields((nx, ny), seed=5)
    # *_init applies the x walls after the halo copy (halo.c, boundary/ux.c),
    # so the halo rows keep the IC's wall entries until the first step;
    # zero them so that the y and z halos agree
    ic2.fields["ux"][:, [0, -1]] = 0.0
    o2 = Oracle(ic2)
    nz = 6
    f3 = {k: np.repeat(v[None], nz, axis=0) for k, v in ic2.fields.items()}
    f3["uz"] = np.zeros_like(f3["uy"])
    ic3 = InitialCondition((ic2.lengths[0], ic2.lengths[1], 0.5), (nx, ny, nz), ic2.xf, ic2.xc, f3,
                           dict(ic2.params))
    o3 = Oracle(ic3)
    assert np.max(np.abs(o3.nusselt() - o2.nusselt())) < 1e-12 * np.max(np.abs(o2.nusselt()))
    # y-only variation vs the same variation in z
    rng = np.random.default_rng(9)
    n = 8
    prof = rng.uniform(-1, 1, size=n)
    base = random_fields((nx, n, n), seed=6)
    fy = {k: np.zeros_like(v) for k, v in base.fields.items()}
    fz = {k: np.zeros_like(v) for k, v in base.fields.items()}
    fy["t"][:] = prof[None, :, None]
    fz["t"][:] = prof[:, None, None]
    fy["ux"][:, :, 1:-1] = prof[None, :, None]
    fz["ux"][:, :, 1:-1] = prof[:, None, None]
    fy["uz"][:] = prof[None, :, None]       # uz(y) mirrors uy(z)
    fz["uy"][:] = prof[:, None, None]
    prm = dict(base.params, t_xm=0.0, t_xp=0.0)
    a = Oracle(InitialCondition(base.lengths, base.glsizes, base.xf, base.xc, fy, prm)).nusselt()
    b = Oracle(InitialCondition(base.lengths, base.glsizes, base.xf, base.xc, fz, prm)).nusselt()
    assert np.max(np.abs(a - b)) < 1e-12 * np.max(np.abs(a))


@pytest.mark.parametrize("gl", [(12, 10), (8, 6, 4)])
def test_statistics_profiles(gl):
    """The reduced profiles are the row sums of the accumulated samples on
    the face (ux1 ux2 uxt: i = 1..nx+1) and cell (i = 0..nx+1) ranges."""
    ic = random_fields(gl, seed=8)
    o = Oracle(ic)
    st = OL.OracleStats(o)
    nd = len(gl)
    samples = []
    try:
        for _ in range(3):
            st.collect()
            samples.append({n: o.get(n) for n in ["ux", "uy", "t"] + (["uz"] if nd == 3 else [])})
            o.integrate()
        assert st.num == 3
        nx, ny = gl[0], gl[1]

        def interior_rows(a):
            # padded arrays: y (and z) halos of one row
            if nd == 2:
                return a.reshape(ny + 2, -1)[1:-1]
            return a.reshape(gl[2] + 2, ny + 2, -1)[1:-1, 1:-1].reshape(-1, a.size // ((gl[2] + 2) * (ny + 2)))

        exp = {k: 0.0 for k in ["ux1", "ux2", "uy1", "uy2", "t1", "t2", "uxt"]}
        for sm in samples:
            ux = interior_rows(sm["ux"])          # (rows, nx+1): faces 1..nx+1
            uy = interior_rows(sm["uy"])          # (rows, nx+2): cells 0..nx+1
            t = interior_rows(sm["t"])
            tf = 0.5 * t[:, :-1] + 0.5 * t[:, 1:]
            for k, v in (("ux1", ux), ("ux2", ux * ux), ("uy1", uy), ("uy2", uy * uy), ("t1", t),
                         ("t2", t * t), ("uxt", ux * tf)):
                exp[k] = exp[k] + v
        for k, v in exp.items():
            assert np.allclose(st.reduce(k), v.sum(axis=0), rtol=1e-13, atol=1e-13), k
        assert st.reduce("ux1").shape == (nx + 1,) and st.reduce("t1").shape == (nx + 2,)
    finally:
        st.close()


# --- THINC face fluxes: upwind cell and face coordinate ----------------------------

_XG = 1.0 / (2.0 * np.sqrt(3.0))   # Gauss points +-1/(2 sqrt 3), weights 1/2 (interface/common.c:4-11)


def _thinc(s):
    """The THINC indicator with beta = 1 (interface/common.c:13, update/main.c:16-26)."""
    return 1.0 / (1.0 + np.exp(-2.0 * s))


def _intercept_normal_to_axis(v):
    """Closed-form intercept of a cell whose normal is (+-1, 0): the
    quadrature 1/2 [H(g + d) + H(-g + d)] = v becomes a quadratic in
    D = exp(-2 d): v D^2 + (v - 1/2)(a + 1/a) D + (v - 1) = 0, a = exp(-2 g)."""
    a = np.exp(-2.0 * _XG)
    B = (v - 0.5) * (a + 1.0 / a)
    D = (-B + np.sqrt(B * B - 4.0 * v * (v - 1.0))) / (2.0 * v)
    return -0.5 * np.log(D)


_PROFILE = np.array([1, 1, 1, 1, 1, 0.9, 0.7, 0.5, 0.3, 0.1, 1e-9, 0, 0, 0, 0, 0.0])


@pytest.mark.parametrize("mirror", [False, True], ids=["n=-x", "n=+x"])
@pytest.mark.parametrize("U", [0.3, -0.3])
def test_thinc_flux_x_upwind_cell_and_face_coordinate(mirror, U):
    """A y-uniform interface (cell normals exactly (+-1, 0)) in a uniform
    ux = U: the face flux between cells i-1 and i (update/flxx.c:22-41) is
    U H(n_x x + d) of the UPWIND cell (i-1 for U > 0, i for U < 0) at its face
    coordinate x = +1/2 (U > 0) or -1/2 (U < 0), d by the closed-form
    intercept, or U vof in single-phase cells.  Mutations caught: the upwind
    cell swapped, the +-1/2 coordinate flipped (each shown to change the
    flux by more than 1e-3 U below), a wrong sign of the normal, a wrong
    Gauss point or weight in the intercept."""
    nx, ny = 16, 6
    prof = _PROFILE[::-1] if mirror else _PROFILE
    vof = np.zeros((ny, nx + 2))
    vof[:, 1:-1] = prof
    ux = np.zeros((ny, nx + 1))
    ux[:, 1:-1] = U
    o = Oracle(_case(nx, ny, {"vof": vof, "ux": ux}))
    o.curvature_tensor()
    o.update_vof(0, 1e-3)
    flux = o.get("flxx").reshape(ny, nx + 1)   # FLXX(i, j) at [j - 1, i - 1]
    assert np.all(flux == flux[0])
    nsign = 1.0 if mirror else -1.0

    def expect(i, swap=False, flip=False):
        up = (i if U < 0 else i - 1) if not swap else (i - 1 if U < 0 else i)
        x = (-0.5 if U < 0 else 0.5) * (-1.0 if flip else 1.0)
        v = prof[up - 1]
        if v < 1e-8 or v > 1 - 1e-8:
            return U * v
        return U * _thinc(nsign * x + _intercept_normal_to_axis(v))

    interface_faces = 0
    for i in range(2, nx + 1):
        assert abs(flux[0, i - 1] - expect(i)) <= 1e-14, i
        if 1e-8 < prof[i - 2] < 1 - 1e-8 and 1e-8 < prof[i - 1] < 1 - 1e-8:
            interface_faces += 1
            assert abs(flux[0, i - 1] - expect(i, swap=True)) > 1e-3 * abs(U)
            assert abs(flux[0, i - 1] - expect(i, flip=True)) > 1e-3 * abs(U)
    assert interface_faces >= 4


@pytest.mark.parametrize("V", [0.3, -0.3])
def test_thinc_flux_y_upwind_cell_and_face_coordinate(V):
    """The same pin for the y faces (update/flxy.c:22-41): an x-uniform
    interface (normals (0, -1)) in a uniform uy = V; the wrap face j = ny+1
    equals face 1 (periodic)."""
    nx, ny = 6, 16
    prof = _PROFILE
    vof = np.zeros((ny, nx + 2))
    vof[:, 1:-1] = prof[:, None]
    uy = np.full((ny, nx + 2), V)
    o = Oracle(_case(nx, ny, {"vof": vof, "uy": uy}, params=dict(uy_xm=V, uy_xp=V)))
    o.curvature_tensor()
    o.update_vof(0, 1e-3)
    # FLXY(i, j) at [j - 1, i - 1]; cells 1 and nx see the x walls (vof = 0
    # there) in their corner normals, so the columns i = 2 .. nx-1 are checked
    flux = o.get("flxy").reshape(ny + 1, nx)[:, 1:-1]
    assert np.all(flux == flux[:, :1])

    def expect(j, swap=False, flip=False):
        up = (j if V < 0 else j - 1) if not swap else (j - 1 if V < 0 else j)
        y = (-0.5 if V < 0 else 0.5) * (-1.0 if flip else 1.0)
        v = prof[(up - 1) % ny]
        if v < 1e-8 or v > 1 - 1e-8:
            return V * v
        return V * _thinc(-y + _intercept_normal_to_axis(v))

    for j in range(1, ny + 2):
        assert abs(flux[j - 1, 0] - expect(j)) <= 1e-14, j
        jj = (j - 1) % ny
        if 1e-8 < prof[jj - 1] < 1 - 1e-8 and 1e-8 < prof[jj] < 1 - 1e-8:
            assert abs(flux[j - 1, 0] - expect(j, swap=True)) > 1e-3 * abs(V)
            assert abs(flux[j - 1, 0] - expect(j, flip=True)) > 1e-3 * abs(V)
    assert flux[0, 0] == flux[ny, 0]


# --- pressure update: the implicit-x correction -------------------------------------

def test_pressure_update_of_a_discrete_eigenmode():
    """psi = cos(pi k x / lx) cos(2 pi m y / ly) at cell centres of a uniform
    grid is an eigenfunction of the Neumann x Laplacian with
    lambda_x = -4 / dx^2 sin^2(pi k / (2 nx)).  fluid_update_pressure
    (update_pressure.c:87-109; x implicit, y explicit, param/implicit.c:3-7)
    must add psi (1 - 1/2 gamma dt nu lambda_x) with gamma = alpha + beta of
    the RK stage (32/60, 8/60, 20/60).  Mutations caught: the implicit term
    dropped or of the wrong sign (a change of 0.6-2.4 % of psi here), gamma of
    another stage, nu replaced by the thermal diffusivity (Pr = 0.5 makes them
    differ), a y term added (the mode has a y dependence)."""
    nx, ny, lx, ly = 32, 16, 1.0, 2.0
    dx = lx / nx
    dy = ly / ny
    k, m = 3, 2
    xc = (np.arange(nx) + 0.5) * dx
    yc = (np.arange(ny) + 0.5) * dy
    psi = np.cos(np.pi * k * xc / lx)[None, :] * np.cos(2 * np.pi * m * yc / ly)[:, None]
    lam = -4.0 / dx ** 2 * np.sin(np.pi * k / (2 * nx)) ** 2
    gammas = (32 / 60, 8 / 60, 20 / 60)
    dt = 0.01
    for rk in range(3):
        o = Oracle(_case(nx, ny, ly=ly, params=dict(Ra=25.0, Pr=0.5)))
        nu = np.sqrt(0.5) / np.sqrt(25.0)       # m_dif = sqrt(Pr / Ra) (fluid/init.c:92)
        assert abs(o.m_dif - nu) < 1e-15
        full = np.zeros((ny + 2, nx + 2))
        full[1:-1, 1:-1] = psi
        o.set("psi", full)
        o.boundaries("psi")
        p0 = _field2d(o, "p", nx, ny).copy()
        o.update_pressure(rk, dt)
        dp = (_field2d(o, "p", nx, ny) - p0)[1:-1, 1:-1]
        factor = 1.0 - 0.5 * gammas[rk] * dt * nu * lam
        assert abs(factor - 1.0) > 5e-3
        assert np.max(np.abs(dp - psi * factor)) < 1e-13
        # the Neumann images of p (boundary/p.c)
        pp = _field2d(o, "p", nx, ny)
        assert np.array_equal(pp[1:-1, 0], pp[1:-1, 1]) and np.array_equal(pp[1:-1, -1], pp[1:-1, -2])


# --- CSF surface tension: the Laplace pressure jump of a static droplet --------------

@pytest.mark.parametrize("uniformx", [True, False])
@pytest.mark.parametrize("R", [0.2, 0.3])
def test_static_droplet_laplace_pressure_jump(uniformx, R):
    """One full step (force, predict, project, pressure) from rest with a
    droplet of radius R in a THINC-like profile vof = 1 / (1 + exp(2 (r - R) / h)):
    the projection turns the CSF force sigma kappa grad(vof) (force.c:11-66)
    into a pressure whose inside-minus-outside jump is the Laplace value
    sigma / R (sigma = 1/We).  Measured: within 1.6 % at 64^2 (0.6-1.1 % at
    128^2); bar 3 %.  Mutations caught: the 1/2 of the face curvature average
    dropped (2x), the sign of the force or of the curvature, tension not
    applied or applied twice, the force left out of the predictor's explicit
    sources.  A sharp profile (exact area fractions) does not give this jump
    in the reference either: the corner normals drop from unit length to zero
    at the edge of the interface band, so the curvature of those cells is of
    order 1/dx and the force there dominates."""
    n = 64
    We = 10.0
    xf, _ = make_xgrid(1.0, n, uniformx)
    xc = 0.5 * (xf[:-1] + xf[1:])
    yc = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(xc, yc)
    r = np.hypot(X - 0.5, Y - 0.5)
    vof = np.zeros((n, n + 2))
    vof[:, 1:-1] = 1.0 / (1.0 + np.exp(2.0 * (r - R) * n))
    o = Oracle(_case(n, n, {"vof": vof}, uniformx=uniformx, params=dict(We=We, Ra=1e4)))
    o.integrate()
    p = _field2d(o, "p", n, n)[1:-1, 1:-1]
    jump = p[r < R - 0.1].mean() - p[r > R + 0.1].mean()
    laplace = 1.0 / We / R
    assert abs(jump / laplace - 1.0) < 0.03, (jump, laplace)
    # the pressure is uniform inside and outside, away from the interface
    assert p[r < R - 0.1].std() < 0.01 * laplace and p[r > R + 0.1].std() < 0.01 * laplace


# --- sensitivity of whole steps on 3D droplet workloads ------------------------------

def test_step_amplifies_transform_rounding_on_droplets():
    """Two oracles that differ only in the rounding of their transforms'
    twiddle tables (octant-reduced vs direct cos / sin; the same DFT, as two
    FFTW plans would be) on the 3D droplet case 64x32x32 (We = 50, velocity
    rms 0.2).  On identical input one Poisson solve agrees to rounding
    (psi within 1e-14 modulo its mean; measured 7.6e-16), yet after one full
    step the velocities differ by 1e-4 relative and the curvature by 0.2:
    where vof varies only at rounding level the corner normals are
    normalised rounding noise (gradient / max(|gradient|, DBL_EPSILON),
    curvature_tensor.c:14-48), and a 1e-16 change turns them around.  So a
    whole-step comparison of two correct implementations of this step on
    droplet workloads is bounded by this amplification, not by 1e-10."""
    from parity_util import demean, interior, rel
    from simplevofsolver_amd.ic import droplets_3d
    ic, _ = droplets_3d(64, 32, 32, lx=2.0, ndrops=2, rmin=0.15, rmax=0.25, urms=0.2, seed=5, Re=1000.0,
                        We=50.0, uniformx=True)
    gl = ic.glsizes
    dxf = np.diff(ic.xf)

    def run(variant, fn):
        OL.set_fft_variant(variant, 3)
        try:
            fn()
        finally:
            OL.set_fft_variant(0, 3)

    a, b = Oracle(ic), Oracle(ic)
    dt = a.decide_dt()
    run(0, lambda: a.potential(0, dt))
    run(1, lambda: b.potential(0, dt))
    e_psi = rel(interior("psi", demean("psi", b.get("psi"), gl, dxf), gl),
                interior("psi", demean("psi", a.get("psi"), gl, dxf), gl))
    assert 0.0 < e_psi < 1e-14, e_psi
    a, b = Oracle(ic), Oracle(ic)
    run(0, a.integrate)
    run(1, b.integrate)
    e_u = max(rel(b.get(n), a.get(n)) for n in ("ux", "uy", "uz"))
    assert e_u > 1e-8, e_u
    assert rel(b.get("curv"), a.get("curv")) > 1e-3
