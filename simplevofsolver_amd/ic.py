"""Seeded initial-condition generator writing the reference's NPY layout.

The layout is the one ``initial_condition/main.py:6-73`` of the reference
writes and ``src/domain.c:21-44`` / ``src/array.c:80-144`` read back:

* ``step.npy`` (<u8, shape ()), ``time.npy`` (<f8, shape ())
* ``glsizes.npy`` (<u8, (NDIMS,)), ``lengths.npy`` (<f8, (NDIMS,))
* ``xf.npy`` (nx+1,), ``xc.npy`` (nx+2,) with xc[0] = 0 and xc[nx+1] = lx
* ``ux`` (..., ny, nx+1); ``uy``, ``uz``, ``p``, ``t``, ``vof`` (..., ny, nx+2),
  3D prepending nz; C order.

Fields are FP64.  Wall entries are overwritten by the boundary update on load
(``src/fluid/init.c:85-88``), so they are written as zeros.

The droplet volume fractions are exact cell-area fractions of discs (closed
form), and for spheres an 8-point Gauss-Legendre integral in z of the exact
disc-rectangle area.  Velocity fields for the 3D and differentiation
workloads come from a discrete vector potential (stream function in 2D), so
they are discretely divergence free and vanish on the x walls.
"""
from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np


# ---------------------------------------------------------------------------
# grids

def make_xgrid(lx: float, nx: int, uniform: bool = True):
    """Face and centre coordinates in x (initial_condition/main.py:15-43)."""
    if uniform:
        xf = np.linspace(0.0, lx, nx + 1, endpoint=True)
    else:
        nclip = 3
        xf = np.arange(0, nx + 1, 1)
        xf = np.cos(np.pi * (xf + 1.0 * nclip) / (nx + 2.0 * nclip))
        xf *= -1.0
        xf = lx * (xf - np.min(xf)) / (np.max(xf) - np.min(xf))
    xc = np.concatenate([[0.0], 0.5 * xf[:-1] + 0.5 * xf[1:], [lx]])
    return np.asarray(xf, dtype=np.float64), np.asarray(xc, dtype=np.float64)


# ---------------------------------------------------------------------------
# exact disc / rectangle intersection areas

def _prim(x, r):
    """int_{-r}^{x} sqrt(r^2 - X^2) dX for x in [-r, r]."""
    x = np.clip(x, -r, r)
    h = np.sqrt(np.maximum(r * r - x * x, 0.0))
    return 0.5 * (x * h + r * r * np.arcsin(np.clip(x / r, -1.0, 1.0))) + 0.25 * np.pi * r * r


def _quad(x, y, r):
    """Area of {X <= x, Y <= y} inside the disc X^2 + Y^2 <= r^2 (centred)."""
    x = np.clip(x, -r, r)
    y = np.clip(y, -r, r)
    xs = np.sqrt(np.maximum(r * r - y * y, 0.0))

    def h_int(a, b):
        b = np.maximum(a, b)
        return _prim(b, r) - _prim(a, r)

    # y >= 0: outer strips |X| >= xs carry the full chord 2h, the middle y + h
    left = 2.0 * h_int(-r * np.ones_like(x), np.minimum(x, -xs))
    mid_hi = np.clip(x, -xs, xs)
    mid = y * np.maximum(mid_hi + xs, 0.0) + h_int(-xs, mid_hi)
    right = 2.0 * h_int(xs, np.maximum(x, xs))
    pos = left + mid + right
    # y < 0: only |X| < xs contributes, with y + h
    neg = mid
    return np.where(y >= 0.0, pos, neg)


def disc_rect_area(x0, x1, y0, y1, cx, cy, r):
    """Exact area of the disc (cx, cy, r) inside [x0, x1] x [y0, y1]."""
    a0, a1 = x0 - cx, x1 - cx
    b0, b1 = y0 - cy, y1 - cy
    area = _quad(a1, b1, r) - _quad(a0, b1, r) - _quad(a1, b0, r) + _quad(a0, b0, r)
    return np.maximum(area, 0.0)


def vof_discs(xf, ny, ly, discs, periodic_y=True):
    """Cell volume fractions (ny, nx) of non-overlapping discs, periodic in y."""
    nx = len(xf) - 1
    dy = ly / ny
    x0 = xf[:-1][None, :]
    x1 = xf[1:][None, :]
    yy = np.arange(ny) * dy
    y0 = yy[:, None]
    y1 = (yy + dy)[:, None]
    cell = (x1 - x0) * (y1 - y0)
    vof = np.zeros((ny, nx))
    for (cx, cy, r) in discs:
        shifts = (-ly, 0.0, ly) if periodic_y else (0.0,)
        for s in shifts:
            vof += disc_rect_area(x0, x1, y0, y1, cx, cy + s, r) / cell
    return np.clip(vof, 0.0, 1.0)


_GL_X, _GL_W = np.polynomial.legendre.leggauss(8)


def vof_spheres(xf, ny, ly, nz, lz, spheres):
    """Cell volume fractions (nz, ny, nx) of non-overlapping spheres, periodic
    in y and z: exact disc areas on 8 Gauss-Legendre z slices per cell."""
    nx = len(xf) - 1
    dy, dz = ly / ny, lz / nz
    x0 = xf[:-1]
    x1 = xf[1:]
    vof = np.zeros((nz, ny, nx))
    for (cx, cy, cz, r) in spheres:
        # columns that meet [cx - r, cx + r] (the others receive exactly 0:
        # both of their x bounds clip to the same side of the disc)
        i0 = max(0, int(np.searchsorted(x1, cx - r, side="left")) - 1)
        i1 = min(nx, int(np.searchsorted(x0, cx + r, side="right")) + 1)
        xa, xb = x0[i0:i1], x1[i0:i1]
        for sy in (-ly, 0.0, ly):
            for sz in (-lz, 0.0, lz):
                ccy, ccz = cy + sy, cz + sz
                ks = np.arange(max(0, int(np.floor((ccz - r) / dz)) - 1),
                               min(nz, int(np.ceil((ccz + r) / dz)) + 1))
                js = np.arange(max(0, int(np.floor((ccy - r) / dy)) - 1),
                               min(ny, int(np.ceil((ccy + r) / dy)) + 1))
                if len(ks) == 0 or len(js) == 0:
                    continue
                for k in ks:
                    za = max(k * dz, ccz - r)
                    zb = min((k + 1) * dz, ccz + r)
                    if zb <= za:
                        continue
                    zq = 0.5 * (za + zb) + 0.5 * (zb - za) * _GL_X
                    wq = 0.5 * (zb - za) * _GL_W
                    y0 = (js * dy)[:, None]
                    y1 = ((js + 1) * dy)[:, None]
                    acc = np.zeros((len(js), i1 - i0))
                    for zz, ww in zip(zq, wq):
                        rz = np.sqrt(max(r * r - (zz - ccz) ** 2, 0.0))
                        if rz <= 0.0:
                            continue
                        acc += ww * disc_rect_area(xa[None, :], xb[None, :], y0, y1, cx, ccy, rz)
                    vof[k, js, i0:i1] += acc / ((xb - xa)[None, :] * dy * dz)
    return np.clip(vof, 0.0, 1.0)


# ---------------------------------------------------------------------------
# divergence-free velocity fields

def velocity_from_streamfunction_2d(xf, ny, ly, psi_fn):
    """ux, uy on the staggered grid from a corner stream function
    psi(xf_i, y_{j-1/2}); discretely divergence free, ux = 0 at the walls
    when psi vanishes there."""
    nx = len(xf) - 1
    dy = ly / ny
    yc = (np.arange(ny + 1)) * dy  # corner rows y_{j-1/2}, j = 0..ny
    X, Y = np.meshgrid(xf, yc)     # (ny+1, nx+1)
    psi = psi_fn(X, Y)
    psi[:, 0] = 0.0
    psi[:, -1] = 0.0
    ux = np.zeros((ny, nx + 1))
    uy = np.zeros((ny, nx + 2))
    ux[:, :] = (psi[1:, :] - psi[:-1, :]) / dy
    dxf = np.diff(xf)
    uy[:, 1:-1] = -(psi[:-1, 1:] - psi[:-1, :-1]) / dxf[None, :]
    ux[:, 0] = 0.0
    ux[:, -1] = 0.0
    return ux, uy


def velocity_from_potential_3d(xf, ny, ly, nz, lz, ay_fn, az_fn, ax_fn):
    """u = curl A on the staggered grid.  A_x on (x_c, y_{j-1/2}, z_{k-1/2}),
    A_y on (x_f, y_c, z_{k-1/2}), A_z on (x_f, y_{j-1/2}, z_c); A_y = A_z = 0
    on the x walls, so ux vanishes there.  Discretely divergence free.
    The potential functions take open (broadcastable) coordinate grids
    z[:, None, None], y[None, :, None], x[None, None, :]."""
    nx = len(xf) - 1
    dy, dz = ly / ny, lz / nz
    xc = 0.5 * (xf[:-1] + xf[1:])
    yf = np.arange(ny) * dy           # y_{j-1/2}, j = 1..ny
    zf = np.arange(nz) * dz
    yc = yf + 0.5 * dy
    zc = zf + 0.5 * dz
    o3 = lambda z, y, x: (z[:, None, None], y[None, :, None], x[None, None, :])  # noqa: E731
    Z, Y, X = o3(zc, yf, xf)
    az = np.broadcast_to(az_fn(X, Y, Z), (nz, ny, nx + 1)).copy()
    Z, Y, X = o3(zf, yc, xf)
    ay = np.broadcast_to(ay_fn(X, Y, Z), (nz, ny, nx + 1)).copy()
    az[:, :, 0] = 0.0
    az[:, :, -1] = 0.0
    ay[:, :, 0] = 0.0
    ay[:, :, -1] = 0.0
    Z, Y, X = o3(zf, yf, xc)
    ax = np.broadcast_to(ax_fn(X, Y, Z), (nz, ny, nx)).copy()
    roll = np.roll
    # ux(i_face, j, k) = d_y A_z - d_z A_y
    ux = (roll(az, -1, axis=1) - az) / dy - (roll(ay, -1, axis=0) - ay) / dz
    # uy(i, j_face, k) = d_z A_x - d_x A_z   (A_z differenced across the cell)
    dxf = np.diff(xf)
    uy = np.zeros((nz, ny, nx + 2))
    uy[:, :, 1:-1] = (roll(ax, -1, axis=0) - ax) / dz - (az[:, :, 1:] - az[:, :, :-1]) / dxf[None, None, :]
    # uz(i, j, k_face) = d_x A_y - d_y A_x
    uz = np.zeros((nz, ny, nx + 2))
    uz[:, :, 1:-1] = (ay[:, :, 1:] - ay[:, :, :-1]) / dxf[None, None, :] - (roll(ax, -1, axis=1) - ax) / dy
    ux[:, :, 0] = 0.0
    ux[:, :, -1] = 0.0
    return ux, uy, uz


def _random_modes(rng, nmodes, lx, ly, lz=None, kmax=4):
    modes = []
    for _ in range(nmodes):
        kx = int(rng.integers(1, kmax + 1))
        ky = int(rng.integers(-kmax, kmax + 1))
        kz = int(rng.integers(-kmax, kmax + 1)) if lz is not None else 0
        amp = rng.standard_normal(3)
        phase = rng.uniform(0.0, 2.0 * np.pi, size=3)
        modes.append((kx, ky, kz, amp, phase))
    return modes


def _modal_fn(modes, comp, lx, ly, lz):
    def fn(X, Y, Z):
        out = 0.0
        for (kx, ky, kz, amp, phase) in modes:
            arg = 2.0 * np.pi * (ky * Y / ly + (kz * Z / lz if lz else 0.0)) + phase[comp]
            out = out + (amp[comp] * np.cos(arg)) * np.sin(np.pi * kx * X / lx)
        return out
    return fn


# ---------------------------------------------------------------------------
# writers

@dataclass
class InitialCondition:
    lengths: tuple
    glsizes: tuple
    xf: np.ndarray
    xc: np.ndarray
    fields: dict = field(default_factory=dict)
    params: dict = field(default_factory=dict)
    step: int = 0
    time: float = 0.0

    @property
    def ndims(self):
        return len(self.glsizes)


def write_ic(dest: str, ic: InitialCondition) -> str:
    """Write ``ic`` in the reference's NPY layout (directory created)."""
    os.makedirs(dest, exist_ok=True)
    np.save(os.path.join(dest, "step.npy"), np.array(ic.step, dtype=np.uint64))
    np.save(os.path.join(dest, "time.npy"), np.array(ic.time, dtype=np.float64))
    np.save(os.path.join(dest, "glsizes.npy"), np.array(ic.glsizes, dtype=np.uint64))
    np.save(os.path.join(dest, "lengths.npy"), np.array(ic.lengths, dtype=np.float64))
    np.save(os.path.join(dest, "xf.npy"), np.asarray(ic.xf, dtype=np.float64))
    np.save(os.path.join(dest, "xc.npy"), np.asarray(ic.xc, dtype=np.float64))
    for name, arr in ic.fields.items():
        np.save(os.path.join(dest, f"{name}.npy"), np.ascontiguousarray(arr, dtype=np.float64))
    return dest


def field_shapes(glsizes):
    nx = glsizes[0]
    tail = tuple(reversed(glsizes[1:]))
    shapes = {"ux": tail + (nx + 1,), "uy": tail + (nx + 2,), "p": tail + (nx + 2,),
              "t": tail + (nx + 2,), "vof": tail + (nx + 2,)}
    if len(glsizes) == 3:
        shapes["uz"] = tail + (nx + 2,)
    return shapes


def _pad_cells(v):
    """(…, nx) cell values -> (…, nx+2) with zero wall entries."""
    out = np.zeros(v.shape[:-1] + (v.shape[-1] + 2,))
    out[..., 1:-1] = v
    return out


# ---------------------------------------------------------------------------
# workloads

def droplet_2d(nx=128, ny=256, lx=1.0, ly=2.0, radius=0.25, centre=(0.5, 0.5),
               amp=1.0e-3, seed=1, uniformx=True, Re=100.0, We=10.0):
    """Plain 2D workload: one circular droplet, exact cell-area fractions,
    zero velocity plus a seeded uniform perturbation of amplitude ``amp`` on
    the interior velocity DOFs; Re -> (Ra = Re^2, Pr = 1), buoyancy off,
    wall temperature 0."""
    rng = np.random.default_rng(seed)
    xf, xc = make_xgrid(lx, nx, uniformx)
    vof = _pad_cells(vof_discs(xf, ny, ly, [(centre[0], centre[1], radius)]))
    ux = np.zeros((ny, nx + 1))
    ux[:, 1:-1] = rng.uniform(-amp, amp, size=(ny, nx - 1))
    uy = np.zeros((ny, nx + 2))
    uy[:, 1:-1] = rng.uniform(-amp, amp, size=(ny, nx))
    p = np.zeros((ny, nx + 2))
    t = np.zeros((ny, nx + 2))
    params = dict(Ra=Re * Re, Pr=1.0, We=We, coef_dt_adv=0.4, coef_dt_dif=0.95,
                  add_buoyancy=0, t_xm=0.0, t_xp=0.0)
    return InitialCondition((lx, ly), (nx, ny), xf, xc,
                            dict(ux=ux, uy=uy, p=p, t=t, vof=vof), params)


def droplets_3d(nx=256, ny=256, nz=256, lx=1.0, ly=1.0, lz=1.0, ndrops=8,
                rmin=0.06, rmax=0.12, urms=0.1, nmodes=16, seed=2, Re=500.0,
                We=20.0, uniformx=True):
    """Plain 3D workload: ``ndrops`` non-overlapping spheres with seeded radii
    in [rmin, rmax] kept 2 cells off the x walls, and a divergence-free
    velocity from ``nmodes`` random-phase Fourier modes scaled to rms
    ``urms``."""
    rng = np.random.default_rng(seed)
    xf, xc = make_xgrid(lx, nx, uniformx)
    dxm = lx / nx
    spheres = []
    tries = 0
    while len(spheres) < ndrops:
        tries += 1
        if tries > 100000:
            raise RuntimeError("could not place non-overlapping droplets")
        r = rng.uniform(rmin, rmax)
        cx = rng.uniform(r + 2 * dxm, lx - r - 2 * dxm)
        cy = rng.uniform(0, ly)
        cz = rng.uniform(0, lz)
        ok = True
        for (ox, oy, oz, orr) in spheres:
            ddy = min(abs(cy - oy), ly - abs(cy - oy))
            ddz = min(abs(cz - oz), lz - abs(cz - oz))
            if (cx - ox) ** 2 + ddy ** 2 + ddz ** 2 < (r + orr + 2 * dxm) ** 2:
                ok = False
                break
        if ok:
            spheres.append((cx, cy, cz, r))
    vof = _pad_cells(vof_spheres(xf, ny, ly, nz, lz, spheres))
    modes = _random_modes(rng, nmodes, lx, ly, lz)
    ux, uy, uz = velocity_from_potential_3d(
        xf, ny, ly, nz, lz,
        _modal_fn(modes, 1, lx, ly, lz), _modal_fn(modes, 2, lx, ly, lz),
        _modal_fn(modes, 0, lx, ly, lz))
    # rms over the interior velocity DOFs
    ms = (np.mean(ux[:, :, 1:-1] ** 2) + np.mean(uy[:, :, 1:-1] ** 2) + np.mean(uz[:, :, 1:-1] ** 2)) / 3.0
    s = urms / np.sqrt(ms) if ms > 0 else 0.0
    ux, uy, uz = ux * s, uy * s, uz * s
    shp = (nz, ny, nx + 2)
    params = dict(Ra=Re * Re, Pr=1.0, We=We, coef_dt_adv=0.4, coef_dt_dif=0.95,
                  add_buoyancy=0, t_xm=0.0, t_xp=0.0)
    return InitialCondition((lx, ly, lz), (nx, ny, nz), xf, xc,
                            dict(ux=ux, uy=uy, uz=uz, p=np.zeros(shp), t=np.zeros(shp),
                                 vof=vof), params), spheres


def droplet_2d_divfree(nx=256, ny=256, lx=1.0, ly=1.0, radius=0.2, urms=0.05,
                       nmodes=8, seed=3, Re=100.0, We=10.0):
    """Differentiation workload: one centred droplet in a divergence-free
    velocity (stream function of random modes, rms ``urms``)."""
    rng = np.random.default_rng(seed)
    xf, xc = make_xgrid(lx, nx, True)
    vof = _pad_cells(vof_discs(xf, ny, ly, [(0.5 * lx, 0.5 * ly, radius)]))
    modes = _random_modes(rng, nmodes, lx, ly, None)
    fn = _modal_fn(modes, 0, lx, ly, None)
    ux, uy = velocity_from_streamfunction_2d(xf, ny, ly, lambda X, Y: fn(X, Y, 0.0 * X))
    ms = 0.5 * (np.mean(ux[:, 1:-1] ** 2) + np.mean(uy[:, 1:-1] ** 2))
    s = urms / np.sqrt(ms) if ms > 0 else 0.0
    ux, uy = ux * s, uy * s
    shp = (ny, nx + 2)
    params = dict(Ra=Re * Re, Pr=1.0, We=We, coef_dt_adv=0.4, coef_dt_dif=0.95,
                  add_buoyancy=0, t_xm=0.0, t_xp=0.0)
    return InitialCondition((lx, ly), (nx, ny), xf, xc,
                            dict(ux=ux, uy=uy, p=np.zeros(shp), t=np.zeros(shp), vof=vof),
                            params)


def random_fields(glsizes, lengths=None, seed=0, uniformx=True, amp=0.1,
                  ndrops=2, tamp=0.5):
    """Small generic case for parity tests: random droplets, random (not
    divergence-free) velocities, random p and t."""
    rng = np.random.default_rng(seed)
    nd = len(glsizes)
    if lengths is None:
        lengths = (1.0,) * nd
    nx = glsizes[0]
    xf, xc = make_xgrid(lengths[0], nx, uniformx)
    if nd == 2:
        ny = glsizes[1]
        discs = []
        for _ in range(ndrops):
            r = rng.uniform(0.1, 0.25) * min(lengths)
            discs.append((rng.uniform(r + 0.05, lengths[0] - r - 0.05), rng.uniform(0, lengths[1]), r))
        vof = _pad_cells(vof_discs(xf, ny, lengths[1], discs))
    else:
        ny, nz = glsizes[1], glsizes[2]
        sph = []
        for _ in range(ndrops):
            r = rng.uniform(0.15, 0.3) * min(lengths)
            sph.append((rng.uniform(r + 0.05, lengths[0] - r - 0.05), rng.uniform(0, lengths[1]),
                        rng.uniform(0, lengths[2]), r))
        vof = _pad_cells(vof_spheres(xf, ny, lengths[1], nz, lengths[2], sph))
    shapes = field_shapes(glsizes)
    fields = {"vof": vof}
    for name, shp in shapes.items():
        if name == "vof":
            continue
        a = rng.uniform(-amp, amp, size=shp)
        if name == "t":
            a = rng.uniform(-tamp, tamp, size=shp)
        fields[name] = a
    params = dict(Ra=1.0e4, Pr=1.0, We=10.0, coef_dt_adv=0.4, coef_dt_dif=0.95,
                  add_buoyancy=1, t_xm=0.5, t_xp=-0.5)
    return InitialCondition(tuple(lengths), tuple(glsizes), xf, xc, fields, params)


def batch_2d_cases(ncases=96, seed=4, sizes=(64, 96, 128, 192, 256), urms=0.05):
    """Batched workload: ``ncases`` independent 2D cases with seeded uneven
    grids (nx from ``sizes``, ny = 2 nx, lx = 1, ly = 2, uniform x), each with
    1-4 non-overlapping droplets of radius 0.05-0.2 kept 2 cells off the x
    walls, a divergence-free velocity of rms ``urms``, Re in [50, 500] and We
    in [5, 50] (Ra = Re^2, Pr = 1, buoyancy off)."""
    rng = np.random.default_rng(seed)
    out = []
    for c in range(ncases):
        nx = int(rng.choice(sizes))
        ny = 2 * nx
        lx, ly = 1.0, 2.0
        xf, xc = make_xgrid(lx, nx, True)
        ndrops = int(rng.integers(1, 5))
        discs = []
        tries = 0
        while len(discs) < ndrops and tries < 10000:
            tries += 1
            r = float(rng.uniform(0.05, 0.2))
            margin = r + 2.0 * lx / nx
            cx = float(rng.uniform(margin, lx - margin))
            cy = float(rng.uniform(0.0, ly))
            ok = True
            for (ox, oy, orad) in discs:
                dy = abs(cy - oy)
                dy = min(dy, ly - dy)
                if (cx - ox) ** 2 + dy ** 2 < (r + orad + 2.0 * lx / nx) ** 2:
                    ok = False
                    break
            if ok:
                discs.append((cx, cy, r))
        vof = _pad_cells(vof_discs(xf, ny, ly, discs))
        modes = _random_modes(rng, 8, lx, ly, None)
        fn = _modal_fn(modes, 0, lx, ly, None)
        ux, uy = velocity_from_streamfunction_2d(xf, ny, ly, lambda X, Y: fn(X, Y, 0.0 * X))
        ms = 0.5 * (np.mean(ux[:, 1:-1] ** 2) + np.mean(uy[:, 1:-1] ** 2))
        sc = urms / np.sqrt(ms) if ms > 0 else 0.0
        ux, uy = ux * sc, uy * sc
        Re = float(rng.uniform(50.0, 500.0))
        We = float(rng.uniform(5.0, 50.0))
        shp = (ny, nx + 2)
        params = dict(Ra=Re * Re, Pr=1.0, We=We, coef_dt_adv=0.4, coef_dt_dif=0.95,
                      add_buoyancy=0, t_xm=0.0, t_xp=0.0)
        out.append(InitialCondition((lx, ly), (nx, ny), xf, xc,
                                    dict(ux=ux, uy=uy, p=np.zeros(shp), t=np.zeros(shp), vof=vof),
                                    params))
    return out


def divfree_direction(ic_obj, rng, nmodes=8):
    """A random discretely divergence-free 2D velocity field (stream function
    of ``nmodes`` random modes, zero normal velocity at the walls) in the
    solver's padded ux / uy layouts, flattened: a perturbation direction for
    derivative checks that keeps the flow incompressible."""
    nx, ny = ic_obj.glsizes
    lx, ly = ic_obj.lengths
    fn = _modal_fn(_random_modes(rng, nmodes, lx, ly, None), 0, lx, ly, None)
    ux, uy = velocity_from_streamfunction_2d(ic_obj.xf, ny, ly, lambda X, Y: fn(X, Y, 0.0 * X))
    px = np.zeros((ny + 2, nx + 1))
    py = np.zeros((ny + 2, nx + 2))
    px[1:-1] = ux
    py[1:-1] = uy
    return px.ravel(), py.ravel()
