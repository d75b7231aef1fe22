"""Helpers shared by the parity tests: paired GPU / oracle cases and the
comparisons the parity bar uses (relative max-norm; pressure and potential
modulo their volume-weighted mean, because the reference's singular k = 0
Poisson mode leaves an additive constant set by roundoff,
src/tdm.c:94 and :38-44)."""
from __future__ import annotations

import numpy as np

from oracle_lib import Oracle


def padded_shape(name, glsizes):
    nx = glsizes[0]
    ny = glsizes[1]
    nd = len(glsizes)
    nz = glsizes[2] if nd == 3 else None
    pz1 = (nz + 2) if nd == 3 else None
    table = {
        "ux": (nx + 1, ny + 2, pz1), "uy": (nx + 2, ny + 2, pz1), "uz": (nx + 2, ny + 2, pz1),
        "p": (nx + 2, ny + 2, pz1), "psi": (nx + 2, ny + 2, pz1), "t": (nx + 2, ny + 2, pz1),
        "curv": (nx + 2, ny + 2, pz1), "vof": (nx + 2, ny + 4, (nz + 4) if nd == 3 else None),
    }
    n0, n1, n2 = table[name]
    return (n2, n1, n0) if nd == 3 else (n1, n0)


def interior(name, arr, glsizes):
    """Interior cells of a padded cell-centred array (x = 1..nx, y, z)."""
    a = np.asarray(arr).reshape(padded_shape(name, glsizes))
    h = 2 if name == "vof" else 1
    if len(glsizes) == 3:
        return a[h:-h, h:-h, 1:-1]
    return a[h:-h, 1:-1]


def read_mask(name, glsizes):
    """Entries a later stage can read: everything except the x-wall entries
    of halo rows / planes, which keep stale values (the halo copy runs before
    the wall update, src/halo.c then fluid/boundary/*.c) and are never read."""
    shp = padded_shape(name, glsizes)
    m = np.ones(shp, dtype=bool)
    h = 2 if name == "vof" else 1
    if len(glsizes) == 3:
        halo = np.zeros(shp[:2], dtype=bool)
        halo[:h, :] = True
        halo[-h:, :] = True
        halo[:, :h] = True
        halo[:, -h:] = True
        m[halo, 0] = False
        m[halo, -1] = False
    else:
        m[:h, 0] = m[:h, -1] = False
        m[-h:, 0] = m[-h:, -1] = False
    return m


def demean(name, arr, glsizes, dxf):
    """Subtract the volume-weighted interior mean (x spacing dxf; y, z uniform)."""
    a = np.asarray(arr, dtype=np.float64).reshape(padded_shape(name, glsizes)).copy()
    inner = interior(name, a, glsizes)
    w = np.broadcast_to(dxf, inner.shape)
    mean = float(np.sum(inner * w) / np.sum(w))
    return a - mean


def rel(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(float(np.max(np.abs(b))), 1e-300)
    return float(np.max(np.abs(a - b)) / scale)


def make_pair(ic, tmpdir, schedule=None):
    import simplevofsolver_amd as sva
    o = Oracle(ic)
    g = sva.solver_from_ic(ic, str(tmpdir), schedule=schedule)
    return o, g


def state_names(ndims):
    names = ["ux", "uy", "p", "psi", "t", "vof", "ifrcx", "ifrcy", "dvof", "normal", "curv",
             "flxx", "flxy", "src_a", "src_b"]
    for f in ("srcux", "srcuy", "srct"):
        names += [f + "_a", f + "_b", f + "_g"]
    if ndims == 3:
        names += ["uz", "ifrcz", "flxz", "srcuz_a", "srcuz_b", "srcuz_g"]
    return names


def copy_state(o, g):
    """Oracle state -> GPU arrays (isolates one entry point per comparison)."""
    for n in state_names(o.ndims):
        g.set(n, o.get(n))
