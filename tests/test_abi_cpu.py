"""CPU checks of the drop-in boundary: the HIP libraries load, export every
entry point include/vof_solver.h declares, and fail loudly (no CPU fallback);
the IC generator writes the reference's NPY layout."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import simplevofsolver_amd as sva
from simplevofsolver_amd.ic import (droplet_2d, droplets_3d, field_shapes, random_fields,
                                    velocity_from_streamfunction_2d, make_xgrid, write_ic)

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols(ndims):
    src = open(os.path.join(ROOT, "include", "vof_solver.h")).read()
    # drop the other dimensionality's conditional blocks
    keep = []
    skip = False
    for line in src.splitlines():
        if line.startswith("#if NDIMS == 3"):
            skip = ndims != 3
            continue
        if line.startswith("#if NDIMS == 2"):
            skip = ndims != 2
            continue
        if line.startswith("#endif") and skip:
            skip = False
            continue
        if not skip:
            keep.append(line)
    text = "\n".join(keep)
    names = set(re.findall(r"^\s*(?:int|const char \*)\s+(\w+)\s*\(", text, re.M))
    names |= set(re.findall(r"^extern const \w+ (\w+);", text, re.M))
    return names


@pytest.mark.parametrize("ndims", [2, 3])
def test_library_exports_every_declared_symbol(ndims):
    lib = sva.load(ndims)
    declared = _declared_symbols(ndims)
    assert "integrate" in declared and "tdm" in declared and len(declared) > 35
    for name in sorted(declared):
        assert hasattr(lib, name), name
    listed = set(sva.ABI_SYMBOLS_COMMON) | (set(sva.ABI_SYMBOLS_3D) if ndims == 3 else set(sva.ABI_SYMBOLS_2D))
    assert declared == listed


def test_domain_init_reports_errors(tmp_path):
    lib = sva.load(2)
    Domain, _, _ = lib._structs
    d = Domain()
    rc = lib.domain_init(str(tmp_path / "missing").encode(), C.byref(d))
    assert rc != 0
    assert b"glsizes.npy" in lib.vof_last_error()


def test_npy_header_mismatch_is_rejected(tmp_path):
    """fileio.c:201-225: dtype and shape are checked before any data is read."""
    ic = random_fields((8, 8), seed=0)
    write_ic(str(tmp_path), ic)
    np.save(tmp_path / "glsizes.npy", np.array([8, 8], dtype=np.int32))
    lib = sva.load(2)
    Domain, _, _ = lib._structs
    d = Domain()
    assert lib.domain_init(str(tmp_path).encode(), C.byref(d)) != 0
    assert b"dtype" in lib.vof_last_error()


@pytest.mark.parametrize("gl", [(16, 24), (8, 12, 10)])
def test_ic_layout(tmp_path, gl):
    ic = random_fields(gl, seed=1)
    write_ic(str(tmp_path), ic)
    assert np.load(tmp_path / "glsizes.npy").dtype == np.dtype("<u8")
    assert np.load(tmp_path / "step.npy").shape == ()
    xf = np.load(tmp_path / "xf.npy")
    xc = np.load(tmp_path / "xc.npy")
    assert xf.shape == (gl[0] + 1,) and xc.shape == (gl[0] + 2,)
    assert xc[0] == 0.0 and xc[-1] == xf[-1]
    for name, shp in field_shapes(gl).items():
        a = np.load(tmp_path / f"{name}.npy")
        assert a.shape == shp and a.dtype == np.dtype("<f8") and a.flags.c_contiguous


def test_droplet_2d_volume_is_exact():
    ic = droplet_2d()
    v = ic.fields["vof"][:, 1:-1]
    assert abs(v.sum() * (1 / 128) * (2 / 256) - np.pi * 0.25 ** 2) < 1e-13
    amp = np.max(np.abs(ic.fields["ux"]))
    assert 0.9e-3 < amp <= 1e-3 and np.all(ic.fields["ux"][:, [0, -1]] == 0.0)


def test_streamfunction_velocity_is_divergence_free():
    xf, _ = make_xgrid(1.0, 32, False)
    ux, uy = velocity_from_streamfunction_2d(xf, 40, 1.0, lambda X, Y: np.sin(3 * X) * np.cos(2 * np.pi * Y))
    dxf = np.diff(xf)
    div = (ux[:, 1:] - ux[:, :-1]) / dxf + (np.roll(uy[:, 1:-1], -1, axis=0) - uy[:, 1:-1]) / (1 / 40)
    assert np.max(np.abs(div)) < 1e-12


def test_droplets_3d_small():
    ic, spheres = droplets_3d(32, 32, 32, seed=2)
    assert len(spheres) == 8
    ux, uy, uz = ic.fields["ux"], ic.fields["uy"], ic.fields["uz"]
    dx = 1 / 32
    div = ((ux[:, :, 1:] - ux[:, :, :-1]) + (np.roll(uy[:, :, 1:-1], -1, axis=1) - uy[:, :, 1:-1])
           + (np.roll(uz[:, :, 1:-1], -1, axis=0) - uz[:, :, 1:-1])) / dx
    assert np.max(np.abs(div)) < 1e-12
    rms = np.sqrt((np.mean(ux[:, :, 1:-1] ** 2) + np.mean(uy[:, :, 1:-1] ** 2) + np.mean(uz[:, :, 1:-1] ** 2)) / 3)
    assert abs(rms - 0.1) < 1e-12
    vol = ic.fields["vof"][:, :, 1:-1].sum() * dx ** 3
    assert abs(vol - sum(4 / 3 * np.pi * r ** 3 for *_, r in spheres)) < 1e-6


@pytest.mark.parametrize("ndims", [2, 3])
def test_tdm_construct_rejects_sizes_the_reference_cannot_solve(ndims):
    """tdm.construct (src/tdm.c:147-183) rejects non-positive sizes and
    counts; sizes the reference's loops would read out of bounds with (gtsv
    reads v[n-2], the periodic path runs gtsv on n - 1 rows) are rejected
    here too.  The checks run before any device call."""
    vt = sva.tdm_vtable(ndims)
    lib = sva.load(ndims)
    info = C.c_void_p(123)
    for size, nrhs, periodic, msg in ((0, 4, False, b"size should be positive"),
                                      (8, 0, False, b"nrhs should be positive"),
                                      (1, 4, False, b"too small"), (2, 4, True, b"too small")):
        assert vt.construct(size, nrhs, periodic, False, C.byref(info)) != 0
        assert not info.value
        assert msg in lib.vof_last_error()
    assert vt.solve(None, None) != 0 and b"info is NULL" in lib.vof_last_error()
    assert vt.destruct(None) != 0
