"""GPU parity: every entry point of the HIP library against the CPU oracle
on the same seeded inputs, through the C ABI.

Bars: FP64 fields within 1e-10 relative after a full
step; per entry point, on identical inputs, results that are bit-identical
where the arithmetic is the same sequence of IEEE operations (stencils,
Thomas with shared factors, reductions by min) and within 1e-13 relative where
libm differs (exp/log in the THINC intercept and indicator) or the transform
differs (FFT, potential compared modulo its mean).
"""
import numpy as np
import pytest

from parity_util import copy_state, demean, interior, make_pair, read_mask, rel
from simplevofsolver_amd.ic import droplet_2d, random_fields

pytestmark = pytest.mark.gpu

CASES = [
    ((32, 48), True, 11),
    ((24, 40), False, 12),
    ((16, 12, 20), True, 13),
    ((12, 16, 8), False, 14),
    # 256-point transforms in y, z (the register radix-16 path)
    ((16, 256), False, 15),
    ((8, 12, 256), True, 16),
    ((10, 256, 6), False, 17),
]


def _ids(c):
    return "x".join(map(str, c[0])) + ("-uniform" if c[1] else "-stretched")


@pytest.mark.parametrize("case", CASES, ids=[_ids(c) for c in CASES])
def test_entry_points_isolated(case, tmp_path):
    glsizes, uniformx, seed = case
    _check_entry_points(random_fields(glsizes, seed=seed, uniformx=uniformx), tmp_path)


@pytest.mark.parametrize("uniformx", [True, False])
def test_entry_points_isolated_on_droplets(uniformx, tmp_path):
    """The 3D droplet workload (surface tension, THINC fringes, We = 50): every
    entry point matches the oracle on identical inputs.  Over whole steps the
    two drift apart: corner normals of rounding-level vof gradients are
    normalised rounding noise (gradient / max(|gradient|, DBL_EPSILON),
    src/interface/curvature_tensor.c:14-48), so the 1e-16 differences of libm
    exp/log in the THINC intercept flip them within one step.  The reference
    does the same against itself under any change of rounding."""
    from simplevofsolver_amd.ic import droplets_3d
    ic, _ = droplets_3d(64, 32, 32, lx=2.0, ndrops=2, rmin=0.15, rmax=0.25, urms=0.2, seed=5, Re=1000.0,
                        We=50.0, uniformx=uniformx)
    _check_entry_points(ic, tmp_path)


def _check_entry_points(ic, tmp_path):
    glsizes = ic.glsizes
    o, g = make_pair(ic, tmp_path)
    nd = len(glsizes)
    dxf = np.diff(ic.xf)
    try:
        # boundary conditions applied by *_init
        for n in ["ux", "uy", "p", "t", "vof"] + (["uz"] if nd == 3 else []):
            assert np.array_equal(g.get(n), o.get(n)), n
        dt = o.decide_dt()
        assert g.decide_dt() == dt
        for rk in range(3):
            copy_state(o, g)
            o.curvature_tensor()
            g.curvature_tensor()
            assert np.array_equal(g.get("dvof"), o.get("dvof"))
            assert np.array_equal(g.get("curv"), o.get("curv"))
            assert rel(g.get("normal"), o.get("normal")) < 1e-13
            copy_state(o, g)
            o.force()
            g.force()
            for n in ["ifrcx", "ifrcy"] + (["ifrcz"] if nd == 3 else []):
                assert np.array_equal(g.get(n), o.get(n)), n
            copy_state(o, g)
            o.update_vof(rk, dt)
            g.update_vof(rk, dt)
            for n in ["flxx", "flxy"] + (["flxz"] if nd == 3 else []) + ["src_a", "src_b", "vof"]:
                assert rel(g.get(n), o.get(n)) < 1e-13, n
            copy_state(o, g)
            o.predict(rk, dt)
            g.predict(rk, dt)
            fl = ["srcux_a", "srcux_g", "srcuy_a", "srcuy_g", "srct_a", "srct_g", "ux", "uy", "t"]
            if nd == 3:
                fl += ["srcuz_a", "srcuz_g", "uz"]
            for n in fl:
                assert np.array_equal(g.get(n), o.get(n)), n
            copy_state(o, g)
            o.potential(rk, dt)
            g.potential(rk, dt, dct=o.x_uniform)
            m = read_mask("psi", glsizes)
            e = rel(demean("psi", g.get("psi"), glsizes, dxf)[m], demean("psi", o.get("psi"), glsizes, dxf)[m])
            assert e < 1e-12, e
            copy_state(o, g)
            o.correct(rk, dt)
            g.correct(rk, dt)
            for n in ["ux", "uy"] + (["uz"] if nd == 3 else []):
                assert np.array_equal(g.get(n), o.get(n)), n
            copy_state(o, g)
            o.update_pressure(rk, dt)
            g.update_pressure(rk, dt)
            assert np.array_equal(g.get("p"), o.get("p"))
    finally:
        g.close()


@pytest.mark.parametrize("case", CASES, ids=[_ids(c) for c in CASES])
def test_integrate_steps(case, tmp_path):
    glsizes, uniformx, seed = case
    ic = random_fields(glsizes, seed=seed, uniformx=uniformx)
    o, g = make_pair(ic, tmp_path)
    nd = len(glsizes)
    dxf = np.diff(ic.xf)
    try:
        for step in range(3):
            dto = o.integrate()
            dtg = g.integrate()
            assert abs(dtg - dto) <= 1e-12 * dto
            for n in ["ux", "uy", "t", "vof"] + (["uz"] if nd == 3 else []):
                assert rel(g.get(n), o.get(n)) < 1e-10, (step, n)
            for n in ("p", "psi"):
                e = rel(interior(n, demean(n, g.get(n), glsizes, dxf), glsizes),
                        interior(n, demean(n, o.get(n), glsizes, dxf), glsizes))
                assert e < 1e-10, (step, n, e)
    finally:
        g.close()


def test_droplet_2d_workload(tmp_path):
    """The 2D 128x256 droplet workload: 5 steps against the oracle, plus
    conservation of the VOF volume and a divergence-free velocity."""
    ic = droplet_2d(seed=1)
    o, g = make_pair(ic, tmp_path)
    dxf = np.diff(ic.xf)
    try:
        vol0 = g.diagnostics()["vof_volume"]
        for step in range(5):
            dto = o.integrate()
            dtg = g.integrate()
            assert abs(dtg - dto) <= 1e-12 * dto
        for n in ("ux", "uy", "vof"):
            assert rel(g.get(n), o.get(n)) < 1e-10, n
        assert rel(interior("p", demean("p", g.get("p"), ic.glsizes, dxf), ic.glsizes),
                   interior("p", demean("p", o.get("p"), ic.glsizes, dxf), ic.glsizes)) < 1e-10
        d = g.diagnostics()
        assert abs(d["vof_volume"] - vol0) < 1e-12 * vol0
        assert d["max_div"] < 1e-10
    finally:
        g.close()


# 63: every stage fused, the predictor as source kernel + x-line kernel;
# 79 and 127: the predictor as one kernel (sources, x solve, update)
FUSION_MASKS = [1, 2, 6, 8, 16, 32, 63, 79, 127]


@pytest.mark.parametrize("mask", FUSION_MASKS)
@pytest.mark.parametrize("case", CASES, ids=[_ids(c) for c in CASES])
def test_fused_schedule_is_bit_identical(case, mask, tmp_path):
    """The fused kernels (with boundary images folded in) give exactly the
    fields of the per-function schedule."""
    import simplevofsolver_amd as sva
    glsizes, uniformx, seed = case
    ic = random_fields(glsizes, seed=seed, uniformx=uniformx)
    g0 = sva.solver_from_ic(ic, str(tmp_path / "a"), schedule=0)
    g1 = sva.solver_from_ic(ic, str(tmp_path / "b"), schedule=1)
    g1.set_fusion(mask)
    g1.set_tuning("iface_overlap", 1)   # the two-stream schedule (default from 2^20 cells on)
    nd = len(glsizes)
    try:
        for step in range(2):
            assert g0.integrate() == g1.integrate()
            for n in ["ux", "uy", "t", "vof", "p", "psi"] + (["uz"] if nd == 3 else []):
                a, b = g1.get(n), g0.get(n)
                assert np.array_equal(a, b), (step, n, float(np.max(np.abs(a - b))))
    finally:
        g0.close()
        g1.close()


@pytest.mark.parametrize("tuning", [("fft_elems", 2048), ("lines_lpb", 1), ("lines_lpb", 7), ("lines_lpb", 15), ("xcd_bands", 1),
                                    ("rhs_rows", 1), ("rhs_rows", 2), ("rhs_rows", 4), ("rhs_rows", 16),
                                    ("iface_planes", 0), ("iface_planes", 4), ("iface_planes", 16),
                                    ("iface_overlap", 0), ("iface_overlap", 1), ("iface_overlap", 2),
                                    ("thomas_rows", 16), ("thomas_rows", 32),
                                    ("pred_rows", 1), ("pred_rows", 2), ("pred_rows", 4)])
@pytest.mark.parametrize("case", CASES, ids=[_ids(c) for c in CASES])
def test_tile_parameters_do_not_change_results(case, tuning, tmp_path):
    """Tile sizes change the work split only: every line and every FFT column
    is processed by the same sequence of operations."""
    import simplevofsolver_amd as sva
    glsizes, uniformx, seed = case
    ic = random_fields(glsizes, seed=seed, uniformx=uniformx)
    g0 = sva.solver_from_ic(ic, str(tmp_path / "a"))
    g1 = sva.solver_from_ic(ic, str(tmp_path / "b"))
    if tuning[0] in ("rhs_rows", "lines_lpb"):
        # parameters of the two-kernel predictor
        g0.set_fusion(63)
        g1.set_fusion(63)
    if tuning[0] == "thomas_rows":
        # a parameter of the sequential x elimination of the Poisson solver
        g0.set_tuning("thomas_part", 0)
        g1.set_tuning("thomas_part", 0)
    g1.set_tuning(*tuning)
    nd = len(glsizes)
    try:
        for step in range(2):
            assert g0.integrate() == g1.integrate()
            for n in ["ux", "uy", "t", "vof", "p", "psi"] + (["uz"] if nd == 3 else []):
                assert np.array_equal(g1.get(n), g0.get(n)), (step, n)
    finally:
        g0.close()
        g1.close()


@pytest.mark.parametrize("planes", [4, 8, 16])
def test_interface_list_overflow(planes, tmp_path):
    """Every cell an interface cell (vof uniform in (0.02, 0.98)): each
    32 x 8 x planes work-group of the 3D interface march meets 256 x planes
    interface cells, more than its 256-entry intercept list, so the cells
    beyond the list are solved after the march.  Curvature equal to the
    oracle's, normals and intercepts within 1e-13 of them, and every field
    bit for bit equal to the block kernel's (one cell per lane, no list)."""
    import simplevofsolver_amd as sva
    from simplevofsolver_amd.ic import _pad_cells
    glsizes = (40, 20, 24)
    ic = random_fields(glsizes, seed=43, uniformx=False)
    rng = np.random.default_rng(43)
    ic.fields["vof"] = _pad_cells(rng.uniform(0.02, 0.98, size=(24, 20, 40)))
    o, g = make_pair(ic, tmp_path)
    b = sva.solver_from_ic(ic, str(tmp_path / "blocks"))
    g.set_tuning("iface_planes", planes)
    b.set_tuning("iface_planes", 0)
    try:
        o.curvature_tensor()
        g.curvature_tensor()
        b.curvature_tensor()
        assert np.array_equal(g.get("curv"), o.get("curv"))
        assert rel(g.get("normal"), o.get("normal")) < 1e-13
        for n in ("curv", "normal"):
            assert np.array_equal(g.get(n), b.get(n)), n
        for step in range(2):
            assert g.integrate() == b.integrate()
            for n in ("ux", "uy", "uz", "t", "vof", "p", "psi", "normal", "curv"):
                assert np.array_equal(g.get(n), b.get(n)), (step, n)
    finally:
        g.close()
        b.close()


PART_CASES = [
    ((256, 64), False, 31),
    ((37, 24), True, 32),
    ((37, 24), False, 33),
    ((5, 16), False, 34),
    ((100, 16, 12), False, 35),
    ((256, 8, 16), True, 36),
]


@pytest.mark.parametrize("case", PART_CASES, ids=[_ids(c) for c in PART_CASES])
def test_partitioned_x_solve(case, tmp_path):
    """The Poisson solver's partitioned x elimination (k_thomas_part: segments
    of 2 .. 8 rows, one lane each, and a reduced system of their end rows; the
    singular (0, 0) line in sequence) against the sequential elimination and
    the oracle: psi within 1e-12 modulo its mean, for nx from 5 (two segments
    of 2 and 3 rows) to 256 (32 segments of 8), uneven segment lengths
    (nx = 37, 100) and stretched x; then a full step of each within 1e-10."""
    import simplevofsolver_amd as sva
    glsizes, uniformx, seed = case
    ic = random_fields(glsizes, seed=seed, uniformx=uniformx)
    dxf = np.diff(ic.xf)
    o, g = make_pair(ic, tmp_path)
    s = sva.solver_from_ic(ic, str(tmp_path / "seq"))
    s.set_tuning("thomas_part", 0)
    try:
        dt = o.decide_dt()
        o.potential(0, dt)
        g.potential(0, dt)
        s.potential(0, dt)
        m = read_mask("psi", glsizes)
        ref = demean("psi", o.get("psi"), glsizes, dxf)[m]
        part = demean("psi", g.get("psi"), glsizes, dxf)[m]
        seq = demean("psi", s.get("psi"), glsizes, dxf)[m]
        assert rel(part, seq) < 1e-12, rel(part, seq)
        assert rel(part, ref) < 1e-12, rel(part, ref)
        assert g.integrate() == s.integrate()
        for n in ["ux", "uy"] + (["uz"] if len(glsizes) == 3 else []):
            assert rel(g.get(n), s.get(n)) < 1e-10, n
    finally:
        g.close()
        s.close()


@pytest.mark.parametrize("uniformx", [True, False])
def test_3d_step_is_symmetric_in_y_and_z(uniformx, tmp_path):
    """The GPU step keeps a y <-> z symmetric 3D state symmetric (every z
    term equals its y counterpart) and agrees with the oracle on it."""
    from test_oracle_pins import yz_asymmetry, yz_symmetric_case
    nx, n = 10, 12
    ic = yz_symmetric_case(nx, n, seed=41, uniformx=uniformx)
    o, g = make_pair(ic, tmp_path)
    try:
        assert yz_asymmetry(g.get, nx, n) == 0.0
        for step in range(3):
            dto = o.integrate()
            dtg = g.integrate()
            assert abs(dtg - dto) <= 1e-12 * dto
            assert yz_asymmetry(g.get, nx, n) < 1e-10, step
        for name in ("ux", "uy", "uz", "t", "vof"):
            assert rel(g.get(name), o.get(name)) < 1e-10, name
    finally:
        g.close()
