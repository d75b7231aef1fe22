"""Driver, save/restart and diagnostics on the GPU.

Restart: no RK state survives a step (beta_0 = 0, runge_kutta.c:8; dt is
recomputed from the fields; psi is recomputed before it is used), so running
2N steps equals N steps + save + restart + N steps exactly.
"""
import os
import subprocess

import numpy as np
import pytest

import simplevofsolver_amd as sva
from parity_util import copy_state, interior, make_pair, rel
from simplevofsolver_amd.build import driver_path
from simplevofsolver_amd.ic import random_fields

pytestmark = pytest.mark.gpu


def _fields(nd):
    return ["ux", "uy", "t", "vof"] + (["uz"] if nd == 3 else [])


@pytest.mark.parametrize("gl", [(24, 32), (12, 16, 10)])
def test_restart_is_bit_identical(gl, tmp_path):
    ic = random_fields(gl, seed=41)
    nd = len(gl)
    a = sva.solver_from_ic(ic, str(tmp_path / "ic"))
    b = sva.solver_from_ic(ic, str(tmp_path / "ic2"))
    try:
        for _ in range(4):
            a.integrate()
        t = 0.0
        for _ in range(2):
            t += b.integrate()
        b.save(str(tmp_path / "save"), step=2, time=t)
    finally:
        b.close()
    c = sva.Solver(str(tmp_path / "save"), nd, dict(ic.params))
    try:
        for _ in range(2):
            c.integrate()
        for n in _fields(nd):
            assert np.array_equal(c.get(n), a.get(n)), n
        assert np.array_equal(interior("p", c.get("p"), gl), interior("p", a.get("p"), gl))
    finally:
        a.close()
        c.close()


def test_save_directory_layout(tmp_path):
    ic = random_fields((16, 12, 8), seed=3)
    g = sva.solver_from_ic(ic, str(tmp_path / "ic"))
    try:
        g.integrate()
        g.save(str(tmp_path / "out"), step=1, time=0.5)
    finally:
        g.close()
    d = tmp_path / "out"
    assert np.load(d / "step.npy") == 1 and np.load(d / "step.npy").dtype == np.uint64
    assert np.load(d / "time.npy") == 0.5
    for n, shp in (("ux", (8, 12, 17)), ("uy", (8, 12, 18)), ("uz", (8, 12, 18)), ("p", (8, 12, 18)),
                   ("t", (8, 12, 18)), ("vof", (8, 12, 18))):
        assert np.load(d / f"{n}.npy").shape == shp
    for n in ("m_dif", "t_dif", "surface_tension", "glsizes", "lengths", "xf", "xc"):
        assert (d / f"{n}.npy").exists()


@pytest.mark.parametrize("nd", [2, 3])
def test_driver_matches_library_steps(nd, tmp_path):
    gl = (16, 20) if nd == 2 else (12, 16, 10)
    ic = random_fields(gl, seed=17)
    g = sva.solver_from_ic(ic, str(tmp_path / "ic"))
    try:
        for _ in range(3):
            g.integrate()
        ref = {n: g.get(n) for n in _fields(nd)}
    finally:
        g.close()
    env = dict(os.environ)
    for k, v in ic.params.items():
        env[k] = repr(float(v))
    exe = driver_path(nd)
    # the driver reads only the environment; compile-time constants stay at
    # the reference defaults, so pass the test case's through the IC params
    # only where they equal those defaults
    assert ic.params["t_xm"] == 0.5 and ic.params["add_buoyancy"] == 1
    res = subprocess.run([exe, str(tmp_path / "ic"), str(tmp_path / "run"), "3"], env=env,
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    out = tmp_path / "run" / "step0000000003"
    g2 = sva.Solver(str(out), nd, dict(ic.params))
    try:
        for n in _fields(nd):
            a = interior(n, g2.get(n), gl) if n != "ux" else g2.get(n)
            b = interior(n, ref[n], gl) if n != "ux" else ref[n]
            assert np.array_equal(a, b), n
    finally:
        g2.close()


@pytest.mark.parametrize("gl", [(32, 24), (16, 12, 20)])
def test_diagnostics_match_oracle(gl, tmp_path):
    ic = random_fields(gl, seed=23)
    o, g = make_pair(ic, tmp_path)
    try:
        for _ in range(2):
            o.integrate()
            g.integrate()
        d = g.diagnostics()
        vmin, vmax, vol = o.vof_stats()
        e = o.energy()
        assert abs(d["vof_min"] - vmin) <= 1e-12 and abs(d["vof_max"] - vmax) <= 1e-12
        assert rel(d["vof_volume"], vol) < 1e-12
        assert rel(d["ke_x"], e[0]) < 1e-10 and rel(d["ke_y"], e[1]) < 1e-10
        assert rel(d["thermal"], e[-1]) < 1e-10
        assert d["max_div"] < 1e-10 and o.max_divergence() < 1e-10
        mom = o.momentum()
        scale = max(abs(d["mom_x"]), abs(d["mom_y"]), 1e-3)
        assert abs(d["mom_x"] - mom[0]) < 1e-12 * scale and abs(d["mom_y"] - mom[1]) < 1e-12 * scale
        if len(gl) == 3:
            assert abs(d["mom_z"] - mom[2]) < 1e-12 * max(abs(d["mom_z"]), 1e-3)
        assert rel(d["interface_area"], o.interface_area()) < 1e-12
    finally:
        g.close()


@pytest.mark.parametrize("gl", [(32, 24), (16, 12, 20)])
def test_nusselt_matches_oracle(gl, tmp_path):
    ic = random_fields(gl, seed=29)
    o, g = make_pair(ic, tmp_path)
    try:
        for _ in range(2):
            o.integrate()
            g.integrate()
        d = g.nusselt()
        ref = o.nusselt()
        got = np.array([d["heat_flux"], d["ke_injection"], d["ke_dissipation"], d["thermal_dissipation"]])
        assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref)), (got, ref)
    finally:
        g.close()


@pytest.mark.parametrize("gl", [(24, 16), (12, 10, 8)])
def test_statistics_match_oracle(gl, tmp_path, monkeypatch):
    """Accumulation is one add per element per sample and the reduction sums
    rows in order, as the oracle does: on identical samples (the oracle's
    state copied in before each collect) the profiles agree bit for bit."""
    import oracle_lib as OL
    monkeypatch.setenv("stat_rate", "0.5")
    monkeypatch.setenv("stat_after", "0.0")
    ic = random_fields(gl, seed=31)
    o, g = make_pair(ic, tmp_path)
    st = OL.OracleStats(o)
    try:
        g.statistics_init(0.0)
        assert g.statistics_next_time() == 0.5
        for _ in range(3):
            st.collect()
            copy_state(o, g)
            g.statistics_collect()
            o.integrate()
        assert g.statistics_next_time() == 2.0
        out = tmp_path / "stat"
        g.statistics_output(str(out))
        assert np.load(out / "num.npy") == 3
        names = ["ux1", "ux2", "uy1", "uy2", "t1", "t2", "uxt"] + (["uz1", "uz2"] if len(gl) == 3 else [])
        for n in names:
            assert np.array_equal(np.load(out / f"{n}.npy"), st.reduce(n)), n
        assert (out / "xf.npy").exists() and (out / "m_dif.npy").exists()
    finally:
        st.close()
        g.close()


def test_driver_logging_and_statistics(tmp_path):
    """With log_rate / stat_rate set the driver writes the reference's log
    files (logging/main.c:82-104) and a statistics directory."""
    gl = (16, 20)
    ic = random_fields(gl, seed=37)
    sva.solver_from_ic(ic, str(tmp_path / "ic")).close()
    env = dict(os.environ)
    for k, v in ic.params.items():
        env[k] = repr(float(v))
    env.update(log_rate="1e-9", stat_rate="1e-9", stat_after="0")
    res = subprocess.run([driver_path(2), str(tmp_path / "ic"), str(tmp_path / "run"), "4"], env=env,
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    log = tmp_path / "run" / "log"
    for name, ncol in (("divergence.dat", 2), ("momentum.dat", 3), ("energy.dat", 4), ("nusselt.dat", 5),
                       ("vof.dat", 4)):
        rows = [ln.split() for ln in (log / name).read_text().splitlines()]
        assert len(rows) == 4 and all(len(r) == ncol for r in rows), name
        float(rows[-1][-1])
    st = tmp_path / "run" / "stat" / "step0000000004"
    assert np.load(st / "num.npy") == 4
    assert np.load(st / "ux1.npy").shape == (gl[0] + 1,)
    final = sva.Solver(str(tmp_path / "run" / "step0000000004"), 2, dict(ic.params))
    try:
        d = final.diagnostics()
        vol = float((log / "vof.dat").read_text().splitlines()[-1].split()[-1])
        assert abs(vol - d["vof_volume"]) < 1e-14
    finally:
        final.close()
