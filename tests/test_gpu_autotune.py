"""Autotuner and FFT radix on the GPU.

* fft_radix = 4 sends 256-point transforms through the mixed-radix path: the
  same discrete transform with another rounding sequence, so fields agree
  with the radix-16 path to the full-step parity bar (1e-10 relative; p and
  psi modulo their mean, whose roundoff-defined constant from the singular
  k = 0 Poisson mode amplifies rounding differences to ~1e-11).
* vof_autotune / vof_batch_autotune time steps from a device snapshot and
  restore it: every array of the case is bit-identical before and after, the
  chosen parameters come from the candidate sets, and the tuned case then
  steps to the fields of an untuned one (bit for bit unless the tuner chose
  the mixed-radix path, then to 1e-10).
"""
import numpy as np
import pytest

import simplevofsolver_amd as sva
from parity_util import demean, interior, rel
from simplevofsolver_amd.ic import random_fields

pytestmark = pytest.mark.gpu

CANDIDATES = {"fft_radix": {4, 16}, "fft_elems": {1024, 2048}, "iface_planes": {0, 4, 8, 16},
              "rhs_rows": {1, 2, 4, 8, 16}, "lines_lpb": {0, 2, 4, 8, 16, 32}, "iface_overlap": {0, 1, 2},
              "thomas_rows": {8, 16, 32}, "pred_rows": {2, 4, 8}}


def _names(nd):
    return ["ux", "uy", "t", "vof", "p", "psi", "normal", "curv", "src_a", "src_b",
            "srcux_a", "srcux_b", "srcuy_a", "srcuy_b", "srct_a", "srct_b"] + (
        ["uz", "srcuz_a", "srcuz_b"] if nd == 3 else [])


def _fields(nd):
    return ["ux", "uy", "t", "vof", "p", "psi"] + (["uz"] if nd == 3 else [])


def _cmp(name, a, b, ic):
    """Relative difference; p and psi modulo their mean (the singular k = 0
    Poisson mode leaves an additive constant set by roundoff)."""
    if name in ("p", "psi"):
        dxf = np.diff(ic.xf)
        g = ic.glsizes
        return rel(interior(name, demean(name, a, g, dxf), g), interior(name, demean(name, b, g, dxf), g))
    return rel(a, b)


def _close(name, a, b, ic, exact):
    if exact:
        return np.array_equal(a, b)
    return _cmp(name, a, b, ic) < 1e-10


@pytest.mark.parametrize("glsizes,seed", [((16, 256), 81), ((8, 12, 256), 82), ((10, 256, 6), 83)])
@pytest.mark.parametrize("mixed", [-1, 1])
def test_fft_radix4_matches_radix16(glsizes, seed, mixed, tmp_path):
    ic = random_fields(glsizes, seed=seed, uniformx=False)
    g0 = sva.solver_from_ic(ic, str(tmp_path / "a"))
    g1 = sva.solver_from_ic(ic, str(tmp_path / "b"))
    g1.set_tuning("fft_radix", 4)
    for g in (g0, g1):
        g.set_poisson_mixed(mixed)
    tol = 1e-10 if mixed < 0 else 1e-6   # FP32 transforms with one refinement: psi within 1e-8
    try:
        for step in range(2):
            d0, d1 = g0.integrate(), g1.integrate()
            assert abs(d0 - d1) <= 1e-14 * d0
            for n in _fields(len(glsizes)):
                assert _cmp(n, g1.get(n), g0.get(n), ic) < tol, (step, n)
    finally:
        g0.close()
        g1.close()


def test_fft_radix_rejects_other_values(tmp_path):
    g = sva.solver_from_ic(random_fields((16, 20), seed=84), str(tmp_path / "a"))
    try:
        with pytest.raises(sva.VofError):
            g.set_tuning("fft_radix", 8)
    finally:
        g.close()


def _chosen(report):
    best = {name[5:]: v for name, v, _ in report if name.startswith("best:")}
    for k, v in best.items():
        assert v in CANDIDATES[k], (k, v)
    return best


@pytest.mark.parametrize("glsizes,seed", [((32, 48), 85), ((16, 256), 86), ((12, 16, 8), 87), ((8, 12, 256), 88)])
def test_autotune_restores_state_and_keeps_results(glsizes, seed, tmp_path):
    nd = len(glsizes)
    ic = random_fields(glsizes, seed=seed, uniformx=(seed % 2 == 0))
    g0 = sva.solver_from_ic(ic, str(tmp_path / "a"))
    g1 = sva.solver_from_ic(ic, str(tmp_path / "b"))
    try:
        # one step first, so that the carried sources and normals hold data
        assert g0.integrate() == g1.integrate()
        before = {n: g1.get(n) for n in _names(nd)}
        report = g1.autotune(steps=1)
        best = _chosen(report)
        expect = {"pred_rows", "rhs_rows", "lines_lpb"} | ({"iface_planes"} if nd == 3 else set())
        assert expect <= set(best), best
        for n in _names(nd):
            assert np.array_equal(g1.get(n), before[n]), n
        exact = best.get("fft_radix", 16) == 16
        for step in range(2):
            d0, d1 = g0.integrate(), g1.integrate()
            assert d0 == d1 if exact else abs(d0 - d1) <= 1e-14 * d0
            for n in _fields(nd):
                assert _close(n, g1.get(n), g0.get(n), ic, exact), (step, n, best)
    finally:
        g0.close()
        g1.close()


def test_batch_autotune_restores_state_and_keeps_results(tmp_path):
    cases = [random_fields((16, 256), seed=91), random_fields((24, 40), seed=92, uniformx=False),
             random_fields((32, 64), seed=93)]
    singles = [sva.solver_from_ic(ic, str(tmp_path / f"a{c}")) for c, ic in enumerate(cases)]
    members = [sva.solver_from_ic(ic, str(tmp_path / f"b{c}")) for c, ic in enumerate(cases)]
    batch = sva.Batch(members)
    try:
        batch.integrate()
        for s in singles:
            s.integrate()
        batch.synchronize()
        before = [{n: m.get(n) for n in _names(2)} for m in members]
        best = _chosen(batch.autotune(steps=1))
        for m, b in zip(members, before):
            for n in _names(2):
                assert np.array_equal(m.get(n), b[n]), n
        exact = best.get("fft_radix", 16) == 16
        for step in range(2):
            dts = batch.integrate()
            ref = [s.integrate() for s in singles]
            for a, b in zip(dts, ref):
                assert a == b if exact else abs(a - b) <= 1e-14 * b
        batch.synchronize()
        for m, s, ic in zip(members, singles, cases):
            for n in _fields(2):
                assert _close(n, m.get(n), s.get(n), ic, exact), (n, best)
    finally:
        batch.close()
        for s in singles + members:
            s.close()
