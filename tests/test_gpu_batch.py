"""Batched form on the GPU: cases of uneven sizes, grids and parameters
advanced together (one launch per kernel of the step for all of them) give
exactly the fields and time steps of integrate() on each case alone."""
import numpy as np
import pytest

import simplevofsolver_amd as sva
from simplevofsolver_amd.ic import droplet_2d, random_fields

pytestmark = pytest.mark.gpu


def _cases_2d():
    cs = [random_fields((32, 48), seed=61), random_fields((24, 40), seed=62, uniformx=False),
          random_fields((64, 32), seed=63), random_fields((16, 20), seed=64, uniformx=False)]
    cs[1].params.update(Ra=2.5e5, We=40.0)
    cs[2].params.update(Ra=2.5e3, We=5.0, add_buoyancy=0)
    d = droplet_2d(nx=32, ny=64, seed=65)
    return cs + [d]


def _cases_3d():
    return [random_fields((16, 12, 20), seed=66), random_fields((12, 16, 8), seed=67, uniformx=False)]


def _cases_mixed_transforms():
    # 256-point y / z transforms select other kernel instantiations than the
    # other lengths: the batch groups those cases into launches of their own
    return [random_fields((16, 256), seed=71), random_fields((12, 40), seed=72, uniformx=False),
            random_fields((8, 256), seed=73, uniformx=False), random_fields((20, 48), seed=74)]


def _cases_mixed_transforms_3d():
    return [random_fields((8, 12, 256), seed=75), random_fields((10, 256, 6), seed=76, uniformx=False),
            random_fields((12, 16, 8), seed=77)]


def _cases_overflow_3d():
    # every cell of the first case an interface cell: each work-group of the
    # 3D interface march meets more cells than its intercept list holds
    from simplevofsolver_amd.ic import _pad_cells
    ic = random_fields((40, 20, 24), seed=79, uniformx=False)
    ic.fields["vof"] = _pad_cells(np.random.default_rng(79).uniform(0.02, 0.98, size=(24, 20, 40)))
    return [ic, random_fields((16, 12, 20), seed=80)]


def _names(nd):
    return ["ux", "uy", "t", "vof", "p", "psi"] + (["uz"] if nd == 3 else [])


@pytest.mark.parametrize("mixed", [-1, 2])
@pytest.mark.parametrize("dim", [2, 3, "2-256", "3-256", "3-overflow"])
def test_batch_is_bit_identical_to_single_cases(dim, mixed, tmp_path):
    cases = {2: _cases_2d, 3: _cases_3d, "2-256": _cases_mixed_transforms,
             "3-256": _cases_mixed_transforms_3d, "3-overflow": _cases_overflow_3d}[dim]()
    dim = len(cases[0].glsizes)
    singles = [sva.solver_from_ic(ic, str(tmp_path / f"a{c}")) for c, ic in enumerate(cases)]
    members = [sva.solver_from_ic(ic, str(tmp_path / f"b{c}")) for c, ic in enumerate(cases)]
    for s in singles + members:
        s.set_poisson_mixed(mixed)
    batch = sva.Batch(members)
    try:
        for step in range(3):
            dts = batch.integrate()
            ref = [s.integrate() for s in singles]
            assert dts == ref, (step, dts, ref)
        batch.synchronize()
        for c, (a, b) in enumerate(zip(members, singles)):
            for n in _names(dim):
                assert np.array_equal(a.get(n), b.get(n)), (c, n)
    finally:
        batch.close()
        for s in singles + members:
            s.close()


def test_batch_rejects_cases_off_the_fused_schedule(tmp_path):
    ic = random_fields((16, 20), seed=70)
    a = sva.solver_from_ic(ic, str(tmp_path / "a"), schedule=0)
    try:
        with pytest.raises(sva.VofError):
            sva.Batch([a])
    finally:
        a.close()


@pytest.mark.parametrize("dim", [2, "2-256", "3-256"])
def test_batch_graph_replays_match_direct_launches(dim, tmp_path):
    """Every replay of the captured phase graphs gives the fields of the
    single cases stepped eagerly in between (and of the same batch with
    graphs off), step after step; the layout is captured once per phase."""
    cases = {2: _cases_2d, "2-256": _cases_mixed_transforms, "3-256": _cases_mixed_transforms_3d}[dim]()
    nd = len(cases[0].glsizes)
    singles = [sva.solver_from_ic(ic, str(tmp_path / f"a{c}")) for c, ic in enumerate(cases)]
    members = [sva.solver_from_ic(ic, str(tmp_path / f"b{c}")) for c, ic in enumerate(cases)]
    direct = [sva.solver_from_ic(ic, str(tmp_path / f"c{c}")) for c, ic in enumerate(cases)]
    batch = sva.Batch(members)
    batch_d = sva.Batch(direct)
    batch_d.set_graphs(False)
    try:
        for step in range(4):
            dts = batch.integrate()
            dts_d = batch_d.integrate()
            ref = [s.integrate() for s in singles]
            assert dts == ref and dts_d == ref, (step, dts, dts_d, ref)
            batch.synchronize()
            batch_d.synchronize()
            for c, (a, d, s) in enumerate(zip(members, direct, singles)):
                for n in _names(nd):
                    want = s.get(n)
                    assert np.array_equal(a.get(n), want), (step, c, n)
                    assert np.array_equal(d.get(n), want), (step, c, n)
        captures, launches = batch.graph_counts()
        assert captures == 4 and launches == 16, (captures, launches)
        assert batch_d.graph_counts() == (0, 0)
    finally:
        batch.close()
        batch_d.close()
        for s in singles + members + direct:
            s.close()


def _cases_split():
    # 10 cases: a batch of 8 or more is run as two batches of alternate cases
    # (5 + 5 here), each on its own stream, stepped in turn
    return _cases_2d() + _cases_mixed_transforms() + [random_fields((40, 24), seed=78, uniformx=False)]


def test_split_batch_is_bit_identical_to_single_cases(tmp_path):
    """A batch of 10 cases (two parts of alternate cases stepped in turn)
    gives the single cases' fields and time steps, with graphs and without,
    returns every case's dt in case order, counts the graphs of both parts,
    and its autotuner (both parts timed together) leaves the state as it was."""
    cases = _cases_split()
    singles = [sva.solver_from_ic(ic, str(tmp_path / f"a{c}")) for c, ic in enumerate(cases)]
    members = [sva.solver_from_ic(ic, str(tmp_path / f"b{c}")) for c, ic in enumerate(cases)]
    direct = [sva.solver_from_ic(ic, str(tmp_path / f"c{c}")) for c, ic in enumerate(cases)]
    batch = sva.Batch(members)
    batch_d = sva.Batch(direct)
    batch_d.set_graphs(False)
    try:
        batch.integrate()
        batch_d.integrate()
        for s in singles:
            s.integrate()
        batch.synchronize()
        before = [[m.get(n) for n in _names(2)] for m in members]
        rep = batch.autotune(1)
        assert any(k.startswith("best:") for k, _, _ in rep)
        # the picks apply to the compared cases too (fft_radix changes results
        # by rounding)
        for k, v, _ in rep:
            if k.startswith("best:"):
                for s in singles + direct:
                    s.set_tuning(k[5:], v)
        for c, m in enumerate(members):
            for n, want in zip(_names(2), before[c]):
                assert np.array_equal(m.get(n), want), ("autotune changed the state", c, n)
        captures0, launches0 = batch.graph_counts()
        for step in range(3):
            dts = batch.integrate()
            dts_d = batch_d.integrate()
            ref = [s.integrate() for s in singles]
            assert dts == ref and dts_d == ref, (step, dts, dts_d, ref)
        batch.synchronize()
        batch_d.synchronize()
        for c, (a, d, s) in enumerate(zip(members, direct, singles)):
            for n in _names(2):
                want = s.get(n)
                assert np.array_equal(a.get(n), want), (c, n)
                assert np.array_equal(d.get(n), want), (c, n)
        captures, launches = batch.graph_counts()
        # 4 phases per part and step, both parts
        assert launches - launches0 == 3 * 4 * 2, (launches0, launches)
        assert captures >= 8, captures
        assert batch_d.graph_counts() == (0, 0)
    finally:
        batch.close()
        batch_d.close()
        for s in singles + members + direct:
            s.close()
