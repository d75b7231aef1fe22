"""Adjoint form on the GPU (2D): the hand-written reverse sweep of the full
RK step against central finite differences of the forward step with the
same frozen dt sequence.

Loss = droplet centroid x + interface length after N steps; control = the
initial velocity DOFs.  The map differentiated is the step with its discrete
decisions held -- the upwind side and single-phase test of every face flux,
the interface test and Newton update count of every cell's intercept -- and
the normals of flat corners (|grad vof| DXC < 1e-2) held constant; the
finite differences replay the decisions of the unperturbed run
(vof_set_freeze) so that they evaluate that map.  Directions: seeded random
divergence-free velocity fields of unit rms.  Bar: every directional
derivative within 1e-6 relative of the central difference at eps = 1e-6."""
import time

import numpy as np
import pytest

import simplevofsolver_amd as sva
from simplevofsolver_amd.ic import divfree_direction, droplet_2d_divfree

pytestmark = pytest.mark.gpu


def _dof_masks(nx, ny):
    mx = np.zeros((ny + 2, nx + 1), dtype=bool)
    mx[1:-1, 1:-1] = True          # faces i = 2..nx
    my = np.zeros((ny + 2, nx + 2), dtype=bool)
    my[1:-1, 1:-1] = True          # cells i = 1..nx
    return mx.ravel(), my.ravel()


def _unit_direction(ic, rng):
    """A divergence-free direction (incompressibility-breaking ones turn
    single-phase cells into interface cells whose normals the perturbation
    itself sets), scaled to unit rms over the velocity DOFs."""
    vx, vy = divfree_direction(ic, rng)
    n, m = ic.glsizes
    mx, my = _dof_masks(n, m)
    r = np.sqrt((np.sum(vx[mx] ** 2) + np.sum(vy[my] ** 2)) / (mx.sum() + my.sum()))
    return vx / r, vy / r


def _setup(tmp_path, n=48, nsteps=12, We=10.0):
    ic = droplet_2d_divfree(nx=n, ny=n, radius=0.2, urms=0.05, nmodes=8, seed=3, We=We)
    ref = sva.solver_from_ic(ic, str(tmp_path / "ref"))
    try:
        dts = [ref.integrate() for _ in range(nsteps)]
    finally:
        ref.close()
    s = sva.solver_from_ic(ic, str(tmp_path / "adj"))
    return ic, s, dts


class Frozen:
    """Losses of the frozen map around the solver's current state: the
    decisions of the unperturbed run are recorded once and replayed."""

    def __init__(self, s, dts):
        self.s, self.dts = s, dts
        self.state = tuple(s.get(k) for k in ("ux", "uy", "p", "vof"))
        self.loss0 = self._run(None, None, 0.0, 1)

    def _run(self, vx, vy, eps, mode):
        s = self.s
        ux0, uy0, p0, vof0 = self.state
        s.set("ux", ux0 + eps * vx if vx is not None else ux0)
        s.set("uy", uy0 + eps * vy if vy is not None else uy0)
        s.boundaries("ux")
        s.boundaries("uy")
        s.set("p", p0)
        s.set("vof", vof0)
        s.set_freeze(mode)
        try:
            for dt in self.dts:
                s.integrate_fixed_dt(dt)
        finally:
            s.set_freeze(0)
        return s.adjoint_loss()

    def central(self, vx, vy, eps):
        return (self._run(vx, vy, eps, 2) - self._run(vx, vy, -eps, 2)) / (2 * eps)


@pytest.mark.parametrize("nsteps,We", [(3, 1e12), (3, 10.0)])
def test_adjoint_matches_central_differences(tmp_path, nsteps, We):
    ic, s, dts = _setup(tmp_path, n=32, nsteps=nsteps, We=We)
    n = ic.glsizes[0]
    try:
        state = tuple(s.get(k) for k in ("ux", "uy", "p", "vof"))
        loss, gux, guy = s.adjoint_gradient(dts, every=2)
        # the call leaves the initial state in place
        for k, a in zip(("ux", "uy", "p", "vof"), state):
            assert np.array_equal(s.get(k), a), k
        mx, my = _dof_masks(n, n)
        assert np.all(gux[~mx] == 0.0) and np.all(guy[~my] == 0.0)
        fr = Frozen(s, dts)
        # the recording run is the unfrozen step: the same loss
        assert fr.loss0 == loss, (fr.loss0, loss)
        rng = np.random.default_rng(7)
        for d in range(4):
            vx, vy = _unit_direction(ic, rng)
            ad = float(gux @ vx + guy @ vy)
            fd = fr.central(vx, vy, 1e-6)
            err = abs(ad - fd) / abs(ad)
            print(f"direction {d}: adjoint {ad:.12e} central FD {fd:.12e} rel {err:.2e}", flush=True)
            assert err <= 1e-6, (d, ad, fd, err)
    finally:
        s.close()


def test_differentiation_workload_256(tmp_path):
    """The differentiation workload: 2D 256 x 256, droplet r = 0.2 in a
    divergence-free velocity of rms 0.05, We = 10, 50 steps, checkpoints
    every 10 steps, 32 seeded directions: 32/32 within 1e-6, and the
    adjoint costs at most 4 forward passes (median of three timings)."""
    ic, s, dts = _setup(tmp_path, n=256, nsteps=50, We=10.0)
    try:
        state = tuple(s.get(k) for k in ("ux", "uy", "p", "vof"))
        loss, gux, guy = s.adjoint_gradient(dts, every=10)
        ratios = []
        for _ in range(3):
            for k, a in zip(("ux", "uy", "p", "vof"), state):
                s.set(k, a)
            s.synchronize()
            t0 = time.perf_counter()
            for dt in dts:
                s.integrate_fixed_dt(dt)
            s.synchronize()
            t_fwd = time.perf_counter() - t0
            for k, a in zip(("ux", "uy", "p", "vof"), state):
                s.set(k, a)
            s.synchronize()
            t0 = time.perf_counter()
            s.adjoint_gradient(dts, every=10)
            t_adj = time.perf_counter() - t0
            ratios.append(t_adj / t_fwd)
        for k, a in zip(("ux", "uy", "p", "vof"), state):
            s.set(k, a)
        fr = Frozen(s, dts)
        assert fr.loss0 == loss
        rng = np.random.default_rng(32)
        errs = []
        for d in range(32):
            vx, vy = _unit_direction(ic, rng)
            ad = float(gux @ vx + guy @ vy)
            fd = fr.central(vx, vy, 1e-6)
            errs.append(abs(ad - fd) / abs(ad))
            print(f"direction {d:2d}: adjoint {ad:+.12e} central FD {fd:+.12e} rel {errs[-1]:.2e}", flush=True)
        errs = np.array(errs)
        print(f"adjoint / forward {np.median(ratios):.2f}; within 1e-6: {(errs <= 1e-6).sum()}/32, "
              f"median {np.median(errs):.2e}, max {errs.max():.2e}", flush=True)
        assert (errs <= 1e-6).all(), errs
        assert np.median(ratios) <= 4.0, ratios
    finally:
        s.close()


def test_adjoint_does_not_depend_on_the_checkpoint_interval(tmp_path):
    ic, s, dts = _setup(tmp_path, n=32, nsteps=6)
    try:
        l1, gx1, gy1 = s.adjoint_gradient(dts, every=1)
        l6, gx6, gy6 = s.adjoint_gradient(dts, every=6)
        assert l1 == l6
        scale = max(np.max(np.abs(gx1)), np.max(np.abs(gy1)))
        # the reverse sweep adds with atomics, so the order of the sums varies
        assert np.max(np.abs(gx1 - gx6)) <= 1e-10 * scale and np.max(np.abs(gy1 - gy6)) <= 1e-10 * scale
    finally:
        s.close()


def test_torch_autograd_binding(tmp_path):
    """step_loss is a torch.autograd.Function whose backward is the adjoint:
    its gradients are the adjoint's, and along a direction they match
    central differences of the frozen map on a small case."""
    import torch
    from simplevofsolver_amd.autograd import step_loss
    ic, s, dts = _setup(tmp_path, n=16, nsteps=2, We=1e12)
    n = ic.glsizes[0]
    try:
        ux = torch.tensor(s.get("ux").reshape(n + 2, n + 1), dtype=torch.float64, requires_grad=True)
        uy = torch.tensor(s.get("uy").reshape(n + 2, n + 2), dtype=torch.float64, requires_grad=True)
        loss = step_loss(s, ux, uy, dts, every=1)
        gx, gy = torch.autograd.grad(loss, (ux, uy))
        l2, gux, guy = s.adjoint_gradient(dts, every=1)
        assert float(loss.detach()) == l2
        scale = max(np.max(np.abs(gux)), np.max(np.abs(guy)))
        assert np.max(np.abs(gx.numpy().ravel() - gux)) <= 1e-10 * scale
        assert np.max(np.abs(gy.numpy().ravel() - guy)) <= 1e-10 * scale
        # the torch gradient along a direction against central differences
        # of the frozen map
        vx, vy = _unit_direction(ic, np.random.default_rng(11))
        tx = torch.tensor(vx.reshape(n + 2, n + 1))
        ty = torch.tensor(vy.reshape(n + 2, n + 2))
        ad = float((gx * tx).sum() + (gy * ty).sum())
        fd = Frozen(s, dts).central(vx, vy, 1e-6)
        assert abs(ad - fd) <= 1e-6 * abs(fd), (ad, fd)
    finally:
        s.close()
