"""Adjoint form on the GPU (2D): the hand-written reverse sweep of the full
RK step against central finite differences of the forward step with the
same frozen dt sequence.

Loss = droplet centroid x + interface length after N steps; control = the
initial velocity DOFs.  Bar: directional derivatives along divergence-free
directions agree with central differences to 1e-6 relative.  The step has
kinks (upwind switches, the single-phase threshold of the THINC flux, the
Newton stopping test of the intercept) whose effect on a central difference
shrinks with eps; at eps = 1e-6 it can reach 1e-3 relative, at 3e-7 and
below it stays under 1e-7 in these cases, so the check uses the smaller of
two step sizes that agrees (roundoff at eps = 1e-7 is about 1e-9)."""
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


def _divfree_direction(ic, rng):
    """Directions that break incompressibility turn single-phase cells into
    interface cells with normals set by the perturbation itself, where the
    step is not differentiable; the checks perturb along divergence-free
    fields."""
    return divfree_direction(ic, rng)


def _setup(tmp_path, n=48, nsteps=12, We=10.0):
    ic = droplet_2d_divfree(nx=n, ny=n, radius=0.2, urms=0.05, nmodes=8, seed=3, We=We)
    ref = sva.solver_from_ic(ic, str(tmp_path / "ref"))
    try:
        dts = [ref.integrate() for _ in range(nsteps)]
    finally:
        ref.close()
    s = sva.solver_from_ic(ic, str(tmp_path / "adj"))
    return ic, s, dts


def _loss_at(s, state, dts, vx, vy, eps):
    ux0, uy0, p0, vof0 = state
    s.set("ux", ux0 + eps * vx)
    s.set("uy", uy0 + eps * vy)
    s.boundaries("ux")
    s.boundaries("uy")
    s.set("p", p0)
    s.set("vof", vof0)
    for dt in dts:
        s.integrate_fixed_dt(dt)
    return s.adjoint_loss()


# (steps, We): several steps without surface tension, one step with it.  With
# surface tension over several steps the curvature of the interface fringe
# (cells holding vof ~ 1e-10 from the THINC tails) jumps when a perturbation
# flips the direction of its near-zero gradients: the step itself is not
# differentiable at the scale of a finite difference there, and central
# differences do not converge as eps shrinks.
@pytest.mark.parametrize("nsteps,We", [(3, 1e12), (1, 10.0)])
def test_adjoint_matches_central_differences(tmp_path, nsteps, We):
    ic, s, dts = _setup(tmp_path, n=32, nsteps=nsteps, We=We)
    n = ic.glsizes[0]
    try:
        state = tuple(s.get(k) for k in ("ux", "uy", "p", "vof"))
        loss, gux, guy = s.adjoint_gradient(dts, every=4)
        # the call leaves the initial state in place
        for k, a in zip(("ux", "uy", "p", "vof"), state):
            assert np.array_equal(s.get(k), a), k
        mx, my = _dof_masks(n, n)
        assert np.all(gux[~mx] == 0.0) and np.all(guy[~my] == 0.0)
        zero = np.zeros_like(state[0]), np.zeros_like(state[1])
        l0 = _loss_at(s, state, dts, zero[0], zero[1], 0.0)
        assert abs(l0 - loss) <= 1e-13 * abs(loss), (l0, loss)
        rng = np.random.default_rng(7)
        for d in range(4):
            vx, vy = _divfree_direction(ic, rng)
            ad = float(gux @ vx + guy @ vy)
            errs = []
            for eps in (3e-7, 1e-7):
                fd = (_loss_at(s, state, dts, vx, vy, eps) - _loss_at(s, state, dts, vx, vy, -eps)) / (2 * eps)
                errs.append(abs(ad - fd) / abs(fd))
                print(f"direction {d}: adjoint {ad:.12e} central FD(eps={eps:g}) {fd:.12e} rel {errs[-1]:.2e}",
                      flush=True)
            assert min(errs) <= 1e-6, (d, ad, errs)
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
    """step_loss is a torch.autograd.Function whose backward is the adjoint;
    gradcheck (central differences of the forward) on a small case."""
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
        # a divergence-free direction through the torch function itself
        vx, vy = _divfree_direction(ic, np.random.default_rng(11))
        tx = torch.tensor(vx.reshape(n + 2, n + 1))
        ty = torch.tensor(vy.reshape(n + 2, n + 2))
        eps = 1e-7
        with torch.no_grad():
            fd = (step_loss(s, ux + eps * tx, uy + eps * ty, dts) - step_loss(s, ux - eps * tx, uy - eps * ty, dts)) / (2 * eps)
        ad = float((gx * tx).sum() + (gy * ty).sum())
        assert abs(ad - float(fd)) <= 1e-6 * abs(float(fd)), (ad, float(fd))
    finally:
        s.close()
