"""Mixed-precision Poisson form on the GPU: FP32 transforms with FP64
eliminations and FP64 residual refinement, against the FP64 oracle.

Bars: after two refinements psi (modulo its mean) within 1e-10 relative of
the FP64 oracle and the pressure within 1e-8; without refinement the FP32
solve alone is within 1e-4 (it is an approximation, not a parity case);
velocity, VOF and temperature after a step are then within the plain
path's 1e-10 bar.
"""
import numpy as np
import pytest

from parity_util import copy_state, demean, interior, make_pair, read_mask, rel
from simplevofsolver_amd.ic import random_fields

pytestmark = pytest.mark.gpu

CASES = [
    ((32, 48), True, 51),
    ((24, 40), False, 52),
    ((16, 12, 20), True, 53),
    ((12, 16, 8), False, 54),
    # 256-point transforms (register radix-16 path) in y, y only, z only
    ((16, 256), True, 55),
    ((8, 256, 6), False, 56),
    ((6, 8, 256), True, 57),
]


def _ids(c):
    return "x".join(map(str, c[0])) + ("-uniform" if c[1] else "-stretched")


@pytest.mark.parametrize("refine,bar", [(0, 1e-4), (1, 1e-8), (2, 1e-10)])
@pytest.mark.parametrize("case", CASES, ids=[_ids(c) for c in CASES])
def test_mixed_potential_against_fp64_oracle(case, refine, bar, tmp_path):
    glsizes, uniformx, seed = case
    ic = random_fields(glsizes, seed=seed, uniformx=uniformx)
    o, g = make_pair(ic, tmp_path)
    g.set_poisson_mixed(refine)
    dxf = np.diff(ic.xf)
    try:
        dt = o.decide_dt()
        for rk in range(3):
            copy_state(o, g)
            o.potential(rk, dt)
            g.potential(rk, dt)
            m = read_mask("psi", glsizes)
            a = demean("psi", g.get("psi"), glsizes, dxf)[m]
            b = demean("psi", o.get("psi"), glsizes, dxf)[m]
            e = rel(a, b)
            print(f"refine={refine} rk={rk} psi rel err {e:.3e}", flush=True)
            assert e < bar, e
    finally:
        g.close()


@pytest.mark.parametrize("case", CASES, ids=[_ids(c) for c in CASES])
def test_mixed_step_against_fp64_oracle(case, tmp_path):
    glsizes, uniformx, seed = case
    ic = random_fields(glsizes, seed=seed, uniformx=uniformx)
    o, g = make_pair(ic, tmp_path)
    g.set_poisson_mixed(2)
    nd = len(glsizes)
    dxf = np.diff(ic.xf)
    try:
        for _ in range(2):
            assert o.integrate() == g.integrate()
        for n in ["ux", "uy", "t", "vof"] + (["uz"] if nd == 3 else []):
            e = rel(interior(n, g.get(n), glsizes), interior(n, o.get(n), glsizes))
            assert e < 1e-10, (n, e)
        e = rel(interior("p", demean("p", g.get("p"), glsizes, dxf), glsizes),
                interior("p", demean("p", o.get("p"), glsizes, dxf), glsizes))
        print(f"pressure rel err {e:.3e}", flush=True)
        assert e < 1e-8, e
    finally:
        g.close()
