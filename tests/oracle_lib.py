"""ctypes wrapper of the CPU oracle (oracle/liboracle{2,3}d.so).

Test infrastructure: the oracle is the checker, never the thing measured or
shipped.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
import this module.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")

_dp = C.POINTER(C.c_double)


class OrcParams(C.Structure):
    _fields_ = [("Ra", C.c_double), ("Pr", C.c_double), ("We", C.c_double),
                ("coef_dt_adv", C.c_double), ("coef_dt_dif", C.c_double),
                ("add_buoyancy", C.c_int),
                ("uy_xm", C.c_double), ("uy_xp", C.c_double),
                ("uz_xm", C.c_double), ("uz_xp", C.c_double),
                ("t_xm", C.c_double), ("t_xp", C.c_double)]


_LIBS = {}


def build_oracle():
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def load(ndims: int):
    if ndims in _LIBS:
        return _LIBS[ndims]
    path = os.path.join(ORACLE_DIR, f"liboracle{ndims}d.so")
    if not os.path.exists(path):
        build_oracle()
    lib = C.CDLL(path)
    vp = C.c_void_p
    lib.orc_create.restype = vp
    lib.orc_create.argtypes = [C.POINTER(C.c_size_t), _dp, _dp, _dp]
    lib.orc_destroy.argtypes = [vp]
    lib.orc_set_params.argtypes = [vp, C.POINTER(OrcParams)]
    lib.orc_get_params.argtypes = [vp, C.POINTER(OrcParams)]
    lib.orc_field_count.restype = C.c_size_t
    lib.orc_field_count.argtypes = [vp, C.c_char_p]
    for fn in ("orc_get_field", "orc_set_field", "orc_load_ic", "orc_dump_slab"):
        getattr(lib, fn).argtypes = [vp, C.c_char_p, _dp]
    for fn in ("orc_m_dif", "orc_t_dif", "orc_tension", "orc_max_divergence"):
        getattr(lib, fn).restype = C.c_double
        getattr(lib, fn).argtypes = [vp]
    lib.orc_x_grid_is_uniform.argtypes = [vp]
    lib.orc_integrate.argtypes = [vp, _dp]
    lib.orc_decide_dt.argtypes = [vp, _dp]
    lib.orc_interface_compute_curvature_tensor.argtypes = [vp]
    lib.orc_interface_compute_force.argtypes = [vp]
    for fn in ("orc_interface_update_vof", "orc_fluid_predict_field",
               "orc_fluid_correct_velocity", "orc_fluid_update_pressure"):
        getattr(lib, fn).argtypes = [vp, C.c_int, C.c_double]
    lib.orc_fluid_compute_potential.argtypes = [vp, C.c_int, C.c_double, C.c_int]
    lib.orc_update_boundaries.argtypes = [vp, C.c_char_p]
    lib.orc_tdm_solve.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp]
    lib.orc_tdm_solve_complex.argtypes = [C.c_int, C.c_int, C.c_int, _dp, _dp, _dp, _dp]
    lib.orc_fft.argtypes = [C.c_int, C.c_int, _dp]
    lib.orc_set_fft_variant.argtypes = [C.c_int]
    lib.orc_rfft.argtypes = [C.c_int, _dp, _dp]
    lib.orc_irfft.argtypes = [C.c_int, _dp, _dp]
    lib.orc_redft10.argtypes = [C.c_int, _dp, _dp]
    lib.orc_redft01.argtypes = [C.c_int, _dp, _dp]
    lib.orc_compute_intercept.restype = C.c_double
    lib.orc_compute_intercept.argtypes = [C.c_double, _dp]
    lib.orc_indicator.restype = C.c_double
    lib.orc_indicator.argtypes = [_dp, _dp]
    lib.orc_vof_stats.argtypes = [vp, _dp, _dp, _dp]
    lib.orc_energy.argtypes = [vp, _dp]
    lib.orc_momentum.argtypes = [vp, _dp]
    lib.orc_interface_area.argtypes = [vp]
    lib.orc_interface_area.restype = C.c_double
    lib.orc_nusselt.argtypes = [vp, _dp]
    lib.orc_stats_create.argtypes = [vp]
    lib.orc_stats_create.restype = vp
    lib.orc_stats_destroy.argtypes = [vp]
    lib.orc_stats_collect.argtypes = [vp, vp]
    lib.orc_stats_num.argtypes = [vp]
    lib.orc_stats_num.restype = C.c_size_t
    lib.orc_stats_reduce.argtypes = [vp, C.c_char_p, _dp]
    _LIBS[ndims] = lib
    return lib


def _ptr(a):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


FIELDS_2D = ["ux", "uy", "p", "psi", "t", "vof", "ifrcx", "ifrcy", "dvof", "normal",
             "curv", "flxx", "flxy", "src_a", "src_b",
             "srcux_a", "srcux_b", "srcux_g", "srcuy_a", "srcuy_b", "srcuy_g",
             "srct_a", "srct_b", "srct_g"]
FIELDS_3D = FIELDS_2D + ["uz", "ifrcz", "flxz", "srcuz_a", "srcuz_b", "srcuz_g"]


class Oracle:
    """One grid held by the CPU oracle."""

    def __init__(self, ic, params=None):
        self.ndims = ic.ndims
        self.lib = load(self.ndims)
        gl = (C.c_size_t * self.ndims)(*ic.glsizes)
        lengths = np.ascontiguousarray(ic.lengths, dtype=np.float64)
        self._xf = np.ascontiguousarray(ic.xf, dtype=np.float64)
        self._xc = np.ascontiguousarray(ic.xc, dtype=np.float64)
        self.h = self.lib.orc_create(gl, _ptr(lengths), _ptr(self._xf), _ptr(self._xc))
        self.glsizes = tuple(ic.glsizes)
        prm = dict(ic.params)
        if params:
            prm.update(params)
        self.set_params(**prm)
        for name, arr in ic.fields.items():
            a = np.ascontiguousarray(arr, dtype=np.float64)
            rc = self.lib.orc_load_ic(self.h, name.encode(), _ptr(a))
            assert rc == 0, name

    def __del__(self):
        try:
            self.lib.orc_destroy(self.h)
        except Exception:
            pass

    def set_params(self, **kw):
        p = OrcParams()
        self.lib.orc_get_params(self.h, C.byref(p))
        for k, v in kw.items():
            setattr(p, k, type(getattr(p, k))(v))
        self.lib.orc_set_params(self.h, C.byref(p))

    def get(self, name):
        n = self.lib.orc_field_count(self.h, name.encode())
        assert n > 0, name
        out = np.empty(n, dtype=np.float64)
        assert self.lib.orc_get_field(self.h, name.encode(), _ptr(out)) == 0
        return out

    def set(self, name, arr):
        a = np.ascontiguousarray(arr, dtype=np.float64).ravel()
        assert a.size == self.lib.orc_field_count(self.h, name.encode())
        assert self.lib.orc_set_field(self.h, name.encode(), _ptr(a)) == 0

    @property
    def m_dif(self):
        return self.lib.orc_m_dif(self.h)

    @property
    def t_dif(self):
        return self.lib.orc_t_dif(self.h)

    @property
    def tension(self):
        return self.lib.orc_tension(self.h)

    @property
    def x_uniform(self):
        return bool(self.lib.orc_x_grid_is_uniform(self.h))

    def integrate(self):
        dt = C.c_double(0.0)
        assert self.lib.orc_integrate(self.h, C.byref(dt)) == 0
        return dt.value

    def decide_dt(self):
        dt = C.c_double(0.0)
        assert self.lib.orc_decide_dt(self.h, C.byref(dt)) == 0
        return dt.value

    def curvature_tensor(self):
        self.lib.orc_interface_compute_curvature_tensor(self.h)

    def force(self):
        self.lib.orc_interface_compute_force(self.h)

    def update_vof(self, rk, dt):
        self.lib.orc_interface_update_vof(self.h, rk, dt)

    def predict(self, rk, dt):
        self.lib.orc_fluid_predict_field(self.h, rk, dt)

    def potential(self, rk, dt, use_dct=None):
        if use_dct is None:
            use_dct = self.x_uniform
        self.lib.orc_fluid_compute_potential(self.h, rk, dt, int(bool(use_dct)))

    def correct(self, rk, dt):
        self.lib.orc_fluid_correct_velocity(self.h, rk, dt)

    def update_pressure(self, rk, dt):
        self.lib.orc_fluid_update_pressure(self.h, rk, dt)

    def boundaries(self, name):
        assert self.lib.orc_update_boundaries(self.h, name.encode()) == 0

    def max_divergence(self):
        return self.lib.orc_max_divergence(self.h)

    def vof_stats(self):
        a, b, c = C.c_double(), C.c_double(), C.c_double()
        self.lib.orc_vof_stats(self.h, C.byref(a), C.byref(b), C.byref(c))
        return a.value, b.value, c.value

    def energy(self):
        out = np.zeros(self.ndims + 1)
        self.lib.orc_energy(self.h, _ptr(out))
        return out

    def momentum(self):
        out = np.zeros(self.ndims)
        self.lib.orc_momentum(self.h, _ptr(out))
        return out

    def interface_area(self):
        return self.lib.orc_interface_area(self.h)

    def nusselt(self):
        out = np.zeros(4)
        self.lib.orc_nusselt(self.h, _ptr(out))
        return out


class OracleStats:
    """statistics.c accumulators on the CPU."""

    def __init__(self, o: Oracle):
        self.o = o
        self.h = o.lib.orc_stats_create(o.h)

    def collect(self):
        self.o.lib.orc_stats_collect(self.h, self.o.h)

    @property
    def num(self):
        return self.o.lib.orc_stats_num(self.h)

    def reduce(self, name):
        nx = self.o.glsizes[0]
        out = np.zeros(nx + 2)
        w = self.o.lib.orc_stats_reduce(self.h, name.encode(), _ptr(out))
        assert w > 0, name
        return out[:w]

    def close(self):
        if self.h:
            self.o.lib.orc_stats_destroy(self.h)
            self.h = None


# free-standing numerics -----------------------------------------------------

def tdm_solve(l, c, u, q, periodic=False, lib_ndims=2):
    lib = load(lib_ndims)
    l, c, u = (np.ascontiguousarray(a, dtype=np.float64) for a in (l, c, u))
    n = l.size
    if np.iscomplexobj(q):
        qq = np.ascontiguousarray(q, dtype=np.complex128).copy()
        nrhs = qq.size // n
        buf = qq.view(np.float64)
        assert lib.orc_tdm_solve_complex(n, nrhs, int(periodic), _ptr(l), _ptr(c), _ptr(u), _ptr(buf)) == 0
        return qq
    qq = np.ascontiguousarray(q, dtype=np.float64).copy()
    nrhs = qq.size // n
    assert lib.orc_tdm_solve(n, nrhs, int(periodic), _ptr(l), _ptr(c), _ptr(u), _ptr(qq)) == 0
    return qq


def set_fft_variant(v, ndims):
    """Twiddle table of the oracle's transforms (0 octant-reduced, the
    default; 1 direct cos / sin): the same DFT up to rounding."""
    load(ndims).orc_set_fft_variant(int(v))


def fft(x, sign=-1):
    lib = load(2)
    z = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib.orc_fft(z.size, sign, _ptr(z.view(np.float64)))
    return z


def rfft(x):
    lib = load(2)
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(x.size // 2 + 1, dtype=np.complex128)
    lib.orc_rfft(x.size, _ptr(x), _ptr(out.view(np.float64)))
    return out


def irfft(z, n):
    lib = load(2)
    z = np.ascontiguousarray(z, dtype=np.complex128)
    out = np.zeros(n)
    lib.orc_irfft(n, _ptr(z.view(np.float64)), _ptr(out))
    return out


def redft10(x):
    lib = load(2)
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros_like(x)
    lib.orc_redft10(x.size, _ptr(x), _ptr(out))
    return out


def redft01(x):
    lib = load(2)
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros_like(x)
    lib.orc_redft01(x.size, _ptr(x), _ptr(out))
    return out


def intercept(vof, normal, ndims=2):
    lib = load(ndims)
    n = np.ascontiguousarray(normal, dtype=np.float64)
    return lib.orc_compute_intercept(float(vof), _ptr(n))


def indicator(normal_d, x, ndims=2):
    lib = load(ndims)
    n = np.ascontiguousarray(normal_d, dtype=np.float64)
    xx = np.ascontiguousarray(x, dtype=np.float64)
    return lib.orc_indicator(_ptr(n), _ptr(xx))
