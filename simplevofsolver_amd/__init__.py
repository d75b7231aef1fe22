"""MI355X (gfx950) SimpleVOFSolver time step behind the reference's C ABI.

The product is the pair of HIP libraries ``libvofhip2d.so`` / ``libvofhip3d.so``
(sources in ``csrc/``, interface in ``include/vof_solver.h``).  This module is
a thin ctypes binding used by the tests, the benchmark and Python callers; it
mirrors the reference's structs field by field and never falls back to a CPU
path: if a library is missing or no GPU is present, construction raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from . import ic  # noqa: F401  (re-export: seeded initial conditions)

PKG = os.path.dirname(os.path.abspath(__file__))

_LIBS: dict = {}


class Array(C.Structure):
    """array_t (include/array.h:10-23); data is a device pointer."""
    _fields_ = [("size", C.c_size_t), ("nadds", C.c_void_p), ("datasize", C.c_size_t),
                ("data", C.c_void_p)]


def _structs(ndims: int):
    dom = [("info", C.c_void_p), ("glsizes", C.c_size_t * ndims), ("mysizes", C.c_size_t * ndims),
           ("offsets", C.c_size_t * ndims), ("lengths", C.c_double * ndims),
           ("xf", C.POINTER(C.c_double)), ("xc", C.POINTER(C.c_double)),
           ("dxf", C.POINTER(C.c_double)), ("dxc", C.POINTER(C.c_double)), ("dy", C.c_double)]
    if ndims == 3:
        dom.append(("dz", C.c_double))
    fl = [("ux", Array), ("uy", Array)]
    if ndims == 3:
        fl.append(("uz", Array))
    fl += [("p", Array), ("psi", Array), ("t", Array), ("srcux", Array * 3), ("srcuy", Array * 3)]
    if ndims == 3:
        fl.append(("srcuz", Array * 3))
    fl += [("srct", Array * 3), ("Ra", C.c_double), ("Pr", C.c_double), ("m_dif", C.c_double),
           ("t_dif", C.c_double)]
    itf = [("vof", Array), ("ifrcx", Array), ("ifrcy", Array)]
    if ndims == 3:
        itf.append(("ifrcz", Array))
    itf += [("dvof", Array), ("normal", Array), ("curv", Array), ("flxx", Array), ("flxy", Array)]
    if ndims == 3:
        itf.append(("flxz", Array))
    itf += [("src", Array * 2), ("tension", C.c_double)]

    class Domain(C.Structure):
        _fields_ = dom

    class Fluid(C.Structure):
        _fields_ = fl

    class Interface(C.Structure):
        _fields_ = itf

    return Domain, Fluid, Interface


# every exported entry point of include/vof_solver.h, per dimensionality
ABI_SYMBOLS_COMMON = [
    "domain_init", "domain_save", "domain_check_x_grid_is_uniform",
    "fluid_init", "fluid_save", "fluid_decide_dt", "fluid_predict_field",
    "fluid_compute_potential_dft", "fluid_compute_potential_dct", "fluid_correct_velocity",
    "fluid_update_pressure", "fluid_update_boundaries_ux", "fluid_update_boundaries_uy",
    "fluid_update_boundaries_p", "fluid_update_boundaries_psi", "fluid_update_boundaries_t",
    "interface_init", "interface_save", "interface_compute_curvature_tensor",
    "interface_compute_force", "interface_update_vof", "interface_update_boundaries_vof",
    "integrate", "tdm", "linear_system_init", "linear_system_finalise",
    "vof_last_error", "vof_set_param", "vof_get_param", "vof_array_to_host",
    "vof_array_from_host", "vof_synchronize", "interface_finalise", "fluid_finalise",
    "domain_finalise", "vof_set_schedule", "vof_diagnostics", "vof_profile",
    "vof_profile_reset", "vof_profile_report", "vof_copy_probe", "vof_set_fusion",
    "vof_read_step_time", "vof_write_step_time", "vof_set_tuning", "vof_nusselt",
    "statistics", "vof_statistics_output_to", "vof_set_poisson_mixed",
    "vof_batch_create", "vof_batch_integrate", "vof_batch_synchronize", "vof_batch_destroy",
    "vof_autotune", "vof_batch_autotune", "vof_batch_set_graphs", "vof_batch_graph_counts",
    "vof_integrate_fixed_dt", "vof_device_alloc", "vof_device_free", "vof_device_copy",
    "tdm_solve_noalloc", "vof_set_freeze",
]
ABI_SYMBOLS_2D = ["vof_adjoint_gradient", "vof_adjoint_loss"]
ABI_SYMBOLS_3D = ["fluid_update_boundaries_uz"]


def lib_path(ndims: int) -> str:
    return os.path.join(PKG, f"libvofhip{ndims}d.so")


def load(ndims: int):
    """Load (never build) the HIP library; raises if it is missing."""
    if ndims in _LIBS:
        return _LIBS[ndims]
    path = lib_path(ndims)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built: run `python -m simplevofsolver_amd.build`")
    lib = C.CDLL(path)
    Domain, Fluid, Interface = _structs(ndims)
    vp, dp, sz = C.c_void_p, C.POINTER(C.c_double), C.c_size_t
    lib.vof_last_error.restype = C.c_char_p
    for fn in ("domain_init",):
        getattr(lib, fn).argtypes = [C.c_char_p, vp]
    for fn in ("fluid_init", "interface_init", "fluid_save", "interface_save"):
        getattr(lib, fn).argtypes = [C.c_char_p, vp, vp]
    lib.domain_save.argtypes = [C.c_char_p, vp]
    lib.domain_check_x_grid_is_uniform.argtypes = [vp, C.POINTER(C.c_bool)]
    lib.fluid_decide_dt.argtypes = [vp, vp, dp]
    lib.fluid_predict_field.argtypes = [vp, sz, C.c_double, vp, vp]
    for fn in ("fluid_compute_potential_dft", "fluid_compute_potential_dct",
               "fluid_correct_velocity", "fluid_update_pressure"):
        getattr(lib, fn).argtypes = [vp, sz, C.c_double, vp]
    lib.interface_update_vof.argtypes = [vp, sz, C.c_double, vp, vp]
    for fn in ("interface_compute_curvature_tensor", "interface_compute_force",
               "interface_update_boundaries_vof", "fluid_update_boundaries_ux",
               "fluid_update_boundaries_uy", "fluid_update_boundaries_p",
               "fluid_update_boundaries_psi", "fluid_update_boundaries_t",
               "interface_finalise", "fluid_finalise"):
        getattr(lib, fn).argtypes = [vp, vp]
    if ndims == 3:
        lib.fluid_update_boundaries_uz.argtypes = [vp, vp]
    lib.integrate.argtypes = [vp, vp, vp, dp]
    lib.vof_set_param.argtypes = [vp, C.c_char_p, C.c_double]
    lib.vof_get_param.argtypes = [vp, C.c_char_p, dp]
    lib.vof_array_to_host.argtypes = [vp, vp, vp]
    lib.vof_array_from_host.argtypes = [vp, vp, vp]
    lib.vof_synchronize.argtypes = [vp]
    lib.domain_finalise.argtypes = [vp]
    lib.vof_set_schedule.argtypes = [vp, C.c_int]
    lib.vof_diagnostics.argtypes = [vp, vp, vp, dp]
    lib.vof_profile.argtypes = [vp, C.c_int]
    lib.vof_profile_reset.argtypes = [vp]
    lib.vof_profile_report.argtypes = [vp, C.c_char_p, C.c_size_t]
    lib.vof_copy_probe.argtypes = [vp, C.c_size_t, C.c_int, dp]
    lib.vof_set_fusion.argtypes = [vp, C.c_int]
    lib.vof_set_tuning.argtypes = [vp, C.c_char_p, C.c_int]
    lib.vof_autotune.argtypes = [vp, vp, vp, C.c_int, C.c_char_p, C.c_size_t]
    lib.vof_batch_autotune.argtypes = [vp, C.c_int, C.c_char_p, C.c_size_t]
    lib.vof_batch_set_graphs.argtypes = [vp, C.c_int]
    lib.vof_batch_graph_counts.argtypes = [vp, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)]
    lib.vof_nusselt.argtypes = [vp, vp, dp]
    lib.vof_set_poisson_mixed.argtypes = [vp, C.c_int]
    lib.vof_batch_create.argtypes = [C.c_int, C.POINTER(vp), C.POINTER(vp), C.POINTER(vp), C.POINTER(vp)]
    lib.vof_batch_integrate.argtypes = [vp, dp]
    lib.vof_batch_synchronize.argtypes = [vp]
    lib.vof_batch_destroy.argtypes = [vp]
    lib.vof_integrate_fixed_dt.argtypes = [vp, vp, vp, C.c_double]
    lib.vof_set_freeze.argtypes = [vp, C.c_int]
    if ndims == 2:
        lib.vof_adjoint_gradient.argtypes = [vp, vp, vp, C.c_int, dp, C.c_int, dp, dp, dp]
        lib.vof_adjoint_loss.argtypes = [vp, vp, dp]
    lib.linear_system_init.argtypes = [vp, C.POINTER(C.c_bool), C.POINTER(C.c_size_t), vp]
    lib.linear_system_finalise.argtypes = [vp]
    lib.vof_statistics_output_to.argtypes = [vp, C.c_char_p]
    lib.vof_read_step_time.argtypes = [C.c_char_p, C.POINTER(C.c_size_t), dp]
    lib.vof_write_step_time.argtypes = [C.c_char_p, C.c_size_t, C.c_double]
    lib._structs = (Domain, Fluid, Interface)
    _LIBS[ndims] = lib
    return lib


class VofError(RuntimeError):
    pass


PARAM_NAMES = ("Ra", "Pr", "We", "coef_dt_adv", "coef_dt_dif", "add_buoyancy", "uy_xm",
               "uy_xp", "uz_xm", "uz_xp", "t_xm", "t_xp")


class _StatisticsVT(C.Structure):
    """statistics_t (include/statistics.h:7-30): four function pointers."""
    _fields_ = [
        ("init", C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_double)),
        ("collect", C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p)),
        ("output", C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_size_t)),
        ("get_next_time", C.CFUNCTYPE(C.c_double)),
    ]


class Solver:
    """One grid on the GPU, initialised from an NPY directory through the
    reference's entry points (domain_init, fluid_init, interface_init)."""

    def __init__(self, dirname: str, ndims: int, params: dict | None = None, schedule: int | None = None):
        self.ndims = ndims
        self.lib = load(ndims)
        Domain, Fluid, Interface = self.lib._structs
        self.domain, self.fluid, self.interface = Domain(), Fluid(), Interface()
        self._ok = False
        self._chk(self.lib.domain_init(dirname.encode(), C.byref(self.domain)), "domain_init")
        for k, v in (params or {}).items():
            self._chk(self.lib.vof_set_param(C.byref(self.domain), k.encode(), float(v)), "vof_set_param")
        if schedule is not None:
            self._chk(self.lib.vof_set_schedule(C.byref(self.domain), int(schedule)), "vof_set_schedule")
        self._chk(self.lib.fluid_init(dirname.encode(), C.byref(self.domain), C.byref(self.fluid)), "fluid_init")
        self._chk(self.lib.interface_init(dirname.encode(), C.byref(self.domain), C.byref(self.interface)),
                  "interface_init")
        self._ok = True

    def _chk(self, rc, what):
        if rc != 0:
            msg = self.lib.vof_last_error().decode(errors="replace")
            raise VofError(f"{what} failed: {msg}")

    # -- arrays ---------------------------------------------------------------
    def _array(self, name: str) -> Array:
        """name: a fluid_t / interface_t member, with [n] for the RK slots
        ('srcux_a' = srcux[0], 'src_b' = src[1], ...)."""
        slot = {"a": 0, "b": 1, "g": 2}
        if "_" in name and name.split("_")[-1] in slot:
            base, s = name.rsplit("_", 1)
            owner = self.interface if base == "src" else self.fluid
            return getattr(owner, base)[slot[s]]
        if hasattr(self.fluid, name) and name not in ("Ra", "Pr", "m_dif", "t_dif"):
            return getattr(self.fluid, name)
        return getattr(self.interface, name)

    def get(self, name: str) -> np.ndarray:
        a = self._array(name)
        out = np.empty(a.datasize // 8, dtype=np.float64)
        self._chk(self.lib.vof_array_to_host(C.byref(self.domain), C.byref(a), out.ctypes.data_as(C.c_void_p)),
                  "vof_array_to_host")
        return out

    def set(self, name: str, values) -> None:
        a = self._array(name)
        v = np.ascontiguousarray(values, dtype=np.float64).ravel()
        if v.size * 8 != a.datasize:
            raise ValueError(f"{name}: {v.size * 8} bytes given, {a.datasize} expected")
        self._chk(self.lib.vof_array_from_host(C.byref(self.domain), C.byref(a), v.ctypes.data_as(C.c_void_p)),
                  "vof_array_from_host")

    # -- solver entry points --------------------------------------------------
    def integrate(self) -> float:
        dt = C.c_double(0.0)
        self._chk(self.lib.integrate(C.byref(self.domain), C.byref(self.fluid), C.byref(self.interface),
                                     C.byref(dt)), "integrate")
        return dt.value

    def decide_dt(self) -> float:
        dt = C.c_double(0.0)
        self._chk(self.lib.fluid_decide_dt(C.byref(self.domain), C.byref(self.fluid), C.byref(dt)),
                  "fluid_decide_dt")
        return dt.value

    def curvature_tensor(self):
        self._chk(self.lib.interface_compute_curvature_tensor(C.byref(self.domain), C.byref(self.interface)),
                  "interface_compute_curvature_tensor")

    def force(self):
        self._chk(self.lib.interface_compute_force(C.byref(self.domain), C.byref(self.interface)),
                  "interface_compute_force")

    def update_vof(self, rk: int, dt: float):
        self._chk(self.lib.interface_update_vof(C.byref(self.domain), rk, dt, C.byref(self.fluid),
                                                C.byref(self.interface)), "interface_update_vof")

    def predict(self, rk: int, dt: float):
        self._chk(self.lib.fluid_predict_field(C.byref(self.domain), rk, dt, C.byref(self.fluid),
                                               C.byref(self.interface)), "fluid_predict_field")

    def potential(self, rk: int, dt: float, dct: bool = False):
        fn = self.lib.fluid_compute_potential_dct if dct else self.lib.fluid_compute_potential_dft
        self._chk(fn(C.byref(self.domain), rk, dt, C.byref(self.fluid)), "fluid_compute_potential")

    def correct(self, rk: int, dt: float):
        self._chk(self.lib.fluid_correct_velocity(C.byref(self.domain), rk, dt, C.byref(self.fluid)),
                  "fluid_correct_velocity")

    def update_pressure(self, rk: int, dt: float):
        self._chk(self.lib.fluid_update_pressure(C.byref(self.domain), rk, dt, C.byref(self.fluid)),
                  "fluid_update_pressure")

    def boundaries(self, name: str):
        if name == "vof":
            self._chk(self.lib.interface_update_boundaries_vof(C.byref(self.domain),
                                                               C.byref(self.interface.vof)), "bc vof")
        else:
            fn = getattr(self.lib, f"fluid_update_boundaries_{name}")
            self._chk(fn(C.byref(self.domain), C.byref(getattr(self.fluid, name))), f"bc {name}")

    def synchronize(self):
        self._chk(self.lib.vof_synchronize(C.byref(self.domain)), "vof_synchronize")

    def diagnostics(self) -> dict:
        out = (C.c_double * 16)()
        self._chk(self.lib.vof_diagnostics(C.byref(self.domain), C.byref(self.fluid), C.byref(self.interface),
                                           out), "vof_diagnostics")
        keys = ["max_div", "vof_min", "vof_max", "vof_volume", "ke_x", "ke_y"]
        if self.ndims == 3:
            keys.append("ke_z")
        keys += ["thermal", "mom_x", "mom_y"]
        if self.ndims == 3:
            keys.append("mom_z")
        keys.append("interface_area")
        return {k: out[i] for i, k in enumerate(keys)}

    def save(self, dirname: str, step: int = 0, time: float = 0.0):
        """save_entrypoint (src/main.c:17-32)."""
        os.makedirs(dirname, exist_ok=True)
        self._chk(self.lib.vof_write_step_time(dirname.encode(), step, time), "vof_write_step_time")
        self._chk(self.lib.domain_save(dirname.encode(), C.byref(self.domain)), "domain_save")
        self._chk(self.lib.fluid_save(dirname.encode(), C.byref(self.domain), C.byref(self.fluid)), "fluid_save")
        self._chk(self.lib.interface_save(dirname.encode(), C.byref(self.domain), C.byref(self.interface)),
                  "interface_save")

    def set_fusion(self, mask: int):
        self._chk(self.lib.vof_set_fusion(C.byref(self.domain), int(mask)), "vof_set_fusion")

    def nusselt(self) -> dict:
        """logging/nusselt/main.c:18-50 on the device."""
        out = (C.c_double * 4)()
        self._chk(self.lib.vof_nusselt(C.byref(self.domain), C.byref(self.fluid), out), "vof_nusselt")
        return dict(zip(["heat_flux", "ke_injection", "ke_dissipation", "thermal_dissipation"], out))

    def statistics_init(self, time: float = 0.0):
        """statistics.init (statistics.c:59-100); reads stat_rate and
        stat_after from the environment."""
        vt = _StatisticsVT.in_dll(self.lib, "statistics")
        self._chk(vt.init(C.byref(self.domain), time), "statistics.init")

    def statistics_collect(self):
        vt = _StatisticsVT.in_dll(self.lib, "statistics")
        self._chk(vt.collect(C.byref(self.domain), C.byref(self.fluid)), "statistics.collect")

    def statistics_next_time(self) -> float:
        return _StatisticsVT.in_dll(self.lib, "statistics").get_next_time()

    def statistics_output(self, dirname: str):
        self._chk(self.lib.vof_statistics_output_to(C.byref(self.domain), dirname.encode()),
                  "vof_statistics_output_to")

    def integrate_fixed_dt(self, dt: float):
        """One step with the given dt (no fluid_decide_dt)."""
        self._chk(self.lib.vof_integrate_fixed_dt(C.byref(self.domain), C.byref(self.fluid),
                                                  C.byref(self.interface), float(dt)), "vof_integrate_fixed_dt")

    def set_freeze(self, mode: int):
        """Decision record (1) / replay (2) / off (0) of the 2D fused step:
        replayed runs keep the recorded upwind sides, single-phase tests,
        Newton update counts and flat corners (vof_set_freeze)."""
        self._chk(self.lib.vof_set_freeze(C.byref(self.domain), int(mode)), "vof_set_freeze")

    def adjoint_loss(self) -> float:
        out = C.c_double()
        self._chk(self.lib.vof_adjoint_loss(C.byref(self.domain), C.byref(self.interface), C.byref(out)),
                  "vof_adjoint_loss")
        return out.value

    def adjoint_gradient(self, dts, every: int = 10):
        """(loss, d loss / d ux0, d loss / d uy0) for len(dts) steps with the
        given dt sequence (2D); the solver is left at its initial state."""
        dts = np.ascontiguousarray(dts, dtype=np.float64)
        gux = np.zeros(self._array("ux").datasize // 8)
        guy = np.zeros(self._array("uy").datasize // 8)
        loss = C.c_double()
        self._chk(self.lib.vof_adjoint_gradient(C.byref(self.domain), C.byref(self.fluid), C.byref(self.interface),
                                                len(dts), dts.ctypes.data_as(C.POINTER(C.c_double)), int(every),
                                                C.byref(loss), gux.ctypes.data_as(C.POINTER(C.c_double)),
                                                guy.ctypes.data_as(C.POINTER(C.c_double))), "vof_adjoint_gradient")
        return loss.value, gux, guy

    def set_poisson_mixed(self, refinements: int):
        """refinements < 0: plain FP64 Poisson solve; >= 0: FP32 transforms
        plus that many FP64 residual corrections."""
        self._chk(self.lib.vof_set_poisson_mixed(C.byref(self.domain), int(refinements)), "vof_set_poisson_mixed")

    def set_tuning(self, name: str, value: int):
        self._chk(self.lib.vof_set_tuning(C.byref(self.domain), name.encode(), int(value)), "vof_set_tuning")

    def autotune(self, steps: int = 3) -> list:
        """vof_autotune: time `steps` steps per candidate tile parameter and
        keep the fastest (the state is restored); returns the report lines
        as (name, value, ms_per_step) tuples."""
        buf = C.create_string_buffer(1 << 14)
        self._chk(self.lib.vof_autotune(C.byref(self.domain), C.byref(self.fluid), C.byref(self.interface),
                                        int(steps), buf, len(buf)), "vof_autotune")
        return _tune_report(buf)

    def profile(self, on: bool = True):
        self._chk(self.lib.vof_profile(C.byref(self.domain), int(on)), "vof_profile")

    def profile_reset(self):
        self._chk(self.lib.vof_profile_reset(C.byref(self.domain)), "vof_profile_reset")

    def profile_report(self) -> dict:
        """{kernel: (launches, total_ms)} from the HIP-event timers."""
        buf = C.create_string_buffer(1 << 16)
        self._chk(self.lib.vof_profile_report(C.byref(self.domain), buf, len(buf)), "vof_profile_report")
        out = {}
        for line in buf.value.decode().splitlines():
            name, cnt, ms = line.split()
            out[name] = (int(cnt), float(ms))
        return out

    def copy_probe(self, n: int, reps: int = 10) -> float:
        g = C.c_double(0.0)
        self._chk(self.lib.vof_copy_probe(C.byref(self.domain), n, reps, C.byref(g)), "vof_copy_probe")
        return g.value

    @property
    def x_uniform(self) -> bool:
        b = C.c_bool(False)
        self.lib.domain_check_x_grid_is_uniform(C.byref(self.domain), C.byref(b))
        return bool(b.value)

    def close(self):
        if getattr(self, "_ok", False):
            self.lib.interface_finalise(C.byref(self.domain), C.byref(self.interface))
            self.lib.fluid_finalise(C.byref(self.domain), C.byref(self.fluid))
            self.lib.domain_finalise(C.byref(self.domain))
            self._ok = False

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """vof_batch_*: many independent cases (Solver objects of one
    dimensionality, any sizes) advanced together, one launch per kernel of
    the step for all of them."""

    def __init__(self, solvers):
        self.solvers = list(solvers)
        if not self.solvers:
            raise ValueError("empty batch")
        nd = self.solvers[0].ndims
        if any(s.ndims != nd for s in self.solvers):
            raise ValueError("a batch holds cases of one dimensionality")
        self.lib = self.solvers[0].lib
        n = len(self.solvers)
        arr = C.c_void_p * n
        doms = arr(*[C.addressof(s.domain) for s in self.solvers])
        flus = arr(*[C.addressof(s.fluid) for s in self.solvers])
        ifs = arr(*[C.addressof(s.interface) for s in self.solvers])
        self.h = C.c_void_p()
        if self.lib.vof_batch_create(n, doms, flus, ifs, C.byref(self.h)) != 0:
            raise VofError("vof_batch_create: " + self.lib.vof_last_error().decode())
        self._dts = (C.c_double * n)()

    def integrate(self):
        """One step of every case; returns the list of dt."""
        if self.lib.vof_batch_integrate(self.h, self._dts) != 0:
            raise VofError("vof_batch_integrate: " + self.lib.vof_last_error().decode())
        return list(self._dts)

    def autotune(self, steps: int = 3) -> list:
        """vof_batch_autotune: candidates applied to every case, timed over
        `steps` batch steps; returns (name, value, ms_per_step) tuples."""
        buf = C.create_string_buffer(1 << 14)
        if self.lib.vof_batch_autotune(self.h, int(steps), buf, len(buf)) != 0:
            raise VofError("vof_batch_autotune: " + self.lib.vof_last_error().decode())
        return _tune_report(buf)

    def set_graphs(self, on: bool):
        """hipGraph replay of the step's phases (default on)."""
        if self.lib.vof_batch_set_graphs(self.h, int(bool(on))) != 0:
            raise VofError("vof_batch_set_graphs: " + self.lib.vof_last_error().decode())

    def graph_counts(self):
        """(graphs captured, graph launches) so far."""
        a, b = C.c_size_t(0), C.c_size_t(0)
        if self.lib.vof_batch_graph_counts(self.h, C.byref(a), C.byref(b)) != 0:
            raise VofError("vof_batch_graph_counts: " + self.lib.vof_last_error().decode())
        return a.value, b.value

    def synchronize(self):
        if self.lib.vof_batch_synchronize(self.h) != 0:
            raise VofError("vof_batch_synchronize: " + self.lib.vof_last_error().decode())

    def close(self):
        if self.h:
            self.lib.vof_batch_destroy(self.h)
            self.h = None


class _TdmVT(C.Structure):
    """tdm_t (include/tdm.h:8-45 of the reference): eight function pointers."""
    _fields_ = [
        ("construct", C.CFUNCTYPE(C.c_int, C.c_int, C.c_int, C.c_bool, C.c_bool, C.POINTER(C.c_void_p))),
        ("get_l", C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.POINTER(C.c_double)))),
        ("get_c", C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.POINTER(C.c_double)))),
        ("get_u", C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.POINTER(C.c_double)))),
        ("get_size", C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_int))),
        ("get_nrhs", C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_int))),
        ("solve", C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p)),
        ("destruct", C.CFUNCTYPE(C.c_int, C.c_void_p)),
    ]


def tdm_vtable(ndims: int = 2) -> _TdmVT:
    """The exported ``tdm`` object of the library of that dimensionality."""
    return _TdmVT.in_dll(load(ndims), "tdm")


class Tdm:
    """One tri-diagonal system, used as the reference uses its ``tdm``
    object (src/tdm.c:147-313): construct, fill l / c / u through the host
    arrays that get_l / get_c / get_u return, solve right-hand sides in
    place, destruct.  ``solve`` takes a DEVICE pointer to ``nrhs`` contiguous
    lines of ``size`` doubles (or complex values as (re, im) pairs)."""

    def __init__(self, size: int, nrhs: int, is_periodic: bool, is_complex: bool, ndims: int = 2):
        self.lib = load(ndims)
        self.vt = tdm_vtable(ndims)
        self.info = C.c_void_p()
        rc = self.vt.construct(int(size), int(nrhs), bool(is_periodic), bool(is_complex), C.byref(self.info))
        if rc != 0:
            raise VofError("tdm.construct: " + self.lib.vof_last_error().decode())
        self.size, self.nrhs = int(size), int(nrhs)
        self.l, self.c, self.u = (self._host(getattr(self.vt, f)) for f in ("get_l", "get_c", "get_u"))

    def _host(self, getter) -> np.ndarray:
        p = C.POINTER(C.c_double)()
        if getter(self.info, C.byref(p)) != 0:
            raise VofError(self.lib.vof_last_error().decode())
        return np.ctypeslib.as_array(p, shape=(self.size,))

    def solve(self, data_ptr) -> None:
        if self.vt.solve(self.info, data_ptr) != 0:
            raise VofError("tdm.solve: " + self.lib.vof_last_error().decode())

    def destruct(self) -> None:
        if self.info:
            self.vt.destruct(self.info)
            self.info = C.c_void_p()

    def __del__(self):
        try:
            self.destruct()
        except Exception:
            pass


def _linear_system_struct(ndims: int):
    class LinearSystem(C.Structure):
        """linear_system_t (include/linear_system.h of the reference)."""
        _fields_ = [("is_initialised", C.c_bool), ("implicit", C.c_bool * ndims),
                    ("x1pncl", C.c_void_p), ("y1pncl", C.c_void_p),
                    ("x1pncl_mysizes", C.c_size_t * ndims), ("y1pncl_mysizes", C.c_size_t * ndims),
                    ("tdm_x", C.c_void_p), ("tdm_y", C.c_void_p),
                    ("transposer_x1_to_y1", C.c_void_p), ("transposer_y1_to_x1", C.c_void_p)]
    return LinearSystem


def _tune_report(buf) -> list:
    out = []
    for line in buf.value.decode().splitlines():
        f = line.split()
        if f[0] == "best":
            out.append(("best:" + f[1], int(f[2]), float(f[3])))
        elif f[0] == "start":
            out.append(("start", None, float(f[2])))
        else:
            out.append((f[0], int(f[1]), float(f[2])))
    return out


def solver_from_ic(ic_obj, dirname: str, schedule: int | None = None, params: dict | None = None) -> "Solver":
    """Write ``ic_obj`` (simplevofsolver_amd.ic.InitialCondition) to ``dirname``
    and initialise a GPU solver from it; its ``params`` are applied through
    vof_set_param."""
    from .ic import write_ic
    write_ic(dirname, ic_obj)
    prm = dict(ic_obj.params)
    if params:
        prm.update(params)
    return Solver(dirname, ic_obj.ndims, prm, schedule)

from . import hip  # noqa: E402,F401  (device buffers through the library's runtime)
