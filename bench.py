"""Benchmark: plain FP64 time step of the 3D 256^3 droplet workload on one
MI355X, reported in Mcell-steps/s.

Contract: ``python bench.py --gpus N --steps K --warmup W`` does W untimed
steps, then times exactly K steps of integrate() (the reference's one-step
entry point) with the device synchronised on both sides, and prints one JSON
line.  A "step" is one full RK3 time step (dt, 3 x [curvature, force, VOF
update, predictor, Poisson, correction, pressure]) over the whole grid with
every field resident in HBM.

Per-kernel times come from HIP events recorded on the library's own stream
around every launch (vof_profile) during a second timed pass of the same K
steps, so the events stay out of the throughput timing; the dominant
kernel's algorithmic bytes per launch divided by its mean launch time gives
``roofline.achieved``.  ``cpu_baseline`` times the CPU oracle (a single-threaded
C restatement of the reference step) on the same workload: one process
(median of 3 steps) and one process per host core at once (at most 16).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mcell-steps/s at 3D 256³ FP64 plain vs batched/mixed-prec; adjoint/forward time"
HBM_PEAK_GBS = 8000.0


def algorithmic_bytes(kernel: str, nx: int, ny: int, nz: int, iface_cells: int | None = None) -> float | None:
    """Compulsory FP64 traffic of one launch: every array element the kernel
    must read once plus every element it must write once (3D layouts).

    Cell normals (nx, ny, nz, d: 32 B) exist only at interface cells
    (1e-8 <= vof <= 1 - 1e-8; curvature_tensor.c:94-145 skips the others), so
    the interface and VOF-update kernels charge them per interface cell:
    `iface_cells` of them (every cell when it is None, an upper bound)."""
    nh = ny // 2 + 1
    c = nx * ny * nz
    s = nx * nh * nz
    nrm = 32 * (c if iface_cells is None else iface_cells)
    table = {
        "k_gradient": (nx + 1) * (ny + 3) * (nz + 3) * (8 + 24),
        "k_normal_curv": nx * (ny + 2) * (nz + 2) * (24 + 8 + 8),
        "k_force": c * (8 + 8 + 24),
        "k_flux": c * (24 + 8 + 24),
        "k_vof_update": c * (24 + 8 + 8 + 8 + 8),
        # fields read (ux: ux uy uz p t ifrcx; uy/uz: ux uy uz p ifrc; t: ux uy
        # uz t) + srcb read, srca + increment written
        "k_rhs_ux": (nx - 1) * ny * nz * (48 + 8 + 16),
        "k_rhs_uy": c * (40 + 8 + 16),
        "k_rhs_uz": c * (40 + 8 + 16),
        "k_rhs_t": c * (32 + 8 + 16),
        # four fields (ux: nx-1 DOFs per line, uy uz t: nx) in one launch:
        # increment read, field read and write
        "k_lines_lds": ((nx - 1) + 3 * nx) * ny * nz * (8 + 8 + 8),
        "k_div_rfft_y": c * 24 + s * 16,
        "k_fft_z": s * 32,
        "k_thomas_spec": s * 32,
        "k_thomas_part": s * 32,
        "k_irfft_y_psi": s * 16 + c * 8,
        "k_correct": c * (8 + 24 + 24),
        "k_pressure": c * (8 + 8 + 8),
        "k_dt_partial": c * 24,
        # fused schedule: vof read and curvature written per cell, the normal
        # written at interface cells
        "k_iface_fused": c * (8 + 8) + nrm,
        "k_iface_march": c * (8 + 8) + nrm,
        # velocity (3) and vof read per cell, the normal of interface cells,
        # the new vof written; srcb read in RK stages 1 and 2, srca written in
        # stages 0 and 1 (the figures are means over the three stages of a
        # step, whose launches alternate)
        "k_vof_fused": c * (24 + 8 + 8) + nrm + c * (8 + 8) * 2 / 3,
        # ux uy uz p t, vof, curvature read and the increment of 4 fields
        # written; srcb of 4 fields read in stages 1 and 2, srca of 4 fields
        # written in stages 0 and 1
        "k_rhs_fused": c * (40 + 8 + 8 + 32) + c * (32 + 32) * 2 / 3,
        "k_rhs_march": c * (40 + 8 + 8 + 32) + c * (32 + 32) * 2 / 3,
        # the same reads and source traffic; the new ux uy uz t written
        # instead of their increments (the x solve runs in LDS)
        "k_pred_lines": c * (40 + 8 + 8 + 32) + c * (32 + 32) * 2 / 3,
        # psi, ux uy uz p read and written back
        "k_corr_press": c * (8 + 32 + 32),
    }
    return table.get(kernel)


def read_traffic(kernel: str, workload: str):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary
    (profiles/traffic.txt: `workload kernel fetch_bytes write_bytes`), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) == 4 and parts[0] == workload and parts[1] == kernel:
                return float(parts[2]) + float(parts[3])
    return None


def cpu_worker(n: int, reps: int, sync_dir: str | None, nworkers: int) -> int:
    """One CPU-baseline process: the oracle (single-threaded C) on the n^3
    benchmark workload, `reps` steps each timed on its own; prints the step
    times as one JSON line.  With sync_dir, every worker of a concurrent
    group has built its oracle before any starts timing."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle
    from simplevofsolver_amd.ic import droplets_3d
    ic, _ = droplets_3d(n, n, n, seed=2)
    o = Oracle(ic)
    del ic
    if sync_dir:
        open(os.path.join(sync_dir, f"ready{os.getpid()}"), "w").close()
        t_wait = time.perf_counter()
        while len([f for f in os.listdir(sync_dir) if f.startswith("ready")]) < nworkers:
            if time.perf_counter() - t_wait > 600:
                print(json.dumps({"error": "workers never all ready"}), flush=True)
                return 1
            time.sleep(0.05)
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        o.integrate()
        times.append(time.perf_counter() - t0)
    print(json.dumps({"times": times}), flush=True)
    return 0


def cpu_baseline(n: int, reps: int = 3, workers: int | None = None):
    """The CPU oracle on the same n^3 workload, timed on the host before the
    GPU is touched: (1) one single-threaded process, median of `reps` steps;
    (2) `workers` such processes at once (one per core, at most 16: the host
    share of one GPU), the aggregate rate of their concurrent steps (median
    step time).  (2) is an upper bound for a parallel CPU run of the step
    on those cores (a decomposed run also exchanges halos and transposes).
    The figure reported is (2); (1) is kept beside it."""
    import statistics
    import subprocess
    cells = n ** 3
    env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1")
    cmd = [sys.executable, os.path.abspath(__file__), "--cpu-worker", "--n", str(n), "--cpu-reps", str(reps)]

    def run(nw):
        with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as sync:
            extra = ["--cpu-sync", sync, "--cpu-workers", str(nw)] if nw > 1 else []
            procs = [subprocess.Popen(cmd + extra, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, env=env)
                     for _ in range(nw)]
            times = []
            for p in procs:
                out, err = p.communicate(timeout=900)
                if p.returncode != 0:
                    raise RuntimeError(f"CPU baseline worker failed: {err[-500:]}")
                times.append(json.loads(out.strip().splitlines()[-1])["times"])
        return times

    ncpu = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    nw = workers or max(1, min(16, ncpu))
    t0 = time.perf_counter()
    single = run(1)[0]
    t1 = time.perf_counter()
    print(f"[bench] CPU oracle, one core: step times {', '.join(f'{t:.2f}' for t in single)} s", file=sys.stderr,
          flush=True)
    many = run(nw) if nw > 1 else [single]
    t2 = time.perf_counter()
    med1 = statistics.median(single)
    medn = statistics.median([t for w in many for t in w])
    print(f"[bench] CPU oracle, {nw} concurrent processes: median step {medn:.2f} s", file=sys.stderr, flush=True)
    return {"value": round(nw * cells / medn / 1e6, 3), "unit": "Mcell-steps/s", "cores": nw, "kind": "port",
            "sample": f"{nw} concurrent single-threaded oracle processes, each {reps} timed steps of the same "
                      f"{n}^3 workload (median step {medn:.2f} s; {t2 - t1:.0f} s wall); host reports {ncpu} "
                      f"CPUs; an upper bound for a parallel CPU step on {nw} cores",
            "single_core": {"value": round(cells / med1 / 1e6, 3), "cores": 1, "reps": reps,
                            "step_s": [round(t, 3) for t in single], "median_step_s": round(med1, 3),
                            "wall_s": round(t1 - t0, 1)}}


def drift_record(d0, d1, steps, ndims):
    """Changes of the conserved / monitored quantities over `steps` steps
    (logging/energy.c, logging/interface.c, logging/divergence.c)."""
    ke = ("ke_x", "ke_y", "ke_z") if ndims == 3 else ("ke_x", "ke_y")
    return {
        "vof_volume_rel": abs(d1["vof_volume"] - d0["vof_volume"]) / d0["vof_volume"],
        "max_div": d1["max_div"],
        "kinetic_energy_rel": sum(d1[k] for k in ke) / sum(d0[k] for k in ke) - 1.0,
        "interface_area_rel": d1["interface_area"] / d0["interface_area"] - 1.0,
        "vof_min": d1["vof_min"],
        "vof_max": d1["vof_max"],
        "steps": steps,
    }


def drift_2d(sva, steps, oracle_steps=100):
    """The 2D workload (128 x 256 on 1 x 2, droplet r = 0.25 at (0.5, 0.5),
    velocity perturbation 1e-3, Re = 100, We = 10, CFL 0.4): drift of volume,
    kinetic energy and interface length over `steps` steps, and after
    `oracle_steps` steps the largest relative difference of ux, uy, vof from
    the CPU oracle stepped from the same state."""
    import numpy as np
    from simplevofsolver_amd.ic import droplet_2d
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from oracle_lib import Oracle
    ic = droplet_2d(seed=1)
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as d:
        g = sva.solver_from_ic(ic, d)
    try:
        o = Oracle(ic)
        d0 = g.diagnostics()
        t0 = time.perf_counter()
        for _ in range(oracle_steps):
            g.integrate()
            o.integrate()
        diff = {}
        for name in ("ux", "uy", "vof"):
            a, b = g.get(name), o.get(name)
            diff[name] = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
        t1 = time.perf_counter()
        ngpu = max(0, steps - oracle_steps)
        for _ in range(ngpu):
            g.integrate()
        g.synchronize()
        t2 = time.perf_counter()
        el = t2 - t0
        rec = drift_record(d0, g.diagnostics(), max(steps, oracle_steps), ndims=2)
        rec["vs_oracle_after"] = {"steps": oracle_steps, **diff}
        # wall_s includes the oracle's steps beside the first oracle_steps
        rec["wall_s"] = round(el, 2)
        if ngpu:
            rec["gpu_ms_per_step"] = round((t2 - t1) / ngpu * 1e3, 4)
            rec["gpu_mcell_steps_per_s"] = round(ic.glsizes[0] * ic.glsizes[1] * ngpu / (t2 - t1) / 1e6, 3)
        rec["workload"] = "2D 128x256 on 1x2, droplet r=0.25, perturbation 1e-3, Re=100, We=10, FP64"
        print(f"[bench] 2D drift over {rec['steps']} steps: KE {rec['kinetic_energy_rel']:+.3e}, interface "
              f"{rec['interface_area_rel']:+.3e}, volume {rec['vof_volume_rel']:.1e}; vs oracle after "
              f"{oracle_steps} steps {diff}", file=sys.stderr, flush=True)
        return rec
    finally:
        g.close()


def plain_rate(g, cells, steps):
    """Mcell-steps/s of the plain FP64 step from the solver's current state
    (the forms below are timed after the drift run, so they are compared
    with the plain step at that same state)."""
    g.integrate()
    g.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.integrate()
    g.synchronize()
    return cells * steps / (time.perf_counter() - t0) / 1e6


def mixed_form(g, cells, steps, plain_value, refine):
    """The same solver with the FP32-transform Poisson solve followed by
    `refine` FP64 residual refinements, timed over `steps` steps."""
    g.set_poisson_mixed(refine)
    g.integrate()
    g.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.integrate()
    g.synchronize()
    el = time.perf_counter() - t0
    g.set_poisson_mixed(-1)
    v = cells * steps / el / 1e6
    print(f"[bench] mixed-precision form, {refine} refinement(s): {v:.1f} Mcell-steps/s "
          f"({el / steps * 1e3:.3f} ms/step)", file=sys.stderr, flush=True)
    # accuracy of the Poisson solution, from the GPU tests against the FP64
    # oracle (tests/test_gpu_mixed.py bars)
    bar = {0: "psi within 1e-4 relative (FP32 solve alone)", 1: "psi within 1e-8 relative",
           2: "psi within 1e-10 relative"}.get(refine, "psi within 1e-10 relative")
    return {"value": round(v, 3), "unit": "Mcell-steps/s", "ms_per_step": round(el / steps * 1e3, 4),
            "vs_plain": round(v / plain_value, 4), "refinements": refine, "accuracy": bar,
            "workload": f"same 3D 256^3 case, FP32 transforms + FP64 x eliminations, {refine} FP64 refinement(s)"}


def autotuned_form(g, cells, steps, plain_value, k):
    """The plain step after vof_autotune (k steps per candidate; the state is
    restored by the tuner), timed over `steps` steps."""
    t0 = time.perf_counter()
    rep = g.autotune(k)
    t_tune = time.perf_counter() - t0
    tuned = {n[5:]: v for n, v, _ in rep if n.startswith("best:")}
    g.integrate()
    g.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        g.integrate()
    g.synchronize()
    el = time.perf_counter() - t0
    v = cells * steps / el / 1e6
    print(f"[bench] autotuned ({t_tune:.1f} s, {len(rep)} trials): {tuned}: {v:.1f} Mcell-steps/s",
          file=sys.stderr, flush=True)
    return {"value": round(v, 3), "unit": "Mcell-steps/s", "ms_per_step": round(el / steps * 1e3, 4),
            "vs_plain": round(v / plain_value, 4), "chosen": tuned, "trials": len(rep),
            "tune_seconds": round(t_tune, 2), "workload": "same 3D 256^3 case, FP64, autotuned tile parameters"}


def _interior(name, a, nx, ny, nz):
    """Interior DOFs of a padded 3D array (x fastest; vof has 2-wide y/z halos)."""
    h = 2 if name == "vof" else 1
    n0 = nx + 1 if name == "ux" else nx + 2
    a = a.reshape(nz + 2 * h, ny + 2 * h, n0)
    return a[h:-h, h:-h, 1:-1]


def precision_form(sva, steps, refine):
    """Precision workload: 3D 512 x 256 x 256 on 2 x 1 x 1 (x walls, y/z
    periodic), 32 seeded droplets r in [0.03, 0.08], divergence-free velocity
    rms 0.2, Re = 1000, We = 50.  Three solvers start from the same NPY and
    step in lockstep: the plain FP64 step, the mixed form (FP32 transforms,
    FP64 x eliminations, `refine` FP64 residual refinements) and an FP64
    twin whose only difference is the rounding of its transforms (the
    mixed-radix path for the 256-point lengths).  At steps 1, 10 and `steps`
    each field of the mixed run and of the twin is compared with the FP64
    run (max-norm relative; p and psi modulo their mean): the twin shows how
    far rounding-level differences alone grow in this flow.  Before stepping,
    one Poisson solve of each solver on the same (initial) input gives the
    error of the mixed solver itself.  Throughput is
    timed per solver over steps 2..`steps` (each step synchronised)."""
    import numpy as np
    from simplevofsolver_amd.ic import droplets_3d, write_ic
    nx, ny, nz = 512, 256, 256
    t0 = time.perf_counter()
    ic, _ = droplets_3d(nx, ny, nz, lx=2.0, ly=1.0, lz=1.0, ndrops=32, rmin=0.03, rmax=0.08, urms=0.2,
                        seed=5, Re=1000.0, We=50.0)
    print(f"[bench] precision IC {nx}x{ny}x{nz} generated in {time.perf_counter() - t0:.1f} s", file=sys.stderr,
          flush=True)
    cells = nx * ny * nz
    names = ("ux", "uy", "uz", "vof", "p", "psi")
    out = {"workload": f"3D {nx}x{ny}x{nz} on 2x1x1, 32 droplets r in [0.03,0.08], velocity rms 0.2, "
                       f"Re=1000, We=50; {steps} steps; mixed = FP32 transforms + FP64 x eliminations + "
                       f"{refine} FP64 refinement(s); fp64_twin = FP64 with mixed-radix 256-point transforms",
           "steps": steps, "refinements": refine}
    params = dict(ic.params)
    checkpoints = sorted({1, min(10, steps), steps})

    def fields(g):
        res = {}
        for n in names:
            a = _interior(n, g.get(n), nx, ny, nz)
            res[n] = a - a.mean() if n in ("p", "psi") else a
        return res

    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as d:
        write_ic(d, ic)
        del ic
        sol = {}
        try:
            for form in ("fp64", "mixed", "fp64_twin"):
                sol[form] = sva.Solver(d, 3, params)
            sol["mixed"].set_poisson_mixed(refine)
            sol["fp64_twin"].set_tuning("fft_radix", 4)
            # one Poisson solve on identical input (the initial state): the
            # accuracy of the mixed solver itself
            dt0 = sol["fp64"].decide_dt()
            for g in sol.values():
                g.potential(0, dt0)
            ref = _interior("psi", sol["fp64"].get("psi"), nx, ny, nz)
            ref = ref - ref.mean()
            solve_err = {}
            for k in ("mixed", "fp64_twin"):
                a = _interior("psi", sol[k].get("psi"), nx, ny, nz)
                a = a - a.mean()
                solve_err[k] = float(np.max(np.abs(a - ref)) / np.max(np.abs(ref)))
            del ref
            out["poisson_solve_psi_error_vs_fp64"] = solve_err
            print(f"[bench] precision workload, one Poisson solve: psi error vs FP64 {solve_err}", file=sys.stderr,
                  flush=True)
            elapsed = {k: 0.0 for k in sol}
            err = {k: {} for k in sol if k != "fp64"}
            for n in range(1, steps + 1):
                for k, g in sol.items():
                    t0 = time.perf_counter()
                    g.integrate()
                    g.synchronize()
                    if n > 1:
                        elapsed[k] += time.perf_counter() - t0
                if n in checkpoints:
                    ref = fields(sol["fp64"])
                    for k in err:
                        f = fields(sol[k])
                        err[k][str(n)] = {m: float(np.max(np.abs(f[m] - ref[m])) /
                                                   max(float(np.max(np.abs(ref[m]))), 1e-300)) for m in names}
                        del f
                    del ref
                    print(f"[bench] precision workload step {n}: mixed {err['mixed'][str(n)]}, "
                          f"twin {err['fp64_twin'][str(n)]}", file=sys.stderr, flush=True)
        finally:
            for g in sol.values():
                g.close()
    for k in sol:
        if steps > 1:
            v = cells * (steps - 1) / elapsed[k] / 1e6
            out[k] = {"value": round(v, 3), "unit": "Mcell-steps/s",
                      "ms_per_step": round(elapsed[k] / (steps - 1) * 1e3, 4)}
    out["error_vs_fp64"] = err
    if steps > 1:
        out["vs_fp64"] = round(out["mixed"]["value"] / out["fp64"]["value"], 4)
        print(f"[bench] precision workload: fp64 {out['fp64']['value']:.1f}, mixed {out['mixed']['value']:.1f} "
              f"Mcell-steps/s", file=sys.stderr, flush=True)
    return out


def batched_form(sva, ncases, steps, autotune=2):
    """`ncases` independent 2D cases of uneven sizes (nx in {64..256},
    ny = 2 nx) advanced together, against the same cases stepped one after
    another with integrate()."""
    from simplevofsolver_amd.ic import batch_2d_cases
    t0 = time.perf_counter()
    cases = batch_2d_cases(ncases, seed=4)
    print(f"[bench] {ncases} batch cases generated in {time.perf_counter() - t0:.1f} s", file=sys.stderr,
          flush=True)
    cells = sum(c.glsizes[0] * c.glsizes[1] for c in cases)
    solvers = []
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as d:
        for c, ic in enumerate(cases):
            solvers.append(sva.solver_from_ic(ic, os.path.join(d, f"c{c}")))
    try:
        # one after another (each integrate() waits for its own dt)
        for s in solvers:
            s.integrate()
        t0 = time.perf_counter()
        nseq = max(1, steps // 5)
        for _ in range(nseq):
            for s in solvers:
                s.integrate()
        for s in solvers:
            s.synchronize()
        seq = cells * nseq / (time.perf_counter() - t0) / 1e6
        b = sva.Batch(solvers)
        try:
            b.integrate()
            b.synchronize()
            tuned = None
            if autotune > 0:
                ta = time.perf_counter()
                rep = b.autotune(autotune)
                tuned = {k[5:]: v for k, v, _ in rep if k.startswith("best:")}
                print(f"[bench] batch autotune ({time.perf_counter() - ta:.1f} s): {tuned}", file=sys.stderr,
                      flush=True)
            t0 = time.perf_counter()
            for _ in range(steps):
                b.integrate()
            b.synchronize()
            el = time.perf_counter() - t0
            # the FP32-Poisson variant of the same batch (FP32 transforms, FP64
            # x eliminations, one FP64 refinement): the phases are captured anew
            for s in solvers:
                s.set_poisson_mixed(1)
            b.integrate()
            b.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                b.integrate()
            b.synchronize()
            el32 = time.perf_counter() - t0
            for s in solvers:
                s.set_poisson_mixed(-1)
        finally:
            b.close()
    finally:
        for s in solvers:
            s.close()
    v = cells * steps / el / 1e6
    v32 = cells * steps / el32 / 1e6
    print(f"[bench] batched form {v:.1f} Mcell-steps/s ({el / steps * 1e3:.3f} ms/step), "
          f"one by one {seq:.1f}; FP32-Poisson variant {v32:.1f}", file=sys.stderr, flush=True)
    return {"value": round(v, 3), "unit": "Mcell-steps/s", "ms_per_step": round(el / steps * 1e3, 4),
            "sequential_value": round(seq, 3), "vs_sequential": round(v / seq, 3), "cases": ncases,
            "cells": cells, "steps": steps, "autotuned": tuned,
            "parts": 4 if ncases >= 16 else 2 if ncases >= 8 else 1,
            "fp32_poisson": {"value": round(v32, 3), "ms_per_step": round(el32 / steps * 1e3, 4),
                             "vs_fp64_batch": round(v32 / v, 4), "refinements": 1},
            "workload": f"{ncases} 2D cases, nx in {{64,96,128,192,256}}, ny = 2 nx, 1-4 droplets, "
                        "Re in [50,500], We in [5,50], FP64"}


def adjoint_form(sva, nsteps=50, every=10, ndirs=32, n=256):
    """Differentiation workload: 2D n x n, one droplet r = 0.2 in a
    divergence-free velocity (rms 0.05); loss = centroid x + interface
    length after nsteps; gradient w.r.t. the initial velocity by the adjoint
    with checkpoints every `every` steps, timed against the forward, and
    checked against central differences (eps = 1e-6) along `ndirs` seeded
    divergence-free directions of unit rms with a frozen dt sequence: of the
    map the adjoint differentiates (the unperturbed run's discrete decisions
    replayed, flat corners held) and, for reference, of the free step; with
    surface tension (We = 10) and without (We -> infinity)."""
    import numpy as np
    from simplevofsolver_amd.ic import divfree_direction, droplet_2d_divfree
    out = {"directions": ndirs, "steps": nsteps, "checkpoint_every": every, "eps": 1e-6,
           "workload": f"2D {n}x{n}, droplet r=0.2, divergence-free velocity rms 0.05, Re=100, loss = centroid x "
                       f"+ interface length after {nsteps} steps, FP64; directions divergence-free, unit rms"}
    mx = np.zeros((n + 2, n + 1), dtype=bool)
    mx[1:-1, 1:-1] = True
    my = np.zeros((n + 2, n + 2), dtype=bool)
    my[1:-1, 1:-1] = True
    mx, my = mx.ravel(), my.ravel()
    for tag, We in (("We10", 10.0), ("no_tension", 1e12)):
        ic = droplet_2d_divfree(nx=n, ny=n, radius=0.2, urms=0.05, nmodes=8, seed=3, We=We)
        with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as d:
            ref = sva.solver_from_ic(ic, os.path.join(d, "ref"))
            s = sva.solver_from_ic(ic, os.path.join(d, "adj"))
        try:
            dts = [ref.integrate() for _ in range(nsteps)]
        finally:
            ref.close()
        try:
            state = tuple(s.get(k) for k in ("ux", "uy", "p", "vof"))

            def set_state(vx=None, vy=None, eps=0.0):
                s.set("ux", state[0] + eps * vx if vx is not None else state[0])
                s.set("uy", state[1] + eps * vy if vy is not None else state[1])
                s.boundaries("ux")
                s.boundaries("uy")
                s.set("p", state[2])
                s.set("vof", state[3])

            def loss_at(vx, vy, eps, mode):
                set_state(vx, vy, eps)
                s.set_freeze(mode)
                for dt in dts:
                    s.integrate_fixed_dt(dt)
                s.set_freeze(0)
                return s.adjoint_loss()

            set_state()
            s.adjoint_gradient(dts, every=every)          # warm-up
            set_state()
            s.synchronize()
            t0 = time.perf_counter()
            for dt in dts:
                s.integrate_fixed_dt(dt)
            s.synchronize()
            t_fwd = time.perf_counter() - t0
            lf = s.adjoint_loss()
            set_state()
            s.synchronize()
            t0 = time.perf_counter()
            loss, gux, guy = s.adjoint_gradient(dts, every=every)
            t_adj = time.perf_counter() - t0
            loss_at(None, None, 0.0, 1)                     # record the decisions
            rng = np.random.default_rng(32)
            errs = {"frozen": [], "free": []}
            for _ in range(ndirs):
                vx, vy = divfree_direction(ic, rng)
                r = np.sqrt((np.sum(vx[mx] ** 2) + np.sum(vy[my] ** 2)) / (mx.sum() + my.sum()))
                vx, vy = vx / r, vy / r
                ad = float(gux @ vx + guy @ vy)
                for kind, mode in (("frozen", 2), ("free", 0)):
                    fd = (loss_at(vx, vy, 1e-6, mode) - loss_at(vx, vy, -1e-6, mode)) / 2e-6
                    errs[kind].append(abs(ad - fd) / max(abs(ad), 1e-300))
        finally:
            s.close()
        rec = {"adjoint_over_forward": round(t_adj / t_fwd, 3), "forward_ms": round(t_fwd * 1e3, 3),
               "adjoint_ms": round(t_adj * 1e3, 3), "loss": loss, "loss_forward": lf}
        for kind, v in errs.items():
            v = np.array(v)
            rec[f"fd_rel_err_{kind}"] = {"median": float(np.median(v)), "max": float(np.nanmax(v)),
                                         "within_1e-6": int((v <= 1e-6).sum())}
        out[tag] = rec
        print(f"[bench] adjoint ({tag}) {t_adj * 1e3:.1f} ms / forward {t_fwd * 1e3:.1f} ms = {t_adj / t_fwd:.2f}; "
              f"FD rel err (eps 1e-6) frozen map: median {np.median(errs['frozen']):.1e}, "
              f"{rec['fd_rel_err_frozen']['within_1e-6']}/{ndirs} within 1e-6; free step: median "
              f"{np.median(errs['free']):.1e}, {rec['fd_rel_err_free']['within_1e-6']}/{ndirs}",
              file=sys.stderr, flush=True)
    out["adjoint_over_forward"] = out["We10"]["adjoint_over_forward"]
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=256, help="cells per direction (3D cube)")
    ap.add_argument("--cpu-reps", type=int, default=3, help="timed oracle steps per CPU-baseline process")
    ap.add_argument("--cpu-workers", type=int, default=0,
                    help="concurrent CPU-baseline processes (0 = one per host CPU, at most 16)")
    ap.add_argument("--cpu-worker", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--cpu-sync", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--schedule", type=int, default=None)
    ap.add_argument("--fusion", type=int, default=None,
                    help="vof_set_fusion mask of the fused schedule (64: one-kernel predictor)")
    ap.add_argument("--kernels", action="store_true", help="print the per-kernel table to stderr")
    ap.add_argument("--poisson-mixed", type=int, default=-1, metavar="K",
                    help="time the mixed-precision form (FP32 transforms + K FP64 refinements) "
                         "instead of the plain FP64 step")
    ap.add_argument("--tuning", action="append", default=[], metavar="NAME=VALUE",
                    help="tile parameter passed to vof_set_tuning (fft_elems, lines_lpb, xcd_bands, rhs_rows)")
    ap.add_argument("--forms", default="mixed,autotuned,batched,precision,adjoint",
                    help="comma-separated forms to run after the plain step")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the mixed-precision and batched forms")
    ap.add_argument("--autotune", type=int, default=3, metavar="K",
                    help="steps per candidate of the autotuner for the autotuned and batched forms (0 = off)")
    ap.add_argument("--precision-steps", type=int, default=100,
                    help="steps of the 512x256x256 precision workload (0 = skip)")
    ap.add_argument("--precision-refine", type=int, default=2,
                    help="FP64 refinements of its mixed-precision run")
    ap.add_argument("--drift-steps", type=int, default=200,
                    help="steps of the 3D workload after which its drift is reported")
    ap.add_argument("--drift-2d-steps", type=int, default=1000,
                    help="steps of the 2D 128x256 droplet workload for its drift report (0 = skip)")
    ap.add_argument("--batch-cases", type=int, default=96)
    ap.add_argument("--batch-steps", type=int, default=30)
    args = ap.parse_args(argv)
    if args.cpu_worker:
        return cpu_worker(args.n, args.cpu_reps, args.cpu_sync, max(1, args.cpu_workers))
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU; --gpus {args.gpus} is not supported "
              f"(the multi-GPU sharding is a design note, not code)", file=sys.stderr)
        return 2

    # the CPU baseline runs first, in child processes started before this
    # process touches the GPU
    cpu = None
    if not args.no_cpu_baseline:
        print("[bench] timing the CPU oracle", file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.n, args.cpu_reps, args.cpu_workers or None)

    import simplevofsolver_amd as sva
    from simplevofsolver_amd.ic import droplets_3d

    n = args.n
    t_ic = time.perf_counter()
    ic, spheres = droplets_3d(n, n, n, seed=2)
    t_ic = time.perf_counter() - t_ic
    print(f"[bench] IC {n}^3 generated in {t_ic:.1f} s", file=sys.stderr, flush=True)
    with tempfile.TemporaryDirectory(dir=os.environ.get("TMPDIR", "/tmp")) as d:
        g = sva.solver_from_ic(ic, d, schedule=args.schedule)
    if args.fusion is not None:
        g.set_fusion(args.fusion)
    for t in args.tuning:
        name, val = t.split("=")
        g.set_tuning(name, int(val))
    if args.poisson_mixed >= 0:
        g.set_poisson_mixed(args.poisson_mixed)
    # the library holds everything on the device from here on
    d0 = g.diagnostics()
    for _ in range(args.warmup):
        g.integrate()
    g.synchronize()
    print(f"[bench] warmup {args.warmup} steps done", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        g.integrate()
    g.synchronize()
    t1 = time.perf_counter()
    el = t1 - t0
    # the same K steps again with HIP events recorded around every launch on
    # the library's stream: per-kernel durations for the roofline (the events
    # themselves are kept out of the throughput timing above)
    g.profile(True)
    g.profile_reset()
    g.synchronize()
    tp0 = time.perf_counter()
    for _ in range(args.steps):
        g.integrate()
    g.synchronize()
    el_prof = time.perf_counter() - tp0
    rep = g.profile_report()
    g.profile(False)
    d1 = g.diagnostics()
    # drift over the workload's length: continue to `drift_steps` steps
    steps_done = args.warmup + 2 * args.steps
    for _ in range(max(0, args.drift_steps - steps_done)):
        g.integrate()
    d2 = g.diagnostics()
    steps_drift = max(steps_done, args.drift_steps)
    # interface cells of the last state (their count moves by well under 1 %
    # over these steps): the cells whose normals the interface kernels write
    vint = _interior("vof", g.get("vof"), n, n, n)
    n_iface = int(((vint >= 1e-8) & (vint <= 1.0 - 1e-8)).sum())
    del vint
    cells = n * n * n
    value = cells * args.steps / el / 1e6
    kern_total = sum(ms for _, ms in rep.values())
    dom = max(rep.items(), key=lambda kv: kv[1][1])
    kname, (kcnt, kms) = dom
    kavg_s = kms / kcnt * 1e-3
    kbytes = algorithmic_bytes(kname, n, n, n, n_iface)
    achieved = kbytes / kavg_s / 1e9 if kbytes else None
    traffic = read_traffic(kname, f"3d{n}")
    copy_gbs = g.copy_probe(1 << 27, 10)
    forms = {}
    want = set() if args.no_forms else set(args.forms.split(","))
    plain_now = value
    if want & {"mixed", "autotuned"} and args.poisson_mixed < 0:
        plain_now = plain_rate(g, cells, args.steps)
        forms["plain_same_state"] = {"value": round(plain_now, 3), "unit": "Mcell-steps/s",
                                     "steps_before": steps_drift + 1,
                                     "note": "the plain step re-timed at the state the mixed and autotuned forms "
                                             "start from; their vs_plain is against this value"}
        print(f"[bench] plain step after the drift run: {plain_now:.1f} Mcell-steps/s", file=sys.stderr,
              flush=True)
    if "mixed" in want and args.poisson_mixed < 0:
        try:
            forms["mixed_precision"] = mixed_form(g, cells, args.steps, plain_now, 1)
            forms["mixed_precision_fp32_solve"] = mixed_form(g, cells, args.steps, plain_now, 0)
        except Exception as e:  # reported, never fatal for the plain line
            forms["mixed_precision"] = {"error": repr(e)}
    if "autotuned" in want and args.poisson_mixed < 0 and args.autotune > 0:
        try:
            forms["autotuned"] = autotuned_form(g, cells, args.steps, plain_now, args.autotune)
        except Exception as e:
            forms["autotuned"] = {"error": repr(e)}
    if args.kernels:
        print(f"[bench] kernels (profiled pass {el_prof * 1e3:.2f} ms, kernel sum {kern_total:.2f} ms; "
              f"timed pass {el * 1e3:.2f} ms):", file=sys.stderr)
        for k, (cnt, ms) in sorted(rep.items(), key=lambda kv: -kv[1][1]):
            b = algorithmic_bytes(k, n, n, n, n_iface)
            bw = f"{b / (ms / cnt * 1e-3) / 1e9:8.1f} GB/s" if b else ""
            print(f"  {k:18s} {cnt:5d} x {ms / cnt * 1e3:9.1f} us  total {ms:9.2f} ms {bw}",
                  file=sys.stderr)
        print(f"[bench] copy probe {copy_gbs:.0f} GB/s", file=sys.stderr, flush=True)
    g.close()
    if "batched" in want:
        try:
            # a batched step takes ~2 ms: the tuner times 10 steps per candidate
            forms["batched"] = batched_form(sva, args.batch_cases, args.batch_steps,
                                            max(args.autotune, 10) if args.autotune > 0 else 0)
        except Exception as e:
            forms["batched"] = {"error": repr(e)}
    if "precision" in want and args.precision_steps > 0:
        try:
            forms["precision"] = precision_form(sva, args.precision_steps, args.precision_refine)
        except Exception as e:
            forms["precision"] = {"error": repr(e)}
    if "adjoint" in want:
        try:
            forms["adjoint"] = adjoint_form(sva)
        except Exception as e:
            forms["adjoint"] = {"error": repr(e)}
    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": f"plain 3D {n}x{n}x{n} FP64 droplets (8 spheres r in [0.06,0.12], "
                        f"16-mode divergence-free velocity rms 0.1, Re=500, We=20), x walls, y/z periodic",
            "cells": cells,
            "form": "plain" if args.poisson_mixed < 0 else f"mixed-precision Poisson, {args.poisson_mixed} refinement(s)",
        },
        "roofline": {
            "bound": "hbm",
            "kernel": kname,
            "achieved": round(achieved, 1) if achieved else None,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
            "traffic": traffic,
            "algorithmic_bytes_per_launch": kbytes,
            "interface_cells": n_iface,
            "mean_launch_us": round(kavg_s * 1e6, 2),
            "profiled_ms_per_step": round(el_prof / args.steps * 1e3, 4),
            "copy_probe_gbs": round(copy_gbs, 1),
        },
        "step_bytes_per_cell": 960,
        "step_hbm_frac": round(960 * value * 1e6 / 1e9 / HBM_PEAK_GBS, 4),
        "drift": drift_record(d0, d2, steps_drift, ndims=3),
    }
    out["forms"] = forms or None
    if args.drift_2d_steps > 0:
        try:
            out["drift_2d"] = drift_2d(sva, args.drift_2d_steps)
        except Exception as e:
            out["drift_2d"] = {"error": repr(e)}
    out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
