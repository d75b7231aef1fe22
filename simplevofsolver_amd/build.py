"""Build the gfx950 libraries in-tree: simplevofsolver_amd/libvofhip{2,3}d.so.

One library per dimensionality because the C structs of the reference's
interface are NDIMS-sized.  Flags:

* ``--offload-arch=gfx950``: MI355X only.
* ``-ffp-contract=off``: no FMA contraction, so that the rounding of every
  stencil matches the reference's ``-std=c99`` build.
* ``-Xclang -target-feature -Xclang -packed-fp32-ops``: keeps packed FP32
  instructions out of the generated code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")

SOURCES = [os.path.join(CSRC, "vof_lib.hip")]
DEPS = SOURCES + [os.path.join(CSRC, "vof_kernels.h"), os.path.join(INCLUDE, "vof_solver.h")]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def lib_path(ndims: int) -> str:
    return os.path.join(PKG, f"libvofhip{ndims}d.so")


def compile_cmd(ndims: int, out: str) -> list[str]:
    return [
        hipcc(), "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
        f"-DNDIMS={ndims}", "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wall", "-Wno-unused-function",
        "-I", INCLUDE, "-I", CSRC,
        *SOURCES, "-o", out,
    ]


def driver_path(ndims: int) -> str:
    return os.path.join(PKG, f"vofdrive{ndims}d")


def build_driver(ndims: int, verbose: bool = False) -> str:
    """The step-count driver (csrc/vof_driver.c), a plain C99 program linked
    against the library of the same dimensionality."""
    out = driver_path(ndims)
    cmd = ["gcc", "-std=c99", "-O2", "-Wall", "-Wextra", f"-DNDIMS={ndims}", "-D_POSIX_C_SOURCE=200809L",
           "-I", INCLUDE, os.path.join(CSRC, "vof_driver.c"), "-L", PKG, f"-lvofhip{ndims}d", "-lm",
           f"-Wl,-rpath,$ORIGIN", "-o", out]
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"driver build failed for NDIMS={ndims}")
    return out


def build(ndims_list=(2, 3), verbose: bool = False) -> list[str]:
    """Compile every library and driver (always, also where built)."""
    outs = []
    for nd in ndims_list:
        out = lib_path(nd)
        tmp = out + ".tmp"
        cmd = compile_cmd(nd, tmp)
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            sys.stderr.write(res.stdout + res.stderr)
            raise RuntimeError(f"hipcc failed for NDIMS={nd}")
        os.replace(tmp, out)
        outs.append(out)
        outs.append(build_driver(nd, verbose))
    return outs


if __name__ == "__main__":
    for p in build(verbose=True):
        print("built", p)
