"""CPU checks of bench.py's contract pieces that need no GPU: the one-GPU
guard, the algorithmic byte counts the roofline divides by, and the
committed traffic table it reads."""
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    res = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                         text=True, timeout=120)
    assert res.returncode == 2
    assert "one GPU" in res.stderr
    assert res.stdout == ""


def test_algorithmic_bytes_per_cell():
    n = 256
    c = n ** 3
    # 88 B/cell every stage + 64 B/cell of carried sources in 2 of 3 stages
    assert abs(bench.algorithmic_bytes("k_rhs_march", n, n, n) / c - (88 + 64 * 2 / 3)) < 1e-9
    # increment read, field read and write of four fields (ux has nx - 1 DOFs per line)
    assert bench.algorithmic_bytes("k_lines_lds", n, n, n) == (4 * n - 1) * n * n * 24
    nh = n // 2 + 1
    assert bench.algorithmic_bytes("k_fft_z", n, n, n) == n * nh * n * 32
    # the spectrum read and written once by either spectral x solver
    for k in ("k_thomas_part", "k_thomas_spec"):
        assert bench.algorithmic_bytes(k, n, n, n) == n * nh * n * 32
    assert bench.algorithmic_bytes("no_such_kernel", n, n, n) is None


def test_interface_kernels_charge_normals_per_interface_cell():
    """The interface kernel reads vof and writes curvature at every cell, and
    writes the 4-component 3D normal only at interface cells; the VOF update
    reads velocity (3) and vof per cell, the normals of interface cells, and
    writes the new vof (+ 16 B of carried sources in 2 of 3 stages)."""
    n, m = 256, 123457
    c = n ** 3
    for k in ("k_iface_march", "k_iface_fused"):
        assert bench.algorithmic_bytes(k, n, n, n, m) == 16 * c + 32 * m
        # without a count, every cell is charged (an upper bound)
        assert bench.algorithmic_bytes(k, n, n, n) == 48 * c
    assert abs(bench.algorithmic_bytes("k_vof_fused", n, n, n, m) - (40 * c + 32 * m + 16 * c * 2 / 3)) < 1e-6
    # the compulsory bytes of the benchmark state lie below what the kernel was
    # measured to move (0.38 GB per launch at 256^3, a few % interface cells)
    assert bench.algorithmic_bytes("k_iface_march", n, n, n, int(0.03 * c)) < 0.38e9


def test_traffic_table_is_read():
    t = bench.read_traffic("k_rhs_march", "3d256")
    assert t is not None and 2e9 < t < 4e9
    assert bench.read_traffic("k_rhs_march", "3d512") is None
