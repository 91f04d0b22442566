"""bench.py on the host: the one-GPU rule of its command line, and its
algorithmic byte count against the per-DOF-update figures and the 32^3 per-step
bytes that define the roofline fraction (152 B FP64, 76 B FP32, 112 B FP32
fields with FP64 LSRK registers; 7.65 GB per FP32 step at 32^3, n1 = 8)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_bytes_per_dof_update():
    npts = 1000
    for s_u, s_k, per in ((8, 8, 152), (4, 4, 76), (4, 8, 112)):
        assert bench.algorithmic_bytes_per_step(npts, 0, s_u, s_k) == 6 * npts * per


def test_per_element_term_and_32cubed_step():
    # five stages, each reading 24 FP64 trilinear coefficients plus 1/eps, 1/mu per element
    assert bench.algorithmic_bytes_per_step(0, 10, 4, 4) == 5 * 10 * 208
    b = bench.algorithmic_bytes_per_step(32 ** 3 * 512, 32 ** 3, 4, 4)
    assert abs(b / 7.65e9 - 1) < 0.01, b
