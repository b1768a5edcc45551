"""Host-side pieces of bench.py (no GPU): the algorithmic bytes per launch
that the roofline divides by, the launch counts of a profiled pass, and the
one-GPU guard."""
import os
import subprocess
import sys
from types import SimpleNamespace

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _prob(dim, C, Nz, N, na, reg="TGV"):
    return SimpleNamespace(dim=dim, C=C, Nz=Nz, na=na, nd=N, regularisation=reg,
                           voxels=lambda: C * Nz * N * N)


def test_kernel_bytes_match_the_per_voxel_figures():
    # 256x256x64 x 4 channels, 60 tilts, 256 bins
    V, B = 4 * 64 * 256 * 256, 4 * 64 * 60 * 256
    kb = bench.kernel_bytes(_prob(3, 4, 64, 256, 60))
    # dual: read ubar 1, pbar 3, v 3, w 6 floats; write v 3, w 6: 88 B
    assert kb["dual_single_kernel"] == 88 * V == 1476395008
    # primal: read u, K^T lam, v 3, w 6, p 3; write u, ubar, p 3, pbar 3: 88 B
    assert kb["primal_single_kernel"] == 88 * V
    # forward + prox: ubar per voxel; lam, f read and lam written per bin
    assert kb["fwd_single_kernel"] == 4 * V + 12 * B == 114294784
    # backprojection: lam per bin in, K^T lam per voxel out
    assert kb["adj_single_kernel"] == 4 * V + 4 * B == 82837504
    # the two stencil steps together: 176 B per voxel-channel against the
    # compulsory 136 B (each re-reads state the other wrote)
    assert kb["dual_single_kernel"] + kb["primal_single_kernel"] == 176 * V


def test_kernel_bytes_tv_and_2d():
    V = 1 * 1 * 183 * 183
    kb = bench.kernel_bytes(_prob(2, 1, 1, 183, 90))
    # 2D TGV dual: ubar 1, pbar 2, v 2, w 3 read; v 2, w 3 written
    assert kb["dual_single_kernel"] == 4 * V * 13
    kt = bench.kernel_bytes(_prob(3, 2, 8, 32, 10, reg="TV"))
    Vt = 2 * 8 * 32 * 32
    # TV: no p, w; dual reads ubar and v, writes v
    assert kt["dual_single_kernel"] == 4 * Vt * 7


def test_launch_counts():
    assert bench.launch_counts(5, False) == (5, 5, 5, 5, 0)
    # fused form: an even number of fused steps after the first dual step
    assert bench.launch_counts(6, True) == (2, 6, 6, 2, 4)
    assert bench.launch_counts(5, True) == (1, 5, 5, 1, 4)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
