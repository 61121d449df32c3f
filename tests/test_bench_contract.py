"""The benchmark's contract on the CPU: the algorithmic byte counts behind
roofline.achieved, and the one-GPU rule.

The byte model is checked against the worked figures for the batched
workload (N = 200k, K = 192, p = 8, 16 RHS, FP32): spreading and
interpolation 1,418.9 MB, each transform direction 2,732 MB (real mesh
K^3 C s plus half spectrum (K/2+1) K^2 C 2s).
"""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_algorithmic_bytes_match_the_worked_figures():
    sys.path.insert(0, ROOT)
    import bench
    N, K, p, R, s = 200_000, 192, 8, 16, 4
    b = bench.algorithmic_bytes(N, K, p, R, s, real_entries=0)
    assert b["spread"] / 1e6 == pytest.approx(1418.9, abs=0.2)
    assert b["interp"] == b["spread"]
    assert b["fft_x_fwd"] / 1e6 == pytest.approx(2732.1, abs=0.2)
    assert b["fft_x_inv"] == b["fft_x_fwd"]
    assert b["fft_total"] == 2 * b["fft_x_fwd"]
    # the z pass reads and writes the half spectrum once each
    assert b["fft_z_infl"] == 2 * (K // 2 + 1) * K * K * 3 * R * 2 * s
    # real space: 2 coefficients + a column index per stored entry
    E = 1000
    assert (bench.algorithmic_bytes(N, K, p, R, s, E)["real"] - b["real"]) == E * (2 * s + 4)


def test_more_than_one_gpu_is_refused_with_a_reason():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_fp64_sum_form_counts_double_outputs():
    """With the final sum in FP64 the output vectors are 8-byte values: the
    interpolation writes 3NR more 4-byte words and the real-space pass reads
    and writes 2 x 3NR more."""
    sys.path.insert(0, ROOT)
    import bench
    N, K, p, R = 200_000, 192, 8, 16
    a = bench.algorithmic_bytes(N, K, p, R, 4, 1000)
    b = bench.algorithmic_bytes(N, K, p, R, 4, 1000, so=8)
    assert b["interp"] - a["interp"] == 3 * N * R * 4
    assert b["real"] - a["real"] == 2 * 3 * N * R * 4
    assert b["spread"] == a["spread"] and b["fft_total"] == a["fft_total"]
