"""bench.py's arithmetic and its one-GPU contract, on the CPU.

The roofline's algorithmic work is the canonical count of field multiplications per key
(7 per extended + affine addition over rows - 2 comb rows, 3 for the last row's fraction,
4 for the batched inversion, and the 265-multiplication inversion shared by the keys of a
lane's chain); the serving percentiles are nearest-rank over the sorted times.
"""
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_canonical_multiplication_counts():
    # W = 16: 16 rows -> 14 additions; W = 24: 11 rows -> 9 additions
    assert bench.field_mults_per_key(16, 256) == 7 * 14 + 3 + 4 + 265 / 256
    assert abs(bench.field_mults_per_key(16, 256) - 106.04) < 0.005
    assert bench.field_mults_per_key(24, 256) == 7 * 9 + 3 + 4 + 265 / 256
    assert abs(bench.field_mults_per_key(24, 256) - 71.035) < 0.001
    # one inversion per key at K = 1; a 4-wave pool shares one over 4 x K keys, for
    # 3 (pool - 1) more multiplications
    assert bench.field_mults_per_key(24, 1) == 7 * 9 + 3 + 4 + 265
    assert bench.field_mults_per_key(24, 8, pool=4) == 7 * 9 + 3 + 4 + (265 + 9) / 32
    # widths the library accepts: ceil(251 / W) rows
    for w in range(2, 29):
        rows = -(-251 // w)
        assert bench.field_mults_per_key(w, 256) == 7 * (rows - 2) + 7 + 265 / 256


def test_percentile_nearest_rank():
    assert bench.percentile([], 0.5) is None
    assert bench.percentile([3.0], 0.99) == 3.0
    xs = list(range(1, 101))
    assert bench.percentile(xs, 0.5) == 51  # round(0.5 * 99) = 50 -> the 51st value
    assert bench.percentile(xs, 0.99) == 99
    assert bench.percentile(list(reversed(xs)), 0.0) == 1
    assert bench.percentile(xs, 1.0) == 100


def test_more_than_one_gpu_is_refused():
    # one GPU per box: --gpus N > 1 exits with status 2 and says why, before any GPU work
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""
