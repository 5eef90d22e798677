"""Host-side pieces of bench.py's measurement line: the algorithmic bytes of
the SpMV storage formats, the HBM traffic read back from the committed
counter summary, and the one-GPU guard.  No GPU is touched."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402

# 3-D diffusion 128^3 (u = 0 on x = 0): active rows, CSR entries, values of
# the symmetric slices (upper triangle at 14 shared offsets, 1.5% padding)
N_ROWS = 2130048
NNZ = 56621950
SYM_VALUES = 29820672


def test_symmetric_slice_bytes_per_launch():
    # values once + p read + q write, 4-byte values and vectors
    assert bench.spmv_bytes(N_ROWS, NNZ, SYM_VALUES, "sym", 4, 4) == 136323072


def test_sell_and_csr_bytes_count_their_metadata():
    n, nnz, vals = 1000, 27000, 28000
    # SELL: padded values + one pattern id per row + p read + q write
    assert bench.spmv_bytes(n, nnz, vals, "sell", 4, 4) == vals * 4 + n * 4 + 2 * n * 4
    # CSR: values and column indices + row pointers + p read + q write
    assert bench.spmv_bytes(n, nnz, vals, "csr", 8, 8) == nnz * 12 + n * 4 + 2 * n * 8


def test_traffic_comes_from_the_committed_counter_summary():
    t = bench.read_traffic("diffusion128", "mixed_sym")
    assert t == 146220841.0
    # measured traffic is within 1.2x of the algorithmic bytes of the launch
    alg = bench.spmv_bytes(N_ROWS, NNZ, SYM_VALUES, "sym", 4, 4)
    assert 1.0 < t / alg < 1.2
    assert bench.read_traffic("no_such_workload", "mixed_sym") is None


def test_more_than_one_gpu_is_refused_before_any_device_work():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""
