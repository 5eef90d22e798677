"""Stencil-line rows of the symmetric-slice SpMV, FP64 and FP32.

On a structured grid every row's offsets are the lines c - 1, c, c + 1
around the centres 0, +-c_m.  The line form loads each line of the vector
once per wavefront and takes the neighbours by wave shifts; the generic row
gathers every offset.  Both sum a row in the same order, so a solve gives the
same bits either way: FP64 PCG on the FP64 symmetric slices and the mixed
solve (FP32 sweeps, FP64 replacement) are run with the line form (default)
and without it (INS_SYM_LINES=0, read once when the library loads), each in a
process of its own, and compared bit for bit.  The solves themselves are
checked against the oracle elsewhere (tests/test_gpu_parity.py)."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

RUN = r"""
import hashlib, json, sys
import numpy as np
from insilico_amd import abi, problems as P
kind, n = sys.argv[1], int(sys.argv[2])
prob = P.diffusion3d(n, seed=21) if kind == "3d" else P.poisson2d(n)
out = {}
with abi.Context(0) as c:
    c.load_problem(prob)
    c.assemble(abi.FP64, abi.ASM_MATRIX | abi.ASM_RHS)
    for name, mode in (("fp64", abi.SOLVE_FP64), ("mixed", abi.SOLVE_MIXED)):
        it, rr, _ = c.solve(mode, abi.FMT_SELL, rtol=1e-9)
        x = np.ascontiguousarray(c.get_solution(), dtype=np.float64)
        out[name] = [it, rr, hashlib.sha256(x.tobytes()).hexdigest()]
print(json.dumps(out), flush=True)
"""

# (grid, size): 3-D 27-point rows (14 stored offsets), 2-D 9-point rows (5);
# the 64^3 grid is large enough for the per-XCD grid of the symmetric SpMV,
# whose block count is fixed for both runs so that the partial sums match
CASES = [("3d", 12, None), ("2d", 40, None), ("3d", 64, "768")]


def _run(kind, n, lines, grid):
    env = dict(os.environ)
    env["INS_SYM_LINES"] = "1" if lines else "0"
    if grid:
        env["INS_SYM_GRID"] = grid
    r = subprocess.run([sys.executable, "-c", RUN, kind, str(n)], cwd=ROOT, env=env,
                       capture_output=True, text=True, timeout=110)
    assert r.returncode == 0, r.stderr[-2000:]
    return json.loads(r.stdout.strip().splitlines()[-1])


@pytest.mark.parametrize("kind,n,grid", CASES)
def test_line_rows_give_the_generic_rows_bits(kind, n, grid):
    print("case", kind, n, "grid", grid, flush=True)
    on = _run(kind, n, True, grid)
    off = _run(kind, n, False, grid)
    for name in ("fp64", "mixed"):
        assert on[name][1] <= 1e-9, (name, on[name])
        assert on[name] == off[name], (name, on[name], off[name])
