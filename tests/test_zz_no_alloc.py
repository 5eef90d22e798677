"""The allocation-free entry point ins_solve_csr_noalloc against the
existing ins_solve(INS_SOLVE_FP64, INS_FMT_CSR) on the same systems: two calls
in a row, the second on a smaller system (its scratch array keeps the first
call's contents), each giving the bits, iteration count and residual of
ins_solve, and a residual the oracle's FP64 SpMV confirms."""
import ctypes as C

import numpy as np
import pytest

from insilico_amd import abi, problems as P

pytestmark = pytest.mark.gpu


def _noalloc():
    L = abi.load()
    f = L.ins_solve_csr_noalloc
    f.argtypes = [C.c_int, C.c_void_p, C.c_int64, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                  C.c_void_p, C.c_void_p, C.c_double, C.c_int, C.POINTER(C.c_int),
                  C.POINTER(C.c_double)]
    f.restype = C.c_int
    return f


def _reference_solve(prob, rtol):
    with abi.Context(0) as c:
        c.load_problem(prob)
        c.assemble(abi.FP64, abi.ASM_MATRIX | abi.ASM_RHS)
        rp, ci = c.get_pattern()
        val, b = c.get_csr_values(), c.get_rhs()
        it, rel, st = c.solve(abi.SOLVE_FP64, abi.FMT_CSR, rtol)
        assert st == abi.INS_OK
        x = c.get_solution()
    return rp, ci, val, b, x, it, rel


class _Hip:
    """Device arrays of the caller, from the HIP runtime the library itself
    runs on (its libamdhip64.so.7, already loaded with it): one runtime in
    the process."""

    def __init__(self):
        abi.load()
        self.L = C.CDLL("libamdhip64.so.7")
        self.L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipSetDevice.argtypes = [C.c_int]
        assert self.L.hipSetDevice(0) == 0
        self.held = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(a.nbytes, 1)) == 0
        self.held.append(p)
        assert self.L.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0   # host to device
        return p.value

    def get(self, p, n, dtype):
        out = np.empty(n, dtype)
        assert self.L.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0   # device to host
        return out

    def free(self):
        for p in self.held:
            self.L.hipFree(p)
        self.held = []


def test_noalloc_solve_equals_ins_solve(oracle):
    f = _noalloc()
    hip = _Hip()
    rtol = 1e-10
    try:
        # a larger scalar system first, then a smaller vector one
        for prob in (P.diffusion3d(22, seed=11), P.elasticity3d(7, seed=4)):
            rp, ci, val, b, x_ref, it_ref, rel_ref = _reference_solve(prob, rtol)
            n, nnz = b.shape[0], ci.shape[0]
            print("noalloc case %s n=%d nnz=%d" % (prob.name, n, nnz), flush=True)
            d_rp, d_ci, d_val, d_b = hip.put(rp), hip.put(ci), hip.put(val), hip.put(b)
            d_x = hip.put(np.full(n, np.nan))
            it, rel = C.c_int(), C.c_double()
            st = f(0, None, n, nnz, d_rp, d_ci, d_val, d_b, d_x, rtol, 2 * n, C.byref(it),
                   C.byref(rel))
            assert st == abi.INS_OK, abi.ERRORS.get(st, st)
            x = hip.get(d_x, n, np.float64)
            assert it.value == it_ref and rel.value == rel_ref
            assert np.array_equal(x, x_ref)
            r = b - oracle.spmv(rp, ci, val, x)
            assert np.linalg.norm(r) <= 1.001 * rtol * np.linalg.norm(b)
            hip.free()
    finally:
        hip.free()


def test_noalloc_rejects_bad_sizes():
    f = _noalloc()
    it, rel = C.c_int(), C.c_double()
    big = abi.load().ins_strerror  # keep the library loaded
    assert big is not None
    # beyond the scratch array: rejected before any device access
    st = f(0, None, (1 << 20) + 1, 27 << 20, 1, 1, 1, 1, 1, 1e-8, 10, C.byref(it), C.byref(rel))
    assert st == abi.ERRORS_INV["ARG"]
    st = f(0, None, 0, 0, 1, 1, 1, 1, 1, 1e-8, 10, C.byref(it), C.byref(rel))
    assert st == abi.ERRORS_INV["ARG"]
    st = f(0, None, 100, 100, None, None, None, None, None, 1e-8, 10, C.byref(it), C.byref(rel))
    assert st == abi.ERRORS_INV["ARG"]
