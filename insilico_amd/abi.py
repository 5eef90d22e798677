"""ctypes binding of the C ABI (include/insilico_hip.h) and a thin Python
mirror of the reference's SOLVER-concept workflow.

The library is loaded from ``insilico_amd/libinsilico_hip.so`` (built in
tree).  If it is missing, loading raises: there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libinsilico_hip.so")

INS_OK = 0
ERRORS = {-1: "ARG", -2: "HIP", -3: "OOM", -4: "RANGE", -5: "PATTERN",
          -6: "BREAKDOWN", -7: "NOCONV", -8: "STATE", -9: "UNSUPPORTED"}
ERRORS_INV = {v: k for k, v in ERRORS.items()}
LAPLACE, STVENANT, NEOHOOKE = 1, 2, 3
FP64, FP32 = 0, 1
ASM_MATRIX, ASM_RHS, ASM_RESIDUAL, ASM_INCREMENTAL = 1, 2, 4, 8
BF_NONE, BF_CONST, BF_SINPROD, BF_COSPROD, BF_TABLE = 0, 1, 2, 3, 4
SOLVE_FP64, SOLVE_MIXED = 0, 1
FMT_CSR, FMT_SELL = 0, 1
DIST_SET, DIST_ADD = 0, 1
DIRECT_CHOLESKY, DIRECT_LU, DIRECT_MAX_N = 0, 1, 32768

# every symbol include/insilico_hip.h declares
SYMBOLS = [
    "ins_create", "ins_destroy", "ins_strerror", "ins_last_error", "ins_get_stats",
    "ins_set_profiling", "ins_sync", "ins_set_mesh", "ins_set_dofs", "ins_set_prescribed",
    "ins_scale_prescribed", "ins_set_dof_values", "ins_get_dof_values", "ins_build_pattern",
    "ins_get_pattern", "ins_set_material", "ins_set_body_force", "ins_set_traction",
    "ins_assemble", "ins_get_csr_values", "ins_get_rhs", "ins_set_rhs", "ins_add_rhs",
    "ins_rhs_norm", "ins_set_csr", "ins_solve", "ins_get_solution", "ins_solution_norm",
    "ins_distribute", "ins_prepare", "ins_server_create", "ins_server_destroy", "ins_server_last_error",
    "ins_server_attach", "ins_server_step", "ins_server_poll", "ins_server_detach",
    "ins_server_running", "ins_set_field", "ins_set_mass", "ins_error_norm",
    "ins_bdf_init", "ins_bdf_step", "ins_set_traction_profile", "ins_cell_eval",
    "ins_set_neumann_flux", "ins_device_memory", "ins_solve_csr_noalloc",
    "ins_quadrature_rule", "ins_set_body_force_values", "ins_set_traction_values",
    "ins_solve_direct",
]
SLOT_FREE, SLOT_RUNNING, SLOT_FINISHED = 0, 1, 2
REF_ZERO, REF_SINPROD, REF_KELVIN, REF_COSPROD = 0, 1, 2, 3
CELL_VALUE, CELL_GRADIENT, CELL_PK2, CELL_CAUCHY = 0, 1, 2, 3


class Stats(C.Structure):
    _fields_ = [("setup_ms", C.c_double), ("assemble_ms", C.c_double),
                ("solve_ms", C.c_double), ("iterations", C.c_int),
                ("refinements", C.c_int), ("relres", C.c_double),
                ("n_active", C.c_int64), ("nnz", C.c_int64),
                ("device_bytes", C.c_int64), ("spmv_ms", C.c_double),
                ("spmv_calls", C.c_int64), ("sell_values", C.c_int64),
                ("sell_patterns", C.c_int), ("sell_sym_offsets", C.c_int),
                ("rep_ms", C.c_double), ("rep_calls", C.c_int64),
                ("element_ms", C.c_double), ("sym_values", C.c_int64),
                ("sym_block", C.c_int), ("assembled_sym", C.c_int),
                ("sym_lines", C.c_int)]


class InsError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s (%d): %s" % (ERRORS.get(status, "?"), status, msg))
        self.status = status


_lib = None


def load():
    """Load the HIP library; raises if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libinsilico_hip.so not built (run __graft_entry__.build())")
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, dbl = C.c_void_p, C.c_int, C.c_int64, C.c_double
    sig = {
        "ins_create": [C.POINTER(vp), i32, C.c_uint64],
        "ins_destroy": [vp],
        "ins_strerror": [i32],
        "ins_last_error": [vp],
        "ins_get_stats": [vp, C.POINTER(Stats)],
        "ins_set_profiling": [vp, i32],
        "ins_sync": [vp],
        "ins_set_mesh": [vp, i32, i64, vp, i64, i32, vp],
        "ins_set_field": [vp, i32, i64, i32, vp],
        "ins_set_mass": [vp, dbl, dbl, vp],
        "ins_error_norm": [vp, i32, i32, vp, vp, C.POINTER(dbl)],
        "ins_bdf_init": [vp, i32, dbl, vp],
        "ins_cell_eval": [vp, i32, vp, vp, vp],
        "ins_set_neumann_flux": [vp, i32, vp],
        "ins_quadrature_rule": [i32, i32, vp, vp],
        "ins_set_body_force_values": [vp, vp],
        "ins_set_traction_values": [vp, vp],
        "ins_set_traction_profile": [vp, i32, dbl],
        "ins_bdf_step": [vp, dbl, i32, i32, i32, dbl, i32, C.POINTER(i32), C.POINTER(dbl)],
        "ins_set_dofs": [vp, i32, vp, i64],
        "ins_set_prescribed": [vp, vp],
        "ins_scale_prescribed": [vp, dbl],
        "ins_set_dof_values": [vp, vp],
        "ins_get_dof_values": [vp, vp],
        "ins_build_pattern": [vp, C.POINTER(i64)],
        "ins_get_pattern": [vp, vp, vp],
        "ins_set_material": [vp, i32, i32, i32, vp],
        "ins_set_body_force": [vp, i32, vp],
        "ins_set_traction": [vp, i64, i32, vp, vp],
        "ins_assemble": [vp, i32, i32],
        "ins_get_csr_values": [vp, vp],
        "ins_get_rhs": [vp, vp],
        "ins_set_rhs": [vp, vp],
        "ins_add_rhs": [vp, vp],
        "ins_rhs_norm": [vp, C.POINTER(dbl)],
        "ins_set_csr": [vp, i64, vp, vp, vp],
        "ins_solve": [vp, i32, i32, dbl, i32, C.POINTER(i32), C.POINTER(dbl)],
        "ins_solve_direct": [vp, i32, C.POINTER(dbl)],
        "ins_get_solution": [vp, vp],
        "ins_solution_norm": [vp, C.POINTER(dbl)],
        "ins_distribute": [vp, i32],
        "ins_prepare": [vp, i32, i32],
        "ins_server_create": [C.POINTER(vp), i32, i32],
        "ins_server_destroy": [vp],
        "ins_server_last_error": [vp],
        "ins_server_attach": [vp, vp, dbl, i32, C.POINTER(i32)],
        "ins_server_step": [vp, i32, C.POINTER(i32)],
        "ins_server_poll": [vp, i32, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32),
                            C.POINTER(dbl)],
        "ins_server_detach": [vp, i32],
        "ins_server_running": [vp, C.POINTER(i32)],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = (C.c_char_p if name in ("ins_strerror", "ins_last_error",
                                            "ins_server_last_error") else C.c_int)
    _lib = L
    return L


def _p(a):
    return None if a is None else a.ctypes.data


def quadrature_rule(dim, npts):
    """(w, xi): the library's tensor Gauss rule, npts per direction on
    [0, 1]^dim, first coordinate fastest (no device needed)"""
    L = load()
    nq = npts ** dim
    w = np.zeros(nq)
    xi = np.zeros(nq * dim)
    got = L.ins_quadrature_rule(dim, npts, _p(w), _p(xi))
    if got != nq:
        raise InsError(got, "ins_quadrature_rule(%d, %d)" % (dim, npts))
    return w, xi.reshape(nq, dim)


class Context:
    """One device context (owns all device state of one problem)."""

    def __init__(self, device=0, budget_bytes=0):
        self.L = load()
        h = C.c_void_p()
        s = self.L.ins_create(C.byref(h), device, budget_bytes)
        if s != INS_OK:
            raise InsError(s, self.L.ins_strerror(s).decode())
        self.h = h
        self._keep = {}

    def close(self):
        if self.h:
            self.L.ins_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, s):
        if s != INS_OK:
            raise InsError(s, self.L.ins_last_error(self.h).decode())
        return s

    # -- discretisation -----------------------------------------------------
    def set_mesh(self, xyz, conn):
        xyz = np.ascontiguousarray(xyz, np.float64)
        conn = np.ascontiguousarray(conn, np.int32)
        self._chk(self.L.ins_set_mesh(self.h, xyz.shape[1], xyz.shape[0], _p(xyz),
                                      conn.shape[0], conn.shape[1], _p(conn)))
        self.n_nodes, self.dim = xyz.shape
        self.n_elem = conn.shape[0]

    def set_field(self, degree, n_field_nodes, fconn):
        """degree-2 field on the mesh just set (field nodes in hierarchic order)"""
        fc = np.ascontiguousarray(fconn, np.int32)
        self._chk(self.L.ins_set_field(self.h, degree, n_field_nodes, fc.shape[1], _p(fc)))
        self.n_nodes = n_field_nodes

    def set_mass(self, density, mass_factor=0.0, hist=None):
        """time-step terms: matrix + mass_factor * M, RHS + M hist"""
        h = None if hist is None else np.ascontiguousarray(hist, np.float64)
        self._chk(self.L.ins_set_mass(self.h, density, mass_factor, _p(h)))

    def error_norm(self, order, kind, params=None, u=None):
        """|u - u_h| (order 0: L2, order 1: H1 semi-norm) on the device; u_h
        from `u` (per field node and component) or the current DoF values"""
        prm = np.zeros(8)
        if params is not None:
            prm[:len(params)] = params
        uu = None if u is None else np.ascontiguousarray(u, np.float64)
        out = C.c_double()
        self._chk(self.L.ins_error_norm(self.h, order, kind, _p(prm), _p(uu), C.byref(out)))
        return out.value

    def set_neumann_flux(self, kind, params=None):
        """Neumann data grad u . n of a built-in function (REF_SINPROD /
        REF_COSPROD) on the traction faces; REF_ZERO = the traction vector"""
        prm = np.zeros(8)
        if params is not None:
            prm[:len(params)] = params
        self._chk(self.L.ins_set_neumann_flux(self.h, kind, _p(prm)))

    def cell_eval(self, quantity, xi=None, u=None):
        """value, gradient, 2nd Piola-Kirchhoff or Cauchy stress at the local
        point xi (None = centroid) of every element, on the device; rows per
        element (value: dpn, gradient: dim*dpn, stresses: 9 = 3x3 row-major)"""
        per = {CELL_VALUE: self.dpn, CELL_GRADIENT: self.dim * self.dpn,
               CELL_PK2: 9, CELL_CAUCHY: 9}.get(quantity, 9)
        x = None if xi is None else np.ascontiguousarray(xi, np.float64)
        uu = None if u is None else np.ascontiguousarray(u, np.float64)
        out = np.zeros(self.n_elem * per)
        self._chk(self.L.ins_cell_eval(self.h, quantity, _p(x), _p(uu), _p(out)))
        return out.reshape(self.n_elem, per)

    def bdf_init(self, order, density=1.0, u0=None):
        u = None if u0 is None else np.ascontiguousarray(u0, np.float64)
        self._chk(self.L.ins_bdf_init(self.h, order, density, _p(u)))

    def bdf_step(self, dt, precision=FP64, mode=SOLVE_MIXED, fmt=FMT_SELL, rtol=1e-8,
                 max_it=None):
        """one device-resident time step; returns (iterations, relres)"""
        it = C.c_int()
        rr = C.c_double()
        self._chk(self.L.ins_bdf_step(self.h, dt, precision, mode, fmt, rtol,
                                      max_it if max_it is not None else max(2 * self.n_active, 100),
                                      C.byref(it), C.byref(rr)))
        return it.value, rr.value

    def set_dofs(self, dpn, dof_id, n_active):
        d = np.ascontiguousarray(dof_id, np.int32)
        self._chk(self.L.ins_set_dofs(self.h, dpn, _p(d), n_active))
        self.dpn, self.n_active = dpn, n_active
        self.dof_n_active = n_active

    def set_prescribed(self, g):
        g = None if g is None else np.ascontiguousarray(g, np.float64)
        self._chk(self.L.ins_set_prescribed(self.h, _p(g)))

    def scale_prescribed(self, f):
        self._chk(self.L.ins_scale_prescribed(self.h, f))

    def set_dof_values(self, u):
        u = None if u is None else np.ascontiguousarray(u, np.float64)
        self._chk(self.L.ins_set_dof_values(self.h, _p(u)))

    def get_dof_values(self):
        u = np.zeros(self.n_nodes * self.dpn)
        self._chk(self.L.ins_get_dof_values(self.h, _p(u)))
        return u

    def build_pattern(self):
        nnz = C.c_int64()
        self._chk(self.L.ins_build_pattern(self.h, C.byref(nnz)))
        self.nnz = nnz.value
        self.n_active = self.dof_n_active
        return nnz.value

    def get_pattern(self):
        rp = np.zeros(self.n_active + 1, np.int32)
        ci = np.zeros(self.nnz, np.int32)
        self._chk(self.L.ins_get_pattern(self.h, _p(rp), _p(ci)))
        return rp, ci

    def set_material(self, kernel, nq, coef):
        coef = np.ascontiguousarray(coef, np.float64)
        if coef.ndim == 1:
            coef = coef[:, None]
        self._chk(self.L.ins_set_material(self.h, kernel, nq, coef.shape[1], _p(coef)))

    def set_body_force(self, kind, params=(0.0, 0.0, 0.0, 0.0)):
        p = np.zeros(4)
        p[:len(params)] = params
        self._chk(self.L.ins_set_body_force(self.h, kind, _p(p)))

    def set_body_force_values(self, f):
        """values of a body force at the quadrature points, (E, nqp, dpn)"""
        v = None if f is None else np.ascontiguousarray(f, np.float64)
        self._chk(self.L.ins_set_body_force_values(self.h, _p(v)))

    def set_traction_values(self, t):
        """values of a Neumann function at the faces' surface points,
        (n_faces, nqf, dpn)"""
        v = None if t is None else np.ascontiguousarray(t, np.float64)
        self._chk(self.L.ins_set_traction_values(self.h, _p(v)))

    def set_traction_profile(self, axis=-1, x0=0.0):
        self._chk(self.L.ins_set_traction_profile(self.h, axis, x0))

    def set_traction(self, faces, traction):
        if faces is None or len(faces) == 0:
            self._chk(self.L.ins_set_traction(self.h, 0, 0, None, None))
            return
        f = np.ascontiguousarray(faces, np.int32)
        t = np.zeros(3)
        t[:len(traction)] = traction
        self._chk(self.L.ins_set_traction(self.h, f.shape[0], f.shape[1], _p(f), _p(t)))

    # -- assembly -------------------------------------------------------------
    def assemble(self, precision=FP64, flags=ASM_MATRIX | ASM_RHS):
        self._chk(self.L.ins_assemble(self.h, precision, flags))

    def get_csr_values(self):
        v = np.zeros(self.nnz)
        self._chk(self.L.ins_get_csr_values(self.h, _p(v)))
        return v

    def get_rhs(self):
        b = np.zeros(self.n_active)
        self._chk(self.L.ins_get_rhs(self.h, _p(b)))
        return b

    def set_rhs(self, b):
        b = np.ascontiguousarray(b, np.float64)
        self._chk(self.L.ins_set_rhs(self.h, _p(b)))

    def add_rhs(self, b):
        b = np.ascontiguousarray(b, np.float64)
        self._chk(self.L.ins_add_rhs(self.h, _p(b)))

    def rhs_norm(self):
        v = C.c_double()
        self._chk(self.L.ins_rhs_norm(self.h, C.byref(v)))
        return v.value

    def set_csr(self, row_ptr, col_idx, val):
        rp = np.ascontiguousarray(row_ptr, np.int32)
        ci = np.ascontiguousarray(col_idx, np.int32)
        v = np.ascontiguousarray(val, np.float64)
        self._chk(self.L.ins_set_csr(self.h, rp.shape[0] - 1, _p(rp), _p(ci), _p(v)))
        self.n_active, self.nnz = rp.shape[0] - 1, ci.shape[0]

    # -- solve ------------------------------------------------------------------
    def solve(self, mode=SOLVE_FP64, fmt=FMT_CSR, rtol=1e-8, max_it=None, check=True):
        it = C.c_int()
        rr = C.c_double()
        s = self.L.ins_solve(self.h, mode, fmt, rtol,
                             max_it if max_it is not None else max(2 * self.n_active, 100),
                             C.byref(it), C.byref(rr))
        if check:
            self._chk(s)
        return it.value, rr.value, s

    def solve_direct(self, kind=DIRECT_CHOLESKY, check=True):
        """dense Cholesky / LU of the FP64 system; returns (relres, status)"""
        rr = C.c_double()
        s = self.L.ins_solve_direct(self.h, kind, C.byref(rr))
        if check:
            self._chk(s)
        return rr.value, s

    def prepare(self, mode=SOLVE_MIXED, fmt=FMT_SELL):
        self._chk(self.L.ins_prepare(self.h, mode, fmt))

    def get_solution(self):
        x = np.zeros(self.n_active)
        self._chk(self.L.ins_get_solution(self.h, _p(x)))
        return x

    def solution_norm(self):
        v = C.c_double()
        self._chk(self.L.ins_solution_norm(self.h, C.byref(v)))
        return v.value

    def distribute(self, op=DIST_SET):
        self._chk(self.L.ins_distribute(self.h, op))

    def stats(self):
        st = Stats()
        self._chk(self.L.ins_get_stats(self.h, C.byref(st)))
        return {k: getattr(st, k) for k, _ in Stats._fields_}

    def set_profiling(self, on):
        self._chk(self.L.ins_set_profiling(self.h, int(on)))

    def sync(self):
        self._chk(self.L.ins_sync(self.h))

    # -- convenience --------------------------------------------------------------
    def load_problem(self, prob):
        """Mesh, DoFs, prescribed values, kernel and loads of a Problem."""
        self.set_mesh(prob.xyz, prob.conn)
        if getattr(prob, "deg", 1) == 2:
            self.set_field(2, prob.n_fnodes, prob.fconn)
        self.set_dofs(prob.dpn, prob.dof_id, prob.n_active)
        self.set_prescribed(prob.g)
        self.set_material(prob.kernel, prob.nq, prob.coef)
        self.set_body_force(prob.bf_kind, prob.bf)
        if getattr(prob, "deg", 1) == 2:
            self.set_traction(prob.ffaces, prob.traction)
        else:
            self.set_traction(prob.faces, prob.traction)
        lin = prob.meta.get("traction_linear")
        self.set_traction_profile(*(lin if lin is not None else (-1, 0.0)))
        flux = prob.meta.get("neumann_flux")
        self.set_neumann_flux(*(flux if flux is not None else (REF_ZERO, None)))
        if prob.meta.get("bf_table") is not None:
            self.set_body_force_values(prob.meta["bf_table"])
        if prob.meta.get("traction_table") is not None:
            self.set_traction_values(prob.meta["traction_table"])
        return self.build_pattern()


class Server:
    """Batched mixed-precision solves of several assembled contexts
    (continuous batching: requests attach and finish between chunks)."""

    def __init__(self, device=0, max_slots=64):
        self.L = load()
        h = C.c_void_p()
        s = self.L.ins_server_create(C.byref(h), device, max_slots)
        if s != INS_OK:
            raise InsError(s, self.L.ins_strerror(s).decode())
        self.h = h

    def close(self):
        if self.h:
            self.L.ins_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _chk(self, s):
        if s != INS_OK:
            raise InsError(s, self.L.ins_server_last_error(self.h).decode())
        return s

    def attach(self, ctx, rtol=1e-8, max_it=None):
        slot = C.c_int()
        self._chk(self.L.ins_server_attach(self.h, ctx.h, rtol,
                                           max_it if max_it is not None else max(2 * ctx.n_active, 100),
                                           C.byref(slot)))
        return slot.value

    def step(self, k=16):
        nf = C.c_int()
        self._chk(self.L.ins_server_step(self.h, k, C.byref(nf)))
        return nf.value

    def poll(self, slot):
        st, status, it, rr = C.c_int(), C.c_int(), C.c_int(), C.c_double()
        self._chk(self.L.ins_server_poll(self.h, slot, C.byref(st), C.byref(status), C.byref(it),
                                         C.byref(rr)))
        return st.value, status.value, it.value, rr.value

    def detach(self, slot):
        self._chk(self.L.ins_server_detach(self.h, slot))

    def running(self):
        n = C.c_int()
        self._chk(self.L.ins_server_running(self.h, C.byref(n)))
        return n.value

    def run_until_done(self, k=16, limit=1000000):
        """step until no request is running"""
        for _ in range(limit):
            if self.running() == 0:
                self.step(k)  # drain the last snapshots
                return
            self.step(k)
        raise RuntimeError("server did not drain")
