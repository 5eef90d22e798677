"""ctypes front-end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by the test-suite, by the smoke check and
by the CPU-baseline leg of bench.py, never by the product package.  See
oracle.c for the reference file:line each function restates.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

LAPLACE, STVENANT, NEOHOOKE = 1, 2, 3
BF_NONE, BF_CONST, BF_SINPROD, BF_COSPROD, BF_TABLE = 0, 1, 2, 3, 4

_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class Problem(C.Structure):
    _fields_ = [
        ("dim", C.c_int), ("n_nodes", C.c_int64), ("xyz", C.c_void_p),
        ("n_elem", C.c_int64), ("npe", C.c_int), ("conn", C.c_void_p),
        ("dpn", C.c_int), ("dof_id", C.c_void_p), ("n_active", C.c_int64),
        ("g", C.c_void_p), ("u", C.c_void_p), ("incremental", C.c_int),
        ("kernel", C.c_int), ("ncoef", C.c_int), ("coef", C.c_void_p),
        ("nq", C.c_int), ("bf_kind", C.c_int), ("bf", C.c_double * 4),
        ("n_faces", C.c_int64), ("nfn", C.c_int), ("faces", C.c_void_p),
        ("traction", C.c_double * 3), ("residual", C.c_int),
        ("nthreads", C.c_int),
        ("deg", C.c_int), ("fnpe", C.c_int), ("fconn", C.c_void_p),
        ("ffn", C.c_int), ("ffaces", C.c_void_p),
        ("tr_kind", C.c_int), ("tr_axis", C.c_int), ("tr_x0", C.c_double),
        ("density", C.c_double), ("mass_factor", C.c_double), ("mass_hist", C.c_void_p),
        ("flux_kind", C.c_int), ("flux_p", C.c_double * 8),
        ("bf_tab", C.c_void_p), ("tr_tab", C.c_void_p),
    ]


def build():
    """Compile oracle/liboracle.so with its Makefile."""
    subprocess.run(["make", "-s", "-B", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_gauss1d.argtypes = [C.c_int, _f64p, _f64p]
        L.orc_tensor_rule.argtypes = [C.c_int, C.c_int, _f64p, _f64p]
        L.orc_q1_shape.argtypes = [C.c_int, _f64p, _f64p, _f64p]
        L.orc_q1_shape.restype = None
        L.orc_lagrange_shape.argtypes = [C.c_int, C.c_int, _f64p, _f64p, _f64p]
        L.orc_lagrange_shape.restype = None
        L.orc_contravariant.argtypes = [C.c_int, _f64p, _f64p, _f64p]
        L.orc_contravariant.restype = C.c_double
        L.orc_element.argtypes = [C.POINTER(Problem), C.c_int64, _f64p, _f64p]
        L.orc_mass_element.argtypes = [C.POINTER(Problem), C.c_int64, _f64p]
        L.orc_bdf_weights.argtypes = [C.c_int, C.c_int, _f64p]
        L.orc_unit_cube.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _f64p, _i32p]
        L.orc_pattern.argtypes = [C.POINTER(Problem), C.POINTER(C.c_int64),
                                  C.c_void_p, C.c_void_p]
        L.orc_assemble.argtypes = [C.POINTER(Problem), _i32p, _i32p, _f64p, _f64p, _f64p]
        L.orc_pcg.argtypes = [C.c_int64, _i32p, _i32p, _f64p, _f64p, _f64p,
                              C.c_double, C.c_int, C.c_int,
                              C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.orc_spmv.argtypes = [C.c_int64, _i32p, _i32p, _f64p, _f64p, _f64p, C.c_int]
        L.orc_spmv.restype = None
        L.orc_l2_error_fundsol.argtypes = [C.POINTER(Problem), _f64p, C.c_double,
                                           C.c_double, _f64p, _f64p]
        L.orc_l2_error_fundsol.restype = C.c_double
        L.orc_fundsol_elasto.argtypes = [C.c_int, C.c_double, C.c_double, _f64p,
                                         _f64p, _f64p, _f64p]
        L.orc_fundsol_elasto.restype = None
        L.orc_error_norm.argtypes = [C.POINTER(Problem), _f64p, C.c_int, C.c_int, _f64p]
        L.orc_error_norm.restype = C.c_double
        L.orc_cell_eval.argtypes = [C.POINTER(Problem), _f64p, C.c_int, _f64p, _f64p]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


class OracleProblem:
    """Holds numpy arrays alive and exposes the C struct."""

    def __init__(self, prob, nthreads=1, u=None, incremental=False, residual=False):
        self.keep = []
        p = Problem()
        p.dim = prob.dim
        xyz = np.ascontiguousarray(prob.xyz, np.float64)
        conn = np.ascontiguousarray(prob.conn, np.int32)
        dof = np.ascontiguousarray(prob.dof_id, np.int32)
        g = np.ascontiguousarray(prob.g, np.float64)
        coef = np.ascontiguousarray(prob.coef, np.float64)
        self.keep += [xyz, conn, dof, g, coef]
        p.n_nodes = xyz.shape[0]
        p.xyz = _ptr(xyz)
        p.n_elem = conn.shape[0]
        p.npe = conn.shape[1]
        p.conn = _ptr(conn)
        p.dpn = prob.dpn
        p.dof_id = _ptr(dof)
        p.n_active = prob.n_active
        p.g = _ptr(g)
        if u is not None:
            uu = np.ascontiguousarray(u, np.float64)
            self.keep.append(uu)
            p.u = _ptr(uu)
        p.incremental = int(incremental)
        p.kernel = prob.kernel
        p.ncoef = coef.shape[1]
        p.coef = _ptr(coef)
        p.nq = prob.nq
        p.bf_kind = prob.bf_kind
        for i, v in enumerate(prob.bf):
            p.bf[i] = v
        meta = getattr(prob, "meta", {})
        if meta.get("bf_table") is not None:
            bt = np.ascontiguousarray(meta["bf_table"], np.float64)
            self.keep.append(bt)
            p.bf_kind = BF_TABLE
            p.bf_tab = _ptr(bt)
        if meta.get("traction_table") is not None:
            tt = np.ascontiguousarray(meta["traction_table"], np.float64)
            self.keep.append(tt)
            p.tr_tab = _ptr(tt)
        if prob.faces is not None and len(prob.faces):
            faces = np.ascontiguousarray(prob.faces, np.int32)
            self.keep.append(faces)
            p.n_faces = faces.shape[0]
            p.nfn = faces.shape[1]
            p.faces = _ptr(faces)
            for i, v in enumerate(prob.traction):
                p.traction[i] = v
            lin = getattr(prob, "meta", {}).get("traction_linear")
            if lin is not None:
                p.tr_kind, p.tr_axis, p.tr_x0 = 1, int(lin[0]), float(lin[1])
            flux = getattr(prob, "meta", {}).get("neumann_flux")
            if flux is not None:
                p.flux_kind = int(flux[0])
                for i, v in enumerate(flux[1]):
                    p.flux_p[i] = v
        p.residual = int(residual)
        p.nthreads = nthreads
        tr = getattr(prob, "meta", {}).get("transient")
        if tr is not None:
            p.density = float(tr["density"])
            p.mass_factor = float(tr["mass_factor"])
            if tr.get("hist") is not None:
                h = np.ascontiguousarray(tr["hist"], np.float64)
                self.keep.append(h)
                p.mass_hist = _ptr(h)
        if getattr(prob, "deg", 1) == 2:
            fconn = np.ascontiguousarray(prob.fconn, np.int32)
            self.keep.append(fconn)
            p.deg = 2
            p.fnpe = fconn.shape[1]
            p.fconn = _ptr(fconn)
            if prob.ffaces is not None and len(prob.ffaces):
                ff = np.ascontiguousarray(prob.ffaces, np.int32)
                self.keep.append(ff)
                p.ffn = ff.shape[1]
                p.ffaces = _ptr(ff)
        self.p = p

    def ref(self):
        return C.byref(self.p)


def pattern(prob, **kw):
    op = OracleProblem(prob, **kw)
    L = lib()
    nnz = C.c_int64()
    L.orc_pattern(op.ref(), C.byref(nnz), None, None)
    rp = np.zeros(prob.n_active + 1, np.int32)
    ci = np.zeros(nnz.value, np.int32)
    L.orc_pattern(op.ref(), C.byref(nnz), rp.ctypes.data, ci.ctypes.data)
    return rp, ci


def assemble(prob, rp=None, ci=None, **kw):
    """Returns (row_ptr, col_idx, val, b, seconds)."""
    if rp is None:
        rp, ci = pattern(prob)
    op = OracleProblem(prob, **kw)
    val = np.zeros(ci.shape[0], np.float64)
    b = np.zeros(prob.n_active, np.float64)
    tm = np.zeros(2, np.float64)
    rc = lib().orc_assemble(op.ref(), rp, ci, val, b, tm)
    if rc != 0:
        raise RuntimeError("oracle assembly failed: %d" % rc)
    return rp, ci, val, b, tm[0]


def element(prob, e, **kw):
    op = OracleProblem(prob, **kw)
    fc = prob.fconn if getattr(prob, "deg", 1) == 2 else prob.conn
    nee = fc.shape[1] * prob.dpn
    Ke = np.zeros(nee * nee, np.float64)
    Fe = np.zeros(nee, np.float64)
    lib().orc_element(op.ref(), e, Ke, Fe)
    return Ke.reshape(nee, nee), Fe


def mass_element(prob, e):
    op = OracleProblem(prob)
    fc = prob.fconn if getattr(prob, "deg", 1) == 2 else prob.conn
    nee = fc.shape[1] * prob.dpn
    Me = np.zeros(nee * nee, np.float64)
    lib().orc_mass_element(op.ref(), e, Me)
    return Me.reshape(nee, nee)


def bdf_weights(order, step):
    a = np.zeros(5)
    k = lib().orc_bdf_weights(order, step, a)
    return a[:k]


def pcg(rp, ci, val, b, tol=1e-8, maxit=None, nthreads=1):
    n = b.shape[0]
    x = np.zeros(n, np.float64)
    it = C.c_int()
    rr = C.c_double()
    lib().orc_pcg(n, rp, ci, val, np.ascontiguousarray(b), x, tol,
                  maxit if maxit is not None else 2 * n, nthreads,
                  C.byref(it), C.byref(rr))
    return x, it.value, rr.value


def spmv(rp, ci, val, x, nthreads=1):
    y = np.zeros(rp.shape[0] - 1, np.float64)
    lib().orc_spmv(y.shape[0], rp, ci, val, np.ascontiguousarray(x, np.float64), y, nthreads)
    return y


def tensor_rule(dim, npts):
    nq = npts ** dim
    w = np.zeros(nq)
    xi = np.zeros(nq * dim)
    lib().orc_tensor_rule(dim, npts, w, xi)
    return w, xi.reshape(nq, dim)


def q1_shape(dim, xi):
    phi = np.zeros(2 ** dim)
    dphi = np.zeros(2 ** dim * dim)
    lib().orc_q1_shape(dim, np.ascontiguousarray(xi, np.float64), phi, dphi)
    return phi, dphi.reshape(-1, dim)


def lagrange_shape(dim, deg, xi):
    nf = (deg + 1) ** dim
    phi = np.zeros(nf)
    dphi = np.zeros(nf * dim)
    lib().orc_lagrange_shape(dim, deg, np.ascontiguousarray(xi, np.float64), phi, dphi)
    return phi, dphi.reshape(-1, dim)


def unit_cube(dim, e1, e2, e3=1):
    nn = (e1 + 1) * (e2 + 1) * ((e3 + 1) if dim == 3 else 1)
    ne = e1 * e2 * (e3 if dim == 3 else 1)
    xyz = np.zeros(nn * dim)
    conn = np.zeros(ne * 2 ** dim, np.int32)
    lib().orc_unit_cube(dim, e1, e2, e3, xyz, conn)
    return xyz.reshape(nn, dim), conn.reshape(ne, 2 ** dim)


def l2_error_fundsol(prob, uh_nodal, lam, mu, src, direc):
    op = OracleProblem(prob)
    return lib().orc_l2_error_fundsol(op.ref(), np.ascontiguousarray(uh_nodal, np.float64),
                                      lam, mu, np.asarray(src, np.float64),
                                      np.asarray(direc, np.float64))


def fundsol_elasto(dim, lam, mu, y, x, direc):
    out = np.zeros(dim)
    lib().orc_fundsol_elasto(dim, lam, mu, np.asarray(y, np.float64),
                             np.asarray(x, np.float64), np.asarray(direc, np.float64), out)
    return out


REF_ZERO, REF_SINPROD, REF_KELVIN, REF_COSPROD = 0, 1, 2, 3


def error_norm(prob, uh, order, kind, params):
    """|u - u_h| (order 0: L2, order 1: H1 semi-norm) with the problem's quadrature."""
    op = OracleProblem(prob)
    prm = np.zeros(8)
    prm[:len(params)] = params
    return lib().orc_error_norm(op.ref(), np.ascontiguousarray(uh, np.float64), order, kind, prm)


CELL_VALUE, CELL_GRADIENT, CELL_PK2, CELL_CAUCHY = 0, 1, 2, 3


def cell_eval(prob, uh, quantity, xi=None):
    """Value, gradient, 2nd Piola-Kirchhoff or Cauchy stress at the local point
    xi (default: the element centroid) of every element; rows per element."""
    op = OracleProblem(prob)
    if xi is None:
        xi = np.full(prob.dim, 0.5)
    per = {CELL_VALUE: prob.dpn, CELL_GRADIENT: prob.dim * prob.dpn,
           CELL_PK2: 9, CELL_CAUCHY: 9}[quantity]
    out = np.zeros(prob.n_elem * per)
    got = lib().orc_cell_eval(op.ref(), np.ascontiguousarray(uh, np.float64), quantity,
                              np.ascontiguousarray(xi, np.float64), out)
    if got != per:
        raise ValueError("orc_cell_eval: quantity not defined for this problem")
    return out.reshape(prob.n_elem, per)
