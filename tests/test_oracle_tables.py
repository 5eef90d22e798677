"""Pins of the oracle's caller-function paths: a body force, a Neumann
function and a conductivity given as their values at the quadrature points
(base/asmb/BodyForce.hpp:172-201, base/asmb/NeumannForce.hpp:132-175,
heat/Laplace.hpp:113-126 evaluate the caller's function there).

* A table of f(x_q) for a function the oracle also has built in (the sine and
  cosine products, the linear traction profile, the co-normal derivative of
  a cosine product) gives the built-in result: the points, their order and
  the (element, point, component) layout agree.  The same table with its
  points permuted does not.
* A conductivity that is 1 at one quadrature point and 0 at the others gives
  that point's share of the element matrix, w_q grad phi_M(xi_q) .
  grad phi_N(xi_q), computed here from the Q1 gradients in closed form at
  the reference's Gauss points (base/quad/GaussLegendre.hpp:122-127).
"""
import numpy as np
import pytest

from insilico_amd import problems as P


def _rhs(oracle, prob, **kw):
    return oracle.assemble(prob, **kw)[3]


def _bf_table(oracle, prob, fun):
    w, xi = oracle.tensor_rule(prob.dim, prob.nq)
    x = P.element_points(prob.xyz, prob.conn, xi)           # (E, nq, dim)
    return np.ascontiguousarray(fun(x)[..., None] if prob.dpn == 1 else fun(x))


def test_body_force_table_equals_builtin_sine(oracle):
    prob = P.poisson2d(12)
    ref = _rhs(oracle, prob)
    tab = _bf_table(oracle, prob, lambda x: 2 * np.pi ** 2 * np.sin(np.pi * x[..., 0]) *
                    np.sin(np.pi * x[..., 1]))
    prob.meta["bf_table"] = tab
    got = _rhs(oracle, prob)
    assert np.max(np.abs(got - ref)) <= 1e-14 * np.max(np.abs(ref))
    prob.meta["bf_table"] = np.ascontiguousarray(tab[:, ::-1, :])   # points reversed
    assert np.max(np.abs(_rhs(oracle, prob) - ref)) > 1e-4 * np.max(np.abs(ref))


def test_body_force_table_equals_builtin_cosine_3d(oracle):
    prob = P.diffusion3d(5, seed=3)
    k = (1.3, 0.7, 2.1)
    amp = sum(v * v for v in k)
    prob.bf_kind, prob.bf = P.BF_COSPROD, (amp, *k)
    ref = _rhs(oracle, prob)
    prob.meta["bf_table"] = _bf_table(
        oracle, prob, lambda x: amp * np.cos(k[0] * x[..., 0]) * np.cos(k[1] * x[..., 1]) *
        np.cos(k[2] * x[..., 2]))
    got = _rhs(oracle, prob)
    assert np.max(np.abs(got - ref)) <= 1e-14 * np.max(np.abs(ref))


def test_vector_body_force_table_components(oracle):
    """Elasticity with a constant body force (0.1, -0.2, 0.3): the table
    of the same constant per point and component equals the built-in."""
    prob = P.elasticity3d(4, seed=2)
    prob.bf_kind, prob.bf = P.BF_CONST, (0.1, -0.2, 0.3, 0.0)
    ref = _rhs(oracle, prob)
    prob.meta["bf_table"] = _bf_table(
        oracle, prob, lambda x: np.broadcast_to(np.array([0.1, -0.2, 0.3]), x.shape).copy())
    assert np.max(np.abs(_rhs(oracle, prob) - ref)) <= 1e-14 * np.max(np.abs(ref))


def _face_table(oracle, prob, fun):
    _, eta = oracle.tensor_rule(prob.dim - 1, prob.nq)
    x, n = P.face_points(prob.xyz, prob.faces, eta)
    return np.ascontiguousarray(fun(x, n))


def test_traction_table_equals_linear_profile(oracle):
    prob = P.elasticity3d(4, seed=5)
    t0, axis, x0 = np.array([0.2, -0.1, 0.4]), 0, 0.3
    prob.traction = tuple(t0)
    prob.meta["traction_linear"] = (axis, x0)
    ref = _rhs(oracle, prob)
    del prob.meta["traction_linear"]
    prob.meta["traction_table"] = _face_table(
        oracle, prob, lambda x, n: t0[None, None, :] * (x[..., axis:axis + 1] - x0))
    got = _rhs(oracle, prob)
    assert np.max(np.abs(got - ref)) <= 1e-14 * np.max(np.abs(ref))


def test_traction_table_equals_conormal_flux(oracle):
    """The Neumann function grad u . n of u = cos product, evaluated at the
    surface points with the normals of face_points, equals the oracle's
    co-normal flux path on a sheared mesh (all six face orientations)."""
    xyz, conn = P.unit_cube(3, 4)
    xyz = xyz @ np.array([[1.0, 0.2, 0.0], [0.1, 1.1, 0.2], [0.0, -0.1, 0.9]]).T
    _, _, bf = P.boundary_faces(conn, 3)
    n = xyz.shape[0]
    a, k = 1.4, np.array([0.9, 1.2, 0.6])
    prob = P.Problem(3, xyz, conn, 1, np.arange(n, dtype=np.int32), n, np.zeros(n), P.LAPLACE,
                     np.ones((conn.shape[0], 1)), 2, faces=bf.astype(np.int32),
                     traction=(0.0, 0.0, 0.0), meta={"neumann_flux": (P.REF_COSPROD, [a, *k])})
    ref = _rhs(oracle, prob)
    del prob.meta["neumann_flux"]

    def flux(x, nrm):
        c = np.cos(k * x)
        grad = np.stack([-a * k[0] * np.sin(k[0] * x[..., 0]) * c[..., 1] * c[..., 2],
                         -a * k[1] * c[..., 0] * np.sin(k[1] * x[..., 1]) * c[..., 2],
                         -a * k[2] * c[..., 0] * c[..., 1] * np.sin(k[2] * x[..., 2])], -1)
        return np.sum(grad * nrm, axis=-1, keepdims=True)

    prob.meta["traction_table"] = _face_table(oracle, prob, flux)
    got = _rhs(oracle, prob)
    assert np.max(np.abs(got - ref)) <= 1e-13 * np.max(np.abs(ref))


def test_per_point_conductivity_closed_form(oracle):
    xyz, conn = P.unit_cube(3, 1)
    prob = P.Problem(3, xyz, conn, 1, np.arange(8, dtype=np.int32), 8, np.zeros(8), P.LAPLACE,
                     np.ones((1, 1)), 2)
    g = (0.788675134594813, 0.211324865405187)   # GaussLegendre.hpp:122-127, descending
    pts = [(g[i], g[j], g[k]) for k in range(2) for j in range(2) for i in range(2)]
    H = (0, 1, 3, 2, 4, 5, 7, 6)
    for q, xi in enumerate(pts):
        kap = np.zeros((1, 8))
        kap[0, q] = 1.0
        prob.coef = kap
        Ke, _ = oracle.element(prob, 0)
        grad = np.zeros((8, 3))
        for lex in range(8):
            bits = [(lex >> d) & 1 for d in range(3)]
            for d in range(3):
                gval = 1.0
                for e in range(3):
                    f = xi[e] if bits[e] else 1.0 - xi[e]
                    gval *= (1.0 if bits[d] else -1.0) if e == d else f
                grad[H[lex], d] = gval
        exp = 0.125 * grad @ grad.T
        np.testing.assert_allclose(Ke, exp, rtol=1e-14, atol=1e-16)
    # a per-point table of one value per element is the per-element conductivity
    prob.coef = np.full((1, 8), 2.5)
    K8, _ = oracle.element(prob, 0)
    prob.coef = np.full((1, 1), 2.5)
    K1, _ = oracle.element(prob, 0)
    assert np.array_equal(K8, K1)
