"""CPU tests of the oracle's cell evaluation (orc_cell_eval): field value and
gradient at a local point of every element, second Piola-Kirchhoff and Cauchy
stress of the solid kernels (solid/Stress.hpp:38-81 with
solid/Deformation.hpp:26-45 and base/post/evaluateField.hpp:166-273).

The reference holds no stress fixture (its elastic drivers write the stress
only into VTK files), so the oracle is pinned by what the mathematics fixes:
affine fields are reproduced exactly by Q1 and Q2 interpolation on a
distorted mesh (value and gradient against a hand-written bilinear map);
uniaxial stretch and simple shear have closed-form stresses for both
material laws; a rigid rotation is stress free; and the Cauchy stress is
objective (sigma(QF) = Q sigma(F) Q^T) and symmetric."""
import numpy as np
import pytest

from insilico_amd import problems as P

VERT = {2: np.array([[0, 0], [1, 0], [1, 1], [0, 1]], float),
        3: np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0],
                     [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)}


def _map(prob, xi):
    """x(xi) of every element by the multilinear map over the vertices in
    hierarchic order (vertex v sits at VERT[dim][v])."""
    V = VERT[prob.dim]
    N = np.prod(np.where(V == 1, xi[None, :], 1.0 - xi[None, :]), axis=1)
    return np.einsum("v,evd->ed", N, prob.xyz[prob.conn])


def _support(prob):
    return prob.fxyz if prob.deg == 2 else prob.xyz


def _affine(prob, A, b):
    """u = A x + b at the field nodes, (n_fnodes * dpn,)"""
    return (_support(prob) @ A.T + b).ravel()


@pytest.mark.parametrize("dim,deg", [(2, 1), (2, 2), (3, 1), (3, 2)])
def test_affine_value_and_gradient(oracle, dim, deg):
    O = oracle
    prob = P.distorted_block(dim, 3, deg=deg)
    rng = np.random.default_rng(11 + dim + deg)
    A = rng.normal(size=(dim, dim))
    b = rng.normal(size=dim)
    u = _affine(prob, A, b)
    for xi in (np.full(dim, 0.5), rng.uniform(size=dim)):
        val = O.cell_eval(prob, u, O.CELL_VALUE, xi)
        np.testing.assert_allclose(val, _map(prob, xi) @ A.T + b, rtol=0, atol=1e-12)
        grad = O.cell_eval(prob, u, O.CELL_GRADIENT, xi).reshape(-1, dim, dim)
        # grad[e, d, c] = du_c / dx_d = A[c, d]
        np.testing.assert_allclose(grad, np.broadcast_to(A.T, grad.shape), rtol=0, atol=1e-11)


def test_scalar_field_value(oracle):
    O = oracle
    prob = P.distorted_block(3, 2, deg=2, kernel=P.LAPLACE)
    u = 0.3 + _support(prob) @ np.array([1.0, -2.0, 0.5])
    xi = np.array([0.2, 0.7, 0.4])
    val = O.cell_eval(prob, u, O.CELL_VALUE, xi)
    np.testing.assert_allclose(val[:, 0], 0.3 + _map(prob, xi) @ np.array([1.0, -2.0, 0.5]),
                               rtol=0, atol=1e-12)
    with pytest.raises(ValueError):
        O.cell_eval(prob, u, O.CELL_CAUCHY)


def _homogeneous(oracle, dim, kernel, Fm, lam=2.0, mu=0.8, deg=1):
    """PK2 and Cauchy stress of the homogeneous deformation x -> Fm x
    (Fm: dim x dim) in every element, as (E, 3, 3) arrays."""
    prob = P.distorted_block(dim, 2, deg=deg, kernel=kernel, lam=lam, mu=mu)
    u = _affine(prob, Fm - np.eye(dim), np.zeros(dim))
    S = oracle.cell_eval(prob, u, oracle.CELL_PK2).reshape(-1, 3, 3)
    sig = oracle.cell_eval(prob, u, oracle.CELL_CAUCHY).reshape(-1, 3, 3)
    return S, sig


@pytest.mark.parametrize("dim", [2, 3])
def test_uniaxial_stretch_closed_form(oracle, dim):
    lam, mu, s = 2.0, 0.8, 1.3
    Fm = np.eye(dim)
    Fm[0, 0] = s
    # StVenant: E11 = (s^2 - 1)/2, S = lam tr(E) I + 2 mu E,
    # sigma = F S F^T / J with J = s
    E11 = 0.5 * (s * s - 1.0)
    S, sig = _homogeneous(oracle, dim, P.STVENANT, Fm, lam, mu)
    S_ref = np.diag([(lam + 2 * mu) * E11, lam * E11, lam * E11])
    sig_ref = np.diag([s * (lam + 2 * mu) * E11, lam * E11 / s, lam * E11 / s])
    np.testing.assert_allclose(S, np.broadcast_to(S_ref, S.shape), rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(sig, np.broadcast_to(sig_ref, sig.shape), rtol=1e-13, atol=1e-13)
    # compressible neo-Hooke: S = (lam ln J - mu) C^-1 + mu I
    S, sig = _homogeneous(oracle, dim, P.NEOHOOKE, Fm, lam, mu)
    L = np.log(s)
    S_ref = np.diag([(lam * L - mu) / s ** 2 + mu, lam * L, lam * L])
    sig_ref = np.diag([(lam * L - mu) / s + mu * s, lam * L / s, lam * L / s])
    np.testing.assert_allclose(S, np.broadcast_to(S_ref, S.shape), rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(sig, np.broadcast_to(sig_ref, sig.shape), rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("deg", [1, 2])
def test_simple_shear_neo_hooke(oracle, deg):
    """x = X + g Y e_x: J = 1, sigma = mu (B - I) with B = F F^T, i.e.
    sigma_xx = mu g^2, sigma_xy = mu g, every other component 0 (the
    textbook Poynting result; a transposed F would put mu g^2 into
    sigma_yy)."""
    g, mu = 0.4, 0.8
    Fm = np.eye(3)
    Fm[0, 1] = g
    _, sig = _homogeneous(oracle, 3, P.NEOHOOKE, Fm, 2.0, mu, deg=deg)
    ref = np.zeros((3, 3))
    ref[0, 0] = mu * g * g
    ref[0, 1] = ref[1, 0] = mu * g
    np.testing.assert_allclose(sig, np.broadcast_to(ref, sig.shape), rtol=0, atol=1e-13)


def _rot(axis, t):
    c, s = np.cos(t), np.sin(t)
    R = np.eye(3)
    i, j = [(1, 2), (2, 0), (0, 1)][axis]
    R[i, i] = R[j, j] = c
    R[i, j], R[j, i] = -s, s
    return R


@pytest.mark.parametrize("kernel", [P.STVENANT, P.NEOHOOKE])
def test_rotation_free_and_objective(oracle, kernel):
    Q = _rot(0, 0.7) @ _rot(2, -0.3)
    _, sig = _homogeneous(oracle, 3, kernel, Q)
    np.testing.assert_allclose(sig, 0.0, atol=1e-13)
    rng = np.random.default_rng(5)
    Fm = np.eye(3) + 0.2 * rng.normal(size=(3, 3))
    assert np.linalg.det(Fm) > 0.5
    S1, sig1 = _homogeneous(oracle, 3, kernel, Fm)
    S2, sig2 = _homogeneous(oracle, 3, kernel, Q @ Fm)
    np.testing.assert_allclose(sig2, np.einsum("ij,ejk,lk->eil", Q, sig1, Q), rtol=0,
                               atol=1e-12)
    np.testing.assert_allclose(S2, S1, rtol=0, atol=1e-12)   # PK2 is rotation invariant
    np.testing.assert_allclose(sig1, np.swapaxes(sig1, 1, 2), rtol=0, atol=1e-13)


def test_small_strain_limit(oracle):
    """|grad u| = O(eps): sigma = lam tr(eps) I + 2 mu eps + O(eps^2) for
    both laws (linear elasticity)."""
    rng = np.random.default_rng(9)
    H = rng.normal(size=(3, 3))
    lam, mu = 2.0, 0.8
    for kernel in (P.STVENANT, P.NEOHOOKE):
        errs = []
        for eps in (1e-3, 1e-4):
            _, sig = _homogeneous(oracle, 3, kernel, np.eye(3) + eps * H, lam, mu)
            e = 0.5 * eps * (H + H.T)
            lin = lam * np.trace(e) * np.eye(3) + 2 * mu * e
            errs.append(np.abs(sig - lin).max())
        assert errs[0] < 50 * 1e-6 and errs[1] < 50 * 1e-8
        assert 50 < errs[0] / errs[1] < 200      # second order in eps
