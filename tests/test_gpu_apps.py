"""The example C++ drivers (insilico_amd/apps), written against the drop-in
C++ layer, run as subprocesses on the GPU and checked against the oracle and
the reference's golden output."""
import os
import subprocess

import numpy as np
import pytest

from insilico_amd import problems as P

pytestmark = pytest.mark.gpu

APPS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "insilico_amd", "apps")


def _run(*args, timeout=120, cwd=None):
    exe = os.path.join(APPS, args[0])
    if not os.path.exists(exe):
        raise RuntimeError("%s not built (run __graft_entry__.build())" % exe)
    out = subprocess.run([exe] + [str(a) for a in args[1:]], capture_output=True, text=True,
                         timeout=timeout, cwd=cwd)
    assert out.returncode == 0, out.stderr
    return out.stdout


def test_poisson_app(oracle):
    """64 x 64 Q1 Poisson through the C++ layer: same DoF count as the Python
    problem, relative residual 1e-10, nodal error equal to the oracle's solve."""
    ndof, it, rel, maxerr, l2 = _run("poisson", 64).split()
    prob = P.poisson2d(64)
    assert int(ndof) == prob.n_active
    assert float(rel) <= 1e-10
    rp, ci, val, b, _ = oracle.assemble(prob)
    x, _, _ = oracle.pcg(rp, ci, val, b, tol=1e-13)
    act = prob.dof_id >= 0
    ex = np.sin(np.pi * prob.xyz[:, 0]) * np.sin(np.pi * prob.xyz[:, 1])
    err_o = np.abs(x[prob.dof_id[act]] - ex[act]).max()
    assert abs(float(maxerr) - err_o) <= 1e-8 * err_o + 1e-11, (maxerr, err_o)
    assert float(l2) < 1e-3


def test_cg_solve_reports_no_convergence():
    """A tolerance below what FP64 round-off reaches: the drop-in cgSolve
    stops at its 2n iteration limit and reports INS_ERR_NOCONV (an exception
    the driver turns into exit status 2) instead of handing back the last
    iterate as a solution (Eigen's ConjugateGradient would return silently,
    base/solver/Eigen3.hpp:265-275)."""
    exe = os.path.join(APPS, "poisson")
    out = subprocess.run([exe, "16", "1e-30"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 2, (out.returncode, out.stderr)
    assert "error -7" in out.stderr and "relres" in out.stderr


@pytest.mark.parametrize("dim,mesh,N", [(2, "quad.002.smf", 2), (2, "quad.005.smf", 5),
                                        (2, "quad.010.smf", 10), (3, "cube.002.smf", 2),
                                        (3, "cube.004.smf", 4), (3, "cube.008.smf", 8)])
def test_linear_elastic_app_golden(oracle, golden, read_vtk, tmp_path, dim, mesh, N):
    """reference/06-elastic/linearElastic{2D,3D}.ref.dat to the printed 6
    digits (L2 error on the device); the VTK file it writes holds the grid,
    the displacement and the centroid Cauchy stress of the oracle's solve of
    the same problem to the printed digits."""
    import shutil
    ref = dict(np.loadtxt(os.path.join(golden, "linearElastic%dD.ref.dat" % dim)))
    shutil.copy(os.path.join(golden, mesh), tmp_path / mesh)
    err = float(_run("linear_elastic", tmp_path / mesh, dim))
    assert err == pytest.approx(ref[N], rel=3e-6), (N, err, ref[N])
    vtk = read_vtk(str(tmp_path / mesh.replace(".smf", ".vtk")))
    xyz, conn = P.read_smf(os.path.join(golden, mesh), dim)
    prob = P.linear_elastic_fundsol(xyz, conn, dim)
    assert vtk["cells"] == prob.conn.tolist()
    rp, ci, val, b, _ = oracle.assemble(prob)
    x, _, _ = oracle.pcg(rp, ci, val, b, tol=1e-13)
    u = prob.g.copy()
    act = prob.dof_id >= 0
    u[act] = x[prob.dof_id[act]]
    sig = oracle.cell_eval(prob, u, oracle.CELL_CAUCHY)
    w = vtk["cell_data"]["sigma"]
    assert np.all(np.abs(sig - w) <= 1e-5 * np.abs(sig).max() + 1e-5 * np.abs(w)), np.abs(sig - w).max()
    disp = np.zeros((prob.n_nodes, 3))
    disp[:, :dim] = u.reshape(-1, dim)
    wd = vtk["point_data"]["disp"]
    assert np.all(np.abs(disp - wd) <= 1e-5 * np.abs(disp).max() + 1e-5 * np.abs(wd))


def newton_oracle(oracle, n, steps, pull, maxit, tol, mu, lam):
    """The load-step / Newton loop of reference/06-elastic/compressible.cpp
    :245-335 on the oracle: returns [(step, iter, |F|, |x| or None)], u."""
    prob = P.neohooke3d(n, mu=mu, lam=lam)
    prob.g[prob.meta["right_x"] * 3] = pull / steps
    act = prob.dof_id >= 0
    u = np.zeros(prob.n_nodes * 3)
    hist = []
    for step in range(steps):
        prob.g *= 1.0 if step == 0 else (step + 1.0) / step
        it = 0
        while it < maxit:
            rp, ci, val, b, _ = oracle.assemble(prob, u=u, incremental=True, residual=True)
            conv1 = np.linalg.norm(b) / prob.n_active
            if conv1 < tol:
                hist.append((step, it, conv1, None))
                break
            x, _, _ = oracle.pcg(rp, ci, val, b, tol=1e-12)
            u[act] += x[prob.dof_id[act]]
            u[~act] = prob.g[~act]
            conv2 = np.linalg.norm(x) / prob.n_active
            hist.append((step, it, conv1, conv2))
            it += 1
            if conv2 < tol:
                break
    return hist, u


def _parse_history(out):
    hist, umax = [], None
    for line in out.splitlines():
        if line.startswith("# max |u_y|"):
            umax = float(line.split()[-1])
        if not line.strip() or line.startswith("#"):
            continue
        f = line.split()
        hist.append((int(f[0]), int(f[1]), float(f[2]), float(f[3]) if len(f) > 3 else None))
    return hist, umax


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_compressible_newton_app(oracle, prec):
    """Newton history (|F|, |x| per iteration) and final displacement of the
    neo-Hookean block against the oracle's Newton loop.  The first residual of
    every Newton sequence is the same FP64 state on both sides (1e-10); the
    later ones carry the CG tolerance and, with the FP32 tangent, its
    rounding, so they are compared where they are above that noise."""
    n, steps, pull, maxit, tol, mu, lam = 5, 3, 0.3, 10, 1e-9, 1.0, 10.0
    args = ["compressible", n, steps, pull, maxit, tol, mu, lam] + ([prec] if prec == "fp32" else [])
    hist, umax = _parse_history(_run(*args))
    ref, u = newton_oracle(oracle, n, steps, pull, maxit, tol, mu, lam)
    assert len(hist) >= steps
    tol_conv = 1e-6 if prec == "fp64" else 1e-3
    by_key = {(s, i): (c1, c2) for s, i, c1, c2 in ref}
    first = {}
    for s, i, c1, c2 in hist:
        first.setdefault(s, (i, c1))
        if (s, i) in by_key and c1 > 1e-7:
            assert c1 == pytest.approx(by_key[(s, i)][0], rel=tol_conv), (s, i, c1, by_key[(s, i)])
    for s, (i, c1) in first.items():
        assert i == 0
    # Newton converges quadratically: each step ends below the tolerance
    last = {}
    for s, i, c1, c2 in hist:
        last[s] = (c1, c2)
    for s, (c1, c2) in last.items():
        assert (c2 is not None and c2 < tol) or c1 < tol, (s, c1, c2)
    assert umax == pytest.approx(np.abs(u[1::3]).max(), rel=1e-6)


INPUTS = {
    # reference/06-elastic/inputCompRef{D,N}.dat
    "D": "E 1000.\nnu 0.3\nmaxIter 30\nloadSteps 5\npull 3.0\ntraction 1000.0\n"
         "dispControlled 1\ntolerance 1.e-12\nisIncompressible 0\n",
    "N": "E 1000.\nnu 0.3\nmaxIter 30\nloadSteps 5\npull 3.0\ntraction 300.0\n"
         "dispControlled 0\ntolerance 1.e-12\nisIncompressible 0\n",
}


def _table(text):
    rows = []
    for ln in text.splitlines():
        if ln.startswith("#") or not ln.strip():
            continue
        v = ln.split()
        rows.append((int(v[0]), int(v[1]), float(v[2]), float(v[3]) if len(v) > 3 else None))
    return rows


@pytest.mark.parametrize("mode", ["D", "N"])
def test_compressible_q2_app_golden(golden, read_vtk, tmp_path, mode):
    """reference/06-elastic/compressible.cpp run as its test harness runs it
    (compressibleTest.mk: quad.020.smf with inputCompRef{D,N}.dat) on the
    drop-in C++ layer with a degree-2 field (generateField, constrainBoundary
    at the Q2 support points, the linear traction profile on the Q2 faces of
    x = 1, device assembly of the Q2 residual and tangent, Jacobi-PCG): every
    |F| and |x| of compRefOut{D,N}.dat down to the round-off floor, and the
    last VTK file (displacement at the nodes, Cauchy stress at the centroids,
    evaluated on the device) against quad.020.{D,N}ref.vtk."""
    import shutil
    shutil.copy(os.path.join(golden, "square_20.smf"), tmp_path / "quad.020.smf")
    (tmp_path / "input.dat").write_text(INPUTS[mode])
    out = _run("compressible_q2", tmp_path / "quad.020.smf", tmp_path / "input.dat", timeout=300)
    with open(os.path.join(golden, "compRefOut%s.dat" % mode)) as fh:
        head = fh.readline()
        ref = _table(fh.read())
    assert out.splitlines()[0].strip() == head.strip()
    hist = _table(out)
    assert [(s, i) for s, i, _, _ in hist] == [(s, i) for s, i, _, _ in ref]
    for (s, i, f, x), (_, _, fr, xr) in zip(hist, ref):
        if fr >= 1e-7:
            assert abs(f - fr) <= 1e-5 * fr, (s, i, f, fr)
        else:
            assert f < 1e-6
        if x is not None and xr > 1e-9:
            assert abs(x - xr) <= 1e-4 * xr, (s, i, x, xr)
    for step in range(6):
        assert (tmp_path / ("quad.020.%04d.vtk" % step)).exists()
    got = read_vtk(str(tmp_path / "quad.020.0005.vtk"))
    want = read_vtk(os.path.join(golden, "quad.020.%sref.vtk" % mode))
    np.testing.assert_array_equal(got["points"], want["points"])
    assert got["cells"] == want["cells"] and got["cell_types"] == want["cell_types"]
    for sec, name in (("point_data", "disp"), ("cell_data", "sigma")):
        g, w = got[sec][name], want[sec][name]
        assert g.shape == w.shape
        tol = 1e-5 * np.maximum(np.abs(w), np.abs(w).max(axis=1, keepdims=True) * 1e-3) + 1e-12
        bad = np.argwhere(np.abs(g - w) > tol)
        assert bad.size == 0, (name, [(tuple(b), g[tuple(b)], w[tuple(b)]) for b in bad[:5]])


def test_mixed_poisson_app_golden(golden, read_vtk, tmp_path):
    """reference/05-mixedPoisson/mixedPoisson.cpp as its test harness runs it
    (appTest.mk: square_020.smf) on the drop-in C++ layer: Q2 field,
    cosine-product body force, co-normal-derivative Neumann data on the
    boundary, device assembly and solve, device L2 / H1 errors and mesh
    volume; stdout against the reference's `ref` and square_020.vtk against
    its ref.vtk (temperature at the nodes, flux at the centroids) to the
    printed digits."""
    import shutil
    shutil.copy(os.path.join(golden, "square_20.smf"), tmp_path / "square_020.smf")
    out = _run("mixed_poisson", tmp_path / "square_020.smf").splitlines()
    with open(os.path.join(golden, "mixedPoisson.ref")) as fh:
        ref = fh.read().splitlines()
    assert len(out) == len(ref)
    assert out[0] == ref[0]
    for a, b in zip(out[1:], ref[1:]):
        ka, va = a.rsplit(" ", 1)
        kb, vb = b.rsplit(" ", 1)
        assert ka == kb and abs(float(va) - float(vb)) <= 1e-5 * abs(float(vb)), (a, b)
    got = read_vtk(str(tmp_path / "square_020.vtk"))
    want = read_vtk(os.path.join(golden, "mixedPoisson.ref.vtk"))
    np.testing.assert_array_equal(got["points"], want["points"])
    assert got["cells"] == want["cells"] and got["cell_types"] == want["cell_types"]
    for sec, name in (("point_data", "temperature"), ("cell_data", "flux")):
        g, w = got[sec][name], want[sec][name]
        assert g.shape == w.shape
        tol = 1e-5 * np.maximum(np.abs(w), np.abs(w).max(axis=1, keepdims=True) * 1e-3) + 1e-9
        bad = np.argwhere(np.abs(g - w) > tol)
        assert bad.size == 0, (name, [(tuple(b), g[tuple(b)], w[tuple(b)]) for b in bad[:5]])


@pytest.mark.parametrize("order", [1, 2])
def test_diffusion_driver_app(oracle, order):
    """heat::DiffusionDriver on the drop-in layer (device time steps): the
    nodal field follows the exact discrete decay of the sin mode, and the
    device errorComputation<0> equals the oracle's error norm of that field
    against exp(-2 pi^2 t) sin(pi x) sin(pi y)."""
    import math
    N, steps, dt = 16, 6, 0.004
    out = _run("diffusion", N, order, steps, dt, timeout=120)
    vals = {}
    for ln in out.splitlines():
        if ln.startswith("# "):
            k, v = ln[2:].rsplit(" ", 1)
            vals[k] = float(v)
    assert vals["discrete deviation"] < 1e-11
    prob = P.poisson2d(N)
    u = vals["y_n"] * np.sin(np.pi * prob.xyz[:, 0]) * np.sin(np.pi * prob.xyz[:, 1])
    t = steps * dt
    ref = oracle.error_norm(prob, u, 0, oracle.REF_SINPROD, [math.exp(-2 * math.pi ** 2 * t), math.pi, math.pi])
    assert abs(vals["L2 error"] - ref) <= 1e-7 * ref


def _spsolve(rp, ci, val, b):
    from scipy.sparse import csr_matrix
    from scipy.sparse.linalg import spsolve
    n = rp.shape[0] - 1
    return spsolve(csr_matrix((val, ci, rp), shape=(n, n)).tocsc(), b)


def _vertex_values(prob, u, dpn):
    """field values at the mesh vertices (the Q2 field's vertex DoF objects)"""
    npe = prob.conn.shape[1]
    fnode = np.zeros(prob.xyz.shape[0], np.int64)
    fnode[prob.conn.ravel()] = prob.fconn[:, :npe].ravel()
    return u.reshape(-1, dpn)[fnode]


@pytest.mark.parametrize("mesh,dim", [("square_20.smf", 2), ("cube.004.smf", 3)])
def test_poisson_driver_app(oracle, golden, read_vtk, tmp_path, mesh, dim):
    """reference/05-mixedPoisson/mixedPoissonWithDriver.cpp's shape on the
    drop-in layer (base::BoundaryValueProblem<Mesh, 1, 2> and
    heat::PoissonDriver, the Dirichlet function on the DoF objects, f(x)
    and ff(x, n) as host functions evaluated by the layer at the quadrature
    and surface points, direct solve): the same discrete problem in the
    oracle -- Q2 field, kernel degree 4 (3 Gauss points per direction,
    volume and surface), its built-in cosine body force and co-normal flux
    -- solved by a sparse direct solve: |u| over the DoF objects to 1e-10,
    the nodal values of test.vtk to its printed digits, in 2-D and 3-D."""
    import shutil
    shutil.copy(os.path.join(golden, mesh), tmp_path / mesh)
    out = _run("mixed_poisson_driver", tmp_path / mesh, dim, cwd=str(tmp_path)).splitlines()
    ndof, unorm = int(out[0].split()[-1]), float(out[1].split()[-1])
    xyz, conn = P.read_smf(os.path.join(golden, mesh), dim)
    prob = P.mixed_poisson_q2(xyz, conn)
    prob.nq = 3
    assert ndof == prob.n_active
    rp, ci, val, b, _ = oracle.assemble(prob)
    x = _spsolve(rp, ci, val, b)
    u = prob.g.copy()
    act = prob.dof_id >= 0
    u[act] = x[prob.dof_id[act]]
    assert abs(unorm - np.linalg.norm(u)) <= 1e-10 * np.linalg.norm(u), (unorm, np.linalg.norm(u))
    vtk = read_vtk(str(tmp_path / "test.vtk"))
    got = np.asarray(vtk["point_data"]["u"], np.float64).reshape(-1)
    want = _vertex_values(prob, u, 1).reshape(-1)
    assert np.max(np.abs(got - want)) <= 1e-5 * np.max(np.abs(want))


def test_compressible_driver_app(oracle, golden, read_vtk, tmp_path):
    """reference/06-elastic/compressibleWithDriver.cpp's shape on the drop-in
    layer (base::BoundaryValueProblem<Mesh, 2, 2>, solid::CompressibleDriver,
    the PulledSheet Dirichlet function on the DoF objects, 5 load steps of
    u_x = 3/5 per step at x = 1): every load step converges within maxIter,
    and the final field equals the oracle's full Newton iteration on the
    same discrete problem (kernel degree 4: 3 Gauss points per direction,
    incremental lifting, sparse direct solves) -- |u| to 1e-9, the nodal
    displacements of the VTK file to its printed digits."""
    import shutil
    shutil.copy(os.path.join(golden, "square_20.smf"), tmp_path / "quad.020.smf")
    (tmp_path / "input.dat").write_text(
        "E 1000.\nnu 0.3\npull 3.0\nmaxIter 30\nloadSteps 5\ntolerance 1.e-12\n")
    out = _run("compressible_driver", tmp_path / "quad.020.smf", tmp_path / "input.dat",
               timeout=300).splitlines()
    steps = [tuple(int(v) for v in ln.split()) for ln in out[:5]]
    assert [s for s, _ in steps] == list(range(5))
    assert all(1 <= it < 30 for _, it in steps), steps
    unorm = float(out[5].split()[-1])
    prob = P.pulled_sheet_q2(20, True)
    prob.nq = 3
    n, dim = prob.n_active, prob.dim
    u = np.zeros(prob.n_fnodes * dim)
    g = np.zeros_like(u)
    con = prob.dof_id < 0
    act = ~con
    for step in range(5):
        g[prob.meta["right_x"] * dim] = 0.6 * (step + 1)
        prob.g = g
        for it in range(30):
            rp, ci, val, b, _ = oracle.assemble(prob, u=u, incremental=True, residual=True)
            if np.linalg.norm(b) / n < 1e-12:
                break
            dx = _spsolve(rp, ci, val, b)
            u[act] += dx[prob.dof_id[act]]
            u[con] = g[con]
            if np.linalg.norm(dx) / n < 1e-12:
                break
    ref = np.linalg.norm(u)
    assert abs(unorm - ref) <= 1e-9 * ref, (unorm, ref)
    vtk = read_vtk(str(tmp_path / "quad.020.vtk"))
    got = np.asarray(vtk["point_data"]["u"], np.float64)[:, :dim]
    want = _vertex_values(prob, u, dim)
    assert np.max(np.abs(got - want)) <= 1e-5 * np.max(np.abs(want))
