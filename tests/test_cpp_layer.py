"""CPU checks of the drop-in C++ layer's host side (insilico/insilico.hpp,
insilico/driver.hpp):

* drivers written in the reference's call shape compile against it:
  base::BoundaryValueProblem + heat::PoissonDriver with a Dirichlet
  function, a body force f(x) and a Neumann function ff(x, n)
  (reference/05-mixedPoisson/mixedPoissonWithDriver.cpp's shape), and
  solid::CompressibleDriver (reference/06-elastic/compressibleWithDriver.cpp);
* the points at which the layer evaluates the caller's functions -- every
  element's quadrature points x(xi_q), every boundary face's surface points
  x(eta_q) and unit normals -- equal the ones formed independently by
  insilico_amd.problems (the points the oracle's table tests use) on the
  shipped SMF meshes, for Q1 and Q2 fields, and the normals point out of
  the domain.  No GPU: only the host code and ins_quadrature_rule run."""
import os
import subprocess

import numpy as np
import pytest

from insilico_amd import abi, problems as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = ["-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "insilico_amd", "include")]
LIBDIR = os.path.dirname(abi.LIB_PATH)


@pytest.mark.parametrize("app", ["mixed_poisson_driver", "compressible_driver"])
def test_reference_shaped_drivers_compile(app):
    src = os.path.join(ROOT, "insilico_amd", "apps", app + ".cpp")
    out = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only"] + INC + [src],
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr


@pytest.fixture(scope="module")
def probe(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("probe") / "points_probe")
    src = os.path.join(ROOT, "tests", "cpp", "points_probe.cpp")
    out = subprocess.run(["g++", "-std=c++17", "-O1", "-Wall"] + INC +
                         [src, "-o", exe, "-L" + LIBDIR, "-linsilico_hip",
                          "-Wl,-rpath," + LIBDIR],
                         capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stderr
    return exe


def _run_probe(exe, mesh, dim, degree, npts):
    out = subprocess.run([exe, mesh, str(dim), str(degree), str(npts)], capture_output=True,
                         text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.splitlines()
    _, ne, nq = lines[0].split()
    ne, nq = int(ne), int(nq)
    ep = np.array([[float(v) for v in ln.split()] for ln in lines[1:1 + ne * nq]])
    k = 1 + ne * nq
    _, nf, nqs, nfn = lines[k].split()
    nf, nqs, nfn = int(nf), int(nqs), int(nfn)
    faces = np.array([[int(v) for v in ln.split()] for ln in lines[k + 1:k + 1 + nf]], np.int64)
    fp = np.array([[float(v) for v in ln.split()] for ln in lines[k + 1 + nf:k + 1 + nf + nf * nqs]])
    return ep.reshape(ne, nq, 3), faces, fp.reshape(nf, nqs, 6)


@pytest.mark.parametrize("mesh,dim,degree,npts", [
    ("quad.005.smf", 2, 1, 2), ("quad.005.smf", 2, 2, 3),
    ("cube.004.smf", 3, 1, 2), ("cube.004.smf", 3, 2, 3), ("cube.002.smf", 3, 1, 1),
])
def test_host_evaluation_points(probe, oracle, golden, mesh, dim, degree, npts):
    path = os.path.join(golden, mesh)
    ep, faces, fp = _run_probe(probe, path, dim, degree, npts)
    xyz, conn = P.read_smf(path, dim)
    _, xi = oracle.tensor_rule(dim, npts)
    want = P.element_points(xyz, conn, xi)
    assert np.max(np.abs(ep[:, :, :dim] - want)) <= 1e-15
    # boundary faces: as many as the oracle's face table finds, corners
    # first; the support points of degree-2 nodes are those of q2_field
    _, _, bf = P.boundary_faces(conn, dim)
    assert faces.shape[0] == bf.shape[0]
    nc = 4 if dim == 3 else 2
    pts = xyz if degree == 1 else P.q2_field(xyz, conn, dim)[1]
    if degree == 2:
        assert faces.shape[1] == (9 if dim == 3 else 3)
    corners = faces[:, :nc]
    assert _face_keys(pts, corners) == _face_keys(xyz, bf)
    _, eta = oracle.tensor_rule(dim - 1, npts)
    x, n = P.face_points(pts, corners, eta)
    assert np.max(np.abs(fp[:, :, :dim] - x)) <= 1e-15
    assert np.max(np.abs(fp[:, :, 3:3 + dim] - n)) <= 1e-14
    # unit normals pointing out of the (box-shaped) domain
    assert np.allclose(np.linalg.norm(fp[:, :, 3:], axis=2), 1.0, rtol=0, atol=1e-14)
    centre = np.zeros(3)
    centre[:dim] = 0.5 * (xyz.min(axis=0) + xyz.max(axis=0))
    assert np.all(np.sum((fp[:, :, :3] - centre) * fp[:, :, 3:], axis=2) > 0)


def _face_keys(points, faces):
    return {tuple(sorted(tuple(np.round(points[v], 12)) for v in f)) for f in faces}
