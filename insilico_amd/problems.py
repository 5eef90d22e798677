"""Host-side problem setup: the reference's mesh, DoF and boundary-condition
conventions, flattened into the arrays the C ABI takes.

This restates the setup half of the reference that defines the index space of
the hot path (it runs once per problem, on the host):

* ``unit_cube``: tools/meshGeneration/unitCube/unitCube.hpp:85-106 (nodes,
  first index fastest, ``x = h*i`` with ``h = 1/e``) and :217-257 (elements,
  first index fastest, lexicographic corners reordered to the hierarchic
  vertex order of base/mesh/HierarchicOrder.hpp:201-272,345-392).
* ``boundary_faces``: faces of the element face table
  (base/mesh/ElementFaces.hpp:284-286 QUAD, :422-434 HEX) that belong to
  exactly one element (base/mesh/MeshBoundary.hpp:57-62,
  base/mesh/createBoundaryFromUnstructured.hpp:51-109).
* ``number_dofs``: iso-parametric DoF = node (base/dof/copyConnectivity.hpp
  :34-60), components interleaved, CONSTRAINED components skipped
  (base/dof/numbering.hpp:45-66).
* random element fields: SplitMix64 + Box-Muller, drawn in element order.
* ``q2_field``: DoF objects of a degree-2 Lagrange field on a Q1 mesh, as
  base/dof/IndexMap.hpp:222-280 delegates to
  base/dof/generateDoFIndicesFromFaces.hpp (vertex, edge, face and cell
  passes over the elements in order; shared n-faces keep the number of their
  first appearance, keyed by the sorted vertex ids).
"""
from dataclasses import dataclass, field
import math

import numpy as np

LAPLACE, STVENANT, NEOHOOKE = 1, 2, 3
BF_NONE, BF_CONST, BF_SINPROD, BF_COSPROD = 0, 1, 2, 3
REF_COSPROD = 3

# lexicographic corner -> hierarchic position (HierarchicOrder.hpp, K = 1)
_H2 = np.array([0, 1, 3, 2])
_H3 = np.array([0, 1, 3, 2, 4, 5, 7, 6])
# element faces in hierarchic vertex numbers (ElementFaces.hpp)
HEX_FACES = np.array([[0, 3, 2, 1], [4, 5, 6, 7], [0, 1, 5, 4],
                      [1, 2, 6, 5], [2, 3, 7, 6], [3, 0, 4, 7]])
QUAD_EDGES = np.array([[0, 1], [1, 2], [2, 3], [3, 0]])


@dataclass
class Problem:
    """Flat description of one boundary-value problem."""
    dim: int
    xyz: np.ndarray            # (n_nodes, dim) float64
    conn: np.ndarray           # (n_elem, 2**dim) int32, hierarchic order
    dpn: int                   # dofs per node
    dof_id: np.ndarray         # (n_nodes*dpn,) int32, -1 = constrained
    n_active: int
    g: np.ndarray              # (n_nodes*dpn,) prescribed values
    kernel: int
    coef: np.ndarray           # (n_elem, ncoef)
    nq: int = 2
    bf_kind: int = BF_NONE
    bf: tuple = (0.0, 0.0, 0.0, 0.0)
    faces: np.ndarray = None   # (n_faces, 2**(dim-1)) int32
    traction: tuple = (0.0, 0.0, 0.0)
    name: str = ""
    meta: dict = field(default_factory=dict)
    # degree-2 field on the Q1 geometry above: dof_id and g are then indexed
    # by field node (fconn), faces stay geometric and ffaces lists the field
    # nodes of each Neumann face in its hierarchic order
    deg: int = 1
    fconn: np.ndarray = None   # (n_elem, 3**dim) int32, hierarchic order
    fxyz: np.ndarray = None    # (n_fnodes, dim) support points
    ffaces: np.ndarray = None  # (n_faces, 3**(dim-1)) int32

    @property
    def n_nodes(self):
        return self.xyz.shape[0]

    @property
    def n_fnodes(self):
        """field nodes (DoF objects): geometry nodes for Q1"""
        return self.fxyz.shape[0] if self.deg == 2 else self.xyz.shape[0]

    @property
    def n_elem(self):
        return self.conn.shape[0]


# ---------------------------------------------------------------------------
# mesh
def unit_cube(dim, n1, n2=None, n3=None):
    """Structured Q1 mesh of the unit square/cube, reference conventions."""
    n2 = n1 if n2 is None else n2
    n3 = n1 if n3 is None else n3
    h = [1.0 / float(n1), 1.0 / float(n2), 1.0 / float(n3)]
    cnt = [n1 + 1, n2 + 1, (n3 + 1) if dim == 3 else 1]
    axes = [h[d] * np.arange(cnt[d], dtype=np.float64) for d in range(dim)]
    if dim == 2:
        Y, X = np.meshgrid(axes[1], axes[0], indexing="ij")
        xyz = np.stack([X.ravel(), Y.ravel()], axis=1)
    else:
        Z, Y, X = np.meshgrid(axes[2], axes[1], axes[0], indexing="ij")
        xyz = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
    N1, N2 = cnt[0], cnt[1]
    if dim == 2:
        i2, i1 = np.meshgrid(np.arange(n2), np.arange(n1), indexing="ij")
        base = (i1 + i2 * N1).ravel()
        lex = [base + d2 * N1 + d1 for d2 in (0, 1) for d1 in (0, 1)]
        H = _H2
    else:
        i3, i2, i1 = np.meshgrid(np.arange(n3), np.arange(n2), np.arange(n1), indexing="ij")
        base = (i1 + i2 * N1 + i3 * N1 * N2).ravel()
        lex = [base + d3 * N1 * N2 + d2 * N1 + d1
               for d3 in (0, 1) for d2 in (0, 1) for d1 in (0, 1)]
        H = _H3
    conn = np.empty((base.shape[0], 2 ** dim), np.int32)
    for v, col in enumerate(lex):
        conn[:, H[v]] = col
    return np.ascontiguousarray(xyz), conn


def round_smf(xyz):
    """Coordinates as the reference's SMF writer prints them: default ostream
    precision, 6 significant digits (tools/meshGeneration/meshGeneration.hpp
    :32-36)."""
    f = np.vectorize(lambda v: float("%g" % v))
    return f(xyz).astype(np.float64)


def read_smf(path, dim):
    """Minimal reader of the SMF mesh format (base/io/smf/Reader.hpp:97-312):
    header lines, ``nNodes nElements``, node coordinates, connectivity."""
    with open(path) as fh:
        lines = [ln.strip() for ln in fh if ln.strip() and not ln.startswith("#")]
    k = 0
    while lines[k].startswith("!"):
        k += 1
    nn, ne = (int(v) for v in lines[k].split())
    k += 1
    xyz = np.array([[float(v) for v in lines[k + i].split()[:dim]] for i in range(nn)])
    k += nn
    conn = np.array([[int(v) for v in lines[k + i].split()] for i in range(ne)], np.int32)
    return xyz, conn


def read_sgf(path):
    """Structured grid file (base/io/sgf/Reader.hpp:155-310): '#' comment
    lines, the element counts per direction (three numbers, trailing zeros
    for fewer dimensions), then the node coordinates in lexicographic order
    (first index fastest), one node per line.  The elements are the grid
    cells in lexicographic order; returned as the Q1 mesh of those cells
    with their vertices in hierarchic order: (xyz, conn, sizes)."""
    with open(path) as fh:
        lines = [ln.split() for ln in fh if ln.strip() and not ln.startswith("#")]
    sizes = [int(v) for v in lines[0][:3]]
    dim = sum(1 for v in sizes if v > 0)
    sizes = sizes[:dim]
    nn = int(np.prod([v + 1 for v in sizes]))
    xyz = np.array([[float(v) for v in ln[:dim]] for ln in lines[1:1 + nn]])
    _, conn = unit_cube(dim, *sizes)
    return xyz, conn, tuple(sizes)


def boundary_faces(conn, dim):
    """(element, face) pairs on the boundary and their face node lists."""
    table = HEX_FACES if dim == 3 else QUAD_EDGES
    nf = table.shape[0]
    fn = conn[:, table].reshape(-1, table.shape[1])          # (E*nf, nfn)
    # a face is on the boundary when its sorted vertex ids occur once; the
    # sorted ids packed two per 64-bit key, the keys sorted together
    key = np.sort(fn, axis=1).astype(np.uint64)
    k1 = (key[:, 0] << np.uint64(32)) | key[:, 1]
    k2 = ((key[:, 2] << np.uint64(32)) | key[:, 3]) if key.shape[1] == 4 else np.zeros_like(k1)
    order = np.lexsort((k2, k1))
    s1, s2 = k1[order], k2[order]
    new = np.ones(order.shape[0], bool)
    new[1:] = (s1[1:] != s1[:-1]) | (s2[1:] != s2[:-1])
    grp = np.cumsum(new) - 1
    on = np.empty(order.shape[0], bool)
    on[order] = np.bincount(grp)[grp] == 1
    idx = np.nonzero(on)[0]
    return idx // nf, idx % nf, fn[idx].astype(np.int32)


# element n-faces in hierarchic vertex numbers, in the order of the element's
# DoF array (base/fe/Policies.hpp:43-77: vertices, edges, faces, cell) and of
# the hierarchic shape functions (base/mesh/HierarchicOrder.hpp:201-452)
HEX_EDGES = np.array([[0, 1], [1, 2], [2, 3], [3, 0], [4, 5], [5, 6],
                      [6, 7], [7, 4], [0, 4], [1, 5], [2, 6], [3, 7]])


def _first_appearance(keys):
    """number distinct rows of `keys` (m, k) by their first appearance"""
    _, first, inv = np.unique(keys, axis=0, return_index=True, return_inverse=True)
    rank = np.empty(first.shape[0], np.int64)
    rank[np.argsort(first, kind="stable")] = np.arange(first.shape[0])
    return rank[inv.ravel()], first.shape[0]


def q2_field(xyz, conn, dim):
    """Degree-2 field nodes of a Q1 mesh: (fconn (E, 3**dim), fxyz).

    generateDoFIndicesFromFaces (base/dof/generateDoFIndicesFromFaces.hpp)
    runs one pass per n-face kind: vertices, then edges (QUAD/HEX edge tables
    of base/mesh/ElementFaces.hpp:282-286,339-347), then faces (HEX face
    table :421-434; the cell of a QUAD), then the HEX cell.  Each pass walks
    the elements in order and their n-faces in table order; an n-face seen
    before (same sorted vertex ids) keeps its number, a new one takes the
    next number.  A pass of co-dimension 0 (the cell) numbers every element
    anew.  Support points: the Q1 geometry at the Lagrange points, i.e. the
    vertex, the edge midpoint, the face and cell centres, summed in vertex
    order (base/dof/constrainBoundary.hpp:77-110 with base/geometry.hpp)."""
    conn = np.asarray(conn, np.int64)
    E = conn.shape[0]
    nv = 2 ** dim
    edges = QUAD_EDGES if dim == 2 else HEX_EDGES
    faces = None if dim == 2 else HEX_FACES
    ncol = 3 ** dim
    fconn = np.empty((E, ncol), np.int64)
    pts = []
    start = 0
    # vertices
    r, m = _first_appearance(conn.reshape(-1, 1))
    fconn[:, :nv] = r.reshape(E, nv)
    vx = np.empty((m, dim))
    vx[r] = xyz[conn.reshape(-1)]
    pts.append(vx)
    start += m
    col = nv
    groups = [edges] + ([faces] if faces is not None else [])
    for tab in groups:
        verts = conn[:, tab]                                   # (E, nf, k)
        key = np.sort(verts, axis=2).reshape(-1, tab.shape[1])
        r, m = _first_appearance(key)
        fconn[:, col:col + tab.shape[0]] = start + r.reshape(E, tab.shape[0])
        c = np.empty((m, dim))
        sm = xyz[verts[..., 0]]
        for j in range(1, tab.shape[1]):
            sm = sm + xyz[verts[..., j]]
        c[r] = (sm * (1.0 / tab.shape[1])).reshape(-1, dim)
        pts.append(c)
        start += m
        col += tab.shape[0]
    # cell
    fconn[:, col] = start + np.arange(E)
    sm = xyz[conn[:, 0]]
    for j in range(1, nv):
        sm = sm + xyz[conn[:, j]]
    pts.append(sm * (1.0 / nv))
    return fconn.astype(np.int32), np.concatenate(pts, axis=0)


def q2_face_nodes(fconn, elem, face, dim):
    """Field nodes of the Neumann faces (element, local face) in the face's
    own hierarchic order: LINE3 = (v0, v1, mid) for a QUAD edge, QUAD9 =
    (w0..w3, edges (w0,w1), (w1,w2), (w2,w3), (w3,w0), centre) for a HEX face
    whose vertices w are in the face table order (ElementFaces.hpp)."""
    elem = np.asarray(elem)
    face = np.asarray(face)
    if dim == 2:
        a, b = QUAD_EDGES[face, 0], QUAD_EDGES[face, 1]
        return np.stack([fconn[elem, a], fconn[elem, b], fconn[elem, 4 + face]], 1).astype(np.int32)
    lut = {}
    for i, (a, b) in enumerate(HEX_EDGES):
        lut[(a, b)] = lut[(b, a)] = i
    out = np.empty((elem.shape[0], 9), np.int32)
    for f in range(6):
        sel = face == f
        if not sel.any():
            continue
        w = HEX_FACES[f]
        cols = list(w) + [8 + lut[(w[i], w[(i + 1) % 4])] for i in range(4)] + [20 + f]
        out[sel] = fconn[elem[sel]][:, cols]
    return out


def q1_shape(dim, xi):
    """Q1 shape functions at the points xi (nq, dim), vertices in
    hierarchic order (base/LagrangeShapeFun.hpp:64-86): (nq, 2^dim)"""
    xi = np.atleast_2d(np.asarray(xi, np.float64))
    H = _H3 if dim == 3 else _H2
    phi = np.ones((xi.shape[0], 1 << dim))
    for lex in range(1 << dim):
        for d in range(dim):
            bit = (lex >> d) & 1
            phi[:, H[lex]] *= xi[:, d] if bit else 1.0 - xi[:, d]
    return phi


def element_points(xyz, conn, xi):
    """x(xi_q) of the Q1 geometry of every element at the local points xi
    (nq, dim): (E, nq, dim)"""
    dim = xyz.shape[1]
    phi = q1_shape(dim, xi)                       # (nq, npe)
    X = xyz[conn]                                 # (E, npe, dim)
    return np.einsum("qn,end->eqd", phi, X)


def face_points(xyz, faces, eta):
    """Surface points x(eta_q) and unit normals of base::SurfaceNormal
    (base/geometry.hpp:255-346: the cross product of the surface Jacobian's
    columns in 3-D, (J10, -J00) in 2-D) of every face at the local points
    eta (nq, dim - 1), faces listed in the face table's order: (x, n), each
    (n_faces, nq, dim)"""
    dim = xyz.shape[1]
    X = xyz[faces]                                # (F, nfn, dim)
    eta = np.atleast_2d(np.asarray(eta, np.float64))
    if dim == 3:
        e0, e1 = eta[:, 0], eta[:, 1]
        phi = np.stack([(1 - e0) * (1 - e1), e0 * (1 - e1), e0 * e1, (1 - e0) * e1], 1)
        d0 = np.stack([-(1 - e1), 1 - e1, e1, -e1], 1)
        d1 = np.stack([-(1 - e0), -e0, e0, 1 - e0], 1)
        J0 = np.einsum("qn,fnd->fqd", d0, X)
        J1 = np.einsum("qn,fnd->fqd", d1, X)
        nrm = np.cross(J0, J1)
    else:
        e0 = eta[:, 0]
        phi = np.stack([1 - e0, e0], 1)
        t = X[:, 1, :] - X[:, 0, :]
        nrm = np.repeat(np.stack([t[:, 1], -t[:, 0]], 1)[:, None, :], eta.shape[0], 1)
    x = np.einsum("qn,fnd->fqd", phi, X)
    return x, nrm / np.linalg.norm(nrm, axis=2, keepdims=True)


def number_dofs(constrained):
    """numberDoFsConsecutively: constrained (n_nodes, dpn) bool -> dof ids."""
    c = np.asarray(constrained, bool).ravel()
    ids = np.full(c.shape[0], -1, np.int64)
    act = ~c
    ids[act] = np.arange(int(act.sum()))
    return ids.astype(np.int32), int(act.sum())


# ---------------------------------------------------------------------------
# random element fields
_GAMMA = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(seed, n, offset=0):
    """Outputs offset..offset+n-1 of SplitMix64 seeded with `seed`."""
    with np.errstate(over="ignore"):
        k = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        z = np.uint64(seed) + k * _GAMMA
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed, n, offset=0):
    return (splitmix64(seed, n, offset) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def normal(seed, n):
    """Box-Muller over consecutive uniform pairs, one N(0,1) per element."""
    u = uniform01(seed, 2 * n)
    u1, u2 = u[0::2], u[1::2]
    return np.sqrt(-2.0 * np.log1p(-u1)) * np.cos(2.0 * np.pi * u2)


def lame(E, nu):
    """mat/Lame.hpp:28-40"""
    lam = E * nu / (1.0 + nu) / (1.0 - 2.0 * nu)
    mu = E / 2.0 / (1.0 + nu)
    return lam, mu


# ---------------------------------------------------------------------------
# benchmark configurations
def poisson2d(n=64):
    """-Lap u = 2 pi^2 sin(pi x) sin(pi y) on the unit square, u = 0 on the
    boundary, Q1, 2x2 Gauss (exact solution sin(pi x) sin(pi y))."""
    xyz, conn = unit_cube(2, n)
    _, _, bf = boundary_faces(conn, 2)
    con = np.zeros(xyz.shape[0], bool)
    con[np.unique(bf)] = True
    dof, na = number_dofs(con[:, None])
    pi = math.pi
    return Problem(2, xyz, conn, 1, dof, na, np.zeros(xyz.shape[0]), LAPLACE,
                   np.ones((conn.shape[0], 1)), 2, BF_SINPROD, (2 * pi * pi, pi, pi, 0.0),
                   name="poisson2d-%d" % n)


def diffusion3d(n=128, seed=12345):
    """-div(k grad u) = 1 on the unit cube, k_e = exp(g_e), g_e ~ N(0,1) per
    element, u = 0 on x = 0, natural elsewhere; Q1, 2x2x2 Gauss."""
    xyz, conn = unit_cube(3, n)
    _, _, bf = boundary_faces(conn, 3)
    bn = np.unique(bf)
    con = np.zeros(xyz.shape[0], bool)
    con[bn[np.abs(xyz[bn, 0]) < 1e-12]] = True
    dof, na = number_dofs(con[:, None])
    k = np.exp(normal(seed, conn.shape[0]))[:, None]
    return Problem(3, xyz, conn, 1, dof, na, np.zeros(xyz.shape[0]), LAPLACE, k, 2,
                   BF_CONST, (1.0, 0.0, 0.0, 0.0), name="diffusion3d-%d" % n,
                   meta={"seed": seed})


def elasticity3d(n=96, seed=12345, nu=0.3):
    """Linear (StVenant at u = 0) elasticity on the unit cube, E_e ~ U[0.5, 2]
    per element, clamped at z = 0, traction (0, 0, -0.01) on z = 1."""
    xyz, conn = unit_cube(3, n)
    _, _, bf = boundary_faces(conn, 3)
    bn = np.unique(bf)
    con = np.zeros((xyz.shape[0], 3), bool)
    con[bn[np.abs(xyz[bn, 2]) < 1e-12], :] = True
    dof, na = number_dofs(con)
    E = 0.5 + 1.5 * uniform01(seed, conn.shape[0])
    lam, mu = lame(E, nu)
    top = bf[np.all(np.abs(xyz[bf, 2] - 1.0) < 1e-12, axis=1)]
    return Problem(3, xyz, conn, 3, dof, na, np.zeros(xyz.shape[0] * 3), STVENANT,
                   np.stack([lam, mu], axis=1), 2, faces=top, traction=(0.0, 0.0, -0.01),
                   name="elasticity3d-%d" % n, meta={"seed": seed})


def neohooke3d(n=64, mu=1.0, lam=10.0):
    """Compressible neo-Hookean block: clamped at x = 0, u_x prescribed on
    x = 1 (reference/06-elastic/PulledSheet.hpp: left face all components,
    right face component 0); the prescribed value is set per load step."""
    xyz, conn = unit_cube(3, n)
    _, _, bf = boundary_faces(conn, 3)
    bn = np.unique(bf)
    con = np.zeros((xyz.shape[0], 3), bool)
    left = bn[np.abs(xyz[bn, 0]) < 1e-5]
    right = bn[np.abs(xyz[bn, 0] - 1.0) < 1e-5]
    con[left, :] = True
    con[right, 0] = True
    dof, na = number_dofs(con)
    coef = np.tile(np.array([[lam, mu]]), (conn.shape[0], 1))
    g = np.zeros(xyz.shape[0] * 3)
    p = Problem(3, xyz, conn, 3, dof, na, g, NEOHOOKE, coef, 2, name="neohooke3d-%d" % n)
    p.meta["right_x"] = right
    return p


def _q2_constrained(fxyz, fnodes_bnd, pred, dpn):
    con = np.zeros((fxyz.shape[0], dpn), bool)
    sel = fnodes_bnd[pred(fxyz[fnodes_bnd])]
    con[sel, :] = True
    return con


def diffusion3d_q2(n=32, seed=12345):
    """The 3-D diffusion problem with a degree-2 field: Q2 (HEX27) on the Q1
    unit-cube mesh, k_e = exp(g_e), u = 0 on x = 0, f = 1, 3 Gauss points per
    direction (kernel degree 2*fieldDeg = 4, heat/PoissonDriver.hpp:191)."""
    xyz, conn = unit_cube(3, n)
    fconn, fxyz = q2_field(xyz, conn, 3)
    el, fa, _ = boundary_faces(conn, 3)
    bn = np.unique(q2_face_nodes(fconn, el, fa, 3))
    con = _q2_constrained(fxyz, bn, lambda x: np.abs(x[:, 0]) < 1e-12, 1)
    dof, na = number_dofs(con)
    k = np.exp(normal(seed, conn.shape[0]))[:, None]
    return Problem(3, xyz, conn, 1, dof, na, np.zeros(fxyz.shape[0]), LAPLACE, k, 3,
                   BF_CONST, (1.0, 0.0, 0.0, 0.0), name="diffusion3dQ2-%d" % n,
                   meta={"seed": seed}, deg=2, fconn=fconn, fxyz=fxyz)


def poisson2d_q2(n=16):
    """-Lap u = 2 pi^2 sin(pi x) sin(pi y), u = 0 on the boundary, Q2 (QUAD9),
    3 Gauss points per direction."""
    xyz, conn = unit_cube(2, n)
    fconn, fxyz = q2_field(xyz, conn, 2)
    el, fa, _ = boundary_faces(conn, 2)
    bn = np.unique(q2_face_nodes(fconn, el, fa, 2))
    con = np.zeros((fxyz.shape[0], 1), bool)
    con[bn] = True
    dof, na = number_dofs(con)
    pi = math.pi
    return Problem(2, xyz, conn, 1, dof, na, np.zeros(fxyz.shape[0]), LAPLACE,
                   np.ones((conn.shape[0], 1)), 3, BF_SINPROD, (2 * pi * pi, pi, pi, 0.0),
                   name="poisson2dQ2-%d" % n, deg=2, fconn=fconn, fxyz=fxyz)


def elasticity3d_q2(n=8, seed=12345, nu=0.3, kernel=STVENANT, traction=None):
    """Elastic block with a degree-2 field: E_e ~ U[0.5, 2], clamped at z = 0,
    body force (0, 0, -1), optionally a traction on z = 1, 2 Gauss points per
    direction."""
    xyz, conn = unit_cube(3, n)
    fconn, fxyz = q2_field(xyz, conn, 3)
    el, fa, bf = boundary_faces(conn, 3)
    fnodes = q2_face_nodes(fconn, el, fa, 3)
    bn = np.unique(fnodes)
    con = _q2_constrained(fxyz, bn, lambda x: np.abs(x[:, 2]) < 1e-12, 3)
    dof, na = number_dofs(con)
    E = 0.5 + 1.5 * uniform01(seed, conn.shape[0])
    lam, mu = lame(E, nu)
    p = Problem(3, xyz, conn, 3, dof, na, np.zeros(fxyz.shape[0] * 3), kernel,
                np.stack([lam, mu], axis=1), 2, BF_CONST, (0.0, 0.0, -1.0, 0.0),
                name="elasticity3dQ2-%d" % n, meta={"seed": seed}, deg=2, fconn=fconn,
                fxyz=fxyz)
    if traction is not None:
        top = np.all(np.abs(xyz[bf, 2] - 1.0) < 1e-12, axis=1)
        p.faces = bf[top].astype(np.int32)
        p.ffaces = fnodes[top]
        p.traction = tuple(traction)
    return p


def pulled_sheet_q2(n=20, disp_controlled=True, E=1000.0, nu=0.3, dim=2):
    """reference/06-elastic/compressible.cpp:108-240 with
    reference/06-elastic/PulledSheet.hpp: neo-Hookean block (0,1)^dim, Q2
    field on the Q1 unit-square/cube mesh, Gauss rule of degree 3 (2 points
    per direction, volume and surface); x = 0 clamped (all components);
    x = 1 either u_x prescribed (displacement control, value set per load
    step) or a traction t_x = value * (y - 1/2) (set per load step)."""
    xyz, conn = unit_cube(dim, n)
    fconn, fxyz = q2_field(xyz, conn, dim)
    el, fa, bf = boundary_faces(conn, dim)
    fnodes = q2_face_nodes(fconn, el, fa, dim)
    bnodes = np.unique(fnodes)
    tol = 1e-5
    con = np.zeros((fxyz.shape[0], dim), bool)
    left = bnodes[np.abs(fxyz[bnodes, 0]) < tol]
    right = bnodes[np.abs(fxyz[bnodes, 0] - 1.0) < tol]
    con[left, :] = True
    if disp_controlled:
        con[right, 0] = True
    dof, na = number_dofs(con)
    lam, mu = lame(E, nu)
    coef = np.tile(np.array([[lam, mu]]), (conn.shape[0], 1))
    on_right = np.all(np.abs(xyz[bf, 0] - 1.0) < tol, axis=1)
    p = Problem(dim, xyz, conn, dim, dof, na, np.zeros(fxyz.shape[0] * dim), NEOHOOKE,
                coef, 2, name="pulledSheetQ2-%d%s" % (n, "D" if disp_controlled else "N"),
                deg=2, fconn=fconn, fxyz=fxyz)
    if not disp_controlled:
        p.faces = bf[on_right].astype(np.int32)
        p.ffaces = fnodes[on_right]
        p.traction = (0.0, 0.0, 0.0)
        p.meta["traction_linear"] = (1, 0.5)
    p.meta.update(right_x=right, E=E, lam=lam, mu=mu)
    return p


def mixed_poisson_q2(xyz, conn, alpha=(3.0, 5.0, 4.0), deg=2):
    """reference/05-mixedPoisson/mixedPoisson.cpp:85-204 with GivenData.hpp
    and ReferenceSolution.hpp: -Lap u = f on the mesh with a degree-2 field
    (deg=1: the iso-parametric Q1 field), u = prod_d cos(alpha_d x_d)
    prescribed at the boundary support points with some x_d = 0, the
    co-normal derivative grad u . n as Neumann data on every boundary face
    (its contributions to constrained rows drop out), f = -Lap u =
    (sum_d alpha_d^2) prod_d cos(alpha_d x_d), kappa = 1, Gauss rule of
    degree 3 (2 points per direction, volume and surface)."""
    xyz = np.ascontiguousarray(xyz, np.float64)
    conn = np.ascontiguousarray(conn, np.int32)
    dim = xyz.shape[1]
    al = np.array(alpha[:dim], np.float64)
    el, fa, bf = boundary_faces(conn, dim)
    if deg == 2:
        fconn, fxyz = q2_field(xyz, conn, dim)
        fnodes = q2_face_nodes(fconn, el, fa, dim)
    else:
        fconn, fxyz, fnodes = None, xyz, bf
    bn = np.unique(fnodes)
    con = np.zeros((fxyz.shape[0], 1), bool)
    dir_nodes = bn[np.any(np.abs(fxyz[bn]) < 1e-5, axis=1)]
    con[dir_nodes, 0] = True
    dof, na = number_dofs(con)
    g = np.zeros(fxyz.shape[0])
    val = np.ones(dir_nodes.shape[0])
    for d in range(dim):
        val = val * np.cos(al[d] * fxyz[dir_nodes, d])
    g[dir_nodes] = val
    amp = float(np.sum(al * al))
    bfp = (amp,) + tuple(al) + (0.0,) * (3 - dim)
    p = Problem(dim, xyz, conn, 1, dof, na, g, LAPLACE, np.ones((conn.shape[0], 1)), 2,
                BF_COSPROD, bfp, faces=bf.astype(np.int32), traction=(0.0, 0.0, 0.0),
                name="mixedPoissonQ%d" % deg)
    if deg == 2:
        p.deg, p.fconn, p.fxyz, p.ffaces = 2, fconn, fxyz, fnodes
    p.meta["neumann_flux"] = (REF_COSPROD, [1.0] + list(al) + [0.0] * (7 - dim))
    p.meta["alpha"] = al
    return p


def fundsol_elasto(dim, lam, mu, y, x, direc):
    """Kelvin solution U(y, x) dir (base/auxi/FundamentalSolution.hpp:160-200),
    evaluated at the points y (m, dim) for the source x."""
    G = mu
    nu = lam / 2.0 / (lam + mu)
    surf = 2.0 * (dim - 1.0) * math.pi
    fac1 = 1.0 / (4.0 * surf * G * (1.0 - nu))
    fac2 = 3.0 - 4.0 * nu
    R = np.asarray(y, np.float64) - np.asarray(x, np.float64)[None, :]
    dist = np.sqrt(np.sum(R * R, axis=1))
    lk = np.log(1.0 / dist) if dim == 2 else 1.0 / dist
    pw = dist ** 2 if dim == 2 else dist ** 3
    d = np.asarray(direc, np.float64)
    out = np.zeros_like(R)
    for i in range(dim):
        s = np.zeros(R.shape[0])
        for j in range(dim):
            Uij = fac1 * ((fac2 * lk if i == j else 0.0) + R[:, i] * R[:, j] / pw)
            s = s + Uij * d[j]
        out[:, i] = s
    return out


def distorted_block(dim, n, deg=1, kernel=STVENANT, lam=1.5, mu=0.75, amp=0.2, seed=7):
    """Unit square/cube of n^dim elements with the interior nodes moved by up
    to amp * h per coordinate (elements stay convex for amp < 0.25), one
    material for all elements, no constraints: a mesh on which the point
    evaluations of a field are checked (a degree-2 field takes its nodes at
    the Lagrange points of the distorted Q1 geometry)."""
    xyz, conn = unit_cube(dim, n)
    inner = np.all((xyz > 1e-12) & (xyz < 1.0 - 1e-12), axis=1)
    shift = (uniform01(seed, xyz.size).reshape(xyz.shape) - 0.5) * (2.0 * amp / n)
    xyz = np.where(inner[:, None], xyz + shift, xyz)
    coef = np.tile(np.array([[lam, mu]]), (conn.shape[0], 1))
    if kernel == LAPLACE:
        coef = np.ones((conn.shape[0], 1))
    dpn = 1 if kernel == LAPLACE else dim
    p = Problem(dim, np.ascontiguousarray(xyz), conn, dpn, None, 0, None, kernel, coef, 2,
                name="distorted%dD-%d" % (dim, n))
    if deg == 2:
        p.deg = 2
        p.fconn, p.fxyz = q2_field(xyz, conn, dim)
    nf = p.n_fnodes
    p.dof_id, p.n_active = number_dofs(np.zeros((nf, dpn), bool))
    p.g = np.zeros(nf * dpn)
    return p


def linear_elastic_fundsol(xyz, conn, dim, E=1000.0, nu=0.25, nq=2):
    """reference/06-elastic/linearElastic.cpp:80-216: Q1 StVenant, the whole
    boundary constrained to the elastostatic fundamental solution with source
    (-0.1, ...) and direction e_1."""
    lam, mu = lame(E, nu)
    _, _, bf = boundary_faces(conn, dim)
    bn = np.unique(bf)
    con = np.zeros((xyz.shape[0], dim), bool)
    con[bn, :] = True
    dof, na = number_dofs(con)
    src = np.full(dim, -0.1)
    direc = np.zeros(dim)
    direc[1] = 1.0
    g = np.zeros((xyz.shape[0], dim))
    g[bn] = fundsol_elasto(dim, lam, mu, xyz[bn], src, direc)
    coef = np.tile(np.array([[lam, mu]]), (conn.shape[0], 1))
    p = Problem(dim, np.ascontiguousarray(xyz, np.float64), np.ascontiguousarray(conn, np.int32),
                dim, dof, na, g.ravel(), STVENANT, coef, nq, name="linearElastic%dD" % dim)
    p.meta.update(lam=lam, mu=mu, src=src, dir=direc)
    return p
