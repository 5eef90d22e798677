// Drop-in C++17 host layer over the C ABI (include/insilico_hip.h).
//
// It re-exposes the names and call order of the reference's template API for
// the supported subset (Q1 QUAD4/HEX8, tensor Gauss-Legendre, heat::Laplace,
// solid::HyperElastic<StVenant | NeoHookeanCompressible>) so that an
// application written against the reference reads the same:
//
//   base::solver::Hip solver(numDoFs);                 // base/solver/Eigen3.hpp:71
//   base::asmb::stiffnessMatrixComputation<FTB>(quadrature, solver,
//                                               fieldBinder, kernel);
//   base::asmb::bodyForceComputation<FTB>(quadrature, solver, fieldBinder, f);
//   solver.finishAssembly();                           // device assembly runs here
//   solver.cgSolve();
//   base::dof::setDoFsFromSolver(solver, field);
//
// The assembly free functions do not loop over elements on the host: they
// record what is to be assembled; finishAssembly() hands mesh, DoFs, kernel
// and loads to the device once and runs the gfx950 kernels (element
// integration, deterministic CSR gather, RHS).  Matrices assembled by hand
// through insertToLHS/insertToRHS take the host triplet path of the reference
// (base/solver/TripletContainer.hpp:305-352) and are uploaded by
// finishAssembly().
//
// Errors: the reference aborts through VERIFY_MSG (base/verify.hpp:63-65);
// this layer throws insilico::Error carrying the C ABI status.
#pragma once

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <functional>
#include <map>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "insilico_hip.h"

namespace insilico {

struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};

inline void check(int s, const ins_ctx* c)
{
    if (s != INS_OK)
        throw Error(s, std::string(ins_strerror(s)) + ": " + (c ? ins_last_error(c) : ""));
}

// ---------------------------------------------------------------------------
// Discretisation (host setup; flattened base::Unstructured + fe::Field)
enum DoFStatus { ACTIVE = 0, CONSTRAINED = 1 };  // base/dof/DegreeOfFreedom.hpp:33-38

struct Mesh {
    int dim = 3;
    std::vector<double> xyz;     // node coordinates, dim per node
    std::vector<int32_t> conn;   // element vertices, hierarchic order
    int npe() const { return 1 << dim; }
    std::size_t numNodes() const { return xyz.size() / dim; }
    std::size_t numElements() const { return conn.size() / npe(); }
};

// tools/meshGeneration/unitCube/unitCube.hpp:85-106,217-257
inline Mesh unitCube(int dim, int e1, int e2, int e3 = 1)
{
    Mesh m;
    m.dim = dim;
    const int n1 = e1 + 1, n2 = e2 + 1, n3 = dim == 3 ? e3 + 1 : 1;
    const double h1 = 1.0 / e1, h2 = 1.0 / e2, h3 = dim == 3 ? 1.0 / e3 : 0.;
    for (int i3 = 0; i3 < n3; i3++)
        for (int i2 = 0; i2 < n2; i2++)
            for (int i1 = 0; i1 < n1; i1++) {
                m.xyz.push_back(h1 * i1);
                m.xyz.push_back(h2 * i2);
                if (dim == 3) m.xyz.push_back(h3 * i3);
            }
    static const int H2[4] = {0, 1, 3, 2};
    static const int H3[8] = {0, 1, 3, 2, 4, 5, 7, 6};
    const int npe = 1 << dim;
    for (int i3 = 0; i3 < (dim == 3 ? e3 : 1); i3++)
        for (int i2 = 0; i2 < e2; i2++)
            for (int i1 = 0; i1 < e1; i1++) {
                const int32_t base = i1 + i2 * n1 + i3 * n1 * n2;
                int32_t lex[8], hier[8];
                int c = 0;
                for (int d3 = 0; d3 <= (dim == 3 ? 1 : 0); d3++)
                    for (int d2 = 0; d2 <= 1; d2++)
                        for (int d1 = 0; d1 <= 1; d1++) lex[c++] = base + d3 * n1 * n2 + d2 * n1 + d1;
                for (int v = 0; v < npe; v++) hier[dim == 3 ? H3[v] : H2[v]] = lex[v];
                m.conn.insert(m.conn.end(), hier, hier + npe);
            }
    return m;
}

// base::mesh::MeshBoundary::create: (element, face) pairs on the boundary,
// faces from the element face table (base/mesh/ElementFaces.hpp:284-286,
// 422-434); returns the face node lists
inline std::vector<std::vector<int32_t>> boundaryFaces(const Mesh& m)
{
    static const int HEX[6][4] = {{0, 3, 2, 1}, {4, 5, 6, 7}, {0, 1, 5, 4},
                                  {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}};
    static const int QUAD[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
    const int nf = m.dim == 3 ? 6 : 4, nfn = m.dim == 3 ? 4 : 2, npe = m.npe();
    std::map<std::vector<int32_t>, std::pair<int, std::vector<int32_t>>> seen;
    for (std::size_t e = 0; e < m.numElements(); e++)
        for (int f = 0; f < nf; f++) {
            std::vector<int32_t> nodes(nfn);
            for (int k = 0; k < nfn; k++)
                nodes[k] = m.conn[e * npe + (m.dim == 3 ? HEX[f][k] : QUAD[f][k])];
            std::vector<int32_t> key = nodes;
            std::sort(key.begin(), key.end());
            auto it = seen.find(key);
            if (it == seen.end()) seen.emplace(key, std::make_pair(1, nodes));
            else it->second.first++;
        }
    std::vector<std::vector<int32_t>> out;
    for (auto& kv : seen)
        if (kv.second.first == 1) out.push_back(kv.second.second);
    return out;
}

// Field of DoF objects with `size` components per node (iso-parametric:
// DoF = node, base/dof/copyConnectivity.hpp:34-60; degree 2: DoF objects of
// the vertices, edges, faces and cells, see generateField)
struct Field {
    int size = 1;
    int degree = 1;
    std::vector<int32_t> conn;      // degree 2: field nodes of every element, hierarchic
    std::vector<double> support;    // degree 2: support point of every field node
    std::vector<int> status;        // per (node, comp)
    std::vector<double> value;      // current value (DegreeOfFreedom values_[0])
    std::vector<double> prescribed; // constraint value
    std::vector<int32_t> index;     // solver row, -1 if constrained

    Field() = default;
    Field(std::size_t numNodes, int s)
        : size(s), status(numNodes * s, ACTIVE), value(numNodes * s, 0.),
          prescribed(numNodes * s, 0.), index(numNodes * s, -1) {}

    // DegreeOfFreedom::constrainValue (base/dof/DegreeOfFreedom.hpp:232-242)
    void constrainValue(std::size_t node, int comp, double v)
    {
        status[node * size + comp] = CONSTRAINED;
        prescribed[node * size + comp] = v;
    }
    bool isActive(std::size_t node, int comp) const { return status[node * size + comp] == ACTIVE; }
    std::size_t numNodes() const { return status.size() / size; }
};

// element n-faces in hierarchic vertex numbers (base/mesh/ElementFaces.hpp)
namespace detail_ {
static const int QUAD_EDGES[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
static const int HEX_EDGES[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                     {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
static const int HEX_FACES[6][4] = {{0, 3, 2, 1}, {4, 5, 6, 7}, {0, 1, 5, 4},
                                    {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}};
}  // namespace detail_

// base::dof::generate<base::fe::Basis<shape, degree>> (base/dof/generate.hpp
// :66-115).  Degree 1: one DoF object per node.  Degree 2: the passes of
// base/dof/IndexMap.hpp:222-280 over generateDoFIndicesFromFaces -- vertices,
// edges, faces (HEX; the cell of a QUAD), cells (HEX) -- each walking the
// elements in order and their n-faces in table order; an n-face seen before
// (same sorted vertex ids) keeps its number.  Support points: the vertex,
// the edge midpoint, the face and cell centres.
inline Field generateField(const Mesh& m, int degree, int size)
{
    if (degree == 1) return Field(m.numNodes(), size);
    const int dim = m.dim, nv = m.npe(), nf = dim == 2 ? 9 : 27;
    const std::size_t E = m.numElements();
    std::vector<int32_t> conn(E * nf, -1);
    std::vector<std::vector<double>> pts;
    std::map<std::vector<int32_t>, int32_t> seen;
    int32_t next = 0;
    auto point = [&](const int32_t* ge, const int* vs, int k) {
        std::vector<double> x(dim, 0.);
        for (int d = 0; d < dim; d++) {
            double sum = 0.;
            for (int v = 0; v < k; v++) sum += m.xyz[(std::size_t)ge[vs[v]] * dim + d];
            x[d] = sum * (1.0 / k);
        }
        return x;
    };
    auto pass = [&](int col, int count, int k, const int* tab, bool shared) {
        seen.clear();
        for (std::size_t e = 0; e < E; e++) {
            const int32_t* ge = &m.conn[e * nv];
            for (int f = 0; f < count; f++) {
                const int* vs = tab + f * k;
                std::vector<int32_t> key(k);
                for (int v = 0; v < k; v++) key[v] = ge[vs[v]];
                std::sort(key.begin(), key.end());
                int32_t id;
                auto it = shared ? seen.find(key) : seen.end();
                if (it != seen.end()) {
                    id = it->second;
                } else {
                    id = next++;
                    if (shared) seen.emplace(key, id);
                    pts.push_back(point(ge, vs, k));
                }
                conn[e * nf + col + f] = id;
            }
        }
    };
    static const int VERT[8] = {0, 1, 2, 3, 4, 5, 6, 7};
    static const int CELL[8] = {0, 1, 2, 3, 4, 5, 6, 7};
    pass(0, nv, 1, VERT, true);
    if (dim == 2) {
        pass(4, 4, 2, &detail_::QUAD_EDGES[0][0], true);
        pass(8, 1, 4, CELL, false);
    } else {
        pass(8, 12, 2, &detail_::HEX_EDGES[0][0], true);
        pass(20, 6, 4, &detail_::HEX_FACES[0][0], true);
        pass(26, 1, 8, CELL, false);
    }
    Field f((std::size_t)next, size);
    f.degree = 2;
    f.conn = std::move(conn);
    f.support.resize((std::size_t)next * dim);
    for (std::size_t i = 0; i < pts.size(); i++)
        for (int d = 0; d < dim; d++) f.support[i * dim + d] = pts[i][d];
    return f;
}

// base/dof/numbering.hpp:45-66
inline std::size_t numberDoFsConsecutively(Field& f)
{
    std::size_t counter = 0;
    for (std::size_t k = 0; k < f.status.size(); k++)
        f.index[k] = f.status[k] == ACTIVE ? (int32_t)counter++ : -1;
    return counter;
}

// base::dof::scaleConstraints (base/dof/Distribute.hpp:76-80)
inline void scaleConstraints(Field& f, double factor)
{
    for (std::size_t k = 0; k < f.status.size(); k++)
        if (f.status[k] == CONSTRAINED) f.prescribed[k] *= factor;
}

// base::dof::constrainBoundary with a functor diriFun(x, node, field)
// applied at the nodes of the boundary faces (base/dof/constrainBoundary.hpp:52-120)
// The DoF objects of every boundary face of a degree-2 field in the face's
// own hierarchic order: LINE3 = (v0, v1, mid) for a QUAD edge; QUAD9 = the 4
// vertices in face-table order, the edges (w0,w1), (w1,w2), (w2,w3),
// (w3,w0), the centre (base/mesh/HierarchicOrder.hpp:131-150, 201-270).
// Faces come in the order of their sorted vertex ids (the boundary is the
// set of faces met by one element, base/mesh/MeshBoundary.hpp).
inline std::vector<std::vector<int32_t>> boundaryFieldFaces(const Mesh& m, const Field& f)
{
    if (f.degree != 2) return boundaryFaces(m);
    const int nf = m.dim == 2 ? 9 : 27, nv = m.npe();
    std::map<std::vector<int32_t>, std::pair<int, std::pair<std::size_t, int>>> cnt;
    const int nfaces = m.dim == 3 ? 6 : 4;
    for (std::size_t e = 0; e < m.numElements(); e++)
        for (int fc = 0; fc < nfaces; fc++) {
            std::vector<int32_t> key;
            if (m.dim == 3)
                for (int k = 0; k < 4; k++) key.push_back(m.conn[e * nv + detail_::HEX_FACES[fc][k]]);
            else
                for (int k = 0; k < 2; k++) key.push_back(m.conn[e * nv + detail_::QUAD_EDGES[fc][k]]);
            std::sort(key.begin(), key.end());
            auto it = cnt.find(key);
            if (it == cnt.end()) cnt.emplace(key, std::make_pair(1, std::make_pair(e, fc)));
            else it->second.first++;
        }
    std::vector<std::vector<int32_t>> out;
    for (auto& kv : cnt) {
        if (kv.second.first != 1) continue;
        const std::size_t e = kv.second.second.first;
        const int fc = kv.second.second.second;
        std::vector<int> locs;
        if (m.dim == 2) {
            locs = {detail_::QUAD_EDGES[fc][0], detail_::QUAD_EDGES[fc][1], 4 + fc};
        } else {
            const int* w = detail_::HEX_FACES[fc];
            for (int k = 0; k < 4; k++) locs.push_back(w[k]);
            for (int k = 0; k < 4; k++) {
                const int a = w[k], b = w[(k + 1) % 4];
                for (int ed = 0; ed < 12; ed++)
                    if ((detail_::HEX_EDGES[ed][0] == a && detail_::HEX_EDGES[ed][1] == b) ||
                        (detail_::HEX_EDGES[ed][0] == b && detail_::HEX_EDGES[ed][1] == a))
                        locs.push_back(8 + ed);
            }
            locs.push_back(20 + fc);
        }
        std::vector<int32_t> nodes;
        for (int loc : locs) nodes.push_back(f.conn[e * nf + loc]);
        out.push_back(std::move(nodes));
    }
    return out;
}

namespace detail_ {
// The library's quadrature rule (ins_quadrature_rule): weights and points,
// the first coordinate fastest
inline void quadratureRule(int dim, int npts, std::vector<double>& w, std::vector<double>& xi)
{
    int nq = 1;
    for (int d = 0; d < dim; d++) nq *= npts;
    w.assign((std::size_t)nq, 0.);
    xi.assign((std::size_t)nq * (dim > 0 ? dim : 1), 0.);
    if (ins_quadrature_rule(dim, npts, w.data(), xi.data()) != nq)
        throw Error(INS_ERR_ARG, "quadrature rule: dim " + std::to_string(dim) + ", " +
                                     std::to_string(npts) + " points per direction");
}

// x(xi) of element e's Q1 geometry: vertices in hierarchic order, vertex v
// at the corner (v0 ^ v1, v1, v2) of the reference cube, v_d the bits of v
// (base/mesh/HierarchicOrder.hpp:201-272)
inline std::array<double, 3> q1Point(const Mesh& m, std::size_t e, const double* xi)
{
    std::array<double, 3> x{0., 0., 0.};
    const int npe = m.npe(), dim = m.dim;
    for (int v = 0; v < npe; v++) {
        const int bit[3] = {(v & 1) ^ ((v >> 1) & 1), (v >> 1) & 1, (v >> 2) & 1};
        double phi = 1.;
        for (int d = 0; d < dim; d++) phi *= bit[d] ? xi[d] : 1. - xi[d];
        const std::size_t node = (std::size_t)m.conn[e * npe + v];
        for (int d = 0; d < dim; d++) x[d] += phi * m.xyz[node * dim + d];
    }
    return x;
}

// Coordinates of a node of a face list: a mesh node (degree 1) or the
// support point of a field node (degree 2, the face's vertices come first)
inline const double* faceNodeX(const Mesh& m, const Field& f, int32_t node)
{
    return f.degree == 2 ? &f.support[(std::size_t)node * m.dim] : &m.xyz[(std::size_t)node * m.dim];
}

// Surface point x(eta) and unit normal of base::SurfaceNormal
// (base/geometry.hpp:255-346) on a face given by its corner nodes in the
// element face-table order: 3-D the cross product of the surface Jacobian's
// columns of the bilinear face map, 2-D (t_y, -t_x) of the edge tangent
inline void facePoint(const Mesh& m, const Field& f, const std::vector<int32_t>& face,
                      const double* eta, std::array<double, 3>& x, std::array<double, 3>& n)
{
    x = {0., 0., 0.};
    n = {0., 0., 0.};
    if (m.dim == 3) {
        const double e0 = eta[0], e1 = eta[1];
        const double phi[4] = {(1 - e0) * (1 - e1), e0 * (1 - e1), e0 * e1, (1 - e0) * e1};
        const double d0[4] = {-(1 - e1), 1 - e1, e1, -e1};
        const double d1[4] = {-(1 - e0), -e0, e0, 1 - e0};
        double J0[3] = {0., 0., 0.}, J1[3] = {0., 0., 0.};
        for (int k = 0; k < 4; k++) {
            const double* X = faceNodeX(m, f, face[k]);
            for (int d = 0; d < 3; d++) {
                x[d] += phi[k] * X[d];
                J0[d] += d0[k] * X[d];
                J1[d] += d1[k] * X[d];
            }
        }
        n = {J0[1] * J1[2] - J0[2] * J1[1], J0[2] * J1[0] - J0[0] * J1[2],
             J0[0] * J1[1] - J0[1] * J1[0]};
    } else {
        const double* X0 = faceNodeX(m, f, face[0]);
        const double* X1 = faceNodeX(m, f, face[1]);
        for (int d = 0; d < 2; d++) x[d] = (1 - eta[0]) * X0[d] + eta[0] * X1[d];
        n = {X1[1] - X0[1], -(X1[0] - X0[0]), 0.};
    }
    const double len = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    for (int d = 0; d < 3; d++) n[d] /= len;
}

// component c of a caller's function value: an indexable vector or a scalar
template <typename R>
inline double component(const R& r, int c)
{
    if constexpr (std::is_arithmetic<R>::value) {
        (void)c;
        return (double)r;
    } else {
        return r[c];
    }
}
}  // namespace detail_

template <typename DIRIFUN>
void constrainBoundary(const Mesh& m, Field& f, DIRIFUN diriFun)
{
    if (f.degree == 2) {
        // the DoF objects of every boundary face: its vertices, edges and
        // (3-D) its face centre, at their support points
        std::vector<char> done(f.numNodes(), 0);
        for (const auto& face : boundaryFieldFaces(m, f)) {
            for (const int32_t node : face) {
                if (done[node]) continue;
                done[node] = 1;
                std::array<double, 3> x{0., 0., 0.};
                for (int d = 0; d < m.dim; d++) x[d] = f.support[(std::size_t)node * m.dim + d];
                diriFun(x, (std::size_t)node, f);
            }
        }
        return;
    }
    std::vector<char> done(m.numNodes(), 0);
    for (const auto& face : boundaryFaces(m))
        for (int32_t node : face) {
            if (done[node]) continue;
            done[node] = 1;
            std::array<double, 3> x{0., 0., 0.};
            for (int d = 0; d < m.dim; d++) x[d] = m.xyz[node * m.dim + d];
            diriFun(x, (std::size_t)node, f);
        }
}

}  // namespace insilico

// ---------------------------------------------------------------------------
// Quadrature, kernels and materials (tags with their parameters)
namespace base {
enum Shape { QUAD = 4, HEX = 8 };
template <Shape SHAPE>
struct ShapeDim {
    static constexpr unsigned value = SHAPE == HEX ? 3 : 2;
};

// Small fixed-size vector for the caller's functions (the reference's
// base::Vector<N>::Type is an Eigen vector, base/linearAlgebra.hpp:45-60):
// element access, the arithmetic and norms such functions use.
template <unsigned N>
struct VecN {
    double v[N] = {};
    double& operator[](std::size_t i) { return v[i]; }
    double operator[](std::size_t i) const { return v[i]; }
    double& operator()(std::size_t i) { return v[i]; }
    double operator()(std::size_t i) const { return v[i]; }
    static constexpr std::size_t size() { return N; }
    static VecN Constant(double a)
    {
        VecN r;
        for (unsigned i = 0; i < N; i++) r.v[i] = a;
        return r;
    }
    static VecN Zero() { return Constant(0.); }
    void setZero() { *this = Zero(); }
    double dot(const VecN& o) const
    {
        double s = 0.;
        for (unsigned i = 0; i < N; i++) s += v[i] * o.v[i];
        return s;
    }
    double squaredNorm() const { return dot(*this); }
    double norm() const { return std::sqrt(squaredNorm()); }
    VecN& operator+=(const VecN& o)
    {
        for (unsigned i = 0; i < N; i++) v[i] += o.v[i];
        return *this;
    }
    VecN& operator-=(const VecN& o)
    {
        for (unsigned i = 0; i < N; i++) v[i] -= o.v[i];
        return *this;
    }
    VecN& operator*=(double a)
    {
        for (unsigned i = 0; i < N; i++) v[i] *= a;
        return *this;
    }
    friend VecN operator+(VecN a, const VecN& b) { return a += b; }
    friend VecN operator-(VecN a, const VecN& b) { return a -= b; }
    friend VecN operator*(VecN a, double s) { return a *= s; }
    friend VecN operator*(double s, VecN a) { return a *= s; }
    friend VecN operator/(VecN a, double s) { return a *= 1. / s; }
    friend VecN operator-(VecN a) { return a *= -1.; }
};
template <unsigned N, typename T = double>
struct Vector {
    typedef VecN<N> Type;
};

// base::Unstructured<SHAPE, GDEG> (base/Unstructured.hpp): the mesh with its
// shape and dimension as compile-time constants
template <Shape SHAPE, unsigned GDEG = 1>
struct Unstructured : insilico::Mesh {
    struct Node {
        static constexpr unsigned dim = ShapeDim<SHAPE>::value;
    };
    struct Element {
        static constexpr Shape shape = SHAPE;
        struct GeomFun {
            static constexpr unsigned degree = GDEG;
        };
    };
    Unstructured() { this->dim = (int)Node::dim; }
};
template <typename MESH>
struct MeshDim {  // a mesh whose dimension is only known at run time: 3
    static constexpr unsigned value = 3;
};
template <Shape S, unsigned G>
struct MeshDim<Unstructured<S, G>> {
    static constexpr unsigned value = ShapeDim<S>::value;
};

// An element as a caller's conductivity function sees it (the reference
// passes the GeomElement pointer of the field tuple, heat/Laplace.hpp:
// 113-126): its mesh and id, and base::Geometry to map a local coordinate
// to the physical point (the Q1 map, base/geometry.hpp:46-90).
template <unsigned DIM>
struct GeomElement {
    const insilico::Mesh* mesh = nullptr;
    std::size_t id = 0;
    std::size_t getID() const { return id; }
};
template <typename ELEMENT>
struct Geometry;
template <unsigned DIM>
struct Geometry<GeomElement<DIM>> {
    VecN<DIM> operator()(const GeomElement<DIM>* ep, const VecN<DIM>& xi) const
    {
        const insilico::Mesh& m = *ep->mesh;
        double loc[3] = {0., 0., 0.};
        for (unsigned d = 0; d < DIM && d < 3; d++) loc[d] = xi[d];
        const std::array<double, 3> x = insilico::detail_::q1Point(m, ep->id, loc);
        VecN<DIM> out;
        for (unsigned d = 0; d < DIM && d < 3; d++) out[d] = x[d];
        return out;
    }
};

// base::Quadrature<DEG,SHAPE>: tensor Gauss-Legendre with (DEG+2)/2 points
// per direction (base/quad/GaussLegendre.hpp:33-36)
template <unsigned DEG, Shape SHAPE>
struct Quadrature {
    static constexpr int numPointsPerDir = (DEG + 2) / 2;
};
// base::SurfaceQuadrature<DEG, SHAPE> (base/Quadrature.hpp): the same rule
// on the faces of SHAPE
template <unsigned DEG, Shape SHAPE>
struct SurfaceQuadrature {
    static constexpr int numPointsPerDir = (DEG + 2) / 2;
};

template <typename MESH, typename FIELD>
struct FieldBinder {  // base/asmb/FieldBinder.hpp:107-188 (one field)
    static constexpr unsigned dim = MeshDim<MESH>::value;
    MESH& mesh;
    FIELD& field;
    FieldBinder(MESH& m, FIELD& f) : mesh(m), field(f) {}
    template <int, int>
    struct TupleBinder {
        using Type = FieldBinder;
    };
};

namespace mesh {
// The surface mesh of a volume mesh's boundary (base::mesh::BoundaryMeshBinder,
// generateBoundaryMesh, base/mesh/generateBoundaryMesh.hpp): the boundary
// faces of the field, as node lists (insilico::boundaryFieldFaces), with the
// volume mesh they belong to.
struct BoundaryMesh {
    const insilico::Mesh* volume = nullptr;
    std::vector<std::vector<int32_t>> faces;
};
}  // namespace mesh

namespace asmb {
// base::asmb::SurfaceFieldBinder (base/asmb/SurfaceFieldBinder.hpp): the
// field over the faces of a boundary mesh
template <typename BMESH, typename FIELD>
struct SurfaceFieldBinder {
    BMESH& boundaryMesh;
    FIELD& field;
    SurfaceFieldBinder(BMESH& b, FIELD& f) : boundaryMesh(b), field(f) {}
    template <int, int>
    struct TupleBinder {
        using Type = SurfaceFieldBinder;
    };
};
}  // namespace asmb
}  // namespace base

namespace mat {
// mat/Lame.hpp:28-40
struct Lame {
    static double lambda(double E, double nu) { return E * nu / (1. + nu) / (1. - 2. * nu); }
    static double mu(double E, double nu) { return E / 2. / (1. + nu); }
};
namespace hypel {
struct StVenant {
    static constexpr int id = INS_STVENANT;
    double lambda, mu;
    StVenant(double l, double m) : lambda(l), mu(m) {}
};
struct NeoHookeanCompressible {
    static constexpr int id = INS_NEOHOOKE;
    double lambda, mu;
    NeoHookeanCompressible(double l, double m) : lambda(l), mu(m) {}
};
}  // namespace hypel
}  // namespace mat

namespace heat {
namespace detail_ {
template <typename FTB>
struct TupleDim {
    static constexpr unsigned value = 3;
};
template <typename M, typename F>
struct TupleDim<base::FieldBinder<M, F>> {
    static constexpr unsigned value = base::FieldBinder<M, F>::dim;
};
}  // namespace detail_

// heat::Laplace (heat/Laplace.hpp:85-97, 113-126): a constant factor, or a
// conductivity function kappa(element, xi) (setConductivityFunction), which
// the assembly evaluates on the host at every element's quadrature points
// and uploads as per-point values; or, as an extension, one conductivity
// per element.
template <typename FTB = void>
struct Laplace {
    static constexpr unsigned dim = detail_::TupleDim<FTB>::value;
    typedef base::GeomElement<dim> GeomElement;
    typedef base::VecN<dim> LocalVecDim;
    typedef std::function<double(const GeomElement*, const LocalVecDim&)> ConductivityFun;
    double factor;
    std::vector<double> perElement;
    ConductivityFun conductivityFun;
    explicit Laplace(double f) : factor(f) {}
    void setConductivityFunction(ConductivityFun& cf)
    {
        conductivityFun = cf;
        perElement.clear();
    }
    void setConductivityFunction(std::function<double(std::size_t)> k, std::size_t numElements)
    {
        conductivityFun = nullptr;
        perElement.resize(numElements);
        for (std::size_t e = 0; e < numElements; e++) perElement[e] = k(e);
    }
};
// heat::Static<MATERIAL> (heat/Static.hpp:98-...), the kernel of
// heat::PoissonDriver (heat/PoissonDriver.hpp:78).  With a constant isotropic
// conductivity K = k I and dK/du = 0 (mat/thermal/IsotropicConstant.hpp:42-52)
// its tangent is k times the Laplace matrix, and it is assembled as such.
template <typename MATERIAL, typename FTB = void>
struct Static {
    MATERIAL material;
    explicit Static(const MATERIAL& m) : material(m) {}
};
}  // namespace heat

namespace mat {
namespace thermal {
// mat::thermal::IsotropicConstant (mat/thermal/IsotropicConstant.hpp:42-52)
struct IsotropicConstant {
    double conductivity;
    explicit IsotropicConstant(double k = 1.0) : conductivity(k) {}
};
}  // namespace thermal
}  // namespace mat

namespace solid {
// solid::HyperElastic<MATERIAL, FTB> (solid/HyperElastic.hpp); per-element
// Lame parameters are an extension (the reference holds one material)
template <typename MATERIAL, typename FTB = void>
struct HyperElastic {
    MATERIAL material;
    std::vector<double> perElement;  // lambda_e, mu_e pairs (optional)
    explicit HyperElastic(const MATERIAL& m) : material(m) {}
};
}  // namespace solid

// ---------------------------------------------------------------------------
// Time integration (base/time/BDF.hpp) and heat::DiffusionDriver
// (heat/DiffusionDriver.hpp:88-118) on the device
namespace base {
namespace time {
template <unsigned Q>
struct BDF {
    static constexpr unsigned order = Q;
};
}  // namespace time
}  // namespace base

namespace heat {
// advanceInTime(numStep, stepSize): assemble K + (a0/dt) M and the RHS with
// the mass history (base/time/ReactionTerms.hpp), solve (the reference's
// choleskySolve: a PCG to 1e-13 here), set the DoFs and push the history --
// all on the device (ins_bdf_step).  Steps must be called with numStep =
// 0, 1, 2, ... (the BDF start-up follows the call count).  The field's DoFs
// must be numbered and its constraints and initial values set.
template <typename MSM = base::time::BDF<1>>
class DiffusionDriver {
public:
    DiffusionDriver(const insilico::Mesh& m, insilico::Field& f, double conductivity = 1.0,
                    double density = 1.0, int device = 0)
        : mesh_(m), field_(f), kappa_(conductivity), rho_(density)
    {
        insilico::check(ins_create(&ctx_, device, 0), nullptr);
    }
    ~DiffusionDriver()
    {
        if (ctx_) ins_destroy(ctx_);
    }
    DiffusionDriver(const DiffusionDriver&) = delete;
    DiffusionDriver& operator=(const DiffusionDriver&) = delete;

    // body force f (bodyForceComputation): INS_BF_CONST or INS_BF_SINPROD
    void setForce(int kind, const std::array<double, 4>& p)
    {
        bfKind_ = kind;
        bf_ = p;
    }
    void advanceInTime(unsigned numStep, double stepSize, int mode = INS_SOLVE_FP64,
                       int format = INS_FMT_SELL, double rtol = 1e-13)
    {
        if (!ready_) setup();
        if (numStep != steps_) throw insilico::Error(INS_ERR_STATE, "time steps must be consecutive");
        int it = 0;
        double rel = 0.;
        const int maxIt = (int)std::max<std::size_t>(2 * n_, 100);
        insilico::check(ins_bdf_step(ctx_, stepSize, INS_FP64, mode, format, rtol, maxIt, &it, &rel), ctx_);
        insilico::check(ins_get_dof_values(ctx_, field_.value.data()), ctx_);
        steps_++;
        lastIterations_ = it;
    }
    int lastIterations() const { return lastIterations_; }
    ins_ctx* context() { return ctx_; }

private:
    void setup()
    {
        const auto& m = mesh_;
        auto& f = field_;
        n_ = 0;
        for (auto id : f.index) n_ += id >= 0;
        insilico::check(ins_set_mesh(ctx_, m.dim, (int64_t)m.numNodes(), m.xyz.data(),
                                     (int64_t)m.numElements(), m.npe(), m.conn.data()), ctx_);
        if (f.degree == 2)
            insilico::check(ins_set_field(ctx_, 2, (int64_t)f.numNodes(), m.dim == 2 ? 9 : 27,
                                          f.conn.data()), ctx_);
        insilico::check(ins_set_dofs(ctx_, f.size, f.index.data(), (int64_t)n_), ctx_);
        insilico::check(ins_set_prescribed(ctx_, f.prescribed.data()), ctx_);
        // kernel degree 2 * fieldDeg (heat/DiffusionDriver.hpp:99)
        const int nq = (2 * f.degree + 2) / 2;
        std::vector<double> coef(m.numElements(), kappa_);
        insilico::check(ins_set_material(ctx_, INS_LAPLACE, nq, 1, coef.data()), ctx_);
        insilico::check(ins_set_body_force(ctx_, bfKind_, bf_.data()), ctx_);
        int64_t nnz = 0;
        insilico::check(ins_build_pattern(ctx_, &nnz), ctx_);
        insilico::check(ins_bdf_init(ctx_, (int)MSM::order, rho_, f.value.data()), ctx_);
        ready_ = true;
    }

    const insilico::Mesh& mesh_;
    insilico::Field& field_;
    double kappa_, rho_;
    ins_ctx* ctx_ = nullptr;
    std::size_t n_ = 0;
    unsigned steps_ = 0;
    bool ready_ = false;
    int lastIterations_ = 0;
    int bfKind_ = INS_BF_NONE;
    std::array<double, 4> bf_{0., 0., 0., 0.};
};
}  // namespace heat

// ---------------------------------------------------------------------------
// SOLVER concept on the device (base/solver/Eigen3.hpp:64-342)
namespace base {
namespace solver {

class Hip {
public:
    explicit Hip(std::size_t size, int device = 0, uint64_t hbmBudget = 0) : n_(size)
    {
        insilico::check(ins_create(&ctx_, device, hbmBudget), nullptr);
    }
    ~Hip()
    {
        if (ctx_) ins_destroy(ctx_);
    }
    Hip(const Hip&) = delete;
    Hip& operator=(const Hip&) = delete;

    ins_ctx* context() { return ctx_; }

    // ---- host triplet path (Eigen3.hpp:82-124, TripletContainer dynamic mode)
    template <typename MATRIX, typename RDOFS, typename CDOFS>
    void insertToLHS(const MATRIX& m, const RDOFS& rows, const CDOFS& cols)
    {
        for (std::size_t i = 0; i < rows.size(); i++)
            for (std::size_t j = 0; j < cols.size(); j++) {
                if ((std::size_t)rows[i] >= n_ || (std::size_t)cols[j] >= n_)
                    throw insilico::Error(INS_ERR_RANGE, "index out of bound");
                triplets_[{(int32_t)rows[i], (int32_t)cols[j]}] += m(i, j);
            }
        hostMatrix_ = true;
    }
    template <typename VECTOR, typename DOFS>
    void insertToRHS(const VECTOR& v, const DOFS& ids)
    {
        hostRhs_.resize(n_, 0.);
        for (std::size_t i = 0; i < ids.size(); i++) {
            if ((std::size_t)ids[i] >= n_) throw insilico::Error(INS_ERR_RANGE, "out of bound");
            hostRhs_[ids[i]] += v[i];
        }
    }

    // ---- device assembly requests (filled by base::asmb free functions)
    template <typename MESH, typename FIELD>
    void bind(const MESH& mesh, FIELD& field)
    {
        mesh_ = &mesh;
        field_ = &field;
    }
    void requestMatrix(int kernel, int nq, std::vector<double> coef, int ncoef, bool incremental)
    {
        kernel_ = kernel;
        nq_ = nq;
        coef_ = std::move(coef);
        ncoef_ = ncoef;
        flags_ |= INS_ASM_MATRIX | INS_ASM_RHS;
        if (incremental) flags_ |= INS_ASM_INCREMENTAL;
    }
    void requestResidual(int kernel, int nq, std::vector<double> coef, int ncoef)
    {
        if (!kernel_) {
            kernel_ = kernel;
            nq_ = nq;
            coef_ = std::move(coef);
            ncoef_ = ncoef;
        }
        flags_ |= INS_ASM_RHS | INS_ASM_RESIDUAL;
    }
    void requestBodyForce(int kind, const std::array<double, 4>& p)
    {
        bfKind_ = kind;
        bf_ = p;
        flags_ |= INS_ASM_RHS;
    }
    // traction t on the faces; with axis >= 0 it is scaled by (x[axis] - x0)
    // at every surface quadrature point
    void requestTraction(std::vector<std::vector<int32_t>> faces, const std::array<double, 3>& t,
                         int axis = -1, double x0 = 0.)
    {
        faces_ = std::move(faces);
        traction_ = t;
        trAxis_ = axis;
        trX0_ = x0;
        flags_ |= INS_ASM_RHS;
    }
    // Neumann data as the co-normal derivative of a built-in function on the
    // faces of requestTraction
    void requestFlux(int refKind, const std::array<double, 8>& p)
    {
        fluxKind_ = refKind;
        fluxP_ = p;
        flags_ |= INS_ASM_RHS;
    }
    // a caller's body force as its values at the quadrature points
    // (ins_set_body_force_values layout)
    void requestBodyForceValues(std::vector<double> f)
    {
        bfTable_ = std::move(f);
        flags_ |= INS_ASM_RHS;
    }
    // a caller's Neumann function as its values at the surface points of the
    // faces, with nqs points per direction (ins_set_traction_values layout)
    void requestTractionValues(std::vector<std::vector<int32_t>> faces, std::vector<double> t,
                               int nqs)
    {
        requestTraction(std::move(faces), {0., 0., 0.});
        trTable_ = std::move(t);
        trNq_ = nqs;
    }
    void setPrecision(int p) { precision_ = p; }
    // forget the recorded assembly requests and the solution (a new Newton
    // iteration re-assembles into the same device context, pattern kept)
    void reset()
    {
        // a fresh solver: no loads requested
        flags_ = 0;
        bfKind_ = INS_BF_NONE;
        faces_.clear();
        traction_ = {0., 0., 0.};
        trAxis_ = -1;
        trX0_ = 0.;
        fluxKind_ = INS_REF_ZERO;
        bfTable_.clear();
        trTable_.clear();
        trNq_ = 0;
        solved_ = false;
        assembled_ = false;
        hostRhs_.clear();
        triplets_.clear();
        hostMatrix_ = false;
    }

    // Eigen3::finishAssembly (base/solver/Eigen3.hpp:142-153)
    void finishAssembly(bool = true)
    {
        if (hostMatrix_) {
            uploadTriplets();
        } else if (flags_) {
            deviceAssembly();
        }
        if (!hostRhs_.empty()) {
            if (!hostMatrix_ && flags_ & INS_ASM_RHS) insilico::check(ins_add_rhs(ctx_, hostRhs_.data()), ctx_);
            else insilico::check(ins_set_rhs(ctx_, hostRhs_.data()), ctx_);
        }
        assembled_ = true;
    }

    // Eigen3::norm(): |b|/N, or |x|/N once solved (the solution overwrites b)
    double norm() const
    {
        double v = 0.;
        if (solved_) insilico::check(ins_solution_norm(ctx_, &v), ctx_);
        else insilico::check(ins_rhs_norm(ctx_, &v), ctx_);
        return v;
    }

    // Eigen3::cgSolve (base/solver/Eigen3.hpp:265-275): Jacobi-PCG from
    // x0 = 0 to |b - A x| <= rtol |b|, at most 2n iterations.  Eigen's
    // ConjugateGradient hands back its last iterate silently when the limit
    // is reached; this call throws insilico::Error(INS_ERR_NOCONV) instead,
    // after storing that iterate (getValue, lastRelres and converged() stay
    // readable), so that an unconverged solve is never taken for a solution.
    int cgSolve(double rtol = 1e-12, int mode = INS_SOLVE_FP64, int format = INS_FMT_SELL)
    {
        int it = 0;
        double rel = 0.;
        const int maxIt = (int)std::max<std::size_t>(2 * n_, 100);
        const int st = ins_solve(ctx_, mode, format, rtol, maxIt, &it, &rel);
        converged_ = st == INS_OK;
        lastRelres_ = rel;
        lastIterations_ = it;
        if (st != INS_ERR_NOCONV) insilico::check(st, ctx_);
        x_.resize(n_);
        insilico::check(ins_get_solution(ctx_, x_.data()), ctx_);
        solved_ = true;
        if (st == INS_ERR_NOCONV)
            throw insilico::Error(st, std::string("cgSolve: ") + ins_last_error(ctx_));
        return it;
    }
    // Direct solves (Eigen3::choleskySolve / luSolve / superLUSolve,
    // base/solver/Eigen3.hpp:157-230) return the factorised solution or
    // abort: a dense device factorisation (ins_solve_direct) up to
    // INS_DIRECT_MAX_N rows, throwing insilico::Error on a failed
    // factorisation; larger systems run the FP64 PCG to 1e-13, which meets
    // that bound or throws INS_ERR_NOCONV (as above).
    void choleskySolve() { directSolve(INS_DIRECT_CHOLESKY); }
    void luSolve() { directSolve(INS_DIRECT_LU); }
    void superLUSolve() { directSolve(INS_DIRECT_LU); }

    void directSolve(int kind)
    {
        if (n_ > INS_DIRECT_MAX_N) {
            cgSolve(1e-13);
            return;
        }
        double rel = 0.;
        converged_ = false;
        insilico::check(ins_solve_direct(ctx_, kind, &rel), ctx_);
        converged_ = true;
        lastRelres_ = rel;
        lastIterations_ = 0;
        x_.resize(n_);
        insilico::check(ins_get_solution(ctx_, x_.data()), ctx_);
        solved_ = true;
    }

    double getValue(std::size_t i) const { return x_.at(i); }  // Eigen3.hpp:293-296
    double lastRelres() const { return lastRelres_; }
    int lastIterations() const { return lastIterations_; }
    bool converged() const { return converged_; }
    std::size_t size() const { return n_; }

    // Eigen3::debugLHS format (row col value), base/solver/Eigen3.hpp:311-317
    template <typename OS>
    void debugLHS(OS& out)
    {
        int64_t nnz = 0;
        ins_stats st;
        insilico::check(ins_get_stats(ctx_, &st), ctx_);
        nnz = st.nnz;
        std::vector<int32_t> rp(n_ + 1), ci(nnz);
        std::vector<double> v(nnz);
        insilico::check(ins_get_pattern(ctx_, rp.data(), ci.data()), ctx_);
        insilico::check(ins_get_csr_values(ctx_, v.data()), ctx_);
        for (std::size_t r = 0; r < n_; r++)
            for (int32_t k = rp[r]; k < rp[r + 1]; k++) out << r << " " << ci[k] << " " << v[k] << "\n";
    }

    // DoF distribution (base/dof/Distribute.hpp:141-215) on the host field
    template <typename FIELD>
    void distribute(FIELD& f, bool add) const
    {
        for (std::size_t k = 0; k < f.status.size(); k++) {
            if (f.status[k] == insilico::ACTIVE) f.value[k] = add ? f.value[k] + x_[f.index[k]] : x_[f.index[k]];
            else f.value[k] = f.prescribed[k];
        }
    }

private:
    void uploadTriplets()
    {
        std::vector<int32_t> rp(n_ + 1, 0), ci;
        std::vector<double> v;
        ci.reserve(triplets_.size());
        v.reserve(triplets_.size());
        for (auto& t : triplets_) {
            rp[t.first.first + 1]++;
            ci.push_back(t.first.second);
            v.push_back(t.second);
        }
        for (std::size_t i = 0; i < n_; i++) rp[i + 1] += rp[i];
        insilico::check(ins_set_csr(ctx_, (int64_t)n_, rp.data(), ci.data(), v.data()), ctx_);
        if (hostRhs_.empty()) hostRhs_.assign(n_, 0.);
    }

    void deviceAssembly()
    {
        if (!mesh_ || !field_) throw insilico::Error(INS_ERR_STATE, "no mesh/field bound");
        const auto& m = *mesh_;
        auto& f = *field_;
        if (!meshSent_) {
            insilico::check(ins_set_mesh(ctx_, m.dim, (int64_t)m.numNodes(), m.xyz.data(),
                                         (int64_t)m.numElements(), m.npe(), m.conn.data()), ctx_);
            if (f.degree == 2)
                insilico::check(ins_set_field(ctx_, 2, (int64_t)f.numNodes(), m.dim == 2 ? 9 : 27,
                                              f.conn.data()), ctx_);
            insilico::check(ins_set_dofs(ctx_, f.size, f.index.data(), (int64_t)n_), ctx_);
            int64_t nnz = 0;
            insilico::check(ins_build_pattern(ctx_, &nnz), ctx_);
            meshSent_ = true;
        }
        insilico::check(ins_set_prescribed(ctx_, f.prescribed.data()), ctx_);
        // all-zero DoF values go over as NULL: the library then knows the
        // displacement is zero and assembles the hyperelastic terms in their
        // zero-displacement form (the same element matrices)
        bool zero = true;
        for (double v : f.value)
            if (v != 0.) {
                zero = false;
                break;
            }
        insilico::check(ins_set_dof_values(ctx_, zero ? nullptr : f.value.data()), ctx_);
        insilico::check(ins_set_material(ctx_, kernel_, nq_, ncoef_, coef_.data()), ctx_);
        insilico::check(ins_set_body_force(ctx_, bfKind_, bf_.data()), ctx_);
        if (!bfTable_.empty()) insilico::check(ins_set_body_force_values(ctx_, bfTable_.data()), ctx_);
        std::vector<int32_t> flat;
        for (auto& fc : faces_) flat.insert(flat.end(), fc.begin(), fc.end());
        const int nfn = faces_.empty() ? 0 : (int)faces_[0].size();
        insilico::check(ins_set_traction(ctx_, (int64_t)faces_.size(), nfn, faces_.empty() ? nullptr : flat.data(),
                                         traction_.data()), ctx_);
        insilico::check(ins_set_traction_profile(ctx_, faces_.empty() ? -1 : trAxis_, trX0_), ctx_);
        insilico::check(ins_set_neumann_flux(ctx_, faces_.empty() ? INS_REF_ZERO : fluxKind_,
                                             fluxP_.data()), ctx_);
        if (!trTable_.empty()) {
            if (trNq_ != nq_)
                throw insilico::Error(INS_ERR_ARG, "the surface quadrature must have the domain "
                                                   "quadrature's points per direction");
            insilico::check(ins_set_traction_values(ctx_, trTable_.data()), ctx_);
        }
        insilico::check(ins_assemble(ctx_, precision_, flags_), ctx_);
    }

    ins_ctx* ctx_ = nullptr;
    std::size_t n_;
    std::map<std::pair<int32_t, int32_t>, double> triplets_;
    std::vector<double> hostRhs_;
    bool hostMatrix_ = false, assembled_ = false, solved_ = false, meshSent_ = false;
    const insilico::Mesh* mesh_ = nullptr;
    insilico::Field* field_ = nullptr;
    int kernel_ = 0, nq_ = 2, ncoef_ = 1, flags_ = 0, precision_ = INS_FP64;
    std::vector<double> coef_;
    int bfKind_ = INS_BF_NONE;
    std::array<double, 4> bf_{0., 0., 0., 0.};
    std::vector<std::vector<int32_t>> faces_;
    std::array<double, 3> traction_{0., 0., 0.};
    int trAxis_ = -1;
    double trX0_ = 0.;
    int fluxKind_ = INS_REF_ZERO;
    std::array<double, 8> fluxP_{0., 0., 0., 0., 0., 0., 0., 0.};
    std::vector<double> bfTable_, trTable_;
    int trNq_ = 0;
    std::vector<double> x_;
    double lastRelres_ = 0.;
    int lastIterations_ = 0;
    bool converged_ = false;
};

}  // namespace solver

// ---------------------------------------------------------------------------
// Assembly free functions (base/asmb/StiffnessMatrix.hpp:52-73,
// ForceIntegrator.hpp:40-56, BodyForce.hpp:67-82, NeumannForce.hpp:36-63)
namespace asmb {

namespace detail_ {
// the conductivity per element, or with a conductivity function its values
// at every element's quadrature points (ncoef = npts^dim)
template <typename FTB>
std::vector<double> laplaceCoef(const heat::Laplace<FTB>& k, const insilico::Mesh& m, int npts,
                                int& ncoef)
{
    const std::size_t ne = m.numElements();
    ncoef = 1;
    if (k.conductivityFun) {
        std::vector<double> w, xi;
        insilico::detail_::quadratureRule(m.dim, npts, w, xi);
        const std::size_t nq = w.size();
        typedef typename heat::Laplace<FTB>::GeomElement GE;
        typedef typename heat::Laplace<FTB>::LocalVecDim LV;
        std::vector<double> c(ne * nq);
        for (std::size_t e = 0; e < ne; e++) {
            const GE ge{&m, e};
            for (std::size_t q = 0; q < nq; q++) {
                LV loc;
                for (int d = 0; d < m.dim && d < (int)LV::size(); d++) loc[d] = xi[q * m.dim + d];
                c[e * nq + q] = k.conductivityFun(&ge, loc);
            }
        }
        ncoef = (int)nq;
        return c;
    }
    if (!k.perElement.empty()) return k.perElement;
    return std::vector<double>(ne, k.factor);
}
template <typename MAT, typename FTB>
std::vector<double> elasticCoef(const solid::HyperElastic<MAT, FTB>& k, std::size_t ne)
{
    if (!k.perElement.empty()) return k.perElement;
    std::vector<double> c(2 * ne);
    for (std::size_t e = 0; e < ne; e++) {
        c[2 * e] = k.material.lambda;
        c[2 * e + 1] = k.material.mu;
    }
    return c;
}
}  // namespace detail_

template <typename FTB, unsigned DEG, Shape SH, typename FB, typename LFTB>
void stiffnessMatrixComputation(const Quadrature<DEG, SH>&, solver::Hip& solver, FB& fieldBinder,
                                const heat::Laplace<LFTB>& kernel, bool incremental = true)
{
    solver.bind(fieldBinder.mesh, fieldBinder.field);
    const int npts = Quadrature<DEG, SH>::numPointsPerDir;
    int ncoef = 1;
    std::vector<double> coef = detail_::laplaceCoef(kernel, fieldBinder.mesh, npts, ncoef);
    solver.requestMatrix(INS_LAPLACE, npts, std::move(coef), ncoef, incremental);
}

template <typename FTB, unsigned DEG, Shape SH, typename FB, typename SFTB>
void stiffnessMatrixComputation(const Quadrature<DEG, SH>& q, solver::Hip& solver, FB& fieldBinder,
                                const heat::Static<mat::thermal::IsotropicConstant, SFTB>& kernel,
                                bool incremental = true)
{
    stiffnessMatrixComputation<FTB>(q, solver, fieldBinder,
                                    heat::Laplace<>(kernel.material.conductivity), incremental);
}

template <typename FTB, unsigned DEG, Shape SH, typename FB, typename MAT, typename KFTB>
void stiffnessMatrixComputation(const Quadrature<DEG, SH>&, solver::Hip& solver, FB& fieldBinder,
                                const solid::HyperElastic<MAT, KFTB>& kernel, bool incremental = true)
{
    solver.bind(fieldBinder.mesh, fieldBinder.field);
    solver.requestMatrix(MAT::id, Quadrature<DEG, SH>::numPointsPerDir,
                         detail_::elasticCoef(kernel, fieldBinder.mesh.numElements()), 2, incremental);
}

template <typename FTB, unsigned DEG, Shape SH, typename FB, typename MAT, typename KFTB>
void computeResidualForces(const Quadrature<DEG, SH>&, solver::Hip& solver, FB& fieldBinder,
                           const solid::HyperElastic<MAT, KFTB>& kernel)
{
    solver.bind(fieldBinder.mesh, fieldBinder.field);
    solver.requestResidual(MAT::id, Quadrature<DEG, SH>::numPointsPerDir,
                           detail_::elasticCoef(kernel, fieldBinder.mesh.numElements()), 2);
}

// computeResidualForces with the heat kernels (ForceIntegrator.hpp:40-56
// with heat/Laplace.hpp residualForce): the kernel is linear, so the
// residual forces -K u vanish while the field is zero -- the state in which
// heat::PoissonDriver assembles (constraints do not change the values,
// base/dof/DegreeOfFreedom.hpp:232-242).  A nonzero field is not supported.
namespace detail_ {
template <typename FB>
void heatResidualAtZero(FB& fieldBinder)
{
    for (double v : fieldBinder.field.value)
        if (v != 0.)
            throw insilico::Error(INS_ERR_UNSUPPORTED,
                                  "computeResidualForces of a heat kernel at a nonzero field");
}
}  // namespace detail_
template <typename FTB, unsigned DEG, Shape SH, typename FB, typename LFTB>
void computeResidualForces(const Quadrature<DEG, SH>&, solver::Hip&, FB& fieldBinder,
                           const heat::Laplace<LFTB>&)
{
    detail_::heatResidualAtZero(fieldBinder);
}
template <typename FTB, unsigned DEG, Shape SH, typename FB, typename MAT, typename SFTB>
void computeResidualForces(const Quadrature<DEG, SH>&, solver::Hip&, FB& fieldBinder,
                           const heat::Static<MAT, SFTB>&)
{
    detail_::heatResidualAtZero(fieldBinder);
}

// built-in body forces: constant vector, or A * prod_d sin(k_d x_d)
struct ConstantForce {
    std::array<double, 3> f;
};
struct SineProductForce {
    double amplitude;
    std::array<double, 3> k;
};

template <typename FTB, unsigned DEG, Shape SH, typename FB>
void bodyForceComputation(const Quadrature<DEG, SH>&, solver::Hip& solver, FB& fieldBinder,
                          const ConstantForce& f)
{
    solver.bind(fieldBinder.mesh, fieldBinder.field);
    solver.requestBodyForce(INS_BF_CONST, {f.f[0], f.f[1], f.f[2], 0.});
}

template <typename FTB, unsigned DEG, Shape SH, typename FB>
void bodyForceComputation(const Quadrature<DEG, SH>&, solver::Hip& solver, FB& fieldBinder,
                          const SineProductForce& f)
{
    solver.bind(fieldBinder.mesh, fieldBinder.field);
    solver.requestBodyForce(INS_BF_SINPROD, {f.amplitude, f.k[0], f.k[1], f.k[2]});
}

// A * prod_d cos(k_d x_d): -Lap u of u = prod_d cos(k_d x_d) with
// A = sum_d k_d^2 (reference/05-mixedPoisson/GivenData.hpp forceFun)
struct CosineProductForce {
    double amplitude;
    std::array<double, 3> k;
};
template <typename FTB, unsigned DEG, Shape SH, typename FB>
void bodyForceComputation(const Quadrature<DEG, SH>&, solver::Hip& solver, FB& fieldBinder,
                          const CosineProductForce& f)
{
    solver.bind(fieldBinder.mesh, fieldBinder.field);
    solver.requestBodyForce(INS_BF_COSPROD, {f.amplitude, f.k[0], f.k[1], f.k[2]});
}

// bodyForceComputation<FTB>(quadrature, solver, fieldBinder, forceFun)
// (base/asmb/BodyForce.hpp:67-82, 172-201) with the caller's function
// f(x) -> VecDoF (or a double for a scalar field): evaluated here at x(xi_q)
// of every element's quadrature points and uploaded as a per-point table
// (ins_set_body_force_values); the device integrates f phi w detJ.
template <typename FTB, unsigned DEG, Shape SH, typename FB, typename FUN>
void bodyForceComputation(const Quadrature<DEG, SH>&, solver::Hip& solver, FB& fieldBinder,
                          const FUN& forceFun)
{
    constexpr unsigned D = ShapeDim<SH>::value;
    const insilico::Mesh& m = fieldBinder.mesh;
    const int dpn = fieldBinder.field.size;
    if (m.dim != (int)D) throw insilico::Error(INS_ERR_ARG, "quadrature shape and mesh dimension differ");
    std::vector<double> w, xi;
    insilico::detail_::quadratureRule(m.dim, Quadrature<DEG, SH>::numPointsPerDir, w, xi);
    const std::size_t nq = w.size(), ne = m.numElements();
    std::vector<double> tab(ne * nq * dpn);
    for (std::size_t e = 0; e < ne; e++)
        for (std::size_t q = 0; q < nq; q++) {
            const std::array<double, 3> p = insilico::detail_::q1Point(m, e, &xi[q * D]);
            VecN<D> x;
            for (unsigned d = 0; d < D; d++) x[d] = p[d];
            const auto f = forceFun(x);
            for (int c = 0; c < dpn; c++)
                tab[(e * nq + q) * dpn + c] = insilico::detail_::component(f, c);
        }
    solver.bind(fieldBinder.mesh, fieldBinder.field);
    solver.requestBodyForceValues(std::move(tab));
}

// neumannForceComputation<SFTB>(surfaceQuadrature, solver, surfaceFieldBinder,
// ff) (base/asmb/NeumannForce.hpp:36-63, 132-175) with the caller's
// function ff(x, n) -> VecDoF: evaluated here at every boundary face's
// surface points x(eta_q) with the unit normal n(eta_q), uploaded as a
// per-point table (ins_set_traction_values); the device integrates
// ff phi w detG.  The surface rule must have the domain rule's points per
// direction (the reference drivers use one kernel degree for both).
template <typename SFTB, unsigned DEG, Shape SH, typename BMESH, typename FIELD, typename FUN>
void neumannForceComputation(const SurfaceQuadrature<DEG, SH>&, solver::Hip& solver,
                             SurfaceFieldBinder<BMESH, FIELD>& sfb, const FUN& ff)
{
    constexpr unsigned D = ShapeDim<SH>::value;
    const insilico::Mesh& m = *sfb.boundaryMesh.volume;
    const insilico::Field& f = sfb.field;
    const int dpn = f.size;
    if (m.dim != (int)D) throw insilico::Error(INS_ERR_ARG, "quadrature shape and mesh dimension differ");
    const int npts = SurfaceQuadrature<DEG, SH>::numPointsPerDir;
    std::vector<double> w, eta;
    insilico::detail_::quadratureRule(m.dim - 1, npts, w, eta);
    const std::size_t nq = w.size();
    const auto& faces = sfb.boundaryMesh.faces;
    std::vector<double> tab(faces.size() * nq * dpn);
    for (std::size_t k = 0; k < faces.size(); k++)
        for (std::size_t q = 0; q < nq; q++) {
            std::array<double, 3> p, nrm;
            insilico::detail_::facePoint(m, f, faces[k], &eta[q * (m.dim - 1)], p, nrm);
            VecN<D> x, n;
            for (unsigned d = 0; d < D; d++) {
                x[d] = p[d];
                n[d] = nrm[d];
            }
            const auto t = ff(x, n);
            for (int c = 0; c < dpn; c++)
                tab[(k * nq + q) * dpn + c] = insilico::detail_::component(t, c);
        }
    solver.requestTractionValues(faces, std::move(tab), npts);
}

// constant traction on the given boundary faces
template <typename FB>
void neumannForceComputation(solver::Hip& solver, FB& fieldBinder,
                             std::vector<std::vector<int32_t>> faces, const std::array<double, 3>& t)
{
    solver.bind(fieldBinder.mesh, fieldBinder.field);
    solver.requestTraction(std::move(faces), t);
}
// a traction growing linearly across the face, t(x) = t * (x[axis] - x0),
// as the Neumann function of reference/06-elastic/PulledSheet.hpp:63-77
// (t_x = value * (y - 1/2) on x = 1)
struct LinearTraction {
    std::array<double, 3> t;
    int axis;
    double x0;
};
template <typename FB>
void neumannForceComputation(solver::Hip& solver, FB& fieldBinder,
                             std::vector<std::vector<int32_t>> faces, const LinearTraction& lt)
{
    solver.bind(fieldBinder.mesh, fieldBinder.field);
    solver.requestTraction(std::move(faces), lt.t, lt.axis, lt.x0);
}
// Neumann data grad u . n of u = A * prod_d cos(k_d x_d), n the outward unit
// normal of the face (reference/05-mixedPoisson/GivenData.hpp neumannBC)
struct CosineProductFlux {
    double amplitude;
    std::array<double, 3> k;
};
template <typename FB>
void neumannForceComputation(solver::Hip& solver, FB& fieldBinder,
                             std::vector<std::vector<int32_t>> faces, const CosineProductFlux& fl)
{
    solver.bind(fieldBinder.mesh, fieldBinder.field);
    solver.requestTraction(std::move(faces), {0., 0., 0.});
    solver.requestFlux(INS_REF_COSPROD, {fl.amplitude, fl.k[0], fl.k[1], fl.k[2], 0., 0., 0., 0.});
}

}  // namespace asmb

namespace dof {
// base/dof/Distribute.hpp:36-50
template <typename FIELD>
void setDoFsFromSolver(const solver::Hip& s, FIELD& f) { s.distribute(f, false); }
template <typename FIELD>
void addToDoFsFromSolver(const solver::Hip& s, FIELD& f) { s.distribute(f, true); }
}  // namespace dof
}  // namespace base
