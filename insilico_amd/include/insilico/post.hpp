// Host-side input/output and post-processing used by the example drivers:
// SMF mesh reader, elastostatic fundamental solution and the L2 error norm.
#pragma once

#include <cmath>
#include <fstream>
#include <sstream>
#include <string>

#include "insilico.hpp"

namespace insilico {

// SMF mesh files (base/io/smf/Reader.hpp:97-312): '#' comments, '!' header
// lines, "numNodes numElements", coordinates, connectivity
inline Mesh readSMF(const std::string& path, int dim)
{
    std::ifstream in(path);
    if (!in) throw Error(INS_ERR_ARG, "cannot open " + path);
    std::string line;
    Mesh m;
    m.dim = dim;
    std::size_t nn = 0, ne = 0;
    while (std::getline(in, line)) {
        if (line.empty() || line[0] == '#' || line[0] == '!') continue;
        std::istringstream ss(line);
        ss >> nn >> ne;
        break;
    }
    for (std::size_t i = 0; i < nn; i++) {
        std::getline(in, line);
        std::istringstream ss(line);
        double x[3] = {0, 0, 0};
        ss >> x[0] >> x[1] >> x[2];
        for (int d = 0; d < dim; d++) m.xyz.push_back(x[d]);
    }
    for (std::size_t e = 0; e < ne; e++) {
        std::getline(in, line);
        std::istringstream ss(line);
        for (int k = 0; k < (1 << dim); k++) {
            int32_t v;
            ss >> v;
            m.conn.push_back(v);
        }
    }
    return m;
}

// Kelvin solution U(y, x) dir (base/auxi/FundamentalSolution.hpp:160-200)
inline std::array<double, 3> fundSolElastoStatic(int dim, double lambda, double mu,
                                                 const double* y, const double* x,
                                                 const double* dir)
{
    const double G = mu, nu = lambda / 2. / (lambda + mu);
    const double surf = 2. * (dim - 1.) * M_PI;
    const double fac1 = 1. / (4. * surf * G * (1. - nu)), fac2 = 3. - 4. * nu;
    double R[3] = {0, 0, 0}, dist = 0.;
    for (int d = 0; d < dim; d++) {
        R[d] = y[d] - x[d];
        dist += R[d] * R[d];
    }
    dist = std::sqrt(dist);
    const double lk = dim == 2 ? std::log(1. / dist) : 1. / dist;
    const double pw = dim == 2 ? dist * dist : dist * dist * dist;
    std::array<double, 3> out{0., 0., 0.};
    for (int i = 0; i < dim; i++)
        for (int j = 0; j < dim; j++)
            out[i] += fac1 * ((i == j ? fac2 * lk : 0.) + R[i] * R[j] / pw) * dir[j];
    return out;
}

// Q1 shape functions in hierarchic order (base/sfun/TensorProduct.hpp)
inline void q1Shape(int dim, const double* xi, double* phi, double* dphi)
{
    static const int H2[4] = {0, 1, 3, 2};
    static const int H3[8] = {0, 1, 3, 2, 4, 5, 7, 6};
    double f[3][2], g[3][2];
    for (int d = 0; d < dim; d++) {
        f[d][0] = 1. - xi[d];
        f[d][1] = xi[d];
        g[d][0] = -1.;
        g[d][1] = 1.;
    }
    int c = 0;
    for (int k = 0; k < (dim == 3 ? 2 : 1); k++)
        for (int j = 0; j < 2; j++)
            for (int i = 0; i < 2; i++) {
                const double fz = dim == 3 ? f[2][k] : 1.;
                const int h = dim == 3 ? H3[c] : H2[c];
                phi[h] = f[0][i] * f[1][j] * fz;
                dphi[h * dim + 0] = g[0][i] * f[1][j] * fz;
                dphi[h * dim + 1] = f[0][i] * g[1][j] * fz;
                if (dim == 3) dphi[h * dim + 2] = f[0][i] * f[1][j] * g[2][k];
                c++;
            }
}

// L2 error of a Q1 field against a reference function (base/post/ErrorNorm.hpp
// :84-125), 2 Gauss points per direction
template <typename REF>
double l2Error(const Mesh& m, const Field& f, REF ref)
{
    const int dim = m.dim, npe = m.npe(), s = f.size;
    const double gp[2] = {0.788675134594813, 0.211324865405187};
    double err = 0.;
    for (std::size_t e = 0; e < m.numElements(); e++) {
        for (int q = 0; q < (1 << dim); q++) {
            double xi[3] = {gp[q & 1], gp[(q >> 1) & 1], gp[(q >> 2) & 1]};
            const double w = dim == 3 ? 0.125 : 0.25;
            double phi[8], dphi[24];
            q1Shape(dim, xi, phi, dphi);
            double J[3][3] = {{0}}, x[3] = {0, 0, 0}, u[3] = {0, 0, 0};
            for (int n = 0; n < npe; n++) {
                const int32_t node = m.conn[e * npe + n];
                for (int i = 0; i < dim; i++) {
                    x[i] += m.xyz[node * dim + i] * phi[n];
                    for (int a = 0; a < dim; a++) J[i][a] += m.xyz[node * dim + i] * dphi[n * dim + a];
                }
                for (int c = 0; c < s; c++) u[c] += f.value[node * s + c] * phi[n];
            }
            const double det = dim == 2 ? J[0][0] * J[1][1] - J[0][1] * J[1][0]
                                        : J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                                              J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                                              J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
            const std::array<double, 3> r = ref(x);
            double d2 = 0.;
            for (int c = 0; c < s; c++) d2 += (u[c] - r[c]) * (u[c] - r[c]);
            err += d2 * w * det;
        }
    }
    return std::sqrt(err);
}

}  // namespace insilico

// ---------------------------------------------------------------------------
// base::post::errorComputation<ORDER> (base/post/ErrorNorm.hpp:60-84) on the
// device, for the reference solutions the library evaluates itself
namespace insilico {
// u = amp * prod_d sin(k_d x_d)
struct SineProductSolution {
    double amp;
    std::array<double, 3> k;
};
// elastostatic fundamental solution (base/auxi/FundamentalSolution.hpp)
struct FundamentalSolutionElastoStatic {
    double lambda, mu;
    std::array<double, 3> source, direction;
};
namespace detail_ {
inline void refParams(const SineProductSolution& r, int& kind, double* p)
{
    kind = INS_REF_SINPROD;
    p[0] = r.amp;
    for (int d = 0; d < 3; d++) p[1 + d] = r.k[d];
}
inline void refParams(const FundamentalSolutionElastoStatic& r, int& kind, double* p)
{
    kind = INS_REF_KELVIN;
    p[0] = r.lambda;
    p[1] = r.mu;
    for (int d = 0; d < 3; d++) {
        p[2 + d] = r.source[d];
        p[5 + d] = r.direction[d];
    }
}
}  // namespace detail_
}  // namespace insilico

namespace base {
namespace post {
// |u - u_h| in the L2 norm (ORDER 0) or the H1 semi-norm (ORDER 1) of the
// field's current values, with the given quadrature
template <unsigned ORDER, unsigned DEG, Shape SHAPE, typename REF>
double errorComputation(const Quadrature<DEG, SHAPE>&, const insilico::Mesh& m,
                        const insilico::Field& f, const REF& ref, int device = 0)
{
    ins_ctx* c = nullptr;
    insilico::check(ins_create(&c, device, 0), nullptr);
    struct Guard {
        ins_ctx* c;
        ~Guard() { ins_destroy(c); }
    } guard{c};
    insilico::check(ins_set_mesh(c, m.dim, (int64_t)m.numNodes(), m.xyz.data(),
                                 (int64_t)m.numElements(), m.npe(), m.conn.data()), c);
    if (f.degree == 2)
        insilico::check(ins_set_field(c, 2, (int64_t)f.numNodes(), m.dim == 2 ? 9 : 27, f.conn.data()), c);
    // all components active: the norm reads the field's values only
    std::vector<int32_t> ids(f.status.size());
    for (std::size_t k = 0; k < ids.size(); k++) ids[k] = (int32_t)k;
    insilico::check(ins_set_dofs(c, f.size, ids.data(), (int64_t)ids.size()), c);
    std::vector<double> one(m.numElements(), 1.0);
    insilico::check(ins_set_material(c, INS_LAPLACE, Quadrature<DEG, SHAPE>::numPointsPerDir, 1,
                                     one.data()), c);
    int kind = INS_REF_ZERO;
    double p[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    insilico::detail_::refParams(ref, kind, p);
    double out = 0.;
    insilico::check(ins_error_norm(c, (int)ORDER, kind, p, f.value.data(), &out), c);
    return out;
}
}  // namespace post
}  // namespace base
