/*
 * CPU oracle: TEST INFRASTRUCTURE ONLY (see oracle.h).
 * Restatement of the reference hot path, function by function, each citing
 * the reference file:line it follows.  Compiled with gcc -O3 (no -march, so
 * no FMA contraction), like the reference release flags
 * config/config.default:102.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* ------------------------------------------------------------------------ */
/* Gauss-Legendre rules on [0,1] with the reference's 15-digit constants and
 * point order: base/quad/GaussLegendre.hpp:33-36 (count), :114-134 (data). */
int orc_gauss1d(int npts, double* w, double* x)
{
    switch (npts) {
    case 1: w[0] = 1.0; x[0] = 0.5; return 0;
    case 2:
        w[0] = 0.5; x[0] = 0.788675134594813;
        w[1] = 0.5; x[1] = 0.211324865405187;
        return 0;
    case 3:
        w[0] = 0.277777777777777; x[0] = 0.887298334620741;
        w[1] = 0.444444444444444; x[1] = 0.5;
        w[2] = 0.277777777777777; x[2] = 0.112701665379259;
        return 0;
    default: return -1;
    }
}

/* Tensor-product rule: base/quad/TensorProduct.hpp:123-175.  Index
 * outer*lowerN+inner, first coordinate fastest; weight = w_lower * w_outer. */
int orc_tensor_rule(int dim, int npts, double* w, double* xi)
{
    double w1[4], x1[4];
    if (orc_gauss1d(npts, w1, x1)) return -1;
    int nq = 1;
    for (int d = 0; d < dim; d++) nq *= npts;
    for (int q = 0; q < nq; q++) {
        int r = q;
        int i0 = r % npts; r /= npts;
        double ww = w1[i0];
        xi[q * dim + 0] = x1[i0];
        for (int d = 1; d < dim; d++) {
            int id = r % npts; r /= npts;
            ww *= w1[id];
            xi[q * dim + d] = x1[id];
        }
        w[q] = ww;
    }
    return nq;
}

/* Q1 tensor-product shape functions and local gradients in hierarchic order.
 * 1-D: base/sfun/Lagrange1D.ipp:46-58; tensor products and their operand
 * order: base/sfun/TensorProduct.hpp:225-320; lexicographic -> hierarchic
 * reorder out[H[n]] = in[n]: base/sfun/TensorProduct.hpp:40-57 with the
 * vertex table of base/mesh/HierarchicOrder.hpp:201-272,345-392. */
void orc_q1_shape(int dim, const double* xi, double* phi, double* dphi)
{
    static const int H2[4] = {0, 1, 3, 2};
    static const int H3[8] = {0, 1, 3, 2, 4, 5, 7, 6};
    double f[3][2], g[3][2];
    for (int d = 0; d < dim; d++) {
        f[d][0] = 1. - xi[d]; f[d][1] = xi[d];
        g[d][0] = -1.;        g[d][1] = 1.;
    }
    double lf[8], lg[8][3];
    /* 2-D lexicographic: ctr over (outer=eta fn, inner=xi fn) */
    double f2[4], g2[4][2];
    int c = 0;
    for (int o = 0; o < 2; o++)
        for (int i = 0; i < 2; i++) {
            f2[c] = f[0][i] * f[1][o];
            g2[c][0] = f[1][o] * g[0][i];
            g2[c][1] = g[1][o] * f[0][i];
            c++;
        }
    if (dim == 2) {
        for (int n = 0; n < 4; n++) {
            lf[n] = f2[n]; lg[n][0] = g2[n][0]; lg[n][1] = g2[n][1];
        }
        for (int n = 0; n < 4; n++) {
            if (phi) phi[H2[n]] = lf[n];
            if (dphi) { dphi[H2[n] * 2 + 0] = lg[n][0]; dphi[H2[n] * 2 + 1] = lg[n][1]; }
        }
        return;
    }
    c = 0;
    for (int o = 0; o < 2; o++)
        for (int i = 0; i < 4; i++) {
            lf[c] = f2[i] * f[2][o];
            lg[c][0] = f[2][o] * g2[i][0];
            lg[c][1] = f[2][o] * g2[i][1];
            lg[c][2] = g[2][o] * f2[i];
            c++;
        }
    for (int n = 0; n < 8; n++) {
        if (phi) phi[H3[n]] = lf[n];
        if (dphi)
            for (int d = 0; d < 3; d++) dphi[H3[n] * 3 + d] = lg[n][d];
    }
}

/* Hierarchic look-up table H (lexicographic position -> hierarchic index) of
 * the (K+1)^dim tensor-product element, K = 1 or 2:
 * base/mesh/HierarchicOrder.hpp:201-270 (QUAD: vertices {0, K, K(K+2),
 * K(K+1)}, then the edges (0,1),(1,2),(2,3),(3,0) walked from their first
 * vertex, then the interior) and :345-452 (HEX: 8 vertices, 12 edges, 6
 * faces spanned by (v1,v2,v3) = (0,1,3),(4,5,7),(0,1,4),(1,2,5),(2,3,6),
 * (3,0,7), then the interior). */
static void hier_table(int dim, int K, int* H)
{
    const int n1 = K + 1, dZ = K * n1 * n1;
    int V[8] = {0, K, K * (K + 2), K * (K + 1), dZ, K + dZ, K * (K + 2) + dZ, K * (K + 1) + dZ};
    static const int E2[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
    static const int E3[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                  {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
    static const int F3[6][3] = {{0, 1, 3}, {4, 5, 7}, {0, 1, 4}, {1, 2, 5}, {2, 3, 6}, {3, 0, 7}};
    int ctr = 0;
    const int nv = dim == 2 ? 4 : 8, ne = dim == 2 ? 4 : 12;
    for (int i = 0; i < nv; i++) H[V[i]] = ctr++;
    for (int i = 0; i < ne; i++) {
        const int a = V[dim == 2 ? E2[i][0] : E3[i][0]], b = V[dim == 2 ? E2[i][1] : E3[i][1]];
        const int delta = (b - a) / K;
        for (int n = 0; n < K - 1; n++) H[a + delta + n * delta] = ctr++;
    }
    if (dim == 2) {
        const int d1 = (V[1] - V[0]) / K, d2 = (V[3] - V[0]) / K;
        for (int m2 = 0; m2 < K - 1; m2++)
            for (int m1 = 0; m1 < K - 1; m1++) H[V[0] + d1 + d2 + m1 * d1 + m2 * d2] = ctr++;
        return;
    }
    for (int f = 0; f < 6; f++) {
        const int a = V[F3[f][0]];
        const int d1 = (V[F3[f][1]] - a) / K, d2 = (V[F3[f][2]] - a) / K;
        for (int m2 = 0; m2 < K - 1; m2++)
            for (int m1 = 0; m1 < K - 1; m1++) H[a + d1 + d2 + m1 * d1 + m2 * d2] = ctr++;
    }
    const int d1 = (V[1] - V[0]) / K, d2 = (V[3] - V[0]) / K, d3 = (V[4] - V[0]) / K;
    for (int m3 = 0; m3 < K - 1; m3++)
        for (int m2 = 0; m2 < K - 1; m2++)
            for (int m1 = 0; m1 < K - 1; m1++)
                H[V[0] + d1 + d2 + d3 + m1 * d1 + m2 * d2 + m3 * d3] = ctr++;
}

/* 1-D Lagrange functions of degree 1 (base/sfun/Lagrange1D.ipp:46-58) and
 * degree 2 (:76-91; support points 0, 1/2, 1 in that order). */
static void lagrange1d(int K, double x, double* f, double* g)
{
    if (K == 1) {
        f[0] = 1. - x; f[1] = x;
        g[0] = -1.;    g[1] = 1.;
        return;
    }
    f[0] = (1. - x) * (1. - 2. * x);
    f[1] = 4. * x * (1. - x);
    f[2] = x * (2. * x - 1.);
    g[0] = 4. * x - 3.;
    g[1] = 4. - 8. * x;
    g[2] = 4. * x - 1.;
}

/* Tensor-product Lagrange shape functions of degree K (1 or 2) on QUAD/HEX
 * in hierarchic order: the recursive products of
 * base/sfun/TensorProduct.hpp:225-320 (value = lower * outer; gradient
 * components below the last = outer * lower-gradient, last = outer-gradient
 * * lower), then out[H[n]] = in[n] (:40-57).  For K = 1 this is
 * orc_q1_shape. */
void orc_lagrange_shape(int dim, int K, const double* xi, double* phi, double* dphi)
{
    const int n1 = K + 1;
    double f[3][3], g[3][3];
    for (int d = 0; d < dim; d++) lagrange1d(K, xi[d], f[d], g[d]);
    double lf[27], lg[27][3];
    /* 1-D */
    int m = n1;
    for (int i = 0; i < n1; i++) { lf[i] = f[0][i]; lg[i][0] = g[0][i]; }
    for (int d = 1; d < dim; d++) {
        double tf[27], tg[27][3];
        int c = 0;
        for (int o = 0; o < n1; o++)
            for (int i = 0; i < m; i++) {
                tf[c] = lf[i] * f[d][o];
                for (int a = 0; a < d; a++) tg[c][a] = f[d][o] * lg[i][a];
                tg[c][d] = g[d][o] * lf[i];
                c++;
            }
        m = c;
        for (int i = 0; i < m; i++) {
            lf[i] = tf[i];
            for (int a = 0; a <= d; a++) lg[i][a] = tg[i][a];
        }
    }
    int H[27];
    hier_table(dim, K, H);
    for (int n = 0; n < m; n++) {
        if (phi) phi[H[n]] = lf[n];
        if (dphi)
            for (int a = 0; a < dim; a++) dphi[H[n] * dim + a] = lg[n][a];
    }
}

/* Jacobi matrix J(i,a) = sum_n X_n[i] dphi_n[a]: base/geometry.hpp:143-177.
 * Contra-variant basis inv(J^T) with det(J^T) by the cofactor (adjugate)
 * method, as Eigen's computeInverseAndDetWithCheck does for 2x2/3x3:
 * base/geometry.hpp:420-445.  grad_x phi_n = inv(J^T) dphi_n:
 * base/LagrangeShapeFun.hpp:182-200.  Returns detJ. */
static double contra_field(int dim, const double* X, const double* dphi, int nf,
                           const double* fdphi, double* gradx);

double orc_contravariant(int dim, const double* X, const double* dphi,
                         double* gradx)
{
    return contra_field(dim, X, dphi, 1 << dim, dphi, gradx);
}

/* as orc_contravariant, with the inverse applied to the local gradients of
 * nf field functions (Q2 on Q1 geometry: base/LagrangeShapeFun.hpp:182-200
 * maps every field gradient by the geometry's contra-variant basis) */
static double contra_field(int dim, const double* X, const double* dphi, int nf,
                           const double* fdphi, double* gradx)
{
    const int npe = 1 << dim;
    double J[3][3] = {{0}};
    for (int n = 0; n < npe; n++)
        for (int i = 0; i < dim; i++)
            for (int a = 0; a < dim; a++)
                J[i][a] += X[n * dim + i] * dphi[n * dim + a];
    double m[3][3] = {{0}}, inv[3][3], det;
    for (int i = 0; i < dim; i++)
        for (int a = 0; a < dim; a++) m[i][a] = J[a][i]; /* m = J^T */
    if (dim == 2) {
        det = m[0][0] * m[1][1] - m[1][0] * m[0][1];
        const double id = 1. / det;
        inv[0][0] = m[1][1] * id;  inv[1][0] = -m[1][0] * id;
        inv[0][1] = -m[0][1] * id; inv[1][1] = m[0][0] * id;
    } else {
        double cof[3][3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                const int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
                const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
                cof[i][j] = m[i1][j1] * m[i2][j2] - m[i1][j2] * m[i2][j1];
            }
        det = cof[0][0] * m[0][0] + cof[1][0] * m[1][0] + cof[2][0] * m[2][0];
        const double id = 1. / det;
        for (int r = 0; r < 3; r++)
            for (int c = 0; c < 3; c++) inv[r][c] = cof[c][r] * id;
    }
    for (int n = 0; n < nf; n++)
        for (int i = 0; i < dim; i++) {
            double s = 0.;
            for (int a = 0; a < dim; a++) s += inv[i][a] * fdphi[n * dim + a];
            gradx[n * dim + i] = s;
        }
    return det;
}

/* field nodes of an element: the geometry nodes (iso-parametric Q1) or the
 * degree-2 field nodes */
static inline int field_npe(const orc_problem* p) { return p->deg == 2 ? p->fnpe : p->npe; }
static inline const int32_t* field_conn(const orc_problem* p)
{
    return p->deg == 2 ? p->fconn : p->conn;
}
#define ORC_MAXF 27            /* field nodes per element */
#define ORC_MAXE (ORC_MAXF * 3) /* element DoFs */

/* 3x3 determinant: mat/TensorAlgebra.hpp:55-60 */
static double det3(const double A[3][3])
{
    return (A[0][0] * A[1][1] * A[2][2] + A[0][1] * A[1][2] * A[2][0] +
            A[0][2] * A[1][0] * A[2][1] - A[0][0] * A[1][2] * A[2][1] -
            A[0][1] * A[1][0] * A[2][2] - A[0][2] * A[1][1] * A[2][0]);
}

/* 3x3 inverse by adjugate (Eigen's Matrix3d::inverse, mat/TensorAlgebra.hpp:67) */
static void inv3(const double m[3][3], double inv[3][3])
{
    double cof[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            const int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
            const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            cof[i][j] = m[i1][j1] * m[i2][j2] - m[i1][j2] * m[i2][j1];
        }
    const double det = cof[0][0] * m[0][0] + cof[1][0] * m[1][0] + cof[2][0] * m[2][0];
    const double id = 1. / det;
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) inv[r][c] = cof[c][r] * id;
}

/* Voigt map: mat/TensorAlgebra.hpp:141-164 */
static int voigt(int i, int j)
{
    static const int map[9] = {0, 3, 4, 1, -1, 5, -1, -1, 2};
    return map[(i + 1) * (j + 1) - 1];
}

/* S and Voigt C for StVenant (mat/hypel/StVenant.hpp:77-114) and compressible
 * neo-Hookean (mat/hypel/NeoHookeanCompressible.hpp:77-172). */
static void material(int kernel, double lam, double mu, const double F[3][3],
                     double S[3][3], double C[6][6])
{
    double CG[3][3];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double s = 0.;
            for (int k = 0; k < 3; k++) s += F[k][i] * F[k][j];
            CG[i][j] = s;
        }
    if (kernel == ORC_STVENANT) {
        double E[3][3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) E[i][j] = 0.5 * (CG[i][j] - (i == j ? 1. : 0.));
        const double trE = E[0][0] + E[1][1] + E[2][2];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                S[i][j] = lam * trE * (i == j ? 1. : 0.) + 2. * mu * E[i][j];
        if (C) {
            for (int i = 0; i < 6; i++)
                for (int j = 0; j < 6; j++) C[i][j] = 0.;
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) C[i][j] = lam;
            for (int i = 0; i < 3; i++) C[i][i] += 2. * mu;
            for (int i = 3; i < 6; i++) C[i][i] += mu;
        }
    } else {
        double Ci[3][3];
        inv3(CG, Ci);
        const double J = det3(F);
        const double logJ = log(J);
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++)
                S[i][j] = (lam * logJ - mu) * Ci[i][j] + mu * (i == j ? 1. : 0.);
        if (C) {
            const double fac2 = mu - lam * log(J);
            for (int A = 0; A < 3; A++)
                for (int B = A; B < 3; B++) {
                    const int AB = voigt(A, B);
                    for (int c = 0; c < 3; c++)
                        for (int D = c; D < 3; D++) {
                            const double v = lam * Ci[A][B] * Ci[c][D] +
                                fac2 * (Ci[A][c] * Ci[B][D] + Ci[A][D] * Ci[B][c]);
                            C[AB][voigt(c, D)] = v;
                        }
                }
        }
    }
}

static double body_force(const orc_problem* p, const double* x, int comp, int64_t e, int q,
                         int nq)
{
    if (p->bf_kind == ORC_BF_TABLE) return p->bf_tab[(e * nq + q) * p->dpn + comp];
    if (p->bf_kind == ORC_BF_CONST) return p->bf[comp];
    if (p->bf_kind == ORC_BF_SINPROD) {
        double f = p->bf[0];
        for (int d = 0; d < p->dim; d++) f *= sin(p->bf[1 + d] * x[d]);
        return f;
    }
    if (p->bf_kind == ORC_BF_COSPROD) {
        /* f = -Laplace(u) of reference/05-mixedPoisson/GivenData.hpp
         * forceFun, amp = sum_d alpha_d^2 */
        double f = p->bf[0];
        for (int d = 0; d < p->dim; d++) f *= cos(p->bf[1 + d] * x[d]);
        return f;
    }
    return 0.;
}

static void ref_solution(int dim, int s, int kind, const double* prm, const double* x,
                         double* val, double* grad);

/* Element matrix Ke (nee x nee, row-major) and element force vector Fe (nee):
 *  - Laplace: heat/Laplace.hpp:113-126 -> base/kernel/Laplace.hpp:100-151
 *  - HyperElastic tangent: solid/HyperElastic.hpp:110-179,282-310, with F
 *    from solid/Deformation.hpp:26-46 and base/post/evaluateField.hpp:229-274
 *  - Fe = body force (base/asmb/BodyForce.hpp:172-201) and, if p->residual,
 *    -residualForce (solid/HyperElastic.hpp:210-257; factor -1 of
 *    base/asmb/ForceIntegrator.hpp:55).
 * Quadrature loop: base/Quadrature.hpp:133-141 over the tensor rule.  Body
 * force and residual are kept as separately ordered sums per element, as
 * their integrators are separate element loops in the reference. */
int orc_element(const orc_problem* p, int64_t e, double* Ke, double* Fe)
{
    const int dim = p->dim, gnpe = p->npe, s = p->dpn;
    const int npe = field_npe(p), nee = npe * s, K = p->deg == 2 ? 2 : 1;
    const int32_t* fc = field_conn(p);
    double X[8 * 3], U[ORC_MAXE];
    for (int n = 0; n < gnpe; n++) {
        const int32_t node = p->conn[e * gnpe + n];
        for (int d = 0; d < dim; d++) X[n * dim + d] = p->xyz[(int64_t)node * dim + d];
    }
    for (int n = 0; n < npe; n++) {
        const int32_t node = fc[e * npe + n];
        for (int c = 0; c < s; c++) U[n * s + c] = p->u ? p->u[(int64_t)node * s + c] : 0.;
    }
    double wq[64], xq[64 * 3];
    const int nq = orc_tensor_rule(dim, p->nq, wq, xq);
    if (nq < 0) return -1;
    if (Ke) for (int i = 0; i < nee * nee; i++) Ke[i] = 0.;
    double fb[ORC_MAXE], fr[ORC_MAXE];
    for (int i = 0; i < nee; i++) { fb[i] = 0.; fr[i] = 0.; }

    for (int q = 0; q < nq; q++) {
        const double* xi = xq + q * dim;
        const double w = wq[q];
        double gphi[8], gdphi[8 * 3];
        double phi[ORC_MAXF], dphi[ORC_MAXF * 3], gx[ORC_MAXF * 3];
        orc_q1_shape(dim, xi, gphi, gdphi);
        double detJ;
        if (K == 1) {
            for (int n = 0; n < gnpe; n++) phi[n] = gphi[n];
            detJ = orc_contravariant(dim, X, gdphi, gx);
        } else {
            orc_lagrange_shape(dim, K, xi, phi, dphi);
            detJ = contra_field(dim, X, gdphi, npe, dphi, gx);
        }
        if (p->kernel == ORC_LAPLACE && Ke) {
            /* kappa(element, xi): one value per element, or per point */
            const double kappa = p->coef ? p->coef[e * p->ncoef + (p->ncoef == nq ? q : 0)] : 1.0;
            const double scalar = kappa * detJ * w;
            for (int M = 0; M < npe; M++)
                for (int N = 0; N < npe; N++) {
                    double entry = 0.;
                    for (int k = 0; k < dim; k++) entry += gx[M * dim + k] * gx[N * dim + k];
                    entry *= scalar;
                    for (int d = 0; d < s; d++) Ke[(M * s + d) * nee + N * s + d] += entry;
                }
        } else if (p->kernel == ORC_STVENANT || p->kernel == ORC_NEOHOOKE) {
            const double lam = p->coef[e * p->ncoef + 0], mu = p->coef[e * p->ncoef + 1];
            /* GradU(J,i) = sum_f gx_f[J] u_f[i]; F = I + GradU^T */
            double F[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
            double GU[3][3] = {{0}};
            for (int f = 0; f < npe; f++)
                for (int J = 0; J < dim; J++)
                    for (int i = 0; i < s; i++) GU[J][i] += gx[f * dim + J] * U[f * s + i];
            for (int J = 0; J < dim; J++)
                for (int i = 0; i < s; i++) F[i][J] += GU[J][i];
            double S[3][3], C[6][6];
            material(p->kernel, lam, mu, F, S, C);
            if (Ke) {
                for (int M = 0; M < npe; M++)
                    for (int N = 0; N < npe; N++)
                        for (int i = 0; i < s; i++)
                            for (int k = 0; k < s; k++) {
                                double sum = 0.;
                                for (int J = 0; J < s; J++)
                                    for (int L = 0; L < s; L++) {
                                        double ce = (i == k ? S[J][L] : 0.);
                                        for (int A = 0; A < s; A++) {
                                            const int v1 = voigt(A, J);
                                            for (int B = 0; B < s; B++)
                                                ce += F[i][A] * C[v1][voigt(B, L)] * F[k][B];
                                        }
                                        sum += gx[M * dim + J] * ce * gx[N * dim + L];
                                    }
                                sum *= detJ * w;
                                Ke[(M * s + i) * nee + N * s + k] += sum;
                            }
            }
            if (p->residual) {
                double P[3][3];
                for (int i = 0; i < 3; i++)
                    for (int J = 0; J < 3; J++) {
                        double t = 0.;
                        for (int k = 0; k < 3; k++) t += F[i][k] * S[k][J];
                        P[i][J] = t;
                    }
                for (int M = 0; M < npe; M++)
                    for (int i = 0; i < s; i++) {
                        double sum = 0.;
                        for (int J = 0; J < s; J++) sum += P[i][J] * gx[M * dim + J];
                        sum *= detJ * w;
                        fr[M * s + i] += sum;
                    }
            }
        }
        if (p->bf_kind != ORC_BF_NONE) {
            /* Jacobian = det(J) (base/geometry.hpp:227-251,480-489); x(xi) by
             * base/geometry.hpp:95-130 */
            double x[3] = {0, 0, 0};
            for (int n = 0; n < gnpe; n++)
                for (int d = 0; d < dim; d++) x[d] += X[n * dim + d] * gphi[n];
            for (int n = 0; n < npe; n++)
                for (int c = 0; c < s; c++)
                    fb[n * s + c] += body_force(p, x, c, e, q, nq) * phi[n] * w * detJ;
        }
    }
    if (Fe)
        for (int i = 0; i < nee; i++) Fe[i] = fb[i] - fr[i];
    return 0;
}

/* Element mass matrix: base/kernel/Mass.hpp:80-110 (scalar = factor * detJ
 * * weight, entry = scalar * phi_M * phi_N on every component), integrated
 * with the element's quadrature (base/Quadrature.hpp:133-141). */
int orc_mass_element(const orc_problem* p, int64_t e, double* Me)
{
    const int dim = p->dim, gnpe = p->npe, s = p->dpn;
    const int npe = field_npe(p), nee = npe * s, K = p->deg == 2 ? 2 : 1;
    double X[8 * 3];
    for (int n = 0; n < gnpe; n++) {
        const int32_t node = p->conn[e * gnpe + n];
        for (int d = 0; d < dim; d++) X[n * dim + d] = p->xyz[(int64_t)node * dim + d];
    }
    double wq[64], xq[64 * 3];
    const int nq = orc_tensor_rule(dim, p->nq, wq, xq);
    if (nq < 0) return -1;
    for (int i = 0; i < nee * nee; i++) Me[i] = 0.;
    for (int q = 0; q < nq; q++) {
        double gphi[8], gdphi[8 * 3], gx[ORC_MAXF * 3];
        double phi[ORC_MAXF], dphi[ORC_MAXF * 3];
        orc_q1_shape(dim, xq + q * dim, gphi, gdphi);
        double detJ;
        if (K == 1) {
            for (int n = 0; n < gnpe; n++) phi[n] = gphi[n];
            detJ = orc_contravariant(dim, X, gdphi, gx);
        } else {
            orc_lagrange_shape(dim, K, xq + q * dim, phi, dphi);
            detJ = contra_field(dim, X, gdphi, npe, dphi, gx);
        }
        const double scalar = p->density * detJ * wq[q];
        for (int M = 0; M < npe; M++)
            for (int N = 0; N < npe; N++) {
                const double entry = scalar * phi[M] * phi[N];
                for (int d = 0; d < s; d++) Me[(M * s + d) * nee + N * s + d] += entry;
            }
    }
    return 0;
}

/* BDF weights (base/time/BDF.hpp:69-99, as written there) and the start-up
 * rule of base/time/MultiStep.hpp:46-63: step + 2 < order + 1 falls back to
 * the method of one order less. */
int orc_bdf_weights(int order, int step, double* a)
{
    while (order > 1 && step + 2 < order + 1) order--;
    static const double w[4][5] = {{1.0, -1.0}, {1.5, -2.0, 0.5},
                                   {11. / 6., -3.0, 1.5, -1. / 3.},
                                   {25. / 12., -4.0, 3.0, 4. / 3., 1. / 4.}};
    if (order < 1 || order > 4) return -1;
    for (int i = 0; i <= order; i++) a[i] = w[order - 1][i];
    return order + 1;
}

/* Structured unit-cube mesh: tools/meshGeneration/unitCube/unitCube.hpp
 * :85-106 (nodes, i1 fastest, x = h*i with h = 1/e) and :217-257 (elements,
 * i1 fastest, lexicographic corner list reordered to hierarchic order by
 * HierarchicCube :187-205). */
int orc_unit_cube(int dim, int e1, int e2, int e3, double* xyz, int32_t* conn)
{
    const int n1 = e1 + 1, n2 = e2 + 1, n3 = (dim == 3 ? e3 + 1 : 1);
    const double h1 = 1.0 / (double)e1, h2 = 1.0 / (double)e2;
    const double h3 = (dim == 3 ? 1.0 / (double)e3 : 0.);
    int64_t k = 0;
    for (int i3 = 0; i3 < n3; i3++)
        for (int i2 = 0; i2 < n2; i2++)
            for (int i1 = 0; i1 < n1; i1++) {
                xyz[k * dim + 0] = h1 * i1;
                xyz[k * dim + 1] = h2 * i2;
                if (dim == 3) xyz[k * dim + 2] = h3 * i3;
                k++;
            }
    static const int H2[4] = {0, 1, 3, 2};
    static const int H3[8] = {0, 1, 3, 2, 4, 5, 7, 6};
    int64_t e = 0;
    const int ne3 = (dim == 3 ? e3 : 1);
    for (int i3 = 0; i3 < ne3; i3++)
        for (int i2 = 0; i2 < e2; i2++)
            for (int i1 = 0; i1 < e1; i1++) {
                const int64_t base = i1 + (int64_t)i2 * n1 + (int64_t)i3 * n1 * n2;
                int32_t cube[8];
                int c = 0;
                for (int d3 = 0; d3 <= (dim == 3 ? 1 : 0); d3++)
                    for (int d2 = 0; d2 <= 1; d2++)
                        for (int d1 = 0; d1 <= 1; d1++)
                            cube[c++] = (int32_t)(base + (int64_t)d3 * n1 * n2 + d2 * n1 + d1);
                const int npe = 1 << dim;
                for (int v = 0; v < npe; v++)
                    conn[e * npe + (dim == 3 ? H3[v] : H2[v])] = cube[v];
                e++;
            }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Effective (active) DoF ids of an element in local order (node-major,
 * component inner): base/asmb/collectFromDoFs.hpp:83-136 and the active-row
 * collection of base/asmb/assembleMatrix.hpp:243-251. */
static int active_ids(const orc_problem* p, int64_t e, int32_t* ids, int* loc)
{
    int na = 0;
    const int npe = field_npe(p);
    const int32_t* fc = field_conn(p);
    for (int n = 0; n < npe; n++) {
        const int64_t node = fc[e * npe + n];
        for (int c = 0; c < p->dpn; c++) {
            const int32_t id = p->dof_id[node * p->dpn + c];
            if (id >= 0) { ids[na] = id; loc[na] = n * p->dpn + c; na++; }
        }
    }
    return na;
}

static int cmp_i32(const void* a, const void* b)
{
    const int32_t x = *(const int32_t*)a, y = *(const int32_t*)b;
    return (x > y) - (x < y);
}

/* Pattern = sorted unique set of (row,col) over all active pairs of every
 * element, explicit zeros kept: base/solver/TripletContainer.hpp:159-301
 * (std::set insert per pair, then copied to a sorted vector).  Rows are
 * bucketed first so the set is built per row; the resulting ordered set is
 * the same. */
int orc_pattern(const orc_problem* p, int64_t* nnz_out, int32_t* row_ptr,
                int32_t* col_idx)
{
    const int64_t n = p->n_active;
    int64_t* cnt = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
    int32_t ids[ORC_MAXE];
    int loc[ORC_MAXE];
    for (int64_t e = 0; e < p->n_elem; e++) {
        const int na = active_ids(p, e, ids, loc);
        for (int r = 0; r < na; r++) cnt[ids[r] + 1] += na;
    }
    for (int64_t i = 0; i < n; i++) cnt[i + 1] += cnt[i];
    const int64_t tot = cnt[n];
    int32_t* buf = (int32_t*)malloc((size_t)(tot ? tot : 1) * sizeof(int32_t));
    int64_t* pos = (int64_t*)malloc((size_t)(n + 1) * sizeof(int64_t));
    memcpy(pos, cnt, (size_t)(n + 1) * sizeof(int64_t));
    for (int64_t e = 0; e < p->n_elem; e++) {
        const int na = active_ids(p, e, ids, loc);
        for (int r = 0; r < na; r++)
            for (int c = 0; c < na; c++) buf[pos[ids[r]]++] = ids[c];
    }
    int64_t nnz = 0;
    if (row_ptr) row_ptr[0] = 0;
    for (int64_t i = 0; i < n; i++) {
        int32_t* b = buf + cnt[i];
        const int64_t m = cnt[i + 1] - cnt[i];
        qsort(b, (size_t)m, sizeof(int32_t), cmp_i32);
        for (int64_t k = 0; k < m; k++) {
            if (k > 0 && b[k] == b[k - 1]) continue;
            if (col_idx) col_idx[nnz] = b[k];
            nnz++;
        }
        if (row_ptr) row_ptr[i + 1] = (int32_t)nnz;
    }
    free(buf); free(pos); free(cnt);
    *nnz_out = nnz;
    return 0;
}

typedef struct { int32_t row, col; double val; } triplet;

/* lower_bound over the sorted triplet array, as
 * base/solver/TripletContainer.hpp:337-351 */
static int64_t triplet_find(const triplet* t, int64_t nnz, int32_t r, int32_t c)
{
    int64_t lo = 0, hi = nnz;
    while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        const int less = (t[mid].row < r) || (!(r < t[mid].row) && t[mid].col < c);
        if (less) lo = mid + 1; else hi = mid;
    }
    if (lo >= nnz || t[lo].row != r || t[lo].col != c) return -1;
    return lo;
}

/* Face traction: base/asmb/NeumannForce.hpp:132-175 with surface metric from
 * base/geometry.hpp:314-346 (|cross(J)| in 3D, |J| in 2D). */
static void face_force(const orc_problem* p, int64_t f, double* out)
{
    const int dim = p->dim, nfn = p->nfn, s = p->dpn;
    const int K = p->deg == 2 ? 2 : 1, ffn = K == 2 ? p->ffn : nfn;
    double X[4 * 3];
    for (int n = 0; n < nfn; n++)
        for (int d = 0; d < dim; d++)
            X[n * dim + d] = p->xyz[(int64_t)p->faces[f * nfn + n] * dim + d];
    double wq[16], xq[16 * 2];
    const int nq = orc_tensor_rule(dim - 1, p->nq, wq, xq);
    for (int i = 0; i < ffn * s; i++) out[i] = 0.;
    for (int q = 0; q < nq; q++) {
        double phi[9], dphi[4 * 2];
        double detG;
        if (dim == 3) {
            orc_q1_shape(2, xq + q * 2, phi, dphi);
            if (K == 2) orc_lagrange_shape(2, 2, xq + q * 2, phi, NULL);
            double J[3][2] = {{0}};
            for (int n = 0; n < 4; n++)
                for (int i = 0; i < 3; i++)
                    for (int a = 0; a < 2; a++) J[i][a] += X[n * 3 + i] * dphi[n * 2 + a];
            const double cx = J[1][0] * J[2][1] - J[2][0] * J[1][1];
            const double cy = J[2][0] * J[0][1] - J[0][0] * J[2][1];
            const double cz = J[0][0] * J[1][1] - J[1][0] * J[0][1];
            detG = sqrt(cx * cx + cy * cy + cz * cz);
        } else {
            const double eta = xq[q];
            phi[0] = 1. - eta; phi[1] = eta;
            if (K == 2) {
                /* LINE3 hierarchic order: ends first, then the midpoint
                 * (base/mesh/HierarchicOrder.hpp:131-150) */
                double f[3], g[3];
                lagrange1d(2, eta, f, g);
                phi[0] = f[0]; phi[1] = f[2]; phi[2] = f[1];
            }
            const double tx = X[2] - X[0], ty = X[3] - X[1];
            detG = sqrt(tx * tx + ty * ty);
        }
        double t[3] = {p->traction[0], p->traction[1], p->traction[2]};
        if (p->tr_tab) {
            for (int c = 0; c < s; c++) t[c] = p->tr_tab[(f * nq + q) * s + c];
        } else if (p->flux_kind != 0) {
            /* surface point and unit normal of base::SurfaceNormal
             * (base/geometry.hpp:255-345; 2-D: (J10, -J00) from
             * base/linearAlgebra.hpp:93-99; 3-D: J.col(0) x J.col(1)) */
            double gp[4], gd[4 * 2], xs[3] = {0, 0, 0}, nrm[3] = {0, 0, 0};
            if (dim == 3) {
                orc_q1_shape(2, xq + q * 2, gp, gd);
                double J[3][2] = {{0}};
                for (int n = 0; n < 4; n++)
                    for (int i = 0; i < 3; i++)
                        for (int a = 0; a < 2; a++) J[i][a] += X[n * 3 + i] * gd[n * 2 + a];
                nrm[0] = J[1][0] * J[2][1] - J[2][0] * J[1][1];
                nrm[1] = J[2][0] * J[0][1] - J[0][0] * J[2][1];
                nrm[2] = J[0][0] * J[1][1] - J[1][0] * J[0][1];
            } else {
                gp[0] = 1. - xq[q];
                gp[1] = xq[q];
                nrm[0] = X[3] - X[1];
                nrm[1] = -(X[2] - X[0]);
            }
            double len = 0.;
            for (int d = 0; d < dim; d++) len += nrm[d] * nrm[d];
            len = sqrt(len);
            for (int n = 0; n < nfn; n++)
                for (int d = 0; d < dim; d++) xs[d] += X[n * dim + d] * gp[n];
            double rv[3], rg[9];
            ref_solution(dim, s, p->flux_kind, p->flux_p, xs, rv, rg);
            for (int c = 0; c < s; c++) {
                double v = 0.;
                for (int d = 0; d < dim; d++) v += rg[d * s + c] * (nrm[d] / len);
                t[c] = v;
            }
        } else if (p->tr_kind == 1) {
            /* surface point x(eta) of the Q1 face geometry */
            double gp[4], gd[4 * 2], xs = 0.;
            if (dim == 3) orc_q1_shape(2, xq + q * 2, gp, gd);
            else { gp[0] = 1. - xq[q]; gp[1] = xq[q]; }
            for (int n = 0; n < nfn; n++) xs += X[n * dim + p->tr_axis] * gp[n];
            for (int c = 0; c < s; c++) t[c] = p->traction[c] * (xs - p->tr_x0);
        }
        for (int n = 0; n < ffn; n++)
            for (int c = 0; c < s; c++) out[n * s + c] += t[c] * phi[n] * wq[q] * detG;
    }
}

/* Assembly: base/asmb/StiffnessMatrix.hpp:159-225 (element loop body),
 * base/asmb/assembleMatrix.hpp:56-130,213-338 (drop constrained rows, lift
 * constrained columns b_r -= g_c K_rc, insert active block), triplet insert
 * base/solver/TripletContainer.hpp:305-352, element loop
 * base/auxi/parallel.hpp:26-60 (omp atomic adds when nthreads > 1).  Forces
 * (residual, body force, Neumann): base/asmb/ForceIntegrator.hpp:126-160 and
 * base/asmb/assembleForces.hpp:59-139, serial loops. */
int orc_assemble(const orc_problem* p, const int32_t* row_ptr,
                 const int32_t* col_idx, double* val, double* b, double* tm)
{
    const int64_t n = p->n_active;
    const int64_t nnz = row_ptr[n];
    const int s = p->dpn, npe = field_npe(p), nee = npe * s;
    const int32_t* fc = field_conn(p);
    triplet* T = (triplet*)malloc((size_t)(nnz ? nnz : 1) * sizeof(triplet));
    for (int64_t i = 0; i < n; i++)
        for (int64_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
            T[k].row = (int32_t)i; T[k].col = col_idx[k]; T[k].val = 0.;
        }
    for (int64_t i = 0; i < n; i++) b[i] = 0.;
    int err = 0;
    const double t0 = now_s();
    const int nth = p->nthreads > 0 ? p->nthreads : 1;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nth) if (nth > 1)
#endif
    for (int64_t e = 0; e < p->n_elem; e++) {
        double Ke[ORC_MAXE * ORC_MAXE];
        orc_element(p, e, Ke, NULL);
        int32_t ids[ORC_MAXE];
        int loc[ORC_MAXE];
        int st[ORC_MAXE];
        double gv[ORC_MAXE];
        for (int nn = 0; nn < npe; nn++) {
            const int64_t node = fc[e * npe + nn];
            for (int c = 0; c < s; c++) {
                const int64_t k = node * s + c;
                st[nn * s + c] = p->dof_id[k] >= 0;
                gv[nn * s + c] = p->g ? (p->incremental && p->u ? p->g[k] - p->u[k] : p->g[k]) : 0.;
            }
        }
        const int na = active_ids(p, e, ids, loc);
        for (int r = 0; r < na; r++) {
            const int lr = loc[r];
            double sv = 0.;
            for (int c = 0; c < nee; c++)
                if (!st[c]) sv -= gv[c] * 1.0 * Ke[lr * nee + c];
            for (int c = 0; c < na; c++) {
                const int64_t k = triplet_find(T, nnz, ids[r], ids[c]);
                if (k < 0) { err = -2; continue; }
#ifdef _OPENMP
#pragma omp atomic
#endif
                T[k].val += Ke[lr * nee + loc[c]];
            }
#ifdef _OPENMP
#pragma omp atomic
#endif
            b[ids[r]] += sv;
        }
    }
    const double t1 = now_s();
    /* force loops, serial as in the reference */
    if (p->residual || p->bf_kind != ORC_BF_NONE) {
        for (int64_t e = 0; e < p->n_elem; e++) {
            double Fe[ORC_MAXE];
            orc_element(p, e, NULL, Fe);
            int32_t ids[ORC_MAXE];
            int loc[ORC_MAXE];
            const int na = active_ids(p, e, ids, loc);
            for (int r = 0; r < na; r++) b[ids[r]] += Fe[loc[r]];
        }
    }
    const int ffn = p->deg == 2 ? p->ffn : p->nfn;
    const int32_t* ffc = p->deg == 2 ? p->ffaces : p->faces;
    for (int64_t f = 0; f < p->n_faces; f++) {
        double ff[9 * 3];
        face_force(p, f, ff);
        for (int nn = 0; nn < ffn; nn++)
            for (int c = 0; c < s; c++) {
                const int32_t id = p->dof_id[(int64_t)ffc[f * ffn + nn] * s + c];
                if (id >= 0) b[id] += ff[nn * s + c];
            }
    }
    /* time integration terms: base/time/ReactionTerms.hpp:174-230, one
     * more serial element loop (lifting of mass_factor * M_e with the
     * prescribed values, then M_e * mass_hist on the active rows) */
    if (p->density > 0.) {
        for (int64_t e = 0; e < p->n_elem; e++) {
            double Me[ORC_MAXE * ORC_MAXE];
            orc_mass_element(p, e, Me);
            int32_t ids[ORC_MAXE];
            int loc[ORC_MAXE];
            int st[ORC_MAXE];
            double gv[ORC_MAXE], hv[ORC_MAXE];
            for (int nn = 0; nn < npe; nn++) {
                const int64_t node = fc[e * npe + nn];
                for (int c = 0; c < s; c++) {
                    const int64_t k = node * s + c;
                    st[nn * s + c] = p->dof_id[k] >= 0;
                    gv[nn * s + c] = p->g ? p->g[k] : 0.;
                    hv[nn * s + c] = p->mass_hist ? p->mass_hist[k] : 0.;
                }
            }
            const int na = active_ids(p, e, ids, loc);
            for (int r = 0; r < na; r++) {
                const int lr = loc[r];
                double sv = 0.;
                for (int c = 0; c < nee; c++)
                    if (!st[c]) sv -= gv[c] * (p->mass_factor * Me[lr * nee + c]);
                for (int c = 0; c < na; c++) {
                    const int64_t k = triplet_find(T, nnz, ids[r], ids[c]);
                    if (k < 0) { err = -2; continue; }
                    T[k].val += p->mass_factor * Me[lr * nee + loc[c]];
                }
                b[ids[r]] += sv;
            }
            for (int r = 0; r < na; r++) {
                double f = 0.;
                for (int c = 0; c < nee; c++) f += Me[loc[r] * nee + c] * hv[c];
                b[ids[r]] += f;
            }
        }
    }
    for (int64_t k = 0; k < nnz; k++) val[k] = T[k].val;
    const double t2 = now_s();
    if (tm) { tm[0] = t2 - t0; tm[1] = t1 - t0; }
    free(T);
    return err;
}

void orc_spmv(int64_t n, const int32_t* rp, const int32_t* ci, const double* v,
              const double* x, double* y, int nthreads)
{
    const int nth = nthreads > 0 ? nthreads : 1;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(nth) if (nth > 1)
#endif
    for (int64_t i = 0; i < n; i++) {
        double s = 0.;
        for (int32_t k = rp[i]; k < rp[i + 1]; k++) s += v[k] * x[ci[k]];
        y[i] = s;
    }
}

static double dot(int64_t n, const double* a, const double* b, int nth)
{
    double s = 0.;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) reduction(+ : s) num_threads(nth) if (nth > 1)
#endif
    for (int64_t i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

/* Jacobi-preconditioned CG following Eigen::ConjugateGradient as called by
 * base/solver/Eigen3.hpp:265-275: DiagonalPreconditioner (1/diag, 1 where the
 * diagonal is 0), stop when |r|^2 < tol^2 |b|^2 on the recursive residual,
 * iteration limit maxit.  Initial guess x0 = 0. */
int orc_pcg(int64_t n, const int32_t* rp, const int32_t* ci, const double* v,
            const double* b, double* x, double tol, int maxit, int nthreads,
            int* iters, double* relres)
{
    const int nth = nthreads > 0 ? nthreads : 1;
    double* r = (double*)malloc((size_t)n * sizeof(double));
    double* z = (double*)malloc((size_t)n * sizeof(double));
    double* pp = (double*)malloc((size_t)n * sizeof(double));
    double* q = (double*)malloc((size_t)n * sizeof(double));
    double* dinv = (double*)malloc((size_t)n * sizeof(double));
    for (int64_t i = 0; i < n; i++) {
        double d = 0.;
        for (int32_t k = rp[i]; k < rp[i + 1]; k++)
            if (ci[k] == i) d = v[k];
        dinv[i] = d != 0. ? 1. / d : 1.;
        x[i] = 0.;
        r[i] = b[i];
    }
    const double bb = dot(n, b, b, nth);
    int it = 0;
    if (bb == 0.) {
        *iters = 0; *relres = 0.;
        free(r); free(z); free(pp); free(q); free(dinv);
        return 0;
    }
    const double thr = tol * tol * bb;
    double rr = dot(n, r, r, nth);
    if (rr >= thr) {
        for (int64_t i = 0; i < n; i++) pp[i] = dinv[i] * r[i];
        double rz = dot(n, r, pp, nth);
        while (it < maxit) {
            orc_spmv(n, rp, ci, v, pp, q, nth);
            const double alpha = rz / dot(n, pp, q, nth);
            for (int64_t i = 0; i < n; i++) { x[i] += alpha * pp[i]; r[i] -= alpha * q[i]; }
            rr = dot(n, r, r, nth);
            if (rr < thr) break;
            for (int64_t i = 0; i < n; i++) z[i] = dinv[i] * r[i];
            const double old = rz;
            rz = dot(n, r, z, nth);
            const double beta = rz / old;
            for (int64_t i = 0; i < n; i++) pp[i] = z[i] + beta * pp[i];
            it++;
        }
    }
    /* true residual */
    orc_spmv(n, rp, ci, v, x, q, nth);
    double tr = 0.;
    for (int64_t i = 0; i < n; i++) { const double d = b[i] - q[i]; tr += d * d; }
    *iters = it;
    *relres = sqrt(tr / bb);
    free(r); free(z); free(pp); free(q); free(dinv);
    return it >= maxit ? 1 : 0;
}

/* Elastostatic fundamental solution U(y,x)*dir: base/auxi/FundamentalSolution
 * .hpp:160-200 (Kelvin tensor), unit-sphere surface and Laplace kernel
 * :27-64. */
void orc_fundsol_elasto(int dim, double lambda, double mu, const double* y,
                        const double* x, const double* dir, double* out)
{
    const double G = mu, nu = lambda / 2. / (lambda + mu);
    const double surf = 2. * (dim - 1.) * M_PI;
    const double fac1 = 1. / (4. * surf * G * (1. - nu));
    const double fac2 = 3. - 4. * nu;
    double R[3], dist = 0.;
    for (int d = 0; d < dim; d++) { R[d] = y[d] - x[d]; dist += R[d] * R[d]; }
    dist = sqrt(dist);
    const double lk = (dim == 2) ? log(1. / dist) : 1. / dist;
    const double pw = (dim == 2) ? dist * dist : dist * dist * dist;
    for (int i = 0; i < dim; i++) {
        double s = 0.;
        for (int j = 0; j < dim; j++) {
            const double Uij = fac1 * ((i == j ? fac2 * lk : 0.) + (R[i] * R[j]) / pw);
            s += Uij * dir[j];
        }
        out[i] = s;
    }
}

/* L2 error norm: base/post/ErrorNorm.hpp:84-125 with the element kernel
 * evaluating |u_h(xi) - u(x(xi))|^2 w detJ per quadrature point. */
double orc_l2_error_fundsol(const orc_problem* p, const double* uh,
                            double lambda, double mu, const double* src,
                            const double* dir)
{
    const int dim = p->dim, npe = p->npe, s = p->dpn;
    double wq[64], xq[64 * 3];
    const int nq = orc_tensor_rule(dim, p->nq, wq, xq);
    double err = 0.;
    for (int64_t e = 0; e < p->n_elem; e++) {
        double X[24], U[24];
        for (int n = 0; n < npe; n++) {
            const int64_t node = p->conn[e * npe + n];
            for (int d = 0; d < dim; d++) X[n * dim + d] = p->xyz[node * dim + d];
            for (int c = 0; c < s; c++) U[n * s + c] = uh[node * s + c];
        }
        for (int q = 0; q < nq; q++) {
            double phi[8], dphi[24], gx[24];
            orc_q1_shape(dim, xq + q * dim, phi, dphi);
            const double detJ = orc_contravariant(dim, X, dphi, gx);
            double x[3] = {0, 0, 0}, v[3] = {0, 0, 0}, ref[3];
            for (int n = 0; n < npe; n++) {
                for (int d = 0; d < dim; d++) x[d] += X[n * dim + d] * phi[n];
                for (int c = 0; c < s; c++) v[c] += U[n * s + c] * phi[n];
            }
            orc_fundsol_elasto(dim, lambda, mu, x, src, dir, ref);
            double d2 = 0.;
            for (int c = 0; c < s; c++) d2 += (v[c] - ref[c]) * (v[c] - ref[c]);
            err += d2 * wq[q] * detJ;
        }
    }
    return sqrt(err);
}

/* Reference solution and its gradient at x (value[c], grad[d*s+c]). */
static void ref_solution(int dim, int s, int kind, const double* prm, const double* x,
                         double* val, double* grad)
{
    for (int c = 0; c < s; c++) val[c] = 0.;
    for (int i = 0; i < dim * s; i++) grad[i] = 0.;
    if (kind == ORC_REF_SINPROD) {
        double v = prm[0];
        for (int d = 0; d < dim; d++) v *= sin(prm[1 + d] * x[d]);
        for (int c = 0; c < s; c++) val[c] = v;
        for (int d = 0; d < dim; d++) {
            double g = prm[0] * prm[1 + d] * cos(prm[1 + d] * x[d]);
            for (int e = 0; e < dim; e++)
                if (e != d) g *= sin(prm[1 + e] * x[e]);
            for (int c = 0; c < s; c++) grad[d * s + c] = g;
        }
    } else if (kind == ORC_REF_COSPROD) {
        /* reference/05-mixedPoisson/ReferenceSolution.hpp evaluate and
         * evaluateGradient (with an amplitude) */
        double v = prm[0];
        for (int d = 0; d < dim; d++) v *= cos(prm[1 + d] * x[d]);
        for (int c = 0; c < s; c++) val[c] = v;
        for (int d = 0; d < dim; d++) {
            double g = -prm[0] * prm[1 + d] * sin(prm[1 + d] * x[d]);
            for (int e = 0; e < dim; e++)
                if (e != d) g *= cos(prm[1 + e] * x[e]);
            for (int c = 0; c < s; c++) grad[d * s + c] = g;
        }
    } else if (kind == ORC_REF_KELVIN) {
        orc_fundsol_elasto(dim, prm[0], prm[1], x, prm + 2, prm + 5, val);
    }
}

/* Error norm: base/post/ErrorNorm.hpp:60-84 (errorComputation: element loop,
 * sqrt of the sum) and :200-224 (per point: x(xi), detJ, ref - approx,
 * errorSquared += dot(error, error) * detJ * weight); approximation by
 * base/post/evaluateField.hpp (value: sum_n phi_n u_n, gradient: sum_n
 * grad_x phi_n u_n). */
double orc_error_norm(const orc_problem* p, const double* uh, int order, int kind,
                      const double* prm)
{
    const int dim = p->dim, gnpe = p->npe, s = p->dpn;
    const int npe = field_npe(p), K = p->deg == 2 ? 2 : 1;
    const int32_t* fc = field_conn(p);
    double wq[64], xq[64 * 3];
    const int nq = orc_tensor_rule(dim, p->nq, wq, xq);
    if (nq < 0 || (order == 1 && kind == ORC_REF_KELVIN)) return -1.;
    double err = 0.;
    for (int64_t e = 0; e < p->n_elem; e++) {
        double X[8 * 3], U[ORC_MAXE];
        for (int n = 0; n < gnpe; n++) {
            const int64_t node = p->conn[e * gnpe + n];
            for (int d = 0; d < dim; d++) X[n * dim + d] = p->xyz[node * dim + d];
        }
        for (int n = 0; n < npe; n++) {
            const int64_t node = fc[e * npe + n];
            for (int c = 0; c < s; c++) U[n * s + c] = uh[node * s + c];
        }
        for (int q = 0; q < nq; q++) {
            double gphi[8], gdphi[8 * 3], phi[ORC_MAXF], dphi[ORC_MAXF * 3], gx[ORC_MAXF * 3];
            orc_q1_shape(dim, xq + q * dim, gphi, gdphi);
            double detJ;
            if (K == 1) {
                for (int n = 0; n < gnpe; n++) phi[n] = gphi[n];
                detJ = orc_contravariant(dim, X, gdphi, gx);
            } else {
                orc_lagrange_shape(dim, K, xq + q * dim, phi, dphi);
                detJ = contra_field(dim, X, gdphi, npe, dphi, gx);
            }
            double x[3] = {0, 0, 0};
            for (int n = 0; n < gnpe; n++)
                for (int d = 0; d < dim; d++) x[d] += X[n * dim + d] * gphi[n];
            double rv[3], rg[9];
            ref_solution(dim, s, kind, prm, x, rv, rg);
            double d2 = 0.;
            if (order == 0) {
                for (int c = 0; c < s; c++) {
                    double v = 0.;
                    for (int n = 0; n < npe; n++) v += U[n * s + c] * phi[n];
                    const double er = rv[c] - v;
                    d2 += er * er;
                }
            } else {
                for (int d = 0; d < dim; d++)
                    for (int c = 0; c < s; c++) {
                        double g = 0.;
                        for (int n = 0; n < npe; n++) g += gx[n * dim + d] * U[n * s + c];
                        const double er = rg[d * s + c] - g;
                        d2 += er * er;
                    }
            }
            err += d2 * detJ * wq[q];
        }
    }
    return sqrt(err);
}

/* Cell evaluation at one local point xi of every element.
 * ORC_CELL_VALUE: u_h(xi) (base/post/evaluateField.hpp:166-206), s values;
 * ORC_CELL_GRADIENT: grad u_h, out[d*s+c] = d u_c / d x_d
 *   (base/post/evaluateField.hpp:231-273), dim*s values;
 * ORC_CELL_PK2 / ORC_CELL_CAUCHY: F = I + (grad u)^T extended by the
 *   identity to 3x3 (solid/Deformation.hpp:26-45), S of the material law
 *   (mat/hypel/StVenant.hpp, NeoHookeanCompressible.hpp) and, for the
 *   Cauchy stress, sigma = F S F^T / det F (solid/Stress.hpp:38-58,
 *   secondPKStress :64-81); 9 values per element, row-major 3x3.
 * Returns the values per element, or -1 for an unknown quantity. */
int orc_cell_eval(const orc_problem* p, const double* uh, int quantity, const double* xi,
                  double* out)
{
    const int dim = p->dim, gnpe = p->npe, s = p->dpn;
    const int npe = field_npe(p), K = p->deg == 2 ? 2 : 1;
    const int32_t* fc = field_conn(p);
    int per;
    switch (quantity) {
    case ORC_CELL_VALUE: per = s; break;
    case ORC_CELL_GRADIENT: per = dim * s; break;
    case ORC_CELL_PK2:
    case ORC_CELL_CAUCHY:
        if (p->kernel == ORC_LAPLACE || s != dim) return -1;
        per = 9;
        break;
    default: return -1;
    }
    double gphi[8], gdphi[8 * 3], phi[ORC_MAXF], dphi[ORC_MAXF * 3];
    orc_q1_shape(dim, xi, gphi, gdphi);
    if (K == 2) orc_lagrange_shape(dim, K, xi, phi, dphi);
    for (int64_t e = 0; e < p->n_elem; e++) {
        double X[8 * 3], U[ORC_MAXE], gx[ORC_MAXF * 3];
        for (int n = 0; n < gnpe; n++) {
            const int64_t node = p->conn[e * gnpe + n];
            for (int d = 0; d < dim; d++) X[n * dim + d] = p->xyz[node * dim + d];
        }
        for (int n = 0; n < npe; n++) {
            const int64_t node = fc[e * npe + n];
            for (int c = 0; c < s; c++) U[n * s + c] = uh[node * s + c];
        }
        if (K == 1) {
            for (int n = 0; n < gnpe; n++) phi[n] = gphi[n];
            (void)orc_contravariant(dim, X, gdphi, gx);
        } else {
            (void)contra_field(dim, X, gdphi, npe, dphi, gx);
        }
        double* o = out + e * per;
        if (quantity == ORC_CELL_VALUE) {
            for (int c = 0; c < s; c++) {
                double v = 0.;
                for (int n = 0; n < npe; n++) v += phi[n] * U[n * s + c];
                o[c] = v;
            }
            continue;
        }
        double G[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};  /* G[d][c] = du_c/dx_d */
        for (int d = 0; d < dim; d++)
            for (int c = 0; c < s; c++) {
                double g = 0.;
                for (int n = 0; n < npe; n++) g += gx[n * dim + d] * U[n * s + c];
                G[d][c] = g;
            }
        if (quantity == ORC_CELL_GRADIENT) {
            for (int d = 0; d < dim; d++)
                for (int c = 0; c < s; c++) o[d * s + c] = G[d][c];
            continue;
        }
        double F[3][3], S[3][3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) F[i][j] = (i == j ? 1. : 0.) + G[j][i];
        material(p->kernel, p->coef[e * p->ncoef + 0], p->coef[e * p->ncoef + 1], F, S, NULL);
        if (quantity == ORC_CELL_PK2) {
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) o[i * 3 + j] = S[i][j];
            continue;
        }
        const double J = det3(F);
        double FS[3][3];
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                double t = 0.;
                for (int k = 0; k < 3; k++) t += F[i][k] * S[k][j];
                FS[i][j] = t;
            }
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                double t = 0.;
                for (int k = 0; k < 3; k++) t += FS[i][k] * F[j][k];
                o[i * 3 + j] = t / J;
            }
    }
    return per;
}
