/*
 * CPU oracle for the finite-element hot path (assembly -> CSR -> Jacobi-CG).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP product
 * path is compared against.  Only the test-suite, the smoke check and the CPU
 * baseline leg of the benchmark may load it; the product library never links,
 * loads or calls it.
 *
 * Every function restates the reference algorithm of thrueberg/inSilico and
 * cites the reference file:line it follows.  Arithmetic order follows the
 * reference wherever it is visible in the reference sources (quadrature order,
 * product order of tensor-product shape functions, per-quadrature-point
 * accumulation, element order of the global sums).  Eigen-internal arithmetic
 * (3x3 inverse, CG loop) is restated from the published algorithms.
 *
 * Pinning (tests/test_oracle.py): Gauss rules by their literal constants and
 * polynomial exactness; shape functions by partition of unity, the Kronecker
 * property and the SMF vertex order; element matrices by closed forms,
 * rigid-body modes, symmetry/PSD and finite differences of the residual; the
 * pattern by sparsity.1.ref.dat and closed-form nnz; assembly and lifting by
 * brute force and the linearElastic{2D,3D}.ref.dat L2 errors; orc_pcg by its
 * residual and a direct solve.  Degree-2 fields on Q1 geometry
 * (tests/test_oracle_q2.py): orc_lagrange_shape by the Kronecker property at
 * the hierarchic support points, reproduction of x^2 and finite differences;
 * Q2 elements by closed-form energies, rigid-body modes and finite differences
 * of the residual; Q2 assembly, lifting and the varying traction by the
 * Newton histories compRefOut{D,N}.dat (every printed digit above the
 * round-off floor).  Parity unpinned: the exact iterate sequence
 * and iteration count of Eigen 3.2's ConjugateGradient (x0 = 0 per Eigen's
 * documentation), and last-bit rounding of Eigen's 3x3
 * computeInverseAndDetWithCheck; no fixture of the reference fixes them.
 */
#ifndef INSILICO_ORACLE_H
#define INSILICO_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* kernel ids (same numbering as the product C-ABI) */
enum { ORC_LAPLACE = 1, ORC_STVENANT = 2, ORC_NEOHOOKE = 3 };
/* body force kinds */
enum { ORC_BF_NONE = 0, ORC_BF_CONST = 1, ORC_BF_SINPROD = 2, ORC_BF_COSPROD = 3,
       ORC_BF_TABLE = 4 };

typedef struct orc_problem {
    int dim;                 /* 2 or 3 */
    int64_t n_nodes;
    const double* xyz;       /* n_nodes*dim */
    int64_t n_elem;
    int npe;                 /* 4 (QUAD4) or 8 (HEX8), hierarchic vertex order */
    const int32_t* conn;     /* n_elem*npe node ids */
    int dpn;                 /* dofs per node (1 scalar, dim for elasticity) */
    const int32_t* dof_id;   /* n_nodes*dpn: >=0 active row, -1 constrained */
    int64_t n_active;
    const double* g;         /* n_nodes*dpn prescribed values (constrained comps) */
    const double* u;         /* n_nodes*dpn current values, NULL = zero */
    int incremental;         /* lifting uses g-u instead of g */
    int kernel;              /* ORC_LAPLACE / ORC_STVENANT / ORC_NEOHOOKE */
    int ncoef;               /* coefficients per element: 1 (kappa) or 2 (lambda,mu) */
    const double* coef;      /* n_elem*ncoef */
    int nq;                  /* Gauss points per direction */
    int bf_kind;
    double bf[4];            /* CONST: f per component; SINPROD: amp, k_x, k_y, k_z */
    int64_t n_faces;         /* Neumann faces with constant traction */
    int nfn;                 /* nodes per face (2 in 2D, 4 in 3D) */
    const int32_t* faces;    /* n_faces*nfn node ids */
    double traction[3];
    int residual;            /* add -F_int(u) (Newton residual) to the RHS */
    int nthreads;            /* 1 = reference default (serial) */
    /* Fields of degree 2 on the Q1 geometry above (deg = 2).  Then conn/npe
     * and faces/nfn are the geometry, and dof_id, g, u are indexed by field
     * node.  deg 0 or 1: iso-parametric Q1 (field node = geometry node). */
    int deg;
    int fnpe;                /* field nodes per element: 9 (QUAD9) / 27 (HEX27) */
    const int32_t* fconn;    /* n_elem*fnpe field node ids, hierarchic order */
    int ffn;                 /* field nodes per Neumann face: 3 (LINE3) / 9 (QUAD9) */
    const int32_t* ffaces;   /* n_faces*ffn field node ids, hierarchic face order */
    /* traction varying along the face: t_c = traction[c] * (x[tr_axis] - tr_x0)
     * when tr_kind = 1 (as the pulled-sheet Neumann function); 0 = constant */
    int tr_kind;
    int tr_axis;
    double tr_x0;
    /* Time integration terms (base/time/ReactionTerms.hpp:140-236): with
     * density > 0 the element mass matrix M_e (base/kernel/Mass.hpp, rho phi
     * phi detJ w) enters as mass_factor * M_e (= a0/(b0 dt) M_e) into the
     * matrix and its lifting, and M_e * mass_hist (mass_hist per node and
     * component = sum_s -a_s/dt u^{n+1-s}, NULL = 0) into the RHS, in a
     * second element loop after the stiffness and the forces. */
    double density;
    double mass_factor;
    const double* mass_hist;
    /* Neumann data as the co-normal derivative t_c = sum_d n_d du_c/dx_d of
     * a built-in function (ORC_REF_SINPROD / ORC_REF_COSPROD with its
     * params, reference/05-mixedPoisson/GivenData.hpp neumannBC) in place of
     * the traction when flux_kind != 0 */
    int flux_kind;
    double flux_p[8];
    /* Values of caller functions at the quadrature points, as the reference
     * evaluates them there: the body force f(x(xi_q)) per element, point and
     * component (bf_kind ORC_BF_TABLE; base/asmb/BodyForce.hpp:172-201),
     * and the Neumann function f(x(eta_q), n(eta_q)) per face, surface point
     * and component (in place of the traction; base/asmb/NeumannForce.hpp
     * :132-175); NULL = not used.  A Laplace problem with ncoef equal to the
     * number of quadrature points takes the conductivity function
     * kappa(element, xi_q) per point (heat/Laplace.hpp:113-126). */
    const double* bf_tab;
    const double* tr_tab;
} orc_problem;

int  orc_gauss1d(int npts, double* w, double* x);
int  orc_tensor_rule(int dim, int npts, double* w, double* xi);
void orc_q1_shape(int dim, const double* xi, double* phi, double* dphi);
void orc_lagrange_shape(int dim, int deg, const double* xi, double* phi, double* dphi);
double orc_contravariant(int dim, const double* X, const double* dphi,
                         double* gradx);
int  orc_element(const orc_problem* p, int64_t e, double* Ke, double* Fe);
/* element mass matrix rho phi_M phi_N detJ w (nee x nee, row-major) */
int  orc_mass_element(const orc_problem* p, int64_t e, double* Me);
/* BDF weights a_0..a_q with the start-up of base/time/MultiStep.hpp (order
 * min(q, step+1)); returns the number of weights */
int  orc_bdf_weights(int order, int step, double* a);
int  orc_unit_cube(int dim, int e1, int e2, int e3, double* xyz, int32_t* conn);

/* pattern as the reference's TripletContainer::registerFields builds it */
int  orc_pattern(const orc_problem* p, int64_t* nnz_out,
                 int32_t* row_ptr /* n_active+1 or NULL */,
                 int32_t* col_idx /* nnz or NULL */);
/* assembly; row_ptr/col_idx must hold the pattern; t[0]=assembly seconds */
int  orc_assemble(const orc_problem* p, const int32_t* row_ptr,
                  const int32_t* col_idx, double* val, double* b, double* t);
/* Jacobi-preconditioned CG as Eigen::ConjugateGradient with x0=0 */
int  orc_pcg(int64_t n, const int32_t* row_ptr, const int32_t* col_idx,
             const double* val, const double* b, double* x, double tol,
             int maxit, int nthreads, int* iters, double* relres);
void orc_spmv(int64_t n, const int32_t* row_ptr, const int32_t* col_idx,
              const double* val, const double* x, double* y, int nthreads);
/* L2 error of a Q1 field against the elastostatic fundamental solution */
double orc_l2_error_fundsol(const orc_problem* p, const double* uh,
                            double lambda, double mu, const double* src,
                            const double* dir);
void orc_fundsol_elasto(int dim, double lambda, double mu, const double* y,
                        const double* x, const double* dir, double* out);
/* reference solutions of the error norm */
enum { ORC_REF_ZERO = 0, ORC_REF_SINPROD = 1, ORC_REF_KELVIN = 2, ORC_REF_COSPROD = 3 };
/* |u - u_h| in the L2 norm (order 0) or the H1 semi-norm (order 1) with the
 * problem's quadrature; uh per field node and component; prm: SINPROD amp,
 * k_x, k_y, k_z (u = amp prod_d sin(k_d x_d)), COSPROD likewise with cos;
 * KELVIN lambda, mu, source[3],
 * direction[3] (order 0 only) */
double orc_error_norm(const orc_problem* p, const double* uh, int order, int kind,
                      const double* prm);
/* quantities of orc_cell_eval */
enum { ORC_CELL_VALUE = 0, ORC_CELL_GRADIENT = 1, ORC_CELL_PK2 = 2, ORC_CELL_CAUCHY = 3 };
/* one local point xi (dim values) of every element: the field value (dpn
 * per element), its gradient (dim*dpn, [d*dpn+c] = du_c/dx_d), or the
 * second Piola-Kirchhoff / Cauchy stress of the solid kernels (9 per
 * element, row-major 3x3); returns the count per element, -1 if invalid */
int orc_cell_eval(const orc_problem* p, const double* uh, int quantity, const double* xi,
                  double* out);

#ifdef __cplusplus
}
#endif
#endif
