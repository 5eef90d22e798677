/*
 * insilico_hip.h -- C ABI of the MI355X finite-element hot path
 * (element assembly -> CSR -> Jacobi-PCG), the drop-in boundary between the
 * reference's template API and hand-written gfx950 HIP kernels.
 *
 * Conventions
 *  - Every function returns an int status: 0 = INS_OK, < 0 = an INS_ERR_*
 *    code.  Nothing aborts, throws or exits across this boundary (the
 *    reference aborts through VERIFY_MSG, base/verify.hpp:63-65,139-149);
 *    ins_last_error() gives the message of the last failure of a context.
 *  - Host arrays are owned by the caller and copied in or out; device memory
 *    is owned by the context.  No torch or HIP types appear here.
 *  - A context is used by one host thread at a time.  Separate contexts are
 *    independent.  All work of a context runs on its own HIP stream.
 *  - Index types: node ids, element connectivity, DoF ids and CSR arrays are
 *    int32 (the reference's Eigen::SparseMatrix<double> uses int indices,
 *    base/solver/Eigen3.hpp:341).  Sizes are int64.
 *
 * Reference seams replaced (file:line in thrueberg/inSilico):
 *  SOLVER concept      base/solver/Eigen3.hpp:64-342
 *  pattern             base/solver/TripletContainer.hpp:159-301
 *  stiffness assembly  base/asmb/StiffnessMatrix.hpp:52-73,159-225
 *  residual forces     base/asmb/ForceIntegrator.hpp:40-56,126-160
 *  body / Neumann      base/asmb/BodyForce.hpp:67-82, NeumannForce.hpp:36-63
 *  DoF distribution    base/dof/Distribute.hpp:36-80,141-215
 */
#ifndef INSILICO_HIP_H
#define INSILICO_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
#define INS_OK                 0
#define INS_ERR_ARG           -1  /* bad argument (null pointer, bad size)      */
#define INS_ERR_HIP           -2  /* a HIP runtime call failed                  */
#define INS_ERR_OOM           -3  /* device allocation failed / over budget     */
#define INS_ERR_RANGE         -4  /* DoF or node id out of range, ids not a
                                     permutation of 0..n_active-1
                                     (Eigen3.hpp:95-98 "index out of bound")   */
#define INS_ERR_PATTERN       -5  /* entry outside the registered pattern
                                     (TripletContainer.hpp:342-346)             */
#define INS_ERR_BREAKDOWN     -6  /* CG breakdown: p.Ap <= 0 or non-finite      */
#define INS_ERR_NOCONV        -7  /* iteration limit reached                    */
#define INS_ERR_STATE         -8  /* call out of order (e.g. solve before
                                     assembly)                                  */
#define INS_ERR_UNSUPPORTED   -9  /* element/kernel combination not provided    */

/* ---- enumerations ------------------------------------------------------ */
/* element kernels (KERNEL concept implementations) */
#define INS_LAPLACE   1  /* heat::Laplace (heat/Laplace.hpp:113-126)           */
#define INS_STVENANT  2  /* solid::HyperElastic<mat::hypel::StVenant>          */
#define INS_NEOHOOKE  3  /* solid::HyperElastic<mat::hypel::NeoHookeanCompr.>  */

/* arithmetic of element integration */
#define INS_FP64 0
#define INS_FP32 1

/* ins_assemble flags */
#define INS_ASM_MATRIX       1  /* tangent / stiffness values into the CSR    */
#define INS_ASM_RHS          2  /* lifting + body + Neumann (+ residual)      */
#define INS_ASM_RESIDUAL     4  /* add -F_int(u) (ForceIntegrator factor -1)  */
#define INS_ASM_INCREMENTAL  8  /* lifting with g - u (StiffnessMatrix.hpp:56) */

/* body force kinds (f evaluated at x(xi), BodyForce.hpp:172-201) */
#define INS_BF_NONE     0
#define INS_BF_CONST    1  /* params[c] = f_c                                */
#define INS_BF_SINPROD  2  /* f = params[0] * prod_d sin(params[1+d] * x_d)  */
#define INS_BF_COSPROD  3  /* f = params[0] * prod_d cos(params[1+d] * x_d)  */
#define INS_BF_TABLE    4  /* values of a caller's f(x) at the quadrature
                              points (ins_set_body_force_values)            */

/* solve modes */
#define INS_SOLVE_FP64   0  /* FP64 Jacobi-PCG on the exported CSR           */
#define INS_SOLVE_MIXED  1  /* FP32 PCG sweeps, FP64 residual/update/dots    */

/* matrix storage used by the solver */
#define INS_FMT_CSR   0  /* plain CSR, int32 column indices                  */
#define INS_FMT_SELL  1  /* 64-row slices, column-major values, per-row column
                            offset pattern ids (indices not re-read per nnz) */

/* ins_distribute operations (base/dof/Distribute.hpp:36-50) */
#define INS_DIST_SET 0
#define INS_DIST_ADD 1

typedef struct ins_ctx ins_ctx;

typedef struct ins_stats {
    double setup_ms;        /* pattern + maps (last ins_build_pattern)       */
    double assemble_ms;     /* last ins_assemble                             */
    double solve_ms;        /* last ins_solve                                */
    int    iterations;      /* last ins_solve: inner PCG iterations          */
    int    refinements;     /* last ins_solve: FP64 residual refreshes       */
    double relres;          /* last ins_solve: |b - A x| / |b| in FP64       */
    int64_t n_active;
    int64_t nnz;
    int64_t device_bytes;   /* device memory held by the context             */
    double spmv_ms;         /* profiled: total time of the dominant SpMV     */
    int64_t spmv_calls;     /* profiled: launches timed                      */
    int64_t sell_values;    /* padded value count of the SELL storage (0 = none) */
    int     sell_patterns;  /* distinct column-offset lists of the SELL storage  */
    int     sell_sym_offsets; /* > 0: the FP32 inner sweeps use the symmetric
                                 slices (upper offsets shared by all rows)       */
    double  rep_ms;         /* profiled: total time of the residual-replacement
                               launches of the dominant kernel (FP64 residual) */
    int64_t rep_calls;
    double  element_ms;     /* profiled: element kernels of the last
                               ins_assemble (integration only, no gather) */
    int64_t sym_values;     /* values of the symmetric slices (0 = none)     */
    int     sym_block;      /* their block size: 1 = scalar rows, 2 or 3 =
                               node blocks of a vector problem (offsets in
                               sell_sym_offsets are then node offsets)        */
    int     assembled_sym;  /* 1: the last matrix assembly wrote the
                               symmetric slices directly (the CSR values are
                               produced from them on demand); 0: the CSR */
    int     sym_lines;      /* 1: the symmetric rows are stencil lines (a
                               structured grid): the sweeps load each line of
                               the vector once per wavefront               */
} ins_stats;

/* ---- context ------------------------------------------------------------ */
/* No reference counterpart (the reference has no device); hbm_budget_bytes
 * caps the context's device memory (0 = no cap). */
int ins_create(ins_ctx** out, int device, uint64_t hbm_budget_bytes);
int ins_destroy(ins_ctx* ctx);
const char* ins_strerror(int status);
const char* ins_last_error(const ins_ctx* ctx);
int ins_get_stats(const ins_ctx* ctx, ins_stats* out);
/* Free and total device memory of GPU `device` (hipMemGetInfo), for
 * admission by free HBM.  No reference counterpart. */
int ins_device_memory(int device, uint64_t* free_bytes, uint64_t* total_bytes);
/* time every dominant-kernel launch with HIP events (adds small overhead) */
int ins_set_profiling(ins_ctx* ctx, int on);
int ins_sync(ins_ctx* ctx);

/* ---- discretisation (host setup of the reference, flattened) ----------- */
/* Mesh: base::Unstructured nodes and elements (base/mesh).  conn lists the
 * element vertices in the reference's hierarchic order
 * (base/mesh/HierarchicOrder.hpp:201-272,345-392); npe = 4 (QUAD4) or
 * 8 (HEX8). */
int ins_set_mesh(ins_ctx* ctx, int dim, int64_t n_nodes, const double* xyz,
                 int64_t n_elem, int npe, const int32_t* conn);

/* Field of degree 2 on the Q1 mesh just set (base::fe::Basis<shape,2> of a
 * Lagrange field on a geomDeg-1 mesh, e.g. reference/06-elastic/
 * compressible.cpp:108-110): field_conn[e*field_npe + k] = field node (DoF
 * object) k of element e in hierarchic order, field_npe = 9 (QUAD9) or 27
 * (HEX27), numbered as base/dof/IndexMap.hpp:222-280 with
 * base/dof/generateDoFIndicesFromFaces.hpp does (vertices, edges, faces,
 * cells).  Call after ins_set_mesh; the DoFs, prescribed and current values
 * are then indexed by field node, and Neumann faces (ins_set_traction,
 * afterwards) list the 3 (LINE3) / 9 (QUAD9) field nodes of each face in its
 * hierarchic order, vertices first. */
int ins_set_field(ins_ctx* ctx, int degree, int64_t n_field_nodes, int field_npe,
                  const int32_t* field_conn);

/* DoFs: dof_id[node*dpn+c] = active row index (base/dof/numbering.hpp:45-66)
 * or -1 for a CONSTRAINED component (base/dof/DegreeOfFreedom.hpp:232-242).
 * Ids must be a permutation of 0..n_active-1. */
int ins_set_dofs(ins_ctx* ctx, int dofs_per_node, const int32_t* dof_id,
                 int64_t n_active);
/* prescribed values g per (node, comp) of constrained components
 * (Constraint values); NULL = all zero */
int ins_set_prescribed(ins_ctx* ctx, const double* g);
/* base::dof::scaleConstraints (base/dof/Distribute.hpp:76-80) */
int ins_scale_prescribed(ins_ctx* ctx, double factor);
/* current DoF values u per (node, comp) (DegreeOfFreedom values_[0]);
 * NULL = zero */
int ins_set_dof_values(ins_ctx* ctx, const double* u);
int ins_get_dof_values(ins_ctx* ctx, double* u);

/* ---- pattern ------------------------------------------------------------- */
/* Replaces Eigen3::registerFields -> TripletContainer::registerFields
 * (base/solver/Eigen3.hpp:331-335, TripletContainer.hpp:159-301): the CSR
 * pattern of all (active row, active col) pairs of every element, explicit
 * zeros kept, columns ascending.  Also builds the node->element map and the
 * row structure used by the device assembly. */
int ins_build_pattern(ins_ctx* ctx, int64_t* nnz_out);
int ins_get_pattern(ins_ctx* ctx, int32_t* row_ptr, int32_t* col_idx);

/* ---- kernels and loads ----------------------------------------------------- */
/* KERNEL object + quadrature: kernel = INS_LAPLACE (ncoef = 1: kappa_e, the
 * conductivity function of heat/Laplace.hpp:94-97 evaluated per element;
 * ncoef = nq^dim: kappa(element, xi_q) at every quadrature point, in the
 * order of ins_quadrature_rule, as heat/Laplace.hpp:113-126 calls it) or
 * INS_STVENANT / INS_NEOHOOKE (ncoef = 2: lambda_e, mu_e, mat/Lame.hpp:28-40);
 * nq = Gauss points per direction of base::Quadrature
 * (base/quad/GaussLegendre.hpp:33-36: nq = (DEG+2)/2). */
int ins_set_material(ins_ctx* ctx, int kernel, int nq, int ncoef,
                     const double* elem_coef);
/* Time-step terms of a linear multi-step method, as
 * BoundaryValueProblem::applyTimeIntegrator adds them
 * (base/BoundaryValueProblem.hpp:337-354 -> base/time/ReactionTerms.hpp
 * :140-236 with base::kernel::Mass, base/kernel/Mass.hpp:80-110): the
 * element mass matrix M_e = int density phi phi enters the matrix as
 * mass_factor * M_e (mass_factor = a0 / (b0 dt)) with its Dirichlet
 * lifting, and the RHS gets M_e * hist, where hist[node*dpn+c] =
 * sum_{s>=1} -a_s/dt u^{n+1-s} (NULL = zero).  heat::Laplace only;
 * density = 0 switches the terms off. */
int ins_set_mass(ins_ctx* ctx, double density, double mass_factor, const double* hist);

/* heat::DiffusionDriver::advanceInTime (heat/DiffusionDriver.hpp:88-118)
 * with BDF(order) (base/time/BDF.hpp, start-up of base/time/MultiStep.hpp)
 * kept on the device: ins_bdf_init sets the order, the density and the
 * initial DoF values u0 (per field node and component, NULL = 0); every
 * ins_bdf_step(dt) forms the history combination,
 * assembles K + (a0/dt) M and the RHS (ins_set_mass semantics), solves,
 * sets the DoF values (setDoFsFromSolver) and pushes them into the history
 * (pushHistory, base/BoundaryValueProblem.hpp:358-363).  ins_get_dof_values
 * reads u^{n+1}. */
int ins_bdf_init(ins_ctx* ctx, int order, double density, const double* u0);
int ins_bdf_step(ins_ctx* ctx, double dt, int precision, int mode, int format, double rtol,
                 int max_it, int* iters, double* relres);

/* bodyForceComputation (base/asmb/BodyForce.hpp:67-82) with a built-in f */
int ins_set_body_force(ins_ctx* ctx, int kind, const double* params);

/* The quadrature rule of base::Quadrature (base/quad/GaussLegendre.hpp:33-36,
 * 114-134; base/quad/TensorProduct.hpp:123-175): npts (1..3) Gauss points
 * per direction on [0,1]^dim, the first coordinate fastest; w[q] and
 * xi[q*dim + d] for q < npts^dim (dim = 1, 2, 3; dim - 1 gives the surface
 * rule of base::SurfaceQuadrature).  Returns npts^dim, or INS_ERR_ARG.  No
 * context: the points at which ins_set_body_force_values,
 * ins_set_traction_values and a per-point conductivity are given. */
int ins_quadrature_rule(int dim, int npts, double* w, double* xi);

/* bodyForceComputation with a caller's function f(x) (BodyForce.hpp:67-82,
 * 172-201: f is evaluated at x(xi_q), the Q1 geometry at each quadrature
 * point): values f[(e * nqp + q) * dpn + c], nqp = nq^dim of
 * ins_set_material, q in the order of ins_quadrature_rule.  Sets the kind
 * INS_BF_TABLE; NULL drops the values (kind INS_BF_NONE).  Call after
 * ins_set_material and ins_set_dofs. */
int ins_set_body_force_values(ins_ctx* ctx, const double* f);
/* neumannForceComputation (base/asmb/NeumannForce.hpp:36-63) with a
 * traction over boundary faces given by their nodes (2 per edge in 2-D, 4
 * per quadrilateral in 3-D, face-local Q1 order; degree-2 fields: 3 / 9
 * field nodes, vertices first) */
int ins_set_traction(ins_ctx* ctx, int64_t n_faces, int nodes_per_face,
                     const int32_t* faces, const double* traction);
/* A traction that varies along the faces, as a Neumann function of the
 * position does (e.g. reference/06-elastic/PulledSheet.hpp neumannBC,
 * value * (x_1 - 1/2) in component 0): the traction vector is scaled by
 * (x[axis] - x0) at every surface point; axis = -1 = constant (default). */
int ins_set_traction_profile(ins_ctx* ctx, int axis, double x0);
/* Neumann data given as the co-normal derivative of a built-in function u
 * (ref_kind INS_REF_SINPROD / INS_REF_COSPROD with its params, see the
 * post-processing section): t_c = sum_d n_d du_c/dx_d at every surface
 * point, as reference/05-mixedPoisson/GivenData.hpp neumannBC computes it,
 * with n the unit normal of base::SurfaceNormal (base/geometry.hpp:314-345:
 * 2-D (dy, -dx) of the face tangent, 3-D the cross product of the surface
 * Jacobian's columns; outward for faces in the element face-table
 * orientation).  Applies to the faces of ins_set_traction in place of its
 * traction vector; ref_kind INS_REF_ZERO switches back to the traction. */
int ins_set_neumann_flux(ins_ctx* ctx, int ref_kind, const double* ref_params);
/* neumannForceComputation with a caller's function f(x, n)
 * (NeumannForce.hpp:36-63,132-175: f at the surface point x(eta_q) and the
 * unit normal n(eta_q) of base::SurfaceNormal, base/geometry.hpp:314-346):
 * values t[(f * nqf + q) * dpn + c] for the faces of ins_set_traction, nqf =
 * nq^(dim-1) surface points in the order of ins_quadrature_rule(dim - 1,
 * nq).  They replace the traction vector and the co-normal flux; NULL
 * drops them.  ins_set_traction with new faces drops them too. */
int ins_set_traction_values(ins_ctx* ctx, const double* t);

/* ---- assembly ------------------------------------------------------------- */
/* stiffnessMatrixComputation + computeResidualForces + bodyForceComputation
 * + neumannForceComputation into the device CSR values and RHS
 * (base/asmb/StiffnessMatrix.hpp:52-73, ForceIntegrator.hpp:40-56,
 * BodyForce.hpp:67-82, NeumannForce.hpp:36-63), deterministic: each CSR
 * entry and RHS entry is summed over elements in ascending element order.
 * precision = INS_FP64 or INS_FP32 (element integration arithmetic). */
int ins_assemble(ins_ctx* ctx, int precision, int flags);
/* Eigen3::debugLHS/debugRHS equivalents (base/solver/Eigen3.hpp:311-323) */
int ins_get_csr_values(ins_ctx* ctx, double* val);
int ins_get_rhs(ins_ctx* ctx, double* b);
int ins_set_rhs(ins_ctx* ctx, const double* b);
/* insertToRHS of a full-length vector (base/solver/Eigen3.hpp:112-124) */
int ins_add_rhs(ins_ctx* ctx, const double* b);
/* Eigen3::norm(): |b|_2 / N (base/solver/Eigen3.hpp:128-138) */
int ins_rhs_norm(ins_ctx* ctx, double* norm);
/* Upload a caller-assembled CSR (the SOLVER-concept insertToLHS path,
 * base/solver/Eigen3.hpp:82-106 + finishAssembly :142-153). */
int ins_set_csr(ins_ctx* ctx, int64_t n, const int32_t* row_ptr,
                const int32_t* col_idx, const double* val);

/* ---- solve ----------------------------------------------------------------- */
/* Eigen3::cgSolve (base/solver/Eigen3.hpp:265-275): Jacobi-preconditioned
 * CG from x0 = 0 until |r|/|b| <= rtol.  mode INS_SOLVE_FP64 or
 * INS_SOLVE_MIXED; format INS_FMT_CSR or INS_FMT_SELL.  Returns
 * INS_ERR_NOCONV if max_it iterations did not reach rtol (the solution of the
 * last iterate is still available). */
int ins_solve(ins_ctx* ctx, int mode, int format, double rtol, int max_it,
              int* iters, double* relres);
/* Eigen3::choleskySolve / luSolve / superLUSolve (base/solver/Eigen3.hpp:
 * 157-230), which return the factorised solution: the FP64 system as a dense
 * n x n matrix on the device, factorised by rocSOLVER -- kind
 * INS_DIRECT_CHOLESKY: potrf + potrs (the lower triangle; SimplicialLDLT's
 * symmetric positive definite case), INS_DIRECT_LU: getrf + getrs with
 * partial pivoting (SparseLU / SuperLU).  relres = |b - A x| / |b| in FP64,
 * as after ins_solve.  n is at most INS_DIRECT_MAX_N (the factor holds 8 n^2
 * bytes; larger systems: INS_ERR_UNSUPPORTED, use ins_solve).
 * INS_ERR_BREAKDOWN when Cholesky meets a non-positive pivot (the matrix is
 * not positive definite) or LU an exactly zero pivot.  Needs an FP64
 * assembly (or ins_set_csr) and the RHS. */
#define INS_DIRECT_CHOLESKY 0
#define INS_DIRECT_LU       1
#define INS_DIRECT_MAX_N    32768
int ins_solve_direct(ins_ctx* ctx, int kind, double* relres);
/* Build what ins_solve(mode, format) needs (SELL form and its values,
 * solver vectors) without solving, so that a later solve or server attach
 * allocates nothing.  With INS_FMT_SELL the SELL form is kept even where its
 * padding exceeds 1/8 of the CSR values (ins_solve alone would fall back to
 * CSR there).  May be called before the first ins_assemble (after
 * ins_build_pattern): the storage it names, like that of the last
 * ins_solve, is the one the next assembly writes; for a problem with
 * symmetric slices and INS_FMT_SELL that is the slices themselves, and the
 * CSR values are rebuilt from them when an export or a CSR solve asks for
 * them.  No reference counterpart (Eigen builds its solver state inside
 * cgSolve, base/solver/Eigen3.hpp:265-275). */
int ins_prepare(ins_ctx* ctx, int mode, int format);
/* Eigen3::getValue for all rows (base/solver/Eigen3.hpp:293-296) */
int ins_get_solution(ins_ctx* ctx, double* x);
/* |solution|_2 / N as Eigen3::norm() reads it after cgSolve */
int ins_solution_norm(ins_ctx* ctx, double* norm);
/* setDoFsFromSolver / addToDoFsFromSolver (base/dof/Distribute.hpp:36-50,
 * 141-215): active components get (or add) the solution, constrained
 * components get their prescribed value. */
int ins_distribute(ins_ctx* ctx, int op);

/* ---- post-processing -------------------------------------------------------- */
#define INS_REF_ZERO     0  /* u = 0: the norm of u_h itself                   */
#define INS_REF_SINPROD  1  /* u = params[0] * prod_d sin(params[1+d] * x_d)   */
#define INS_REF_KELVIN   2  /* elastostatic fundamental solution
                               (base/auxi/FundamentalSolution.hpp:160-200):
                               params = lambda, mu, source[3], direction[3]   */
#define INS_REF_COSPROD  3  /* u = params[0] * prod_d cos(params[1+d] * x_d)
                               (reference/05-mixedPoisson/
                               ReferenceSolution.hpp)                          */
/* base::post::errorComputation<order> (base/post/ErrorNorm.hpp:60-84,
 * 200-224): |u - u_h| in the L2 norm (order 0) or the H1 semi-norm (order
 * 1; not for INS_REF_KELVIN), integrated with the context's quadrature.
 * u_h: values per field node and component (NULL = the context's current
 * DoF values, as ins_distribute leaves them). */
int ins_error_norm(ins_ctx* ctx, int order, int ref_kind, const double* ref_params,
                   const double* u, double* out);

#define INS_CELL_VALUE     0  /* u_h(xi): dpn values per element              */
#define INS_CELL_GRADIENT  1  /* grad u_h(xi): dim*dpn, [d*dpn+c] = du_c/dx_d  */
#define INS_CELL_PK2       2  /* 2nd Piola-Kirchhoff stress S: 9, row-major   */
#define INS_CELL_CAUCHY    3  /* Cauchy stress F S F^T / det F: 9, row-major  */
/* One local point xi of every element, as the reference's cell data are
 * formed for base::io::vtk::writeCellData (base/io/vtk/LegacyWriter.hpp
 * :314-328): base::post::evaluateField / evaluateFieldGradient
 * (base/post/evaluateField.hpp:166-206, 231-273) and solid::secondPKStress /
 * solid::cauchy (solid/Stress.hpp:38-81; F = I + grad u^T extended by the
 * identity to 3x3, solid/Deformation.hpp:26-45) with the context's material
 * (INS_STVENANT / INS_NEOHOOKE and their per-element lambda, mu; the stresses
 * need dpn = dim).  xi: dim local coordinates in [0,1] (NULL = the centroid,
 * base::ShapeCentroid); u: values per field node and component (NULL = the
 * context's current DoF values); out: host array of n_elem * (values per
 * element), element-major. */
int ins_cell_eval(ins_ctx* ctx, int quantity, const double* xi, const double* u,
                  double* out);

/* ---- allocation-free solve ------------------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory: code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc.  The FP64 Jacobi-PCG of ins_solve(INS_SOLVE_FP64,
 * INS_FMT_CSR) -- same kernels, grids and reduction order, so the same
 * iterates -- on a CSR system held in caller-owned DEVICE arrays of GPU
 * `device`: row_ptr (n+1, row_ptr[n] = nnz), col_idx (nnz, ascending per
 * row), val (nnz), b (n) are read, x (n) receives the solution; stream is a
 * hipStream_t (NULL = the default stream) and the call returns after the
 * stream has finished the solve.  Its scratch (scalars, block partials,
 * Jacobi diagonal, iterates r, p, q) is a static device array of
 * 32 MiB + 80 KiB that each GPU the call runs on holds until the process
 * ends; the call allocates nothing.  Every call on one GPU shares that
 * array: calls that overlap on one GPU race.  n > INS_NOALLOC_MAX_ROWS,
 * null pointers, rtol <= 0 or row_ptr[n] != nnz give INS_ERR_ARG; otherwise
 * the statuses of ins_solve (INS_ERR_NOCONV with x the last iterate,
 * INS_ERR_BREAKDOWN). */
#define INS_NOALLOC_MAX_ROWS (1 << 20)
int ins_solve_csr_noalloc(int device, void* stream, int64_t n, int64_t nnz,
                          const int32_t* row_ptr, const int32_t* col_idx, const double* val,
                          const double* b, double* x, double rtol, int max_it, int* iters,
                          double* relres);

/* ---- serving: many solves batched into shared launches --------------------
 * No single reference counterpart: the reference solves one system per
 * Eigen3::cgSolve call (base/solver/Eigen3.hpp:265-275).  A server runs the
 * mixed-precision Jacobi-PCG of INS_SOLVE_MIXED / INS_FMT_SELL for up to
 * max_slots assembled contexts at once: every iteration is three launches
 * shared by all running requests, each request keeps its matrix, diagonal
 * and iterates resident in its own context, and requests join and leave
 * between chunks (continuous batching).  A request's iterates are the same
 * bits alone, in any mix, and through ins_solve(ctx, INS_SOLVE_MIXED,
 * INS_FMT_SELL, ...).  A server and the contexts attached to it are used by
 * one host thread. */
typedef struct ins_server ins_server;

#define INS_SLOT_FREE      0
#define INS_SLOT_RUNNING   1
#define INS_SLOT_FINISHED  2

int ins_server_create(ins_server** out, int device, int max_slots /* 1..64 */);
int ins_server_destroy(ins_server* srv);
const char* ins_server_last_error(const ins_server* srv);
/* Admit an assembled context (ins_assemble done, its own stream may still be
 * busy: the request starts after that work).  *slot receives its slot. */
int ins_server_attach(ins_server* srv, ins_ctx* ctx, double rtol, int max_it,
                      int* slot);
/* Queue k batched iterations of every running request, then act on the
 * stop / replacement flags of the chunk queued before (one chunk stays in
 * flight).  *n_finished = requests that finished during this call. */
int ins_server_step(ins_server* srv, int k, int* n_finished);
/* state INS_SLOT_*; status INS_OK, INS_ERR_NOCONV or INS_ERR_BREAKDOWN */
int ins_server_poll(ins_server* srv, int slot, int* state, int* status,
                    int* iters, double* relres);
/* Free the slot.  A finished request's context then holds its solution
 * (ins_get_solution); detaching a running request cancels it.  The chunk
 * still queued on the server stream may name the context: its next call
 * (any ins_* on it, ins_destroy included) first waits for the server
 * stream's work queued up to the detach, so it can be reused or destroyed
 * at once. */
int ins_server_detach(ins_server* srv, int slot);
int ins_server_running(const ins_server* srv, int* n_running);

#ifdef __cplusplus
}
#endif
#endif
