// Element integration and deterministic assembly on gfx950.
//
//  1. element kernels write every element matrix (and force vector) once
//     into a staging buffer:
//       - heat::Laplace: one element per lane, the Q1 local-gradient table of
//         all quadrature points in LDS, 36 (3-D) / 10 (2-D) unique entries of
//         the symmetric element matrix kept in registers, written element-major
//         ([e][entry]) so the gather reads whole sectors;
//       - solid::HyperElastic: one wavefront per element; lanes < nqp build
//         geometry, F, S and the Voigt C of their quadrature point, all 64
//         lanes then form the effective elasticity A_iJkL of every point, and
//         lane (M,N) accumulates the 3x3 block K_MN over the points.
//  2. a gather kernel sums, for every CSR entry, the staged element
//     contributions of the elements shared by the row node and the column
//     node in ascending element order: the summation order of the
//     reference's serial element loop (base/auxi/parallel.hpp:26-60 with
//     base/solver/TripletContainer.hpp:305-352), with no float atomics.
//  3. an RHS kernel does the same per active row for the Dirichlet lifting
//     (base/asmb/assembleMatrix.hpp:91-95), element forces and face tractions.
#include <cmath>
#include <cstring>

#include "common.hpp"

namespace ins {

namespace {

constexpr int BLK = 256;

// body force: a built-in function of x, or (INS_BF_TABLE) the values of a
// caller's function at the quadrature points, tab[(e * nqp + q) * s + c]
struct Bf {
    int kind;
    double p[4];
    const double* tab;
    int nqp, s;
};

template <int B, int E>
struct IPow {
    static constexpr int value = B * IPow<B, E - 1>::value;
};
template <int B>
struct IPow<B, 0> {
    static constexpr int value = 1;
};

// table layout: dphi[nq][npe][dim], phi[nq][npe], w[nq]
template <int DIM, int NQ>
struct Tab {
    static constexpr int NPE = 1 << DIM;
    static constexpr int DPHI = 0;
    static constexpr int PHI = NQ * NPE * DIM;
    static constexpr int W = PHI + NQ * NPE;
    static constexpr int SIZE = W + NQ;
};

template <typename T>
__device__ inline T body_force(const Bf& bf, const T* x, int dim, int comp, int64_t e, int q)
{
    if (bf.kind == INS_BF_TABLE) return (T)bf.tab[(e * bf.nqp + q) * bf.s + comp];
    if (bf.kind == INS_BF_CONST) return (T)bf.p[comp];
    if (bf.kind == INS_BF_SINPROD) {
        T f = (T)bf.p[0];
        for (int d = 0; d < dim; d++) f *= sin((T)bf.p[1 + d] * x[d]);
        return f;
    }
    if (bf.kind == INS_BF_COSPROD) {
        T f = (T)bf.p[0];
        for (int d = 0; d < dim; d++) f *= cos((T)bf.p[1 + d] * x[d]);
        return f;
    }
    return (T)0;
}

// inverse of m (= J^T) by cofactors and det (Eigen computeInverseAndDetWithCheck)
template <int DIM, typename T>
__device__ inline T inverse_det(const T (&m)[DIM][DIM], T (&inv)[DIM][DIM])
{
    if constexpr (DIM == 2) {
        const T det = m[0][0] * m[1][1] - m[1][0] * m[0][1];
        const T id = (T)1 / det;
        inv[0][0] = m[1][1] * id;
        inv[1][0] = -m[1][0] * id;
        inv[0][1] = -m[0][1] * id;
        inv[1][1] = m[0][0] * id;
        return det;
    } else {
        T cof[3][3];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                const int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
                const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
                cof[i][j] = m[i1][j1] * m[i2][j2] - m[i1][j2] * m[i2][j1];
            }
        const T det = cof[0][0] * m[0][0] + cof[1][0] * m[1][0] + cof[2][0] * m[2][0];
        const T id = (T)1 / det;
#pragma unroll
        for (int r = 0; r < 3; r++)
#pragma unroll
            for (int c = 0; c < 3; c++) inv[r][c] = cof[c][r] * id;
        return det;
    }
}

// physical gradients at one quadrature point; returns detJ
// (base/geometry.hpp:143-177,420-445; base/LagrangeShapeFun.hpp:182-200)
template <int DIM, int NPE, typename T>
__device__ inline T phys_grad(const T (&X)[NPE][DIM], const T* dphi, T (&gx)[NPE][DIM])
{
    T J[DIM][DIM];
#pragma unroll
    for (int i = 0; i < DIM; i++)
#pragma unroll
        for (int a = 0; a < DIM; a++) J[i][a] = (T)0;
#pragma unroll
    for (int n = 0; n < NPE; n++)
#pragma unroll
        for (int i = 0; i < DIM; i++)
#pragma unroll
            for (int a = 0; a < DIM; a++) J[i][a] += X[n][i] * dphi[n * DIM + a];
    T m[DIM][DIM], inv[DIM][DIM];
#pragma unroll
    for (int i = 0; i < DIM; i++)
#pragma unroll
        for (int a = 0; a < DIM; a++) m[i][a] = J[a][i];
    const T det = inverse_det<DIM, T>(m, inv);
#pragma unroll
    for (int n = 0; n < NPE; n++)
#pragma unroll
        for (int i = 0; i < DIM; i++) {
            T s = (T)0;
#pragma unroll
            for (int a = 0; a < DIM; a++) s += inv[i][a] * dphi[n * DIM + a];
            gx[n][i] = s;
        }
    return det;
}

// ---------------------------------------------------------------------------
// heat::Laplace element kernel (heat/Laplace.hpp:113-126 ->
// base/kernel/Laplace.hpp:100-151): one element per lane.
template <int DIM, int NQ1, typename T, bool WANT_M = false>
__global__ __launch_bounds__(BLK) void k_elem_laplace(
    int64_t E, const int32_t* __restrict__ conn, const double* __restrict__ xyz,
    const double* __restrict__ coef, int ncoef, const T* __restrict__ tab, Bf bf,
    T* __restrict__ ke, double* __restrict__ fe, T* __restrict__ me = nullptr, T rho = 0)
{
    constexpr int NPE = 1 << DIM;
    constexpr int NQ = IPow<NQ1, DIM>::value;
    constexpr int NK = NPE * (NPE + 1) / 2;
    using TB = Tab<DIM, NQ>;
    __shared__ T s_tab[TB::SIZE];
    for (int i = threadIdx.x; i < TB::SIZE; i += blockDim.x) s_tab[i] = tab[i];
    __syncthreads();
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
         e += (int64_t)gridDim.x * blockDim.x) {
        T X[NPE][DIM];
#pragma unroll
        for (int n = 0; n < NPE; n++) {
            const int64_t node = conn[e * NPE + n];
#pragma unroll
            for (int d = 0; d < DIM; d++) X[n][d] = (T)xyz[node * DIM + d];
        }
        const T kappa = (T)coef[e * ncoef];
        T K[NK], Mm[WANT_M ? NK : 1];
#pragma unroll
        for (int k = 0; k < NK; k++) K[k] = (T)0;
        if constexpr (WANT_M) {
#pragma unroll
            for (int k = 0; k < NK; k++) Mm[k] = (T)0;
        }
        T fb[NPE];
#pragma unroll
        for (int n = 0; n < NPE; n++) fb[n] = (T)0;
        for (int q = 0; q < NQ; q++) {
            const T* dphi = s_tab + TB::DPHI + q * NPE * DIM;
            T gx[NPE][DIM];
            const T det = phys_grad<DIM, NPE, T>(X, dphi, gx);
            const T w = s_tab[TB::W + q];
            // kappa(element, xi): one value per element, or one per point
            const T scal = (ncoef == NQ ? (T)coef[e * ncoef + q] : kappa) * det * w;
            int k = 0;
#pragma unroll
            for (int M = 0; M < NPE; M++)
#pragma unroll
                for (int N = M; N < NPE; N++) {
                    T entry = (T)0;
#pragma unroll
                    for (int d = 0; d < DIM; d++) entry += gx[M][d] * gx[N][d];
                    entry *= scal;
                    K[k++] += entry;
                }
            if constexpr (WANT_M) {
                // base/kernel/Mass.hpp:93-104: scalar = rho detJ w, entry =
                // scalar phi_M phi_N
                const T* phi = s_tab + TB::PHI + q * NPE;
                const T sm_ = rho * det * w;
                int km = 0;
#pragma unroll
                for (int M = 0; M < NPE; M++)
#pragma unroll
                    for (int N = M; N < NPE; N++) Mm[km++] += sm_ * phi[M] * phi[N];
            }
            if (fe && bf.kind != INS_BF_NONE) {
                const T* phi = s_tab + TB::PHI + q * NPE;
                T x[3] = {0, 0, 0};
#pragma unroll
                for (int n = 0; n < NPE; n++)
#pragma unroll
                    for (int d = 0; d < DIM; d++) x[d] += X[n][d] * phi[n];
                const T f = body_force<T>(bf, x, DIM, 0, e, q);
#pragma unroll
                for (int n = 0; n < NPE; n++) fb[n] += f * phi[n] * w * det;
            }
        }
        // 64-element slabs, entry-major inside a slab (entry k of element e at
        // (e/64)*64*NK + k*64 + e%64): a wavefront stores 64 consecutive
        // values per entry, and the gather, whose lanes hold consecutive
        // nodes and so read the same entry of consecutive elements, loads
        // contiguous runs
        const int64_t sb = (e >> 6) * 64 * NK + (e & 63);
#pragma unroll
        for (int k = 0; k < NK; k++) ke[sb + k * 64] = K[k];
        if constexpr (WANT_M) {
#pragma unroll
            for (int k = 0; k < NK; k++) me[sb + k * 64] = Mm[k];
        }
        if (fe) {
            const int64_t fb0 = (e >> 6) * 64 * NPE + (e & 63);
#pragma unroll
            for (int n = 0; n < NPE; n++) fe[fb0 + n * 64] = (double)fb[n];
        }
    }
}

// ---------------------------------------------------------------------------
// Material laws at one quadrature point: S and the Voigt C.
// mat/hypel/StVenant.hpp:77-114, mat/hypel/NeoHookeanCompressible.hpp:77-172,
// mat/TensorAlgebra.hpp:55-60 (det), :141-164 (Voigt).
__device__ inline int voigt_idx(int i, int j)
{
    // map[(i+1)*(j+1)-1] with map = {0,3,4,1,-,5,-,-,2}
    const int d = (i + 1) * (j + 1);
    return d == 1 ? 0 : d == 2 ? 3 : d == 3 ? 4 : d == 4 ? 1 : d == 6 ? 5 : 2;
}

template <typename T>
__device__ inline void inv3(const T m[3][3], T inv[3][3])
{
    T cof[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            const int i1 = (i + 1) % 3, i2 = (i + 2) % 3;
            const int j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            cof[i][j] = m[i1][j1] * m[i2][j2] - m[i1][j2] * m[i2][j1];
        }
    const T det = cof[0][0] * m[0][0] + cof[1][0] * m[1][0] + cof[2][0] * m[2][0];
    const T id = (T)1 / det;
#pragma unroll
    for (int r = 0; r < 3; r++)
#pragma unroll
        for (int c = 0; c < 3; c++) inv[r][c] = cof[c][r] * id;
}

template <int MAT, typename T>
__device__ inline void material(T lam, T mu, const T F[3][3], T S[3][3], T C[6][6],
                                bool wantC)
{
    T CG[3][3];
#pragma unroll
    for (int i = 0; i < 3; i++)
#pragma unroll
        for (int j = 0; j < 3; j++) {
            T s = (T)0;
#pragma unroll
            for (int k = 0; k < 3; k++) s += F[k][i] * F[k][j];
            CG[i][j] = s;
        }
    if constexpr (MAT == INS_STVENANT) {
        T E[3][3];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) E[i][j] = (T)0.5 * (CG[i][j] - (i == j ? (T)1 : (T)0));
        const T trE = E[0][0] + E[1][1] + E[2][2];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++)
                S[i][j] = lam * trE * (i == j ? (T)1 : (T)0) + (T)2 * mu * E[i][j];
        if (wantC) {
#pragma unroll
            for (int i = 0; i < 6; i++)
#pragma unroll
                for (int j = 0; j < 6; j++) C[i][j] = (T)0;
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int j = 0; j < 3; j++) C[i][j] = lam;
#pragma unroll
            for (int i = 0; i < 3; i++) C[i][i] += (T)2 * mu;
#pragma unroll
            for (int i = 3; i < 6; i++) C[i][i] += mu;
        }
    } else {
        T Ci[3][3];
        inv3<T>(CG, Ci);
        const T J = (F[0][0] * F[1][1] * F[2][2] + F[0][1] * F[1][2] * F[2][0] +
                     F[0][2] * F[1][0] * F[2][1] - F[0][0] * F[1][2] * F[2][1] -
                     F[0][1] * F[1][0] * F[2][2] - F[0][2] * F[1][1] * F[2][0]);
        const T logJ = log(J);
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++)
                S[i][j] = (lam * logJ - mu) * Ci[i][j] + mu * (i == j ? (T)1 : (T)0);
        if (wantC) {
            const T fac2 = mu - lam * logJ;
#pragma unroll
            for (int A = 0; A < 3; A++)
#pragma unroll
                for (int B = A; B < 3; B++) {
                    const int AB = voigt_idx(A, B);
#pragma unroll
                    for (int c = 0; c < 3; c++)
#pragma unroll
                        for (int D = c; D < 3; D++)
                            C[AB][voigt_idx(c, D)] =
                                lam * Ci[A][B] * Ci[c][D] +
                                fac2 * (Ci[A][c] * Ci[B][D] + Ci[A][D] * Ci[B][c]);
                }
        }
    }
}

// ---------------------------------------------------------------------------
// Degree-2 Lagrange fields on Q1 geometry (QUAD9 / HEX27 on QUAD4 / HEX8).
// Table layout: geometry dphi [NQ][GN][DIM], geometry phi [NQ][GN], field
// dphi [NQ][NF][DIM], field phi [NQ][NF], weights [NQ].
template <int DIM, int NQ>
struct TabQ2 {
    static constexpr int GN = 1 << DIM;
    static constexpr int NF = IPow<3, DIM>::value;
    static constexpr int GD = 0;
    static constexpr int GP = GD + NQ * GN * DIM;
    static constexpr int FD = GP + NQ * GN;
    static constexpr int FP = FD + NQ * NF * DIM;
    static constexpr int W = FP + NQ * NF;
    static constexpr int SIZE = W + NQ;
};

// J from the Q1 geometry, inv(J^T) and detJ, with the arithmetic of phys_grad
// (base/geometry.hpp:143-177,420-445)
template <int DIM, typename T>
__device__ inline T geo_inverse(const T* X, const T* gd, T (&inv)[DIM][DIM])
{
    constexpr int GN = 1 << DIM;
    T J[DIM][DIM];
#pragma unroll
    for (int i = 0; i < DIM; i++)
#pragma unroll
        for (int a = 0; a < DIM; a++) J[i][a] = (T)0;
#pragma unroll
    for (int n = 0; n < GN; n++)
#pragma unroll
        for (int i = 0; i < DIM; i++)
#pragma unroll
            for (int a = 0; a < DIM; a++) J[i][a] += X[n * DIM + i] * gd[n * DIM + a];
    T m[DIM][DIM];
#pragma unroll
    for (int i = 0; i < DIM; i++)
#pragma unroll
        for (int a = 0; a < DIM; a++) m[i][a] = J[a][i];
    return inverse_det<DIM, T>(m, inv);
}

// 16x16x4 matrix-core products of the element kernels: lane l supplies
// A[l%16][l/16] and B[l/16][l%16]; element r of the result holds column l%16
// of row row(l, r): 4*(l/16) + r for FP32, l/16 + 4*r for FP64 (measured on
// gfx950 with a rank-one product)
template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    static __device__ inline v4 op(double a, double b, v4 c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct Mfma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    static __device__ inline v4 op(float a, float b, v4 c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// heat::Laplace on a degree-2 field as one matrix product per element:
// with G[(q,i)][M] = (grad_x phi_M)_i at quadrature point q and s_q =
// kappa detJ_q w_q, Ke = G^T diag(s) G (the sum over q and i of
// base/kernel/Laplace.hpp:100-151).  One wavefront per element, four per
// block sharing the reference-gradient table in LDS.  Lanes < NQ build
// inv(J^T) and s_q of their point; then for each group of 4 rows of G every
// lane forms its G entries of the 16-wide column tiles (3 multiply-adds on
// the table) and the upper tiles of Ke advance by one 16x16x4 matrix-core
// step each, with independent accumulators.  The symmetric half is written
// element-major straight from the accumulators (16 consecutive entries per
// row segment).
template <int DIM, int NQ1, typename T>
struct LapQ2 {
    static constexpr int NQ = IPow<NQ1, DIM>::value;
    static constexpr int GN = 1 << DIM;
    static constexpr int NF = IPow<3, DIM>::value;
    static constexpr int KD = NQ * DIM;
    static constexpr int KP = (KD + 3) / 4 * 4;
    static constexpr int NT = (NF + 15) / 16;
    static constexpr int NFP = NT * 16;
    static constexpr int NK = NF * (NF + 1) / 2;
    static constexpr int INV = 0;                // [NQ][DIM][DIM]
    static constexpr int SC = INV + NQ * DIM * DIM;
    static constexpr int DET = SC + NQ;
    static constexpr int FQ = DET + NQ;
    static constexpr int X = FQ + NQ;            // [GN][DIM]
    static constexpr int SIZE = X + GN * DIM;
    static constexpr int WPB = 4;                // wavefronts (elements) per block
};

template <int DIM, int NQ1, typename T>
__global__ __launch_bounds__(256) void k_elem_laplace_q2(
    int64_t E, const int32_t* __restrict__ gconn, const double* __restrict__ xyz,
    const double* __restrict__ coef, int ncoef, const T* __restrict__ tab, Bf bf,
    T* __restrict__ ke, double* __restrict__ fe, T* __restrict__ me, T rho)
{
    using L = LapQ2<DIM, NQ1, T>;
    using TB = TabQ2<DIM, L::NQ>;
    constexpr int NQ = L::NQ, GN = L::GN, NF = L::NF, NT = L::NT, WPB = L::WPB;
    __shared__ T s_fd[NQ * NF * DIM];            // reference gradients of the field
    __shared__ T s_all[WPB * L::SIZE];
    for (int i = threadIdx.x; i < NQ * NF * DIM; i += blockDim.x) s_fd[i] = tab[TB::FD + i];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r16 = lane & 15, k4 = lane >> 4;
    T* sm = s_all + wave * L::SIZE;
    __syncthreads();
    for (int64_t base = (int64_t)blockIdx.x * WPB; base < E; base += (int64_t)gridDim.x * WPB) {
        const int64_t e = base + wave;
        const bool valid = e < E;
        if (valid && lane < GN * DIM) {
            const int64_t node = gconn[e * GN + lane / DIM];
            sm[L::X + lane] = (T)xyz[node * DIM + lane % DIM];
        }
        __syncthreads();
        if (valid && lane < NQ) {
            const int q = lane;
            T inv[DIM][DIM];
            const T det = geo_inverse<DIM, T>(sm + L::X, tab + TB::GD + q * GN * DIM, inv);
#pragma unroll
            for (int i = 0; i < DIM; i++)
#pragma unroll
                for (int a = 0; a < DIM; a++) sm[L::INV + (q * DIM + i) * DIM + a] = inv[i][a];
            const T w = tab[TB::W + q];
            sm[L::SC + q] = (T)coef[e * ncoef + (ncoef == NQ ? q : 0)] * det * w;
            sm[L::DET + q] = det;
            if (fe && bf.kind != INS_BF_NONE) {
                T x[3] = {0, 0, 0};
#pragma unroll
                for (int n = 0; n < GN; n++)
#pragma unroll
                    for (int d = 0; d < DIM; d++) x[d] += sm[L::X + n * DIM + d] * tab[TB::GP + q * GN + n];
                sm[L::FQ + q] = body_force<T>(bf, x, DIM, 0, e, q);
            }
        }
        __syncthreads();
        if (valid) {
            typename Mfma<T>::v4 acc[NT * (NT + 1) / 2];
#pragma unroll
            for (int t = 0; t < NT * (NT + 1) / 2; t++) acc[t] = (typename Mfma<T>::v4){0, 0, 0, 0};
#pragma clang loop unroll_count(3)
            for (int s0 = 0; s0 < L::KP; s0 += 4) {
                const int k = s0 + k4;
                const bool kin = k < L::KD;
                const int q = kin ? k / DIM : 0, i = kin ? k - q * DIM : 0;
                const T* iv = sm + L::INV + (q * DIM + i) * DIM;
                const T sq = sm[L::SC + q];
                T g[NT];
#pragma unroll
                for (int t = 0; t < NT; t++) {
                    const int M = t * 16 + r16;
                    T v = (T)0;
                    if (kin && M < NF) {
                        const T* fd = s_fd + (q * NF + M) * DIM;
#pragma unroll
                        for (int a = 0; a < DIM; a++) v += iv[a] * fd[a];
                    }
                    g[t] = v;
                }
                int t = 0;
#pragma unroll
                for (int ti = 0; ti < NT; ti++)
#pragma unroll
                    for (int tj = ti; tj < NT; tj++) {
                        acc[t] = Mfma<T>::op(g[ti], g[tj] * sq, acc[t]);
                        t++;
                    }
            }
            // rows M of tile ti, columns N of tile tj, upper half only
            T* dst = ke + e * (int64_t)L::NK;
            int t = 0;
#pragma unroll
            for (int ti = 0; ti < NT; ti++)
#pragma unroll
                for (int tj = ti; tj < NT; tj++) {
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int M = ti * 16 + Mfma<T>::row(lane, r);
                        const int N = tj * 16 + r16;
                        if (M < NF && N < NF && M <= N)
                            dst[M * NF - (M * (M - 1)) / 2 + (N - M)] = acc[t][r];
                    }
                    t++;
                }
            if (me) {
                // mass matrix Phi^T diag(rho detJ w) Phi over the points
                // (base/kernel/Mass.hpp), same tiles
                typename Mfma<T>::v4 accm[NT * (NT + 1) / 2];
#pragma unroll
                for (int t2 = 0; t2 < NT * (NT + 1) / 2; t2++) accm[t2] = (typename Mfma<T>::v4){0, 0, 0, 0};
                for (int s0 = 0; s0 < NQ; s0 += 4) {
                    const int q = s0 + k4;
                    const bool qin = q < NQ;
                    const T wq = qin ? rho * sm[L::DET + q] * tab[TB::W + q] : (T)0;
                    T ph[NT];
#pragma unroll
                    for (int tt = 0; tt < NT; tt++) {
                        const int M = tt * 16 + r16;
                        ph[tt] = (qin && M < NF) ? tab[TB::FP + q * NF + M] : (T)0;
                    }
                    int t2 = 0;
#pragma unroll
                    for (int ti = 0; ti < NT; ti++)
#pragma unroll
                        for (int tj = ti; tj < NT; tj++) {
                            accm[t2] = Mfma<T>::op(ph[ti], ph[tj] * wq, accm[t2]);
                            t2++;
                        }
                }
                T* dm = me + e * (int64_t)L::NK;
                int t2 = 0;
#pragma unroll
                for (int ti = 0; ti < NT; ti++)
#pragma unroll
                    for (int tj = ti; tj < NT; tj++) {
#pragma unroll
                        for (int r = 0; r < 4; r++) {
                            const int M = ti * 16 + Mfma<T>::row(lane, r);
                            const int N = tj * 16 + r16;
                            if (M < NF && N < NF && M <= N)
                                dm[M * NF - (M * (M - 1)) / 2 + (N - M)] = accm[t2][r];
                        }
                        t2++;
                    }
            }
            if (fe && lane < NF) {
                T fb = (T)0;
                if (bf.kind != INS_BF_NONE)
#pragma unroll 3
                    for (int q = 0; q < NQ; q++)
                        fb += sm[L::FQ + q] * tab[TB::FP + q * NF + lane] * tab[TB::W + q] *
                              sm[L::DET + q];
                fe[e * NF + lane] = (double)fb;
            }
        }
        __syncthreads();
    }
}

// solid::HyperElastic element matrices on the matrix cores, for Q1 fields and
// for degree-2 fields on the Q1 geometry (solid/HyperElastic.hpp:110-179
// tangent, :210-257 residual, :282-310 effective elasticity).  With
// beta_{M,i}[a] = sum over the index pairs (A, J) of Voigt row a of
// F_iA dphi_{M,J} (the Voigt strain of a unit displacement of DoF (M, i)),
//   K[(M,i)][(N,k)] = sum_q dw_q ( sum_ab beta_{M,i}[a] C_q[a][b] beta_{N,k}[b]
//                                  + delta_ik dphi_M . S_q dphi_N ),
// which is the reference's sum_JL dphi_{M,J} A_iJkL dphi_{N,L}, A_iJkL =
// delta_ik S_JL + sum_AB F_iA C_{V(A,J)V(B,L)} F_kB, with the sums over
// (A, J) and (B, L) grouped by Voigt rows.  Both terms are matrix products
// over (point, row): Bt (dw C B) with NV rows per point and dphi^T (dw S
// dphi) with DIM rows per point (padded to 4).
//
// Two kernels.  k_hyper_points (one thread per element and point) evaluates
// the geometry, F, the material law and the residual integrand of every
// point into a point record; k_elem_hyper (one wavefront per element) loads
// its element's records into LDS, and for each chunk of QC points writes
// beta and dw C beta of every DoF (and dphi, dw S dphi of every field
// function) to LDS and advances the upper tiles of both products by 16x16x4
// matrix-core steps; the element matrix is written as its packed upper
// triangle, element-major.  Splitting the point work from the matrix work
// keeps every lane of the point kernel busy (a wavefront per element would
// leave 56 of 64 lanes idle during it) at the price of one write and one
// read of the records.
template <int DIM>
struct VoigtRows;
template <>
struct VoigtRows<3> {
    static constexpr int NV = 6;
    // index pair (J, K) of row a, and its row of the 6x6 Voigt C
    static __device__ inline int j(int a) { return a < 3 ? a : (a == 5 ? 1 : 0); }
    static __device__ inline int k(int a) { return a < 3 ? a : (a == 3 ? 1 : 2); }
    static __device__ inline int c(int a) { return a; }
};
template <>
struct VoigtRows<2> {
    static constexpr int NV = 3;   // xx, yy, xy: the z rows of a plane problem vanish
    static __device__ inline int j(int a) { return a < 2 ? a : 0; }
    static __device__ inline int k(int a) { return a < 2 ? a : 1; }
    static __device__ inline int c(int a) { return a < 2 ? a : 3; }
};

// upper-triangle index of the symmetric 6x6 Voigt matrix (a <= b)
__device__ inline int c21(int a, int b) { return a * 6 - (a * (a - 1)) / 2 + (b - a); }

// Point record (elements of T): inv(J^T) [9], detJ w, F [9], S [6, Voigt
// order], C [21, upper triangle], detJ, P = F S [9], body force [3]; stored
// field-major over all points (field f of point p at f * n_points + p), so
// that the point kernel's stores and the element kernel's loads of the
// consecutive points of an element are contiguous runs
struct PtRec {
    static constexpr int INV = 0, DW = 9, F = 10, S = 19, C = 25, DET = 46, P = 47, FQ = 56;
    static constexpr int K = 46;      // what the matrix needs
    static constexpr int SIZE = 60;   // with the force integrands (16-byte multiple for FP64)
    // At zero displacement F = I, S = 0 and P = 0 exactly, and the tangent of
    // both laws is the isotropic lambda I(x)I + 2 mu I, whose Voigt form holds
    // three distinct values (lambda + 2 mu, lambda, mu at C(0,0), C(0,1),
    // C(3,3)).  Such records carry only these fields (LIN_K of them for the
    // matrix, LIN_ALL with the body-force integrand).
    static constexpr int LIN_K = 13, LIN_ALL = 17;
    static __host__ __device__ constexpr int lin_field(int j)
    {
        return j < 10 ? j : j == 10 ? C + 0 : j == 11 ? C + 1 : j == 12 ? C + 15 : j == 13 ? DET
                                                                                       : FQ + (j - 14);
    }
};

// The Voigt tangent C of the material law at F, upper triangle (the values
// of material(..., wantC = true))
template <int MAT, typename T>
__device__ inline void store_voigt_c21(T lam, T mu, const T F[3][3], T* __restrict__ C)
{
    if constexpr (MAT == INS_STVENANT) {
#pragma unroll
        for (int i = 0; i < 6; i++)
#pragma unroll
            for (int j = i; j < 6; j++)
                C[c21(i, j)] = (i < 3 && j < 3 ? lam : (T)0) + (i == j ? (i < 3 ? (T)2 * mu : mu) : (T)0);
    } else {
        T CG[3][3];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) {
                T s = (T)0;
#pragma unroll
                for (int k = 0; k < 3; k++) s += F[k][i] * F[k][j];
                CG[i][j] = s;
            }
        T Ci[3][3];
        inv3<T>(CG, Ci);
        const T J = (F[0][0] * F[1][1] * F[2][2] + F[0][1] * F[1][2] * F[2][0] +
                     F[0][2] * F[1][0] * F[2][1] - F[0][0] * F[1][2] * F[2][1] -
                     F[0][1] * F[1][0] * F[2][2] - F[0][2] * F[1][1] * F[2][0]);
        const T fac2 = mu - lam * log(J);
#pragma unroll
        for (int A = 0; A < 3; A++)
#pragma unroll
            for (int B = A; B < 3; B++) {
                const int AB = voigt_idx(A, B);
#pragma unroll
                for (int c = 0; c < 3; c++)
#pragma unroll
                    for (int D = c; D < 3; D++) {
                        const int cD = voigt_idx(c, D);
                        if (AB <= cD)
                            C[c21(AB, cD)] = lam * Ci[A][B] * Ci[c][D] +
                                             fac2 * (Ci[A][c] * Ci[B][D] + Ci[A][D] * Ci[B][c]);
                    }
            }
    }
}

// tables of the element kernels: geometry dphi [NQ][GN][DIM] and phi
// [NQ][GN], field dphi [NQ][NF][DIM] and phi [NQ][NF], weights [NQ] (for Q1
// fields the field tables are the geometry ones)
template <typename T>
struct ElTab {
    const T *gd, *gp, *fd, *fp, *w;
};

template <int DIM, int NF, int NQ1, typename T, int MAT, bool LIN>
__global__ __launch_bounds__(256) void k_hyper_points(
    int64_t E, const int32_t* __restrict__ gconn, const int32_t* __restrict__ fconn,
    const double* __restrict__ xyz, const double* __restrict__ coef,
    const double* __restrict__ u, int ncoef, ElTab<T> tb, Bf bf, int want_k, int want_f,
    T* __restrict__ pts)
{
    constexpr int GN = 1 << DIM, NQ = IPow<NQ1, DIM>::value, S = DIM;
    for (int64_t pt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; pt < E * NQ;
         pt += (int64_t)gridDim.x * blockDim.x) {
        const int64_t e = pt / NQ;
        const int q = (int)(pt - e * NQ);
        T X[GN * DIM];
#pragma unroll
        for (int n = 0; n < GN; n++) {
            const int64_t node = gconn[e * GN + n];
#pragma unroll
            for (int d = 0; d < DIM; d++) X[n * DIM + d] = (T)xyz[node * DIM + d];
        }
        T inv[DIM][DIM];
        const T det = geo_inverse<DIM, T>(X, tb.gd + q * GN * DIM, inv);
        const T dw = det * tb.w[q];
        const int64_t np = E * NQ;
        T* rec = pts + pt;   // field f at rec[f * np]
        if constexpr (LIN) {
            // zero displacement: the geometry, the three values of the
            // isotropic tangent (as store_voigt_c21 rounds them) and the
            // body-force integrand
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int a = 0; a < 3; a++) rec[(int64_t)(PtRec::INV + i * 3 + a) * np] = (i < DIM && a < DIM) ? inv[i][a] : (T)0;
            rec[(int64_t)(PtRec::DW) * np] = dw;
            const T lam = (T)coef[e * ncoef], mu = (T)coef[e * ncoef + 1];
            if (want_k) {
                rec[(int64_t)(PtRec::C + 0) * np] = lam + (T)2 * mu;
                rec[(int64_t)(PtRec::C + 1) * np] = lam + (T)0;
                rec[(int64_t)(PtRec::C + 15) * np] = (T)0 + mu;
            }
            if (want_f && bf.kind != INS_BF_NONE) {
                rec[(int64_t)(PtRec::DET) * np] = det;
                T x[3] = {0, 0, 0};
#pragma unroll
                for (int n = 0; n < GN; n++)
#pragma unroll
                    for (int d = 0; d < DIM; d++) x[d] += X[n * DIM + d] * tb.gp[q * GN + n];
#pragma unroll
                for (int c = 0; c < 3; c++)
                    rec[(int64_t)(PtRec::FQ + c) * np] = c < S ? body_force<T>(bf, x, DIM, c, e, q) : (T)0;
            }
            continue;
        }
        // F = I + (sum_f dphi_f u_f^T)^T (solid/Deformation.hpp:26-46)
        T GU[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
        if (u) {
#pragma unroll 1
            for (int f = 0; f < NF; f++) {
                const T* fd = tb.fd + (q * NF + f) * DIM;
                const int64_t node = fconn[e * NF + f];
                T uf[S];
#pragma unroll
                for (int i = 0; i < S; i++) uf[i] = (T)u[node * S + i];
#pragma unroll
                for (int J = 0; J < DIM; J++) {
                    T g = (T)0;
#pragma unroll
                    for (int a = 0; a < DIM; a++) g += inv[J][a] * fd[a];
#pragma unroll
                    for (int i = 0; i < S; i++) GU[J][i] += g * uf[i];
                }
            }
        }
        T Fm[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
#pragma unroll
        for (int J = 0; J < DIM; J++)
#pragma unroll
            for (int i = 0; i < S; i++) Fm[i][J] += GU[J][i];
        const T lam = (T)coef[e * ncoef], mu = (T)coef[e * ncoef + 1];
        T Sm[3][3], Cm[6][6];
        material<MAT, T>(lam, mu, Fm, Sm, Cm, false);
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int a = 0; a < 3; a++) rec[(int64_t)(PtRec::INV + i * 3 + a) * np] = (i < DIM && a < DIM) ? inv[i][a] : (T)0;
        rec[(int64_t)(PtRec::DW) * np] = dw;
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = 0; j < 3; j++) rec[(int64_t)(PtRec::F + i * 3 + j) * np] = Fm[i][j];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int j = i; j < 3; j++) rec[(int64_t)(PtRec::S + voigt_idx(i, j)) * np] = Sm[i][j];
        if (want_k) {
            T c[21];
            store_voigt_c21<MAT, T>(lam, mu, Fm, c);
#pragma unroll
            for (int k = 0; k < 21; k++) rec[(int64_t)(PtRec::C + k) * np] = c[k];
        }
        if (want_f) {
            rec[(int64_t)(PtRec::DET) * np] = det;
#pragma unroll
            for (int i = 0; i < 3; i++)
#pragma unroll
                for (int J = 0; J < 3; J++) {
                    T t = (T)0;
#pragma unroll
                    for (int k = 0; k < 3; k++) t += Fm[i][k] * Sm[k][J];
                    rec[(int64_t)(PtRec::P + i * 3 + J) * np] = t;
                }
            T x[3] = {0, 0, 0};
            if (bf.kind != INS_BF_NONE) {
#pragma unroll
                for (int n = 0; n < GN; n++)
#pragma unroll
                    for (int d = 0; d < DIM; d++) x[d] += X[n * DIM + d] * tb.gp[q * GN + n];
            }
#pragma unroll
            for (int c = 0; c < 3; c++)
                rec[(int64_t)(PtRec::FQ + c) * np] = (bf.kind != INS_BF_NONE && c < S) ? body_force<T>(bf, x, DIM, c, e, q) : (T)0;
        }
    }
}

// Ordering of LDS accesses among the lanes of one wavefront (a block of 64
// threads): a wavefront's LDS operations are performed in issue order, so a
// compiler barrier for memory operations is all that is needed (wavefront-
// scope fences emit no waits).  Unlike __syncthreads() it does not wait for
// the wavefront's outstanding global loads and stores.
__device__ inline void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int DIM, int NF, int NQ1, int QC>
struct Hyp {
    static constexpr int NQ = IPow<NQ1, DIM>::value;
    static constexpr int S = DIM;
    static constexpr int NEE = NF * S;
    static constexpr int NV = VoigtRows<DIM>::NV;
    static constexpr int NT = (NEE + 15) / 16;   // 16-wide tiles of the element matrix
    static constexpr int NTG = (NF + 15) / 16;   // ... of the geometric term
    static constexpr int KC = QC * NV;           // Voigt rows of a chunk
    static constexpr int KCP = (KC + 3) / 4 * 4;
    static constexpr int KG = QC * 4;            // gradient rows of a chunk (DIM used)
    static constexpr int NK = NEE * (NEE + 1) / 2;
    // LDS layout (elements of T)
    static constexpr int R = 0;                  // [NQ][PtRec::SIZE] the element's point records
    static constexpr int BV = R + NQ * PtRec::SIZE;   // [KCP][NEE] beta
    static constexpr int CB = BV + KCP * NEE;    // [KCP][NEE] dw C beta
    static constexpr int GG = CB + KCP * NEE;    // [KG][NF] dphi
    static constexpr int SG = GG + KG * NF;      // [KG][NF] dw S dphi
    static constexpr int KGO = SG + KG * NF;     // [NF][NF] geometric term
    // the reference gradients of the field [NQ][NF][DIM] in LDS when small
    static constexpr bool FDL = NQ * NF * DIM <= 512;
    static constexpr int FD = KGO + NF * NF;
    static constexpr int SIZE = FD + (FDL ? NQ * NF * DIM : 0);
    // the packed upper triangle is assembled in the product buffers when it
    // fits there and then written out contiguously; otherwise every lane
    // writes its tile entries straight to memory
    static constexpr bool PACK = NK <= 2 * KCP * NEE;
    static constexpr int PK = BV;
    // upper tiles of the element matrix, taken in passes of at most TPP
    // (the accumulators of one pass stay in registers; every pass sweeps the
    // points again)
    static constexpr int NTT = NT * (NT + 1) / 2;
    static constexpr int TPP = 12;
    static constexpr int NPASS = (NTT + TPP - 1) / TPP;
    static constexpr int TP = NTT < TPP ? NTT : TPP;
};

template <int DIM, int NF, int NQ1, int QC, typename T, bool WANT_K, bool LIN>
__global__ __launch_bounds__(64) void k_elem_hyper(int64_t E, const T* __restrict__ pts,
                                                   ElTab<T> tb, int want_f, int want_resid,
                                                   int has_bf, T* __restrict__ ke,
                                                   double* __restrict__ fe)
{
    using L = Hyp<DIM, NF, NQ1, QC>;
    using V = VoigtRows<DIM>;
    using MF = Mfma<T>;
    constexpr int NQ = L::NQ, S = L::S, NEE = L::NEE, NV = L::NV;
    constexpr int NT = L::NT, NTG = L::NTG, RS = PtRec::SIZE;
    __shared__ T sm[L::SIZE];
    const int lane = threadIdx.x;
    const int r16 = lane & 15, k4 = lane >> 4;
    // rows of the chunk buffers that no point fills stay zero
    for (int i = L::KC * NEE + lane; i < L::KCP * NEE; i += 64) {
        sm[L::BV + i] = (T)0;
        sm[L::CB + i] = (T)0;
    }
    for (int i = lane; i < L::KG * NF; i += 64) {
        sm[L::GG + i] = (T)0;
        sm[L::SG + i] = (T)0;
    }
    if (L::FDL)
        for (int i = lane; i < NQ * NF * DIM; i += 64) sm[L::FD + i] = tb.fd[i];
    const T* __restrict__ fdt = L::FDL ? sm + L::FD : tb.fd;
    // the point records of an element (field f of its point q at f * np +
    // e * NQ + q) go to LDS as [q][f]; those of the next element are loaded
    // into registers while this one is processed (when they take at most 8
    // values per lane).  Only the fields in use are read: the first K (the
    // matrix) or all of them (with the force integrands); at zero
    // displacement the LIN_K or LIN_ALL fields of PtRec::lin_field.
    // (The zero-displacement records take the direct loads: with the
    // register prefetch their elements after the first of a wave came out
    // as zero matrices, a fault whose cause was not found.  At the sizes
    // that matter every wave takes one element, so the prefetch does not
    // pay there anyway.)
    constexpr int NFR = LIN ? PtRec::LIN_ALL : RS;   // fields a record may carry
    constexpr int NR = NQ * NFR;
    constexpr int RPL = (NR + 63) / 64;
    constexpr bool PRE = RPL <= 8 && !LIN;
    const int nfl = LIN ? (want_f && has_bf ? PtRec::LIN_ALL : PtRec::LIN_K)
                        : (want_f ? RS : PtRec::K);
    const int nrl = NQ * nfl;
    const int64_t np = E * NQ;
    // value r of this lane: point q of field f, at f * np + q (+ e * NQ) in
    // memory and at q * RS + f in LDS (-1: the lane has no value r)
    int64_t goff[PRE ? RPL : 1];
    int loff[PRE ? RPL : 1];
    if constexpr (PRE) {
#pragma unroll
        for (int r = 0; r < RPL; r++) {
            const int i = lane + 64 * r, j = i / NQ, q = i - j * NQ;
            const int f = LIN ? PtRec::lin_field(j < NFR ? j : 0) : j;
            goff[r] = i < nrl ? (int64_t)f * np + q : 0;
            loff[r] = i < nrl ? L::R + q * RS + f : -1;
        }
    }
    T pre[PRE ? RPL : 1];
    auto fetch = [&](int64_t ee) {
        if constexpr (PRE) {
#pragma unroll
            for (int r = 0; r < RPL; r++) {
                pre[r] = (ee < E && loff[r] >= 0) ? pts[goff[r] + ee * NQ] : (T)0;
            }
        }
    };
    // every XCD (blocks are placed round-robin) takes a contiguous eighth of
    // the elements, so that the 64 elements of one slab of the output are
    // written through one L2
    const int xcd = blockIdx.x % 8, nloc = gridDim.x / 8;
    const int64_t per = ((E + 7) / 8 + 63) / 64 * 64;
    const int64_t lo = xcd * per, hi = lo + per < E ? lo + per : E;
    fetch(lo + blockIdx.x / 8 < hi ? lo + blockIdx.x / 8 : E);
    for (int64_t e = lo + blockIdx.x / 8; e < hi; e += nloc) {
        if constexpr (PRE) {
#pragma unroll
            for (int r = 0; r < RPL; r++)
                if (loff[r] >= 0) sm[loff[r]] = pre[r];
            fetch(e + nloc < hi ? e + nloc : E);
        } else {
            for (int i = lane; i < nrl; i += 64) {
                const int j = i / NQ, q = i - j * NQ;
                const int f = LIN ? PtRec::lin_field(j) : j;
                sm[L::R + q * RS + f] = pts[(int64_t)f * np + e * NQ + q];
            }
        }
        wave_lds_sync();
        if (WANT_K) {
            constexpr int TP = L::TP;
#pragma unroll
            for (int pass = 0; pass < L::NPASS; pass++) {
                typename MF::v4 acc[TP], accg[NTG * (NTG + 1) / 2];
#pragma unroll
                for (int t = 0; t < TP; t++) acc[t] = (typename MF::v4){0, 0, 0, 0};
#pragma unroll
                for (int t = 0; t < NTG * (NTG + 1) / 2; t++) accg[t] = (typename MF::v4){0, 0, 0, 0};
#pragma unroll 1
                for (int q0 = 0; q0 < NQ; q0 += QC) {
                    // beta and dw C beta of every DoF, dphi and dw S dphi of
                    // every field function, at the chunk's points
                    for (int it = lane; it < QC * NF; it += 64) {
                        const int qc = it / NF, M = it - qc * NF, q = q0 + qc;
                        const bool qin = q < NQ;
                        const T* rc = sm + L::R + (qin ? q : 0) * RS;
                        T g[3] = {0, 0, 0};
                        if (qin) {
                            const T* fd = fdt + (q * NF + M) * DIM;
#pragma unroll
                            for (int J = 0; J < DIM; J++) {
                                T s = (T)0;
#pragma unroll
                                for (int a = 0; a < DIM; a++) s += rc[PtRec::INV + J * 3 + a] * fd[a];
                                g[J] = s;
                            }
                        }
                        const T dw = qin ? rc[PtRec::DW] : (T)0;
                        if constexpr (LIN) {
                            // F = I: beta_a of component i is g[K] if i = J,
                            // g[J] if i = K, else 0; C is isotropic, so the
                            // products with its zero entries are left out
                            // (the sums keep the order and values of the
                            // general form below)
                            const T cd = rc[PtRec::C + 0], cl = rc[PtRec::C + 1],
                                    cm = rc[PtRec::C + 15];
#pragma unroll
                            for (int i = 0; i < S; i++) {
                                T be[NV];
#pragma unroll
                                for (int a = 0; a < NV; a++) {
                                    const int J = V::j(a), K = V::k(a);
                                    be[a] = i == J ? g[K] : (J != K && i == K) ? g[J] : (T)0;
                                }
#pragma unroll
                                for (int a = 0; a < NV; a++) {
                                    T cb;
                                    if (V::c(a) < 3) {
                                        cb = (a == 0 ? cd : cl) * be[0];
                                        cb += (a == 1 ? cd : cl) * be[1];
                                        if (DIM == 3) cb += (a == 2 ? cd : cl) * be[2];
                                    } else {
                                        cb = cm * be[a];
                                    }
                                    sm[L::BV + (qc * NV + a) * NEE + M * S + i] = be[a];
                                    sm[L::CB + (qc * NV + a) * NEE + M * S + i] = dw * cb;
                                }
                            }
                        } else {
#pragma unroll 1
                        for (int i = 0; i < S; i++) {
                            T be[NV];
#pragma unroll
                            for (int a = 0; a < NV; a++) {
                                const int J = V::j(a), K = V::k(a);
                                T b = rc[PtRec::F + i * 3 + J] * g[K];
                                if (J != K) b += rc[PtRec::F + i * 3 + K] * g[J];
                                be[a] = b;
                            }
#pragma unroll
                            for (int a = 0; a < NV; a++) {
                                T cb = (T)0;
#pragma unroll
                                for (int b = 0; b < NV; b++) {
                                    const int ca = V::c(a), cbb = V::c(b);
                                    cb += rc[PtRec::C + (ca <= cbb ? c21(ca, cbb) : c21(cbb, ca))] * be[b];
                                }
                                sm[L::BV + (qc * NV + a) * NEE + M * S + i] = be[a];
                                sm[L::CB + (qc * NV + a) * NEE + M * S + i] = dw * cb;
                            }
                        }
                        }
                        if (!LIN && pass == 0) {
#pragma unroll
                            for (int J = 0; J < DIM; J++) {
                                T sg = (T)0;
#pragma unroll
                                for (int Lx = 0; Lx < DIM; Lx++) sg += rc[PtRec::S + voigt_idx(J, Lx)] * g[Lx];
                                sm[L::GG + (qc * 4 + J) * NF + M] = g[J];
                                sm[L::SG + (qc * 4 + J) * NF + M] = dw * sg;
                            }
                        }
                    }
                    wave_lds_sync();
                    // this pass's upper tiles of Bt (dw C B) over the chunk's
                    // Voigt rows
#pragma unroll 2
                    for (int s0 = 0; s0 < L::KCP; s0 += 4) {
                        const T* br = sm + L::BV + (s0 + k4) * NEE;
                        const T* cr = sm + L::CB + (s0 + k4) * NEE;
                        T av[NT], bv[NT];
#pragma unroll
                        for (int t = 0; t < NT; t++) {
                            const int p = t * 16 + r16;
                            av[t] = p < NEE ? br[p] : (T)0;
                            bv[t] = p < NEE ? cr[p] : (T)0;
                        }
                        int t = 0;
#pragma unroll
                        for (int ti = 0; ti < NT; ti++)
#pragma unroll
                            for (int tj = ti; tj < NT; tj++) {
                                if (t >= pass * TP && t < pass * TP + TP)
                                    acc[t - pass * TP] = MF::op(av[ti], bv[tj], acc[t - pass * TP]);
                                t++;
                            }
                    }
                    // ... and of dphi^T (dw S dphi), one step per point
                    // (S = 0 at zero displacement)
                    if (!LIN && pass == 0) {
#pragma unroll 1
                        for (int s0 = 0; s0 < L::KG; s0 += 4) {
                            const T* gr = sm + L::GG + (s0 + k4) * NF;
                            const T* sr = sm + L::SG + (s0 + k4) * NF;
                            T av[NTG], bv[NTG];
#pragma unroll
                            for (int t = 0; t < NTG; t++) {
                                const int M = t * 16 + r16;
                                av[t] = M < NF ? gr[M] : (T)0;
                                bv[t] = M < NF ? sr[M] : (T)0;
                            }
                            int t = 0;
#pragma unroll
                            for (int ti = 0; ti < NTG; ti++)
#pragma unroll
                                for (int tj = ti; tj < NTG; tj++) {
                                    accg[t] = MF::op(av[ti], bv[tj], accg[t]);
                                    t++;
                                }
                        }
                    }
                    wave_lds_sync();
                }
                if (!LIN && pass == 0) {
                    // the geometric term to LDS
                    int t = 0;
#pragma unroll
                    for (int ti = 0; ti < NTG; ti++)
#pragma unroll
                        for (int tj = ti; tj < NTG; tj++) {
#pragma unroll
                            for (int r = 0; r < 4; r++) {
                                const int M = ti * 16 + MF::row(lane, r), N = tj * 16 + r16;
                                if (M < NF && N < NF) sm[L::KGO + M * NF + N] = accg[t][r];
                            }
                            t++;
                        }
                    wave_lds_sync();
                }
                // the packed upper triangle: material part plus the geometric
                // part on the diagonal of every node block; in memory entry k
                // of element e is at (e/64)*64*NK + k*64 + e%64
                constexpr int OS = L::PACK ? 1 : 64;
                T* dst = L::PACK ? sm + L::PK : ke + (e >> 6) * (64 * (int64_t)L::NK) + (e & 63);
                int t = 0;
#pragma unroll
                for (int ti = 0; ti < NT; ti++)
#pragma unroll
                    for (int tj = ti; tj < NT; tj++) {
                        if (t >= pass * TP && t < pass * TP + TP) {
#pragma unroll
                            for (int r = 0; r < 4; r++) {
                                const int p = ti * 16 + MF::row(lane, r), qq = tj * 16 + r16;
                                if (p < NEE && qq < NEE && p <= qq) {
                                    const int M = p / S, i = p - M * S, N = qq / S, k = qq - N * S;
                                    T v = acc[t - pass * TP][r];
                                    if (!LIN && i == k) v += sm[L::KGO + M * NF + N];
                                    dst[(p * NEE - (p * (p - 1)) / 2 + (qq - p)) * OS] = v;
                                }
                            }
                        }
                        t++;
                    }
            }
            if (L::PACK) {
                wave_lds_sync();
                T* dst = ke + (e >> 6) * (64 * (int64_t)L::NK) + (e & 63);
                for (int k = lane; k < L::NK; k += 64) dst[k * 64] = sm[L::PK + k];
            }
        }
        // element force vector, one DoF per lane, each sum over the points in
        // order: body force f phi w detJ, internal force (P grad phi) detJ w
        if (want_f) {
            for (int j = lane; j < NEE; j += 64) {
                const int M = j / S, i = j - M * S;
                T sb = (T)0, sr = (T)0;
                if (has_bf)
#pragma unroll 1
                    for (int q = 0; q < NQ; q++) {
                        const T* rc = sm + L::R + q * RS;
                        sb += rc[PtRec::FQ + i] * tb.fp[q * NF + M] * tb.w[q] * rc[PtRec::DET];
                    }
                if (want_resid && !LIN)   // P = 0 at zero displacement
#pragma unroll 1
                    for (int q = 0; q < NQ; q++) {
                        const T* rc = sm + L::R + q * RS;
                        const T* fd = tb.fd + (q * NF + M) * DIM;
                        T t = (T)0;
#pragma unroll
                        for (int J = 0; J < DIM; J++) {
                            T g = (T)0;
#pragma unroll
                            for (int a = 0; a < DIM; a++) g += rc[PtRec::INV + J * 3 + a] * fd[a];
                            t += rc[PtRec::P + i * 3 + J] * g;
                        }
                        t *= rc[PtRec::DW];
                        sr += t;
                    }
                fe[e * NEE + j] = (double)sb - (double)sr;
            }
        }
        wave_lds_sync();
    }
}

// ---------------------------------------------------------------------------
// staged element-matrix accessors
// heat::Laplace: the upper triangle of the symmetric element matrix; Q1
// elements in 64-element slabs (entry k of element e at (e/64)*64*NK + k*64
// + e%64), degree-2 elements element-major
template <typename T, int NPE>
struct KeLaplace {
    const T* p;
    __device__ inline double get(int64_t e, int a, int cr, int b, int cc) const
    {
        if (cr != cc) return 0.;
        const int i = a < b ? a : b, j = a < b ? b : a;
        const int k = i * NPE - (i * (i - 1)) / 2 + (j - i);
        constexpr int NK = NPE * (NPE + 1) / 2;
        if constexpr (NPE <= 8) return (double)p[(e >> 6) * (64 * NK) + k * 64 + (e & 63)];
        else return (double)p[e * NK + k];
    }
};

// solid::HyperElastic: the packed upper triangle of the symmetric element
// matrix over DoFs (node, component) in 64-element slabs (entry k of element
// e at (e/64)*64*NK + k*64 + e%64)
template <typename T, int NPE, int S>
struct KeElastic {
    const T* p;
    __device__ inline double get(int64_t e, int a, int cr, int b, int cc) const
    {
        constexpr int NEE = NPE * S, NK = NEE * (NEE + 1) / 2;
        const int r = a * S + cr, c = b * S + cc;
        const int i = r < c ? r : c, j = r < c ? c : r;
        const int k = i * NEE - (i * (i - 1)) / 2 + (j - i);
        return (double)p[(e >> 6) * (64 * (int64_t)NK) + k * 64 + (e & 63)];
    }
};

// v + f * m with the product rounded on its own (never fused into a
// multiply-add), so that every gather kernel adds the scaled mass terms with
// the same bits
template <typename TA>
__device__ inline TA add_scaled(TA v, double f, double m)
{
#pragma clang fp contract(off)
    return v + (TA)(f * m);
}

// Gather: G lanes per node, lane j handles neighbour j of the node and the
// S x S block of CSR entries coupling the two nodes.
template <int NPE, int S, class KE, typename TA, class ME>
__global__ __launch_bounds__(BLK) void k_gather(
    int64_t n_nodes, int lgG, int lg_npe, const int32_t* __restrict__ adj_ptr,
    const int32_t* __restrict__ adj, const int32_t* __restrict__ nb_ptr,
    const int32_t* __restrict__ nb, const int32_t* __restrict__ nboff,
    const int32_t* __restrict__ dof_id, const int32_t* __restrict__ row_ptr,
    const int32_t* __restrict__ col, const int32_t* __restrict__ conn,
    const typename PairWord<NPE>::type* __restrict__ pinfo, int monotone, KE ke, ME me,
    double mfac, double* __restrict__ v64, float* __restrict__ v32, int* __restrict__ err)
{
    constexpr int BITS = PairWord<NPE>::bits, FLAG = PairWord<NPE>::flag;
    const int G = 1 << lgG;
    const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = gt >> lgG;
    const int j0 = (int)(gt & (G - 1));
    if (n >= n_nodes) return;
    bool any = false;
#pragma unroll
    for (int c = 0; c < S; c++) any |= dof_id[n * S + c] >= 0;
    if (!any) return;
    const int32_t nbs = nb_ptr[n], nbe = nb_ptr[n + 1];
    const int32_t as = adj_ptr[n], ae = adj_ptr[n + 1];
    const int mask = (1 << lg_npe) - 1;
    for (int j = j0; j < nbe - nbs; j += G) {
        const int64_t m = nb[nbs + j];
        TA v[S][S];
#pragma unroll
        for (int a = 0; a < S; a++)
#pragma unroll
            for (int b = 0; b < S; b++) v[a][b] = (TA)0;
        // the elements shared by n and m, in ascending element order: walk
        // n's (sorted) elements and find m among each element's vertices;
        // the loads are independent of each other and the same for all the
        // lanes of node n.  A second pass adds the scaled mass matrices of
        // a time step after all stiffness contributions, as the reference's
        // separate element loop does (base/time/ReactionTerms.hpp:174-186).
        const int npass = me.p ? 2 : 1;
        for (int pass = 0; pass < npass; pass++)
        if (pinfo) {
            // precomputed at pattern set-up: one word per pair
            const auto w = pinfo[nbs + j];
            for (int32_t i = 0; i < ae - as; i++) {
                const int nib = (int)((w >> (BITS * i)) & (2 * FLAG - 1));
                if (!nib) continue;
                const int32_t xa = adj[as + i];
                const int64_t e = xa >> lg_npe;
                const int la = xa & mask, lb = nib & (FLAG - 1);
#pragma unroll
                for (int cr = 0; cr < S; cr++)
#pragma unroll
                    for (int cc = 0; cc < S; cc++)
                        v[cr][cc] = pass == 0 ? v[cr][cc] + (TA)ke.get(e, la, cr, lb, cc)
                                              : add_scaled<TA>(v[cr][cc], mfac,
                                                               me.get(e, la, cr, lb, cc));
            }
        } else {
            for (int32_t ia = as; ia < ae; ia++) {
                const int32_t xa = adj[ia];
                const int64_t e = xa >> lg_npe;
                const int la = xa & mask;
                const int32_t* __restrict__ ce = conn + e * NPE;
                int lb = -1;
#pragma unroll
                for (int t = 0; t < NPE; t++)
                    if (ce[t] == (int32_t)m) lb = t;
                if (lb < 0) continue;
#pragma unroll
                for (int cr = 0; cr < S; cr++)
#pragma unroll
                    for (int cc = 0; cc < S; cc++)
                        v[cr][cc] = pass == 0 ? v[cr][cc] + (TA)ke.get(e, la, cr, lb, cc)
                                              : add_scaled<TA>(v[cr][cc], mfac,
                                                               me.get(e, la, cr, lb, cc));
            }
        }
        const int32_t off = nboff[nbs + j];
#pragma unroll
        for (int cr = 0; cr < S; cr++) {
            const int32_t r = dof_id[n * S + cr];
            if (r < 0) continue;
            const int32_t rb = row_ptr[r], re = row_ptr[r + 1];
            int k = 0;
#pragma unroll
            for (int cc = 0; cc < S; cc++) {
                const int32_t cid = dof_id[m * S + cc];
                if (cid < 0) continue;
                int32_t pos;
                if (monotone) {
                    pos = rb + off + k;
                } else {
                    int32_t lo = rb, hi = re;
                    while (lo < hi) {
                        const int32_t mid = (lo + hi) >> 1;
                        if (col[mid] < cid) lo = mid + 1; else hi = mid;
                    }
                    pos = lo;
                    if (pos >= re || col[pos] != cid) {
                        atomicOr(err, 1);
                        continue;
                    }
                }
                k++;
                if (v64) v64[pos] = (double)v[cr][cc];
                if (v32) v32[pos] = (float)v[cr][cc];
            }
        }
    }
}

// RHS per node: lifting (per element: t = -sum_c g_c K_rc, then b += t),
// element forces, face forces; each in ascending element / face order.
// emask (when set) names the prescribed components of every element, so the
// lifting visits only those; the sums are the same as over all components.
template <int NPE, int S, class KE, class ME>
__global__ __launch_bounds__(BLK) void k_rhs(
    int64_t n_nodes, int lg_npe, const int32_t* __restrict__ adj_ptr,
    const int32_t* __restrict__ adj, const int32_t* __restrict__ conn,
    const int32_t* __restrict__ dof_id, const double* __restrict__ g,
    const double* __restrict__ u, int incremental, int lift, KE ke,
    const double* __restrict__ fe, int lg_nfn, const int32_t* __restrict__ fadj_ptr,
    const int32_t* __restrict__ fadj, const double* __restrict__ ff, int nfn, ME me,
    double mfac, const double* __restrict__ hist, const uint32_t* __restrict__ emask,
    int fe_slab, double* __restrict__ b)
{
    for (int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; n < n_nodes;
         n += (int64_t)gridDim.x * blockDim.x) {
        const int32_t as = adj_ptr[n], ae = adj_ptr[n + 1];
        const int mask = (1 << lg_npe) - 1;
        for (int cr = 0; cr < S; cr++) {
            const int32_t r = dof_id[n * S + cr];
            if (r < 0) continue;
            double acc = 0.;
            if (lift) {
                for (int32_t k = as; k < ae; k++) {
                    const int64_t e = adj[k] >> lg_npe;
                    const int a = adj[k] & mask;
                    double t = 0.;
                    if (emask) {
                        // the set bits ascend as (bnode, cc) do below
                        for (uint32_t m = emask[e]; m; m &= m - 1) {
                            const int bit = __builtin_ctz(m);
                            const int bnode = bit / S, cc = bit - bnode * S;
                            const int64_t idx = (int64_t)conn[e * NPE + bnode] * S + cc;
                            const double gv = incremental ? g[idx] - u[idx] : g[idx];
                            t -= gv * ke.get(e, a, cr, bnode, cc);
                        }
                    } else {
                        for (int bnode = 0; bnode < NPE; bnode++) {
                            const int64_t mnode = conn[e * NPE + bnode];
#pragma unroll
                            for (int cc = 0; cc < S; cc++) {
                                const int64_t idx = mnode * S + cc;
                                if (dof_id[idx] >= 0) continue;
                                const double gv = incremental ? g[idx] - u[idx] : g[idx];
                                t -= gv * ke.get(e, a, cr, bnode, cc);
                            }
                        }
                    }
                    acc += t;
                }
            }
            if (fe) {
                for (int32_t k = as; k < ae; k++) {
                    const int64_t e = adj[k] >> lg_npe;
                    const int a = adj[k] & mask;
                    // element-major, or in 64-element slabs (heat::Laplace Q1)
                    acc += fe_slab ? fe[(e >> 6) * (64 * NPE * S) + (a * S + cr) * 64 + (e & 63)]
                                   : fe[e * (NPE * S) + a * S + cr];
                }
            }
            if (ff) {
                const int fmask = (1 << lg_nfn) - 1;
                for (int32_t k = fadj_ptr[n]; k < fadj_ptr[n + 1]; k++) {
                    const int64_t f = fadj[k] >> lg_nfn;
                    const int a = fadj[k] & fmask;
                    acc += ff[f * (nfn * S) + a * S + cr];
                }
            }
            // time-step terms (base/time/ReactionTerms.hpp:174-230): per
            // element the lifting of mfac * M_e with the prescribed values,
            // then M_e times the history combination
            if (me.p) {
                for (int32_t k = as; k < ae; k++) {
                    const int64_t e = adj[k] >> lg_npe;
                    const int a = adj[k] & mask;
                    double t = 0., f = 0.;
                    for (int bnode = 0; bnode < NPE; bnode++) {
                        const int64_t mnode = conn[e * NPE + bnode];
#pragma unroll
                        for (int cc = 0; cc < S; cc++) {
                            const int64_t idx = mnode * S + cc;
                            const double mv = me.get(e, a, cr, bnode, cc);
                            if (lift && dof_id[idx] < 0) t -= g[idx] * (mfac * mv);
                            if (hist) f += mv * hist[idx];
                        }
                    }
                    acc += t;
                    acc += f;
                }
            }
            b[r] = acc;
        }
    }
}

// prescribed components of every element (bit b * S + c: component c of local
// node b), for the RHS lifting
template <int NPE, int S>
__global__ void k_emask(int64_t E, const int32_t* __restrict__ conn,
                        const int32_t* __restrict__ dof_id, uint32_t* __restrict__ emask)
{
    static_assert(NPE * S <= 32, "one 32-bit word per element");
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
         e += (int64_t)gridDim.x * blockDim.x) {
        uint32_t m = 0;
#pragma unroll
        for (int bnode = 0; bnode < NPE; bnode++) {
            const int64_t node = conn[e * NPE + bnode];
#pragma unroll
            for (int cc = 0; cc < S; cc++)
                if (dof_id[node * S + cc] < 0) m |= 1u << (bnode * S + cc);
        }
        emask[e] = m;
    }
}

// Assembly straight into the symmetric slices: one lane per active node a
// (a wavefront holds one 64-node slice), for every stored offset u_k the
// D x D block (a, a + u_k) summed over the elements the slot's word names,
// in node a's adjacency order (ascending element id, as k_gather sums the
// same entries, so the values are the bits of the CSR upper entries);
// stored k-major and lane-contiguous (a wavefront writes 64 consecutive
// values per slot) in FP64 and FP32.  The diagonal slot gives the Jacobi
// diagonal (1/a_ii, or 1 where a_ii = 0) of rows a * D + i.  A second pass
// adds the scaled mass matrices of a time step after all stiffness
// contributions, as k_gather does.
template <int NPE, int D, class KE, typename TA, class ME>
__global__ __launch_bounds__(BLK) void k_sym_gather(
    int64_t N, int nu, const int32_t* __restrict__ anode, const int32_t* __restrict__ adj_ptr,
    const int32_t* __restrict__ adj, int lg_npe, const uint32_t* __restrict__ spw, KE ke, ME me,
    double mfac, double* __restrict__ vs64, float* __restrict__ vs32, double* __restrict__ d64,
    float* __restrict__ d32)
{
    constexpr int BITS = PairWord<NPE>::bits, FLAG = PairWord<NPE>::flag;
    static_assert(NPE <= 8, "32-bit words, at most 8 elements per node");
    const SymLayout L{D, nu};
    for (int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; a < N;
         a += (int64_t)gridDim.x * blockDim.x) {
    const int32_t A = anode[a];
    const int32_t as = adj_ptr[A];
    const int cnt = adj_ptr[A + 1] - as;
    const int mask = (1 << lg_npe) - 1;
    int32_t xa[8];
#pragma unroll
    for (int i = 0; i < 8; i++) xa[i] = i < cnt ? adj[as + i] : 0;
    const int64_t base = L.pos(a, 0);   // slot s at base + L.pos(0, s)
    const uint32_t* __restrict__ wp = spw + (a >> 6) * 64 * (int64_t)nu + (a & 63);
    const int npass = me.p ? 2 : 1;
    for (int k = 0; k < nu; k++) {
        const uint32_t w = wp[(int64_t)k * 64];
        TA v[D][D];
#pragma unroll
        for (int i = 0; i < D; i++)
#pragma unroll
            for (int j = 0; j < D; j++) v[i][j] = (TA)0;
        for (int pass = 0; pass < npass; pass++)
#pragma unroll
            for (int i = 0; i < 8; i++) {
                const int nib = (int)((w >> (BITS * i)) & (2 * FLAG - 1));
                if (i >= cnt || !nib) continue;
                const int64_t e = xa[i] >> lg_npe;
                const int la = xa[i] & mask, lb = nib & (FLAG - 1);
#pragma unroll
                for (int cr = 0; cr < D; cr++)
#pragma unroll
                    for (int cc = 0; cc < D; cc++)
                        v[cr][cc] = pass == 0 ? v[cr][cc] + (TA)ke.get(e, la, cr, lb, cc)
                                              : add_scaled<TA>(v[cr][cc], mfac,
                                                               me.get(e, la, cr, lb, cc));
            }
        if (k == 0) {
#pragma unroll
            for (int i = 0; i < D; i++) {
#pragma unroll
                for (int j = i; j < D; j++) {
                    const int64_t o = base + L.pos(0, L.slot(0, i, j));
                    if (vs64) vs64[o] = (double)v[i][j];
                    vs32[o] = (float)v[i][j];
                }
                const double d = vs64 ? (double)v[i][i] : (double)(float)v[i][i];
                const double id = d != 0. ? 1. / d : 1.;
                d64[a * D + i] = id;
                d32[a * D + i] = (float)id;
            }
        } else {
#pragma unroll
            for (int i = 0; i < D; i++)
#pragma unroll
                for (int j = 0; j < D; j++) {
                    const int64_t o = base + L.pos(0, L.slot(k, i, j));
                    if (vs64) vs64[o] = (double)v[i][j];
                    vs32[o] = (float)v[i][j];
                }
        }
    }
    }
}

// ---------------------------------------------------------------------------
// Built-in functions of the position: reference solutions of the error
// norms and the Neumann data given as their co-normal derivative.
// v[c] = u_c(x), g[d*s+c] = du_c/dx_d.
struct RefFn {
    int kind;
    double p[8];
};

template <int DIM>
__device__ inline void ref_value(const RefFn& r, const double* x, int s, double* v, double* g)
{
    for (int c = 0; c < 3; c++) v[c] = 0.;
    for (int i = 0; i < 9; i++) g[i] = 0.;
    if (r.kind == INS_REF_SINPROD) {
        double val = r.p[0];
#pragma unroll
        for (int d = 0; d < DIM; d++) val *= sin(r.p[1 + d] * x[d]);
        for (int c = 0; c < s; c++) v[c] = val;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            double gd = r.p[0] * r.p[1 + d] * cos(r.p[1 + d] * x[d]);
#pragma unroll
            for (int e = 0; e < DIM; e++)
                if (e != d) gd *= sin(r.p[1 + e] * x[e]);
            for (int c = 0; c < s; c++) g[d * s + c] = gd;
        }
    } else if (r.kind == INS_REF_COSPROD) {
        double val = r.p[0];
#pragma unroll
        for (int d = 0; d < DIM; d++) val *= cos(r.p[1 + d] * x[d]);
        for (int c = 0; c < s; c++) v[c] = val;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            double gd = -r.p[0] * r.p[1 + d] * sin(r.p[1 + d] * x[d]);
#pragma unroll
            for (int e = 0; e < DIM; e++)
                if (e != d) gd *= cos(r.p[1 + e] * x[e]);
            for (int c = 0; c < s; c++) g[d * s + c] = gd;
        }
    } else if (r.kind == INS_REF_KELVIN) {
        // base/auxi/FundamentalSolution.hpp:160-200
        const double lam = r.p[0], mu = r.p[1];
        const double nu = lam / 2. / (lam + mu);
        const double surf = 2. * (DIM - 1.) * M_PI;
        const double fac1 = 1. / (4. * surf * mu * (1. - nu));
        const double fac2 = 3. - 4. * nu;
        double R[3], dist = 0.;
#pragma unroll
        for (int d = 0; d < DIM; d++) {
            R[d] = x[d] - r.p[2 + d];
            dist += R[d] * R[d];
        }
        dist = sqrt(dist);
        const double lk = (DIM == 2) ? log(1. / dist) : 1. / dist;
        const double pw = (DIM == 2) ? dist * dist : dist * dist * dist;
#pragma unroll
        for (int i = 0; i < DIM; i++) {
            double sum = 0.;
#pragma unroll
            for (int j = 0; j < DIM; j++) {
                const double Uij = fac1 * ((i == j ? fac2 * lk : 0.) + (R[i] * R[j]) / pw);
                sum += Uij * r.p[5 + j];
            }
            v[i] = sum;
        }
    }
}


// Neumann data on faces (base/asmb/NeumannForce.hpp:132-175 with the
// surface normal and metric of base/geometry.hpp:314-346): one face per
// lane, S field components.  A face lists its FFN field nodes in its own
// hierarchic order, the first GNF of them its vertices (gmap maps them to
// geometry nodes; null = the field nodes are the geometry nodes).  The data
// t_c is the traction (constant, or with axis >= 0 scaled by
// (x[axis] - x0) at the surface point), or with flux.kind != 0 the
// co-normal derivative sum_d n_d du_c/dx_d of a built-in function.
template <int DIM, int FFN, int S>
__global__ __launch_bounds__(BLK) void k_face(int64_t nf, const int32_t* __restrict__ faces,
                                              const int32_t* __restrict__ gmap,
                                              const double* __restrict__ xyz,
                                              const double* __restrict__ gdt,
                                              const double* __restrict__ gpt,
                                              const double* __restrict__ fpt,
                                              const double* __restrict__ wt, int nqf,
                                              double t0, double t1, double t2, int axis,
                                              double x0, RefFn flux,
                                              const double* __restrict__ ttab,
                                              double* __restrict__ ff)
{
    constexpr int GNF = (DIM == 3) ? 4 : 2;
    constexpr int LD = DIM - 1;
    const double tr[3] = {t0, t1, t2};
    for (int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; f < nf;
         f += (int64_t)gridDim.x * blockDim.x) {
        double X[GNF][DIM];
        for (int n = 0; n < GNF; n++) {
            const int64_t fn = faces[f * FFN + n];
            const int64_t node = gmap ? gmap[fn] : fn;
            for (int d = 0; d < DIM; d++) X[n][d] = xyz[node * DIM + d];
        }
        double out[FFN * S];
        for (int i = 0; i < FFN * S; i++) out[i] = 0.;
        for (int q = 0; q < nqf; q++) {
            const double* dphi = gdt + q * GNF * LD;
            const double* gphi = gpt + q * GNF;
            const double* phi = fpt + q * FFN;
            const double w = wt[q];
            double detG, nrm[3];
            if constexpr (DIM == 3) {
                double J[3][2] = {{0, 0}, {0, 0}, {0, 0}};
                for (int n = 0; n < 4; n++)
                    for (int i = 0; i < 3; i++)
                        for (int a = 0; a < 2; a++) J[i][a] += X[n][i] * dphi[n * 2 + a];
                nrm[0] = J[1][0] * J[2][1] - J[2][0] * J[1][1];
                nrm[1] = J[2][0] * J[0][1] - J[0][0] * J[2][1];
                nrm[2] = J[0][0] * J[1][1] - J[1][0] * J[0][1];
                detG = sqrt(nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2]);
            } else {
                const double tx = X[1][0] - X[0][0], ty = X[1][1] - X[0][1];
                nrm[0] = ty;
                nrm[1] = -tx;
                nrm[2] = 0.;
                detG = sqrt(tx * tx + ty * ty);
            }
            double t[3] = {tr[0], tr[1], tr[2]};
            if (ttab) {
                // a caller's Neumann function, evaluated at this point
                for (int c = 0; c < S; c++) t[c] = ttab[(f * nqf + q) * S + c];
            } else if (flux.kind != 0) {
                double xs[3] = {0., 0., 0.};
                for (int n = 0; n < GNF; n++)
                    for (int d = 0; d < DIM; d++) xs[d] += X[n][d] * gphi[n];
                double rv[3], rg[9];
                ref_value<DIM>(flux, xs, S, rv, rg);
                for (int c = 0; c < S; c++) {
                    double v = 0.;
                    for (int d = 0; d < DIM; d++) v += rg[d * S + c] * (nrm[d] / detG);
                    t[c] = v;
                }
            } else if (axis >= 0) {
                double xs = 0.;
                for (int n = 0; n < GNF; n++) xs += X[n][axis] * gphi[n];
                for (int c = 0; c < S; c++) t[c] = tr[c] * (xs - x0);
            }
            for (int n = 0; n < FFN; n++)
                for (int c = 0; c < S; c++) out[n * S + c] += t[c] * phi[n] * w * detG;
        }
        for (int i = 0; i < FFN * S; i++) ff[f * FFN * S + i] = out[i];
    }
}

// ---------------------------------------------------------------------------
// Error norms (base/post/ErrorNorm.hpp:60-84,200-224): one element per lane;
// per quadrature point x(xi) and detJ from the Q1 geometry, u_h or grad u_h
// from the field functions, errorSquared += (ref - approx).(ref - approx)
// detJ w.  Lanes sum their elements in order, blocks reduce in a fixed tree,
// the host adds the block partials in block order.
template <int DIM, int NF, int S>
__global__ __launch_bounds__(256) void k_error_norm(
    int64_t E, const int32_t* __restrict__ gconn, const int32_t* __restrict__ fconn,
    const double* __restrict__ xyz, const double* __restrict__ u,
    const double* __restrict__ gd, const double* __restrict__ gp,
    const double* __restrict__ fd, const double* __restrict__ fp,
    const double* __restrict__ wq, int nq, int order, RefFn ref, double* __restrict__ part)
{
    constexpr int GN = 1 << DIM;
    double acc = 0.;
    for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < E;
         e += (int64_t)gridDim.x * blockDim.x) {
        double X[GN * DIM];
#pragma unroll
        for (int n = 0; n < GN; n++) {
            const int64_t node = gconn[e * GN + n];
#pragma unroll
            for (int d = 0; d < DIM; d++) X[n * DIM + d] = xyz[node * DIM + d];
        }
        double err = 0.;
        for (int q = 0; q < nq; q++) {
            double inv[DIM][DIM];
            const double det = geo_inverse<DIM, double>(X, gd + q * GN * DIM, inv);
            double x[3] = {0, 0, 0};
#pragma unroll
            for (int n = 0; n < GN; n++)
#pragma unroll
                for (int d = 0; d < DIM; d++) x[d] += X[n * DIM + d] * gp[q * GN + n];
            double rv[3], rg[9];
            ref_value<DIM>(ref, x, S, rv, rg);
            double d2 = 0.;
            if (order == 0) {
#pragma unroll
                for (int c = 0; c < S; c++) {
                    double v = 0.;
                    for (int n = 0; n < NF; n++)
                        v += u[(int64_t)fconn[e * NF + n] * S + c] * fp[q * NF + n];
                    const double er = rv[c] - v;
                    d2 += er * er;
                }
            } else {
                double g[DIM][S];
#pragma unroll
                for (int d = 0; d < DIM; d++)
#pragma unroll
                    for (int c = 0; c < S; c++) g[d][c] = 0.;
                for (int n = 0; n < NF; n++) {
                    const double* f = fd + (q * NF + n) * DIM;
                    double gx[DIM];
#pragma unroll
                    for (int i = 0; i < DIM; i++) {
                        double t = 0.;
#pragma unroll
                        for (int a = 0; a < DIM; a++) t += inv[i][a] * f[a];
                        gx[i] = t;
                    }
                    const int64_t node = fconn[e * NF + n];
#pragma unroll
                    for (int d = 0; d < DIM; d++)
#pragma unroll
                        for (int c = 0; c < S; c++) g[d][c] += gx[d] * u[node * S + c];
                }
#pragma unroll
                for (int d = 0; d < DIM; d++)
#pragma unroll
                    for (int c = 0; c < S; c++) {
                        const double er = rg[d * S + c] - g[d][c];
                        d2 += er * er;
                    }
            }
            err += d2 * det * wq[q];
        }
        acc += err;
    }
    __shared__ double red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Jacobi: inverse diagonal of each row (Eigen DiagonalPreconditioner:
// 1/a_ii, or 1 where a_ii = 0)
__global__ void k_dinv(int64_t n, const int32_t* __restrict__ row_ptr,
                       const int32_t* __restrict__ col, const double* __restrict__ v64,
                       const float* __restrict__ v32, double* __restrict__ d64,
                       float* __restrict__ d32)
{
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
         r += (int64_t)gridDim.x * blockDim.x) {
        int32_t lo = row_ptr[r], hi = row_ptr[r + 1];
        const int32_t end = hi;
        while (lo < hi) {
            const int32_t mid = (lo + hi) >> 1;
            if (col[mid] < r) lo = mid + 1; else hi = mid;
        }
        double d = 0.;
        if (lo < end && col[lo] == r) d = v64 ? v64[lo] : (double)v32[lo];
        const double id = d != 0. ? 1. / d : 1.;
        d64[r] = id;
        d32[r] = (float)id;
    }
}

// ---- host-side restatement of the rules for the tables -----------------------
// base/quad/GaussLegendre.hpp:114-134 constants and order;
// base/quad/TensorProduct.hpp:123-175 tensor index (first coordinate fastest).
int gauss1d(int np, double* w, double* x)
{
    if (np == 1) { w[0] = 1.0; x[0] = 0.5; return 0; }
    if (np == 2) {
        w[0] = 0.5; x[0] = 0.788675134594813;
        w[1] = 0.5; x[1] = 0.211324865405187;
        return 0;
    }
    if (np == 3) {
        w[0] = 0.277777777777777; x[0] = 0.887298334620741;
        w[1] = 0.444444444444444; x[1] = 0.5;
        w[2] = 0.277777777777777; x[2] = 0.112701665379259;
        return 0;
    }
    return -1;
}

int tensor_rule(int dim, int np, std::vector<double>& w, std::vector<double>& xi)
{
    double w1[4], x1[4];
    if (gauss1d(np, w1, x1)) return -1;
    int nq = 1;
    for (int d = 0; d < dim; d++) nq *= np;
    w.assign(nq, 0.);
    xi.assign((size_t)nq * dim, 0.);
    for (int q = 0; q < nq; q++) {
        int r = q;
        const int i0 = r % np;
        r /= np;
        double ww = w1[i0];
        xi[q * dim] = x1[i0];
        for (int d = 1; d < dim; d++) {
            const int id = r % np;
            r /= np;
            ww *= w1[id];
            xi[q * dim + d] = x1[id];
        }
        w[q] = ww;
    }
    return nq;
}

// Q1 shape functions in hierarchic order: base/sfun/Lagrange1D.ipp:46-58,
// base/sfun/TensorProduct.hpp:40-57,225-320, HierarchicOrder.hpp:201-272.
void q1_shape(int dim, const double* xi, double* phi, double* dphi)
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
    double f2[4], g2[4][2];
    int c = 0;
    for (int o = 0; o < 2; o++)
        for (int i = 0; i < 2; i++) {
            f2[c] = f[0][i] * f[1][o];
            g2[c][0] = f[1][o] * g[0][i];
            g2[c][1] = g[1][o] * f[0][i];
            c++;
        }
    if (dim == 1) {
        phi[0] = f[0][0]; phi[1] = f[0][1];
        dphi[0] = g[0][0]; dphi[1] = g[0][1];
        return;
    }
    if (dim == 2) {
        for (int n = 0; n < 4; n++) {
            phi[H2[n]] = f2[n];
            dphi[H2[n] * 2] = g2[n][0];
            dphi[H2[n] * 2 + 1] = g2[n][1];
        }
        return;
    }
    c = 0;
    double lf[8], lg[8][3];
    for (int o = 0; o < 2; o++)
        for (int i = 0; i < 4; i++) {
            lf[c] = f2[i] * f[2][o];
            lg[c][0] = f[2][o] * g2[i][0];
            lg[c][1] = f[2][o] * g2[i][1];
            lg[c][2] = g[2][o] * f2[i];
            c++;
        }
    for (int n = 0; n < 8; n++) {
        phi[H3[n]] = lf[n];
        for (int d = 0; d < 3; d++) dphi[H3[n] * 3 + d] = lg[n][d];
    }
}

// Lagrange shape functions of degree 2 in hierarchic order
// (base/sfun/Lagrange1D.ipp:76-91 in 1-D, tensor products of
// base/sfun/TensorProduct.hpp:225-320, reorder out[H[n]] = in[n] with the
// table of base/mesh/HierarchicOrder.hpp:201-270 (QUAD) and :345-452 (HEX):
// vertices, edges, faces, interior).
void lagrange2_shape(int dim, const double* xi, double* phi, double* dphi)
{
    static const int H2[9] = {0, 4, 1, 7, 8, 5, 3, 6, 2};
    static const int H3[27] = {0, 8, 1, 11, 20, 9, 3, 10, 2, 16, 22, 17, 25, 26, 23, 19, 24, 18,
                               4, 12, 5, 15, 21, 13, 7, 14, 6};
    double f[3][3], g[3][3];
    for (int d = 0; d < dim; d++) {
        const double x = xi[d];
        f[d][0] = (1. - x) * (1. - 2. * x);
        f[d][1] = 4. * x * (1. - x);
        f[d][2] = x * (2. * x - 1.);
        g[d][0] = 4. * x - 3.;
        g[d][1] = 4. - 8. * x;
        g[d][2] = 4. * x - 1.;
    }
    double lf[27], lg[27][3];
    int m = 3;
    for (int i = 0; i < 3; i++) {
        lf[i] = f[0][i];
        lg[i][0] = g[0][i];
    }
    for (int d = 1; d < dim; d++) {
        double tf[27], tg[27][3];
        int c = 0;
        for (int o = 0; o < 3; o++)
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
    const int* H = dim == 2 ? H2 : H3;
    for (int n = 0; n < m; n++) {
        phi[H[n]] = lf[n];
        for (int a = 0; a < dim; a++) dphi[H[n] * dim + a] = lg[n][a];
    }
}

// degree-2 field table: geometry dphi, geometry phi, field dphi, field phi, w;
// then the face table: geometry dphi, geometry phi, field phi (LINE3 ends
// first then the midpoint, base/mesh/HierarchicOrder.hpp:131-150; QUAD9
// hierarchic), w
void fill_tables_q2(int dim, int nq1, std::vector<double>& t, int* face_off, int* nqf)
{
    std::vector<double> w, xi;
    const int nq = tensor_rule(dim, nq1, w, xi);
    const int gn = 1 << dim, nf = dim == 2 ? 9 : 27;
    const size_t GD = 0, GP = GD + (size_t)nq * gn * dim, FD = GP + (size_t)nq * gn,
                 FP = FD + (size_t)nq * nf * dim, W = FP + (size_t)nq * nf;
    t.assign(W + nq, 0.);
    for (int q = 0; q < nq; q++) {
        double phi[8], dphi[24], fphi[27], fdphi[81];
        q1_shape(dim, &xi[q * dim], phi, dphi);
        lagrange2_shape(dim, &xi[q * dim], fphi, fdphi);
        for (int n = 0; n < gn; n++) {
            for (int d = 0; d < dim; d++) t[GD + (q * gn + n) * dim + d] = dphi[n * dim + d];
            t[GP + q * gn + n] = phi[n];
        }
        for (int n = 0; n < nf; n++) {
            for (int d = 0; d < dim; d++) t[FD + (q * nf + n) * dim + d] = fdphi[n * dim + d];
            t[FP + q * nf + n] = fphi[n];
        }
        t[W + q] = w[q];
    }
    *face_off = (int)t.size();
    std::vector<double> wf, xf;
    const int ld = dim - 1, gnf = 1 << ld, ffn = dim == 2 ? 3 : 9;
    const int nqfv = tensor_rule(ld, nq1, wf, xf);
    const size_t b0 = t.size(), bGP = b0 + (size_t)nqfv * gnf * ld, bFP = bGP + (size_t)nqfv * gnf,
                 bW = bFP + (size_t)nqfv * ffn;
    t.resize(bW + nqfv, 0.);
    for (int q = 0; q < nqfv; q++) {
        double phi[4], dphi[8];
        q1_shape(ld, &xf[q * ld], phi, dphi);
        for (int n = 0; n < gnf; n++) {
            for (int d = 0; d < ld; d++) t[b0 + (q * gnf + n) * ld + d] = dphi[n * ld + d];
            t[bGP + q * gnf + n] = phi[n];
        }
        if (ld == 1) {
            const double x = xf[q];
            t[bFP + q * 3 + 0] = (1. - x) * (1. - 2. * x);
            t[bFP + q * 3 + 1] = x * (2. * x - 1.);
            t[bFP + q * 3 + 2] = 4. * x * (1. - x);
        } else {
            double fphi[9], fdphi[18];
            lagrange2_shape(2, &xf[q * ld], fphi, fdphi);
            for (int n = 0; n < 9; n++) t[bFP + q * 9 + n] = fphi[n];
        }
        t[bW + q] = wf[q];
    }
    *nqf = nqfv;
}

// volume table followed by the face table (dphi, phi, w on faces)
void fill_tables(int dim, int nq1, std::vector<double>& t, int* face_off, int* nqf)
{
    std::vector<double> w, xi;
    const int nq = tensor_rule(dim, nq1, w, xi);
    const int npe = 1 << dim;
    t.assign((size_t)nq * npe * (dim + 1) + nq, 0.);
    for (int q = 0; q < nq; q++) {
        double phi[8], dphi[24];
        q1_shape(dim, &xi[q * dim], phi, dphi);
        for (int n = 0; n < npe; n++) {
            for (int d = 0; d < dim; d++) t[(q * npe + n) * dim + d] = dphi[n * dim + d];
            t[(size_t)nq * npe * dim + q * npe + n] = phi[n];
        }
        t[(size_t)nq * npe * (dim + 1) + q] = w[q];
    }
    *face_off = (int)t.size();
    std::vector<double> wf, xf;
    const int ld = dim - 1;
    const int nqfv = tensor_rule(ld, nq1, wf, xf);
    const int nfn = 1 << ld;
    const size_t base = t.size();
    t.resize(base + (size_t)nqfv * nfn * (ld + 1) + nqfv, 0.);
    for (int q = 0; q < nqfv; q++) {
        double phi[4], dphi[8];
        q1_shape(ld, &xf[q * ld], phi, dphi);
        for (int n = 0; n < nfn; n++) {
            for (int d = 0; d < ld; d++) t[base + (q * nfn + n) * ld + d] = dphi[n * ld + d];
            t[base + (size_t)nqfv * nfn * ld + q * nfn + n] = phi[n];
        }
        t[base + (size_t)nqfv * nfn * (ld + 1) + q] = wf[q];
    }
    *nqf = nqfv;
}

template <int DIM, int NQ1, typename T>
void launch_laplace_nq(Ctx* c, unsigned g, const T* tab, Bf bf, T* ke, double* fe, T* me)
{
    if (me)
        k_elem_laplace<DIM, NQ1, T, true><<<g, BLK, 0, c->stream>>>(
            c->n_elem, c->conn.p, c->xyz.p, c->coef.p, c->ncoef, tab, bf, ke, fe, me, (T)c->density);
    else
        k_elem_laplace<DIM, NQ1, T, false><<<g, BLK, 0, c->stream>>>(
            c->n_elem, c->conn.p, c->xyz.p, c->coef.p, c->ncoef, tab, bf, ke, fe);
}

template <int DIM, typename T>
int launch_laplace(Ctx* c, const T* tab, Bf bf, T* ke, double* fe, T* me)
{
    const unsigned g = grid_for(c->n_elem, BLK);
    switch (c->nq) {
    case 1: launch_laplace_nq<DIM, 1, T>(c, g, tab, bf, ke, fe, me); break;
    case 2: launch_laplace_nq<DIM, 2, T>(c, g, tab, bf, ke, fe, me); break;
    case 3: launch_laplace_nq<DIM, 3, T>(c, g, tab, bf, ke, fe, me); break;
    default: return fail(c, INS_ERR_UNSUPPORTED, "quadrature with more than 3 points per direction");
    }
    INS_HIP_CHECK(c, hipGetLastError());
    return INS_OK;
}

template <int DIM, typename T>
int launch_laplace_q2(Ctx* c, const T* tab, Bf bf, T* ke, double* fe, T* me)
{
    const unsigned g = grid_for(c->n_elem, 4, 1u << 20);
    switch (c->nq) {
    case 1: k_elem_laplace_q2<DIM, 1, T><<<g, 256, 0, c->stream>>>(c->n_elem, c->gconn.p, c->xyz.p, c->coef.p, c->ncoef, tab, bf, ke, fe, me, (T)c->density); break;
    case 2: k_elem_laplace_q2<DIM, 2, T><<<g, 256, 0, c->stream>>>(c->n_elem, c->gconn.p, c->xyz.p, c->coef.p, c->ncoef, tab, bf, ke, fe, me, (T)c->density); break;
    case 3: k_elem_laplace_q2<DIM, 3, T><<<g, 256, 0, c->stream>>>(c->n_elem, c->gconn.p, c->xyz.p, c->coef.p, c->ncoef, tab, bf, ke, fe, me, (T)c->density); break;
    default: return fail(c, INS_ERR_UNSUPPORTED, "quadrature with more than 3 points per direction");
    }
    INS_HIP_CHECK(c, hipGetLastError());
    return INS_OK;
}

// the hyperelastic element kernels for a field of NF functions per element:
// the point records, then the element matrices with chunks of points that
// fill the 64 lanes (HEX8: 4 points of 8 functions, HEX27: 2 points of 27)
template <int DIM, int NF, int NQ1, typename T, int MAT>
int launch_hyper_nq(Ctx* c, const T* tab, Bf bf, int want_k, int want_r, T* ke, double* fe)
{
    constexpr int NQ = IPow<NQ1, DIM>::value;
    constexpr int QC0 = 64 / NF < 4 ? 64 / NF : 4;
    constexpr int QC = QC0 < NQ ? QC0 : NQ;
    ElTab<T> tb;
    if constexpr (NF == (1 << DIM)) {
        using TB = Tab<DIM, NQ>;
        tb = ElTab<T>{tab + TB::DPHI, tab + TB::PHI, tab + TB::DPHI, tab + TB::PHI, tab + TB::W};
    } else {
        using TB = TabQ2<DIM, NQ>;
        tb = ElTab<T>{tab + TB::GD, tab + TB::GP, tab + TB::FD, tab + TB::FP, tab + TB::W};
    }
    const int want_f = fe != nullptr;
    const int64_t E = c->n_elem;
    INS_TRY(ensure(c, c->stage_pt, (size_t)E * NQ * PtRec::SIZE * sizeof(T)));
    T* pts = (T*)c->stage_pt.p;
    // a whole number of blocks per XCD (the kernel gives each XCD an eighth)
    const unsigned g = (unsigned)std::max<int64_t>(8, std::min<int64_t>((E + 7) / 8 * 8, 1 << 22));
    const int has_bf = bf.kind != INS_BF_NONE;
    auto run = [&](auto lin) {
        constexpr bool LIN = decltype(lin)::value;
        k_hyper_points<DIM, NF, NQ1, T, MAT, LIN><<<grid_for(E * NQ, 256), 256, 0, c->stream>>>(
            E, c->geo_conn(), c->conn.p, c->xyz.p, c->coef.p, c->u.p, c->ncoef, tb, bf, want_k,
            want_f, pts);
        if (want_k)
            k_elem_hyper<DIM, NF, NQ1, QC, T, true, LIN><<<g, 64, 0, c->stream>>>(
                E, pts, tb, want_f, want_r, has_bf, ke, fe);
        else
            k_elem_hyper<DIM, NF, NQ1, QC, T, false, LIN><<<g, 64, 0, c->stream>>>(
                E, pts, tb, want_f, want_r, has_bf, ke, fe);
    };
    // the displacement is known to be zero until something writes it
    // (ins_set_dof_values, ins_distribute, time stepping): then F = I
    // exactly and the point records reduce to the geometry and the
    // isotropic tangent (INS_LIN=0 takes the general form throughout)
    static const int lin_on = [] { const char* v = getenv("INS_LIN"); return v ? atoi(v) : 1; }();
    if (c->u_zero && lin_on) run(std::true_type{});
    else run(std::false_type{});
    INS_HIP_CHECK(c, hipGetLastError());
    return INS_OK;
}

template <int DIM, int NF, typename T, int MAT>
int launch_hyper(Ctx* c, const T* tab, Bf bf, int want_k, int want_r, T* ke, double* fe)
{
    switch (c->nq) {
    case 1: return launch_hyper_nq<DIM, NF, 1, T, MAT>(c, tab, bf, want_k, want_r, ke, fe);
    case 2: return launch_hyper_nq<DIM, NF, 2, T, MAT>(c, tab, bf, want_k, want_r, ke, fe);
    case 3: return launch_hyper_nq<DIM, NF, 3, T, MAT>(c, tab, bf, want_k, want_r, ke, fe);
    default: return fail(c, INS_ERR_UNSUPPORTED, "quadrature with more than 3 points per direction");
    }
}

template <typename T, int DIM>
int launch_elastic_mat(Ctx* c, const T* tab, Bf bf, int want_k, int want_r, T* ke, double* fe)
{
    constexpr int NF1 = 1 << DIM, NF2 = DIM == 3 ? 27 : 9;
    if (c->deg == 2) {
        if (c->kernel == INS_STVENANT)
            return launch_hyper<DIM, NF2, T, INS_STVENANT>(c, tab, bf, want_k, want_r, ke, fe);
        return launch_hyper<DIM, NF2, T, INS_NEOHOOKE>(c, tab, bf, want_k, want_r, ke, fe);
    }
    if (c->kernel == INS_STVENANT)
        return launch_hyper<DIM, NF1, T, INS_STVENANT>(c, tab, bf, want_k, want_r, ke, fe);
    return launch_hyper<DIM, NF1, T, INS_NEOHOOKE>(c, tab, bf, want_k, want_r, ke, fe);
}

int lg2(int v)
{
    int l = 0;
    while ((1 << l) < v) l++;
    return l;
}

template <int NPE, int S, class KE, class ME>
int launch_gather_rhs(Ctx* c, KE ke, ME me, int do_matrix, int do_rhs, int lift,
                      int incremental, const double* fe, const double* ff, int precision,
                      int* derr, bool direct = false)
{
    if constexpr (NPE <= 8) {
        if (do_matrix && direct) {
            // straight into the symmetric slices (and the Jacobi diagonal)
            Sell& S_ = c->sell;
            const int64_t N = S_.snodes;
            const unsigned g = grid_for(N, BLK);
            double* o64 = precision == INS_FP64 ? S_.vs64.p : nullptr;
            if (precision == INS_FP64)
                k_sym_gather<NPE, S, KE, double, ME><<<g, BLK, 0, c->stream>>>(
                    N, S_.nu, S_.anode.p, c->adj_ptr.p, c->adj.p, c->lg_npe, S_.spw.p, ke, me,
                    c->mass_factor, o64, S_.vs.p, c->dinv64.p, c->dinv32.p);
            else
                k_sym_gather<NPE, S, KE, float, ME><<<g, BLK, 0, c->stream>>>(
                    N, S_.nu, S_.anode.p, c->adj_ptr.p, c->adj.p, c->lg_npe, S_.spw.p, ke, me,
                    c->mass_factor, o64, S_.vs.p, c->dinv64.p, c->dinv32.p);
            INS_HIP_CHECK(c, hipGetLastError());
            do_matrix = 0;
        }
    }
    if (do_matrix) {
        int G = 1;
        while (G < c->max_nb && G < 64) G <<= 1;
        const int lgG = lg2(G);
        const int64_t threads = c->n_nodes << lgG;
        const unsigned g = (unsigned)((threads + BLK - 1) / BLK);
        double* v64 = (precision == INS_FP64) ? c->val64.p : nullptr;
        float* v32 = c->val32.p;
        const typename PairWord<NPE>::type* pw = nullptr;
        if (c->have_pinfo) {
            if constexpr (NPE <= 8) pw = c->pinfo.p;
            else pw = c->pinfo64.p;
        }
        if (precision == INS_FP64)
            k_gather<NPE, S, KE, double, ME><<<g, BLK, 0, c->stream>>>(
                c->n_nodes, lgG, c->lg_npe, c->adj_ptr.p, c->adj.p, c->nb_ptr.p, c->nb.p,
                c->nboff.p, c->dof_id.p, c->row_ptr.p, c->col.p, c->conn.p, pw,
                c->monotone ? 1 : 0, ke, me, c->mass_factor, v64, v32, derr);
        else
            k_gather<NPE, S, KE, float, ME><<<g, BLK, 0, c->stream>>>(
                c->n_nodes, lgG, c->lg_npe, c->adj_ptr.p, c->adj.p, c->nb_ptr.p, c->nb.p,
                c->nboff.p, c->dof_id.p, c->row_ptr.p, c->col.p, c->conn.p, pw,
                c->monotone ? 1 : 0, ke, me, c->mass_factor, nullptr, v32, derr);
        INS_HIP_CHECK(c, hipGetLastError());
    }
    if (do_rhs) {
        k_rhs<NPE, S, KE, ME><<<grid_for(c->n_nodes, BLK), BLK, 0, c->stream>>>(
            c->n_nodes, c->lg_npe, c->adj_ptr.p, c->adj.p, c->conn.p, c->dof_id.p, c->g.p, c->u.p,
            incremental, lift, ke, fe, lg2(c->nfn > 0 ? c->nfn : 1), c->fadj_ptr.p, c->fadj.p, ff,
            c->nfn, me, c->mass_factor, c->have_mhist ? c->mhist.p : nullptr,
            c->have_emask ? c->emask.p : nullptr, c->fe_slab ? 1 : 0, c->b.p);
        INS_HIP_CHECK(c, hipGetLastError());
    }
    return INS_OK;
}

template <typename T, int NF>
int gather_q2_t(Ctx* c, bool elastic, int do_matrix, int do_rhs, int lift, int incremental,
                const double* fe, const double* ff, int precision, int* derr)
{
    const T* st = (const T*)c->stage_ke.p;
    const KeLaplace<T, NF> me{(!elastic && c->density > 0.) ? (const T*)c->stage_me.p : nullptr};
    const int s = c->dpn;
    constexpr int D = NF == 27 ? 3 : 2;
    if (elastic)
        return launch_gather_rhs<NF, D>(c, KeElastic<T, NF, D>{st}, me, do_matrix, do_rhs, lift,
                                        incremental, fe, ff, precision, derr);
    if (s == 1)
        return launch_gather_rhs<NF, 1>(c, KeLaplace<T, NF>{st}, me, do_matrix, do_rhs, lift,
                                        incremental, fe, ff, precision, derr);
    if (s == D)
        return launch_gather_rhs<NF, D>(c, KeLaplace<T, NF>{st}, me, do_matrix, do_rhs, lift,
                                        incremental, fe, ff, precision, derr);
    return fail(c, INS_ERR_UNSUPPORTED, "element/dof combination");
}

int gather_q2(Ctx* c, bool elastic, int do_matrix, int do_rhs, int lift, int incremental,
              const double* fe, const double* ff, int precision, int* derr)
{
    if (c->npe == 27)
        return precision == INS_FP64
                   ? gather_q2_t<double, 27>(c, elastic, do_matrix, do_rhs, lift, incremental, fe, ff, precision, derr)
                   : gather_q2_t<float, 27>(c, elastic, do_matrix, do_rhs, lift, incremental, fe, ff, precision, derr);
    return precision == INS_FP64
               ? gather_q2_t<double, 9>(c, elastic, do_matrix, do_rhs, lift, incremental, fe, ff, precision, derr)
               : gather_q2_t<float, 9>(c, elastic, do_matrix, do_rhs, lift, incremental, fe, ff, precision, derr);
}

}  // namespace

int quadrature_rule(int dim, int npts, double* w, double* xi)
{
    std::vector<double> ww, xx;
    const int nq = tensor_rule(dim, npts, ww, xx);
    if (nq < 0) return -1;
    for (int q = 0; q < nq; q++) {
        w[q] = ww[q];
        for (int d = 0; d < dim; d++) xi[q * dim + d] = xx[(size_t)q * dim + d];
    }
    return nq;
}

int make_tables(Ctx* c)
{
    std::vector<double> t;
    if (c->deg == 2) {
        fill_tables_q2(c->dim, c->nq, t, &c->tab_face_off, &c->tab_nqf);
    } else {
        fill_tables(c->dim, c->nq, t, &c->tab_face_off, &c->tab_nqf);
    }
    c->tab_deg = c->deg;
    std::vector<float> tf(t.begin(), t.end());
    INS_TRY(ensure(c, c->tab64, t.size()));
    INS_TRY(ensure(c, c->tab32, t.size()));
    INS_HIP_CHECK(c, hipMemcpyAsync(c->tab64.p, t.data(), t.size() * sizeof(double),
                                    hipMemcpyHostToDevice, c->stream));
    INS_HIP_CHECK(c, hipMemcpyAsync(c->tab32.p, tf.data(), tf.size() * sizeof(float),
                                    hipMemcpyHostToDevice, c->stream));
    INS_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return INS_OK;
}

int compute_dinv(Ctx* c)
{
    INS_TRY(ensure(c, c->dinv64, c->n_active));
    INS_TRY(ensure(c, c->dinv32, c->n_active));
    const double* v64 = (c->matrix_precision == INS_FP64) ? c->val64.p : nullptr;
    k_dinv<<<grid_for(c->n_active, BLK), BLK, 0, c->stream>>>(
        c->n_active, c->row_ptr.p, c->col.p, v64, c->val32.p, c->dinv64.p, c->dinv32.p);
    INS_HIP_CHECK(c, hipGetLastError());
    return INS_OK;
}

int assemble(Ctx* c, int precision, int flags)
{
    if (!c->have_pattern) return fail(c, INS_ERR_STATE, "ins_assemble before ins_build_pattern");
    if (c->csr_from_host)
        return fail(c, INS_ERR_STATE,
                    "ins_assemble after ins_set_csr: call ins_build_pattern to assemble into the "
                    "DoF pattern again");
    if (c->kernel == 0) return fail(c, INS_ERR_STATE, "ins_assemble before ins_set_material");
    const bool elastic = (c->kernel == INS_STVENANT || c->kernel == INS_NEOHOOKE);
    if (elastic && c->dpn != c->dim)
        return fail(c, INS_ERR_UNSUPPORTED, "elasticity needs dofs_per_node == dim");
    if (!elastic && c->dpn != 1 && c->dpn != c->dim)
        return fail(c, INS_ERR_UNSUPPORTED, "Laplace supports 1 or dim components");
    // ASM_STAGED: the RHS again from the element data staged by the last full
    // assembly (the same matrix and loads, a new history or new prescribed
    // values); no element kernels run
    const bool staged = (flags & ASM_STAGED) != 0;
    const int do_matrix = staged ? 0 : (flags & INS_ASM_MATRIX);
    const int do_rhs = staged ? 1 : (flags & INS_ASM_RHS);
    const int want_r = (flags & INS_ASM_RESIDUAL) && elastic;
    const int incremental = (flags & INS_ASM_INCREMENTAL) ? 1 : 0;
    const int npe = c->npe, s = c->dpn;
    const int64_t E = c->n_elem;
    Bf bf;
    bf.kind = c->bf_kind;
    for (int i = 0; i < 4; i++) bf.p[i] = c->bf[i];
    bf.tab = c->bftab.p;
    bf.nqp = 1;
    for (int d = 0; d < c->dim; d++) bf.nqp *= c->nq;
    bf.s = s;
    if (bf.kind == INS_BF_TABLE && !c->have_bftab)
        return fail(c, INS_ERR_STATE, "body force values not set");
    if (c->tab_deg != c->deg) INS_TRY(make_tables(c));

    INS_HIP_CHECK(c, hipEventRecord(c->ev0, c->stream));
    int* derr = nullptr;
    INS_TRY(ensure(c, c->dflag, 4));
    derr = c->dflag.p;
    INS_HIP_CHECK(c, hipMemsetAsync(derr, 0, sizeof(int), c->stream));

    // staging
    const size_t esize = (precision == INS_FP64) ? sizeof(double) : sizeof(float);
    size_t ke_elems;
    if (!elastic)
        ke_elems = (size_t)((E + 63) / 64) * 64 * (npe * (npe + 1) / 2);
    else
        ke_elems = (size_t)((E + 63) / 64) * 64 * (npe * s) * (npe * s + 1) / 2;
    if (do_matrix) INS_TRY(ensure(c, c->stage_ke, ke_elems * esize));
    const bool need_fe = do_rhs && (c->bf_kind != INS_BF_NONE || want_r);
    if (need_fe && !staged) INS_TRY(ensure(c, c->stage_fe, (size_t)((E + 63) / 64) * 64 * npe * s));
    // element force vectors in 64-element slabs: those of the Q1 heat::Laplace kernel
    if (!staged) c->fe_slab = !elastic && c->deg == 1;
    // time-step mass terms (heat::Laplace only)
    const bool mass = c->density > 0.;
    if (mass && elastic)
        return fail(c, INS_ERR_UNSUPPORTED, "mass terms are provided for heat::Laplace");
    if (mass && !staged) INS_TRY(ensure(c, c->stage_me, ke_elems * esize));
    const void* mep = mass ? (const void*)c->stage_me.p : nullptr;
    if (do_matrix) {
        if (precision == INS_FP64) INS_TRY(ensure(c, c->val64, c->nnz));
        INS_TRY(ensure(c, c->val32, c->nnz));
    }
    if (do_rhs) {
        INS_TRY(ensure(c, c->b, c->n_active));
        INS_TRY(ensure(c, c->g, (size_t)c->n_nodes * s));
        INS_TRY(ensure(c, c->u, (size_t)c->n_nodes * s));
    }
    double* fe = need_fe ? c->stage_fe.p : nullptr;
    // a SELL solve of a problem with symmetric slices reads only those: the
    // gather writes them directly, and the CSR values follow on demand
    const Sell& SS = c->sell;
    const bool direct = do_matrix && c->want_format == INS_FMT_SELL && SS.built && SS.ok &&
                        SS.sym_on && SS.direct && c->deg == 1 && c->npe <= 8 &&
                        SS.sdpn == s && c->dofs_match();
    if (direct) {
        const int prec_was = c->matrix_precision;
        c->matrix_precision = precision;   // zero_sym sizes the FP64 copy by it
        const int zs = zero_sym(c);
        c->matrix_precision = prec_was;
        INS_TRY(zs);
        INS_TRY(ensure(c, c->dinv64, c->n_active));
        INS_TRY(ensure(c, c->dinv32, c->n_active));
    }
    if (do_rhs && !c->have_emask && c->deg == 1 && npe * s <= 32) {
        INS_TRY(ensure(c, c->emask, (size_t)E));
        if (npe == 8 && s == 1)
            k_emask<8, 1><<<grid_for(E, BLK), BLK, 0, c->stream>>>(E, c->conn.p, c->dof_id.p, c->emask.p);
        else if (npe == 8 && s == 3)
            k_emask<8, 3><<<grid_for(E, BLK), BLK, 0, c->stream>>>(E, c->conn.p, c->dof_id.p, c->emask.p);
        else if (npe == 4 && s == 1)
            k_emask<4, 1><<<grid_for(E, BLK), BLK, 0, c->stream>>>(E, c->conn.p, c->dof_id.p, c->emask.p);
        else if (npe == 4 && s == 2)
            k_emask<4, 2><<<grid_for(E, BLK), BLK, 0, c->stream>>>(E, c->conn.p, c->dof_id.p, c->emask.p);
        INS_HIP_CHECK(c, hipGetLastError());
        c->have_emask = (npe == 8 && (s == 1 || s == 3)) || (npe == 4 && (s == 1 || s == 2));
    }

    // 1. element kernels
    if (c->profiling) INS_HIP_CHECK(c, hipEventRecord(c->ev_el[0], c->stream));
    if (staged) {
    } else if (!elastic) {
        if (do_matrix || need_fe || mass) {
            if (!do_matrix) INS_TRY(ensure(c, c->stage_ke, ke_elems * esize));
            if (c->deg == 2) {
                if (precision == INS_FP64) {
                    if (c->dim == 3) INS_TRY((launch_laplace_q2<3, double>(c, c->tab64.p, bf, (double*)c->stage_ke.p, fe, (double*)mep)));
                    else INS_TRY((launch_laplace_q2<2, double>(c, c->tab64.p, bf, (double*)c->stage_ke.p, fe, (double*)mep)));
                } else {
                    if (c->dim == 3) INS_TRY((launch_laplace_q2<3, float>(c, c->tab32.p, bf, (float*)c->stage_ke.p, fe, (float*)mep)));
                    else INS_TRY((launch_laplace_q2<2, float>(c, c->tab32.p, bf, (float*)c->stage_ke.p, fe, (float*)mep)));
                }
            } else if (precision == INS_FP64) {
                if (c->dim == 3) INS_TRY((launch_laplace<3, double>(c, c->tab64.p, bf, (double*)c->stage_ke.p, fe, (double*)mep)));
                else INS_TRY((launch_laplace<2, double>(c, c->tab64.p, bf, (double*)c->stage_ke.p, fe, (double*)mep)));
            } else {
                if (c->dim == 3) INS_TRY((launch_laplace<3, float>(c, c->tab32.p, bf, (float*)c->stage_ke.p, fe, (float*)mep)));
                else INS_TRY((launch_laplace<2, float>(c, c->tab32.p, bf, (float*)c->stage_ke.p, fe, (float*)mep)));
            }
        }
    } else {
        if (do_matrix) {
            if (precision == INS_FP64) {
                if (c->dim == 3) INS_TRY((launch_elastic_mat<double, 3>(c, c->tab64.p, bf, 1, want_r && need_fe, (double*)c->stage_ke.p, fe)));
                else INS_TRY((launch_elastic_mat<double, 2>(c, c->tab64.p, bf, 1, want_r && need_fe, (double*)c->stage_ke.p, fe)));
            } else {
                if (c->dim == 3) INS_TRY((launch_elastic_mat<float, 3>(c, c->tab32.p, bf, 1, 0, (float*)c->stage_ke.p, nullptr)));
                else INS_TRY((launch_elastic_mat<float, 2>(c, c->tab32.p, bf, 1, 0, (float*)c->stage_ke.p, nullptr)));
            }
        }
        // forces always in FP64 (the Newton residual stays FP64); a
        // residual-only pass when the tangent was FP32 or not requested
        if (need_fe && (precision != INS_FP64 || !do_matrix)) {
            if (c->dim == 3) INS_TRY((launch_elastic_mat<double, 3>(c, c->tab64.p, bf, 0, want_r, nullptr, fe)));
            else INS_TRY((launch_elastic_mat<double, 2>(c, c->tab64.p, bf, 0, want_r, nullptr, fe)));
        }
    }
    if (c->profiling) INS_HIP_CHECK(c, hipEventRecord(c->ev_el[1], c->stream));
    // faces
    double* ff = nullptr;
    if (staged && c->n_faces > 0) {
        ff = c->stage_ff.p;
    } else if (do_rhs && c->n_faces > 0) {
        INS_TRY(ensure(c, c->stage_ff, (size_t)c->n_faces * c->nfn * s));
        ff = c->stage_ff.p;
        const unsigned g = grid_for(c->n_faces, BLK);
        // face table: geometry dphi [nqf][gnf][dim-1], geometry phi [nqf][gnf],
        // field phi [nqf][nfn], weights [nqf] (for Q1 the field phi is the
        // geometry phi)
        const int nqf = c->tab_nqf, ld = c->dim - 1, gnf = 1 << ld;
        const double* gd = c->tab64.p + c->tab_face_off;
        const double* gp = gd + (size_t)nqf * gnf * ld;
        const double* fp = c->deg == 2 ? gp + (size_t)nqf * gnf : gp;
        const double* w = fp + (size_t)nqf * c->nfn;
        const int32_t* gmap = c->deg == 2 ? c->fvert.p : nullptr;
        const double* T = c->traction;
        const int ax = c->tr_axis;
        const double x0 = c->tr_x0;
        RefFn flux;
        flux.kind = c->flux_kind;
        for (int i = 0; i < 8; i++) flux.p[i] = c->flux_p[i];
        // ff holds nfn * s values per face (the gather reads that stride)
        const int key = c->dim * 100 + c->nfn * 10 + s;
#define INS_FACE(D, N, S_)                                                                      \
    k_face<D, N, S_><<<g, BLK, 0, c->stream>>>(c->n_faces, c->faces.p, gmap, c->xyz.p, gd, gp, \
                                               fp, w, nqf, T[0], T[1], T[2], ax, x0, flux,    \
                                               c->have_ttab ? c->ttab.p : nullptr, ff)
        switch (key) {
        case 393: INS_FACE(3, 9, 3); break;
        case 391: INS_FACE(3, 9, 1); break;
        case 343: INS_FACE(3, 4, 3); break;
        case 341: INS_FACE(3, 4, 1); break;
        case 232: INS_FACE(2, 3, 2); break;
        case 231: INS_FACE(2, 3, 1); break;
        case 222: INS_FACE(2, 2, 2); break;
        case 221: INS_FACE(2, 2, 1); break;
        default: return fail(c, INS_ERR_UNSUPPORTED, "Neumann faces: dofs per node must be 1 or dim");
        }
#undef INS_FACE
        INS_HIP_CHECK(c, hipGetLastError());
    }

    // 2./3. gather into CSR (or straight into the symmetric slices of a
    // SELL solve) and RHS
    const int lift = do_matrix || staged;
    if (c->deg == 2) {
        INS_TRY(gather_q2(c, elastic, do_matrix, do_rhs, lift, incremental, fe, ff, precision, derr));
    } else if (!elastic) {
        if (precision == INS_FP64) {
            if (npe == 8 && s == 1) INS_TRY((launch_gather_rhs<8, 1>(c, KeLaplace<double, 8>{(const double*)c->stage_ke.p}, KeLaplace<double, 8>{(const double*)mep}, do_matrix, do_rhs, lift, incremental, fe, ff, precision, derr, direct)));
            else if (npe == 8 && s == 3) INS_TRY((launch_gather_rhs<8, 3>(c, KeLaplace<double, 8>{(const double*)c->stage_ke.p}, KeLaplace<double, 8>{(const double*)mep}, do_matrix, do_rhs, lift, incremental, fe, ff, precision, derr, direct)));
            else if (npe == 4 && s == 1) INS_TRY((launch_gather_rhs<4, 1>(c, KeLaplace<double, 4>{(const double*)c->stage_ke.p}, KeLaplace<double, 4>{(const double*)mep}, do_matrix, do_rhs, lift, incremental, fe, ff, precision, derr, direct)));
            else if (npe == 4 && s == 2) INS_TRY((launch_gather_rhs<4, 2>(c, KeLaplace<double, 4>{(const double*)c->stage_ke.p}, KeLaplace<double, 4>{(const double*)mep}, do_matrix, do_rhs, lift, incremental, fe, ff, precision, derr, direct)));
            else return fail(c, INS_ERR_UNSUPPORTED, "element/dof combination");
        } else {
            if (npe == 8 && s == 1) INS_TRY((launch_gather_rhs<8, 1>(c, KeLaplace<float, 8>{(const float*)c->stage_ke.p}, KeLaplace<float, 8>{(const float*)mep}, do_matrix, do_rhs, lift, incremental, fe, ff, precision, derr, direct)));
            else if (npe == 8 && s == 3) INS_TRY((launch_gather_rhs<8, 3>(c, KeLaplace<float, 8>{(const float*)c->stage_ke.p}, KeLaplace<float, 8>{(const float*)mep}, do_matrix, do_rhs, lift, incremental, fe, ff, precision, derr, direct)));
            else if (npe == 4 && s == 1) INS_TRY((launch_gather_rhs<4, 1>(c, KeLaplace<float, 4>{(const float*)c->stage_ke.p}, KeLaplace<float, 4>{(const float*)mep}, do_matrix, do_rhs, lift, incremental, fe, ff, precision, derr, direct)));
            else if (npe == 4 && s == 2) INS_TRY((launch_gather_rhs<4, 2>(c, KeLaplace<float, 4>{(const float*)c->stage_ke.p}, KeLaplace<float, 4>{(const float*)mep}, do_matrix, do_rhs, lift, incremental, fe, ff, precision, derr, direct)));
            else return fail(c, INS_ERR_UNSUPPORTED, "element/dof combination");
        }
    } else {
        if (precision == INS_FP64) {
            if (npe == 8) INS_TRY((launch_gather_rhs<8, 3>(c, KeElastic<double, 8, 3>{(const double*)c->stage_ke.p}, KeLaplace<double, 8>{(const double*)mep}, do_matrix, do_rhs, lift, incremental, fe, ff, precision, derr, direct)));
            else INS_TRY((launch_gather_rhs<4, 2>(c, KeElastic<double, 4, 2>{(const double*)c->stage_ke.p}, KeLaplace<double, 4>{(const double*)mep}, do_matrix, do_rhs, lift, incremental, fe, ff, precision, derr, direct)));
        } else {
            if (npe == 8) INS_TRY((launch_gather_rhs<8, 3>(c, KeElastic<float, 8, 3>{(const float*)c->stage_ke.p}, KeLaplace<float, 8>{(const float*)mep}, do_matrix, do_rhs, lift, incremental, fe, ff, precision, derr, direct)));
            else INS_TRY((launch_gather_rhs<4, 2>(c, KeElastic<float, 4, 2>{(const float*)c->stage_ke.p}, KeLaplace<float, 4>{(const float*)mep}, do_matrix, do_rhs, lift, incremental, fe, ff, precision, derr, direct)));
        }
    }
    if (do_matrix) {
        c->matrix_precision = precision;
        c->stats.assembled_sym = direct ? 1 : 0;
        if (direct) {
            c->csr_stale = true;    // the CSR values follow on demand
            c->sell.stale = false;
        } else {
            INS_TRY(compute_dinv(c));
            c->csr_stale = false;
            c->sell.stale = true;
        }
    }
    INS_HIP_CHECK(c, hipEventRecord(c->ev1, c->stream));
    int herr = 0;
    INS_HIP_CHECK(c, hipMemcpyAsync(&herr, derr, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    INS_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    float ms = 0.f;
    INS_HIP_CHECK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->stats.assemble_ms = ms;
    c->stats.element_ms = 0.;
    if (c->profiling) {
        float me = 0.f;
        INS_HIP_CHECK(c, hipEventElapsedTime(&me, c->ev_el[0], c->ev_el[1]));
        c->stats.element_ms = me;
    }
    if (herr) return fail(c, INS_ERR_PATTERN, "element entry outside the pattern");
    if (do_matrix) c->have_matrix = true;
    if (do_rhs) c->have_rhs = true;
    c->have_solution = false;
    return INS_OK;
}

template <int DIM, int NF, int S>
static void launch_error_norm(Ctx* c, unsigned g, const double* u, int order, const RefFn& ref,
                              double* part)
{
    const double* t = c->tab64.p;
    int nq = 1;
    for (int d = 0; d < DIM; d++) nq *= c->nq;
    const int gn = 1 << DIM;
    const double *gd, *gp, *fd, *fp, *w;
    if (c->deg == 2) {
        gd = t;
        gp = gd + (size_t)nq * gn * DIM;
        fd = gp + (size_t)nq * gn;
        fp = fd + (size_t)nq * NF * DIM;
        w = fp + (size_t)nq * NF;
    } else {
        gd = fd = t;
        gp = fp = t + (size_t)nq * gn * DIM;
        w = gp + (size_t)nq * gn;
    }
    k_error_norm<DIM, NF, S><<<g, 256, 0, c->stream>>>(c->n_elem, c->geo_conn(), c->conn.p,
                                                        c->xyz.p, u, gd, gp, fd, fp, w, nq,
                                                        order, ref, part);
}

int error_norm(Ctx* c, int order, int kind, const double* prm, const double* hu, double* out)
{
    if (c->kernel == 0) return fail(c, INS_ERR_STATE, "ins_error_norm before ins_set_material");
    if (!c->have_dofs) return fail(c, INS_ERR_STATE, "ins_error_norm before ins_set_dofs");
    if (order != 0 && order != 1) return fail(c, INS_ERR_ARG, "order must be 0 (L2) or 1 (H1)");
    if (kind != INS_REF_ZERO && kind != INS_REF_SINPROD && kind != INS_REF_KELVIN &&
        kind != INS_REF_COSPROD)
        return fail(c, INS_ERR_ARG, "unknown reference solution");
    if (kind == INS_REF_KELVIN && (order != 0 || c->dpn != c->dim))
        return fail(c, INS_ERR_UNSUPPORTED, "the fundamental solution is a vector L2 reference");
    if ((kind != INS_REF_ZERO && !prm) || !out) return fail(c, INS_ERR_ARG, "ins_error_norm arguments");
    if (c->tab_deg != c->deg) INS_TRY(make_tables(c));
    RefFn ref;
    ref.kind = kind;
    for (int i = 0; i < 8; i++) ref.p[i] = (kind != INS_REF_ZERO) ? prm[i] : 0.;
    const int64_t nd = c->n_nodes * c->dpn;
    const double* u = c->u.p;
    double* tmp = nullptr;
    if (hu) {
        INS_HIP_CHECK(c, hipMallocAsync((void**)&tmp, nd * sizeof(double), c->stream));
        INS_HIP_CHECK(c, hipMemcpyAsync(tmp, hu, nd * sizeof(double), hipMemcpyHostToDevice,
                                        c->stream));
        u = tmp;
    }
    const unsigned g = grid_for(c->n_elem, 256, 1024);
    double* part = nullptr;
    INS_HIP_CHECK(c, hipMallocAsync((void**)&part, g * sizeof(double), c->stream));
    const int key = c->dim * 1000 + c->npe * 10 + c->dpn;
    switch (key) {
    case 2041: launch_error_norm<2, 4, 1>(c, g, u, order, ref, part); break;
    case 2042: launch_error_norm<2, 4, 2>(c, g, u, order, ref, part); break;
    case 3081: launch_error_norm<3, 8, 1>(c, g, u, order, ref, part); break;
    case 3083: launch_error_norm<3, 8, 3>(c, g, u, order, ref, part); break;
    case 2091: launch_error_norm<2, 9, 1>(c, g, u, order, ref, part); break;
    case 2092: launch_error_norm<2, 9, 2>(c, g, u, order, ref, part); break;
    case 3271: launch_error_norm<3, 27, 1>(c, g, u, order, ref, part); break;
    case 3273: launch_error_norm<3, 27, 3>(c, g, u, order, ref, part); break;
    default:
        (void)hipFreeAsync(part, c->stream);
        if (tmp) (void)hipFreeAsync(tmp, c->stream);
        return fail(c, INS_ERR_UNSUPPORTED, "element/dof combination");
    }
    INS_HIP_CHECK(c, hipGetLastError());
    std::vector<double> h(g);
    INS_HIP_CHECK(c, hipMemcpyAsync(h.data(), part, g * sizeof(double), hipMemcpyDeviceToHost,
                                    c->stream));
    INS_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    (void)hipFreeAsync(part, c->stream);
    if (tmp) (void)hipFreeAsync(tmp, c->stream);
    double sum = 0.;
    for (unsigned i = 0; i < g; i++) sum += h[i];
    *out = std::sqrt(sum);
    return INS_OK;
}

// ---------------------------------------------------------------------------
// Cell evaluation at one local point of every element (the cell data of the
// reference's VTK output): field value, field gradient
// (base/post/evaluateField.hpp:166-206, 231-273), second Piola-Kirchhoff and
// Cauchy stress (solid/Stress.hpp:38-81 with F = I + grad u^T extended by
// the identity, solid/Deformation.hpp:26-45).  One element per lane; the
// shape functions at the point are uniform over the grid and come in as a
// kernel argument; the per-element results go through LDS so that each
// block writes one contiguous run of out.
struct CellTab {
    double gd[8 * 3];    // geometry (Q1) local gradients
    double fp[27];       // field shape values
    double fd[27 * 3];   // field local gradients
};

template <int DIM, int NF, int S, int QTY>
struct CellPer {
    static constexpr int value = QTY == INS_CELL_VALUE ? S : QTY == INS_CELL_GRADIENT ? DIM * S : 9;
};

template <int DIM, int NF, int S, int QTY, int MAT>
__global__ __launch_bounds__(256) void k_cell_eval(
    int64_t E, const int32_t* __restrict__ gconn, const int32_t* __restrict__ fconn,
    const double* __restrict__ xyz, const double* __restrict__ u,
    const double* __restrict__ coef, int ncoef, CellTab tab, double* __restrict__ out)
{
    constexpr int GN = 1 << DIM;
    constexpr int PER = CellPer<DIM, NF, S, QTY>::value;
    __shared__ double stage[256 * PER];
    const int64_t first = (int64_t)blockIdx.x * 256;
    const int64_t e = first + threadIdx.x;
    if (e < E) {
        double o[PER];
        if constexpr (QTY == INS_CELL_VALUE) {
#pragma unroll
            for (int c = 0; c < S; c++) o[c] = 0.;
            for (int n = 0; n < NF; n++) {
                const int64_t node = fconn[e * NF + n];
#pragma unroll
                for (int c = 0; c < S; c++) o[c] += tab.fp[n] * u[node * S + c];
            }
        } else {
            double X[GN * DIM];
#pragma unroll
            for (int n = 0; n < GN; n++) {
                const int64_t node = gconn[e * GN + n];
#pragma unroll
                for (int d = 0; d < DIM; d++) X[n * DIM + d] = xyz[node * DIM + d];
            }
            double inv[DIM][DIM];
            (void)geo_inverse<DIM, double>(X, tab.gd, inv);
            double G[3][3] = {{0., 0., 0.}, {0., 0., 0.}, {0., 0., 0.}};   // du_c/dx_d
            for (int n = 0; n < NF; n++) {
                double gx[DIM];
#pragma unroll
                for (int i = 0; i < DIM; i++) {
                    double t = 0.;
#pragma unroll
                    for (int a = 0; a < DIM; a++) t += inv[i][a] * tab.fd[n * DIM + a];
                    gx[i] = t;
                }
                const int64_t node = fconn[e * NF + n];
#pragma unroll
                for (int c = 0; c < S; c++) {
                    const double v = u[node * S + c];
#pragma unroll
                    for (int d = 0; d < DIM; d++) G[d][c] += gx[d] * v;
                }
            }
            if constexpr (QTY == INS_CELL_GRADIENT) {
#pragma unroll
                for (int d = 0; d < DIM; d++)
#pragma unroll
                    for (int c = 0; c < S; c++) o[d * S + c] = G[d][c];
            } else {
                double F[3][3], Sg[3][3], Cd[6][6];
#pragma unroll
                for (int i = 0; i < 3; i++)
#pragma unroll
                    for (int j = 0; j < 3; j++) F[i][j] = (i == j ? 1. : 0.) + G[j][i];
                material<MAT, double>(coef[e * ncoef], coef[e * ncoef + 1], F, Sg, Cd, false);
                if constexpr (QTY == INS_CELL_PK2) {
#pragma unroll
                    for (int i = 0; i < 3; i++)
#pragma unroll
                        for (int j = 0; j < 3; j++) o[i * 3 + j] = Sg[i][j];
                } else {
                    const double J = F[0][0] * (F[1][1] * F[2][2] - F[1][2] * F[2][1]) -
                                     F[0][1] * (F[1][0] * F[2][2] - F[1][2] * F[2][0]) +
                                     F[0][2] * (F[1][0] * F[2][1] - F[1][1] * F[2][0]);
                    double FS[3][3];
#pragma unroll
                    for (int i = 0; i < 3; i++)
#pragma unroll
                        for (int j = 0; j < 3; j++) {
                            double t = 0.;
#pragma unroll
                            for (int k = 0; k < 3; k++) t += F[i][k] * Sg[k][j];
                            FS[i][j] = t;
                        }
#pragma unroll
                    for (int i = 0; i < 3; i++)
#pragma unroll
                        for (int j = 0; j < 3; j++) {
                            double t = 0.;
#pragma unroll
                            for (int k = 0; k < 3; k++) t += FS[i][k] * F[j][k];
                            o[i * 3 + j] = t / J;
                        }
                }
            }
        }
#pragma unroll
        for (int k = 0; k < PER; k++) stage[threadIdx.x * PER + k] = o[k];
    }
    __syncthreads();
    const int64_t cnt = (E - first < 256 ? E - first : 256) * PER;
    for (int64_t i = threadIdx.x; i < cnt; i += 256) out[first * PER + i] = stage[i];
}

template <int DIM, int NF, int S, int QTY, int MAT>
static void launch_cell(Ctx* c, const CellTab& tab, const double* u, double* out)
{
    const unsigned g = (unsigned)((c->n_elem + 255) / 256);
    k_cell_eval<DIM, NF, S, QTY, MAT><<<g, 256, 0, c->stream>>>(
        c->n_elem, c->geo_conn(), c->conn.p, c->xyz.p, u, c->coef.p, c->ncoef, tab, out);
}

template <int DIM, int NF, int S>
static bool dispatch_cell(Ctx* c, int qty, const CellTab& tab, const double* u, double* out)
{
    switch (qty) {
    case INS_CELL_VALUE: launch_cell<DIM, NF, S, INS_CELL_VALUE, 0>(c, tab, u, out); return true;
    case INS_CELL_GRADIENT:
        launch_cell<DIM, NF, S, INS_CELL_GRADIENT, 0>(c, tab, u, out);
        return true;
    default: break;
    }
    if constexpr (S == DIM) {
        const bool sv = c->kernel == INS_STVENANT;
        if (qty == INS_CELL_PK2) {
            if (sv) launch_cell<DIM, NF, S, INS_CELL_PK2, INS_STVENANT>(c, tab, u, out);
            else launch_cell<DIM, NF, S, INS_CELL_PK2, INS_NEOHOOKE>(c, tab, u, out);
        } else {
            if (sv) launch_cell<DIM, NF, S, INS_CELL_CAUCHY, INS_STVENANT>(c, tab, u, out);
            else launch_cell<DIM, NF, S, INS_CELL_CAUCHY, INS_NEOHOOKE>(c, tab, u, out);
        }
        return true;
    }
    return false;
}

int cell_eval(Ctx* c, int qty, const double* hxi, const double* hu, double* hout)
{
    if (!c->have_dofs) return fail(c, INS_ERR_STATE, "ins_cell_eval before ins_set_dofs");
    if (qty < INS_CELL_VALUE || qty > INS_CELL_CAUCHY || !hout)
        return fail(c, INS_ERR_ARG, "ins_cell_eval arguments");
    const int dim = c->dim, nf = c->npe, s = c->dpn;
    if (qty >= INS_CELL_PK2) {
        if (c->kernel != INS_STVENANT && c->kernel != INS_NEOHOOKE)
            return fail(c, INS_ERR_STATE, "stresses need a solid material (ins_set_material)");
        if (s != dim || c->ncoef < 2 || !c->coef.p)
            return fail(c, INS_ERR_ARG, "stresses need dpn = dim and lambda, mu per element");
    }
    double xi[3] = {0.5, 0.5, 0.5};
    if (hxi)
        for (int d = 0; d < dim; d++) {
            if (!(hxi[d] >= 0. && hxi[d] <= 1.)) return fail(c, INS_ERR_ARG, "xi outside [0,1]");
            xi[d] = hxi[d];
        }
    CellTab tab;
    std::memset(&tab, 0, sizeof(tab));
    {
        double phi[8], dphi[24];
        q1_shape(dim, xi, phi, dphi);
        for (int i = 0; i < (1 << dim) * dim; i++) tab.gd[i] = dphi[i];
        if (c->deg == 2) {
            lagrange2_shape(dim, xi, tab.fp, tab.fd);
        } else {
            for (int n = 0; n < (1 << dim); n++) tab.fp[n] = phi[n];
            for (int i = 0; i < (1 << dim) * dim; i++) tab.fd[i] = dphi[i];
        }
    }
    const int per = qty == INS_CELL_VALUE ? s : qty == INS_CELL_GRADIENT ? dim * s : 9;
    const int64_t nd = c->n_nodes * s;
    const double* u = c->u.p;
    double* tmp = nullptr;
    double* dout = nullptr;
    INS_HIP_CHECK(c, hipMallocAsync((void**)&dout, (size_t)c->n_elem * per * sizeof(double),
                                    c->stream));
    if (hu) {
        INS_HIP_CHECK(c, hipMallocAsync((void**)&tmp, nd * sizeof(double), c->stream));
        INS_HIP_CHECK(c, hipMemcpyAsync(tmp, hu, nd * sizeof(double), hipMemcpyHostToDevice,
                                        c->stream));
        u = tmp;
    }
    bool ok = false;
    const int key = dim * 1000 + nf * 10 + s;
    switch (key) {
    case 2041: ok = dispatch_cell<2, 4, 1>(c, qty, tab, u, dout); break;
    case 2042: ok = dispatch_cell<2, 4, 2>(c, qty, tab, u, dout); break;
    case 3081: ok = dispatch_cell<3, 8, 1>(c, qty, tab, u, dout); break;
    case 3083: ok = dispatch_cell<3, 8, 3>(c, qty, tab, u, dout); break;
    case 2091: ok = dispatch_cell<2, 9, 1>(c, qty, tab, u, dout); break;
    case 2092: ok = dispatch_cell<2, 9, 2>(c, qty, tab, u, dout); break;
    case 3271: ok = dispatch_cell<3, 27, 1>(c, qty, tab, u, dout); break;
    case 3273: ok = dispatch_cell<3, 27, 3>(c, qty, tab, u, dout); break;
    default: break;
    }
    if (ok) {
        INS_HIP_CHECK(c, hipGetLastError());
        INS_HIP_CHECK(c, hipMemcpyAsync(hout, dout, (size_t)c->n_elem * per * sizeof(double),
                                        hipMemcpyDeviceToHost, c->stream));
    }
    (void)hipFreeAsync(dout, c->stream);
    if (tmp) (void)hipFreeAsync(tmp, c->stream);
    INS_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    if (!ok) return fail(c, INS_ERR_UNSUPPORTED, "element/dof combination");
    return INS_OK;
}

}  // namespace ins
