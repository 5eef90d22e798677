// Jacobi-preconditioned CG on gfx950 (Eigen::ConjugateGradient semantics as
// called by base/solver/Eigen3.hpp:265-275: x0 = 0, z = D^-1 r, stop when
// |r|^2 < rtol^2 |b|^2 on the recursive residual).
//
// Every iteration is three kernels that stream HBM once each:
//   K1  q = A p fused with the p.q reduction            -> alpha
//   K2  x += alpha p, r -= alpha q fused with r.r, r.D^-1 r -> beta, stop test
//   K3  p = D^-1 r + beta p
// Reductions are deterministic: each block reduces its rows in a fixed tree,
// writes one partial, and the last block to finish (integer ticket) sums the
// partials in block order and updates the device-side scalars.  No float
// atomics anywhere.  The host only synchronises once per chunk of iterations.
//
// Mixed precision (INS_SOLVE_MIXED): FP32 matrix values and FP32 vectors in
// K1-K3 with FP32 accumulation in the SpMV and FP64 accumulation of all dot
// products; at the end of every chunk the FP32 iterate is folded into the
// FP64 solution and the residual is replaced by the FP64 residual b - A x
// (FP64 values, FP64 arithmetic).  Convergence is only declared on that FP64
// residual.
#include <algorithm>
#include <cmath>

#include "common.hpp"

namespace ins {

namespace {

constexpr int RB = 256;       // threads per block of vector kernels
constexpr int MAXG = 2048;    // max blocks of a reducing kernel
constexpr int CHUNK = 40;     // iterations between host checks

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// sums NV values over the block; result valid in thread 0
template <int NV>
__device__ inline void block_sum(double (&v)[NV])
{
    __shared__ double sh[NV][RB / 64];
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] = wave_sum(v[k]);
    const int wave = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int k = 0; k < NV; k++) sh[k][wave] = v[k];
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) {
            double s = 0.;
            for (int w = 0; w < (int)(blockDim.x >> 6); w++) s += sh[k][w];
            v[k] = s;
        }
    }
}

// Deterministic grid reduction: returns true in thread 0 of the last block,
// with tot[] = sum over blocks in block order.
template <int NV>
__device__ inline bool grid_sum(double (&v)[NV], double* partial, unsigned* counter,
                                double (&tot)[NV])
{
    block_sum<NV>(v);
    __shared__ int amlast;
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) partial[k * MAXG + blockIdx.x] = v[k];
        __threadfence();
        const unsigned t = atomicAdd(counter, 1u);
        amlast = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!amlast) return false;
    __threadfence();
    double s[NV];
#pragma unroll
    for (int k = 0; k < NV; k++) s[k] = 0.;
    for (int i = threadIdx.x; i < (int)gridDim.x; i += blockDim.x)
#pragma unroll
        for (int k = 0; k < NV; k++) s[k] += partial[k * MAXG + i];
    block_sum<NV>(s);
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NV; k++) tot[k] = s[k];
        *counter = 0u;
    }
    return threadIdx.x == 0;
}

// ---------------------------------------------------------------------------
// CSR SpMV, LPR lanes per row; optional fused p.q reduction
template <int LPR, typename TV, typename TX, typename TA>
__device__ inline TA csr_row(int64_t row, int sub, const int32_t* __restrict__ rp,
                             const int32_t* __restrict__ col, const TV* __restrict__ val,
                             const TX* __restrict__ x)
{
    TA s = (TA)0;
    const int32_t e = rp[row + 1];
    for (int32_t k = rp[row] + sub; k < e; k += LPR) s += (TA)val[k] * (TA)x[col[k]];
#pragma unroll
    for (int o = LPR / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, LPR);
    return s;
}

// SELL-64 row: one lane per row, 64 rows per wavefront
template <typename TV, typename TX, typename TA>
__device__ inline TA sell_row(int64_t row, int64_t n, int lane, int64_t slice,
                              const int64_t* __restrict__ sptr, const int32_t* __restrict__ slen,
                              const int32_t* __restrict__ rpat, const int32_t* __restrict__ pptr,
                              const int32_t* __restrict__ plen, const int32_t* __restrict__ pdel,
                              const TV* __restrict__ v, const TX* __restrict__ x)
{
    TA s = (TA)0;
    const int64_t base = sptr[slice] + lane;
    const int L = slen[slice];
    const int pid = row < n ? rpat[row] : 0;
    const int uni = __builtin_amdgcn_readfirstlane(pid);
    if (__all(row >= n || pid == uni)) {
        // uniform slice: the offset list is wave-uniform
        const int32_t* d = pdel + pptr[uni];
        const int len = plen[uni];
        if (row < n)
            for (int k = 0; k < len; k++) s += (TA)v[base + (int64_t)k * 64] * (TA)x[row + d[k]];
    } else if (row < n) {
        const int32_t* d = pdel + pptr[pid];
        const int len = plen[pid];
        for (int k = 0; k < len; k++) s += (TA)v[base + (int64_t)k * 64] * (TA)x[row + d[k]];
    }
    (void)L;
    return s;
}

struct SellView {
    const int64_t* sptr;
    const int32_t* slen;
    const int32_t* rpat;
    const int32_t* pptr;
    const int32_t* plen;
    const int32_t* pdel;
};

// K1: q = A p ; pq -> alpha
template <int LPR, typename TV, typename TX, typename TA>
__global__ __launch_bounds__(RB) void k_spmv_pq_csr(int64_t n, const int32_t* __restrict__ rp,
                                                    const int32_t* __restrict__ col,
                                                    const TV* __restrict__ val,
                                                    const TX* __restrict__ p, TX* __restrict__ q,
                                                    double* partial, unsigned* counter, Scal* sc)
{
    if (sc->done) return;
    const int sub = threadIdx.x & (LPR - 1);
    const int64_t stride = (int64_t)gridDim.x * (RB / LPR);
    double pq = 0.;
    for (int64_t row = ((int64_t)blockIdx.x * RB + threadIdx.x) / LPR; row < n; row += stride) {
        const TA s = csr_row<LPR, TV, TX, TA>(row, sub, rp, col, val, p);
        if (sub == 0) {
            const TX qs = (TX)s;
            q[row] = qs;
            pq += (double)p[row] * (double)qs;
        }
    }
    double v[1] = {pq}, tot[1];
    if (grid_sum<1>(v, partial, counter, tot)) {
        sc->pq = tot[0];
        if (!(tot[0] > 0.) || !isfinite(tot[0])) {
            sc->breakdown = 1;
            sc->done = 1;
        } else {
            sc->alpha = sc->rho / tot[0];
        }
    }
}

template <typename TV, typename TX, typename TA>
__global__ __launch_bounds__(RB) void k_spmv_pq_sell(int64_t n, SellView S,
                                                     const TV* __restrict__ val,
                                                     const TX* __restrict__ p, TX* __restrict__ q,
                                                     double* partial, unsigned* counter, Scal* sc)
{
    if (sc->done) return;
    const int lane = threadIdx.x & 63;
    const int64_t nsl = (n + 63) / 64;
    const int64_t wstride = (int64_t)gridDim.x * (RB / 64);
    double pq = 0.;
    for (int64_t slice = ((int64_t)blockIdx.x * RB + threadIdx.x) / 64; slice < nsl;
         slice += wstride) {
        const int64_t row = slice * 64 + lane;
        const TA s = sell_row<TV, TX, TA>(row, n, lane, slice, S.sptr, S.slen, S.rpat, S.pptr,
                                          S.plen, S.pdel, val, p);
        if (row < n) {
            const TX qs = (TX)s;
            q[row] = qs;
            pq += (double)p[row] * (double)qs;
        }
    }
    double v[1] = {pq}, tot[1];
    if (grid_sum<1>(v, partial, counter, tot)) {
        sc->pq = tot[0];
        if (!(tot[0] > 0.) || !isfinite(tot[0])) {
            sc->breakdown = 1;
            sc->done = 1;
        } else {
            sc->alpha = sc->rho / tot[0];
        }
    }
}

// K2: x += alpha p ; r -= alpha q ; rr, rz -> stop test, beta
template <typename TX, typename TD>
__global__ __launch_bounds__(RB) void k_update_xr(int64_t n, TX* __restrict__ x, TX* __restrict__ r,
                                                  const TX* __restrict__ p,
                                                  const TX* __restrict__ q,
                                                  const TD* __restrict__ dinv, double* partial,
                                                  unsigned* counter, Scal* sc)
{
    if (sc->done) return;
    const TX alpha = (TX)sc->alpha;
    double rr = 0., rz = 0.;
    for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * RB) {
        x[i] += alpha * p[i];
        const TX ri = r[i] - alpha * q[i];
        r[i] = ri;
        const TX zi = (TX)dinv[i] * ri;
        rr += (double)ri * (double)ri;
        rz += (double)ri * (double)zi;
    }
    double v[2] = {rr, rz}, tot[2];
    if (grid_sum<2>(v, partial, counter, tot)) {
        sc->rr = tot[0];
        sc->iters += 1;
        if (!isfinite(tot[0])) {
            sc->breakdown = 1;
            sc->done = 1;
        } else if (tot[0] < sc->thr) {
            sc->done = 1;
        } else {
            sc->beta = tot[1] / sc->rho;
            sc->rho = tot[1];
        }
    }
}

// K3: p = D^-1 r + beta p
template <typename TX, typename TD>
__global__ __launch_bounds__(RB) void k_update_p(int64_t n, TX* __restrict__ p,
                                                 const TX* __restrict__ r,
                                                 const TD* __restrict__ dinv, const Scal* sc)
{
    if (sc->done) return;
    const TX beta = (TX)sc->beta;
    for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * RB)
        p[i] = (TX)dinv[i] * r[i] + beta * p[i];
}

// start: x = 0, r = b, p = D^-1 b, rho = r.p, rr, bb
template <typename TX, typename TD>
__global__ __launch_bounds__(RB) void k_init(int64_t n, const double* __restrict__ b,
                                             TX* __restrict__ x, TX* __restrict__ r,
                                             TX* __restrict__ p, const TD* __restrict__ dinv,
                                             double rtol2, double* partial, unsigned* counter,
                                             Scal* sc)
{
    double rz = 0., rr = 0.;
    for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * RB) {
        const TX bi = (TX)b[i];
        x[i] = (TX)0;
        r[i] = bi;
        const TX zi = (TX)dinv[i] * bi;
        p[i] = zi;
        rr += (double)b[i] * (double)b[i];
        rz += (double)bi * (double)zi;
    }
    double v[2] = {rr, rz}, tot[2];
    if (grid_sum<2>(v, partial, counter, tot)) {
        sc->bb = tot[0];
        sc->rr = tot[0];
        sc->rho = tot[1];
        sc->thr = rtol2 * tot[0];
        sc->iters = 0;
        sc->breakdown = 0;
        sc->done = (tot[0] == 0. || tot[0] < sc->thr) ? 1 : 0;
    }
}

// ---- mixed precision: residual replacement ------------------------------------
// x64 += x32, x32 = 0
__global__ __launch_bounds__(RB) void k_fold(int64_t n, double* __restrict__ x64,
                                             float* __restrict__ x32)
{
    for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * RB) {
        x64[i] += (double)x32[i];
        x32[i] = 0.f;
    }
}

// t = b - A x (FP64), tt = |t|^2; optionally r32 = (float)t and rz
template <int LPR, typename TV>
__global__ __launch_bounds__(RB) void k_resid_csr(int64_t n, const int32_t* __restrict__ rp,
                                                  const int32_t* __restrict__ col,
                                                  const TV* __restrict__ val,
                                                  const double* __restrict__ x,
                                                  const double* __restrict__ b,
                                                  double* __restrict__ t, float* __restrict__ r32,
                                                  const float* __restrict__ dinv32,
                                                  double* partial, unsigned* counter, Scal* sc,
                                                  int replace)
{
    const int sub = threadIdx.x & (LPR - 1);
    const int64_t stride = (int64_t)gridDim.x * (RB / LPR);
    double tt = 0., rz = 0.;
    for (int64_t row = ((int64_t)blockIdx.x * RB + threadIdx.x) / LPR; row < n; row += stride) {
        const double s = csr_row<LPR, TV, double, double>(row, sub, rp, col, val, x);
        if (sub == 0) {
            const double ti = b[row] - s;
            if (t) t[row] = ti;
            tt += ti * ti;
            if (replace) {
                const float rf = (float)ti;
                r32[row] = rf;
                rz += (double)rf * (double)(dinv32[row] * rf);
            }
        }
    }
    double v[2] = {tt, rz}, tot[2];
    if (grid_sum<2>(v, partial, counter, tot)) {
        sc->tt = tot[0];
        if (replace) {
            if (tot[0] < sc->thr) {
                sc->done = 1;
            } else {
                sc->done = 0;
                sc->rho = tot[1];
            }
        }
    }
}

template <typename TV>
__global__ __launch_bounds__(RB) void k_resid_sell(int64_t n, SellView S, const TV* __restrict__ val,
                                                   const double* __restrict__ x,
                                                   const double* __restrict__ b,
                                                   double* __restrict__ t, float* __restrict__ r32,
                                                   const float* __restrict__ dinv32,
                                                   double* partial, unsigned* counter, Scal* sc,
                                                   int replace)
{
    const int lane = threadIdx.x & 63;
    const int64_t nsl = (n + 63) / 64;
    const int64_t wstride = (int64_t)gridDim.x * (RB / 64);
    double tt = 0., rz = 0.;
    for (int64_t slice = ((int64_t)blockIdx.x * RB + threadIdx.x) / 64; slice < nsl;
         slice += wstride) {
        const int64_t row = slice * 64 + lane;
        const double s = sell_row<TV, double, double>(row, n, lane, slice, S.sptr, S.slen, S.rpat,
                                                      S.pptr, S.plen, S.pdel, val, x);
        if (row < n) {
            const double ti = b[row] - s;
            if (t) t[row] = ti;
            tt += ti * ti;
            if (replace) {
                const float rf = (float)ti;
                r32[row] = rf;
                rz += (double)rf * (double)(dinv32[row] * rf);
            }
        }
    }
    double v[2] = {tt, rz}, tot[2];
    if (grid_sum<2>(v, partial, counter, tot)) {
        sc->tt = tot[0];
        if (replace) {
            if (tot[0] < sc->thr) {
                sc->done = 1;
            } else {
                sc->done = 0;
                sc->rho = tot[1];
            }
        }
    }
}

// ---- SELL construction --------------------------------------------------------
__device__ inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

__global__ void k_row_hash(int64_t n, const int32_t* __restrict__ rp,
                           const int32_t* __restrict__ col, uint64_t* __restrict__ h)
{
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
         r += (int64_t)gridDim.x * blockDim.x) {
        uint64_t z = mix64(0x9e3779b97f4a7c15ULL + (uint64_t)(rp[r + 1] - rp[r]));
        for (int32_t k = rp[r]; k < rp[r + 1]; k++)
            z = mix64(z ^ (uint64_t)(uint32_t)(col[k] - (int32_t)r) * 0x100000001b3ULL);
        h[r] = z | 1ULL;  // never 0 (0 marks an empty slot)
    }
}

// open-addressing table of distinct row hashes; value = a representative row
__global__ void k_hash_insert(int64_t n, const uint64_t* __restrict__ h, uint64_t* keys,
                              int32_t* reps, uint64_t mask, int* err)
{
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
         r += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t key = h[r];
        uint64_t slot = key & mask;
        for (uint64_t probe = 0; probe <= mask; probe++) {
            const unsigned long long prev =
                atomicCAS((unsigned long long*)&keys[slot], 0ULL, (unsigned long long)key);
            if (prev == 0ULL) {
                reps[slot] = (int32_t)r;
                break;
            }
            if (prev == key) break;
            slot = (slot + 1) & mask;
            if (probe == mask) atomicOr(err, 1);
        }
    }
}

__global__ void k_hash_find(int64_t n, const uint64_t* __restrict__ h,
                            const uint64_t* __restrict__ keys, uint64_t mask,
                            int32_t* __restrict__ slot_of)
{
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
         r += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t key = h[r];
        uint64_t slot = key & mask;
        while (keys[slot] != key) slot = (slot + 1) & mask;
        slot_of[r] = (int32_t)slot;
    }
}

__global__ void k_mark_slots(uint64_t cap, const uint64_t* __restrict__ keys,
                             int32_t* __restrict__ used)
{
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < (int64_t)cap;
         s += (int64_t)gridDim.x * blockDim.x)
        used[s] = keys[s] != 0ULL;
}

// pattern id of every row and check that the row equals its representative
__global__ void k_row_pattern(int64_t n, const int32_t* __restrict__ rp,
                              const int32_t* __restrict__ col, const int32_t* __restrict__ slot_of,
                              const int32_t* __restrict__ slot_id,
                              const int32_t* __restrict__ reps, int32_t* __restrict__ row_pat,
                              int* err)
{
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int32_t s = slot_of[r];
        const int64_t rep = reps[s];
        row_pat[r] = slot_id[s];
        const int32_t len = rp[r + 1] - rp[r];
        if (rp[rep + 1] - rp[rep] != len) {
            atomicOr(err, 2);
            continue;
        }
        for (int32_t k = 0; k < len; k++)
            if (col[rp[r] + k] - (int32_t)r != col[rp[rep] + k] - (int32_t)rep) {
                atomicOr(err, 2);
                break;
            }
    }
}

__global__ void k_pat_len(uint64_t cap, const uint64_t* __restrict__ keys,
                          const int32_t* __restrict__ slot_id, const int32_t* __restrict__ reps,
                          const int32_t* __restrict__ rp, int32_t* __restrict__ plen)
{
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < (int64_t)cap;
         s += (int64_t)gridDim.x * blockDim.x)
        if (keys[s]) {
            const int64_t rep = reps[s];
            plen[slot_id[s]] = rp[rep + 1] - rp[rep];
        }
}

__global__ void k_pat_fill(uint64_t cap, const uint64_t* __restrict__ keys,
                           const int32_t* __restrict__ slot_id, const int32_t* __restrict__ reps,
                           const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                           const int32_t* __restrict__ pptr, int32_t* __restrict__ pdel)
{
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < (int64_t)cap;
         s += (int64_t)gridDim.x * blockDim.x)
        if (keys[s]) {
            const int64_t rep = reps[s];
            const int32_t o = pptr[slot_id[s]];
            for (int32_t k = rp[rep]; k < rp[rep + 1]; k++) pdel[o + (k - rp[rep])] = col[k] - (int32_t)rep;
        }
}

__global__ void k_slice_len(int64_t n, const int32_t* __restrict__ rp, int32_t* __restrict__ slen)
{
    const int64_t nsl = (n + 63) / 64;
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < nsl;
         s += (int64_t)gridDim.x * blockDim.x) {
        int32_t m = 0;
        for (int64_t r = s * 64; r < n && r < s * 64 + 64; r++) m = max(m, rp[r + 1] - rp[r]);
        slen[s] = m;
    }
}

__global__ void k_slice_ptr(int64_t nsl, const int32_t* __restrict__ slen, int64_t* __restrict__ sptr)
{
    // single-thread scan over slices (nsl <= n/64 is small)
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        int64_t acc = 0;
        for (int64_t s = 0; s < nsl; s++) {
            sptr[s] = acc;
            acc += (int64_t)slen[s] * 64;
        }
        sptr[nsl] = acc;
    }
}

// scatter CSR values into slices (zero padding)
__global__ void k_sell_fill(int64_t n, const int32_t* __restrict__ rp,
                            const double* __restrict__ v64, const float* __restrict__ v32,
                            const int64_t* __restrict__ sptr, const int32_t* __restrict__ slen,
                            double* __restrict__ s64, float* __restrict__ s32)
{
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t slice = r / 64;
        const int lane = (int)(r % 64);
        const int64_t base = sptr[slice] + lane;
        const int32_t len = rp[r + 1] - rp[r];
        const int32_t L = slen[slice];
        for (int32_t k = 0; k < L; k++) {
            const int64_t o = base + (int64_t)k * 64;
            if (k < len) {
                if (s64) s64[o] = v64 ? v64[rp[r] + k] : (double)v32[rp[r] + k];
                if (s32) s32[o] = v32[rp[r] + k];
            } else {
                if (s64) s64[o] = 0.;
                if (s32) s32[o] = 0.f;
            }
        }
    }
}

// setDoFsFromSolver / addToDoFsFromSolver (base/dof/Distribute.hpp:141-215)
__global__ void k_distribute(int64_t nd, const int32_t* __restrict__ dof_id,
                             const double* __restrict__ x, const double* __restrict__ g,
                             double* __restrict__ u, int add)
{
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nd;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int32_t id = dof_id[k];
        if (id >= 0) u[k] = add ? u[k] + x[id] : x[id];
        else u[k] = g[k];
    }
}

__global__ __launch_bounds__(RB) void k_norm2(int64_t n, const double* __restrict__ v,
                                              double* partial, unsigned* counter, Scal* sc)
{
    double s = 0.;
    for (int64_t i = (int64_t)blockIdx.x * RB + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * RB)
        s += v[i] * v[i];
    double a[1] = {s}, tot[1];
    if (grid_sum<1>(a, partial, counter, tot)) sc->tt = tot[0];
}

unsigned red_grid(int64_t n, int rows_per_block)
{
    int64_t g = (n + rows_per_block - 1) / rows_per_block;
    if (g < 1) g = 1;
    if (g > MAXG) g = MAXG;
    return (unsigned)g;
}

int ensure_solver_state(Ctx* c, int mode)
{
    const int64_t n = c->n_active;
    INS_TRY(ensure(c, c->x, n));
    INS_TRY(ensure(c, c->r, n));
    INS_TRY(ensure(c, c->p, n));
    INS_TRY(ensure(c, c->q, n));
    if (mode == INS_SOLVE_MIXED) {
        INS_TRY(ensure(c, c->x32, n));
        INS_TRY(ensure(c, c->r32, n));
        INS_TRY(ensure(c, c->p32, n));
        INS_TRY(ensure(c, c->q32, n));
    }
    if (!c->partial.p) INS_TRY(ensure(c, c->partial, (size_t)4 * MAXG));
    if (!c->counter.p) {
        INS_TRY(ensure(c, c->counter, 4));
        INS_HIP_CHECK(c, hipMemsetAsync(c->counter.p, 0, 4 * sizeof(unsigned), c->stream));
    }
    if (!c->scal.p) INS_TRY(ensure(c, c->scal, 1));
    if (!c->h_scal) INS_HIP_CHECK(c, hipHostMalloc((void**)&c->h_scal, sizeof(Scal), 0));
    return INS_OK;
}

SellView sell_view(Ctx* c)
{
    SellView v;
    v.sptr = c->sell.slice_ptr.p;
    v.slen = c->sell.slice_len.p;
    v.rpat = c->sell.row_pat.p;
    v.pptr = c->sell.pat_ptr.p;
    v.plen = c->sell.pat_len.p;
    v.pdel = c->sell.pat_delta.p;
    return v;
}

// lanes per row of the CSR kernels, from the mean row length
int pick_lpr(Ctx* c)
{
    const double avg = c->n_active ? (double)c->nnz / (double)c->n_active : 1.;
    if (avg > 48) return 16;
    if (avg > 12) return 8;
    return 4;
}

template <typename TV, typename TX, typename TA>
int launch_spmv_pq(Ctx* c, int format, const TV* csr_val, const TV* sell_val, const TX* p, TX* q)
{
    const int64_t n = c->n_active;
    if (format == INS_FMT_SELL) {
        const unsigned g = red_grid((n + 63) / 64, RB / 64);
        k_spmv_pq_sell<TV, TX, TA><<<g, RB, 0, c->stream>>>(n, sell_view(c), sell_val, p, q,
                                                           c->partial.p, c->counter.p, c->scal.p);
    } else {
        const int lpr = pick_lpr(c);
        const unsigned g = red_grid(n, RB / lpr);
        if (lpr == 16)
            k_spmv_pq_csr<16, TV, TX, TA><<<g, RB, 0, c->stream>>>(n, c->row_ptr.p, c->col.p, csr_val, p, q, c->partial.p, c->counter.p, c->scal.p);
        else if (lpr == 8)
            k_spmv_pq_csr<8, TV, TX, TA><<<g, RB, 0, c->stream>>>(n, c->row_ptr.p, c->col.p, csr_val, p, q, c->partial.p, c->counter.p, c->scal.p);
        else
            k_spmv_pq_csr<4, TV, TX, TA><<<g, RB, 0, c->stream>>>(n, c->row_ptr.p, c->col.p, csr_val, p, q, c->partial.p, c->counter.p, c->scal.p);
    }
    INS_HIP_CHECK(c, hipGetLastError());
    return INS_OK;
}

int launch_resid(Ctx* c, int format, int replace, double* t)
{
    const int64_t n = c->n_active;
    if (format == INS_FMT_SELL) {
        const unsigned g = red_grid((n + 63) / 64, RB / 64);
        if (c->matrix_precision == INS_FP64)
            k_resid_sell<double><<<g, RB, 0, c->stream>>>(n, sell_view(c), c->sell.v64.p, c->x.p, c->b.p, t,
                                                          c->r32.p, c->dinv32.p, c->partial.p,
                                                          c->counter.p, c->scal.p, replace);
        else
            k_resid_sell<float><<<g, RB, 0, c->stream>>>(n, sell_view(c), c->sell.v32.p, c->x.p, c->b.p, t,
                                                         c->r32.p, c->dinv32.p, c->partial.p,
                                                         c->counter.p, c->scal.p, replace);
    } else {
        const int lpr = pick_lpr(c);
        const unsigned g = red_grid(n, RB / lpr);
        if (c->matrix_precision == INS_FP64) {
            if (lpr == 16) k_resid_csr<16, double><<<g, RB, 0, c->stream>>>(n, c->row_ptr.p, c->col.p, c->val64.p, c->x.p, c->b.p, t, c->r32.p, c->dinv32.p, c->partial.p, c->counter.p, c->scal.p, replace);
            else if (lpr == 8) k_resid_csr<8, double><<<g, RB, 0, c->stream>>>(n, c->row_ptr.p, c->col.p, c->val64.p, c->x.p, c->b.p, t, c->r32.p, c->dinv32.p, c->partial.p, c->counter.p, c->scal.p, replace);
            else k_resid_csr<4, double><<<g, RB, 0, c->stream>>>(n, c->row_ptr.p, c->col.p, c->val64.p, c->x.p, c->b.p, t, c->r32.p, c->dinv32.p, c->partial.p, c->counter.p, c->scal.p, replace);
        } else {
            if (lpr == 16) k_resid_csr<16, float><<<g, RB, 0, c->stream>>>(n, c->row_ptr.p, c->col.p, c->val32.p, c->x.p, c->b.p, t, c->r32.p, c->dinv32.p, c->partial.p, c->counter.p, c->scal.p, replace);
            else if (lpr == 8) k_resid_csr<8, float><<<g, RB, 0, c->stream>>>(n, c->row_ptr.p, c->col.p, c->val32.p, c->x.p, c->b.p, t, c->r32.p, c->dinv32.p, c->partial.p, c->counter.p, c->scal.p, replace);
            else k_resid_csr<4, float><<<g, RB, 0, c->stream>>>(n, c->row_ptr.p, c->col.p, c->val32.p, c->x.p, c->b.p, t, c->r32.p, c->dinv32.p, c->partial.p, c->counter.p, c->scal.p, replace);
        }
    }
    INS_HIP_CHECK(c, hipGetLastError());
    return INS_OK;
}

int read_scal(Ctx* c)
{
    INS_HIP_CHECK(c, hipMemcpyAsync(c->h_scal, c->scal.p, sizeof(Scal), hipMemcpyDeviceToHost,
                                    c->stream));
    INS_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return INS_OK;
}

// profiling of the dominant kernel (the SpMV of K1)
struct ProfScope {
    Ctx* c;
    bool on;
    ProfScope(Ctx* c_) : c(c_), on(c_->profiling)
    {
        if (on) hipEventRecord(c->evp0, c->stream);
    }
    ~ProfScope()
    {
        if (on) {
            hipEventRecord(c->evp1, c->stream);
            hipEventSynchronize(c->evp1);
            float ms = 0.f;
            hipEventElapsedTime(&ms, c->evp0, c->evp1);
            c->stats.spmv_ms += ms;
            c->stats.spmv_calls += 1;
        }
    }
};

}  // namespace

int build_sell(Ctx* c)
{
    Sell& S = c->sell;
    const int64_t n = c->n_active;
    S.n = n;
    uint64_t* h = nullptr;
    uint64_t* keys = nullptr;
    int32_t *reps = nullptr, *slot_of = nullptr, *used = nullptr, *slot_id = nullptr;
    int* err = nullptr;
    uint64_t cap = 1024;
    while (cap < (uint64_t)(2 * n + 2)) cap <<= 1;
    INS_HIP_CHECK(c, hipMalloc((void**)&h, n * sizeof(uint64_t) + 8));
    INS_HIP_CHECK(c, hipMalloc((void**)&keys, cap * sizeof(uint64_t)));
    INS_HIP_CHECK(c, hipMalloc((void**)&reps, cap * sizeof(int32_t)));
    INS_HIP_CHECK(c, hipMalloc((void**)&slot_of, n * sizeof(int32_t) + 8));
    INS_HIP_CHECK(c, hipMalloc((void**)&used, (cap + 1) * sizeof(int32_t)));
    INS_HIP_CHECK(c, hipMalloc((void**)&slot_id, (cap + 1) * sizeof(int32_t)));
    INS_HIP_CHECK(c, hipMalloc((void**)&err, sizeof(int)));
    INS_HIP_CHECK(c, hipMemsetAsync(keys, 0, cap * sizeof(uint64_t), c->stream));
    INS_HIP_CHECK(c, hipMemsetAsync(err, 0, sizeof(int), c->stream));
    const unsigned g = grid_for(n, 256);
    k_row_hash<<<g, 256, 0, c->stream>>>(n, c->row_ptr.p, c->col.p, h);
    k_hash_insert<<<g, 256, 0, c->stream>>>(n, h, keys, reps, cap - 1, err);
    k_hash_find<<<g, 256, 0, c->stream>>>(n, h, keys, cap - 1, slot_of);
    k_mark_slots<<<grid_for(cap, 256), 256, 0, c->stream>>>(cap, keys, used);
    INS_HIP_CHECK(c, hipGetLastError());
    int64_t npat = 0;
    int s = exclusive_scan_i32(c, used, slot_id, (int64_t)cap, &npat);
    if (s != INS_OK) return s;
    S.npat = (int)npat;
    INS_TRY(ensure(c, S.row_pat, n));
    k_row_pattern<<<g, 256, 0, c->stream>>>(n, c->row_ptr.p, c->col.p, slot_of, slot_id, reps,
                                            S.row_pat.p, err);
    INS_TRY(ensure(c, S.pat_len, npat + 1));
    INS_TRY(ensure(c, S.pat_ptr, npat + 1));
    k_pat_len<<<grid_for(cap, 256), 256, 0, c->stream>>>(cap, keys, slot_id, reps, c->row_ptr.p,
                                                         S.pat_len.p);
    INS_HIP_CHECK(c, hipGetLastError());
    int64_t ndel = 0;
    s = exclusive_scan_i32(c, S.pat_len.p, S.pat_ptr.p, npat, &ndel);
    if (s != INS_OK) return s;
    INS_TRY(ensure(c, S.pat_delta, ndel));
    k_pat_fill<<<grid_for(cap, 256), 256, 0, c->stream>>>(cap, keys, slot_id, reps, c->row_ptr.p,
                                                          c->col.p, S.pat_ptr.p, S.pat_delta.p);
    const int64_t nsl = (n + 63) / 64;
    S.nslices = nsl;
    INS_TRY(ensure(c, S.slice_len, nsl + 1));
    INS_TRY(ensure(c, S.slice_ptr, nsl + 1));
    k_slice_len<<<grid_for(nsl, 256), 256, 0, c->stream>>>(n, c->row_ptr.p, S.slice_len.p);
    k_slice_ptr<<<1, 1, 0, c->stream>>>(nsl, S.slice_len.p, S.slice_ptr.p);
    INS_HIP_CHECK(c, hipGetLastError());
    int herr = 0;
    int64_t nvals = 0;
    INS_HIP_CHECK(c, hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    INS_HIP_CHECK(c, hipMemcpyAsync(&nvals, S.slice_ptr.p + nsl, sizeof(int64_t),
                                    hipMemcpyDeviceToHost, c->stream));
    INS_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    hipFree(h);
    hipFree(keys);
    hipFree(reps);
    hipFree(slot_of);
    hipFree(used);
    hipFree(slot_id);
    hipFree(err);
    S.nvals = nvals;
    // worth it only if the offset dictionary is small and padding modest
    S.ok = (herr == 0) && (ndel * 8 <= c->nnz + 64) && (nvals <= c->nnz + c->nnz / 4 + 64 * 64);
    if (!S.ok) {
        S.built = true;
        return INS_OK;
    }
    if (c->matrix_precision == INS_FP64) INS_TRY(ensure(c, S.v64, nvals));
    else release(c, S.v64);
    INS_TRY(ensure(c, S.v32, nvals));
    k_sell_fill<<<grid_for(n, 256), 256, 0, c->stream>>>(
        n, c->row_ptr.p, (c->matrix_precision == INS_FP64) ? c->val64.p : nullptr, c->val32.p,
        S.slice_ptr.p, S.slice_len.p, (c->matrix_precision == INS_FP64) ? S.v64.p : nullptr,
        S.v32.p);
    INS_HIP_CHECK(c, hipGetLastError());
    INS_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    S.built = true;
    return INS_OK;
}

int solve(Ctx* c, int mode, int format, double rtol, int max_it, int* iters, double* relres)
{
    if (!c->have_matrix || !c->have_rhs)
        return fail(c, INS_ERR_STATE, "ins_solve before ins_assemble (matrix and rhs)");
    if (mode == INS_SOLVE_FP64 && c->matrix_precision != INS_FP64)
        return fail(c, INS_ERR_STATE, "FP64 solve needs an FP64 assembly");
    if (format == INS_FMT_SELL && !c->sell.built) INS_TRY(build_sell(c));
    if (format == INS_FMT_SELL && !c->sell.ok) format = INS_FMT_CSR;
    INS_TRY(ensure_solver_state(c, mode));
    const int64_t n = c->n_active;
    const unsigned gv = red_grid(n, RB);
    c->stats.spmv_ms = 0.;
    c->stats.spmv_calls = 0;
    INS_HIP_CHECK(c, hipEventRecord(c->ev0, c->stream));
    int refinements = 0;
    int status = INS_OK;
    if (mode == INS_SOLVE_FP64) {
        const double* sv = (format == INS_FMT_SELL) ? c->sell.v64.p : nullptr;
        k_init<double, double><<<gv, RB, 0, c->stream>>>(n, c->b.p, c->x.p, c->r.p, c->p.p,
                                                         c->dinv64.p, rtol * rtol, c->partial.p,
                                                         c->counter.p, c->scal.p);
        INS_HIP_CHECK(c, hipGetLastError());
        INS_TRY(read_scal(c));
        while (!c->h_scal->done && c->h_scal->iters < max_it) {
            const int k = std::min(CHUNK, max_it - c->h_scal->iters);
            for (int it = 0; it < k; it++) {
                {
                    ProfScope ps(c);
                    INS_TRY((launch_spmv_pq<double, double, double>(c, format, c->val64.p, sv, c->p.p, c->q.p)));
                }
                k_update_xr<double, double><<<gv, RB, 0, c->stream>>>(n, c->x.p, c->r.p, c->p.p, c->q.p, c->dinv64.p, c->partial.p, c->counter.p, c->scal.p);
                k_update_p<double, double><<<gv, RB, 0, c->stream>>>(n, c->p.p, c->r.p, c->dinv64.p, c->scal.p);
            }
            INS_HIP_CHECK(c, hipGetLastError());
            INS_TRY(read_scal(c));
        }
        // true residual |b - A x|
        INS_TRY(launch_resid(c, format, 0, nullptr));
    } else {
        // mixed: FP32 sweeps, FP64 residual replacement at every chunk end
        const float* sv = (format == INS_FMT_SELL) ? c->sell.v32.p : nullptr;
        INS_HIP_CHECK(c, hipMemsetAsync(c->x.p, 0, n * sizeof(double), c->stream));
        k_init<float, float><<<gv, RB, 0, c->stream>>>(n, c->b.p, c->x32.p, c->r32.p, c->p32.p,
                                                       c->dinv32.p, rtol * rtol, c->partial.p,
                                                       c->counter.p, c->scal.p);
        INS_HIP_CHECK(c, hipGetLastError());
        INS_TRY(read_scal(c));
        bool converged = c->h_scal->done && !c->h_scal->breakdown;
        while (!converged && c->h_scal->iters < max_it && !c->h_scal->breakdown) {
            const int k = std::min(CHUNK, max_it - c->h_scal->iters);
            for (int it = 0; it < k; it++) {
                {
                    ProfScope ps(c);
                    INS_TRY((launch_spmv_pq<float, float, float>(c, format, c->val32.p, sv, c->p32.p, c->q32.p)));
                }
                k_update_xr<float, float><<<gv, RB, 0, c->stream>>>(n, c->x32.p, c->r32.p, c->p32.p, c->q32.p, c->dinv32.p, c->partial.p, c->counter.p, c->scal.p);
                k_update_p<float, float><<<gv, RB, 0, c->stream>>>(n, c->p32.p, c->r32.p, c->dinv32.p, c->scal.p);
            }
            INS_HIP_CHECK(c, hipGetLastError());
            // residual replacement: x64 += x32; r = b - A x64 (FP64)
            k_fold<<<gv, RB, 0, c->stream>>>(n, c->x.p, c->x32.p);
            INS_HIP_CHECK(c, hipGetLastError());
            INS_TRY(launch_resid(c, format, 1, nullptr));
            refinements++;
            INS_TRY(read_scal(c));
            converged = c->h_scal->done && !c->h_scal->breakdown;
        }
    }
    INS_TRY(read_scal(c));
    INS_HIP_CHECK(c, hipEventRecord(c->ev1, c->stream));
    INS_HIP_CHECK(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    INS_HIP_CHECK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    const Scal& S = *c->h_scal;
    const double rel = S.bb > 0. ? std::sqrt(S.tt / S.bb) : 0.;
    c->stats.solve_ms = ms;
    c->stats.iterations = S.iters;
    c->stats.refinements = refinements;
    c->stats.relres = rel;
    if (iters) *iters = S.iters;
    if (relres) *relres = rel;
    c->have_solution = true;
    if (S.breakdown) status = fail(c, INS_ERR_BREAKDOWN, "CG breakdown (p.Ap <= 0 or non-finite)");
    else if (!(rel <= rtol)) status = fail(c, INS_ERR_NOCONV, "iteration limit reached before rtol");
    return status;
}

int distribute(Ctx* c, int op)
{
    if (!c->have_solution) return fail(c, INS_ERR_STATE, "ins_distribute before ins_solve");
    const int64_t nd = c->n_nodes * c->dpn;
    k_distribute<<<grid_for(nd, 256), 256, 0, c->stream>>>(nd, c->dof_id.p, c->x.p, c->g.p,
                                                            c->u.p, op == INS_DIST_ADD);
    INS_HIP_CHECK(c, hipGetLastError());
    INS_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return INS_OK;
}

int vec_norm2(Ctx* c, const double* v, int64_t n, double* out)
{
    INS_TRY(ensure_solver_state(c, INS_SOLVE_FP64));
    k_norm2<<<red_grid(n, RB), RB, 0, c->stream>>>(n, v, c->partial.p, c->counter.p, c->scal.p);
    INS_HIP_CHECK(c, hipGetLastError());
    INS_TRY(read_scal(c));
    *out = std::sqrt(c->h_scal->tt);
    return INS_OK;
}

}  // namespace ins
