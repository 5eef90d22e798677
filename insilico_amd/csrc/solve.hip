// Jacobi-preconditioned CG on gfx950 (Eigen::ConjugateGradient semantics as
// called by base/solver/Eigen3.hpp:265-275: x0 = 0, z = D^-1 r, stop when
// |r|^2 < rtol^2 |b|^2 on the recursive residual).
//
// Every iteration is three kernels that stream HBM once each:
//   K1  q = A p, per-block partials of p.q
//   K2  alpha = rho / sum(p.q); x += alpha p; r -= alpha q; partials of r.r, r.D^-1 r
//   K3  stop test on sum(r.r); beta = sum(r.z) / rho; p = D^-1 r + beta p
// Reductions are deterministic and atomic-free: a producer kernel writes one
// partial per block (fixed tree inside the block); every block of the
// consumer kernel sums the same partials in the same order in its prologue,
// so all blocks agree on the scalar without a grid barrier, a float atomic or
// an extra launch.  Block 0 records the scalars for the host.  rho is kept in
// two slots (iteration parity) so that no block reads a slot being written.
// The host synchronises once per chunk of iterations.
//
// Mixed precision (INS_SOLVE_MIXED): FP32 matrix values and FP32 vectors in
// K1-K3 with FP32 accumulation in the SpMV and FP64 accumulation of all dot
// products; at the end of every chunk the FP32 iterate is folded into the
// FP64 solution and the residual is replaced by the FP64 residual b - A x
// (FP64 values, FP64 arithmetic).  Convergence is only declared on that FP64
// residual.
#include <algorithm>
#include <climits>
#include <deque>
#include <mutex>
#include <cmath>
#include <cstdlib>

#include <hip/hip_ext.h>
#include <rocsolver/rocsolver.h>

#include "common.hpp"

// Multiply-adds fuse only within one source expression (s += a * b is one
// FMA everywhere).  With contraction left to the backend, whether a product
// fuses depended on its other uses inside each kernel, so that the same
// inline row sum rounded differently in the solve's kernel and in the
// serving batch's kernel; a request must give the same bits in both.
#pragma clang fp contract(on)

namespace ins {

namespace {

constexpr int VB = 512;       // threads per block of the vector kernels
constexpr int SB = 256;       // threads per block of the SpMV kernels
constexpr int MAXG = 2048;    // max blocks of a partial-producing kernel
constexpr int CHUNK = 32;     // FP64 iterations between host checks
constexpr int MCHUNK = 16;    // mixed: iterations between host checks
// partial slots
constexpr int P_PQ = 0, P_RR = 1, P_RZ = 2, P_TT = 3, P_BB = 4, NSLOT = 5;

__device__ inline double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// sum of v over the block, returned to every thread (fixed order)
template <int NV>
__device__ inline void block_allsum(double (&v)[NV])
{
    __shared__ double sh[NV][16];
#pragma unroll
    for (int k = 0; k < NV; k++) v[k] = wave_sum(v[k]);
    const int nw = blockDim.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int k = 0; k < NV; k++) sh[k][threadIdx.x >> 6] = v[k];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double s = 0.;
        for (int w = 0; w < nw; w++) s += sh[k][w];
        v[k] = s;
    }
}

// A block's place in the decomposition of one solve: its index and the
// number of blocks.  A plain launch uses blockIdx/gridDim; a batched launch
// (several solves in one grid) gives every solve the decomposition it has
// alone, so its sums, and hence its iterates, do not depend on the batch.
struct Blk {
    int id, n;
};
__device__ inline Blk grid_blk() { return Blk{(int)blockIdx.x, (int)gridDim.x}; }

// one partial per block
template <int NV>
__device__ inline void block_partials(double (&v)[NV], double* part, const int (&slot)[NV],
                                      Blk b = grid_blk())
{
    block_allsum<NV>(v);
    if (threadIdx.x == 0)
#pragma unroll
        for (int k = 0; k < NV; k++) part[slot[k] * MAXG + b.id] = v[k];
}

// every block: the same total of cnt partials of the given slots
template <int NV>
__device__ inline void totals(const double* part, int cnt, const int (&slot)[NV], double (&t)[NV])
{
#pragma unroll
    for (int k = 0; k < NV; k++) t[k] = 0.;
    for (int i = threadIdx.x; i < cnt; i += blockDim.x)
#pragma unroll
        for (int k = 0; k < NV; k++) t[k] += part[slot[k] * MAXG + i];
    block_allsum<NV>(t);
}

__device__ inline bool leader(Blk b = grid_blk()) { return b.id == 0 && threadIdx.x == 0; }

// this thread's share of totals<NV> (the loads only; block_allsum finishes
// it), so that a kernel can issue them together with its other loads
template <int NV>
__device__ inline void totals_share(const double* part, int cnt, const int (&slot)[NV],
                                    double (&t)[NV])
{
#pragma unroll
    for (int k = 0; k < NV; k++) t[k] = 0.;
    for (int i = threadIdx.x; i < cnt; i += blockDim.x)
#pragma unroll
        for (int k = 0; k < NV; k++) t[k] += part[slot[k] * MAXG + i];
}

// index of the calling wavefront within its block, as a wave-uniform value:
// slice indices derived from it stay in scalar registers, so the per-slice
// offset, pattern and column-offset reads go through the scalar cache rather
// than through one vector load per lane
__device__ inline int wave_in_block() { return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6); }

// Workgroups are placed on the 8 XCDs round-robin by block id.  Each XCD
// owns one contiguous eighth of the items (rows or slices) and its blocks
// sweep that eighth together in grid-stride order: 8 sequential streams,
// one per XCD, with the x gathers of the SpMV and the vectors reused by the
// next kernel local to that XCD's L2.
constexpr int NXCD = 8;

struct Sweep {
    int64_t first;   // first item of this block
    int64_t end;     // end of the XCD's range
    int64_t stride;  // items between two visits of this block
};

// items are visited as first + j*stride (+ per-block offsets < step)
__device__ int g_xcd_mode = 0;  // 1: per-XCD ranges for every kernel (INS_XCD)
__device__ int g_nt = 0;        // 1: streaming (non-temporal) vector stores and value loads
__device__ int g_bsym_bands = 4;   // banded slice order of the node-block SpMV (INS_BSYM_BANDS)
// mixed: a replacement restarts the direction when |b - A x|^2 exceeds this
// multiple of the recursive |r|^2 (INS_RESTART_RATIO)
__device__ double g_restart_ratio = 64.;

template <typename T>
__device__ inline void st(T* p, T v, bool nt)
{
    if (nt) __builtin_nontemporal_store(v, p);
    else *p = v;
}

__device__ inline Sweep xcd_sweep(int64_t total, int64_t step, Blk b = grid_blk(),
                                  bool per_xcd = false)
{
    const int g = b.n;
    Sweep w;
    if ((g % NXCD) || !(g_xcd_mode || per_xcd)) {
        w.first = (int64_t)b.id * step;
        w.end = total;
        w.stride = (int64_t)g * step;
        return w;
    }
    const int xcd = b.id % NXCD, local = b.id / NXCD, nloc = g / NXCD;
    int64_t per = (total + NXCD - 1) / NXCD;
    per = (per + step - 1) / step * step;
    const int64_t lo = (int64_t)xcd * per;
    w.first = lo + (int64_t)local * step;
    w.end = lo + per < total ? lo + per : total;
    w.stride = (int64_t)nloc * step;
    return w;
}

// ---------------------------------------------------------------------------
// SpMV rows
struct SellView {
    const int32_t* perm;   // slice row -> matrix row (null: identity)
    const int64_t* sptr;   // value offset of every slice
    const int32_t* spat;   // column-offset list (pattern) of every slice
    const int32_t* pptr;
    const int32_t* plen;
    const int32_t* pdel;
};

// Values of a slice are stored in groups of 4 consecutive entries per row:
// position = slice_base + (k/4)*256 + lane*4 + k%4, so a lane reads 4 entries
// of its row with one 16-byte load (32 bytes for FP64) and a wavefront reads
// 1 KB contiguous.
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

template <typename TV>
__device__ inline void load4(const TV* p, TV (&o)[4])
{
    if constexpr (sizeof(TV) == 4) {
        const f32x4 v = *reinterpret_cast<const f32x4*>(p);
        o[0] = v.x; o[1] = v.y; o[2] = v.z; o[3] = v.w;
    } else {
        const f64x2 a = *reinterpret_cast<const f64x2*>(p);
        const f64x2 b = *reinterpret_cast<const f64x2*>(p + 2);
        o[0] = a.x; o[1] = a.y; o[2] = b.x; o[3] = b.y;
    }
}


// one lane per row of a 64-row slice; all rows of a slice share one
// column-offset list (the union of their own lists, values zero-padded), so
// the offsets are wave-uniform and read with scalar loads.  Padded entries
// may point outside 0..n-1: their index is clamped (their value is 0).
template <typename TV, typename TX, typename TA, int SG = 3>
__device__ inline TA sell_row(int64_t row, bool live, int64_t n, int lane, int64_t slice,
                              const SellView& S, const TV* __restrict__ v,
                              const TX* __restrict__ x)
{
    TA s = (TA)0;
    const int64_t base = S.sptr[slice] + lane * 4;
    const int pid = S.spat[slice];
    const int32_t* d = S.pdel + S.pptr[pid];
    const int len = S.plen[pid];
    const int64_t last = n - 1;
    for (int k0 = 0; k0 < len; k0 += 4 * SG) {
        TV vv[SG][4];
        TX xx[SG][4];
#pragma unroll
        for (int g = 0; g < SG; g++) {
            const int kg = k0 + 4 * g;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                vv[g][j] = (TV)0;
                xx[g][j] = (TX)0;
            }
            if (kg < len && live) {
                load4<TV>(v + base + (int64_t)(kg >> 2) * 256, vv[g]);
#pragma unroll
                for (int j = 0; j < 4; j++)
                    if (kg + j < len) {
                        int64_t idx = row + d[kg + j];
                        idx = idx < 0 ? 0 : (idx > last ? last : idx);
                        xx[g][j] = x[idx];
                    }
            }
        }
#pragma unroll
        for (int g = 0; g < SG; g++)
#pragma unroll
            for (int j = 0; j < 4; j++) s += (TA)vv[g][j] * (TA)xx[g][j];
    }
    return s;
}

// Symmetric FP32 slices (scalar problems, identity row order): the values
// of the offsets u_k >= 0 shared by all rows, k-major within a slice (lane
// contiguous, 256 bytes per wave and offset); the entry (r, r - u_k) below
// the diagonal is the mirror (r - u_k, r), one k-row of the slice holding
// row r - u_k.  Every stored value is read twice per SpMV, once for its own
// row and once for its mirror, the second time from the caches.
struct SymView {
    const float* v;     // null: the full slices
    const int32_t* u;   // nu offsets, ascending (wave-uniform, scalar loads)
    int nu;
    int dpn;            // 1: scalar rows; 2, 3: node blocks (see bsym_node)
    int64_t nn;         // node blocks: nodes per component
    int lines;          // scalar rows of stencil lines (see sym_lines_sweep)
    const double* v64;  // the same slices in FP64 (FP64 sweeps, residual replacement)
};

// the symmetric slices of the value type of a sweep
template <typename TV>
__device__ inline const TV* sym_vals(const SymView& Y)
{
    if constexpr (sizeof(TV) == 8) return Y.v64;
    else return Y.v;
}

template <typename TV, typename TX, typename TA, int G = 8>
__device__ inline TA sym_row_v(int64_t row, bool live, int64_t n, int lane, int64_t slice,
                               const TV* __restrict__ vals, const int32_t* __restrict__ u,
                               int nu, const TX* __restrict__ x)
{
    TA s = (TA)0;
    if (!live) return s;
    const int64_t ld = (int64_t)64 * nu;   // values per slice
    const TV* __restrict__ own = vals + slice * ld + lane;
    const int64_t last = n - 1;
    for (int k0 = 0; k0 < nu; k0 += G) {
        TV vu[G], vl[G];
        TX xu[G], xl[G];
#pragma unroll
        for (int j = 0; j < G; j++) {
            vu[j] = (TV)0;
            vl[j] = (TV)0;
            xu[j] = (TX)0;
            xl[j] = (TX)0;
            const int k = k0 + j;
            if (k < nu) {
                const int d = u[k];
                int64_t iu = row + d;
                iu = iu > last ? last : iu;   // padded entries: value 0
                vu[j] = own[(int64_t)k * 64];
                xu[j] = x[iu];
                const int64_t jl = row - d;
                if (d > 0 && jl >= 0) {
                    vl[j] = vals[(jl >> 6) * ld + (int64_t)k * 64 + (jl & 63)];
                    xl[j] = x[jl];
                }
            }
        }
#pragma unroll
        for (int j = 0; j < G; j++) s += (TA)vu[j] * (TA)xu[j];
#pragma unroll
        for (int j = 0; j < G; j++) s += (TA)vl[j] * (TA)xl[j];
    }
    return s;
}

template <typename TX, typename TA, int G = 8>
__device__ inline TA sym_row(int64_t row, bool live, int64_t n, int lane, int64_t slice,
                             const SymView& Y, const TX* __restrict__ x)
{
    return sym_row_v<float, TX, TA, G>(row, live, n, lane, slice, Y.v, Y.u, Y.nu, x);
}

// Wave-wide shifts by one lane (the GFX9 wavefront-shift DPP controls):
// lane i takes lane i - 1's value (shr) or lane i + 1's (shl); the lane
// without a source keeps `edge`.
__device__ inline float wave_shr1(float v, float edge)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, edge),
                                                                 __builtin_bit_cast(int, v), 0x138,
                                                                 0xf, 0xf, false));
}
__device__ inline float wave_shl1(float v, float edge)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, edge),
                                                                 __builtin_bit_cast(int, v), 0x130,
                                                                 0xf, 0xf, false));
}

__device__ inline double wave_shr1(double v, double edge)
{
    const uint64_t a = __builtin_bit_cast(uint64_t, v), e = __builtin_bit_cast(uint64_t, edge);
    const uint32_t lo = __builtin_amdgcn_update_dpp((int)(uint32_t)e, (int)(uint32_t)a, 0x138,
                                                    0xf, 0xf, false);
    const uint32_t hi = __builtin_amdgcn_update_dpp((int)(uint32_t)(e >> 32),
                                                    (int)(uint32_t)(a >> 32), 0x138, 0xf, 0xf,
                                                    false);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
__device__ inline double wave_shl1(double v, double edge)
{
    const uint64_t a = __builtin_bit_cast(uint64_t, v), e = __builtin_bit_cast(uint64_t, edge);
    const uint32_t lo = __builtin_amdgcn_update_dpp((int)(uint32_t)e, (int)(uint32_t)a, 0x130,
                                                    0xf, 0xf, false);
    const uint32_t hi = __builtin_amdgcn_update_dpp((int)(uint32_t)(e >> 32),
                                                    (int)(uint32_t)(a >> 32), 0x130, 0xf, 0xf,
                                                    false);
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Symmetric slices whose offsets are stencil lines: u = {0, 1, c1 - 1, c1,
// c1 + 1, c2 - 1, ...} (NU = 14: the 27-point rows of a 3-D grid; NU = 5:
// the 9-point rows of a 2-D grid), so that every row reads x on the lines
// c - 1, c, c + 1 around the centres 0, +-c_m.  A wave loads each line of x
// once (64 consecutive values, coalesced) and takes its two neighbours by
// wave shifts, with the two values past the wave's ends as edge values: 9
// x loads per row instead of 27.  Every address is a wave-uniform base plus
// a lane offset fixed for the whole sweep (the mirror offsets are computed
// once).
//
// One row of an interior slice (own: the slice's values, mb: the values of
// the slice `back` slices earlier, pb: x at the slice's first row).  The
// offsets are loaded LG at a time (LG a multiple of 8: 16 keeps all FP32
// loads of a row in flight, 8 holds FP64 rows to the registers of the
// generic row) and summed eight at a time in sym_row_v's order, so the row
// sums are the same bits.
template <int NU, int LG, typename TV, typename TX, typename TA>
__device__ inline TA sym_lines_row(int lane, const TV* __restrict__ own, const TV* __restrict__ mb,
                                   const uint32_t (&mo)[NU], const int (&du)[NU],
                                   const TX* __restrict__ pb)
{
    static_assert(LG % 8 == 0, "loads in whole groups of the sums");
    // every later group starts a line (k0 = LG = 2 + 3m), so a line's
    // offsets lie in one group
    static_assert(LG >= NU || (LG % 3 == 2 && NU <= 2 * LG), "a line within one load group");
    TA s = (TA)0;
#pragma unroll
    for (int k0 = 0; k0 < NU; k0 += LG) {
        TV vu[LG], vl[LG];
        TX xu[LG], xl[LG];
#pragma unroll
        for (int j = 0; j < LG; j++) {
            const int k = k0 + j;
            vu[j] = (TV)0;
            vl[j] = (TV)0;
            xu[j] = (TX)0;
            xl[j] = (TX)0;
            if (k < NU) vu[j] = own[k * 64 + lane];
            if (k > 0 && k < NU) vl[j] = mb[mo[k]];
        }
        // x line by line: line 0 (offsets 0, 1), line m (c - 1, c, c + 1
        // at k = 2 + 3m, 3 + 3m, 4 + 3m)
#pragma unroll
        for (int j = 0; j < LG; j++) {
            const int k = k0 + j;
            if (k == 0 && j + 1 < LG) {
                const TX a = pb[lane];
                xu[j] = a;
                xu[j + 1] = wave_shl1(a, pb[64]);
                xl[j + 1] = wave_shr1(a, pb[-1]);
            } else if (k >= 2 && k < NU && (k - 2) % 3 == 0 && j + 2 < LG) {
                const int c = du[k + 1];
                const TX au = pb[c + lane];
                xu[j] = wave_shr1(au, pb[c - 1]);
                xu[j + 1] = au;
                xu[j + 2] = wave_shl1(au, pb[c + 64]);
                const TX al = pb[lane - c];
                xl[j] = wave_shl1(al, pb[64 - c]);
                xl[j + 1] = al;
                xl[j + 2] = wave_shr1(al, pb[-c - 1]);
            }
        }
#pragma unroll
        for (int g0 = 0; g0 < LG; g0 += 8) {
#pragma unroll
            for (int j = g0; j < g0 + 8; j++)
                if (k0 + j < NU) s += (TA)vu[j] * (TA)xu[j];
#pragma unroll
            for (int j = g0; j < g0 + 8; j++)
                if (k0 + j < NU) s += (TA)vl[j] * (TA)xl[j];
        }
        // the next group's loads start after these sums (registers of one
        // group live at a time)
        if (k0 + LG < NU) __builtin_amdgcn_sched_barrier(0);
    }
    return s;
}

// Every row of the slices of one block (per-XCD ranges) with its sum
// (A x)_row: f(row, s) for the rows below n.  The slices whose lines reach
// past row 0 or row n - 1 take sym_row_v.
template <int NU, int LG, typename TV, typename TX, typename TA, class F>
__device__ inline void sym_lines_visit(Blk b, int64_t n, const SymView& Y,
                                       const TV* __restrict__ vals, const TX* __restrict__ x, F f)
{
    const int lane = threadIdx.x & 63;
    const int64_t nsl = (n + 63) / 64;
    const Sweep w = xcd_sweep(nsl, SB / 64, b, true);
    int du[NU];
#pragma unroll
    for (int k = 0; k < NU; k++) du[k] = __builtin_amdgcn_readfirstlane(Y.u[k]);
    const int dmax = du[NU - 1];
    const int64_t ld = (int64_t)64 * NU;
    const int back = 1 + (dmax >> 6);        // slices a mirror entry lies back, at most
    uint32_t mo[NU];                          // mirror of (row, row - u_k) from slice - back
#pragma unroll
    for (int k = 0; k < NU; k++) {
        const int t = lane - du[k];
        mo[k] = (uint32_t)((int64_t)(back + (t >> 6)) * ld + k * 64 + (t & 63));
    }
    const int64_t s_lo = back, s_hi = (n - 1 - dmax - 64) / 64;
    for (int64_t slice = w.first + wave_in_block(); slice < w.end; slice += w.stride) {
        const int64_t row = slice * 64 + lane;
        const bool live = row < n;
        TA s;
        if (slice < s_lo || slice > s_hi)
            s = sym_row_v<TV, TX, TA, 8>(live ? row : 0, live, n, lane, slice, vals, Y.u, Y.nu, x);
        else
            s = sym_lines_row<NU, LG, TV, TX, TA>(lane, vals + slice * ld,
                                                  vals + (slice - back) * ld, mo, du,
                                                  x + slice * 64);
        if (live) f(row, s);
    }
}

// q = A p over the stencil-line slices; returns this thread's share of p.q
template <int NU, typename TV, typename TX, typename TA>
__device__ inline double sym_lines_sweep(Blk b, int64_t n, const SymView& Y,
                                         const TV* __restrict__ vals, const TX* __restrict__ p,
                                         TX* __restrict__ q)
{
    double pq = 0.;
    sym_lines_visit<NU, sizeof(TV) == 4 ? 16 : 8, TV, TX, TA>(
        b, n, Y, vals, p, [&](int64_t row, TA s) {
            const TX qs = (TX)s;
            st(&q[row], qs, g_nt);
            pq += (double)p[row] * (double)qs;
        });
    return pq;
}

// CSR row with LPR lanes; each lane keeps 4 entries in flight
template <int LPR, typename TV, typename TX, typename TA>
__device__ inline TA csr_row(int64_t row, int sub, const int32_t* __restrict__ rp,
                             const int32_t* __restrict__ col, const TV* __restrict__ val,
                             const TX* __restrict__ x)
{
    TA s = (TA)0;
    const int32_t e = rp[row + 1];
    for (int32_t k0 = rp[row] + sub; k0 < e; k0 += 4 * LPR) {
        int32_t cc[4];
        TV vv[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int32_t k = k0 + j * LPR;
            cc[j] = (k < e) ? col[k] : -1;
            vv[j] = (k < e) ? val[k] : (TV)0;
        }
        TX xx[4];
#pragma unroll
        for (int j = 0; j < 4; j++) xx[j] = (cc[j] >= 0) ? x[cc[j]] : (TX)0;
#pragma unroll
        for (int j = 0; j < 4; j++) s += (TA)vv[j] * (TA)xx[j];
    }
#pragma unroll
    for (int o = LPR / 2; o > 0; o >>= 1) s += __shfl_xor(s, o, LPR);
    return s;
}

// Symmetric node-block slices (vector problems, one lane per node of a
// 64-node slice): out[i] = sum over the node's D x D blocks of the rows of
// its D components.  Blocks (a, a + u_k), u_k >= 0, are stored at node a in
// the slots of SymLayout, slot-major and lane contiguous.  Block
// (a, a - u_k) is block (a - u_k, a) transposed, read at node a - u_k.
// Vectors are component-major: component c of node b at c * nn + b.
// Offsets are taken G at a time so that G blocks of loads are in flight per
// wave.  (Slots grouped four by four for 16-byte loads, one access per
// group and lane, were measured slower: 126.9 against 117.0 us per FP32
// SpMV at 96^3 elasticity.)
template <int D>
__device__ inline constexpr int tri_slot(int i, int j)
{
    return i * D - i * (i - 1) / 2 + (j - i);
}

// the D*D entries (row-major) of block k >= 1 of the node whose slot 0 is
// at p (its lane included)
template <int D, typename TV>
__device__ inline void bsym_block(const TV* __restrict__ p, int k, TV (&b)[D * D])
{
    constexpr int T = D * (D + 1) / 2, D2 = D * D;
    const int64_t s0 = T + (int64_t)(k - 1) * D2;
#pragma unroll
    for (int e = 0; e < D2; e++) b[e] = p[(s0 + e) * 64];
}

template <int D, typename TV, typename TX, typename TA, int G = 3>
__device__ inline void bsym_node(int64_t a, bool live, int lane, int64_t slice, const SymView& Y,
                                 const TV* __restrict__ vals, const TX* __restrict__ x,
                                 TA (&out)[D])
{
    constexpr int T = D * (D + 1) / 2, D2 = D * D;
    const int nu = Y.nu;
    const int64_t nn = Y.nn, last = nn - 1;
    const int64_t ld = 64 * SymLayout{D, nu}.width();
#pragma unroll
    for (int i = 0; i < D; i++) out[i] = (TA)0;
    if (!live) return;
    const TV* __restrict__ own = vals + slice * ld + lane;
    {
        TV bv[T];
#pragma unroll
        for (int e = 0; e < T; e++) bv[e] = own[(int64_t)e * 64];
        TX xa[D];
#pragma unroll
        for (int j = 0; j < D; j++) xa[j] = x[j * nn + a];
#pragma unroll
        for (int i = 0; i < D; i++)
#pragma unroll
            for (int j = 0; j < D; j++)
                out[i] += (TA)bv[i <= j ? tri_slot<D>(i, j) : tri_slot<D>(j, i)] * (TA)xa[j];
    }
    for (int k0 = 1; k0 < nu; k0 += G) {
        TV bu[G][D2], bl[G][D2];
        TX xu[G][D], xl[G][D];
#pragma unroll
        for (int g = 0; g < G; g++) {
            const int k = k0 + g;
#pragma unroll
            for (int e = 0; e < D2; e++) {
                bu[g][e] = (TV)0;
                bl[g][e] = (TV)0;
            }
#pragma unroll
            for (int j = 0; j < D; j++) {
                xu[g][j] = (TX)0;
                xl[g][j] = (TX)0;
            }
            if (k < nu) {
                const int d = Y.u[k];
                int64_t bn = a + d;
                bn = bn > last ? last : bn;   // padded blocks hold zeros
                bsym_block<D, TV>(own, k, bu[g]);
#pragma unroll
                for (int j = 0; j < D; j++) xu[g][j] = x[j * nn + bn];
                const int64_t m = a - d;
                if (m >= 0) {
                    const TV* __restrict__ mp = vals + (m >> 6) * ld + (m & 63);
                    bsym_block<D, TV>(mp, k, bl[g]);
#pragma unroll
                    for (int j = 0; j < D; j++) xl[g][j] = x[j * nn + m];
                }
            }
        }
#pragma unroll
        for (int g = 0; g < G; g++) {
#pragma unroll
            for (int i = 0; i < D; i++)
#pragma unroll
                for (int j = 0; j < D; j++) out[i] += (TA)bu[g][i * D + j] * (TA)xu[g][j];
#pragma unroll
            for (int i = 0; i < D; i++)
#pragma unroll
                for (int j = 0; j < D; j++) out[i] += (TA)bl[g][j * D + i] * (TA)xl[g][j];
        }
    }
}

// Node blocks whose node offsets are stencil lines (u = {0, 1, c1 - 1, c1,
// c1 + 1, ...}, NU = 14 in 3-D, 5 in 2-D), FP32: as sym_lines_sweep does for
// scalar rows, a wave loads each line of p of each component once (64
// consecutive nodes) and takes the neighbouring nodes by wave shifts, 3D
// instead of D * (2 NU - 1) vector loads of p per node.  The blocks of a
// line (u = c - 1, c, c + 1) are taken together.  Only for slices whose
// lines stay inside the nodes (the caller checks).
template <int D, int NU>
__device__ inline void bsym_lines_node(int64_t a, int lane, int64_t slice, const SymView& Y,
                                       const float* __restrict__ vals,
                                       const float* __restrict__ p, const int (&du)[NU],
                                       float (&out)[D])
{
    constexpr int T = D * (D + 1) / 2, D2 = D * D;
    const int64_t nn = Y.nn;
    const int64_t ld = 64 * SymLayout{D, NU}.width();
    const float* __restrict__ own = vals + slice * ld + lane;
#pragma unroll
    for (int i = 0; i < D; i++) out[i] = 0.f;
    // line 0: the diagonal block and the block of u = 1 (and its mirror)
    {
        float bv[T], bu[D2], bl[D2], x0[D], xu[D], xl[D];
#pragma unroll
        for (int e = 0; e < T; e++) bv[e] = own[(int64_t)e * 64];
        bsym_block<D, float>(own, 1, bu);
        const int64_t m = a - 1;
        bsym_block<D, float>(vals + (m >> 6) * ld + (m & 63), 1, bl);
#pragma unroll
        for (int j = 0; j < D; j++) {
            const float* __restrict__ pb = p + j * nn + slice * 64;
            const float c0 = pb[lane];
            x0[j] = c0;
            xu[j] = wave_shl1(c0, pb[64]);
            xl[j] = wave_shr1(c0, pb[-1]);
        }
#pragma unroll
        for (int i = 0; i < D; i++)
#pragma unroll
            for (int j = 0; j < D; j++) out[i] += bv[i <= j ? tri_slot<D>(i, j) : tri_slot<D>(j, i)] * x0[j];
#pragma unroll
        for (int i = 0; i < D; i++)
#pragma unroll
            for (int j = 0; j < D; j++) out[i] += bu[i * D + j] * xu[j];
#pragma unroll
        for (int i = 0; i < D; i++)
#pragma unroll
            for (int j = 0; j < D; j++) out[i] += bl[j * D + i] * xl[j];
    }
    // lines m: blocks k = 2 + 3m, 3 + 3m, 4 + 3m (u = c - 1, c, c + 1)
#pragma unroll 1
    for (int mline = 0; mline < (NU - 2) / 3; mline++) {
        const int k0 = 2 + 3 * mline;
        const int c = du[k0 + 1];
        float bu[3][D2], bl[3][D2], xu[3][D], xl[3][D];
#pragma unroll
        for (int t = 0; t < 3; t++) {
            bsym_block<D, float>(own, k0 + t, bu[t]);
            const int64_t m = a - du[k0 + t];
            bsym_block<D, float>(vals + (m >> 6) * ld + (m & 63), k0 + t, bl[t]);
        }
#pragma unroll
        for (int j = 0; j < D; j++) {
            const float* __restrict__ pb = p + j * nn + slice * 64;
            const float au = pb[c + lane];
            xu[0][j] = wave_shr1(au, pb[c - 1]);
            xu[1][j] = au;
            xu[2][j] = wave_shl1(au, pb[c + 64]);
            const float al = pb[lane - c];
            xl[0][j] = wave_shl1(al, pb[64 - c]);
            xl[1][j] = al;
            xl[2][j] = wave_shr1(al, pb[-c - 1]);
        }
#pragma unroll
        for (int t = 0; t < 3; t++) {
#pragma unroll
            for (int i = 0; i < D; i++)
#pragma unroll
                for (int j = 0; j < D; j++) out[i] += bu[t][i * D + j] * xu[t][j];
#pragma unroll
            for (int i = 0; i < D; i++)
#pragma unroll
                for (int j = 0; j < D; j++) out[i] += bl[t][j * D + i] * xl[t][j];
        }
    }
}

// q = A p on the node blocks: every lane writes the D rows of its node and
// adds their p.q to the block partial
template <int D, typename TV, typename TX, typename TA>
__device__ inline double bsym_sweep(Blk b, const SymView& Y, const TX* __restrict__ p,
                                    TX* __restrict__ q)
{
    const TV* __restrict__ vals = sym_vals<TV>(Y);
    constexpr int G = sizeof(TV) == 8 ? 1 : 3;   // FP64 blocks: one offset at a time
    const int lane = threadIdx.x & 63;
    const int64_t nn = Y.nn;
    const int64_t nsl = (nn + 63) / 64;
    double pq = 0.;
    // stencil lines (FP32): the line form for the slices whose lines stay
    // inside the nodes
    constexpr int NUL = D == 3 ? 14 : 5;
    const bool lines = sizeof(TV) == 4 && sizeof(TX) == 4 && sizeof(TA) == 4 && Y.lines &&
                       Y.nu == NUL;
    int du[NUL];
    int64_t s_lo = 1, s_hi = -1;
    if (lines) {
#pragma unroll
        for (int k = 0; k < NUL; k++) du[k] = __builtin_amdgcn_readfirstlane(Y.u[k]);
        s_lo = 1 + (du[NUL - 1] >> 6);
        s_hi = (nn - 1 - du[NUL - 1] - 64) / 64;
    }
    auto visit = [&](int64_t slice) {
        const int64_t a = slice * 64 + lane;
        const bool live = a < nn;
        TA out[D];
        if constexpr (sizeof(TV) == 4 && sizeof(TX) == 4 && sizeof(TA) == 4) {
            if (lines && slice >= s_lo && slice <= s_hi) {
                float o[D];
                bsym_lines_node<D, NUL>(a, lane, slice, Y, (const float*)vals, (const float*)p,
                                        du, o);
#pragma unroll
                for (int i = 0; i < D; i++) {
                    const float qs = o[i];
                    st(&q[i * nn + a], qs, g_nt);
                    pq += (double)p[i * nn + a] * (double)qs;
                }
                return;
            }
        }
        bsym_node<D, TV, TX, TA, G>(live ? a : 0, live, lane, slice, Y, vals, p, out);
        if (live) {
#pragma unroll
            for (int i = 0; i < D; i++) {
                const TX qs = (TX)out[i];
                st(&q[i * nn + a], qs, g_nt);
                pq += (double)p[i * nn + a] * (double)qs;
            }
        }
    };
    const int bands = g_bsym_bands;
    if (bands > 0 && b.n % NXCD == 0) {
        // banded order: the XCD's slices as rows of P (the largest node
        // offset in slices, about one node plane), taken band by band of
        // P / bands columns, row after row, so that a slice and the slices
        // one plane back holding its mirrors are visited about P / bands
        // slices apart instead of P.  At 96^3 elasticity with 0 / 2 / 4 / 8
        // / 16 bands: FP32 SpMV 118.9 / 117.8 / 117.1 / 120.2 / 117.3 us,
        // FP64 SELL solve 397 / 379 / 372 / 375 / 368 ms per step
        const int step = SB / 64, xcd = b.id % NXCD, local = b.id / NXCD, nloc = b.n / NXCD;
        int per = (int)((nsl + NXCD - 1) / NXCD);
        per = (per + step - 1) / step * step;
        const int lo = xcd * per, hi = lo + per < (int)nsl ? lo + per : (int)nsl;
        const int P = Y.u[Y.nu - 1] / 64 > 1 ? Y.u[Y.nu - 1] / 64 : 1;
        const int R = hi > lo ? (hi - lo + P - 1) / P : 0;
        const int Wb = (P + bands - 1) / bands, RW = R * Wb;
        for (int v = local * step + wave_in_block(); v < bands * RW; v += nloc * step) {
            const int h = v / RW, t = v - h * RW, r = t / Wb, c = h * Wb + (t - r * Wb);
            const int slice = lo + r * P + c;
            if (c < P && slice < hi) visit(slice);
        }
        return pq;
    }
    const Sweep w = xcd_sweep(nsl, SB / 64, b, true);
    for (int64_t slice = w.first + wave_in_block(); slice < w.end; slice += w.stride) visit(slice);
    return pq;
}

// ---------------------------------------------------------------------------
// Operands of the residual replacement of mixed precision, carried by K1:
// t = b - A x64 with the FP64 (or, after an FP32 assembly, FP32) values in
// FP64 arithmetic; r32 = (float)t.
template <typename TVR>
struct Rep {
    const TVR* val;
    const double* x64;
    const double* b;
    float* r32;
    const float* dinv32;
    int sym;   // 1: val holds the symmetric slices of the SymView's layout
};

// the rows of a true residual t = b - A x64 (FP64 arithmetic) visited by one
// block, round-robin over the slices; f(row, t) takes each row's value.
// RS: the storages a kernel may meet (0 full slices, 1 scalar symmetric
// slices, 4 node-block slices, 2 any, chosen by sym and Y.dpn), so that a
// kernel carries the registers of its own forms only.
template <typename TVR, int RS, class F>
__device__ inline void resid_rows(Blk b, int64_t n, const SellView& S, const TVR* __restrict__ val,
                                  int sym, const SymView& Y, const double* __restrict__ x,
                                  const double* __restrict__ bv, F f)
{
    const int lane = threadIdx.x & 63;
    if (RS == 0) sym = 0;
    if (RS == 1 || RS == 4) sym = 1;
    if constexpr (RS == 4 || RS == 2) {
      if (RS == 4 || (sym && Y.dpn > 1)) {
        // node blocks: D rows per lane, component-major vectors
        const int64_t nn = Y.nn;
        const Sweep w = xcd_sweep((nn + 63) / 64, SB / 64, b);
        for (int64_t slice = w.first + wave_in_block(); slice < w.end; slice += w.stride) {
            const int64_t a = slice * 64 + lane;
            const bool live = a < nn;
            if (Y.dpn == 3) {
                double out[3];
                bsym_node<3, TVR, double, double, 1>(live ? a : 0, live, lane, slice, Y, val, x,
                                                     out);
                if (live)
#pragma unroll
                    for (int i = 0; i < 3; i++) f(i * nn + a, bv[i * nn + a] - out[i]);
            } else {
                double out[2];
                bsym_node<2, TVR, double, double, 1>(live ? a : 0, live, lane, slice, Y, val, x,
                                                     out);
                if (live)
#pragma unroll
                    for (int i = 0; i < 2; i++) f(i * nn + a, bv[i * nn + a] - out[i]);
            }
        }
        return;
      }
    }
    if constexpr (RS != 4) {
        // the symmetric slices sweep per-XCD ranges, as their SpMV does (a
        // mirror read then finds its slice in the same XCD's L2)
        const Sweep w = xcd_sweep((n + 63) / 64, SB / 64, b, RS != 0 && sym);
        for (int64_t slice = w.first + wave_in_block(); slice < w.end; slice += w.stride) {
            const int64_t rho = slice * 64 + lane;
            const bool live = rho < n;
            const int64_t row = live ? (S.perm ? (int64_t)S.perm[rho] : rho) : 0;
            double s;
            if (RS != 0 && sym)
                s = sym_row_v<TVR, double, double, 8>(row, live, n, lane, slice, val, Y.u, Y.nu,
                                                      x);
            else
                s = sell_row<TVR, double, double, 2>(row, live, n, lane, slice, S, val, x);
            if (live) f(row, bv[row] - s);
        }
    }
}

// K1: q = A p with per-block partials of p.q; at the replacement iteration
// (it == rep_at) the FP64 true residual instead, with partials of t.t and
// of r32.z
// SYM: 0 = full slices only, 1 = scalar symmetric slices (generic rows),
// 3 = scalar symmetric slices of stencil lines, 4 = node-block symmetric
// slices, 2 = any of them, chosen per call by Y (the batched serving
// kernel).  Separate instantiations keep each row form's registers and
// schedule out of the others' kernels (all share the replacement path).
// FP64 sweeps (TV = double) of problems with symmetric slices read their
// FP64 copy (SymView::v64), in the same row forms.
template <typename TV, typename TX, typename TA, int SGK, typename TVR, int SYM = 0>
__device__ inline void spmv_sell_body(Blk b, int64_t n, const SellView& S,
                                      const TV* __restrict__ val, const TX* __restrict__ p,
                                      TX* __restrict__ q, double* part, const Scal* sc, int it,
                                      const Rep<TVR>& R, const SymView& Y)
{
    const int lane = threadIdx.x & 63;
    const int64_t nsl = (n + 63) / 64;
    if (it == sc->rep_at) {
        // FP64 true residual, round-robin placement
        double tt = 0., rz = 0.;
        // (the line form in FP64 holds about 150 registers, which would cost
        // the FP32 sweep of this kernel its occupancy: generic rows here)
        resid_rows<TVR, SYM == 3 ? 1 : SYM>(b, n, S, R.val, R.sym, Y, R.x64, R.b,
                                            [&](int64_t row, double ti) {
            tt += ti * ti;
            const float rf = (float)ti;
            R.r32[row] = rf;
            rz += (double)rf * (double)(R.dinv32[row] * rf);
        });
        double v[2] = {tt, rz};
        const int sl[2] = {P_TT, P_RZ};
        block_partials<2>(v, part, sl, b);
        return;
    }
    if (sc->done || sc->stop) return;
    if constexpr (SYM == 4 || (sizeof(TV) == 4 && SYM == 2)) {
        if (SYM == 4 || (Y.v && Y.dpn > 1)) {
            double v[1] = {Y.dpn == 3 ? bsym_sweep<3, TV, TX, TA>(b, Y, p, q)
                                      : bsym_sweep<2, TV, TX, TA>(b, Y, p, q)};
            const int sl[1] = {P_PQ};
            block_partials<1>(v, part, sl, b);
            return;
        }
    }
    if constexpr (SYM == 3 ||
                  (SYM == 2 && sizeof(TV) == 4 && sizeof(TX) == 4 && sizeof(TA) == 4)) {
        if (SYM == 3 || (Y.v && Y.lines)) {
            const TV* __restrict__ sv = sym_vals<TV>(Y);
            double v[1] = {Y.nu == 14 ? sym_lines_sweep<14, TV, TX, TA>(b, n, Y, sv, p, q)
                                      : sym_lines_sweep<5, TV, TX, TA>(b, n, Y, sv, p, q)};
            const int sl[1] = {P_PQ};
            block_partials<1>(v, part, sl, b);
            return;
        }
    }
    if constexpr (SYM != 3 && SYM != 4) {
        // the symmetric slices sweep per-XCD ranges: a mirror read then finds
        // the slice it needs in the same XCD's L2
        const bool symv = SYM == 1 || (SYM == 2 && Y.v != nullptr);
        const Sweep w = xcd_sweep(nsl, SB / 64, b, symv);
        double pq = 0.;
        for (int64_t slice = w.first + wave_in_block(); slice < w.end; slice += w.stride) {
            const int64_t rho = slice * 64 + lane;
            const bool live = rho < n;
            const int64_t row = live ? (S.perm ? (int64_t)S.perm[rho] : rho) : 0;
            TA s;
            if constexpr (SYM == 1) {
                s = sym_row_v<TV, TX, TA>(row, live, n, lane, slice, sym_vals<TV>(Y), Y.u, Y.nu, p);
            } else if constexpr (sizeof(TV) == 4 && SYM == 2) {
                if (Y.v) s = sym_row<TX, TA>(row, live, n, lane, slice, Y, p);
                else s = sell_row<TV, TX, TA, SGK>(row, live, n, lane, slice, S, val, p);
            } else {
                s = sell_row<TV, TX, TA, SGK>(row, live, n, lane, slice, S, val, p);
            }
            if (live) {
                const TX qs = (TX)s;
                st(&q[row], qs, g_nt);
                pq += (double)p[row] * (double)qs;
            }
        }
        double v[1] = {pq};
        const int sl[1] = {P_PQ};
        block_partials<1>(v, part, sl, b);
    }
}

template <typename TV, typename TX, typename TA, int SGK, typename TVR, int SYM = 0>
__global__ __launch_bounds__(SB) void k_spmv_sell(int64_t n, SellView S, const TV* __restrict__ val,
                                                  const TX* __restrict__ p, TX* __restrict__ q,
                                                  double* part, const Scal* sc, int it, Rep<TVR> R,
                                                  SymView Y)
{
    spmv_sell_body<TV, TX, TA, SGK, TVR, SYM>(grid_blk(), n, S, val, p, q, part, sc, it, R, Y);
}

// The FP32 stencil-line SpMV held to the registers of WPE waves per SIMD.
// Uncapped it takes 97 VGPRs, allocated in granules of 8 as 104 (4 waves
// per SIMD); at 96 (WPE = 5) the 128^3 sweep takes 26.7 instead of 28.5 us
// on the same grid of 4 blocks per CU.  (The FP64 line rows held to 128
// registers for 4 waves spill, 92 bytes per lane, and run slower.)
template <int WPE, typename TV, typename TX, typename TA, int SGK, typename TVR, int SYM>
__global__ __launch_bounds__(SB) __attribute__((amdgpu_waves_per_eu(WPE))) void k_spmv_cap(
    int64_t n, SellView S, const TV* __restrict__ val, const TX* __restrict__ p,
    TX* __restrict__ q, double* part, const Scal* sc, int it, Rep<TVR> R, SymView Y)
{
    spmv_sell_body<TV, TX, TA, SGK, TVR, SYM>(grid_blk(), n, S, val, p, q, part, sc, it, R, Y);
}

template <int LPR, typename TV, typename TX, typename TA, typename TVR>
__global__ __launch_bounds__(SB) void k_spmv_csr(int64_t n, const int32_t* __restrict__ rp,
                                                 const int32_t* __restrict__ col,
                                                 const TV* __restrict__ val,
                                                 const TX* __restrict__ p, TX* __restrict__ q,
                                                 double* part, const Scal* sc, int it, Rep<TVR> R)
{
    const int sub = threadIdx.x & (LPR - 1);
    const Sweep w = xcd_sweep(n, SB / LPR);
    if (it == sc->rep_at) {
        double tt = 0., rz = 0.;
        for (int64_t row = w.first + threadIdx.x / LPR; row < w.end; row += w.stride) {
            const double s = csr_row<LPR, TVR, double, double>(row, sub, rp, col, R.val, R.x64);
            if (sub == 0) {
                const double ti = R.b[row] - s;
                tt += ti * ti;
                const float rf = (float)ti;
                R.r32[row] = rf;
                rz += (double)rf * (double)(R.dinv32[row] * rf);
            }
        }
        double v[2] = {tt, rz};
        const int sl[2] = {P_TT, P_RZ};
        block_partials<2>(v, part, sl);
        return;
    }
    if (sc->done || sc->stop) return;
    double pq = 0.;
    for (int64_t row = w.first + threadIdx.x / LPR; row < w.end; row += w.stride) {
        const TA s = csr_row<LPR, TV, TX, TA>(row, sub, rp, col, val, p);
        if (sub == 0) {
            const TX qs = (TX)s;
            st(&q[row], qs, g_nt);
            pq += (double)p[row] * (double)qs;
        }
    }
    double v[1] = {pq};
    const int sl[1] = {P_PQ};
    block_partials<1>(v, part, sl);
}

// 16-byte vector of a scalar type (4 floats or 2 doubles)
template <typename T>
struct V16;
template <>
struct V16<float> {
    typedef float type __attribute__((ext_vector_type(4)));
    static constexpr int W = 4;
};
template <>
struct V16<double> {
    typedef double type __attribute__((ext_vector_type(2)));
    static constexpr int W = 2;
};

// K2: alpha from the p.q partials; r -= alpha q; r.r, r.z partials (the
// x update waits for K3)
template <typename TX, typename TD>
__device__ inline void update_xr_body(Blk b, int64_t n, TX* __restrict__ x, TX* __restrict__ r,
                                      const TX* __restrict__ p, const TX* __restrict__ q,
                                      const TD* __restrict__ dinv, double* part, int g_pq,
                                      Scal* sc, int it)
{
    // everything this kernel waits on before its sweep -- the scalars of
    // the last iterations and this thread's share of K1's p.q partials --
    // is loaded at once (one memory latency instead of three in a row)
    const int rep_at = sc->rep_at, done = sc->done, stop = sc->stop;
    const double rho_it = sc->rho[it & 1];
    double tpq[1];
    {
        const int slq[1] = {P_PQ};
        totals_share<1>(part, g_pq, slq, tpq);
    }
    if (it == rep_at) {
        // replacement iteration: accept or continue from the true residual
        // of K1 (one block; the others have nothing to do)
        if (b.id != 0) return;
        double t[2];
        const int sl[2] = {P_TT, P_RZ};
        totals<2>(part, g_pq, sl, t);
        if (threadIdx.x == 0) {
            sc->tt = t[0];
            sc->nrep += 1;
            if (!isfinite(t[0])) {
                sc->breakdown = 1;
                sc->done = 1;
            } else if (t[0] < sc->thr) {
                sc->done = 1;
                sc->fin = 1;
            } else {
                // keep the direction while the recursion still tracks the
                // true residual; once the FP32 recursion has converged or
                // drifted, restart from p = D^-1 r
                sc->restart = (sc->conv32 || t[0] > g_restart_ratio * sc->rr) ? 1 : 0;
                sc->rho[(it + 1) & 1] = t[1];
                sc->rr = t[0];
                sc->rref = t[0];
            }
            sc->conv32 = 0;
        }
        return;
    }
    if (done) return;
    if (stop) {
        // converged in the last K3: done from here on
        if (leader(b)) sc->done = 1;
        return;
    }
    block_allsum<1>(tpq);
    const double pq = tpq[0];
    if (!(pq > 0.) || !isfinite(pq)) {
        if (leader(b)) {
            sc->pq = pq;
            sc->breakdown = 1;
            sc->done = 1;
        }
        return;
    }
    const double ad = rho_it / pq;
    if (leader(b)) {
        sc->pq = pq;
        sc->alpha = ad;
        sc->iters += 1;
    }
    static_assert(sizeof(TX) == sizeof(TD), "vectors and diagonal share a type");
    typedef typename V16<TX>::type VT;
    constexpr int W = V16<TX>::W;
    const TX alpha = (TX)ad;
    double rr = 0., rz = 0.;
    // 16-byte accesses over the full vectors, then the n % W tail
    const int64_t nv = n / W;
    const Sweep w = xcd_sweep(nv, VB, b);
    (void)x;  // x += alpha p is done by K3, which reads p anyway
    (void)p;
    VT* __restrict__ rv = reinterpret_cast<VT*>(r);
    const VT* __restrict__ qv = reinterpret_cast<const VT*>(q);
    const VT* __restrict__ dv = reinterpret_cast<const VT*>(dinv);
    for (int64_t i = w.first + threadIdx.x; i < w.end; i += w.stride) {
        const VT ri = rv[i] - alpha * qv[i];
        rv[i] = ri;
        const VT zi = dv[i] * ri;
#pragma unroll
        for (int k = 0; k < W; k++) {
            rr += (double)ri[k] * (double)ri[k];
            rz += (double)ri[k] * (double)zi[k];
        }
    }
    for (int64_t i = nv * W + (int64_t)b.id * VB + threadIdx.x; i < n; i += (int64_t)b.n * VB) {
        const TX ri = r[i] - alpha * q[i];
        r[i] = ri;
        const TX zi = (TX)dinv[i] * ri;
        rr += (double)ri * (double)ri;
        rz += (double)ri * (double)zi;
    }
    double v[2] = {rr, rz};
    const int sl[2] = {P_RR, P_RZ};
    block_partials<2>(v, part, sl, b);
}

template <typename TX, typename TD>
__global__ __launch_bounds__(VB) void k_update_xr(int64_t n, TX* __restrict__ x, TX* __restrict__ r,
                                                  const TX* __restrict__ p,
                                                  const TX* __restrict__ q,
                                                  const TD* __restrict__ dinv, double* part,
                                                  int g_pq, Scal* sc, int it)
{
    update_xr_body<TX, TD>(grid_blk(), n, x, r, p, q, dinv, part, g_pq, sc, it);
}

// K3: stop test; x += alpha p; beta; p = D^-1 r + beta p.  In mixed precision, once the
// recursive residual fell below delta^2 of the last true one, or below the
// tolerance, the next iteration is a replacement: every block folds its rows
// of x32 into x64 here and the leader sets rep_at.  At the replacement
// iteration K3 restarts the direction if K2 asked for it.
template <typename TX, typename TD>
__device__ inline void update_p_body(Blk b, int64_t n, TX* __restrict__ p,
                                     const TX* __restrict__ r, const TD* __restrict__ dinv,
                                     const double* part, int g, Scal* sc, int it,
                                     double* __restrict__ x64, TX* __restrict__ x)
{
    typedef typename V16<TX>::type VT;
    constexpr int W = V16<TX>::W;
    constexpr bool mixed = sizeof(TX) == 4;
    const int64_t nv = n / W;
    const Sweep w = xcd_sweep(nv, VB, b);
    VT* __restrict__ xv = reinterpret_cast<VT*>(x);
    VT* __restrict__ pv = reinterpret_cast<VT*>(p);
    const VT* __restrict__ rv = reinterpret_cast<const VT*>(r);
    const VT* __restrict__ dv = reinterpret_cast<const VT*>(dinv);
    // the scalars and this thread's share of K2's partials, loaded at once
    const int rep_at = sc->rep_at, restart = sc->restart, done = sc->done;
    const int par = it & 1;
    const double rho = sc->rho[par], thr = sc->thr, delta2 = sc->delta2, rref = sc->rref;
    const double alpha_d = sc->alpha;
    double t[2];
    {
        const int sl[2] = {P_RR, P_RZ};
        totals_share<2>(part, g, sl, t);
    }
    if (it == rep_at) {
        if (!restart || done) return;
        for (int64_t i = w.first + threadIdx.x; i < w.end; i += w.stride) pv[i] = dv[i] * rv[i];
        for (int64_t i = nv * W + (int64_t)b.id * VB + threadIdx.x; i < n; i += (int64_t)b.n * VB)
            p[i] = (TX)dinv[i] * r[i];
        return;
    }
    if (done) return;
    block_allsum<2>(t);
    const double rr = t[0], rz = t[1];
    const bool bad = !isfinite(rr);
    const bool conv = rr < thr;
    const bool pause = mixed && !bad && !conv && rr < delta2 * rref;
    if (leader(b)) {
        sc->rr = rr;
        if (bad) {
            sc->breakdown = 1;
            sc->done = 1;
        } else if (conv && !mixed) {
            sc->stop = 1;
        } else if (conv || pause) {
            sc->conv32 = conv ? 1 : 0;
            sc->rep_at = it + 1;
        }
        if (!bad && !conv) {
            sc->rho[par ^ 1] = rz;
            sc->beta = rz / rho;
        }
    }
    if (bad) return;
    // x += alpha p with the direction of this iteration's SpMV (alpha from
    // K2), done here where p is read anyway; then fold x32 into x64 before
    // a replacement, and p = D^-1 r + beta p
    const TX alpha = (TX)alpha_d;
    if (mixed && (conv || pause)) {
        // every item is read and written by one thread in one visit (the p
        // update of another thread's item must not precede its x update)
        const TX beta = conv ? (TX)0 : (TX)(rz / rho);
        for (int64_t i = w.first + threadIdx.x; i < w.end; i += w.stride) {
            const VT pi = pv[i];
            const VT xi = xv[i] + alpha * pi;
#pragma unroll
            for (int k = 0; k < W; k++) x64[i * W + k] += (double)xi[k];
            xv[i] = VT{};
            if (!conv) pv[i] = dv[i] * rv[i] + beta * pi;
        }
        for (int64_t i = nv * W + (int64_t)b.id * VB + threadIdx.x; i < n; i += (int64_t)b.n * VB) {
            const TX pi = p[i];
            const TX xi = x[i] + alpha * pi;
            x64[i] += (double)xi;
            x[i] = (TX)0;
            if (!conv) p[i] = (TX)dinv[i] * r[i] + beta * pi;
        }
        return;
    } else if (conv) {
        for (int64_t i = w.first + threadIdx.x; i < w.end; i += w.stride) xv[i] = xv[i] + alpha * pv[i];
        for (int64_t i = nv * W + (int64_t)b.id * VB + threadIdx.x; i < n; i += (int64_t)b.n * VB)
            x[i] = x[i] + alpha * p[i];
        return;
    } else {
        const TX beta = (TX)(rz / rho);
        for (int64_t i = w.first + threadIdx.x; i < w.end; i += w.stride) {
            const VT pi = pv[i];
            xv[i] = xv[i] + alpha * pi;
            pv[i] = dv[i] * rv[i] + beta * pi;
        }
        for (int64_t i = nv * W + (int64_t)b.id * VB + threadIdx.x; i < n; i += (int64_t)b.n * VB) {
            const TX pi = p[i];
            x[i] = x[i] + alpha * pi;
            p[i] = (TX)dinv[i] * r[i] + beta * pi;
        }
    }
}

template <typename TX, typename TD>
__global__ __launch_bounds__(VB) void k_update_p(int64_t n, TX* __restrict__ p,
                                                 const TX* __restrict__ r,
                                                 const TD* __restrict__ dinv, const double* part,
                                                 int g, Scal* sc, int it, double* x64, TX* x)
{
    update_p_body<TX, TD>(grid_blk(), n, p, r, dinv, part, g, sc, it, x64, x);
}

// start: x = 0, r = b, p = D^-1 b; partials of b.b and r.z
template <typename TX, typename TD>
__global__ __launch_bounds__(VB) void k_init(int64_t n, const double* __restrict__ b,
                                             TX* __restrict__ x, TX* __restrict__ r,
                                             TX* __restrict__ p, const TD* __restrict__ dinv,
                                             double* part)
{
    double bb = 0., rz = 0.;
    const Sweep w = xcd_sweep(n, VB);
    for (int64_t i = w.first + threadIdx.x; i < w.end; i += w.stride) {
        const TX bi = (TX)b[i];
        x[i] = (TX)0;
        r[i] = bi;
        const TX zi = (TX)dinv[i] * bi;
        p[i] = zi;
        bb += b[i] * b[i];
        rz += (double)bi * (double)zi;
    }
    double v[2] = {bb, rz};
    const int sl[2] = {P_BB, P_RZ};
    block_partials<2>(v, part, sl);
}

__global__ __launch_bounds__(VB) void k_init_fin(const double* part, int g, double rtol2,
                                                 double delta2, Scal* sc)
{
    double t[2];
    const int sl[2] = {P_BB, P_RZ};
    totals<2>(part, g, sl, t);
    if (threadIdx.x == 0) {
        sc->bb = t[0];
        sc->rr = t[0];
        sc->tt = t[0];
        sc->thr = rtol2 * t[0];
        sc->rho[0] = t[1];
        sc->rho[1] = t[1];
        sc->iters = 0;
        sc->breakdown = 0;
        sc->delta2 = delta2;
        sc->done = (t[0] == 0.) ? 1 : 0;
        sc->stop = 0;
        sc->fin = sc->done;
        sc->restart = 0;
        sc->rep_at = INT_MAX;
        sc->conv32 = 0;
        sc->nrep = 0;
        sc->rref = t[0];
    }
}

// ---- true residual at the end of a solve -------------------------------------------
// x64 += x32 (mixed precision, when a solve stops at its iteration limit)
__global__ __launch_bounds__(VB) void k_fold(int64_t n, double* __restrict__ x64,
                                             float* __restrict__ x32)
{
    const Sweep w = xcd_sweep(n, VB);
    for (int64_t i = w.first + threadIdx.x; i < w.end; i += w.stride) {
        x64[i] += (double)x32[i];
        x32[i] = 0.f;
    }
}

// t = b - A x in FP64; partials of t.t (on the symmetric slices when sym)
template <typename TV>
__global__ __launch_bounds__(SB) void k_resid_sell(int64_t n, SellView S, const TV* __restrict__ val,
                                                   int sym, SymView Y, const double* __restrict__ x,
                                                   const double* __restrict__ b, double* part)
{
    double tt = 0.;
    resid_rows<TV, 2>(grid_blk(), n, S, val, sym, Y, x, b,
                      [&](int64_t, double ti) { tt += ti * ti; });
    double v[1] = {tt};
    const int sl[1] = {P_TT};
    block_partials<1>(v, part, sl);
}

template <int LPR, typename TV>
__global__ __launch_bounds__(SB) void k_resid_csr(int64_t n, const int32_t* __restrict__ rp,
                                                  const int32_t* __restrict__ col,
                                                  const TV* __restrict__ val,
                                                  const double* __restrict__ x,
                                                  const double* __restrict__ b, double* part)
{
    const int sub = threadIdx.x & (LPR - 1);
    double tt = 0.;
    const Sweep w = xcd_sweep(n, SB / LPR);
    for (int64_t row = w.first + threadIdx.x / LPR; row < w.end; row += w.stride) {
        const double s = csr_row<LPR, TV, double, double>(row, sub, rp, col, val, x);
        if (sub == 0) {
            const double ti = b[row] - s;
            tt += ti * ti;
        }
    }
    double v[1] = {tt};
    const int sl[1] = {P_TT};
    block_partials<1>(v, part, sl);
}

// |b - A x|^2 into the scalars
__global__ __launch_bounds__(VB) void k_resid_fin(const double* part, int g, Scal* sc)
{
    double t[1];
    const int sl[1] = {P_TT};
    totals<1>(part, g, sl, t);
    if (threadIdx.x == 0) sc->tt = t[0];
}

__global__ __launch_bounds__(VB) void k_norm2(int64_t n, const double* __restrict__ v, double* part)
{
    double s = 0.;
    const Sweep w = xcd_sweep(n, VB);
    for (int64_t i = w.first + threadIdx.x; i < w.end; i += w.stride)
        s += v[i] * v[i];
    double a[1] = {s};
    const int sl[1] = {P_TT};
    block_partials<1>(a, part, sl);
}

__global__ __launch_bounds__(VB) void k_norm2_fin(const double* part, int g, Scal* sc)
{
    double t[1];
    const int sl[1] = {P_TT};
    totals<1>(part, g, sl, t);
    if (threadIdx.x == 0) sc->tt = t[0];
}

// ---- SELL construction --------------------------------------------------------
constexpr int MAXU = 512;  // longest per-slice offset list kept

__device__ inline int wave_min_i(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}

// union of the column-offset lists (col - row) of the 64 rows of a slice,
// ascending: one wavefront per slice, repeated wave minimum of list heads
__global__ __launch_bounds__(256) void k_slice_union(int64_t n, const int32_t* __restrict__ perm,
                                                     const int32_t* __restrict__ rp,
                                                     const int32_t* __restrict__ col,
                                                     int32_t* __restrict__ udel,
                                                     int32_t* __restrict__ ulen, int* err)
{
    const int64_t nsl = (n + 63) / 64;
    const int lane = threadIdx.x & 63;
    for (int64_t slice = (int64_t)blockIdx.x * 4 + wave_in_block(); slice < nsl;
         slice += (int64_t)gridDim.x * 4) {
        const int64_t rho = slice * 64 + lane;
        int64_t row = 0;
        int32_t k = 0, e = 0;
        if (rho < n) {
            row = perm ? (int64_t)perm[rho] : rho;
            k = rp[row];
            e = rp[row + 1];
        }
        int cnt = 0;
        while (true) {
            const int head = (k < e) ? col[k] - (int32_t)row : INT32_MAX;
            const int m = wave_min_i(head);
            if (m == INT32_MAX) break;
            if (lane == 0 && cnt < MAXU) udel[slice * MAXU + cnt] = m;
            cnt++;
            if (head == m) k++;
        }
        if (lane == 0) {
            ulen[slice] = cnt;
            if (cnt > MAXU) atomicOr(err, 4);
        }
    }
}

// component of every row (vector problems)
__global__ void k_row_comp(int64_t nd, const int32_t* __restrict__ dof_id, int dpn,
                           int32_t* __restrict__ comp)
{
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nd;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int32_t id = dof_id[k];
        if (id >= 0) comp[id] = (int32_t)(k % dpn);
    }
}

__global__ void k_comp_flag(int64_t n, const int32_t* __restrict__ comp, int c,
                            int32_t* __restrict__ flag)
{
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
         r += (int64_t)gridDim.x * blockDim.x)
        flag[r] = comp[r] == c;
}

// rows of component c, ascending, after `off` earlier rows
__global__ void k_comp_place(int64_t n, const int32_t* __restrict__ comp, int c,
                             const int32_t* __restrict__ pos, int64_t off,
                             int32_t* __restrict__ perm)
{
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
         r += (int64_t)gridDim.x * blockDim.x)
        if (comp[r] == c) perm[off + pos[r]] = (int32_t)r;
}

// rows of component c in the solver order of the field nodes
__global__ void k_ord_flag(int64_t nn, const int32_t* __restrict__ order,
                           const int32_t* __restrict__ dof_id, int dpn, int c,
                           int32_t* __restrict__ flag)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn;
         i += (int64_t)gridDim.x * blockDim.x)
        flag[i] = dof_id[(int64_t)order[i] * dpn + c] >= 0;
}

__global__ void k_ord_place(int64_t nn, const int32_t* __restrict__ order,
                            const int32_t* __restrict__ dof_id, int dpn, int c,
                            const int32_t* __restrict__ pos, int64_t off,
                            int32_t* __restrict__ perm)
{
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nn;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t id = dof_id[(int64_t)order[i] * dpn + c];
        if (id >= 0) perm[off + pos[i]] = id;
    }
}

// columns of the permuted matrix for any permutation: one wavefront per row,
// each entry's rank among the row's permuted columns by counting (columns of
// a row are distinct, rows hold a few hundred entries at most)
__global__ __launch_bounds__(256) void k_prow_fill_sorted(int64_t n,
                                                          const int32_t* __restrict__ perm,
                                                          const int32_t* __restrict__ iperm,
                                                          const int32_t* __restrict__ rp,
                                                          const int32_t* __restrict__ col,
                                                          const int32_t* __restrict__ prp,
                                                          int32_t* __restrict__ pcol,
                                                          int32_t* __restrict__ pent)
{
    const int lane = threadIdx.x & 63;
    for (int64_t r = (int64_t)blockIdx.x * 4 + wave_in_block(); r < n;
         r += (int64_t)gridDim.x * 4) {
        const int32_t row = perm[r];
        const int32_t k0 = rp[row], k1 = rp[row + 1];
        const int32_t o = prp[r];
        for (int32_t kb = k0; kb < k1; kb += 64) {
            const int32_t k = kb + lane;
            const int32_t pc = (k < k1) ? iperm[col[k]] : INT32_MAX;
            int32_t rank = 0;
            for (int32_t i = k0; i < k1; i++) rank += iperm[col[i]] < pc;
            if (k < k1) {
                pcol[o + rank] = pc;
                pent[o + rank] = k;
            }
        }
    }
}

// inverse of the slice-row permutation
__global__ void k_iperm(int64_t n, const int32_t* __restrict__ perm, int32_t* __restrict__ iperm)
{
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
         r += (int64_t)gridDim.x * blockDim.x)
        iperm[perm[r]] = (int32_t)r;
}

// row lengths of the permuted matrix P A P^T
__global__ void k_prow_len(int64_t n, const int32_t* __restrict__ perm,
                           const int32_t* __restrict__ rp, int32_t* __restrict__ len)
{
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int32_t row = perm[r];
        len[r] = rp[row + 1] - rp[row];
    }
}

// columns of the permuted matrix, ascending, with the original CSR entry of
// each.  Permuted indices order the rows by component, then by row, so a
// row's permuted columns are its columns of component 0 (in their order),
// then of component 1, ...  One wavefront per row: ballots give every entry
// its rank among the entries of its component.
__global__ __launch_bounds__(256) void k_prow_fill(int64_t n, int dpn,
                                                   const int32_t* __restrict__ perm,
                                                   const int32_t* __restrict__ iperm,
                                                   const int32_t* __restrict__ comp,
                                                   const int32_t* __restrict__ rp,
                                                   const int32_t* __restrict__ col,
                                                   const int32_t* __restrict__ prp,
                                                   int32_t* __restrict__ pcol,
                                                   int32_t* __restrict__ pent)
{
    const int lane = threadIdx.x & 63;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    for (int64_t r = (int64_t)blockIdx.x * 4 + wave_in_block(); r < n;
         r += (int64_t)gridDim.x * 4) {
        const int32_t row = perm[r];
        const int32_t k0 = rp[row], k1 = rp[row + 1];
        const int32_t o = prp[r];
        int cnt[4] = {0, 0, 0, 0};
        for (int32_t kb = k0; kb < k1; kb += 64) {
            const int32_t k = kb + lane;
            const int cc = (k < k1) ? comp[col[k]] : -1;
            for (int c = 0; c < dpn; c++) cnt[c] += __popcll(__ballot(cc == c));
        }
        int base[4] = {0, 0, 0, 0};
        for (int c = 1; c < dpn; c++) base[c] = base[c - 1] + cnt[c - 1];
        for (int32_t kb = k0; kb < k1; kb += 64) {
            const int32_t k = kb + lane;
            const int32_t cl = (k < k1) ? col[k] : 0;
            const int cc = (k < k1) ? comp[cl] : -1;
            for (int c = 0; c < dpn; c++) {
                const uint64_t m = __ballot(cc == c);
                if (cc == c) {
                    const int32_t d = o + base[c] + __popcll(m & below);
                    pcol[d] = iperm[cl];
                    pent[d] = k;
                }
                base[c] += __popcll(m);
            }
        }
    }
}

// x[perm[r]] = xp[r] / xp[r] = x[perm[r]]
template <typename T>
__global__ void k_perm_out(int64_t n, const int32_t* __restrict__ perm, const T* __restrict__ xp,
                           T* __restrict__ x)
{
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
         r += (int64_t)gridDim.x * blockDim.x)
        x[perm[r]] = xp[r];
}

template <typename T>
__global__ void k_perm_in(int64_t n, const int32_t* __restrict__ perm, const T* __restrict__ x,
                          T* __restrict__ xp)
{
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
         r += (int64_t)gridDim.x * blockDim.x)
        xp[r] = x[perm[r]];
}

__device__ inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
    return z ^ (z >> 31);
}

__global__ void k_list_hash(int64_t m, const int32_t* __restrict__ udel,
                            const int32_t* __restrict__ ulen, uint64_t* __restrict__ h)
{
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < m;
         s += (int64_t)gridDim.x * blockDim.x) {
        const int len = min(ulen[s], MAXU);
        uint64_t z = mix64(0x9e3779b97f4a7c15ULL + (uint64_t)len);
        for (int k = 0; k < len; k++)
            z = mix64(z ^ (uint64_t)(uint32_t)udel[s * MAXU + k] * 0x100000001b3ULL);
        h[s] = z | 1ULL;  // never 0 (0 marks an empty slot)
    }
}

// open-addressing table of distinct list hashes; value = a representative
__global__ void k_hash_insert(int64_t m, const uint64_t* __restrict__ h, uint64_t* keys,
                              int32_t* reps, uint64_t mask, int* err)
{
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < m;
         r += (int64_t)gridDim.x * blockDim.x) {
        const uint64_t key = h[r];
        uint64_t slot = key & mask;
        for (uint64_t probe = 0; probe <= mask; probe++) {
            // most lists repeat: read first, CAS only empty slots
            const uint64_t cur = keys[slot];
            if (cur == key) break;
            if (cur != 0ULL) {
                slot = (slot + 1) & mask;
                continue;
            }
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

__global__ void k_hash_find(int64_t m, const uint64_t* __restrict__ h,
                            const uint64_t* __restrict__ keys, uint64_t mask,
                            int32_t* __restrict__ slot_of)
{
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < m;
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

// pattern id of every slice; check that the list equals its representative
__global__ void k_list_pattern(int64_t m, const int32_t* __restrict__ udel,
                               const int32_t* __restrict__ ulen,
                               const int32_t* __restrict__ slot_of,
                               const int32_t* __restrict__ slot_id,
                               const int32_t* __restrict__ reps, int32_t* __restrict__ pat,
                               int* err)
{
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < m;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int32_t s = slot_of[r];
        const int64_t rep = reps[s];
        pat[r] = slot_id[s];
        const int len = ulen[r];
        if (ulen[rep] != len) {
            atomicOr(err, 2);
            continue;
        }
        for (int k = 0; k < min(len, MAXU); k++)
            if (udel[r * MAXU + k] != udel[rep * MAXU + k]) {
                atomicOr(err, 2);
                break;
            }
    }
}

__global__ void k_pat_len(uint64_t cap, const uint64_t* __restrict__ keys,
                          const int32_t* __restrict__ slot_id, const int32_t* __restrict__ reps,
                          const int32_t* __restrict__ ulen, int32_t* __restrict__ plen)
{
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < (int64_t)cap;
         s += (int64_t)gridDim.x * blockDim.x)
        if (keys[s]) plen[slot_id[s]] = min(ulen[reps[s]], MAXU);
}

__global__ void k_pat_fill(uint64_t cap, const uint64_t* __restrict__ keys,
                           const int32_t* __restrict__ slot_id, const int32_t* __restrict__ reps,
                           const int32_t* __restrict__ udel, const int32_t* __restrict__ ulen,
                           const int32_t* __restrict__ pptr, int32_t* __restrict__ pdel)
{
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < (int64_t)cap;
         s += (int64_t)gridDim.x * blockDim.x)
        if (keys[s]) {
            const int64_t rep = reps[s];
            const int32_t o = pptr[slot_id[s]];
            const int len = min(ulen[rep], MAXU);
            for (int k = 0; k < len; k++) pdel[o + k] = udel[rep * MAXU + k];
        }
}

__global__ void k_slice_vals(int64_t nsl, const int32_t* __restrict__ ulen, int32_t* __restrict__ sv)
{
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < nsl;
         s += (int64_t)gridDim.x * blockDim.x)
        sv[s] = ((ulen[s] + 3) & ~3) * 64;
}

__global__ void k_widen(int64_t n, const int32_t* __restrict__ a, int64_t* __restrict__ b)
{
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < n;
         s += (int64_t)gridDim.x * blockDim.x)
        b[s] = a[s];
}

// position of every CSR entry inside the slices: lane = row, walk the
// slice's offset list and the row's columns together (setup, once)
// Slice position of every CSR entry: one wavefront per row, lanes over the
// row's entries, each finding its column offset in the slice's (ascending)
// offset list by bisection.
__global__ __launch_bounds__(256) void k_sell_map(int64_t n, SellView S,
                                                  const int32_t* __restrict__ rp,
                                                  const int32_t* __restrict__ col,
                                                  const int32_t* __restrict__ ent,
                                                  int32_t* __restrict__ map, int* err)
{
    const int lane = threadIdx.x & 63;
    for (int64_t rho = (int64_t)blockIdx.x * 4 + wave_in_block(); rho < n;
         rho += (int64_t)gridDim.x * 4) {
        const int64_t slice = rho >> 6;
        const int lp = (int)(rho & 63);
        const int pid = S.spat[slice];
        const int32_t* d = S.pdel + S.pptr[pid];
        const int len = S.plen[pid];
        const int64_t base = S.sptr[slice] + lp * 4;
        const int32_t k0 = rp[rho], k1 = rp[rho + 1];
        for (int32_t k = k0 + lane; k < k1; k += 64) {
            const int32_t delta = col[k] - (int32_t)rho;
            int lo = 0, hi = len;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (d[mid] < delta) lo = mid + 1;
                else hi = mid;
            }
            if (lo < len && d[lo] == delta)
                map[ent ? ent[k] : k] = (int32_t)(base + (int64_t)(lo >> 2) * 256 + (lo & 3));
            else
                atomicOr(err, 8);
        }
    }
}

// position in the symmetric slices of every CSR entry (row, col >= row):
// offset col - row looked up in the shared ascending list; -1 below the
// diagonal.  err: an offset missing from the list.
__global__ void k_sym_map(int64_t n, const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                          const int32_t* __restrict__ su, int nu, int32_t* __restrict__ smap,
                          int* err)
{
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
         r += (int64_t)gridDim.x * blockDim.x) {
        for (int32_t e = rp[r]; e < rp[r + 1]; e++) {
            const int64_t d = (int64_t)col[e] - r;
            if (d < 0) {
                smap[e] = -1;
                continue;
            }
            int lo = 0, hi = nu - 1, k = -1;
            while (lo <= hi) {
                const int mid = (lo + hi) >> 1;
                if (su[mid] == d) {
                    k = mid;
                    break;
                }
                if (su[mid] < d) lo = mid + 1;
                else hi = mid - 1;
            }
            if (k < 0) {
                smap[e] = -1;
                atomicOr(err, 1);
                continue;
            }
            smap[e] = (int32_t)((r >> 6) * 64 * (int64_t)nu + (int64_t)k * 64 + (r & 63));
        }
    }
}

// Shared-element words of the stored slots of the symmetric slices, for the
// assembly straight into them: slot (a, k) couples active node a with active
// node a + u_k; bits [4i, 4i + 4) of its word hold, for the i-th element of
// node a's adjacency list (ascending element order), a presence bit (8) and
// the local index of node a + u_k in that element, as the pair words of the
// CSR gather do (0 for padding slots).  Words are k-major within a slice of
// 64 nodes, like the values.
template <int NPE>
__global__ void k_sym_words(int64_t N, int nu, const int32_t* __restrict__ su,
                            const int32_t* __restrict__ anode, const int32_t* __restrict__ adj_ptr,
                            const int32_t* __restrict__ adj, const int32_t* __restrict__ conn,
                            int lg_npe, uint32_t* __restrict__ spw)
{
    constexpr int BITS = PairWord<NPE>::bits, FLAG = PairWord<NPE>::flag;
    static_assert(NPE <= 8, "32-bit words");
    const int64_t total = (N + 63) / 64 * 64 * nu;
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < total;
         p += (int64_t)gridDim.x * blockDim.x) {
        const int64_t sl = p / (64 * (int64_t)nu);
        const int k = (int)((p / 64) % nu);
        const int64_t a = sl * 64 + (p & 63);
        uint32_t w = 0;
        if (a < N && a + su[k] < N) {
            const int32_t A = anode[a], B = anode[a + su[k]];
            const int32_t as = adj_ptr[A], ae = adj_ptr[A + 1];
            for (int32_t i = 0; i < ae - as; i++) {
                const int64_t e = adj[as + i] >> lg_npe;
#pragma unroll
                for (int t = 0; t < NPE; t++)
                    if (conn[e * NPE + t] == B) w |= (uint32_t)(FLAG + t) << (BITS * i);
            }
        }
        spw[p] = w;
    }
}

// mesh node of every active node (all its D components active, ids a*D+c)
__global__ void k_anode(int64_t n_nodes, int D, const int32_t* __restrict__ dof_id,
                        int32_t* __restrict__ anode)
{
    for (int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; m < n_nodes;
         m += (int64_t)gridDim.x * blockDim.x) {
        const int32_t id = dof_id[m * D];
        if (id >= 0) anode[id / D] = (int32_t)m;
    }
}

// slice position of every CSR entry in the node-block slices (-1: below
// the diagonal, read at its mirror); rows interleaved, r = a * D + i
__global__ void k_bsym_map(int64_t n, int D, const int32_t* __restrict__ rp,
                           const int32_t* __restrict__ col, const int32_t* __restrict__ su, int nu,
                           int32_t* __restrict__ smap, int* err)
{
    const SymLayout L{D, nu};
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t a = r / D;
        const int i = (int)(r - a * D);
        for (int32_t e = rp[r]; e < rp[r + 1]; e++) {
            const int64_t bn = col[e] / D;
            const int j = (int)(col[e] - bn * D);
            const int64_t d = bn - a;
            if (d < 0 || (d == 0 && j < i)) {
                smap[e] = -1;
                continue;
            }
            int lo = 0, hi = nu - 1, k = -1;
            while (lo <= hi) {
                const int mid = (lo + hi) >> 1;
                if (su[mid] == d) {
                    k = mid;
                    break;
                }
                if (su[mid] < d) lo = mid + 1;
                else hi = mid - 1;
            }
            if (k < 0) {
                smap[e] = -1;
                atomicOr(err, 1);
                continue;
            }
            smap[e] = (int32_t)L.pos(a, L.slot(k, i, j));
        }
    }
}

__global__ void k_sym_scatter(int64_t nnz, const int32_t* __restrict__ smap,
                              const double* __restrict__ v64, const float* __restrict__ v32,
                              double* __restrict__ vs64, float* __restrict__ vs)
{
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nnz;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int32_t o = smap[k];
        if (o < 0) continue;
        if (vs64) vs64[o] = v64[k];
        vs[o] = v32[k];
    }
}

// CSR values from the symmetric slices (the assembly wrote only those): row
// r = a * D + i of node a, column c = b * D + j; the block (a, b) is stored
// at node a when b - a = u_k >= 0 (for b = a only its upper triangle), else
// transposed at node b
template <int D>
__global__ void k_sym_to_csr(int64_t n, const int32_t* __restrict__ rp,
                             const int32_t* __restrict__ col, const int32_t* __restrict__ su,
                             int nu, const double* __restrict__ vs64,
                             const float* __restrict__ vs, double* __restrict__ v64,
                             float* __restrict__ v32)
{
    const SymLayout L{D, nu};
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int64_t a = r / D;
        const int i = (int)(r - a * D);
        for (int32_t e = rp[r]; e < rp[r + 1]; e++) {
            const int64_t bn = col[e] / D;
            const int j = (int)(col[e] - bn * D);
            const bool up = bn > a || (bn == a && j >= i);
            const int64_t owner = up ? bn - a : a - bn;   // the offset u_k
            int lo = 0, hi = nu - 1, k = 0;
            while (lo <= hi) {
                const int mid = (lo + hi) >> 1;
                if (su[mid] == owner) {
                    k = mid;
                    break;
                }
                if (su[mid] < owner) lo = mid + 1;
                else hi = mid - 1;
            }
            const int64_t node = up ? a : bn;
            const int ii = up ? i : j, jj = up ? j : i;
            const int64_t pos = L.pos(node, L.slot(k, ii, jj));
            if (vs64) v64[e] = vs64[pos];
            v32[e] = vs64 ? (float)vs64[pos] : vs[pos];
        }
    }
}

// copy CSR values to their slice positions (padding stays zero)
__global__ void k_sell_scatter(int64_t nnz, const int32_t* __restrict__ map,
                               const double* __restrict__ v64, const float* __restrict__ v32,
                               double* __restrict__ s64, float* __restrict__ s32)
{
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nnz;
         k += (int64_t)gridDim.x * blockDim.x) {
        const int32_t o = map[k];
        if (s64) s64[o] = v64[k];
        s32[o] = v32[k];
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

unsigned round8(int64_t g)
{
    if (g <= 8) return (unsigned)std::max<int64_t>(1, g);
    return (unsigned)((g + 7) / 8 * 8);
}

int env_int(const char* name, int dflt);

// blocks of the vector kernels: at most INS_VEC_BLOCKS (256; 8..2048).  At
// 128^3 the mixed step takes 35.3 ms with 192-256 blocks, 35.8 with 512,
// 36.7 with 128 and 42.7 with 2048 (every block sums the partials of the
// kernel before)
unsigned vec_grid(int64_t n)
{
    static const int cap = std::max(8, std::min(MAXG, env_int("INS_VEC_BLOCKS", 256)));
    int64_t g = (n + VB - 1) / VB;
    return round8(std::max<int64_t>(1, std::min<int64_t>(g, cap)));
}

unsigned sell_grid(int64_t n)
{
    const int64_t nsl = (n + 63) / 64;
    int64_t g = (nsl + SB / 64 - 1) / (SB / 64);
    return round8(std::max<int64_t>(1, std::min<int64_t>(g, MAXG)));
}

// Blocks of the symmetric SpMV: each XCD sweeps its eighth of the slices
// with its blocks' 4 waves per step, bpc blocks on every CU of the XCD.
// Scalar rows take 4 per CU: measured at 128^3 (n = 2.13M, 1024 vs 1280
// blocks on one box), the line-sharing rows 24.2 vs 31.5 us, the generic
// rows 30.9 vs 40.7 us; 768, 1152, 1536 and 2048 blocks were all slower
// than 1024.  Node blocks take 2 per CU in FP32 and 1 in FP64: at 96^3
// elasticity the FP32 SpMV takes 131.5 / 119.1 / 125.1 / 135.0 / 122.9 us
// with 1 / 2 / 3 / 4 / 8 blocks per CU (a fractional number per CU, 1.5 or
// 2.5, 143-140 us), and the FP64 SELL solve 389 / 422 / 436 / 450 ms per
// step with 1 / 2 / 3 / 4.  Small problems (fewer slices than 16 per CU)
// take the SELL grid.
unsigned sym_grid(int64_t n, int cus, int bpc)
{
    const int64_t nsl = (n + 63) / 64;
    const int wpb = SB / 64;
    const int cpx = std::max(1, cus / NXCD);   // CUs per XCD
    if (nsl < (int64_t)4 * cpx * NXCD * wpb) return sell_grid(n);
    return (unsigned)std::min(bpc * cpx * NXCD, MAXG);
}

unsigned csr_grid(int64_t n, int lpr)
{
    int64_t g = (n * lpr + SB - 1) / SB;
    return round8(std::max<int64_t>(1, std::min<int64_t>(g, MAXG)));
}

// tuning knobs (environment, read once): INS_CSR_LPR = 4/8/16/32 lanes per
// CSR row, INS_MIXED_CHUNK, INS_XCD, INS_NT
int env_int(const char* name, int dflt)
{
    const char* v = getenv(name);
    return v ? atoi(v) : dflt;
}

double env_double(const char* name, double dflt)
{
    const char* v = std::getenv(name);
    return (v && *v) ? std::atof(v) : dflt;
}

// mixed: replace the residual once |r|^2 fell below delta^2 of the last
// true one (INS_DELTA2, 0.01: every 10x reduction of |r|)
double mixed_delta2()
{
    static const double d = env_double("INS_DELTA2", 0.01);
    return d;
}

int ensure_solver_state(Ctx* c, int mode)
{
    const int64_t n = c->n_active;
    INS_TRY(ensure(c, c->x, n));
    if (mode == INS_SOLVE_FP64) {
        INS_TRY(ensure(c, c->r, n));
        INS_TRY(ensure(c, c->p, n));
        INS_TRY(ensure(c, c->q, n));
    } else {
        INS_TRY(ensure(c, c->x32, n));
        INS_TRY(ensure(c, c->r32, n));
        INS_TRY(ensure(c, c->p32, n));
        INS_TRY(ensure(c, c->q32, n));
    }
    if (!c->partial.p) INS_TRY(ensure(c, c->partial, (size_t)NSLOT * MAXG));
    if (!c->scal.p) INS_TRY(ensure(c, c->scal, 1));
    if (!c->h_scal) INS_HIP_CHECK(c, hipHostMalloc((void**)&c->h_scal, sizeof(Scal), 0));
    if (!c->h_snap) {
        INS_HIP_CHECK(c, hipHostMalloc((void**)&c->h_snap, 2 * sizeof(Scal), 0));
        const int mode = env_int("INS_XCD", 0);
        INS_HIP_CHECK(c, hipMemcpyToSymbol(HIP_SYMBOL(g_xcd_mode), &mode, sizeof(int)));
        const int nt = env_int("INS_NT", 0);
        INS_HIP_CHECK(c, hipMemcpyToSymbol(HIP_SYMBOL(g_nt), &nt, sizeof(int)));
        const int bands = env_int("INS_BSYM_BANDS", 4);
        INS_HIP_CHECK(c, hipMemcpyToSymbol(HIP_SYMBOL(g_bsym_bands), &bands, sizeof(int)));
        const double ratio = env_double("INS_RESTART_RATIO", 64.);
        INS_HIP_CHECK(c, hipMemcpyToSymbol(HIP_SYMBOL(g_restart_ratio), &ratio, sizeof(double)));
    }
    for (int i = 0; i < 2; i++)
        if (!c->evs[i]) INS_HIP_CHECK(c, hipEventCreateWithFlags(&c->evs[i], hipEventDisableTiming));
    return INS_OK;
}

// the slices index rows of the (permuted, for vector problems) solve
// numbering directly
SellView sell_view(Ctx* c)
{
    SellView v;
    v.perm = nullptr;
    v.sptr = c->sell.slice_ptr.p;
    v.spat = c->sell.slice_pat.p;
    v.pptr = c->sell.pat_ptr.p;
    v.plen = c->sell.pat_len.p;
    v.pdel = c->sell.pat_delta.p;
    return v;
}

// The vectors of a solve in its numbering: the context's own for CSR solves
// and scalar problems; for SELL solves of vector problems, copies of b and
// of the Jacobi diagonal in the component-major numbering of the slices, and
// a separate FP64 solution (solve_vectors_out writes it back).
int solve_vectors(Ctx* c, int format, int mode, hipStream_t st)
{
    const int64_t n = c->n_active;
    c->sperm = format == INS_FMT_SELL && c->sell.perm_on;
    if (!c->sperm) {
        c->sx64 = c->x.p;
        c->sb = c->b.p;
        c->sd64 = c->dinv64.p;
        c->sd32 = c->dinv32.p;
        return INS_OK;
    }
    INS_TRY(ensure(c, c->xp, n));
    INS_TRY(ensure(c, c->bp, n));
    const int32_t* perm = c->sell.perm.p;
    k_perm_in<double><<<grid_for(n, 256), 256, 0, st>>>(n, perm, c->b.p, c->bp.p);
    if (mode == INS_SOLVE_MIXED) {
        INS_TRY(ensure(c, c->d32p, n));
        k_perm_in<float><<<grid_for(n, 256), 256, 0, st>>>(n, perm, c->dinv32.p, c->d32p.p);
    } else {
        INS_TRY(ensure(c, c->d64p, n));
        k_perm_in<double><<<grid_for(n, 256), 256, 0, st>>>(n, perm, c->dinv64.p, c->d64p.p);
    }
    INS_HIP_CHECK(c, hipGetLastError());
    c->sx64 = c->xp.p;
    c->sb = c->bp.p;
    c->sd64 = c->d64p.p;
    c->sd32 = c->d32p.p;
    return INS_OK;
}

int solve_vectors_out(Ctx* c, hipStream_t st)
{
    if (!c->sperm) return INS_OK;
    const int64_t n = c->n_active;
    k_perm_out<double><<<grid_for(n, 256), 256, 0, st>>>(n, c->sell.perm.p, c->xp.p, c->x.p);
    INS_HIP_CHECK(c, hipGetLastError());
    return INS_OK;
}

// lanes per row of the CSR kernels, from the mean row length
int lpr_for(int64_t n, int64_t nnz)
{
    static const int forced = env_int("INS_CSR_LPR", 0);
    if (forced == 4 || forced == 8 || forced == 16 || forced == 32) return forced;
    const double avg = n ? (double)nnz / (double)n : 1.;
    if (avg > 48) return 16;
    if (avg > 12) return 8;
    return 4;
}

int pick_lpr(Ctx* c) { return lpr_for(c->n_active, c->nnz); }

// K1; returns the number of partials written
// Launch of the dominant kernel on the context stream.  With profiling on,
// the launch itself records an event pair around the kernel (no separate
// marker packets in the stream), read once the solve has finished.
template <typename F, typename... Args>
int launch_timed(Ctx* c, F kernel, unsigned grid, unsigned block, Args... args)
{
    if (!c->profiling) {
        kernel<<<grid, block, 0, c->stream>>>(args...);
        return INS_OK;
    }
    const size_t k = c->prof_n;
    while (c->prof_ev.size() < 2 * (k + 1)) {
        hipEvent_t e = nullptr;
        INS_HIP_CHECK(c, hipEventCreate(&e));
        c->prof_ev.push_back(e);
    }
    c->prof_n++;
    hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(block), 0, c->stream, c->prof_ev[2 * k],
                          c->prof_ev[2 * k + 1], 0, args...);
    return INS_OK;
}

// The FP32 sweeps of a SELL solve use the symmetric slices when they were
// built for the vectors in use: the scalar slices in the identity order, the
// node-block slices on the component-major vectors
bool use_sym(const Ctx* c, bool permuted)
{
    return c->sell.sym_on && (c->sell.sdpn > 1) == permuted;
}

SymView sym_view(const Ctx* c)
{
    return SymView{c->sell.vs.p, c->sell.su.p, c->sell.nu, c->sell.sdpn, c->sell.snodes,
                   c->sell.slines ? 1 : 0, c->sell.vs64.p};
}

// Blocks of the SpMV on the full slices (sym = false) or on the symmetric
// slices, the same for a solve alone and inside a server batch, so that a
// request's partial sums, and with them its iterates, do not depend on where
// it runs
// the FP32 stencil-line SpMV in its register-capped instantiation (INS_SYM_CAP)
static const int g_cap = env_int("INS_SYM_CAP", 1);

unsigned spmv_grid(const Ctx* c, bool sym, bool fp64)
{
    const int64_t n = c->n_active;
    if (!sym) return sell_grid(n);
    static const int forced = env_int("INS_SYM_GRID", 0);
    if (forced >= 8) return round8(std::min(forced, MAXG));
    if (c->sell.sdpn > 1) return sym_grid(c->sell.snodes, c->n_cu, fp64 ? 1 : 2);
    // the FP64 stencil-line rows hold about 150 registers: 3 waves per SIMD
    return sym_grid(n, c->n_cu, fp64 && c->sell.slines ? 3 : 4);
}

template <typename TV, typename TX, typename TA, typename TVR>
int launch_spmv_r(Ctx* c, int format, const TV* csr_val, const TV* sell_val, const TX* p, TX* q,
                  int* g_out, int it, const TVR* rep_val, int rep_sym)
{
    const int64_t n = c->n_active;
    const Scal* sc = c->scal.p;
    const Rep<TVR> R{rep_val, c->sx64, c->sb, c->r32.p, c->sd32, rep_sym};
    if (format == INS_FMT_SELL) {
        const bool sym = use_sym(c, c->sperm);
        const unsigned g = spmv_grid(c, sym, sizeof(TV) == 8);
        SymView Y{nullptr, nullptr, 0, 1, 0, 0, nullptr};
        if (sym) Y = sym_view(c);
        if constexpr (sizeof(TV) == 8) {
            // FP64 sweeps: the FP64 symmetric slices (stencil lines, generic
            // scalar rows or node blocks)
            if (sym && Y.dpn > 1)
                INS_TRY(launch_timed(c, k_spmv_sell<TV, TX, TA, 4, TVR, 4>, g, SB, n, sell_view(c),
                                     sell_val, p, q, c->partial.p, sc, it, R, Y));
            else if (sym && Y.lines)
                INS_TRY(launch_timed(c, k_spmv_sell<TV, TX, TA, 4, TVR, 3>, g, SB, n, sell_view(c),
                                     sell_val, p, q, c->partial.p, sc, it, R, Y));
            else if (sym)
                INS_TRY(launch_timed(c, k_spmv_sell<TV, TX, TA, 4, TVR, 1>, g, SB, n, sell_view(c),
                                     sell_val, p, q, c->partial.p, sc, it, R, Y));
            else
                INS_TRY(launch_timed(c, k_spmv_sell<TV, TX, TA, 4, TVR, 0>, g, SB, n, sell_view(c),
                                     sell_val, p, q, c->partial.p, sc, it, R, Y));
        } else if (Y.v && Y.dpn > 1)
            INS_TRY(launch_timed(c, k_spmv_sell<TV, TX, TA, 4, TVR, 4>, g, SB, n, sell_view(c),
                                 sell_val, p, q, c->partial.p, sc, it, R, Y));
        else if (Y.v && Y.lines && g_cap)
            INS_TRY(launch_timed(c, k_spmv_cap<5, TV, TX, TA, 4, TVR, 3>, g, SB, n,
                                 sell_view(c), sell_val, p, q, c->partial.p, sc, it, R, Y));
        else if (Y.v && Y.lines)
            INS_TRY(launch_timed(c, k_spmv_sell<TV, TX, TA, 4, TVR, 3>, g, SB, n, sell_view(c),
                                 sell_val, p, q, c->partial.p, sc, it, R, Y));
        else if (Y.v)
            INS_TRY(launch_timed(c, k_spmv_sell<TV, TX, TA, 4, TVR, 1>, g, SB, n, sell_view(c),
                                 sell_val, p, q, c->partial.p, sc, it, R, Y));
        else
            INS_TRY(launch_timed(c, k_spmv_sell<TV, TX, TA, 4, TVR, 0>, g, SB, n, sell_view(c),
                                 sell_val, p, q, c->partial.p, sc, it, R, Y));
        *g_out = (int)g;
    } else {
        const int lpr = pick_lpr(c);
        const unsigned g = csr_grid(n, lpr);
        const int32_t* rp = c->row_ptr.p;
        const int32_t* ci = c->col.p;
        if (lpr == 32)
            INS_TRY(launch_timed(c, k_spmv_csr<32, TV, TX, TA, TVR>, g, SB, n, rp, ci, csr_val, p, q, c->partial.p, sc, it, R));
        else if (lpr == 16)
            INS_TRY(launch_timed(c, k_spmv_csr<16, TV, TX, TA, TVR>, g, SB, n, rp, ci, csr_val, p, q, c->partial.p, sc, it, R));
        else if (lpr == 8)
            INS_TRY(launch_timed(c, k_spmv_csr<8, TV, TX, TA, TVR>, g, SB, n, rp, ci, csr_val, p, q, c->partial.p, sc, it, R));
        else
            INS_TRY(launch_timed(c, k_spmv_csr<4, TV, TX, TA, TVR>, g, SB, n, rp, ci, csr_val, p, q, c->partial.p, sc, it, R));
        *g_out = (int)g;
    }
    INS_HIP_CHECK(c, hipGetLastError());
    return INS_OK;
}

// K1 launch; the values of the residual replacement are the FP64 ones when
// the assembly was FP64 (the FP64 solve never replaces), in the storage of
// the solve: CSR, the full slices, or the symmetric slices where those are
// the ones in use
template <typename TV, typename TX, typename TA>
int launch_spmv(Ctx* c, int format, const TV* csr_val, const TV* sell_val, const TX* p, TX* q,
                int* g_out, int it)
{
    const bool sell = format == INS_FMT_SELL;
    const int sym = sell && use_sym(c, c->sperm) ? 1 : 0;
    if constexpr (sizeof(TX) == 4) {
        if (c->matrix_precision != INS_FP64)
            return launch_spmv_r<TV, TX, TA, float>(
                c, format, csr_val, sell_val, p, q, g_out, it,
                sym ? c->sell.vs.p : (sell ? c->sell.v32.p : c->val32.p), sym);
    }
    return launch_spmv_r<TV, TX, TA, double>(
        c, format, csr_val, sell_val, p, q, g_out, it,
        sym ? c->sell.vs64.p : (sell ? c->sell.v64.p : c->val64.p), sym);
}

// |b - A x64| (and the residual replacement); then the finalising kernel
// true residual |b - A x64|^2 into the scalars
int launch_resid(Ctx* c, int format, hipStream_t stream = nullptr, Scal* scal = nullptr)
{
    const int64_t n = c->n_active;
    hipStream_t st = stream ? stream : c->stream;
    Scal* sc = scal ? scal : c->scal.p;
    unsigned g;
    const bool v64 = c->matrix_precision == INS_FP64;
    if (format == INS_FMT_SELL) {
        const int sym = use_sym(c, c->sperm) ? 1 : 0;
        const SymView Y = sym ? sym_view(c) : SymView{nullptr, nullptr, 0, 1, 0, 0, nullptr};
        g = sell_grid(sym && Y.dpn > 1 ? Y.nn : n);
        if (v64)
            k_resid_sell<double><<<g, SB, 0, st>>>(n, sell_view(c), sym ? c->sell.vs64.p : c->sell.v64.p,
                                                   sym, Y, c->sx64, c->sb, c->partial.p);
        else
            k_resid_sell<float><<<g, SB, 0, st>>>(n, sell_view(c), sym ? c->sell.vs.p : c->sell.v32.p,
                                                  sym, Y, c->sx64, c->sb, c->partial.p);
    } else {
        const int lpr = pick_lpr(c);
        g = csr_grid(n, lpr);
        if (v64) {
            if (lpr == 16) k_resid_csr<16, double><<<g, SB, 0, st>>>(n, c->row_ptr.p, c->col.p, c->val64.p, c->sx64, c->sb, c->partial.p);
            else if (lpr == 8) k_resid_csr<8, double><<<g, SB, 0, st>>>(n, c->row_ptr.p, c->col.p, c->val64.p, c->sx64, c->sb, c->partial.p);
            else k_resid_csr<4, double><<<g, SB, 0, st>>>(n, c->row_ptr.p, c->col.p, c->val64.p, c->sx64, c->sb, c->partial.p);
        } else {
            if (lpr == 16) k_resid_csr<16, float><<<g, SB, 0, st>>>(n, c->row_ptr.p, c->col.p, c->val32.p, c->sx64, c->sb, c->partial.p);
            else if (lpr == 8) k_resid_csr<8, float><<<g, SB, 0, st>>>(n, c->row_ptr.p, c->col.p, c->val32.p, c->sx64, c->sb, c->partial.p);
            else k_resid_csr<4, float><<<g, SB, 0, st>>>(n, c->row_ptr.p, c->col.p, c->val32.p, c->sx64, c->sb, c->partial.p);
        }
    }
    INS_HIP_CHECK(c, hipGetLastError());
    k_resid_fin<<<1, VB, 0, st>>>(c->partial.p, (int)g, sc);
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

// Sum of the recorded K1 durations, in two classes: the SpMV launches, and
// the replacement launches (FP64 true residual: about twice the bytes and
// time, more than 1.6x the median).  Launches enqueued after the stop find
// the flag set and return within a few microseconds (under a quarter of the
// median); they are left out.
int prof_collect(Ctx* c)
{
    std::vector<float> ms(c->prof_n, 0.f);
    for (size_t k = 0; k < c->prof_n; k++)
        INS_HIP_CHECK(c, hipEventElapsedTime(&ms[k], c->prof_ev[2 * k], c->prof_ev[2 * k + 1]));
    std::vector<float> sorted(ms);
    std::sort(sorted.begin(), sorted.end());
    const float med = sorted.empty() ? 0.f : sorted[sorted.size() / 2];
    for (float v : ms) {
        if (v < 0.25f * med) continue;
        if (v > 1.6f * med) {
            c->stats.rep_ms += v;
            c->stats.rep_calls += 1;
        } else {
            c->stats.spmv_ms += v;
            c->stats.spmv_calls += 1;
        }
    }
    c->prof_n = 0;
    return INS_OK;
}

// one chunk of PCG iterations (K1, K2, K3 per iteration)
template <typename TV, typename TX>
int run_chunk(Ctx* c, int format, const TV* csr_val, const TV* sell_val, TX* x, TX* r, TX* p,
              TX* q, const TX* dinv, int k, int64_t* it_launched)
{
    const int64_t n = c->n_active;
    const unsigned gv = vec_grid(n);
    for (int i = 0; i < k; i++) {
        const int it = (int)*it_launched;
        int g1 = 0;
        INS_TRY((launch_spmv<TV, TX, TX>(c, format, csr_val, sell_val, p, q, &g1, it)));
        k_update_xr<TX, TX><<<gv, VB, 0, c->stream>>>(n, x, r, p, q, dinv, c->partial.p, g1,
                                                      c->scal.p, it);
        k_update_p<TX, TX><<<gv, VB, 0, c->stream>>>(n, p, r, dinv, c->partial.p, (int)gv,
                                                     c->scal.p, it, c->sx64, x);
        (*it_launched)++;
    }
    INS_HIP_CHECK(c, hipGetLastError());
    return INS_OK;
}

// Symmetric FP32 slices (see SymView): built when the matrix is scalar in
// the identity row order, the upper column offsets of all slice lists
// together number at most SYM_MAXU, and the shared list pads the upper
// triangle by at most 1/8.  INS_SYM=0 switches them off.
constexpr int SYM_MAXU = 64;

// Node-block symmetric slices of a vector problem (see Sell::sdpn): built
// when every node carries all or none of its dpn components, the active rows
// are interleaved per node in node order (row a * dpn + c), the node offsets
// of the pattern are one list of at most SYM_MAXU offsets shared by all
// nodes whose lower offsets mirror upper ones, and the stored blocks pad the
// upper triangle by at most 1/8.  The offsets are read off the full slices'
// column-offset lists, which hold (c' - c) N + delta in the component-major
// numbering; every CSR entry is then placed by k_bsym_map, and an entry that
// finds no place leaves the full slices in use.
// The words of the direct assembly (k_sym_words) for Q1 elements of up to 8
// nodes with at most 8 elements per node whose symmetric slices hold the
// diagonal as offset 0
int build_sym_words(Ctx* c, int64_t N, int D)
{
    Sell& S = c->sell;
    S.direct = false;
    S.zeros = false;
    S.zeros64 = false;
    static const int enabled = env_int("INS_SYM_DIRECT", 1);
    if (!enabled || c->deg != 1 || c->npe > 8 || !c->have_pinfo || c->csr_from_host ||
        !c->dofs_match())
        return INS_OK;
    int32_t u0 = -1;
    INS_HIP_CHECK(c, hipMemcpyAsync(&u0, S.su.p, sizeof(int32_t), hipMemcpyDeviceToHost,
                                    c->stream));
    INS_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    if (u0 != 0) return INS_OK;
    INS_TRY(ensure(c, S.anode, N));
    k_anode<<<grid_for(c->n_nodes, 256), 256, 0, c->stream>>>(c->n_nodes, D, c->dof_id.p,
                                                            S.anode.p);
    const int64_t total = (N + 63) / 64 * 64 * S.nu;
    INS_TRY(ensure(c, S.spw, total));
    const unsigned g = grid_for(total, 256);
    if (c->npe == 8)
        k_sym_words<8><<<g, 256, 0, c->stream>>>(N, S.nu, S.su.p, S.anode.p, c->adj_ptr.p,
                                                 c->adj.p, c->conn.p, c->lg_npe, S.spw.p);
    else
        k_sym_words<4><<<g, 256, 0, c->stream>>>(N, S.nu, S.su.p, S.anode.p, c->adj_ptr.p,
                                                 c->adj.p, c->conn.p, c->lg_npe, S.spw.p);
    INS_HIP_CHECK(c, hipGetLastError());
    S.direct = true;
    return INS_OK;
}

int build_bsym(Ctx* c)
{
    Sell& S = c->sell;
    const int D = c->dpn;
    if (D < 2 || D > 3 || c->have_order || !c->dofs_match() || !S.perm_on) return INS_OK;
    const int64_t n = c->n_active;
    if (n % D) return INS_OK;
    const int64_t N = n / D;
    {
        std::vector<int32_t> dof((size_t)(c->n_nodes * D));
        INS_HIP_CHECK(c, hipMemcpyAsync(dof.data(), c->dof_id.p, dof.size() * sizeof(int32_t),
                                        hipMemcpyDeviceToHost, c->stream));
        INS_HIP_CHECK(c, hipStreamSynchronize(c->stream));
        int64_t a = 0;
        for (int64_t node = 0; node < c->n_nodes; node++) {
            const int32_t* id = &dof[(size_t)(node * D)];
            int act = 0;
            for (int k = 0; k < D; k++) act += id[k] >= 0;
            if (act == 0) continue;
            if (act != D) return INS_OK;
            for (int k = 0; k < D; k++)
                if (id[k] != a * D + k) return INS_OK;
            a++;
        }
        if (a != N) return INS_OK;
    }
    std::vector<int32_t> plen(S.npat), pptr(S.npat + 1);
    INS_HIP_CHECK(c, hipMemcpyAsync(plen.data(), S.pat_len.p, S.npat * sizeof(int32_t),
                                    hipMemcpyDeviceToHost, c->stream));
    INS_HIP_CHECK(c, hipMemcpyAsync(pptr.data(), S.pat_ptr.p, (S.npat + 1) * sizeof(int32_t),
                                    hipMemcpyDeviceToHost, c->stream));
    INS_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    const int64_t ndel = (int64_t)pptr[S.npat - 1] + plen[S.npat - 1];
    std::vector<int32_t> del(ndel > 0 ? ndel : 1);
    INS_HIP_CHECK(c, hipMemcpyAsync(del.data(), S.pat_delta.p, ndel * sizeof(int32_t),
                                    hipMemcpyDeviceToHost, c->stream));
    INS_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    std::vector<int32_t> up, lowmag;
    for (int64_t i = 0; i < ndel; i++) {
        int64_t dd = ((int64_t)del[i] % N + N) % N;   // the node offset, in (-N/2, N/2]
        if (dd > N / 2) dd -= N;
        if (dd >= 0) up.push_back((int32_t)dd);
        else lowmag.push_back((int32_t)-dd);
    }
    std::sort(up.begin(), up.end());
    up.erase(std::unique(up.begin(), up.end()), up.end());
    std::sort(lowmag.begin(), lowmag.end());
    lowmag.erase(std::unique(lowmag.begin(), lowmag.end()), lowmag.end());
    for (int32_t d : lowmag)
        if (!std::binary_search(up.begin(), up.end(), d)) return INS_OK;
    const int nu = (int)up.size();
    if (nu == 0 || nu > SYM_MAXU || up[0] != 0) return INS_OK;
    const int64_t nvs = (N + 63) / 64 * 64 * SymLayout{D, nu}.width();
    const int64_t upper = (c->nnz + c->n_active) / 2;
    if (nvs > upper + upper / 8 + 64 * 64 || nvs >= INT32_MAX) return INS_OK;
    INS_TRY(ensure(c, S.su, nu));
    INS_HIP_CHECK(c, hipMemcpyAsync(S.su.p, up.data(), nu * sizeof(int32_t),
                                    hipMemcpyHostToDevice, c->stream));
    INS_TRY(ensure(c, S.smap, c->nnz));
    int* err = nullptr;
    INS_HIP_CHECK(c, hipMallocAsync((void**)&err, sizeof(int), c->stream));
    INS_HIP_CHECK(c, hipMemsetAsync(err, 0, sizeof(int), c->stream));
    k_bsym_map<<<grid_for(n, 256), 256, 0, c->stream>>>(n, D, c->row_ptr.p, c->col.p, S.su.p, nu,
                                                       S.smap.p, err);
    INS_HIP_CHECK(c, hipGetLastError());
    int herr = 0;
    INS_HIP_CHECK(c, hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    INS_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    (void)hipFreeAsync(err, c->stream);
    if (herr) return INS_OK;
    S.nu = nu;
    S.nvs = nvs;
    S.sdpn = D;
    S.snodes = N;
    // node offsets on stencil lines (as for scalar rows): {0, 1, c1 - 1, c1,
    // c1 + 1, ...}, 14 of them in 3-D and 5 in 2-D
    S.slines = nu == (D == 3 ? 14 : 5) && up[0] == 0 && up[1] == 1;
    for (int m = 0; S.slines && m < (nu - 2) / 3; m++) {
        const int32_t c0 = up[3 + 3 * m];
        S.slines = up[2 + 3 * m] == c0 - 1 && up[4 + 3 * m] == c0 + 1;
    }
    static const int lines_on = env_int("INS_SYM_LINES", 1);
    S.slines = S.slines && lines_on;
    S.sym_on = true;
    return build_sym_words(c, N, D);
}

int build_sym(Ctx* c)
{
    Sell& S = c->sell;
    S.sym_on = false;
    S.direct = false;
    S.zeros = false;
    S.sdpn = 1;
    S.snodes = 0;
    S.slines = false;
    static const int enabled = env_int("INS_SYM", 1);
    if (!enabled || !S.ok || S.npat <= 0 || c->csr_from_host) return INS_OK;
    if (c->dpn > 1) return build_bsym(c);
    if (S.perm_on) return INS_OK;
    std::vector<int32_t> plen(S.npat), pptr(S.npat + 1);
    INS_HIP_CHECK(c, hipMemcpyAsync(plen.data(), S.pat_len.p, S.npat * sizeof(int32_t),
                                    hipMemcpyDeviceToHost, c->stream));
    INS_HIP_CHECK(c, hipMemcpyAsync(pptr.data(), S.pat_ptr.p, (S.npat + 1) * sizeof(int32_t),
                                    hipMemcpyDeviceToHost, c->stream));
    INS_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    const int64_t ndel = (int64_t)pptr[S.npat - 1] + plen[S.npat - 1];
    std::vector<int32_t> del(ndel > 0 ? ndel : 1);
    INS_HIP_CHECK(c, hipMemcpyAsync(del.data(), S.pat_delta.p, ndel * sizeof(int32_t),
                                    hipMemcpyDeviceToHost, c->stream));
    INS_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    std::vector<int32_t> up, lowmag;
    for (int64_t i = 0; i < ndel; i++) {
        if (del[i] >= 0) up.push_back(del[i]);
        else lowmag.push_back(-del[i]);
    }
    std::sort(up.begin(), up.end());
    up.erase(std::unique(up.begin(), up.end()), up.end());
    std::sort(lowmag.begin(), lowmag.end());
    lowmag.erase(std::unique(lowmag.begin(), lowmag.end()), lowmag.end());
    // a symmetric pattern mirrors every lower offset into the upper list
    for (int32_t d : lowmag)
        if (!std::binary_search(up.begin(), up.end(), d)) return INS_OK;
    const int nu = (int)up.size();
    if (nu == 0 || nu > SYM_MAXU) return INS_OK;
    const int64_t nsl = S.nslices;
    const int64_t nvs = nsl * 64 * (int64_t)nu;
    const int64_t upper = (c->nnz + c->n_active) / 2;
    if (nvs > upper + upper / 8 + 64 * 64 || nvs >= INT32_MAX) return INS_OK;
    INS_TRY(ensure(c, S.su, nu));
    INS_HIP_CHECK(c, hipMemcpyAsync(S.su.p, up.data(), nu * sizeof(int32_t),
                                    hipMemcpyHostToDevice, c->stream));
    INS_TRY(ensure(c, S.smap, c->nnz));
    int* err = nullptr;
    INS_HIP_CHECK(c, hipMallocAsync((void**)&err, sizeof(int), c->stream));
    INS_HIP_CHECK(c, hipMemsetAsync(err, 0, sizeof(int), c->stream));
    k_sym_map<<<grid_for(c->n_active, 256), 256, 0, c->stream>>>(
        c->n_active, c->row_ptr.p, c->col.p, S.su.p, nu, S.smap.p, err);
    INS_HIP_CHECK(c, hipGetLastError());
    int herr = 0;
    INS_HIP_CHECK(c, hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    INS_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    (void)hipFreeAsync(err, c->stream);
    if (herr) return INS_OK;
    S.nu = nu;
    S.nvs = nvs;
    // stencil lines: {0, 1, c1 - 1, c1, c1 + 1, ...} with 1 < c1 - 1, the
    // centres ascending and every line at least one wave (64 rows) apart
    S.slines = (nu == 14 || nu == 5) && up[0] == 0 && up[1] == 1;
    for (int m = 0; S.slines && m < (nu - 2) / 3; m++) {
        const int32_t c0 = up[3 + 3 * m];
        S.slines = up[2 + 3 * m] == c0 - 1 && up[4 + 3 * m] == c0 + 1;
    }
    static const int lines_on = env_int("INS_SYM_LINES", 1);
    S.slines = S.slines && lines_on;
    S.sym_on = true;
    S.snodes = c->n_active;
    return build_sym_words(c, c->n_active, 1);
}

}  // namespace
// force: keep the SELL form even when its padding makes it larger than the
// CSR values by more than 1/8 (the serving kernels need it)
int build_sell(Ctx* c, bool force)
{
    Sell& S = c->sell;
    const int64_t n = c->n_active;
    const int64_t nsl = (n + 63) / 64;
    S.n = n;
    S.nslices = nsl;
    uint64_t* h = nullptr;
    uint64_t* keys = nullptr;
    int32_t *udel = nullptr, *ulen = nullptr, *reps = nullptr, *slot_of = nullptr;
    int32_t *used = nullptr, *slot_id = nullptr;
    int* err = nullptr;
    uint64_t cap = 1024;
    while (cap < (uint64_t)(2 * nsl + 2)) cap <<= 1;
    INS_HIP_CHECK(c, hipMallocAsync((void**)&udel, (size_t)nsl * MAXU * sizeof(int32_t) + 8, c->stream));
    INS_HIP_CHECK(c, hipMallocAsync((void**)&ulen, (nsl + 1) * sizeof(int32_t), c->stream));
    INS_HIP_CHECK(c, hipMallocAsync((void**)&h, nsl * sizeof(uint64_t) + 8, c->stream));
    INS_HIP_CHECK(c, hipMallocAsync((void**)&keys, cap * sizeof(uint64_t), c->stream));
    INS_HIP_CHECK(c, hipMallocAsync((void**)&reps, cap * sizeof(int32_t), c->stream));
    INS_HIP_CHECK(c, hipMallocAsync((void**)&slot_of, nsl * sizeof(int32_t) + 8, c->stream));
    INS_HIP_CHECK(c, hipMallocAsync((void**)&used, (cap + 1) * sizeof(int32_t), c->stream));
    INS_HIP_CHECK(c, hipMallocAsync((void**)&slot_id, (cap + 1) * sizeof(int32_t), c->stream));
    INS_HIP_CHECK(c, hipMallocAsync((void**)&err, sizeof(int), c->stream));
    INS_HIP_CHECK(c, hipMemsetAsync(keys, 0, cap * sizeof(uint64_t), c->stream));
    INS_HIP_CHECK(c, hipMemsetAsync(err, 0, sizeof(int), c->stream));
    // Vector problems: the slices are built on P A P^T, the matrix in the
    // numbering that orders the rows by component (then by row), so that the
    // 64 rows of a slice are consecutive nodes of one component, their
    // column offsets are shared, and the gathers and stores of a slice are
    // contiguous.  SELL solves of such problems run on permuted vectors.
    // Degree-2 fields: the slices are built in the solver order of the field
    // nodes (grouped by n-face class, then by support point), in which the
    // rows of a slice are neighbours of one kind with shared column offsets;
    // the face-based DoF numbering itself scatters them.
    // (a CSR uploaded by ins_set_csr with another row count has no DoF
    // structure: its slices keep the identity row order)
    const bool ordered = c->have_order && c->dofs_match();
    S.perm_on = c->dofs_match() && (c->dpn > 1 || ordered);
    const int32_t* brp = c->row_ptr.p;   // CSR the slices are built from
    const int32_t* bcol = c->col.p;
    int32_t *prp = nullptr, *pcol = nullptr, *pent = nullptr;
    if (S.perm_on) {
        INS_TRY(ensure(c, S.perm, n + 1));
        const int64_t nn = c->n_nodes;
        const int64_t nf = (nn > n ? nn : n) + 1;
        int32_t *comp = nullptr, *flag = nullptr, *pos = nullptr, *iperm = nullptr;
        INS_HIP_CHECK(c, hipMallocAsync((void**)&comp, (n + 1) * sizeof(int32_t), c->stream));
        INS_HIP_CHECK(c, hipMallocAsync((void**)&flag, nf * sizeof(int32_t), c->stream));
        INS_HIP_CHECK(c, hipMallocAsync((void**)&pos, nf * sizeof(int32_t), c->stream));
        INS_HIP_CHECK(c, hipMallocAsync((void**)&iperm, (n + 1) * sizeof(int32_t), c->stream));
        const int64_t nd = c->n_nodes * c->dpn;
        k_row_comp<<<grid_for(nd, 256), 256, 0, c->stream>>>(nd, c->dof_id.p, c->dpn, comp);
        int64_t off = 0;
        for (int cc = 0; cc < c->dpn; cc++) {
            int64_t tot = 0;
            if (ordered) {
                k_ord_flag<<<grid_for(nn, 256), 256, 0, c->stream>>>(nn, c->node_order.p,
                                                                     c->dof_id.p, c->dpn, cc, flag);
                int st2 = exclusive_scan_i32(c, flag, pos, nn, &tot);
                if (st2 != INS_OK) return st2;
                k_ord_place<<<grid_for(nn, 256), 256, 0, c->stream>>>(
                    nn, c->node_order.p, c->dof_id.p, c->dpn, cc, pos, off, S.perm.p);
            } else {
                k_comp_flag<<<grid_for(n, 256), 256, 0, c->stream>>>(n, comp, cc, flag);
                int st2 = exclusive_scan_i32(c, flag, pos, n, &tot);
                if (st2 != INS_OK) return st2;
                k_comp_place<<<grid_for(n, 256), 256, 0, c->stream>>>(n, comp, cc, pos, off,
                                                                      S.perm.p);
            }
            off += tot;
        }
        k_iperm<<<grid_for(n, 256), 256, 0, c->stream>>>(n, S.perm.p, iperm);
        INS_HIP_CHECK(c, hipMallocAsync((void**)&prp, (n + 1) * sizeof(int32_t), c->stream));
        INS_HIP_CHECK(c, hipMallocAsync((void**)&pcol, (size_t)c->nnz * sizeof(int32_t) + 8, c->stream));
        INS_HIP_CHECK(c, hipMallocAsync((void**)&pent, (size_t)c->nnz * sizeof(int32_t) + 8, c->stream));
        k_prow_len<<<grid_for(n, 256), 256, 0, c->stream>>>(n, S.perm.p, c->row_ptr.p, flag);
        int64_t pnnz = 0;
        int st3 = exclusive_scan_i32(c, flag, prp, n, &pnnz);
        if (st3 != INS_OK) return st3;
        if (ordered)
            k_prow_fill_sorted<<<grid_for(n, 4, 1u << 16), 256, 0, c->stream>>>(
                n, S.perm.p, iperm, c->row_ptr.p, c->col.p, prp, pcol, pent);
        else
            k_prow_fill<<<grid_for(n, 4, 1u << 16), 256, 0, c->stream>>>(
                n, c->dpn, S.perm.p, iperm, comp, c->row_ptr.p, c->col.p, prp, pcol, pent);
        INS_HIP_CHECK(c, hipGetLastError());
        INS_HIP_CHECK(c, hipStreamSynchronize(c->stream));
        (void)hipFreeAsync(comp, c->stream);
        (void)hipFreeAsync(flag, c->stream);
        (void)hipFreeAsync(pos, c->stream);
        (void)hipFreeAsync(iperm, c->stream);
        brp = prp;
        bcol = pcol;
    }
    k_slice_union<<<grid_for(nsl, 4, 1u << 16), 256, 0, c->stream>>>(n, nullptr, brp, bcol,
                                                                      udel, ulen, err);
    const unsigned g = grid_for(nsl, 256);
    k_list_hash<<<g, 256, 0, c->stream>>>(nsl, udel, ulen, h);
    k_hash_insert<<<g, 256, 0, c->stream>>>(nsl, h, keys, reps, cap - 1, err);
    k_hash_find<<<g, 256, 0, c->stream>>>(nsl, h, keys, cap - 1, slot_of);
    k_mark_slots<<<grid_for(cap, 256), 256, 0, c->stream>>>(cap, keys, used);
    INS_HIP_CHECK(c, hipGetLastError());
    int64_t npat = 0;
    int s = exclusive_scan_i32(c, used, slot_id, (int64_t)cap, &npat);
    if (s != INS_OK) return s;
    S.npat = (int)npat;
    INS_TRY(ensure(c, S.slice_pat, nsl + 1));
    k_list_pattern<<<g, 256, 0, c->stream>>>(nsl, udel, ulen, slot_of, slot_id, reps,
                                             S.slice_pat.p, err);
    INS_TRY(ensure(c, S.pat_len, npat + 1));
    INS_TRY(ensure(c, S.pat_ptr, npat + 1));
    k_pat_len<<<grid_for(cap, 256), 256, 0, c->stream>>>(cap, keys, slot_id, reps, ulen,
                                                         S.pat_len.p);
    INS_HIP_CHECK(c, hipGetLastError());
    int64_t ndel = 0;
    s = exclusive_scan_i32(c, S.pat_len.p, S.pat_ptr.p, npat, &ndel);
    if (s != INS_OK) return s;
    INS_TRY(ensure(c, S.pat_delta, ndel + 1));
    k_pat_fill<<<grid_for(cap, 256), 256, 0, c->stream>>>(cap, keys, slot_id, reps, udel, ulen,
                                                          S.pat_ptr.p, S.pat_delta.p);
    INS_TRY(ensure(c, S.slice_ptr, nsl + 1));
    {
        int32_t* sv = nullptr;
        int32_t* sp32 = nullptr;
        INS_HIP_CHECK(c, hipMallocAsync((void**)&sv, (nsl + 1) * sizeof(int32_t), c->stream));
        INS_HIP_CHECK(c, hipMallocAsync((void**)&sp32, (nsl + 1) * sizeof(int32_t), c->stream));
        k_slice_vals<<<grid_for(nsl, 256), 256, 0, c->stream>>>(nsl, ulen, sv);
        int64_t tv = 0;
        s = exclusive_scan_i32(c, sv, sp32, nsl, &tv);
        if (s != INS_OK) return s;
        k_widen<<<grid_for(nsl + 1, 256), 256, 0, c->stream>>>(nsl + 1, sp32, S.slice_ptr.p);
        INS_HIP_CHECK(c, hipGetLastError());
        INS_HIP_CHECK(c, hipStreamSynchronize(c->stream));
        (void)hipFreeAsync(sv, c->stream);
        (void)hipFreeAsync(sp32, c->stream);
    }
    int herr = 0;
    int64_t nvals = 0;
    INS_HIP_CHECK(c, hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    INS_HIP_CHECK(c, hipMemcpyAsync(&nvals, S.slice_ptr.p + nsl, sizeof(int64_t),
                                    hipMemcpyDeviceToHost, c->stream));
    INS_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    (void)hipFreeAsync(udel, c->stream);
    (void)hipFreeAsync(ulen, c->stream);
    (void)hipFreeAsync(h, c->stream);
    (void)hipFreeAsync(keys, c->stream);
    (void)hipFreeAsync(reps, c->stream);
    (void)hipFreeAsync(slot_of, c->stream);
    (void)hipFreeAsync(used, c->stream);
    (void)hipFreeAsync(slot_id, c->stream);
    (void)hipFreeAsync(err, c->stream);
    S.nvals = nvals;
    // worth it only when the shared lists stay close to the rows' own lists
    S.ok = (herr == 0) && (force || nvals <= c->nnz + c->nnz / 8 + 64 * 64) && nvals < INT32_MAX;
    if (S.ok) {
        INS_TRY(ensure(c, S.map, c->nnz));
        int* e2 = nullptr;
        INS_HIP_CHECK(c, hipMallocAsync((void**)&e2, sizeof(int), c->stream));
        INS_HIP_CHECK(c, hipMemsetAsync(e2, 0, sizeof(int), c->stream));
        k_sell_map<<<grid_for(n, 4, 1u << 16), 256, 0, c->stream>>>(n, sell_view(c), brp, bcol,
                                                                     pent, S.map.p, e2);
        INS_HIP_CHECK(c, hipGetLastError());
        int he = 0;
        INS_HIP_CHECK(c, hipMemcpyAsync(&he, e2, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        INS_HIP_CHECK(c, hipStreamSynchronize(c->stream));
        (void)hipFreeAsync(e2, c->stream);
        if (he) S.ok = false;
    }
    if (prp) {
        (void)hipFreeAsync(prp, c->stream);
        (void)hipFreeAsync(pcol, c->stream);
        (void)hipFreeAsync(pent, c->stream);
    }
    S.built = true;
    S.stale = true;
    S.zero64 = false;
    S.zero32 = false;
    INS_TRY(build_sym(c));
    return INS_OK;
}

// copy the current CSR values into the slices (after every assembly)
int ensure_csr(Ctx* c);

int fill_sell(Ctx* c)
{
    Sell& S = c->sell;
    const bool f64 = c->matrix_precision == INS_FP64;
    INS_TRY(ensure_csr(c));
    if (S.sym_on) {
        // the symmetric slices are the only ones the SELL solves read
        INS_TRY(zero_sym(c));
        k_sym_scatter<<<grid_for(c->nnz, 256), 256, 0, c->stream>>>(
            c->nnz, S.smap.p, f64 ? c->val64.p : nullptr, c->val32.p, f64 ? S.vs64.p : nullptr,
            S.vs.p);
        INS_HIP_CHECK(c, hipGetLastError());
        S.stale = false;
        return INS_OK;
    }
    // a new slice structure: size the value arrays to it and zero them once
    // (the scatter writes the stored entries only; padding must read 0)
    if (f64 && !S.zero64) {
        INS_TRY(ensure(c, S.v64, S.nvals));
        INS_HIP_CHECK(c, hipMemsetAsync(S.v64.p, 0, S.nvals * sizeof(double), c->stream));
        S.zero64 = true;
    }
    if (!S.zero32) {
        INS_TRY(ensure(c, S.v32, S.nvals));
        INS_HIP_CHECK(c, hipMemsetAsync(S.v32.p, 0, S.nvals * sizeof(float), c->stream));
        S.zero32 = true;
    }
    k_sell_scatter<<<grid_for(c->nnz, 256), 256, 0, c->stream>>>(
        c->nnz, S.map.p, f64 ? c->val64.p : nullptr, c->val32.p, f64 ? S.v64.p : nullptr,
        S.v32.p);
    INS_HIP_CHECK(c, hipGetLastError());
    S.stale = false;
    return INS_OK;
}

// The symmetric slices sized to the current structure and zeroed once
// (writers fill the stored slots; the padding must read 0)
int zero_sym(Ctx* c)
{
    Sell& S = c->sell;
    if (!S.zeros) {
        INS_TRY(ensure(c, S.vs, S.nvs));
        INS_HIP_CHECK(c, hipMemsetAsync(S.vs.p, 0, S.nvs * sizeof(float), c->stream));
        S.zeros = true;
    }
    if (c->matrix_precision == INS_FP64 && !S.zeros64) {
        INS_TRY(ensure(c, S.vs64, S.nvs));
        INS_HIP_CHECK(c, hipMemsetAsync(S.vs64.p, 0, S.nvs * sizeof(double), c->stream));
        S.zeros64 = true;
    }
    return INS_OK;
}

// CSR values current again after an assembly that wrote the symmetric
// slices only (export, CSR solves, direct solves)
int ensure_csr(Ctx* c)
{
    if (!c->csr_stale) return INS_OK;
    Sell& S = c->sell;
    const bool f64 = c->matrix_precision == INS_FP64;
    if (f64) INS_TRY(ensure(c, c->val64, c->nnz));
    INS_TRY(ensure(c, c->val32, c->nnz));
    const int64_t n = c->n_active;
    const unsigned g = grid_for(n, 256);
    const double* s64 = f64 ? S.vs64.p : nullptr;
    double* d64 = f64 ? c->val64.p : nullptr;
    if (S.sdpn == 3)
        k_sym_to_csr<3><<<g, 256, 0, c->stream>>>(n, c->row_ptr.p, c->col.p, S.su.p, S.nu, s64,
                                                  S.vs.p, d64, c->val32.p);
    else if (S.sdpn == 2)
        k_sym_to_csr<2><<<g, 256, 0, c->stream>>>(n, c->row_ptr.p, c->col.p, S.su.p, S.nu, s64,
                                                  S.vs.p, d64, c->val32.p);
    else
        k_sym_to_csr<1><<<g, 256, 0, c->stream>>>(n, c->row_ptr.p, c->col.p, S.su.p, S.nu, s64,
                                                  S.vs.p, d64, c->val32.p);
    INS_HIP_CHECK(c, hipGetLastError());
    c->csr_stale = false;
    return INS_OK;
}

// Host loop.  Chunks of iterations are enqueued two ahead, each followed by
// an asynchronous copy of the device scalars into a pinned snapshot; the host
// waits only for the older snapshot, so the GPU never idles while the host
// decides.  Once a snapshot shows the stop, the iterations already enqueued
// behind it exit at once.
struct Pipe {
    Ctx* c;
    int inflight = 0;
    int head = 0;  // oldest in-flight snapshot slot
    int enq = 0;   // iterations enqueued so far (skipped ones included)

    int issue_snapshot()
    {
        const int slot = (head + inflight) & 1;
        INS_HIP_CHECK(c, hipMemcpyAsync(&c->h_snap[slot], c->scal.p, sizeof(Scal),
                                        hipMemcpyDeviceToHost, c->stream));
        INS_HIP_CHECK(c, hipEventRecord(c->evs[slot], c->stream));
        inflight++;
        return INS_OK;
    }
    // waits for the oldest snapshot
    int pop(Scal* out)
    {
        INS_HIP_CHECK(c, hipEventSynchronize(c->evs[head]));
        *out = c->h_snap[head];
        head ^= 1;
        inflight--;
        return INS_OK;
    }
    int drain()
    {
        Scal tmp;
        while (inflight) INS_TRY(pop(&tmp));
        return INS_OK;
    }
};

int solve(Ctx* c, int mode, int format, double rtol, int max_it, int* iters, double* relres)
{
    if (!c->have_matrix || !c->have_rhs)
        return fail(c, INS_ERR_STATE, "ins_solve before ins_assemble (matrix and rhs)");
    if (mode == INS_SOLVE_FP64 && c->matrix_precision != INS_FP64)
        return fail(c, INS_ERR_STATE, "FP64 solve needs an FP64 assembly");
    if (format == INS_FMT_SELL && !c->sell.built) INS_TRY(build_sell(c, false));
    if (format == INS_FMT_SELL && !c->sell.ok) format = INS_FMT_CSR;
    if (format == INS_FMT_SELL && c->sell.stale) INS_TRY(fill_sell(c));
    if (format == INS_FMT_CSR) INS_TRY(ensure_csr(c));
    c->want_format = format;
    INS_TRY(ensure_solver_state(c, mode));
    const int64_t n = c->n_active;
    const unsigned gv = vec_grid(n);
    const bool mixed = (mode == INS_SOLVE_MIXED);
    static const int mchunk = std::max(1, env_int("INS_MIXED_CHUNK", MCHUNK));
    const int chunk = mixed ? mchunk : CHUNK;
    c->stats.spmv_ms = 0.;
    c->stats.spmv_calls = 0;
    c->stats.rep_ms = 0.;
    c->stats.rep_calls = 0;
    c->prof_n = 0;
    INS_HIP_CHECK(c, hipEventRecord(c->ev0, c->stream));
    INS_TRY(solve_vectors(c, format, mode, c->stream));
    int64_t launched = 0;
    const double delta2 = mixed_delta2();
    if (!mixed) {
        k_init<double, double><<<gv, VB, 0, c->stream>>>(n, c->sb, c->sx64, c->r.p, c->p.p,
                                                         c->sd64, c->partial.p);
        k_init_fin<<<1, VB, 0, c->stream>>>(c->partial.p, (int)gv, rtol * rtol, 0., c->scal.p);
    } else {
        INS_HIP_CHECK(c, hipMemsetAsync(c->sx64, 0, n * sizeof(double), c->stream));
        k_init<float, float><<<gv, VB, 0, c->stream>>>(n, c->sb, c->x32.p, c->r32.p, c->p32.p,
                                                       c->sd32, c->partial.p);
        k_init_fin<<<1, VB, 0, c->stream>>>(c->partial.p, (int)gv, rtol * rtol, delta2,
                                            c->scal.p);
    }
    INS_HIP_CHECK(c, hipGetLastError());
    INS_TRY(read_scal(c));
    // The device decides everything, replacements included; the host keeps
    // two chunks queued and stops queueing once a snapshot shows the stop.
    // The iteration limit counts SpMV iterations (replacement iterations,
    // one per replacement, come on top).
    Pipe pipe{c};
    bool stop = c->h_scal->done != 0;
    int reps_seen = 0;
    while (!stop) {
        while (pipe.inflight < 2 && pipe.enq < max_it + reps_seen) {
            const int k = std::min(chunk, max_it + reps_seen - pipe.enq);
            if (mixed)
                INS_TRY((run_chunk<float, float>(c, format, c->val32.p,
                                                 format == INS_FMT_SELL ? c->sell.v32.p : nullptr,
                                                 c->x32.p, c->r32.p, c->p32.p, c->q32.p,
                                                 c->sd32, k, &launched)));
            else
                INS_TRY((run_chunk<double, double>(c, format, c->val64.p,
                                                   format == INS_FMT_SELL ? c->sell.v64.p : nullptr,
                                                   c->sx64, c->r.p, c->p.p, c->q.p, c->sd64, k,
                                                   &launched)));
            pipe.enq += k;
            INS_TRY(pipe.issue_snapshot());
        }
        if (pipe.inflight == 0) break;  // iteration limit
        Scal S;
        INS_TRY(pipe.pop(&S));
        reps_seen = S.nrep;
        if (S.done || S.stop) stop = true;
    }
    INS_TRY(pipe.drain());
    if (mixed) {
        INS_TRY(read_scal(c));
        if (!c->h_scal->fin && !c->h_scal->breakdown) {
            // iteration limit: fold what was computed and measure it
            k_fold<<<gv, VB, 0, c->stream>>>(n, c->sx64, c->x32.p);
            INS_TRY(launch_resid(c, format));
        }
    } else {
        INS_TRY(launch_resid(c, format));
    }
    INS_TRY(solve_vectors_out(c, c->stream));
    INS_TRY(read_scal(c));
    INS_HIP_CHECK(c, hipEventRecord(c->ev1, c->stream));
    INS_HIP_CHECK(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    INS_HIP_CHECK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    if (c->profiling) INS_TRY(prof_collect(c));
    const Scal& S = *c->h_scal;
    const double rel = S.bb > 0. ? std::sqrt(S.tt / S.bb) : 0.;
    c->stats.solve_ms = ms;
    c->stats.iterations = S.iters;
    c->stats.refinements = S.nrep;
    c->stats.relres = rel;
    if (iters) *iters = S.iters;
    if (relres) *relres = rel;
    c->have_solution = true;
    int status = INS_OK;
    if (S.breakdown)
        status = fail(c, INS_ERR_BREAKDOWN, "CG breakdown (p.Ap <= 0 or non-finite) after " +
                                                std::to_string(S.iters) + " iterations");
    else if (!(rel <= rtol)) {
        char msg[192];
        if (!mixed && S.stop)
            snprintf(msg, sizeof msg,
                     "the recursive residual met rtol %.3g after %d iterations, the true "
                     "one does not (relres %.3g: round-off floor)", rtol, (int)S.iters, rel);
        else
            snprintf(msg, sizeof msg,
                     "iteration limit reached before rtol %.3g (%d iterations, relres %.3g)",
                     rtol, (int)S.iters, rel);
        status = fail(c, INS_ERR_NOCONV, msg);
    }
    return status;
}

int prepare(Ctx* c, int mode, int format)
{
    if (!c->have_pattern) return fail(c, INS_ERR_STATE, "ins_prepare before ins_build_pattern");
    if (format == INS_FMT_SELL) {
        if (!c->sell.built || !c->sell.ok) INS_TRY(build_sell(c, true));
        if (c->have_matrix && c->sell.ok && c->sell.stale) INS_TRY(fill_sell(c));
    } else if (c->have_matrix) {
        INS_TRY(ensure_csr(c));
    }
    c->want_format = format;
    INS_TRY(ensure_solver_state(c, mode));
    INS_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return INS_OK;
}

int distribute(Ctx* c, int op)
{
    if (!c->have_solution) return fail(c, INS_ERR_STATE, "ins_distribute before ins_solve");
    if (!c->dofs_match())
        return fail(c, INS_ERR_STATE,
                    "ins_distribute: the solved system (ins_set_csr) does not have one row per "
                    "active DoF");
    const int64_t nd = c->n_nodes * c->dpn;
    c->u_zero = false;
    k_distribute<<<grid_for(nd, 256), 256, 0, c->stream>>>(nd, c->dof_id.p, c->x.p, c->g.p,
                                                            c->u.p, op == INS_DIST_ADD);
    INS_HIP_CHECK(c, hipGetLastError());
    INS_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return INS_OK;
}

namespace {
// out = sum_s w[s] * h_s over up to 4 history vectors
__global__ void k_hist_comb(int64_t nd, int ns, const double* __restrict__ h0,
                            const double* __restrict__ h1, const double* __restrict__ h2,
                            const double* __restrict__ h3, double w0, double w1, double w2,
                            double w3, double* __restrict__ out)
{
    for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nd;
         k += (int64_t)gridDim.x * blockDim.x) {
        double v = w0 * h0[k];
        if (ns > 1) v += w1 * h1[k];
        if (ns > 2) v += w2 * h2[k];
        if (ns > 3) v += w3 * h3[k];
        out[k] = v;
    }
}
}  // namespace

// base/time/BDF.hpp:69-99 as written there, start-up of
// base/time/MultiStep.hpp:46-63
static int bdf_weights(int order, int step, double* a)
{
    while (order > 1 && step + 2 < order + 1) order--;
    static const double w[4][5] = {{1.0, -1.0}, {1.5, -2.0, 0.5},
                                   {11. / 6., -3.0, 1.5, -1. / 3.},
                                   {25. / 12., -4.0, 3.0, 4. / 3., 1. / 4.}};
    for (int i = 0; i <= order; i++) a[i] = w[order - 1][i];
    return order + 1;
}

int bdf_step(Ctx* c, double dt, int precision, int mode, int format, double rtol, int max_it,
             int* iters, double* relres)
{
    if (c->bdf_order < 1) return fail(c, INS_ERR_STATE, "ins_bdf_step before ins_bdf_init");
    if (!(dt > 0.)) return fail(c, INS_ERR_ARG, "dt must be positive");
    double a[5];
    const int na = bdf_weights(c->bdf_order, c->bdf_step, a);
    const int64_t nd = c->n_nodes * c->dpn;
    INS_TRY(ensure(c, c->mhist, (size_t)nd));
    const double* hp[4] = {nullptr, nullptr, nullptr, nullptr};
    double w[4] = {0., 0., 0., 0.};
    for (int s = 1; s < na; s++) {
        hp[s - 1] = c->bdf_hist[c->bdf_ring[s - 1]].p;
        w[s - 1] = -a[s] / dt;
    }
    k_hist_comb<<<grid_for(nd, 256), 256, 0, c->stream>>>(
        nd, na - 1, hp[0], hp[1] ? hp[1] : hp[0], hp[2] ? hp[2] : hp[0], hp[3] ? hp[3] : hp[0],
        w[0], w[1], w[2], w[3], c->mhist.p);
    INS_HIP_CHECK(c, hipGetLastError());
    c->density = c->bdf_density;
    c->mass_factor = a[0] / dt;
    c->have_mhist = true;
    // the matrix K + (a0/dt) M changes only with a0/dt: a step with the same
    // factor rebuilds the RHS from the staged element data (the values are
    // the same bits a full re-assembly gives)
    if (c->bdf_staged && c->have_matrix && c->bdf_factor == c->mass_factor &&
        c->bdf_precision == precision) {
        INS_TRY(assemble(c, precision, ASM_STAGED));
    } else {
        c->bdf_staged = false;
        INS_TRY(assemble(c, precision, INS_ASM_MATRIX | INS_ASM_RHS));
        c->bdf_staged = true;
        c->bdf_factor = c->mass_factor;
        c->bdf_precision = precision;
    }
    const int st = solve(c, mode, format, rtol, max_it, iters, relres);
    if (st != INS_OK) return st;
    INS_TRY(distribute(c, INS_DIST_SET));
    // push: the oldest slot takes u^{n+1}
    const int q = c->bdf_order;
    const int last = c->bdf_ring[q - 1];
    for (int s = q - 1; s > 0; s--) c->bdf_ring[s] = c->bdf_ring[s - 1];
    c->bdf_ring[0] = last;
    INS_HIP_CHECK(c, hipMemcpyAsync(c->bdf_hist[last].p, c->u.p, nd * sizeof(double),
                                    hipMemcpyDeviceToDevice, c->stream));
    c->bdf_step++;
    return INS_OK;
}

int vec_norm2(Ctx* c, const double* v, int64_t n, double* out)
{
    INS_TRY(ensure_solver_state(c, INS_SOLVE_FP64));
    const unsigned g = vec_grid(n);
    k_norm2<<<g, VB, 0, c->stream>>>(n, v, c->partial.p);
    k_norm2_fin<<<1, VB, 0, c->stream>>>(c->partial.p, (int)g, c->scal.p);
    INS_HIP_CHECK(c, hipGetLastError());
    INS_TRY(read_scal(c));
    *out = std::sqrt(c->h_scal->tt);
    return INS_OK;
}

// ---------------------------------------------------------------------------
// Direct solve: the FP64 system as a dense column-major matrix, factorised by
// rocSOLVER (potrf/potrs or getrf/getrs) on the context's stream.
namespace {

// CSR -> dense column-major: entry (r, j) at j * n + r; the rest is zero
__global__ void k_csr_to_dense(int64_t n, const int32_t* __restrict__ rp,
                               const int32_t* __restrict__ col, const double* __restrict__ val,
                               double* __restrict__ A)
{
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
         r += (int64_t)gridDim.x * blockDim.x)
        for (int32_t k = rp[r]; k < rp[r + 1]; k++) A[(int64_t)col[k] * n + r] = val[k];
}

struct DenseScratch {
    double* A = nullptr;
    rocblas_int* info = nullptr;
    rocblas_int* ipiv = nullptr;
    rocblas_handle h = nullptr;
    ~DenseScratch()
    {
        if (h) (void)rocblas_destroy_handle(h);
        if (A) (void)hipFree(A);
        if (info) (void)hipFree(info);
        if (ipiv) (void)hipFree(ipiv);
    }
};

int rocsolver_status(Ctx* c, rocblas_status st, const char* what)
{
    if (st == rocblas_status_success) return INS_OK;
    return fail(c, INS_ERR_HIP, std::string(what) + ": rocBLAS status " + std::to_string((int)st));
}

}  // namespace

int solve_direct(Ctx* c, int kind, double* relres)
{
    if (!c->have_matrix || !c->have_rhs)
        return fail(c, INS_ERR_STATE, "ins_solve_direct before ins_assemble (matrix and rhs)");
    if (c->matrix_precision != INS_FP64)
        return fail(c, INS_ERR_STATE, "ins_solve_direct needs an FP64 assembly");
    const int64_t n = c->n_active;
    if (n > INS_DIRECT_MAX_N)
        return fail(c, INS_ERR_UNSUPPORTED,
                    "ins_solve_direct: " + std::to_string(n) + " rows exceed INS_DIRECT_MAX_N (" +
                        std::to_string(INS_DIRECT_MAX_N) + "); solve with ins_solve");
    INS_TRY(ensure_csr(c));
    INS_TRY(ensure_solver_state(c, INS_SOLVE_FP64));
    INS_TRY(solve_vectors(c, INS_FMT_CSR, INS_SOLVE_FP64, c->stream));
    c->have_solution = false;
    c->stats.iterations = 0;
    c->stats.refinements = 0;
    INS_HIP_CHECK(c, hipEventRecord(c->ev0, c->stream));
    double bnorm = 0.;
    INS_TRY(vec_norm2(c, c->sb, n, &bnorm));
    double rel = 0.;
    if (n > 0 && bnorm > 0.) {
        DenseScratch d;
        const size_t bytes = (size_t)n * (size_t)n * sizeof(double);
        if (hipMalloc((void**)&d.A, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return fail(c, INS_ERR_OOM, "ins_solve_direct: dense factor of " +
                                            std::to_string(bytes) + " bytes");
        }
        INS_HIP_CHECK(c, hipMalloc((void**)&d.info, sizeof(rocblas_int)));
        INS_HIP_CHECK(c, hipMemsetAsync(d.A, 0, bytes, c->stream));
        k_csr_to_dense<<<grid_for(n, 256), 256, 0, c->stream>>>(n, c->row_ptr.p, c->col.p,
                                                                c->val64.p, d.A);
        INS_HIP_CHECK(c, hipGetLastError());
        INS_HIP_CHECK(c, hipMemcpyAsync(c->sx64, c->sb, (size_t)n * sizeof(double),
                                        hipMemcpyDeviceToDevice, c->stream));
        INS_TRY(rocsolver_status(c, rocblas_create_handle(&d.h), "rocblas_create_handle"));
        INS_TRY(rocsolver_status(c, rocblas_set_stream(d.h, c->stream), "rocblas_set_stream"));
        const rocblas_int m = (rocblas_int)n;
        if (kind == INS_DIRECT_CHOLESKY) {
            INS_TRY(rocsolver_status(
                c, rocsolver_dpotrf(d.h, rocblas_fill_lower, m, d.A, m, d.info), "rocsolver_dpotrf"));
        } else {
            INS_HIP_CHECK(c, hipMalloc((void**)&d.ipiv, (size_t)n * sizeof(rocblas_int)));
            INS_TRY(rocsolver_status(c, rocsolver_dgetrf(d.h, m, m, d.A, m, d.ipiv, d.info),
                                     "rocsolver_dgetrf"));
        }
        rocblas_int info = 0;
        INS_HIP_CHECK(c, hipMemcpyAsync(&info, d.info, sizeof(info), hipMemcpyDeviceToHost,
                                        c->stream));
        INS_HIP_CHECK(c, hipStreamSynchronize(c->stream));
        if (info != 0)
            return fail(c, INS_ERR_BREAKDOWN,
                        std::string(kind == INS_DIRECT_CHOLESKY
                                        ? "Cholesky factorisation: non-positive pivot "
                                        : "LU factorisation: zero pivot ") +
                            std::to_string((int)info) + " of " + std::to_string(n));
        if (kind == INS_DIRECT_CHOLESKY)
            INS_TRY(rocsolver_status(
                c, rocsolver_dpotrs(d.h, rocblas_fill_lower, m, 1, d.A, m, c->sx64, m),
                "rocsolver_dpotrs"));
        else
            INS_TRY(rocsolver_status(c, rocsolver_dgetrs(d.h, rocblas_operation_none, m, 1, d.A, m,
                                                         d.ipiv, c->sx64, m),
                                     "rocsolver_dgetrs"));
        INS_TRY(launch_resid(c, INS_FMT_CSR));
        INS_TRY(read_scal(c));
        rel = std::sqrt(c->h_scal->tt) / bnorm;
    } else {
        INS_HIP_CHECK(c, hipMemsetAsync(c->sx64, 0, (size_t)n * sizeof(double), c->stream));
    }
    INS_HIP_CHECK(c, hipEventRecord(c->ev1, c->stream));
    INS_HIP_CHECK(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    INS_HIP_CHECK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->stats.solve_ms = ms;
    c->stats.relres = rel;
    if (relres) *relres = rel;
    c->have_solution = true;
    return INS_OK;
}

// ===========================================================================
// Serving: many mixed-precision solves batched into shared launches.
//
// Every running request keeps its matrix (SELL, FP32 and FP64 values), its
// Jacobi diagonal and its iterates resident in its own context.  One chunk of
// the server runs k iterations of all of them with three launches per
// iteration (kb_spmv, kb_update_xr, kb_update_p); block b of a launch serves
// the request whose block range holds b, with the block index and count that
// request has when it is solved alone, so its sums and iterates are the same
// alone and in any mix.  Residual replacements and the final residual run per
// request, on the server stream, between chunks.
namespace {

constexpr int MAXS = 64;  // requests in one batch

// Device pointers in the table are declared as global memory in the device
// pass, so the kernels that read them use global and scalar loads (as for
// kernel arguments) rather than flat ones.
#if defined(__HIP_DEVICE_COMPILE__)
#define INS_GLOBAL __attribute__((address_space(1)))
#else
#define INS_GLOBAL
#endif

struct GSellView {
    const int32_t INS_GLOBAL* perm;
    const int64_t INS_GLOBAL* sptr;
    const int32_t INS_GLOBAL* spat;
    const int32_t INS_GLOBAL* pptr;
    const int32_t INS_GLOBAL* plen;
    const int32_t INS_GLOBAL* pdel;
};

struct SlotDesc {
    int64_t n;
    GSellView S;
    const float INS_GLOBAL* val;
    const float INS_GLOBAL* dinv;
    float INS_GLOBAL* x;
    float INS_GLOBAL* r;
    float INS_GLOBAL* p;
    float INS_GLOBAL* q;
    double INS_GLOBAL* part;
    Scal INS_GLOBAL* sc;
    const double INS_GLOBAL* v64;  // FP64 slice values of the residual replacement (the
                                   // symmetric ones when vs is set)
    double INS_GLOBAL* x64;
    const double INS_GLOBAL* b;
    const float INS_GLOBAL* vs;     // symmetric slices of the FP32 sweeps (null: none)
    const int32_t INS_GLOBAL* su;
    int nu;
    int sdpn;
    int64_t snodes;
    int slines;
    int64_t join;  // batched iteration of the request's first launch
    int g_spmv, g_vec;
};

__device__ inline SellView sell_of(const GSellView& g)
{
    SellView v;
    v.perm = (const int32_t*)g.perm;
    v.sptr = (const int64_t*)g.sptr;
    v.spat = (const int32_t*)g.spat;
    v.pptr = (const int32_t*)g.pptr;
    v.plen = (const int32_t*)g.plen;
    v.pdel = (const int32_t*)g.pdel;
    return v;
}

struct BatchTab {
    int nslot, pad;
    int pre_spmv[MAXS + 1];  // first block of every request in kb_spmv
    int pre_vec[MAXS + 1];   // ... in kb_update_xr / kb_update_p
    SlotDesc d[MAXS];
};

// request of block bid: the last s with pre[s] <= bid
__device__ inline int find_slot(const int* pre, int ns, int bid)
{
    int lo = 0, hi = ns - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (pre[mid] <= bid) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__global__ __launch_bounds__(SB) void kb_spmv(const BatchTab* __restrict__ bt, int64_t it)
{
    const int s = find_slot(bt->pre_spmv, bt->nslot, blockIdx.x);
    const SlotDesc& d = bt->d[s];
    const Rep<double> R{(const double*)d.v64, (const double*)d.x64, (const double*)d.b,
                        (float*)d.r, (const float*)d.dinv, d.vs != nullptr ? 1 : 0};
    spmv_sell_body<float, float, float, 4, double, 2>(Blk{(int)blockIdx.x - bt->pre_spmv[s], d.g_spmv},
                                                   d.n, sell_of(d.S), (const float*)d.val,
                                                   (const float*)d.p, (float*)d.q, (double*)d.part,
                                                   (const Scal*)d.sc, (int)(it - d.join), R,
                                                   SymView{(const float*)d.vs,
                                                           (const int32_t*)d.su, d.nu, d.sdpn,
                                                           d.snodes, d.slines, nullptr});
}

__global__ __launch_bounds__(VB) void kb_update_xr(const BatchTab* __restrict__ bt, int64_t it)
{
    const int s = find_slot(bt->pre_vec, bt->nslot, blockIdx.x);
    const SlotDesc& d = bt->d[s];
    update_xr_body<float, float>(Blk{(int)blockIdx.x - bt->pre_vec[s], d.g_vec}, d.n, (float*)d.x,
                                 (float*)d.r, (const float*)d.p, (const float*)d.q,
                                 (const float*)d.dinv, (double*)d.part, d.g_spmv, (Scal*)d.sc,
                                 (int)(it - d.join));
}

__global__ __launch_bounds__(VB) void kb_update_p(const BatchTab* __restrict__ bt, int64_t it)
{
    const int s = find_slot(bt->pre_vec, bt->nslot, blockIdx.x);
    const SlotDesc& d = bt->d[s];
    update_p_body<float, float>(Blk{(int)blockIdx.x - bt->pre_vec[s], d.g_vec}, d.n, (float*)d.p,
                                (const float*)d.r, (const float*)d.dinv, (const double*)d.part,
                                d.g_vec, (Scal*)d.sc, (int)(it - d.join), (double*)d.x64,
                                (float*)d.x);
}

}  // namespace

struct SlotH {
    Ctx* c = nullptr;
    int state = INS_SLOT_FREE;
    int status = INS_OK;
    double rtol = 0.;
    int max_it = 0;
    int64_t join_it = 0;      // index of the request's first batched iteration
    int64_t join_chunk = 0;   // first chunk whose snapshot includes its start
    int refinements = 0;
    int iters = 0;
    double relres = 0.;
    bool relres_known = false;
};

struct Server {
    int device = 0;
    hipStream_t stream = nullptr;
    int max_slots = 0;
    std::string err;
    std::vector<SlotH> slots;
    Scal* scal = nullptr;                 // device, one per slot
    static constexpr int NSNAP = 4;
    Scal* h_snap = nullptr;               // pinned, NSNAP x max_slots
    hipEvent_t snap_ev[NSNAP] = {};
    std::deque<std::pair<int64_t, int>> inflight;  // (chunk, snapshot ring entry)
    static constexpr int NTAB = 4;
    BatchTab* h_tab = nullptr;            // pinned staging ring
    BatchTab* d_tab = nullptr;            // device ring
    hipEvent_t tab_ev[NTAB] = {};
    int tab_cur = 0;
    bool dirty = true;
    int nrun = 0;
    int grid_spmv = 0, grid_vec = 0;
    int64_t it = 0;      // batched iterations launched
    int64_t chunks = 0;  // chunks launched
    hipEvent_t ev_sync = nullptr;
};

namespace {

int sfail(Server* s, int code, const std::string& msg)
{
    s->err = msg;
    return code;
}

#define SRV_HIP(s, call)                                                          \
    do {                                                                          \
        hipError_t e_ = (call);                                                   \
        if (e_ != hipSuccess)                                                     \
            return sfail((s), INS_ERR_HIP, std::string(#call ": ") + hipGetErrorString(e_)); \
    } while (0)

// work of request s finished on the server stream before later work of its context
int hand_back(Server* sv, SlotH& h)
{
    SRV_HIP(sv, hipEventRecord(sv->ev_sync, sv->stream));
    SRV_HIP(sv, hipStreamWaitEvent(h.c->stream, sv->ev_sync, 0));
    return INS_OK;
}

int finish(Server* sv, int s, int status, const Scal* S)
{
    SlotH& h = sv->slots[s];
    if (solve_vectors_out(h.c, sv->stream) != INS_OK) return sfail(sv, INS_ERR_HIP, h.c->err);
    h.state = INS_SLOT_FINISHED;
    h.status = status;
    h.c->have_solution = true;
    if (S) {
        h.refinements = S->nrep;
        h.iters = S->iters;
        h.relres = S->bb > 0. ? std::sqrt(S->tt / S->bb) : 0.;
        h.relres_known = true;
        h.c->stats.iterations = h.iters;
        h.c->stats.relres = h.relres;
    }
    h.c->stats.refinements = h.refinements;
    sv->nrun--;
    sv->dirty = true;
    return hand_back(sv, h);
}

int process_snapshot(Server* sv, int64_t chunk, const Scal* snap)
{
    for (int s = 0; s < sv->max_slots; s++) {
        SlotH& h = sv->slots[s];
        if (h.state != INS_SLOT_RUNNING || chunk < h.join_chunk) continue;
        const Scal& S = snap[s];
        if (S.fin) {
            INS_TRY(finish(sv, s, INS_OK, &S));
        } else if (S.breakdown) {
            INS_TRY(finish(sv, s, INS_ERR_BREAKDOWN, &S));
        } else if (S.iters >= h.max_it) {
            // iteration limit: fold what was computed and measure it
            Ctx* c = h.c;
            k_fold<<<vec_grid(c->n_active), VB, 0, sv->stream>>>(c->n_active, c->sx64, c->x32.p);
            SRV_HIP(sv, hipGetLastError());
            if (launch_resid(c, INS_FMT_SELL, sv->stream, sv->scal + s) != INS_OK)
                return sfail(sv, INS_ERR_HIP, c->err);
            h.iters = S.iters;
            h.relres_known = false;
            INS_TRY(finish(sv, s, INS_ERR_NOCONV, nullptr));
        }
    }
    return INS_OK;
}

int upload_table(Server* sv)
{
    const int t = (sv->tab_cur + 1) % Server::NTAB;
    SRV_HIP(sv, hipEventSynchronize(sv->tab_ev[t]));
    BatchTab& T = sv->h_tab[t];
    int ns = 0, gs = 0, gvv = 0;
    for (int s = 0; s < sv->max_slots; s++) {
        const SlotH& h = sv->slots[s];
        if (h.state != INS_SLOT_RUNNING) continue;
        Ctx* c = h.c;
        SlotDesc& d = T.d[ns];
        d.n = c->n_active;
        const SellView sv0 = sell_view(c);
#define INS_AS(field, ptr) field = (decltype(field))(ptr)
        INS_AS(d.S.perm, sv0.perm);
        INS_AS(d.S.sptr, sv0.sptr);
        INS_AS(d.S.spat, sv0.spat);
        INS_AS(d.S.pptr, sv0.pptr);
        INS_AS(d.S.plen, sv0.plen);
        INS_AS(d.S.pdel, sv0.pdel);
        INS_AS(d.val, c->sell.v32.p);
        INS_AS(d.dinv, c->sd32);
        INS_AS(d.x, c->x32.p);
        INS_AS(d.r, c->r32.p);
        INS_AS(d.p, c->p32.p);
        INS_AS(d.q, c->q32.p);
        INS_AS(d.part, c->partial.p);
        INS_AS(d.sc, sv->scal + s);
        const bool sym = use_sym(c, c->sperm);
        INS_AS(d.v64, sym ? c->sell.vs64.p : c->sell.v64.p);
        INS_AS(d.x64, c->sx64);
        INS_AS(d.b, c->sb);
        INS_AS(d.vs, sym ? c->sell.vs.p : nullptr);
        INS_AS(d.su, sym ? c->sell.su.p : nullptr);
        d.nu = sym ? c->sell.nu : 0;
        d.sdpn = sym ? c->sell.sdpn : 1;
        d.snodes = sym ? c->sell.snodes : 0;
        d.slines = sym && c->sell.slines ? 1 : 0;
#undef INS_AS
        d.g_spmv = (int)spmv_grid(c, sym, false);
        d.g_vec = (int)vec_grid(d.n);
        d.join = h.join_it;
        T.pre_spmv[ns] = gs;
        T.pre_vec[ns] = gvv;
        gs += d.g_spmv;
        gvv += d.g_vec;
        ns++;
    }
    T.nslot = ns;
    T.pre_spmv[ns] = gs;
    T.pre_vec[ns] = gvv;
    sv->grid_spmv = gs;
    sv->grid_vec = gvv;
    SRV_HIP(sv, hipMemcpyAsync(sv->d_tab + t, &T, sizeof(BatchTab), hipMemcpyHostToDevice,
                               sv->stream));
    SRV_HIP(sv, hipEventRecord(sv->tab_ev[t], sv->stream));
    sv->tab_cur = t;
    sv->dirty = false;
    return INS_OK;
}

int pop_snapshot(Server* sv)
{
    const auto f = sv->inflight.front();
    sv->inflight.pop_front();
    SRV_HIP(sv, hipEventSynchronize(sv->snap_ev[f.second]));
    return process_snapshot(sv, f.first, sv->h_snap + (size_t)f.second * sv->max_slots);
}

}  // namespace

int server_create(Server** out, int device, int max_slots)
{
    if (!out || max_slots < 1 || max_slots > MAXS) return INS_ERR_ARG;
    *out = nullptr;
    if (hipSetDevice(device) != hipSuccess) return INS_ERR_HIP;
    Server* sv = new (std::nothrow) Server();
    if (!sv) return INS_ERR_OOM;
    sv->device = device;
    sv->max_slots = max_slots;
    sv->slots.resize(max_slots);
    bool ok = hipStreamCreateWithFlags(&sv->stream, hipStreamNonBlocking) == hipSuccess &&
              hipMalloc((void**)&sv->scal, max_slots * sizeof(Scal)) == hipSuccess &&
              hipHostMalloc((void**)&sv->h_snap, Server::NSNAP * max_slots * sizeof(Scal), 0) ==
                  hipSuccess &&
              hipHostMalloc((void**)&sv->h_tab, Server::NTAB * sizeof(BatchTab), 0) == hipSuccess &&
              hipMalloc((void**)&sv->d_tab, Server::NTAB * sizeof(BatchTab)) == hipSuccess &&
              hipEventCreateWithFlags(&sv->ev_sync, hipEventDisableTiming) == hipSuccess;
    for (int i = 0; ok && i < Server::NSNAP; i++)
        ok = hipEventCreateWithFlags(&sv->snap_ev[i], hipEventDisableTiming) == hipSuccess;
    for (int i = 0; ok && i < Server::NTAB; i++)
        ok = hipEventCreateWithFlags(&sv->tab_ev[i], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        server_destroy(sv);
        return INS_ERR_HIP;
    }
    *out = sv;
    return INS_OK;
}

void server_destroy(Server* sv)
{
    if (!sv) return;
    (void)hipSetDevice(sv->device);
    if (sv->stream) (void)hipStreamSynchronize(sv->stream);
    for (auto& e : sv->snap_ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : sv->tab_ev)
        if (e) (void)hipEventDestroy(e);
    if (sv->ev_sync) (void)hipEventDestroy(sv->ev_sync);
    if (sv->scal) (void)hipFree(sv->scal);
    if (sv->d_tab) (void)hipFree(sv->d_tab);
    if (sv->h_snap) (void)hipHostFree(sv->h_snap);
    if (sv->h_tab) (void)hipHostFree(sv->h_tab);
    if (sv->stream) (void)hipStreamDestroy(sv->stream);
    delete sv;
}

const char* server_error(Server* sv) { return sv ? sv->err.c_str() : "null server"; }

int server_attach(Server* sv, Ctx* c, double rtol, int max_it, int* slot)
{
    if (!sv || !c || !slot || !(rtol > 0.) || max_it < 1) return INS_ERR_ARG;
    if (c->device != sv->device) return sfail(sv, INS_ERR_ARG, "context on another device");
    if (!c->have_matrix || !c->have_rhs)
        return sfail(sv, INS_ERR_STATE, "request context not assembled (matrix and rhs)");
    for (int s = 0; s < sv->max_slots; s++)
        if (sv->slots[s].state != INS_SLOT_FREE && sv->slots[s].c == c)
            return sfail(sv, INS_ERR_STATE, "context already attached");
    int s = -1;
    for (int k = 0; k < sv->max_slots && s < 0; k++)
        if (sv->slots[k].state == INS_SLOT_FREE) s = k;
    if (s < 0) return sfail(sv, INS_ERR_STATE, "no free slot");
    auto ctx_fail = [&](int st) { return sfail(sv, st, c->err); };
    int st;
    if ((!c->sell.built || !c->sell.ok) && (st = build_sell(c, true)) != INS_OK)
        return ctx_fail(st);
    if (!c->sell.ok)
        return sfail(sv, INS_ERR_UNSUPPORTED, "matrix has no SELL form (irregular rows)");
    if (c->matrix_precision != INS_FP64)
        return sfail(sv, INS_ERR_UNSUPPORTED, "serving needs an FP64 assembly (FP64 residual)");
    if (c->sell.stale && (st = fill_sell(c)) != INS_OK) return ctx_fail(st);
    c->want_format = INS_FMT_SELL;
    if ((st = ensure_solver_state(c, INS_SOLVE_MIXED)) != INS_OK) return ctx_fail(st);
    // the request's own setup (assembly, SELL) precedes its start
    SRV_HIP(sv, hipEventRecord(sv->ev_sync, c->stream));
    SRV_HIP(sv, hipStreamWaitEvent(sv->stream, sv->ev_sync, 0));
    const int64_t n = c->n_active;
    const unsigned gv = vec_grid(n);
    if ((st = solve_vectors(c, INS_FMT_SELL, INS_SOLVE_MIXED, sv->stream)) != INS_OK)
        return ctx_fail(st);
    SRV_HIP(sv, hipMemsetAsync(c->sx64, 0, n * sizeof(double), sv->stream));
    k_init<float, float><<<gv, VB, 0, sv->stream>>>(n, c->sb, c->x32.p, c->r32.p, c->p32.p,
                                                    c->sd32, c->partial.p);
    k_init_fin<<<1, VB, 0, sv->stream>>>(c->partial.p, (int)gv, rtol * rtol, mixed_delta2(),
                                 sv->scal + s);
    SRV_HIP(sv, hipGetLastError());
    SlotH& h = sv->slots[s];
    h = SlotH();
    h.c = c;
    h.state = INS_SLOT_RUNNING;
    h.rtol = rtol;
    h.max_it = max_it;
    h.join_it = sv->it;
    h.join_chunk = sv->chunks;
    c->have_solution = false;
    sv->nrun++;
    sv->dirty = true;
    *slot = s;
    return INS_OK;
}

int server_step(Server* sv, int k, int* n_finished)
{
    if (!sv || k < 1) return INS_ERR_ARG;
    int before = 0;
    for (const auto& h : sv->slots) before += (h.state == INS_SLOT_FINISHED);
    if (sv->nrun > 0) {
        if (sv->dirty) INS_TRY(upload_table(sv));
        const BatchTab* tab = sv->d_tab + sv->tab_cur;
        for (int i = 0; i < k; i++) {
            kb_spmv<<<sv->grid_spmv, SB, 0, sv->stream>>>(tab, sv->it);
            kb_update_xr<<<sv->grid_vec, VB, 0, sv->stream>>>(tab, sv->it);
            kb_update_p<<<sv->grid_vec, VB, 0, sv->stream>>>(tab, sv->it);
            sv->it++;
        }
        SRV_HIP(sv, hipGetLastError());
        const int r = (int)(sv->chunks % Server::NSNAP);
        SRV_HIP(sv, hipMemcpyAsync(sv->h_snap + (size_t)r * sv->max_slots, sv->scal,
                                   sv->max_slots * sizeof(Scal), hipMemcpyDeviceToHost,
                                   sv->stream));
        SRV_HIP(sv, hipEventRecord(sv->snap_ev[r], sv->stream));
        sv->inflight.emplace_back(sv->chunks, r);
        sv->chunks++;
    }
    // keep one chunk queued behind the one being examined
    while (sv->inflight.size() >= 2 || (!sv->inflight.empty() && sv->nrun == 0))
        INS_TRY(pop_snapshot(sv));
    int after = 0;
    for (const auto& h : sv->slots) after += (h.state == INS_SLOT_FINISHED);
    if (n_finished) *n_finished = after - before;
    return INS_OK;
}

int server_poll(Server* sv, int s, int* state, int* status, int* iters, double* relres)
{
    if (!sv || s < 0 || s >= sv->max_slots) return INS_ERR_ARG;
    SlotH& h = sv->slots[s];
    if (h.state == INS_SLOT_FINISHED && !h.relres_known) {
        Scal S;
        SRV_HIP(sv, hipStreamSynchronize(sv->stream));
        SRV_HIP(sv, hipMemcpy(&S, sv->scal + s, sizeof(Scal), hipMemcpyDeviceToHost));
        h.relres = S.bb > 0. ? std::sqrt(S.tt / S.bb) : 0.;
        h.relres_known = true;
        h.c->stats.iterations = h.iters;
        h.c->stats.relres = h.relres;
    }
    if (state) *state = h.state;
    if (status) *status = h.status;
    if (iters) *iters = h.state == INS_SLOT_FREE ? 0 : h.iters;
    if (relres) *relres = h.relres;
    return INS_OK;
}

int server_detach(Server* sv, int s)
{
    if (!sv || s < 0 || s >= sv->max_slots) return INS_ERR_ARG;
    SlotH& h = sv->slots[s];
    if (h.state == INS_SLOT_FREE) return sfail(sv, INS_ERR_STATE, "slot not in use");
    // the chunk queued behind the one that showed the request finished still
    // has the request in its table: the context's next call waits for the
    // server stream's work up to here before its buffers change hands
    if (!h.c->srv_ev)
        SRV_HIP(sv, hipEventCreateWithFlags(&h.c->srv_ev, hipEventDisableTiming));
    SRV_HIP(sv, hipEventRecord(h.c->srv_ev, sv->stream));
    h.c->srv_pending = true;
    if (h.state == INS_SLOT_RUNNING) {
        // cancel: nothing of this request may still be queued when it leaves
        h.state = INS_SLOT_FREE;
        sv->nrun--;
        sv->dirty = true;
        SRV_HIP(sv, hipStreamSynchronize(sv->stream));
        sv->inflight.clear();
    } else if (h.status != INS_OK && h.status != INS_ERR_BREAKDOWN) {
        // queued chunks may still iterate on its vectors (its stop flag is not set)
        SRV_HIP(sv, hipStreamSynchronize(sv->stream));
    }
    h = SlotH();
    return INS_OK;
}

int server_running(Server* sv) { return sv ? sv->nrun : 0; }

// ===========================================================================
// Allocation-free FP64 CSR solve (ins_solve_csr_noalloc).
//
// The Jacobi-PCG of ins_solve(INS_SOLVE_FP64, INS_FMT_CSR) on a CSR system in
// caller-owned device arrays, with the same kernels, grids and reduction
// order (the iterates are the bits ins_solve computes for that system).  Its
// scratch -- the device scalars, the block partials, the Jacobi diagonal and
// the iterates r, p, q -- lives in `scratch`, a static device array of fixed
// size: the call allocates nothing (no hipMalloc, hipFree, pinned host
// memory or events).  Every GPU holds its own copy of the array from the
// first call on it until the process ends, and every call on that GPU uses
// it: calls that overlap on one GPU race.  Nothing in it has to start at
// zero (each quantity is written before it is read).
constexpr size_t NA_SCAL = 256;                                  // Scal, padded
constexpr size_t NA_PART = (size_t)NSLOT * MAXG * sizeof(double);
constexpr size_t NA_HEAD = NA_SCAL + NA_PART;
constexpr size_t NA_BYTES = NA_HEAD + (size_t)4 * INS_NOALLOC_MAX_ROWS * sizeof(double);
static_assert(sizeof(Scal) <= NA_SCAL, "scalars fit their slot");
static_assert(NA_PART % 256 == 0, "diagonal and iterates stay 256-byte aligned");

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

namespace {

constexpr int NA_MAXDEV = 64;
std::once_flag na_once[NA_MAXDEV];
void* na_ptr[NA_MAXDEV];
hipError_t na_rc[NA_MAXDEV];

// Jacobi diagonal 1 / a_rr of a CSR with ascending columns (1 where the
// diagonal is absent or zero), as the context's assembly computes it
__global__ void k_na_dinv(int64_t n, const int32_t* __restrict__ row_ptr,
                          const int32_t* __restrict__ col, const double* __restrict__ val,
                          double* __restrict__ d)
{
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
         r += (int64_t)gridDim.x * blockDim.x) {
        int32_t lo = row_ptr[r], hi = row_ptr[r + 1];
        const int32_t end = hi;
        while (lo < hi) {
            const int32_t mid = (lo + hi) >> 1;
            if (col[mid] < r) lo = mid + 1; else hi = mid;
        }
        const double a = (lo < end && col[lo] == r) ? val[lo] : 0.;
        d[r] = a != 0. ? 1. / a : 1.;
    }
}

template <int LPR>
void na_spmv(hipStream_t st, unsigned g, int64_t n, const int32_t* rp, const int32_t* col,
             const double* val, const double* p, double* q, double* part, const Scal* sc, int it,
             const Rep<double>& R)
{
    k_spmv_csr<LPR, double, double, double, double><<<g, SB, 0, st>>>(n, rp, col, val, p, q, part,
                                                                       sc, it, R);
}

template <int LPR>
void na_resid(hipStream_t st, unsigned g, int64_t n, const int32_t* rp, const int32_t* col,
              const double* val, const double* x, const double* b, double* part)
{
    k_resid_csr<LPR, double><<<g, SB, 0, st>>>(n, rp, col, val, x, b, part);
}

int na_read(hipStream_t st, const Scal* dsc, Scal* h)
{
    if (hipMemcpyAsync(h, dsc, sizeof(Scal), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return INS_ERR_HIP;
    return INS_OK;
}

}  // namespace

int solve_csr_noalloc(int device, void* stream, int64_t n, int64_t nnz, const int32_t* rp,
                      const int32_t* col, const double* val, const double* b, double* x,
                      double rtol, int max_it, int* iters, double* relres)
{
    if (n <= 0 || n > INS_NOALLOC_MAX_ROWS || nnz < n || nnz > INT32_MAX || !rp || !col ||
        !val || !b || !x || !(rtol > 0.) || max_it < 0 || device < 0 || device >= NA_MAXDEV)
        return INS_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device >= ndev) return INS_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return INS_ERR_HIP;
    // the device address of this GPU's copy of the scratch array, resolved
    // once per device with that device current
    std::call_once(na_once[device], [device] {
        na_rc[device] = hipGetSymbolAddress(&na_ptr[device], HIP_SYMBOL(scratch));
    });
    if (na_rc[device] != hipSuccess || !na_ptr[device]) return INS_ERR_HIP;
    hipStream_t st = static_cast<hipStream_t>(stream);
    // the CSR must end at nnz (a 4-byte read of row_ptr[n])
    int32_t last = -1;
    if (hipMemcpyAsync(&last, rp + n, sizeof(int32_t), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return INS_ERR_HIP;
    if ((int64_t)last != nnz) return INS_ERR_ARG;

    unsigned char* base = static_cast<unsigned char*>(na_ptr[device]);
    Scal* sc = reinterpret_cast<Scal*>(base);
    double* part = reinterpret_cast<double*>(base + NA_SCAL);
    double* dinv = reinterpret_cast<double*>(base + NA_HEAD);
    double* r = dinv + n;
    double* p = r + n;
    double* q = p + n;

    const unsigned gv = vec_grid(n);
    const int lpr = lpr_for(n, nnz);
    const unsigned g1 = csr_grid(n, lpr);
    const Rep<double> R{val, x, b, nullptr, nullptr};   // FP64: no replacement
    k_na_dinv<<<grid_for(n, 256), 256, 0, st>>>(n, rp, col, val, dinv);
    k_init<double, double><<<gv, VB, 0, st>>>(n, b, x, r, p, dinv, part);
    k_init_fin<<<1, VB, 0, st>>>(part, (int)gv, rtol * rtol, 0., sc);
    if (hipGetLastError() != hipSuccess) return INS_ERR_HIP;
    Scal S;
    INS_TRY(na_read(st, sc, &S));
    int enq = 0;
    while (!S.done && !S.stop && enq < max_it) {
        const int k = std::min(CHUNK, max_it - enq);
        for (int i = 0; i < k; i++) {
            const int it = enq + i;
            if (lpr == 32) na_spmv<32>(st, g1, n, rp, col, val, p, q, part, sc, it, R);
            else if (lpr == 16) na_spmv<16>(st, g1, n, rp, col, val, p, q, part, sc, it, R);
            else if (lpr == 8) na_spmv<8>(st, g1, n, rp, col, val, p, q, part, sc, it, R);
            else na_spmv<4>(st, g1, n, rp, col, val, p, q, part, sc, it, R);
            k_update_xr<double, double><<<gv, VB, 0, st>>>(n, x, r, p, q, dinv, part, (int)g1,
                                                            sc, it);
            k_update_p<double, double><<<gv, VB, 0, st>>>(n, p, r, dinv, part, (int)gv, sc, it, x,
                                                           x);
        }
        if (hipGetLastError() != hipSuccess) return INS_ERR_HIP;
        enq += k;
        INS_TRY(na_read(st, sc, &S));
    }
    // |b - A x| in FP64 over the same rows
    if (lpr == 32) na_resid<32>(st, g1, n, rp, col, val, x, b, part);
    else if (lpr == 16) na_resid<16>(st, g1, n, rp, col, val, x, b, part);
    else if (lpr == 8) na_resid<8>(st, g1, n, rp, col, val, x, b, part);
    else na_resid<4>(st, g1, n, rp, col, val, x, b, part);
    k_resid_fin<<<1, VB, 0, st>>>(part, (int)g1, sc);
    if (hipGetLastError() != hipSuccess) return INS_ERR_HIP;
    INS_TRY(na_read(st, sc, &S));
    const double rel = S.bb > 0. ? std::sqrt(S.tt / S.bb) : 0.;
    if (iters) *iters = S.iters;
    if (relres) *relres = rel;
    if (S.breakdown) return INS_ERR_BREAKDOWN;
    if (!(rel <= rtol)) return INS_ERR_NOCONV;
    return INS_OK;
}

}  // namespace ins
