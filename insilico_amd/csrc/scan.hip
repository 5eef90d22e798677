// Device-wide exclusive prefix sum of int32 counts (row lengths, adjacency
// degrees).  Three passes: chunk sums, a scan of the chunk sums (recursive
// for large inputs), and a chunk-local scan with the chunk offset.  The
// result is deterministic (integer arithmetic).
#include "common.hpp"

namespace ins {

namespace {

constexpr int SB = 256;          // threads per block
constexpr int SPT = 4;           // items per thread
constexpr int SCHUNK = SB * SPT; // items per block

__device__ inline int64_t block_exclusive_scan(int64_t v, int64_t* total)
{
    __shared__ int64_t s[SB];
    const int t = threadIdx.x;
    s[t] = v;
    __syncthreads();
    for (int off = 1; off < SB; off <<= 1) {
        int64_t add = (t >= off) ? s[t - off] : 0;
        __syncthreads();
        s[t] += add;
        __syncthreads();
    }
    const int64_t incl = s[t];
    *total = s[SB - 1];
    __syncthreads();
    return incl - v;
}

template <typename TI>
__global__ __launch_bounds__(SB) void k_chunk_sums(const TI* in, int64_t n,
                                                   int64_t* sums)
{
    const int64_t base = (int64_t)blockIdx.x * SCHUNK;
    int64_t v = 0;
    for (int k = 0; k < SPT; k++) {
        const int64_t i = base + (int64_t)k * SB + threadIdx.x;
        if (i < n) v += (int64_t)in[i];
    }
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    __shared__ int64_t w[SB / 64];
    if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        int64_t s = 0;
        for (int i = 0; i < SB / 64; i++) s += w[i];
        sums[blockIdx.x] = s;
    }
}

// scans `in` chunk-wise, adding offsets[blockIdx] (may be null); writes
// exclusive results to out (TO) and the grand total to out[n] when last.
template <typename TI, typename TO>
__global__ __launch_bounds__(SB) void k_chunk_scan(const TI* in, int64_t n,
                                                   const int64_t* offsets,
                                                   TO* out, int write_total)
{
    const int64_t base = (int64_t)blockIdx.x * SCHUNK;
    const int t = threadIdx.x;
    int64_t vals[SPT];
    int64_t mine = 0;
    for (int k = 0; k < SPT; k++) {
        const int64_t i = base + (int64_t)t * SPT + k;  // contiguous per thread
        vals[k] = (i < n) ? (int64_t)in[i] : 0;
        mine += vals[k];
    }
    int64_t tot;
    int64_t ex = block_exclusive_scan(mine, &tot);
    const int64_t off = offsets ? offsets[blockIdx.x] : 0;
    int64_t run = off + ex;
    for (int k = 0; k < SPT; k++) {
        const int64_t i = base + (int64_t)t * SPT + k;
        if (i < n) out[i] = (TO)run;
        run += vals[k];
    }
    if (write_total && blockIdx.x == gridDim.x - 1 && t == SB - 1) out[n] = (TO)(off + tot);
}

// exclusive scan of int64 values in place into `out` (n+1 entries)
int scan_i64(Ctx* c, const int64_t* in, int64_t n, int64_t* out)
{
    const int64_t nb = (n + SCHUNK - 1) / SCHUNK;
    if (nb <= 1) {
        k_chunk_scan<int64_t, int64_t><<<1, SB, 0, c->stream>>>(in, n, nullptr, out, 1);
        INS_HIP_CHECK(c, hipGetLastError());
        return INS_OK;
    }
    int64_t* sums = nullptr;
    int64_t* soff = nullptr;
    INS_HIP_CHECK(c, hipMallocAsync((void**)&sums, nb * sizeof(int64_t), c->stream));
    INS_HIP_CHECK(c, hipMallocAsync((void**)&soff, (nb + 1) * sizeof(int64_t), c->stream));
    k_chunk_sums<int64_t><<<(unsigned)nb, SB, 0, c->stream>>>(in, n, sums);
    INS_HIP_CHECK(c, hipGetLastError());
    int s = scan_i64(c, sums, nb, soff);
    if (s == INS_OK) {
        k_chunk_scan<int64_t, int64_t><<<(unsigned)nb, SB, 0, c->stream>>>(in, n, soff, out, 1);
        INS_HIP_CHECK(c, hipGetLastError());
    }
    INS_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    (void)hipFreeAsync(sums, c->stream);
    (void)hipFreeAsync(soff, c->stream);
    return s;
}

}  // namespace

// out has n+1 entries; *total receives out[n].  Fails if the total does not
// fit an int32 (indices are int32 throughout, as Eigen's int indices).
int exclusive_scan_i32(Ctx* c, const int32_t* in, int32_t* out, int64_t n,
                       int64_t* total)
{
    const int64_t nb = (n + SCHUNK - 1) / SCHUNK;
    int64_t* sums = nullptr;
    int64_t* soff = nullptr;
    INS_HIP_CHECK(c, hipMallocAsync((void**)&sums, (nb + 1) * sizeof(int64_t), c->stream));
    INS_HIP_CHECK(c, hipMallocAsync((void**)&soff, (nb + 1) * sizeof(int64_t), c->stream));
    if (nb > 0) {
        k_chunk_sums<int32_t><<<(unsigned)nb, SB, 0, c->stream>>>(in, n, sums);
        INS_HIP_CHECK(c, hipGetLastError());
    }
    int s = scan_i64(c, sums, nb, soff);
    if (s != INS_OK) return s;
    int64_t tot = 0;
    INS_HIP_CHECK(c, hipMemcpyAsync(&tot, soff + nb, sizeof(int64_t),
                                    hipMemcpyDeviceToHost, c->stream));
    INS_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    if (tot > INT32_MAX) {
        (void)hipFreeAsync(sums, c->stream);
        (void)hipFreeAsync(soff, c->stream);
        return fail(c, INS_ERR_UNSUPPORTED, "index total exceeds int32 range");
    }
    if (nb > 0) {
        k_chunk_scan<int32_t, int32_t><<<(unsigned)nb, SB, 0, c->stream>>>(in, n, soff, out, 0);
        INS_HIP_CHECK(c, hipGetLastError());
    }
    int32_t t32 = (int32_t)tot;
    INS_HIP_CHECK(c, hipMemcpyAsync(out + n, &t32, sizeof(int32_t), hipMemcpyHostToDevice,
                                    c->stream));
    INS_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    (void)hipFreeAsync(sums, c->stream);
    (void)hipFreeAsync(soff, c->stream);
    *total = tot;
    return INS_OK;
}

}  // namespace ins
