// Setup kernels: node->item adjacency (items are elements or Neumann faces),
// node neighbour lists and the CSR pattern of the reduced system.
//
// The pattern is the one base/solver/TripletContainer.hpp:159-301 builds:
// every (active row, active col) pair of every element, explicit zeros kept,
// columns ascending.  With iso-parametric Q1 DoFs (DoF id = node id,
// base/dof/copyConnectivity.hpp:34-60) two components couple exactly when
// their nodes share an element, so the row of (node n, comp c) is the list of
// active components of the sorted neighbour nodes of n.
#include <type_traits>

#include "common.hpp"

namespace ins {

namespace {

constexpr int BLK = 256;
constexpr int NB_BLK = 64;    // threads per block of the neighbour kernels

// neighbour nodes per node kept in LDS: 27 for HEX8, 25 for QUAD9, 125 for
// HEX27 on a structured mesh; room for unstructured meshes up to the limit
template <int NPE>
struct MaxNb {
    static constexpr int value = NPE > 9 ? 128 : 96;
};

__global__ void k_adj_count(const int32_t* items, int64_t nent, int64_t n_nodes,
                            int32_t* cnt, int* err)
{
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nent;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int32_t v = items[t];
        if (v < 0 || v >= n_nodes) {
            atomicOr(err, 1);
            continue;
        }
        atomicAdd(&cnt[v], 1);
    }
}

__global__ void k_adj_fill(const int32_t* items, int64_t nent, int64_t n_nodes, int per_item,
                           int lg_stride, const int32_t* ptr, int32_t* fillpos, int32_t* list)
{
    for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < nent;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int32_t v = items[t];
        if (v < 0 || v >= n_nodes) continue;
        const int32_t slot = atomicAdd(&fillpos[v], 1);
        const int32_t item = (int32_t)(t / per_item);
        const int32_t local = (int32_t)(t - (int64_t)item * per_item);
        list[ptr[v] + slot] = (item << lg_stride) | local;
    }
}

// sort every node's list ascending (entries are unique); lists are short
__global__ void k_adj_sort(int64_t n_nodes, const int32_t* ptr, int32_t* list,
                           int* maxdeg)
{
    for (int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; n < n_nodes;
         n += (int64_t)gridDim.x * blockDim.x) {
        const int32_t b = ptr[n], e = ptr[n + 1];
        for (int32_t i = b + 1; i < e; i++) {
            const int32_t key = list[i];
            int32_t j = i - 1;
            while (j >= b && list[j] > key) {
                list[j + 1] = list[j];
                j--;
            }
            list[j + 1] = key;
        }
        atomicMax(maxdeg, e - b);
    }
}

// Collect the sorted unique neighbour nodes of node n into buf (LDS).
template <int NPE, int MAXNB = MaxNb<NPE>::value>
__device__ inline int collect_neighbours(int64_t n, const int32_t* adj_ptr,
                                         const int32_t* adj, const int32_t* conn,
                                         int32_t* buf, int lg_npe, int* err)
{
    int cnt = 0;
    const int32_t b = adj_ptr[n], e = adj_ptr[n + 1];
    for (int32_t k = b; k < e; k++) {
        const int32_t el = adj[k] >> lg_npe;
        for (int v = 0; v < NPE; v++) {
            const int32_t m = conn[(int64_t)el * NPE + v];
            // insert m keeping buf sorted and unique
            int pos = cnt;
            while (pos > 0 && buf[pos - 1] > m) pos--;
            if (pos > 0 && buf[pos - 1] == m) continue;
            if (cnt >= MAXNB) {
                atomicOr(err, 2);
                return cnt;
            }
            for (int i = cnt; i > pos; i--) buf[i] = buf[i - 1];
            buf[pos] = m;
            cnt++;
        }
    }
    return cnt;
}

__device__ inline int n_active_comps(const int32_t* dof_id, int64_t node, int dpn)
{
    int a = 0;
    for (int c = 0; c < dpn; c++) a += dof_id[node * dpn + c] >= 0;
    return a;
}

// pass 1: neighbour count and row length of every node
template <int NPE, int MAXNB = MaxNb<NPE>::value>
__global__ __launch_bounds__(NB_BLK) void k_nb_count(
    int64_t n_nodes, const int32_t* adj_ptr, const int32_t* adj, const int32_t* conn,
    const int32_t* dof_id, int dpn, int lg_npe, int32_t* nb_cnt, int32_t* row_len,
    int* err, int* maxnb)
{
    __shared__ int32_t sbuf[NB_BLK * MAXNB];
    int32_t* buf = sbuf + threadIdx.x * MAXNB;
    for (int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; n < n_nodes;
         n += (int64_t)gridDim.x * blockDim.x) {
        const int cnt = collect_neighbours<NPE>(n, adj_ptr, adj, conn, buf, lg_npe, err);
        int len = 0;
        for (int j = 0; j < cnt; j++) len += n_active_comps(dof_id, buf[j], dpn);
        nb_cnt[n] = cnt;
        atomicMax(maxnb, cnt);
        for (int c = 0; c < dpn; c++) {
            const int32_t r = dof_id[n * dpn + c];
            if (r >= 0) row_len[r] = len;
        }
    }
}

// pass 2: neighbour lists, per-neighbour column offsets, column indices
template <int NPE, int MAXNB = MaxNb<NPE>::value>
__global__ __launch_bounds__(NB_BLK) void k_nb_fill(
    int64_t n_nodes, const int32_t* adj_ptr, const int32_t* adj, const int32_t* conn,
    const int32_t* dof_id, int dpn, int lg_npe, const int32_t* nb_ptr, int32_t* nb,
    int32_t* nboff, const int32_t* row_ptr, int32_t* col, int* err)
{
    __shared__ int32_t sbuf[NB_BLK * MAXNB];
    int32_t* buf = sbuf + threadIdx.x * MAXNB;
    for (int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; n < n_nodes;
         n += (int64_t)gridDim.x * blockDim.x) {
        const int cnt = collect_neighbours<NPE>(n, adj_ptr, adj, conn, buf, lg_npe, err);
        const int32_t nbs = nb_ptr[n];
        int off = 0;
        for (int j = 0; j < cnt; j++) {
            nb[nbs + j] = buf[j];
            nboff[nbs + j] = off;
            off += n_active_comps(dof_id, buf[j], dpn);
        }
        for (int c = 0; c < dpn; c++) {
            const int32_t r = dof_id[n * dpn + c];
            if (r < 0) continue;
            int32_t pos = row_ptr[r];
            for (int j = 0; j < cnt; j++) {
                const int64_t m = buf[j];
                for (int d = 0; d < dpn; d++) {
                    const int32_t cid = dof_id[m * dpn + d];
                    if (cid >= 0) col[pos++] = cid;
                }
            }
        }
    }
}

// rows whose column ids are not ascending (non-monotone numbering): sort
__global__ void k_sort_rows(int64_t n, const int32_t* row_ptr, int32_t* col)
{
    for (int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; r < n;
         r += (int64_t)gridDim.x * blockDim.x) {
        const int32_t b = row_ptr[r], e = row_ptr[r + 1];
        for (int32_t i = b + 1; i < e; i++) {
            const int32_t key = col[i];
            int32_t j = i - 1;
            while (j >= b && col[j] > key) {
                col[j + 1] = col[j];
                j--;
            }
            col[j + 1] = key;
        }
    }
}

}  // namespace

int build_adjacency(Ctx* c, const int32_t* items, int64_t n_items, int per_item,
                    int lg_stride, int64_t n_nodes, DArr<int32_t>& ptr,
                    DArr<int32_t>& list, int* max_deg)
{
    const int64_t nent = n_items * per_item;
    INS_TRY(ensure(c, ptr, n_nodes + 1));
    INS_TRY(ensure(c, list, nent));
    int32_t* cnt = nullptr;
    int* flags = nullptr;
    INS_HIP_CHECK(c, hipMallocAsync((void**)&cnt, (n_nodes + 1) * sizeof(int32_t), c->stream));
    INS_HIP_CHECK(c, hipMallocAsync((void**)&flags, 2 * sizeof(int), c->stream));
    INS_HIP_CHECK(c, hipMemsetAsync(cnt, 0, (n_nodes + 1) * sizeof(int32_t), c->stream));
    INS_HIP_CHECK(c, hipMemsetAsync(flags, 0, 2 * sizeof(int), c->stream));
    k_adj_count<<<grid_for(nent, BLK), BLK, 0, c->stream>>>(items, nent, n_nodes, cnt, flags);
    INS_HIP_CHECK(c, hipGetLastError());
    int64_t total = 0;
    int s = exclusive_scan_i32(c, cnt, ptr.p, n_nodes, &total);
    if (s != INS_OK) return s;
    INS_HIP_CHECK(c, hipMemsetAsync(cnt, 0, (n_nodes + 1) * sizeof(int32_t), c->stream));
    k_adj_fill<<<grid_for(nent, BLK), BLK, 0, c->stream>>>(items, nent, n_nodes, per_item,
                                                          lg_stride, ptr.p, cnt, list.p);
    INS_HIP_CHECK(c, hipGetLastError());
    k_adj_sort<<<grid_for(n_nodes, BLK), BLK, 0, c->stream>>>(n_nodes, ptr.p, list.p,
                                                             flags + 1);
    INS_HIP_CHECK(c, hipGetLastError());
    int hflags[2];
    INS_HIP_CHECK(c, hipMemcpyAsync(hflags, flags, sizeof(hflags), hipMemcpyDeviceToHost,
                                    c->stream));
    INS_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    (void)hipFreeAsync(cnt, c->stream);
    (void)hipFreeAsync(flags, c->stream);
    if (hflags[0]) return fail(c, INS_ERR_RANGE, "node id out of range in connectivity");
    if (max_deg) *max_deg = hflags[1];
    return INS_OK;
}

// Shared elements of every (node, neighbour) pair, for nodes of at most 8
// elements: for element i of the node (adjacency order, ascending element
// id) field i holds the presence bit + the neighbour's local index in it, or
// 0 (4-bit fields 8 + index for up to 8 nodes per element, 6-bit fields
// 32 + index for the degree-2 elements).  Lets the assembly gather skip the
// search on every assembly.
template <int NPE>
__global__ __launch_bounds__(256) void k_pair_info(int64_t n_nodes, int lgG, int lg_npe,
                                                   const int32_t* __restrict__ adj_ptr,
                                                   const int32_t* __restrict__ adj,
                                                   const int32_t* __restrict__ conn,
                                                   const int32_t* __restrict__ nb_ptr,
                                                   const int32_t* __restrict__ nb,
                                                   typename PairWord<NPE>::type* __restrict__ pinfo)
{
    typedef typename PairWord<NPE>::type W;
    constexpr int BITS = PairWord<NPE>::bits, FLAG = PairWord<NPE>::flag;
    const int G = 1 << lgG;
    const int64_t gt = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t n = gt >> lgG;
    const int j0 = (int)(gt & (G - 1));
    if (n >= n_nodes) return;
    const int32_t nbs = nb_ptr[n], nbe = nb_ptr[n + 1];
    const int32_t as = adj_ptr[n], ae = adj_ptr[n + 1];
    for (int j = j0; j < nbe - nbs; j += G) {
        const int32_t m = nb[nbs + j];
        W w = 0;
        for (int32_t i = 0; i < ae - as; i++) {
            const int64_t e = adj[as + i] >> lg_npe;
#pragma unroll
            for (int t = 0; t < NPE; t++)
                if (conn[e * NPE + t] == m) w |= (W)(FLAG + t) << (BITS * i);
        }
        pinfo[nbs + j] = w;
    }
}

template <int NPE>
int pattern_kernels(Ctx* c, int64_t nn, int32_t* nb_cnt, int32_t* row_len, int* flags,
                    int64_t* nbtot_out, int64_t* nnz_out)
{
    const unsigned g = grid_for(nn, NB_BLK, 1u << 20);
    k_nb_count<NPE><<<g, NB_BLK, 0, c->stream>>>(nn, c->adj_ptr.p, c->adj.p, c->conn.p,
                                                  c->dof_id.p, c->dpn, c->lg_npe, nb_cnt,
                                                  row_len, flags, flags + 1);
    INS_HIP_CHECK(c, hipGetLastError());
    int64_t nbtot = 0, nnz = 0;
    INS_TRY(ensure(c, c->nb_ptr, nn + 1));
    INS_TRY(exclusive_scan_i32(c, nb_cnt, c->nb_ptr.p, nn, &nbtot));
    INS_TRY(ensure(c, c->row_ptr, c->n_active + 1));
    INS_TRY(exclusive_scan_i32(c, row_len, c->row_ptr.p, c->n_active, &nnz));
    INS_TRY(ensure(c, c->nb, nbtot));
    INS_TRY(ensure(c, c->nboff, nbtot));
    INS_TRY(ensure(c, c->col, nnz));
    k_nb_fill<NPE><<<g, NB_BLK, 0, c->stream>>>(nn, c->adj_ptr.p, c->adj.p, c->conn.p,
                                                 c->dof_id.p, c->dpn, c->lg_npe, c->nb_ptr.p,
                                                 c->nb.p, c->nboff.p, c->row_ptr.p, c->col.p,
                                                 flags);
    INS_HIP_CHECK(c, hipGetLastError());
    *nbtot_out = nbtot;
    *nnz_out = nnz;
    return INS_OK;
}

template <int NPE>
int pair_info(Ctx* c, int64_t nn, int64_t nbtot)
{
    int G = 1;
    while (G < c->max_nb && G < 64) G <<= 1;
    int lgG = 0;
    while ((1 << lgG) < G) lgG++;
    const int64_t threads = nn << lgG;
    const unsigned gp = (unsigned)((threads + 255) / 256);
    typename PairWord<NPE>::type* w;
    if constexpr (NPE <= 8) {
        INS_TRY(ensure(c, c->pinfo, (size_t)nbtot + 1));
        w = c->pinfo.p;
    } else {
        INS_TRY(ensure(c, c->pinfo64, (size_t)nbtot + 1));
        w = c->pinfo64.p;
    }
    k_pair_info<NPE><<<gp, 256, 0, c->stream>>>(nn, lgG, c->lg_npe, c->adj_ptr.p, c->adj.p,
                                                 c->conn.p, c->nb_ptr.p, c->nb.p, w);
    INS_HIP_CHECK(c, hipGetLastError());
    return INS_OK;
}

int build_pattern(Ctx* c)
{
    const int64_t nn = c->n_nodes;
    const int npe = c->npe;
    if (npe != 4 && npe != 8 && npe != 9 && npe != 27)
        return fail(c, INS_ERR_UNSUPPORTED, "element with this number of nodes");
    int maxadj = 0;
    INS_TRY(build_adjacency(c, c->conn.p, c->n_elem, npe, c->lg_npe, nn, c->adj_ptr, c->adj,
                            &maxadj));
    const int maxnb = npe > 9 ? MaxNb<27>::value : MaxNb<8>::value;
    if (maxadj * npe > 4 * maxnb && npe <= 9)
        return fail(c, INS_ERR_UNSUPPORTED, "node belongs to too many elements");

    int32_t* nb_cnt = nullptr;
    int32_t* row_len = nullptr;
    int* flags = nullptr;
    INS_HIP_CHECK(c, hipMallocAsync((void**)&nb_cnt, (nn + 1) * sizeof(int32_t), c->stream));
    INS_HIP_CHECK(c, hipMallocAsync((void**)&row_len, (c->n_active + 1) * sizeof(int32_t), c->stream));
    INS_HIP_CHECK(c, hipMallocAsync((void**)&flags, 2 * sizeof(int), c->stream));
    INS_HIP_CHECK(c, hipMemsetAsync(flags, 0, 2 * sizeof(int), c->stream));
    INS_HIP_CHECK(c, hipMemsetAsync(row_len, 0, (c->n_active + 1) * sizeof(int32_t), c->stream));
    int64_t nbtot = 0, nnz = 0;
    switch (npe) {
    case 4: INS_TRY(pattern_kernels<4>(c, nn, nb_cnt, row_len, flags, &nbtot, &nnz)); break;
    case 8: INS_TRY(pattern_kernels<8>(c, nn, nb_cnt, row_len, flags, &nbtot, &nnz)); break;
    case 9: INS_TRY(pattern_kernels<9>(c, nn, nb_cnt, row_len, flags, &nbtot, &nnz)); break;
    default: INS_TRY(pattern_kernels<27>(c, nn, nb_cnt, row_len, flags, &nbtot, &nnz)); break;
    }
    if (!c->monotone) {
        k_sort_rows<<<grid_for(c->n_active, BLK), BLK, 0, c->stream>>>(c->n_active,
                                                                       c->row_ptr.p, c->col.p);
        INS_HIP_CHECK(c, hipGetLastError());
    }
    int hflags[2];
    INS_HIP_CHECK(c, hipMemcpyAsync(hflags, flags, sizeof(hflags), hipMemcpyDeviceToHost,
                                    c->stream));
    INS_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    (void)hipFreeAsync(nb_cnt, c->stream);
    (void)hipFreeAsync(row_len, c->stream);
    (void)hipFreeAsync(flags, c->stream);
    if (hflags[0] & 2)
        return fail(c, INS_ERR_UNSUPPORTED,
                    "more neighbour nodes than supported (" + std::to_string(maxnb) + ")");
    c->max_nb = hflags[1];
    c->max_adj = maxadj;
    c->nnz = nnz;
    c->have_pinfo = false;
    if (maxadj <= 8) {
        switch (npe) {
        case 4: INS_TRY(pair_info<4>(c, nn, nbtot)); break;
        case 8: INS_TRY(pair_info<8>(c, nn, nbtot)); break;
        case 9: INS_TRY(pair_info<9>(c, nn, nbtot)); break;
        default: INS_TRY(pair_info<27>(c, nn, nbtot)); break;
        }
        c->have_pinfo = true;
    }
    c->have_pattern = true;
    return INS_OK;
}

}  // namespace ins
