// C-ABI entry points (include/insilico_hip.h).  Host-side validation,
// uploads and orchestration; all compute is in the device kernels of
// setup.hip, assemble.hip and solve.hip.  No entry point aborts or throws:
// failures return an INS_ERR_* code and leave a message in the context.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>

#include "common.hpp"

namespace ins {

int fail(Ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    return code;
}

int dalloc_bytes(Ctx* c, void** p, size_t* cap_bytes, size_t bytes)
{
    if (c->budget && (uint64_t)(c->held + (int64_t)bytes) > c->budget)
        return fail(c, INS_ERR_OOM, "device memory budget exceeded");
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return fail(c, INS_ERR_OOM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    *p = q;
    *cap_bytes = bytes;
    c->held += (int64_t)bytes;
    return INS_OK;
}

void dfree_bytes(Ctx* c, void** p, size_t* cap_bytes)
{
    if (*p) {
        hipFree(*p);
        c->held -= (int64_t)*cap_bytes;
    }
    *p = nullptr;
    *cap_bytes = 0;
}

template <typename T>
static int upload(Ctx* c, DArr<T>& a, const T* h, size_t n)
{
    INS_TRY(ensure(c, a, n));
    if (n) INS_HIP_CHECK(c, hipMemcpyAsync(a.p, h, n * sizeof(T), hipMemcpyHostToDevice, c->stream));
    INS_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return INS_OK;
}

template <typename T>
static int download(Ctx* c, T* h, const DArr<T>& a, size_t n)
{
    if (n) INS_HIP_CHECK(c, hipMemcpyAsync(h, a.p, n * sizeof(T), hipMemcpyDeviceToHost, c->stream));
    INS_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return INS_OK;
}

static void release_all(Ctx* c)
{
    release(c, c->xyz); release(c, c->conn); release(c, c->gconn); release(c, c->pinfo64); release(c, c->node_order); release(c, c->fvert); release(c, c->mhist); release(c, c->stage_me); for (auto& hh : c->bdf_hist) release(c, hh); release(c, c->dof_id); release(c, c->rowmap);
    release(c, c->g); release(c, c->u); release(c, c->adj_ptr); release(c, c->adj);
    release(c, c->nb_ptr); release(c, c->nb); release(c, c->nboff); release(c, c->row_ptr);
    release(c, c->col); release(c, c->coef); release(c, c->tab64); release(c, c->tab32);
    release(c, c->faces); release(c, c->fadj_ptr); release(c, c->fadj); release(c, c->val64);
    release(c, c->val32); release(c, c->b); release(c, c->dinv64); release(c, c->dinv32);
    release(c, c->stage_ke); release(c, c->stage_fe); release(c, c->stage_ff); release(c, c->stage_pt);
    release(c, c->sell.slice_ptr); release(c, c->sell.slice_pat);
    release(c, c->sell.pat_ptr); release(c, c->sell.pat_delta); release(c, c->sell.pat_len);
    release(c, c->sell.v64); release(c, c->sell.v32); release(c, c->sell.map); release(c, c->sell.perm);
    release(c, c->x); release(c, c->r); release(c, c->p); release(c, c->q); release(c, c->t64);
    release(c, c->x32); release(c, c->r32); release(c, c->p32); release(c, c->q32);
    release(c, c->partial); release(c, c->scal); release(c, c->dflag);
    release(c, c->xp); release(c, c->bp); release(c, c->d64p); release(c, c->d32p);
    release(c, c->pinfo);
    release(c, c->bftab); release(c, c->ttab);
    release(c, c->sell.vs64); release(c, c->sell.spw); release(c, c->sell.anode);
    release(c, c->sell.vs); release(c, c->sell.su); release(c, c->sell.smap);
    release(c, c->emask);
}

// A context detached from a server may still be named by the batched work
// queued on the server stream before the detach: wait for that work before
// the context's buffers are reused, freed or read
int wait_server_release(Ctx* c)
{
    if (!c->srv_pending) return INS_OK;
    c->srv_pending = false;
    INS_HIP_CHECK(c, hipEventSynchronize(c->srv_ev));
    return INS_OK;
}

static void invalidate_pattern(Ctx* c)
{
    c->have_pattern = false;
    c->have_matrix = false;
    c->have_rhs = false;
    c->have_solution = false;
    c->sell.built = false;
    c->sell.direct = false;
    c->csr_stale = false;
    c->have_emask = false;
}

}  // namespace ins

using namespace ins;

#define CTX_CHECK(ctx)                                            \
    do {                                                          \
        if (!(ctx)) return INS_ERR_ARG;                           \
        hipError_t de_ = hipSetDevice((ctx)->device);             \
        if (de_ != hipSuccess)                                    \
            return fail((ctx), INS_ERR_HIP, "hipSetDevice failed"); \
        INS_TRY(::ins::wait_server_release(ctx));                 \
    } while (0)

extern "C" {

int ins_create(ins_ctx** out, int device, uint64_t hbm_budget_bytes)
{
    if (!out) return INS_ERR_ARG;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return INS_ERR_HIP;
    if (device < 0 || device >= ndev) return INS_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return INS_ERR_HIP;
    Ctx* c = new (std::nothrow) Ctx();
    if (!c) return INS_ERR_OOM;
    c->device = device;
    {
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) ==
                hipSuccess && cus > 0)
            c->n_cu = cus;
    }
    c->budget = hbm_budget_bytes;
    // temporaries of the setup steps are stream-ordered allocations on the
    // context stream (no device-wide synchronisation); keep their memory in
    // the pool between calls
    hipMemPool_t pool = nullptr;
    if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess && pool) {
        uint64_t keep = UINT64_MAX;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        hipEventCreate(&c->ev_el[0]) != hipSuccess || hipEventCreate(&c->ev_el[1]) != hipSuccess) {
        delete c;
        return INS_ERR_HIP;
    }
    *out = reinterpret_cast<ins_ctx*>(c);
    return INS_OK;
}

int ins_destroy(ins_ctx* h)
{
    Ctx* c = reinterpret_cast<Ctx*>(h);
    if (!c) return INS_ERR_ARG;
    hipSetDevice(c->device);
    (void)wait_server_release(c);
    hipStreamSynchronize(c->stream);
    release_all(c);
    if (c->srv_ev) hipEventDestroy(c->srv_ev);
    if (c->h_scal) hipHostFree(c->h_scal);
    if (c->h_snap) hipHostFree(c->h_snap);
    for (int i = 0; i < 2; i++)
        if (c->evs[i]) hipEventDestroy(c->evs[i]);
    hipEventDestroy(c->ev0);
    hipEventDestroy(c->ev1);
    if (c->ev_el[0]) hipEventDestroy(c->ev_el[0]);
    if (c->ev_el[1]) hipEventDestroy(c->ev_el[1]);
    for (hipEvent_t e : c->prof_ev) hipEventDestroy(e);
    hipStreamDestroy(c->stream);
    delete c;
    return INS_OK;
}

const char* ins_strerror(int s)
{
    switch (s) {
    case INS_OK: return "ok";
    case INS_ERR_ARG: return "invalid argument";
    case INS_ERR_HIP: return "HIP runtime error";
    case INS_ERR_OOM: return "out of device memory or over budget";
    case INS_ERR_RANGE: return "index out of range";
    case INS_ERR_PATTERN: return "entry outside the sparsity pattern";
    case INS_ERR_BREAKDOWN: return "CG breakdown (matrix not SPD?)";
    case INS_ERR_NOCONV: return "no convergence within the iteration limit";
    case INS_ERR_STATE: return "call out of order";
    case INS_ERR_UNSUPPORTED: return "unsupported configuration";
    default: return "unknown status";
    }
}

const char* ins_last_error(const ins_ctx* h)
{
    const Ctx* c = reinterpret_cast<const Ctx*>(h);
    return c ? c->err.c_str() : "null context";
}

int ins_device_memory(int device, uint64_t* free_bytes, uint64_t* total_bytes)
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return INS_ERR_HIP;
    if (device < 0 || device >= ndev || !free_bytes || !total_bytes) return INS_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return INS_ERR_HIP;
    size_t f = 0, t = 0;
    if (hipMemGetInfo(&f, &t) != hipSuccess) return INS_ERR_HIP;
    *free_bytes = f;
    *total_bytes = t;
    return INS_OK;
}

int ins_get_stats(const ins_ctx* h, ins_stats* out)
{
    const Ctx* c = reinterpret_cast<const Ctx*>(h);
    if (!c || !out) return INS_ERR_ARG;
    *out = c->stats;
    out->n_active = c->n_active;
    out->nnz = c->nnz;
    out->device_bytes = c->held;
    out->sell_values = (c->sell.built && c->sell.ok) ? c->sell.nvals : 0;
    out->sell_patterns = (c->sell.built && c->sell.ok) ? c->sell.npat : 0;
    const bool sym = c->sell.built && c->sell.ok && c->sell.sym_on;
    out->sell_sym_offsets = sym ? c->sell.nu : 0;
    out->sym_values = sym ? c->sell.nvs : 0;
    out->sym_block = sym ? c->sell.sdpn : 0;
    out->sym_lines = sym && c->sell.slines ? 1 : 0;
    return INS_OK;
}

int ins_set_profiling(ins_ctx* h, int on)
{
    Ctx* c = reinterpret_cast<Ctx*>(h);
    if (!c) return INS_ERR_ARG;
    c->profiling = on != 0;
    return INS_OK;
}

int ins_sync(ins_ctx* h)
{
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_CHECK(c);
    INS_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return INS_OK;
}

int ins_set_mesh(ins_ctx* h, int dim, int64_t n_nodes, const double* xyz, int64_t n_elem,
                 int npe, const int32_t* conn)
{
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_CHECK(c);
    c->bdf_staged = false;
    if ((dim != 2 && dim != 3) || n_nodes <= 0 || n_elem <= 0 || !xyz || !conn)
        return fail(c, INS_ERR_ARG, "ins_set_mesh: bad arguments");
    if (npe != (1 << dim))
        return fail(c, INS_ERR_UNSUPPORTED, "only Q1 QUAD4/HEX8 elements are provided");
    if (n_nodes > INT32_MAX || n_elem * npe > INT32_MAX)
        return fail(c, INS_ERR_UNSUPPORTED, "mesh exceeds int32 indexing");
    for (int64_t i = 0; i < n_elem * npe; i++)
        if (conn[i] < 0 || conn[i] >= n_nodes)
            return fail(c, INS_ERR_RANGE, "connectivity refers to a node out of range");
    c->dim = dim;
    c->npe = npe;
    c->lg_npe = (npe == 8) ? 3 : 2;
    c->n_nodes = n_nodes;
    c->n_elem = n_elem;
    c->deg = 1;
    c->gnpe = npe;
    c->n_gnodes = n_nodes;
    c->have_order = false;
    INS_TRY(upload(c, c->xyz, xyz, (size_t)n_nodes * dim));
    INS_TRY(upload(c, c->conn, conn, (size_t)n_elem * npe));
    invalidate_pattern(c);
    c->have_dofs = false;
    c->n_faces = 0;
    return INS_OK;
}

// Solver storage order of the degree-2 field nodes: each node takes the
// class of its first appearance in an element (vertex; edge along local axis
// 0, 1, 2; face normal to local axis 2, 1, 0; cell) and the support point
// there (the mean of the element vertices spanning its n-face); nodes are
// ordered by class, then by (z, y, x) quantised to 1e-7 of the bounding-box
// diagonal.  On a structured mesh the rows of one class and line then have
// equal column-offset lists, which the SELL slices share.
static int order_field_nodes(Ctx* c, const int32_t* fconn)
{
    const int dim = c->dim, gn = c->gnpe, nf = c->npe;
    std::vector<double> X((size_t)c->n_gnodes * dim);
    std::vector<int32_t> G((size_t)c->n_elem * gn);
    INS_HIP_CHECK(c, hipMemcpyAsync(X.data(), c->xyz.p, X.size() * sizeof(double),
                                    hipMemcpyDeviceToHost, c->stream));
    INS_HIP_CHECK(c, hipMemcpyAsync(G.data(), c->gconn.p, G.size() * sizeof(int32_t),
                                    hipMemcpyDeviceToHost, c->stream));
    INS_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    // n-faces of the hierarchic element and their class
    static const int E2[4][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}};
    static const int E3[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                                  {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};
    static const int F3[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 1, 5, 4},
                                 {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}};
    static const int EC3[12] = {1, 2, 1, 2, 1, 2, 1, 2, 3, 3, 3, 3};
    static const int FC3[6] = {4, 4, 5, 6, 5, 6};
    const int64_t nn = c->n_nodes;
    std::vector<int8_t> cls((size_t)nn, -1);
    std::vector<int32_t> fvert((size_t)nn, -1);
    std::vector<double> pt((size_t)nn * 3, 0.);
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (size_t i = 0; i < (size_t)c->n_gnodes; i++)
        for (int d = 0; d < dim; d++) {
            lo[d] = std::min(lo[d], X[i * dim + d]);
            hi[d] = std::max(hi[d], X[i * dim + d]);
        }
    for (int64_t e = 0; e < c->n_elem; e++) {
        const int32_t* ge = &G[(size_t)e * gn];
        for (int k = 0; k < nf; k++) {
            const int32_t node = fconn[e * nf + k];
            if (k < gn) fvert[node] = ge[k];
            if (cls[node] >= 0) continue;
            int vs[8], nv = 0, cl = 0;
            if (k < gn) {
                vs[nv++] = k;
            } else if (dim == 2) {
                if (k < 8) {
                    vs[nv++] = E2[k - 4][0];
                    vs[nv++] = E2[k - 4][1];
                    cl = 1 + (k - 4) % 2;
                } else {
                    for (int v = 0; v < 4; v++) vs[nv++] = v;
                    cl = 3;
                }
            } else if (k < 20) {
                vs[nv++] = E3[k - 8][0];
                vs[nv++] = E3[k - 8][1];
                cl = EC3[k - 8];
            } else if (k < 26) {
                for (int v = 0; v < 4; v++) vs[nv++] = F3[k - 20][v];
                cl = FC3[k - 20];
            } else {
                for (int v = 0; v < 8; v++) vs[nv++] = v;
                cl = 7;
            }
            cls[node] = (int8_t)cl;
            for (int d = 0; d < dim; d++) {
                double sum = 0.;
                for (int v = 0; v < nv; v++) sum += X[(size_t)ge[vs[v]] * dim + d];
                pt[(size_t)node * 3 + d] = sum / nv;
            }
        }
    }
    double diag = 0.;
    for (int d = 0; d < dim; d++) diag += (hi[d] - lo[d]) * (hi[d] - lo[d]);
    const double tol = 1e-7 * (diag > 0. ? std::sqrt(diag) : 1.);
    struct Key {
        int64_t k[4];
        int32_t node;
    };
    std::vector<Key> keys((size_t)nn);
    for (int64_t i = 0; i < nn; i++) {
        Key& K = keys[(size_t)i];
        K.node = (int32_t)i;
        K.k[0] = cls[i] >= 0 ? cls[i] : 8;
        for (int d = 0; d < 3; d++) {
            const int dd = 2 - d;  // z, y, x
            K.k[1 + d] = dd < dim ? (int64_t)std::llround((pt[(size_t)i * 3 + dd] - lo[dd]) / tol) : 0;
        }
    }
    std::sort(keys.begin(), keys.end(), [](const Key& a, const Key& b) {
        for (int j = 0; j < 4; j++)
            if (a.k[j] != b.k[j]) return a.k[j] < b.k[j];
        return a.node < b.node;
    });
    std::vector<int32_t> order((size_t)nn);
    for (int64_t i = 0; i < nn; i++) order[(size_t)i] = keys[(size_t)i].node;
    INS_TRY(upload(c, c->node_order, order.data(), order.size()));
    INS_TRY(upload(c, c->fvert, fvert.data(), fvert.size()));
    c->fvert_host.swap(fvert);
    c->have_order = true;
    return INS_OK;
}

int ins_set_field(ins_ctx* h, int degree, int64_t n_field_nodes, int field_npe,
                  const int32_t* field_conn)
{
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_CHECK(c);
    c->bdf_staged = false;
    if (c->n_elem == 0) return fail(c, INS_ERR_STATE, "ins_set_field before ins_set_mesh");
    if (degree != 2) return fail(c, INS_ERR_UNSUPPORTED, "only fields of degree 2 on a Q1 mesh");
    const int want = (c->dim == 2) ? 9 : 27;
    if (field_npe != want || n_field_nodes <= 0 || !field_conn)
        return fail(c, INS_ERR_ARG, "ins_set_field: field_npe must be 9 (2-D) or 27 (3-D)");
    if (c->deg == 2) return fail(c, INS_ERR_STATE, "ins_set_field called twice for one mesh");
    if (n_field_nodes > INT32_MAX || c->n_elem * field_npe > INT32_MAX ||
        c->n_elem > (INT32_MAX >> 5))
        return fail(c, INS_ERR_UNSUPPORTED, "field exceeds int32 indexing");
    for (int64_t i = 0; i < c->n_elem * field_npe; i++)
        if (field_conn[i] < 0 || field_conn[i] >= n_field_nodes)
            return fail(c, INS_ERR_RANGE, "field connectivity refers to a node out of range");
    c->n_faces = 0;  // Neumann faces are given in field nodes: set them after the field
    // the Q1 connectivity becomes the geometry
    std::swap(c->gconn, c->conn);
    c->gnpe = c->npe;
    c->n_gnodes = c->n_nodes;
    c->deg = 2;
    c->npe = field_npe;
    c->lg_npe = ceil_lg2(field_npe);
    c->n_nodes = n_field_nodes;
    INS_TRY(upload(c, c->conn, field_conn, (size_t)c->n_elem * field_npe));
    INS_TRY(order_field_nodes(c, field_conn));
    invalidate_pattern(c);
    c->have_dofs = false;
    return INS_OK;
}

int ins_set_dofs(ins_ctx* h, int dpn, const int32_t* dof_id, int64_t n_active)
{
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_CHECK(c);
    c->bdf_staged = false;
    if (c->n_nodes == 0) return fail(c, INS_ERR_STATE, "ins_set_dofs before ins_set_mesh");
    if (dpn < 1 || dpn > 3 || !dof_id || n_active < 0)
        return fail(c, INS_ERR_ARG, "ins_set_dofs: bad arguments");
    const int64_t nd = c->n_nodes * dpn;
    // ids must be a permutation of 0..n_active-1 (base/dof/numbering.hpp:45-66)
    std::vector<uint8_t> seen((size_t)n_active, 0);
    bool mono = true;
    int32_t last = -1;
    for (int64_t k = 0; k < nd; k++) {
        const int32_t id = dof_id[k];
        if (id < -1 || id >= n_active)
            return fail(c, INS_ERR_RANGE, "DoF id out of range: " + std::to_string(id));
        if (id < 0) continue;
        if (seen[id]) return fail(c, INS_ERR_RANGE, "DoF id used twice: " + std::to_string(id));
        seen[id] = 1;
        if (id < last) mono = false;
        last = id;
    }
    for (int64_t i = 0; i < n_active; i++)
        if (!seen[i]) return fail(c, INS_ERR_RANGE, "DoF id missing: " + std::to_string(i));
    c->dpn = dpn;
    c->n_active = n_active;
    c->dof_n_active = n_active;
    c->csr_from_host = false;
    c->monotone = mono;
    INS_TRY(upload(c, c->dof_id, dof_id, (size_t)nd));
    INS_TRY(ensure(c, c->g, (size_t)nd));
    INS_TRY(ensure(c, c->u, (size_t)nd));
    INS_HIP_CHECK(c, hipMemsetAsync(c->g.p, 0, nd * sizeof(double), c->stream));
    INS_HIP_CHECK(c, hipMemsetAsync(c->u.p, 0, nd * sizeof(double), c->stream));
    INS_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    c->u_zero = true;
    invalidate_pattern(c);
    c->have_dofs = true;
    return INS_OK;
}

int ins_set_prescribed(ins_ctx* h, const double* g)
{
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_CHECK(c);
    if (!c->have_dofs) return fail(c, INS_ERR_STATE, "ins_set_prescribed before ins_set_dofs");
    const int64_t nd = c->n_nodes * c->dpn;
    if (!g) {
        INS_HIP_CHECK(c, hipMemsetAsync(c->g.p, 0, nd * sizeof(double), c->stream));
        INS_HIP_CHECK(c, hipStreamSynchronize(c->stream));
        return INS_OK;
    }
    return upload(c, c->g, g, (size_t)nd);
}

int ins_scale_prescribed(ins_ctx* h, double factor)
{
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_CHECK(c);
    if (!c->have_dofs) return fail(c, INS_ERR_STATE, "ins_scale_prescribed before ins_set_dofs");
    const int64_t nd = c->n_nodes * c->dpn;
    std::vector<double> g((size_t)nd);
    INS_TRY(download(c, g.data(), c->g, (size_t)nd));
    for (auto& v : g) v *= factor;
    return upload(c, c->g, g.data(), (size_t)nd);
}

int ins_set_dof_values(ins_ctx* h, const double* u)
{
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_CHECK(c);
    if (!c->have_dofs) return fail(c, INS_ERR_STATE, "ins_set_dof_values before ins_set_dofs");
    const int64_t nd = c->n_nodes * c->dpn;
    if (!u) {
        INS_HIP_CHECK(c, hipMemsetAsync(c->u.p, 0, nd * sizeof(double), c->stream));
        INS_HIP_CHECK(c, hipStreamSynchronize(c->stream));
        c->u_zero = true;
        return INS_OK;
    }
    c->u_zero = false;
    return upload(c, c->u, u, (size_t)nd);
}

int ins_get_dof_values(ins_ctx* h, double* u)
{
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_CHECK(c);
    if (!c->have_dofs || !u) return fail(c, INS_ERR_STATE, "no DoF values");
    return download(c, u, c->u, (size_t)(c->n_nodes * c->dpn));
}

int ins_build_pattern(ins_ctx* h, int64_t* nnz_out)
{
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_CHECK(c);
    if (!c->have_dofs) return fail(c, INS_ERR_STATE, "ins_build_pattern before ins_set_dofs");
    // a CSR installed by ins_set_csr may have had another size: the pattern
    // is rebuilt for the DoFs
    c->n_active = c->dof_n_active;
    invalidate_pattern(c);
    INS_HIP_CHECK(c, hipEventRecord(c->ev0, c->stream));
    INS_TRY(build_pattern(c));
    c->csr_from_host = false;
    INS_HIP_CHECK(c, hipEventRecord(c->ev1, c->stream));
    INS_HIP_CHECK(c, hipEventSynchronize(c->ev1));
    float ms = 0.f;
    INS_HIP_CHECK(c, hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->stats.setup_ms = ms;
    c->have_matrix = false;
    c->sell.built = false;
    if (nnz_out) *nnz_out = c->nnz;
    return INS_OK;
}

int ins_get_pattern(ins_ctx* h, int32_t* row_ptr, int32_t* col_idx)
{
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_CHECK(c);
    if (!c->have_pattern) return fail(c, INS_ERR_STATE, "no pattern");
    if (row_ptr) INS_TRY(download(c, row_ptr, c->row_ptr, (size_t)c->n_active + 1));
    if (col_idx) INS_TRY(download(c, col_idx, c->col, (size_t)c->nnz));
    return INS_OK;
}

int ins_set_material(ins_ctx* h, int kernel, int nq, int ncoef, const double* coef)
{
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_CHECK(c);
    c->bdf_staged = false;
    if (c->n_elem == 0) return fail(c, INS_ERR_STATE, "ins_set_material before ins_set_mesh");
    if (kernel != INS_LAPLACE && kernel != INS_STVENANT && kernel != INS_NEOHOOKE)
        return fail(c, INS_ERR_UNSUPPORTED, "unknown kernel");
    if (nq < 1 || nq > 3) return fail(c, INS_ERR_UNSUPPORTED, "1..3 Gauss points per direction");
    const int need = (kernel == INS_LAPLACE) ? 1 : 2;
    if (ncoef < need || !coef) return fail(c, INS_ERR_ARG, "too few element coefficients");
    int nqp = 1;
    for (int d = 0; d < c->dim; d++) nqp *= nq;
    if (kernel == INS_LAPLACE && ncoef != 1 && ncoef != nqp)
        return fail(c, INS_ERR_ARG,
                    "heat::Laplace takes 1 conductivity per element or 1 per quadrature point");
    c->kernel = kernel;
    c->nq = nq;
    c->ncoef = ncoef;
    INS_TRY(upload(c, c->coef, coef, (size_t)c->n_elem * ncoef));
    INS_TRY(make_tables(c));
    c->have_matrix = false;
    return INS_OK;
}

int ins_set_mass(ins_ctx* h, double density, double mass_factor, const double* hist)
{
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_CHECK(c);
    c->bdf_staged = false;
    if (!(density >= 0.) || !std::isfinite(mass_factor))
        return fail(c, INS_ERR_ARG, "ins_set_mass: density >= 0 and a finite factor");
    if (density > 0. && !c->have_dofs)
        return fail(c, INS_ERR_STATE, "ins_set_mass before ins_set_dofs");
    c->density = density;
    c->mass_factor = density > 0. ? mass_factor : 0.;
    c->have_mhist = false;
    if (density > 0. && hist) {
        INS_TRY(upload(c, c->mhist, hist, (size_t)(c->n_nodes * c->dpn)));
        c->have_mhist = true;
    }
    c->have_matrix = false;
    return INS_OK;
}

int ins_set_body_force(ins_ctx* h, int kind, const double* params)
{
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_CHECK(c);
    c->bdf_staged = false;
    if (kind != INS_BF_NONE && kind != INS_BF_CONST && kind != INS_BF_SINPROD &&
        kind != INS_BF_COSPROD && kind != INS_BF_TABLE)
        return fail(c, INS_ERR_ARG, "unknown body force kind");
    if (kind == INS_BF_TABLE) {
        if (!c->have_bftab) return fail(c, INS_ERR_STATE, "no body force values set");
        c->bf_kind = kind;
        return INS_OK;
    }
    if (kind != INS_BF_NONE && !params) return fail(c, INS_ERR_ARG, "body force parameters");
    c->bf_kind = kind;
    for (int i = 0; i < 4; i++) c->bf[i] = (kind != INS_BF_NONE) ? params[i] : 0.;
    return INS_OK;
}

int ins_quadrature_rule(int dim, int npts, double* w, double* xi)
{
    if (dim < 1 || dim > 3 || npts < 1 || npts > 3 || !w || !xi) return INS_ERR_ARG;
    const int nq = quadrature_rule(dim, npts, w, xi);
    return nq > 0 ? nq : INS_ERR_ARG;
}

int ins_set_body_force_values(ins_ctx* h, const double* f)
{
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_CHECK(c);
    c->bdf_staged = false;
    if (!f) {
        c->have_bftab = false;
        if (c->bf_kind == INS_BF_TABLE) c->bf_kind = INS_BF_NONE;
        return INS_OK;
    }
    if (c->kernel == 0 || !c->have_dofs)
        return fail(c, INS_ERR_STATE, "ins_set_body_force_values before ins_set_material / ins_set_dofs");
    int64_t nqp = 1;
    for (int d = 0; d < c->dim; d++) nqp *= c->nq;
    INS_TRY(upload(c, c->bftab, f, (size_t)(c->n_elem * nqp * c->dpn)));
    c->have_bftab = true;
    c->bf_kind = INS_BF_TABLE;
    return INS_OK;
}

int ins_set_traction_values(ins_ctx* h, const double* t)
{
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_CHECK(c);
    c->bdf_staged = false;
    if (!t) {
        c->have_ttab = false;
        return INS_OK;
    }
    if (c->n_faces <= 0 || c->kernel == 0 || !c->have_dofs)
        return fail(c, INS_ERR_STATE,
                    "ins_set_traction_values before ins_set_traction / ins_set_material");
    int64_t nqf = 1;
    for (int d = 0; d + 1 < c->dim; d++) nqf *= c->nq;
    INS_TRY(upload(c, c->ttab, t, (size_t)(c->n_faces * nqf * c->dpn)));
    c->have_ttab = true;
    return INS_OK;
}

int ins_set_traction(ins_ctx* h, int64_t n_faces, int nfn, const int32_t* faces,
                     const double* traction)
{
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_CHECK(c);
    c->bdf_staged = false;
    c->have_ttab = false;   // values of the faces set before
    if (c->n_nodes == 0) return fail(c, INS_ERR_STATE, "ins_set_traction before ins_set_mesh");
    if (n_faces == 0) {
        c->n_faces = 0;
        return INS_OK;
    }
    const int want = c->deg == 2 ? (c->dim == 2 ? 3 : 9) : (1 << (c->dim - 1));
    if (n_faces < 0 || !faces || !traction || nfn != want)
        return fail(c, INS_ERR_ARG, "ins_set_traction: bad arguments");
    for (int64_t i = 0; i < n_faces * nfn; i++)
        if (faces[i] < 0 || faces[i] >= c->n_nodes)
            return fail(c, INS_ERR_RANGE, "face refers to a node out of range");
    if (c->deg == 2) {
        // the first 2 (4) field nodes of a face are its vertices
        const int gnf = 1 << (c->dim - 1);
        for (int64_t f = 0; f < n_faces; f++)
            for (int k = 0; k < gnf; k++)
                if (c->fvert_host[(size_t)faces[f * nfn + k]] < 0)
                    return fail(c, INS_ERR_ARG, "a face must list its vertices first");
    }
    c->n_faces = n_faces;
    c->nfn = nfn;
    for (int i = 0; i < 3; i++) c->traction[i] = i < c->dim ? traction[i] : 0.;
    INS_TRY(upload(c, c->faces, faces, (size_t)n_faces * nfn));
    INS_TRY(build_adjacency(c, c->faces.p, n_faces, nfn, ceil_lg2(nfn), c->n_nodes,
                            c->fadj_ptr, c->fadj, nullptr));
    return INS_OK;
}

int ins_set_traction_profile(ins_ctx* h, int axis, double x0)
{
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_CHECK(c);
    c->bdf_staged = false;
    if (axis < -1 || axis >= 3 || (c->dim && axis >= c->dim))
        return fail(c, INS_ERR_ARG, "ins_set_traction_profile: axis -1..dim-1");
    c->tr_axis = axis;
    c->tr_x0 = axis >= 0 ? x0 : 0.;
    return INS_OK;
}

int ins_set_neumann_flux(ins_ctx* h, int ref_kind, const double* ref_params)
{
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_CHECK(c);
    c->bdf_staged = false;
    if (ref_kind != INS_REF_ZERO && ref_kind != INS_REF_SINPROD && ref_kind != INS_REF_COSPROD)
        return fail(c, INS_ERR_ARG, "ins_set_neumann_flux: INS_REF_ZERO, _SINPROD or _COSPROD");
    if (ref_kind != INS_REF_ZERO && !ref_params)
        return fail(c, INS_ERR_ARG, "ins_set_neumann_flux: params");
    c->flux_kind = ref_kind;
    for (int i = 0; i < 8; i++) c->flux_p[i] = ref_kind != INS_REF_ZERO ? ref_params[i] : 0.;
    return INS_OK;
}

int ins_assemble(ins_ctx* h, int precision, int flags)
{
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_CHECK(c);
    c->bdf_staged = false;
    if (precision != INS_FP64 && precision != INS_FP32)
        return fail(c, INS_ERR_ARG, "precision must be INS_FP64 or INS_FP32");
    if (!(flags & (INS_ASM_MATRIX | INS_ASM_RHS)))
        return fail(c, INS_ERR_ARG, "nothing to assemble");
    if (c->kernel == INS_LAPLACE && c->dpn > 1 && c->bf_kind != INS_BF_NONE && (flags & INS_ASM_RHS))
        return fail(c, INS_ERR_UNSUPPORTED, "vector Laplace with a body force");
    return assemble(c, precision, flags);
}

int ins_get_csr_values(ins_ctx* h, double* val)
{
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_CHECK(c);
    if (!c->have_matrix || !val) return fail(c, INS_ERR_STATE, "no assembled matrix");
    INS_TRY(ensure_csr(c));
    if (c->matrix_precision == INS_FP64) return download(c, val, c->val64, (size_t)c->nnz);
    std::vector<float> tmp((size_t)c->nnz);
    INS_TRY(download(c, tmp.data(), c->val32, (size_t)c->nnz));
    for (int64_t i = 0; i < c->nnz; i++) val[i] = tmp[i];
    return INS_OK;
}

int ins_get_rhs(ins_ctx* h, double* b)
{
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_CHECK(c);
    if (!c->have_rhs || !b) return fail(c, INS_ERR_STATE, "no assembled rhs");
    return download(c, b, c->b, (size_t)c->n_active);
}

int ins_set_rhs(ins_ctx* h, const double* b)
{
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_CHECK(c);
    if (c->n_active <= 0 || !b) return fail(c, INS_ERR_STATE, "ins_set_rhs before ins_set_dofs");
    INS_TRY(upload(c, c->b, b, (size_t)c->n_active));
    c->have_rhs = true;
    c->have_solution = false;
    return INS_OK;
}

int ins_add_rhs(ins_ctx* h, const double* b)
{
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_CHECK(c);
    if (!c->have_rhs || !b) return fail(c, INS_ERR_STATE, "ins_add_rhs before the rhs exists");
    std::vector<double> cur((size_t)c->n_active);
    INS_TRY(download(c, cur.data(), c->b, (size_t)c->n_active));
    for (int64_t i = 0; i < c->n_active; i++) cur[i] += b[i];
    return upload(c, c->b, cur.data(), (size_t)c->n_active);
}

int ins_rhs_norm(ins_ctx* h, double* norm)
{
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_CHECK(c);
    if (!c->have_rhs || !norm) return fail(c, INS_ERR_STATE, "no rhs");
    double v = 0.;
    INS_TRY(vec_norm2(c, c->b.p, c->n_active, &v));
    *norm = c->n_active ? v / (double)c->n_active : 0.;
    return INS_OK;
}

int ins_set_csr(ins_ctx* h, int64_t n, const int32_t* row_ptr, const int32_t* col_idx,
                const double* val)
{
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_CHECK(c);
    c->bdf_staged = false;
    if (n <= 0 || !row_ptr || !col_idx || !val) return fail(c, INS_ERR_ARG, "ins_set_csr: bad arguments");
    if (row_ptr[0] != 0) return fail(c, INS_ERR_ARG, "row_ptr[0] must be 0");
    for (int64_t i = 0; i < n; i++)
        if (row_ptr[i + 1] < row_ptr[i]) return fail(c, INS_ERR_ARG, "row_ptr not ascending");
    const int64_t nnz = row_ptr[n];
    for (int64_t i = 0; i < n; i++)
        for (int32_t k = row_ptr[i]; k < row_ptr[i + 1]; k++) {
            if (col_idx[k] < 0 || col_idx[k] >= n)
                return fail(c, INS_ERR_RANGE, "column index out of bound");
            if (k > row_ptr[i] && col_idx[k] <= col_idx[k - 1])
                return fail(c, INS_ERR_ARG, "columns must be strictly ascending per row");
        }
    // the element assembly, the RHS kernels and ins_distribute work on the
    // DoF maps of ins_set_dofs; they refuse (INS_ERR_STATE) until
    // ins_build_pattern rebuilds the pattern, or, for ins_distribute, unless
    // the uploaded system has as many rows as there are active DoFs.  The
    // right-hand side must be set again for the uploaded system.
    invalidate_pattern(c);
    c->n_active = n;
    c->nnz = nnz;
    INS_TRY(upload(c, c->row_ptr, row_ptr, (size_t)n + 1));
    INS_TRY(upload(c, c->col, col_idx, (size_t)nnz));
    INS_TRY(upload(c, c->val64, val, (size_t)nnz));
    std::vector<float> v32(val, val + nnz);
    INS_TRY(upload(c, c->val32, v32.data(), (size_t)nnz));
    c->matrix_precision = INS_FP64;
    c->have_pattern = true;
    c->have_matrix = true;
    c->have_solution = false;
    c->csr_from_host = true;
    c->csr_stale = false;
    c->sell.built = false;
    c->sell.stale = true;
    INS_TRY(compute_dinv(c));
    INS_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    return INS_OK;
}

int ins_solve(ins_ctx* h, int mode, int format, double rtol, int max_it, int* iters,
              double* relres)
{
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_CHECK(c);
    if (mode != INS_SOLVE_FP64 && mode != INS_SOLVE_MIXED) return fail(c, INS_ERR_ARG, "solve mode");
    if (format != INS_FMT_CSR && format != INS_FMT_SELL) return fail(c, INS_ERR_ARG, "format");
    if (!(rtol > 0.) || max_it < 0) return fail(c, INS_ERR_ARG, "rtol/max_it");
    return solve(c, mode, format, rtol, max_it, iters, relres);
}

int ins_solve_direct(ins_ctx* h, int kind, double* relres)
{
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_CHECK(c);
    if (kind != INS_DIRECT_CHOLESKY && kind != INS_DIRECT_LU)
        return fail(c, INS_ERR_ARG, "ins_solve_direct: kind");
    return solve_direct(c, kind, relres);
}

int ins_prepare(ins_ctx* h, int mode, int format)
{
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_CHECK(c);
    if (mode != INS_SOLVE_FP64 && mode != INS_SOLVE_MIXED) return fail(c, INS_ERR_ARG, "solve mode");
    if (format != INS_FMT_CSR && format != INS_FMT_SELL) return fail(c, INS_ERR_ARG, "format");
    return prepare(c, mode, format);
}

int ins_get_solution(ins_ctx* h, double* x)
{
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_CHECK(c);
    if (!c->have_solution || !x) return fail(c, INS_ERR_STATE, "no solution");
    return download(c, x, c->x, (size_t)c->n_active);
}

int ins_solution_norm(ins_ctx* h, double* norm)
{
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_CHECK(c);
    if (!c->have_solution || !norm) return fail(c, INS_ERR_STATE, "no solution");
    double v = 0.;
    INS_TRY(vec_norm2(c, c->x.p, c->n_active, &v));
    *norm = c->n_active ? v / (double)c->n_active : 0.;
    return INS_OK;
}

int ins_distribute(ins_ctx* h, int op)
{
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_CHECK(c);
    if (op != INS_DIST_SET && op != INS_DIST_ADD) return fail(c, INS_ERR_ARG, "distribute op");
    return distribute(c, op);
}

// ---- serving ------------------------------------------------------------------
int ins_server_create(ins_server** out, int device, int max_slots)
{
    if (!out) return INS_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return INS_ERR_HIP;
    if (device < 0 || device >= ndev) return INS_ERR_ARG;
    Server* sv = nullptr;
    const int s = server_create(&sv, device, max_slots);
    *out = reinterpret_cast<ins_server*>(sv);
    return s;
}

int ins_server_destroy(ins_server* h)
{
    if (!h) return INS_ERR_ARG;
    server_destroy(reinterpret_cast<Server*>(h));
    return INS_OK;
}

const char* ins_server_last_error(const ins_server* h)
{
    return server_error(reinterpret_cast<Server*>(const_cast<ins_server*>(h)));
}

int ins_server_attach(ins_server* h, ins_ctx* ctx, double rtol, int max_it, int* slot)
{
    Ctx* c = reinterpret_cast<Ctx*>(ctx);
    if (!h || !c) return INS_ERR_ARG;
    CTX_CHECK(c);
    return server_attach(reinterpret_cast<Server*>(h), c, rtol, max_it, slot);
}

int ins_server_step(ins_server* h, int k, int* n_finished)
{
    return server_step(reinterpret_cast<Server*>(h), k, n_finished);
}

int ins_server_poll(ins_server* h, int slot, int* state, int* status, int* iters,
                    double* relres)
{
    return server_poll(reinterpret_cast<Server*>(h), slot, state, status, iters, relres);
}

int ins_server_detach(ins_server* h, int slot)
{
    return server_detach(reinterpret_cast<Server*>(h), slot);
}

int ins_server_running(const ins_server* h, int* n)
{
    if (!h || !n) return INS_ERR_ARG;
    *n = server_running(reinterpret_cast<Server*>(const_cast<ins_server*>(h)));
    return INS_OK;
}

int ins_error_norm(ins_ctx* h, int order, int ref_kind, const double* ref_params,
                   const double* u, double* out)
{
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_CHECK(c);
    return error_norm(c, order, ref_kind, ref_params, u, out);
}

int ins_cell_eval(ins_ctx* h, int quantity, const double* xi, const double* u, double* out)
{
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_CHECK(c);
    return cell_eval(c, quantity, xi, u, out);
}

int ins_solve_csr_noalloc(int device, void* stream, int64_t n, int64_t nnz,
                          const int32_t* row_ptr, const int32_t* col_idx, const double* val,
                          const double* b, double* x, double rtol, int max_it, int* iters,
                          double* relres)
{
    return solve_csr_noalloc(device, stream, n, nnz, row_ptr, col_idx, val, b, x, rtol, max_it,
                             iters, relres);
}

int ins_bdf_init(ins_ctx* h, int order, double density, const double* u0)
{
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_CHECK(c);
    c->bdf_staged = false;
    if (order < 1 || order > 4 || !(density > 0.))
        return fail(c, INS_ERR_ARG, "ins_bdf_init: order 1..4 and a positive density");
    if (!c->have_dofs) return fail(c, INS_ERR_STATE, "ins_bdf_init before ins_set_dofs");
    const int64_t nd = c->n_nodes * c->dpn;
    c->u_zero = false;
    if (u0) INS_TRY(upload(c, c->u, u0, (size_t)nd));
    else INS_HIP_CHECK(c, hipMemsetAsync(c->u.p, 0, nd * sizeof(double), c->stream));
    c->u_zero = u0 == nullptr;
    for (int s = 0; s < order; s++) {
        INS_TRY(ensure(c, c->bdf_hist[s], (size_t)nd));
        INS_HIP_CHECK(c, hipMemcpyAsync(c->bdf_hist[s].p, c->u.p, nd * sizeof(double),
                                        hipMemcpyDeviceToDevice, c->stream));
        c->bdf_ring[s] = s;
    }
    INS_HIP_CHECK(c, hipStreamSynchronize(c->stream));
    c->bdf_order = order;
    c->bdf_step = 0;
    c->bdf_density = density;
    return INS_OK;
}

int ins_bdf_step(ins_ctx* h, double dt, int precision, int mode, int format, double rtol,
                 int max_it, int* iters, double* relres)
{
    Ctx* c = reinterpret_cast<Ctx*>(h);
    CTX_CHECK(c);
    if (precision != INS_FP64 && precision != INS_FP32)
        return fail(c, INS_ERR_ARG, "precision must be INS_FP64 or INS_FP32");
    if (mode != INS_SOLVE_FP64 && mode != INS_SOLVE_MIXED) return fail(c, INS_ERR_ARG, "solve mode");
    if (format != INS_FMT_CSR && format != INS_FMT_SELL) return fail(c, INS_ERR_ARG, "format");
    if (!(rtol > 0.) || max_it < 0) return fail(c, INS_ERR_ARG, "rtol/max_it");
    return bdf_step(c, dt, precision, mode, format, rtol, max_it, iters, relres);
}

}  // extern "C"
