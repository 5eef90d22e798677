// Internal declarations of the gfx950 finite-element library: the context
// that owns all device state, device buffers with budget accounting, and the
// launchers each translation unit provides.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <type_traits>
#include <vector>

#include "insilico_hip.h"

namespace ins {

// -------------------------------------------------------------------------
// status helpers
struct Ctx;
int fail(Ctx* c, int code, const std::string& msg);
int wait_server_release(Ctx* c);

#define INS_HIP_CHECK(ctx, call)                                              \
    do {                                                                      \
        hipError_t e_ = (call);                                               \
        if (e_ != hipSuccess)                                                 \
            return ::ins::fail((ctx), INS_ERR_HIP,                            \
                               std::string(#call ": ") + hipGetErrorString(e_)); \
    } while (0)

#define INS_TRY(expr)                                                         \
    do {                                                                      \
        int s_ = (expr);                                                      \
        if (s_ != INS_OK) return s_;                                          \
    } while (0)

// -------------------------------------------------------------------------
// device array owned by a context; grows on demand, counts against budget
template <typename T>
struct DArr {
    T* p = nullptr;
    size_t cap = 0;  // elements
};

// device-side scalars of one PCG solve
struct Scal {
    double rho[2];   // r.z of the current direction, by iteration parity
    double pq;       // p.Ap
    double alpha;
    double beta;
    double rr;       // |r|^2 of the recursive residual
    double bb;       // |b|^2
    double thr;      // rtol^2 |b|^2
    double tt;       // |b - A x|^2 (true residual)
    int done;        // 1 = stop (converged, breakdown or empty rhs)
    int breakdown;
    int iters;       // SpMVs performed inside the loop
    int restart;     // mixed: the replacement restarts the direction
    int fin;         // converged on the true (FP64) residual, or empty rhs
    int rep_at;      // mixed: iteration that replaces the residual (set by the
                     // iteration before it, so no block of the launch that
                     // sets it sees it)
    int conv32;      // mixed: the FP32 recursion reached the tolerance
    int nrep;        // mixed: residual replacements so far
    double rref;     // mixed: |b - A x|^2 at the last replacement
    double delta2;   // mixed: replace when |r|^2 < delta2 * rref
    int stop;        // FP64: converged in K3 of the last iteration; the next
                     // K1 turns it into done (K3 itself never reads it, so
                     // every block of that K3 still does its x update)
    int pad_;
};

// SELL-64 matrix: 64-row slices, values column-major inside a slice; every
// slice uses one column-offset list (col = row + delta) from a dictionary.
struct Sell {
    int64_t n = 0;
    int64_t nslices = 0;
    int64_t nvals = 0;       // padded value count
    int npat = 0;
    DArr<int64_t> slice_ptr; // nslices+1 value offsets (multiples of 64)
    DArr<int32_t> slice_pat; // nslices: id of the slice's column-offset list
    DArr<int32_t> pat_ptr;   // npat+1
    DArr<int32_t> pat_delta; // concatenated column offsets
    DArr<int32_t> pat_len;   // npat
    DArr<int32_t> map;       // nnz: slice position of every CSR entry
    DArr<int32_t> perm;      // slice row -> matrix row (vector problems)
    bool perm_on = false;
    DArr<double> v64;
    DArr<float> v32;
    bool ok = false;         // dictionary small enough to be worth it
    bool built = false;
    bool stale = true;       // values older than the last assembly
    bool zero64 = false;     // v64 / v32 sized and zeroed for the current slices
    bool zero32 = false;
    // Symmetric FP32 slices for the inner sweeps of scalar problems in the
    // identity row order: only the entries (r, r + u_k) for the offsets
    // u_k >= 0 of one list shared by all rows, stored at
    // (r / 64) * 64 * nu + k * 64 + r % 64; an entry below the diagonal,
    // (r, r - u_k), is read at its mirror (r - u_k, r).  Half the values of
    // the full slices.
    bool sym_on = false;
    bool zeros = false;      // vs sized and zeroed
    int nu = 0;
    DArr<int32_t> su;        // nu offsets, ascending
    DArr<float> vs;
    DArr<int32_t> smap;      // nnz: vs position of every CSR entry (-1 below the diagonal)
    int64_t nvs = 0;
    // Vector problems whose nodes carry all or none of their dpn components,
    // rows interleaved per node: the same idea on dpn x dpn node blocks
    // (sdpn = dpn).  su lists the node offsets u_k >= 0 shared by all
    // nodes; node a stores its blocks (a, a + u_k), the diagonal block as its
    // upper triangle, at (a / 64) * 64 * nbv + slot * 64 + a % 64 with
    // nbv = dpn (dpn + 1) / 2 + (nu - 1) dpn^2 values per node; the block
    // (a, a - u_k) is read transposed at node a - u_k.  The FP32 sweeps then
    // run on the component-major vectors of the full slices (component c of
    // node a at c * snodes + a).  sdpn = 1: the scalar slices above.
    int sdpn = 1;
    int64_t snodes = 0;
    // scalar slices whose offsets form stencil lines (c - 1, c, c + 1): the
    // SpMV loads p once per line and shifts it across the wave
    bool slines = false;
    // the same slices in FP64: the FP64 sweeps, the residual replacement and
    // the final residual of SELL solves read them (the full slices are not
    // built where the symmetric ones are)
    DArr<double> vs64;
    bool zeros64 = false;
    // Assembly straight into the symmetric slices (Q1 elements of up to 8
    // nodes, at most 8 elements per node): per stored slot (node a, offset
    // k) a word naming the elements node a shares with node a + u_k and the
    // local index of the latter in each (the pinfo layout, in node a's
    // adjacency order); anode[a] = mesh node of active node a
    bool direct = false;
    DArr<uint32_t> spw;
    DArr<int32_t> anode;
};

struct Ctx {
    int device = 0;
    int n_cu = 256;         // compute units of the device
    hipStream_t stream = nullptr;
    uint64_t budget = 0;
    int64_t held = 0;
    std::string err;
    bool profiling = false;
    ins_stats stats{};

    // mesh.  conn/npe/n_nodes describe the field nodes (DoF objects) the
    // pattern, the gather and the RHS work on; gconn/gnpe the Q1 geometry of
    // the element kernels.  Iso-parametric Q1 (deg 1): the two coincide and
    // gconn is empty.  adj entries are element << lg_npe | local.
    int dim = 0, npe = 0, lg_npe = 0;
    int64_t n_nodes = 0, n_elem = 0;
    DArr<double> xyz;
    DArr<int32_t> conn;
    int deg = 1, gnpe = 0;
    int64_t n_gnodes = 0;
    DArr<int32_t> gconn;
    const int32_t* geo_conn() const { return deg == 2 ? gconn.p : conn.p; }
    // solver storage order of the field nodes (degree 2): by n-face class,
    // then by support point; used to build the SELL slices
    DArr<int32_t> node_order;
    bool have_order = false;
    // degree 2: geometry node of every vertex field node (-1 for the others)
    DArr<int32_t> fvert;
    std::vector<int32_t> fvert_host;

    // dofs
    int dpn = 0;
    int64_t n_active = 0;      // rows of the system being solved
    int64_t dof_n_active = 0;  // active DoFs of ins_set_dofs (ins_set_csr may
                               // install a system of another size)
    bool monotone = true;
    DArr<int32_t> dof_id;   // n_nodes*dpn
    DArr<int32_t> rowmap;   // n_active: node*dpn + comp of each row
    DArr<double> g;         // prescribed
    DArr<double> u;         // current values
    bool have_dofs = false;

    // the DoF-derived maps (dof_id, rowmap, the component order of the SELL
    // slices) describe the system being solved
    bool dofs_match() const { return have_dofs && n_active == dof_n_active; }

    // pattern
    bool have_pattern = false;
    // the CSR came from ins_set_csr (values of unknown symmetry) rather than
    // from the element assembly (symmetric pattern and values)
    bool csr_from_host = false;
    int64_t nnz = 0;
    int max_nb = 0;         // max neighbour nodes of a node
    int max_adj = 0;        // max elements of a node
    // per (node, neighbour) pair, when every node has at most 8 elements:
    // one field per element of the node (in adjacency order) holding the
    // neighbour's local index there plus a presence bit; 4-bit fields in a
    // 32-bit word for up to 8 nodes per element (pinfo), 6-bit fields in a
    // 64-bit word for the degree-2 elements (pinfo64)
    DArr<uint32_t> pinfo;
    DArr<uint64_t> pinfo64;
    bool have_pinfo = false;
    DArr<int32_t> adj_ptr, adj;   // node -> element*npe + local, by element
    DArr<int32_t> nb_ptr, nb, nboff;
    DArr<int32_t> row_ptr, col;

    // material / loads
    int kernel = 0, nq = 2, ncoef = 0;
    DArr<double> coef;
    DArr<double> tab64;     // quadrature tables (dphi, phi, w)
    DArr<float> tab32;
    int tab_face_off = 0, tab_nqf = 0;
    int tab_deg = 0;        // field degree the tables were built for
    int bf_kind = INS_BF_NONE;
    double bf[4] = {0, 0, 0, 0};
    int64_t n_faces = 0;
    int nfn = 0;
    double traction[3] = {0, 0, 0};
    int tr_axis = -1;       // >= 0: traction scaled by (x[tr_axis] - tr_x0)
    double tr_x0 = 0.;
    // Neumann data as the co-normal derivative of a built-in function
    // (INS_REF_* kind, 0 = the traction vector above)
    int flux_kind = 0;
    double flux_p[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    DArr<int32_t> faces, fadj_ptr, fadj;
    // values of caller functions at the quadrature points: the body force
    // (INS_BF_TABLE) per element, point and component, and the Neumann
    // function per face, surface point and component
    DArr<double> bftab, ttab;
    bool have_bftab = false, have_ttab = false;
    // time-step terms: density of the mass matrix (0 = none), its factor in
    // the matrix (a0 / (b0 dt)) and the history combination of the RHS
    double density = 0., mass_factor = 0.;
    DArr<double> mhist;
    bool have_mhist = false;
    // device-resident BDF stepping: hist[ring[s]] = u^{n-s}
    int bdf_order = 0, bdf_step = 0, bdf_ring[4] = {0, 1, 2, 3};
    double bdf_density = 0.;
    // the staged element data and the CSR values of the last step's full
    // assembly stay valid for the next step with the same factor a0/dt
    bool bdf_staged = false;
    double bdf_factor = 0.;
    int bdf_precision = 0;
    DArr<double> bdf_hist[4];

    // assembly products
    bool have_matrix = false, have_rhs = false;
    int matrix_precision = INS_FP64;
    // storage of the last solve or ins_prepare: an assembly of a problem
    // with symmetric slices writes them directly when it is INS_FMT_SELL
    int want_format = INS_FMT_CSR;
    // the CSR values predate the last assembly, which wrote only the
    // symmetric slices (ensure_csr rebuilds them on demand)
    bool csr_stale = false;
    // per element: bit b * dpn + c set where component c of local node b is
    // prescribed (elements of at most 32 such bits; the RHS lifting visits
    // only those)
    DArr<uint32_t> emask;
    bool have_emask = false;
    // u (the DoF values) is all zeros: set where it is cleared, dropped by
    // every write of it (upload, distribute, time stepping)
    bool u_zero = true;
    DArr<double> val64;
    DArr<float> val32;
    DArr<double> b;
    DArr<double> dinv64;
    DArr<float> dinv32;
    DArr<uint8_t> stage_ke;  // element matrices (type depends on precision)
    DArr<uint8_t> stage_me;  // element mass matrices (time steps)
    DArr<uint8_t> stage_pt;  // point records of the hyperelastic kernels
    DArr<double> stage_fe;   // element force vectors (FP64)
    bool fe_slab = false;    // ... in 64-element slabs (heat::Laplace Q1) or element-major
    DArr<double> stage_ff;   // face force vectors
    Sell sell;

    // solver state
    DArr<double> x, r, p, q, t64;
    DArr<float> x32, r32, p32, q32;
    DArr<double> partial;
    // SELL solves of vector problems run in the component-major numbering:
    // permuted copies of b and the Jacobi diagonal, and the FP64 solution
    DArr<double> xp, bp, d64p;
    DArr<float> d32p;
    // vectors of the current solve in its numbering (set by solve_vectors)
    double* sx64 = nullptr;
    const double* sb = nullptr;
    const double* sd64 = nullptr;
    const float* sd32 = nullptr;
    bool sperm = false;
    DArr<Scal> scal;
    Scal* h_scal = nullptr;  // pinned
    Scal* h_snap = nullptr;  // pinned, 2 snapshots for the pipelined host loop
    hipEvent_t evs[2] = {nullptr, nullptr};
    DArr<int> dflag;
    bool have_solution = false;

    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    hipEvent_t ev_el[2] = {nullptr, nullptr};  // profiled element kernels
    // set by ins_server_detach: the server stream's work queued up to the
    // detach, which may still name this context's buffers; every later call
    // on the context waits for it before touching device memory
    hipEvent_t srv_ev = nullptr;
    bool srv_pending = false;
    // profiling: one event pair per SpMV launch of a solve (grown on demand)
    std::vector<hipEvent_t> prof_ev;
    size_t prof_n = 0;
};

// word of the per-(node, neighbour) shared-element record: 4-bit fields
// (presence bit 8 + local index) for elements of up to 8 nodes, 6-bit fields
// (presence bit 32 + local index) for the 9- and 27-node elements
template <int NPE>
struct PairWord {
    typedef typename std::conditional<(NPE <= 8), uint32_t, uint64_t>::type type;
    static constexpr int bits = NPE <= 8 ? 4 : 6;
    static constexpr int flag = 1 << (bits - 1);
};

// allocation with budget accounting (count in elements)
int dalloc_bytes(Ctx* c, void** p, size_t* cap_bytes, size_t bytes);
void dfree_bytes(Ctx* c, void** p, size_t* cap_bytes);

template <typename T>
int ensure(Ctx* c, DArr<T>& a, size_t n)
{
    if (n <= a.cap && a.p) return INS_OK;
    size_t capb = a.cap * sizeof(T);
    void* p = a.p;
    dfree_bytes(c, &p, &capb);
    a.p = nullptr;
    a.cap = 0;
    size_t want = (n ? n : 1) * sizeof(T);
    int s = dalloc_bytes(c, &p, &capb, want);
    if (s != INS_OK) return s;
    a.p = static_cast<T*>(p);
    a.cap = capb / sizeof(T);
    return INS_OK;
}

template <typename T>
void release(Ctx* c, DArr<T>& a)
{
    size_t capb = a.cap * sizeof(T);
    void* p = a.p;
    dfree_bytes(c, &p, &capb);
    a.p = nullptr;
    a.cap = 0;
}

// grid helper
// Positions in the symmetric slices (64 rows or nodes per slice), slot-major
// and lane contiguous: slot s of row or node a at
// (a/64)*64*width + s*64 + a%64.  Scalar rows (D = 1): slot k is the entry
// of offset u_k.  Node blocks (D = 2, 3): slots 0..T-1 the upper triangle
// of the diagonal block (row-major, i <= j), slot T + (k-1)D^2 + iD + j the
// entry (i, j) of block k >= 1, T = D(D+1)/2.
struct SymLayout {
    int D, nu;
    __host__ __device__ int64_t width() const
    {
        return (int64_t)D * (D + 1) / 2 + (int64_t)(nu - 1) * D * D;
    }
    __host__ __device__ int slot(int k, int i, int j) const
    {
        if (k == 0) return i * D - i * (i - 1) / 2 + (j - i);
        return D * (D + 1) / 2 + (k - 1) * D * D + i * D + j;
    }
    __host__ __device__ int64_t pos(int64_t a, int s) const
    {
        return (a >> 6) * 64 * width() + (int64_t)s * 64 + (a & 63);
    }
};

inline unsigned grid_for(int64_t work, int block, unsigned cap = 65535u * 8)
{
    int64_t g = (work + block - 1) / block;
    if (g < 1) g = 1;
    if (g > (int64_t)cap) g = cap;
    return (unsigned)g;
}

// ---- launchers provided by the translation units ---------------------------
// solve.hip: serving (batched mixed-precision solves)
struct Server;
int server_create(Server** out, int device, int max_slots);
void server_destroy(Server* sv);
const char* server_error(Server* sv);
int server_attach(Server* sv, Ctx* c, double rtol, int max_it, int* slot);
int server_step(Server* sv, int k, int* n_finished);
int server_poll(Server* sv, int s, int* state, int* status, int* iters, double* relres);
int server_detach(Server* sv, int s);
int server_running(Server* sv);
// scan.hip
int exclusive_scan_i32(Ctx* c, const int32_t* in, int32_t* out, int64_t n,
                       int64_t* total);
// setup.hip
int build_adjacency(Ctx* c, const int32_t* items, int64_t n_items, int per_item,
                    int lg_stride, int64_t n_nodes, DArr<int32_t>& ptr,
                    DArr<int32_t>& list, int* max_deg);
inline int ceil_lg2(int v)
{
    int l = 0;
    while ((1 << l) < v) l++;
    return l;
}
int build_pattern(Ctx* c);
// assemble.hip
int make_tables(Ctx* c);
// the tensor Gauss rule of the element kernels: returns npts^dim, or -1
int quadrature_rule(int dim, int npts, double* w, double* xi);
constexpr int ASM_STAGED = 1 << 8;  // internal: RHS from the staged element data
int assemble(Ctx* c, int precision, int flags);
int compute_dinv(Ctx* c);
int error_norm(Ctx* c, int order, int kind, const double* prm, const double* u, double* out);
int cell_eval(Ctx* c, int quantity, const double* xi, const double* u, double* out);
// solve.hip
int solve(Ctx* c, int mode, int format, double rtol, int max_it, int* iters,
          double* relres);
int build_sell(Ctx* c, bool force);
int fill_sell(Ctx* c);
int zero_sym(Ctx* c);
int ensure_csr(Ctx* c);
int prepare(Ctx* c, int mode, int format);
int solve_direct(Ctx* c, int kind, double* relres);
int distribute(Ctx* c, int op);
int bdf_step(Ctx* c, double dt, int precision, int mode, int format, double rtol, int max_it,
             int* iters, double* relres);
int vec_norm2(Ctx* c, const double* v, int64_t n, double* out);
int solve_csr_noalloc(int device, void* stream, int64_t n, int64_t nnz, const int32_t* rp,
                      const int32_t* col, const double* val, const double* b, double* x,
                      double rtol, int max_it, int* iters, double* relres);

}  // namespace ins
