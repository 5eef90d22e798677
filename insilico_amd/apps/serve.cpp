// Serving stream: Poisson arrivals of 3-D diffusion and elasticity solves of
// uneven sizes, admitted by free device memory, solved by continuous batching
// on one GPU (ins_server_*), with arrival-to-result latency per request.
//
// Admission is first come first served while the device reports enough free
// HBM (hipMemGetInfo through ins_device_memory, less a margin) for the
// request's estimated footprint, within an optional budget.
//
// Requests: 50/50 diffusion (-div(k grad u) = 1, k_e = exp(g_e), g_e ~ N(0,1),
// u = 0 on x = 0) and linear elasticity (E_e ~ U[0.5, 2], nu = 0.3, clamped
// z = 0, traction (0, 0, -0.01) on z = 1) on n^3 Q1 hex meshes with
// n = round(exp(U[ln nmin, ln nmax])), coefficient field per request seed,
// mixed-precision Jacobi-PCG to 1e-8 (FP32 sweeps, FP64 residual
// replacement).  Mesh, DoF and boundary conventions are those of the
// reference's unitCube generator and numbering (unitCube.hpp:85-106,217-257;
// base/dof/numbering.hpp:45-66), as in insilico_amd/problems.py.
//
// Phases: (1) capacity: a batch of requests all present at t = 0, solved
// back to back, gives requests/s; (2) the stream: exponential inter-arrival
// times at load x capacity; (3) check: sampled requests are solved again
// alone through ins_solve and compared bit for bit with their result in the
// stream.  Prints one JSON line.
//
// Host work is split: worker threads generate a request and run its setup
// and assembly on its own context (own HIP stream); the main thread admits,
// attaches, steps the server and collects results.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "insilico_hip.h"

namespace {

using Clock = std::chrono::steady_clock;

double now_s(Clock::time_point t0)
{
    return std::chrono::duration<double>(Clock::now() - t0).count();
}

void die(const char* what, int st, const char* msg)
{
    std::fprintf(stderr, "serve: %s failed (%d): %s\n", what, st, msg ? msg : "");
    std::exit(1);
}

#define CK(ctx, call)                                                    \
    do {                                                                 \
        int s_ = (call);                                                 \
        if (s_ != INS_OK) die(#call, s_, ins_last_error(ctx));           \
    } while (0)
#define SK(srv, call)                                                    \
    do {                                                                 \
        int s_ = (call);                                                 \
        if (s_ != INS_OK) die(#call, s_, ins_server_last_error(srv));    \
    } while (0)

// ---- random numbers: SplitMix64 (as problems.splitmix64) --------------------
const uint64_t GAMMA = 0x9E3779B97F4A7C15ULL;
uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
// output k (0-based) of SplitMix64 seeded with seed
uint64_t splitmix(uint64_t seed, uint64_t k) { return mix64(seed + (k + 1) * GAMMA); }
double u01(uint64_t seed, uint64_t k) { return (double)(splitmix(seed, k) >> 11) * 0x1.0p-53; }

// ---- requests ------------------------------------------------------------------
enum { DIFFUSION = 0, ELASTICITY = 1 };

struct Spec {
    int id;
    int phys;
    int n;
    uint64_t seed;
    double t_arrival = 0.;
    int64_t n_dof = 0;
    int64_t nnz = 0;
    double est_bytes = 0.;
};

struct Problem {
    int dpn = 1;
    std::vector<double> xyz;
    std::vector<int32_t> conn;
    std::vector<int32_t> dof;
    int64_t n_active = 0;
    int kernel = INS_LAPLACE;
    std::vector<double> coef;
    int bf_kind = INS_BF_NONE;
    double bf[4] = {0, 0, 0, 0};
    std::vector<int32_t> faces;
    double traction[3] = {0, 0, 0};
};

// structured unit cube, nodes and elements first index fastest, hierarchic
// vertex order {0,1,3,2,4,5,7,6} of the lexicographic corners
void generate(const Spec& s, Problem& P)
{
    const int n = s.n, N1 = n + 1;
    const int64_t nn = (int64_t)N1 * N1 * N1, ne = (int64_t)n * n * n;
    const double h = 1.0 / (double)n;
    P.xyz.resize(nn * 3);
    for (int64_t k = 0, i3 = 0; i3 < N1; i3++)
        for (int64_t i2 = 0; i2 < N1; i2++)
            for (int64_t i1 = 0; i1 < N1; i1++, k++) {
                P.xyz[3 * k] = h * (double)i1;
                P.xyz[3 * k + 1] = h * (double)i2;
                P.xyz[3 * k + 2] = h * (double)i3;
            }
    static const int H[8] = {0, 1, 3, 2, 4, 5, 7, 6};
    P.conn.resize(ne * 8);
    for (int64_t e = 0, i3 = 0; i3 < n; i3++)
        for (int64_t i2 = 0; i2 < n; i2++)
            for (int64_t i1 = 0; i1 < n; i1++, e++) {
                const int64_t base = i1 + i2 * N1 + i3 * N1 * N1;
                for (int v = 0; v < 8; v++) {
                    const int d1 = v & 1, d2 = (v >> 1) & 1, d3 = v >> 2;
                    P.conn[e * 8 + H[v]] = (int32_t)(base + d3 * N1 * N1 + d2 * N1 + d1);
                }
            }
    if (s.phys == DIFFUSION) {
        P.dpn = 1;
        P.kernel = INS_LAPLACE;
        P.dof.assign(nn, -1);
        int32_t cnt = 0;
        for (int64_t k = 0; k < nn; k++)
            if (k % N1 != 0) P.dof[k] = cnt++;  // u = 0 on x = 0
        P.n_active = cnt;
        P.coef.resize(ne);
        for (int64_t e = 0; e < ne; e++) {
            const double a = u01(s.seed, 2 * e), b = u01(s.seed, 2 * e + 1);
            P.coef[e] = std::exp(std::sqrt(-2.0 * std::log1p(-a)) * std::cos(2.0 * M_PI * b));
        }
        P.bf_kind = INS_BF_CONST;
        P.bf[0] = 1.0;
        P.faces.clear();
    } else {
        P.dpn = 3;
        P.kernel = INS_STVENANT;
        P.dof.assign(nn * 3, -1);
        int32_t cnt = 0;
        for (int64_t k = 0; k < nn; k++)
            if (k >= (int64_t)N1 * N1)  // clamped on z = 0
                for (int c = 0; c < 3; c++) P.dof[3 * k + c] = cnt++;
        P.n_active = cnt;
        P.coef.resize(ne * 2);
        const double nu = 0.3;
        for (int64_t e = 0; e < ne; e++) {
            const double E = 0.5 + 1.5 * u01(s.seed, e);
            P.coef[2 * e] = E * nu / (1.0 + nu) / (1.0 - 2.0 * nu);
            P.coef[2 * e + 1] = E / 2.0 / (1.0 + nu);
        }
        P.bf_kind = INS_BF_NONE;
        // top faces (hierarchic vertices 4..7 of the top element layer)
        P.faces.resize((int64_t)n * n * 4);
        for (int64_t f = 0; f < (int64_t)n * n; f++) {
            const int64_t e = (int64_t)(n - 1) * n * n + f;
            for (int v = 0; v < 4; v++) P.faces[f * 4 + v] = P.conn[e * 8 + 4 + v];
        }
        P.traction[2] = -0.01;
    }
}

void size_spec(Spec& s)
{
    const int64_t N = s.n;
    const int64_t nodes = (N + 1) * (N + 1) * (N + 1);
    if (s.phys == DIFFUSION) {
        s.n_dof = nodes - (N + 1) * (N + 1);
        s.nnz = (3 * N - 2) * (3 * N + 1) * (3 * N + 1);
    } else {
        s.n_dof = 3 * (nodes - (N + 1) * (N + 1));
        s.nnz = 9 * (3 * N - 2) * (3 * N + 1) * (3 * N + 1);
    }
    const double E = (double)N * N * N;
    // values (FP64 + FP32), columns, SELL map and slices, element staging,
    // vectors; refined by the bytes the contexts really hold
    s.est_bytes = 40.0 * s.nnz + (s.phys == DIFFUSION ? 300.0 : 2400.0) * E + 200.0 * s.n_dof +
                  64.0 * nodes;
}

void setup(ins_ctx* c, const Problem& P)
{
    const int64_t nn = (int64_t)P.xyz.size() / 3, ne = (int64_t)P.conn.size() / 8;
    CK(c, ins_set_mesh(c, 3, nn, P.xyz.data(), ne, 8, P.conn.data()));
    CK(c, ins_set_dofs(c, P.dpn, P.dof.data(), P.n_active));
    CK(c, ins_set_prescribed(c, nullptr));
    CK(c, ins_set_material(c, P.kernel, 2, P.kernel == INS_LAPLACE ? 1 : 2, P.coef.data()));
    CK(c, ins_set_body_force(c, P.bf_kind, P.bf));
    if (!P.faces.empty())
        CK(c, ins_set_traction(c, (int64_t)P.faces.size() / 4, 4, P.faces.data(), P.traction));
    else
        CK(c, ins_set_traction(c, 0, 0, nullptr, nullptr));
    int64_t nnz = 0;
    CK(c, ins_build_pattern(c, &nnz));
    // the SELL form first: the assembly then writes the slices the batched
    // solve reads (the symmetric ones where the problem has them)
    CK(c, ins_prepare(c, INS_SOLVE_MIXED, INS_FMT_SELL));
    CK(c, ins_assemble(c, INS_FP64, INS_ASM_MATRIX | INS_ASM_RHS));
}

// ---- worker pool: generation + setup + assembly on a request's context ----------
struct Job {
    Spec spec;
    ins_ctx* ctx;
};

struct Pool {
    std::mutex m;
    std::condition_variable cv_job, cv_done;
    std::deque<Job> jobs, done;
    bool stop = false;
    std::vector<std::thread> threads;

    void start(int nthreads)
    {
        for (int i = 0; i < nthreads; i++)
            threads.emplace_back([this] {
                Problem P;
                for (;;) {
                    Job j;
                    {
                        std::unique_lock<std::mutex> lk(m);
                        cv_job.wait(lk, [this] { return stop || !jobs.empty(); });
                        if (stop && jobs.empty()) return;
                        j = jobs.front();
                        jobs.pop_front();
                    }
                    generate(j.spec, P);
                    setup(j.ctx, P);
                    {
                        std::lock_guard<std::mutex> lk(m);
                        done.push_back(j);
                    }
                    cv_done.notify_all();
                }
            });
    }
    void submit(const Job& j)
    {
        {
            std::lock_guard<std::mutex> lk(m);
            jobs.push_back(j);
        }
        cv_job.notify_one();
    }
    bool take(Job* j)
    {
        std::lock_guard<std::mutex> lk(m);
        if (done.empty()) return false;
        *j = done.front();
        done.pop_front();
        return true;
    }
    void finish()
    {
        {
            std::lock_guard<std::mutex> lk(m);
            stop = true;
        }
        cv_job.notify_all();
        for (auto& t : threads) t.join();
        threads.clear();
    }
};

struct Result {
    double t_arrival = 0., t_admit = 0., t_ready = 0., t_done = 0.;
    int iters = 0, status = 0;
    double relres = 0.;
    std::vector<double> x;  // kept for the requests checked afterwards
};

struct Opts {
    int requests = 2000;
    uint64_t seed = 12345;
    int nmin = 16, nmax = 96;
    double load = 0.7;
    double budget_gb = 0.;
    int slots = 32;
    int chunk = 8;
    int calib = 150;
    int verify = 4;
    int workers = 4;
    double rtol = 1e-8;
};

// Runs the specs (arrival times set) through the server; returns the makespan.
// device memory left to the runtime and the server when admitting
const double FREE_MARGIN = 4e9;

double run_stream(const Opts& o, std::vector<Spec>& specs, std::vector<Result>& res,
                  const std::vector<char>& keep, uint64_t budget, const char* phase,
                  int* max_resident, double* max_bytes, double* min_free)
{
    ins_server* srv = nullptr;
    {
        const int st = ins_server_create(&srv, 0, o.slots);
        if (st != INS_OK) die("ins_server_create", st, ins_strerror(st));
    }
    struct CtxSlot {
        ins_ctx* c = nullptr;
        double held = 0.;
        bool busy = false;
    };
    std::vector<CtxSlot> ctxs(o.slots);
    Pool pool;
    pool.start(o.workers);
    res.assign(specs.size(), Result());
    std::deque<int> queue;         // arrived, not admitted
    std::vector<int> slot_req(o.slots, -1);
    std::vector<int> req_ctx(specs.size(), -1);
    size_t next = 0, finished = 0;
    int in_setup = 0, resident = 0;
    *max_resident = 0;
    *max_bytes = 0.;
    double reserved = 0.;          // estimates of requests in setup
    const Clock::time_point t0 = Clock::now();
    double last_log = 0.;
    std::vector<double> x;

    auto held_total = [&]() {
        double s = 0.;
        for (auto& cs : ctxs) s += cs.held;
        return s;
    };
    auto refresh_held = [&](int k) {
        ins_stats st;
        if (ctxs[k].c && ins_get_stats(ctxs[k].c, &st) == INS_OK) ctxs[k].held = (double)st.device_bytes;
    };

    while (finished < specs.size()) {
        const double t = now_s(t0);
        while (next < specs.size() && specs[next].t_arrival <= t) queue.push_back((int)next++);
        // admission: first come first served, by free memory and free contexts
        while (!queue.empty()) {
            const Spec& s = specs[queue.front()];
            int k = -1;
            double freeable = 0.;
            for (int i = 0; i < o.slots; i++)
                if (!ctxs[i].busy && (k < 0 || ctxs[i].held > ctxs[k].held)) k = i;
            if (k < 0) break;
            freeable = ctxs[k].held;
            // free HBM as the device reports it (the context about to be
            // reused gives its buffers back), less the estimates of the
            // requests still in setup and a margin for the runtime
            uint64_t fr = 0, tot = 0;
            {
                const int st = ins_device_memory(0, &fr, &tot);
                if (st != INS_OK) die("ins_device_memory", st, ins_strerror(st));
            }
            *min_free = std::min(*min_free, (double)fr);
            const bool no_room = (double)fr + freeable - reserved - s.est_bytes < FREE_MARGIN;
            if (no_room || held_total() - freeable + reserved + s.est_bytes > (double)budget) {
                // give memory of other idle contexts back before waiting
                bool freed = false;
                for (int i = 0; i < o.slots; i++)
                    if (i != k && !ctxs[i].busy && ctxs[i].c) {
                        ins_destroy(ctxs[i].c);
                        ctxs[i].c = nullptr;
                        ctxs[i].held = 0.;
                        freed = true;
                        break;
                    }
                if (freed) continue;
                if (resident + in_setup > 0) break;  // wait for memory
            }
            if (!ctxs[k].c) {
                const int st = ins_create(&ctxs[k].c, 0, 0);
                if (st != INS_OK) die("ins_create", st, ins_strerror(st));
            }
            ctxs[k].busy = true;
            const int id = queue.front();
            queue.pop_front();
            req_ctx[id] = k;
            res[id].t_arrival = specs[id].t_arrival;
            res[id].t_admit = now_s(t0);
            reserved += specs[id].est_bytes;
            in_setup++;
            pool.submit(Job{specs[id], ctxs[k].c});
        }
        // requests whose setup finished join the batch
        Job j;
        while (pool.take(&j)) {
            const int id = j.spec.id;
            int slot = -1;
            SK(srv, ins_server_attach(srv, j.ctx, o.rtol, 20000, &slot));
            slot_req[slot] = id;
            res[id].t_ready = now_s(t0);
            reserved -= specs[id].est_bytes;
            in_setup--;
            resident++;
            refresh_held(req_ctx[id]);
        }
        *max_resident = std::max(*max_resident, resident);
        *max_bytes = std::max(*max_bytes, held_total());
        int nrun = 0;
        ins_server_running(srv, &nrun);
        if (nrun == 0 && resident == 0) {
            // idle: wait for an arrival or a setup
            std::this_thread::sleep_for(std::chrono::microseconds(50));
            continue;
        }
        int nfin = 0;
        SK(srv, ins_server_step(srv, o.chunk, &nfin));
        if (nfin == 0 && nrun > 0) continue;
        for (int sl = 0; sl < o.slots; sl++) {
            const int id = slot_req[sl];
            if (id < 0) continue;
            int state = 0, status = 0, it = 0;
            double rr = 0.;
            SK(srv, ins_server_poll(srv, sl, &state, &status, &it, &rr));
            if (state != INS_SLOT_FINISHED) continue;
            const int k = req_ctx[id];
            const int64_t n = specs[id].n_dof;
            x.resize(n);
            CK(ctxs[k].c, ins_get_solution(ctxs[k].c, x.data()));
            Result& r = res[id];
            r.t_done = now_s(t0);
            r.iters = it;
            r.status = status;
            r.relres = rr;
            if (keep[id]) r.x = x;
            SK(srv, ins_server_detach(srv, sl));
            slot_req[sl] = -1;
            ctxs[k].busy = false;
            refresh_held(k);
            resident--;
            finished++;
        }
        const double tl = now_s(t0);
        if (tl - last_log > 5.0) {
            last_log = tl;
            std::fprintf(stderr, "[serve] %s t=%.1fs arrived=%zu done=%zu resident=%d setup=%d queued=%zu\n",
                         phase, tl, next, finished, resident, in_setup, queue.size());
        }
    }
    const double makespan = now_s(t0);
    pool.finish();
    ins_server_destroy(srv);
    for (auto& cs : ctxs)
        if (cs.c) ins_destroy(cs.c);
    return makespan;
}

double pct(std::vector<double> v, double p)
{
    if (v.empty()) return 0.;
    std::sort(v.begin(), v.end());
    const double r = p / 100.0 * (double)(v.size() - 1);
    const size_t lo = (size_t)std::floor(r), hi = std::min(v.size() - 1, lo + 1);
    return v[lo] + (r - (double)lo) * (v[hi] - v[lo]);
}

}  // namespace

int main(int argc, char** argv)
{
    Opts o;
    for (int i = 1; i < argc; i++) {
        const std::string a = argv[i];
        auto val = [&](void) -> const char* {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "missing value for %s\n", a.c_str());
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "--requests") o.requests = std::atoi(val());
        else if (a == "--seed") o.seed = std::strtoull(val(), nullptr, 10);
        else if (a == "--nmin") o.nmin = std::atoi(val());
        else if (a == "--nmax") o.nmax = std::atoi(val());
        else if (a == "--load") o.load = std::atof(val());
        else if (a == "--budget-gb") o.budget_gb = std::atof(val());
        else if (a == "--slots") o.slots = std::atoi(val());
        else if (a == "--chunk") o.chunk = std::atoi(val());
        else if (a == "--calib") o.calib = std::atoi(val());
        else if (a == "--verify") o.verify = std::atoi(val());
        else if (a == "--workers") o.workers = std::atoi(val());
        else if (a == "--rtol") o.rtol = std::atof(val());
        else {
            std::printf("Usage: %s [--requests N] [--seed S] [--nmin A] [--nmax B] [--load F]\n"
                        "          [--budget-gb G] [--slots K] [--chunk C] [--calib M]\n"
                        "          [--verify V] [--workers W] [--rtol T]\n", argv[0]);
            return a == "--help" ? 0 : 2;
        }
    }
    // request specs
    std::vector<Spec> specs(o.requests);
    for (int i = 0; i < o.requests; i++) {
        Spec& s = specs[i];
        s.id = i;
        s.phys = u01(o.seed, 3 * (uint64_t)i) < 0.5 ? DIFFUSION : ELASTICITY;
        const double ln = std::log((double)o.nmin) +
                          u01(o.seed, 3 * (uint64_t)i + 1) * (std::log((double)o.nmax) - std::log((double)o.nmin));
        s.n = std::max(o.nmin, std::min(o.nmax, (int)std::lround(std::exp(ln))));
        s.seed = splitmix(o.seed ^ 0x5EEDULL, (uint64_t)i);
        size_spec(s);
    }
    // admission budget: device memory the resident requests may hold
    // (MI355X: 288 GB of HBM; the rest is left to the runtime and the server)
    const uint64_t budget = (uint64_t)((o.budget_gb > 0 ? o.budget_gb : 200.0) * 1e9);

    // (1) capacity: the first calib requests present at t = 0
    const int nc = std::min(o.calib, o.requests);
    std::vector<Spec> cal(specs.begin(), specs.begin() + nc);
    for (auto& s : cal) s.t_arrival = 0.;
    std::vector<Result> rc;
    int mr = 0;
    double mb = 0., mf = 1e300;
    const double t_cal = run_stream(o, cal, rc, std::vector<char>(nc, 0), budget, "capacity", &mr,
                                    &mb, &mf);
    const double capacity = nc / t_cal;
    const double rate = o.load * capacity;

    // (2) the stream
    double t = 0.;
    for (int i = 0; i < o.requests; i++) {
        const double u = u01(o.seed ^ 0xA7712ULL, (uint64_t)i);
        t += -std::log1p(-u) / rate;
        specs[i].t_arrival = t;
    }
    std::vector<char> keep(o.requests, 0);
    std::vector<int> check;
    for (int k = 0; k < o.verify && o.requests > 0; k++) {
        const int id = (int)(splitmix(o.seed ^ 0xC4EC4ULL, (uint64_t)k) % (uint64_t)o.requests);
        if (!keep[id]) check.push_back(id);
        keep[id] = 1;
    }
    std::vector<Result> res;
    int max_res = 0;
    double max_bytes = 0., min_free = 1e300;
    std::fprintf(stderr, "[serve] capacity %.2f req/s from %d requests; stream of %d at %.2f req/s\n",
                 capacity, nc, o.requests, rate);
    const double makespan = run_stream(o, specs, res, keep, budget, "stream", &max_res, &max_bytes,
                                       &min_free);

    // (3) the sampled requests alone through ins_solve: bit-identical?
    int equal = 0;
    Problem P;
    std::vector<double> x;
    for (int id : check) {
        ins_ctx* c = nullptr;
        if (ins_create(&c, 0, 0) != INS_OK) die("ins_create", -2, "");
        generate(specs[id], P);
        setup(c, P);
        int it = 0;
        double rr = 0.;
        CK(c, ins_solve(c, INS_SOLVE_MIXED, INS_FMT_SELL, o.rtol, 20000, &it, &rr));
        x.resize(specs[id].n_dof);
        CK(c, ins_get_solution(c, x.data()));
        if (x.size() == res[id].x.size() &&
            std::memcmp(x.data(), res[id].x.data(), x.size() * sizeof(double)) == 0 &&
            it == res[id].iters)
            equal++;
        ins_destroy(c);
    }

    // requests that did not converge: the same request alone, mixed (as in
    // the stream) and plain FP64, to tell the batch from the arithmetic
    int reported = 0;
    for (int id = 0; id < o.requests && reported < 8; id++) {
        if (res[id].status == INS_OK) continue;
        reported++;
        const Spec& s = specs[id];
        std::fprintf(stderr, "[serve] not converged: id %d %s n %d seed %llu: status %d, %d iterations, relres %.3e\n",
                     id, s.phys == DIFFUSION ? "diffusion" : "elasticity", s.n,
                     (unsigned long long)s.seed, res[id].status, res[id].iters, res[id].relres);
        for (int mode = 0; mode < 2; mode++) {
            ins_ctx* c = nullptr;
            if (ins_create(&c, 0, 0) != INS_OK) die("ins_create", -2, "");
            generate(s, P);
            setup(c, P);
            int it = 0;
            double rr = 0.;
            const int st = ins_solve(c, mode == 0 ? INS_SOLVE_MIXED : INS_SOLVE_FP64,
                                     mode == 0 ? INS_FMT_SELL : INS_FMT_CSR, o.rtol, 20000, &it, &rr);
            std::fprintf(stderr, "[serve]   alone %s: status %d, %d iterations, relres %.3e\n",
                         mode == 0 ? "mixed/SELL" : "FP64/CSR", st, it, rr);
            ins_destroy(c);
        }
    }

    std::vector<double> lat, wait, setup_t;
    double dofs = 0.;
    int ok = 0, worst_status = 0;
    double worst_rel = 0.;
    int n_diff = 0;
    for (int i = 0; i < o.requests; i++) {
        lat.push_back(res[i].t_done - res[i].t_arrival);
        wait.push_back(res[i].t_admit - res[i].t_arrival);
        setup_t.push_back(res[i].t_ready - res[i].t_admit);
        dofs += (double)specs[i].n_dof;
        ok += res[i].status == INS_OK;
        if (res[i].status != INS_OK) worst_status = res[i].status;
        worst_rel = std::max(worst_rel, res[i].relres);
        n_diff += specs[i].phys == DIFFUSION;
    }
    std::printf("{\"metric\": \"serving stream: p50/p99 arrival-to-result latency\", "
                "\"requests\": %d, \"diffusion\": %d, \"elasticity\": %d, \"n_range\": [%d, %d], "
                "\"rtol\": %g, \"capacity_req_per_s\": %.3f, \"load\": %.2f, \"rate_req_per_s\": %.3f, "
                "\"makespan_s\": %.3f, \"throughput_req_per_s\": %.3f, \"throughput_dof_per_s\": %.1f, "
                "\"p50_latency_ms\": %.2f, \"p99_latency_ms\": %.2f, \"mean_latency_ms\": %.2f, "
                "\"p50_queue_wait_ms\": %.2f, \"p99_queue_wait_ms\": %.2f, \"p50_setup_ms\": %.2f, "
                "\"max_resident\": %d, \"max_device_gb\": %.2f, \"budget_gb\": %.1f, "
                "\"admission\": \"free HBM (hipMemGetInfo) less %.0f GB margin, and the budget\", "
                "\"min_free_gb\": %.2f, \"slots\": %d, "
                "\"converged\": %d, \"worst_status\": %d, \"max_relres\": %.3e, "
                "\"alone_vs_mix_checked\": %zu, \"alone_vs_mix_bitwise_equal\": %d}\n",
                o.requests, n_diff, o.requests - n_diff, o.nmin, o.nmax, o.rtol, capacity, o.load,
                rate, makespan, o.requests / makespan, dofs / makespan, 1e3 * pct(lat, 50),
                1e3 * pct(lat, 99), 1e3 * (lat.empty() ? 0. : [&] { double s = 0; for (double v : lat) s += v; return s / lat.size(); }()),
                1e3 * pct(wait, 50), 1e3 * pct(wait, 99), 1e3 * pct(setup_t, 50), max_res,
                max_bytes / 1e9, budget / 1e9, FREE_MARGIN / 1e9, min_free / 1e9, o.slots, ok,
                worst_status, worst_rel, check.size(),
                equal);
    return 0;
}
