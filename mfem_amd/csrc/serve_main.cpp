// serve_main.cpp -- h1pa_serve: a request loop that serves Poisson solves on
// one GPU through the C ABI (include/h1pa.h) only.  No MFEM counterpart.
//
// Requests arrive as a Poisson process; each names an n^3 Cartesian hex mesh
// (n in [nmin, nmax], P(n) proportional to 1/n), an order p (uniform in
// [pmin, pmax]), a coefficient seed (k = exp of 8 seeded Fourier modes) and
// an RHS seed (uniform [-1, 1]); interior vertices are perturbed by 0.15 h,
// the boundary is homogeneous Dirichlet, the stop test is CGSolver's
// relative measure at rel_tol.
//
//   --mode batch   continuous batching: arrived requests are admitted in
//                  arrival order while the HBM budget allows and set up on
//                  --setup-threads host threads (own streams) beside the
//                  replays;
//                  all admitted solves advance together in graph replays that
//                  run back to back (the loop checks h1pa_batch_query,
//                  retires the finished solves, releases their memory, adds
//                  the finished setups and launches the next replay at once).
//   --mode serial  plain FP64 GPU PCG: requests are solved one at a time, in
//                  arrival order (h1pa_pcg_solve); the same setup thread
//                  prepares the next request during each solve.
//   --mode mg      continuous serving with p-multigrid: arrived requests are
//                  admitted in arrival order while the HBM budget allows, set
//                  up (with their multigrid levels: orders halved down to 1,
//                  Chebyshev(2) smoothing, coarse CG) on --setup-threads host
//                  threads, and solved by --workers solver threads, each on a
//                  stream of its own, so the solves of several requests run on
//                  the GPU at once (CG preconditioned by the V-cycle, FP64;
//                  p = 1 requests, which have no coarser order, take
//                  Jacobi-PCG; the coarse CG is Jacobi-preconditioned).
//                  --precision bf16: the cycle's applies read
//                  BF16 copies of the PA data (FP64 vectors), the outer CG
//                  stays FP64.  The stop test is CGSolver's at rel_tol on the
//                  solver's own (multigrid) measure; each result's Jacobi
//                  measure -- the serial mode's stop measure -- is checked
//                  after the solve and reported (max_jacobi_rel_residual).
//   --solo         (batch mode) admit a request only when no other is live:
//                  every request is solved alone through the batch.
//   --dump DIR     write each request's solution X (the constrained system's,
//                  raw little-endian doubles) to DIR/x_<id>.bin and one line
//                  "id n p coeff_seed rhs_seed iterations converged" per
//                  request to DIR/requests.txt.
//
// Prints one JSON line: solves/s over the whole run, latency percentiles
// (arrival to result), and the request mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <random>
#include <string>
#include <thread>
#include <vector>

#include "../../include/h1pa.h"

namespace {

struct Opts {
    std::string mode = "batch";
    int precision = H1PA_PREC_BF16_PA;
    double rate = 200.0;      // requests per second
    double seconds = 2.0;     // arrival window
    double budget_gb = 200.0;
    int nmin = 6, nmax = 32, pmin = 1, pmax = 8;
    double rel_tol = 1e-10;
    int iters_per_replay = 8;
    int use_graph = 1;        // --no-graph: plain launches (per-kernel profiling)
    uint64_t seed = 1;
    int max_requests = 1 << 30;
    bool solo = false;
    std::string dump;
    int workers = 4;          // mg mode: solver threads (one stream each)
    int setup_threads = 2;    // mg and batch modes: set-up threads
};

struct Request {
    int id = 0;
    double t_arrive = 0.0;
    int n = 0, p = 0;
    uint64_t coeff_seed = 0, rhs_seed = 0;
    double t_done = -1.0;
    int iterations = 0;
    int converged = 0;
    double jacobi_rel = -1.0;  // mg mode: the Jacobi measure of the result
};

struct Live {
    h1pa_mesh *mesh = nullptr;
    h1pa_space *space = nullptr;
    h1pa_form *form = nullptr;
    h1pa_multigrid *mg = nullptr;
    double *b = nullptr, *x = nullptr, *B = nullptr, *X = nullptr;
    int64_t bytes = 0;
};

void check(int rc, const char *what) {
    if (rc != H1PA_OK) {
        std::fprintf(stderr, "h1pa_serve: %s failed (%d): %s\n", what, rc, h1pa_last_error());
        std::exit(1);
    }
}

void hcheck(hipError_t e, const char *what) {
    if (e != hipSuccess) {
        std::fprintf(stderr, "h1pa_serve: %s: %s\n", what, hipGetErrorString(e));
        std::exit(1);
    }
}

// Device bytes a request will hold once admitted (form, space, mesh, batch
// slot and the four request vectors); used before it is built.
int64_t estimate_bytes(int n, int p, int precision, bool batch) {
    const int64_t ne = (int64_t)n * n * n, D = p + 1, Q = p + 2;
    const int64_t N = (int64_t)(n * p + 1) * (n * p + 1) * (n * p + 1);
    const int64_t D3 = D * D * D, Q3 = Q * Q * Q;
    int64_t b = 0;
    b += ne * 6 * Q3 * 8;                    // PA data
    b += ne * D3 * 8;                        // form E scratch
    b += ne * D3 * 4 * 4 + N * 4 * 2;        // gather, indices, indices_nat, elem_dof, offsets, perm
    b += ne * (29 * 4 + 18 + 8 * 3 * 8);     // topology, orientations, nodes
    b += N;                                  // boundary mask
    b += 4 * N * 8;                          // b, x, B, X
    if (batch) {
        if (precision == H1PA_PREC_FP64) {
            b += 4 * N * 8 + ne * D3 * 8 + N;
        } else {
            const int sd = precision == H1PA_PREC_BF16_PA ? 2 : 4;
            b += ne * 6 * Q3 * sd + ne * D3 * 4;     // shadow PA data, FP32 E scratch
            b += 5 * N * 4 + 2 * N * 8 + ne * D3 * 8 + N;
        }
    } else {
        b += 3 * N * 8 + 2 * N * 8;          // PCG vectors and the diagonal
    }
    return b;
}

double now_s(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

std::vector<Request> make_requests(const Opts &o) {
    std::mt19937_64 rng(o.seed);
    std::exponential_distribution<double> gap(o.rate);
    // P(n) proportional to 1/n on [nmin, nmax]
    std::vector<double> w;
    for (int n = o.nmin; n <= o.nmax; ++n) w.push_back(1.0 / n);
    std::discrete_distribution<int> pick_n(w.begin(), w.end());
    std::uniform_int_distribution<int> pick_p(o.pmin, o.pmax);
    std::vector<Request> r;
    double t = 0.0;
    for (int i = 0; i < o.max_requests; ++i) {
        t += gap(rng);
        if (t >= o.seconds) break;
        Request q;
        q.id = i;
        q.t_arrive = t;
        q.n = o.nmin + pick_n(rng);
        q.p = pick_p(rng);
        q.coeff_seed = rng();
        q.rhs_seed = rng();
        r.push_back(q);
    }
    return r;
}

// Multigrid orders of a p-request: p halved (rounding up) down to 1.
std::vector<int> mg_orders(int p) {
    std::vector<int> o{p};
    while (o.back() > 1) o.push_back((o.back() + 1) / 2);
    std::reverse(o.begin(), o.end());
    return o;
}

// Mesh, space, form (assembled), RHS and the constrained system B, X; with
// `mg`, also the request's multigrid (p >= 2).
Live setup(const Request &q, hipStream_t s, bool mg = false,
           int mg_precision = H1PA_PREC_FP64) {
    Live L;
    check(h1pa_mesh_create_cartesian(q.n, q.n, q.n, 1, &L.mesh), "mesh");
    check(h1pa_mesh_perturb(L.mesh, 0.15, q.coeff_seed), "perturb");
    check(h1pa_space_create(L.mesh, q.p, &L.space), "space");
    h1pa_form_desc d{};
    d.integrators = H1PA_INTEG_DIFFUSION;
    d.coeff_kind = H1PA_COEFF_FOURIER;
    d.coeff_seed = q.coeff_seed;
    check(h1pa_form_create(L.space, &d, &L.form), "form");
    check(h1pa_form_assemble(L.form, s), "assemble");
    int64_t info[4];
    check(h1pa_space_info(L.space, info), "space_info");
    const int64_t N = info[3];
    // stream-ordered: the release (on another stream, after every use has
    // been waited for) does not synchronize the device
    hcheck(hipMallocAsync(reinterpret_cast<void **>(&L.b), 4 * N * sizeof(double), s),
           "hipMallocAsync vectors");
    L.x = L.b + N;
    L.B = L.x + N;
    L.X = L.B + N;
    check(h1pa_vector_randomize(L.b, N, q.rhs_seed, -1.0, 1.0, s), "randomize");
    hcheck(hipMemsetAsync(L.x, 0, N * sizeof(double), s), "memset");
    const int *ess;
    int64_t nEss;
    check(h1pa_space_boundary_dofs(L.space, &ess, &nEss), "boundary");
    check(h1pa_form_linear_system(L.form, ess, nEss, L.x, L.b, L.X, L.B, s), "linear_system");
    int64_t fb = 0;
    check(h1pa_form_device_bytes(L.form, &fb), "device_bytes");
    L.bytes = fb + 4 * N * 8;
    if (mg && q.p >= 2) {
        const std::vector<int> ord = mg_orders(q.p);
        // ex26's settings: Chebyshev(2), V(1, 1), coarse CG to sqrt(1e-4)
        // Jacobi on the coarse level: fewer coarse CG iterations (launch-bound)
        const h1pa_mg_opts mo{2, 1, 1, 0, 1e-2, 200, 10, 1e-8, 12345, mg_precision, 1};
        check(h1pa_multigrid_create(L.form, (int)ord.size(), ord.data(), &mo, &L.mg, s),
              "multigrid_create");
        L.bytes += L.bytes / 2;  // coarser levels, smoothers, transfers, work vectors
    }
    return L;
}

// Writes the request's solution X to dir/x_<id>.bin (after its solve).
void dump_solution(const std::string &dir, int id, const Live &L) {
    int64_t info[4];
    check(h1pa_space_info(L.space, info), "space_info");
    std::vector<double> h((size_t)info[3]);
    hcheck(hipMemcpy(h.data(), L.X, h.size() * sizeof(double), hipMemcpyDeviceToHost),
           "dump copy");
    const std::string path = dir + "/x_" + std::to_string(id) + ".bin";
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f || std::fwrite(h.data(), sizeof(double), h.size(), f) != h.size()) {
        std::fprintf(stderr, "h1pa_serve: cannot write %s\n", path.c_str());
        std::exit(1);
    }
    std::fclose(f);
}

// Frees a request's objects; every device use of them has been waited for.
// With the library's block cache on and the vectors stream-ordered, nothing
// here synchronizes the device (the solves of other requests run on).
void release(Live &L, hipStream_t s) {
    if (L.mg) h1pa_multigrid_destroy(L.mg);
    if (L.form) h1pa_form_destroy(L.form);
    if (L.space) h1pa_space_destroy(L.space);
    if (L.mesh) h1pa_mesh_destroy(L.mesh);
    if (L.b) (void)hipFreeAsync(L.b, s);
    L = Live{};
}

// Sets requests up on a host thread of its own, on its own stream, so that
// the loop's replays and solves never wait for a setup: give() queues a
// request, take() returns the finished setups (their device work complete).
class SetupWorker {
  public:
    explicit SetupWorker(const std::vector<Request> &reqs, bool mg = false, int threads = 1,
                         int mg_precision = H1PA_PREC_FP64)
        : reqs_(reqs), mg_(mg), mg_precision_(mg_precision) {
        for (int i = 0; i < std::max(1, threads); ++i) th_.emplace_back([this] { run(); });
    }
    void give(int id) {
        {
            std::lock_guard<std::mutex> g(m_);
            todo_.push_back(id);
        }
        cv_.notify_all();
    }
    std::vector<std::pair<int, Live>> take() {
        std::lock_guard<std::mutex> g(m_);
        std::vector<std::pair<int, Live>> out(ready_.begin(), ready_.end());
        ready_.clear();
        return out;
    }
    void wait_ready(double seconds) {
        std::unique_lock<std::mutex> g(m_);
        cv_.wait_for(g, std::chrono::duration<double>(seconds), [this] { return !ready_.empty(); });
    }
    // Joins the thread; returns its total setup time (seconds).
    double stop() {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : th_)
            if (t.joinable()) t.join();
        return setup_s_;
    }
    ~SetupWorker() { stop(); }

  private:
    void run() {
        int dev = 0;
        hcheck(hipGetDevice(&dev), "get device");
        hcheck(hipSetDevice(dev), "set device");
        hipStream_t ws;
        hcheck(hipStreamCreateWithFlags(&ws, hipStreamNonBlocking), "stream");
        for (;;) {
            int id;
            {
                std::unique_lock<std::mutex> g(m_);
                cv_.wait(g, [this] { return stop_ || !todo_.empty(); });
                if (todo_.empty()) break;  // stop_ and nothing left
                id = todo_.front();
                todo_.pop_front();
            }
            const auto ts = std::chrono::steady_clock::now();
            Live L = setup(reqs_[id], ws, mg_, mg_precision_);
            hcheck(hipStreamSynchronize(ws), "setup sync");
            const double dt =
                std::chrono::duration<double>(std::chrono::steady_clock::now() - ts).count();
            {
                std::lock_guard<std::mutex> g(m_);
                setup_s_ += dt;
                ready_.emplace_back(id, L);
            }
            cv_.notify_all();
        }
        hcheck(hipStreamDestroy(ws), "stream destroy");
    }
    const std::vector<Request> &reqs_;
    const bool mg_;
    const int mg_precision_;
    std::mutex m_;
    std::condition_variable cv_;
    std::deque<int> todo_;
    std::deque<std::pair<int, Live>> ready_;
    bool stop_ = false;
    double setup_s_ = 0.0;
    std::vector<std::thread> th_;
};

// mg mode: solver threads, each with a stream of its own, take set-up
// requests and solve them; finished ones go back to the loop (which releases
// their memory, so no solver thread waits in a free).
struct Solved {
    int id;
    Live L;
    double t_done;
    int iterations, converged;
};

// mg mode: finished solves are checked on the serial mode's measure, dumped
// and released on a host thread of their own (a free waits for the whole
// device, which must stall neither the loop nor the solver threads).
class Reaper {
  public:
    Reaper(const std::string &dump) : dump_(dump), th_([this] { run(); }) {}
    void give(const Solved &sv) {
        {
            std::lock_guard<std::mutex> g(m_);
            todo_.push_back(sv);
        }
        cv_.notify_all();
    }
    // ids released since the last call, with their Jacobi measure
    std::vector<std::pair<int, double>> take() {
        std::lock_guard<std::mutex> g(m_);
        std::vector<std::pair<int, double>> out(done_.begin(), done_.end());
        done_.clear();
        return out;
    }
    void stop() {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
        }
        cv_.notify_all();
        if (th_.joinable()) th_.join();
    }
    ~Reaper() { stop(); }

  private:
    void run();
    const std::string dump_;
    std::mutex m_;
    std::condition_variable cv_;
    std::deque<Solved> todo_;
    std::deque<std::pair<int, double>> done_;
    bool stop_ = false;
    std::thread th_;
};

// The serial mode's stop measure on a result: sqrt((B r, r) / (B b, b)),
// r = B - A X on the constrained operator, B = Jacobi (ess rows 1).
double jacobi_measure(const Live &L, hipStream_t s) {
    const int *ess;
    int64_t nEss;
    check(h1pa_space_boundary_dofs(L.space, &ess, &nEss), "boundary");
    int64_t info[4];
    check(h1pa_space_info(L.space, info), "space_info");
    const int64_t N = info[3];
    double *r = nullptr;
    hcheck(hipMallocAsync(reinterpret_cast<void **>(&r), 2 * N * sizeof(double), s),
           "hipMallocAsync");
    double *dg = r + N;
    check(h1pa_form_constrained_mult(L.form, ess, nEss, L.X, r, s), "residual");
    check(h1pa_vector_add(r, -1.0, L.B, N, s), "residual");
    check(h1pa_form_assemble_diagonal(L.form, dg, s), "diagonal");
    hcheck(hipStreamSynchronize(s), "sync");
    std::vector<double> hr(N), hd(N), hb(N);
    std::vector<int> he(nEss);
    hcheck(hipMemcpy(hr.data(), r, N * 8, hipMemcpyDeviceToHost), "copy");
    hcheck(hipMemcpy(hd.data(), dg, N * 8, hipMemcpyDeviceToHost), "copy");
    hcheck(hipMemcpy(hb.data(), L.B, N * 8, hipMemcpyDeviceToHost), "copy");
    if (nEss) hcheck(hipMemcpy(he.data(), ess, nEss * 4, hipMemcpyDeviceToHost), "copy");
    hcheck(hipFreeAsync(r, s), "hipFreeAsync");
    for (int i : he) hd[i] = 1.0;
    double rr = 0.0, bb = 0.0;
    for (int64_t i = 0; i < N; ++i) {
        rr += hr[i] * hr[i] / hd[i];
        bb += hb[i] * hb[i] / hd[i];
    }
    return bb > 0 ? std::sqrt(rr / bb) : 0.0;
}

void Reaper::run() {
    int dev = 0;
    hcheck(hipGetDevice(&dev), "get device");
    hcheck(hipSetDevice(dev), "set device");
    hipStream_t rs;
    hcheck(hipStreamCreateWithFlags(&rs, hipStreamNonBlocking), "stream");
    for (;;) {
        Solved sv;
        {
            std::unique_lock<std::mutex> g(m_);
            cv_.wait(g, [this] { return stop_ || !todo_.empty(); });
            if (todo_.empty()) break;
            sv = todo_.front();
            todo_.pop_front();
        }
        const double meas = jacobi_measure(sv.L, rs);
        if (!dump_.empty()) dump_solution(dump_, sv.id, sv.L);
        release(sv.L, rs);
        std::lock_guard<std::mutex> g(m_);
        done_.emplace_back(sv.id, meas);
    }
    hcheck(hipStreamDestroy(rs), "stream destroy");
}
class SolverPool {
  public:
    SolverPool(int workers, double rel_tol, std::chrono::steady_clock::time_point t0)
        : rel_tol_(rel_tol), t0_(t0) {
        for (int i = 0; i < std::max(1, workers); ++i) th_.emplace_back([this] { run(); });
    }
    void give(int id, const Live &L) {
        {
            std::lock_guard<std::mutex> g(m_);
            todo_.emplace_back(id, L);
        }
        cv_.notify_all();
    }
    std::vector<Solved> take() {
        std::lock_guard<std::mutex> g(m_);
        std::vector<Solved> out(done_.begin(), done_.end());
        done_.clear();
        return out;
    }
    void wait_done(double seconds) {
        std::unique_lock<std::mutex> g(m_);
        done_cv_.wait_for(g, std::chrono::duration<double>(seconds),
                          [this] { return !done_.empty(); });
    }
    void stop() {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : th_)
            if (t.joinable()) t.join();
    }
    ~SolverPool() { stop(); }

  private:
    void run() {
        int dev = 0;
        hcheck(hipGetDevice(&dev), "get device");
        hcheck(hipSetDevice(dev), "set device");
        hipStream_t ws;
        hcheck(hipStreamCreateWithFlags(&ws, hipStreamNonBlocking), "stream");
        for (;;) {
            std::pair<int, Live> job;
            {
                std::unique_lock<std::mutex> g(m_);
                cv_.wait(g, [this] { return stop_ || !todo_.empty(); });
                if (todo_.empty()) break;
                job = todo_.front();
                todo_.pop_front();
            }
            Live &L = job.second;
            const int *ess;
            int64_t nEss;
            check(h1pa_space_boundary_dofs(L.space, &ess, &nEss), "boundary");
            h1pa_pcg_opts po{rel_tol_, 0.0, 200000, 8};
            h1pa_pcg_stats st{};
            const int rc = L.mg ? h1pa_pcg_solve_multigrid(L.mg, L.B, L.X, &po, &st, ws)
                                : h1pa_pcg_solve(L.form, ess, nEss, L.B, L.X, &po, &st, ws);
            if (rc != H1PA_OK && rc != H1PA_ENOTCONV) check(rc, "solve");
            hcheck(hipStreamSynchronize(ws), "sync");
            const double t_done = std::chrono::duration<double>(
                                      std::chrono::steady_clock::now() - t0_).count();
            Solved sv{job.first, L, t_done, st.iterations, st.converged};
            {
                std::lock_guard<std::mutex> g(m_);
                done_.push_back(sv);
            }
            done_cv_.notify_all();
        }
        hcheck(hipStreamDestroy(ws), "stream destroy");
    }
    const double rel_tol_;
    const std::chrono::steady_clock::time_point t0_;
    std::mutex m_;
    std::condition_variable cv_, done_cv_;
    std::deque<std::pair<int, Live>> todo_;
    std::deque<Solved> done_;
    bool stop_ = false;
    std::vector<std::thread> th_;
};

double pct(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double pos = q * (v.size() - 1);
    const size_t i = (size_t)pos;
    const double f = pos - i;
    return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
}

Opts parse(int argc, char **argv) {
    Opts o;
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        auto val = [&]() -> std::string {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "missing value for %s\n", a.c_str());
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "--mode") o.mode = val();
        else if (a == "--precision") {
            const std::string p = val();
            o.precision = p == "fp64" ? H1PA_PREC_FP64 : p == "fp32" ? H1PA_PREC_FP32
                                                                     : H1PA_PREC_BF16_PA;
        } else if (a == "--rate") o.rate = std::atof(val().c_str());
        else if (a == "--seconds") o.seconds = std::atof(val().c_str());
        else if (a == "--budget-gb") o.budget_gb = std::atof(val().c_str());
        else if (a == "--nmin") o.nmin = std::atoi(val().c_str());
        else if (a == "--nmax") o.nmax = std::atoi(val().c_str());
        else if (a == "--pmin") o.pmin = std::atoi(val().c_str());
        else if (a == "--pmax") o.pmax = std::atoi(val().c_str());
        else if (a == "--rel-tol") o.rel_tol = std::atof(val().c_str());
        else if (a == "--iters-per-replay") o.iters_per_replay = std::atoi(val().c_str());
        else if (a == "--seed") o.seed = std::strtoull(val().c_str(), nullptr, 10);
        else if (a == "--no-graph") o.use_graph = 0;
        else if (a == "--max-requests") o.max_requests = std::atoi(val().c_str());
        else if (a == "--solo") o.solo = true;
        else if (a == "--dump") o.dump = val();
        else if (a == "--workers") o.workers = std::atoi(val().c_str());
        else if (a == "--setup-threads") o.setup_threads = std::atoi(val().c_str());
        else {
            std::fprintf(stderr, "unknown option %s\n", a.c_str());
            std::exit(2);
        }
    }
    if (o.mode != "batch" && o.mode != "serial" && o.mode != "mg") {
        std::fprintf(stderr, "--mode must be batch, serial or mg\n");
        std::exit(2);
    }
    if (o.workers < 1 || o.workers > 16 || o.setup_threads < 1 || o.setup_threads > 8) {
        std::fprintf(stderr, "--workers must be in [1, 16], --setup-threads in [1, 8]\n");
        std::exit(2);
    }
    if (o.nmin < 1 || o.nmax < o.nmin || o.pmin < 1 || o.pmax > 8 || o.pmax < o.pmin ||
        o.rate <= 0 || o.seconds <= 0) {
        std::fprintf(stderr, "bad request-mix options\n");
        std::exit(2);
    }
    return o;
}

}  // namespace

int main(int argc, char **argv) {
    const Opts o = parse(argc, argv);
    std::vector<Request> reqs = make_requests(o);
    const int R = (int)reqs.size();
    hipStream_t s;
    hcheck(hipStreamCreateWithFlags(&s, hipStreamNonBlocking), "stream");
    const int64_t budget = (int64_t)(o.budget_gb * 1e9);
    // Releases must not synchronize the device while other requests solve:
    // the library keeps freed blocks (up to 64 GB) instead of hipFree, and the
    // default pool behind hipMallocAsync keeps what it is given back.
    check(h1pa_set_device_cache(int64_t(64) << 30), "device cache");
    {
        int dev = 0;
        hipMemPool_t pool = nullptr;
        hcheck(hipGetDevice(&dev), "device");
        hcheck(hipDeviceGetDefaultMemPool(&pool, dev), "default pool");
        uint64_t threshold = UINT64_MAX;
        hcheck(hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &threshold),
               "pool threshold");
    }
    const bool batch = o.mode == "batch";
    const bool mgmode = o.mode == "mg";
    // batch: the batch's precision; mg: the cycle's (FP64, or BF16-PA applies)
    const int prec = (batch || mgmode) ? o.precision : H1PA_PREC_FP64;
    if (mgmode && prec == H1PA_PREC_FP32) {
        std::fprintf(stderr, "--mode mg takes --precision fp64 or bf16\n");
        return 2;
    }

    // warm-up: one small solve (first-use costs, code object loading)
    {
        Request w;
        w.n = 4;
        w.p = 2;
        Live L = setup(w, s);
        const int *ess;
        int64_t nEss;
        check(h1pa_space_boundary_dofs(L.space, &ess, &nEss), "boundary");
        h1pa_pcg_opts po{o.rel_tol, 0.0, 10000, 8};
        h1pa_pcg_stats st{};
        check(h1pa_pcg_solve(L.form, ess, nEss, L.B, L.X, &po, &st, s), "warmup");
        hcheck(hipStreamSynchronize(s), "sync");
        release(L, s);
    }
    // The apply-variant autotuner runs on the first form of each (p, Q) with
    // >= 20000 elements.  Run it here, for every order of the mix, before the
    // setup thread and the clock start: the choice is then made on an idle
    // GPU, not timed against replays, and no setup on the worker thread tunes.
    for (int p = o.pmin; p <= o.pmax; ++p) {
        Request w;
        w.n = 28;  // 21952 elements
        w.p = p;
        Live L = setup(w, s, mgmode, prec);
        if (batch && prec != H1PA_PREC_FP64) {
            int64_t info[4];
            check(h1pa_space_info(L.space, info), "space_info");
            float *v = nullptr;
            hcheck(hipMalloc(&v, 2 * info[3] * sizeof(float)), "hipMalloc tune");
            hcheck(hipMemsetAsync(v, 0, 2 * info[3] * sizeof(float), s), "memset");
            check(h1pa_form_mult_f32(L.form, prec, v, v + info[3], s), "tune reduced precision");
            hcheck(hipStreamSynchronize(s), "sync");
            hcheck(hipFree(v), "hipFree tune");
        }
        release(L, s);
    }

    h1pa_batch *B = nullptr;
    if (batch) {
        h1pa_batch_opts bo{};
        bo.precision = prec;
        bo.iters_per_replay = o.iters_per_replay;
        bo.use_graph = o.use_graph;
        bo.autotune = 1;
        bo.delta = 0.1;
        bo.max_outer = 1000;
        bo.max_inner = 200000;
        check(h1pa_batch_create(&bo, &B), "batch_create");
    }

    std::map<int64_t, Live> live;
    std::map<int64_t, int64_t> live_est;
    std::deque<int> queue;
    int64_t used = 0, peak_used = 0;
    int next = 0, done = 0, peak_live = 0, pending = 0;
    double launch_s = 0.0;  // host time inside h1pa_batch_launch (the GPU has no replay queued)
    bool inflight = false;  // batch mode: a replay is running on the GPU
    std::vector<int64_t> tags(4096);
    std::vector<h1pa_batch_result> res(4096);
    SetupWorker W(reqs, mgmode, (mgmode || batch) ? o.setup_threads : 1, prec);
    const auto t0 = std::chrono::steady_clock::now();
    std::unique_ptr<SolverPool> pool;
    std::unique_ptr<Reaper> reaper;
    if (mgmode) {
        pool.reset(new SolverPool(o.workers, o.rel_tol, t0));
        reaper.reset(new Reaper(o.dump));
    }
    int in_solve = 0;  // mg mode: requests handed to the solver threads
    while (done < R) {
        double t = now_s(t0);
        while (next < R && reqs[next].t_arrive <= t) queue.push_back(next++);
        if (batch) {
            // Retire a finished replay, add the solves the setup thread has
            // finished, hand queued requests to it while the HBM budget allows,
            // and start the next replay, so the GPU runs replays back to back
            // while requests are set up beside them.
            auto finish_replay = [&]() {
                int running = 0;
                check(h1pa_batch_wait(B, &running), "batch_wait");
                inflight = false;
                int nd = 0;
                check(h1pa_batch_poll(B, (int)tags.size(), tags.data(), res.data(), &nd),
                      "poll");
                const double tf = now_s(t0);
                for (int i = 0; i < nd; ++i) {
                    Request &q = reqs[tags[i]];
                    q.t_done = tf;
                    q.iterations = res[i].iterations;
                    q.converged = res[i].converged;
                    if (!o.dump.empty()) dump_solution(o.dump, (int)tags[i], live[tags[i]]);
                    release(live[tags[i]], s);
                    live.erase(tags[i]);
                    used -= live_est[tags[i]];
                    live_est.erase(tags[i]);
                    ++done;
                }
            };
            if (inflight) {
                int idle = 0;
                check(h1pa_batch_query(B, &idle), "batch_query");
                if (idle) finish_replay();
            }
            for (auto &r : W.take()) {
                Live &L = r.second;
                const int *ess;
                int64_t nEss;
                check(h1pa_space_boundary_dofs(L.space, &ess, &nEss), "boundary");
                h1pa_pcg_opts po{o.rel_tol, 0.0, 200000, 8};
                int slot = 0;
                check(h1pa_batch_add(B, L.form, ess, nEss, L.B, L.X, &po, r.first, &slot, s),
                      "batch_add");
                const int64_t acc = std::max(live_est[r.first], L.bytes);
                used += acc - live_est[r.first];
                live_est[r.first] = acc;
                live[r.first] = L;
                --pending;
                peak_used = std::max(peak_used, used);
                peak_live = std::max(peak_live, (int)live.size());
            }
            while (!queue.empty()) {
                const Request &q = reqs[queue.front()];
                const int64_t est = estimate_bytes(q.n, q.p, prec, true);
                const bool idle = live.empty() && pending == 0;
                if (o.solo && !idle) break;
                if (!(idle || used + est <= budget)) break;
                used += est;
                live_est[q.id] = est;
                ++pending;
                W.give(queue.front());
                queue.pop_front();
            }
            if (!inflight && !live.empty()) {
                int launched = 0;
                const double tl = now_s(t0);
                check(h1pa_batch_launch(B, &launched), "batch_launch");
                launch_s += now_s(t0) - tl;
                inflight = launched != 0;
                if (!launched) finish_replay();  // every live solve has finished: retire them
                continue;
            }
            if (inflight) {  // wait for the replay (the setup thread keeps working)
                finish_replay();
                continue;
            }
            if (pending > 0) {
                W.wait_ready(0.001);
            } else if (live.empty() && queue.empty() && next < R) {
                const double wait = reqs[next].t_arrive - now_s(t0);
                if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
            }
        } else if (mgmode) {
            // Hand finished setups to the solver threads, retire finished
            // solves (their memory is released here, not on a solver
            // thread), and admit queued requests while the HBM budget allows.
            bool progress = false;
            for (auto &r : W.take()) {
                pool->give(r.first, r.second);
                live[r.first] = r.second;
                --pending;
                ++in_solve;
                progress = true;
            }
            for (auto &sv : pool->take()) {
                Request &q = reqs[sv.id];
                q.t_done = sv.t_done;
                q.iterations = sv.iterations;
                q.converged = sv.converged;
                live.erase(sv.id);
                reaper->give(sv);
                --in_solve;
                progress = true;
            }
            for (auto &r : reaper->take()) {  // released: the budget frees up
                reqs[r.first].jacobi_rel = r.second;
                used -= live_est[r.first];
                live_est.erase(r.first);
                ++done;
                progress = true;
            }
            while (!queue.empty()) {
                const Request &q = reqs[queue.front()];
                const int64_t est = estimate_bytes(q.n, q.p, H1PA_PREC_FP64, false) * 3 / 2;
                const bool idle = pending == 0 && in_solve == 0;
                if (!(idle || used + est <= budget)) break;
                used += est;
                live_est[q.id] = est;
                ++pending;
                W.give(queue.front());
                queue.pop_front();
                progress = true;
            }
            peak_used = std::max(peak_used, used);
            peak_live = std::max(peak_live, pending + in_solve);
            if (!progress) {
                if (pending + in_solve > 0 || !live_est.empty()) {
                    pool->wait_done(0.0005);
                } else if (next < R) {
                    const double wait = reqs[next].t_arrive - now_s(t0);
                    if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
                }
            }
        } else {
            // One solve at a time; the setup thread prepares the next request
            // while the current one solves.
            if (pending == 0 && !queue.empty()) {
                W.give(queue.front());
                queue.pop_front();
                ++pending;
            }
            if (pending == 0) {
                if (next < R) {
                    const double wait = reqs[next].t_arrive - now_s(t0);
                    if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
                }
                continue;
            }
            auto got = W.take();
            if (got.empty()) {
                W.wait_ready(0.01);
                continue;
            }
            --pending;
            const int id = got.front().first;
            Live L = got.front().second;
            if (!queue.empty()) {  // set the next request up during this solve
                W.give(queue.front());
                queue.pop_front();
                ++pending;
            }
            Request &q = reqs[id];
            peak_used = std::max(peak_used, L.bytes);
            peak_live = 1;
            const int *ess;
            int64_t nEss;
            check(h1pa_space_boundary_dofs(L.space, &ess, &nEss), "boundary");
            h1pa_pcg_opts po{o.rel_tol, 0.0, 200000, 8};
            h1pa_pcg_stats st{};
            const int rc = h1pa_pcg_solve(L.form, ess, nEss, L.B, L.X, &po, &st, s);
            if (rc != H1PA_OK && rc != H1PA_ENOTCONV) check(rc, "pcg_solve");
            hcheck(hipStreamSynchronize(s), "sync");
            q.t_done = now_s(t0);
            q.iterations = st.iterations;
            q.converged = st.converged;
            if (!o.dump.empty()) dump_solution(o.dump, id, L);
            release(L, s);
            ++done;
        }
    }
    if (pool) pool->stop();
    if (reaper) reaper->stop();
    const double setup_s = W.stop();
    const double makespan = now_s(t0);
    if (B) h1pa_batch_destroy(B);
    hcheck(hipStreamDestroy(s), "stream destroy");

    if (!o.dump.empty()) {
        const std::string path = o.dump + "/requests.txt";
        FILE *f = std::fopen(path.c_str(), "w");
        if (!f) {
            std::fprintf(stderr, "h1pa_serve: cannot write %s\n", path.c_str());
            return 1;
        }
        for (const Request &q : reqs)
            std::fprintf(f, "%d %d %d %llu %llu %d %d\n", q.id, q.n, q.p,
                         (unsigned long long)q.coeff_seed, (unsigned long long)q.rhs_seed,
                         q.iterations, q.converged);
        std::fclose(f);
    }
    std::vector<double> lat;
    int64_t dofs = 0, conv = 0, iters = 0;
    double max_jacobi = -1.0;  // mg mode: worst result on the serial mode's measure
    for (const Request &q : reqs) max_jacobi = std::max(max_jacobi, q.jacobi_rel);
    double dof_iters = 0.0;  // sum of DOFs x iterations: the solver work of the run
    for (const Request &q : reqs) {
        lat.push_back(q.t_done - q.t_arrive);
        const int64_t m = (int64_t)q.n * q.p + 1;
        dofs += m * m * m;
        conv += q.converged;
        iters += q.iterations;
        dof_iters += (double)(m * m * m) * q.iterations;
    }
    std::printf(
        "{\"mode\": \"%s\", \"precision\": \"%s\", \"rate\": %.1f, \"window_s\": %.2f, "
        "\"requests\": %d, \"converged\": %lld, \"makespan_s\": %.3f, \"solves_per_s\": %.3f, "
        "\"latency_p50_s\": %.4f, \"latency_p90_s\": %.4f, \"latency_p99_s\": %.4f, "
        "\"latency_max_s\": %.4f, \"mean_dofs\": %.0f, \"mean_iterations\": %.1f, "
        "\"peak_live\": %d, \"peak_bytes\": %lld, \"budget_bytes\": %lld, "
        "\"setup_s\": %.3f, \"launch_s\": %.3f, \"dof_iterations\": %.4e, "
        "\"max_jacobi_rel_residual\": %.3e, \"workers\": %d}\n",
        o.mode.c_str(), prec == H1PA_PREC_FP64 ? "fp64" : prec == H1PA_PREC_FP32 ? "fp32" : "bf16",
        o.rate, o.seconds, R, (long long)conv, makespan, R / makespan, pct(lat, 0.5),
        pct(lat, 0.9), pct(lat, 0.99), pct(lat, 1.0), R ? (double)dofs / R : 0.0,
        R ? (double)iters / R : 0.0, peak_live, (long long)peak_used, (long long)budget, setup_s, launch_s,
        dof_iters, max_jacobi, mgmode ? o.workers : 1);
    return 0;
}
