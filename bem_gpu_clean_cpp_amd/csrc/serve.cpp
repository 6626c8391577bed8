// Multi-frequency batches and the request server.
//
// Both run several resident solvers at once, each driven by its own host
// thread on its own HIP stream.  The assembly of one problem (VALU-bound) then
// overlaps the solve of another (HBM-bound matrix products plus host-side
// Givens rotations), and the host round trips of one solver's Krylov loop hide
// behind the device work of the others.
//
// bem_batch_solve: one mesh, many wavenumbers, right-hand sides shared by all
// frequencies or given per frequency; the workers keep their mesh and matrix
// buffers for all frequencies they take.
//
// bem_server: a request loop with continuous admission.  Requests (own mesh,
// wavenumber, right-hand sides) queue in arrival order; a free slot takes the
// head of the queue as soon as its HBM footprint fits in what the running
// requests leave of the budget, so requests that arrive while others run start
// without waiting for a batch boundary.  Every slot frees its device buffers
// (stream-ordered, into the memory pool) when its request ends, so the budget
// accounts for exactly the running requests.  Results do not depend on what
// else runs: every kernel's launch shape and reduction order follow from the
// request's own sizes.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "bem_internal.hpp"
#include "handles.hpp"

using namespace bem;

namespace {

double now_ms(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

bem_solve_params default_params() {
    bem_solve_params p;
    p.method = BEM_SOLVE_GMRES;
    p.restart = 30;
    p.tol = 1e-6f;
    p.maxit = 3000;
    p.refine_tol = 1e-8;
    p.max_refine = 50;
    return p;
}

// one solve on a resident solver: assemble, right-hand sides, solve, copy out
int run_one(bem_solver *sv, float k, int precision, int kind, const float *params, int nrhs, float amp,
            const bem_solve_params &sp, bem_zcomplex *x, int ldx, bem_solve_stats *st) {
    HOST_CALL(bem_solver_assemble(sv, k, precision));
    HOST_CALL(bem_solver_set_rhs(sv, kind, params, nrhs, amp));
    HOST_CALL(bem_solver_solve(sv, sp.method, sp.restart, sp.tol, sp.maxit, sp.refine_tol, sp.max_refine, st));
    if (x) HOST_CALL(bem_solver_get_solution_z(sv, x, ldx));
    return EXIT_SUCCESS;
}

// Frequencies fused into one lockstep GMRES (opt-in, BEM_BATCH_FUSE = group
// size): square FP32 systems (no CHIEF rows) solved with GMRES take the
// frequencies in groups of consecutive ones of the batch order, all group
// members' matrices resident side by side, so each Arnoldi step is one product
// launch over all of them and one host wait per chunk for all their systems.
// The groups depend only on the wavenumbers and nrhs, not on the worker count.
// Not the default: on the 64-frequency s4 batch it measured 0.163-0.168 s
// against 0.142 s for one frequency per worker (a group runs as many steps as
// its slowest member, and each worker's host loop rotates 4x the systems).
constexpr int kFuseMax = 4;
constexpr int kMaxBatchWorkers = 8;

int fuse_width(const Mesh &m, int precision, int nrhs, const bem_solve_params &p, int workers) {
    const char *env = getenv("BEM_BATCH_FUSE");
    const int want = env ? std::max(1, std::min(kFuseMax, atoi(env))) : 1;
    if (want < 2 || precision != BEM_PREC_FP32 || p.method != BEM_SOLVE_GMRES || nrhs < 2 || m.C() > 0) return 1;
    int fuse = std::max(1, std::min(want, kMaxRhs / nrhs));
    // every worker holds fuse matrices: keep that within half the free HBM
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return 1;
    const size_t ld = ((size_t)m.N() + 63) / 64 * 64;
    while (fuse > 1 && (size_t)fuse * workers * sizeof(float2) * ld * (size_t)m.N() > free_b / 2) fuse--;
    return fuse;
}

// a worker's buffers for fused groups
struct FusedWS {
    float2 *A = nullptr;
    double2 *B = nullptr, *X = nullptr;
    GmresCache *gc = nullptr;
    hipEvent_t ev[2 * kFuseMax + 2] = {};
    hipStream_t st = nullptr;
    ~FusedWS() {
        if (gc) gmres_cache_free(gc);
        sfree(A, st);
        sfree(B, st);
        sfree(X, st);
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

// frequencies fs[0..nf) on the worker's solver: assemble each into its slot,
// its right-hand sides, one GMRES over all nf * nrhs systems, copy out
int run_fused(bem_solver *sv, FusedWS &fw, const float *k, const int *fs, int nf, int kind, const float *params,
              int params_per_freq, int nrhs, float amp, const bem_solve_params &p, bem_zcomplex *X, int ldx,
              bem_solve_stats *stats) {
    std::lock_guard<std::mutex> lk(sv->mu);
    DeviceProblem &dp = sv->dp;
    hipStream_t s = sv->stream;
    fw.st = s;
    const long ld = dp.ldA, astride = ld * (long)dp.N;
    const int R = nf * nrhs;
    if (!fw.A) {
        HOST_CALL(salloc(&fw.A, (size_t)kFuseMax * astride, s));
        HOST_CALL(salloc(&fw.B, (size_t)ld * kFuseMax * nrhs, s));
        HOST_CALL(salloc(&fw.X, (size_t)ld * kFuseMax * nrhs, s));
        for (auto &e : fw.ev) HIP_CALL(hipEventCreate(&e));
    }
    for (int q = 0; q < nf; q++) {
        const int f = fs[q];
        HIP_CALL(hipEventRecord(fw.ev[2 * q], s));
        HOST_CALL(assemble_fp32(dp, k[f], fw.A + (size_t)q * astride, s));
        HIP_CALL(hipEventRecord(fw.ev[2 * q + 1], s));
        RhsSpec rs;
        rs.kind = kind;
        rs.nrhs = nrhs;
        rs.amp = amp;
        const float *pf = params_per_freq ? params + (size_t)f * 3 * nrhs : params;
        rs.params.assign(pf, pf + 3 * (size_t)nrhs);
        HOST_CALL(rhs_fp64(dp, (double)k[f], fw.B + (size_t)q * nrhs * ld, ld, rs, s));
    }
    HIP_CALL(hipMemsetAsync(fw.X, 0, sizeof(double2) * ld * R, s));
    bem_solve_stats st{};
    st.method = BEM_SOLVE_GMRES;
    const double outer = p.refine_tol > 0 ? p.refine_tol : (double)p.tol;
    const int cycles = p.max_refine > 0 ? p.max_refine : (1 << 20);
    HIP_CALL(hipEventRecord(fw.ev[2 * kFuseMax], s));
    HOST_CALL(gmres_device(fw.A, nullptr, ld, dp.M, dp.N, fw.B, ld, fw.X, ld, true, R, p.restart, (double)p.tol,
                           p.maxit > 0 ? p.maxit : 100000, outer, cycles, &st, s, &fw.gc, nf > 1 ? astride : 0, nf));
    HIP_CALL(hipEventRecord(fw.ev[2 * kFuseMax + 1], s));
    if (X) {
        for (int q = 0; q < nf; q++)
            HIP_CALL(hipMemcpy2DAsync(X + (size_t)fs[q] * nrhs * ldx, sizeof(bem_zcomplex) * ldx,
                                      fw.X + (size_t)q * nrhs * ld, sizeof(double2) * ld, sizeof(double2) * dp.N,
                                      nrhs, hipMemcpyDeviceToHost, s));
    }
    HIP_CALL(hipStreamSynchronize(s));
    float sms = 0;
    HIP_CALL(hipEventElapsedTime(&sms, fw.ev[2 * kFuseMax], fw.ev[2 * kFuseMax + 1]));
    st.solve_ms = sms;
    if (stats) {
        for (int q = 0; q < nf; q++) {
            float ams = 0;
            HIP_CALL(hipEventElapsedTime(&ams, fw.ev[2 * q], fw.ev[2 * q + 1]));
            stats[fs[q]] = st;
            stats[fs[q]].assemble_ms = ams;
        }
    }
    return EXIT_SUCCESS;
}

}  // namespace

extern "C" {

double bem_request_footprint(int M, int N, int E, int nrhs, int precision, int method, int restart) {
    if (M < N || N <= 0 || E <= 0 || nrhs <= 0) return -1.0;
    const size_t ld = ((size_t)M + 63) / 64 * 64;
    const size_t ldv = ((size_t)std::max(M, N) + 63) / 64 * 64;
    const size_t R = (size_t)nrhs;
    const size_t ea = precision == BEM_PREC_FP64 ? sizeof(double2) : sizeof(float2);
    size_t b = device_problem_bytes(M, N, E);
    b += ea * ld * N;                                               // A
    b += (2 * sizeof(double2) + sizeof(float2)) * ld * R;           // B, X (FP64), X (FP32)
    if (method == BEM_SOLVE_QR) {
        b += ea * ld * N;                                           // factored copy of A
        b += sizeof(float2) * 32 * (ld + 32 * ((N + 31) / 32));     // block reflectors, T factors
        b += 2 * sizeof(double2) * ld + ws_gemv_n(M, N, 1);
    } else {
        const size_t K = (size_t)std::max(1, std::min(restart, N));
        b += sizeof(double2) * ldv * (K + 1) * R;                   // Krylov basis
        b += (3 * sizeof(double2) + 3 * sizeof(float2)) * ldv * R;  // work vectors
        b += std::max(std::max(ws_gemv_n(M, N, 1), ws_gemm(M, N, nrhs)), ws_mres(M, N, nrhs));
        b += sizeof(double2) * (K + 2) * (2 + ((size_t)std::max(M, N) + 1023) / 1024) * R;
        if (method == BEM_SOLVE_GMRES && precision != BEM_PREC_FP64 && nrhs == 1 && M > N && ne_fused_fits(M, N))
            b += sizeof(float2) * (size_t)M * ne_ldr(N) + std::max(ws_ne_resid(M, N), ws_ne_rhs(M, N));  // row-major copy of A
        // conjugate-transposed copy of A (block CG makes it for its FP16-split
        // products at any count above one)
        if (precision != BEM_PREC_FP64 && nrhs > 1 &&
            ((M > N && adjoint_copy_pays(M, N, nrhs)) || method == BEM_SOLVE_BLOCK_CG))
            b += adjoint_copy_bytes(M, N);
        if (method == BEM_SOLVE_BLOCK_CG)  // block vectors, Gram partials, the split's planes and bound
            b += (6 * sizeof(double2) + sizeof(float2)) * ldv * R + sizeof(double2) * 4096 * ((N + 63) / 64 + 21) +
                 ws_split_x(std::max(M, N), nrhs) + ws_adjoint_absmax(M, N);
    }
    b += sizeof(float2) * R * ld;  // right-hand-side partials
    return (double)b * 1.05 + (1 << 20);
}

// ---------------------------------------------------------- batch solve ---
// A batch context: `workers` resident solvers of one mesh, kept between runs,
// so each run finds the mesh's device records, the matrices and the solvers'
// workspaces allocated (at the s4 batching size, creating 4 solvers took
// 45 ms, their first solves 12.6 ms against 4.9 ms warm, and destroying them
// 18 ms: a quarter of a 64-frequency run).  One run at a time.
struct bem_batch {
    std::vector<bem_solver *> sv;
    std::mutex mu;
    ~bem_batch() {
        for (auto *s : sv) bem_solver_destroy(s);
    }
};

bem_batch *bem_batch_create(const bem_mesh *m, int workers) {
    if (!m) {
        set_error(__FILE__, __LINE__, "null mesh");
        return nullptr;
    }
    if (workers < 1 || workers > kMaxBatchWorkers) {
        set_error(__FILE__, __LINE__, "batch: %d workers (1 to %d)", workers, kMaxBatchWorkers);
        return nullptr;
    }
    auto *b = new bem_batch();
    for (int w = 0; w < workers; w++) {
        bem_solver *s = bem_solver_create(m);
        if (!s) {
            const std::string e = last_error();
            delete b;
            set_error(__FILE__, __LINE__, "batch: solver creation failed: %s", e.c_str());
            return nullptr;
        }
        b->sv.push_back(s);
    }
    return b;
}

void bem_batch_destroy(bem_batch *b) { delete b; }

int bem_batch_run(bem_batch *b, const float *k, int nk, int precision, int rhs_kind, const float *params, int nrhs,
                  int params_per_freq, float amp, const bem_solve_params *sp, bem_zcomplex *X, int ldx,
                  bem_solve_stats *stats) {
    if (!b || b->sv.empty() || !k || nk <= 0 || !params || nrhs <= 0) BEM_FAIL("bad batch arguments");
    std::lock_guard<std::mutex> lk(b->mu);
    const Mesh &mesh = b->sv[0]->mesh;
    if (X && ldx < (int)mesh.N()) BEM_FAIL("ldx must be >= N");
    const bem_solve_params p = sp ? *sp : default_params();
    const int workers = std::min((int)b->sv.size(), nk);
    const std::vector<bem_solver *> &sv = b->sv;
    // highest wavenumbers first: their solves take the most iterations, so
    // starting them first keeps the workers' finishing times close
    std::vector<int> order(nk);
    for (int f = 0; f < nk; f++) order[f] = f;
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) { return k[a] > k[c]; });
    const int fuse = fuse_width(mesh, precision, nrhs, p, workers);
    std::atomic<int> next{0};
    std::atomic<int> failed{0};
    std::string first_error;
    std::mutex emu;
    auto work = [&](int w) {
        if (fuse > 1) {
            FusedWS fw;
            for (;;) {
                const int g = next.fetch_add(1);
                const int q0 = g * fuse;
                if (q0 >= nk || failed.load()) break;
                const int nf = std::min(fuse, nk - q0);
                if (run_fused(sv[w], fw, k, order.data() + q0, nf, rhs_kind, params, params_per_freq, nrhs, amp, p, X,
                              ldx, stats) != EXIT_SUCCESS) {
                    std::lock_guard<std::mutex> lk2(emu);
                    if (!failed.exchange(1)) first_error = last_error();
                    break;
                }
            }
            return;
        }
        for (;;) {
            const int q = next.fetch_add(1);
            if (q >= nk || failed.load()) break;
            const int f = order[q];
            bem_zcomplex *xf = X ? X + (size_t)f * nrhs * ldx : nullptr;
            bem_solve_stats st{};
            const float *pf = params_per_freq ? params + (size_t)f * 3 * nrhs : params;
            if (run_one(sv[w], k[f], precision, rhs_kind, pf, nrhs, amp, p, xf, ldx, &st) != EXIT_SUCCESS) {
                std::lock_guard<std::mutex> lk2(emu);
                if (!failed.exchange(1)) first_error = last_error();
                break;
            }
            if (stats) stats[f] = st;
        }
    };
    std::vector<std::thread> th;
    for (int w = 1; w < workers; w++) th.emplace_back(work, w);
    work(0);
    for (auto &t : th) t.join();
    if (failed.load()) BEM_FAIL("batch: %s", first_error.c_str());
    return EXIT_SUCCESS;
}

int bem_batch_solve(const bem_mesh *m, const float *k, int nk, int precision, int rhs_kind, const float *params,
                    int nrhs, int params_per_freq, float amp, const bem_solve_params *sp, int workers,
                    bem_zcomplex *X, int ldx, bem_solve_stats *stats) {
    if (!m || !k || nk <= 0 || !params || nrhs <= 0) BEM_FAIL("bad batch arguments");
    if (X && ldx < (int)m->m.N()) BEM_FAIL("ldx must be >= N");
    workers = std::max(1, std::min(std::min(workers, nk), kMaxBatchWorkers));
    std::unique_ptr<bem_batch> b(bem_batch_create(m, workers));
    if (!b) return EXIT_FAILURE;
    return bem_batch_run(b.get(), k, nk, precision, rhs_kind, params, nrhs, params_per_freq, amp, sp, X, ldx, stats);
}

}  // extern "C"

// --------------------------------------------------------- request server ---
namespace {

struct Request {
    long id = 0;
    bem_mesh mesh;
    float k = 0.f;
    int precision = BEM_PREC_FP32;
    int kind = BEM_RHS_PLANE;
    std::vector<float> params;
    int nrhs = 0;
    float amp = 1.f;
    bem_solve_params sp{};
    double footprint = 0;
    int state = 0;  // 0 queued, 1 running, 2 done, 3 failed, 4 cancelled
    int slot = -1;
    double t_submit = 0, t_start = 0, t_end = 0;
    std::vector<bem_zcomplex> x;
    bem_solve_stats st{};
    std::string error;
};

}  // namespace

struct bem_server {
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<std::shared_ptr<Request>> queue;
    std::unordered_map<long, std::shared_ptr<Request>> reqs;  // until collected by bem_server_wait
    double budget = 0;
    double in_use = 0;
    double peak_in_use = 0;
    long next_id = 1;
    int running = 0;
    int max_running = 0;
    int waiters = 0;  // threads inside bem_server_wait
    bool stop = false;
    std::vector<std::thread> threads;
    std::vector<bem_solver *> slots;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    void worker(int i) {
        bem_solver *&sv = slots[i];
        for (;;) {
            std::shared_ptr<Request> r;
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] {
                    return stop || (!queue.empty() && in_use + queue.front()->footprint <= budget);
                });
                if (stop) return;
                r = queue.front();
                queue.pop_front();
                in_use += r->footprint;
                peak_in_use = std::max(peak_in_use, in_use);
                running++;
                max_running = std::max(max_running, running);
                r->state = 1;
                r->slot = i;
                r->t_start = now_ms(t0);
            }
            int rc = EXIT_SUCCESS;
            if (!sv) {
                sv = bem_solver_create(&r->mesh);
                if (!sv) rc = EXIT_FAILURE;
            } else {
                rc = bem_solver_set_mesh(sv, &r->mesh);
            }
            if (rc == EXIT_SUCCESS) {
                r->x.assign((size_t)r->mesh.m.N() * r->nrhs, bem_zcomplex{0, 0});
                rc = run_one(sv, r->k, r->precision, r->kind, r->params.data(), r->nrhs, r->amp, r->sp, r->x.data(),
                             r->mesh.m.N(), &r->st);
            }
            if (rc != EXIT_SUCCESS) r->error = last_error();
            if (sv) (void)bem_solver_release(sv);  // the footprint goes back to the budget
            {
                std::lock_guard<std::mutex> lk(mu);
                in_use -= r->footprint;
                running--;
                r->t_end = now_ms(t0);
                r->state = rc == EXIT_SUCCESS ? 2 : 3;
            }
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }
};

extern "C" {

bem_server *bem_server_create(int slots, double hbm_budget_bytes) {
    Context *c = nullptr;
    if (ctx_init(&c) != EXIT_SUCCESS) return nullptr;
    if (slots <= 0 || slots > 16) {
        set_error(__FILE__, __LINE__, "slots must be in [1, 16]");
        return nullptr;
    }
    size_t free_b = 0, total_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) {
        set_error(__FILE__, __LINE__, "hipMemGetInfo failed");
        return nullptr;
    }
    auto *s = new bem_server();
    s->budget = hbm_budget_bytes > 0 ? std::min(hbm_budget_bytes, 0.95 * (double)free_b) : 0.9 * (double)free_b;
    s->slots.assign(slots, nullptr);
    for (int i = 0; i < slots; i++) s->threads.emplace_back(&bem_server::worker, s, i);
    return s;
}

long bem_server_submit(bem_server *s, const bem_mesh *m, float k, int precision, int rhs_kind, const float *params,
                       int nrhs, float amp, const bem_solve_params *sp) {
    if (!s || !m || !params || nrhs <= 0 || nrhs > kMaxRhs) {
        set_error(__FILE__, __LINE__, "bad request arguments (nrhs must be in [1, %d])", kMaxRhs);
        return -1;
    }
    if (precision != BEM_PREC_FP32 && precision != BEM_PREC_FP64) {
        set_error(__FILE__, __LINE__, "precision must be 32 or 64");
        return -1;
    }
    if (validate_mesh(m->m) != EXIT_SUCCESS) return -1;
    auto r = std::make_shared<Request>();
    r->mesh.m = m->m;
    r->k = k;
    r->precision = precision;
    r->kind = rhs_kind;
    r->params.assign(params, params + 3 * (size_t)nrhs);
    r->nrhs = nrhs;
    r->amp = amp;
    r->sp = sp ? *sp : default_params();
    const int N = m->m.N(), M = N + m->m.C(), E = m->m.E();
    r->footprint = bem_request_footprint(M, N, E, nrhs, precision, r->sp.method, r->sp.restart);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (r->footprint > s->budget) {
            set_error(__FILE__, __LINE__, "request needs %.3g bytes of HBM, more than the budget %.3g", r->footprint,
                      s->budget);
            return -1;
        }
        r->id = s->next_id++;
        r->t_submit = now_ms(s->t0);
        s->reqs[r->id] = r;
        s->queue.push_back(r);
    }
    s->cv_work.notify_all();
    return r->id;
}

int bem_server_wait(bem_server *s, long id, bem_zcomplex *x, int ldx, bem_request_info *info) {
    if (!s) BEM_FAIL("null server");
    std::shared_ptr<Request> r;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        auto it = s->reqs.find(id);
        if (it == s->reqs.end()) BEM_FAIL("unknown request %ld", id);
        r = it->second;
        // counted while inside: bem_server_destroy frees the server only
        // after every waiter has left it
        s->waiters++;
        s->cv_done.wait(lk, [&] { return r->state >= 2; });
        s->reqs.erase(id);  // by key: submissions during the wait may rehash
        if (--s->waiters == 0) s->cv_done.notify_all();
    }
    // nothing below touches the server
    const int N = r->mesh.m.N();
    if (info) {
        info->id = r->id;
        info->status = r->state == 2 ? 0 : 1;
        info->slot = r->slot;
        info->N = N;
        info->M = N + r->mesh.m.C();
        info->nrhs = r->nrhs;
        info->footprint_bytes = r->footprint;
        info->submit_ms = r->t_submit;
        info->start_ms = r->t_start;
        info->end_ms = r->t_end;
        info->solve = r->st;
    }
    if (r->state != 2) BEM_FAIL("request %ld failed: %s", id, r->error.c_str());
    if (x) {
        if (ldx < N) BEM_FAIL("ldx must be >= N (%d)", N);
        for (int q = 0; q < r->nrhs; q++)
            std::copy(r->x.begin() + (size_t)q * N, r->x.begin() + (size_t)(q + 1) * N, x + (size_t)q * ldx);
    }
    return EXIT_SUCCESS;
}

int bem_server_status(bem_server *s, bem_server_info *info) {
    if (!s || !info) BEM_FAIL("null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    info->queued = (int)s->queue.size();
    info->running = s->running;
    info->max_running = s->max_running;
    info->slots = (int)s->slots.size();
    info->budget_bytes = s->budget;
    info->in_use_bytes = s->in_use;
    info->peak_in_use_bytes = s->peak_in_use;
    return EXIT_SUCCESS;
}

void bem_server_destroy(bem_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        // requests still queued are cancelled; running ones finish first
        for (auto &r : s->queue) r->state = 4;
        s->queue.clear();
    }
    s->cv_done.notify_all();  // waiters of cancelled requests return
    {
        // the server is freed only after the running requests have ended and
        // every thread inside bem_server_wait has left it
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_done.wait(lk, [&] { return s->running == 0 && s->waiters == 0; });
        s->stop = true;
    }
    s->cv_work.notify_all();
    s->cv_done.notify_all();
    for (auto &t : s->threads) t.join();
    for (auto *sv : s->slots) bem_solver_destroy(sv);
    delete s;
}

}  // extern "C"
