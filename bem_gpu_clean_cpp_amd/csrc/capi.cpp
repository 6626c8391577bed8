// extern "C" boundary of libbem (declared in include/bem.h).
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <mutex>

#include "bem_internal.hpp"
#include "handles.hpp"

using namespace bem;

namespace {

std::mutex g_mu;  // the library runs on one stream; serialise host entry points

int stream(hipStream_t *s) {
    Context *c = nullptr;
    HOST_CALL(ctx_init(&c));
    *s = c->stream;
    return EXIT_SUCCESS;
}
int stream_of_library(hipStream_t *s) { return stream(s); }

int mesh_device(bem_mesh *m, hipStream_t s) {
    if (m->dp_version != m->m.version || m->dp.M == 0) {
        HOST_CALL(device_problem_build(m->m, m->dp, s));
        m->dp_version = m->m.version;
    }
    return EXIT_SUCCESS;
}

// grow-only device scratch for the host-buffer entry points
struct Scratch {
    void *p = nullptr;
    size_t n = 0;
    int get(size_t bytes, void **out) {
        if (bytes > n) {
            if (p) (void)hipFree(p);
            p = nullptr;
            n = 0;
            HIP_CALL(hipMalloc(&p, bytes));
            n = bytes;
        }
        *out = p;
        return EXIT_SUCCESS;
    }
};
Scratch g_sa, g_sb, g_sc;

long round64(long m) { return (m + 63) / 64 * 64; }

template <typename C2>
__global__ void k_gather_rows(const C2 *A, long lda, int N, const int *rows, int nr, double2 *out) {
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nr * N) return;
    const int q = (int)(t / N), j = (int)(t % N);
    const C2 v = A[(size_t)j * lda + rows[q]];
    out[(size_t)q * N + j] = make_double2(v.x, v.y);
}

}  // namespace

extern "C" {

// ------------------------------------------------------------------ mesh ---
bem_mesh *bem_mesh_create(void) { return new bem_mesh(); }

void bem_mesh_destroy(bem_mesh *m) {
    if (!m) return;
    device_problem_free(m->dp);
    delete m;
}

int bem_mesh_read_obj(bem_mesh *m, const char *file) {
    if (!m) BEM_FAIL("null mesh");
    return read_obj(m->m, file);
}

int bem_mesh_set(bem_mesh *m, const bem_cart_coord *pnts, int numPnts, const bem_tri_elem *elems, int numElems) {
    if (!m || (numPnts > 0 && !pnts) || (numElems > 0 && !elems) || numPnts < 0 || numElems < 0)
        BEM_FAIL("bad mesh arguments");
    m->m.pnts.assign((const float *)pnts, (const float *)pnts + 3 * (size_t)numPnts);
    m->m.elems.assign(elems, elems + numElems);
    m->m.chief.clear();
    m->m.bb_set = false;
    m->m.version++;
    return validate_mesh(m->m);
}

int bem_mesh_set_bc(bem_mesh *m, bem_complex ca, bem_complex cb, bem_complex cc) {
    if (!m) BEM_FAIL("null mesh");
    if (cb.x == 0.f && cb.y == 0.f) BEM_FAIL("cb must be nonzero (the reference divides by it)");
    for (auto &e : m->m.elems) {
        e.bc[0] = ca;
        e.bc[1] = cb;
        e.bc[2] = cc;
    }
    m->m.version++;
    return EXIT_SUCCESS;
}

int bem_mesh_find_bb(bem_mesh *m, float threshold) {
    if (!m) BEM_FAIL("null mesh");
    return find_bb(m->m, threshold);
}

int bem_mesh_get_bb(const bem_mesh *m, float *bb6) {
    if (!m || !bb6) BEM_FAIL("null argument");
    if (!m->m.bb_set) BEM_FAIL("bounding box not computed (call bem_mesh_find_bb)");
    memcpy(bb6, m->m.bb, sizeof(float) * 6);
    return EXIT_SUCCESS;
}

int bem_mesh_gen_chief(bem_mesh *m, int num, float threshold, unsigned long long seed) {
    if (!m) BEM_FAIL("null mesh");
    return gen_chief(m->m, num, threshold, seed);
}

int bem_mesh_set_chief(bem_mesh *m, const bem_cart_coord *pnts, int num) {
    if (!m || num < 0 || (num > 0 && !pnts)) BEM_FAIL("bad CHIEF arguments");
    m->m.chief.assign((const float *)pnts, (const float *)pnts + 3 * (size_t)num);
    m->m.version++;
    return EXIT_SUCCESS;
}

int bem_mesh_num_pnts(const bem_mesh *m) { return m ? m->m.N() : 0; }
int bem_mesh_num_elems(const bem_mesh *m) { return m ? m->m.E() : 0; }
int bem_mesh_num_chief(const bem_mesh *m) { return m ? m->m.C() : 0; }

int bem_mesh_get_pnts(const bem_mesh *m, bem_cart_coord *out) {
    if (!m || !out) BEM_FAIL("null argument");
    memcpy(out, m->m.pnts.data(), m->m.pnts.size() * sizeof(float));
    return EXIT_SUCCESS;
}
int bem_mesh_get_elems(const bem_mesh *m, bem_tri_elem *out) {
    if (!m || !out) BEM_FAIL("null argument");
    memcpy(out, m->m.elems.data(), m->m.elems.size() * sizeof(bem_tri_elem));
    return EXIT_SUCCESS;
}
int bem_mesh_get_chief(const bem_mesh *m, bem_cart_coord *out) {
    if (!m || (!out && m->m.C() > 0)) BEM_FAIL("null argument");
    if (m->m.C() > 0) memcpy(out, m->m.chief.data(), m->m.chief.size() * sizeof(float));
    return EXIT_SUCCESS;
}

int bem_mesh_gen_icosphere(bem_mesh *m, int subdiv, float radius) {
    if (!m) BEM_FAIL("null mesh");
    return gen_icosphere(m->m, subdiv, radius);
}

int bem_mesh_gen_ellipsoid(bem_mesh *m, int subdiv, float ax, float ay, float az, int nbumps, int lmax, float amp,
                           unsigned long long seed) {
    if (!m) BEM_FAIL("null mesh");
    return gen_ellipsoid(m->m, subdiv, ax, ay, az, nbumps, lmax, amp, seed);
}

double bem_mesh_signed_volume(const bem_mesh *m) { return m ? signed_volume(m->m) : 0.0; }

int bem_mesh_fix_orientation(bem_mesh *m, int *flipped) {
    if (!m) BEM_FAIL("null mesh");
    return fix_orientation(m->m, flipped);
}

int bem_mesh_points_inside(const bem_mesh *m, const bem_cart_coord *pts, int n, unsigned long long seed, int *inside) {
    if (!m) BEM_FAIL("null mesh");
    return points_inside(m->m, (const float *)pts, n, seed, inside);
}
float bem_mesh_surf_area(const bem_mesh *m) { return m ? (float)surf_area(m->m) : 0.f; }

// ----------------------------------------------------------- global set-up ---
int bem_quad_init(int order) { return quad_init(order); }

int bem_quad_get(double *pnts, double *wgts) {
    HOST_CALL(ensure_quad());
    const Config &c = config();
    if (pnts) memcpy(pnts, c.qp, sizeof(double) * c.qorder);
    if (wgts) memcpy(wgts, c.qw, sizeof(double) * c.qorder);
    return EXIT_SUCCESS;
}

int bem_set_medium(float density, float speed) {
    if (!(density > 0.f) || !(speed > 0.f)) BEM_FAIL("density and speed must be positive");
    config().density = density;
    config().speed = speed;
    return EXIT_SUCCESS;
}

int bem_set_source_strength(float q) {
    config().strength = q;
    return EXIT_SUCCESS;
}

int bem_set_near_field(float factor, int max_level) {
    if (!(factor >= 0.f) || factor > 16.f) BEM_FAIL("near-field factor must be in [0, 16], got %g", factor);
    if (max_level < 1 || max_level > 24) BEM_FAIL("near-field subdivision depth must be in [1, 24], got %d", max_level);
    config().near_factor = factor;
    config().near_levels = max_level;
    return EXIT_SUCCESS;
}

// ------------------------------------------------ reference entry points ---
static int system_impl(bem_mesh *mh, float k, const bem_cart_coord *srcs, int numSrcs, bem_complex *A, int lda,
                       bem_complex *B, int ldb) {
    std::lock_guard<std::mutex> lk(g_mu);
    // arguments first: a bad call fails without touching the device
    HOST_CALL(validate_mesh(mh->m));
    const int M = mh->m.N() + mh->m.C(), N = mh->m.N();
    if (!A || lda < M) BEM_FAIL("lda (%d) must be >= M (%d)", lda, M);
    if (numSrcs < 0 || (numSrcs > 0 && (!B || !srcs || ldb < M))) BEM_FAIL("bad B / sources (ldb %d, M %d)", ldb, M);
    if (numSrcs > kMaxRhs) BEM_FAIL("at most %d sources", kMaxRhs);
    hipStream_t s;
    HOST_CALL(stream(&s));
    HOST_CALL(mesh_device(mh, s));
    DeviceProblem &dp = mh->dp;
    void *pa, *pb;
    HOST_CALL(g_sa.get(sizeof(float2) * dp.ldA * N, &pa));
    HOST_CALL(g_sb.get(sizeof(float2) * dp.ldA * std::max(1, numSrcs), &pb));
    float2 *Ad = (float2 *)pa, *Bd = (float2 *)pb;
    HOST_CALL(assemble_fp32(dp, k, Ad, s));
    RhsSpec rhs;
    rhs.kind = BEM_RHS_POINT;
    rhs.nrhs = numSrcs;
    rhs.amp = config().strength;
    rhs.params.assign((const float *)srcs, (const float *)srcs + 3 * (size_t)numSrcs);
    HOST_CALL(rhs_fp32(dp, (double)k, Bd, dp.ldA, rhs, s));
    HIP_CALL(hipMemcpy2DAsync(A, sizeof(bem_complex) * lda, Ad, sizeof(float2) * dp.ldA, sizeof(float2) * M, N,
                              hipMemcpyDeviceToHost, s));
    if (numSrcs > 0)
        HIP_CALL(hipMemcpy2DAsync(B, sizeof(bem_complex) * ldb, Bd, sizeof(float2) * dp.ldA, sizeof(float2) * M,
                                  numSrcs, hipMemcpyDeviceToHost, s));
    HIP_CALL(hipStreamSynchronize(s));
    return EXIT_SUCCESS;
}

int bem_system(const bem_mesh *m, float k, const bem_cart_coord *srcs, int numSrcs, bem_complex *A, int lda,
               bem_complex *B, int ldb) {
    if (!m) BEM_FAIL("null mesh");
    return system_impl(const_cast<bem_mesh *>(m), k, srcs, numSrcs, A, lda, B, ldb);
}

int bem_gen_system(float k, const bem_tri_elem *elems, int numElems, const bem_cart_coord *pnts, int numNods,
                   int numCHIEF, const bem_cart_coord *srcs, int numSrcs, bem_complex *A, int lda, bem_complex *B,
                   int ldb) {
    if (!elems || !pnts || numNods <= 0 || numElems <= 0 || numCHIEF < 0) BEM_FAIL("bad system arguments");
    bem_mesh tmp;
    tmp.m.pnts.assign((const float *)pnts, (const float *)pnts + 3 * (size_t)numNods);
    tmp.m.chief.assign((const float *)pnts + 3 * (size_t)numNods,
                       (const float *)pnts + 3 * (size_t)(numNods + numCHIEF));
    tmp.m.elems.assign(elems, elems + numElems);
    tmp.m.version = 1;
    int rc = system_impl(&tmp, k, srcs, numSrcs, A, lda, B, ldb);
    device_problem_free(tmp.dp);
    return rc;
}

static int qr_host(bem_complex *A, int mA, int nA, int ldA, bem_complex *B, int nB, int ldB, bem_complex *Q,
                   bool write_r) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (mA < nA || nA <= 0) BEM_FAIL("Senario not included yet: least squares needs mA >= nA > 0 (%d x %d)", mA, nA);
    if (!A || ldA < mA || nB < 0 || (nB > 0 && (!B || ldB < mA))) BEM_FAIL("bad QR arguments");
    hipStream_t s;
    HOST_CALL(stream(&s));
    const long ld = round64(mA);
    void *pa, *pb, *pq = nullptr;
    HOST_CALL(g_sa.get(sizeof(float2) * ld * nA, &pa));
    HOST_CALL(g_sb.get(sizeof(float2) * ld * std::max(1, nB), &pb));
    if (Q) HOST_CALL(g_sc.get(sizeof(float2) * (size_t)mA * mA, &pq));
    float2 *Ad = (float2 *)pa, *Bd = (float2 *)pb;
    HIP_CALL(hipMemcpy2DAsync(Ad, sizeof(float2) * ld, A, sizeof(bem_complex) * ldA, sizeof(float2) * mA, nA,
                              hipMemcpyHostToDevice, s));
    if (nB > 0)
        HIP_CALL(hipMemcpy2DAsync(Bd, sizeof(float2) * ld, B, sizeof(bem_complex) * ldB, sizeof(float2) * mA, nB,
                                  hipMemcpyHostToDevice, s));
    if (Q || write_r) {
        HOST_CALL(qr_lsq_with_q_f32(Ad, ld, mA, nA, Bd, ld, nB, (float2 *)pq, mA, s));
    } else {
        HOST_CALL(qr_lsq_f32(Ad, ld, mA, nA, Bd, ld, nB, s));
    }
    if (nB > 0)
        HIP_CALL(hipMemcpy2DAsync(B, sizeof(bem_complex) * ldB, Bd, sizeof(float2) * ld, sizeof(float2) * mA, nB,
                                  hipMemcpyDeviceToHost, s));
    if (write_r) {
        HIP_CALL(hipMemcpy2DAsync(A, sizeof(bem_complex) * ldA, Ad, sizeof(float2) * ld, sizeof(float2) * mA, nA,
                                  hipMemcpyDeviceToHost, s));
    }
    if (Q) HIP_CALL(hipMemcpyAsync(Q, pq, sizeof(float2) * (size_t)mA * mA, hipMemcpyDeviceToHost, s));
    HIP_CALL(hipStreamSynchronize(s));
    if (write_r) {
        // keep R, zero the reflectors below the diagonal
        for (int j = 0; j < nA; j++)
            for (int i = j + 1; i < mA; i++) A[(size_t)j * ldA + i] = {0.f, 0.f};
    }
    return EXIT_SUCCESS;
}

int bem_qr_solve(const bem_complex *A, int mA, int nA, int ldA, bem_complex *B, int nB, int ldB) {
    return qr_host(const_cast<bem_complex *>(A), mA, nA, ldA, B, nB, ldB, nullptr, false);
}

int bem_lsq_solve(bem_complex *A, int m, int n, int lda, bem_complex *B, int nrhs, int ldb, bem_complex *Q) {
    if (m <= n) BEM_FAIL("Senario not included yet: lsqSolver needs m > n (%d x %d)", m, n);
    return qr_host(A, m, n, lda, B, nrhs, ldb, Q, true);
}

int bem_gmres(const bem_complex *A, const bem_complex *b, int m, int maxIter, float threshold, bem_complex *x) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (!A || !b || !x || m <= 0) BEM_FAIL("bad GMRES arguments");
    if (maxIter <= 0) return EXIT_SUCCESS;
    hipStream_t s;
    HOST_CALL(stream(&s));
    const long ld = round64(m);
    void *pa, *pb;
    HOST_CALL(g_sa.get(sizeof(float2) * ld * m, &pa));
    HOST_CALL(g_sb.get(sizeof(double2) * ld * 2, &pb));
    float2 *Ad = (float2 *)pa;
    double2 *bd = (double2 *)pb, *xd = bd + ld;
    HIP_CALL(hipMemcpy2DAsync(Ad, sizeof(float2) * ld, A, sizeof(bem_complex) * m, sizeof(float2) * m, m,
                              hipMemcpyHostToDevice, s));
    std::vector<double2> hb(m), hx(m);
    for (int i = 0; i < m; i++) {
        hb[i] = make_double2(b[i].x, b[i].y);
        hx[i] = make_double2(x[i].x, x[i].y);
    }
    HIP_CALL(hipMemcpyAsync(bd, hb.data(), sizeof(double2) * m, hipMemcpyHostToDevice, s));
    HIP_CALL(hipMemcpyAsync(xd, hx.data(), sizeof(double2) * m, hipMemcpyHostToDevice, s));
    // unrestarted, like the reference: one Krylov space of up to maxIter
    // steps (at most m).  Only when its FP64 basis would take more than half
    // of the free HBM is the space restarted at the largest length that fits.
    int restart = std::min(maxIter, m);
    {
        size_t free_b = 0, total_b = 0;
        HIP_CALL(hipMemGetInfo(&free_b, &total_b));
        const long fit = (long)(free_b / 2 / (sizeof(double2) * (size_t)ld)) - 1;
        if ((long)restart > fit) restart = (int)std::max(1L, fit);
    }
    HOST_CALL(gmres_device(Ad, nullptr, ld, m, m, bd, ld, xd, ld, false, 1, restart, (double)threshold, maxIter,
                           (double)threshold, 1 << 20, nullptr, s));
    HIP_CALL(hipMemcpyAsync(hx.data(), xd, sizeof(double2) * m, hipMemcpyDeviceToHost, s));
    HIP_CALL(hipStreamSynchronize(s));
    for (int i = 0; i < m; i++) x[i] = {(float)hx[i].x, (float)hx[i].y};
    return EXIT_SUCCESS;
}

bem_complex bem_ext_pressure(float k, const bem_mesh *mc, bem_cart_coord src, bem_cart_coord x,
                             const bem_complex *surfPressure) {
    bem_complex out = {NAN, NAN};
    auto fail = [&]() { return out; };
    if (!mc || !surfPressure) {
        set_error(__FILE__, __LINE__, "null argument");
        return fail();
    }
    std::lock_guard<std::mutex> lk(g_mu);
    bem_mesh *m = const_cast<bem_mesh *>(mc);
    hipStream_t s;
    if (stream(&s) != EXIT_SUCCESS || validate_mesh(m->m) != EXIT_SUCCESS || mesh_device(m, s) != EXIT_SUCCESS)
        return fail();
    const int N = m->m.N();
    void *pa, *pb;
    if (g_sa.get(sizeof(float2) * N + 64, &pa) || g_sb.get(sizeof(float4) + sizeof(float2) + 64, &pb)) return fail();
    float2 *Xd = (float2 *)pa;
    float4 *xd = (float4 *)pb;
    float2 *od = (float2 *)((char *)pb + 32);
    float4 xh = make_float4(x.coords[0], x.coords[1], x.coords[2], 0.f);
    if (hipMemcpyAsync(Xd, surfPressure, sizeof(float2) * N, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(xd, &xh, sizeof(float4), hipMemcpyHostToDevice, s) != hipSuccess) {
        set_error(__FILE__, __LINE__, "copy failed");
        return fail();
    }
    RhsSpec rhs;
    rhs.kind = BEM_RHS_POINT;
    rhs.nrhs = 1;
    rhs.amp = config().strength;
    rhs.params = {src.coords[0], src.coords[1], src.coords[2]};
    if (field_fp32(m->dp, k, xd, 1, Xd, N, 1, rhs, od, 1, s) != EXIT_SUCCESS) return fail();
    float2 o;
    if (hipMemcpy(&o, od, sizeof(float2), hipMemcpyDeviceToHost) != hipSuccess) {
        set_error(__FILE__, __LINE__, "copy failed");
        return fail();
    }
    out = {o.x, o.y};
    return out;
}

int bem_ext_pressure_noalloc(float k, const bem_mesh *mc, bem_cart_coord src, const bem_cart_coord *xs, int P,
                             const bem_complex *surfPressure, bem_complex *out, void *stream) {
    if (!mc || !xs || !surfPressure || !out || P <= 0) BEM_FAIL("bad field arguments");
    std::lock_guard<std::mutex> lk(g_mu);
    const bem_mesh *m = mc;
    // the mesh's device records come from an earlier call that could
    // allocate them (bem_ext_pressure or bem_system on this mesh)
    if (m->dp.M == 0 || m->dp_version != m->m.version)
        BEM_FAIL("mesh is not resident on the device: call bem_ext_pressure or bem_system on it once first");
    const size_t need = field_noalloc_bytes(m->dp.E, P);
    if (need > field_noalloc_capacity())
        BEM_FAIL("%d field points on %d elements need %zu bytes of scratch, more than its %zu", P, m->dp.E, need,
                 field_noalloc_capacity());
    hipStream_t s = (hipStream_t)stream;
    if (!s) HOST_CALL(stream_of_library(&s));
    return field_fp32_noalloc(m->dp, k, (const float *)xs, P, (const float2 *)surfPressure, config().strength,
                              src.coords, (float2 *)out, s);
}

int bem_wrt_cplx_mat(const bem_complex *mat, int numRows, int numCols, int ldm, const char *file) {
    FILE *f = fopen(file, "w");
    if (!f) BEM_FAIL("Failed to open file.");
    for (int i = 0; i < numRows; i++) {
        for (int j = 0; j < numCols; j++) {
            if (fprintf(f, "(%f,%f) ", mat[(size_t)j * ldm + i].x, mat[(size_t)j * ldm + i].y) < 0) {
                fclose(f);
                BEM_FAIL("Failed to write!");
            }
        }
        fprintf(f, "\n");
    }
    fclose(f);
    return EXIT_SUCCESS;
}

void bem_print_complex_matrix(const bem_complex *A, int row, int col, int lda) {
    for (int i = 0; i < row; i++) {
        for (int j = 0; j < col; j++) printf("(%f,%f) ", A[(size_t)j * lda + i].x, A[(size_t)j * lda + i].y);
        printf("\n");
    }
}

// --------------------------------------------------- device-resident solver ---
bem_solver *bem_solver_create(const bem_mesh *m) {
    if (!m) {
        set_error(__FILE__, __LINE__, "null mesh");
        return nullptr;
    }
    Context *c = nullptr;
    if (ctx_init(&c) != EXIT_SUCCESS) return nullptr;  // device selection
    auto *sv = new bem_solver();
    sv->mesh = m->m;
    if (hipStreamCreateWithFlags(&sv->stream, hipStreamNonBlocking) != hipSuccess) {
        set_error(__FILE__, __LINE__, "stream creation failed");
        delete sv;
        return nullptr;
    }
    if (device_problem_build(sv->mesh, sv->dp, sv->stream) != EXIT_SUCCESS) {
        (void)hipStreamDestroy(sv->stream);
        delete sv;
        return nullptr;
    }
    return sv;
}

// buffers are stream-ordered on the solver's stream: freeing them never waits
// for other solvers' work
static void solver_free_mats(bem_solver *s) {
    sfree(s->A32, s->stream);
    sfree(s->A64, s->stream);
    sfree(s->B64, s->stream);
    sfree(s->X64, s->stream);
    sfree(s->X32, s->stream);
    sfree(s->F, s->stream);
    sfree(s->fbterm, s->stream);
    gmres_cache_free(s->gcache);
    device_problem_free(s->fdp);
    s->a_cap = s->x_cap = s->f_cap = 0;
    s->fP = 0;
    s->xr = 0;
    s->prec = 0;
}

void bem_solver_destroy(bem_solver *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        (void)hipStreamSynchronize(s->stream);
        solver_free_mats(s);
        device_problem_free(s->dp);
        for (hipEvent_t e : {s->ae0, s->ae1, s->se0, s->se1})
            if (e) (void)hipEventDestroy(e);
        (void)hipStreamDestroy(s->stream);
    }
    delete s;
}

int bem_solver_set_mesh(bem_solver *sv, const bem_mesh *m) {
    if (!sv || !m) BEM_FAIL("null argument");
    std::lock_guard<std::mutex> lk(sv->mu);
    solver_free_mats(sv);
    sv->mesh = m->m;
    return device_problem_build(sv->mesh, sv->dp, sv->stream);
}

int bem_solver_release(bem_solver *sv) {
    if (!sv) BEM_FAIL("null solver");
    std::lock_guard<std::mutex> lk(sv->mu);
    solver_free_mats(sv);
    device_problem_free(sv->dp);
    HIP_CALL(hipStreamSynchronize(sv->stream));
    return EXIT_SUCCESS;
}

// the last assembly's time into the stats, once its events have completed
static int resolve_assembly_time(const bem_solver *sv) {
    if (!sv->asm_pending) return EXIT_SUCCESS;
    HIP_CALL(hipEventSynchronize(sv->ae1));
    float ms = 0;
    HIP_CALL(hipEventElapsedTime(&ms, sv->ae0, sv->ae1));
    sv->stats.assemble_ms = ms;
    sv->asm_pending = false;
    return EXIT_SUCCESS;
}

int bem_solver_assemble(bem_solver *sv, float k, int precision) {
    if (!sv) BEM_FAIL("null solver");
    if (precision != BEM_PREC_FP32 && precision != BEM_PREC_FP64) BEM_FAIL("precision must be 32 or 64");
    std::lock_guard<std::mutex> lk(sv->mu);
    hipStream_t s = sv->stream;
    DeviceProblem &dp = sv->dp;
    if (dp.M == 0) BEM_FAIL("solver has no mesh (released)");
    if (!sv->ae0) HIP_CALL(hipEventCreate(&sv->ae0));
    if (!sv->ae1) HIP_CALL(hipEventCreate(&sv->ae1));
    HIP_CALL(hipEventRecord(sv->ae0, s));
    const size_t need = (size_t)dp.ldA * dp.N;
    if (precision == BEM_PREC_FP32) {
        if (sv->A64) {
            sfree(sv->A64, s);
            sv->a_cap = 0;
        }
        if (sv->a_cap < need) {
            sfree(sv->A32, s);
            HOST_CALL(salloc(&sv->A32, need, s));
            sv->a_cap = need;
        }
        HOST_CALL(assemble_fp32(dp, k, sv->A32, s));
    } else {
        if (sv->A32) {
            sfree(sv->A32, s);
            sv->a_cap = 0;
        }
        if (sv->a_cap < need) {
            sfree(sv->A64, s);
            HOST_CALL(salloc(&sv->A64, need, s));
            sv->a_cap = need;
        }
        HOST_CALL(assemble_fp64(dp, (double)k, sv->A64, s));
    }
    // no wait here: the assembly's time is read from the events later
    HIP_CALL(hipEventRecord(sv->ae1, s));
    sv->asm_pending = true;
    sv->k = k;
    sv->prec = precision;
    sv->xr = 0;
    return EXIT_SUCCESS;
}

int bem_solver_set_rhs(bem_solver *sv, int kind, const float *params, int nrhs, float amp) {
    if (!sv) BEM_FAIL("null solver");
    if (kind != BEM_RHS_POINT && kind != BEM_RHS_PLANE) BEM_FAIL("unknown right-hand-side kind %d", kind);
    if (nrhs <= 0 || nrhs > kMaxRhs || !params) BEM_FAIL("nrhs must be in [1, %d]", kMaxRhs);
    std::lock_guard<std::mutex> lk(sv->mu);
    sv->rhs.kind = kind;
    sv->rhs.nrhs = nrhs;
    sv->rhs.amp = amp;
    sv->rhs.params.assign(params, params + 3 * (size_t)nrhs);
    sv->xr = 0;  // the previous solution belongs to other right-hand sides
    return EXIT_SUCCESS;
}

int bem_solver_solve(bem_solver *sv, int method, int restart, float tol, int maxit, double refine_tol, int max_refine,
                     bem_solve_stats *stats) {
    if (!sv) BEM_FAIL("null solver");
    if (!sv->prec) BEM_FAIL("solver not assembled");
    if (sv->rhs.nrhs <= 0) BEM_FAIL("no right-hand sides set");
    if (method != BEM_SOLVE_QR && method != BEM_SOLVE_GMRES && method != BEM_SOLVE_CGLS &&
        method != BEM_SOLVE_BLOCK_CG)
        BEM_FAIL("unknown method %d", method);
    std::lock_guard<std::mutex> lk(sv->mu);
    hipStream_t s = sv->stream;
    DeviceProblem &dp = sv->dp;
    const int R = sv->rhs.nrhs;
    const long ld = dp.ldA;
    if (sv->x_cap < (size_t)ld * R) {
        sfree(sv->B64, s);
        sfree(sv->X64, s);
        sfree(sv->X32, s);
        HOST_CALL(salloc(&sv->B64, (size_t)ld * R, s));
        HOST_CALL(salloc(&sv->X64, (size_t)ld * R, s));
        HOST_CALL(salloc(&sv->X32, (size_t)ld * R, s));
        sv->x_cap = (size_t)ld * R;
    }
    sv->xr = 0;
    if (!sv->se0) HIP_CALL(hipEventCreate(&sv->se0));
    if (!sv->se1) HIP_CALL(hipEventCreate(&sv->se1));
    HIP_CALL(hipEventRecord(sv->se0, s));
    HOST_CALL(rhs_fp64(dp, (double)sv->k, sv->B64, ld, sv->rhs, s));
    bem_solve_stats st{};
    st.method = method;
    if (method == BEM_SOLVE_CGLS || method == BEM_SOLVE_BLOCK_CG) {
        const double outer = refine_tol > 0 ? refine_tol : (double)tol;
        const int cycles = max_refine > 0 ? max_refine : (1 << 20);
        if (maxit <= 0) maxit = 100000;
        HOST_CALL(cgls_device(sv->A32, sv->A64, ld, dp.M, dp.N, sv->B64, ld, sv->X64, ld, R, (double)tol, maxit,
                              outer, cycles, &st, s, method == BEM_SOLVE_BLOCK_CG));
    } else if (method == BEM_SOLVE_GMRES) {
        HIP_CALL(hipMemsetAsync(sv->X64, 0, sizeof(double2) * ld * R, s));
        const double outer = refine_tol > 0 ? refine_tol : (double)tol;
        const int cycles = max_refine > 0 ? max_refine : (1 << 20);
        if (maxit <= 0) maxit = 100000;
        HOST_CALL(gmres_device(sv->A32, sv->A64, ld, dp.M, dp.N, sv->B64, ld, sv->X64, ld, true, R, restart,
                               (double)tol, maxit, outer, cycles, &st, s, &sv->gcache));
    } else {
        if (sv->A32) {
            HOST_CALL(qr_refined_f32(sv->A32, ld, dp.M, dp.N, sv->B64, ld, sv->X64, ld, R,
                                     refine_tol > 0 ? refine_tol : 1e-12, max_refine > 0 ? max_refine : 5, &st, s));
        } else {
            double2 *Af = nullptr;
            HOST_CALL(salloc(&Af, (size_t)ld * dp.N, s));
            HIP_CALL(hipMemcpyAsync(Af, sv->A64, sizeof(double2) * ld * dp.N, hipMemcpyDeviceToDevice, s));
            HIP_CALL(hipMemcpyAsync(sv->X64, sv->B64, sizeof(double2) * ld * R, hipMemcpyDeviceToDevice, s));
            int rc = qr_lsq_f64(Af, ld, dp.M, dp.N, sv->X64, ld, R, s);
            sfree(Af, s);
            HOST_CALL(rc);
            st.converged = 1;
        }
    }
    HIP_CALL(hipEventRecord(sv->se1, s));
    HIP_CALL(hipStreamSynchronize(s));
    sv->xr = R;
    // device time from the solve's first work to its last (an assembly still
    // running when the solve was called is not counted)
    float sms = 0;
    HIP_CALL(hipEventElapsedTime(&sms, sv->se0, sv->se1));
    st.solve_ms = sms;
    HOST_CALL(resolve_assembly_time(sv));
    st.assemble_ms = sv->stats.assemble_ms;
    sv->stats = st;
    if (stats) *stats = st;
    return EXIT_SUCCESS;
}

int bem_solver_get_solution_z(const bem_solver *sv, bem_zcomplex *x, int ldx) {
    if (!sv || !x) BEM_FAIL("null argument");
    if (sv->xr <= 0) BEM_FAIL("no solution");
    if (ldx < sv->dp.N) BEM_FAIL("ldx must be >= N");
    std::lock_guard<std::mutex> lk(sv->mu);
    hipStream_t s = sv->stream;
    HIP_CALL(hipMemcpy2DAsync(x, sizeof(bem_zcomplex) * ldx, sv->X64, sizeof(double2) * sv->dp.ldA,
                              sizeof(double2) * sv->dp.N, sv->rhs.nrhs, hipMemcpyDeviceToHost, s));
    HIP_CALL(hipStreamSynchronize(s));
    return EXIT_SUCCESS;
}

int bem_solver_get_solution(const bem_solver *sv, bem_complex *x, int ldx) {
    if (!sv || !x) BEM_FAIL("null argument");
    if (ldx < sv->dp.N) BEM_FAIL("ldx must be >= N");
    std::vector<bem_zcomplex> z((size_t)sv->dp.N * std::max(1, sv->rhs.nrhs));
    HOST_CALL(bem_solver_get_solution_z(sv, z.data(), sv->dp.N));
    for (int r = 0; r < sv->rhs.nrhs; r++)
        for (int i = 0; i < sv->dp.N; i++) {
            const bem_zcomplex v = z[(size_t)r * sv->dp.N + i];
            x[(size_t)r * ldx + i] = {(float)v.x, (float)v.y};
        }
    return EXIT_SUCCESS;
}

__global__ void k_z2c_copy(const double2 *in, float2 *out, long n) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = make_float2((float)in[i].x, (float)in[i].y);
}

int bem_solver_get_system(const bem_solver *sv, bem_complex *A, int lda, bem_complex *B, int ldb) {
    if (!sv) BEM_FAIL("null solver");
    if (!sv->prec) BEM_FAIL("solver not assembled");
    const DeviceProblem &dp = sv->dp;
    if (A && lda < dp.M) BEM_FAIL("lda must be >= M");
    if (B && ldb < dp.M) BEM_FAIL("ldb must be >= M");
    std::lock_guard<std::mutex> lk(sv->mu);
    hipStream_t s = sv->stream;
    if (A) {
        const float2 *src = sv->A32;
        float2 *tmp = nullptr;
        if (!src) {
            const long n = dp.ldA * (long)dp.N;
            HOST_CALL(salloc(&tmp, (size_t)n, s));
            k_z2c_copy<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(sv->A64, tmp, n);
            HIP_CALL(hipGetLastError());
            src = tmp;
        }
        hipError_t e = hipMemcpy2DAsync(A, sizeof(bem_complex) * lda, src, sizeof(float2) * dp.ldA,
                                        sizeof(float2) * dp.M, dp.N, hipMemcpyDeviceToHost, s);
        (void)hipStreamSynchronize(s);
        sfree(tmp, s);
        HIP_CALL(e);
    }
    if (B && sv->rhs.nrhs > 0) {
        float2 *Bd = nullptr;
        HOST_CALL(salloc(&Bd, (size_t)dp.ldA * sv->rhs.nrhs, s));
        int rc = rhs_fp32(const_cast<DeviceProblem &>(dp), (double)sv->k, Bd, dp.ldA, sv->rhs, s);
        hipError_t e = hipMemcpy2DAsync(B, sizeof(bem_complex) * ldb, Bd, sizeof(float2) * dp.ldA,
                                        sizeof(float2) * dp.M, sv->rhs.nrhs, hipMemcpyDeviceToHost, s);
        (void)hipStreamSynchronize(s);
        sfree(Bd, s);
        HOST_CALL(rc);
        HIP_CALL(e);
    }
    return EXIT_SUCCESS;
}

int bem_solver_get_rows(const bem_solver *sv, const int *rows, int nrows, bem_zcomplex *out) {
    if (!sv || !rows || !out || nrows <= 0) BEM_FAIL("bad arguments");
    if (!sv->prec) BEM_FAIL("solver not assembled");
    const DeviceProblem &dp = sv->dp;
    for (int q = 0; q < nrows; q++)
        if (rows[q] < 0 || rows[q] >= dp.M) BEM_FAIL("row %d out of range [0, %d)", rows[q], dp.M);
    std::lock_guard<std::mutex> lk(sv->mu);
    hipStream_t s = sv->stream;
    int *rd = nullptr;
    double2 *od = nullptr;
    HOST_CALL(salloc(&rd, nrows, s));
    HOST_CALL(salloc(&od, (size_t)nrows * dp.N, s));
    HIP_CALL(hipMemcpyAsync(rd, rows, sizeof(int) * nrows, hipMemcpyHostToDevice, s));
    const long n = (long)nrows * dp.N;
    const unsigned nb = (unsigned)((n + 255) / 256);
    if (sv->A32) k_gather_rows<float2><<<nb, 256, 0, s>>>(sv->A32, dp.ldA, dp.N, rd, nrows, od);
    else k_gather_rows<double2><<<nb, 256, 0, s>>>(sv->A64, dp.ldA, dp.N, rd, nrows, od);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, od, sizeof(double2) * n, hipMemcpyDeviceToHost, s);
    (void)hipStreamSynchronize(s);
    sfree(rd, s);
    sfree(od, s);
    HIP_CALL(e);
    return EXIT_SUCCESS;
}

int bem_solver_field(bem_solver *sv, const bem_cart_coord *pts, int P, bem_complex *out, int ldo) {
    if (!sv || !pts || !out || P <= 0) BEM_FAIL("bad field arguments");
    if (sv->xr <= 0) BEM_FAIL("no solution");
    if (ldo < P) BEM_FAIL("ldo must be >= P");
    std::lock_guard<std::mutex> lk(sv->mu);
    hipStream_t s = sv->stream;
    DeviceProblem &dp = sv->dp;
    const int R = sv->rhs.nrhs;
    const long n = dp.ldA * (long)R;
    k_z2c_copy<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(sv->X64, sv->X32, n);
    HIP_CALL(hipGetLastError());
    std::vector<float4> xh(P);
    for (int p = 0; p < P; p++) xh[p] = make_float4(pts[p].coords[0], pts[p].coords[1], pts[p].coords[2], 0.f);
    float4 *xd = nullptr;
    float2 *od = nullptr;
    HOST_CALL(salloc(&xd, P, s));
    HOST_CALL(salloc(&od, (size_t)P * R, s));
    HIP_CALL(hipMemcpyAsync(xd, xh.data(), sizeof(float4) * P, hipMemcpyHostToDevice, s));
    int rc = field_fp32(dp, sv->k, xd, P, sv->X32, dp.ldA, R, sv->rhs, od, P, s);
    hipError_t e = hipMemcpy2DAsync(out, sizeof(bem_complex) * ldo, od, sizeof(float2) * P, sizeof(float2) * P, R,
                                    hipMemcpyDeviceToHost, s);
    (void)hipStreamSynchronize(s);
    sfree(xd, s);
    sfree(od, s);
    HOST_CALL(rc);
    HIP_CALL(e);
    return EXIT_SUCCESS;
}

int bem_solver_field_matrix(bem_solver *sv, const bem_cart_coord *pts, int P) {
    if (!sv || !pts || P <= 0) BEM_FAIL("bad field-matrix arguments");
    if (!sv->prec) BEM_FAIL("solver not assembled");
    std::lock_guard<std::mutex> lk(sv->mu);
    hipStream_t s = sv->stream;
    sv->fmesh = sv->mesh;
    sv->fmesh.chief.assign((const float *)pts, (const float *)pts + 3 * (size_t)P);
    sv->fmesh.version++;
    sv->fP = 0;
    HOST_CALL(device_problem_build(sv->fmesh, sv->fdp, s));
    const long ldF = ((long)P + 63) / 64 * 64;
    const size_t need = (size_t)ldF * sv->dp.N;
    if (sv->f_cap < need) {
        sfree(sv->F, s);
        HOST_CALL(salloc(&sv->F, need, s));
        sv->f_cap = need;
    }
    sfree(sv->fbterm, s);
    HOST_CALL(salloc(&sv->fbterm, (size_t)P, s));
    sv->ldF = ldF;
    HOST_CALL(field_matrix_fp32(sv->fdp, sv->k, sv->F, ldF, sv->fbterm, s));
    HIP_CALL(hipStreamSynchronize(s));
    sv->fP = P;
    sv->fk = sv->k;
    return EXIT_SUCCESS;
}

int bem_solver_field_apply(bem_solver *sv, bem_complex *out, int ldo) {
    if (!sv || !out) BEM_FAIL("bad field-apply arguments");
    if (sv->fP <= 0) BEM_FAIL("no field matrix (bem_solver_field_matrix)");
    if (sv->fk != sv->k) BEM_FAIL("field matrix assembled at k = %g, solver at k = %g", sv->fk, sv->k);
    if (sv->xr <= 0) BEM_FAIL("no solution");
    if (ldo < sv->fP) BEM_FAIL("ldo must be >= P");
    std::lock_guard<std::mutex> lk(sv->mu);
    hipStream_t s = sv->stream;
    const DeviceProblem &dp = sv->dp;
    const int R = sv->rhs.nrhs, P = sv->fP;
    const long n = dp.ldA * (long)R;
    k_z2c_copy<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(sv->X64, sv->X32, n);
    HIP_CALL(hipGetLastError());
    float2 *Y = nullptr, *od = nullptr;
    char *ws = nullptr;
    HOST_CALL(salloc(&Y, (size_t)sv->ldF * R, s));
    HOST_CALL(salloc(&od, (size_t)sv->ldF * R, s));
    HOST_CALL(salloc(&ws, std::max(ws_gemv_n(P, dp.N, 1), ws_gemm(P, dp.N, R)), s));
    int rc = field_apply_fp32(sv->fdp, (double)sv->k, sv->F, sv->ldF, sv->fbterm, sv->X32, dp.ldA, sv->rhs, Y, od,
                              sv->ldF, ws, s);
    hipError_t e = hipSuccess;
    if (rc == EXIT_SUCCESS)
        e = hipMemcpy2DAsync(out, sizeof(bem_complex) * ldo, od, sizeof(float2) * sv->ldF, sizeof(float2) * P, R,
                             hipMemcpyDeviceToHost, s);
    (void)hipStreamSynchronize(s);
    sfree(Y, s);
    sfree(od, s);
    sfree(ws, s);
    HOST_CALL(rc);
    HIP_CALL(e);
    return EXIT_SUCCESS;
}

int bem_solver_get_field_matrix(const bem_solver *sv, bem_complex *F, int ldf) {
    if (!sv || !F) BEM_FAIL("null argument");
    if (sv->fP <= 0) BEM_FAIL("no field matrix (bem_solver_field_matrix)");
    if (ldf < sv->fP) BEM_FAIL("ldf must be >= P");
    std::lock_guard<std::mutex> lk(sv->mu);
    hipStream_t s = sv->stream;
    HIP_CALL(hipMemcpy2DAsync(F, sizeof(bem_complex) * ldf, sv->F, sizeof(float2) * sv->ldF, sizeof(float2) * sv->fP,
                              sv->dp.N, hipMemcpyDeviceToHost, s));
    HIP_CALL(hipStreamSynchronize(s));
    return EXIT_SUCCESS;
}

int bem_solver_dims(const bem_solver *sv, int *M, int *N, int *nrhs) {
    if (!sv) BEM_FAIL("null solver");
    if (M) *M = sv->dp.M;
    if (N) *N = sv->dp.N;
    if (nrhs) *nrhs = sv->rhs.nrhs;
    return EXIT_SUCCESS;
}

int bem_solver_stats(const bem_solver *sv, bem_solve_stats *st) {
    if (!sv || !st) BEM_FAIL("null argument");
    std::lock_guard<std::mutex> lk(sv->mu);
    HOST_CALL(resolve_assembly_time(sv));
    *st = sv->stats;
    return EXIT_SUCCESS;
}

// ---------------------------------------------------------------- tuning ---
int bem_autotune_assembly(const bem_mesh *m, float k, int precision, bem_tune_result *out) {
    if (!m) BEM_FAIL("null mesh");
    if (precision != BEM_PREC_FP32 && precision != BEM_PREC_FP64) BEM_FAIL("precision must be 32 or 64");
    std::lock_guard<std::mutex> lk(g_mu);
    hipStream_t s;
    HOST_CALL(stream(&s));
    bem_tune_result r{};
    HOST_CALL(autotune_assembly(m->m, (double)k, precision, &r.lanes_per_row, &r.waves, &r.best_ms, &r.default_ms,
                                s));
    if (out) *out = r;
    return EXIT_SUCCESS;
}
int bem_autotune_products(int M, int N, int nrhs, bem_product_tune_result *out) {
    if (M < N || N <= 0 || nrhs <= 0 || nrhs > kMaxRhs) BEM_FAIL("bad product sizes %d x %d, %d right-hand sides", M, N,
                                                                   nrhs);
    std::lock_guard<std::mutex> lk(g_mu);
    hipStream_t s;
    HOST_CALL(stream(&s));
    int shape[4] = {0, 0, 0, 0};
    double ms[4] = {0, 0, 0, 0};
    HOST_CALL(autotune_products(M, N, nrhs, shape, ms, s));
    if (out) {
        out->rhs_per_wave_n = shape[0];
        out->chunks_n = shape[1];
        out->rhs_per_wave_h = shape[2];
        out->chunks_h = shape[3];
        out->best_ms_n = ms[0];
        out->default_ms_n = ms[1];
        out->best_ms_h = ms[2];
        out->default_ms_h = ms[3];
    }
    return EXIT_SUCCESS;
}
int bem_tune_save(const char *file) {
    if (!file) BEM_FAIL("null file");
    return tune_save(file);
}
int bem_tune_load(const char *file) {
    if (!file) BEM_FAIL("null file");
    return tune_load(file);
}
int bem_tune_clear(void) {
    tune_clear();
    return EXIT_SUCCESS;
}

// ------------------------------------------------------------------ timing ---
int bem_timing_enable(int on) { return timing_enable(on); }
int bem_timing_reset(void) { return timing_reset(); }
int bem_timing_get(const char *kernel, double *total_ms, long *count) {
    if (!kernel) BEM_FAIL("null kernel name");
    return timing_get(kernel, total_ms, count);
}
int bem_sync(void) {
    hipStream_t s;
    HOST_CALL(stream(&s));
    HIP_CALL(hipDeviceSynchronize());  // every solver stream and the library stream
    return EXIT_SUCCESS;
}

const char *bem_last_error(void) { return last_error(); }
const char *bem_version(void) { return "bem-mi355x 0.1 (gfx950)"; }

}  // extern "C"
