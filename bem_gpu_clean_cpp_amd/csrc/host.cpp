// Host-side runtime of libbem: error reporting, global configuration,
// Gauss-Legendre tables, the HIP stream/context, event timing, and the mesh
// front-end (OBJ reader, bounding box, CHIEF generator, icosphere and perturbed
// ellipsoid generators, column-tile planner for the row-owned assembly).
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <unordered_map>

#include "bem_internal.hpp"

namespace bem {

// ---------------------------------------------------------------- errors ---
static thread_local std::string g_last_error;

static bool quiet_errors() {
    static const bool quiet = getenv("BEM_QUIET") && atoi(getenv("BEM_QUIET")) != 0;
    return quiet;
}

void set_error(const char *file, int line, const char *fmt, ...) {
    char msg[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(msg, sizeof msg, fmt, ap);
    va_end(ap);
    char full[1400];
    const char *base = strrchr(file, '/');
    snprintf(full, sizeof full, "%s:%d: %s", base ? base + 1 : file, line, msg);
    g_last_error = full;
    // The reference prints "Error at file:line" before returning EXIT_FAILURE
    // (numerical.h:60-122); BEM_QUIET=1 silences it.
    if (!quiet_errors()) fprintf(stderr, "Error at %s:%d: %s\n", base ? base + 1 : file, line, msg);
}
// Every failing path sets a message before it returns, so the callee's
// message is the current one; the caller's site goes in front of it
// ("capi.cpp:188: validate_mesh(mh->m) failed: host.cpp:246: Not enough ...").
void chain_error(const char *file, int line, const char *what) {
    const char *base = strrchr(file, '/');
    char head[512];
    snprintf(head, sizeof head, "%s:%d: %s failed", base ? base + 1 : file, line, what);
    g_last_error = g_last_error.empty() ? std::string(head) : std::string(head) + ": " + g_last_error;
    if (!quiet_errors()) fprintf(stderr, "Error at %s\n", head);
}
const char *last_error() { return g_last_error.c_str(); }
void clear_error() { g_last_error.clear(); }

// ---------------------------------------------------------------- config ---
Config &config() {
    static Config c;
    return c;
}

// Gauss-Legendre nodes and weights by the Golub-Welsch method (the method of
// gaussQuad::genGaussParams, numerical.cu:186-225): the nodes are the
// eigenvalues of the symmetric tridiagonal Jacobi matrix with off-diagonal
// b_i = (i+1)/sqrt((2i+2)^2-1), the weights 2 (first component of the unit
// eigenvector)^2.  The eigenproblem is solved by the cyclic Jacobi method
// (Jacobi 1846; Golub & Van Loan, Matrix Computations, 4th ed., section 8.5):
// sweep over the off-diagonal pairs (p, q) in row order, zero each with a
// plane rotation J (J^T A J), and accumulate the rotations into the eigenvector
// matrix, until the off-diagonal mass is at rounding level.  Jacobi
// rotations converge quadratically once the off-diagonal part is small and
// give eigenvectors orthogonal to working precision; n <= 32 keeps the
// O(n^3) sweeps negligible.
int gauss_legendre(int n, double *x, double *w) {
    if (n < 1 || n > 32) BEM_FAIL("quadrature order %d out of range [1, 32]", n);
    std::vector<double> a((size_t)n * n, 0.0), v((size_t)n * n, 0.0);
    auto A = [&](int i, int j) -> double & { return a[(size_t)i * n + j]; };
    auto V = [&](int i, int j) -> double & { return v[(size_t)i * n + j]; };
    for (int i = 0; i < n; i++) V(i, i) = 1.0;
    double total = 0.0;
    for (int i = 0; i + 1 < n; i++) {
        const double b = (i + 1) / std::sqrt(4.0 * (i + 1) * (i + 1) - 1.0);
        A(i, i + 1) = A(i + 1, i) = b;
        total += 2.0 * b * b;
    }
    auto offdiag = [&]() {
        double o = 0.0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < n; j++)
                if (i != j) o += A(i, j) * A(i, j);
        return o;
    };
    int sweeps = 0;
    while (offdiag() > 1e-32 * total) {
        if (++sweeps > 64) BEM_FAIL("Gauss-Legendre eigen-iteration did not converge");
        for (int p = 0; p + 1 < n; p++)
            for (int q = p + 1; q < n; q++) {
                const double apq = A(p, q);
                if (apq == 0.0) continue;
                // rotation angle phi with tan(2 phi) = 2 a_pq / (a_qq - a_pp):
                // t = tan(phi) is the root of t^2 + 2 tau t - 1 = 0 of smaller size
                const double tau = (A(q, q) - A(p, p)) / (2.0 * apq);
                const double t = (tau >= 0.0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
                const double c = 1.0 / std::sqrt(1.0 + t * t), sn = t * c;
                for (int k = 0; k < n; k++) {  // A J: columns p and q
                    const double kp = A(k, p), kq = A(k, q);
                    A(k, p) = c * kp - sn * kq;
                    A(k, q) = sn * kp + c * kq;
                }
                for (int k = 0; k < n; k++) {  // J^T (A J): rows p and q
                    const double pk = A(p, k), qk = A(q, k);
                    A(p, k) = c * pk - sn * qk;
                    A(q, k) = sn * pk + c * qk;
                }
                for (int k = 0; k < n; k++) {  // V J
                    const double kp = V(k, p), kq = V(k, q);
                    V(k, p) = c * kp - sn * kq;
                    V(k, q) = sn * kp + c * kq;
                }
            }
    }
    std::vector<int> idx(n);
    std::iota(idx.begin(), idx.end(), 0);
    std::sort(idx.begin(), idx.end(), [&](int i, int j) { return A(i, i) < A(j, j); });
    for (int i = 0; i < n; i++) {
        x[i] = A(idx[i], idx[i]);
        w[i] = 2.0 * V(0, idx[i]) * V(0, idx[i]);
    }
    return EXIT_SUCCESS;
}

int quad_init(int order) {
    Config &c = config();
    if (order != kQOrder) BEM_FAIL("the GPU kernels are built for the %d-point rule, got %d", kQOrder, order);
    HOST_CALL(gauss_legendre(order, c.qp, c.qw));
    c.qorder = order;
    return EXIT_SUCCESS;
}

int ensure_quad() {
    static std::mutex mu;  // solvers on several host threads may get here first together
    std::lock_guard<std::mutex> lk(mu);
    if (config().qorder == 0) return quad_init(kQOrder);
    return EXIT_SUCCESS;
}

// Point q = n*6 + m of the collapsed rule: theta from node n, rho from node m
// (mesh.cu:623-636), regular map xi1 = rho(1-theta), xi2 = rho - xi1.
void quad_tables(QuadTables &t) {
    const Config &c = config();
    for (int n = 0; n < kQOrder; n++) {
        double theta = 0.5 + 0.5 * c.qp[n];
        for (int m = 0; m < kQOrder; m++) {
            double rho = 0.5 + 0.5 * c.qp[m];
            int q = n * kQOrder + m;
            double xi1 = rho * (1 - theta), xi2 = rho * theta;
            t.dxi1[q] = xi1;
            t.dxi2[q] = xi2;
            t.dw[q] = 0.25 * c.qw[n] * c.qw[m] * rho;
            t.xi1[q] = (float)xi1;
            t.xi2[q] = (float)xi2;
            t.w[q] = (float)t.dw[q];
        }
    }
}

// --------------------------------------------------------------- context ---
int ctx_init(Context **out) {
    static Context ctx;
    static std::mutex mu;
    std::lock_guard<std::mutex> lk(mu);
    if (!ctx.stream) {
        int dev = 0;
        HIP_CALL(hipGetDevice(&dev));
        // default-priority stream
        HIP_CALL(hipStreamCreateWithFlags(&ctx.stream, hipStreamNonBlocking));
        ctx.device = dev;
        // Stream-ordered buffers (matrices, Krylov state, temporaries) come from the
        // device's default pool; keep freed memory in the pool so the next
        // solve's allocations do not map fresh pages.
        hipMemPool_t pool;
        if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
            uint64_t keep = UINT64_MAX;
            (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        }
    }
    *out = &ctx;
    return EXIT_SUCCESS;
}

// ---------------------------------------------------------------- timing ---
struct TimerRec {
    std::string name;
    hipEvent_t a, b;
};
static std::mutex g_tmu;
static bool g_timing = false;
static std::vector<TimerRec> g_timers;
static std::vector<hipEvent_t> g_event_pool;

bool timing_on() { return g_timing; }

static hipEvent_t take_event() {
    if (!g_event_pool.empty()) {
        hipEvent_t e = g_event_pool.back();
        g_event_pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void timing_begin(TimedRegion &r, const char *name, hipStream_t s) {
    r.name = nullptr;
    if (!g_timing) return;
    std::lock_guard<std::mutex> lk(g_tmu);
    r.a = take_event();
    r.b = take_event();
    if (!r.a || !r.b) return;
    r.name = name;
    (void)hipEventRecord(r.a, s);
}

void timing_end(TimedRegion &r, hipStream_t s) {
    if (!r.name) return;
    (void)hipEventRecord(r.b, s);
    std::lock_guard<std::mutex> lk(g_tmu);
    g_timers.push_back({r.name, r.a, r.b});
    r.name = nullptr;
}

int timing_enable(int on) {
    g_timing = on != 0;
    return EXIT_SUCCESS;
}

int timing_reset() {
    std::lock_guard<std::mutex> lk(g_tmu);
    for (auto &t : g_timers) {
        (void)hipEventSynchronize(t.b);
        g_event_pool.push_back(t.a);
        g_event_pool.push_back(t.b);
    }
    g_timers.clear();
    return EXIT_SUCCESS;
}

int timing_get(const char *name, double *total_ms, long *count) {
    std::lock_guard<std::mutex> lk(g_tmu);
    double tot = 0;
    long cnt = 0;
    for (auto &t : g_timers) {
        if (t.name != name) continue;
        HIP_CALL(hipEventSynchronize(t.b));
        float ms = 0;
        HIP_CALL(hipEventElapsedTime(&ms, t.a, t.b));
        tot += ms;
        cnt++;
    }
    if (total_ms) *total_ms = tot;
    if (count) *count = cnt;
    return EXIT_SUCCESS;
}

// ------------------------------------------------------------------ mesh ---
int validate_mesh(const Mesh &m) {
    if (m.N() == 0 || m.E() == 0) BEM_FAIL("Not enough mesh information! (%d nodes, %d elements)", m.N(), m.E());
    for (int e = 0; e < m.E(); e++)
        for (int j = 0; j < 3; j++) {
            int v = m.elems[e].nodes[j];
            if (v < 0 || v >= m.N()) BEM_FAIL("element %d references node %d outside [0, %d)", e, v, m.N());
        }
    return EXIT_SUCCESS;
}

// OBJ reader.  The semantics of mesh::readObj (mesh.cu:1215-1265) for the
// lines it understands: "v x y z" is a vertex, "f a b c" a triangle with
// 1-based indices, every element gets bc (0, 1, -1) (mesh.cu:1240-1242).
// Robustness the reference lacks: "vn"/"vt" lines are not vertices, face
// tokens may be "a/b/c" or negative (relative), polygons are fanned.
int read_obj(Mesh &m, const char *file) {
    FILE *fp = fopen(file, "r");
    if (!fp) BEM_FAIL("Failed to open file %s.", file ? file : "(null)");
    std::vector<float> pn;
    std::vector<bem_tri_elem> el;
    char buf[8192];
    long lineno = 0;
    while (fgets(buf, sizeof buf, fp)) {
        lineno++;
        if (buf[0] == 'v' && (buf[1] == ' ' || buf[1] == '\t')) {
            float c[3];
            if (sscanf(buf + 1, "%f %f %f", &c[0], &c[1], &c[2]) != 3) {
                fclose(fp);
                BEM_FAIL("%s:%ld: malformed vertex line", file, lineno);
            }
            pn.insert(pn.end(), c, c + 3);
        } else if (buf[0] == 'f' && (buf[1] == ' ' || buf[1] == '\t')) {
            std::vector<int> idx;
            char *p = buf + 1;
            while (*p) {
                while (*p == ' ' || *p == '\t') p++;
                if (!*p || *p == '\n' || *p == '\r' || *p == '#') break;
                char *endp = nullptr;
                long v = strtol(p, &endp, 10);
                if (endp == p) {
                    fclose(fp);
                    BEM_FAIL("%s:%ld: malformed face line", file, lineno);
                }
                long nv = (long)(pn.size() / 3);
                idx.push_back((int)(v < 0 ? nv + v : v - 1));
                p = endp;
                while (*p && *p != ' ' && *p != '\t' && *p != '\n' && *p != '\r') p++;  // skip /t/n
            }
            if (idx.size() < 3) {
                fclose(fp);
                BEM_FAIL("%s:%ld: face with fewer than 3 vertices", file, lineno);
            }
            for (size_t t = 1; t + 1 < idx.size(); t++) {
                bem_tri_elem e;
                e.nodes[0] = idx[0];
                e.nodes[1] = idx[t];
                e.nodes[2] = idx[t + 1];
                e.bc[0] = {0.f, 0.f};
                e.bc[1] = {1.f, 0.f};
                e.bc[2] = {-1.f, 0.f};
                el.push_back(e);
            }
        }
    }
    fclose(fp);
    m.pnts.swap(pn);
    m.elems.swap(el);
    m.chief.clear();
    m.bb_set = false;
    m.version++;
    return validate_mesh(m);
}

int find_bb(Mesh &m, float threshold) {
    if (m.N() == 0 || m.E() == 0) BEM_FAIL("Not enough mesh information!");
    float lo[3], hi[3];
    for (int d = 0; d < 3; d++) lo[d] = hi[d] = m.pnts[d];
    for (int i = 1; i < m.N(); i++)
        for (int d = 0; d < 3; d++) {
            lo[d] = std::min(lo[d], m.pnts[3 * i + d]);
            hi[d] = std::max(hi[d], m.pnts[3 * i + d]);
        }
    for (int d = 0; d < 3; d++) {
        m.bb[2 * d] = lo[d] - threshold;
        m.bb[2 * d + 1] = hi[d] + threshold;
    }
    m.bb_set = true;
    return EXIT_SUCCESS;
}

// SplitMix64 (Steele, Lea & Flood, OOPSLA 2014).
struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
};

// Moller-Trumbore ray/triangle test in double; counts hits with t > 0.
static bool ray_hits(const double *o, const double *d, const float *a, const float *b, const float *c) {
    double e1[3] = {b[0] - (double)a[0], b[1] - (double)a[1], b[2] - (double)a[2]};
    double e2[3] = {c[0] - (double)a[0], c[1] - (double)a[1], c[2] - (double)a[2]};
    double p[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2], d[0] * e2[1] - d[1] * e2[0]};
    double det = e1[0] * p[0] + e1[1] * p[1] + e1[2] * p[2];
    if (std::fabs(det) < 1e-300) return false;
    double inv = 1.0 / det;
    double t0[3] = {o[0] - a[0], o[1] - a[1], o[2] - a[2]};
    double u = (t0[0] * p[0] + t0[1] * p[1] + t0[2] * p[2]) * inv;
    if (u < 0.0 || u > 1.0) return false;
    double q[3] = {t0[1] * e1[2] - t0[2] * e1[1], t0[2] * e1[0] - t0[0] * e1[2], t0[0] * e1[1] - t0[1] * e1[0]};
    double v = (d[0] * q[0] + d[1] * q[1] + d[2] * q[2]) * inv;
    if (v < 0.0 || u + v > 1.0) return false;
    double t = (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]) * inv;
    return t > 0.0;
}

// inObj (numerical.cu:119-131): a point is inside when a ray from it crosses
// the surface an odd number of times.  A ray through a shared edge or vertex
// can count one crossing twice or miss it, so three rays in independent
// random directions (uniform on the sphere) vote and the majority decides:
// one unlucky ray no longer places a CHIEF point outside the body.
bool point_inside(const Mesh &m, const double *x, SplitMix64 &rng) {
    int votes = 0;
    for (int ray = 0; ray < 3; ray++) {
        double z = 2.0 * rng.uniform() - 1.0, ph = 2.0 * kPi * rng.uniform();
        double rr = std::sqrt(std::max(0.0, 1.0 - z * z));
        double d[3] = {rr * std::cos(ph), rr * std::sin(ph), z};
        int cnt = 0;
        for (const auto &e : m.elems)
            cnt += ray_hits(x, d, &m.pnts[3 * e.nodes[0]], &m.pnts[3 * e.nodes[1]], &m.pnts[3 * e.nodes[2]]);
        votes += cnt & 1;
    }
    return votes >= 2;
}

int points_inside(const Mesh &m, const float *pts, int n, uint64_t seed, int *out) {
    if (m.N() == 0 || m.E() == 0) BEM_FAIL("Not enough mesh information!");
    if (n < 0 || (n > 0 && (!pts || !out))) BEM_FAIL("bad point arguments");
    SplitMix64 rng(seed);
    for (int i = 0; i < n; i++) {
        const double x[3] = {pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]};
        out[i] = point_inside(m, x, rng) ? 1 : 0;
    }
    return EXIT_SUCCESS;
}

// mesh::genCHIEF (mesh.cu:1430-1483): rejection sampling in the bounding box.
int gen_chief(Mesh &m, int num, float threshold, uint64_t seed) {
    if (num < 0) BEM_FAIL("negative CHIEF count");
    if (m.N() == 0 || m.E() == 0) BEM_FAIL("Not enough mesh information!");
    if (!m.bb_set) HOST_CALL(find_bb(m, 0.0001f));
    SplitMix64 rng(seed);
    std::vector<float> out;
    long tries = 0;
    const long max_tries = 200000L + 1000L * num;
    while ((int)(out.size() / 3) < num) {
        if (++tries > max_tries) BEM_FAIL("CHIEF generation found only %d of %d points", (int)(out.size() / 3), num);
        double x[3];
        for (int d = 0; d < 3; d++) x[d] = m.bb[2 * d] + (m.bb[2 * d + 1] - m.bb[2 * d]) * rng.uniform();
        double dmin = 1e300;
        for (int i = 0; i < m.N(); i++) {
            double dx = x[0] - m.pnts[3 * i], dy = x[1] - m.pnts[3 * i + 1], dz = x[2] - m.pnts[3 * i + 2];
            dmin = std::min(dmin, dx * dx + dy * dy + dz * dz);
        }
        if (std::sqrt(dmin) < threshold) continue;
        if (!point_inside(m, x, rng)) continue;
        for (int d = 0; d < 3; d++) out.push_back((float)x[d]);
    }
    m.chief.swap(out);
    m.version++;
    return EXIT_SUCCESS;
}

// Icosphere: icosahedron subdivided s times by edge midpoints, every vertex
// projected to the sphere at each level; wound inward like the reference's
// meshes (negative signed volume).
int gen_icosphere(Mesh &m, int s, float radius) {
    if (s < 0 || s > 8) BEM_FAIL("icosphere subdivision %d out of range [0, 8]", s);
    const double t = (1.0 + std::sqrt(5.0)) / 2.0;
    std::vector<double> v = {-1, t, 0, 1, t, 0, -1, -t, 0, 1, -t, 0, 0, -1, t, 0, 1, t,
                             0, -1, -t, 0, 1, -t, t, 0, -1, t, 0, 1, -t, 0, -1, -t, 0, 1};
    std::vector<int> f = {0, 11, 5, 0, 5, 1, 0, 1, 7, 0, 7, 10, 0, 10, 11, 1, 5, 9, 5, 11, 4,
                          11, 10, 2, 10, 7, 6, 7, 1, 8, 3, 9, 4, 3, 4, 2, 3, 2, 6, 3, 6, 8,
                          3, 8, 9, 4, 9, 5, 2, 4, 11, 6, 2, 10, 8, 6, 7, 9, 8, 1};
    auto proj = [](double *p) {
        double n = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
        p[0] /= n; p[1] /= n; p[2] /= n;
    };
    for (size_t i = 0; i < v.size(); i += 3) proj(&v[i]);
    for (int lev = 0; lev < s; lev++) {
        std::unordered_map<uint64_t, int> mid;
        std::vector<int> nf;
        nf.reserve(f.size() * 4);
        auto midpoint = [&](int a, int b) {
            uint64_t key = a < b ? ((uint64_t)a << 32) | (uint32_t)b : ((uint64_t)b << 32) | (uint32_t)a;
            auto it = mid.find(key);
            if (it != mid.end()) return it->second;
            double p[3] = {(v[3 * a] + v[3 * b]) / 2, (v[3 * a + 1] + v[3 * b + 1]) / 2, (v[3 * a + 2] + v[3 * b + 2]) / 2};
            proj(p);
            int id = (int)(v.size() / 3);
            v.insert(v.end(), p, p + 3);
            mid[key] = id;
            return id;
        };
        for (size_t i = 0; i < f.size(); i += 3) {
            int a = f[i], b = f[i + 1], c = f[i + 2];
            int ab = midpoint(a, b), bc = midpoint(b, c), ca = midpoint(c, a);
            int tri[12] = {a, ab, ca, b, bc, ab, c, ca, bc, ab, bc, ca};
            nf.insert(nf.end(), tri, tri + 12);
        }
        f.swap(nf);
    }
    m.pnts.resize(v.size());
    for (size_t i = 0; i < v.size(); i++) m.pnts[i] = (float)(v[i] * radius);
    m.elems.resize(f.size() / 3);
    for (size_t e = 0; e < f.size() / 3; e++) {
        bem_tri_elem &el = m.elems[e];
        // the table above is wound outward; swap two vertices to point the
        // element normals (p1-p3)x(p2-p3) into the body
        el.nodes[0] = f[3 * e];
        el.nodes[1] = f[3 * e + 2];
        el.nodes[2] = f[3 * e + 1];
        el.bc[0] = {0.f, 0.f};
        el.bc[1] = {1.f, 0.f};
        el.bc[2] = {-1.f, 0.f};
    }
    m.chief.clear();
    m.bb_set = false;
    m.version++;
    return EXIT_SUCCESS;
}

// Real spherical harmonic of degree l, order m (unnormalised):
// P_l^|m|(cos th) * (cos m ph | sin |m| ph).
static double real_ylm(int l, int mm, double ct, double ph) {
    int am = std::abs(mm);
    double st = std::sqrt(std::max(0.0, 1 - ct * ct));
    double pmm = 1.0;
    for (int i = 1; i <= am; i++) pmm *= -(2 * i - 1) * st;
    double plm;
    if (l == am) {
        plm = pmm;
    } else {
        double pm1 = ct * (2 * am + 1) * pmm;
        if (l == am + 1) {
            plm = pm1;
        } else {
            double a = pmm, b = pm1, c = 0;
            for (int ll = am + 2; ll <= l; ll++) {
                c = (ct * (2 * ll - 1) * b - (ll + am - 1) * a) / (ll - am);
                a = b;
                b = c;
            }
            plm = c;
        }
    }
    if (mm > 0) return plm * std::cos(am * ph);
    if (mm < 0) return plm * std::sin(am * ph);
    return plm;
}

int gen_ellipsoid(Mesh &m, int s, float ax, float ay, float az, int nb, int lmax, float amp, uint64_t seed) {
    HOST_CALL(gen_icosphere(m, s, 1.0f));
    if (nb < 0 || lmax < 1) BEM_FAIL("bad perturbation parameters");
    SplitMix64 rng(seed);
    std::vector<int> L(nb), Mo(nb);
    std::vector<double> a(nb), scale(nb);
    for (int b = 0; b < nb; b++) {
        L[b] = 1 + (int)(rng.uniform() * lmax);
        if (L[b] > lmax) L[b] = lmax;
        Mo[b] = -L[b] + (int)(rng.uniform() * (2 * L[b] + 1));
        if (Mo[b] > L[b]) Mo[b] = L[b];
        a[b] = (2 * rng.uniform() - 1) * amp / std::max(nb, 1);
        // normalise to max |Y| = 1 on a fine grid
        double mx = 0;
        for (int i = 0; i <= 180; i++) {
            double ct = std::cos(kPi * i / 180.0);
            for (int j = 0; j < 360; j++) mx = std::max(mx, std::fabs(real_ylm(L[b], Mo[b], ct, 2 * kPi * j / 360.0)));
        }
        scale[b] = mx > 0 ? 1.0 / mx : 0.0;
    }
    for (int i = 0; i < m.N(); i++) {
        double u[3] = {m.pnts[3 * i], m.pnts[3 * i + 1], m.pnts[3 * i + 2]};
        double nrm = std::sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
        for (double &c : u) c /= nrm;
        double ct = u[2], ph = std::atan2(u[1], u[0]);
        double f = 1.0;
        for (int b = 0; b < nb; b++) f += a[b] * scale[b] * real_ylm(L[b], Mo[b], ct, ph);
        m.pnts[3 * i] = (float)(ax * u[0] * f);
        m.pnts[3 * i + 1] = (float)(ay * u[1] * f);
        m.pnts[3 * i + 2] = (float)(az * u[2] * f);
    }
    m.version++;
    return EXIT_SUCCESS;
}

double signed_volume(const Mesh &m) {
    double vol = 0;
    for (const auto &e : m.elems) {
        const float *a = &m.pnts[3 * e.nodes[0]], *b = &m.pnts[3 * e.nodes[1]], *c = &m.pnts[3 * e.nodes[2]];
        vol += a[0] * ((double)b[1] * c[2] - (double)b[2] * c[1]) - a[1] * ((double)b[0] * c[2] - (double)b[2] * c[0]) +
               a[2] * ((double)b[0] * c[1] - (double)b[1] * c[0]);
    }
    return vol / 6.0;
}

// Orientation repair.  Within each edge-connected component, neighbouring
// triangles must run through their shared edge in opposite directions; a
// breadth-first walk from the component's first triangle flips every
// neighbour that runs through it in the same direction.  Each component is
// then flipped as a whole if its signed volume is positive, so its normals
// (p1 - p3) x (p2 - p3) point into the body like the reference meshes'.
// Edges shared by more than two triangles (non-manifold) propagate from the
// first two only.  Returns the number of elements whose winding changed.
int fix_orientation(Mesh &m, int *flipped) {
    HOST_CALL(validate_mesh(m));
    const int E = m.E();
    std::unordered_map<uint64_t, std::vector<int>> edge_elems;
    auto key = [](int a, int b) {
        return a < b ? ((uint64_t)(uint32_t)a << 32) | (uint32_t)b : ((uint64_t)(uint32_t)b << 32) | (uint32_t)a;
    };
    for (int e = 0; e < E; e++)
        for (int j = 0; j < 3; j++) edge_elems[key(m.elems[e].nodes[j], m.elems[e].nodes[(j + 1) % 3])].push_back(e);
    // does element e run from a to b along one of its edges?
    auto runs = [&](int e, int a, int b) {
        const int *n = m.elems[e].nodes;
        for (int j = 0; j < 3; j++)
            if (n[j] == a && n[(j + 1) % 3] == b) return true;
        return false;
    };
    auto flip = [&](int e) { std::swap(m.elems[e].nodes[1], m.elems[e].nodes[2]); };
    std::vector<int> comp(E, -1);
    std::vector<char> changed(E, 0);
    int ncomp = 0;
    for (int s = 0; s < E; s++) {
        if (comp[s] >= 0) continue;
        std::vector<int> members{s}, queue{s};
        comp[s] = ncomp;
        for (size_t h = 0; h < queue.size(); h++) {
            const int e = queue[h];
            for (int j = 0; j < 3; j++) {
                const int a = m.elems[e].nodes[j], b = m.elems[e].nodes[(j + 1) % 3];
                const std::vector<int> &nb = edge_elems[key(a, b)];
                for (size_t t = 0; t < nb.size() && t < 2; t++) {
                    const int f = nb[t];
                    if (f == e || comp[f] >= 0) continue;
                    comp[f] = ncomp;
                    if (runs(f, a, b)) {  // same direction as e: inconsistent
                        flip(f);
                        changed[f] ^= 1;
                    }
                    queue.push_back(f);
                    members.push_back(f);
                }
            }
        }
        double vol = 0;
        for (int e : members) {
            const float *a = &m.pnts[3 * m.elems[e].nodes[0]], *b = &m.pnts[3 * m.elems[e].nodes[1]],
                        *c = &m.pnts[3 * m.elems[e].nodes[2]];
            vol += a[0] * ((double)b[1] * c[2] - (double)b[2] * c[1]) -
                   a[1] * ((double)b[0] * c[2] - (double)b[2] * c[0]) + a[2] * ((double)b[0] * c[1] - (double)b[1] * c[0]);
        }
        if (vol > 0)
            for (int e : members) {
                flip(e);
                changed[e] ^= 1;
            }
        ncomp++;
    }
    int n = 0;
    for (int e = 0; e < E; e++) n += changed[e];
    if (n > 0) m.version++;
    if (flipped) *flipped = n;
    return EXIT_SUCCESS;
}

// trnglArea (mesh.cu:1075-1080) summed like cmptSurfArea (mesh.cu:1091-1117).
double surf_area(const Mesh &m) {
    double area = 0;
    for (const auto &e : m.elems) {
        const float *p1 = &m.pnts[3 * e.nodes[0]], *p2 = &m.pnts[3 * e.nodes[1]], *p3 = &m.pnts[3 * e.nodes[2]];
        double a[3] = {(double)p1[0] - p3[0], (double)p1[1] - p3[1], (double)p1[2] - p3[2]};
        double b[3] = {(double)p2[0] - p3[0], (double)p2[1] - p3[1], (double)p2[2] - p3[2]};
        double c[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]};
        area += std::sqrt(c[0] * c[0] + c[1] * c[1] + c[2] * c[2]) / 2;
    }
    return area;
}

// ------------------------------------------------------------ tile plan ---
static uint64_t spread3(uint64_t x) {
    x &= 0x1fffff;
    x = (x | x << 32) & 0x1f00000000ffffull;
    x = (x | x << 16) & 0x1f0000ff0000ffull;
    x = (x | x << 8) & 0x100f00f00f00f00full;
    x = (x | x << 4) & 0x10c30c30c30c30c3ull;
    x = (x | x << 2) & 0x1249249249249249ull;
    return x;
}

// Recursive coordinate bisection of nodes [lo, hi) of `ids` into `parts`
// parts; part ids are written to `part` starting at `first`.
static void rcb(const Mesh &m, std::vector<int> &ids, int lo, int hi, int parts, int first, std::vector<int> &part) {
    if (parts <= 1 || hi - lo <= 1) {
        for (int i = lo; i < hi; i++) part[ids[i]] = first;
        return;
    }
    float mn[3] = {1e30f, 1e30f, 1e30f}, mx[3] = {-1e30f, -1e30f, -1e30f};
    for (int i = lo; i < hi; i++)
        for (int d = 0; d < 3; d++) {
            mn[d] = std::min(mn[d], m.pnts[3 * ids[i] + d]);
            mx[d] = std::max(mx[d], m.pnts[3 * ids[i] + d]);
        }
    int ax = 0;
    for (int d = 1; d < 3; d++)
        if (mx[d] - mn[d] > mx[ax] - mn[ax]) ax = d;
    int pl = parts / 2;
    int mid = lo + (int)((long)(hi - lo) * pl / parts);
    std::nth_element(ids.begin() + lo, ids.begin() + mid, ids.begin() + hi,
                     [&](int a, int b) { return m.pnts[3 * a + ax] < m.pnts[3 * b + ax]; });
    rcb(m, ids, lo, mid, pl, first, part);
    rcb(m, ids, mid, hi, parts - pl, first + pl, part);
}

void build_tile_plan(const Mesh &m, int P, TilePlan &plan) {
    const int N = m.N(), E = m.E();
    P = std::max(1, std::min(P, N));
    std::vector<int> ids(N), part(N, 0);
    std::iota(ids.begin(), ids.end(), 0);
    rcb(m, ids, 0, N, P, 0, part);
    // rank of each node inside its part: Morton order of the coordinates
    float mn[3] = {1e30f, 1e30f, 1e30f}, mx[3] = {-1e30f, -1e30f, -1e30f};
    for (int i = 0; i < N; i++)
        for (int d = 0; d < 3; d++) {
            mn[d] = std::min(mn[d], m.pnts[3 * i + d]);
            mx[d] = std::max(mx[d], m.pnts[3 * i + d]);
        }
    std::vector<uint64_t> key(N);
    for (int i = 0; i < N; i++) {
        uint64_t k = 0;
        for (int d = 0; d < 3; d++) {
            double u = (m.pnts[3 * i + d] - mn[d]) / std::max(1e-30f, mx[d] - mn[d]);
            k |= spread3((uint64_t)(u * 2097151.0)) << d;
        }
        key[i] = k;
    }
    std::vector<std::vector<std::pair<uint64_t, int>>> lists(P);
    std::vector<char> used(N, 0);
    for (int e = 0; e < E; e++) {
        const int *nd = m.elems[e].nodes;
        int pe[3] = {part[nd[0]], part[nd[1]], part[nd[2]]};
        for (int j = 0; j < 3; j++) used[nd[j]] = 1;
        for (int j = 0; j < 3; j++) {
            bool seen = false;
            for (int i = 0; i < j; i++) seen |= (pe[i] == pe[j]);
            if (seen) continue;
            int flags = 0;
            uint64_t kmin = ~0ull;
            for (int i = 0; i < 3; i++)
                if (pe[i] == pe[j]) {
                    flags |= 1 << i;
                    kmin = std::min(kmin, key[nd[i]]);
                }
            if (pe[j] == pe[0]) flags |= 1 << 6;  // owner: part of the first vertex
            lists[pe[j]].push_back({kmin, (e << 8) | flags});
        }
    }
    plan.P = P;
    plan.ptr.assign(P + 1, 0);
    plan.entries.clear();
    plan.entries.reserve((size_t)E * 2);
    std::vector<char> touched(N, 0);
    for (int p = 0; p < P; p++) {
        auto &L = lists[p];
        std::stable_sort(L.begin(), L.end(), [](const std::pair<uint64_t, int> &a, const std::pair<uint64_t, int> &b) {
            return a.first < b.first;
        });
        for (auto &kv : L) {
            int ent = kv.second;
            int e = ent >> 8;
            for (int j = 0; j < 3; j++)
                if ((ent >> j) & 1) {
                    int nd = m.elems[e].nodes[j];
                    if (!touched[nd]) {
                        touched[nd] = 1;
                        ent |= 1 << (3 + j);
                    }
                }
            plan.entries.push_back(ent);
        }
        plan.ptr[p + 1] = (int)plan.entries.size();
    }
    plan.isolated.clear();
    for (int i = 0; i < N; i++)
        if (!used[i]) plan.isolated.push_back(i);
}

void build_ghost_plan(const Mesh &m, int P, GhostPlan &g) {
    const int N = m.N(), E = m.E();
    P = std::max(1, std::min(P, N));
    std::vector<int> ids(N), part(N, 0);
    std::iota(ids.begin(), ids.end(), 0);
    rcb(m, ids, 0, N, P, 0, part);
    float mn[3] = {1e30f, 1e30f, 1e30f}, mx[3] = {-1e30f, -1e30f, -1e30f};
    for (int i = 0; i < N; i++)
        for (int d = 0; d < 3; d++) {
            mn[d] = std::min(mn[d], m.pnts[3 * i + d]);
            mx[d] = std::max(mx[d], m.pnts[3 * i + d]);
        }
    std::vector<uint64_t> key(N);
    for (int i = 0; i < N; i++) {
        uint64_t k = 0;
        for (int d = 0; d < 3; d++) {
            double u = (m.pnts[3 * i + d] - mn[d]) / std::max(1e-30f, mx[d] - mn[d]);
            k |= spread3((uint64_t)(u * 2097151.0)) << d;
        }
        key[i] = k;
    }
    std::vector<std::vector<std::pair<uint64_t, int>>> lists(P);
    long halo = 0;
    for (int e = 0; e < E; e++) {
        const int *nd = m.elems[e].nodes;
        const int pe[3] = {part[nd[0]], part[nd[1]], part[nd[2]]};
        halo += 1 + (pe[1] != pe[0]) + (pe[2] != pe[0] && pe[2] != pe[1]);
        const int owner = (pe[1] == pe[2]) ? pe[1] : pe[0];
        uint64_t kmin = ~0ull;
        for (int j = 0; j < 3; j++)
            if (pe[j] == owner) kmin = std::min(kmin, key[nd[j]]);
        lists[owner].push_back({kmin, e});
    }
    g.dup = E > 0 ? (double)halo / E : 1.0;
    g.P = P;
    g.ptr.assign(P + 1, 0);
    g.entries.clear();
    g.slots.clear();
    g.entries.reserve(E);
    g.slots.reserve(3 * (size_t)E);
    g.nslots = 0;
    std::vector<char> touched(N, 0);
    std::vector<int> slot_node;           // node of each side column
    std::vector<int> local(N, -1);        // side column of a node in the current tile
    std::vector<int> local_nodes;
    for (int p = 0; p < P; p++) {
        auto &L = lists[p];
        std::stable_sort(L.begin(), L.end(), [](const std::pair<uint64_t, int> &a, const std::pair<uint64_t, int> &b) {
            return a.first < b.first;
        });
        for (auto &kv : L) {
            const int e = kv.second;
            int ent = (e << 8) | (1 << 6);
            int sl[3];
            for (int j = 0; j < 3; j++) {
                const int nd = m.elems[e].nodes[j];
                if (part[nd] == p) {
                    ent |= 1 << j;
                    sl[j] = -1;
                    if (!touched[nd]) {
                        touched[nd] = 1;
                        ent |= 1 << (3 + j);
                    }
                } else {
                    if (local[nd] < 0) {
                        local[nd] = g.nslots++;
                        slot_node.push_back(nd);
                        local_nodes.push_back(nd);
                        ent |= 1 << (3 + j);
                    }
                    sl[j] = local[nd];
                }
            }
            g.entries.push_back(ent);
            for (int j = 0; j < 3; j++) g.slots.push_back(sl[j]);
        }
        g.ptr[p + 1] = (int)g.entries.size();
        for (int nd : local_nodes) local[nd] = -1;
        local_nodes.clear();
    }
    // side columns per node in the order they were made (tile order)
    std::vector<std::vector<int>> per(N);
    for (int sidx = 0; sidx < g.nslots; sidx++) per[slot_node[sidx]].push_back(sidx);
    g.gptr.assign(1, 0);
    g.gnode.clear();
    g.gown.clear();
    g.gslot.clear();
    for (int nd = 0; nd < N; nd++) {
        if (per[nd].empty()) continue;
        g.gnode.push_back(nd);
        g.gown.push_back(touched[nd] ? 1 : 0);
        for (int sidx : per[nd]) g.gslot.push_back(sidx);
        g.gptr.push_back((int)g.gslot.size());
    }
}

}  // namespace bem
