/*
 * bem.h -- C-ABI of the MI355X boundary-element Helmholtz solver.
 *
 * Drop-in boundary for the hot path of ZiqiFrancisFan/BEM_GPU_Clean_CPP:
 * OBJ mesh load -> per-frequency collocation assembly (optional CHIEF rows)
 * -> least-squares QR or GMRES solve -> exterior field.  Each entry point below
 * names the reference function it replaces (file:line).  All functions use
 * plain pointers and sizes; the POD structs are bit-compatible with the
 * reference's cuFloatComplex (8 B), cartCoord (12 B) and triElem (40 B).
 *
 * Conventions kept from the reference:
 *   - matrices are column-major, A is M x N with lda >= M, rows are the N mesh
 *     nodes followed by the C CHIEF points, columns are the nodes
 *     (numerical.h:44-46, mesh.cu:993-999);
 *   - the caller owns every host buffer; nothing keeps the caller's pointers;
 *   - functions return EXIT_SUCCESS (0) or EXIT_FAILURE (1) and print
 *     "Error at file:line" on failure (numerical.h:60-122); bem_last_error()
 *     returns the message of the last failure in this thread, each call
 *     site in front of the message of the call that failed inside it;
 *     arguments are checked before any device work;
 *   - time convention e^{+i omega t}, G = exp(-ikr)/(4 pi r) (numerical.cu:70-74);
 *   - air density 1.2041 and speed of sound 343.21 by default (numerical.cu:12,14),
 *     the source strength of the incident point source is 0 by default
 *     (numerical.h:128-130).
 * Extensions the reference does not have are marked "extension".
 */
#ifndef BEM_H
#define BEM_H

#ifdef __cplusplus
extern "C" {
#endif

/* cuFloatComplex (cuComplex.h): interleaved float {x = real, y = imag}. */
typedef struct __attribute__((aligned(8))) bem_complex {
    float x, y;
} bem_complex;

/* Double-precision complex, extension (the FP64 solve path). */
typedef struct __attribute__((aligned(16))) bem_zcomplex {
    double x, y;
} bem_zcomplex;

/* cartCoord (mesh.h:26-99): 3 floats, 12 bytes. */
typedef struct bem_cart_coord {
    float coords[3];
} bem_cart_coord;

/* triElem (mesh.h:140-209): 0-based node indices and the boundary condition
 * ca*p + cb*v = cc as bc[0..2] = (ca, cb, cc); 40 bytes. */
typedef struct bem_tri_elem {
    int nodes[3];
    bem_complex bc[3];
} bem_tri_elem;

/* mesh (mesh.h:216-259): nodes, elements, CHIEF points, bounding box. */
typedef struct bem_mesh bem_mesh;

/* ---------------------------------------------------------------- mesh --- */

/* mesh::mesh() (mesh.h:244). */
bem_mesh *bem_mesh_create(void);
/* mesh::~mesh() (mesh.cu:1352-1356). */
void bem_mesh_destroy(bem_mesh *m);
/* mesh::readObj (mesh.cu:1215-1265): "v x y z" vertices and "f a b c" faces
 * (1-based), default bc (ca, cb, cc) = (0, 1, -1).  Unlike the reference, "vn"
 * and "vt" lines are not read as vertices and "f a/b/c" faces are accepted. */
int bem_mesh_read_obj(bem_mesh *m, const char *file);
/* Extension: set nodes and elements from arrays (copies them). */
int bem_mesh_set(bem_mesh *m, const bem_cart_coord *pnts, int numPnts, const bem_tri_elem *elems,
                 int numElems);
/* Extension: set the same boundary condition ca*p + cb*v = cc on every element. */
int bem_mesh_set_bc(bem_mesh *m, bem_complex ca, bem_complex cb, bem_complex cc);
/* mesh::findBB (mesh.cu:1369-1409): bounding box widened by threshold. */
int bem_mesh_find_bb(bem_mesh *m, float threshold);
/* Extension: the bounding box found by bem_mesh_find_bb as {xl,xu,yl,yu,zl,zu}. */
int bem_mesh_get_bb(const bem_mesh *m, float *bb6);
/* mesh::genCHIEF (mesh.cu:1430-1483): rejection-sample num points inside the
 * bounding box that are inside the surface (odd ray-crossing count) and at least
 * threshold away from every node.  The reference reseeds cuRAND per candidate
 * and casts the ray along the same uniforms; here a SplitMix64 stream seeded by
 * `seed` draws the position and three independent ray directions on the
 * sphere, and the majority of their crossing parities decides. */
int bem_mesh_gen_chief(bem_mesh *m, int num, float threshold, unsigned long long seed);
/* Extension: set the CHIEF points explicitly (copies them). */
int bem_mesh_set_chief(bem_mesh *m, const bem_cart_coord *pnts, int num);
/* mesh::getNumPnts / getNumElems / getNumChief (mesh.h:256-258). */
int bem_mesh_num_pnts(const bem_mesh *m);
int bem_mesh_num_elems(const bem_mesh *m);
int bem_mesh_num_chief(const bem_mesh *m);
/* Extension: copy out nodes / elements / CHIEF points (arrays sized by the getters). */
int bem_mesh_get_pnts(const bem_mesh *m, bem_cart_coord *out);
int bem_mesh_get_elems(const bem_mesh *m, bem_tri_elem *out);
int bem_mesh_get_chief(const bem_mesh *m, bem_cart_coord *out);
/* Extension: icosphere of radius r at subdivision s (10*4^s+2 nodes), wound
 * like the reference meshes (normals into the body, negative signed volume). */
int bem_mesh_gen_icosphere(bem_mesh *m, int subdiv, float radius);
/* Extension: perturbed ellipsoid: icosphere s mapped to semi-axes (ax,ay,az),
 * radial factor 1 + sum_b a_b Y_b(u) over nbumps seeded real spherical
 * harmonics up to degree lmax, sum |a_b| <= amp. */
int bem_mesh_gen_ellipsoid(bem_mesh *m, int subdiv, float ax, float ay, float az, int nbumps,
                           int lmax, float amp, unsigned long long seed);
/* Extension: signed volume (negative for the reference's winding). */
double bem_mesh_signed_volume(const bem_mesh *m);
/* Extension: make the winding consistent across every edge-connected part
 * of the surface and point each part's normals (p1-p3)x(p2-p3) into the body
 * (negative signed volume), as the reference's meshes are wound; *flipped
 * (may be NULL) receives the number of elements whose winding changed. */
int bem_mesh_fix_orientation(bem_mesh *m, int *flipped);
/* Extension: the CHIEF generator's inside test (inObj, numerical.cu:119-131)
 * for n points: inside[i] = 1 when the majority of three rays in seeded
 * random directions cross the surface an odd number of times. */
int bem_mesh_points_inside(const bem_mesh *m, const bem_cart_coord *pts, int n, unsigned long long seed,
                           int *inside);
/* cmptSurfArea (mesh.cu:1091-1117), computed on the host. */
float bem_mesh_surf_area(const bem_mesh *m);

/* ------------------------------------------------------- global set-up --- */

/* gaussQuad::gaussQuad(order) + sendToDevice (numerical.cu:152-156,227-231):
 * Gauss-Legendre tables for the 6x6 collapsed rule.  Order 6 (INTORDER,
 * numerical.h:40-42) is the only order the GPU kernels are built for. */
int bem_quad_init(int order);
/* Extension: copy the current nodes/weights (order values each). */
int bem_quad_get(double *pnts, double *wgts);
/* Extension: medium constants (the reference's __constant__ density/speed). */
int bem_set_medium(float density, float speed);
/* Extension: strength q of the incident point source (STRENGTH, numerical.h:128-130). */
int bem_set_source_strength(float q);
/* Extension: FP64 near-singular integration.  The reference integrates every
 * element that does not hold the collocation point with the 6x6 rule
 * (g_h_c_l_nsgl, mesh.cu:597-672; atomicFuncs.cu:75-76 excludes only the
 * element's own vertices), whose error reaches 1e-3 for a point a quarter of
 * the element size away.  With factor > 0, every (row, element) pair whose
 * distance is below factor times the element's longest edge is integrated in
 * FP64 by adaptive subdivision (midpoint splits, at most max_level deep) and
 * its entries of A, the diagonal and the boundary-condition term are
 * corrected after assembly, in a fixed order per row.  Applies to every later
 * assembly (FP32 and FP64, collocation matrix and field operator; not to
 * bem_ext_pressure and bem_solver_field).  factor = 0 (default) keeps the
 * reference's rule everywhere; max_level default 12. */
int bem_set_near_field(float factor, int max_level);

/* -------------------------------------------------- reference entry points --- */

/* bemSystem (mesh.h:222-223, mesh.cu:985-1006) == atomicGenSystem
 * (atomicFuncs.cu:218-298): assemble A (M x N, lda) and B (M x numSrcs, ldb)
 * for wavenumber k; B[:, s] = q*G(x, srcs[s]) - sum_e (cc/cb) g_e(x).
 * A and B are host buffers.  Assembly is deterministic (no atomics). */
int bem_system(const bem_mesh *m, float k, const bem_cart_coord *srcs, int numSrcs, bem_complex *A,
               int lda, bem_complex *B, int ldb);
/* atomicGenSystem (atomicFuncs.h:39-42): the same on raw arrays; pnts holds
 * the numNods nodes followed by the numCHIEF CHIEF points. */
int bem_gen_system(float k, const bem_tri_elem *elems, int numElems, const bem_cart_coord *pnts,
                   int numNods, int numCHIEF, const bem_cart_coord *srcs, int numSrcs, bem_complex *A,
                   int lda, bem_complex *B, int ldb);
/* qrSolver (numerical.h:196-197, numerical.cu:580-644): least squares
 * min ||A X - B|| for mA >= nA by Householder QR; the solution overwrites
 * B[0:nA, 0:nB].  A is not modified. */
int bem_qr_solve(const bem_complex *A, int mA, int nA, int ldA, bem_complex *B, int nB, int ldB);
/* lsqSolver (numerical.h:189-190, numerical.cu:506-554): the same least
 * squares for m > n; A is overwritten with R (upper triangle), and if Q is not
 * NULL it receives the explicit m x m unitary factor (ld m). */
int bem_lsq_solve(bem_complex *A, int m, int n, int lda, bem_complex *B, int nrhs, int ldb,
                  bem_complex *Q);
/* GMRES (GMRES.h:104-105, GMRES.cu:145-279): square m x m (lda = m) system,
 * unrestarted, initial guess in x, stops when the residual estimate
 * |err|/||b|| < threshold or after maxIter steps; solution in x.  The Krylov
 * basis is kept in HBM in FP64 ((min(maxIter, m) + 1) * m * 16 bytes) and the
 * packed Hessenberg matrix on the host (about min(maxIter, m)^2 * 8 bytes);
 * only when the basis would exceed half of the free HBM is the space
 * restarted at the largest length that fits. */
int bem_gmres(const bem_complex *A, const bem_complex *b, int m, int maxIter, float threshold,
              bem_complex *x);
/* genExtPressure (mesh.h:425-426, mesh.cu:1038-1073): pressure at one
 * exterior point x from the nodal surface pressure:
 * q*G(x, src) - sum_e [sum_j (h_j - (ca/cb) g_j) p_j + (cc/cb) sum_j g_j]. */
bem_complex bem_ext_pressure(float k, const bem_mesh *m, bem_cart_coord src, bem_cart_coord x,
                             const bem_complex *surfPressure);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): bem_ext_pressure for P points at once with every
 * temporary in a static device scratch array of 32 MiB per GPU, held until
 * the process ends; nothing is allocated or freed.  xs (P points), the nodal
 * surfPressure (N values) and out (P values) are DEVICE pointers; the work is
 * queued on `stream` (a hipStream_t; NULL: the library's stream) and the call
 * returns without waiting.  The incident point source at src has the strength
 * set by bem_set_source_strength.  The mesh must already be resident: call
 * bem_ext_pressure or bem_system on it once first (those allocate its device
 * records).  P must satisfy 16 P + 8 P ceil(E / 64) + 512 <= 32 MiB (about
 * 3200 points at E = 81920), else EXIT_FAILURE with a message.  Every call on
 * a GPU shares that GPU's scratch: calls that overlap on one GPU race.  Per
 * point, the result is bit-identical to bem_ext_pressure's. */
int bem_ext_pressure_noalloc(float k, const bem_mesh *m, bem_cart_coord src, const bem_cart_coord *xs, int P,
                             const bem_complex *surfPressure, bem_complex *out, void *stream);
/* wrtCplxMat (numerical.h:193-194, numerical.cu:556-578): "(%f,%f) " rows. */
int bem_wrt_cplx_mat(const bem_complex *mat, int numRows, int numCols, int ldm, const char *file);
/* printComplexMatrix (numerical.cu:81-93) on stdout. */
void bem_print_complex_matrix(const bem_complex *A, int row, int col, int lda);

/* ------------------------------------ device-resident solver (extension) --- */
/* A solver keeps the mesh, the assembled matrix and the Krylov state in HBM
 * between calls; inputs/outputs cross PCIe only when asked. */
typedef struct bem_solver bem_solver;

enum { BEM_PREC_FP32 = 32, BEM_PREC_FP64 = 64 };
enum { BEM_RHS_POINT = 0, BEM_RHS_PLANE = 1 };
enum { BEM_SOLVE_QR = 0, BEM_SOLVE_GMRES = 1, BEM_SOLVE_CGLS = 2, BEM_SOLVE_BLOCK_CG = 3 };

typedef struct bem_solve_stats {
    int method;           /* BEM_SOLVE_QR, BEM_SOLVE_GMRES, BEM_SOLVE_CGLS or BEM_SOLVE_BLOCK_CG */
    int iterations;       /* total inner Krylov iterations (all refinement steps) */
    int refinements;      /* outer FP64 refinement steps */
    int converged;        /* 1 when the requested tolerance was reached */
    double residual;      /* final relative FP64 residual of the (normal) equations */
    double assemble_ms;   /* last assembly, device time */
    double solve_ms;      /* last solve, device time from its first to its last work */
} bem_solve_stats;

/* Create a solver for the mesh (nodes, elements, CHIEF points are copied to HBM). */
bem_solver *bem_solver_create(const bem_mesh *m);
void bem_solver_destroy(bem_solver *s);
/* Assemble A and the boundary-condition part of B at wavenumber k in FP32
 * (complex64 A) or FP64 (complex128 A).  Returns once the work is queued on
 * the solver's stream; every later call on the solver is ordered after it,
 * and assemble_ms is filled in when the stats are read or a solve ends. */
int bem_solver_assemble(bem_solver *s, float k, int precision);
/* Right-hand sides: kind BEM_RHS_POINT (params = x,y,z per source, incident
 * q*G(x, src)) or BEM_RHS_PLANE (params = unit direction per wave, incident
 * amp*exp(-i k d.x)); nrhs <= 64. */
int bem_solver_set_rhs(bem_solver *s, int kind, const float *params, int nrhs, float amp);
/* Solve for all right-hand sides.  method BEM_SOLVE_GMRES: restarted GMRES
 * (restart, tol relative) on the square system, or on the normal equations
 * A^H A x = A^H b when CHIEF rows make it overdetermined; the FP32 path is
 * wrapped in FP64 residual refinement until refine_tol (<= 0: one pass; a
 * refinement cycle stops at 0.1 refine_tol when that is looser than tol).
 * method BEM_SOLVE_QR: blocked Householder least squares (FP32 path also
 * refined in FP64).  method BEM_SOLVE_CGLS: conjugate gradients on the
 * normal equations (restart unused), for the overdetermined CHIEF systems,
 * with the same FP64 refinement, each right-hand side on its own.
 * method BEM_SOLVE_BLOCK_CG: the same normal equations solved by block
 * conjugate gradients over all right-hand sides at once (one shared block
 * Krylov space, orthonormalised residual block), with the same refinement;
 * falls back to BEM_SOLVE_CGLS's iteration when the block breaks down (e.g.
 * two equal right-hand sides), and takes that iteration from the start when
 * M * N * nrhs < 4e8, where a block step's fixed cost outweighs its fewer
 * steps (environment BEM_BLOCK_CG_ALWAYS: the block at every size).
 * iterations counts block steps. */
int bem_solver_solve(bem_solver *s, int method, int restart, float tol, int maxit, double refine_tol,
                     int max_refine, bem_solve_stats *stats);
/* Copy the nodal solution (N x nrhs, ldx) to the host. */
int bem_solver_get_solution(const bem_solver *s, bem_complex *x, int ldx);
int bem_solver_get_solution_z(const bem_solver *s, bem_zcomplex *x, int ldx);
/* Copy the assembled A (M x N) / B (M x nrhs) to the host as complex64. */
int bem_solver_get_system(const bem_solver *s, bem_complex *A, int lda, bem_complex *B, int ldb);
/* Rows rows[0..nrows) of the assembled A as complex128, row-major (nrows x N):
 * spot checks of very large systems without copying all of A. */
int bem_solver_get_rows(const bem_solver *s, const int *rows, int nrows, bem_zcomplex *out);
/* Exterior field at P points for every right-hand side: out is P x nrhs (ldo). */
int bem_solver_field(bem_solver *s, const bem_cart_coord *pts, int P, bem_complex *out, int ldo);
/* Exterior field as a matrix (genExtPressure, mesh.cu:1038-1073, for many
 * points and solutions at once): assemble the P x N radiation operator F of
 * the field points at the solver's current wavenumber with the collocation
 * kernel (the field points are rows off the surface, like CHIEF rows), and
 * keep it in HBM.  bem_solver_field_apply then evaluates
 * out (P x nrhs, ldo) = incident - boundary term - F x for every current
 * solution with one MFMA product; bem_solver_get_field_matrix copies F
 * (complex64, column-major P x N, ldf >= P) to the host. */
int bem_solver_field_matrix(bem_solver *s, const bem_cart_coord *pts, int P);
int bem_solver_field_apply(bem_solver *s, bem_complex *out, int ldo);
int bem_solver_get_field_matrix(const bem_solver *s, bem_complex *F, int ldf);
/* Sizes: M rows, N columns, current nrhs. */
int bem_solver_dims(const bem_solver *s, int *M, int *N, int *nrhs);
/* Last solve statistics. */
int bem_solver_stats(const bem_solver *s, bem_solve_stats *stats);
/* Point the solver at another mesh, keeping its stream; device buffers are
 * reallocated (stream-ordered) to the new sizes. */
int bem_solver_set_mesh(bem_solver *s, const bem_mesh *m);
/* Free every device buffer of the solver but keep its stream; a later
 * bem_solver_set_mesh makes it usable again. */
int bem_solver_release(bem_solver *s);

/* ------------------------------------- batches and serving (extension) --- */
/* Solve settings shared by the batch and server calls (the arguments of
 * bem_solver_solve). */
typedef struct bem_solve_params {
    int method;         /* BEM_SOLVE_GMRES, BEM_SOLVE_QR, BEM_SOLVE_CGLS or BEM_SOLVE_BLOCK_CG */
    int restart;
    float tol;
    int maxit;
    double refine_tol;
    int max_refine;
} bem_solve_params;

/* One mesh at nk wavenumbers k[f], each with nrhs right-hand sides: the same
 * 3 * nrhs params for every frequency, or (params_per_freq != 0) nk blocks of
 * 3 * nrhs, one per frequency.  X (may be null) receives N x (nrhs * nk) complex128 solutions, column
 * f * nrhs + r at X + (f * nrhs + r) * ldx; stats (nk entries, may be null)
 * each frequency's solve statistics.  `workers` resident solvers (<= 8), each
 * on its own stream and host thread, take frequencies in turn, highest
 * wavenumber first.  sp == NULL: GMRES(30, 1e-6) refined to 1e-8.
 * Results equal those of one solver running the frequencies one by one.
 * Environment BEM_BATCH_FUSE = g (2..4): square FP32 systems (no CHIEF points)
 * solved with GMRES and at least 2 right-hand sides go in groups of g
 * frequencies (at most 64 systems, and within half the free HBM for the
 * workers' matrices) that run one lockstep GMRES, every Arnoldi step one
 * product launch over the group's matrices; a group's stats are those of its
 * joint solve, and its results agree with the one-by-one solves to the
 * refinement tolerance, independent of the worker count. */
int bem_batch_solve(const bem_mesh *m, const float *k, int nk, int precision, int rhs_kind, const float *params,
                    int nrhs, int params_per_freq, float amp, const bem_solve_params *sp, int workers,
                    bem_zcomplex *X, int ldx, bem_solve_stats *stats);

/* A batch context: the `workers` (1..8) resident solvers of bem_batch_solve,
 * created once and kept, so repeated batches of one mesh skip the solvers'
 * set-up (the mesh's device records, the matrices, the solve workspaces).
 * The mesh is copied.  bem_batch_run takes the arguments of bem_batch_solve
 * and gives the same results; min(workers, nk) solvers take part; one run at
 * a time per context (a second caller waits).  bem_batch_solve is
 * bem_batch_create + bem_batch_run + bem_batch_destroy.  Extension: the
 * reference solves frequencies one at a time in its driver loop. */
typedef struct bem_batch bem_batch;
bem_batch *bem_batch_create(const bem_mesh *m, int workers);
int bem_batch_run(bem_batch *b, const float *k, int nk, int precision, int rhs_kind, const float *params, int nrhs,
                  int params_per_freq, float amp, const bem_solve_params *sp, bem_zcomplex *X, int ldx,
                  bem_solve_stats *stats);
void bem_batch_destroy(bem_batch *b);

/* Request server: `slots` resident solvers (<= 16) on their own streams take
 * queued requests in arrival order as soon as a slot is free and the request's
 * HBM footprint fits in the budget left by the running ones
 * (hbm_budget_bytes <= 0: 90% of the free device memory). */
typedef struct bem_server bem_server;

typedef struct bem_request_info {
    long id;
    int status;                 /* 0 solved, 1 failed or cancelled */
    int slot;                   /* slot that ran it */
    int N, M, nrhs;
    double footprint_bytes;     /* HBM admitted for it */
    double submit_ms, start_ms, end_ms;  /* ms since the server was created */
    bem_solve_stats solve;
} bem_request_info;

typedef struct bem_server_info {
    int queued, running, max_running, slots;
    double budget_bytes, in_use_bytes, peak_in_use_bytes;
} bem_server_info;

bem_server *bem_server_create(int slots, double hbm_budget_bytes);
/* Queue one request: the mesh is copied (the caller may free it at once).
 * Returns the request id (> 0), or -1 (bad arguments, or a footprint above
 * the budget). */
long bem_server_submit(bem_server *s, const bem_mesh *m, float k, int precision, int rhs_kind, const float *params,
                       int nrhs, float amp, const bem_solve_params *sp);
/* Block until request `id` ends; copy its N x nrhs solution (x may be null)
 * and its record, and forget it. */
int bem_server_wait(bem_server *s, long id, bem_zcomplex *x, int ldx, bem_request_info *info);
int bem_server_status(bem_server *s, bem_server_info *info);
/* Cancel queued requests (their waiters return EXIT_FAILURE), let running
 * ones finish, wait until every thread inside bem_server_wait has returned,
 * then free everything.  No call may start on the server once this one has. */
void bem_server_destroy(bem_server *s);
/* HBM bytes a request of these sizes is admitted with (upper bound of what
 * its solve allocates). */
double bem_request_footprint(int M, int N, int E, int nrhs, int precision, int method, int restart);

/* ------------------------------------------------ launch tuning (extension) --- */
/* The assembly's launch shape: lanes per collocation row (1 or 4) and the
 * number of column tiles, given as the total wave count of the launch.  The
 * defaults suit most sizes; bem_autotune_assembly times candidate shapes for
 * this mesh at wavenumber k on the device and keeps the fastest for meshes of
 * its size class (rows and elements in half-octave buckets) in this process;
 * solvers and device problems created afterwards use it.  bem_tune_save /
 * bem_tune_load keep the table in a text file. */
typedef struct bem_tune_result {
    int lanes_per_row;
    int waves;
    double best_ms;      /* assembly time with the chosen shape */
    double default_ms;   /* assembly time with the default shape */
} bem_tune_result;
int bem_autotune_assembly(const bem_mesh *m, float k, int precision, bem_tune_result *out);
/* The multi-right-hand-side MFMA products the Krylov and CG solvers run
 * (Y = A X and Z = A^H T, FP32, refined): for an M x N matrix and nrhs
 * right-hand sides, times every candidate shape -- the kernel family (16 or
 * 32 right-hand sides per wave) and the split-K chunk count (1-16) -- on the
 * device with a synthetic matrix, and keeps the fastest per product for the
 * size class (M and N in half-octave buckets, nrhs in steps of 16) in this
 * process; saved and loaded with the assembly shapes by bem_tune_save /
 * bem_tune_load. */
typedef struct bem_product_tune_result {
    int rhs_per_wave_n, chunks_n;  /* Y = A X */
    int rhs_per_wave_h, chunks_h;  /* Z = A^H T */
    double best_ms_n, default_ms_n, best_ms_h, default_ms_h;
} bem_product_tune_result;
int bem_autotune_products(int M, int N, int nrhs, bem_product_tune_result *out);
int bem_tune_save(const char *file);
int bem_tune_load(const char *file);
int bem_tune_clear(void);

/* ----------------------------------------------------- timing (extension) --- */
/* HIP-event timing of the library's kernels on the stream they run on.
 * bem_timing_enable(1) starts recording; bem_timing_get returns the total
 * milliseconds and launch count recorded for a kernel name since the last
 * bem_timing_reset. */
int bem_timing_enable(int on);
int bem_timing_reset(void);
int bem_timing_get(const char *kernel, double *total_ms, long *count);
/* Synchronise the library's stream. */
int bem_sync(void);

/* Message of the last failure in this thread ("" if none). */
const char *bem_last_error(void);
/* Library version string. */
const char *bem_version(void);

#ifdef __cplusplus
}
#endif
#endif /* BEM_H */
