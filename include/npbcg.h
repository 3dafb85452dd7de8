/*
 * npbcg — MI355X (gfx950) implementation of the NPB CG conj_grad path.
 *
 * C ABI.  Plain pointers and sizes; no C++ or torch types.  Every entry point
 * returns 0 (NPBCG_OK) or a negative npbcg_status and never exits the process
 * (the reference calls exit() on every failure).  A context is used by one
 * host thread at a time and owns one HIP stream.
 *
 * Reference citations are GMAP/NPB-GPU file:line (CUDA/CG/cg.cu; HIP/CG/cg.cpp
 * is the same program with hip* names).
 */
#ifndef NPBCG_H
#define NPBCG_H

#ifdef __cplusplus
extern "C" {
#endif

#define NPBCG_ABI_VERSION 3

typedef enum npbcg_status {
    NPBCG_OK = 0,
    NPBCG_ERR_INVALID = -1,   /* bad argument: NULL pointer, n < 1, malformed CSR */
    NPBCG_ERR_NOMEM = -2,     /* host or device allocation failed */
    NPBCG_ERR_DEVICE = -3,    /* HIP runtime error; text via npbcg_last_error() */
    NPBCG_ERR_NODEVICE = -4,  /* no GPU visible (cg.cu:1424-1426 exits instead) */
    NPBCG_ERR_CAPACITY = -5,  /* nz capacity exceeded (cg.cu:1670-1674 exits instead) */
    NPBCG_ERR_INTERNAL = -6,  /* generator insertion failure (cg.cu:1748-1751 exits) */
    NPBCG_ERR_MISMATCH = -7,  /* arrays differ from the matrix the context holds */
    NPBCG_ERR_NOCONV = -8     /* a solve stopped before reaching its tolerance */
} npbcg_status;

/* Inner-loop engines for the FP64 conj_grad (npbcg_conj_grad,
 * npbcg_conj_grad_device, npbcg_power).  Both run the one-reduction CG step
 * (npbcg_opts.cg_form) and keep every scalar on the GPU; the reference issues
 * 6 launches and 2 blocking copies per CG step (cg.cu:814-837). */
enum {
    NPBCG_ENGINE_AUTO = 0,        /* KERNELS (measured faster: 8.5 ms against 14.6 ms for
                                     10 outer iterations at n = 75000) */
    NPBCG_ENGINE_KERNELS = 1,     /* fused per-phase kernels (two a CG step), replayed
                                     from a captured HIP graph per outer iteration */
    NPBCG_ENGINE_PERSISTENT = 2   /* one cooperative kernel runs whole outer iterations
                                     (all cgitmax steps, the residual, zeta and the
                                     normalisation; npbcg_power: every iteration in one
                                     launch) with grid barriers between its phases.
                                     Needs the LDS-staged layout with every workgroup
                                     resident: npbcg_create returns NPBCG_ERR_INVALID
                                     otherwise.  Solves ignore the engine. */
};

/* Vectors held by a context (npbcg_get_vector / npbcg_set_vector). */
enum {
    NPBCG_VEC_X = 0, NPBCG_VEC_Z = 1, NPBCG_VEC_P = 2, NPBCG_VEC_Q = 3,
    NPBCG_VEC_R = 4
};

typedef struct npbcg_opts {
    int device;       /* GPU ordinal; out of range -> 0, as GPU_DEVICE (cg.cu:1427-1432) */
    int cgitmax;      /* inner CG iterations per conj_grad; 0 -> 25 (cg.cu:666) */
    int engine;       /* NPBCG_ENGINE_* */
    int spmv_variant; /* 0 = auto; 1-4 CSR row kernels (64/32/16 lanes per row,
                         64 with 2 loads in flight); 5 = LDS-staged column blocks */
    int use_graph;    /* nonzero: replay each outer iteration as a captured HIP graph */
    int tile_rows;    /* LDS-staged SpMV: rows per tile, 0 = auto */
    int tile_width;   /* LDS-staged SpMV: vector entries per LDS block, 0 = auto */
    int fuse_p_update;/* LDS-staged SpMV: form p = r + beta p inside the next SpMV
                         instead of a separate launch (default 0) */
    int solve_batch;  /* npbcg_solve: CG steps launched before the first host check
                         (later checks use half); 0 -> 10 */
    int cg_form;      /* conj_grad / npbcg_power on the LDS-staged layout: 0 = auto (the
                         one-reduction form), 1 = the reference's step order (q = A p;
                         alpha; z, r; rho; beta; p, three launches a step), 2 = the
                         one-reduction form: s = A r, then one update forming p, q = A p,
                         p.q = r.s - beta rho / alpha_prev, z and r (two launches a step).
                         Solves always use the reference's step order. */
    int matrix_check; /* npbcg_conj_grad / npbcg_compat_conj_grad: how the host arrays
                         are checked against the context's matrix (NPBCG_CHECK_*) */
    int dot_pass;     /* one-reduction form: where r.s is formed.  0 = auto (in its own
                         pass from 24 column blocks up), 1 = in the SpMV, 2 = in its own
                         pass (the SpMV stores the block sums of s = A r only; a combine
                         kernel adds them into s and forms r.s; three launches a step) */
    int sort_window;  /* LDS-staged SpMV: rows of a workgroup's range sorted together by
                         their length in a column block before they are cut into 64-row
                         slices (a multiple of 64).  0 = auto: the whole range, within
                         2048 .. 8192 rows.  Wider windows pad fewer entries; narrower
                         ones keep a slice's block sums closer together. */
    int reserved[3];
} npbcg_opts;

/* Check of the host arrays handed to npbcg_conj_grad (npbcg_opts.matrix_check). */
enum {
    NPBCG_CHECK_FULL = 0,    /* rowstr compared and colidx/a digested on every call: an
                                in-place edit is seen (the serial conj_grad reads its
                                arrays on every call, cg.cu:703-709) */
    NPBCG_CHECK_SAMPLED = 1  /* rowstr compared, colidx/a by pointer or 62 sampled
                                entries: the reference GPU path's contract (the matrix is
                                uploaded once by setup_gpu, cg.cu:1390-1603, and the
                                caller does not edit it in place) */
};

/* Timing of the last npbcg_power / npbcg_conj_grad_device call, from HIP
 * events recorded on the context's stream. */
typedef struct npbcg_timing {
    double total_ms;        /* events around the whole call */
    double spmv_avg_ms;     /* mean duration of sampled q = A p SpMV launches */
    int spmv_samples;       /* how many SpMV launches were bracketed (0 = none) */
    int launches;           /* kernel launches (or graph launches) issued */
    long long nnz;          /* rowstr[n] of the context's matrix */
    int n;
    int engine;             /* engine actually used */
} npbcg_timing;

void npbcg_default_opts(npbcg_opts *opts);
/* Number of visible GPUs (the count setup_gpu reads, cg.cu:1421).  A host
 * without a GPU yields 0 and NPBCG_OK. */
int npbcg_device_count(int *count);
int npbcg_abi_version(void);
/* Text of the most recent failure in this thread ("" if none). */
const char *npbcg_last_error(void);
const char *npbcg_status_string(int status);

/* ---- problem generator (host) -------------------------------------------
 * Replaces makea (cg.cu:1303-1368) with sparse (:1611-1785), sprnvc
 * (:1798-1833), vecset (:1841-1857) and the RNG set-up of main (:406-408:
 * tran = seed, one discarded draw).  Output is the 0-based CSR after the
 * column shift of :434-438.  rowstr/colidx/a match the reference bit for bit:
 * duplicates are summed in the reference's order.  colidx/a need nz_cap
 * entries; the reference's capacity is n*(nonzer+1)^2 (cg.cu:81).
 * Returns nnz (>= 0) or a negative npbcg_status.  *ratio_out (nullable)
 * receives pow(rcond, 1/n) (cg.cu:1695). */
long npbcg_makea(int n, int nonzer, double rcond, double shift, double seed,
                 long nz_cap, int *rowstr, int *colidx, double *a,
                 double *ratio_out);

/* ---- context: device residency -------------------------------------------
 * Replaces setup_gpu (cg.cu:1390-1603): selects the device, uploads the CSR
 * once (nnz = rowstr[n]) and allocates x, z, p, q, r (n+2 entries each, as
 * cg.cu:105-109) on the device; x starts as all ones (cg.cu:441-443).
 * opts may be NULL (defaults). */
typedef struct npbcg_ctx npbcg_ctx;
int npbcg_create(npbcg_ctx **out, int n, const int *rowstr, const int *colidx,
                 const double *a, const npbcg_opts *opts);
/* Replaces release_gpu (cg.cu:1370-1388).  NULL is accepted. */
void npbcg_destroy(npbcg_ctx *ctx);

/* ---- the drop-in conj_grad -----------------------------------------------
 * Same arguments and post-conditions as the reference's
 *   static void conj_grad(int colidx[], int rowstr[], double x[], double z[],
 *       double a[], double p[], double q[], double r[], double* rnorm);
 * (cg.cu:225-233, body :653-801), with the context in front.  Host arrays:
 * x is read (n+1 entries); z (the solution of A z = x after cgitmax CG
 * steps), p, q, r (r = A z on exit) and *rnorm = ||x - A z|| are written,
 * n+1 entries each; index n gets q=z=0, r=p=x[n] as cg.cu:671-676 leaves it.
 * colidx/rowstr/a must describe the matrix the context was created with:
 * rowstr is compared in full and colidx/a digested in full (or, with
 * NPBCG_CHECK_SAMPLED, by pointer or sampled content); a mismatch, an
 * in-place edit included, returns NPBCG_ERR_MISMATCH. */
int npbcg_conj_grad(npbcg_ctx *ctx, int colidx[], int rowstr[], double x[],
                    double z[], double a[], double p[], double q[], double r[],
                    double *rnorm);

/* Replaces conj_grad_gpu (cg.cu:803-846): the same computation on the
 * context's device-resident x, z, p, q, r; only *rnorm returns to the host. */
int npbcg_conj_grad_device(npbcg_ctx *ctx, double *rnorm);

/* Replaces gpu_kernel_ten (cg.cu:1182-1201): x.z and z.z of the device
 * vectors.  Either output pointer may be NULL. */
int npbcg_norms(npbcg_ctx *ctx, double *norm_temp1, double *norm_temp2);

/* Replaces gpu_kernel_eleven (cg.cu:1247-1265): x = scale * z on the device. */
int npbcg_scale_x(npbcg_ctx *ctx, double scale);

/* The whole timed loop of cg.cu:500-520 on the device, from the context's
 * current x: niter times { conj_grad; zeta = shift + 1/(x.z);
 * x = z/||z|| }.  zeta_out/rnorm_out (niter entries each, nullable) receive
 * the per-iteration values printed at cg.cu:515-516; they are copied back
 * once at the end.  Nothing crosses to the host inside the loop. */
int npbcg_power(npbcg_ctx *ctx, int niter, double shift, double *zeta_out,
                double *rnorm_out);

/* Device vector access (n+1 entries; index n included). */
int npbcg_set_vector(npbcg_ctx *ctx, int which, const double *host);
int npbcg_get_vector(npbcg_ctx *ctx, int which, double *host);
/* x[0..n] = value, as cg.cu:441-443 and :485-487 do with 1.0. */
int npbcg_fill_x(npbcg_ctx *ctx, double value);

/* Timing of the last device call (see npbcg_timing).  Sampling SpMV launch
 * durations adds event records to the stream; it is off unless enabled. */
int npbcg_set_spmv_sampling(npbcg_ctx *ctx, int every_nth_call);
int npbcg_get_timing(const npbcg_ctx *ctx, npbcg_timing *out);

/* Device time of one kernel of the path, summed over a profiled run. */
typedef struct {
    char name[64];   /* kernel (and the product it forms, for the SpMV) */
    int threads;     /* workgroup size */
    int calls;       /* launches */
    double ms;       /* total device time */
} npbcg_kernel_time;

/* Profiled npbcg_power: the same niter outer iterations from the current x
 * (same results; zeta_out nullable), launched kernel by kernel without the
 * captured graph, with a HIP event recorded after every launch.  Each
 * launch's time is the gap between its event and the one before, so the
 * kernel times add up to *total_ms, the whole run.  out[0..*nkinds) receives
 * one entry per kernel in order of first launch (at most max_kinds; a larger
 * count returns NPBCG_ERR_INVALID).  Replaces the reference's PROFILING build
 * (cg.cu:83-94 timers around each launch, table printed at cg.cu:565-589).
 * Library calls also open roctx ranges named after the call
 * ("npbcg_power", "npbcg_solve", ...) for rocprofv3 --marker-trace. */
int npbcg_profile_kernels(npbcg_ctx *ctx, int niter, double shift, double *zeta_out,
                          npbcg_kernel_time *out, int max_kinds, int *nkinds,
                          double *total_ms);

/* The context's kernel configuration as a one-line JSON object (SpMV variant,
 * block/tile shape, workgroup counts, stored entries) written to buf
 * (NUL-terminated, truncated to len).  Returns the full length. */
int npbcg_describe(const npbcg_ctx *ctx, char *buf, int len);

/* One complete q = A p product with the context's kernels on its device
 * vectors (for the LDS-staged layout: the SpMV and the combine of its
 * per-block row sums), repeated reps times; *avg_ms gets the mean duration
 * from HIP events.  Used to measure SpMV bandwidth in isolation. */
int npbcg_spmv_bench(npbcg_ctx *ctx, int reps, double *avg_ms);

/* The CG SpMV kernel alone (no combine of the per-block row sums), reps
 * launches back to back as successive CG steps issue them (the LDS-staged
 * walk alternates direction with the step parity), each bracketed by HIP
 * events recorded by its own dispatch; *avg_ms gets the mean kernel duration.
 * precision 0: the FP64 kernel of conj_grad and FP64 solves (the reference's
 * k3, cg.cu:937-961); 1: the FP32 kernel of the mixed solve's inner CG (needs
 * the LDS-staged layout; builds the FP32 copy on first use). */
int npbcg_spmv_kernel_bench(npbcg_ctx *ctx, int precision, int reps, double *avg_ms);

/* Single-system autotuner, in place of the reference's compile-time launch
 * configuration (config/gpu.config:32-42, setparams.cpp:256-332): times the
 * q = A p SpMV of this matrix through temporary contexts (created with *opts,
 * other fields unchanged) for the LDS-staged layout at its default shape,
 * with twice the row ranges and with half-width column blocks, and for the
 * four CSR row kernels; writes the fastest into opts->spmv_variant,
 * tile_width and tile_rows, for the caller's npbcg_create.  The text file
 * cache_path (may be NULL) keeps one line per tuning, "<key> <variant>
 * <width> <rows> <ms>", keyed by n, nnz, a digest of the row lengths and the
 * CU count; a later call with the same key reads it instead of timing.  ms
 * (may be NULL) receives the chosen SpMV time. */
int npbcg_tune(int n, const int *rowstr, const int *colidx, const double *a, npbcg_opts *opts,
               const char *cache_path, double *ms);

/* A plain device-to-device copy of `bytes` (16-byte vector loads and
 * stores, grid-stride, four loads in flight per thread) repeated reps times on the context's device;
 * *avg_ms is the mean duration.  It has no counterpart in the reference: it
 * measures the achievable streaming bandwidth the SpMV is compared with, and
 * calibrates memory-traffic counters with a kernel of known traffic.  The
 * two buffers are allocated and freed inside the call. */
int npbcg_copy_bench(npbcg_ctx *ctx, long long bytes, int reps, double *avg_ms);
/* npbcg_copy_bench with a choice of stream: mode 0 the copy (bytes read and
 * bytes written), mode 1 a read of `bytes` only (one 1024-thread workgroup per
 * CU, eight 16-byte loads in flight per thread, no stores).  A size well above the 256 MB memory-side cache
 * measures HBM; one that fits measures that cache. */
int npbcg_stream_bench(npbcg_ctx *ctx, int mode, long long bytes, int reps, double *avg_ms);

/* ---- solve to a tolerance ---------------------------------------------------
 * A x = b from x = 0 until the FP64 true relative residual ||b - A x|| / ||b||
 * (recomputed with the FP64 matrix) is <= rtol.  The reference has no such
 * entry: it always runs 25 steps (cg.cu:666).  The CG steps are those of
 * conj_grad (cg.cu:689-772); the stop test and the restarts are this
 * library's.  The context's p, q, r are used as workspace; x and z of the
 * power loop are left alone.
 *   NPBCG_SOLVE_FP64   plain CG in FP64 with a device-side stop test; if the
 *                      true residual misses rtol, CG restarts from it.
 *   NPBCG_SOLVE_MIXED  FP64 iterative refinement around an inner CG whose
 *                      matrix, vectors and SpMV are FP32 (rho/alpha/beta, dot
 *                      products and the stop test FP64); the correction is
 *                      accumulated and the residual recomputed in FP64.
 *                      Needs the LDS-staged layout (spmv_variant 5). */
enum { NPBCG_SOLVE_FP64 = 0, NPBCG_SOLVE_MIXED = 1 };

/* Preconditioners M^-1 for npbcg_solve (PCG: rho = r.M^-1 r, p = M^-1 r +
 * beta p; the stop test uses |r.M^-1 r| <= rtol^2 |r0.M^-1 r0|, the final
 * check is always the FP64 true residual).  NEUMANN and LEARNED share one
 * sparse form on the K = 4 largest off-diagonals |a_jk| of every row:
 *   M^-1 = S D^-1 S - Q (N + N^T) Q,  N_jk = a_jk / (a_jj a_kk),
 *   S = diag(exp g), Q = diag(exp(g + h)),
 * applied as one CSR product (8 lanes per row) per CG step; the stop test
 * of these two stays on the recursive r.r, as in plain CG. */
enum {
    NPBCG_PRECOND_NONE = 0,     /* plain CG */
    NPBCG_PRECOND_JACOBI = 1,   /* M^-1 = 1 / diag(A) */
    NPBCG_PRECOND_LEARNED = 2,  /* g, h per row from npbcg_net_apply; falls back
                                   to Jacobi-CG if the solve fails */
    NPBCG_PRECOND_NEUMANN = 3   /* g = h = 0: D^-1 - D^-1 (E + E^T) D^-1, E the top-K
                                   off-diagonals (a Neumann-type approximation
                                   of A^-1 and the network's base) */
};

typedef struct npbcg_solve_stats {
    int converged;     /* 1 if the FP64 true relative residual <= rtol */
    int iterations;    /* CG steps (inner FP32 steps for MIXED) */
    int sweeps;        /* refinement sweeps, or restarts + 1 for FP64 */
    int fp64_spmvs;    /* FP64 matrix products: CG steps and true residuals */
    int fp32_spmvs;    /* FP32 matrix products (MIXED) */
    int precond;       /* preconditioner used (NPBCG_PRECOND_*) */
    int fallback;      /* 1 if a learned-preconditioner solve was redone with Jacobi */
    int reserved_i;
    double relres;     /* final ||b - A x|| / ||b|| in FP64 */
    double time_ms;    /* HIP events around the device work of the solve */
    double reserved_d[4];
} npbcg_solve_stats;

/* Uploads the right-hand side b (n entries) and records ||b||. */
int npbcg_set_rhs(npbcg_ctx *ctx, const double *b);
/* Solves with the resident b; maxit bounds the CG steps (inner steps for
 * MIXED).  Returns NPBCG_ERR_NOCONV (stats filled) if rtol was not reached. */
int npbcg_solve(npbcg_ctx *ctx, int mode, double rtol, int maxit,
                npbcg_solve_stats *stats);
/* Downloads the last solution (n entries). */
int npbcg_get_solution(npbcg_ctx *ctx, double *x);
/* Selects the preconditioner of later solves.  LEARNED needs npbcg_net_apply
 * on this context first (NPBCG_ERR_INVALID otherwise). */
int npbcg_set_preconditioner(npbcg_ctx *ctx, int kind);
/* Downloads the diagonal of the active M^-1 (n entries; plain CG: ones). */
int npbcg_get_preconditioner(npbcg_ctx *ctx, double *minv);
/* Applies the active M^-1 to v (n entries, host) into z: z = M^-1 v in FP64. */
int npbcg_apply_preconditioner(npbcg_ctx *ctx, const double *v, double *z);

/* ---- learned preconditioner ------------------------------------------------
 * Training data are residual/correction pairs the solver generates itself:
 * plain FP64 CG on A x = b to x* (rtol 1e-12), then, from x_0 = 0, the CG
 * residuals r_k and corrections e_k = x* - x_k of its first P steps.  The
 * network (3 message-passing layers over the top-K row graph, hidden 64,
 * BF16 MFMA with FP32 accumulation) predicts g, h per row and is trained with
 * Adam on L = sum_k |M^-1 r_k - e_k|^2 / (P |e_k|^2); npbcg_net_train weights
 * each system's gradient by 1 / L(g = h = 0), the loss of the Neumann base. */
#define NPBCG_MAX_PAIRS 8
#define NPBCG_NET_FEATURES 8

/* Plain CG to x* for the resident right-hand side, then the first `pairs`
 * residual/correction pairs, kept on the device for npbcg_net_loss_grad;
 * r and e (pairs * n floats each, pair-major) receive copies when not NULL. */
int npbcg_log_pairs(npbcg_ctx *ctx, int pairs, float *r, float *e);
/* Replaces the device pairs with host ones (pairs * n floats each). */
int npbcg_set_pairs(npbcg_ctx *ctx, int pairs, const float *r, const float *e);
/* The top-K graph of the context's matrix: neighbours nbr[n*K], message
 * weights wgt[n*K], N values nval[n*K] and features feat[n*8] (any may be
 * NULL); K is returned. */
int npbcg_graph(npbcg_ctx *ctx, int *nbr, float *wgt, float *nval, float *feat);

typedef struct npbcg_net npbcg_net;

typedef struct npbcg_train_opts {
    int systems;             /* training systems */
    int n_min, n_max;        /* order n, uniform */
    int nonzer_min, nonzer_max;   /* nonzeros per generator vector, uniform */
    double shift_min, shift_max;  /* diagonal shift = (nonzer+1)^2 * U[shift_min, shift_max]
                                     (the NPB classes sit near 0.43 .. 0.2; the range keeps
                                     every system definite) */
    double rcond;
    unsigned long long seed; /* seeds every system's (n, nonzer, shift, matrix seed, b) */
    int pairs;               /* residual/correction pairs per system (<= 8) */
    int steps_per_system;    /* Adam steps per system */
    float lr;                /* Adam learning rate */
    int gen_threads;         /* host generator threads (0: up to 16) */
    int first_system;        /* index of the first system (runs may be split in chunks) */
    int reserved[5];
} npbcg_train_opts;

typedef struct npbcg_train_stats {
    int systems;             /* systems trained on */
    int steps;
    int skipped;             /* systems whose plain CG missed 1e-12 (not used) */
    int reserved_i;
    double loss_first;       /* mean loss of the first 10% of systems, before their step */
    double loss_last;        /* same, last 10% */
    double base_first;       /* Neumann (g = h = 0) loss of the same systems */
    double base_last;
    double wait_ms;          /* time the device side waited for the generator */
    double total_ms;
    double reserved[3];
} npbcg_train_stats;

int npbcg_net_param_count(void);
/* A network with He-initialised layers and a zero readout (g = h = 0). */
int npbcg_net_create(npbcg_net **net, int device, unsigned long long seed);
void npbcg_net_destroy(npbcg_net *net);
/* Parameters (npbcg_net_param_count floats) and feature normalisation
 * (2 * NPBCG_NET_FEATURES floats: mean then std; all-zero if never set). */
int npbcg_net_get_params(npbcg_net *net, float *params, float *feature_norm);
int npbcg_net_set_params(npbcg_net *net, const float *params, const float *feature_norm);
void npbcg_net_default_train_opts(npbcg_train_opts *opts);
/* Trains on opts->systems generated systems.  loss_hist and base_hist (may be
 * NULL) receive, per system, the loss before its step and the Neumann loss. */
int npbcg_net_train(npbcg_net *net, const npbcg_train_opts *opts, npbcg_train_stats *stats,
                    double *loss_hist, double *base_hist);
/* Loss of the context's resident pairs; grad (param_count floats) and gh
 * (g then h, 2n floats) when not NULL.  Does not change the parameters. */
int npbcg_net_loss_grad(npbcg_net *net, npbcg_ctx *ctx, double *loss, float *grad, float *gh);
/* One Adam step on the gradient of the context's resident pairs. */
int npbcg_net_step(npbcg_net *net, npbcg_ctx *ctx, float lr, double *loss);
/* Inference: g, h for the context's matrix, installed as its LEARNED
 * preconditioner; ms (may be NULL) receives the device time. */
int npbcg_net_apply(npbcg_net *net, npbcg_ctx *ctx, double *ms);

/* ---- batched solves ----------------------------------------------------------
 * Many independent systems, each with nrhs right-hand sides, solved together:
 * the matrices are concatenated block-diagonally on the device and one CG step
 * of every unfinished system is one launch sequence (SpMM over all systems with
 * the nrhs vectors read together, x/r update, p update).  MIXED runs the inner
 * CG in FP32 inside FP64 refinement sweeps (FP64 true residual per sweep);
 * FP64 runs the same kernels in FP64 with one sweep.  There is no counterpart
 * in the reference, which solves one system per process. */
typedef struct npbcg_batch npbcg_batch;

typedef struct npbcg_batch_opts {
    int device;
    int nrhs;            /* 1, 2 or 4 */
    int mode;            /* NPBCG_SOLVE_FP64 or NPBCG_SOLVE_MIXED */
    int lanes_per_row;   /* lanes per row of the row-gather kernels: 4, 8, 16, 32, 64;
                            0 = tuned or default */
    int rows_per_tile;   /* rows per SpMM workgroup: 64 .. 4096 (layout 3: one
                            2048-row sorting window per item, whatever is given);
                            0 = tuned or default */
    int inner_steps;     /* MIXED: inner CG steps per sweep at most (0: 60) */
    int layout;          /* SpMM: 0 = tuned or the default (3), 1 = row gathers,
                            2 = column blocks (vector blocks staged in LDS, rows
                            in CSR order), 3 = SELL slices over LDS-staged column
                            blocks (one lane per row, 16-bit columns) */
    int reserved[9];
} npbcg_batch_opts;

typedef struct npbcg_batch_stats {
    int solves;          /* systems * nrhs */
    int converged;       /* solves whose FP64 true relative residual <= rtol */
    int sweeps;          /* FP64 residual checks */
    int max_steps;       /* largest inner CG step count of a solve */
    int lanes_per_row;
    int rows_per_tile;
    int layout;          /* 1 row gathers, 2 column blocks, 3 SELL slices */
    int reserved_i;
    double time_ms;      /* HIP events around the device work */
    double max_relres;
    double reserved_d[4];
} npbcg_batch_stats;

void npbcg_batch_default_opts(npbcg_batch_opts *opts);
/* Uploads nsys CSR matrices (0-based, n[s] rows each). */
int npbcg_batch_create(npbcg_batch **batch, int nsys, const int *n, const int *const *rowstr,
                       const int *const *colidx, const double *const *a,
                       const npbcg_batch_opts *opts);
void npbcg_batch_destroy(npbcg_batch *batch);
/* Right-hand sides: system after system, each n[s] rows of nrhs values. */
int npbcg_batch_set_rhs(npbcg_batch *batch, const double *b);
/* Solves every system to rtol (FP64 true residual); maxit bounds the CG steps
 * of each solve.  steps and relres (nsys*nrhs each, may be NULL) receive the
 * per-solve results.  NPBCG_ERR_NOCONV if a solve missed rtol. */
int npbcg_batch_solve(npbcg_batch *batch, double rtol, int maxit, npbcg_batch_stats *stats,
                      int *steps, double *relres);
int npbcg_batch_get_solution(npbcg_batch *batch, double *x);
/* Autotuner: times the SpMM of this batch for every (layout, lanes_per_row,
 * rows_per_tile) candidate of layouts 1 and 3, keeps the fastest, and
 * records it in the text
 * file cache_path (may be NULL) under the batch's shape; a later call with the
 * same shape and file reuses the record without timing.  ms (may be NULL)
 * receives the chosen SpMM time. */
int npbcg_batch_tune(npbcg_batch *batch, const char *cache_path, int *lanes_per_row,
                     int *rows_per_tile, double *ms);
/* Mean device time of the batch's SpMM (current launch shape) over reps
 * launches, every system marked running. */
int npbcg_batch_spmm_bench(npbcg_batch *batch, int reps, double *avg_ms);
/* Generator threads of npbcg_makea calls made from the calling thread
 * (0: the default, NPBCG_GEN_THREADS or up to 16). */
void npbcg_set_generator_threads(int threads);

/* ---- reference-shaped shim ------------------------------------------------
 * For a caller that keeps the reference's exact call
 *   conj_grad(colidx, rowstr, x, z, a, p, q, r, &rnorm);
 * with no context argument: the order n comes from npbcg_compat_set_n (the
 * reference's global naa, cg.cu:124,402), and a context is created on first
 * use and re-created if rowstr/colidx/a change.  Errors set *rnorm to NaN
 * and leave the text in npbcg_last_error(). */
int npbcg_compat_set_n(int naa);
void npbcg_compat_conj_grad(int colidx[], int rowstr[], double x[], double z[],
                            double a[], double p[], double q[], double r[],
                            double *rnorm);
void npbcg_compat_release(void);

/* ---- allocation-free conj_grad -------------------------------------------
 * No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc): the computation of npbcg_conj_grad_device (conj_grad_gpu,
 * cg.cu:803-846) with no context, on device arrays the caller owns:
 *   rowstr (n+1), colidx and a (rowstr[n] entries): the CSR, columns ascending;
 *   x (n+2) read for n+1 entries; z, p, q, r (n+2) written for n+1 entries,
 *   index n as npbcg_conj_grad leaves it (q = z = 0, r = p = x[n]);
 *   rnorm: one double of device memory receiving ||x - A z||.
 * cgitmax CG steps (25 in the reference) run with the 64-lane CSR SpMV in the
 * reference's step order, so results are bit-identical to npbcg_conj_grad_device
 * on a context created with spmv_variant = 1.  Work is enqueued on `stream`
 * (a hipStream_t; NULL = the null stream) and the call returns without
 * synchronising; it calls no hipMalloc, hipFree or memset, so it can be
 * captured into a graph.  The partial sums and CG scalars live in a
 * file-scope device array of NPBCG_NOALLOC_SCRATCH_BYTES per GPU, held until
 * the process ends and shared by every call on that GPU: calls that overlap
 * on one GPU race.  A launch shape needing more scratch than that (more than
 * 256 CUs) returns NPBCG_ERR_CAPACITY. */
#define NPBCG_NOALLOC_SCRATCH_BYTES (160 * 1024)
int npbcg_conj_grad_device_noalloc(int n, const int *rowstr, const int *colidx,
                                   const double *a, double *x, double *z, double *p,
                                   double *q, double *r, int cgitmax, double *rnorm,
                                   void *stream);

#ifdef __cplusplus
}
#endif
#endif /* NPBCG_H */
