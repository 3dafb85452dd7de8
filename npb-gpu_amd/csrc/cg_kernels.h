// Host-visible interface of the npbcg device kernels.
#ifndef NPBCG_CG_KERNELS_H
#define NPBCG_CG_KERNELS_H

#include <cstdint>

#include <hip/hip_runtime.h>

#include <vector>

namespace npbcg_dev {

struct Scalars;

// SpMV kernel shapes.
enum SpmvVariant {
    kSpmvAuto = 0,
    kSpmvSeg64 = 1,    // CSR, one wavefront per row, 4 loads in flight per lane
    kSpmvSeg32 = 2,    // CSR, half-wavefront per row
    kSpmvSeg16 = 3,    // CSR, quarter-wavefront per row
    kSpmvSeg64U2 = 4,  // CSR, one wavefront per row, 2 loads in flight per lane
    kSpmvTiled = 5     // column-blocked copy, vector blocks staged in LDS
};

// Device copy of the column-blocked SELL matrix (tiled_layout.h) and its
// launch shape.
struct SellDev {
    const int *wg_slice = nullptr;
    const int *slice_off = nullptr;
    const int *slice_row = nullptr;
    const double *val = nullptr;
    const uint16_t *col = nullptr;
    const int *row_slot = nullptr;  // nblk * n: slot of row j in column block b
    double *rows_part = nullptr;  // one block row sum per slot, stored in slot order
    long long slots = 0;          // 64 per slice
    int W = 0, nblk = 0, R = 0, nrange = 0;
    int grid = 0;       // nblk * nrange workgroups
    int block = 1024;   // threads per workgroup: 1024 or 512
    int lds_bytes = 0;  // dynamic LDS: W vector entries + one double per slice
};

// Per-workgroup partial sums handed from one kernel to the next.
struct Partials {
    double *xx;      // x.x of cg_init            [vec_blocks]
    double *rr[2];   // r.r of update_zr, alternating by CG step [vec_blocks]
    double *pq;      // p.q of the SpMV           [spmv_grid]
    double *res;     // ||x-Az||^2, x.z, z.z of the residual [3 * max(spmv_grid, vec_blocks)]
    double *nrm;     // x.z, z.z of norms_kernel  [2 * vec_blocks]
    double *rz[2];   // r.M^-1 r of the sparse preconditioner, by CG step; [1] also
                     // holds the start's r0.z0 [vec_blocks]
};

struct PrecondOp;  // sparse preconditioner operator (learned_kernels.h)

// Events stamped on the stream after each launch of a profiled run: launch
// i's device time is the gap between stamp i and stamp i + 1 (stamp 0 comes
// before the first launch).  Used with direct launches only.
struct KernelMarks {
    std::vector<hipEvent_t> ev;
    std::vector<const char *> name;  // kernel of launch i
    std::vector<int> threads;        // its workgroup size
    hipError_t err = hipSuccess;
    void stamp(hipStream_t st);
    void note(const char *kname, int nthreads, hipStream_t st);
    ~KernelMarks();
};

// Device pointers and launch shapes of one FP64 problem.
struct Fp64Problem {
    int n;
    long long nnz;
    const int *rowstr;
    const int *colidx;
    const double *a;
    double *x, *z, *p, *q, *r;  // n+2 entries each
    double *p_alt;              // second direction buffer (SELL path ping-pong)
    Partials part;
    Scalars *sc;
    double *zeta_hist, *rnorm_hist;
    int hist_cap;
    int vec_blocks;             // workgroups of the element-wise kernels
    int spmv_blocks;            // workgroups of the CSR SpMV kernels
    int spmv_variant;
    int fuse_p_update;          // SELL path: p update folded into the next SpMV
    int cg_form;                // SELL path, conj_grad only: 1 = one-reduction form
                                // (s = A r, then one update kernel per step)
    int dot_apart;              // with cg_form: the SpMV stores s = A r without r.s;
                                // combine_rs adds the block sums into s (in p_alt)
                                // and forms r.s, then update_cg reads s directly
    const double *minv;         // diagonal preconditioner M^-1 (nullptr: plain CG)
    const PrecondOp *pop;       // sparse preconditioner (nullptr: none); z = M^-1 r in pz
    double *pz;
    SellDev sell;
    KernelMarks *marks = nullptr;  // profiled run: a stamp after every launch
};

// Workgroups launched by the SpMV of this problem (= partials it writes).
int spmv_grid(const Fp64Problem &pb);

// Workgroups of the element-wise kernels for order n on `cus` CUs: one per
// 256 entries, at most 8 per CU, rounded up to a multiple of the 8 XCDs so
// the kernels that combine block row sums can give each XCD a contiguous
// run of rows (xcd_chunk in cg_device.h).
int vector_blocks(int n, int cus);

// Optional HIP events bracketing the first `count` SpMV launches of a
// conj_grad call (used to time the SpMV kernel inside a timed run).
struct SpmvProbe {
    hipEvent_t *start;
    hipEvent_t *stop;
    int count;
};

int spmv_rows_per_block(int variant);

// Raises the dynamic-LDS limit of the SELL SpMV kernels to lds_bytes.
hipError_t prepare_sell_kernels(int lds_bytes);

// A complete q = A p for every layout (SELL: SpMV then block combine).
hipError_t launch_spmv_q(const Fp64Problem &pb, hipStream_t st);
// The CG SpMV kernel alone as CG step `step` launches it (SELL: row
// partials, walk direction by step parity), bracketed by e0/e1 when non-null.
hipError_t launch_spmv_kernel(const Fp64Problem &pb, int step, hipEvent_t e0, hipEvent_t e1,
                              hipStream_t st);
// r = A z (partials ||x - r||^2, x.z, z.z), then a one-workgroup finish that
// forms rnorm, zeta, xscale and, with record != 0, appends to the history.
hipError_t launch_resid(const Fp64Problem &pb, int record, hipStream_t st);
hipError_t launch_cg_init(const Fp64Problem &pb, int scale_x, hipStream_t st);
// CG step `step` of a sequence; with sets_tol, step 0 also fixes the solve's
// stop level from its r0.r0.
hipError_t launch_update_zr(const Fp64Problem &pb, int step, hipStream_t st,
                            int sets_tol = 1);
hipError_t launch_update_p(const Fp64Problem &pb, int step, hipStream_t st);
hipError_t launch_norms(const Fp64Problem &pb, hipStream_t st);
// x = s * z, or x = xscale * z (the last residual finish) if from_scalars.
hipError_t launch_scale_x(const Fp64Problem &pb, double s, int from_scalars,
                          hipStream_t st);
// Resets the outer-loop and solve bookkeeping; stop != 0 enables the
// relative-residual stop test r.r <= rtol^2 * r0.r0 inside the CG steps.
hipError_t launch_reset_scalars(const Fp64Problem &pb, double shift, hipStream_t st,
                                double rtol = 0.0, int stop = 0);
// Clears the stop flag; end_of_solve also returns the scalars to conj_grad /
// power use (no stop test, no skipped residual passes).
hipError_t launch_clear_done(const Fp64Problem &pb, hipStream_t st, int end_of_solve = 0);
hipError_t launch_fill(double *v, int count, double value, int blocks, hipStream_t st);
// Copy of bytes (a multiple of 16) from src to dst.
hipError_t launch_copy(const void *src, void *dst, long long bytes, int blocks,
                       hipStream_t st);
// Read of bytes (a multiple of 16) from src with `blocks` 1024-thread
// workgroups; sink (one double) is written only on an impossible sum.
hipError_t launch_read(const void *src, long long bytes, double *sink, int blocks,
                       hipStream_t st);

// One conj_grad: init (optionally applying x = xscale z first), cgitmax x
// (spmv, update_zr, update_p), residual SpMV and finish.
hipError_t enqueue_conj_grad(const Fp64Problem &pb, int cgitmax, int scale_x, int record,
                             hipStream_t st, const SpmvProbe *probe);
// Kernel launches of one enqueue_conj_grad.
int conj_grad_launches(const Fp64Problem &pb, int cgitmax);

// `steps` CG steps (SpMV, update_zr, update_p) of a solve on pb (x holds b,
// z the iterate); the stop flag makes steps after convergence return at once.
// Step `it` of a solve sequence (it selects the partial-sum buffers).
hipError_t enqueue_solve_step(const Fp64Problem &pb, int it, int sets_tol, hipStream_t st);
hipError_t enqueue_solve_steps(const Fp64Problem &pb, int steps, int sets_tol,
                               hipStream_t st);

// FP32 side of the mixed-precision solve (cg_mixed.hip).
struct MixedProblem {
    const float *val = nullptr;   // FP32 values in the SELL order of the FP64 copy
    const float *minv = nullptr;  // FP32 diagonal preconditioner (nullptr: none)
    const PrecondOp *pop = nullptr;  // sparse preconditioner (FP32 values); z in pz
    float *pz = nullptr;
    float *d = nullptr, *r = nullptr, *p = nullptr, *q = nullptr;  // n+2 each
    float *p_alt = nullptr;       // second p buffer of the fused p update
    float *rows_part = nullptr;   // nblk * n per-block row sums, by row
    double *part_xx = nullptr;    // r0.r0 partials of an inner start [vec_blocks]
    double *part_rr[2] = {nullptr, nullptr};  // r.r partials [vec_blocks]
    double *part_rz[2] = {nullptr, nullptr};  // r.M^-1 r partials (sparse preconditioner)
    double *part_pq = nullptr;    // p.q partials [sell grid]
    Scalars *sc = nullptr;        // inner-solve scalars
    int lds_bytes = 0;
    int stage_rows = 0;           // rows per range collected in LDS by the SpMV (0: none)
};
hipError_t prepare_mixed_kernels(int lds_bytes);
// The FP32 SELL SpMV kernel of an inner CG step (q = A p with p.q partials,
// p in mp.p, walk direction by step parity), bracketed by e0/e1.
hipError_t launch_spmv32_kernel(const Fp64Problem &pb, const MixedProblem &mp, int step,
                                hipEvent_t e0, hipEvent_t e1, hipStream_t st);
hipError_t launch_to_fp32(const double *src, float *dst, long long count, hipStream_t st);
// Inner FP32 CG on A d = b - A x (or b on the first sweep): init and `steps`
// steps; pb is the solve view (x holds b, z the FP64 iterate, r = A z).
hipError_t enqueue_inner32(const Fp64Problem &pb, const MixedProblem &mp, int first_sweep,
                           int steps, hipStream_t st);
// z += d (the FP64 iterate takes the FP32 correction).
hipError_t launch_refine(const Fp64Problem &pb, const MixedProblem &mp, hipStream_t st);
// r = b - r (r held A z), p = r, r.r partials into part.xx.
hipError_t launch_residual_restart(const Fp64Problem &pb, hipStream_t st);

hipError_t launch_narrow(const double *in, float *out, int n, int blocks, hipStream_t st);
// minv = 1/diag(A) from the CSR.
hipError_t launch_jacobi(const Fp64Problem &pb, double *minv, hipStream_t st);

// Start of a refinement sweep of the mixed solve, decided on the device from
// the FP64 true residual of the previous sweep (outer->rnorm): if it met rtol
// (or maxit inner steps are used up) the inner solve and this sweep's residual
// pass are skipped; else the inner stop level max(1e-6, rtol/(2 relres))^2 is
// set.  Lets several sweeps be enqueued with one host synchronisation.
hipError_t launch_sweep_prepare(Scalars *outer, Scalars *inner, double rhs_norm, double rtol,
                                int maxit, int first, hipStream_t st);

size_t scalars_bytes();

// Persistent engine (cg_persistent.hip): one cooperative launch runs whole
// outer iterations on the SELL layout, with grid barriers between phases.
struct PersistentState {
    double *part = nullptr;    // persistent_partials(pb) doubles
    unsigned *bar = nullptr;   // [0] barrier arrivals, [1] abort word (nonzero: a
                               // barrier waited past spin_limit; results invalid)
    long long spin_limit = 0;  // wall-clock ticks one barrier may wait
    int cgitmax = 25;
    void *args = nullptr;      // device copy of the kernel's argument block
                               // (persistent_args_bytes())
};
size_t persistent_args_bytes();
// Workgroups of the persistent kernel that fit on one CU (0 off the SELL layout).
hipError_t persistent_capacity(const Fp64Problem &pb, int *per_cu);
int persistent_partials(const Fp64Problem &pb);
// Grid barriers passed by one outer iteration.
long long persistent_barriers(int cgitmax);
// niter outer iterations {conj_grad; rnorm, zeta, xscale} from the context's
// x (scale_first: x = xscale z of the previous call applied first); record
// appends zeta and rnorm to the history; final_scale applies x = xscale z after
// the last one.
hipError_t launch_persistent(const Fp64Problem &pb, const PersistentState &ps, int niter,
                             int scale_first, int record, int final_scale, hipStream_t st);

}  // namespace npbcg_dev

#endif
