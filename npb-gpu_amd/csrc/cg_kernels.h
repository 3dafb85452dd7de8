// Host-visible interface of the npbcg device kernels.
#ifndef NPBCG_CG_KERNELS_H
#define NPBCG_CG_KERNELS_H

#include <hip/hip_runtime.h>

namespace npbcg_dev {

struct Scalars;

// SpMV kernel shapes: lanes per CSR row and loads in flight per lane.
enum SpmvVariant {
    kSpmvAuto = 0,
    kSpmvSeg64 = 1,    // one wavefront per row, 4 loads in flight per lane
    kSpmvSeg32 = 2,    // half-wavefront per row
    kSpmvSeg16 = 3,    // quarter-wavefront per row
    kSpmvSeg64U2 = 4,  // one wavefront per row, 2 loads in flight per lane
    kSpmvTiled = 5     // column-blocked copy, vector blocks staged in LDS
};

constexpr int kTiledBlock = 1024;  // threads per workgroup of the tiled SpMV

// Device copy of the column-blocked matrix (tiled_layout.h) and its launch
// shape.
struct TiledDev {
    const int *seg = nullptr;
    const double *val = nullptr;
    const uint16_t *col = nullptr;
    int T = 0, W = 0, nblk = 0, ntile = 0;
    int grid = 0;       // workgroups launched (grid-stride over tiles)
    int seg_lanes = 32; // lanes per row segment: 16, 32 or 64
    int lds_bytes = 0;  // dynamic LDS: (W + T) doubles
};

// Device pointers and launch shapes of one FP64 problem.
struct Fp64Problem {
    int n;
    long long nnz;
    const int *rowstr;
    const int *colidx;
    const double *a;
    double *x, *z, *p, *q, *r;  // n+2 entries each
    double *part;               // reduction partials, 3 * max(blocks)
    Scalars *sc;
    double *zeta_hist, *rnorm_hist;
    int hist_cap;
    int vec_blocks;             // workgroups of the element-wise kernels
    int spmv_blocks;            // workgroups of the SpMV kernels
    int spmv_variant;
    TiledDev tiled;
};

// Optional HIP events bracketing the first `count` SpMV launches of a
// conj_grad call (used to time the SpMV kernel inside a timed run).
struct SpmvProbe {
    hipEvent_t *start;
    hipEvent_t *stop;
    int count;
};

int spmv_rows_per_block(int variant);

// Raises the dynamic-LDS limit of the tiled SpMV kernels to lds_bytes.
hipError_t prepare_tiled_kernels(int lds_bytes);

hipError_t launch_spmv_pq(const Fp64Problem &pb, hipStream_t st);
hipError_t launch_resid(const Fp64Problem &pb, int record, hipStream_t st);
hipError_t launch_update_zr(const Fp64Problem &pb, hipStream_t st);
hipError_t launch_update_p(const Fp64Problem &pb, hipStream_t st);
hipError_t launch_cg_init(const Fp64Problem &pb, hipStream_t st);
hipError_t launch_norms(const Fp64Problem &pb, hipStream_t st);
hipError_t launch_scale_x(const Fp64Problem &pb, double s, int from_pending,
                          hipStream_t st);
hipError_t launch_clear_pending(const Fp64Problem &pb, hipStream_t st);
hipError_t launch_reset_scalars(const Fp64Problem &pb, double shift, hipStream_t st);
hipError_t launch_fill(double *v, int count, double value, int blocks, hipStream_t st);

// One conj_grad: init, cgitmax x (spmv_pq, update_zr, update_p), resid.
// record != 0 stores (zeta, rnorm) in the history and leaves x's
// normalisation pending for the next init.
hipError_t enqueue_conj_grad(const Fp64Problem &pb, int cgitmax, int record,
                             hipStream_t st, const SpmvProbe *probe);

size_t scalars_bytes();

}  // namespace npbcg_dev

#endif
