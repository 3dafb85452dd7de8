// FP64 conj_grad kernels for gfx950 (multi-kernel engine).
//
// One inner CG iteration of the reference (CUDA/CG/cg.cu:814-837) is six
// launches and two blocking device-to-host copies there (k3 SpMV, k4 p.q,
// k5 two axpys, k6 r.r, k7 p update, with alpha/beta formed on the host).
// Here it is three launches with every scalar kept on the device:
//   spmv_pq    q = A p (one wavefront or sub-wavefront per row) and p.q;
//              the last workgroup forms alpha = rho/d            (k3+k4)
//   update_zr  z += alpha p, r -= alpha q and r.r; the last workgroup
//              forms beta = rho/rho0                               (k5+k6)
//   update_p   p = r + beta p                                      (k7)
// Around the inner loop:
//   cg_init    q = z = 0, r = p = x through index n, rho = x.x; applies a
//              pending x = z/||z|| of the previous outer iteration (k11+k1+k2)
//   resid      r = A z, ||x - r||, x.z, z.z and zeta            (k8+k9+k10)
// Reductions are two-level and in a fixed order (workgroup partials, then the
// last workgroup to finish sums them in index order), so results are
// reproducible run to run.
#include "cg_kernels.h"
#include "cg_device.h"

namespace npbcg_dev {

// Dot product of one CSR row with v, by a group of SEG lanes.  Each lane
// walks the row with stride SEG; UNR independent (colidx, a) pairs are loaded
// before their gathers so several misses are in flight per lane.
template <int SEG, int UNR>
__device__ __forceinline__ double row_dot(const int *__restrict__ colidx,
                                          const double *__restrict__ a,
                                          const double *__restrict__ v, int s,
                                          int e, int sl) {
    double sum = 0.0;
    for (int k0 = s + sl; k0 < e; k0 += SEG * UNR) {
        int c[UNR];
        double av[UNR];
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            const int k = k0 + u * SEG;
            const bool ok = k < e;
            c[u] = ok ? colidx[k] : -1;
            av[u] = ok ? a[k] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < UNR; u++)
            if (c[u] >= 0) sum = fma(av[u], v[c[u]], sum);
    }
    return segment_sum<SEG>(sum);
}

// q = A p and d = p.q; last workgroup: alpha = rho/d, rho0 = rho.
template <int SEG, int UNR>
__global__ __launch_bounds__(kBlock) void spmv_pq_kernel(
    const int *__restrict__ rowstr, const int *__restrict__ colidx,
    const double *__restrict__ a, const double *__restrict__ p,
    double *__restrict__ q, int n, double *__restrict__ part, Scalars *sc) {
    constexpr int kGroups = kBlock / SEG;
    const int sl = threadIdx.x % SEG;
    const int g = blockIdx.x * kGroups + threadIdx.x / SEG;
    const int ng = gridDim.x * kGroups;
    double acc[1] = {0.0};
    for (int row = g; row < n; row += ng) {
        const int s = rowstr[row], e = rowstr[row + 1];
        const double sum = row_dot<SEG, UNR>(colidx, a, p, s, e, sl);
        if (sl == 0) {
            q[row] = sum;
            acc[0] += p[row] * sum;
        }
    }
    double tot[1];
    if (grid_reduce<1>(acc, part, &sc->ticket[kTicketSpmv], tot) && threadIdx.x == 0) {
        sc->d = tot[0];
        sc->alpha = sc->rho / tot[0];
        sc->rho0 = sc->rho;
    }
}

// z += alpha p; r -= alpha q; rho = r.r; last workgroup: beta = rho/rho0.
__global__ __launch_bounds__(kBlock) void update_zr_kernel(
    const double *__restrict__ p, const double *__restrict__ q,
    double *__restrict__ z, double *__restrict__ r, int n,
    double *__restrict__ part, Scalars *sc) {
    const double alpha = sc->alpha;
    double acc[1] = {0.0};
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock) {
        z[j] = z[j] + alpha * p[j];
        const double rj = r[j] - alpha * q[j];
        r[j] = rj;
        acc[0] += rj * rj;
    }
    double tot[1];
    if (grid_reduce<1>(acc, part, &sc->ticket[kTicketZR], tot) && threadIdx.x == 0) {
        sc->rho = tot[0];
        sc->beta = tot[0] / sc->rho0;
    }
}

// p = r + beta p.
__global__ __launch_bounds__(kBlock) void update_p_kernel(
    const double *__restrict__ r, double *__restrict__ p, int n,
    const Scalars *__restrict__ sc) {
    const double beta = sc->beta;
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock)
        p[j] = r[j] + beta * p[j];
}

// Start of conj_grad (cg.cu:671-686): q = z = 0, r = p = x for j <= n and
// rho = x.x over j < n.  If the previous outer iteration left a pending
// normalisation, x = xscale * z (j < n) is applied first (cg.cu:479-481).
__global__ __launch_bounds__(kBlock) void cg_init_kernel(
    double *__restrict__ x, double *__restrict__ z, double *__restrict__ p,
    double *__restrict__ q, double *__restrict__ r, int n,
    double *__restrict__ part, Scalars *sc) {
    const bool scale = sc->pending_scale != 0;
    const double xs = sc->xscale;
    double acc[1] = {0.0};
    for (int j = blockIdx.x * kBlock + threadIdx.x; j <= n; j += gridDim.x * kBlock) {
        double xv = x[j];
        if (scale && j < n) {
            xv = xs * z[j];
            x[j] = xv;
        }
        q[j] = 0.0;
        z[j] = 0.0;
        r[j] = xv;
        p[j] = xv;
        if (j < n) acc[0] += xv * xv;
    }
    double tot[1];
    if (grid_reduce<1>(acc, part, &sc->ticket[kTicketInit], tot) && threadIdx.x == 0) {
        sc->rho = tot[0];
        sc->pending_scale = 0;
    }
}

// End of conj_grad and the outer step (cg.cu:782-800, :505-514):
// r = A z, rnorm = ||x - r||, x.z, z.z, zeta = shift + 1/(x.z),
// xscale = 1/sqrt(z.z).  With record != 0 the (zeta, rnorm) pair is stored
// at history index sc->iter and a normalisation of x is left pending.
template <int SEG, int UNR>
__global__ __launch_bounds__(kBlock) void resid_kernel(
    const int *__restrict__ rowstr, const int *__restrict__ colidx,
    const double *__restrict__ a, const double *__restrict__ x,
    const double *__restrict__ z, double *__restrict__ r, int n,
    double *__restrict__ part, Scalars *sc, double *__restrict__ zeta_hist,
    double *__restrict__ rnorm_hist, int hist_cap, int record) {
    constexpr int kGroups = kBlock / SEG;
    const int sl = threadIdx.x % SEG;
    const int g = blockIdx.x * kGroups + threadIdx.x / SEG;
    const int ng = gridDim.x * kGroups;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int row = g; row < n; row += ng) {
        const int s = rowstr[row], e = rowstr[row + 1];
        const double sum = row_dot<SEG, UNR>(colidx, a, z, s, e, sl);
        if (sl == 0) {
            r[row] = sum;
            const double xv = x[row], zv = z[row];
            const double dlt = xv - sum;
            acc[0] += dlt * dlt;
            acc[1] += xv * zv;
            acc[2] += zv * zv;
        }
    }
    double tot[3];
    if (grid_reduce<3>(acc, part, &sc->ticket[kTicketResid], tot) && threadIdx.x == 0) {
        const double rnorm = sqrt(tot[0]);
        sc->rnorm = rnorm;
        sc->xz = tot[1];
        sc->zz = tot[2];
        sc->xscale = 1.0 / sqrt(tot[2]);
        const double zeta = sc->shift + 1.0 / tot[1];
        sc->zeta = zeta;
        if (record) {
            const int it = sc->iter;
            if (it < hist_cap) {
                zeta_hist[it] = zeta;
                rnorm_hist[it] = rnorm;
            }
            sc->iter = it + 1;
            sc->pending_scale = 1;
        }
    }
}

// x.z and z.z (cg.cu:1182-1245, gpu_kernel_ten).
__global__ __launch_bounds__(kBlock) void norms_kernel(
    const double *__restrict__ x, const double *__restrict__ z, int n,
    double *__restrict__ part, Scalars *sc) {
    double acc[2] = {0.0, 0.0};
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock) {
        const double zv = z[j];
        acc[0] += x[j] * zv;
        acc[1] += zv * zv;
    }
    double tot[2];
    if (grid_reduce<2>(acc, part, &sc->ticket[kTicketNorms], tot) && threadIdx.x == 0) {
        sc->xz = tot[0];
        sc->zz = tot[1];
    }
}

// x = s * z for j < n (cg.cu:1261-1265, gpu_kernel_eleven).  With
// from_pending != 0 the scale is sc->xscale and only applied if pending.
__global__ __launch_bounds__(kBlock) void scale_x_kernel(
    double *__restrict__ x, const double *__restrict__ z, int n, double s,
    int from_pending, Scalars *sc) {
    if (from_pending) {
        if (sc->pending_scale == 0) return;
        s = sc->xscale;
    }
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock)
        x[j] = s * z[j];
}

__global__ void clear_pending_kernel(Scalars *sc) {
    if (threadIdx.x == 0 && blockIdx.x == 0) sc->pending_scale = 0;
}

// Resets the outer-loop bookkeeping: history index, shift, pending flag and
// all tickets.  A kernel rather than a memset so it may sit in a graph.
__global__ void reset_scalars_kernel(Scalars *sc, double shift) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        sc->iter = 0;
        sc->shift = shift;
        sc->pending_scale = 0;
        for (int i = 0; i < 8; i++) sc->ticket[i] = 0u;
    }
}

__global__ void fill_kernel(double *__restrict__ v, int count, double value) {
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < count; j += gridDim.x * kBlock)
        v[j] = value;
}

// ---------------------------------------------------------------------------
// LDS-staged SpMV over the column-blocked copy (tiled_layout.h).
//
// One 1024-thread workgroup owns a tile of T rows.  For each column block it
// stages the W entries of the input vector the block covers into LDS with
// coalesced 16-byte loads, then groups of SEG lanes sweep the tile's row
// segments in that block, streaming (value, 16-bit column) pairs from HBM and
// gathering the vector from LDS.  Each row's running sum lives in LDS and is
// owned by one lane group, so the per-row summation order (block by block,
// then lane-strided with a fixed DPP tree) is fixed.  The random 8-byte
// gathers that bound the CSR kernel (one 128-byte L2 line per gather) become
// LDS reads; the vector crosses from L2 once per tile and block.
// ---------------------------------------------------------------------------

template <int SEG, int UNR>
__device__ __forceinline__ double seg_dot_lds(const double *__restrict__ val,
                                              const uint16_t *__restrict__ col,
                                              const double *vblk, int s, int e, int sl) {
    double sum = 0.0;
    for (int k0 = s + sl; k0 < e; k0 += SEG * UNR) {
        double av[UNR];
        int c[UNR];
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            const int k = k0 + u * SEG;
            const bool ok = k < e;
            av[u] = ok ? val[k] : 0.0;
            c[u] = ok ? (int)col[k] : -1;
        }
#pragma unroll
        for (int u = 0; u < UNR; u++)
            if (c[u] >= 0) sum = fma(av[u], vblk[c[u]], sum);
    }
    return segment_sum<SEG>(sum);
}

// MODE 0: y = A v with v = p, y = q; epilogue d = p.q, alpha = rho/d.
// MODE 1: y = A v with v = z, y = r; epilogue ||x - r||, x.z, z.z, zeta.
template <int SEG, int UNR, int MODE>
__global__ __launch_bounds__(kTiledBlock) void spmv_tiled_kernel(
    const int *__restrict__ seg, const double *__restrict__ val,
    const uint16_t *__restrict__ col, const double *__restrict__ v,
    double *__restrict__ y, const double *__restrict__ x, int n, int T, int W, int nblk,
    int ntile, double *__restrict__ part, Scalars *sc, double *__restrict__ zeta_hist,
    double *__restrict__ rnorm_hist, int hist_cap, int record) {
    extern __shared__ double2 lds_v2[];  // 16-byte aligned dynamic LDS
    double *lds = reinterpret_cast<double *>(lds_v2);
    double *vblk = lds;       // W entries of the input vector
    double *rowsum = lds + W; // T running row sums
    constexpr int G = kTiledBlock / SEG;
    const int g = threadIdx.x / SEG, sl = threadIdx.x % SEG;
    double acc[MODE == 0 ? 1 : 3] = {};
    for (int t = blockIdx.x; t < ntile; t += gridDim.x) {
        const int row0 = t * T;
        const int nrows = min(T, n - row0);
        for (int i = threadIdx.x; i < T; i += kTiledBlock) rowsum[i] = 0.0;
        for (int b = 0; b < nblk; b++) {
            const int c0 = b * W;
            const int wb = min(W, n - c0);
            __syncthreads();  // previous block's gathers are done with vblk
            const double2 *v2 = reinterpret_cast<const double2 *>(v + c0);
            double2 *l2 = reinterpret_cast<double2 *>(vblk);
#pragma unroll 4
            for (int i = threadIdx.x; i < wb / 2; i += kTiledBlock) l2[i] = v2[i];
            if ((wb & 1) && threadIdx.x == 0) vblk[wb - 1] = v[c0 + wb - 1];
            __syncthreads();
            const int *sp = seg + (size_t)(t * nblk + b) * (T + 1);
            for (int r = g; r < nrows; r += G) {
                const double s = seg_dot_lds<SEG, UNR>(val, col, vblk, sp[r], sp[r + 1], sl);
                if (sl == 0) rowsum[r] += s;
            }
        }
        __syncthreads();
        for (int r = threadIdx.x; r < nrows; r += kTiledBlock) {
            const int row = row0 + r;
            const double yv = rowsum[r];
            y[row] = yv;
            if constexpr (MODE == 0) {
                acc[0] += v[row] * yv;
            } else {
                const double xv = x[row], zv = v[row];
                const double dlt = xv - yv;
                acc[0] += dlt * dlt;
                acc[1] += xv * zv;
                acc[2] += zv * zv;
            }
        }
        __syncthreads();  // rowsum is reset for the next tile
    }
    double tot[MODE == 0 ? 1 : 3];
    if constexpr (MODE == 0) {
        if (grid_reduce<1, kTiledBlock>(acc, part, &sc->ticket[kTicketSpmv], tot) &&
            threadIdx.x == 0) {
            sc->d = tot[0];
            sc->alpha = sc->rho / tot[0];
            sc->rho0 = sc->rho;
        }
    } else {
        if (grid_reduce<3, kTiledBlock>(acc, part, &sc->ticket[kTicketResid], tot) &&
            threadIdx.x == 0) {
            const double rnorm = sqrt(tot[0]);
            sc->rnorm = rnorm;
            sc->xz = tot[1];
            sc->zz = tot[2];
            sc->xscale = 1.0 / sqrt(tot[2]);
            const double zeta = sc->shift + 1.0 / tot[1];
            sc->zeta = zeta;
            if (record) {
                const int it = sc->iter;
                if (it < hist_cap) {
                    zeta_hist[it] = zeta;
                    rnorm_hist[it] = rnorm;
                }
                sc->iter = it + 1;
                sc->pending_scale = 1;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Host-side launchers
// ---------------------------------------------------------------------------

namespace {

template <int SEG, int UNR>
hipError_t launch_spmv_pq_t(const Fp64Problem &pb, hipStream_t st) {
    spmv_pq_kernel<SEG, UNR><<<pb.spmv_blocks, kBlock, 0, st>>>(
        pb.rowstr, pb.colidx, pb.a, pb.p, pb.q, pb.n, pb.part, pb.sc);
    return hipGetLastError();
}

template <int SEG, int UNR>
hipError_t launch_resid_t(const Fp64Problem &pb, int record, hipStream_t st) {
    resid_kernel<SEG, UNR><<<pb.spmv_blocks, kBlock, 0, st>>>(
        pb.rowstr, pb.colidx, pb.a, pb.x, pb.z, pb.r, pb.n, pb.part, pb.sc,
        pb.zeta_hist, pb.rnorm_hist, pb.hist_cap, record);
    return hipGetLastError();
}

template <int SEG, int MODE>
hipError_t launch_tiled_t(const Fp64Problem &pb, int record, hipStream_t st) {
    const TiledDev &td = pb.tiled;
    const double *v = MODE == 0 ? pb.p : pb.z;
    double *y = MODE == 0 ? pb.q : pb.r;
    spmv_tiled_kernel<SEG, 2, MODE><<<td.grid, kTiledBlock, td.lds_bytes, st>>>(
        td.seg, td.val, td.col, v, y, pb.x, pb.n, td.T, td.W, td.nblk, td.ntile, pb.part,
        pb.sc, pb.zeta_hist, pb.rnorm_hist, pb.hist_cap, record);
    return hipGetLastError();
}

template <int MODE>
hipError_t launch_tiled(const Fp64Problem &pb, int record, hipStream_t st) {
    switch (pb.tiled.seg_lanes) {
    case 16: return launch_tiled_t<16, MODE>(pb, record, st);
    case 64: return launch_tiled_t<64, MODE>(pb, record, st);
    default: return launch_tiled_t<32, MODE>(pb, record, st);
    }
}

template <int SEG, int MODE>
hipError_t set_tiled_lds_t(int bytes) {
    return hipFuncSetAttribute(
        reinterpret_cast<const void *>(&spmv_tiled_kernel<SEG, 2, MODE>),
        hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

}  // namespace

hipError_t prepare_tiled_kernels(int lds_bytes) {
    hipError_t e = hipSuccess;
    hipError_t r[6] = {set_tiled_lds_t<16, 0>(lds_bytes), set_tiled_lds_t<32, 0>(lds_bytes),
                       set_tiled_lds_t<64, 0>(lds_bytes), set_tiled_lds_t<16, 1>(lds_bytes),
                       set_tiled_lds_t<32, 1>(lds_bytes), set_tiled_lds_t<64, 1>(lds_bytes)};
    for (hipError_t x : r)
        if (x != hipSuccess) e = x;
    return e;
}

int spmv_rows_per_block(int variant) {
    switch (variant) {
    case kSpmvSeg16: return kBlock / 16;
    case kSpmvSeg32: return kBlock / 32;
    default: return kBlock / 64;
    }
}

hipError_t launch_spmv_pq(const Fp64Problem &pb, hipStream_t st) {
    switch (pb.spmv_variant) {
    case kSpmvTiled: return launch_tiled<0>(pb, 0, st);
    case kSpmvSeg16: return launch_spmv_pq_t<16, 4>(pb, st);
    case kSpmvSeg32: return launch_spmv_pq_t<32, 4>(pb, st);
    case kSpmvSeg64U2: return launch_spmv_pq_t<64, 2>(pb, st);
    default: return launch_spmv_pq_t<64, 4>(pb, st);
    }
}

hipError_t launch_resid(const Fp64Problem &pb, int record, hipStream_t st) {
    switch (pb.spmv_variant) {
    case kSpmvTiled: return launch_tiled<1>(pb, record, st);
    case kSpmvSeg16: return launch_resid_t<16, 4>(pb, record, st);
    case kSpmvSeg32: return launch_resid_t<32, 4>(pb, record, st);
    case kSpmvSeg64U2: return launch_resid_t<64, 2>(pb, record, st);
    default: return launch_resid_t<64, 4>(pb, record, st);
    }
}

hipError_t launch_update_zr(const Fp64Problem &pb, hipStream_t st) {
    update_zr_kernel<<<pb.vec_blocks, kBlock, 0, st>>>(pb.p, pb.q, pb.z, pb.r, pb.n,
                                                       pb.part, pb.sc);
    return hipGetLastError();
}

hipError_t launch_update_p(const Fp64Problem &pb, hipStream_t st) {
    update_p_kernel<<<pb.vec_blocks, kBlock, 0, st>>>(pb.r, pb.p, pb.n, pb.sc);
    return hipGetLastError();
}

hipError_t launch_cg_init(const Fp64Problem &pb, hipStream_t st) {
    cg_init_kernel<<<pb.vec_blocks, kBlock, 0, st>>>(pb.x, pb.z, pb.p, pb.q, pb.r,
                                                     pb.n, pb.part, pb.sc);
    return hipGetLastError();
}

hipError_t launch_norms(const Fp64Problem &pb, hipStream_t st) {
    norms_kernel<<<pb.vec_blocks, kBlock, 0, st>>>(pb.x, pb.z, pb.n, pb.part, pb.sc);
    return hipGetLastError();
}

hipError_t launch_scale_x(const Fp64Problem &pb, double s, int from_pending,
                          hipStream_t st) {
    scale_x_kernel<<<pb.vec_blocks, kBlock, 0, st>>>(pb.x, pb.z, pb.n, s, from_pending,
                                                     pb.sc);
    return hipGetLastError();
}

hipError_t launch_clear_pending(const Fp64Problem &pb, hipStream_t st) {
    clear_pending_kernel<<<1, 64, 0, st>>>(pb.sc);
    return hipGetLastError();
}

hipError_t launch_reset_scalars(const Fp64Problem &pb, double shift, hipStream_t st) {
    reset_scalars_kernel<<<1, 64, 0, st>>>(pb.sc, shift);
    return hipGetLastError();
}

hipError_t launch_fill(double *v, int count, double value, int blocks, hipStream_t st) {
    fill_kernel<<<blocks, kBlock, 0, st>>>(v, count, value);
    return hipGetLastError();
}

// One conj_grad (cg.cu:653-801) as a launch sequence on `st`.
hipError_t enqueue_conj_grad(const Fp64Problem &pb, int cgitmax, int record,
                             hipStream_t st, const SpmvProbe *probe) {
    hipError_t e = launch_cg_init(pb, st);
    for (int it = 0; it < cgitmax && e == hipSuccess; it++) {
        const bool timed = probe && probe->start && it < probe->count;
        if (timed) e = hipEventRecord(probe->start[it], st);
        if (e == hipSuccess) e = launch_spmv_pq(pb, st);
        if (timed && e == hipSuccess) e = hipEventRecord(probe->stop[it], st);
        if (e == hipSuccess) e = launch_update_zr(pb, st);
        if (e == hipSuccess) e = launch_update_p(pb, st);
    }
    if (e == hipSuccess) e = launch_resid(pb, record, st);
    return e;
}

}  // namespace npbcg_dev

namespace npbcg_dev {
size_t scalars_bytes() { return sizeof(Scalars); }
}  // namespace npbcg_dev
