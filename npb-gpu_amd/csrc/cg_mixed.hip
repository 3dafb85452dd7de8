// Mixed-precision CG kernels (FP32 operator and vectors, FP64 scalars).
//
// The mixed solve is FP64 iterative refinement around an inner CG that runs
// on an FP32 copy of the matrix:
//   x = 0, r = b                                       (FP64)
//   repeat: inner CG on A d = r in FP32 (matrix, vectors, SpMV accumulation),
//           with rho, alpha, beta, the dot products and the stop test in FP64;
//           x += d and r = b - A x with the FP64 matrix   (true residual)
//   until ||r|| <= rtol ||b||.
// The FP32 matrix shares the FP64 SELL layout's slices and 16-bit columns;
// only the values are FP32 (6 bytes per stored entry instead of 10).  Kernel
// structure and the partial-sum handoff are those of the FP64 path.
#include <hip/hip_ext.h>

#include "cg_device.h"
#include "cg_kernels.h"
#include "learned_kernels.h"

namespace npbcg_dev {

// Entries per lane in flight in the FP32 SELL SpMV stream (one chunk).
constexpr int kSellUnr32 = 8;

// val32 = (float) val64 (the FP32 operator).
__global__ __launch_bounds__(kBlock) void to_fp32_kernel(const double *__restrict__ src,
                                                         float *__restrict__ dst,
                                                         long long count) {
    for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < count;
         i += (long long)gridDim.x * kBlock)
        dst[i] = (float)src[i];
}

// q = A p in FP32 over the SELL layout (see spmv_sell_kernel): per-block row
// sums into rows32[b][row], FP64 partials of p.q.
// stage_rows = R (rows per range) collects the sums of the workgroup's rows,
// all of row range wg % nrange, in LDS and stores them once the slices are
// done, as one coalesced run by row; a slice's 64 rows lie anywhere in their
// sort window, so storing them straight away writes 4 bytes into each of up
// to 64 cache lines.  stage_rows = 0 stores them directly (ranges whose rows
// do not fit in the LDS beside the vector block).
// PUPD: the previous step's p update is done here.  beta = (r.r) / (old r.r)
// from the partials and the stop test on r.r (one workgroup records it; all
// workgroups reach the same decision and return); else p = r + beta p_old is
// formed while the vector block is staged, and the workgroups of row range 0
// store their block of it in p_new (p_old and p_new are different buffers).
template <int BLOCK, int UNR, bool PUPD>
__global__ __launch_bounds__(BLOCK) void spmv_sell32_kernel(
    const int *__restrict__ wg_slice, const int *__restrict__ slice_off,
    const int *__restrict__ slice_row, const float *__restrict__ val,
    const uint16_t *__restrict__ col, const float *__restrict__ v,
    float *__restrict__ rows_part, double *__restrict__ dot_part, int n, int W, int nrange,
    Scalars *sc, const float *__restrict__ r, float *__restrict__ p_new,
    const double *__restrict__ rr_part, const double *__restrict__ rho_part, int n_vec,
    int rev, int stage_rows) {
    if (sc->done) return;
    float b32 = 0.0f;
    if constexpr (PUPD) {
        double rho, rho0;
        reduce_partials2<BLOCK>(rr_part, n_vec, rho_part, n_vec, rho, rho0);
        const bool stop = fabs(rho) <= sc->tol2;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            sc->rho = rho;
            sc->beta = rho / rho0;
            if (stop) sc->done = 1;
        }
        if (stop) return;
        b32 = (float)(rho / rho0);
    }
    extern __shared__ double2 lds_v2[];  // 16-byte aligned
    float *vblk = reinterpret_cast<float *>(lds_v2);
    double *slice_dot = reinterpret_cast<double *>(vblk + ((W + 3) & ~3));
    constexpr int kWaves = BLOCK / kWave;
    const int wg = blockIdx.x;
    const int b = wg / nrange;
    const int r0 = (wg % nrange) * stage_rows;  // first row of the range (staged)
    const int c0 = b * W;
    const int wb = min(W, n - c0);
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int s0 = wg_slice[wg];
    const int nsl = wg_slice[wg + 1] - s0;
    float *rows_lds = reinterpret_cast<float *>(slice_dot + nsl);  // stage_rows entries
    // Slice of round rd for this wave (snake order, slices placed for even
    // widths per wavefront), or -1; the index grows with rd, so after the
    // first -1 there is none.
    auto slice_of = [&](int rd) {
        const int k = rd * kWaves + ((rd & 1) ? kWaves - 1 - wave : wave);
        return k < nsl ? k : -1;
    };
    // This wave's rounds forward or (rev) backward: successive SpMVs
    // alternate so each starts on the slices the last one streamed last (held
    // by the memory-side cache); no sum changes.
    int nrw = 0;
    while (slice_of(nrw) >= 0) nrw++;
    auto slice_at = [&](int i) { return i < nrw ? slice_of(rev ? nrw - 1 - i : i) : -1; };
    // Software-pipelined entry stream (sell_chunk): the next chunk's loads, or
    // the next slice's first chunk and lane rows, go out before the current
    // chunk's LDS gathers; the first chunk is loaded before the LDS fill.
    float na[UNR];
    int nc[UNR];
    int cur = slice_at(0), cbase = 0, cwidth = 0, nrow = -1;
    if (cur >= 0) {
        cbase = slice_off[s0 + cur];
        cwidth = (slice_off[s0 + cur + 1] - cbase) / kWave;
        nrow = slice_row[(size_t)(s0 + cur) * kWave + lane];
        if (cwidth > 0) sell_chunk<float, UNR>(val, col, cbase, cwidth, 0, lane, na, nc);
    }
    if constexpr (PUPD) {
        const float2 *v2 = reinterpret_cast<const float2 *>(v + c0);
        const float2 *r2 = reinterpret_cast<const float2 *>(r + c0);
        float2 *o2 = reinterpret_cast<float2 *>(p_new + c0);
        float2 *l2 = reinterpret_cast<float2 *>(vblk);
        const bool store = (wg % nrange) == 0;
#pragma unroll 4
        for (int i = threadIdx.x; i < wb / 2; i += BLOCK) {
            const float2 pv = v2[i], rv = r2[i];
            float2 nv;
            nv.x = fmaf(b32, pv.x, rv.x);
            nv.y = fmaf(b32, pv.y, rv.y);
            l2[i] = nv;
            if (store) o2[i] = nv;
        }
        if ((wb & 1) && threadIdx.x == 0) {
            const float nv = fmaf(b32, v[c0 + wb - 1], r[c0 + wb - 1]);
            vblk[wb - 1] = nv;
            if (store) p_new[c0 + wb - 1] = nv;
        }
    } else {
        // c0 is even (W even); stage pairs (up to 10 in flight per thread),
        // then a possible odd tail.
        stage_block<BLOCK, 10240 / BLOCK>(reinterpret_cast<const float2 *>(v + c0),
                                          reinterpret_cast<float2 *>(vblk), wb / 2);
        if ((wb & 1) && threadIdx.x == 0) vblk[wb - 1] = v[c0 + wb - 1];
    }
    __syncthreads();
    float *part_b = rows_part + (size_t)b * n;
    for (int i = 0; cur >= 0; i++) {
        const int k = cur, width = cwidth, base = cbase, row = nrow;
        const int nxt = slice_at(i + 1);
        int xbase = 0, xwidth = 0;
        if (nxt >= 0) {
            xbase = slice_off[s0 + nxt];
            xwidth = (slice_off[s0 + nxt + 1] - xbase) / kWave;
        }
        // p of the lane's row for p.q, loaded while the slice streams
        float pr = 0.0f;
        if (row >= 0) pr = PUPD ? fmaf(b32, v[row], r[row]) : v[row];
        float acc = 0.0f;
        const int nch = (width + UNR - 1) / UNR;
        for (int ch = 0; ch < nch; ch++) {
            float av[UNR];
            int c[UNR];
#pragma unroll
            for (int u = 0; u < UNR; u++) {
                av[u] = na[u];
                c[u] = nc[u];
            }
            sell_trim<float, UNR>(width - ch * UNR, av, c);
            if (ch + 1 < nch) {
                sell_chunk<float, UNR>(val, col, base, width, (ch + 1) * UNR, lane, na, nc);
            } else if (nxt >= 0) {
                nrow = slice_row[(size_t)(s0 + nxt) * kWave + lane];
                if (xwidth > 0) sell_chunk<float, UNR>(val, col, xbase, xwidth, 0, lane, na, nc);
            }
#pragma unroll
            for (int u = 0; u < UNR; u++) acc = fmaf(av[u], vblk[c[u]], acc);
        }
        if (nch == 0 && nxt >= 0) {
            nrow = slice_row[(size_t)(s0 + nxt) * kWave + lane];
            if (xwidth > 0) sell_chunk<float, UNR>(val, col, xbase, xwidth, 0, lane, na, nc);
        }
        if (row >= 0) {
            if (stage_rows)
                rows_lds[row - r0] = acc;
            else
                part_b[row] = acc;
        }
        const double d = segment_sum<64>((double)pr * (double)acc);
        if (lane == 0) slice_dot[k] = d;
        cur = nxt;
        cbase = xbase;
        cwidth = xwidth;
    }
    __syncthreads();
    if (stage_rows) {
        // every row of the range has a slot in each block, so all are set
        const int nr = min(stage_rows, n - r0);
        for (int i = threadIdx.x; i < nr; i += BLOCK) part_b[r0 + i] = rows_lds[i];
    }
    double dot = 0.0;
    for (int i = threadIdx.x; i < nsl; i += BLOCK) dot += slice_dot[i];
    store_block_partial<BLOCK>(dot, dot_part);
}

// Start of an inner solve: d = 0, r = p = fp32(b - Ax) (or fp32(b) on the
// first sweep, x = 0), FP64 partials of r.r.
__global__ __launch_bounds__(kBlock) void init32_kernel(
    const double *__restrict__ b, const double *__restrict__ ax, int first, int n,
    float *__restrict__ d, float *__restrict__ r, float *__restrict__ p,
    double *__restrict__ rr_out, const float *__restrict__ minv) {
    double acc = 0.0;
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock) {
        const float rj = (float)(first ? b[j] : b[j] - ax[j]);
        const float sj = minv ? minv[j] * rj : rj;  // s = M^-1 r
        d[j] = 0.0f;
        r[j] = rj;
        p[j] = sj;
        acc += (double)rj * (double)sj;
    }
    store_block_partial<kBlock>(acc, rr_out);
}

// alpha = rho / (p.q) from FP64 partials; q = sum of the block row sums;
// d += alpha p; r -= alpha q; FP64 partials of the new r.r.  The first step
// of an inner solve sets its stop level tol2 = rtol2 * r0.r0.
__global__ __launch_bounds__(kBlock) void update_zr32_kernel(
    const float *__restrict__ p, float *__restrict__ q, float *__restrict__ d,
    float *__restrict__ r, int n, const double *__restrict__ rho_part,
    const double *__restrict__ pq_part, int n_pq, double *__restrict__ rr_out,
    const float *__restrict__ rows_part, int nblk, Scalars *sc, int first_step,
    const float *__restrict__ minv, const double *__restrict__ tol_part) {
    // The stop flag is read with the prefetch below and tested after it (no
    // kernel of this launch writes it).  This thread's first row (block
    // partials of q, p, d, r) is loaded before the scalar reductions, which it
    // does not depend on, so the memory round trips overlap; the sums keep
    // their order.
    const int done = sc->done;
    const int j0 = blockIdx.x * kBlock + threadIdx.x;
    constexpr int kPre = 8;  // block partials of q prefetched
    float qb[kPre], p0 = 0.0f, d0 = 0.0f, r0 = 0.0f, m0 = 0.0f;
    if (j0 < n) {
#pragma unroll
        for (int b = 0; b < kPre; b++)
            if (b == 0 || b < nblk) qb[b] = rows_part[(size_t)b * n + j0];
        p0 = p[j0];
        d0 = d[j0];
        r0 = r[j0];
        if (minv) m0 = minv[j0];
    }
    if (done) return;  // the same value for the whole grid
    double rho, dd;
    reduce_partials2<kBlock>(rho_part, gridDim.x, pq_part, n_pq, rho, dd);
    const double alpha = rho / dd;
    // stop level from r0.r0 (tol_part, sparse preconditioner) or rho
    const double rho_tol = (first_step && tol_part) ? reduce_partials<kBlock>(tol_part, gridDim.x) : rho;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        sc->rho0 = rho;
        sc->d = dd;
        sc->alpha = alpha;
        if (first_step) {
            sc->tol2 = sc->rtol2 * fabs(rho_tol);
            sc->bb = rho_tol;
        }
        sc->steps += 1;
    }
    const float a32 = (float)alpha;
    double acc = 0.0;
    if (j0 < n) {
        float qj = qb[0];
#pragma unroll
        for (int b = 1; b < kPre; b++)
            if (b < nblk) qj += qb[b];
#pragma unroll 8
        for (int b = kPre; b < nblk; b++) qj += rows_part[(size_t)b * n + j0];
        q[j0] = qj;
        d[j0] = fmaf(a32, p0, d0);
        const float rj = fmaf(-a32, qj, r0);
        r[j0] = rj;
        acc += (double)rj * (double)(minv ? m0 * rj : rj);
    }
    for (int j = j0 + gridDim.x * kBlock; j < n; j += gridDim.x * kBlock) {
        float qj = rows_part[j];
#pragma unroll 8
        for (int b = 1; b < nblk; b++) qj += rows_part[(size_t)b * n + j];
        q[j] = qj;
        d[j] = fmaf(a32, p[j], d[j]);
        const float rj = fmaf(-a32, qj, r[j]);
        r[j] = rj;
        acc += (double)rj * (double)(minv ? minv[j] * rj : rj);
    }
    store_block_partial<kBlock>(acc, rr_out);
}

// beta = (new r.r) / (old r.r) in FP64; p = r + beta p; stop test (on
// stop_part, the r.r of a sparse-preconditioned solve, when given).
__global__ __launch_bounds__(kBlock) void update_p32_kernel(
    const float *__restrict__ r, float *__restrict__ p, int n,
    const double *__restrict__ rr_part, const double *__restrict__ rho_part, Scalars *sc,
    const float *__restrict__ minv, const double *__restrict__ stop_part) {
    if (sc->done) return;
    // first row loaded ahead of the reductions (as in update_zr32_kernel)
    const int j0 = blockIdx.x * kBlock + threadIdx.x;
    float r0 = 0.0f, p0 = 0.0f, m0 = 0.0f;
    if (j0 < n) {
        r0 = r[j0];
        p0 = p[j0];
        if (minv) m0 = minv[j0];
    }
    double rho, rho0;
    reduce_partials2<kBlock>(rr_part, gridDim.x, rho_part, gridDim.x, rho, rho0);
    const double beta = rho / rho0;
    const double level = stop_part ? reduce_partials<kBlock>(stop_part, gridDim.x) : rho;
    const bool stop = fabs(level) <= sc->tol2;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        sc->rho = rho;
        sc->beta = beta;
        if (stop) sc->done = 1;
    }
    if (stop) return;
    const float b32 = (float)beta;
    if (j0 < n) p[j0] = minv ? fmaf(b32, p0, m0 * r0) : fmaf(b32, p0, r0);
    if (minv) {
        for (int j = j0 + gridDim.x * kBlock; j < n; j += gridDim.x * kBlock)
            p[j] = fmaf(b32, p[j], minv[j] * r[j]);
    } else {
        for (int j = j0 + gridDim.x * kBlock; j < n; j += gridDim.x * kBlock)
            p[j] = fmaf(b32, p[j], r[j]);
    }
}

__global__ void sweep_prepare_kernel(Scalars *outer, Scalars *inner, double rhs_norm,
                                     double rtol, int maxit, int first) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        if (first) {
            outer->inner_total = 0;
            outer->sweeps = 0;
            outer->sweep1_steps = 0;
            outer->sweep2_steps = 0;
        } else {
            outer->inner_total += inner->steps;
            if (outer->sweeps == 1) outer->sweep1_steps = inner->steps;
            if (outer->sweeps == 2) outer->sweep2_steps = inner->steps;
        }
        const double rel = first ? 1.0 : outer->rnorm / rhs_norm;
        const bool stop = (!first && rel <= rtol) || outer->inner_total >= maxit;
        const double t = fmax(1e-6, 0.5 * rtol / fmax(rel, 1e-300));
        inner->rtol2 = t * t;
        inner->tol2 = 0.0;
        inner->stop_enabled = 1;
        inner->steps = 0;
        inner->done = stop ? 1 : 0;
        outer->skip_resid = stop ? 1 : 0;
        if (!stop) outer->sweeps += 1;
    }
}

// x += d (FP64 accumulation of the FP32 correction).
__global__ __launch_bounds__(kBlock) void refine_kernel(double *__restrict__ x,
                                                       const float *__restrict__ d, int n) {
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock)
        x[j] += (double)d[j];
}

// r = b - A z (in place over r = A z), p = r, partials of r.r: restart of an
// FP64 solve from its true residual.
__global__ __launch_bounds__(kBlock) void residual_restart_kernel(
    const double *__restrict__ b, double *__restrict__ r, double *__restrict__ p, int n,
    double *__restrict__ rr_out, const double *__restrict__ minv) {
    double acc = 0.0;
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock) {
        const double rj = b[j] - r[j];
        const double sj = minv ? minv[j] * rj : rj;
        r[j] = rj;
        p[j] = sj;
        acc += rj * sj;
    }
    store_block_partial<kBlock>(acc, rr_out);
}

// out = (float) in (the FP32 copy of the preconditioner).
__global__ __launch_bounds__(kBlock) void narrow_kernel(const double *__restrict__ in,
                                                       float *__restrict__ out, int n) {
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock)
        out[j] = (float)in[j];
}

// minv = 1 / diag(A) from the CSR (Jacobi).
__global__ __launch_bounds__(kBlock) void jacobi_kernel(const int *__restrict__ rowstr,
                                                       const int *__restrict__ colidx,
                                                       const double *__restrict__ a,
                                                       double *__restrict__ minv, int n) {
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock) {
        double d = 0.0;
        for (int k = rowstr[j]; k < rowstr[j + 1]; k++)
            if (colidx[k] == j) d = a[k];
        minv[j] = d != 0.0 ? 1.0 / d : 1.0;
    }
}

// ---------------------------------------------------------------------------

hipError_t prepare_mixed_kernels(int lds_bytes) {
    hipError_t last = hipSuccess;
    for (const void *f : {reinterpret_cast<const void *>(&spmv_sell32_kernel<1024, kSellUnr32, false>),
                          reinterpret_cast<const void *>(&spmv_sell32_kernel<512, kSellUnr32, false>),
                          reinterpret_cast<const void *>(&spmv_sell32_kernel<1024, kSellUnr32, true>),
                          reinterpret_cast<const void *>(&spmv_sell32_kernel<512, kSellUnr32, true>)}) {
        hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, lds_bytes);
        if (e != hipSuccess) last = e;
    }
    return last;
}

template <bool PUPD>
static hipError_t launch_spmv32(const Fp64Problem &pb, const MixedProblem &mp, const float *v,
                                float *p_new, const double *rr_part, const double *rho_part,
                                int rev, hipStream_t st) {
    const SellDev &sd = pb.sell;
    if (sd.block == 512)
        spmv_sell32_kernel<512, kSellUnr32, PUPD><<<sd.grid, 512, mp.lds_bytes, st>>>(
            sd.wg_slice, sd.slice_off, sd.slice_row, mp.val, sd.col, v, mp.rows_part, mp.part_pq,
            pb.n, sd.W, sd.nrange, mp.sc, mp.r, p_new, rr_part, rho_part, pb.vec_blocks, rev,
            mp.stage_rows);
    else
        spmv_sell32_kernel<1024, kSellUnr32, PUPD><<<sd.grid, 1024, mp.lds_bytes, st>>>(
            sd.wg_slice, sd.slice_off, sd.slice_row, mp.val, sd.col, v, mp.rows_part, mp.part_pq,
            pb.n, sd.W, sd.nrange, mp.sc, mp.r, p_new, rr_part, rho_part, pb.vec_blocks, rev,
            mp.stage_rows);
    return hipGetLastError();
}

hipError_t launch_spmv32_kernel(const Fp64Problem &pb, const MixedProblem &mp, int step,
                                hipEvent_t e0, hipEvent_t e1, hipStream_t st) {
    const SellDev &sd = pb.sell;
    const int rev = step & 1;
    const float *v = mp.p;
    const float *r_none = nullptr;
    float *p_none = nullptr;
    const double *d_none = nullptr;
    if (sd.block == 512)
        hipExtLaunchKernelGGL(spmv_sell32_kernel<512, kSellUnr32, false>, dim3(sd.grid), dim3(512),
                              (uint32_t)mp.lds_bytes, st, e0, e1, 0u, sd.wg_slice, sd.slice_off,
                              sd.slice_row, mp.val, sd.col, v, mp.rows_part, mp.part_pq, pb.n, sd.W,
                              sd.nrange, mp.sc, r_none, p_none, d_none, d_none, pb.vec_blocks, rev,
                              mp.stage_rows);
    else
        hipExtLaunchKernelGGL(spmv_sell32_kernel<1024, kSellUnr32, false>, dim3(sd.grid),
                              dim3(1024), (uint32_t)mp.lds_bytes, st, e0, e1, 0u, sd.wg_slice,
                              sd.slice_off, sd.slice_row, mp.val, sd.col, v, mp.rows_part,
                              mp.part_pq, pb.n, sd.W, sd.nrange, mp.sc, r_none, p_none, d_none,
                              d_none, pb.vec_blocks, rev, mp.stage_rows);
    return hipGetLastError();
}

hipError_t launch_to_fp32(const double *src, float *dst, long long count, hipStream_t st) {
    to_fp32_kernel<<<2048, kBlock, 0, st>>>(src, dst, count);
    return hipGetLastError();
}

static const double *rho32_source(const MixedProblem &mp, int step) {
    return step == 0 ? mp.part_xx : mp.part_rr[(step - 1) & 1];
}

hipError_t enqueue_inner32(const Fp64Problem &pb, const MixedProblem &mp, int first_sweep,
                           int steps, hipStream_t st) {
    const int vb = pb.vec_blocks;
    init32_kernel<<<vb, kBlock, 0, st>>>(pb.x, pb.r, first_sweep, pb.n, mp.d, mp.r, mp.p,
                                         mp.part_xx, mp.minv);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && mp.pop)  // p = z = M^-1 r; r0.z0 into rz[1], r0.r0 stays in xx
        e = launch_precond_apply32(*mp.pop, mp.r, mp.pz, mp.p, mp.part_rz[1], vb, nullptr, st);
    const SellDev &sd = pb.sell;
    // Plain CG: the p update of step it-1 runs inside step it's SpMV (two
    // launches per step), p alternating between mp.p and mp.p_alt.
    const bool fused = !mp.pop && !mp.minv && mp.p_alt;
    float *pbuf[2] = {mp.p, mp.p_alt};
    for (int it = 0; it < steps && e == hipSuccess; it++) {
        float *p_it = fused ? pbuf[it & 1] : mp.p;
        if (fused && it > 0)
            e = launch_spmv32<true>(pb, mp, pbuf[(it - 1) & 1], p_it, mp.part_rr[(it - 1) & 1],
                                    rho32_source(mp, it - 1), it & 1, st);
        else
            e = launch_spmv32<false>(pb, mp, p_it, nullptr, nullptr, nullptr, it & 1, st);
        if (e != hipSuccess) break;
        // sparse preconditioner: alpha from the previous step's r.z (the
        // start's in rz[1]); beta from r.z; the stop test on r.r
        const double *rho_prev = mp.pop ? mp.part_rz[(it + 1) & 1] : rho32_source(mp, it);
        update_zr32_kernel<<<vb, kBlock, 0, st>>>(p_it, mp.q, mp.d, mp.r, pb.n, rho_prev,
                                                  mp.part_pq, sd.grid, mp.part_rr[it & 1],
                                                  mp.rows_part, sd.nblk, mp.sc, it == 0 ? 1 : 0,
                                                  mp.minv, mp.pop ? mp.part_xx : nullptr);
        e = hipGetLastError();
        if (e != hipSuccess || fused) continue;
        if (mp.pop) {
            e = launch_precond_apply32(*mp.pop, mp.r, mp.pz, nullptr, mp.part_rz[it & 1], vb,
                                       mp.sc, st);
            if (e != hipSuccess) break;
            update_p32_kernel<<<vb, kBlock, 0, st>>>(mp.pz, mp.p, pb.n, mp.part_rz[it & 1],
                                                     rho_prev, mp.sc, nullptr,
                                                     mp.part_rr[it & 1]);
        } else {
            update_p32_kernel<<<vb, kBlock, 0, st>>>(mp.r, mp.p, pb.n, mp.part_rr[it & 1],
                                                     rho_prev, mp.sc, mp.minv, nullptr);
        }
        e = hipGetLastError();
    }
    return e;
}

hipError_t launch_sweep_prepare(Scalars *outer, Scalars *inner, double rhs_norm, double rtol,
                                int maxit, int first, hipStream_t st) {
    sweep_prepare_kernel<<<1, 64, 0, st>>>(outer, inner, rhs_norm, rtol, maxit, first);
    return hipGetLastError();
}

hipError_t launch_refine(const Fp64Problem &pb, const MixedProblem &mp, hipStream_t st) {
    refine_kernel<<<pb.vec_blocks, kBlock, 0, st>>>(pb.z, mp.d, pb.n);
    return hipGetLastError();
}

hipError_t launch_residual_restart(const Fp64Problem &pb, hipStream_t st) {
    residual_restart_kernel<<<pb.vec_blocks, kBlock, 0, st>>>(pb.x, pb.r, pb.p, pb.n,
                                                              pb.part.xx, pb.minv);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !pb.pop) return e;
    return launch_precond_apply(*pb.pop, pb.r, pb.pz, pb.p, pb.part.rz[1], pb.vec_blocks,
                                nullptr, st);
}

hipError_t launch_narrow(const double *in, float *out, int n, int blocks, hipStream_t st) {
    narrow_kernel<<<blocks, kBlock, 0, st>>>(in, out, n);
    return hipGetLastError();
}

hipError_t launch_jacobi(const Fp64Problem &pb, double *minv, hipStream_t st) {
    jacobi_kernel<<<pb.vec_blocks, kBlock, 0, st>>>(pb.rowstr, pb.colidx, pb.a, minv, pb.n);
    return hipGetLastError();
}

}  // namespace npbcg_dev
