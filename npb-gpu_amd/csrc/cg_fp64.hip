// FP64 conj_grad kernels for gfx950 (multi-kernel engine).
//
// One inner CG iteration of the reference (CUDA/CG/cg.cu:814-837) is six
// launches and two blocking device-to-host copies there (k3 SpMV, k4 p.q,
// k5 two axpys, k6 r.r, k7 p update, with alpha/beta formed on the host).
// Here nothing returns to the host.  In the reference's step order it is
// three launches:
//   spmv       q = A p, per-workgroup partials of p.q                 (k3+k4)
//   update_zr  reduces rho and p.q partials -> alpha; z += alpha p,
//              r -= alpha q; partials of the new r.r                  (k5+k6)
//   update_p   reduces old and new r.r -> beta; p = r + beta p        (k7)
// In the one-reduction form (conj_grad's default on the SELL layout) it is
// two: spmv of r (s = A r, partials of r.s) and update_cg, which forms beta,
// p, q = A p by recursion, p.q from r.s, alpha, z and r (k3..k7).
// Around the inner loop:
//   cg_init    optionally x = xscale z (the previous outer step's
//              normalisation); q = z = 0, r = p = x through index n;
//              partials of x.x                                   (k11+k1+k2)
//   resid      r = A z with partials of ||x - r||^2, x.z, z.z, then a
//              one-workgroup finish forming rnorm, zeta, xscale (k8+k9+k10)
//
// No kernel hands values to other workgroups of its own launch, so there are
// no grid-wide fences or atomics: each workgroup stores one partial and the
// next launch reduces all partials in a fixed order (every consumer
// workgroup computes the same bits).  Kernel boundaries give the ordering.
#include <hip/hip_ext.h>

#include <algorithm>

#include "cg_device.h"
#include "cg_kernels.h"
#include "learned_kernels.h"

namespace npbcg_dev {

// Entries per lane in flight in the SELL SpMV stream (one chunk): 12 x 10 B
// per lane, the most that fits 128 VGPRs with the prefetched chunk.
constexpr int kSellUnr = 12;

// ---------------------------------------------------------------------------
// CSR SpMV (one group of SEG lanes per row)
// ---------------------------------------------------------------------------

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

// MODE 0: y = q = A p, partial p.q.
// MODE 1: y = r = A z, partials ||x - r||^2, x.z, z.z (stride gridDim.x).
template <int SEG, int UNR, int MODE>
__global__ __launch_bounds__(kBlock) void spmv_csr_kernel(
    const int *__restrict__ rowstr, const int *__restrict__ colidx,
    const double *__restrict__ a, const double *__restrict__ v,
    double *__restrict__ y, const double *__restrict__ x, int n,
    double *__restrict__ part, const Scalars *__restrict__ sc) {
    if (MODE == 0 && sc->done) return;  // a solve already met its tolerance
    if (MODE == 1 && sc->skip_resid) return;  // a refinement sweep that is not needed
    constexpr int kGroups = kBlock / SEG;
    const int sl = threadIdx.x % SEG;
    // With a whole wavefront per row the row index is wave-uniform; saying so
    // (readfirstlane) keeps the row loop's exit a uniform branch around the
    // cross-lane reduction.
    const int group = SEG == kWave ? __builtin_amdgcn_readfirstlane(threadIdx.x / SEG)
                                   : (int)(threadIdx.x / SEG);
    const int g = blockIdx.x * kGroups + group;
    const int ng = gridDim.x * kGroups;
    double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0;
    for (int row = g; row < n; row += ng) {
        const int s = rowstr[row], e = rowstr[row + 1];
        const double sum = row_dot<SEG, UNR>(colidx, a, v, s, e, sl);
        if (sl == 0) {
            y[row] = sum;
            if constexpr (MODE == 0) {
                acc0 += v[row] * sum;
            } else {
                const double xv = x[row], zv = v[row];
                const double dlt = xv - sum;
                acc0 += dlt * dlt;
                acc1 += xv * zv;
                acc2 += zv * zv;
            }
        }
    }
    store_block_partial<kBlock>(acc0, part);
    if constexpr (MODE == 1) {
        __syncthreads();
        store_block_partial<kBlock>(acc1, part + gridDim.x);
        __syncthreads();
        store_block_partial<kBlock>(acc2, part + 2 * gridDim.x);
    }
}

// ---------------------------------------------------------------------------
// LDS-staged SpMV over the column-blocked SELL copy (tiled_layout.h).
//
// Workgroup wg owns column block b = wg / nrange of row range wg % nrange.
// It stages the block's W entries of the input vector into LDS once, with
// coalesced 16-byte loads, then each wavefront streams 64-row slices: lane l
// accumulates its own row down the slice (value and 16-bit column loads are
// contiguous across the wavefront; UNR entries are in flight per lane), so no
// cross-lane reduction is needed.  The row's sum over this block goes to
// the lane's slot, part[64 * slice + lane], so each slice stores its 64 sums
// as one coalesced run (scattered 8-byte stores by row were written back as
// partial lines, several times their size); update_zr (or resid_combine)
// finds row j's sums through row_slot and adds the blocks in order
// b = 0, 1, ... .  The random 8-byte gathers that bound the CSR kernel (one
// 128-byte L2 line per gather) become LDS reads, and the vector crosses from
// L2 into LDS once per workgroup.  For q = A p the workgroup also forms its
// partial of p.q = sum_b sum_rows p[row] * (block b's sum of the row).
// ---------------------------------------------------------------------------

// PUPD: the input vector is p_new = r + beta p_old, built here instead of by
// a separate update_p launch.  beta comes from the r.r partials (rr_part new,
// rho_part old); the workgroups of row range 0 store their block of p_new.
template <int BLOCK, int UNR, int MODE, bool PUPD>
__global__ __launch_bounds__(BLOCK) void spmv_sell_kernel(
    const int *__restrict__ wg_slice, const int *__restrict__ slice_off,
    const int *__restrict__ slice_row, const double *__restrict__ val,
    const uint16_t *__restrict__ col, const double *__restrict__ v,
    double *__restrict__ rows_part, double *__restrict__ dot_part, int n, int W,
    int nrange, const double *__restrict__ r, double *__restrict__ p_new,
    const double *__restrict__ rr_part, const double *__restrict__ rho_part, int n_vec,
    const Scalars *__restrict__ sc, int rev) {
    if (MODE == 0 && sc->done) return;  // a solve already met its tolerance
    if (MODE == 1 && sc->skip_resid) return;  // a refinement sweep that is not needed
    // Dynamic LDS: W vector entries, then one p.q slot per slice of this
    // workgroup (the p.q partial is summed in slice order).
    extern __shared__ double2 lds_v2[];  // 16-byte aligned
    double *vblk = reinterpret_cast<double *>(lds_v2);
    double *slice_dot = vblk + W;
    constexpr int kWaves = BLOCK / kWave;
    const int wg = blockIdx.x;
    const int b = wg / nrange;
    const int c0 = b * W;
    const int wb = min(W, n - c0);
    double beta = 0.0;
    if constexpr (PUPD) {
        double rr, rho;
        reduce_partials2<BLOCK>(rr_part, n_vec, rho_part, n_vec, rr, rho);
        beta = rr / rho;
    }
    if constexpr (PUPD) {
        const double2 *v2 = reinterpret_cast<const double2 *>(v + c0);
        const double2 *r2 = reinterpret_cast<const double2 *>(r + c0);
        double2 *o2 = reinterpret_cast<double2 *>(p_new + c0);
        const bool store = (wg % nrange) == 0;
#pragma unroll 4
        for (int i = threadIdx.x; i < wb / 2; i += BLOCK) {
            const double2 pv = v2[i], rv = r2[i];
            double2 nv;
            nv.x = rv.x + beta * pv.x;
            nv.y = rv.y + beta * pv.y;
            lds_v2[i] = nv;
            if (store) o2[i] = nv;
        }
        if ((wb & 1) && threadIdx.x == 0) {
            const double nv = r[c0 + wb - 1] + beta * v[c0 + wb - 1];
            vblk[wb - 1] = nv;
            if (store) p_new[c0 + wb - 1] = nv;
        }
    }
    const int lane = threadIdx.x & (kWave - 1);
    // Wave-uniform indices (readfirstlane): every branch below that depends on
    // them is a scalar branch, so the wavefront stays converged around the
    // cross-lane reduction.
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int s0 = wg_slice[wg];
    const int nsl = wg_slice[wg + 1] - s0;
    // Slice (index in this workgroup) of round rd for this wave, or -1, in
    // "snake" order (the direction alternates by round).  The layout places
    // the slices so that every wavefront's total width under this order is
    // close to the others' (build_sell_blocks).  The index grows with rd, so
    // after the first -1 there is none.
    auto slice_of = [&](int rd) {
        const int k = rd * kWaves + ((rd & 1) ? kWaves - 1 - wave : wave);
        return k < nsl ? k : -1;
    };
    // This wave's rounds, walked forward or (rev) backward: successive SpMVs
    // alternate, so each pass starts on the slices the previous one streamed
    // last, which the memory-side cache still holds when the matrix is about
    // its size.  Only the order of the slices changes, not any sum.
    int nrw = 0;
    while (slice_of(nrw) >= 0) nrw++;
    auto slice_at = [&](int i) { return i < nrw ? slice_of(rev ? nrw - 1 - i : i) : -1; };
    // The entry stream is software-pipelined in chunks of UNR entries
    // (sell_chunk): the next chunk's loads, or the next slice's first chunk
    // and lane rows, go out before the current chunk's LDS gathers.  The first
    // chunk of this wave's first slice is loaded before the vector block is
    // staged, so its memory latency overlaps the LDS fill.
    double na[UNR];
    int nc[UNR];
    int cur = slice_at(0), cbase = 0, cwidth = 0, nrow = -1;
    if (cur >= 0) {
        cbase = slice_off[s0 + cur];
        cwidth = (slice_off[s0 + cur + 1] - cbase) / kWave;
        nrow = slice_row[(size_t)(s0 + cur) * kWave + lane];
        if (cwidth > 0) sell_chunk<double, UNR>(val, col, cbase, cwidth, 0, lane, na, nc);
    }
    if constexpr (!PUPD) {
        // a block of up to 20,480 doubles: 10 (or 20) 16-byte loads per thread
        stage_block<BLOCK, 10240 / BLOCK>(reinterpret_cast<const double2 *>(v + c0), lds_v2,
                                          wb / 2);
        if ((wb & 1) && threadIdx.x == 0) vblk[wb - 1] = v[c0 + wb - 1];
    }
    __syncthreads();
    for (int i = 0; cur >= 0; i++) {
        const int k = cur, width = cwidth, base = cbase, row = nrow;
        const int nxt = slice_at(i + 1);
        int xbase = 0, xwidth = 0;
        if (nxt >= 0) {
            xbase = slice_off[s0 + nxt];
            xwidth = (slice_off[s0 + nxt + 1] - xbase) / kWave;
        }
        // p of the lane's row for p.q, loaded while the slice streams
        double pr = 0.0;
        if constexpr (MODE == 0)
            if (row >= 0) pr = PUPD ? r[row] + beta * v[row] : v[row];
        double acc = 0.0;
        const int nch = (width + UNR - 1) / UNR;
        for (int ch = 0; ch < nch; ch++) {
            double av[UNR];
            int c[UNR];
#pragma unroll
            for (int u = 0; u < UNR; u++) {
                av[u] = na[u];
                c[u] = nc[u];
            }
            sell_trim<double, UNR>(width - ch * UNR, av, c);
            if (ch + 1 < nch) {
                sell_chunk<double, UNR>(val, col, base, width, (ch + 1) * UNR, lane, na, nc);
            } else if (nxt >= 0) {
                nrow = slice_row[(size_t)(s0 + nxt) * kWave + lane];
                if (xwidth > 0) sell_chunk<double, UNR>(val, col, xbase, xwidth, 0, lane, na, nc);
            }
#pragma unroll
            for (int u = 0; u < UNR; u++) acc = fma(av[u], vblk[c[u]], acc);
        }
        if (nch == 0 && nxt >= 0) {
            nrow = slice_row[(size_t)(s0 + nxt) * kWave + lane];
            if (xwidth > 0) sell_chunk<double, UNR>(val, col, xbase, xwidth, 0, lane, na, nc);
        }
        // the slice's 64 block sums, one coalesced run in slot order
        rows_part[(size_t)(s0 + k) * kWave + lane] = acc;
        if constexpr (MODE == 0) {
            const double d = segment_sum<64>(pr * acc);
            if (lane == 0) slice_dot[k] = d;
        }
        cur = nxt;
        cbase = xbase;
        cwidth = xwidth;
    }
    if constexpr (MODE == 0) {
        __syncthreads();
        double dot = 0.0;
        for (int i = threadIdx.x; i < nsl; i += BLOCK) dot += slice_dot[i];
        store_block_partial<BLOCK>(dot, dot_part);
    }
}

// y = sum over blocks of the SELL SpMV's row partials (blocks in order),
// each found at its slot.
__global__ __launch_bounds__(kBlock) void combine_rows_kernel(
    const double *__restrict__ rows_part, const int *__restrict__ row_slot, int nblk,
    double *__restrict__ y, int n) {
    for (int j = xcd_chunk(blockIdx.x, gridDim.x) * kBlock + threadIdx.x; j < n;
         j += gridDim.x * kBlock) {
        double s = rows_part[row_slot[j]];
#pragma unroll 8
        for (int b = 1; b < nblk; b++) s += rows_part[row_slot[(size_t)b * n + j]];
        y[j] = s;
    }
}

// r = sum over blocks of the residual SpMV's row partials, with partials of
// ||x - r||^2, x.z and z.z (stride gridDim.x).
__global__ __launch_bounds__(kBlock) void resid_combine_kernel(
    const double *__restrict__ rows_part, const int *__restrict__ row_slot, int nblk,
    const double *__restrict__ x, const double *__restrict__ z, double *__restrict__ r, int n,
    double *__restrict__ res, const Scalars *__restrict__ sc) {
    if (sc->skip_resid) return;
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    for (int j = xcd_chunk(blockIdx.x, gridDim.x) * kBlock + threadIdx.x; j < n;
         j += gridDim.x * kBlock) {
        double rj = rows_part[row_slot[j]];
#pragma unroll 8
        for (int b = 1; b < nblk; b++) rj += rows_part[row_slot[(size_t)b * n + j]];
        r[j] = rj;
        const double xv = x[j], zv = z[j];
        const double dlt = xv - rj;
        a0 += dlt * dlt;
        a1 += xv * zv;
        a2 += zv * zv;
    }
    store_block_partial<kBlock>(a0, res);
    __syncthreads();
    store_block_partial<kBlock>(a1, res + gridDim.x);
    __syncthreads();
    store_block_partial<kBlock>(a2, res + 2 * gridDim.x);
}

// ---------------------------------------------------------------------------
// Element-wise kernels
// ---------------------------------------------------------------------------

// z += alpha p; r -= alpha q with alpha = rho / (p.q); partials of new r.r.
// rho comes from the partials of the previous r.r (or of x.x for the first
// step), p.q from the SpMV's partials.
// With nblk > 0, q is first assembled from the SELL SpMV's per-block row
// partials (blocks added in order, each found at its slot) and stored.
__global__ __launch_bounds__(kBlock) void update_zr_kernel(
    const double *__restrict__ p, double *__restrict__ q, double *__restrict__ z,
    double *__restrict__ r, int n, const double *__restrict__ rho_part, int n_rho,
    const double *__restrict__ pq_part, int n_pq, double *__restrict__ rr_out,
    const double *__restrict__ rows_part, const int *__restrict__ row_slot, int nblk,
    Scalars *sc, int first_step, const double *__restrict__ minv,
    const double *__restrict__ tol_part) {
    // The solve's stop flag is read with the prefetch below and tested after
    // it (no kernel of this launch writes it), so its round trip overlaps
    // the others.  This thread's first row (its block partials of q, p, z, r)
    // is loaded before the scalar reductions, which it does not depend on, so
    // the memory round trips overlap; the sums keep their order.
    const int done = sc->done;
    const int j0 = xcd_chunk(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    constexpr int kPre = 8;  // block partials of q prefetched
    double qb[kPre], p0 = 0.0, z0 = 0.0, r0 = 0.0, m0 = 0.0;
    if (j0 < n) {
        if (nblk > 0) {
            int sl[kPre];
#pragma unroll
            for (int b = 0; b < kPre; b++)
                if (b < nblk) sl[b] = row_slot[(size_t)b * n + j0];
#pragma unroll
            for (int b = 0; b < kPre; b++)
                if (b < nblk) qb[b] = rows_part[sl[b]];
        } else {
            qb[0] = q[j0];
        }
        p0 = p[j0];
        z0 = z[j0];
        r0 = r[j0];
        if (minv) m0 = minv[j0];
    }
    if (done) return;  // the same value for the whole grid
    double rho, d;
    reduce_partials2<kBlock>(rho_part, n_rho, pq_part, n_pq, rho, d);
    const double alpha = rho / d;
    // the stop level: rtol^2 times this start's r0.r0 (tol_part), or times
    // rho itself (plain CG, diagonal PCG)
    const double rho_tol = (first_step && tol_part) ? reduce_partials<kBlock>(tol_part, n_rho) : rho;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        sc->rho0 = rho;
        sc->d = d;
        sc->alpha = alpha;
        if (first_step) {
            sc->tol2 = sc->rtol2 * fabs(rho_tol);
            sc->bb = rho_tol;
        }
        sc->steps += 1;
    }
    double acc = 0.0;
    if (j0 < n) {
        double qj;
        if (nblk > 0) {
            qj = qb[0];
#pragma unroll
            for (int b = 1; b < kPre; b++)
                if (b < nblk) qj += qb[b];
#pragma unroll 8
            for (int b = kPre; b < nblk; b++) qj += rows_part[row_slot[(size_t)b * n + j0]];
            q[j0] = qj;
        } else {
            qj = qb[0];
        }
        z[j0] = z0 + alpha * p0;
        const double rj = r0 - alpha * qj;
        r[j0] = rj;
        acc += minv ? rj * (m0 * rj) : rj * rj;
    }
    for (int j = j0 + gridDim.x * kBlock; j < n; j += gridDim.x * kBlock) {
        double qj;
        if (nblk > 0) {
            qj = rows_part[row_slot[j]];
#pragma unroll 8
            for (int b = 1; b < nblk; b++) qj += rows_part[row_slot[(size_t)b * n + j]];
            q[j] = qj;
        } else {
            qj = q[j];
        }
        z[j] = z[j] + alpha * p[j];
        const double rj = r[j] - alpha * qj;
        r[j] = rj;
        acc += minv ? rj * (minv[j] * rj) : rj * rj;
    }
    store_block_partial<kBlock>(acc, rr_out);
}

// p_out = r + beta p_in with beta = (new r.r) / (old r.r), both from
// partials (p_in may equal p_out).  The stop test uses stop_part (the r.r of a
// sparse-preconditioned solve) when given, else the new rho.
__global__ __launch_bounds__(kBlock) void update_p_kernel(
    const double *__restrict__ r, const double *p_in, double *p_out, int n,
    const double *__restrict__ rr_part, const double *__restrict__ rho_part, int n_vec,
    Scalars *sc, const double *__restrict__ minv, const double *__restrict__ stop_part) {
    if (sc->done) return;
    // first row loaded ahead of the reductions (as in update_zr_kernel)
    const int j0 = blockIdx.x * kBlock + threadIdx.x;
    double r0 = 0.0, p0 = 0.0, m0 = 0.0;
    if (j0 < n) {
        r0 = r[j0];
        p0 = p_in[j0];
        if (minv) m0 = minv[j0];
    }
    double rho, rho0;
    reduce_partials2<kBlock>(rr_part, n_vec, rho_part, n_vec, rho, rho0);
    const double beta = rho / rho0;
    const double level = stop_part ? reduce_partials<kBlock>(stop_part, n_vec) : rho;
    const bool stop = sc->stop_enabled && fabs(level) <= sc->tol2;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        sc->rho = rho;
        sc->beta = beta;
        // Solve mode: the recursive residual met the tolerance; the remaining
        // steps of the launched sequence return at entry.  p is left as is
        // (the solution z is final).
        if (stop) sc->done = 1;
    }
    if (stop) return;
    if (j0 < n) p_out[j0] = minv ? m0 * r0 + beta * p0 : r0 + beta * p0;
    if (minv) {
        for (int j = j0 + gridDim.x * kBlock; j < n; j += gridDim.x * kBlock)
            p_out[j] = minv[j] * r[j] + beta * p_in[j];
    } else {
        for (int j = j0 + gridDim.x * kBlock; j < n; j += gridDim.x * kBlock)
            p_out[j] = r[j] + beta * p_in[j];
    }
}

// One CG step of the one-reduction form (Chronopoulos and Gear, 1989), after
// s = A r (the SELL SpMV of r, its block row partials and r.s partials):
//   beta = rho / rho_prev, p = r + beta p, q = s + beta q (= A p),
//   delta = p.q = r.s - beta rho / alpha_prev, alpha = rho / delta,
//   z += alpha p, r -= alpha q, partials of the new r.r.
// It is the reference's step (cg.cu:703-760) with p.q taken from r.s and the
// previous step's scalars (r_k orthogonal to r_(k-1), A symmetric), so a step
// is two launches; the first step has beta = 0 (p = r = x, q = s).  rho_prev
// and alpha_prev sit in slot (step - 1) & 1 of the scalars; workgroup 0
// writes this step's into slot step & 1.
__global__ __launch_bounds__(kBlock) void update_cg_kernel(
    double *__restrict__ p, double *__restrict__ q, double *__restrict__ z,
    double *__restrict__ r, int n, const double *__restrict__ rho_part, int n_rho,
    const double *__restrict__ rs_part, int n_rs, double *__restrict__ rr_out,
    const double *__restrict__ rows_part, const int *__restrict__ row_slot, int nblk,
    Scalars *sc, int step) {
    // this thread's first row, loaded ahead of the reductions.  Without
    // row_slot, rows_part holds s itself (nblk = 1, row j at j).
    const int j0 = xcd_chunk(blockIdx.x, gridDim.x) * kBlock + threadIdx.x;
    constexpr int kPre = 8;
    double sb[kPre], p0 = 0.0, q0 = 0.0, z0 = 0.0, r0 = 0.0;
    if (j0 < n) {
        if (row_slot) {
            int sl[kPre];
#pragma unroll
            for (int b = 0; b < kPre; b++)
                if (b == 0 || b < nblk) sl[b] = row_slot[(size_t)b * n + j0];
#pragma unroll
            for (int b = 0; b < kPre; b++)
                if (b == 0 || b < nblk) sb[b] = rows_part[sl[b]];
        } else {
            sb[0] = rows_part[j0];
        }
        p0 = p[j0];
        q0 = q[j0];
        z0 = z[j0];
        r0 = r[j0];
    }
    // the previous step's scalars (slot (step - 1) & 1; workgroup 0 writes
    // slot step & 1 below) are loaded before the reductions as well
    double rho_prev = 1.0, alpha_prev = 1.0;
    if (step > 0) {
        rho_prev = sc->cg_rho[(step - 1) & 1];
        alpha_prev = sc->cg_alpha[(step - 1) & 1];
    }
    double rho, rs;
    reduce_partials2<kBlock>(rho_part, n_rho, rs_part, n_rs, rho, rs);
    double beta = 0.0, delta = rs;
    if (step > 0) {
        beta = rho / rho_prev;
        delta = rs - beta * rho / alpha_prev;
    }
    const double alpha = rho / delta;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        sc->cg_rho[step & 1] = rho;
        sc->cg_alpha[step & 1] = alpha;
        sc->rho0 = rho;
        sc->beta = beta;
        sc->d = delta;
        sc->alpha = alpha;
        sc->steps += 1;
    }
    double acc = 0.0;
    for (int j = j0; j < n; j += gridDim.x * kBlock) {
        double sj, pj, qj, zj, rj;
        if (j == j0) {
            sj = sb[0];
#pragma unroll
            for (int b = 1; b < kPre; b++)
                if (b < nblk) sj += sb[b];
#pragma unroll 8
            for (int b = kPre; b < nblk; b++) sj += rows_part[row_slot[(size_t)b * n + j]];
            pj = p0;
            qj = q0;
            zj = z0;
            rj = r0;
        } else if (!row_slot) {
            sj = rows_part[j];
            pj = p[j];
            qj = q[j];
            zj = z[j];
            rj = r[j];
        } else {
            sj = rows_part[row_slot[j]];
#pragma unroll 8
            for (int b = 1; b < nblk; b++) sj += rows_part[row_slot[(size_t)b * n + j]];
            pj = p[j];
            qj = q[j];
            zj = z[j];
            rj = r[j];
        }
        pj = rj + beta * pj;
        qj = sj + beta * qj;
        p[j] = pj;
        q[j] = qj;
        z[j] = zj + alpha * pj;
        const double rn = rj - alpha * qj;
        r[j] = rn;
        acc += rn * rn;
    }
    store_block_partial<kBlock>(acc, rr_out);
}

// s = sum of the column blocks' row sums of A r (blocks in order, each found
// at its slot, the same sum update_cg forms), stored; partials of r.s.  Used
// when the SpMV leaves out r.s: with many column blocks each slice holds only
// a few entries per row, and gathering r[row] for the dot in every block cost
// more than this pass.
__global__ __launch_bounds__(kBlock) void combine_rs_kernel(
    const double *__restrict__ r, const double *__restrict__ rows_part,
    const int *__restrict__ row_slot, int nblk, int n, double *__restrict__ s_out,
    double *__restrict__ rs_out) {
    double acc = 0.0;
    for (int j = xcd_chunk(blockIdx.x, gridDim.x) * kBlock + threadIdx.x; j < n;
         j += gridDim.x * kBlock) {
        double sj = rows_part[row_slot[j]];
#pragma unroll 8
        for (int b = 1; b < nblk; b++) sj += rows_part[row_slot[(size_t)b * n + j]];
        s_out[j] = sj;
        acc += r[j] * sj;
    }
    store_block_partial<kBlock>(acc, rs_out);
}

// Start of conj_grad (cg.cu:671-686): q = z = 0, r = p = x for j <= n and
// partials of x.x over j < n.  With scale_x, x = xscale * z (j < n) is
// applied first: the normalisation of the previous outer step (:479-481).
__global__ __launch_bounds__(kBlock) void cg_init_kernel(
    double *__restrict__ x, double *__restrict__ z, double *__restrict__ p,
    double *__restrict__ q, double *__restrict__ r, int n, int scale_x,
    const Scalars *__restrict__ sc, double *__restrict__ xx_out,
    const double *__restrict__ minv) {
    const double xs = scale_x ? sc->xscale : 1.0;
    double acc = 0.0;
    for (int j = blockIdx.x * kBlock + threadIdx.x; j <= n; j += gridDim.x * kBlock) {
        double xv = x[j];
        if (scale_x && j < n) {
            xv = xs * z[j];
            x[j] = xv;
        }
        q[j] = 0.0;
        z[j] = 0.0;
        r[j] = xv;
        const double sv = (minv && j < n) ? minv[j] * xv : xv;  // s = M^-1 r
        p[j] = sv;
        if (j < n) acc += xv * sv;
    }
    store_block_partial<kBlock>(acc, xx_out);
}

// One workgroup: reduce the residual SpMV's partials and form
// rnorm = ||x - A z|| (cg.cu:795-800), norm_temp1 = x.z, norm_temp2 =
// 1/sqrt(z.z) and zeta = shift + 1/(x.z) (:505-514).
__global__ __launch_bounds__(kBlock) void resid_finish_kernel(
    const double *__restrict__ res, int count, Scalars *sc,
    double *__restrict__ zeta_hist, double *__restrict__ rnorm_hist, int hist_cap,
    int record) {
    if (sc->skip_resid) return;
    const double s0 = reduce_partials<kBlock>(res, count);
    const double s1 = reduce_partials<kBlock>(res + count, count);
    const double s2 = reduce_partials<kBlock>(res + 2 * count, count);
    if (threadIdx.x == 0) {
        const double rnorm = sqrt(s0);
        sc->rnorm = rnorm;
        sc->xz = s1;
        sc->zz = s2;
        sc->xscale = 1.0 / sqrt(s2);
        const double zeta = sc->shift + 1.0 / s1;
        sc->zeta = zeta;
        if (record) {
            const int it = sc->iter;
            if (it < hist_cap) {
                zeta_hist[it] = zeta;
                rnorm_hist[it] = rnorm;
            }
            sc->iter = it + 1;
        }
    }
}

// Partials of x.z and z.z (cg.cu:1182-1245, gpu_kernel_ten).
__global__ __launch_bounds__(kBlock) void norms_kernel(
    const double *__restrict__ x, const double *__restrict__ z, int n,
    double *__restrict__ part) {
    double a0 = 0.0, a1 = 0.0;
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock) {
        const double zv = z[j];
        a0 += x[j] * zv;
        a1 += zv * zv;
    }
    store_block_partial<kBlock>(a0, part);
    __syncthreads();
    store_block_partial<kBlock>(a1, part + gridDim.x);
}

__global__ __launch_bounds__(kBlock) void norms_finish_kernel(
    const double *__restrict__ part, int count, Scalars *sc) {
    const double s0 = reduce_partials<kBlock>(part, count);
    const double s1 = reduce_partials<kBlock>(part + count, count);
    if (threadIdx.x == 0) {
        sc->xz = s0;
        sc->zz = s1;
    }
}

// x = s * z for j < n (cg.cu:1261-1265, gpu_kernel_eleven); with
// from_scalars the scale is the last residual finish's xscale.
__global__ __launch_bounds__(kBlock) void scale_x_kernel(
    double *__restrict__ x, const double *__restrict__ z, int n, double s,
    int from_scalars, const Scalars *__restrict__ sc) {
    if (from_scalars) s = sc->xscale;
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock)
        x[j] = s * z[j];
}

// dst = src over `count` 16-byte elements (bandwidth reference): each thread
// issues four loads before its four stores.
__global__ __launch_bounds__(kBlock) void copy_kernel(const double2 *__restrict__ src,
                                                     double2 *__restrict__ dst, long long count) {
    const long long stride = (long long)gridDim.x * kBlock;
    long long i = blockIdx.x * (long long)kBlock + threadIdx.x;
    for (; i + 3 * stride < count; i += 4 * stride) {
        const double2 a = src[i], b = src[i + stride], c = src[i + 2 * stride],
                      d = src[i + 3 * stride];
        dst[i] = a;
        dst[i + stride] = b;
        dst[i + 2 * stride] = c;
        dst[i + 3 * stride] = d;
    }
    for (; i < count; i += stride) dst[i] = src[i];
}

// Reads `count` 16-byte elements (read-bandwidth reference): 1024-thread
// workgroups, eight loads in flight per thread, the grid sweeping the buffer
// in order (of the shapes measured on MI355X, the fastest read from both HBM
// and the memory-side cache); the sum is stored only on an impossible value,
// which keeps every load live without write traffic.
constexpr int kReadBlock = 1024;
constexpr int kReadUnroll = 8;
__global__ __launch_bounds__(kReadBlock) void read_kernel(const double2 *__restrict__ src,
                                                          long long count,
                                                          double *__restrict__ sink) {
    const long long stride = (long long)gridDim.x * kReadBlock;
    long long i = blockIdx.x * (long long)kReadBlock + threadIdx.x;
    double s = 0.0;
    for (; i + (kReadUnroll - 1) * stride < count; i += kReadUnroll * stride) {
        double2 v[kReadUnroll];
#pragma unroll
        for (int u = 0; u < kReadUnroll; u++) v[u] = src[i + u * stride];
#pragma unroll
        for (int u = 0; u < kReadUnroll; u++) s += v[u].x + v[u].y;
    }
    for (; i < count; i += stride) s += src[i].x + src[i].y;
    if (s == -1.2345e300) sink[0] = s;
}

// Resets the outer-loop bookkeeping.  A kernel rather than a memset so it may
// sit in a captured graph.
__global__ void reset_scalars_kernel(Scalars *sc, double shift, double rtol2, int stop) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        sc->iter = 0;
        sc->shift = shift;
        sc->rtol2 = rtol2;
        sc->tol2 = 0.0;
        sc->stop_enabled = stop;
        sc->done = 0;
        sc->steps = 0;
        sc->skip_resid = 0;
        sc->sweeps = 0;
        sc->inner_total = 0;
    }
}

// Clears the stop flag (a solve restarting from its true residual).
__global__ void clear_done_kernel(Scalars *sc, int end_of_solve) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        sc->done = 0;
        if (end_of_solve) {  // back to conj_grad / power use of the same scalars
            sc->stop_enabled = 0;
            sc->skip_resid = 0;
        }
    }
}

__global__ void fill_kernel(double *__restrict__ v, int count, double value) {
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < count; j += gridDim.x * kBlock)
        v[j] = value;
}

// ---------------------------------------------------------------------------
// Host-side launchers
// ---------------------------------------------------------------------------

namespace {

// Kernel launch on `st`; with events, they are recorded by the dispatch
// itself (hipExtLaunchKernelGGL), so their interval is the kernel's own run.
template <typename K, typename... Args>
hipError_t launch_k(K kernel, int grid, int block, int lds, hipStream_t st, hipEvent_t e0,
                    hipEvent_t e1, Args... args) {
    if (e0)
        hipExtLaunchKernelGGL(kernel, dim3(grid), dim3(block), (uint32_t)lds, st, e0, e1, 0u,
                              args...);
    else
        kernel<<<grid, block, lds, st>>>(args...);
    return hipGetLastError();
}

// The launch's error, after a profiling stamp when the run is profiled.
hipError_t noted(const Fp64Problem &pb, hipError_t e, const char *kname, int threads,
                 hipStream_t st) {
    if (e == hipSuccess && pb.marks) pb.marks->note(kname, threads, st);
    return e;
}

template <int SEG, int UNR, int MODE>
hipError_t launch_csr_t(const Fp64Problem &pb, hipStream_t st, hipEvent_t e0, hipEvent_t e1) {
    const double *v = MODE == 0 ? pb.p : pb.z;
    double *y = MODE == 0 ? pb.q : pb.r;
    double *part = MODE == 0 ? pb.part.pq : pb.part.res;
    return noted(pb,
                 launch_k(spmv_csr_kernel<SEG, UNR, MODE>, pb.spmv_blocks, kBlock, 0, st, e0, e1,
                          pb.rowstr, pb.colidx, pb.a, v, y, pb.x, pb.n, part,
                          (const Scalars *)pb.sc),
                 MODE == 0 ? "spmv_csr_kernel (q = A p)" : "spmv_csr_kernel (r = A z)", kBlock, st);
}

bool fused_p(const Fp64Problem &pb) {
    return pb.spmv_variant == kSpmvTiled && pb.fuse_p_update;
}

// The p buffer holding CG step `step`'s direction on the SELL path, where p
// alternates between p and p_alt (the fused update writes the other one).
double *p_of_step(const Fp64Problem &pb, int step) {
    return (fused_p(pb) && (step & 1)) ? pb.p_alt : pb.p;
}

// The rho partials feeding CG step `step`: x.x from init for step 0, else
// the r.r written by the previous step's update_zr.
const double *rho_source(const Fp64Problem &pb, int step) {
    return step == 0 ? pb.part.xx : pb.part.rr[(step - 1) & 1];
}

// MODE 0 at `step` (q = A p): for step > 0 the kernel forms p of this step
// from r and the previous step's p.  MODE 1: r = A z.  Slice order: forward
// on even steps, backward on odd ones and for r = A z, which follows the odd
// count of CG steps of an outer iteration.
template <int BLOCK, int MODE>
hipError_t launch_sell_t(const Fp64Problem &pb, int step, hipStream_t st, hipEvent_t e0,
                         hipEvent_t e1) {
    const SellDev &d = pb.sell;
    const int rev = MODE == 0 ? (step & 1) : 1;
    if (MODE == 0 && step > 0 && fused_p(pb)) {
        return noted(pb,
                     launch_k(spmv_sell_kernel<BLOCK, kSellUnr, 0, true>, d.grid, BLOCK, d.lds_bytes,
                              st, e0, e1, d.wg_slice, d.slice_off, d.slice_row, d.val, d.col,
                              (const double *)p_of_step(pb, step - 1), d.rows_part, pb.part.pq,
                              pb.n, d.W, d.nrange, (const double *)pb.r, p_of_step(pb, step),
                              (const double *)pb.part.rr[(step - 1) & 1],
                              rho_source(pb, step - 1), pb.vec_blocks, (const Scalars *)pb.sc,
                              rev),
                     "spmv_sell_kernel (p = r + beta p; q = A p)", BLOCK, st);
    }
    if (MODE == 0 && pb.cg_form && pb.dot_apart) {  // s = A r, block sums only
        return noted(pb,
                     launch_k(spmv_sell_kernel<BLOCK, kSellUnr, 1, false>, d.grid, BLOCK,
                              d.lds_bytes, st, e0, e1, d.wg_slice, d.slice_off, d.slice_row,
                              d.val, d.col, (const double *)pb.r, d.rows_part, pb.part.pq, pb.n,
                              d.W, d.nrange, (const double *)nullptr, (double *)nullptr,
                              (const double *)nullptr, (const double *)nullptr, 0,
                              (const Scalars *)pb.sc, step & 1),
                     "spmv_sell_kernel (s = A r, r.s apart)", BLOCK, st);
    }
    // one-reduction form: the CG SpMV is s = A r (its dot partials are r.s)
    const double *v = MODE == 0 ? (pb.cg_form ? pb.r : p_of_step(pb, step)) : pb.z;
    return noted(pb,
                 launch_k(spmv_sell_kernel<BLOCK, kSellUnr, MODE, false>, d.grid, BLOCK, d.lds_bytes,
                          st, e0, e1, d.wg_slice, d.slice_off, d.slice_row, d.val, d.col, v,
                          d.rows_part, pb.part.pq, pb.n, d.W, d.nrange, (const double *)nullptr,
                          (double *)nullptr, (const double *)nullptr, (const double *)nullptr,
                          0, (const Scalars *)pb.sc, rev),
                 MODE == 1    ? "spmv_sell_kernel (r = A z)"
                 : pb.cg_form ? "spmv_sell_kernel (s = A r)"
                              : "spmv_sell_kernel (q = A p)",
                 BLOCK, st);
}

template <int MODE>
hipError_t launch_spmv(const Fp64Problem &pb, hipStream_t st, hipEvent_t e0 = nullptr,
                       hipEvent_t e1 = nullptr, int step = 0) {
    switch (pb.spmv_variant) {
    case kSpmvTiled:
        return pb.sell.block == 512 ? launch_sell_t<512, MODE>(pb, step, st, e0, e1)
                                    : launch_sell_t<1024, MODE>(pb, step, st, e0, e1);
    case kSpmvSeg16: return launch_csr_t<16, 4, MODE>(pb, st, e0, e1);
    case kSpmvSeg32: return launch_csr_t<32, 4, MODE>(pb, st, e0, e1);
    case kSpmvSeg64U2: return launch_csr_t<64, 2, MODE>(pb, st, e0, e1);
    default: return launch_csr_t<64, 4, MODE>(pb, st, e0, e1);
    }
}

template <int BLOCK, int MODE, bool PUPD>
hipError_t set_sell_lds_t(int bytes) {
    return hipFuncSetAttribute(
        reinterpret_cast<const void *>(&spmv_sell_kernel<BLOCK, kSellUnr, MODE, PUPD>),
        hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

}  // namespace

int vector_blocks(int n, int cus) {
    const int want = std::max(1, std::min((n + 1 + kBlock - 1) / kBlock, 8 * cus));
    return want >= 8 ? (want + 7) & ~7 : want;
}

int spmv_grid(const Fp64Problem &pb) {
    return pb.spmv_variant == kSpmvTiled ? pb.sell.grid : pb.spmv_blocks;
}

int resid_partials(const Fp64Problem &pb) {
    return pb.spmv_variant == kSpmvTiled ? pb.vec_blocks : pb.spmv_blocks;
}

hipError_t prepare_sell_kernels(int lds_bytes) {
    hipError_t e = hipSuccess;
    hipError_t r[6] = {set_sell_lds_t<512, 0, false>(lds_bytes),
                       set_sell_lds_t<1024, 0, false>(lds_bytes),
                       set_sell_lds_t<512, 0, true>(lds_bytes),
                       set_sell_lds_t<1024, 0, true>(lds_bytes),
                       set_sell_lds_t<512, 1, false>(lds_bytes),
                       set_sell_lds_t<1024, 1, false>(lds_bytes)};
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

hipError_t launch_spmv_q(const Fp64Problem &pb, hipStream_t st) {
    Fp64Problem pq = pb;
    pq.cg_form = 0;  // q = A p, whatever form the CG loop uses
    hipError_t e = launch_spmv<0>(pq, st);
    if (e != hipSuccess || pb.spmv_variant != kSpmvTiled) return e;
    combine_rows_kernel<<<pb.vec_blocks, kBlock, 0, st>>>(pb.sell.rows_part, pb.sell.row_slot,
                                                          pb.sell.nblk, pb.q, pb.n);
    return hipGetLastError();
}

hipError_t launch_spmv_kernel(const Fp64Problem &pb, int step, hipEvent_t e0, hipEvent_t e1,
                              hipStream_t st) {
    return launch_spmv<0>(pb, st, e0, e1, step);
}

hipError_t launch_resid(const Fp64Problem &pb, int record, hipStream_t st) {
    hipError_t e = launch_spmv<1>(pb, st);
    if (e != hipSuccess) return e;
    if (pb.spmv_variant == kSpmvTiled) {
        resid_combine_kernel<<<pb.vec_blocks, kBlock, 0, st>>>(
            pb.sell.rows_part, pb.sell.row_slot, pb.sell.nblk, pb.x, pb.z, pb.r, pb.n,
            pb.part.res, pb.sc);
        e = noted(pb, hipGetLastError(), "resid_combine_kernel", kBlock, st);
        if (e != hipSuccess) return e;
    }
    resid_finish_kernel<<<1, kBlock, 0, st>>>(pb.part.res, resid_partials(pb), pb.sc,
                                              pb.zeta_hist, pb.rnorm_hist, pb.hist_cap,
                                              record);
    return noted(pb, hipGetLastError(), "resid_finish_kernel", kBlock, st);
}

hipError_t launch_cg_init(const Fp64Problem &pb, int scale_x, hipStream_t st) {
    cg_init_kernel<<<pb.vec_blocks, kBlock, 0, st>>>(pb.x, pb.z, pb.p, pb.q, pb.r, pb.n,
                                                     scale_x, pb.sc, pb.part.xx, pb.minv);
    hipError_t e = noted(pb, hipGetLastError(), "cg_init_kernel", kBlock, st);
    if (e != hipSuccess || !pb.pop) return e;
    // sparse M^-1: p = z = M^-1 r; r.z into rz[1] (read by step 0), r.r stays in xx
    return launch_precond_apply(*pb.pop, pb.r, pb.pz, pb.p, pb.part.rz[1], pb.vec_blocks,
                                nullptr, st);
}

hipError_t launch_update_zr(const Fp64Problem &pb, int step, hipStream_t st,
                            int sets_tol) {
    // sparse preconditioner: alpha = (r.z of the previous step) / p.q, where
    // the start's r0.z0 sits in rz[1]; the stop level comes from r0.r0 (xx)
    update_zr_kernel<<<pb.vec_blocks, kBlock, 0, st>>>(
        p_of_step(pb, step), pb.q, pb.z, pb.r, pb.n,
        pb.pop ? pb.part.rz[(step + 1) & 1] : rho_source(pb, step), pb.vec_blocks, pb.part.pq,
        spmv_grid(pb), pb.part.rr[step & 1], pb.sell.rows_part, pb.sell.row_slot,
        pb.spmv_variant == kSpmvTiled ? pb.sell.nblk : 0, pb.sc,
        (step == 0 && sets_tol) ? 1 : 0, pb.minv, pb.pop ? pb.part.xx : nullptr);
    return noted(pb, hipGetLastError(), "update_zr_kernel", kBlock, st);
}

// p = r + beta p closing CG step `step`; the result always lands in pb.p.
hipError_t launch_update_p(const Fp64Problem &pb, int step, hipStream_t st) {
    if (pb.pop)  // beta from r.z; stop test on r.r
        update_p_kernel<<<pb.vec_blocks, kBlock, 0, st>>>(
            pb.pz, p_of_step(pb, step), pb.p, pb.n, pb.part.rz[step & 1],
            pb.part.rz[(step + 1) & 1], pb.vec_blocks, pb.sc, nullptr, pb.part.rr[step & 1]);
    else
        update_p_kernel<<<pb.vec_blocks, kBlock, 0, st>>>(
            pb.r, p_of_step(pb, step), pb.p, pb.n, pb.part.rr[step & 1], rho_source(pb, step),
            pb.vec_blocks, pb.sc, pb.minv, nullptr);
    return noted(pb, hipGetLastError(), "update_p_kernel", kBlock, st);
}

hipError_t launch_norms(const Fp64Problem &pb, hipStream_t st) {
    norms_kernel<<<pb.vec_blocks, kBlock, 0, st>>>(pb.x, pb.z, pb.n, pb.part.nrm);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    norms_finish_kernel<<<1, kBlock, 0, st>>>(pb.part.nrm, pb.vec_blocks, pb.sc);
    return hipGetLastError();
}

hipError_t launch_scale_x(const Fp64Problem &pb, double s, int from_scalars,
                          hipStream_t st) {
    scale_x_kernel<<<pb.vec_blocks, kBlock, 0, st>>>(pb.x, pb.z, pb.n, s, from_scalars,
                                                     pb.sc);
    return noted(pb, hipGetLastError(), "scale_x_kernel", kBlock, st);
}

hipError_t launch_reset_scalars(const Fp64Problem &pb, double shift, hipStream_t st,
                                double rtol, int stop) {
    reset_scalars_kernel<<<1, 64, 0, st>>>(pb.sc, shift, rtol * rtol, stop);
    return hipGetLastError();
}

hipError_t launch_clear_done(const Fp64Problem &pb, hipStream_t st, int end_of_solve) {
    clear_done_kernel<<<1, 64, 0, st>>>(pb.sc, end_of_solve);
    return hipGetLastError();
}

hipError_t launch_copy(const void *src, void *dst, long long bytes, int blocks,
                       hipStream_t st) {
    copy_kernel<<<blocks, kBlock, 0, st>>>(static_cast<const double2 *>(src),
                                           static_cast<double2 *>(dst), bytes / 16);
    return hipGetLastError();
}

hipError_t launch_read(const void *src, long long bytes, double *sink, int blocks,
                       hipStream_t st) {
    read_kernel<<<blocks, kReadBlock, 0, st>>>(static_cast<const double2 *>(src), bytes / 16,
                                               sink);
    return hipGetLastError();
}

hipError_t launch_fill(double *v, int count, double value, int blocks, hipStream_t st) {
    fill_kernel<<<blocks, kBlock, 0, st>>>(v, count, value);
    return hipGetLastError();
}

hipError_t enqueue_conj_grad(const Fp64Problem &pb, int cgitmax, int scale_x, int record,
                             hipStream_t st, const SpmvProbe *probe) {
    if (pb.cg_form) {
        // one-reduction form: s = A r, then update_cg, two launches a step;
        // a last p = r + beta p leaves p as the reference's loop does
        hipError_t e = launch_cg_init(pb, scale_x, st);
        for (int it = 0; it < cgitmax && e == hipSuccess; it++) {
            const bool timed = probe && probe->start && it < probe->count;
            e = launch_spmv<0>(pb, st, timed ? probe->start[it] : nullptr,
                               timed ? probe->stop[it] : nullptr, it);
            if (e != hipSuccess) break;
            if (pb.dot_apart) {  // s into p_alt with its r.s partials, then the update
                combine_rs_kernel<<<pb.vec_blocks, kBlock, 0, st>>>(
                    pb.r, pb.sell.rows_part, pb.sell.row_slot, pb.sell.nblk, pb.n, pb.p_alt,
                    pb.part.pq);
                e = noted(pb, hipGetLastError(), "combine_rs_kernel", kBlock, st);
                if (e != hipSuccess) break;
                update_cg_kernel<<<pb.vec_blocks, kBlock, 0, st>>>(
                    pb.p, pb.q, pb.z, pb.r, pb.n, rho_source(pb, it), pb.vec_blocks, pb.part.pq,
                    pb.vec_blocks, pb.part.rr[it & 1], pb.p_alt, nullptr, 1, pb.sc, it);
            } else {
                update_cg_kernel<<<pb.vec_blocks, kBlock, 0, st>>>(
                    pb.p, pb.q, pb.z, pb.r, pb.n, rho_source(pb, it), pb.vec_blocks, pb.part.pq,
                    spmv_grid(pb), pb.part.rr[it & 1], pb.sell.rows_part, pb.sell.row_slot,
                    pb.sell.nblk, pb.sc, it);
            }
            e = noted(pb, hipGetLastError(), "update_cg_kernel", kBlock, st);
        }
        if (e == hipSuccess && cgitmax > 0) e = launch_update_p(pb, cgitmax - 1, st);
        if (e == hipSuccess) e = launch_resid(pb, record, st);
        return e;
    }
    // With the fused SELL path, p = r + beta p of step it-1 is formed inside
    // step it's SpMV, so a step is two launches; one update_p closes the last.
    const bool fused = fused_p(pb);
    hipError_t e = launch_cg_init(pb, scale_x, st);
    for (int it = 0; it < cgitmax && e == hipSuccess; it++) {
        const bool timed = probe && probe->start && it < probe->count;
        e = launch_spmv<0>(pb, st, timed ? probe->start[it] : nullptr,
                           timed ? probe->stop[it] : nullptr, it);
        if (e == hipSuccess) e = launch_update_zr(pb, it, st);
        if (e == hipSuccess && (!fused || it == cgitmax - 1)) e = launch_update_p(pb, it, st);
    }
    if (e == hipSuccess) e = launch_resid(pb, record, st);
    return e;
}

int conj_grad_launches(const Fp64Problem &pb, int cgitmax) {
    const int resid = pb.spmv_variant == kSpmvTiled ? 3 : 2;
    if (pb.cg_form) return 1 + (pb.dot_apart ? 3 : 2) * cgitmax + (cgitmax > 0 ? 1 : 0) + resid;
    if (fused_p(pb)) return 1 + 2 * cgitmax + (cgitmax > 0 ? 1 : 0) + resid;
    return 1 + 3 * cgitmax + resid;
}

hipError_t enqueue_solve_step(const Fp64Problem &pb, int it, int sets_tol, hipStream_t st) {
    hipError_t e = launch_spmv<0>(pb, st, nullptr, nullptr, it);
    if (e == hipSuccess) e = launch_update_zr(pb, it, st, sets_tol);
    if (e == hipSuccess && pb.pop)  // z = M^-1 r and its r.z partials
        e = launch_precond_apply(*pb.pop, pb.r, pb.pz, nullptr, pb.part.rz[it & 1],
                                 pb.vec_blocks, pb.sc, st);
    if (e == hipSuccess) e = launch_update_p(pb, it, st);
    return e;
}

hipError_t enqueue_solve_steps(const Fp64Problem &pb, int steps, int sets_tol,
                               hipStream_t st) {
    hipError_t e = hipSuccess;
    for (int it = 0; it < steps && e == hipSuccess; it++)
        e = enqueue_solve_step(pb, it, sets_tol, st);
    return e;
}

size_t scalars_bytes() { return sizeof(Scalars); }

void KernelMarks::stamp(hipStream_t st) {
    if (err != hipSuccess) return;
    hipEvent_t e = nullptr;
    err = hipEventCreateWithFlags(&e, hipEventDisableSystemFence);
    if (err != hipSuccess) return;
    ev.push_back(e);
    err = hipEventRecord(e, st);
}

void KernelMarks::note(const char *kname, int nthreads, hipStream_t st) {
    stamp(st);
    if (err != hipSuccess) return;
    name.push_back(kname);
    threads.push_back(nthreads);
}

KernelMarks::~KernelMarks() {
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
}

}  // namespace npbcg_dev
