// Batched CG over many independent systems with several right-hand sides
// each (gfx950).
//
// All systems are concatenated block-diagonally: one CSR with global row and
// column indices, vectors stored row-major with the NR right-hand sides of a
// row next to each other ([row][NR]).  Work is cut into tiles of consecutive
// rows that never cross a system boundary, one workgroup per tile, so every
// per-system reduction is a fixed-order sum over that system's tile partials
// and no atomics are needed.  One CG step for the whole batch is three
// launches (SpMM + p.q, x/r update + r.r, p update + stop test); a system
// whose right-hand sides have all converged costs its workgroups one flag
// read.
//
// The SpMM reads each matrix entry once for all NR vectors.  Three layouts:
// row gathers (LPR lanes share a CSR row, every lane gathers the NR values of
// its column as one vector load), column blocks in CSR order, and SELL slices
// over LDS-staged column blocks (one lane per row; the default).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "batch_kernels.h"
#include "cg_device.h"

namespace npbcg_dev {

namespace {

template <typename T, int NR>
struct VecT {
    T v[NR];
};

template <typename T, int NR>
__device__ __forceinline__ VecT<T, NR> load_vec(const T *__restrict__ p) {
    VecT<T, NR> r;
    if constexpr (sizeof(T) * NR == 16) {
        typedef __attribute__((__vector_size__(16))) T v16;
        const v16 x = *reinterpret_cast<const v16 *>(p);
#pragma unroll
        for (int k = 0; k < NR; k++) r.v[k] = x[k];
    } else if constexpr (sizeof(T) * NR == 32) {
        typedef __attribute__((__vector_size__(16))) T v16;
        const v16 x0 = *reinterpret_cast<const v16 *>(p);
        const v16 x1 = *reinterpret_cast<const v16 *>(p + NR / 2);
#pragma unroll
        for (int k = 0; k < NR / 2; k++) {
            r.v[k] = x0[k];
            r.v[NR / 2 + k] = x1[k];
        }
    } else {
#pragma unroll
        for (int k = 0; k < NR; k++) r.v[k] = p[k];
    }
    return r;
}

// LDS copy of one vector row (NR values), as one store where it is 16 bytes.
template <typename T, int NR>
__device__ __forceinline__ void store_vec(T *p, const VecT<T, NR> &v) {
    if constexpr (sizeof(T) * NR == 16) {
        typedef __attribute__((__vector_size__(16))) T v16;
        v16 x;
#pragma unroll
        for (int k = 0; k < NR; k++) x[k] = v.v[k];
        *reinterpret_cast<v16 *>(p) = x;
    } else {
#pragma unroll
        for (int k = 0; k < NR; k++) p[k] = v.v[k];
    }
}

// Block sums of NR per-thread doubles; thread 0 writes out[0..NR).
template <int NR, int BLOCK = kBlock>
__device__ __forceinline__ void block_sum_store(const double (&acc)[NR], double *out) {
    constexpr int kWavesPerBlock = BLOCK / kWave;
    __shared__ double ws[NR][kWavesPerBlock];
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
#pragma unroll
    for (int k = 0; k < NR; k++) {
        const double s = segment_sum<64>(acc[k]);
        if (lane == 0) ws[k][wid] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < NR; k++) {
            double t = ws[k][0];
            for (int w = 1; w < kWavesPerBlock; w++) t += ws[k][w];
            out[k] = t;
        }
    }
}

// out[k] = sum over tiles t in [t0, t1) of part[t * NR + k], in a fixed order
// (every workgroup of the system gets the same bits).  All threads call it.
template <int NR>
__device__ __forceinline__ void system_sum(const double *__restrict__ part, int t0, int t1,
                                           double (&out)[NR]) {
    __shared__ double ws[NR][kWavesPerBlock];
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    double acc[NR];
#pragma unroll
    for (int k = 0; k < NR; k++) acc[k] = 0.0;
    for (int t = t0 + threadIdx.x; t < t1; t += kBlock)
#pragma unroll
        for (int k = 0; k < NR; k++) acc[k] += part[(size_t)t * NR + k];
#pragma unroll
    for (int k = 0; k < NR; k++) {
        const double s = segment_sum<64>(acc[k]);
        if (lane == 0) ws[k][wid] = s;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NR; k++) {
        double t = ws[k][0];
        for (int w = 1; w < kWavesPerBlock; w++) t += ws[k][w];
        out[k] = t;
    }
    __syncthreads();
}

template <int NR>
__device__ __forceinline__ bool system_done(const int *__restrict__ done, int s) {
    bool all = true;
#pragma unroll
    for (int k = 0; k < NR; k++) all = all && done[s * NR + k] != 0;
    return all;
}

// acc[k] += sum over this lane's entries of row `row` of a_e * x[col_e][k]; the
// lane takes entries sub, sub + LPR, ... with four loads in flight.
template <typename T, int NR, int LPR>
__device__ __forceinline__ void row_products(const BatchDev &bd, const T *__restrict__ val,
                                             const T *__restrict__ x, int row, int sub,
                                             T (&acc)[NR]) {
    int e = bd.rowptr[row] + sub;
    const int e1 = bd.rowptr[row + 1];
    for (; e + 3 * LPR < e1; e += 4 * LPR) {
        T a[4];
        int c[4];
#pragma unroll
        for (int u = 0; u < 4; u++) {
            a[u] = val[e + u * LPR];
            c[u] = bd.col[e + u * LPR];
        }
        VecT<T, NR> xv[4];
#pragma unroll
        for (int u = 0; u < 4; u++) xv[u] = load_vec<T, NR>(x + (size_t)c[u] * NR);
#pragma unroll
        for (int u = 0; u < 4; u++)
#pragma unroll
            for (int k = 0; k < NR; k++) acc[k] = fma(a[u], xv[u].v[k], acc[k]);
    }
    for (; e < e1; e += LPR) {
        const T a = val[e];
        const VecT<T, NR> xv = load_vec<T, NR>(x + (size_t)bd.col[e] * NR);
#pragma unroll
        for (int k = 0; k < NR; k++) acc[k] = fma(a, xv.v[k], acc[k]);
    }
}

}  // namespace

// q = A p over one tile (LPR lanes per row), q stored, partials of p.q per RHS.
template <typename T, int NR, int LPR>
__global__ __launch_bounds__(kBlock) void batch_spmm_kernel(BatchDev bd, const T *__restrict__ val,
                                                            const T *__restrict__ p,
                                                            T *__restrict__ q,
                                                            double *__restrict__ pq_part) {
    const int t = bd.block_tile[blockIdx.x];
    if (t < 0) return;
    const int s = bd.tile_sys[t];
    if (system_done<NR>(bd.done, s)) return;
    const int row0 = bd.tile_row[t], row1 = bd.tile_row[t + 1];
    const int grp = threadIdx.x / LPR, sub = threadIdx.x % LPR;
    constexpr int kGroups = kBlock / LPR;
    double dot[NR];
#pragma unroll
    for (int k = 0; k < NR; k++) dot[k] = 0.0;
    for (int row = row0 + grp; row < row1; row += kGroups) {
        T acc[NR];
#pragma unroll
        for (int k = 0; k < NR; k++) acc[k] = T(0);
        row_products<T, NR, LPR>(bd, val, p, row, sub, acc);
#pragma unroll
        for (int m = LPR / 2; m >= 1; m >>= 1)
#pragma unroll
            for (int k = 0; k < NR; k++) acc[k] += __shfl_xor(acc[k], m, kWave);
        if (sub == 0) {
#pragma unroll
            for (int k = 0; k < NR; k++) {
                q[(size_t)row * NR + k] = acc[k];
                dot[k] += (double)p[(size_t)row * NR + k] * (double)acc[k];
            }
        }
    }
    block_sum_store<NR>(dot, pq_part + (size_t)t * NR);
}

// Column-blocked SpMM: item (system s, column block b, rows [row0, row1)).
// The block's vector entries (cbw rows of p, NR values each) are staged in
// LDS, then LPR lanes per row accumulate the row's entries of block b and the
// row partial is stored; p.q partials follow by linearity,
// p.q = sum_b sum_rows p_row (A_b p)_row.
template <typename T, int NR, int LPR>
__global__ __launch_bounds__(kBlock) void batch_spmm_cb_kernel(BatchDev bd,
                                                               const T *__restrict__ val,
                                                               const T *__restrict__ p,
                                                               T *__restrict__ rows_part,
                                                               double *__restrict__ pq_part) {
    extern __shared__ double2 lds_cb[];  // 16-byte aligned
    T *vlds = reinterpret_cast<T *>(lds_cb);
    const int it = blockIdx.x;
    const int s = bd.item[4 * it];
    if (system_done<NR>(bd.done, s)) return;
    const int b = bd.item[4 * it + 1], row0 = bd.item[4 * it + 2], row1 = bd.item[4 * it + 3];
    const int srow = bd.sys_row[s], ns = bd.sys_row[s + 1] - srow;
    const int nb = bd.sys_nb[s];
    const int c0 = b * bd.cbw;
    const int w = min(bd.cbw, ns - c0);
    {
        const T *src = p + (size_t)(srow + c0) * NR;
        for (int i = threadIdx.x; i < w; i += kBlock) {
            const VecT<T, NR> v = load_vec<T, NR>(src + (size_t)i * NR);
#pragma unroll
            for (int k = 0; k < NR; k++) vlds[i * NR + k] = v.v[k];
        }
    }
    __syncthreads();
    const int grp = threadIdx.x / LPR, sub = threadIdx.x % LPR;
    constexpr int kGroups = kBlock / LPR;
    const int *segs = bd.seg + bd.sys_seg[s];
    T *part = rows_part + bd.sys_part[s] + (size_t)b * ns * NR;
    const int col0 = srow + c0;  // global column of the block's first entry
    double dot[NR];
#pragma unroll
    for (int k = 0; k < NR; k++) dot[k] = 0.0;
    for (int row = row0 + grp; row < row1; row += kGroups) {
        const int local = row - srow;
        const int e0 = segs[(size_t)local * (nb + 1) + b], e1 = segs[(size_t)local * (nb + 1) + b + 1];
        T acc[NR];
#pragma unroll
        for (int k = 0; k < NR; k++) acc[k] = T(0);
        int e = e0 + sub;
        for (; e + 3 * LPR < e1; e += 4 * LPR) {
            T a[4];
            int c[4];
#pragma unroll
            for (int u = 0; u < 4; u++) {
                a[u] = val[e + u * LPR];
                c[u] = bd.col[e + u * LPR] - col0;
            }
#pragma unroll
            for (int u = 0; u < 4; u++)
#pragma unroll
                for (int k = 0; k < NR; k++) acc[k] = fma(a[u], vlds[c[u] * NR + k], acc[k]);
        }
        for (; e < e1; e += LPR) {
            const T a = val[e];
            const int c = bd.col[e] - col0;
#pragma unroll
            for (int k = 0; k < NR; k++) acc[k] = fma(a, vlds[c * NR + k], acc[k]);
        }
#pragma unroll
        for (int m = LPR / 2; m >= 1; m >>= 1)
#pragma unroll
            for (int k = 0; k < NR; k++) acc[k] += __shfl_xor(acc[k], m, kWave);
        if (sub == 0) {
#pragma unroll
            for (int k = 0; k < NR; k++) {
                part[(size_t)local * NR + k] = acc[k];
                dot[k] += (double)p[(size_t)row * NR + k] * (double)acc[k];
            }
        }
    }
    block_sum_store<NR>(dot, pq_part + (size_t)it * NR);
}

// SELL SpMM over one 2048-row sorting window per item (system s, window
// rows [r0, r0 + nrows)).  For each column block of the system in turn, the
// block's vector entries (cbw rows of p, NR values each) are staged in LDS
// and each wavefront streams whole slices of the window's block, one row per
// lane, so the (value, 16-bit column) loads are contiguous across the
// wavefront and a row needs no cross-lane reduction.  The rows are sorted
// by segment length per block, so a lane's row changes from block to block:
// the row's block sum is added to the window's row sums, kept in LDS beside
// the vector block (the first block stores, later blocks add, in block
// order).  After the last block the window's rows are complete: q = A p is
// written once, with the p.q partial of the item.  No block partials go
// through memory.  Slices go to wavefronts in a fixed snake order, so the
// sums are reproducible.
//
// RESID = true computes the FP64 true residual instead: res = b - A x for the
// window's rows and the partials of |res|^2.
//
// The entry stream is software-pipelined with sell_chunk: the loads of the
// next chunk (or of the next slice's first chunk and its lane rows) are issued
// before the current chunk's LDS gathers and FMAs, and a block's first chunk
// is issued before its LDS fill, whose loads all go out before its stores.
//
// One workgroup per CU walks its own list of items (block_item: item_blocks
// + 1 list offsets, then the lists), with a barrier after each block.
// rev = 1 walks the list backwards: successive SpMMs alternate, so each starts
// on the items the last one streamed last, which the memory-side cache still
// holds.
template <typename T, int NR, int BLOCK, int UNR, bool RESID>
__global__ __launch_bounds__(BLOCK) void batch_spmm_sell_kernel(BatchDev bd,
                                                                const T *__restrict__ sval,
                                                                const T *__restrict__ p,
                                                                const double *__restrict__ rhs,
                                                                T *__restrict__ out,
                                                                double *__restrict__ part,
                                                                int rev) {
    extern __shared__ double2 lds_sell[];  // 16-byte aligned
    T *vlds = reinterpret_cast<T *>(lds_sell);
    T *acc_lds = vlds + (size_t)bd.cbw * NR;  // [kSellWindow][NR] row sums of the window
    constexpr int kWaves = BLOCK / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    // wave-uniform (scalar) index: every branch on the slice loop is scalar
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int *__restrict__ off = bd.sell_off;
    const uint16_t *__restrict__ col = bd.sell_col;
    const int *__restrict__ lane_row = bd.sell_row;
    const int *__restrict__ wl = bd.block_item + bd.item_blocks + 1;
    const int k0 = bd.block_item[blockIdx.x], k1 = bd.block_item[blockIdx.x + 1];
    for (int li = k0; li < k1; li++) {
        const int it = wl[rev ? k0 + k1 - 1 - li : li];
        const int s = bd.item[4 * it];
        if (!RESID && system_done<NR>(bd.done, s)) continue;  // the same for every thread
        const int r0 = bd.item[4 * it + 1], nrows = bd.item[4 * it + 2];
        const int *__restrict__ ranges = bd.item_sl + bd.item[4 * it + 3];
        const int srow = bd.sys_row[s], ns = bd.sys_row[s + 1] - srow, nb = bd.sys_nb[s];
        const int wrow = srow + r0;  // first global row of the window
        for (int b = 0; b < nb; b++) {
            const int sl0 = ranges[2 * b], nsl = ranges[2 * b + 1] - sl0;
            const int c0 = b * bd.cbw;
            const int w = min(bd.cbw, ns - c0);
            // Slice of round rd for this wave (snake order: slices of a window
            // are sorted widest first, and alternating the direction of each
            // round evens out the width per wavefront), or -1.  Rounds are
            // increasing in k, so after the first -1 there is none.
            auto slice_of = [&](int rd) {
                const int k = rd * kWaves + ((rd & 1) ? kWaves - 1 - wave : wave);
                return k < nsl ? sl0 + k : -1;
            };
            T na[UNR];
            int nc[UNR];
            int cur = slice_of(0), cbase = 0, cwidth = 0, nrow = -1;
            if (cur >= 0) {
                cbase = off[cur];
                cwidth = (off[cur + 1] - cbase) / kWave;
                nrow = lane_row[(size_t)cur * kWave + lane];
                if (cwidth > 0) sell_chunk<T, UNR>(sval, col, cbase, cwidth, 0, lane, na, nc);
            }
            if constexpr (NR * sizeof(T) % 16 == 0) {  // rows of whole 16-byte units, aligned
                constexpr int kUnits = NR * (int)sizeof(T) / 16;
                stage_block<BLOCK, (128 * 1024 / 16 + BLOCK - 1) / BLOCK>(
                    reinterpret_cast<const double2 *>(p + (size_t)(srow + c0) * NR), lds_sell,
                    w * kUnits);
            } else {
                stage_block<BLOCK, (128 * 1024 / (int)sizeof(T) + BLOCK - 1) / BLOCK>(
                    p + (size_t)(srow + c0) * NR, vlds, w * NR);
            }
            __syncthreads();
            for (int rd = 0; cur >= 0; rd++) {
                const int width = cwidth, base = cbase, row = nrow;
                const int nxt = slice_of(rd + 1);
                int xbase = 0, xwidth = 0;
                if (nxt >= 0) {
                    xbase = off[nxt];
                    xwidth = (off[nxt + 1] - xbase) / kWave;
                }
                T acc[NR];
#pragma unroll
                for (int q = 0; q < NR; q++) acc[q] = T(0);
                const int nch = (width + UNR - 1) / UNR;
                for (int ch = 0; ch < nch; ch++) {
                    T a[UNR];
                    int c[UNR];
#pragma unroll
                    for (int u = 0; u < UNR; u++) {
                        a[u] = na[u];
                        c[u] = nc[u];
                    }
                    sell_trim<T, UNR>(width - ch * UNR, a, c);
                    if (ch + 1 < nch) {
                        sell_chunk<T, UNR>(sval, col, base, width, (ch + 1) * UNR, lane, na, nc);
                    } else if (nxt >= 0) {
                        nrow = lane_row[(size_t)nxt * kWave + lane];
                        if (xwidth > 0) sell_chunk<T, UNR>(sval, col, xbase, xwidth, 0, lane, na, nc);
                    }
#pragma unroll
                    for (int u = 0; u < UNR; u++) {
                        const VecT<T, NR> xv = load_vec<T, NR>(vlds + c[u] * NR);
#pragma unroll
                        for (int q = 0; q < NR; q++) acc[q] = fma(a[u], xv.v[q], acc[q]);
                    }
                }
                if (nch == 0 && nxt >= 0) {
                    nrow = lane_row[(size_t)nxt * kWave + lane];
                    if (xwidth > 0) sell_chunk<T, UNR>(sval, col, xbase, xwidth, 0, lane, na, nc);
                }
                if (row >= 0) {  // every row of the window is in one slice of each block
                    T *dst = acc_lds + (size_t)(row - wrow) * NR;
                    VecT<T, NR> sum;
                    if (b == 0) {
#pragma unroll
                        for (int q = 0; q < NR; q++) sum.v[q] = acc[q];
                    } else {
                        sum = load_vec<T, NR>(dst);
#pragma unroll
                        for (int q = 0; q < NR; q++) sum.v[q] += acc[q];
                    }
                    store_vec<T, NR>(dst, sum);
                }
                cur = nxt;
                cbase = xbase;
                cwidth = xwidth;
            }
            __syncthreads();  // the vector block is free again; the row sums are visible
        }
        double dot[NR];
#pragma unroll
        for (int q = 0; q < NR; q++) dot[q] = 0.0;
        for (int i = threadIdx.x; i < nrows; i += BLOCK) {
            const VecT<T, NR> sum = load_vec<T, NR>(acc_lds + (size_t)i * NR);
            const size_t at = (size_t)(wrow + i) * NR;
            if constexpr (RESID) {
#pragma unroll
                for (int q = 0; q < NR; q++) {
                    const double rv = rhs[at + q] - (double)sum.v[q];
                    out[at + q] = (T)rv;
                    dot[q] += rv * rv;
                }
            } else {
                const VecT<T, NR> pv = load_vec<T, NR>(p + at);
                store_vec<T, NR>(out + at, sum);
#pragma unroll
                for (int q = 0; q < NR; q++) dot[q] += (double)pv.v[q] * (double)sum.v[q];
            }
        }
        block_sum_store<NR, BLOCK>(dot, part + (size_t)it * NR);
        __syncthreads();  // the row sums and the reduction slots are free again
    }
}

// Residual of x = 0: res = b; |b|^2 partials per tile.
template <int NR>
__global__ __launch_bounds__(kBlock) void batch_resid_zero_kernel(BatchDev bd,
                                                                  const double *__restrict__ b,
                                                                  double *__restrict__ res,
                                                                  double *__restrict__ part) {
    const int t = blockIdx.x;
    const int row0 = bd.tile_row[t], row1 = bd.tile_row[t + 1];
    double nrm[NR];
#pragma unroll
    for (int k = 0; k < NR; k++) nrm[k] = 0.0;
    for (int row = row0 + threadIdx.x; row < row1; row += kBlock)
#pragma unroll
        for (int k = 0; k < NR; k++) {
            const size_t at = (size_t)row * NR + k;
            const double v = b[at];
            res[at] = v;
            nrm[k] += v * v;
        }
    block_sum_store<NR>(nrm, part + (size_t)t * NR);
}

// alpha = rho / p.q per RHS of the tile's system; d += alpha p, r -= alpha q;
// partials of the new r.r.  The first step of a solve sets tol2 = rtol2 * rho.
template <typename T, int NR>
__global__ __launch_bounds__(kBlock) void batch_update_kernel(BatchDev bd, const T *__restrict__ p,
                                                              const T *__restrict__ q,
                                                              const T *__restrict__ rows_part,
                                                              T *__restrict__ d, T *__restrict__ r,
                                                              const double *__restrict__ rho_part,
                                                              const double *__restrict__ pq_part,
                                                              double *__restrict__ rr_part,
                                                              int first_step) {
    const int t = blockIdx.x;
    const int s = bd.tile_sys[t];
    if (system_done<NR>(bd.done, s)) return;
    const int t0 = bd.sys_tile[s], t1 = bd.sys_tile[s + 1];
    double rho[NR], pq[NR];
    system_sum<NR>(rho_part, t0, t1, rho);
    if (bd.cb)
        system_sum<NR>(pq_part, bd.sys_item[s], bd.sys_item[s + 1], pq);
    else
        system_sum<NR>(pq_part, t0, t1, pq);
    T alpha[NR];
    bool act[NR];
#pragma unroll
    for (int k = 0; k < NR; k++) {
        act[k] = bd.done[s * NR + k] == 0;
        alpha[k] = act[k] && pq[k] != 0.0 ? (T)(rho[k] / pq[k]) : T(0);
    }
    if (t == t0)
#pragma unroll
        for (int k = 0; k < NR; k++)
            if (threadIdx.x == k) {
                BatchScalars &sc = bd.sc[s * NR + k];
                if (first_step) sc.tol2 = sc.rtol2 * fabs(rho[k]);
                sc.rho = rho[k];
                if (act[k]) {
                    sc.steps += 1;
                    sc.sweep_steps += 1;
                }
            }
    const int row0 = bd.tile_row[t], row1 = bd.tile_row[t + 1];
    double acc[NR];
#pragma unroll
    for (int k = 0; k < NR; k++) acc[k] = 0.0;
    // column blocks in CSR order: q = sum of the blocks' row partials; the
    // SELL SpMM writes q itself
    const bool combine = bd.cb && !bd.sell;
    const int srow = combine ? bd.sys_row[s] : 0;
    const int ns = combine ? bd.sys_row[s + 1] - srow : 0;
    const int nb = combine ? bd.sys_nb[s] : 0;
    const T *part = combine ? rows_part + bd.sys_part[s] : nullptr;
    for (int row = row0 + threadIdx.x; row < row1; row += kBlock) {
        T qv[NR];
        if (combine) {  // blocks in order
            const size_t local = (size_t)(row - srow) * NR;
#pragma unroll
            for (int k = 0; k < NR; k++) qv[k] = part[local + k];
#pragma unroll 8
            for (int b = 1; b < nb; b++)
#pragma unroll
                for (int k = 0; k < NR; k++) qv[k] += part[(size_t)b * ns * NR + local + k];
        } else {
#pragma unroll
            for (int k = 0; k < NR; k++) qv[k] = q[(size_t)row * NR + k];
        }
#pragma unroll
        for (int k = 0; k < NR; k++) {
            const size_t at = (size_t)row * NR + k;
            d[at] = fma(alpha[k], p[at], d[at]);
            const T rv = fma(-alpha[k], qv[k], r[at]);
            r[at] = rv;
            acc[k] += (double)rv * (double)rv;
        }
    }
    block_sum_store<NR>(acc, rr_part + (size_t)t * NR);
}

// beta = r.r / rho per RHS; stop test r.r <= tol2 (the tile's system's first
// tile sets the flag); p = r + beta p for the right-hand sides still running.
// The flag is written during the launch: another tile of the system that
// reads it already set takes the right-hand side as inactive, which skips the
// same p update the stop test skips, so every tile does the same work either
// way and the next step's kernels see the flag.
template <typename T, int NR>
__global__ __launch_bounds__(kBlock) void batch_pupdate_kernel(BatchDev bd,
                                                               const T *__restrict__ r,
                                                               T *__restrict__ p,
                                                               const double *__restrict__ rr_part,
                                                               const double *__restrict__ rho_part) {
    const int t = blockIdx.x;
    const int s = bd.tile_sys[t];
    if (system_done<NR>(bd.done, s)) return;
    const int t0 = bd.sys_tile[s], t1 = bd.sys_tile[s + 1];
    double rr[NR], rho[NR];
    system_sum<NR>(rr_part, t0, t1, rr);
    system_sum<NR>(rho_part, t0, t1, rho);
    T beta[NR];
    bool run[NR];
#pragma unroll
    for (int k = 0; k < NR; k++) {
        const BatchScalars &sc = bd.sc[s * NR + k];
        const bool act = bd.done[s * NR + k] == 0;
        const bool stop = act && (fabs(rr[k]) <= sc.tol2 || sc.sweep_steps >= bd.maxit);
        run[k] = act && !stop;
        beta[k] = run[k] && rho[k] != 0.0 ? (T)(rr[k] / rho[k]) : T(0);
    }
    if (t == t0)
#pragma unroll
        for (int k = 0; k < NR; k++)
            if (threadIdx.x == k && bd.done[s * NR + k] == 0 && !run[k]) bd.done[s * NR + k] = 1;
    const int row0 = bd.tile_row[t], row1 = bd.tile_row[t + 1];
    for (int row = row0 + threadIdx.x; row < row1; row += kBlock)
#pragma unroll
        for (int k = 0; k < NR; k++)
            if (run[k]) {
                const size_t at = (size_t)row * NR + k;
                p[at] = fma(beta[k], p[at], r[at]);
            }
}

// Start of an inner solve: r = p = (T)(b - A x) (given in res), d = 0,
// partials of r.r.  Right-hand sides already done keep r = p = 0.
template <typename T, int NR>
__global__ __launch_bounds__(kBlock) void batch_init_kernel(BatchDev bd,
                                                            const double *__restrict__ res,
                                                            T *__restrict__ d, T *__restrict__ r,
                                                            T *__restrict__ p,
                                                            double *__restrict__ rr_part) {
    const int t = blockIdx.x;
    const int s = bd.tile_sys[t];
    const int row0 = bd.tile_row[t], row1 = bd.tile_row[t + 1];
    bool act[NR];
#pragma unroll
    for (int k = 0; k < NR; k++) act[k] = bd.done[s * NR + k] == 0;
    double acc[NR];
#pragma unroll
    for (int k = 0; k < NR; k++) acc[k] = 0.0;
    for (int row = row0 + threadIdx.x; row < row1; row += kBlock)
#pragma unroll
        for (int k = 0; k < NR; k++) {
            const size_t at = (size_t)row * NR + k;
            const T v = act[k] ? (T)res[at] : T(0);
            d[at] = T(0);
            r[at] = v;
            p[at] = v;
            acc[k] += (double)v * (double)v;
        }
    block_sum_store<NR>(acc, rr_part + (size_t)t * NR);
}

// res = b - A x in FP64 (LPR lanes per row), partials of |res|^2.
template <int NR, int LPR>
__global__ __launch_bounds__(kBlock) void batch_residual_kernel(BatchDev bd,
                                                                const double *__restrict__ val,
                                                                const double *__restrict__ b,
                                                                const double *__restrict__ x,
                                                                double *__restrict__ res,
                                                                double *__restrict__ part) {
    const int t = bd.block_tile[blockIdx.x];
    if (t < 0) return;
    const int row0 = bd.tile_row[t], row1 = bd.tile_row[t + 1];
    const int grp = threadIdx.x / LPR, sub = threadIdx.x % LPR;
    constexpr int kGroups = kBlock / LPR;
    double nrm[NR];
#pragma unroll
    for (int k = 0; k < NR; k++) nrm[k] = 0.0;
    for (int row = row0 + grp; row < row1; row += kGroups) {
        double acc[NR];
#pragma unroll
        for (int k = 0; k < NR; k++) acc[k] = 0.0;
        row_products<double, NR, LPR>(bd, val, x, row, sub, acc);
#pragma unroll
        for (int m = LPR / 2; m >= 1; m >>= 1)
#pragma unroll
            for (int k = 0; k < NR; k++) acc[k] += __shfl_xor(acc[k], m, kWave);
        if (sub == 0)
#pragma unroll
            for (int k = 0; k < NR; k++) {
                const size_t at = (size_t)row * NR + k;
                const double rv = b[at] - acc[k];
                res[at] = rv;
                nrm[k] += rv * rv;
            }
    }
    block_sum_store<NR>(nrm, part + (size_t)t * NR);
}

// One workgroup per system: out[s*NR + k] = sum of its tiles' partials.
template <int NR>
__global__ __launch_bounds__(kBlock) void batch_system_sum_kernel(BatchDev bd,
                                                                  const double *__restrict__ part,
                                                                  double *__restrict__ out) {
    const int s = blockIdx.x;
    double v[NR];
    system_sum<NR>(part, bd.sys_tile[s], bd.sys_tile[s + 1], v);
#pragma unroll
    for (int k = 0; k < NR; k++)
        if (threadIdx.x == k) out[s * NR + k] = v[k];
}

// x += (double) d.
template <typename T>
__global__ __launch_bounds__(kBlock) void batch_refine_kernel(double *__restrict__ x,
                                                              const T *__restrict__ d,
                                                              long long count) {
    for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < count;
         i += (long long)gridDim.x * kBlock)
        x[i] += (double)d[i];
}

__global__ __launch_bounds__(kBlock) void batch_narrow_kernel(const double *__restrict__ in,
                                                              float *__restrict__ out,
                                                              long long count) {
    for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < count;
         i += (long long)gridDim.x * kBlock)
        out[i] = (float)in[i];
}

// ---------------------------------------------------------------------------
// Host launchers (templates dispatched on NR in {1, 2, 4} and LPR in {4..64})
// ---------------------------------------------------------------------------

namespace {

int grid_for(long long items) {
    return (int)std::max<long long>(1, std::min<long long>((items + kBlock - 1) / kBlock, 8192));
}

template <typename F>
hipError_t with_lpr(int lpr, F &&f) {
    switch (lpr) {
        case 4: return f(std::integral_constant<int, 4>{});
        case 8: return f(std::integral_constant<int, 8>{});
        case 16: return f(std::integral_constant<int, 16>{});
        case 32: return f(std::integral_constant<int, 32>{});
        case 64: return f(std::integral_constant<int, 64>{});
        default: return hipErrorInvalidValue;
    }
}

template <typename F>
hipError_t with_nr(int nr, F &&f) {
    switch (nr) {
        case 1: return f(std::integral_constant<int, 1>{});
        case 2: return f(std::integral_constant<int, 2>{});
        case 4: return f(std::integral_constant<int, 4>{});
        default: return hipErrorInvalidValue;
    }
}

// Entries per pipelined chunk of the SELL SpMM: 8, except FP64 with four
// right-hand sides (1024-thread workgroups leave 128 registers a lane).
template <typename T, int NR>
constexpr int sell_unr() {
    return sizeof(T) == 8 && NR == 4 ? 4 : 8;
}

// q = A p (row sums, or block row partials for cb) and p.q partials; rev
// selects the SELL SpMM's item order (alternated by the callers).
template <typename T, int NR, int LPR>
void launch_spmm(const BatchDev &bd, const BatchVecs<T> &v, int rev, hipStream_t st) {
    if (bd.sell) {
        batch_spmm_sell_kernel<T, NR, kSellBlock, sell_unr<T, NR>(), false>
            <<<bd.item_blocks, kSellBlock, sell_lds_bytes(bd.cbw, NR, sizeof(T)), st>>>(
                bd, v.sval, v.p, nullptr, v.q, v.part_pq, rev);
    } else if (bd.cb) {
        batch_spmm_cb_kernel<T, NR, LPR><<<bd.items, kBlock, (size_t)bd.cbw * NR * sizeof(T), st>>>(
            bd, v.val, v.p, v.rows_part, v.part_pq);
    } else {
        batch_spmm_kernel<T, NR, LPR><<<bd.blocks, kBlock, 0, st>>>(bd, v.val, v.p, v.q, v.part_pq);
    }
}

}  // namespace

template <typename T>
hipError_t batch_cg_step(const BatchDev &bd, const BatchVecs<T> &v, int step, int first,
                         hipStream_t st) {
    const int tiles = bd.tiles;
    const double *rho = step == 0 ? v.part_init : v.part_rr[(step - 1) & 1];
    double *rr = v.part_rr[step & 1];
    hipError_t e = with_nr(bd.nr, [&](auto nrc) {
        constexpr int NR = decltype(nrc)::value;
        return with_lpr(bd.lpr, [&](auto lc) {
            constexpr int LPR = decltype(lc)::value;
            launch_spmm<T, NR, LPR>(bd, v, step & 1, st);
            return hipGetLastError();
        });
    });
    if (e != hipSuccess) return e;
    e = with_nr(bd.nr, [&](auto nrc) {
        constexpr int NR = decltype(nrc)::value;
        batch_update_kernel<T, NR><<<tiles, kBlock, 0, st>>>(bd, v.p, v.q, v.rows_part, v.d, v.r,
                                                             rho, v.part_pq, rr, first);
        hipError_t e2 = hipGetLastError();
        if (e2 != hipSuccess) return e2;
        batch_pupdate_kernel<T, NR><<<tiles, kBlock, 0, st>>>(bd, v.r, v.p, rr, rho);
        return hipGetLastError();
    });
    return e;
}

template <typename T>
hipError_t batch_spmm(const BatchDev &bd, const BatchVecs<T> &v, int rev, hipStream_t st) {
    return with_nr(bd.nr, [&](auto nrc) {
        constexpr int NR = decltype(nrc)::value;
        return with_lpr(bd.lpr, [&](auto lc) {
            constexpr int LPR = decltype(lc)::value;
            launch_spmm<T, NR, LPR>(bd, v, rev, st);
            return hipGetLastError();
        });
    });
}

// Raises the dynamic LDS limit of the column-blocked SpMM instances.
template <typename T>
hipError_t batch_prepare_cb(int bytes) {
    hipError_t last = hipSuccess;
    auto set = [&](const void *f) {
        hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        if (e != hipSuccess) last = e;
    };
#define NPBCG_CB_SET(NR)                                                          \
    set(reinterpret_cast<const void *>(&batch_spmm_cb_kernel<T, NR, 4>));        \
    set(reinterpret_cast<const void *>(&batch_spmm_cb_kernel<T, NR, 8>));        \
    set(reinterpret_cast<const void *>(&batch_spmm_cb_kernel<T, NR, 16>));       \
    set(reinterpret_cast<const void *>(&batch_spmm_cb_kernel<T, NR, 32>));       \
    set(reinterpret_cast<const void *>(&batch_spmm_cb_kernel<T, NR, 64>));        \
    set(reinterpret_cast<const void *>(                                          \
        &batch_spmm_sell_kernel<T, NR, kSellBlock, sell_unr<T, NR>(), false>));
    NPBCG_CB_SET(1)
    NPBCG_CB_SET(2)
    NPBCG_CB_SET(4)
#undef NPBCG_CB_SET
    return last;
}

hipError_t batch_prepare_resid(int bytes) {
    hipError_t last = hipSuccess;
    auto set = [&](const void *f) {
        hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
        if (e != hipSuccess) last = e;
    };
    set(reinterpret_cast<const void *>(
        &batch_spmm_sell_kernel<double, 1, kSellBlock, sell_unr<double, 1>(), true>));
    set(reinterpret_cast<const void *>(
        &batch_spmm_sell_kernel<double, 2, kSellBlock, sell_unr<double, 2>(), true>));
    set(reinterpret_cast<const void *>(
        &batch_spmm_sell_kernel<double, 4, kSellBlock, sell_unr<double, 4>(), true>));
    return last;
}

template <typename T>
hipError_t batch_init(const BatchDev &bd, const BatchVecs<T> &v, const double *res,
                      hipStream_t st) {
    return with_nr(bd.nr, [&](auto nrc) {
        constexpr int NR = decltype(nrc)::value;
        batch_init_kernel<T, NR><<<bd.tiles, kBlock, 0, st>>>(bd, res, v.d, v.r, v.p, v.part_init);
        return hipGetLastError();
    });
}

template <typename T>
hipError_t batch_refine(const BatchDev &bd, double *x, const T *d, hipStream_t st) {
    const long long count = (long long)bd.rows * bd.nr;
    batch_refine_kernel<T><<<grid_for(count), kBlock, 0, st>>>(x, d, count);
    return hipGetLastError();
}

hipError_t batch_residual(const BatchDev &bd, const double *val, const double *b, const double *x,
                          double *res, double *part, double *sys_out, hipStream_t st) {
    hipError_t e = with_nr(bd.nr, [&](auto nrc) {
        constexpr int NR = decltype(nrc)::value;
        return with_lpr(bd.lpr, [&](auto lc) {
            constexpr int LPR = decltype(lc)::value;
            batch_residual_kernel<NR, LPR><<<bd.blocks, kBlock, 0, st>>>(bd, val, b, x, res, part);
            return hipGetLastError();
        });
    });
    if (e != hipSuccess) return e;
    return with_nr(bd.nr, [&](auto nrc) {
        constexpr int NR = decltype(nrc)::value;
        batch_system_sum_kernel<NR><<<bd.nsys, kBlock, 0, st>>>(bd, part, sys_out);
        return hipGetLastError();
    });
}

hipError_t batch_residual_sell(const BatchDev &bd, const double *sval, const double *b,
                               const double *x, double *res, double *part, double *sys_out,
                               hipStream_t st) {
    // |res|^2 partials are per item: the system sums run over sys_item
    BatchDev by_item = bd;
    by_item.sys_tile = bd.sys_item;
    return with_nr(bd.nr, [&](auto nrc) {
        constexpr int NR = decltype(nrc)::value;
        batch_spmm_sell_kernel<double, NR, kSellBlock, sell_unr<double, NR>(), true>
            <<<bd.item_blocks, kSellBlock, sell_lds_bytes(bd.cbw, NR, sizeof(double)), st>>>(
                bd, sval, x, b, res, part, 0);
        hipError_t e2 = hipGetLastError();
        if (e2 != hipSuccess) return e2;
        batch_system_sum_kernel<NR><<<bd.nsys, kBlock, 0, st>>>(by_item, part, sys_out);
        return hipGetLastError();
    });
}

hipError_t batch_residual_zero(const BatchDev &bd, const double *b, double *res, double *part,
                               double *sys_out, hipStream_t st) {
    return with_nr(bd.nr, [&](auto nrc) {
        constexpr int NR = decltype(nrc)::value;
        batch_resid_zero_kernel<NR><<<bd.tiles, kBlock, 0, st>>>(bd, b, res, part);
        hipError_t e2 = hipGetLastError();
        if (e2 != hipSuccess) return e2;
        batch_system_sum_kernel<NR><<<bd.nsys, kBlock, 0, st>>>(bd, part, sys_out);
        return hipGetLastError();
    });
}

hipError_t batch_narrow(const double *in, float *out, long long count, hipStream_t st) {
    batch_narrow_kernel<<<grid_for(count), kBlock, 0, st>>>(in, out, count);
    return hipGetLastError();
}

template hipError_t batch_cg_step<float>(const BatchDev &, const BatchVecs<float> &, int, int,
                                         hipStream_t);
template hipError_t batch_cg_step<double>(const BatchDev &, const BatchVecs<double> &, int, int,
                                          hipStream_t);
template hipError_t batch_prepare_cb<float>(int);
template hipError_t batch_prepare_cb<double>(int);
template hipError_t batch_spmm<float>(const BatchDev &, const BatchVecs<float> &, int, hipStream_t);
template hipError_t batch_spmm<double>(const BatchDev &, const BatchVecs<double> &, int,
                                       hipStream_t);
template hipError_t batch_init<float>(const BatchDev &, const BatchVecs<float> &, const double *,
                                      hipStream_t);
template hipError_t batch_init<double>(const BatchDev &, const BatchVecs<double> &,
                                       const double *, hipStream_t);
template hipError_t batch_refine<float>(const BatchDev &, double *, const float *, hipStream_t);
template hipError_t batch_refine<double>(const BatchDev &, double *, const double *, hipStream_t);

}  // namespace npbcg_dev
