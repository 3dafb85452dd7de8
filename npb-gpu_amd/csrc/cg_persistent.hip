// Persistent FP64 conj_grad engine for gfx950: one cooperative launch runs
// whole outer iterations of the inverse-power loop (cg.cu:500-520), each a
// conj_grad of cgitmax CG steps (cg.cu:803-846), with grid-wide barriers
// between its phases instead of kernel boundaries.
//
// The CG step is the one-reduction form of the multi-kernel engine
// (cg_fp64.hip, update_cg_kernel): s = A r with r.s partials, then one update
// forming beta, p = r + beta p, q = s + beta q (= A p), delta = p.q =
// r.s - beta rho / alpha_prev, alpha = rho / delta, z += alpha p, r -= alpha q
// and the new r.r partials.  Phases of one outer iteration (G workgroups, one
// per CU on the column-blocked SELL layout, all resident: cooperative launch):
//   SpMV(k)    stage the k-th multiplied vector's column block in LDS, stream
//              the workgroup's slices, row sums to their slots (64 per slice,
//              one coalesced run), r.s partial; at k = 0 the vector is x, or xscale * z (the previous
//              iteration's normalisation, cg.cu:479-481) which the row-range-0
//              workgroups also store into x with its x.x partial
//   barrier
//   update(k)  every workgroup reduces the same partials in the same order
//              (same bits everywhere), then updates its grid-stride rows
//   barrier    ... k = 0 .. cgitmax-1
//   resid      p = r + beta p (the reference leaves p so); r = A z streamed
//   barrier
//   combine    r = sum of the blocks' row sums; partials of ||x - r||^2, x.z, z.z
//   barrier
//   finish     rnorm, zeta = shift + 1/(x.z), xscale = 1/sqrt(z.z) (cg.cu:505-514),
//              reduced by every workgroup; workgroup 0 records them.
// The per-step scalars live in registers of every workgroup; nothing returns
// to the host inside the launch.
//
// Barrier: one thread per workgroup releases its writes at device scope,
// adds 1 to an arrival counter and polls it until every workgroup of the
// barrier has arrived, then acquires.  Every wait is bounded: a barrier that
// waits longer than spin_limit wall-clock ticks sets an abort word, after
// which every barrier returns at once, so the grid always drains; the host
// reports the abort as an error.
#include <hip/hip_runtime.h>

#include "cg_device.h"
#include "cg_kernels.h"
#include "sell_walk.h"

namespace npbcg_dev {

namespace {

constexpr int kPersistUnr = 12;  // entries per lane in flight (as the standalone SpMV)

struct PersistArgs {
    const int *wg_slice, *slice_off, *slice_row;
    const double *val;
    const uint16_t *col;
    const int *row_slot;  // nblk * n: slot of row j in column block b
    double *rows_part;    // block row sums by slot
    int n, W, nrange, nblk;
    double *x, *z, *p, *q, *r;
    double *part_rs;     // [G] r.s of the SpMV phase
    double *part_xx;     // [G] x.x of step 0
    double *part_rr[2];  // [G] r.r of the update phases, alternating by step
    double *part_res;    // [3 G] ||x - r||^2, x.z, z.z of the combine phase
    Scalars *sc;
    double *zeta_hist, *rnorm_hist;
    int hist_cap;
    unsigned *bar;       // [0] arrivals, [1] abort word
    long long spin_limit;
    int cgitmax, niter, scale_first, record, final_scale;
};


// p, with the compiler told that it points to global memory (a pointer read
// from memory is generic, and generic accesses compile to flat loads and
// stores; the address-space round trip lets them be global ones).
#define NPBCG_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ T *global_ptr(T *p) {
    return (T *)(T NPBCG_GLOBAL *)p;
}

// The argument block, read at the start of a phase through an empty asm
// statement on its address: the fields are loaded where the phase uses them
// instead of being held in registers across the whole kernel.  The block is
// read through the constant address space (it is written before the launch
// and never during it), so the loads are scalar and the values uniform.
__device__ __forceinline__ PersistArgs phase_args(const PersistArgs *args) {
    asm volatile("" : "+s"(args));
    typedef const PersistArgs __attribute__((address_space(4))) ConstArgs;
    ConstArgs *c = (ConstArgs *)args;
    PersistArgs a;
    a.wg_slice = global_ptr(c->wg_slice);
    a.slice_off = global_ptr(c->slice_off);
    a.slice_row = global_ptr(c->slice_row);
    a.val = global_ptr(c->val);
    a.col = global_ptr(c->col);
    a.row_slot = global_ptr(c->row_slot);
    a.rows_part = global_ptr(c->rows_part);
    a.n = c->n;
    a.W = c->W;
    a.nrange = c->nrange;
    a.nblk = c->nblk;
    a.x = global_ptr(c->x);
    a.z = global_ptr(c->z);
    a.p = global_ptr(c->p);
    a.q = global_ptr(c->q);
    a.r = global_ptr(c->r);
    a.part_rs = global_ptr(c->part_rs);
    a.part_xx = global_ptr(c->part_xx);
    a.part_rr[0] = global_ptr(c->part_rr[0]);
    a.part_rr[1] = global_ptr(c->part_rr[1]);
    a.part_res = global_ptr(c->part_res);
    a.sc = global_ptr(c->sc);
    a.zeta_hist = global_ptr(c->zeta_hist);
    a.rnorm_hist = global_ptr(c->rnorm_hist);
    a.hist_cap = c->hist_cap;
    a.bar = global_ptr(c->bar);
    a.spin_limit = c->spin_limit;
    a.cgitmax = c->cgitmax;
    a.niter = c->niter;
    a.scale_first = c->scale_first;
    a.record = c->record;
    a.final_scale = c->final_scale;
    return a;
}

// threadIdx.x through an empty asm statement, taken at the start of a phase:
// the compiler cannot hoist the per-thread addresses derived from it out of
// the phase loops, where they would hold vector registers across the whole
// kernel and push the SpMV stream into spills.
__device__ __forceinline__ int phase_tid() {
    int t = threadIdx.x;
    asm volatile("" : "+v"(t));
    return t;
}

// A value every lane of the wavefront holds, kept in scalar registers (the
// compiler cannot see that sums read back from LDS are uniform; without this
// the loop-carried CG scalars occupy vector registers the SpMV stream needs).
__device__ __forceinline__ double uniform(double v) { return read_lane(v, 0); }

__device__ __forceinline__ void grid_sync(unsigned *bar, unsigned target, long long limit) {
    __syncthreads();
    if (threadIdx.x == 0) {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const long long t0 = wall_clock64();
        for (int spin = 0;; spin++) {
            if (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= target)
                break;
            if ((spin & 31) == 31) {
                if (__hip_atomic_load(bar + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) break;
                if (wall_clock64() - t0 > limit) {
                    __hip_atomic_store(bar + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
            }
            __builtin_amdgcn_s_sleep(1);
        }
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
}

// The SpMV phase: y_b[row] = sum over block b of a * v (v = scale * src), for
// every row of this workgroup's range.  With dot: the r.s partial
// sum_rows v[row] * y_b[row] into part_dot[wg] (summed in slice order).  With
// store_x: row-range-0 workgroups store v over their block into x and the x.x
// partial into part_xx[wg] (other workgroups store 0).  The flags are runtime
// values (uniform branches), so the stream loop is compiled once.
template <int BLOCK>
__device__ __forceinline__ void spmv_phase(const PersistArgs &a, const double *src, double scale,
                                           int rev, bool dot, bool store_x) {
    extern __shared__ double2 lds_v2[];  // 16-byte aligned
    double *vblk = reinterpret_cast<double *>(lds_v2);
    double *slice_dot = vblk + a.W;
    const int wg = blockIdx.x;
    const int b = wg / a.nrange;
    const int c0 = b * a.W;
    const int wb = min(a.W, a.n - c0);
    // stage scale * src[c0 .. c0 + wb): every thread's loads go out before any
    // of its stores (indices clamped, as stage_block)
    constexpr int kFill = 10;  // 16-byte loads in flight per thread before the stores
    const int tid = phase_tid();
    const bool own_x = store_x && (wg % a.nrange) == 0;
    const double2 *s2 = reinterpret_cast<const double2 *>(src + c0);
    double2 *x2 = reinterpret_cast<double2 *>(a.x + c0);
    const int half = wb / 2;
    double xx = 0.0;
    if (half > 0) {
        double2 t[kFill];
#pragma unroll
        for (int k = 0; k < kFill; k++) t[k] = s2[min(tid + k * BLOCK, half - 1)];
#pragma unroll
        for (int k = 0; k < kFill; k++) {
            const int i = tid + k * BLOCK;
            double2 v;
            v.x = scale * t[k].x;
            v.y = scale * t[k].y;
            lds_v2[min(i, half - 1)] = v;
            if (own_x && i < half) {
                xx += v.x * v.x + v.y * v.y;
                x2[i] = v;
            }
        }
        for (int i = tid + kFill * BLOCK; i < half; i += BLOCK) {
            double2 v = s2[i];
            v.x *= scale;
            v.y *= scale;
            lds_v2[i] = v;
            if (own_x) {
                xx += v.x * v.x + v.y * v.y;
                x2[i] = v;
            }
        }
    }
    if ((wb & 1) && tid == 0) {
        const double v = scale * src[c0 + wb - 1];
        vblk[wb - 1] = v;
        if (own_x) {
            xx += v * v;
            a.x[c0 + wb - 1] = v;
        }
    }
    if (store_x) store_block_partial<BLOCK>(xx, a.part_xx);
    SellWalk<double, BLOCK, kPersistUnr> walk;
    walk.begin(a.wg_slice, a.slice_off, a.slice_row, a.val, a.col, wg, rev);
    __syncthreads();
    double *part_wg = a.rows_part + (size_t)walk.s0 * kWave;
    walk.run(
        vblk, [&](int row) -> double { return (dot && row >= 0) ? scale * src[row] : 0.0; },
        [&](int k, int row, double acc, double pr) {
            part_wg[(size_t)k * kWave + (threadIdx.x & (kWave - 1))] = acc;
            if (dot) {
                const double d = segment_sum<64>(pr * acc);
                if ((threadIdx.x & (kWave - 1)) == 0) slice_dot[k] = d;
            }
        });
    if (dot) {
        __syncthreads();
        double sum = 0.0;
        for (int i = threadIdx.x; i < walk.nsl; i += BLOCK) sum += slice_dot[i];
        store_block_partial<BLOCK>(sum, a.part_rs);
    }
}

template <int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(4))) void cg_persistent_kernel(const PersistArgs *__restrict__ args) {
    const int G = gridDim.x;
    const int gs = G * BLOCK;
    unsigned nbar = 0;
    auto sync = [&]() {
        const PersistArgs a = phase_args(args);
        grid_sync(a.bar, ++nbar * (unsigned)G, a.spin_limit);
    };
    int niter, cgitmax, scale_first, hist;
    double shift, xscale;
    {
        const PersistArgs a = phase_args(args);
        niter = a.niter;
        cgitmax = a.cgitmax;
        scale_first = a.scale_first;
        // index n as cg.cu:671-676 leaves it: q = z = 0, r = p = x[n]
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            const double xn = a.x[a.n];
            a.q[a.n] = 0.0;
            a.z[a.n] = 0.0;
            a.r[a.n] = xn;
            a.p[a.n] = xn;
        }
        shift = uniform(a.sc->shift);
        xscale = uniform(scale_first ? a.sc->xscale : 1.0);
        hist = __builtin_amdgcn_readfirstlane(a.sc->iter);
    }
    for (int it = 0; it < niter; it++) {
        const bool scale = it > 0 || scale_first;
        double rho_prev = 0.0, alpha_prev = 0.0;
        for (int k = 0; k < cgitmax; k++) {
            {  // ---- s = A r (r = x, or xscale z_old, at k = 0)
                const PersistArgs a = phase_args(args);
                const double *src = k > 0 ? a.r : scale ? a.z : a.x;
                spmv_phase<BLOCK>(a, src, (k == 0 && scale) ? xscale : 1.0, k & 1, true, k == 0);
            }
            sync();
            {  // ---- the update of step k
                const PersistArgs a = phase_args(args);
                const int n = a.n, nblk = a.nblk;
                double rho, rs;
                reduce_partials2<BLOCK>(k == 0 ? a.part_xx : ((k & 1) ? a.part_rr[0] : a.part_rr[1]), G,
                                        a.part_rs, G, rho, rs);
                rho = uniform(rho);
                rs = uniform(rs);
                double beta = 0.0, delta = rs;
                if (k > 0) {
                    beta = uniform(rho / rho_prev);
                    delta = uniform(rs - beta * rho / alpha_prev);
                }
                const double alpha = uniform(rho / delta);
                double acc = 0.0;
                for (int j = xcd_chunk(blockIdx.x, G) * BLOCK + phase_tid(); j < n; j += gs) {
                    double sj = a.rows_part[a.row_slot[j]];
#pragma unroll 8
                    for (int bb = 1; bb < nblk; bb++)
                        sj += a.rows_part[a.row_slot[(size_t)bb * n + j]];
                    double pj, qj, zj, rj;
                    if (k == 0) {  // p = r = x, q = s, z = 0 (cg.cu:671-686)
                        rj = a.x[j];
                        pj = rj;
                        qj = sj;
                        zj = alpha * pj;
                    } else {
                        rj = a.r[j];
                        pj = rj + beta * a.p[j];
                        qj = sj + beta * a.q[j];
                        zj = a.z[j] + alpha * pj;
                    }
                    const double rn = rj - alpha * qj;
                    a.p[j] = pj;
                    a.q[j] = qj;
                    a.z[j] = zj;
                    a.r[j] = rn;
                    acc += rn * rn;
                }
                store_block_partial<BLOCK>(acc, (k & 1) ? a.part_rr[1] : a.part_rr[0]);
                rho_prev = rho;
                alpha_prev = alpha;
            }
            sync();
        }
        {  // ---- p = r + beta p closing the last step (cg.cu:762-771), then r = A z
            const PersistArgs a = phase_args(args);
            if (cgitmax > 0) {
                const double rr = uniform(reduce_partials<BLOCK>(((cgitmax - 1) & 1) ? a.part_rr[1] : a.part_rr[0], G));
                const double beta = uniform(rr / rho_prev);
                for (int j = blockIdx.x * BLOCK + phase_tid(); j < a.n; j += gs)
                    a.p[j] = a.r[j] + beta * a.p[j];
            }
            spmv_phase<BLOCK>(a, a.z, 1.0, 1, false, false);
        }
        sync();
        {  // ---- r = sum of the block row sums; ||x - r||^2, x.z, z.z partials
            const PersistArgs a = phase_args(args);
            const int n = a.n, nblk = a.nblk;
            double e0 = 0.0, e1 = 0.0, e2 = 0.0;
            for (int j = xcd_chunk(blockIdx.x, G) * BLOCK + phase_tid(); j < n; j += gs) {
                double rj = a.rows_part[a.row_slot[j]];
#pragma unroll 8
                for (int bb = 1; bb < nblk; bb++) rj += a.rows_part[a.row_slot[(size_t)bb * n + j]];
                a.r[j] = rj;
                const double xv = a.x[j], zv = a.z[j];
                const double dlt = xv - rj;
                e0 += dlt * dlt;
                e1 += xv * zv;
                e2 += zv * zv;
            }
            store_block_partial<BLOCK>(e0, a.part_res);
            __syncthreads();
            store_block_partial<BLOCK>(e1, a.part_res + G);
            __syncthreads();
            store_block_partial<BLOCK>(e2, a.part_res + 2 * G);
        }
        sync();
        {  // ---- rnorm, zeta, xscale: the same bits in every workgroup
            const PersistArgs a = phase_args(args);
            const double s0 = uniform(reduce_partials<BLOCK>(a.part_res, G));
            const double s1 = uniform(reduce_partials<BLOCK>(a.part_res + G, G));
            const double s2 = uniform(reduce_partials<BLOCK>(a.part_res + 2 * G, G));
            xscale = uniform(1.0 / sqrt(s2));
            if (blockIdx.x == 0 && threadIdx.x == 0) {
                const double rnorm = sqrt(s0);
                const double zeta = shift + 1.0 / s1;
                a.sc->rnorm = rnorm;
                a.sc->xz = s1;
                a.sc->zz = s2;
                a.sc->xscale = xscale;
                a.sc->zeta = zeta;
                if (a.record) {
                    if (hist < a.hist_cap) {
                        a.zeta_hist[hist] = zeta;
                        a.rnorm_hist[hist] = rnorm;
                    }
                    a.sc->iter = hist + 1;
                }
            }
            hist++;
        }
    }
    // x = z / ||z|| after the last iteration (cg.cu:519, applied once more)
    const PersistArgs a = phase_args(args);
    if (a.final_scale && niter > 0)
        for (int j = blockIdx.x * BLOCK + threadIdx.x; j < a.n; j += gs) a.x[j] = xscale * a.z[j];
}

// Stores the argument block the persistent kernel reads and clears the
// barrier words (a kernel, so the block travels by value with the launch).
__global__ void persist_setup_kernel(PersistArgs a, PersistArgs *dst) {
    if (threadIdx.x == 0) {
        *dst = a;
        a.bar[0] = 0u;
        a.bar[1] = 0u;
    }
}

template <int BLOCK>
hipError_t launch_t(const PersistArgs *dev_args, int grid, int lds, hipStream_t st) {
    void *args[] = {&dev_args};
    return hipLaunchCooperativeKernel(reinterpret_cast<const void *>(&cg_persistent_kernel<BLOCK>),
                                      dim3(grid), dim3(BLOCK), args, (unsigned)lds, st);
}

template <int BLOCK>
hipError_t occupancy_t(int lds, int *per_cu) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void *>(&cg_persistent_kernel<BLOCK>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, lds);
    if (e != hipSuccess) return e;
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, cg_persistent_kernel<BLOCK>,
                                                        BLOCK, (size_t)lds);
}

}  // namespace

hipError_t persistent_capacity(const Fp64Problem &pb, int *per_cu) {
    *per_cu = 0;
    if (pb.spmv_variant != kSpmvTiled) return hipSuccess;
    return pb.sell.block == 512 ? occupancy_t<512>(pb.sell.lds_bytes, per_cu)
                                : occupancy_t<1024>(pb.sell.lds_bytes, per_cu);
}

int persistent_partials(const Fp64Problem &pb) { return 7 * pb.sell.grid; }

size_t persistent_args_bytes() { return sizeof(PersistArgs); }

// Barriers one outer iteration passes: two per CG step, two around the
// residual.
long long persistent_barriers(int cgitmax) { return 2LL * cgitmax + 2; }

hipError_t launch_persistent(const Fp64Problem &pb, const PersistentState &ps, int niter,
                             int scale_first, int record, int final_scale, hipStream_t st) {
    const SellDev &d = pb.sell;
    PersistArgs a{};
    a.wg_slice = (d.wg_slice);
    a.slice_off = (d.slice_off);
    a.slice_row = (d.slice_row);
    a.val = (d.val);
    a.col = (d.col);
    a.row_slot = (d.row_slot);
    a.rows_part = (d.rows_part);
    a.n = pb.n;
    a.W = d.W;
    a.nrange = d.nrange;
    a.nblk = d.nblk;
    a.x = (pb.x);
    a.z = (pb.z);
    a.p = (pb.p);
    a.q = (pb.q);
    a.r = (pb.r);
    const int G = d.grid;
    a.part_rs = (ps.part);
    a.part_xx = (ps.part + G);
    a.part_rr[0] = (ps.part + 2 * G);
    a.part_rr[1] = (ps.part + 3 * G);
    a.part_res = (ps.part + 4 * G);
    a.sc = (pb.sc);
    a.zeta_hist = (pb.zeta_hist);
    a.rnorm_hist = (pb.rnorm_hist);
    a.hist_cap = pb.hist_cap;
    a.bar = (ps.bar);
    a.spin_limit = ps.spin_limit;
    a.cgitmax = ps.cgitmax;
    a.niter = niter;
    a.scale_first = scale_first;
    a.record = record;
    a.final_scale = final_scale;
    PersistArgs *dev = static_cast<PersistArgs *>(ps.args);
    persist_setup_kernel<<<1, 64, 0, st>>>(a, dev);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    return d.block == 512 ? launch_t<512>(dev, G, d.lds_bytes, st)
                          : launch_t<1024>(dev, G, d.lds_bytes, st);
}

}  // namespace npbcg_dev
