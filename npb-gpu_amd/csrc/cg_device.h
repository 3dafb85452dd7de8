// Device-side building blocks shared by the npbcg kernels (gfx950, wave64).
#ifndef NPBCG_CG_DEVICE_H
#define NPBCG_CG_DEVICE_H

#include <hip/hip_runtime.h>

#include <cstdint>

namespace npbcg_dev {

constexpr int kWave = 64;
constexpr int kBlock = 256;             // 4 waves per workgroup
constexpr int kWavesPerBlock = kBlock / kWave;

// DPP lane moves on a double (two 32-bit halves moved with one control).
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
    int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
    int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double read_lane(double v, int lane) {
    int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

// Sum over aligned groups of SEG lanes (SEG in 4..64).  Every lane of a group
// receives the group's sum; the order of additions is fixed, so results are
// reproducible run to run.  Within a 16-lane row the butterfly uses DPP
// quad_perm [1,0,3,2], [2,3,0,1], row_half_mirror and row_mirror; the 16-lane
// row sums are combined with ds_bpermute (SEG = 32) or readlane (SEG = 64).
template <int SEG>
__device__ __forceinline__ double segment_sum(double v) {
    static_assert(SEG == 4 || SEG == 8 || SEG == 16 || SEG == 32 || SEG == 64, "SEG");
    v += dpp_move<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_move<0x4E>(v);   // quad_perm [2,3,0,1]
    if constexpr (SEG >= 8) v += dpp_move<0x141>(v);  // row_half_mirror
    if constexpr (SEG >= 16) v += dpp_move<0x140>(v); // row_mirror
    if constexpr (SEG == 32) v += __shfl_xor(v, 16, 64);
    if constexpr (SEG == 64)
        v = (read_lane(v, 0) + read_lane(v, 16)) + (read_lane(v, 32) + read_lane(v, 48));
    return v;
}

// The chunk of kBlock rows (per grid sweep) that workgroup bid of nb takes
// in the kernels that read block row sums by slot.  With nb a multiple of 8,
// the workgroups dispatched to one XCD (every 8th) take consecutive chunks,
// so the slots of their rows, which lie in runs of slices per sort window,
// are fetched into that XCD's L2 once rather than by every XCD.
__device__ __forceinline__ int xcd_chunk(int bid, int nb) {
    return (nb & 7) ? bid : (bid & 7) * (nb >> 3) + (bid >> 3);
}

// Copies count elements src[0..count) to LDS dst[0..count) with BLOCK
// threads.  Every thread first issues its loads for the first FILL * BLOCK
// elements (FILL in flight at once, held in registers), then writes them: the
// copy costs one memory round trip instead of one per few elements.  Indices
// past the end are clamped to the last element on both sides (those lanes
// write its own value again), so neither loop branches and the compiler
// cannot sink a load next to its store.  Elements past FILL * BLOCK (none for
// the block sizes used) follow in a plain loop.
template <int BLOCK, int FILL, typename V>
__device__ __forceinline__ void stage_block(const V *__restrict__ src, V *dst, int count) {
    if (count <= 0) return;
    V t[FILL];
#pragma unroll
    for (int k = 0; k < FILL; k++) t[k] = src[min((int)threadIdx.x + k * BLOCK, count - 1)];
#pragma unroll
    for (int k = 0; k < FILL; k++) dst[min((int)threadIdx.x + k * BLOCK, count - 1)] = t[k];
    for (int i = threadIdx.x + FILL * BLOCK; i < count; i += BLOCK) dst[i] = src[i];
}

// Loads entries j0 .. j0+UNR-1 of lane `lane` of the SELL slice at `base`
// (entry j at base + 64 j + lane; j0 < width).  A full chunk loads at a
// constant stride, so the offsets are instruction immediates.  In the
// slice's last chunk (a wave-uniform case) chunks of more than 8 entries
// load only the entries left and zero the rest; shorter chunks clamp the
// index to width - 1, so no load leaves the slice, and the entries at or past
// the width are cleared by sell_trim where the chunk is used, so the loads
// never wait on a mask.
template <typename T, int UNR>
__device__ __forceinline__ void sell_chunk(const T *__restrict__ val,
                                           const uint16_t *__restrict__ col, int base, int width,
                                           int j0, int lane, T (&a)[UNR], int (&c)[UNR]) {
    if constexpr (UNR > 8) {
        // base, width and j0 are the same for the whole wavefront; saying so
        // keeps the clamped path's offsets scalar (without it a 12-entry FP64
        // chunk exceeds 128 VGPRs and spills; at 8 entries it is slower)
        base = __builtin_amdgcn_readfirstlane(base);
        width = __builtin_amdgcn_readfirstlane(width);
        j0 = __builtin_amdgcn_readfirstlane(j0);
    }
    const size_t at = (size_t)base + (size_t)j0 * kWave + lane;
    const T *vp = val + at;
    const uint16_t *cp = col + at;
    if (j0 + UNR <= width) {
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            a[u] = vp[u * kWave];
            c[u] = cp[u * kWave];
        }
    } else if constexpr (UNR > 8) {
        // cnt = width - j0 entries are left (1 .. UNR-1, wave-uniform as
        // marked above): only those are loaded, under scalar branches, and
        // the rest of the chunk is (value 0, column 0).  With a few entries
        // per row and block (1.5 M rows: about 5) the clamped loads below
        // issued up to twice the load instructions the entries need: FP64
        // SpMV 2.27 against 2.00 ms there, 68.4 against 67.7 us at
        // n = 150000.  The FP32 SpMV's 8-entry chunks keep the clamped
        // loads: skipping measured 18.6-18.7 against 18.4-18.5 us at
        // n = 75000 and 2.78 against 2.69 ms at 1.5 M rows, with or without
        // the uniform marking.
        const int cnt = width - j0;
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            if (u < cnt) {
                a[u] = vp[u * kWave];
                c[u] = cp[u * kWave];
            } else {
                a[u] = T(0);
                c[u] = 0;
            }
        }
    } else {
        const int last = width - 1 - j0;
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            const int k = min(u, last) * kWave;
            a[u] = vp[k];
            c[u] = cp[k];
        }
    }
}

// Value 0 and column 0 for the entries of a chunk at or past `valid` (the
// slice width less the chunk's first entry; the same for the whole wavefront).
template <typename T, int UNR>
__device__ __forceinline__ void sell_trim(int valid, T (&a)[UNR], int (&c)[UNR]) {
    if (valid >= UNR) return;
#pragma unroll
    for (int u = 0; u < UNR; u++)
        if (u >= valid) {
            a[u] = T(0);
            c[u] = 0;
        }
}

// Device-resident CG scalars of one context.  Kernels never pass values to
// other workgroups of the same launch: each workgroup stores a partial sum and
// the next kernel in the stream reduces the partials (every workgroup in the
// same fixed order, so all obtain the same bits).  Scalars that outlive a
// launch (written by workgroup 0) are read only by later launches.
struct Scalars {
    double rho;      // r.r of the current residual
    double rho0;     // previous rho
    double d;        // p.q
    double alpha;    // rho / d
    double beta;     // rho / rho0
    double rnorm;    // ||x - A z||
    double xz;       // x.z  (norm_temp1)
    double zz;       // z.z
    double xscale;   // 1/sqrt(z.z)  (norm_temp2)
    double zeta;     // shift + 1/(x.z)
    double shift;
    double tol2;     // solve: stop when r.r <= tol2 (= rtol^2 * r0.r0)
    double rtol2;    // solve: rtol^2, set before the solve
    double bb;       // solve: r0.r0 of the step that set tol2 (b.b for x0 = 0)
    int iter;        // outer iterations recorded so far in the history
    int stop_enabled;  // solve mode: CG steps stop once r.r <= tol2
    int done;        // set when the stop test passed; later CG steps return at once
    int steps;       // CG steps executed since the last reset
    int skip_resid;  // mixed solve: the residual pass of a sweep that is not needed
    int sweeps;      // mixed solve: refinement sweeps run
    int inner_total; // mixed solve: inner steps of the finished sweeps
    int sweep1_steps;  // mixed solve: inner steps of the first two sweeps
    int sweep2_steps;
    int pad;
    // one-reduction CG form: rho and alpha of CG step k in slot k & 1 (a step
    // reads the previous step's slot while its workgroup 0 writes its own)
    double cg_rho[2];
    double cg_alpha[2];
};

// Sum of this workgroup's per-thread values, written to part[blockIdx.x]
// (one store per workgroup; fixed order: DPP tree per wave, waves in order).
template <int BLOCK>
__device__ __forceinline__ void store_block_partial(double acc, double *part) {
    constexpr int kWaves = BLOCK / kWave;
    __shared__ double wsum[kWaves];
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    const double s = segment_sum<64>(acc);
    if (lane == 0) wsum[wid] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = wsum[0];
        for (int w = 1; w < kWaves; w++) t += wsum[w];
        part[blockIdx.x] = t;
    }
}

// Sum of part[0..count) by the whole workgroup, in a fixed order that
// depends only on count: the first kBlock threads sum strided slices, their
// four wavefront sums are added in order.  Any BLOCK >= kBlock therefore gets
// the same bits, so kernels of different shapes agree on alpha and beta.
// Every thread receives the result.
template <int BLOCK>
__device__ __forceinline__ double reduce_partials(const double *__restrict__ part, int count) {
    static_assert(BLOCK % kBlock == 0, "BLOCK must be a multiple of kBlock");
    __shared__ double wsum[kWavesPerBlock];
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    if (__builtin_amdgcn_readfirstlane(threadIdx.x / kWave) < kWavesPerBlock) {
        double s = 0.0;
        for (int i = threadIdx.x; i < count; i += kBlock) s += part[i];
        s = segment_sum<64>(s);
        if (lane == 0) wsum[wid] = s;
    }
    __syncthreads();
    double t = wsum[0];
    for (int w = 1; w < kWavesPerBlock; w++) t += wsum[w];
    __syncthreads();  // wsum may be reused by the caller's next reduction
    return t;
}

// reduce_partials of two arrays at once: both arrays' loads are issued
// before either sum is formed, so the two reductions cost one memory round
// trip.  Each result has the bits reduce_partials gives for its array.
template <int BLOCK>
__device__ __forceinline__ void reduce_partials2(const double *__restrict__ pa, int na,
                                                 const double *__restrict__ pb, int nb,
                                                 double &ra, double &rb) {
    static_assert(BLOCK % kBlock == 0, "BLOCK must be a multiple of kBlock");
    __shared__ double wsum2[2][kWavesPerBlock];
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    if (__builtin_amdgcn_readfirstlane(threadIdx.x / kWave) < kWavesPerBlock) {
        const int t = threadIdx.x;
        const double a0 = t < na ? pa[t] : 0.0;
        const double b0 = t < nb ? pb[t] : 0.0;
        double sa = 0.0, sb = 0.0;
        sa += a0;
        sb += b0;
        for (int i = t + kBlock; i < na; i += kBlock) sa += pa[i];
        for (int i = t + kBlock; i < nb; i += kBlock) sb += pb[i];
        sa = segment_sum<64>(sa);
        sb = segment_sum<64>(sb);
        if (lane == 0) {
            wsum2[0][wid] = sa;
            wsum2[1][wid] = sb;
        }
    }
    __syncthreads();
    double ta = wsum2[0][0], tb = wsum2[1][0];
    for (int w = 1; w < kWavesPerBlock; w++) {
        ta += wsum2[0][w];
        tb += wsum2[1][w];
    }
    __syncthreads();  // wsum2 may be reused by the caller's next reduction
    ra = ta;
    rb = tb;
}

}  // namespace npbcg_dev

#endif
