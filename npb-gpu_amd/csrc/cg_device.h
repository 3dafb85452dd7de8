// Device-side building blocks shared by the npbcg kernels (gfx950, wave64).
#ifndef NPBCG_CG_DEVICE_H
#define NPBCG_CG_DEVICE_H

#include <hip/hip_runtime.h>

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

// Device-resident CG scalars of one context.  Written by the last workgroup
// of the kernel that completes each reduction and read by later kernels, so
// no scalar crosses to the host inside the solver loops.
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
    int iter;        // outer iterations recorded so far in the history
    int pending_scale;  // 1: x must become xscale * z before the next conj_grad
    unsigned ticket[8]; // last-workgroup counters, one per reduction site
};

enum TicketSlot { kTicketInit = 0, kTicketSpmv = 1, kTicketZR = 2, kTicketResid = 3,
                  kTicketNorms = 4 };

// Reduce NV per-thread accumulators over the workgroup (fixed order) and,
// in the last workgroup to finish, over all workgroups' partials.  Returns
// true in the last workgroup, where tot[] then holds the grid totals (valid
// in every thread of that workgroup).  part holds NV*gridDim.x doubles.
//
// Ordering: each workgroup stores its partial, issues an agent-scope RELEASE
// fence (L2 write-back only) and bumps the ticket with a relaxed atomic; the
// last workgroup issues an agent-scope ACQUIRE fence before reading the
// partials.  A full __threadfence() in every workgroup would also invalidate
// the XCD's L2 each time and evict the gathered vector mid-kernel.
template <int NV, int BLOCK = kBlock>
__device__ __forceinline__ bool grid_reduce(double (&acc)[NV], double *part,
                                            unsigned *ticket, double (&tot)[NV]) {
    constexpr int kWaves = BLOCK / kWave;
    __shared__ double wsum[NV][kWaves];
    __shared__ bool is_last;
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
#pragma unroll
    for (int v = 0; v < NV; v++) {
        double s = segment_sum<64>(acc[v]);
        if (lane == 0) wsum[v][wid] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int v = 0; v < NV; v++) {
            double s = wsum[v][0];
            for (int w = 1; w < kWaves; w++) s += wsum[v][w];
            part[(size_t)v * gridDim.x + blockIdx.x] = s;
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        unsigned t = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT);
        is_last = (t == gridDim.x - 1);
    }
    __syncthreads();
    if (!is_last) return false;
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    const int nb = gridDim.x;
#pragma unroll
    for (int v = 0; v < NV; v++) {
        double s = 0.0;
        for (int b = threadIdx.x; b < nb; b += BLOCK)
            s += part[(size_t)v * nb + b];
        s = segment_sum<64>(s);
        if (lane == 0) wsum[v][wid] = s;
    }
    __syncthreads();
#pragma unroll
    for (int v = 0; v < NV; v++) {
        double s = wsum[v][0];
        for (int w = 1; w < kWaves; w++) s += wsum[v][w];
        tot[v] = s;
    }
    if (threadIdx.x == 0)
        __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return true;
}

}  // namespace npbcg_dev

#endif
