// Device kernels of the learned sparse preconditioner (gfx950).
//
// A message-passing network over the CSR row graph predicts, for every row j,
// two scale corrections g_j, h_j of a sparse approximate inverse:
//   M^-1 = S D^-1 S - Q (N + N^T) Q,  S = diag(e^g), Q = diag(e^(g+h)),
//   N_jk = a_jk / (a_jj a_kk) on the top-K off-diagonals of row j.
//
//   graph     top-K off-diagonals of each row by |a_jk| (ties: lower column);
//             message weights |a_jk| / sqrt|a_jj a_kk|, normalised to sum 1
//   features  8 dimensionless row statistics, standardised per model
//   network   H0 = ReLU(X0 W0 + b0)                     X0: features, 32 wide
//             H_l = ReLU([H_{l-1} | sum_k w_jk H_{l-1,k}] W_l + b_l), l = 1..3
//             g = G tanh(ug . H3 + cg),  h = G tanh(uh . H3 + ch)
//   loss      residual/correction pairs (r_p, e_p = A^-1 r_p) logged from plain
//             CG: L = sum_p |M^-1 r_p - e_p|^2 / (P |e_p|^2)
// GEMMs run on v_mfma_f32_16x16x32_bf16 (BF16 operands, FP32 accumulation).
// Activations are stored BF16; gradients are FP32 until they feed an MFMA.
// Every reduction has a fixed order, so a run repeats bit for bit.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "cg_device.h"
#include "learned_kernels.h"

namespace npbcg_dev {

typedef __attribute__((__vector_size__(8 * sizeof(__bf16)))) __bf16 bf16x8;
typedef __attribute__((__vector_size__(4 * sizeof(__bf16)))) __bf16 bf16x4;
typedef __attribute__((__vector_size__(4 * sizeof(float)))) float f32x4;

namespace {

// Workgroups of a grid-stride kernel (capped; the kernel loops).
int blocks_for(long long items, int per_block = kBlock, int cap = 65535) {
    long long b = (items + per_block - 1) / per_block;
    return (int)std::max<long long>(1, std::min<long long>(b, cap));
}

// Workgroups of a kernel that covers exactly one item per thread group (no
// loop, so no cap).
int exact_blocks(long long items, int per_block) {
    return (int)std::max<long long>(1, (items + per_block - 1) / per_block);
}

}  // namespace

int row_blocks(int n) { return exact_blocks((long long)n * 16, kBlock); }

// ---------------------------------------------------------------------------
// Graph and features (once per matrix)
// ---------------------------------------------------------------------------

// a_jj of every row, 16 lanes per row (0 when the row stores no diagonal).
__global__ __launch_bounds__(kBlock) void diag_kernel(const int *__restrict__ rowstr,
                                                     const int *__restrict__ colidx,
                                                     const double *__restrict__ a,
                                                     double *__restrict__ adiag, int n) {
    const long long gid = blockIdx.x * (long long)kBlock + threadIdx.x;
    const int j = (int)(gid / 16), sub = (int)(gid % 16);
    double d = 0.0;
    if (j < n)
        for (int k = rowstr[j] + sub; k < rowstr[j + 1]; k += 16)
            if (colidx[k] == j) d = a[k];
    d = segment_sum<16>(d);  // one lane holds the diagonal, the others 0
    if (j < n && sub == 0) adiag[j] = d;
}

__device__ __forceinline__ double wave_max(double v) {
    for (int m = 32; m >= 1; m >>= 1) v = fmax(v, __shfl_xor(v, m, kWave));
    return v;
}

__device__ __forceinline__ double safe_abs(double d) { return d != 0.0 ? fabs(d) : 1.0; }

// One wavefront per row: the features and the top-K neighbours.  Selection
// round t takes the best candidate (|a| descending, column ascending) strictly
// after round t-1's; a row with fewer than K off-diagonals is padded with
// zero-weight self edges.  A row of at most kFeatCache * 64 entries is loaded
// into registers once (lane l holds entries l, l + 64, ...; absent entries
// carry column j, which every pass skips as it skips the diagonal) and the
// feature pass and the K selection rounds read it from there, in the same
// order as from memory; longer rows are read from memory in every pass.
__global__ __launch_bounds__(kBlock) void graph_features_kernel(
    const int *__restrict__ rowstr, const int *__restrict__ colidx,
    const double *__restrict__ a, const double *__restrict__ adiag, int n,
    float *__restrict__ feat, int *__restrict__ nbr, float *__restrict__ wgt,
    float *__restrict__ nval) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    constexpr int kFeatCache = 8;
    const int nw = gridDim.x * kWavesPerBlock;
    for (int j = blockIdx.x * kWavesPerBlock + wave; j < n; j += nw) {
        const int s = rowstr[j], e = rowstr[j + 1];
        const double dj = adiag[j];
        const double aj = safe_abs(dj);
        const bool cached = e - s <= kFeatCache * kWave;  // the same for the whole wavefront
        int cc[kFeatCache];
        double cv[kFeatCache], cd[kFeatCache];
        if (cached) {
#pragma unroll
            for (int u = 0; u < kFeatCache; u++) {
                const int k = s + lane + u * kWave;
                cc[u] = k < e ? colidx[k] : j;
                cv[u] = k < e ? a[k] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < kFeatCache; u++) cd[u] = cc[u] != j ? adiag[cc[u]] : 1.0;
        }
        double s_abs = 0, s_sq = 0, s_sgn = 0, s_nb = 0, s_w = 0, mx = 0;
        auto feature_terms = [&](int c, double v, double adc) {
            const double av = fabs(v);
            const double ak = safe_abs(adc);
            s_abs += av;
            s_sq += v * v;
            s_sgn += v;
            s_nb += v * v / (aj * ak);
            s_w += av / sqrt(aj * ak);
            mx = fmax(mx, av);
        };
        if (cached) {
#pragma unroll
            for (int u = 0; u < kFeatCache; u++)
                if (cc[u] != j) feature_terms(cc[u], cv[u], cd[u]);
        } else {
            for (int k = s + lane; k < e; k += kWave) {
                const int c = colidx[k];
                if (c == j) continue;
                feature_terms(c, a[k], adiag[c]);
            }
        }
        s_abs = segment_sum<64>(s_abs);
        s_sq = segment_sum<64>(s_sq);
        s_sgn = segment_sum<64>(s_sgn);
        s_nb = segment_sum<64>(s_nb);
        s_w = segment_sum<64>(s_w);
        mx = wave_max(mx);
        if (lane == 0) {
            float *f = feat + (size_t)j * kFeat;
            f[0] = (float)(s_abs / aj);
            f[1] = (float)(s_sq / (aj * aj));
            f[2] = (float)(mx / aj);
            f[3] = (float)(s_sgn / aj);
            f[4] = (float)s_nb;
            f[5] = (float)s_w;
            f[6] = (float)(log2((double)(e - s)) / 8.0);
            f[7] = (float)(log10(aj) / 3.0);
        }
        double pv = INFINITY;
        int pc = -1;
        double wk[kNbr], nk[kNbr];
        int ck[kNbr];
        double wsum = 0.0;
#pragma unroll
        for (int t = 0; t < kNbr; t++) {
            double bv = -1.0, bval = 0.0;
            int bc = 0x7fffffff;
            auto candidate = [&](int c, double v) {
                const double av = fabs(v);
                const bool after = av < pv || (av == pv && c > pc);
                const bool better = av > bv || (av == bv && c < bc);
                if (after && better) {
                    bv = av;
                    bc = c;
                    bval = v;
                }
            };
            if (cached) {
#pragma unroll
                for (int u = 0; u < kFeatCache; u++)
                    if (cc[u] != j) candidate(cc[u], cv[u]);
            } else {
                for (int k = s + lane; k < e; k += kWave) {
                    const int c = colidx[k];
                    if (c == j) continue;
                    candidate(c, a[k]);
                }
            }
            for (int m = 32; m >= 1; m >>= 1) {
                const double ov = __shfl_xor(bv, m, kWave);
                const int oc = __shfl_xor(bc, m, kWave);
                const double oval = __shfl_xor(bval, m, kWave);
                if (ov > bv || (ov == bv && oc < bc)) {
                    bv = ov;
                    bc = oc;
                    bval = oval;
                }
            }
            if (bv >= 0.0 && bc != 0x7fffffff) {
                const double ak = adiag[bc];
                ck[t] = bc;
                wk[t] = bv / sqrt(aj * safe_abs(ak));
                nk[t] = (dj != 0.0 && ak != 0.0) ? bval / (dj * ak) : 0.0;
            } else {
                ck[t] = j;
                wk[t] = 0.0;
                nk[t] = 0.0;
            }
            wsum += wk[t];
            pv = bv;
            pc = bc;
        }
        if (lane == 0) {
            const double inv = wsum > 0.0 ? 1.0 / wsum : 0.0;
#pragma unroll
            for (int t = 0; t < kNbr; t++) {
                const size_t at = (size_t)j * kNbr + t;
                nbr[at] = ck[t];
                wgt[at] = (float)(wk[t] * inv);
                nval[at] = (float)nk[t];
            }
        }
    }
}

// In-degree of every node over the edges with nonzero weight.
__global__ __launch_bounds__(kBlock) void indegree_kernel(const int *__restrict__ nbr,
                                                          const float *__restrict__ wgt, int n,
                                                          int *__restrict__ cnt) {
    const long long m = (long long)n * kNbr;
    for (long long e = blockIdx.x * (long long)kBlock + threadIdx.x; e < m;
         e += (long long)gridDim.x * kBlock)
        if (wgt[e] != 0.0f) atomicAdd(&cnt[nbr[e]], 1);
}

// ptr[j] = sum of cnt[0..j) for j <= n (exact, so any order gives the same
// values): one 1024-thread workgroup walks the counts in tiles of 8192, each
// thread taking 8 consecutive counts (coalesced across the wavefront), a
// wavefront scan of the thread sums, then the wavefront totals through LDS,
// with a running carry between tiles.
__global__ __launch_bounds__(1024) void scan_kernel(const int *__restrict__ cnt,
                                                    int *__restrict__ ptr, int n) {
    constexpr int kPer = 8, kTile = 1024 * kPer, kWaves = 1024 / kWave;
    __shared__ int wsum[kWaves];
    const int t = threadIdx.x, lane = t & (kWave - 1), w = t / kWave;
    int carry = 0;
    for (int base = 0; base < n; base += kTile) {
        int v[kPer], s = 0;
#pragma unroll
        for (int k = 0; k < kPer; k++) {
            const int i = base + t * kPer + k;
            v[k] = i < n ? cnt[i] : 0;
            s += v[k];
        }
        int x = s;  // inclusive scan over the wavefront
#pragma unroll
        for (int off = 1; off < kWave; off <<= 1) {
            const int y = __shfl_up(x, off, kWave);
            if (lane >= off) x += y;
        }
        if (lane == kWave - 1) wsum[w] = x;
        __syncthreads();
        int before = 0, total = 0;
#pragma unroll
        for (int i = 0; i < kWaves; i++) {
            if (i < w) before += wsum[i];
            total += wsum[i];
        }
        int e = carry + before + x - s;  // exclusive prefix of this thread's first count
#pragma unroll
        for (int k = 0; k < kPer; k++) {
            const int i = base + t * kPer + k;
            if (i < n) ptr[i] = e;
            e += v[k];
        }
        carry += total;
        __syncthreads();  // wsum is rewritten by the next tile
    }
    if (t == 0) ptr[n] = carry;
}

__global__ __launch_bounds__(kBlock) void transpose_fill_kernel(
    const int *__restrict__ nbr, const float *__restrict__ wgt, const float *__restrict__ nval,
    int n, const int *__restrict__ ptr, int *__restrict__ cursor, int *__restrict__ tsrc,
    float *__restrict__ tw, float *__restrict__ tval) {
    const long long m = (long long)n * kNbr;
    for (long long e = blockIdx.x * (long long)kBlock + threadIdx.x; e < m;
         e += (long long)gridDim.x * kBlock) {
        if (wgt[e] == 0.0f) continue;
        const int k = nbr[e];
        const int at = ptr[k] + atomicAdd(&cursor[k], 1);
        tsrc[at] = (int)(e / kNbr);
        tw[at] = wgt[e];
        tval[at] = nval[e];
    }
}

// Orders each in-list by source row (insertion sort; lists are short), so
// sums over in-edges do not depend on the atomic fill order.
__global__ __launch_bounds__(kBlock) void transpose_sort_kernel(const int *__restrict__ ptr,
                                                                int *__restrict__ tsrc,
                                                                float *__restrict__ tw,
                                                                float *__restrict__ tval,
                                                                int n) {
    for (int k = blockIdx.x * kBlock + threadIdx.x; k < n; k += gridDim.x * kBlock) {
        const int lo = ptr[k], hi = ptr[k + 1];
        for (int i = lo + 1; i < hi; i++) {
            const int sv = tsrc[i];
            const float wv = tw[i], nv = tval[i];
            int p = i - 1;
            while (p >= lo && tsrc[p] > sv) {
                tsrc[p + 1] = tsrc[p];
                tw[p + 1] = tw[p];
                tval[p + 1] = tval[p];
                p--;
            }
            tsrc[p + 1] = sv;
            tw[p + 1] = wv;
            tval[p + 1] = nv;
        }
    }
}

// Mean and standard deviation of each feature over the rows (one workgroup,
// fixed-order tree).  stats = [mean[kFeat], std[kFeat]].
__global__ __launch_bounds__(1024) void feature_stats_kernel(const float *__restrict__ feat,
                                                             int n, float *__restrict__ stats) {
    __shared__ double acc[2 * kFeat][1024 / kWave];
    const int lane = threadIdx.x & (kWave - 1);
    const int wid = threadIdx.x / kWave;
    double s[2 * kFeat] = {};
    for (int j = threadIdx.x; j < n; j += 1024)
#pragma unroll
        for (int c = 0; c < kFeat; c++) {
            const double v = feat[(size_t)j * kFeat + c];
            s[c] += v;
            s[kFeat + c] += v * v;
        }
#pragma unroll
    for (int c = 0; c < 2 * kFeat; c++) {
        const double t = segment_sum<64>(s[c]);
        if (lane == 0) acc[c][wid] = t;
    }
    __syncthreads();
    if (threadIdx.x < kFeat) {
        const int c = threadIdx.x;
        double m = 0.0, q = 0.0;
        for (int w = 0; w < 1024 / kWave; w++) {
            m += acc[c][w];
            q += acc[kFeat + c][w];
        }
        m /= n;
        const double var = fmax(q / n - m * m, 0.0);
        stats[c] = (float)m;
        stats[kFeat + c] = (float)fmax(sqrt(var), 1e-3);
    }
}

// ---------------------------------------------------------------------------
// Training data: right-hand sides and residual/correction pairs
// ---------------------------------------------------------------------------

// Uniform(-1, 1) right-hand side from a 64-bit mix of (seed, j).
__global__ __launch_bounds__(kBlock) void random_rhs_kernel(double *__restrict__ b, int n,
                                                            unsigned long long seed) {
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock) {
        unsigned long long x = seed * 0x9E3779B97F4A7C15ull + (unsigned long long)j;
        x ^= x >> 30;
        x *= 0xBF58476D1CE4E5B9ull;
        x ^= x >> 27;
        x *= 0x94D049BB133111EBull;
        x ^= x >> 31;
        b[j] = 2.0 * ((double)(x >> 11) * (1.0 / 9007199254740992.0)) - 1.0;
    }
}

// r_p = r, e_p = xstar - z (FP32), partials of |e_p|^2.
__global__ __launch_bounds__(kBlock) void log_pair_kernel(
    const double *__restrict__ r, const double *__restrict__ z,
    const double *__restrict__ xstar, int n, float *__restrict__ r_out,
    float *__restrict__ e_out, double *__restrict__ part) {
    double acc = 0.0;
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock) {
        const float ej = (float)(xstar[j] - z[j]);
        r_out[j] = (float)r[j];
        e_out[j] = ej;
        acc += (double)ej * (double)ej;
    }
    store_block_partial<kBlock>(acc, part);
}

__global__ __launch_bounds__(kBlock) void pair_weights_kernel(const double *__restrict__ part,
                                                              int parts, int pairs,
                                                              float *__restrict__ omega) {
    for (int p = 0; p < pairs; p++) {
        const double s = reduce_partials<kBlock>(part + (size_t)p * parts, parts);
        if (threadIdx.x == 0) omega[p] = s > 0.0 ? (float)(1.0 / (pairs * s)) : 0.0f;
    }
}

__global__ __launch_bounds__(kBlock) void sum_kernel(const double *__restrict__ part, int count,
                                                     double *__restrict__ out) {
    const double s = reduce_partials<kBlock>(part, count);
    if (threadIdx.x == 0) out[0] = s;
}

// ---------------------------------------------------------------------------
// Network forward
// ---------------------------------------------------------------------------

// X0 = standardised features, zero-padded to kK0 columns (BF16).
__global__ __launch_bounds__(kBlock) void build_x0_kernel(const float *__restrict__ feat,
                                                          const float *__restrict__ norm, int n,
                                                          __bf16 *__restrict__ x0) {
    const long long m = (long long)n * kK0;
    for (long long i = blockIdx.x * (long long)kBlock + threadIdx.x; i < m;
         i += (long long)gridDim.x * kBlock) {
        const int j = (int)(i / kK0), c = (int)(i % kK0);
        float v = 0.0f;
        if (c < kFeat) v = (feat[(size_t)j * kFeat + c] - norm[c]) / norm[kFeat + c];
        x0[i] = (__bf16)v;
    }
}

// Y[n x 64] = ReLU(X[n x K] W + b): a wavefront computes the four 16x16
// output blocks of kFwdTiles consecutive 16-row tiles, K in steps of 32.  Wt
// is W transposed (64 x K, BF16); the wavefront loads its B fragments (and
// the bias) once and keeps them in registers for all its tiles, so the
// weights are read once per wavefront rather than once per tile (at K = 128
// they are 16 KB against the tile's 4 KB of X).  The next tile's A fragments
// are loaded before the current tile's products.  Lane l supplies
// A[l%16][8(l/16)..+7] and B[8(l/16)..+7][l%16] and receives
// D[4(l/16)..+3][l%16] of each 16x16x32 product.
constexpr int kFwdTiles = 4;
template <int K>
__global__ __launch_bounds__(kBlock) void gemm_fwd_kernel(const __bf16 *__restrict__ x, int ldx,
                                                          int n, const __bf16 *__restrict__ wt,
                                                          const float *__restrict__ bias,
                                                          __bf16 *__restrict__ y, int ldy) {
    constexpr int KS = K / 32;
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int tile0 = (blockIdx.x * kWavesPerBlock + wave) * kFwdTiles;
    if (tile0 * 16 >= n) return;
    const int kq = 8 * (lane >> 4);
    bf16x8 bw[KS][4];
    float bj[4];
#pragma unroll
    for (int jb = 0; jb < 4; jb++) {
#pragma unroll
        for (int k = 0; k < KS; k++)
            bw[k][jb] = *reinterpret_cast<const bf16x8 *>(
                wt + (size_t)(jb * 16 + (lane & 15)) * K + 32 * k + kq);
        bj[jb] = bias[jb * 16 + (lane & 15)];
    }
    auto load_a = [&](int row0, bf16x8 (&av)[KS]) {
        const int ar = row0 + (lane & 15);
#pragma unroll
        for (int k = 0; k < KS; k++) {
            av[k] = bf16x8{};
            if (ar < n) av[k] = *reinterpret_cast<const bf16x8 *>(x + (size_t)ar * ldx + 32 * k + kq);
        }
    };
    bf16x8 av[KS];
    load_a(tile0 * 16, av);
    for (int t = 0; t < kFwdTiles; t++) {
        const int row0 = (tile0 + t) * 16;
        if (row0 >= n) break;  // the same for the whole wavefront
        bf16x8 cur[KS];
#pragma unroll
        for (int k = 0; k < KS; k++) cur[k] = av[k];
        if (t + 1 < kFwdTiles) load_a(row0 + 16, av);
        f32x4 acc[4] = {};
#pragma unroll
        for (int k = 0; k < KS; k++)
#pragma unroll
            for (int jb = 0; jb < 4; jb++)
                acc[jb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(cur[k], bw[k][jb], acc[jb], 0, 0, 0);
#pragma unroll
        for (int jb = 0; jb < 4; jb++) {
            const int j = jb * 16 + (lane & 15);
#pragma unroll
            for (int q = 0; q < 4; q++) {
                const int row = row0 + 4 * (lane >> 4) + q;
                if (row < n) y[(size_t)row * ldy + j] = (__bf16)fmaxf(acc[jb][q] + bj[jb], 0.0f);
            }
        }
    }
}

// X[:, 64:128] = sum_t w[j,t] X[nbr[j,t], 0:64]: 16 lanes per row, 4 columns each.
__global__ __launch_bounds__(kBlock) void aggregate_kernel(__bf16 *__restrict__ x,
                                                           const int *__restrict__ nbr,
                                                           const float *__restrict__ wgt,
                                                           int n) {
    const long long gid = blockIdx.x * (long long)kBlock + threadIdx.x;
    const int j = (int)(gid / 16), f = (int)(gid % 16) * 4;
    if (j >= n) return;
    float acc[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int t = 0; t < kNbr; t++) {
        const int k = nbr[(size_t)j * kNbr + t];
        const float w = wgt[(size_t)j * kNbr + t];
        const bf16x4 h = *reinterpret_cast<const bf16x4 *>(x + (size_t)k * kXW + f);
#pragma unroll
        for (int c = 0; c < 4; c++) acc[c] = fmaf(w, (float)h[c], acc[c]);
    }
    bf16x4 o;
#pragma unroll
    for (int c = 0; c < 4; c++) o[c] = (__bf16)acc[c];
    *reinterpret_cast<bf16x4 *>(x + (size_t)j * kXW + kHid + f) = o;
}

// g = G tanh(ug . H3 + cg), h = G tanh(uh . H3 + ch): 16 lanes per row.
__global__ __launch_bounds__(kBlock) void readout_fwd_kernel(const __bf16 *__restrict__ h3,
                                                             const float *__restrict__ ro,
                                                             float gmax, int n,
                                                             float *__restrict__ gh) {
    const long long gid = blockIdx.x * (long long)kBlock + threadIdx.x;
    const int j = (int)(gid / 16), f = (int)(gid % 16) * 4;
    const bool valid = j < n;
    double sg = 0.0, sh = 0.0;
    if (valid) {
        const bf16x4 hv = *reinterpret_cast<const bf16x4 *>(h3 + (size_t)j * kHid + f);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            sg += (double)((float)hv[q] * ro[f + q]);
            sh += (double)((float)hv[q] * ro[kHid + 1 + f + q]);
        }
    }
    sg = segment_sum<16>(sg) + ro[kHid];
    sh = segment_sum<16>(sh) + ro[2 * kHid + 1];
    if (valid && (threadIdx.x & 15) == 0) {
        gh[j] = (float)(gmax * tanh(sg));
        gh[n + j] = (float)(gmax * tanh(sh));
    }
}

// ---------------------------------------------------------------------------
// Pair loss under the preconditioner of (g, h)
// ---------------------------------------------------------------------------

// out[p] += sum over the symmetric pattern of row j of C_jk q_k v[p][k],
// C = N + N^T, q_k = exp(g_k + h_k): out-edges of j first, then in-edges.
__device__ __forceinline__ void gather_sym(const NetSystem &sys, const float *__restrict__ gh,
                                           int n, int j, const float *__restrict__ v, int pairs,
                                           float (&out)[kPairs]) {
#pragma unroll
    for (int t = 0; t < kNbr; t++) {
        const float c = sys.nval[(size_t)j * kNbr + t];
        if (c == 0.0f) continue;
        const int k = sys.nbr[(size_t)j * kNbr + t];
        const float cq = c * expf(gh[k] + gh[n + k]);
        for (int p = 0; p < pairs; p++) out[p] = fmaf(cq, v[(size_t)p * n + k], out[p]);
    }
    for (int e = sys.tptr[j]; e < sys.tptr[j + 1]; e++) {
        const float c = sys.tval[e];
        if (c == 0.0f) continue;
        const int k = sys.tsrc[e];
        const float cq = c * expf(gh[k] + gh[n + k]);
        for (int p = 0; p < pairs; p++) out[p] = fmaf(cq, v[(size_t)p * n + k], out[p]);
    }
}

// y_p = s2_j r_p - q_j w_p, w_p = (C Q r_p)_j; loss partials; pw = w, pd = dL/dy.
__global__ __launch_bounds__(kBlock) void pair_loss_fwd_kernel(NetSystem sys,
                                                               const float *__restrict__ gh,
                                                               PairSet ps,
                                                               float *__restrict__ pw,
                                                               float *__restrict__ pd,
                                                               double *__restrict__ loss_part) {
    const int n = sys.n;
    double acc = 0.0;
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock) {
        const float g = gh[j], h = gh[n + j];
        const float s2 = (float)(exp(2.0 * g) / sys.adiag[j]);
        const float q = expf(g + h);
        float w[kPairs] = {};
        gather_sym(sys, gh, n, j, ps.r, ps.count, w);
        for (int p = 0; p < ps.count; p++) {
            const size_t at = (size_t)p * n + j;
            const float y = s2 * ps.r[at] - q * w[p];
            const float diff = y - ps.e[at];
            acc += (double)ps.omega[p] * (double)diff * (double)diff;
            pw[at] = w[p];
            pd[at] = 2.0f * ps.omega[p] * diff;
        }
    }
    store_block_partial<kBlock>(acc, loss_part);
}

// dL/dq_j = sum_p (-d_pj w_pj - r_pj (C Q d_p)_j); dL/dg = 2 s2 sum_p d r + q dL/dq;
// dL/dh = q dL/dq.
__global__ __launch_bounds__(kBlock) void pair_loss_bwd_kernel(NetSystem sys,
                                                               const float *__restrict__ gh,
                                                               PairSet ps,
                                                               const float *__restrict__ pw,
                                                               const float *__restrict__ pd,
                                                               float *__restrict__ dgh) {
    const int n = sys.n;
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock) {
        const float g = gh[j], h = gh[n + j];
        const float s2 = (float)(exp(2.0 * g) / sys.adiag[j]);
        const float q = expf(g + h);
        float phi[kPairs] = {};
        gather_sym(sys, gh, n, j, pd, ps.count, phi);
        float dq = 0.0f, ds = 0.0f;
        for (int p = 0; p < ps.count; p++) {
            const size_t at = (size_t)p * n + j;
            dq -= pd[at] * pw[at] + ps.r[at] * phi[p];
            ds = fmaf(pd[at], ps.r[at], ds);
        }
        dgh[j] = 2.0f * s2 * ds + q * dq;
        dgh[n + j] = q * dq;
    }
}

// ---------------------------------------------------------------------------
// Network backward
// ---------------------------------------------------------------------------

// From dL/dg, dL/dh: dZ3 = (dtg ug + dth uh) * (H3 > 0) in BF16, and this
// workgroup's partial readout gradient [ug(64), cg, uh(64), ch].
__global__ __launch_bounds__(kBlock) void readout_bwd_kernel(const __bf16 *__restrict__ h3,
                                                             const float *__restrict__ ro,
                                                             float gmax, int n,
                                                             const float *__restrict__ dgh,
                                                             __bf16 *__restrict__ dz,
                                                             float *__restrict__ rpart) {
    constexpr int kRows = kBlock / 16;
    __shared__ float lds[kRows][kReadout];
    const long long gid = blockIdx.x * (long long)kBlock + threadIdx.x;
    const int j = (int)(gid / 16), f = (int)(gid % 16) * 4;
    const int grp = threadIdx.x / 16;
    const bool valid = j < n;
    float hv[4] = {0.f, 0.f, 0.f, 0.f};
    double sg = 0.0, sh = 0.0;
    if (valid) {
        const bf16x4 h4 = *reinterpret_cast<const bf16x4 *>(h3 + (size_t)j * kHid + f);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            hv[q] = (float)h4[q];
            sg += (double)(hv[q] * ro[f + q]);
            sh += (double)(hv[q] * ro[kHid + 1 + f + q]);
        }
    }
    sg = segment_sum<16>(sg) + ro[kHid];
    sh = segment_sum<16>(sh) + ro[2 * kHid + 1];
    float dtg = 0.f, dth = 0.f;
    if (valid) {
        const double tg = tanh(sg), th = tanh(sh);
        dtg = (float)(dgh[j] * gmax * (1.0 - tg * tg));
        dth = (float)(dgh[n + j] * gmax * (1.0 - th * th));
        bf16x4 o;
#pragma unroll
        for (int q = 0; q < 4; q++)
            o[q] = (__bf16)(hv[q] > 0.f ? dtg * ro[f + q] + dth * ro[kHid + 1 + f + q] : 0.f);
        *reinterpret_cast<bf16x4 *>(dz + (size_t)j * kHid + f) = o;
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
        lds[grp][f + q] = dtg * hv[q];
        lds[grp][kHid + 1 + f + q] = dth * hv[q];
    }
    if ((threadIdx.x & 15) == 0) {
        lds[grp][kHid] = dtg;
        lds[grp][2 * kHid + 1] = dth;
    }
    __syncthreads();
    if (threadIdx.x < kReadout) {
        float s = 0.f;
        for (int r = 0; r < kRows; r++) s += lds[r][threadIdx.x];
        rpart[(size_t)blockIdx.x * kReadout + threadIdx.x] = s;
    }
}

// dX[n x 128] = dZ[n x 64] W^T with W row-major [128 x 64] (BF16); FP32 out.
__global__ __launch_bounds__(kBlock) void gemm_bwd_data_kernel(const __bf16 *__restrict__ dz,
                                                               int n,
                                                               const __bf16 *__restrict__ w,
                                                               float *__restrict__ dx) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    const int row0 = (blockIdx.x * kWavesPerBlock + wave) * 16;
    if (row0 >= n) return;
    const int ar = row0 + (lane & 15);
    const int kq = 8 * (lane >> 4);
    bf16x8 av[2] = {};
    if (ar < n) {
        av[0] = *reinterpret_cast<const bf16x8 *>(dz + (size_t)ar * kHid + kq);
        av[1] = *reinterpret_cast<const bf16x8 *>(dz + (size_t)ar * kHid + 32 + kq);
    }
#pragma unroll
    for (int jb = 0; jb < kXW / 16; jb++) {
        f32x4 acc = {};
#pragma unroll
        for (int ks = 0; ks < 2; ks++) {
            const bf16x8 bv = *reinterpret_cast<const bf16x8 *>(
                w + (size_t)(jb * 16 + (lane & 15)) * kHid + ks * 32 + kq);
            acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[ks], bv, acc, 0, 0, 0);
        }
        const int jcol = jb * 16 + (lane & 15);
#pragma unroll
        for (int t = 0; t < 4; t++) {
            const int row = row0 + 4 * (lane >> 4) + t;
            if (row < n) dx[(size_t)row * kXW + jcol] = acc[t];
        }
    }
}

// dH[k] = dX[k, 0:64] + sum over in-edges (j -> k) of w dX[j, 64:128];
// dZprev = dH * (Hprev > 0) in BF16, Hprev = X[:, 0:64] of this layer's input.
__global__ __launch_bounds__(kBlock) void aggregate_bwd_kernel(
    const float *__restrict__ dx, const int *__restrict__ tptr, const int *__restrict__ tsrc,
    const float *__restrict__ tw, const __bf16 *__restrict__ x, int n,
    __bf16 *__restrict__ dzprev) {
    const long long gid = blockIdx.x * (long long)kBlock + threadIdx.x;
    const int k = (int)(gid / 16), f = (int)(gid % 16) * 4;
    if (k >= n) return;
    const f32x4 own = *reinterpret_cast<const f32x4 *>(dx + (size_t)k * kXW + f);
    float acc[4] = {own[0], own[1], own[2], own[3]};
    for (int e = tptr[k]; e < tptr[k + 1]; e++) {
        const int src = tsrc[e];
        const float w = tw[e];
        const f32x4 g = *reinterpret_cast<const f32x4 *>(dx + (size_t)src * kXW + kHid + f);
#pragma unroll
        for (int c = 0; c < 4; c++) acc[c] = fmaf(w, g[c], acc[c]);
    }
    const bf16x4 hv = *reinterpret_cast<const bf16x4 *>(x + (size_t)k * kXW + f);
    bf16x4 o;
#pragma unroll
    for (int c = 0; c < 4; c++) o[c] = (__bf16)((float)hv[c] > 0.f ? acc[c] : 0.f);
    *reinterpret_cast<bf16x4 *>(dzprev + (size_t)k * kHid + f) = o;
}

// Per-workgroup partial dW[K x 64] = X^T dZ and db[64] = column sums of dZ
// over this workgroup's rows, in chunks of 32 rows staged in LDS; the MFMA
// reduction dimension is the row index.  part holds K*64 + 64 floats per
// workgroup.
template <int K>
__global__ __launch_bounds__(kBlock) void gemm_bwd_weight_kernel(const __bf16 *__restrict__ x,
                                                                 int ldx, int n,
                                                                 const __bf16 *__restrict__ dz,
                                                                 float *__restrict__ part) {
    constexpr int RB = K / 16;  // 16-row blocks of dW
    constexpr int RBW = (RB + kWavesPerBlock - 1) / kWavesPerBlock;
    __shared__ __bf16 xs[32][K + 8];
    __shared__ __bf16 gs[32][kHid + 8];
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    f32x4 acc[RBW][4] = {};
    float db = 0.f;
    for (int r0 = blockIdx.x * 32; r0 < n; r0 += gridDim.x * 32) {
        __syncthreads();
        for (int i = threadIdx.x; i < 32 * K / 8; i += kBlock) {
            const int rr = i / (K / 8), cc = (i % (K / 8)) * 8, row = r0 + rr;
            bf16x8 v = {};
            if (row < n) v = *reinterpret_cast<const bf16x8 *>(x + (size_t)row * ldx + cc);
#pragma unroll
            for (int t = 0; t < 8; t++) xs[rr][cc + t] = v[t];
        }
        for (int i = threadIdx.x; i < 32 * kHid / 8; i += kBlock) {
            const int rr = i / (kHid / 8), cc = (i % (kHid / 8)) * 8, row = r0 + rr;
            bf16x8 v = {};
            if (row < n) v = *reinterpret_cast<const bf16x8 *>(dz + (size_t)row * kHid + cc);
#pragma unroll
            for (int t = 0; t < 8; t++) gs[rr][cc + t] = v[t];
        }
        __syncthreads();
        if (threadIdx.x < kHid)
            for (int rr = 0; rr < 32; rr++) db += (float)gs[rr][threadIdx.x];
        const int kq = 8 * (lane >> 4);
#pragma unroll
        for (int q = 0; q < RBW; q++) {
            const int rb = wave + q * kWavesPerBlock;
            if (rb < RB) {
                bf16x8 av;
#pragma unroll
                for (int t = 0; t < 8; t++) av[t] = xs[kq + t][rb * 16 + (lane & 15)];
#pragma unroll
                for (int jb = 0; jb < 4; jb++) {
                    bf16x8 bv;
#pragma unroll
                    for (int t = 0; t < 8; t++) bv[t] = gs[kq + t][jb * 16 + (lane & 15)];
                    acc[q][jb] =
                        __builtin_amdgcn_mfma_f32_16x16x32_bf16(av, bv, acc[q][jb], 0, 0, 0);
                }
            }
        }
    }
    float *out = part + (size_t)blockIdx.x * (K * kHid + kHid);
#pragma unroll
    for (int q = 0; q < RBW; q++) {
        const int rb = wave + q * kWavesPerBlock;
        if (rb < RB) {
#pragma unroll
            for (int jb = 0; jb < 4; jb++)
#pragma unroll
                for (int t = 0; t < 4; t++)
                    out[(size_t)(rb * 16 + 4 * (lane >> 4) + t) * kHid + jb * 16 + (lane & 15)] =
                        acc[q][jb][t];
        }
    }
    if (threadIdx.x < kHid) out[K * kHid + threadIdx.x] = db;
}

// out[i] = sum over p of part[p * stride + i], i < count, p in increasing
// order within each of kRedWaves interleaved groups, the groups then added in
// order (a fixed order: the same bits every run).  A 1024-thread workgroup
// takes 64 consecutive columns; wavefront w sums the parts p = w, w + 16, ...
// of its lane's column, so the loads are coalesced and 16 wavefronts share a
// column's parts.
constexpr int kRedWaves = 16;
__global__ __launch_bounds__(kRedWaves * kWave) void reduce_parts_kernel(
    const float *__restrict__ part, int parts, int stride, int count, float *__restrict__ out) {
    __shared__ float ws[kRedWaves][kWave];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int i = blockIdx.x * kWave + lane;
    float s = 0.f;
    if (i < count)
        for (int p = wave; p < parts; p += kRedWaves) s += part[(size_t)p * stride + i];
    ws[wave][lane] = s;
    __syncthreads();
    if (wave == 0 && i < count) {
        float t = ws[0][lane];
        for (int w = 1; w < kRedWaves; w++) t += ws[w][lane];
        out[i] = t;
    }
}

hipError_t launch_reduce_parts(const float *part, int parts, int stride, int count, float *out,
                               hipStream_t st) {
    reduce_parts_kernel<<<(count + kWave - 1) / kWave, kRedWaves * kWave, 0, st>>>(part, parts,
                                                                                stride, count, out);
    return hipGetLastError();
}

// Adam step on the FP32 master parameters (bias-corrected moments).
__global__ __launch_bounds__(kBlock) void adam_kernel(float *__restrict__ p,
                                                      const float *__restrict__ g,
                                                      float *__restrict__ m1,
                                                      float *__restrict__ m2, int count, float lr,
                                                      float b1, float b2, float eps, float c1,
                                                      float c2) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < count; i += gridDim.x * kBlock) {
        const float gi = g[i];
        const float a = b1 * m1[i] + (1.f - b1) * gi;
        const float v = b2 * m2[i] + (1.f - b2) * gi * gi;
        m1[i] = a;
        m2[i] = v;
        p[i] -= lr * (a / c1) / (sqrtf(v / c2) + eps);
    }
}

// g *= 1 / den[0] (per-system weighting of a gradient; den > 0).
__global__ __launch_bounds__(kBlock) void scale_inverse_kernel(float *__restrict__ g, int count,
                                                               const double *__restrict__ den) {
    const double d = den[0];
    const float s = d > 0.0 ? (float)(1.0 / d) : 1.0f;
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < count; i += gridDim.x * kBlock) g[i] *= s;
}

// BF16 copies of the layer weights: W (K x 64, row-major) and W^T (64 x K).
__global__ __launch_bounds__(kBlock) void pack_weights_kernel(const float *__restrict__ params,
                                                              __bf16 *__restrict__ wb) {
    for (int i = blockIdx.x * kBlock + threadIdx.x; i < kWtElems; i += gridDim.x * kBlock) {
        int K, base, src;
        if (i < kK0 * kHid) {
            K = kK0;
            base = 0;
            src = kOffW0;
        } else {
            const int l = (i - kK0 * kHid) / (kXW * kHid);
            K = kXW;
            base = kK0 * kHid + l * kXW * kHid;
            src = kOffW1 + l * kLayerStride;
        }
        const int local = i - base;
        const int kin = local / kHid, jo = local % kHid;
        const __bf16 v = (__bf16)params[src + local];
        wb[kWtRowMajor + base + local] = v;           // W[kin][jo]
        wb[kWtTransposed + base + jo * K + kin] = v;  // W^T[jo][kin]
    }
}

// ---------------------------------------------------------------------------
// Preconditioner operator
// ---------------------------------------------------------------------------

// One thread per row: entry 0 is the diagonal exp(2 g_j) / a_jj, the others
// -exp(g_j + h_j + g_k + h_k) N (0 for padding).
__global__ __launch_bounds__(kBlock) void precond_values_kernel(
    int n, const int *__restrict__ off, const int *__restrict__ col,
    const float *__restrict__ src_val, const double *__restrict__ adiag,
    const float *__restrict__ gh, double *__restrict__ val, float *__restrict__ val32) {
    for (int j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock) {
        const double d = adiag[j];
        const double gj = (double)gh[j] + gh[n + j];
        double v = d != 0.0 ? exp(2.0 * (double)gh[j]) / d : 0.0;
        val[off[j]] = v;
        val32[off[j]] = (float)v;
        for (int e = off[j] + 1; e < off[j + 1]; e++) {
            const float sv = src_val[e];
            const int k = col[e];
            v = sv != 0.0f ? -exp(gj + gh[k] + gh[n + k]) * sv : 0.0;
            val[e] = v;
            val32[e] = (float)v;
        }
    }
}

// z = M^-1 r, 4 lanes per row; p = z when p is given; r.z partials.  A row
// holds the diagonal, 4 out-neighbours and its in-neighbours (9 entries on
// average), so a lane's entries e, e + 4, e + 8 are loaded together and one
// round of dependent loads (row bounds, then columns and values, then the
// gathered r) serves the whole row.  The gathers are latency-bound, so the
// workgroups are 1024 wide and the vector kernels' grid (the partial count
// the next kernel expects) holds every row of n = 150000 in one pass.
constexpr int kApplyBlock = 1024;
template <typename T>
__global__ __launch_bounds__(kApplyBlock) void precond_apply_kernel(
    int n, const int *__restrict__ off, const int *__restrict__ col, const T *__restrict__ val,
    const T *__restrict__ r, T *__restrict__ z, T *__restrict__ p, double *__restrict__ rz_part,
    const Scalars *__restrict__ sc) {
    constexpr int kLanes = 4;
    constexpr int kUnroll = 3;
    constexpr int kRows = kApplyBlock / kLanes;
    if (sc && sc->done) return;
    const int sub = threadIdx.x % kLanes;
    double acc = 0.0;
    const long long stride = (long long)gridDim.x * kRows;
    for (long long j = blockIdx.x * (long long)kRows + threadIdx.x / kLanes; j < n;
         j += stride) {
        T zj = 0;
        const int e1 = off[j + 1];
        for (int e = off[j] + sub; e < e1; e += kLanes * kUnroll) {
            T v[kUnroll];
            int c[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; u++) {
                const bool in = e + u * kLanes < e1;
                v[u] = in ? val[e + u * kLanes] : T(0);
                c[u] = in ? col[e + u * kLanes] : (int)j;
            }
#pragma unroll
            for (int u = 0; u < kUnroll; u++) zj += v[u] * r[c[u]];
        }
#pragma unroll
        for (int m = kLanes / 2; m >= 1; m >>= 1) zj += __shfl_xor(zj, m, kWave);
        if (sub == 0) {
            z[j] = zj;
            if (p) p[j] = zj;
            acc += (double)r[j] * (double)zj;
        }
    }
    store_block_partial<kApplyBlock>(acc, rz_part);
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------

hipError_t launch_graph_features(const int *rowstr, const int *colidx, const double *a, int n,
                                 double *adiag, float *feat, int *nbr, float *wgt, float *nval,
                                 hipStream_t st) {
    diag_kernel<<<row_blocks(n), kBlock, 0, st>>>(rowstr, colidx, a, adiag, n);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    graph_features_kernel<<<blocks_for(n, kWavesPerBlock), kBlock, 0, st>>>(
        rowstr, colidx, a, adiag, n, feat, nbr, wgt, nval);
    return hipGetLastError();
}

hipError_t launch_transpose(const int *nbr, const float *wgt, const float *nval, int n, int *cnt,
                            int *tptr, int *tsrc, float *tw, float *tval, hipStream_t st) {
    hipError_t e = hipMemsetAsync(cnt, 0, sizeof(int) * (size_t)n, st);
    if (e != hipSuccess) return e;
    indegree_kernel<<<blocks_for((long long)n * kNbr), kBlock, 0, st>>>(nbr, wgt, n, cnt);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    scan_kernel<<<1, 1024, 0, st>>>(cnt, tptr, n);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = hipMemsetAsync(cnt, 0, sizeof(int) * (size_t)n, st)) != hipSuccess) return e;
    transpose_fill_kernel<<<blocks_for((long long)n * kNbr), kBlock, 0, st>>>(
        nbr, wgt, nval, n, tptr, cnt, tsrc, tw, tval);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    transpose_sort_kernel<<<blocks_for(n), kBlock, 0, st>>>(tptr, tsrc, tw, tval, n);
    return hipGetLastError();
}

// Structure of the operator M^-1 = D (I + G N + N^T H) D as a CSR: row j is
// [j | its K out-neighbours | its in-neighbours (transpose order)], so its
// first entry sits at j (1 + K) + tptr[j] and no scan is needed.  src holds
// the neighbour values (0 for the diagonal slot).  One thread per row, grid
// stride (row n only writes off[n]).
__global__ __launch_bounds__(kBlock) void build_op_kernel(const int *__restrict__ nbr,
                                                          const float *__restrict__ nval,
                                                          const int *__restrict__ tptr,
                                                          const int *__restrict__ tsrc,
                                                          const float *__restrict__ tval, int n,
                                                          int *__restrict__ off,
                                                          int *__restrict__ col,
                                                          float *__restrict__ src) {
    for (int j = blockIdx.x * kBlock + threadIdx.x; j <= n; j += gridDim.x * kBlock) {
        const int t0 = tptr[j];
        const int base = j * (1 + kNbr) + t0;
        off[j] = base;
        if (j == n) break;
        col[base] = j;
        src[base] = 0.0f;
#pragma unroll
        for (int t = 0; t < kNbr; t++) {
            col[base + 1 + t] = nbr[(size_t)j * kNbr + t];
            src[base + 1 + t] = nval[(size_t)j * kNbr + t];
        }
        const int t1 = tptr[j + 1];
        for (int e = t0; e < t1; e++) {
            col[base + 1 + kNbr + e - t0] = tsrc[e];
            src[base + 1 + kNbr + e - t0] = tval[e];
        }
    }
}

hipError_t launch_build_op(const int *nbr, const float *nval, const int *tptr, const int *tsrc,
                           const float *tval, int n, int *off, int *col, float *src,
                           hipStream_t st) {
    build_op_kernel<<<blocks_for((long long)n + 1), kBlock, 0, st>>>(nbr, nval, tptr, tsrc, tval,
                                                                    n, off, col, src);
    return hipGetLastError();
}

hipError_t launch_feature_stats(const float *feat, int n, float *stats, hipStream_t st) {
    feature_stats_kernel<<<1, 1024, 0, st>>>(feat, n, stats);
    return hipGetLastError();
}

hipError_t launch_log_pair(const double *r, const double *z, const double *xstar, int n,
                           float *r_out, float *e_out, double *part, int blocks,
                           hipStream_t st) {
    log_pair_kernel<<<blocks, kBlock, 0, st>>>(r, z, xstar, n, r_out, e_out, part);
    return hipGetLastError();
}

hipError_t launch_pair_weights(const double *part, int parts, int pairs, float *omega,
                               hipStream_t st) {
    pair_weights_kernel<<<1, kBlock, 0, st>>>(part, parts, pairs, omega);
    return hipGetLastError();
}

hipError_t launch_random_rhs(double *b, int n, unsigned long long seed, hipStream_t st) {
    random_rhs_kernel<<<blocks_for(n), kBlock, 0, st>>>(b, n, seed);
    return hipGetLastError();
}

hipError_t launch_sum(const double *part, int count, double *out, hipStream_t st) {
    sum_kernel<<<1, kBlock, 0, st>>>(part, count, out);
    return hipGetLastError();
}

hipError_t launch_net_forward(const NetBuffers &nb, const NetSystem &sys, const float *params,
                              const __bf16 *wb, float gmax, hipStream_t st) {
    const int n = sys.n;
    hipError_t e;
    build_x0_kernel<<<blocks_for((long long)n * kK0), kBlock, 0, st>>>(sys.feat, nb.norm, n,
                                                                      nb.x0);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    const int gb = exact_blocks(n, 16 * kFwdTiles * kWavesPerBlock);
    gemm_fwd_kernel<kK0><<<gb, kBlock, 0, st>>>(nb.x0, kK0, n, wb + kWtTransposed,
                                                params + kOffB0, nb.x[0], kXW);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    for (int l = 0; l < kLayers; l++) {
        aggregate_kernel<<<row_blocks(n), kBlock, 0, st>>>(nb.x[l], sys.nbr, sys.wgt, n);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        __bf16 *out = l + 1 < kLayers ? nb.x[l + 1] : nb.h3;
        const int ldy = l + 1 < kLayers ? kXW : kHid;
        gemm_fwd_kernel<kXW><<<gb, kBlock, 0, st>>>(
            nb.x[l], kXW, n, wb + kWtTransposed + kK0 * kHid + l * kXW * kHid,
            params + kOffB1 + l * kLayerStride, out, ldy);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    readout_fwd_kernel<<<row_blocks(n), kBlock, 0, st>>>(nb.h3, params + kOffR, gmax, n, nb.gh);
    return hipGetLastError();
}

hipError_t launch_pair_loss(const NetBuffers &nb, const NetSystem &sys, const PairSet &ps,
                            int want_grad, hipStream_t st) {
    const int blocks = row_blocks(sys.n);
    pair_loss_fwd_kernel<<<blocks, kBlock, 0, st>>>(sys, nb.gh, ps, nb.pw, nb.pd, nb.loss_part);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || !want_grad) return e;
    pair_loss_bwd_kernel<<<blocks, kBlock, 0, st>>>(sys, nb.gh, ps, nb.pw, nb.pd, nb.dgh);
    return hipGetLastError();
}

hipError_t launch_net_backward(const NetBuffers &nb, const NetSystem &sys, const float *params,
                               const __bf16 *wb, float gmax, float *grad, hipStream_t st) {
    const int n = sys.n;
    hipError_t e;
    const int rb = row_blocks(n);
    readout_bwd_kernel<<<rb, kBlock, 0, st>>>(nb.h3, params + kOffR, gmax, n, nb.dgh, nb.dz,
                                              nb.rpart);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if ((e = launch_reduce_parts(nb.rpart, rb, kReadout, kReadout, grad + kOffR, st)) != hipSuccess)
        return e;
    const int gb = exact_blocks(n, 16 * kWavesPerBlock);  // one 16-row tile per wavefront
    const int wgb = std::min(nb.wparts, blocks_for(n, 32));
    for (int l = kLayers - 1; l >= 0; l--) {
        gemm_bwd_weight_kernel<kXW><<<wgb, kBlock, 0, st>>>(nb.x[l], kXW, n, nb.dz, nb.wpart);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        if ((e = launch_reduce_parts(nb.wpart, wgb, kLayerStride, kLayerStride,
                                     grad + kOffW1 + l * kLayerStride, st)) != hipSuccess)
            return e;
        gemm_bwd_data_kernel<<<gb, kBlock, 0, st>>>(
            nb.dz, n, wb + kWtRowMajor + kK0 * kHid + l * kXW * kHid, nb.dx);
        if ((e = hipGetLastError()) != hipSuccess) return e;
        aggregate_bwd_kernel<<<row_blocks(n), kBlock, 0, st>>>(nb.dx, sys.tptr, sys.tsrc, sys.tw,
                                                               nb.x[l], n, nb.dz);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    gemm_bwd_weight_kernel<kK0><<<wgb, kBlock, 0, st>>>(nb.x0, kK0, n, nb.dz, nb.wpart);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    return launch_reduce_parts(nb.wpart, wgb, kK0 * kHid + kHid, kK0 * kHid + kHid, grad + kOffW0,
                               st);
}

hipError_t launch_adam(float *p, const float *g, float *m1, float *m2, int count, float lr,
                       float b1, float b2, float eps, int step, hipStream_t st) {
    const float c1 = 1.f - powf(b1, (float)step), c2 = 1.f - powf(b2, (float)step);
    adam_kernel<<<blocks_for(count), kBlock, 0, st>>>(p, g, m1, m2, count, lr, b1, b2, eps, c1,
                                                      c2);
    return hipGetLastError();
}

hipError_t launch_scale_inverse(float *g, int count, const double *den, hipStream_t st) {
    scale_inverse_kernel<<<blocks_for(count), kBlock, 0, st>>>(g, count, den);
    return hipGetLastError();
}

hipError_t launch_pack_weights(const float *params, __bf16 *wb, hipStream_t st) {
    pack_weights_kernel<<<blocks_for(kWtElems), kBlock, 0, st>>>(params, wb);
    return hipGetLastError();
}

hipError_t launch_precond_values(const PrecondOp &op, const double *adiag, const float *gh,
                                 hipStream_t st) {
    precond_values_kernel<<<blocks_for(op.n), kBlock, 0, st>>>(op.n, op.off, op.col, op.src_val,
                                                               adiag, gh, op.val, op.val32);
    return hipGetLastError();
}

hipError_t launch_precond_apply(const PrecondOp &op, const double *r, double *z, double *p,
                                double *rz_part, int blocks, const Scalars *sc,
                                hipStream_t st) {
    precond_apply_kernel<double><<<blocks, kApplyBlock, 0, st>>>(op.n, op.off, op.col, op.val, r,
                                                                 z, p, rz_part, sc);
    return hipGetLastError();
}

hipError_t launch_precond_apply32(const PrecondOp &op, const float *r, float *z, float *p,
                                  double *rz_part, int blocks, const Scalars *sc,
                                  hipStream_t st) {
    precond_apply_kernel<float><<<blocks, kApplyBlock, 0, st>>>(op.n, op.off, op.col, op.val32,
                                                                r, z, p, rz_part, sc);
    return hipGetLastError();
}

}  // namespace npbcg_dev
