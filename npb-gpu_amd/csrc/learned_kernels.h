// Learned sparse preconditioner: network shapes, parameter layout, device
// buffers and kernel launchers (learned_kernels.hip).
//
// Preconditioner family (symmetric by construction):
//   M^-1 = S D^-1 S - Q (N + N^T) Q
//   D = diag(a_jj); S = diag(exp g); Q = diag(exp(g + h));
//   N_jk = a_jk / (a_jj a_kk) for the K largest |a_jk| of row j (k != j).
// g = h = 0 is the first-order Neumann approximation of A^-1 restricted to the
// top-K graph; the network predicts the per-row scale corrections g, h.
#ifndef NPBCG_LEARNED_KERNELS_H
#define NPBCG_LEARNED_KERNELS_H

#include <hip/hip_runtime.h>

#include "cg_device.h"

namespace npbcg_dev {

constexpr int kNbr = 4;       // graph neighbours per row (top-|a_jk|)
constexpr int kFeat = 8;      // row features
constexpr int kK0 = 32;       // input width of the embedding GEMM (features, zero-padded)
constexpr int kHid = 64;      // hidden width
constexpr int kXW = 2 * kHid; // [H | aggregated H] input of a message-passing layer
constexpr int kLayers = 3;    // message-passing layers
constexpr int kPairs = 8;     // residual/correction pairs per training system

// FP32 master parameters (one array).
constexpr int kOffW0 = 0;                          // [kK0][kHid]
constexpr int kOffB0 = kOffW0 + kK0 * kHid;        // [kHid]
constexpr int kOffW1 = kOffB0 + kHid;              // layer l: W [kXW][kHid], b [kHid]
constexpr int kLayerStride = kXW * kHid + kHid;
constexpr int kOffB1 = kOffW1 + kXW * kHid;
constexpr int kOffR = kOffW1 + kLayers * kLayerStride;  // readout: ug[64] cg uh[64] ch
constexpr int kReadout = 2 * (kHid + 1);
constexpr int kParams = kOffR + kReadout;

// BF16 weight copies: W row-major (backward data GEMM) and W^T (forward GEMM).
constexpr int kWtElems = kK0 * kHid + kLayers * kXW * kHid;
constexpr int kWtRowMajor = 0;
constexpr int kWtTransposed = kWtElems;

// One matrix as the network and the loss see it (device pointers).
struct NetSystem {
    int n = 0;
    const double *adiag = nullptr;   // a_jj [n]
    const float *feat = nullptr;     // [n][kFeat]
    const int *nbr = nullptr;        // [n][kNbr] (self with zero weight when short)
    const float *wgt = nullptr;      // message weights |a_jk| / sqrt|a_jj a_kk|
    const float *nval = nullptr;     // N_jk = a_jk / (a_jj a_kk) per [n][kNbr] slot
    const int *tptr = nullptr;       // in-edges of every node [n+1] (nonzero weight)
    const int *tsrc = nullptr;       // source row of each in-edge, ascending per node
    const float *tw = nullptr;       // its message weight
    const float *tval = nullptr;     // its N value (N_src,node)
};

// Residual/correction pairs of one system: r_p, e_p = x* - x_p (FP32, [p][n])
// and their loss weights 1 / (P |e_p|^2).
struct PairSet {
    int count = 0;
    const float *r = nullptr;
    const float *e = nullptr;
    const float *omega = nullptr;
};

// Activations, gradients and partial-sum space of one network evaluation.
struct NetBuffers {
    int cap = 0;                 // rows allocated
    __bf16 *x0 = nullptr;        // [cap][kK0]
    __bf16 *x[kLayers] = {};     // [cap][kXW]: [H_l | agg H_l]
    __bf16 *h3 = nullptr;        // [cap][kHid]
    __bf16 *dz = nullptr;        // [cap][kHid]
    float *dx = nullptr;         // [cap][kXW]
    float *gh = nullptr;         // g [cap], h [cap]
    float *dgh = nullptr;        // dL/dg [cap], dL/dh [cap]
    float *pw = nullptr;         // loss forward: (N + N^T) Q r_p per pair [kPairs][cap]
    float *pd = nullptr;         // loss forward: dL/dy per pair [kPairs][cap]
    const float *norm = nullptr; // feature mean [kFeat], std [kFeat]
    float *wpart = nullptr;      // weight-gradient partials [wparts][kXW*kHid + kHid]
    int wparts = 0;
    float *rpart = nullptr;      // readout-gradient partials [row blocks][kReadout]
    double *loss_part = nullptr; // loss partials [row blocks]
};

// Graph, features and transpose (once per matrix).  cnt is n ints of scratch.
hipError_t launch_graph_features(const int *rowstr, const int *colidx, const double *a, int n,
                                 double *adiag, float *feat, int *nbr, float *wgt, float *nval,
                                 hipStream_t st);
hipError_t launch_transpose(const int *nbr, const float *wgt, const float *nval, int n, int *cnt,
                            int *tptr, int *tsrc, float *tw, float *tval, hipStream_t st);

// CSR structure of the operator from the graph: off[n + 1], and per row
// [diagonal | K out-neighbours | in-neighbours] in col, neighbour values in
// src (0 for the diagonal).  off[n] = n (1 + K) + tptr[n].
hipError_t launch_build_op(const int *nbr, const float *nval, const int *tptr, const int *tsrc,
                           const float *tval, int n, int *off, int *col, float *src,
                           hipStream_t st);

// Feature mean/std over the rows of one system: stats[2*kFeat].
hipError_t launch_feature_stats(const float *feat, int n, float *stats, hipStream_t st);

// Pair logging: r_p = r (FP64 CG residual), e_p = xstar - z, |e_p|^2 partials.
hipError_t launch_log_pair(const double *r, const double *z, const double *xstar, int n,
                           float *r_out, float *e_out, double *part, int blocks,
                           hipStream_t st);
// omega_p = 1 / (P * sum of the |e_p|^2 partials of pair p).
hipError_t launch_pair_weights(const double *part, int parts, int pairs, float *omega,
                               hipStream_t st);
hipError_t launch_random_rhs(double *b, int n, unsigned long long seed, hipStream_t st);
int row_blocks(int n);   // workgroups of the 16-lanes-per-row kernels

// Forward through the readout: gh = (g, h).
hipError_t launch_net_forward(const NetBuffers &nb, const NetSystem &sys, const float *params,
                              const __bf16 *wb, float gmax, hipStream_t st);
// Loss of the pairs under the preconditioner of gh; loss_part, pw, pd; then
// dL/dg, dL/dh into dgh when want_grad.
hipError_t launch_pair_loss(const NetBuffers &nb, const NetSystem &sys, const PairSet &ps,
                            int want_grad, hipStream_t st);
// Backward from dgh to the parameter gradient grad[kParams].
hipError_t launch_net_backward(const NetBuffers &nb, const NetSystem &sys, const float *params,
                               const __bf16 *wb, float gmax, float *grad, hipStream_t st);
hipError_t launch_adam(float *p, const float *g, float *m1, float *m2, int count, float lr,
                       float b1, float b2, float eps, int step, hipStream_t st);
hipError_t launch_pack_weights(const float *params, __bf16 *wb, hipStream_t st);
// g[0..count) *= 1 / den[0]
hipError_t launch_scale_inverse(float *g, int count, const double *den, hipStream_t st);
// sum of count doubles in fixed order (one workgroup) -> out[0]
hipError_t launch_sum(const double *part, int count, double *out, hipStream_t st);

// The preconditioner as a CSR operator: row j holds its diagonal first, then
// its K out-neighbours, then its in-neighbours (rows whose top-K contain j).
// src_val holds N for off-diagonal entries (0 for padding self edges).
struct PrecondOp {
    int n = 0;
    const int *off = nullptr;      // [n+1]
    const int *col = nullptr;      // [off[n]]
    const float *src_val = nullptr;
    double *val = nullptr;         // FP64 values
    float *val32 = nullptr;        // FP32 values
};
// Values from gh: diagonal exp(2 g_j) / a_jj, off-diagonal -exp(g_j+h_j+g_k+h_k) N.
hipError_t launch_precond_values(const PrecondOp &op, const double *adiag, const float *gh,
                                 hipStream_t st);
// z = M^-1 r: writes z (and p when p != nullptr) and r.z partials [blocks];
// returns at entry once sc->done is set (sc may be nullptr).
hipError_t launch_precond_apply(const PrecondOp &op, const double *r, double *z, double *p,
                                double *rz_part, int blocks, const Scalars *sc, hipStream_t st);
hipError_t launch_precond_apply32(const PrecondOp &op, const float *r, float *z, float *p,
                                  double *rz_part, int blocks, const Scalars *sc,
                                  hipStream_t st);

}  // namespace npbcg_dev

#endif
