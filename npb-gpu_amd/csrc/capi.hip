// C ABI of npbcg: contexts, device residency, the drop-in conj_grad and the
// device-resident inverse-power loop.  Declarations and reference citations
// are in include/npbcg.h.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <functional>
#include <limits>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "cg_device.h"
#include "cg_kernels.h"
#include "learned_kernels.h"
#include "npbcg.h"
#include "npbcg_internal.h"
#include "trace_range.h"
#include "tiled_layout.h"

using npbcg_dev::Fp64Problem;
using npbcg_dev::Scalars;

namespace npbcg_detail {

namespace {
thread_local std::string g_last_error;
}

int set_error(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

void clear_error() { g_last_error.clear(); }

}  // namespace npbcg_detail

using npbcg_detail::clear_error;
using npbcg_detail::set_error;

#define NPBCG_HIP(call)                                                          \
    do {                                                                         \
        hipError_t _e = (call);                                                  \
        if (_e != hipSuccess)                                                    \
            return set_error(NPBCG_ERR_DEVICE, std::string(#call) + ": " +       \
                                                   hipGetErrorString(_e));       \
    } while (0)

struct npbcg_ctx {
    int device = 0;
    int n = 0;
    long long nnz = 0;
    npbcg_opts opts{};
    int cgitmax = 25;
    int engine = NPBCG_ENGINE_KERNELS;
    hipStream_t stream = nullptr;

    int *d_rowstr = nullptr;
    int *d_colidx = nullptr;
    double *d_a = nullptr;
    int *d_sell_idx = nullptr;     // column-blocked SELL copy: wg_slice, slice_off, slice_row
    double *d_sell_val = nullptr;
    uint16_t *d_sell_col = nullptr;
    double *d_rows_part = nullptr; // nblk * n per-block row sums
    bool cols_sorted = true;
    long long sell_stored = 0;     // SELL entries, padding included
    int sort_window = 0;           // rows sorted together per column block
    double *d_vec = nullptr;   // x, z, p, q, r, p_alt, b, xsol back to back, stride vstride
    size_t vstride = 0;
    double *d_part = nullptr;  // all partial-sum buffers, carved by refresh_problem
    int part_vec = 0, part_spmv = 0;
    Scalars *d_sc = nullptr;
    double *d_hist = nullptr;  // zeta then rnorm, hist_cap each
    int hist_cap = 0;
    Fp64Problem pb{};

    // persistent engine: partials, barrier words, argument block (lazy)
    npbcg_dev::PersistentState ps{};

    // Captured outer iterations: [0] first (x used as is), [1] later ones
    // (x = z/||z|| of the previous iteration applied first).
    hipGraphExec_t outer_graph[2] = {nullptr, nullptr};
    hipGraph_t outer_graph_src[2] = {nullptr, nullptr};

    hipEvent_t ev_start = nullptr, ev_stop = nullptr;
    std::vector<hipEvent_t> probe_start, probe_stop;
    int sample_every = 0;
    long long device_calls = 0;
    npbcg_timing timing{};

    double *h_stage = nullptr;  // pinned host staging, 5 * (n+1) doubles + scalars
    std::vector<int> h_rowstr;  // for the drop-in argument check
    const int *h_colidx_ptr = nullptr;
    const double *h_a_ptr = nullptr;
    uint64_t matrix_digest = 0;  // of colidx and a at create (full check)
    std::vector<int> colidx_sample;
    std::vector<double> a_sample;

    // solve to a tolerance
    double rhs_norm = -1.0;        // ||b|| of the resident right-hand side
    Fp64Problem pbs{};             // solve view: x -> b, z -> solution
    npbcg_dev::MixedProblem mp{};  // FP32 side of the mixed solve (lazy)
    float *d_val32 = nullptr;
    float *d_vec32 = nullptr;      // d, r, p, q (FP32), stride vstride
    int precond = NPBCG_PRECOND_NONE;
    bool learned_ready = false;
    double *d_minv = nullptr;      // Jacobi diagonal (FP64)
    float *d_minv32 = nullptr;     // its FP32 copy
    float *d_rows32 = nullptr;
    double *d_part32 = nullptr;
    Scalars *d_sc32 = nullptr;

    // top-K graph and the sparse preconditioner operator (lazy)
    struct Graph {
        bool graph_built = false;  // device graph (ensure_graph_device)
        bool built = false;        // plus the operator (ensure_row_graph)
        double *adiag = nullptr;
        float *feat = nullptr;
        int *nbr = nullptr;
        float *wgt = nullptr, *nval = nullptr;
        int *tptr = nullptr, *tsrc = nullptr;
        float *tw = nullptr, *tval = nullptr;
        int *cnt = nullptr;
        int *op_off = nullptr, *op_col = nullptr;
        float *op_src = nullptr;
        double *op_val = nullptr;
        float *op_val32 = nullptr;
        float *gh_zero = nullptr, *gh_learned = nullptr;  // (g, h), 2n each
        double *pz = nullptr;    // z = M^-1 r work vectors
        float *pz32 = nullptr;
        size_t slots = 0;             // entries of the operator
        npbcg_dev::PrecondOp op;
        int loaded = -1;         // preconditioner kind whose values op holds
    } g;

    // the preconditioner the solve views point at (select_precond); differs
    // from `precond` while a learned solve is redone with Jacobi
    int active_precond = NPBCG_PRECOND_NONE;

    // step budgets learned from the last converged solve of each (mode,
    // preconditioner in effect) (0: the defaults): FP64 steps of the first
    // batch; mixed inner steps of sweeps 1, 2
    int budget[2][4][2] = {};

    // captured solve sequences, one per mode; replayed while the key matches
    struct SolveGraph {
        hipGraph_t g = nullptr;
        hipGraphExec_t ex = nullptr;
        int kind = -1, maxit = 0, batch = 0;
        double rtol = 0.0, rhs = 0.0;
    } solve_graph[2];

    // residual/correction pairs (lazy)
    struct Pairs {
        int count = 0;
        float *r = nullptr, *e = nullptr, *omega = nullptr;
        double *part = nullptr, *xstar = nullptr;
    } pairs;
};

// Device vectors: x, z, p, q, r of the reference plus p_alt (fused p
// update), b and the solution of npbcg_solve.
constexpr int kNumVectors = 8;
// Column blocks from which the one-reduction form forms r.s in its own pass
// (npbcg_opts.dot_pass = 0).
constexpr int kDotApartBlocks = 24;
// Widest default sort window of the SELL copy (npbcg_opts.sort_window = 0).
constexpr int kMaxSortWindow = 8192;
constexpr int kVecPAlt = 5, kVecB = 6, kVecXsol = 7;

namespace {

int pick_variant(int requested, int n, long long nnz, bool sorted) {
    if (requested >= npbcg_dev::kSpmvSeg64 && requested <= npbcg_dev::kSpmvTiled) {
        if (requested == npbcg_dev::kSpmvTiled && !sorted) return npbcg_dev::kSpmvSeg64;
        return requested;
    }
    const double avg = n > 0 ? (double)nnz / n : 0.0;
    if (sorted && n >= 16384 && avg >= 16.0) return npbcg_dev::kSpmvTiled;
    if (avg >= 96.0) return npbcg_dev::kSpmvSeg64;
    if (avg >= 40.0) return npbcg_dev::kSpmvSeg32;
    return npbcg_dev::kSpmvSeg16;
}

double *vec_ptr(npbcg_ctx *c, int which) { return c->d_vec + (size_t)which * c->vstride; }

void refresh_problem(npbcg_ctx *c) {
    Fp64Problem &pb = c->pb;
    pb.n = c->n;
    pb.nnz = c->nnz;
    pb.rowstr = c->d_rowstr;
    pb.colidx = c->d_colidx;
    pb.a = c->d_a;
    pb.x = vec_ptr(c, NPBCG_VEC_X);
    pb.z = vec_ptr(c, NPBCG_VEC_Z);
    pb.p = vec_ptr(c, NPBCG_VEC_P);
    pb.q = vec_ptr(c, NPBCG_VEC_Q);
    pb.r = vec_ptr(c, NPBCG_VEC_R);
    pb.p_alt = vec_ptr(c, kVecPAlt);
    double *pp = c->d_part;
    const size_t nv = c->part_vec, ns = c->part_spmv;
    pb.part.xx = pp;
    pb.part.rr[0] = pp + nv;
    pb.part.rr[1] = pp + 2 * nv;
    pb.part.nrm = pp + 3 * nv;           // 2 * nv
    pb.part.pq = pp + 5 * nv;            // ns
    pb.part.res = pp + 5 * nv + ns;      // 3 * ns
    pb.part.rz[0] = pp + 5 * nv + 4 * ns;
    pb.part.rz[1] = pp + 6 * nv + 4 * ns;
    pb.sc = c->d_sc;
    pb.zeta_hist = c->d_hist;
    pb.rnorm_hist = c->d_hist ? c->d_hist + c->hist_cap : nullptr;
    pb.hist_cap = c->hist_cap;
}

// The solve view of the problem: the right-hand side sits where conj_grad
// keeps x and the iterate where it keeps z.
void refresh_solve_view(npbcg_ctx *c) {
    c->pbs = c->pb;
    c->pbs.x = vec_ptr(c, kVecB);
    c->pbs.z = vec_ptr(c, kVecXsol);
    c->pbs.fuse_p_update = 0;  // a solve may stop after any step; p stays in place
    c->pbs.cg_form = 0;        // the stop test sits in the reference-order p update
}

void drop_graph(npbcg_ctx *c) {
    for (int i = 0; i < 2; i++) {
        if (c->outer_graph[i]) (void)hipGraphExecDestroy(c->outer_graph[i]);
        if (c->outer_graph_src[i]) (void)hipGraphDestroy(c->outer_graph_src[i]);
        c->outer_graph[i] = nullptr;
        c->outer_graph_src[i] = nullptr;
    }
}

int check_csr(int n, const int *rowstr, const int *colidx, long long *nnz_out,
              bool *sorted_out) {
    if (rowstr[0] != 0) return set_error(NPBCG_ERR_INVALID, "rowstr[0] must be 0");
    for (int j = 0; j < n; j++)
        if (rowstr[j + 1] < rowstr[j])
            return set_error(NPBCG_ERR_INVALID, "rowstr must be non-decreasing");
    const long long nnz = rowstr[n];
    bool sorted = true;
    for (int j = 0; j < n; j++)
        for (long long k = rowstr[j]; k < rowstr[j + 1]; k++) {
            if (colidx[k] < 0 || colidx[k] >= n)
                return set_error(NPBCG_ERR_INVALID, "colidx entry outside [0, n)");
            if (k > rowstr[j] && colidx[k] <= colidx[k - 1]) sorted = false;
        }
    *nnz_out = nnz;
    *sorted_out = sorted;
    return NPBCG_OK;
}

// Digest of `bytes` bytes at p (whole 8-byte words, then the tail): four
// independent lanes h = ((h ^ w) * K) ^ (h >> 29) over the words i = lane
// (mod 4), folded in lane order.  Each step is a bijection of h for a fixed
// word and an injection of the word for a fixed h, so a change of any single
// word always changes the digest.
uint64_t digest_bytes(const unsigned char *p, size_t bytes, uint64_t seed) {
    constexpr uint64_t K = 0x9E3779B97F4A7C15ull;
    uint64_t h[4] = {seed ^ 0x243F6A8885A308D3ull, seed ^ 0x13198A2E03707344ull,
                     seed ^ 0xA4093822299F31D0ull, seed ^ 0x082EFA98EC4E6C89ull};
    auto mix = [](uint64_t x, uint64_t w) {
        x = (x ^ w) * K;
        return x ^ (x >> 29);
    };
    const size_t words = bytes / 8;
    size_t i = 0;
    for (; i + 4 <= words; i += 4) {
        uint64_t w[4];
        std::memcpy(w, p + 8 * i, 32);
        h[0] = mix(h[0], w[0]);
        h[1] = mix(h[1], w[1]);
        h[2] = mix(h[2], w[2]);
        h[3] = mix(h[3], w[3]);
    }
    for (; i < words; i++) {
        uint64_t w;
        std::memcpy(&w, p + 8 * i, 8);
        h[i & 3] = mix(h[i & 3], w);
    }
    uint64_t tail = 0;
    std::memcpy(&tail, p + 8 * words, bytes - 8 * words);
    uint64_t out = mix(mix(0, bytes), tail);
    for (int k = 0; k < 4; k++) out = mix(out, h[k]);
    return out;
}

// Digest of colidx[0..nnz) and a[0..nnz): the arrays are cut into up to 16
// equal chunks hashed by parallel host threads (one thread below 4 MB); the
// chunk digests are folded in chunk order, so the value does not depend on
// the thread count.
uint64_t matrix_digest(const int *colidx, const double *a, long long nnz) {
    if (nnz <= 0) return 0;
    const unsigned char *src[2] = {reinterpret_cast<const unsigned char *>(colidx),
                                   reinterpret_cast<const unsigned char *>(a)};
    const size_t len[2] = {sizeof(int) * (size_t)nnz, sizeof(double) * (size_t)nnz};
    constexpr int kChunks = 16;
    uint64_t part[2][kChunks] = {};
    auto chunk = [&](int arr, int k) {
        const size_t per = ((len[arr] + kChunks - 1) / kChunks + 7) & ~(size_t)7;
        const size_t b = std::min(len[arr], per * (size_t)k);
        const size_t e = std::min(len[arr], b + per);
        part[arr][k] = digest_bytes(src[arr] + b, e - b, (uint64_t)(arr * kChunks + k));
    };
    const size_t total = len[0] + len[1];
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int nthreads = total < ((size_t)4 << 20) ? 1 : (int)std::min<unsigned>(16, hw);
    if (nthreads == 1) {
        for (int arr = 0; arr < 2; arr++)
            for (int k = 0; k < kChunks; k++) chunk(arr, k);
    } else {
        std::vector<std::thread> pool;
        for (int t = 0; t < nthreads; t++)
            pool.emplace_back([&, t]() {
                for (int w = t; w < 2 * kChunks; w += nthreads) chunk(w / kChunks, w % kChunks);
            });
        for (auto &th : pool) th.join();
    }
    uint64_t out = 0;
    for (int arr = 0; arr < 2; arr++)
        for (int k = 0; k < kChunks; k++) {
            out = (out ^ part[arr][k]) * 0x9E3779B97F4A7C15ull;
            out ^= out >> 31;
        }
    return out;
}

// Samples used to check that the arrays handed to npbcg_conj_grad are the
// ones the context holds (rowstr is compared in full).
void take_samples(npbcg_ctx *c, const int *colidx, const double *a) {
    c->colidx_sample.clear();
    c->a_sample.clear();
    const long long step = std::max<long long>(1, c->nnz / 61);
    for (long long k = 0; k < c->nnz; k += step) {
        c->colidx_sample.push_back(colidx[k]);
        c->a_sample.push_back(a[k]);
    }
    if (c->nnz > 0) {
        c->colidx_sample.push_back(colidx[c->nnz - 1]);
        c->a_sample.push_back(a[c->nnz - 1]);
    }
}

// Whether the host arrays describe the matrix the context holds.  rowstr is
// compared in full.  With the full check (the default) colidx and a are
// digested on every call, so an in-place edit between calls is seen, as the
// reference's serial conj_grad reads its arrays on every call (cg.cu:703-709);
// with the sampled check the caller keeps the reference GPU path's contract
// (the matrix is uploaded once by setup_gpu, cg.cu:1390-1603, and never
// edited) and only the pointers and 62 sampled entries are compared.
bool same_matrix(const npbcg_ctx *c, const int *rowstr, const int *colidx,
                 const double *a) {
    if (std::memcmp(rowstr, c->h_rowstr.data(), sizeof(int) * (c->n + 1)) != 0)
        return false;
    if (c->opts.matrix_check == NPBCG_CHECK_FULL)
        return matrix_digest(colidx, a, c->nnz) == c->matrix_digest;
    if (colidx == c->h_colidx_ptr && a == c->h_a_ptr) return true;
    const long long step = std::max<long long>(1, c->nnz / 61);
    size_t i = 0;
    for (long long k = 0; k < c->nnz; k += step, i++)
        if (colidx[k] != c->colidx_sample[i] ||
            std::memcmp(&a[k], &c->a_sample[i], sizeof(double)) != 0)
            return false;
    return true;
}

int ensure_history(npbcg_ctx *c, int niter) {
    if (niter <= c->hist_cap) return NPBCG_OK;
    drop_graph(c);  // the captured graph holds the old history pointers
    if (c->d_hist) (void)hipFree(c->d_hist);
    c->d_hist = nullptr;
    int cap = std::max(niter, 128);
    NPBCG_HIP(hipMalloc(&c->d_hist, sizeof(double) * 2 * (size_t)cap));
    c->hist_cap = cap;
    refresh_problem(c);
    refresh_solve_view(c);
    return NPBCG_OK;
}

int capture_outer(npbcg_ctx *c, int scale_x, hipGraph_t *g_out, hipGraphExec_t *ex_out) {
    NPBCG_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e =
        npbcg_dev::enqueue_conj_grad(c->pb, c->cgitmax, scale_x, 1, c->stream, nullptr);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(c->stream, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return set_error(NPBCG_ERR_DEVICE,
                         std::string("graph capture failed: ") +
                             hipGetErrorString(e != hipSuccess ? e : e2));
    }
    hipGraphExec_t ex = nullptr;
    hipError_t e3 = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    if (e3 != hipSuccess) {
        (void)hipGraphDestroy(g);
        return set_error(NPBCG_ERR_DEVICE,
                         std::string("hipGraphInstantiate: ") + hipGetErrorString(e3));
    }
    *g_out = g;
    *ex_out = ex;
    return NPBCG_OK;
}

int ensure_graph(npbcg_ctx *c) {
    for (int i = 0; i < 2; i++) {
        if (c->outer_graph[i]) continue;
        int rc = capture_outer(c, i, &c->outer_graph_src[i], &c->outer_graph[i]);
        if (rc != NPBCG_OK) return rc;
    }
    return NPBCG_OK;
}

int ensure_probes(npbcg_ctx *c) {
    const size_t want = (size_t)c->cgitmax;
    while (c->probe_start.size() < want) {
        hipEvent_t a, b;
        NPBCG_HIP(hipEventCreateWithFlags(&a, hipEventDisableSystemFence));
        NPBCG_HIP(hipEventCreateWithFlags(&b, hipEventDisableSystemFence));
        c->probe_start.push_back(a);
        c->probe_stop.push_back(b);
    }
    return NPBCG_OK;
}

// niter outer iterations through the persistent kernel, in launches short
// enough that the barrier counter cannot wrap, then the abort word: nonzero
// means a grid barrier timed out and the results are not valid.
int run_persistent(npbcg_ctx *c, int niter, int scale_first, int record, int final_scale) {
    const long long per_iter = npbcg_dev::persistent_barriers(c->cgitmax) * c->pb.sell.grid;
    const int chunk = (int)std::max<long long>(1, std::min<long long>(niter, (1LL << 31) / per_iter));
    for (int done = 0; done < niter; done += chunk) {
        const int k = std::min(chunk, niter - done);
        NPBCG_HIP(npbcg_dev::launch_persistent(c->pb, c->ps, k, (done > 0 || scale_first) ? 1 : 0,
                                               record, (final_scale && done + k == niter) ? 1 : 0,
                                               c->stream));
    }
    unsigned words[2] = {0, 0};
    NPBCG_HIP(hipMemcpyAsync(c->h_stage, c->ps.bar, sizeof(words), hipMemcpyDeviceToHost,
                             c->stream));
    NPBCG_HIP(hipStreamSynchronize(c->stream));
    std::memcpy(words, c->h_stage, sizeof(words));
    if (words[1] != 0)
        return set_error(NPBCG_ERR_DEVICE,
                         "persistent conj_grad: a grid barrier timed out (results invalid)");
    return NPBCG_OK;
}

int read_scalars(npbcg_ctx *c, Scalars *out) {
    NPBCG_HIP(hipMemcpyAsync(c->h_stage, c->d_sc, sizeof(Scalars), hipMemcpyDeviceToHost,
                             c->stream));
    NPBCG_HIP(hipStreamSynchronize(c->stream));
    std::memcpy(out, c->h_stage, sizeof(Scalars));
    return NPBCG_OK;
}

}  // namespace

extern "C" {

void npbcg_default_opts(npbcg_opts *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->device = 0;
    o->cgitmax = 25;
    o->engine = NPBCG_ENGINE_AUTO;
    o->spmv_variant = 0;
    o->use_graph = 1;
}

int npbcg_abi_version(void) { return NPBCG_ABI_VERSION; }

int npbcg_device_count(int *count) {
    if (!count) return set_error(NPBCG_ERR_INVALID, "npbcg_device_count: NULL argument");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess) ndev = 0;
    *count = ndev;
    return NPBCG_OK;
}

const char *npbcg_last_error(void) { return npbcg_detail::g_last_error.c_str(); }

const char *npbcg_status_string(int s) {
    switch (s) {
    case NPBCG_OK: return "ok";
    case NPBCG_ERR_INVALID: return "invalid argument";
    case NPBCG_ERR_NOMEM: return "out of memory";
    case NPBCG_ERR_DEVICE: return "HIP runtime error";
    case NPBCG_ERR_NODEVICE: return "no GPU device";
    case NPBCG_ERR_CAPACITY: return "matrix capacity exceeded";
    case NPBCG_ERR_INTERNAL: return "internal generator error";
    case NPBCG_ERR_MISMATCH: return "arrays do not match the context's matrix";
    case NPBCG_ERR_NOCONV: return "solve did not converge";
    default: return "unknown status";
    }
}

int npbcg_create(npbcg_ctx **out, int n, const int *rowstr, const int *colidx,
                 const double *a, const npbcg_opts *opts) {
    npbcg_detail::TraceRange trace("npbcg_create");
    if (!out) return set_error(NPBCG_ERR_INVALID, "npbcg_create: out is NULL");
    *out = nullptr;
    if (n < 1 || !rowstr || !colidx || !a)
        return set_error(NPBCG_ERR_INVALID, "npbcg_create: invalid argument");
    long long nnz = 0;
    bool sorted = true;
    int rc = check_csr(n, rowstr, colidx, &nnz, &sorted);
    if (rc != NPBCG_OK) return rc;

    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return set_error(NPBCG_ERR_NODEVICE, "npbcg_create: no GPU device visible");

    npbcg_ctx *c = new (std::nothrow) npbcg_ctx();
    if (!c) return set_error(NPBCG_ERR_NOMEM, "npbcg_create: host allocation");
    npbcg_default_opts(&c->opts);
    if (opts) c->opts = *opts;
    c->device = (c->opts.device >= 0 && c->opts.device < ndev) ? c->opts.device : 0;
    c->cgitmax = c->opts.cgitmax > 0 ? c->opts.cgitmax : 25;
    if (c->opts.engine < NPBCG_ENGINE_AUTO || c->opts.engine > NPBCG_ENGINE_PERSISTENT) {
        delete c;
        return set_error(NPBCG_ERR_INVALID, "npbcg_create: unknown engine");
    }
    if (c->opts.dot_pass < 0 || c->opts.dot_pass > 2) {
        delete c;
        return set_error(NPBCG_ERR_INVALID, "npbcg_create: dot_pass must be 0, 1 or 2");
    }
    if (c->opts.sort_window < 0 || c->opts.sort_window % npbcg_detail::kSliceRows) {
        delete c;
        return set_error(NPBCG_ERR_INVALID,
                         "npbcg_create: sort_window must be 0 or a positive multiple of 64");
    }
    // AUTO takes the graph-replayed kernels: 8.5 ms against 14.6 ms for the
    // persistent kernel over 10 outer iterations at n = 75000, whose grid
    // barriers each write back and invalidate every XCD's L2
    c->engine = c->opts.engine == NPBCG_ENGINE_PERSISTENT ? NPBCG_ENGINE_PERSISTENT
                                                          : NPBCG_ENGINE_KERNELS;
    c->n = n;
    c->nnz = nnz;
    c->cols_sorted = sorted;

    auto fail = [&](int code) {
        npbcg_destroy(c);
        return code;
    };
#define CREATE_HIP(call)                                                          \
    do {                                                                          \
        hipError_t _e = (call);                                                   \
        if (_e != hipSuccess)                                                     \
            return fail(set_error(_e == hipErrorOutOfMemory ? NPBCG_ERR_NOMEM     \
                                                            : NPBCG_ERR_DEVICE,   \
                                  std::string(#call) + ": " + hipGetErrorString(_e))); \
    } while (0)

    CREATE_HIP(hipSetDevice(c->device));
    CREATE_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    CREATE_HIP(hipEventCreateWithFlags(&c->ev_start, hipEventDisableSystemFence));
    CREATE_HIP(hipEventCreateWithFlags(&c->ev_stop, hipEventDisableSystemFence));

    hipDeviceProp_t prop;
    CREATE_HIP(hipGetDeviceProperties(&prop, c->device));
    const int cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;

    int variant = pick_variant(c->opts.spmv_variant, n, nnz, sorted);
    const int rpb = npbcg_dev::spmv_rows_per_block(variant);
    const int vec_blocks = npbcg_dev::vector_blocks(n, cus);
    const int spmv_blocks = std::max(1, std::min((n + rpb - 1) / rpb, 4 * cus));

    // Column-blocked SELL copy for the LDS-staged SpMV.  One column block of
    // the vector must fit the CU's LDS: the widest blocks (fewest partial
    // sums) take one 1024-thread workgroup per CU; blocks that fit twice use
    // 512-thread workgroups, two per CU.  Row ranges spread the
    // nblk * nrange workgroups evenly over the CUs.
    npbcg_detail::SellBlocks<double> sm;
    npbcg_dev::SellDev sd;
    if (variant == npbcg_dev::kSpmvTiled) {
        int lds_cap = 0;
        if (hipDeviceGetAttribute(&lds_cap, hipDeviceAttributeSharedMemPerBlockOptin,
                                  c->device) != hipSuccess || lds_cap <= 0)
            lds_cap = 64 * 1024;
        const int lds_budget = lds_cap - 2048;  // static LDS of the reduction epilogue
        int W;
        if (c->opts.tile_width > 0) {
            W = c->opts.tile_width;
        } else {
            const int wmax = std::min(65536, lds_budget / 8);
            const int nblk = (n + wmax - 1) / wmax;
            W = (n + nblk - 1) / nblk;
        }
        W = std::max(2, (W + 1) & ~1);
        if (W > n) W = (n + 1) & ~1;
        int block = (W * 8 <= lds_budget / 2 - 1024) ? 512 : 1024;
        int nblk = (n + W - 1) / W;
        auto default_rows = [&]() {
            const int slots = cus * (block == 512 ? 2 : 1);
            const int nrange = std::max(1, slots / nblk);
            return (int)(((long long)n + nrange - 1) / nrange);
        };
        int R = c->opts.tile_rows > 0 ? c->opts.tile_rows : default_rows();
        // The per-slice p.q slots (one double per 64-row slice of a workgroup)
        // share the LDS with the vector block.  When the range's slices do not
        // fit beside the block (large n: few ranges of many rows), the block
        // is narrowed to 7/8 of the LDS and the range capped at the slices
        // the rest holds: more, shorter workgroups instead of the CSR kernel.
        const int kSlice = npbcg_detail::kSliceRows;
        if ((long long)(W + (R + kSlice - 1) / kSlice) * 8 > lds_budget) {
            if (c->opts.tile_width <= 0) {
                const int wcap = (lds_budget / 8) * 7 / 8;
                if (W > wcap) {
                    const int nb = (n + wcap - 1) / wcap;
                    W = std::max(2, ((n + nb - 1) / nb + 1) & ~1);
                    block = (W * 8 <= lds_budget / 2 - 1024) ? 512 : 1024;
                    nblk = (n + W - 1) / W;
                    if (c->opts.tile_rows <= 0) R = default_rows();
                }
            }
            if (c->opts.tile_rows <= 0) {
                const int rmax = kSlice * (lds_budget / 8 - W - 1);
                if (rmax >= kSlice && R > rmax) {
                    // The capped ranges give nblk * nrange workgroups for
                    // `slots` CU slots; of the range counts from the least
                    // the cap allows up to twice that, take the one whose
                    // last round of workgroups fills the most slots (equal
                    // workgroups: a part-filled round leaves CUs idle).  At
                    // 1.5 M rows: 85 blocks x 12 ranges = 3.98 rounds of 256
                    // instead of 85 x 10 = 3.32.
                    const int slots = cus * (block == 512 ? 2 : 1);
                    const int nr0 = (n + rmax - 1) / rmax;
                    int best = nr0;
                    double best_idle = 2.0;
                    for (int nr = nr0; nr <= 2 * nr0; nr++) {
                        const double rounds = (double)nblk * nr / slots;
                        const double idle = (std::ceil(rounds) - rounds) / std::ceil(rounds);
                        if (idle < best_idle - 1e-9) {
                            best_idle = idle;
                            best = nr;
                        }
                    }
                    R = (int)(((long long)n + best - 1) / best);
                }
            }
        }
        // Rows are sorted by their length in a block over the whole range,
        // from 2048 up to 8192 rows: at n = 150000 (ranges of 4688 rows) that
        // stores 1.019 entries per nonzero instead of 1.048 with 2048-row
        // windows, at 1.5 M rows 1.016 instead of 1.058.
        const int window =
            c->opts.sort_window > 0
                ? c->opts.sort_window
                : std::min(kMaxSortWindow, std::max(npbcg_detail::kSortWindow,
                                                    (R + kSlice - 1) / kSlice * kSlice));
        if (W * 8 > lds_budget || !sorted ||
            !npbcg_detail::build_sell_blocks(n, rowstr, colidx, a, W, R, &sm, window,
                                             block / 64)) {
            variant = pick_variant(npbcg_dev::kSpmvSeg64, n, nnz, false);
        } else {
            c->sort_window = window;
            sd.W = sm.W;
            sd.nblk = sm.nblk;
            sd.R = sm.R;
            sd.nrange = sm.nrange;
            sd.grid = sm.nblk * sm.nrange;
            sd.block = block;
            int max_slices = 0;
            for (size_t w = 0; w + 1 < sm.wg_slice.size(); w++)
                max_slices = std::max(max_slices, sm.wg_slice[w + 1] - sm.wg_slice[w]);
            sd.lds_bytes = (sm.W + max_slices) * 8;
            if (sd.lds_bytes > lds_budget) variant = pick_variant(npbcg_dev::kSpmvSeg64, n, nnz, false);
        }
    }

    c->vstride = ((size_t)n + 2 + 63) & ~(size_t)63;
    CREATE_HIP(hipMalloc(&c->d_rowstr, sizeof(int) * ((size_t)n + 1)));
    CREATE_HIP(hipMalloc(&c->d_colidx, sizeof(int) * std::max<long long>(1, nnz)));
    CREATE_HIP(hipMalloc(&c->d_a, sizeof(double) * std::max<long long>(1, nnz)));
    CREATE_HIP(hipMalloc(&c->d_vec, sizeof(double) * kNumVectors * c->vstride));
    c->part_vec = vec_blocks;
    c->part_spmv = variant == npbcg_dev::kSpmvTiled ? std::max(sd.grid, vec_blocks)
                                                    : spmv_blocks;
    CREATE_HIP(hipMalloc(&c->d_part,
                         sizeof(double) * (7 * (size_t)c->part_vec + 4 * (size_t)c->part_spmv)));
    if (variant == npbcg_dev::kSpmvTiled) {
        CREATE_HIP(npbcg_dev::prepare_sell_kernels(sd.lds_bytes));
        const size_t nwg1 = sm.wg_slice.size(), nsl1 = sm.slice_off.size();
        const size_t nrow = sm.slice_row.size();
        const size_t idx = nwg1 + nsl1 + nrow + sm.row_slot.size();
        CREATE_HIP(hipMalloc(&c->d_sell_idx, sizeof(int) * idx));
        CREATE_HIP(hipMalloc(&c->d_sell_val, sizeof(double) * std::max<size_t>(1, sm.val.size())));
        CREATE_HIP(hipMalloc(&c->d_sell_col,
                             sizeof(uint16_t) * std::max<size_t>(1, sm.col.size())));
        // one block sum per slot (64 per slice), stored by slot
        CREATE_HIP(hipMalloc(&c->d_rows_part, sizeof(double) * std::max<size_t>(1, nrow)));
        CREATE_HIP(hipMemcpy(c->d_sell_idx, sm.wg_slice.data(), sizeof(int) * nwg1,
                             hipMemcpyHostToDevice));
        CREATE_HIP(hipMemcpy(c->d_sell_idx + nwg1, sm.slice_off.data(), sizeof(int) * nsl1,
                             hipMemcpyHostToDevice));
        CREATE_HIP(hipMemcpy(c->d_sell_idx + nwg1 + nsl1, sm.slice_row.data(),
                             sizeof(int) * nrow, hipMemcpyHostToDevice));
        CREATE_HIP(hipMemcpy(c->d_sell_idx + nwg1 + nsl1 + nrow, sm.row_slot.data(),
                             sizeof(int) * sm.row_slot.size(), hipMemcpyHostToDevice));
        if (!sm.val.empty()) {
            CREATE_HIP(hipMemcpy(c->d_sell_val, sm.val.data(), sizeof(double) * sm.val.size(),
                                 hipMemcpyHostToDevice));
            CREATE_HIP(hipMemcpy(c->d_sell_col, sm.col.data(),
                                 sizeof(uint16_t) * sm.col.size(), hipMemcpyHostToDevice));
        }
        sd.wg_slice = c->d_sell_idx;
        sd.slice_off = c->d_sell_idx + nwg1;
        sd.slice_row = c->d_sell_idx + nwg1 + nsl1;
        sd.row_slot = c->d_sell_idx + nwg1 + nsl1 + nrow;
        sd.slots = (long long)nrow;
        sd.val = c->d_sell_val;
        sd.col = c->d_sell_col;
        sd.rows_part = c->d_rows_part;
        c->sell_stored = sm.stored;
    }
    CREATE_HIP(hipMalloc(&c->d_sc, npbcg_dev::scalars_bytes()));
    CREATE_HIP(hipHostMalloc(&c->h_stage,
                             sizeof(double) * (5 * ((size_t)n + 1)) + 4096));

    CREATE_HIP(hipMemcpyAsync(c->d_rowstr, rowstr, sizeof(int) * ((size_t)n + 1),
                              hipMemcpyHostToDevice, c->stream));
    if (nnz > 0) {
        CREATE_HIP(hipMemcpyAsync(c->d_colidx, colidx, sizeof(int) * nnz,
                                  hipMemcpyHostToDevice, c->stream));
        CREATE_HIP(hipMemcpyAsync(c->d_a, a, sizeof(double) * nnz, hipMemcpyHostToDevice,
                                  c->stream));
    }
    CREATE_HIP(hipMemsetAsync(c->d_vec, 0, sizeof(double) * kNumVectors * c->vstride, c->stream));
    CREATE_HIP(hipMemsetAsync(c->d_sc, 0, npbcg_dev::scalars_bytes(), c->stream));

    refresh_problem(c);
    c->pb.vec_blocks = vec_blocks;
    c->pb.spmv_blocks = spmv_blocks;
    c->pb.spmv_variant = variant;
    c->pb.fuse_p_update = c->opts.fuse_p_update ? 1 : 0;
    // one-reduction CG form: LDS-staged layout (s = A r goes to its block row
    // partials) and no fused p update
    c->pb.cg_form = (variant == npbcg_dev::kSpmvTiled && !c->pb.fuse_p_update &&
                     c->opts.cg_form != 1) ? 1 : 0;
    // r.s in a pass of its own once each row has only a few entries per
    // column block (NPB class D: 85 blocks of about 5 entries per row)
    c->pb.dot_apart = c->pb.cg_form && (c->opts.dot_pass == 2 ||
                                        (c->opts.dot_pass == 0 && sd.nblk >= kDotApartBlocks))
                          ? 1 : 0;
    c->pb.sell = sd;
    refresh_solve_view(c);
    if (c->engine == NPBCG_ENGINE_PERSISTENT) {
        // one cooperative launch per call: every workgroup of the SELL grid must
        // be resident at once, or the launch is refused here rather than later
        if (variant != npbcg_dev::kSpmvTiled)
            return fail(set_error(NPBCG_ERR_INVALID,
                                  "npbcg_create: the persistent engine needs the LDS-staged "
                                  "layout (spmv_variant 5 with sorted columns)"));
        int coop = 0, per_cu = 0, clock_khz = 0;
        CREATE_HIP(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, c->device));
        CREATE_HIP(npbcg_dev::persistent_capacity(c->pb, &per_cu));
        if (!coop || (long long)per_cu * cus < sd.grid)
            return fail(set_error(NPBCG_ERR_INVALID,
                                  "npbcg_create: the persistent engine's " +
                                      std::to_string(sd.grid) +
                                      " workgroups cannot all be resident (" +
                                      std::to_string(per_cu) + " per CU)"));
        CREATE_HIP(hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeWallClockRate, c->device));
        CREATE_HIP(hipMalloc(&c->ps.part, sizeof(double) * npbcg_dev::persistent_partials(c->pb)));
        CREATE_HIP(hipMalloc(&c->ps.bar, 2 * sizeof(unsigned)));
        CREATE_HIP(hipMalloc(&c->ps.args, npbcg_dev::persistent_args_bytes()));
        // a barrier that waits 2 s has lost a workgroup: the kernel gives up
        c->ps.spin_limit = (long long)std::max(clock_khz, 1000) * 2000;
        c->ps.cgitmax = c->cgitmax;
    }
    CREATE_HIP(npbcg_dev::launch_fill(c->pb.x, n + 1, 1.0, vec_blocks, c->stream));
    CREATE_HIP(hipStreamSynchronize(c->stream));
#undef CREATE_HIP

    c->h_rowstr.assign(rowstr, rowstr + n + 1);
    c->h_colidx_ptr = colidx;
    c->h_a_ptr = a;
    take_samples(c, colidx, a);
    if (c->opts.matrix_check == NPBCG_CHECK_FULL) c->matrix_digest = matrix_digest(colidx, a, nnz);
    c->timing.n = n;
    c->timing.nnz = nnz;
    c->timing.engine = c->engine;
    *out = c;
    clear_error();
    return NPBCG_OK;
}

void npbcg_destroy(npbcg_ctx *c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graph(c);
    for (auto &sg : c->solve_graph) {
        if (sg.ex) (void)hipGraphExecDestroy(sg.ex);
        if (sg.g) (void)hipGraphDestroy(sg.g);
    }
    for (auto e : c->probe_start) (void)hipEventDestroy(e);
    for (auto e : c->probe_stop) (void)hipEventDestroy(e);
    if (c->ev_start) (void)hipEventDestroy(c->ev_start);
    if (c->ev_stop) (void)hipEventDestroy(c->ev_stop);
    (void)hipFree(c->d_rowstr);
    (void)hipFree(c->d_colidx);
    (void)hipFree(c->d_a);
    (void)hipFree(c->d_sell_idx);
    (void)hipFree(c->d_sell_val);
    (void)hipFree(c->d_sell_col);
    (void)hipFree(c->d_rows_part);
    (void)hipFree(c->d_val32);
    (void)hipFree(c->d_minv);
    (void)hipFree(c->d_minv32);
    for (void *p : {(void *)c->g.adiag, (void *)c->g.feat, (void *)c->g.nbr, (void *)c->g.wgt,
                    (void *)c->g.nval, (void *)c->g.tptr, (void *)c->g.tsrc, (void *)c->g.tw,
                    (void *)c->g.tval, (void *)c->g.cnt, (void *)c->g.op_off, (void *)c->g.op_col,
                    (void *)c->g.op_src, (void *)c->g.op_val, (void *)c->g.op_val32,
                    (void *)c->g.gh_zero, (void *)c->g.gh_learned, (void *)c->g.pz,
                    (void *)c->g.pz32, (void *)c->pairs.r, (void *)c->pairs.e,
                    (void *)c->pairs.omega, (void *)c->pairs.part, (void *)c->pairs.xstar})
        (void)hipFree(p);
    (void)hipFree(c->d_vec32);
    (void)hipFree(c->d_rows32);
    (void)hipFree(c->d_part32);
    (void)hipFree(c->d_sc32);
    (void)hipFree(c->d_vec);
    (void)hipFree(c->d_part);
    (void)hipFree(c->d_sc);
    (void)hipFree(c->d_hist);
    (void)hipFree(c->ps.part);
    (void)hipFree(c->ps.bar);
    (void)hipFree(c->ps.args);
    if (c->h_stage) (void)hipHostFree(c->h_stage);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int npbcg_conj_grad_device(npbcg_ctx *c, double *rnorm) {
    npbcg_detail::TraceRange trace("npbcg_conj_grad");
    if (!c) return set_error(NPBCG_ERR_INVALID, "npbcg_conj_grad_device: NULL context");
    NPBCG_HIP(hipSetDevice(c->device));
    NPBCG_HIP(hipEventRecord(c->ev_start, c->stream));
    if (c->engine == NPBCG_ENGINE_PERSISTENT) {
        int rc = run_persistent(c, 1, 0, 0, 0);
        if (rc != NPBCG_OK) return rc;
    } else {
        NPBCG_HIP(npbcg_dev::enqueue_conj_grad(c->pb, c->cgitmax, 0, 0, c->stream, nullptr));
    }
    NPBCG_HIP(hipEventRecord(c->ev_stop, c->stream));
    Scalars s;
    int rc = read_scalars(c, &s);
    if (rc != NPBCG_OK) return rc;
    float ms = 0.f;
    NPBCG_HIP(hipEventElapsedTime(&ms, c->ev_start, c->ev_stop));
    c->timing.total_ms = ms;
    c->timing.spmv_samples = 0;
    c->timing.spmv_avg_ms = 0.0;
    c->timing.launches = c->engine == NPBCG_ENGINE_PERSISTENT
                             ? 2
                             : npbcg_dev::conj_grad_launches(c->pb, c->cgitmax);
    c->timing.engine = c->engine;
    if (rnorm) *rnorm = s.rnorm;
    clear_error();
    return NPBCG_OK;
}

int npbcg_conj_grad(npbcg_ctx *c, int colidx[], int rowstr[], double x[], double z[],
                    double a[], double p[], double q[], double r[], double *rnorm) {
    if (!c || !colidx || !rowstr || !x || !z || !a || !p || !q || !r || !rnorm)
        return set_error(NPBCG_ERR_INVALID, "npbcg_conj_grad: NULL argument");
    if (!same_matrix(c, rowstr, colidx, a))
        return set_error(NPBCG_ERR_MISMATCH,
                         "npbcg_conj_grad: rowstr/colidx/a differ from the context's matrix");
    NPBCG_HIP(hipSetDevice(c->device));
    const size_t bytes = sizeof(double) * ((size_t)c->n + 1);
    std::memcpy(c->h_stage, x, bytes);
    NPBCG_HIP(hipMemcpyAsync(c->pb.x, c->h_stage, bytes, hipMemcpyHostToDevice, c->stream));
    double rn = 0.0;
    int rc = npbcg_conj_grad_device(c, &rn);
    if (rc != NPBCG_OK) return rc;
    double *outs[4] = {z, p, q, r};
    const int which[4] = {NPBCG_VEC_Z, NPBCG_VEC_P, NPBCG_VEC_Q, NPBCG_VEC_R};
    for (int i = 0; i < 4; i++)
        NPBCG_HIP(hipMemcpyAsync(c->h_stage + (size_t)i * (c->n + 1), vec_ptr(c, which[i]),
                                 bytes, hipMemcpyDeviceToHost, c->stream));
    NPBCG_HIP(hipStreamSynchronize(c->stream));
    for (int i = 0; i < 4; i++)
        std::memcpy(outs[i], c->h_stage + (size_t)i * (c->n + 1), bytes);
    *rnorm = rn;
    clear_error();
    return NPBCG_OK;
}

int npbcg_norms(npbcg_ctx *c, double *norm_temp1, double *norm_temp2) {
    if (!c) return set_error(NPBCG_ERR_INVALID, "npbcg_norms: NULL context");
    NPBCG_HIP(hipSetDevice(c->device));
    NPBCG_HIP(npbcg_dev::launch_norms(c->pb, c->stream));
    Scalars s;
    int rc = read_scalars(c, &s);
    if (rc != NPBCG_OK) return rc;
    if (norm_temp1) *norm_temp1 = s.xz;
    if (norm_temp2) *norm_temp2 = s.zz;
    clear_error();
    return NPBCG_OK;
}

int npbcg_scale_x(npbcg_ctx *c, double scale) {
    if (!c) return set_error(NPBCG_ERR_INVALID, "npbcg_scale_x: NULL context");
    NPBCG_HIP(hipSetDevice(c->device));
    NPBCG_HIP(npbcg_dev::launch_scale_x(c->pb, scale, 0, c->stream));
    NPBCG_HIP(hipStreamSynchronize(c->stream));
    clear_error();
    return NPBCG_OK;
}

int npbcg_power(npbcg_ctx *c, int niter, double shift, double *zeta_out,
                double *rnorm_out) {
    npbcg_detail::TraceRange trace("npbcg_power");
    if (!c || niter < 0) return set_error(NPBCG_ERR_INVALID, "npbcg_power: invalid argument");
    NPBCG_HIP(hipSetDevice(c->device));
    int rc = ensure_history(c, std::max(niter, 1));
    if (rc != NPBCG_OK) return rc;
    if (c->engine == NPBCG_ENGINE_PERSISTENT) {
        // the whole loop in cooperative launches; no SpMV sampling inside one
        c->device_calls++;
        NPBCG_HIP(npbcg_dev::launch_reset_scalars(c->pb, shift, c->stream));
        NPBCG_HIP(hipEventRecord(c->ev_start, c->stream));
        if (niter > 0) {
            rc = run_persistent(c, niter, 0, 1, 1);
            if (rc != NPBCG_OK) return rc;
        }
        NPBCG_HIP(hipEventRecord(c->ev_stop, c->stream));
        NPBCG_HIP(hipStreamSynchronize(c->stream));
        if (niter > 0 && zeta_out)
            NPBCG_HIP(hipMemcpy(zeta_out, c->d_hist, sizeof(double) * niter, hipMemcpyDeviceToHost));
        if (niter > 0 && rnorm_out)
            NPBCG_HIP(hipMemcpy(rnorm_out, c->d_hist + c->hist_cap, sizeof(double) * niter,
                                hipMemcpyDeviceToHost));
        float ms = 0.f;
        NPBCG_HIP(hipEventElapsedTime(&ms, c->ev_start, c->ev_stop));
        c->timing.total_ms = ms;
        c->timing.launches = 2 * std::max(niter, 1) + 1;
        c->timing.spmv_samples = 0;
        c->timing.spmv_avg_ms = 0.0;
        c->timing.engine = c->engine;
        clear_error();
        return NPBCG_OK;
    }
    const bool use_graph = c->opts.use_graph != 0;
    if (use_graph) {
        rc = ensure_graph(c);
        if (rc != NPBCG_OK) return rc;
    }
    const bool sample = c->sample_every > 0 && (c->device_calls % c->sample_every) == 0;
    c->device_calls++;
    npbcg_dev::SpmvProbe probe{nullptr, nullptr, 0};
    if (sample) {
        rc = ensure_probes(c);
        if (rc != NPBCG_OK) return rc;
        probe = {c->probe_start.data(), c->probe_stop.data(), c->cgitmax};
    }

    NPBCG_HIP(npbcg_dev::launch_reset_scalars(c->pb, shift, c->stream));
    NPBCG_HIP(hipEventRecord(c->ev_start, c->stream));
    int launches = 0;
    for (int it = 0; it < niter; it++) {
        const int scale = it > 0 ? 1 : 0;
        if (sample && it == 0) {
            NPBCG_HIP(npbcg_dev::enqueue_conj_grad(c->pb, c->cgitmax, scale, 1, c->stream,
                                                   &probe));
            launches += npbcg_dev::conj_grad_launches(c->pb, c->cgitmax);
        } else if (use_graph) {
            NPBCG_HIP(hipGraphLaunch(c->outer_graph[scale], c->stream));
            launches += 1;
        } else {
            NPBCG_HIP(npbcg_dev::enqueue_conj_grad(c->pb, c->cgitmax, scale, 1, c->stream,
                                                   nullptr));
            launches += npbcg_dev::conj_grad_launches(c->pb, c->cgitmax);
        }
    }
    if (niter > 0) NPBCG_HIP(npbcg_dev::launch_scale_x(c->pb, 0.0, 1, c->stream));
    NPBCG_HIP(hipEventRecord(c->ev_stop, c->stream));
    NPBCG_HIP(hipStreamSynchronize(c->stream));
    if (niter > 0 && zeta_out)
        NPBCG_HIP(hipMemcpy(zeta_out, c->d_hist, sizeof(double) * niter,
                            hipMemcpyDeviceToHost));
    if (niter > 0 && rnorm_out)
        NPBCG_HIP(hipMemcpy(rnorm_out, c->d_hist + c->hist_cap, sizeof(double) * niter,
                            hipMemcpyDeviceToHost));
    float ms = 0.f;
    NPBCG_HIP(hipEventElapsedTime(&ms, c->ev_start, c->ev_stop));
    c->timing.total_ms = ms;
    c->timing.launches = launches + 2;
    c->timing.engine = c->engine;
    c->timing.spmv_samples = 0;
    c->timing.spmv_avg_ms = 0.0;
    if (sample && niter > 0) {
        double sum = 0.0;
        for (int i = 0; i < probe.count; i++) {
            float t = 0.f;
            NPBCG_HIP(hipEventElapsedTime(&t, c->probe_start[i], c->probe_stop[i]));
            sum += t;
        }
        c->timing.spmv_samples = probe.count;
        c->timing.spmv_avg_ms = sum / probe.count;
    }
    clear_error();
    return NPBCG_OK;
}

int npbcg_profile_kernels(npbcg_ctx *c, int niter, double shift, double *zeta_out,
                          npbcg_kernel_time *out, int max_kinds, int *nkinds,
                          double *total_ms) {
    npbcg_detail::TraceRange trace("npbcg_profile_kernels");
    if (!c || niter < 1 || !out || max_kinds < 1 || !nkinds)
        return set_error(NPBCG_ERR_INVALID, "npbcg_profile_kernels: invalid argument");
    NPBCG_HIP(hipSetDevice(c->device));
    int rc = ensure_history(c, niter);
    if (rc != NPBCG_OK) return rc;
    NPBCG_HIP(npbcg_dev::launch_reset_scalars(c->pb, shift, c->stream));
    npbcg_dev::KernelMarks marks;
    Fp64Problem pb = c->pb;
    pb.marks = &marks;
    marks.stamp(c->stream);
    hipError_t e = marks.err;
    for (int it = 0; it < niter && e == hipSuccess; it++)
        e = npbcg_dev::enqueue_conj_grad(pb, c->cgitmax, it > 0 ? 1 : 0, 1, c->stream, nullptr);
    if (e == hipSuccess) e = npbcg_dev::launch_scale_x(pb, 0.0, 1, c->stream);
    if (e == hipSuccess) e = marks.err;
    NPBCG_HIP(e);
    NPBCG_HIP(hipStreamSynchronize(c->stream));
    if (zeta_out)
        NPBCG_HIP(hipMemcpy(zeta_out, c->d_hist, sizeof(double) * niter, hipMemcpyDeviceToHost));
    // one entry per kernel name, in order of first launch
    std::vector<npbcg_kernel_time> kinds;
    double total = 0.0;
    for (size_t i = 0; i < marks.name.size(); i++) {
        float t = 0.f;
        NPBCG_HIP(hipEventElapsedTime(&t, marks.ev[i], marks.ev[i + 1]));
        total += t;
        size_t k = 0;
        while (k < kinds.size() && std::strcmp(kinds[k].name, marks.name[i]) != 0) k++;
        if (k == kinds.size()) {
            npbcg_kernel_time kt{};
            std::snprintf(kt.name, sizeof(kt.name), "%s", marks.name[i]);
            kt.threads = marks.threads[i];
            kinds.push_back(kt);
        }
        kinds[k].calls++;
        kinds[k].ms += t;
    }
    *nkinds = (int)kinds.size();
    if ((int)kinds.size() > max_kinds)
        return set_error(NPBCG_ERR_INVALID, "npbcg_profile_kernels: more kernels than max_kinds");
    std::copy(kinds.begin(), kinds.end(), out);
    if (total_ms) *total_ms = total;
    c->timing.total_ms = total;
    c->timing.launches = (int)marks.name.size() + 1;
    c->timing.spmv_samples = 0;
    c->timing.spmv_avg_ms = 0.0;
    clear_error();
    return NPBCG_OK;
}

int npbcg_set_vector(npbcg_ctx *c, int which, const double *host) {
    if (!c || !host || which < NPBCG_VEC_X || which > NPBCG_VEC_R)
        return set_error(NPBCG_ERR_INVALID, "npbcg_set_vector: invalid argument");
    NPBCG_HIP(hipSetDevice(c->device));
    const size_t bytes = sizeof(double) * ((size_t)c->n + 1);
    std::memcpy(c->h_stage, host, bytes);
    NPBCG_HIP(hipMemcpyAsync(vec_ptr(c, which), c->h_stage, bytes, hipMemcpyHostToDevice,
                             c->stream));
    NPBCG_HIP(hipStreamSynchronize(c->stream));
    clear_error();
    return NPBCG_OK;
}

int npbcg_get_vector(npbcg_ctx *c, int which, double *host) {
    if (!c || !host || which < NPBCG_VEC_X || which > NPBCG_VEC_R)
        return set_error(NPBCG_ERR_INVALID, "npbcg_get_vector: invalid argument");
    NPBCG_HIP(hipSetDevice(c->device));
    const size_t bytes = sizeof(double) * ((size_t)c->n + 1);
    NPBCG_HIP(hipMemcpyAsync(c->h_stage, vec_ptr(c, which), bytes, hipMemcpyDeviceToHost,
                             c->stream));
    NPBCG_HIP(hipStreamSynchronize(c->stream));
    std::memcpy(host, c->h_stage, bytes);
    clear_error();
    return NPBCG_OK;
}

int npbcg_fill_x(npbcg_ctx *c, double value) {
    if (!c) return set_error(NPBCG_ERR_INVALID, "npbcg_fill_x: NULL context");
    NPBCG_HIP(hipSetDevice(c->device));
    NPBCG_HIP(npbcg_dev::launch_fill(c->pb.x, c->n + 1, value, c->pb.vec_blocks, c->stream));
    NPBCG_HIP(hipStreamSynchronize(c->stream));
    clear_error();
    return NPBCG_OK;
}

int npbcg_set_spmv_sampling(npbcg_ctx *c, int every_nth_call) {
    if (!c || every_nth_call < 0)
        return set_error(NPBCG_ERR_INVALID, "npbcg_set_spmv_sampling: invalid argument");
    c->sample_every = every_nth_call;
    c->device_calls = 0;
    return NPBCG_OK;
}

int npbcg_get_timing(const npbcg_ctx *c, npbcg_timing *out) {
    if (!c || !out) return set_error(NPBCG_ERR_INVALID, "npbcg_get_timing: NULL argument");
    *out = c->timing;
    return NPBCG_OK;
}

int npbcg_describe(const npbcg_ctx *c, char *buf, int len) {
    if (!c || (!buf && len > 0) || len < 0)
        return set_error(NPBCG_ERR_INVALID, "npbcg_describe: invalid argument");
    const npbcg_dev::SellDev &sd = c->pb.sell;
    hipDeviceProp_t prop{};
    if (hipGetDeviceProperties(&prop, c->device) != hipSuccess) std::strcpy(prop.name, "unknown");
    prop.name[sizeof(prop.name) - 1] = '\0';
    prop.gcnArchName[sizeof(prop.gcnArchName) - 1] = '\0';
    if (prop.name[0] == '\0') std::strncpy(prop.name, prop.gcnArchName, sizeof(prop.name) - 1);
    for (char *q = prop.name; *q; q++)
        if (*q == '"' || *q == '\\') *q = ' ';
    char tmp[768];
    int k = std::snprintf(
        tmp, sizeof tmp,
        "{\"n\": %d, \"nnz\": %lld, \"spmv_variant\": %d, \"vec_blocks\": %d, "
        "\"csr_blocks\": %d, \"block_width\": %d, \"col_blocks\": %d, "
        "\"range_rows\": %d, \"row_ranges\": %d, \"sell_workgroups\": %d, "
        "\"sell_threads\": %d, \"sell_stored\": %lld, \"sort_window\": %d, \"cgitmax\": %d, "
        "\"use_graph\": %d, \"fuse_p_update\": %d, \"cg_form\": %d, \"dot_pass\": %d, "
        "\"engine\": %d, \"device_name\": \"%s\", \"compute_units\": %d}",
        c->n, c->nnz, c->pb.spmv_variant, c->pb.vec_blocks, c->pb.spmv_blocks, sd.W,
        sd.nblk, sd.R, sd.nrange, sd.grid, sd.block, c->sell_stored, c->sort_window, c->cgitmax,
        c->opts.use_graph, c->pb.fuse_p_update, c->pb.cg_form ? 2 : 1,
        c->pb.dot_apart ? 2 : 1, c->engine, prop.name,
        prop.multiProcessorCount);
    if (len > 0) {
        std::strncpy(buf, tmp, (size_t)len - 1);
        buf[len - 1] = '\0';
    }
    return k;
}

int npbcg_spmv_bench(npbcg_ctx *c, int reps, double *avg_ms) {
    if (!c || reps < 1 || !avg_ms)
        return set_error(NPBCG_ERR_INVALID, "npbcg_spmv_bench: invalid argument");
    NPBCG_HIP(hipSetDevice(c->device));
    NPBCG_HIP(npbcg_dev::launch_spmv_q(c->pb, c->stream));  // warm
    NPBCG_HIP(hipEventRecord(c->ev_start, c->stream));
    for (int i = 0; i < reps; i++) NPBCG_HIP(npbcg_dev::launch_spmv_q(c->pb, c->stream));
    NPBCG_HIP(hipEventRecord(c->ev_stop, c->stream));
    NPBCG_HIP(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    NPBCG_HIP(hipEventElapsedTime(&ms, c->ev_start, c->ev_stop));
    *avg_ms = ms / reps;
    clear_error();
    return NPBCG_OK;
}

}  // extern "C"

namespace {

// Record key of a single-system tuning: the order, the nonzeros, an FNV-1a
// digest of the row lengths and the CU count.
std::string tune_key(int n, const int *rowstr, int cus) {
    uint64_t h = 1469598103934665603ull;
    for (int j = 0; j < n; j++) {
        uint32_t len = (uint32_t)(rowstr[j + 1] - rowstr[j]);
        for (int k = 0; k < 4; k++) {
            h ^= (len >> (8 * k)) & 0xFFu;
            h *= 1099511628211ull;
        }
    }
    char buf[160];
    std::snprintf(buf, sizeof buf, "npbcg1:n=%d:nnz=%d:rows=%016llx:cus=%d", n, rowstr[n],
                  (unsigned long long)h, cus);
    return buf;
}

// "<key> <variant> <width> <rows> <ms>" per line; the last record of a key wins.
bool tune_lookup(const char *path, const std::string &key, int *variant, int *w, int *r,
                 double *ms) {
    if (!path) return false;
    FILE *f = std::fopen(path, "r");
    if (!f) return false;
    char k[512];
    int v = 0, a = 0, b = 0;
    double t = 0.0;
    bool found = false;
    while (std::fscanf(f, "%511s %d %d %d %lf", k, &v, &a, &b, &t) == 5)
        if (key == k) {
            *variant = v;
            *w = a;
            *r = b;
            *ms = t;
            found = true;
        }
    std::fclose(f);
    return found;
}

void tune_store(const char *path, const std::string &key, int variant, int w, int r, double ms) {
    if (!path) return;
    FILE *f = std::fopen(path, "a");
    if (!f) return;
    std::fprintf(f, "%s %d %d %d %.6f\n", key.c_str(), variant, w, r, ms);
    std::fclose(f);
}

}  // namespace

extern "C" {

int npbcg_tune(int n, const int *rowstr, const int *colidx, const double *a, npbcg_opts *opts,
               const char *cache_path, double *ms_out) {
    npbcg_detail::TraceRange trace("npbcg_tune");
    if (n < 1 || !rowstr || !colidx || !a || !opts)
        return set_error(NPBCG_ERR_INVALID, "npbcg_tune: invalid argument");
    const npbcg_opts base = *opts;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1)
        return set_error(NPBCG_ERR_NODEVICE, "npbcg_tune: no GPU");
    const int device = (base.device >= 0 && base.device < count) ? base.device : 0;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        cus = 0;
    const std::string key = tune_key(n, rowstr, cus);
    int best_v = 0, best_w = 0, best_r = 0;
    double best = 0.0;
    if (!tune_lookup(cache_path, key, &best_v, &best_w, &best_r, &best)) {
        best = std::numeric_limits<double>::infinity();
        // Times one shape through a temporary context; a request the context
        // does not honour (e.g. the LDS-staged layout on unsorted rows) is
        // skipped.  W and R receive the LDS-staged shape actually built.
        auto time_one = [&](int variant, int w, int r, int *W, int *R) -> int {
            npbcg_opts o = base;
            o.spmv_variant = variant;
            o.tile_width = w;
            o.tile_rows = r;
            npbcg_ctx *c = nullptr;
            int rc = npbcg_create(&c, n, rowstr, colidx, a, &o);
            if (rc != NPBCG_OK) return rc;
            double t = std::numeric_limits<double>::infinity();
            if (c->pb.spmv_variant == variant) {
                for (int trial = 0; trial < 2 && rc == NPBCG_OK; trial++) {
                    double ms = 0.0;
                    rc = npbcg_spmv_bench(c, 20, &ms);
                    t = std::min(t, ms);
                }
            }
            const int bw = variant == npbcg_dev::kSpmvTiled ? c->pb.sell.W : 0;
            const int br = variant == npbcg_dev::kSpmvTiled ? c->pb.sell.R : 0;
            npbcg_destroy(c);
            if (rc != NPBCG_OK) return rc;
            if (W) *W = bw;
            if (R) *R = br;
            if (t < best) {
                best = t;
                best_v = variant;
                best_w = bw;
                best_r = br;
            }
            return NPBCG_OK;
        };
        // the LDS-staged layout at its default shape, with twice the row
        // ranges, and with half-width column blocks (two workgroups per CU)
        int W0 = 0, R0 = 0;
        int rc = time_one(npbcg_dev::kSpmvTiled, 0, 0, &W0, &R0);
        if (rc != NPBCG_OK) return rc;
        if (W0 > 0) {
            if (R0 > 1) {
                rc = time_one(npbcg_dev::kSpmvTiled, W0, (R0 + 1) / 2, nullptr, nullptr);
                if (rc != NPBCG_OK) return rc;
            }
            const int nblk = (n + W0 - 1) / W0;
            const int w2 = (((n + 2 * nblk - 1) / (2 * nblk)) + 1) & ~1;
            if (w2 >= 2 && w2 < W0) {
                rc = time_one(npbcg_dev::kSpmvTiled, w2, 0, nullptr, nullptr);
                if (rc != NPBCG_OK) return rc;
            }
        }
        for (int v : {npbcg_dev::kSpmvSeg64, npbcg_dev::kSpmvSeg32, npbcg_dev::kSpmvSeg16,
                      npbcg_dev::kSpmvSeg64U2}) {
            rc = time_one(v, 0, 0, nullptr, nullptr);
            if (rc != NPBCG_OK) return rc;
        }
        tune_store(cache_path, key, best_v, best_w, best_r, best);
    }
    opts->spmv_variant = best_v;
    opts->tile_width = best_w;
    opts->tile_rows = best_r;
    if (ms_out) *ms_out = best;
    clear_error();
    return NPBCG_OK;
}


int npbcg_copy_bench(npbcg_ctx *c, long long bytes, int reps, double *avg_ms) {
    return npbcg_stream_bench(c, 0, bytes, reps, avg_ms);
}

int npbcg_stream_bench(npbcg_ctx *c, int mode, long long bytes, int reps, double *avg_ms) {
    if (!c || bytes < 16 || reps < 1 || !avg_ms || (mode != 0 && mode != 1))
        return set_error(NPBCG_ERR_INVALID, "npbcg_stream_bench: invalid argument");
    NPBCG_HIP(hipSetDevice(c->device));
    bytes &= ~15LL;
    void *src = nullptr, *dst = nullptr;
    hipError_t e = hipMalloc(&src, (size_t)bytes);
    if (e == hipSuccess) e = hipMalloc(&dst, (size_t)bytes);
    if (e == hipSuccess) e = hipMemsetAsync(src, 0, (size_t)bytes, c->stream);
    hipDeviceProp_t prop;
    if (e == hipSuccess) e = hipGetDeviceProperties(&prop, c->device);
    const int cus = (e == hipSuccess && prop.multiProcessorCount > 0) ? prop.multiProcessorCount
                                                                       : 256;
    auto run = [&]() {
        return mode == 0 ? npbcg_dev::launch_copy(src, dst, bytes, 8 * cus, c->stream)
                         : npbcg_dev::launch_read(src, bytes, static_cast<double *>(dst), cus,
                                                  c->stream);
    };
    if (e == hipSuccess) e = run();
    if (e == hipSuccess) e = hipEventRecord(c->ev_start, c->stream);
    for (int i = 0; i < reps && e == hipSuccess; i++) e = run();
    if (e == hipSuccess) e = hipEventRecord(c->ev_stop, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, c->ev_start, c->ev_stop);
    (void)hipFree(src);
    (void)hipFree(dst);
    if (e != hipSuccess)
        return set_error(e == hipErrorOutOfMemory ? NPBCG_ERR_NOMEM : NPBCG_ERR_DEVICE,
                         std::string("npbcg_stream_bench: ") + hipGetErrorString(e));
    *avg_ms = ms / reps;
    clear_error();
    return NPBCG_OK;
}

// ---- solve to a tolerance ---------------------------------------------------

namespace {

int ensure_mixed(npbcg_ctx *c) {
    if (c->d_val32) return NPBCG_OK;
    const npbcg_dev::SellDev &sd = c->pb.sell;
    const size_t stored = (size_t)std::max<long long>(1, c->sell_stored);
    const size_t vb = (size_t)c->pb.vec_blocks;
    NPBCG_HIP(hipMalloc(&c->d_val32, sizeof(float) * stored));
    NPBCG_HIP(hipMalloc(&c->d_vec32, sizeof(float) * 5 * c->vstride));
    NPBCG_HIP(hipMalloc(&c->d_rows32, sizeof(float) * (size_t)sd.nblk * c->n));
    NPBCG_HIP(hipMalloc(&c->d_part32, sizeof(double) * (5 * vb + (size_t)sd.grid)));
    NPBCG_HIP(hipMalloc(&c->d_sc32, npbcg_dev::scalars_bytes()));
    NPBCG_HIP(hipMemsetAsync(c->d_vec32, 0, sizeof(float) * 5 * c->vstride, c->stream));
    NPBCG_HIP(hipMemsetAsync(c->d_sc32, 0, npbcg_dev::scalars_bytes(), c->stream));
    NPBCG_HIP(npbcg_dev::launch_to_fp32(c->d_sell_val, c->d_val32, c->sell_stored, c->stream));
    // FP32 vector block (padded to 4 entries) + one FP64 slot per slice, and
    // the FP32 row sums of one row range when the range is wider than one
    // 2048-row sort window and they fit beside them.  A narrower range keeps
    // each slice's rows within 2048 rows, and its sums are stored straight
    // away (staging them: 18.3 against 18.0 us at n = 75000; at n = 150000,
    // sorted over its 4688-row ranges, 35.9 against 41.0 us).
    const int slices_lds = (sd.lds_bytes / 8) - sd.W;  // max slices per workgroup
    int lds32 = ((sd.W + 3) & ~3) * 4 + slices_lds * 8;
    int lds_cap = 0;
    if (hipDeviceGetAttribute(&lds_cap, hipDeviceAttributeSharedMemPerBlockOptin, c->device) !=
            hipSuccess ||
        lds_cap <= 0)
        lds_cap = 64 * 1024;
    const int range_rows = std::min(sd.R, c->n);
    int stage_rows = 0;
    if (range_rows > npbcg_detail::kSortWindow &&
        (long long)lds32 + 4LL * range_rows <= lds_cap - 2048) {
        stage_rows = sd.R;
        lds32 += 4 * range_rows;
    }
    NPBCG_HIP(npbcg_dev::prepare_mixed_kernels(lds32));
    npbcg_dev::MixedProblem &mp = c->mp;
    mp.val = c->d_val32;
    mp.d = c->d_vec32;
    mp.r = c->d_vec32 + c->vstride;
    mp.p = c->d_vec32 + 2 * c->vstride;
    mp.q = c->d_vec32 + 3 * c->vstride;
    mp.p_alt = c->d_vec32 + 4 * c->vstride;
    mp.rows_part = c->d_rows32;
    mp.part_xx = c->d_part32;
    mp.part_rr[0] = c->d_part32 + vb;
    mp.part_rr[1] = c->d_part32 + 2 * vb;
    mp.part_pq = c->d_part32 + 3 * vb;
    mp.part_rz[0] = c->d_part32 + 3 * vb + sd.grid;
    mp.part_rz[1] = c->d_part32 + 4 * vb + sd.grid;
    mp.sc = c->d_sc32;
    mp.lds_bytes = lds32;
    mp.stage_rows = stage_rows;
    NPBCG_HIP(hipStreamSynchronize(c->stream));
    return NPBCG_OK;
}

int read_scalars_from(npbcg_ctx *c, const Scalars *dev, Scalars *out) {
    NPBCG_HIP(hipMemcpyAsync(c->h_stage, dev, sizeof(Scalars), hipMemcpyDeviceToHost,
                             c->stream));
    NPBCG_HIP(hipStreamSynchronize(c->stream));
    std::memcpy(out, c->h_stage, sizeof(Scalars));
    return NPBCG_OK;
}

}  // namespace

int npbcg_set_rhs(npbcg_ctx *c, const double *b) {
    if (!c || !b) return set_error(NPBCG_ERR_INVALID, "npbcg_set_rhs: NULL argument");
    NPBCG_HIP(hipSetDevice(c->device));
    double s = 0.0;
    for (int j = 0; j < c->n; j++) s += b[j] * b[j];
    c->rhs_norm = std::sqrt(s);
    const size_t bytes = sizeof(double) * (size_t)c->n;
    std::memcpy(c->h_stage, b, bytes);
    NPBCG_HIP(hipMemcpyAsync(c->pbs.x, c->h_stage, bytes, hipMemcpyHostToDevice, c->stream));
    NPBCG_HIP(hipStreamSynchronize(c->stream));
    clear_error();
    return NPBCG_OK;
}

int npbcg_get_solution(npbcg_ctx *c, double *x) {
    if (!c || !x) return set_error(NPBCG_ERR_INVALID, "npbcg_get_solution: NULL argument");
    NPBCG_HIP(hipSetDevice(c->device));
    const size_t bytes = sizeof(double) * (size_t)c->n;
    NPBCG_HIP(hipMemcpyAsync(c->h_stage, c->pbs.z, bytes, hipMemcpyDeviceToHost, c->stream));
    NPBCG_HIP(hipStreamSynchronize(c->stream));
    std::memcpy(x, c->h_stage, bytes);
    clear_error();
    return NPBCG_OK;
}

namespace {

int ensure_precond_buffers(npbcg_ctx *c) {
    if (c->d_minv) return NPBCG_OK;
    NPBCG_HIP(hipMalloc(&c->d_minv, sizeof(double) * c->vstride));
    NPBCG_HIP(hipMalloc(&c->d_minv32, sizeof(float) * c->vstride));
    NPBCG_HIP(npbcg_dev::launch_jacobi(c->pb, c->d_minv, c->stream));
    NPBCG_HIP(npbcg_dev::launch_narrow(c->d_minv, c->d_minv32, c->n, c->pb.vec_blocks,
                                       c->stream));
    NPBCG_HIP(hipStreamSynchronize(c->stream));
    return NPBCG_OK;
}

// The top-K graph (device) and the CSR layout of the sparse operator M^-1:
// row j holds its diagonal, its K out-neighbours and its in-neighbours (rows
// whose top-K contain j).
// The top-K graph on the device (features, neighbours, in-edge lists) and the
// zero (g, h) of the Neumann form: all the network and its loss need.
int ensure_graph_device(npbcg_ctx *c) {
    auto &g = c->g;
    if (g.graph_built) return NPBCG_OK;
    const int n = c->n;
    const int K = npbcg_dev::kNbr;
    const size_t nk = (size_t)n * K;
    NPBCG_HIP(hipMalloc(&g.adiag, sizeof(double) * n));
    NPBCG_HIP(hipMalloc(&g.feat, sizeof(float) * n * npbcg_dev::kFeat));
    NPBCG_HIP(hipMalloc(&g.nbr, sizeof(int) * nk));
    NPBCG_HIP(hipMalloc(&g.wgt, sizeof(float) * nk));
    NPBCG_HIP(hipMalloc(&g.nval, sizeof(float) * nk));
    NPBCG_HIP(hipMalloc(&g.tptr, sizeof(int) * (n + 1)));
    NPBCG_HIP(hipMalloc(&g.tsrc, sizeof(int) * nk));
    NPBCG_HIP(hipMalloc(&g.tw, sizeof(float) * nk));
    NPBCG_HIP(hipMalloc(&g.tval, sizeof(float) * nk));
    NPBCG_HIP(hipMalloc(&g.cnt, sizeof(int) * n));
    NPBCG_HIP(npbcg_dev::launch_graph_features(c->d_rowstr, c->d_colidx, c->d_a, n, g.adiag,
                                               g.feat, g.nbr, g.wgt, g.nval, c->stream));
    NPBCG_HIP(npbcg_dev::launch_transpose(g.nbr, g.wgt, g.nval, n, g.cnt, g.tptr, g.tsrc, g.tw,
                                          g.tval, c->stream));
    NPBCG_HIP(hipMalloc(&g.gh_zero, sizeof(float) * 2 * n));
    NPBCG_HIP(hipMemsetAsync(g.gh_zero, 0, sizeof(float) * 2 * n, c->stream));
    g.graph_built = true;
    return NPBCG_OK;
}

// The CSR operator of M^-1 (structure built on the device from the graph;
// the host reads only the slot count).
int ensure_row_graph(npbcg_ctx *c) {
    auto &g = c->g;
    if (g.built) return NPBCG_OK;
    int rc = ensure_graph_device(c);
    if (rc != NPBCG_OK) return rc;
    const int n = c->n;
    const int K = npbcg_dev::kNbr;
    int edges = 0;
    NPBCG_HIP(hipMemcpyAsync(&edges, g.tptr + n, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    NPBCG_HIP(hipStreamSynchronize(c->stream));
    // row j: [diagonal | K out-neighbours | in-neighbours], first slot j (1 + K) + tptr[j]
    const long long slots_ll = (long long)n * (1 + K) + edges;
    if (slots_ll > (long long)INT32_MAX)
        return set_error(NPBCG_ERR_CAPACITY, "preconditioner operator exceeds 32-bit offsets");
    const size_t slots = (size_t)slots_ll;
    g.slots = slots;
    NPBCG_HIP(hipMalloc(&g.op_off, sizeof(int) * (n + 1)));
    NPBCG_HIP(hipMalloc(&g.op_col, sizeof(int) * slots));
    NPBCG_HIP(hipMalloc(&g.op_src, sizeof(float) * slots));
    NPBCG_HIP(hipMalloc(&g.op_val, sizeof(double) * slots));
    NPBCG_HIP(hipMalloc(&g.op_val32, sizeof(float) * slots));
    NPBCG_HIP(hipMalloc(&g.gh_learned, sizeof(float) * 2 * n));
    NPBCG_HIP(hipMalloc(&g.pz, sizeof(double) * c->vstride));
    NPBCG_HIP(hipMalloc(&g.pz32, sizeof(float) * c->vstride));
    NPBCG_HIP(npbcg_dev::launch_build_op(g.nbr, g.nval, g.tptr, g.tsrc, g.tval, n, g.op_off,
                                         g.op_col, g.op_src, c->stream));
    NPBCG_HIP(hipMemsetAsync(g.gh_learned, 0, sizeof(float) * 2 * n, c->stream));
    NPBCG_HIP(hipMemsetAsync(g.pz, 0, sizeof(double) * c->vstride, c->stream));
    NPBCG_HIP(hipMemsetAsync(g.pz32, 0, sizeof(float) * c->vstride, c->stream));
    g.op.n = n;
    g.op.off = g.op_off;
    g.op.col = g.op_col;
    g.op.src_val = g.op_src;
    g.op.val = g.op_val;
    g.op.val32 = g.op_val32;
    g.loaded = -1;
    g.built = true;
    return NPBCG_OK;
}

// Points the solve views at the preconditioner of `kind`.
int select_precond(npbcg_ctx *c, int kind) {
    c->pbs.minv = nullptr;
    c->mp.minv = nullptr;
    c->pbs.pop = nullptr;
    c->pbs.pz = nullptr;
    c->mp.pop = nullptr;
    c->mp.pz = nullptr;
    c->active_precond = NPBCG_PRECOND_NONE;
    if (kind == NPBCG_PRECOND_NONE) return NPBCG_OK;
    if (kind == NPBCG_PRECOND_JACOBI) {
        int rc = ensure_precond_buffers(c);
        if (rc != NPBCG_OK) return rc;
        c->pbs.minv = c->d_minv;
        c->mp.minv = c->d_minv32;
        c->active_precond = kind;
        return NPBCG_OK;
    }
    int rc = ensure_row_graph(c);
    if (rc != NPBCG_OK) return rc;
    auto &g = c->g;
    if (g.loaded != kind) {
        const float *gh = kind == NPBCG_PRECOND_LEARNED ? g.gh_learned : g.gh_zero;
        NPBCG_HIP(npbcg_dev::launch_precond_values(g.op, g.adiag, gh, c->stream));
        g.loaded = kind;
    }
    c->pbs.pop = &g.op;
    c->pbs.pz = g.pz;
    c->mp.pop = &g.op;
    c->mp.pz = g.pz32;
    c->active_precond = kind;
    return NPBCG_OK;
}

int ensure_pairs(npbcg_ctx *c) {
    auto &pr = c->pairs;
    if (pr.r) return NPBCG_OK;
    const size_t pn = (size_t)NPBCG_MAX_PAIRS * c->n;
    NPBCG_HIP(hipMalloc(&pr.r, sizeof(float) * pn));
    NPBCG_HIP(hipMalloc(&pr.e, sizeof(float) * pn));
    NPBCG_HIP(hipMalloc(&pr.omega, sizeof(float) * NPBCG_MAX_PAIRS));
    NPBCG_HIP(hipMalloc(&pr.part, sizeof(double) * NPBCG_MAX_PAIRS * c->pb.vec_blocks));
    NPBCG_HIP(hipMalloc(&pr.xstar, sizeof(double) * c->vstride));
    return NPBCG_OK;
}

int solve_once(npbcg_ctx *c, int mode, double rtol, int maxit, npbcg_solve_stats *st);

}  // namespace

}  // extern "C"

namespace npbcg_detail {

// The context's matrix as the network sees it (builds the graph if needed).
int ctx_net_system(npbcg_ctx *c, npbcg_dev::NetSystem *sys) {
    int rc = ensure_graph_device(c);
    if (rc != NPBCG_OK) return rc;
    const auto &g = c->g;
    sys->n = c->n;
    sys->adiag = g.adiag;
    sys->feat = g.feat;
    sys->nbr = g.nbr;
    sys->wgt = g.wgt;
    sys->nval = g.nval;
    sys->tptr = g.tptr;
    sys->tsrc = g.tsrc;
    sys->tw = g.tw;
    sys->tval = g.tval;
    return NPBCG_OK;
}

// The resident pairs (count 0 before npbcg_log_pairs / npbcg_set_pairs).
npbcg_dev::PairSet ctx_pairs(const npbcg_ctx *c) {
    npbcg_dev::PairSet ps;
    ps.count = c->pairs.count;
    ps.r = c->pairs.r;
    ps.e = c->pairs.e;
    ps.omega = c->pairs.omega;
    return ps;
}

const float *ctx_zero_gh(const npbcg_ctx *c) { return c->g.gh_zero; }

// Installs (g, h) from the device array gh as the LEARNED preconditioner.
int ctx_install_learned(npbcg_ctx *c, const float *gh) {
    int rc = ensure_row_graph(c);
    if (rc != NPBCG_OK) return rc;
    NPBCG_HIP(hipMemcpyAsync(c->g.gh_learned, gh, sizeof(float) * 2 * c->n,
                             hipMemcpyDeviceToDevice, c->stream));
    if (c->g.loaded == NPBCG_PRECOND_LEARNED) c->g.loaded = -1;
    c->learned_ready = true;
    return NPBCG_OK;
}

hipStream_t ctx_stream(const npbcg_ctx *c) { return c->stream; }
int ctx_device(const npbcg_ctx *c) { return c->device; }
int ctx_n(const npbcg_ctx *c) { return c->n; }

}  // namespace npbcg_detail

extern "C" {

int npbcg_set_preconditioner(npbcg_ctx *c, int kind) {
    if (!c || kind < NPBCG_PRECOND_NONE || kind > NPBCG_PRECOND_NEUMANN)
        return set_error(NPBCG_ERR_INVALID, "npbcg_set_preconditioner: invalid argument");
    if (kind == NPBCG_PRECOND_LEARNED && !c->learned_ready)
        return set_error(NPBCG_ERR_INVALID,
                         "npbcg_set_preconditioner: no learned preconditioner (npbcg_net_apply)");
    NPBCG_HIP(hipSetDevice(c->device));
    c->precond = kind;
    clear_error();
    return NPBCG_OK;
}

int npbcg_get_preconditioner(npbcg_ctx *c, double *minv) {
    if (!c || !minv) return set_error(NPBCG_ERR_INVALID, "npbcg_get_preconditioner: NULL");
    NPBCG_HIP(hipSetDevice(c->device));
    int rc = select_precond(c, c->precond);
    if (rc != NPBCG_OK) return rc;
    const size_t bytes = sizeof(double) * c->n;
    if (c->pbs.minv) {
        NPBCG_HIP(hipMemcpyAsync(c->h_stage, c->pbs.minv, bytes, hipMemcpyDeviceToHost,
                                 c->stream));
        NPBCG_HIP(hipStreamSynchronize(c->stream));
        std::memcpy(minv, c->h_stage, bytes);
    } else if (c->pbs.pop) {  // entry 0 of every row of the operator
        const auto &g = c->g;
        std::vector<double> val(g.slots);
        std::vector<int> off(c->n + 1);
        NPBCG_HIP(hipMemcpyAsync(val.data(), g.op_val, sizeof(double) * val.size(),
                                 hipMemcpyDeviceToHost, c->stream));
        NPBCG_HIP(hipMemcpyAsync(off.data(), g.op_off, sizeof(int) * off.size(),
                                 hipMemcpyDeviceToHost, c->stream));
        NPBCG_HIP(hipStreamSynchronize(c->stream));
        for (int j = 0; j < c->n; j++)
            minv[j] = val[(size_t)off[j]];
    } else {
        for (int j = 0; j < c->n; j++) minv[j] = 1.0;
    }
    clear_error();
    return NPBCG_OK;
}

int npbcg_apply_preconditioner(npbcg_ctx *c, const double *v, double *z) {
    if (!c || !v || !z) return set_error(NPBCG_ERR_INVALID, "npbcg_apply_preconditioner: NULL");
    NPBCG_HIP(hipSetDevice(c->device));
    int rc = select_precond(c, c->precond);
    if (rc != NPBCG_OK) return rc;
    const int n = c->n;
    const size_t bytes = sizeof(double) * n;
    double *dv = c->pbs.q, *dz = c->pbs.p;  // solve-view scratch (rebuilt by every solve)
    std::memcpy(c->h_stage, v, bytes);
    NPBCG_HIP(hipMemcpyAsync(dv, c->h_stage, bytes, hipMemcpyHostToDevice, c->stream));
    if (c->pbs.pop) {
        NPBCG_HIP(npbcg_dev::launch_precond_apply(*c->pbs.pop, dv, dz, nullptr, c->pbs.part.xx,
                                                  c->pb.vec_blocks, nullptr, c->stream));
        NPBCG_HIP(hipMemcpyAsync(c->h_stage, dz, bytes, hipMemcpyDeviceToHost, c->stream));
        NPBCG_HIP(hipStreamSynchronize(c->stream));
        std::memcpy(z, c->h_stage, bytes);
    } else {
        NPBCG_HIP(hipStreamSynchronize(c->stream));
        std::vector<double> m(n);
        rc = npbcg_get_preconditioner(c, m.data());
        if (rc != NPBCG_OK) return rc;
        for (int j = 0; j < n; j++) z[j] = m[j] * v[j];
    }
    clear_error();
    return NPBCG_OK;
}

int npbcg_log_pairs(npbcg_ctx *c, int pairs, float *r, float *e) {
    if (!c || pairs < 1 || pairs > NPBCG_MAX_PAIRS)
        return set_error(NPBCG_ERR_INVALID, "npbcg_log_pairs: invalid argument");
    if (c->rhs_norm <= 0.0)
        return set_error(NPBCG_ERR_INVALID, "npbcg_log_pairs: needs a nonzero right-hand side");
    NPBCG_HIP(hipSetDevice(c->device));
    int rc = ensure_pairs(c);
    if (rc != NPBCG_OK) return rc;
    // x* from plain FP64 CG
    const int kind = c->precond;
    c->precond = NPBCG_PRECOND_NONE;
    npbcg_solve_stats st{};
    rc = npbcg_solve(c, NPBCG_SOLVE_FP64, 1e-12, 5000, &st);
    c->precond = kind;
    if (rc != NPBCG_OK) return rc;
    auto &pr = c->pairs;
    const int n = c->n, vb = c->pb.vec_blocks;
    NPBCG_HIP(hipMemcpyAsync(pr.xstar, c->pbs.z, sizeof(double) * n, hipMemcpyDeviceToDevice,
                             c->stream));
    // the same CG from x = 0, logging (r_k, x* - x_k) before step k
    Fp64Problem lp = c->pbs;
    lp.minv = nullptr;
    lp.pop = nullptr;
    lp.pz = nullptr;
    NPBCG_HIP(npbcg_dev::launch_reset_scalars(lp, 0.0, c->stream, 0.0, 0));
    NPBCG_HIP(npbcg_dev::launch_cg_init(lp, 0, c->stream));
    for (int k = 0; k < pairs; k++) {
        NPBCG_HIP(npbcg_dev::launch_log_pair(lp.r, lp.z, pr.xstar, n, pr.r + (size_t)k * n,
                                             pr.e + (size_t)k * n, pr.part + (size_t)k * vb, vb,
                                             c->stream));
        if (k + 1 < pairs) NPBCG_HIP(npbcg_dev::enqueue_solve_step(lp, k, 0, c->stream));
    }
    NPBCG_HIP(npbcg_dev::launch_pair_weights(pr.part, vb, pairs, pr.omega, c->stream));
    NPBCG_HIP(hipMemcpyAsync(c->pbs.z, pr.xstar, sizeof(double) * n, hipMemcpyDeviceToDevice,
                             c->stream));  // npbcg_get_solution returns x*
    pr.count = pairs;
    if (r)
        NPBCG_HIP(hipMemcpyAsync(r, pr.r, sizeof(float) * pairs * (size_t)n,
                                 hipMemcpyDeviceToHost, c->stream));
    if (e)
        NPBCG_HIP(hipMemcpyAsync(e, pr.e, sizeof(float) * pairs * (size_t)n,
                                 hipMemcpyDeviceToHost, c->stream));
    NPBCG_HIP(hipStreamSynchronize(c->stream));
    clear_error();
    return NPBCG_OK;
}

int npbcg_set_pairs(npbcg_ctx *c, int pairs, const float *r, const float *e) {
    if (!c || !r || !e || pairs < 1 || pairs > NPBCG_MAX_PAIRS)
        return set_error(NPBCG_ERR_INVALID, "npbcg_set_pairs: invalid argument");
    NPBCG_HIP(hipSetDevice(c->device));
    int rc = ensure_pairs(c);
    if (rc != NPBCG_OK) return rc;
    const int n = c->n;
    std::vector<float> omega(pairs);
    for (int k = 0; k < pairs; k++) {
        double s = 0.0;
        for (int j = 0; j < n; j++) s += (double)e[(size_t)k * n + j] * e[(size_t)k * n + j];
        omega[k] = s > 0.0 ? (float)(1.0 / (pairs * s)) : 0.0f;
    }
    auto &pr = c->pairs;
    NPBCG_HIP(hipMemcpyAsync(pr.r, r, sizeof(float) * pairs * (size_t)n, hipMemcpyHostToDevice,
                             c->stream));
    NPBCG_HIP(hipMemcpyAsync(pr.e, e, sizeof(float) * pairs * (size_t)n, hipMemcpyHostToDevice,
                             c->stream));
    NPBCG_HIP(hipMemcpyAsync(pr.omega, omega.data(), sizeof(float) * pairs,
                             hipMemcpyHostToDevice, c->stream));
    NPBCG_HIP(hipStreamSynchronize(c->stream));
    pr.count = pairs;
    clear_error();
    return NPBCG_OK;
}

int npbcg_graph(npbcg_ctx *c, int *nbr, float *wgt, float *nval, float *feat) {
    if (!c) return set_error(NPBCG_ERR_INVALID, "npbcg_graph: NULL context");
    NPBCG_HIP(hipSetDevice(c->device));
    int rc = ensure_graph_device(c);
    if (rc != NPBCG_OK) return rc;
    const size_t nk = (size_t)c->n * npbcg_dev::kNbr;
    const auto &g = c->g;
    if (nbr) NPBCG_HIP(hipMemcpyAsync(nbr, g.nbr, sizeof(int) * nk, hipMemcpyDeviceToHost, c->stream));
    if (wgt) NPBCG_HIP(hipMemcpyAsync(wgt, g.wgt, sizeof(float) * nk, hipMemcpyDeviceToHost, c->stream));
    if (nval)
        NPBCG_HIP(hipMemcpyAsync(nval, g.nval, sizeof(float) * nk, hipMemcpyDeviceToHost, c->stream));
    if (feat)
        NPBCG_HIP(hipMemcpyAsync(feat, g.feat, sizeof(float) * c->n * npbcg_dev::kFeat,
                                 hipMemcpyDeviceToHost, c->stream));
    NPBCG_HIP(hipStreamSynchronize(c->stream));
    clear_error();
    return npbcg_dev::kNbr;
}

int npbcg_solve(npbcg_ctx *c, int mode, double rtol, int maxit, npbcg_solve_stats *st) {
    npbcg_detail::TraceRange trace("npbcg_solve");
    if (!c) return set_error(NPBCG_ERR_INVALID, "npbcg_solve: NULL context");
    NPBCG_HIP(hipSetDevice(c->device));
    int rc = select_precond(c, c->precond);
    if (rc != NPBCG_OK) return rc;
    npbcg_solve_stats s{};
    rc = solve_once(c, mode, rtol, maxit, &s);
    s.precond = c->active_precond;
    if (c->precond == NPBCG_PRECOND_LEARNED &&
        (rc == NPBCG_ERR_NOCONV || !(s.relres == s.relres))) {
        // The learned operator did not get there: Jacobi-CG redoes the solve
        // from x = 0.  The solve views now point at the Jacobi diagonal, and the
        // captured solve sequences are keyed by the preconditioner in effect,
        // so nothing of the learned sequence is replayed.
        const double first_ms = s.time_ms;
        rc = select_precond(c, NPBCG_PRECOND_JACOBI);
        if (rc != NPBCG_OK) return rc;
        rc = solve_once(c, mode, rtol, maxit, &s);
        s.precond = c->active_precond;
        s.fallback = 1;
        s.time_ms += first_ms;
    }
    if (st) *st = s;
    return rc;
}

namespace {

// Runs the first stretch of a solve: replays its captured graph when the
// key (mode, preconditioner in effect, rtol, maxit, batch, ||b||) matches the
// last capture, else captures `enqueue` anew (or, with use_graph = 0,
// launches it directly).  The preconditioner in effect decides which
// buffers the captured kernels read (Jacobi diagonal or sparse operator).
int run_solve_sequence(npbcg_ctx *c, int mode, double rtol, int maxit, int batch,
                       const std::function<hipError_t()> &enqueue) {
    if (!c->opts.use_graph) {
        NPBCG_HIP(enqueue());
        return NPBCG_OK;
    }
    auto &sg = c->solve_graph[mode];
    if (!(sg.ex && sg.kind == c->active_precond && sg.rtol == rtol && sg.maxit == maxit &&
          sg.batch == batch && sg.rhs == c->rhs_norm)) {
        if (sg.ex) (void)hipGraphExecDestroy(sg.ex);
        if (sg.g) (void)hipGraphDestroy(sg.g);
        sg = npbcg_ctx::SolveGraph{};
        NPBCG_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        hipError_t e = enqueue();
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(c->stream, &g);
        if (e != hipSuccess || e2 != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return set_error(NPBCG_ERR_DEVICE, std::string("solve graph capture failed: ") +
                                                   hipGetErrorString(e != hipSuccess ? e : e2));
        }
        hipGraphExec_t ex = nullptr;
        hipError_t e3 = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
        if (e3 != hipSuccess) {
            (void)hipGraphDestroy(g);
            return set_error(NPBCG_ERR_DEVICE, std::string("hipGraphInstantiate: ") +
                                                   hipGetErrorString(e3));
        }
        sg.g = g;
        sg.ex = ex;
        sg.kind = c->active_precond;
        sg.rtol = rtol;
        sg.maxit = maxit;
        sg.batch = batch;
        sg.rhs = c->rhs_norm;
    }
    NPBCG_HIP(hipGraphLaunch(sg.ex, c->stream));
    return NPBCG_OK;
}

int solve_once(npbcg_ctx *c, int mode, double rtol, int maxit, npbcg_solve_stats *st) {
    if (!c || !(rtol > 0.0) || maxit < 1 ||
        (mode != NPBCG_SOLVE_FP64 && mode != NPBCG_SOLVE_MIXED))
        return set_error(NPBCG_ERR_INVALID, "npbcg_solve: invalid argument");
    if (c->rhs_norm < 0.0)
        return set_error(NPBCG_ERR_INVALID, "npbcg_solve: no right-hand side (npbcg_set_rhs)");
    if (mode == NPBCG_SOLVE_MIXED && c->pb.spmv_variant != npbcg_dev::kSpmvTiled)
        return set_error(NPBCG_ERR_INVALID,
                         "npbcg_solve: mixed precision needs the LDS-staged layout");
    NPBCG_HIP(hipSetDevice(c->device));
    npbcg_solve_stats s{};
    if (c->rhs_norm == 0.0) {  // x = 0 solves b = 0
        NPBCG_HIP(npbcg_dev::launch_fill(c->pbs.z, c->n + 1, 0.0, c->pb.vec_blocks, c->stream));
        NPBCG_HIP(hipStreamSynchronize(c->stream));
        s.converged = 1;
        if (st) *st = s;
        clear_error();
        return NPBCG_OK;
    }
    if (mode == NPBCG_SOLVE_MIXED) {
        int rc = ensure_mixed(c);
        if (rc != NPBCG_OK) return rc;
    }
    const Fp64Problem &pbs = c->pbs;
    const int batch0 = c->opts.solve_batch > 0 ? c->opts.solve_batch : 10;
    float ms_total = 0.f;
    Scalars sc{};
    double relres = 1.0;
    if (mode == NPBCG_SOLVE_FP64) {
        // a repeated solve on this matrix launches one step more than the last
        // one needed instead of a full batch (steps past the stop return early
        // but still cost a launch each)
        int *bud = c->budget[0][c->active_precond];
        int batch = std::min(bud[0] > 0 ? bud[0] : batch0, maxit);
        // first batch: reset, init, `batch` steps, residual (captured once per key)
        auto first_batch = [&]() -> hipError_t {
            hipError_t e = npbcg_dev::launch_reset_scalars(pbs, 0.0, c->stream, rtol, 1);
            if (e == hipSuccess) e = npbcg_dev::launch_cg_init(pbs, 0, c->stream);
            if (e == hipSuccess) e = npbcg_dev::enqueue_solve_steps(pbs, batch, 1, c->stream);
            if (e == hipSuccess) e = npbcg_dev::launch_resid(pbs, 0, c->stream);
            return e;
        };
        NPBCG_HIP(hipEventRecord(c->ev_start, c->stream));
        int rc = run_solve_sequence(c, 0, rtol, maxit, batch, first_batch);
        if (rc != NPBCG_OK) return rc;
        bool first = true;
        for (;;) {
            if (!first) {
                NPBCG_HIP(npbcg_dev::enqueue_solve_steps(pbs, batch, 0, c->stream));
                NPBCG_HIP(npbcg_dev::launch_resid(pbs, 0, c->stream));
            }
            NPBCG_HIP(hipEventRecord(c->ev_stop, c->stream));
            int rc = read_scalars_from(c, pbs.sc, &sc);
            if (rc != NPBCG_OK) return rc;
            s.sweeps++;
            s.fp64_spmvs = sc.steps + s.sweeps;
            relres = sc.rnorm / c->rhs_norm;
            if (relres <= rtol || sc.steps >= maxit) break;
            // the true residual missed rtol: restart CG from it
            NPBCG_HIP(npbcg_dev::launch_clear_done(pbs, c->stream));
            NPBCG_HIP(npbcg_dev::launch_residual_restart(pbs, c->stream));
            batch = std::min(std::max(batch0 / 2, 2), maxit - sc.steps);
            first = false;
        }
        s.iterations = sc.steps;
        // a converged solve sets the first batch of the next one to its own
        // step count plus one (at most the default): an easier right-hand side
        // lowers it, a harder one (which needed a restart) raises it again
        if (relres <= rtol) bud[0] = std::min(batch0, sc.steps + 1);
    } else {
        const npbcg_dev::MixedProblem &mp = c->mp;
        Fp64Problem inner = pbs;  // same launch shapes, inner scalars
        inner.sc = mp.sc;
        // Sweeps are decided on the device (launch_sweep_prepare), so two are
        // enqueued before the first host synchronisation (FP32 holds about
        // 1e-6 per sweep, so two reach 1e-10 from x = 0 on well-conditioned
        // systems); later ones go one at a time.  maxit is checked per sweep.
        int *bud = c->budget[1][c->active_precond];
        const int b1 = bud[0] > 0 ? bud[0] : batch0 - 2;
        const int b2 = bud[1] > 0 ? bud[1] : batch0 / 2 + 1;
        auto sweep = [&](int index) -> hipError_t {
            const int steps = index == 0 ? b1 : index == 1 ? b2 : batch0 / 2 + 1;
            hipError_t e = npbcg_dev::launch_sweep_prepare(pbs.sc, mp.sc, c->rhs_norm, rtol,
                                                           maxit, index == 0 ? 1 : 0, c->stream);
            if (e == hipSuccess)
                e = npbcg_dev::enqueue_inner32(pbs, mp, index == 0 ? 1 : 0, steps, c->stream);
            if (e == hipSuccess) e = npbcg_dev::launch_refine(pbs, mp, c->stream);
            if (e == hipSuccess) e = npbcg_dev::launch_resid(pbs, 0, c->stream);
            return e;
        };
        auto first_two = [&]() -> hipError_t {
            hipError_t e = npbcg_dev::launch_reset_scalars(pbs, 0.0, c->stream, rtol, 0);
            if (e == hipSuccess) e = npbcg_dev::launch_reset_scalars(inner, 0.0, c->stream, 1e-6, 1);
            if (e == hipSuccess)
                e = npbcg_dev::launch_fill(pbs.z, c->n + 1, 0.0, c->pb.vec_blocks, c->stream);
            if (e == hipSuccess) e = sweep(0);
            if (e == hipSuccess) e = sweep(1);
            return e;
        };
        NPBCG_HIP(hipEventRecord(c->ev_start, c->stream));
        int rc0 = run_solve_sequence(c, 1, rtol, maxit, b1 * 1000 + b2, first_two);
        if (rc0 != NPBCG_OK) return rc0;
        int enqueued = 2, inner_steps = 0, si_last_steps = 0;
        for (bool first = true;; first = false) {
            if (!first) NPBCG_HIP(sweep(enqueued++));
            NPBCG_HIP(hipEventRecord(c->ev_stop, c->stream));
            Scalars si{};
            int rc = read_scalars_from(c, mp.sc, &si);
            if (rc != NPBCG_OK) return rc;
            rc = read_scalars_from(c, pbs.sc, &sc);
            if (rc != NPBCG_OK) return rc;
            inner_steps = sc.inner_total + si.steps;
            si_last_steps = si.steps;
            relres = sc.rnorm / c->rhs_norm;
            if (relres <= rtol || inner_steps >= maxit || enqueued >= 1000) break;
        }
        s.sweeps = sc.sweeps;
        s.iterations = inner_steps;
        if (relres <= rtol && s.sweeps == 2) {  // the common case: learn both budgets
            bud[0] = std::min(batch0 - 2, std::max(2, sc.sweep1_steps + 1));
            bud[1] = std::min(batch0 / 2 + 1, std::max(2, si_last_steps + 1));
        } else if (relres <= rtol && s.sweeps == 1) {
            bud[0] = std::min(batch0 - 2, std::max(2, sc.sweep1_steps + 1));
        } else if (relres <= rtol) {  // needed more sweeps: back to the defaults
            bud[0] = bud[1] = 0;
        }
        s.fp32_spmvs = inner_steps;
        s.fp64_spmvs = s.sweeps;
    }
    NPBCG_HIP(npbcg_dev::launch_clear_done(pbs, c->stream, 1));
    NPBCG_HIP(hipEventElapsedTime(&ms_total, c->ev_start, c->ev_stop));
    s.relres = relres;
    s.time_ms = ms_total;
    s.converged = relres <= rtol ? 1 : 0;
    if (st) *st = s;
    if (!s.converged)
        return set_error(NPBCG_ERR_NOCONV, "npbcg_solve: tolerance not reached within maxit");
    clear_error();
    return NPBCG_OK;
}

}  // namespace

// ---- reference-shaped shim -------------------------------------------------

namespace {
std::mutex g_compat_mu;
int g_compat_n = 0;
npbcg_ctx *g_compat_ctx = nullptr;
}  // namespace

int npbcg_compat_set_n(int naa) {
    if (naa < 1) return set_error(NPBCG_ERR_INVALID, "npbcg_compat_set_n: naa < 1");
    std::lock_guard<std::mutex> lk(g_compat_mu);
    if (naa != g_compat_n && g_compat_ctx) {
        npbcg_destroy(g_compat_ctx);
        g_compat_ctx = nullptr;
    }
    g_compat_n = naa;
    return NPBCG_OK;
}

void npbcg_compat_conj_grad(int colidx[], int rowstr[], double x[], double z[], double a[],
                            double p[], double q[], double r[], double *rnorm) {
    std::lock_guard<std::mutex> lk(g_compat_mu);
    const double nan = std::numeric_limits<double>::quiet_NaN();
    if (g_compat_n < 1 || !rowstr) {
        set_error(NPBCG_ERR_INVALID, "npbcg_compat_conj_grad: call npbcg_compat_set_n first");
        if (rnorm) *rnorm = nan;
        return;
    }
    if (g_compat_ctx && !same_matrix(g_compat_ctx, rowstr, colidx, a)) {
        npbcg_destroy(g_compat_ctx);
        g_compat_ctx = nullptr;
    }
    if (!g_compat_ctx) {
        if (npbcg_create(&g_compat_ctx, g_compat_n, rowstr, colidx, a, nullptr) != NPBCG_OK) {
            g_compat_ctx = nullptr;
            if (rnorm) *rnorm = nan;
            return;
        }
    }
    if (npbcg_conj_grad(g_compat_ctx, colidx, rowstr, x, z, a, p, q, r, rnorm) != NPBCG_OK)
        if (rnorm) *rnorm = nan;
}

void npbcg_compat_release(void) {
    std::lock_guard<std::mutex> lk(g_compat_mu);
    npbcg_destroy(g_compat_ctx);
    g_compat_ctx = nullptr;
}

}  // extern "C"

extern "C" {

int npbcg_spmv_kernel_bench(npbcg_ctx *c, int precision, int reps, double *avg_ms) {
    npbcg_detail::TraceRange trace("npbcg_spmv_kernel_bench");
    if (!c || reps < 1 || !avg_ms || (precision != 0 && precision != 1))
        return set_error(NPBCG_ERR_INVALID, "npbcg_spmv_kernel_bench: invalid argument");
    if (precision == 1 && c->pb.spmv_variant != npbcg_dev::kSpmvTiled)
        return set_error(NPBCG_ERR_INVALID,
                         "npbcg_spmv_kernel_bench: the FP32 SpMV needs the LDS-staged layout");
    NPBCG_HIP(hipSetDevice(c->device));
    if (precision == 1) {
        int rc = ensure_mixed(c);
        if (rc != NPBCG_OK) return rc;
        // the multiplied vector: x in FP32 (finite values); no stop flag set
        NPBCG_HIP(npbcg_dev::launch_to_fp32(c->pb.x, c->mp.p, c->n, c->stream));
        Fp64Problem inner = c->pb;
        inner.sc = c->mp.sc;
        NPBCG_HIP(npbcg_dev::launch_clear_done(inner, c->stream, 1));
    }
    NPBCG_HIP(npbcg_dev::launch_clear_done(c->pb, c->stream, 1));
    std::vector<hipEvent_t> ev(2 * (size_t)reps, nullptr);
    hipError_t e = hipSuccess;
    for (size_t i = 0; i < ev.size() && e == hipSuccess; i++)
        e = hipEventCreateWithFlags(&ev[i], hipEventDisableSystemFence);
    auto launch = [&](int step, hipEvent_t e0, hipEvent_t e1) {
        return precision == 1
                   ? npbcg_dev::launch_spmv32_kernel(c->pb, c->mp, step, e0, e1, c->stream)
                   : npbcg_dev::launch_spmv_kernel(c->pb, step, e0, e1, c->stream);
    };
    for (int i = 0; i < 2 && e == hipSuccess; i++) e = launch(i, nullptr, nullptr);  // warm
    for (int i = 0; i < reps && e == hipSuccess; i++) e = launch(i, ev[2 * i], ev[2 * i + 1]);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    double sum = 0.0;
    for (int i = 0; i < reps && e == hipSuccess; i++) {
        float ms = 0.f;
        e = hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]);
        sum += ms;
    }
    for (hipEvent_t x : ev)
        if (x) (void)hipEventDestroy(x);
    if (e != hipSuccess)
        return set_error(NPBCG_ERR_DEVICE,
                         std::string("npbcg_spmv_kernel_bench: ") + hipGetErrorString(e));
    *avg_ms = sum / reps;
    clear_error();
    return NPBCG_OK;
}

}  // extern "C"
