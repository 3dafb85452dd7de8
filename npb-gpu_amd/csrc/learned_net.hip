// Host side of the learned preconditioner: the network object (FP32 master
// parameters, Adam moments, BF16 weight copies, activation buffers), the
// training loop over generated systems and inference into a context.
//
// Training streams systems: host threads run the NPB generator (one system
// per thread, single-threaded each) ahead of the device, which for every
// system uploads the CSR, solves it with plain FP64 CG, logs the first P
// residual/correction pairs, builds the top-K graph and takes Adam steps on
// the pair loss.  Systems are consumed in index order, so a run is
// reproducible for a given seed whatever the thread count.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "cg_kernels.h"
#include "learned_kernels.h"
#include "npbcg.h"
#include "npbcg_internal.h"
#include "trace_range.h"

using namespace npbcg_dev;
using npbcg_detail::clear_error;
using npbcg_detail::set_error;

namespace npbcg_detail {
int ctx_net_system(npbcg_ctx *c, NetSystem *sys);
PairSet ctx_pairs(const npbcg_ctx *c);
const float *ctx_zero_gh(const npbcg_ctx *c);
int ctx_install_learned(npbcg_ctx *c, const float *gh);
hipStream_t ctx_stream(const npbcg_ctx *c);
int ctx_device(const npbcg_ctx *c);
int ctx_n(const npbcg_ctx *c);
}  // namespace npbcg_detail

#define NET_HIP(call)                                                            \
    do {                                                                         \
        hipError_t _e = (call);                                                  \
        if (_e != hipSuccess)                                                    \
            return set_error(NPBCG_ERR_DEVICE, std::string(#call) + ": " +       \
                                                   hipGetErrorString(_e));       \
    } while (0)

namespace {

constexpr float kGmax = 1.0f;      // |g|, |h| <= 1: scales within [1/e, e]
constexpr int kWeightParts = 512;  // workgroups of the weight-gradient GEMMs

uint64_t splitmix64(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

double uniform01(uint64_t &s) { return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0); }

}  // namespace

struct npbcg_net {
    int device = 0;
    hipStream_t stream = nullptr;
    float *params = nullptr, *grad = nullptr, *m1 = nullptr, *m2 = nullptr;
    __bf16 *wb = nullptr;
    float *norm = nullptr;       // feature mean/std
    bool norm_set = false;
    int adam_step = 0;
    double *d_loss = nullptr;    // [0] learned loss, [1] base loss
    double *h_loss = nullptr;    // pinned copy
    NetBuffers nb;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

namespace {

void free_buffers(NetBuffers &nb) {
    for (void *p : {(void *)nb.x0, (void *)nb.x[0], (void *)nb.x[1], (void *)nb.x[2],
                    (void *)nb.h3, (void *)nb.dz, (void *)nb.dx, (void *)nb.gh, (void *)nb.dgh,
                    (void *)nb.pw, (void *)nb.pd, (void *)nb.wpart, (void *)nb.rpart,
                    (void *)nb.loss_part})
        (void)hipFree(p);
    const float *norm = nb.norm;
    nb = NetBuffers{};
    nb.norm = norm;
}

int ensure_buffers(npbcg_net *net, int n) {
    NetBuffers &nb = net->nb;
    if (nb.cap >= n) return NPBCG_OK;
    free_buffers(nb);
    const size_t m = (size_t)n;
    NET_HIP(hipMalloc(&nb.x0, sizeof(__bf16) * m * kK0));
    for (int l = 0; l < kLayers; l++) NET_HIP(hipMalloc(&nb.x[l], sizeof(__bf16) * m * kXW));
    NET_HIP(hipMalloc(&nb.h3, sizeof(__bf16) * m * kHid));
    NET_HIP(hipMalloc(&nb.dz, sizeof(__bf16) * m * kHid));
    NET_HIP(hipMalloc(&nb.dx, sizeof(float) * m * kXW));
    NET_HIP(hipMalloc(&nb.gh, sizeof(float) * 2 * m));
    NET_HIP(hipMalloc(&nb.dgh, sizeof(float) * 2 * m));
    NET_HIP(hipMalloc(&nb.pw, sizeof(float) * kPairs * m));
    NET_HIP(hipMalloc(&nb.pd, sizeof(float) * kPairs * m));
    nb.wparts = kWeightParts;
    NET_HIP(hipMalloc(&nb.wpart, sizeof(float) * kWeightParts * kLayerStride));
    NET_HIP(hipMalloc(&nb.rpart, sizeof(float) * row_blocks(n) * kReadout));
    NET_HIP(hipMalloc(&nb.loss_part, sizeof(double) * row_blocks(n)));
    nb.cap = n;
    return NPBCG_OK;
}

// He-normal layer weights (Box-Muller over splitmix64), zero biases and a zero
// readout, so an untrained network gives g = h = 0.
void init_params(std::vector<float> &p, unsigned long long seed) {
    p.assign(kParams, 0.0f);
    uint64_t s = seed ^ 0x5DEECE66Dull;
    auto normal = [&]() {
        double u1 = std::max(uniform01(s), 1e-300), u2 = uniform01(s);
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    };
    const double s0 = std::sqrt(2.0 / kFeat);
    for (int k = 0; k < kFeat; k++)  // rows >= kFeat multiply zero padding
        for (int j = 0; j < kHid; j++) p[kOffW0 + k * kHid + j] = (float)(s0 * normal());
    const double s1 = std::sqrt(2.0 / kXW);
    for (int l = 0; l < kLayers; l++)
        for (int i = 0; i < kXW * kHid; i++) p[kOffW1 + l * kLayerStride + i] = (float)(s1 * normal());
}

// Forward, pair loss (and, when grad is set, backward into grad) of the
// context's resident pairs.  d_loss[0] receives the loss; with base, d_loss[1]
// receives the loss at g = h = 0.  relative divides the gradient by d_loss[1]
// (the loss relative to the Neumann base, so that systems of very different
// conditioning weigh alike in training).
int evaluate(npbcg_net *net, npbcg_ctx *ctx, bool grad, bool base, bool relative = false) {
    NetSystem sys;
    int rc = npbcg_detail::ctx_net_system(ctx, &sys);
    if (rc != NPBCG_OK) return rc;
    const PairSet ps = npbcg_detail::ctx_pairs(ctx);
    if (ps.count < 1) return set_error(NPBCG_ERR_INVALID, "no residual/correction pairs");
    rc = ensure_buffers(net, sys.n);
    if (rc != NPBCG_OK) return rc;
    hipStream_t st = npbcg_detail::ctx_stream(ctx);
    NetBuffers &nb = net->nb;
    if (!net->norm_set) {
        NET_HIP(launch_feature_stats(sys.feat, sys.n, net->norm, st));
        net->norm_set = true;
    }
    const int rb = row_blocks(sys.n);
    if (base) {
        NetBuffers b0 = nb;
        b0.gh = const_cast<float *>(npbcg_detail::ctx_zero_gh(ctx));
        NET_HIP(launch_pair_loss(b0, sys, ps, 0, st));
        NET_HIP(launch_sum(nb.loss_part, rb, net->d_loss + 1, st));
    }
    NET_HIP(launch_net_forward(nb, sys, net->params, net->wb, kGmax, st));
    NET_HIP(launch_pair_loss(nb, sys, ps, grad ? 1 : 0, st));
    NET_HIP(launch_sum(nb.loss_part, rb, net->d_loss, st));
    if (grad) {
        NET_HIP(launch_net_backward(nb, sys, net->params, net->wb, kGmax, net->grad, st));
        if (relative) NET_HIP(launch_scale_inverse(net->grad, kParams, net->d_loss + 1, st));
    }
    return NPBCG_OK;
}

int adam_step(npbcg_net *net, hipStream_t st, float lr) {
    net->adam_step++;
    NET_HIP(launch_adam(net->params, net->grad, net->m1, net->m2, kParams, lr, 0.9f, 0.999f,
                        1e-8f, net->adam_step, st));
    NET_HIP(launch_pack_weights(net->params, net->wb, st));
    return NPBCG_OK;
}

int read_losses(npbcg_net *net, hipStream_t st, double *loss, double *base) {
    NET_HIP(hipMemcpyAsync(net->h_loss, net->d_loss, 2 * sizeof(double), hipMemcpyDeviceToHost, st));
    NET_HIP(hipStreamSynchronize(st));
    if (loss) *loss = net->h_loss[0];
    if (base) *base = net->h_loss[1];
    return NPBCG_OK;
}

// One generated training system.
struct GenSystem {
    int n = 0, nonzer = 0;
    double shift = 0.0, seed = 0.0;
    uint64_t rhs_seed = 0;
    std::vector<int> rowstr, colidx;
    std::vector<double> a;
    int status = 0;
};

void system_params(const npbcg_train_opts &o, int i, GenSystem &g) {
    uint64_t s = o.seed * 0x2545F4914F6CDD1Dull + (uint64_t)i * 0x9E3779B97F4A7C15ull + 1;
    g.n = o.n_min + (int)(uniform01(s) * (o.n_max - o.n_min + 1));
    g.n = std::min(g.n, o.n_max);
    g.nonzer = o.nonzer_min + (int)(uniform01(s) * (o.nonzer_max - o.nonzer_min + 1));
    g.nonzer = std::min(g.nonzer, o.nonzer_max);
    const double w = (double)(g.nonzer + 1) * (g.nonzer + 1);
    g.shift = w * (o.shift_min + uniform01(s) * (o.shift_max - o.shift_min));
    g.seed = (double)(1 + splitmix64(s) % ((1ull << 46) - 2));  // integer in (0, 2^46)
    g.rhs_seed = splitmix64(s);
}

void generate(const npbcg_train_opts &o, int i, GenSystem &g) {
    system_params(o, i, g);
    const long cap = (long)g.n * (g.nonzer + 1) * (g.nonzer + 1);
    g.rowstr.resize(g.n + 1);
    g.colidx.resize(cap);
    g.a.resize(cap);
    npbcg_detail::set_generator_threads(1);
    const long nnz = npbcg_makea(g.n, g.nonzer, o.rcond, g.shift, g.seed, cap, g.rowstr.data(),
                                 g.colidx.data(), g.a.data(), nullptr);
    g.status = nnz < 0 ? (int)nnz : 0;
}

void uniform_rhs(uint64_t seed, int n, std::vector<double> &b) {
    b.resize(n);
    uint64_t s = seed;
    for (int j = 0; j < n; j++) b[j] = 2.0 * uniform01(s) - 1.0;
}

}  // namespace

extern "C" {

int npbcg_net_param_count(void) { return kParams; }

void npbcg_net_default_train_opts(npbcg_train_opts *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->systems = 2000;
    o->n_min = 10000;
    o->n_max = 75000;
    o->nonzer_min = 7;
    o->nonzer_max = 15;
    o->shift_min = 0.3;
    o->shift_max = 0.5;
    o->rcond = 0.1;
    o->seed = 20260101ull;
    o->pairs = NPBCG_MAX_PAIRS;
    o->steps_per_system = 2;
    o->lr = 1e-3f;
    o->gen_threads = 0;
}

int npbcg_net_create(npbcg_net **out, int device, unsigned long long seed) {
    if (!out) return set_error(NPBCG_ERR_INVALID, "npbcg_net_create: NULL");
    *out = nullptr;
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1)
        return set_error(NPBCG_ERR_NODEVICE, "npbcg_net_create: no GPU");
    if (device < 0 || device >= count) device = 0;
    npbcg_net *net = new npbcg_net();
    net->device = device;
    auto fail = [&](int rc) {
        npbcg_net_destroy(net);
        return rc;
    };
#define NET_TRY(call)                                                            \
    do {                                                                         \
        hipError_t _e = (call);                                                  \
        if (_e != hipSuccess)                                                    \
            return fail(set_error(NPBCG_ERR_DEVICE, std::string(#call) + ": " +  \
                                                        hipGetErrorString(_e))); \
    } while (0)
    NET_TRY(hipSetDevice(device));
    NET_TRY(hipStreamCreate(&net->stream));
    NET_TRY(hipMalloc(&net->params, sizeof(float) * kParams));
    NET_TRY(hipMalloc(&net->grad, sizeof(float) * kParams));
    NET_TRY(hipMalloc(&net->m1, sizeof(float) * kParams));
    NET_TRY(hipMalloc(&net->m2, sizeof(float) * kParams));
    NET_TRY(hipMalloc(&net->wb, sizeof(__bf16) * 2 * kWtElems));
    NET_TRY(hipMalloc(&net->norm, sizeof(float) * 2 * kFeat));
    NET_TRY(hipMalloc(&net->d_loss, sizeof(double) * 2));
    NET_TRY(hipHostMalloc(&net->h_loss, sizeof(double) * 2, hipHostMallocDefault));
    NET_TRY(hipEventCreateWithFlags(&net->ev0, hipEventDisableSystemFence));
    NET_TRY(hipEventCreateWithFlags(&net->ev1, hipEventDisableSystemFence));
    net->nb.norm = net->norm;
    std::vector<float> p;
    init_params(p, seed);
    NET_TRY(hipMemcpy(net->params, p.data(), sizeof(float) * kParams, hipMemcpyHostToDevice));
    NET_TRY(hipMemset(net->m1, 0, sizeof(float) * kParams));
    NET_TRY(hipMemset(net->m2, 0, sizeof(float) * kParams));
    NET_TRY(hipMemset(net->grad, 0, sizeof(float) * kParams));
    NET_TRY(hipMemset(net->norm, 0, sizeof(float) * 2 * kFeat));
    NET_TRY(launch_pack_weights(net->params, net->wb, net->stream));
    NET_TRY(hipStreamSynchronize(net->stream));
#undef NET_TRY
    *out = net;
    clear_error();
    return NPBCG_OK;
}

void npbcg_net_destroy(npbcg_net *net) {
    if (!net) return;
    (void)hipSetDevice(net->device);
    if (net->stream) (void)hipStreamSynchronize(net->stream);
    free_buffers(net->nb);
    for (void *p : {(void *)net->params, (void *)net->grad, (void *)net->m1, (void *)net->m2,
                    (void *)net->wb, (void *)net->norm, (void *)net->d_loss})
        (void)hipFree(p);
    if (net->h_loss) (void)hipHostFree(net->h_loss);
    if (net->ev0) (void)hipEventDestroy(net->ev0);
    if (net->ev1) (void)hipEventDestroy(net->ev1);
    if (net->stream) (void)hipStreamDestroy(net->stream);
    delete net;
}

int npbcg_net_get_params(npbcg_net *net, float *params, float *feature_norm) {
    if (!net) return set_error(NPBCG_ERR_INVALID, "npbcg_net_get_params: NULL");
    NET_HIP(hipSetDevice(net->device));
    NET_HIP(hipDeviceSynchronize());
    if (params) NET_HIP(hipMemcpy(params, net->params, sizeof(float) * kParams, hipMemcpyDeviceToHost));
    if (feature_norm) {
        if (net->norm_set)
            NET_HIP(hipMemcpy(feature_norm, net->norm, sizeof(float) * 2 * kFeat,
                              hipMemcpyDeviceToHost));
        else
            std::memset(feature_norm, 0, sizeof(float) * 2 * kFeat);
    }
    clear_error();
    return NPBCG_OK;
}

int npbcg_net_set_params(npbcg_net *net, const float *params, const float *feature_norm) {
    if (!net) return set_error(NPBCG_ERR_INVALID, "npbcg_net_set_params: NULL");
    NET_HIP(hipSetDevice(net->device));
    NET_HIP(hipDeviceSynchronize());
    if (params) {
        NET_HIP(hipMemcpy(net->params, params, sizeof(float) * kParams, hipMemcpyHostToDevice));
        NET_HIP(launch_pack_weights(net->params, net->wb, net->stream));
    }
    if (feature_norm) {
        bool any = false;
        for (int i = 0; i < 2 * kFeat; i++) any |= feature_norm[i] != 0.0f;
        for (int i = kFeat; i < 2 * kFeat && any; i++)
            if (!(feature_norm[i] > 0.0f))
                return set_error(NPBCG_ERR_INVALID, "npbcg_net_set_params: std must be > 0");
        NET_HIP(hipMemcpy(net->norm, feature_norm, sizeof(float) * 2 * kFeat,
                          hipMemcpyHostToDevice));
        net->norm_set = any;
    }
    NET_HIP(hipStreamSynchronize(net->stream));
    clear_error();
    return NPBCG_OK;
}

int npbcg_net_loss_grad(npbcg_net *net, npbcg_ctx *ctx, double *loss, float *grad, float *gh) {
    if (!net || !ctx) return set_error(NPBCG_ERR_INVALID, "npbcg_net_loss_grad: NULL");
    if (npbcg_detail::ctx_device(ctx) != net->device)
        return set_error(NPBCG_ERR_INVALID, "npbcg_net_loss_grad: context on another device");
    NET_HIP(hipSetDevice(net->device));
    NET_HIP(hipStreamSynchronize(net->stream));
    int rc = evaluate(net, ctx, grad != nullptr, false);
    if (rc != NPBCG_OK) return rc;
    hipStream_t st = npbcg_detail::ctx_stream(ctx);
    const int n = npbcg_detail::ctx_n(ctx);
    if (grad) NET_HIP(hipMemcpyAsync(grad, net->grad, sizeof(float) * kParams, hipMemcpyDeviceToHost, st));
    if (gh) NET_HIP(hipMemcpyAsync(gh, net->nb.gh, sizeof(float) * 2 * n, hipMemcpyDeviceToHost, st));
    rc = read_losses(net, st, loss, nullptr);
    if (rc != NPBCG_OK) return rc;
    clear_error();
    return NPBCG_OK;
}

int npbcg_net_step(npbcg_net *net, npbcg_ctx *ctx, float lr, double *loss) {
    npbcg_detail::TraceRange trace("npbcg_net_step");
    if (!net || !ctx || !(lr > 0.0f)) return set_error(NPBCG_ERR_INVALID, "npbcg_net_step: invalid argument");
    if (npbcg_detail::ctx_device(ctx) != net->device)
        return set_error(NPBCG_ERR_INVALID, "npbcg_net_step: context on another device");
    NET_HIP(hipSetDevice(net->device));
    NET_HIP(hipStreamSynchronize(net->stream));
    int rc = evaluate(net, ctx, true, false);
    if (rc != NPBCG_OK) return rc;
    hipStream_t st = npbcg_detail::ctx_stream(ctx);
    rc = adam_step(net, st, lr);
    if (rc != NPBCG_OK) return rc;
    rc = read_losses(net, st, loss, nullptr);
    if (rc != NPBCG_OK) return rc;
    clear_error();
    return NPBCG_OK;
}

int npbcg_net_apply(npbcg_net *net, npbcg_ctx *ctx, double *ms) {
    npbcg_detail::TraceRange trace("npbcg_net_apply");
    if (!net || !ctx) return set_error(NPBCG_ERR_INVALID, "npbcg_net_apply: NULL");
    if (npbcg_detail::ctx_device(ctx) != net->device)
        return set_error(NPBCG_ERR_INVALID, "npbcg_net_apply: context on another device");
    NET_HIP(hipSetDevice(net->device));
    NET_HIP(hipStreamSynchronize(net->stream));
    NetSystem sys;
    int rc = npbcg_detail::ctx_net_system(ctx, &sys);
    if (rc != NPBCG_OK) return rc;
    rc = ensure_buffers(net, sys.n);
    if (rc != NPBCG_OK) return rc;
    hipStream_t st = npbcg_detail::ctx_stream(ctx);
    NET_HIP(hipEventRecord(net->ev0, st));
    if (!net->norm_set) {
        NET_HIP(launch_feature_stats(sys.feat, sys.n, net->norm, st));
        net->norm_set = true;
    }
    NET_HIP(launch_net_forward(net->nb, sys, net->params, net->wb, kGmax, st));
    NET_HIP(hipEventRecord(net->ev1, st));
    rc = npbcg_detail::ctx_install_learned(ctx, net->nb.gh);
    if (rc != NPBCG_OK) return rc;
    NET_HIP(hipStreamSynchronize(st));
    float t = 0.f;
    NET_HIP(hipEventElapsedTime(&t, net->ev0, net->ev1));
    if (ms) *ms = t;
    clear_error();
    return NPBCG_OK;
}

int npbcg_net_train(npbcg_net *net, const npbcg_train_opts *opts, npbcg_train_stats *stats,
                    double *loss_hist, double *base_hist) {
    npbcg_detail::TraceRange trace("npbcg_net_train");
    if (!net || !opts) return set_error(NPBCG_ERR_INVALID, "npbcg_net_train: NULL");
    const npbcg_train_opts o = *opts;
    if (o.systems < 1 || o.n_min < 64 || o.n_max < o.n_min || o.nonzer_min < 1 ||
        o.nonzer_max < o.nonzer_min || o.nonzer_max >= o.n_min || !(o.shift_min > 0.0) ||
        o.shift_max < o.shift_min || !(o.rcond > 0.0) || o.pairs < 1 ||
        o.pairs > NPBCG_MAX_PAIRS || o.steps_per_system < 1 || !(o.lr > 0.0f) ||
        o.first_system < 0)
        return set_error(NPBCG_ERR_INVALID, "npbcg_net_train: invalid options");
    NET_HIP(hipSetDevice(net->device));
    const auto t0 = std::chrono::steady_clock::now();
    int threads = o.gen_threads > 0 ? o.gen_threads
                                    : (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency()));
    threads = std::min(threads, o.systems);
    const int window = 2 * threads;

    // generator pipeline: system i lands in slot i % window
    std::vector<GenSystem> slot(window);
    std::vector<int> slot_of(window, -1);
    std::mutex mu;
    std::condition_variable cv;
    std::atomic<int> next{0};
    int consumed = 0;
    bool stop = false;
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++)
        pool.emplace_back([&]() {
            for (;;) {
                const int i = next.fetch_add(1);
                if (i >= o.systems) return;
                {
                    std::unique_lock<std::mutex> lk(mu);
                    cv.wait(lk, [&] { return stop || i < consumed + window; });
                    if (stop) return;
                }
                GenSystem g;
                generate(o, o.first_system + i, g);
                {
                    std::lock_guard<std::mutex> lk(mu);
                    slot[i % window] = std::move(g);
                    slot_of[i % window] = i;
                }
                cv.notify_all();
            }
        });
    auto shutdown = [&]() {
        {
            std::lock_guard<std::mutex> lk(mu);
            stop = true;
        }
        cv.notify_all();
        for (auto &th : pool) th.join();
    };

    npbcg_train_stats s{};
    double wait_ms = 0.0;
    const int win = std::max(1, o.systems / 10);
    double lf = 0, ll = 0, bf = 0, bl = 0;
    int nf = 0, nl = 0;
    int rc = NPBCG_OK;
    std::vector<double> b;
    for (int i = 0; i < o.systems && rc == NPBCG_OK; i++) {
        GenSystem g;
        {
            const auto w0 = std::chrono::steady_clock::now();
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] { return slot_of[i % window] == i; });
            g = std::move(slot[i % window]);
            slot_of[i % window] = -1;
            consumed = i + 1;
            wait_ms += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
        }
        cv.notify_all();
        if (g.status != 0) {
            rc = set_error(NPBCG_ERR_INTERNAL, "npbcg_net_train: generator failed");
            break;
        }
        npbcg_opts co;
        npbcg_default_opts(&co);
        co.device = net->device;
        co.spmv_variant = 1;  // CSR rows: no host-side layout build per system
        co.use_graph = 0;
        npbcg_ctx *ctx = nullptr;
        rc = npbcg_create(&ctx, g.n, g.rowstr.data(), g.colidx.data(), g.a.data(), &co);
        if (rc != NPBCG_OK) break;
        uniform_rhs(g.rhs_seed, g.n, b);
        rc = npbcg_set_rhs(ctx, b.data());
        if (rc == NPBCG_OK) rc = npbcg_log_pairs(ctx, o.pairs, nullptr, nullptr);
        if (rc == NPBCG_ERR_NOCONV) {  // no x* for this system: leave it out
            npbcg_destroy(ctx);
            rc = NPBCG_OK;
            s.skipped++;
            if (loss_hist) loss_hist[i] = std::nan("");
            if (base_hist) base_hist[i] = std::nan("");
            continue;
        }
        double loss = 0, base = 0;
        for (int k = 0; k < o.steps_per_system && rc == NPBCG_OK; k++) {
            rc = evaluate(net, ctx, true, k == 0, true);
            if (rc != NPBCG_OK) break;
            if (k == 0) {
                rc = read_losses(net, npbcg_detail::ctx_stream(ctx), &loss, &base);
                if (rc != NPBCG_OK) break;
            }
            rc = adam_step(net, npbcg_detail::ctx_stream(ctx), o.lr);
            s.steps++;
        }
        if (rc == NPBCG_OK) {
            hipError_t e = hipStreamSynchronize(npbcg_detail::ctx_stream(ctx));
            if (e != hipSuccess) rc = set_error(NPBCG_ERR_DEVICE, hipGetErrorString(e));
        }
        npbcg_destroy(ctx);
        if (rc != NPBCG_OK) break;
        if (loss_hist) loss_hist[i] = loss;
        if (base_hist) base_hist[i] = base;
        if (i < win) {
            lf += loss;
            bf += base;
            nf++;
        }
        if (i >= o.systems - win) {
            ll += loss;
            bl += base;
            nl++;
        }
        s.systems++;
    }
    shutdown();
    if (rc != NPBCG_OK) return rc;
    s.loss_first = nf ? lf / nf : 0.0;
    s.loss_last = nl ? ll / nl : 0.0;
    s.base_first = nf ? bf / nf : 0.0;
    s.base_last = nl ? bl / nl : 0.0;
    s.wait_ms = wait_ms;
    s.total_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (stats) *stats = s;
    clear_error();
    return NPBCG_OK;
}

}  // extern "C"
