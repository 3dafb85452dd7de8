// Host side of batched solves: block-diagonal upload, tiling, the FP64
// refinement loop around the batched inner CG, and the SpMM autotuner with
// its persisted record file.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "batch_kernels.h"
#include "npbcg.h"
#include "npbcg_internal.h"

using namespace npbcg_dev;
using npbcg_detail::clear_error;
using npbcg_detail::set_error;

#define B_HIP(call)                                                              \
    do {                                                                         \
        hipError_t _e = (call);                                                  \
        if (_e != hipSuccess)                                                    \
            return set_error(NPBCG_ERR_DEVICE, std::string(#call) + ": " +       \
                                                   hipGetErrorString(_e));       \
    } while (0)

struct npbcg_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    npbcg_batch_opts opts{};
    int nsys = 0, nr = 1, rows = 0;
    long long nnz = 0;
    std::vector<int> n, row_off;
    std::vector<long long> sys_nnz;
    bool rhs_set = false;

    int *d_rowptr = nullptr, *d_col = nullptr;
    double *d_val = nullptr;
    float *d_val32 = nullptr;
    int *d_tiles = nullptr;          // tile_sys | tile_row | sys_tile
    int tiles_cap = 0;
    int *d_block_tile = nullptr;     // workgroup -> tile of the gathering kernels
    size_t block_tile_cap = 0;
    // column-blocked SpMM
    int cbw = 0;                     // column block width
    std::vector<int> sys_nb;         // blocks per system
    int *d_items = nullptr;          // [items][4]
    int items_cap = 0;
    int *d_sysinfo = nullptr;        // sys_item | sys_row | sys_nb
    long long *d_sys64 = nullptr;    // sys_seg | sys_part
    int *d_seg = nullptr;            // per-row block offsets
    void *d_rows_part = nullptr;     // per-block row partials (inner type)
    int *d_flags = nullptr;          // done | done_next
    BatchScalars *d_sc = nullptr;
    double *d_b = nullptr, *d_x = nullptr, *d_res = nullptr;   // [rows][nr]
    void *d_inner = nullptr;         // d, r, p, q in the inner type
    double *d_part = nullptr;        // 4 * part_cap * nr (init, pq, rr0, rr1)
    int part_cap = 0;                // max(tiles, items) at the smallest tile
    double *d_sysout = nullptr;      // nsys * nr
    BatchDev bd;
};

namespace {

// Cuts every system into tiles of `rt` rows (a system's last tile is shorter).
void make_tiles(const npbcg_batch *b, int rt, std::vector<int> &tile_sys, std::vector<int> &tile_row,
                std::vector<int> &sys_tile) {
    tile_sys.clear();
    tile_row.clear();
    sys_tile.assign(b->nsys + 1, 0);
    for (int s = 0; s < b->nsys; s++) {
        sys_tile[s] = (int)tile_sys.size();
        for (int r = 0; r < b->n[s]; r += rt) {
            tile_sys.push_back(s);
            tile_row.push_back(b->row_off[s] + r);
        }
    }
    sys_tile[b->nsys] = (int)tile_sys.size();
    tile_row.push_back(b->rows);
}

int max_tiles(const npbcg_batch *b, int rt) {
    int t = 0;
    for (int s = 0; s < b->nsys; s++) t += (b->n[s] + rt - 1) / rt;
    return t;
}

constexpr int kMinTile = 64;
constexpr int kXcds = 8;  // MI355X: workgroup b of a launch runs on XCD b % 8

// Workgroup -> tile order of the gathering kernels.  With at least kXcds
// systems, whole systems are placed on XCDs (greedy by nonzeros, largest
// first) so that a system's vector is gathered from one XCD's L2; the lists
// are interleaved to match the round-robin dispatch and padded with -1.
// Fewer systems keep the plain order.
std::vector<int> xcd_order(const npbcg_batch *b, const std::vector<int> &sys_tile,
                           const std::vector<long long> &sys_nnz) {
    const int tiles = sys_tile[b->nsys];
    std::vector<int> out;
    if (b->nsys < kXcds) {
        out.resize(tiles);
        for (int t = 0; t < tiles; t++) out[t] = t;
        return out;
    }
    std::vector<int> order(b->nsys);
    for (int s = 0; s < b->nsys; s++) order[s] = s;
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return sys_nnz[x] > sys_nnz[y]; });
    std::vector<long long> load(kXcds, 0);
    std::vector<std::vector<int>> lists(kXcds);
    for (int s : order) {
        const int x = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        load[x] += sys_nnz[s];
        for (int t = sys_tile[s]; t < sys_tile[s + 1]; t++) lists[x].push_back(t);
    }
    size_t len = 0;
    for (auto &l : lists) len = std::max(len, l.size());
    out.assign(len * kXcds, -1);
    for (int x = 0; x < kXcds; x++)
        for (size_t i = 0; i < lists[x].size(); i++) out[i * kXcds + x] = lists[x][i];
    return out;
}

int apply_shape(npbcg_batch *b, int lpr, int rt, int layout) {
    std::vector<int> ts, tr, st;
    make_tiles(b, rt, ts, tr, st);
    if (layout == 2) {  // items: (system, column block, row range) in system order
        std::vector<int> items, sys_item(b->nsys + 1, 0);
        for (int s = 0; s < b->nsys; s++) {
            sys_item[s] = (int)(items.size() / 4);
            for (int blk = 0; blk < b->sys_nb[s]; blk++)
                for (int r = 0; r < b->n[s]; r += rt) {
                    items.push_back(s);
                    items.push_back(blk);
                    items.push_back(b->row_off[s] + r);
                    items.push_back(b->row_off[s] + std::min(b->n[s], r + rt));
                }
        }
        sys_item[b->nsys] = (int)(items.size() / 4);
        const int ni = sys_item[b->nsys];
        if (ni > b->items_cap) return set_error(NPBCG_ERR_INTERNAL, "batch: item capacity");
        B_HIP(hipMemcpyAsync(b->d_items, items.data(), sizeof(int) * items.size(),
                             hipMemcpyHostToDevice, b->stream));
        B_HIP(hipMemcpyAsync(b->d_sysinfo, sys_item.data(), sizeof(int) * (b->nsys + 1),
                             hipMemcpyHostToDevice, b->stream));
        B_HIP(hipStreamSynchronize(b->stream));
        b->bd.items = ni;
    }
    b->bd.cb = layout == 2 ? 1 : 0;
    b->opts.layout = layout;
    const int tiles = (int)ts.size();
    if (tiles > b->tiles_cap) return set_error(NPBCG_ERR_INTERNAL, "batch: tile capacity");
    std::vector<int> all;
    all.reserve(2 * tiles + 1 + b->nsys + 1);
    all.insert(all.end(), ts.begin(), ts.end());
    all.insert(all.end(), tr.begin(), tr.end());
    all.insert(all.end(), st.begin(), st.end());
    B_HIP(hipMemcpyAsync(b->d_tiles, all.data(), sizeof(int) * all.size(), hipMemcpyHostToDevice,
                         b->stream));
    B_HIP(hipStreamSynchronize(b->stream));
    const std::vector<int> bt = xcd_order(b, st, b->sys_nnz);
    if (bt.size() > b->block_tile_cap) {
        (void)hipFree(b->d_block_tile);
        b->d_block_tile = nullptr;
        b->block_tile_cap = 0;
        B_HIP(hipMalloc(&b->d_block_tile, sizeof(int) * bt.size()));
        b->block_tile_cap = bt.size();
    }
    B_HIP(hipMemcpyAsync(b->d_block_tile, bt.data(), sizeof(int) * bt.size(), hipMemcpyHostToDevice,
                         b->stream));
    B_HIP(hipStreamSynchronize(b->stream));
    b->bd.blocks = (int)bt.size();
    b->bd.block_tile = b->d_block_tile;
    b->bd.tiles = tiles;
    b->bd.tile_sys = b->d_tiles;
    b->bd.tile_row = b->d_tiles + tiles;
    b->bd.sys_tile = b->d_tiles + 2 * tiles + 1;
    b->bd.lpr = lpr;
    b->opts.lanes_per_row = lpr;
    b->opts.rows_per_tile = rt;
    return NPBCG_OK;
}

// Sizes of the column-blocked layout for the batch's inner type: blocks of
// cbw columns whose nrhs-wide vector entries fill lds bytes.
int cb_width(const npbcg_batch *b, int device) {
    int optin = 0;
    if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, device) !=
            hipSuccess || optin <= 0)
        optin = 64 * 1024;
    const int elem = b->opts.mode == NPBCG_SOLVE_MIXED ? 4 : 8;
    const int want = b->opts.mode == NPBCG_SOLVE_MIXED ? 64 * 1024 : 128 * 1024;
    const int lds = std::min(want, optin - 2048);
    return std::max(64, lds / (elem * b->nr));
}

template <typename T>
BatchVecs<T> inner_vecs(const npbcg_batch *b) {
    BatchVecs<T> v;
    const size_t m = (size_t)b->rows * b->nr;
    T *base = static_cast<T *>(b->d_inner);
    v.val = std::is_same<T, float>::value ? (const T *)(const void *)b->d_val32
                                          : (const T *)(const void *)b->d_val;
    v.d = base;
    v.r = base + m;
    v.p = base + 2 * m;
    v.q = base + 3 * m;
    const size_t pt = (size_t)b->part_cap * b->nr;
    v.rows_part = static_cast<T *>(b->d_rows_part);
    v.part_init = b->d_part;
    v.part_pq = b->d_part + pt;
    v.part_rr[0] = b->d_part + 2 * pt;
    v.part_rr[1] = b->d_part + 3 * pt;
    return v;
}

std::string shape_key(const npbcg_batch *b) {
    char buf[256];
    std::snprintf(buf, sizeof buf, "batch:sys=%d:rows=%d:nnz=%lld:nrhs=%d:mode=%d", b->nsys,
                  b->rows, b->nnz, b->nr, b->opts.mode);
    return buf;
}

// Record file: one line per tuning, "<shape key> <lanes> <rows> <layout> <ms>";
// the last record of a key wins.
bool cache_lookup(const char *path, const std::string &key, int *lpr, int *rt, int *layout,
                  double *ms) {
    if (!path) return false;
    FILE *f = std::fopen(path, "r");
    if (!f) return false;
    char k[512];
    int a = 0, c = 0, l = 0;
    double t = 0;
    bool found = false;
    while (std::fscanf(f, "%511s %d %d %d %lf", k, &a, &c, &l, &t) == 5)
        if (key == k) {
            *lpr = a;
            *rt = c;
            *layout = l;
            *ms = t;
            found = true;
        }
    std::fclose(f);
    return found;
}

void cache_store(const char *path, const std::string &key, int lpr, int rt, int layout,
                 double ms) {
    if (!path) return;
    FILE *f = std::fopen(path, "a");
    if (!f) return;
    std::fprintf(f, "%s %d %d %d %.6f\n", key.c_str(), lpr, rt, layout, ms);
    std::fclose(f);
}

template <typename T>
int time_spmm(npbcg_batch *b, int reps, double *ms) {
    const BatchVecs<T> v = inner_vecs<T>(b);
    B_HIP(batch_spmm<T>(b->bd, v, b->stream));  // warm
    B_HIP(hipEventRecord(b->ev0, b->stream));
    for (int i = 0; i < reps; i++) B_HIP(batch_spmm<T>(b->bd, v, b->stream));
    B_HIP(hipEventRecord(b->ev1, b->stream));
    B_HIP(hipEventSynchronize(b->ev1));
    float t = 0;
    B_HIP(hipEventElapsedTime(&t, b->ev0, b->ev1));
    *ms = t / reps;
    return NPBCG_OK;
}

// Defaults without tuning: lanes by average row length, 256-row tiles.
void default_shape(const npbcg_batch *b, int *lpr, int *rt) {
    const double avg = b->rows ? (double)b->nnz / b->rows : 0.0;
    *lpr = avg >= 96 ? 32 : avg >= 40 ? 16 : avg >= 12 ? 8 : 4;
    *rt = 256;
}

template <typename T>
int run_inner(npbcg_batch *b, int maxsteps) {
    const BatchVecs<T> v = inner_vecs<T>(b);
    B_HIP(batch_init<T>(b->bd, v, b->d_res, b->stream));
    const int chunk = 8;
    for (int step = 0; step < maxsteps;) {
        const int end = std::min(maxsteps, step + chunk);
        for (; step < end; step++) B_HIP(batch_cg_step<T>(b->bd, v, step, step == 0, b->stream));
        // stop early when every solve has finished
        std::vector<int> done(b->nsys * b->nr);
        B_HIP(hipMemcpyAsync(done.data(), b->bd.done, sizeof(int) * done.size(),
                             hipMemcpyDeviceToHost, b->stream));
        B_HIP(hipStreamSynchronize(b->stream));
        bool all = true;
        for (int f : done) all = all && f;
        if (all) break;
    }
    B_HIP(batch_refine<T>(b->bd, b->d_x, v.d, b->stream));
    return NPBCG_OK;
}

}  // namespace

extern "C" {

void npbcg_batch_default_opts(npbcg_batch_opts *o) {
    if (!o) return;
    std::memset(o, 0, sizeof(*o));
    o->nrhs = 1;
    o->mode = NPBCG_SOLVE_MIXED;
}

void npbcg_batch_destroy(npbcg_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (void *p : {(void *)b->d_items, (void *)b->d_sysinfo, (void *)b->d_sys64, (void *)b->d_seg,
                    b->d_rows_part, (void *)b->d_rowptr, (void *)b->d_col, (void *)b->d_val,
                    (void *)b->d_val32,
                    (void *)b->d_tiles, (void *)b->d_block_tile, (void *)b->d_flags, (void *)b->d_sc, (void *)b->d_b,
                    (void *)b->d_x, (void *)b->d_res, b->d_inner, (void *)b->d_part,
                    (void *)b->d_sysout})
        (void)hipFree(p);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int npbcg_batch_create(npbcg_batch **out, int nsys, const int *n, const int *const *rowstr,
                       const int *const *colidx, const double *const *a,
                       const npbcg_batch_opts *opts) {
    if (!out || nsys < 1 || !n || !rowstr || !colidx || !a)
        return set_error(NPBCG_ERR_INVALID, "npbcg_batch_create: invalid argument");
    *out = nullptr;
    npbcg_batch_opts o;
    npbcg_batch_default_opts(&o);
    if (opts) o = *opts;
    if ((o.nrhs != 1 && o.nrhs != 2 && o.nrhs != 4) ||
        (o.mode != NPBCG_SOLVE_FP64 && o.mode != NPBCG_SOLVE_MIXED) || o.layout < 0 ||
        o.layout > 2)
        return set_error(NPBCG_ERR_INVALID,
                         "npbcg_batch_create: nrhs must be 1, 2 or 4; mode and layout valid");
    long long rows = 0, nnz = 0;
    for (int s = 0; s < nsys; s++) {
        if (n[s] < 1 || !rowstr[s] || !colidx[s] || !a[s] || rowstr[s][0] != 0)
            return set_error(NPBCG_ERR_INVALID, "npbcg_batch_create: malformed system");
        for (int j = 0; j < n[s]; j++)
            if (rowstr[s][j + 1] < rowstr[s][j])
                return set_error(NPBCG_ERR_INVALID, "npbcg_batch_create: rowstr not monotone");
        for (long long k = 0; k < rowstr[s][n[s]]; k++)
            if (colidx[s][k] < 0 || colidx[s][k] >= n[s])
                return set_error(NPBCG_ERR_INVALID, "npbcg_batch_create: column outside the system");
        rows += n[s];
        nnz += rowstr[s][n[s]];
    }
    if (rows > (1ll << 30) / o.nrhs || nnz >= (1ll << 31))
        return set_error(NPBCG_ERR_CAPACITY, "npbcg_batch_create: batch too large for 32-bit indices");
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count < 1)
        return set_error(NPBCG_ERR_NODEVICE, "npbcg_batch_create: no GPU");
    npbcg_batch *b = new npbcg_batch();
    b->device = (o.device >= 0 && o.device < count) ? o.device : 0;
    b->opts = o;
    b->nsys = nsys;
    b->nr = o.nrhs;
    b->rows = (int)rows;
    b->nnz = nnz;
    b->n.assign(n, n + nsys);
    b->row_off.assign(nsys + 1, 0);
    for (int s = 0; s < nsys; s++) b->row_off[s + 1] = b->row_off[s] + n[s];
    b->sys_nnz.resize(nsys);
    for (int s = 0; s < nsys; s++) b->sys_nnz[s] = rowstr[s][n[s]];
    auto fail = [&](int rc) {
        npbcg_batch_destroy(b);
        return rc;
    };
#define TRY(call)                                                                \
    do {                                                                         \
        hipError_t _e = (call);                                                  \
        if (_e != hipSuccess)                                                    \
            return fail(set_error(NPBCG_ERR_DEVICE, std::string(#call) + ": " +  \
                                                        hipGetErrorString(_e))); \
    } while (0)
    TRY(hipSetDevice(b->device));
    TRY(hipStreamCreate(&b->stream));
    TRY(hipEventCreateWithFlags(&b->ev0, hipEventDisableSystemFence));
    TRY(hipEventCreateWithFlags(&b->ev1, hipEventDisableSystemFence));
    // global CSR: rows concatenated, columns offset by the system's first
    // row, each row's entries in column order (the column blocks need it)
    std::vector<int> rp((size_t)rows + 1), col((size_t)nnz);
    std::vector<double> val((size_t)nnz);
    std::vector<float> val32((size_t)nnz);
    b->cbw = cb_width(b, b->device);
    b->sys_nb.resize(nsys);
    std::vector<long long> sys_seg(nsys), sys_part(nsys);
    long long seg_total = 0, part_total = 0;
    for (int s = 0; s < nsys; s++) {
        b->sys_nb[s] = (n[s] + b->cbw - 1) / b->cbw;
        sys_seg[s] = seg_total;
        sys_part[s] = part_total;
        seg_total += (long long)n[s] * (b->sys_nb[s] + 1);
        part_total += (long long)n[s] * b->sys_nb[s] * b->nr;
    }
    std::vector<int> seg((size_t)seg_total);
    long long e = 0;
    std::vector<int> order;
    for (int s = 0; s < nsys; s++) {
        const int off = b->row_off[s], nb = b->sys_nb[s];
        for (int j = 0; j < n[s]; j++) {
            rp[off + j] = (int)e;
            const int k0 = rowstr[s][j], k1 = rowstr[s][j + 1];
            order.resize(k1 - k0);
            for (int k = k0; k < k1; k++) order[k - k0] = k;
            std::stable_sort(order.begin(), order.end(),
                             [&](int x, int y) { return colidx[s][x] < colidx[s][y]; });
            int* sg = &seg[(size_t)(sys_seg[s] + (long long)j * (nb + 1))];
            int blk = 0;
            for (int k : order) {
                const int c = colidx[s][k];
                while (blk < nb && c >= blk * b->cbw) sg[blk++] = (int)e;
                col[e] = c + off;
                val[e] = a[s][k];
                val32[e] = (float)a[s][k];
                e++;
            }
            while (blk <= nb) sg[blk++] = (int)e;
        }
    }
    rp[rows] = (int)e;
    const size_t m = (size_t)rows * b->nr;
    b->tiles_cap = max_tiles(b, kMinTile);
    b->items_cap = 0;
    for (int s = 0; s < nsys; s++) b->items_cap += b->sys_nb[s] * ((n[s] + kMinTile - 1) / kMinTile);
    b->part_cap = std::max(b->tiles_cap, b->items_cap);
    std::vector<int> sysinfo(3 * (size_t)nsys + 2);  // sys_item (filled per shape) | sys_row | sys_nb
    for (int s = 0; s <= nsys; s++) sysinfo[nsys + 1 + s] = b->row_off[s];
    for (int s = 0; s < nsys; s++) sysinfo[2 * nsys + 2 + s] = b->sys_nb[s];
    std::vector<long long> sys64(2 * (size_t)nsys);
    for (int s = 0; s < nsys; s++) {
        sys64[s] = sys_seg[s];
        sys64[nsys + s] = sys_part[s];
    }
    TRY(hipMalloc(&b->d_rowptr, sizeof(int) * (rows + 1)));
    TRY(hipMalloc(&b->d_col, sizeof(int) * std::max<long long>(nnz, 1)));
    TRY(hipMalloc(&b->d_val, sizeof(double) * std::max<long long>(nnz, 1)));
    if (o.mode == NPBCG_SOLVE_MIXED)
        TRY(hipMalloc(&b->d_val32, sizeof(float) * std::max<long long>(nnz, 1)));
    TRY(hipMalloc(&b->d_tiles, sizeof(int) * (2 * (size_t)b->tiles_cap + nsys + 2)));
    TRY(hipMalloc(&b->d_flags, sizeof(int) * 2 * nsys * b->nr));
    TRY(hipMalloc(&b->d_sc, sizeof(BatchScalars) * nsys * b->nr));
    TRY(hipMalloc(&b->d_b, sizeof(double) * m));
    TRY(hipMalloc(&b->d_x, sizeof(double) * m));
    TRY(hipMalloc(&b->d_res, sizeof(double) * m));
    const size_t inner_elem = o.mode == NPBCG_SOLVE_MIXED ? sizeof(float) : sizeof(double);
    TRY(hipMalloc(&b->d_inner, inner_elem * 4 * m));
    TRY(hipMalloc(&b->d_part, sizeof(double) * 4 * (size_t)b->part_cap * b->nr));
    TRY(hipMalloc(&b->d_items, sizeof(int) * 4 * (size_t)std::max(b->items_cap, 1)));
    TRY(hipMalloc(&b->d_sysinfo, sizeof(int) * sysinfo.size()));
    TRY(hipMalloc(&b->d_sys64, sizeof(long long) * sys64.size()));
    TRY(hipMalloc(&b->d_seg, sizeof(int) * std::max<long long>(seg_total, 1)));
    TRY(hipMalloc(&b->d_rows_part, inner_elem * std::max<long long>(part_total, 1)));
    TRY(hipMemcpyAsync(b->d_sysinfo, sysinfo.data(), sizeof(int) * sysinfo.size(),
                       hipMemcpyHostToDevice, b->stream));
    TRY(hipMemcpyAsync(b->d_sys64, sys64.data(), sizeof(long long) * sys64.size(),
                       hipMemcpyHostToDevice, b->stream));
    TRY(hipMemcpyAsync(b->d_seg, seg.data(), sizeof(int) * seg_total, hipMemcpyHostToDevice,
                       b->stream));
    TRY(hipMalloc(&b->d_sysout, sizeof(double) * nsys * b->nr));
    TRY(hipMemcpyAsync(b->d_rowptr, rp.data(), sizeof(int) * (rows + 1), hipMemcpyHostToDevice,
                       b->stream));
    TRY(hipMemcpyAsync(b->d_col, col.data(), sizeof(int) * nnz, hipMemcpyHostToDevice, b->stream));
    TRY(hipMemcpyAsync(b->d_val, val.data(), sizeof(double) * nnz, hipMemcpyHostToDevice, b->stream));
    if (b->d_val32)
        TRY(hipMemcpyAsync(b->d_val32, val32.data(), sizeof(float) * nnz, hipMemcpyHostToDevice,
                           b->stream));
    TRY(hipMemsetAsync(b->d_x, 0, sizeof(double) * m, b->stream));
    TRY(hipMemsetAsync(b->d_inner, 0, inner_elem * 4 * m, b->stream));
    TRY(hipMemsetAsync(b->d_flags, 0, sizeof(int) * 2 * nsys * b->nr, b->stream));
    TRY(hipMemsetAsync(b->d_sc, 0, sizeof(BatchScalars) * nsys * b->nr, b->stream));
    TRY(hipStreamSynchronize(b->stream));  // host staging goes out of scope
#undef TRY
    b->bd.nsys = nsys;
    b->bd.nr = b->nr;
    b->bd.rows = b->rows;
    b->bd.rowptr = b->d_rowptr;
    b->bd.col = b->d_col;
    b->bd.done = b->d_flags;
    b->bd.done_next = b->d_flags + nsys * b->nr;
    b->bd.sc = b->d_sc;
    b->bd.cbw = b->cbw;
    b->bd.item = b->d_items;
    b->bd.sys_item = b->d_sysinfo;
    b->bd.sys_row = b->d_sysinfo + nsys + 1;
    b->bd.sys_nb = b->d_sysinfo + 2 * nsys + 2;
    b->bd.sys_seg = b->d_sys64;
    b->bd.sys_part = b->d_sys64 + nsys;
    b->bd.seg = b->d_seg;
    {
        const int lds = b->cbw * b->nr * (o.mode == NPBCG_SOLVE_MIXED ? 4 : 8);
        hipError_t e2 = o.mode == NPBCG_SOLVE_MIXED ? batch_prepare_cb<float>(lds)
                                                    : batch_prepare_cb<double>(lds);
        if (e2 != hipSuccess)
            return fail(set_error(NPBCG_ERR_DEVICE, std::string("batch_prepare_cb: ") +
                                                        hipGetErrorString(e2)));
    }
    int lpr = o.lanes_per_row, rt = o.rows_per_tile;
    int dl, dr;
    default_shape(b, &dl, &dr);
    if (lpr == 0) lpr = dl;
    if (rt == 0) rt = dr;
    if ((lpr & (lpr - 1)) || lpr < 4 || lpr > 64 || rt < kMinTile || rt > 4096)
        return fail(set_error(NPBCG_ERR_INVALID, "npbcg_batch_create: invalid launch shape"));
    int rc = apply_shape(b, lpr, rt, o.layout == 1 ? 1 : 2);
    if (rc != NPBCG_OK) return fail(rc);
    *out = b;
    clear_error();
    return NPBCG_OK;
}

int npbcg_batch_set_rhs(npbcg_batch *b, const double *rhs) {
    if (!b || !rhs) return set_error(NPBCG_ERR_INVALID, "npbcg_batch_set_rhs: NULL");
    B_HIP(hipSetDevice(b->device));
    B_HIP(hipMemcpyAsync(b->d_b, rhs, sizeof(double) * (size_t)b->rows * b->nr,
                         hipMemcpyHostToDevice, b->stream));
    B_HIP(hipStreamSynchronize(b->stream));
    b->rhs_set = true;
    clear_error();
    return NPBCG_OK;
}

int npbcg_batch_tune(npbcg_batch *b, const char *cache_path, int *lanes_per_row,
                     int *rows_per_tile, double *ms_out) {
    if (!b) return set_error(NPBCG_ERR_INVALID, "npbcg_batch_tune: NULL");
    B_HIP(hipSetDevice(b->device));
    const std::string key = shape_key(b);
    int lpr = 0, rt = 0, layout = 2;
    double best = 0.0;
    if (!cache_lookup(cache_path, key, &lpr, &rt, &layout, &best)) {
        // All solves marked running so every tile does its work; vectors hold
        // whatever they hold (timing only); state is reset afterwards.
        B_HIP(hipMemsetAsync(b->d_flags, 0, sizeof(int) * 2 * b->nsys * b->nr, b->stream));
        best = 1e30;
        for (int cy : {1, 2})
            for (int cl : {4, 8, 16, 32, 64})
                for (int cr : {64, 128, 256, 512, 1024, 2048, 4096}) {
                    int rc = apply_shape(b, cl, cr, cy);
                    if (rc != NPBCG_OK) return rc;
                    double t = 0;
                    rc = b->opts.mode == NPBCG_SOLVE_MIXED ? time_spmm<float>(b, 5, &t)
                                                           : time_spmm<double>(b, 5, &t);
                    if (rc != NPBCG_OK) return rc;
                    if (t < best) {
                        best = t;
                        lpr = cl;
                        rt = cr;
                        layout = cy;
                    }
                }
        cache_store(cache_path, key, lpr, rt, layout, best);
    }
    int rc = apply_shape(b, lpr, rt, layout);
    if (rc != NPBCG_OK) return rc;
    if (lanes_per_row) *lanes_per_row = lpr;
    if (rows_per_tile) *rows_per_tile = rt;
    if (ms_out) *ms_out = best;
    clear_error();
    return NPBCG_OK;
}

int npbcg_batch_solve(npbcg_batch *b, double rtol, int maxit, npbcg_batch_stats *stats,
                      int *steps_out, double *relres_out) {
    if (!b || !(rtol > 0.0) || maxit < 1)
        return set_error(NPBCG_ERR_INVALID, "npbcg_batch_solve: invalid argument");
    if (!b->rhs_set) return set_error(NPBCG_ERR_INVALID, "npbcg_batch_solve: no right-hand sides");
    B_HIP(hipSetDevice(b->device));
    const int ns = b->nsys * b->nr;
    const bool mixed = b->opts.mode == NPBCG_SOLVE_MIXED;
    const int inner_max = b->opts.inner_steps > 0 ? b->opts.inner_steps : 60;
    std::vector<BatchScalars> sc(ns);
    std::vector<int> done(ns, 0);
    std::vector<double> bb(ns), rr(ns);
    // |b|^2 per solve: the residual of x = 0
    B_HIP(hipMemsetAsync(b->d_x, 0, sizeof(double) * (size_t)b->rows * b->nr, b->stream));
    B_HIP(hipMemsetAsync(b->d_sc, 0, sizeof(BatchScalars) * ns, b->stream));
    B_HIP(hipEventRecord(b->ev0, b->stream));
    B_HIP(batch_residual(b->bd, b->d_val, b->d_b, b->d_x, b->d_res, b->d_part, b->d_sysout,
                         b->stream));
    B_HIP(hipMemcpyAsync(bb.data(), b->d_sysout, sizeof(double) * ns, hipMemcpyDeviceToHost,
                         b->stream));
    B_HIP(hipStreamSynchronize(b->stream));
    rr = bb;
    int sweeps = 1;
    for (;;) {
        B_HIP(hipMemcpyAsync(sc.data(), b->d_sc, sizeof(BatchScalars) * ns, hipMemcpyDeviceToHost,
                             b->stream));
        B_HIP(hipStreamSynchronize(b->stream));
        bool any = false;
        int budget = 0;
        for (int i = 0; i < ns; i++) {
            const double rel = bb[i] > 0.0 ? std::sqrt(rr[i] / bb[i]) : 0.0;
            done[i] = (rel <= rtol || sc[i].steps >= maxit) ? 1 : 0;
            // inner target relative to this sweep's residual; FP32 holds about 1e-6 of it
            const double t = mixed ? std::max(1e-6, 0.5 * rtol / std::max(rel, 1e-300)) : 0.5 * rtol / std::max(rel, 1e-300);
            sc[i].rtol2 = t * t;
            sc[i].tol2 = 0.0;
            sc[i].sweep_steps = 0;
            if (!done[i]) {
                any = true;
                budget = std::max(budget, maxit - sc[i].steps);
            }
        }
        if (!any) break;
        b->bd.maxit = mixed ? std::min(inner_max, budget) : budget;
        B_HIP(hipMemcpyAsync(b->d_sc, sc.data(), sizeof(BatchScalars) * ns, hipMemcpyHostToDevice,
                             b->stream));
        B_HIP(hipMemcpyAsync(b->bd.done, done.data(), sizeof(int) * ns, hipMemcpyHostToDevice,
                             b->stream));
        B_HIP(hipMemcpyAsync(b->bd.done_next, done.data(), sizeof(int) * ns, hipMemcpyHostToDevice,
                             b->stream));
        int rc = mixed ? run_inner<float>(b, b->bd.maxit) : run_inner<double>(b, b->bd.maxit);
        if (rc != NPBCG_OK) return rc;
        B_HIP(batch_residual(b->bd, b->d_val, b->d_b, b->d_x, b->d_res, b->d_part, b->d_sysout,
                             b->stream));
        B_HIP(hipMemcpyAsync(rr.data(), b->d_sysout, sizeof(double) * ns, hipMemcpyDeviceToHost,
                             b->stream));
        B_HIP(hipStreamSynchronize(b->stream));
        sweeps++;
    }
    B_HIP(hipEventRecord(b->ev1, b->stream));
    B_HIP(hipEventSynchronize(b->ev1));
    float ms = 0;
    B_HIP(hipEventElapsedTime(&ms, b->ev0, b->ev1));
    npbcg_batch_stats st{};
    st.solves = ns;
    st.sweeps = sweeps;
    st.lanes_per_row = b->bd.lpr;
    st.rows_per_tile = b->opts.rows_per_tile;
    st.layout = b->opts.layout;
    st.time_ms = ms;
    for (int i = 0; i < ns; i++) {
        const double rel = bb[i] > 0.0 ? std::sqrt(rr[i] / bb[i]) : 0.0;
        if (rel <= rtol) st.converged++;
        st.max_relres = std::max(st.max_relres, rel);
        st.max_steps = std::max(st.max_steps, sc[i].steps);
        if (steps_out) steps_out[i] = sc[i].steps;
        if (relres_out) relres_out[i] = rel;
    }
    if (stats) *stats = st;
    if (st.converged < ns)
        return set_error(NPBCG_ERR_NOCONV, "npbcg_batch_solve: " + std::to_string(ns - st.converged) +
                                               " solve(s) missed the tolerance within maxit");
    clear_error();
    return NPBCG_OK;
}

int npbcg_batch_spmm_bench(npbcg_batch *b, int reps, double *avg_ms) {
    if (!b || reps < 1 || !avg_ms) return set_error(NPBCG_ERR_INVALID, "npbcg_batch_spmm_bench: invalid argument");
    B_HIP(hipSetDevice(b->device));
    B_HIP(hipMemsetAsync(b->d_flags, 0, sizeof(int) * 2 * b->nsys * b->nr, b->stream));
    int rc = b->opts.mode == NPBCG_SOLVE_MIXED ? time_spmm<float>(b, reps, avg_ms)
                                               : time_spmm<double>(b, reps, avg_ms);
    if (rc != NPBCG_OK) return rc;
    clear_error();
    return NPBCG_OK;
}

int npbcg_batch_get_solution(npbcg_batch *b, double *x) {
    if (!b || !x) return set_error(NPBCG_ERR_INVALID, "npbcg_batch_get_solution: NULL");
    B_HIP(hipSetDevice(b->device));
    B_HIP(hipMemcpyAsync(x, b->d_x, sizeof(double) * (size_t)b->rows * b->nr, hipMemcpyDeviceToHost,
                         b->stream));
    B_HIP(hipStreamSynchronize(b->stream));
    clear_error();
    return NPBCG_OK;
}

}  // extern "C"
