// Host side of batched solves: block-diagonal upload, tiling, the SELL copy,
// the FP64 refinement loop around the batched inner CG, and the SpMM
// autotuner with its persisted record file.
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
#include "trace_range.h"
#include "tiled_layout.h"

using namespace npbcg_dev;
using npbcg_detail::build_sell_blocks;
using npbcg_detail::clear_error;
using npbcg_detail::SellBlocks;
using npbcg_detail::set_error;

#define B_HIP(call)                                                              \
    do {                                                                         \
        hipError_t _e = (call);                                                  \
        if (_e != hipSuccess)                                                    \
            return set_error(NPBCG_ERR_DEVICE, std::string(#call) + ": " +       \
                                                   hipGetErrorString(_e));       \
    } while (0)

// Host metadata of one SELL copy of the batch.
struct SellHost {
    int W = 0;                          // column block width
    std::vector<int> nb;                // blocks per system
    std::vector<std::vector<int>> wg;   // per system: global first slice of each
                                        // (block, window), block-major, + end
    std::vector<int> nrange;            // windows per system
    std::vector<long long> part;        // per system: first row partial (elements)
    long long part_total = 0;
    std::vector<int> off;               // [slices + 1] first entry of each slice
};

struct npbcg_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    npbcg_batch_opts opts{};
    int nsys = 0, nr = 1, rows = 0;
    int cus = 256;                   // compute units of the device
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
    // SELL copy (layout 3), inner type values
    bool sell_ok = false;
    SellHost sell_h;
    void *d_sell_val = nullptr;
    uint16_t *d_sell_col = nullptr;
    int *d_sell_off = nullptr, *d_sell_row = nullptr;
    int *d_block_item = nullptr;     // workgroup lists of the SELL SpMM (sell_wg_lists)
    size_t block_item_cap = 0;
    int *d_item_sl = nullptr;        // slice ranges of the SELL items (make_sell_items)
    size_t item_sl_cap = 0;
    // FP64 SELL copy of the true residual: in FP64 mode the inner copy itself
    // (aliased), in MIXED mode a copy of its own with blocks sized for FP64
    // vectors.  Items are one sorting window each.
    bool res_ok = false, res_owns = false;
    SellHost res_h;
    double *d_res_val = nullptr;
    uint16_t *d_res_col = nullptr;
    int *d_res_off = nullptr, *d_res_row = nullptr;
    int *d_res_items = nullptr;      // [items][4]
    int *d_res_sys = nullptr;        // sys_item | sys_row | sys_nb
    long long *d_res_part = nullptr; // sys_part
    int *d_res_block_item = nullptr;
    int *d_res_item_sl = nullptr;
    int res_items = 0, res_item_blocks = 0;
    int *d_flags = nullptr;          // done
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

// Workgroup -> work unit (tile or item) order of the gathering kernels;
// system s owns units [sys_first[s], sys_first[s+1]).  With at least kXcds
// systems, whole systems are placed on XCDs (greedy by nonzeros, largest
// first) so that a system's vector is read from one XCD's L2; the lists are
// interleaved to match the round-robin dispatch.  Fewer systems keep the
// plain order.  Either way the length is a multiple of kXcds, padded with -1.
std::vector<int> xcd_order(const npbcg_batch *b, const std::vector<int> &sys_tile,
                           const std::vector<long long> &sys_nnz) {
    const int tiles = sys_tile[b->nsys];
    std::vector<int> out;
    if (b->nsys < kXcds) {  // padded to a multiple of kXcds (the SELL SpMM's
                            // reversed order maps slots within groups of 8)
        out.assign(((size_t)tiles + kXcds - 1) / kXcds * kXcds, -1);
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

// Uploads a workgroup -> unit order into *buf (grown as needed).
int upload_order(npbcg_batch *b, const std::vector<int> &order, int **buf, size_t *cap) {
    if (order.size() > *cap) {
        (void)hipFree(*buf);
        *buf = nullptr;
        *cap = 0;
        B_HIP(hipMalloc(buf, sizeof(int) * order.size()));
        *cap = order.size();
    }
    B_HIP(hipMemcpyAsync(*buf, order.data(), sizeof(int) * order.size(), hipMemcpyHostToDevice,
                         b->stream));
    B_HIP(hipStreamSynchronize(b->stream));
    return NPBCG_OK;
}

// Rows per tile of the vector kernels under the SELL layout.
constexpr int kSellTileRows = 256;

// SELL items, one per sorting window: (system, first local row, rows, offset
// into *item_sl), in system order, and the first item of each system.
// item_sl holds, for each item, the window's slice range [first, end) in
// each column block of its system.
void make_sell_items(const SellHost &h, const std::vector<int> &n, std::vector<int> *items,
                     std::vector<int> *sys_item, std::vector<int> *item_sl) {
    const int nsys = (int)n.size();
    items->clear();
    item_sl->clear();
    sys_item->assign(nsys + 1, 0);
    for (int s = 0; s < nsys; s++) {
        (*sys_item)[s] = (int)(items->size() / 4);
        const int nw = h.nrange[s];
        const std::vector<int> &wg = h.wg[s];
        for (int r = 0; r < nw; r++) {
            items->push_back(s);
            items->push_back(r * kSellWindow);
            items->push_back(std::min(kSellWindow, n[s] - r * kSellWindow));
            items->push_back((int)item_sl->size());
            for (int blk = 0; blk < h.nb[s]; blk++) {
                item_sl->push_back(wg[(size_t)blk * nw + r]);
                item_sl->push_back(wg[(size_t)blk * nw + r + 1]);
            }
        }
    }
    (*sys_item)[nsys] = (int)(items->size() / 4);
}

// Workgroup lists of the SELL SpMM.  The launch has one workgroup per CU
// (each holds a vector block and a window's row sums), kXcds * per_xcd in all; workgroup g
// walks the items first[g] .. first[g + 1] - 1 of the list (backwards when
// the launch reverses its order), so no workgroup is dispatched per item and
// no dispatch slot is left empty.  With at least kXcds systems, whole
// systems go to XCDs (greedy by cost, largest first; workgroup g runs on XCD
// g % kXcds), so a system's vector blocks are read through one XCD's L2.
// Inside an XCD (over all workgroups with fewer systems) the items go,
// heaviest first, to the least loaded workgroup (longest processing time
// first).  An item costs its stored entries (value and 16-bit column), the
// vector blocks it stages, its rows' q and p, and a fixed 16 KB per block for
// the fill and the start of the block's pipeline.
// *out = first[0 .. G] | items.
void sell_wg_lists(const npbcg_batch *b, const SellHost &h, const std::vector<int> &items,
                   const std::vector<int> &sys_item, const std::vector<int> &item_sl, int elem,
                   std::vector<int> *out) {
    const int nsys = b->nsys, per_xcd = std::max(1, b->cus / kXcds), G = kXcds * per_xcd;
    const int ni = sys_item[nsys];
    std::vector<double> cost(ni), sys_cost(nsys, 0.0);
    for (int s = 0; s < nsys; s++)
        for (int it = sys_item[s]; it < sys_item[s + 1]; it++) {
            const int nrows = items[4 * it + 2], so = items[4 * it + 3];
            double stored = 0.0;
            for (int blk = 0; blk < h.nb[s]; blk++)
                stored += (double)h.off[item_sl[so + 2 * blk + 1]] - (double)h.off[item_sl[so + 2 * blk]];
            cost[it] = stored * (elem + 2.0) + (double)(b->n[s] + 2 * nrows) * b->nr * elem +
                       16384.0 * h.nb[s];
            sys_cost[s] += cost[it];
        }
    std::vector<int> order(nsys);
    for (int s = 0; s < nsys; s++) order[s] = s;
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return sys_cost[x] > sys_cost[y]; });
    // items of each XCD (one pool over all workgroups with fewer systems)
    const int pools = nsys >= kXcds ? kXcds : 1;
    std::vector<std::vector<int>> pool(pools);
    std::vector<double> xload(kXcds, 0.0);
    for (int s : order) {
        int x = 0;
        if (pools > 1) {
            x = (int)(std::min_element(xload.begin(), xload.end()) - xload.begin());
            xload[x] += sys_cost[s];
        }
        for (int it = sys_item[s]; it < sys_item[s + 1]; it++) pool[x].push_back(it);
    }
    std::vector<std::vector<int>> lists(G);
    std::vector<double> load(G, 0.0);
    const int stride = pools > 1 ? kXcds : 1, count = pools > 1 ? per_xcd : G;
    for (int x = 0; x < pools; x++) {
        std::vector<int> &its = pool[x];
        std::stable_sort(its.begin(), its.end(), [&](int a, int c) { return cost[a] > cost[c]; });
        for (int it : its) {
            int best = x;
            for (int j = 1; j < count; j++) {
                const int g = x + j * stride;
                if (load[g] < load[best]) best = g;
            }
            load[best] += cost[it];
            lists[best].push_back(it);
        }
    }
    out->assign((size_t)G + 1, 0);
    out->reserve((size_t)G + 1 + ni);
    for (int g = 0; g < G; g++) {
        (*out)[g + 1] = (*out)[g] + (int)lists[g].size();
        out->insert(out->end(), lists[g].begin(), lists[g].end());
    }
}

int apply_shape(npbcg_batch *b, int lpr, int rt, int layout) {
    if (layout == 3 && !b->sell_ok)
        return set_error(NPBCG_ERR_CAPACITY, "batch: no SELL copy (more than 2^31 stored entries)");
    std::vector<int> ts, tr, st;
    make_tiles(b, layout == 3 ? kSellTileRows : rt, ts, tr, st);
    if (layout == 3) {  // items: one sorting window each, in system order
        rt = kSellWindow;
        std::vector<int> items, sys_item, item_sl;
        make_sell_items(b->sell_h, b->n, &items, &sys_item, &item_sl);
        const int ni = sys_item[b->nsys];
        if (ni > b->items_cap) return set_error(NPBCG_ERR_INTERNAL, "batch: item capacity");
        B_HIP(hipMemcpyAsync(b->d_items, items.data(), sizeof(int) * items.size(),
                             hipMemcpyHostToDevice, b->stream));
        B_HIP(hipMemcpyAsync(b->d_sysinfo, sys_item.data(), sizeof(int) * (b->nsys + 1),
                             hipMemcpyHostToDevice, b->stream));
        B_HIP(hipStreamSynchronize(b->stream));
        int rc = upload_order(b, item_sl, &b->d_item_sl, &b->item_sl_cap);
        if (rc != NPBCG_OK) return rc;
        std::vector<int> wl;
        sell_wg_lists(b, b->sell_h, items, sys_item, item_sl,
                      b->opts.mode == NPBCG_SOLVE_MIXED ? 4 : 8, &wl);
        rc = upload_order(b, wl, &b->d_block_item, &b->block_item_cap);
        if (rc != NPBCG_OK) return rc;
        b->bd.items = ni;
        b->bd.item_blocks = (int)wl.size() - ni - 1;
        b->bd.block_item = b->d_block_item;
        b->bd.item_sl = b->d_item_sl;
    }
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
    b->bd.cb = layout >= 2 ? 1 : 0;
    b->bd.sell = layout == 3 ? 1 : 0;
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
    int rc = upload_order(b, bt, &b->d_block_tile, &b->block_tile_cap);
    if (rc != NPBCG_OK) return rc;
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

// Column block width of a SELL copy with elem-byte vector entries: the
// block's nrhs-wide entries and the row sums of one 2048-row window fill the
// workgroup's LDS (160 KB on gfx950, less 2 KB for the static reduction
// slots): one SELL workgroup per CU.  Four FP64 right-hand sides get 3008
// columns, four FP32 ones 8064.
int sell_width(const npbcg_batch *b, int device, int elem) {
    int optin = 0;
    if (hipDeviceGetAttribute(&optin, hipDeviceAttributeSharedMemPerBlockOptin, device) !=
            hipSuccess || optin <= 0)
        optin = 64 * 1024;
    const int lds = optin - 2048;
    return std::min(65536, std::max(64, lds / (elem * b->nr) - kSellWindow));
}

// Column blocks of the inner type (layouts 2 and 3 share them).
int cb_width(const npbcg_batch *b, int device) {
    return sell_width(b, device, b->opts.mode == NPBCG_SOLVE_MIXED ? 4 : 8);
}

// SELL copy of every system (local CSR with ascending columns per row, as
// in the global CSR), concatenated: slice offsets shifted by the entries
// before the system, lane rows made global.  False if the stored entries
// exceed 32-bit offsets.
template <class V>
bool build_batch_sell(const npbcg_batch *b, int W, const std::vector<int> &rp,
                      const std::vector<int> &col, const std::vector<double> &val, SellHost *h,
                      std::vector<V> *sval, std::vector<uint16_t> *scol, std::vector<int> *soff,
                      std::vector<int> *srow) {
    long long entries = 0;
    int slices = 0;
    h->W = W;
    h->nb.assign(b->nsys, 0);
    h->wg.assign(b->nsys, {});
    h->nrange.assign(b->nsys, 0);
    h->part.assign(b->nsys, 0);
    h->part_total = 0;
    for (int s = 0; s < b->nsys; s++) {
        h->nb[s] = (b->n[s] + W - 1) / W;
        h->part[s] = h->part_total;
        h->part_total += (long long)b->n[s] * h->nb[s] * b->nr;
    }
    std::vector<int> lrs, lci;
    for (int s = 0; s < b->nsys; s++) {
        const int off = b->row_off[s], ns = b->n[s];
        const int e0 = rp[off], nnz_s = rp[off + ns] - e0;
        lrs.resize((size_t)ns + 1);
        lci.resize((size_t)nnz_s);
        for (int j = 0; j <= ns; j++) lrs[j] = rp[off + j] - e0;
        for (int k = 0; k < nnz_s; k++) lci[k] = col[(size_t)e0 + k] - off;
        SellBlocks<V> sb;
        if (!build_sell_blocks(ns, lrs.data(), lci.data(), val.data() + e0, W, kSellWindow, &sb) ||
            sb.nblk != h->nb[s] || entries + sb.stored > (long long)INT32_MAX)
            return false;
        h->nrange[s] = sb.nrange;
        std::vector<int> &wg = h->wg[s];
        wg.resize(sb.wg_slice.size());
        for (size_t i = 0; i < wg.size(); i++) wg[i] = slices + sb.wg_slice[i];
        const int nsl = (int)sb.slice_off.size() - 1;
        for (int i = 0; i < nsl; i++) soff->push_back((int)(entries + sb.slice_off[i]));
        for (int r : sb.slice_row) srow->push_back(r >= 0 ? r + off : -1);
        sval->insert(sval->end(), sb.val.begin(), sb.val.end());
        scol->insert(scol->end(), sb.col.begin(), sb.col.end());
        entries += sb.stored;
        slices += nsl;
    }
    soff->push_back((int)entries);
    h->off = *soff;
    return true;
}

// Column block width of the FP64 residual copy.
int res_width(const npbcg_batch *b, int device) { return sell_width(b, device, 8); }

template <typename T>
BatchVecs<T> inner_vecs(const npbcg_batch *b) {
    BatchVecs<T> v;
    const size_t m = (size_t)b->rows * b->nr;
    T *base = static_cast<T *>(b->d_inner);
    v.val = std::is_same<T, float>::value ? (const T *)(const void *)b->d_val32
                                          : (const T *)(const void *)b->d_val;
    v.sval = static_cast<const T *>(b->d_sell_val);
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

// The batch as seen by the FP64 SELL residual: its own column blocks,
// items and partial offsets; rows, tiles and flags of the batch.
BatchDev resid_dev(const npbcg_batch *b) {
    BatchDev d = b->bd;
    const int ns = b->nsys;
    d.cb = 1;
    d.sell = 1;
    d.cbw = b->res_h.W;
    d.items = b->res_items;
    d.item = b->d_res_items;
    d.sys_item = b->d_res_sys;
    d.sys_row = b->d_res_sys + ns + 1;
    d.sys_nb = b->d_res_sys + 2 * ns + 2;
    d.sys_part = b->d_res_part;
    d.item_blocks = b->res_item_blocks;
    d.block_item = b->d_res_block_item;
    d.sell_off = b->d_res_off;
    d.sell_row = b->d_res_row;
    d.sell_col = b->d_res_col;
    d.item_sl = b->d_res_item_sl;
    return d;
}

// res = b - A x (FP64), |res|^2 per solve into d_sysout.
int true_residual(npbcg_batch *b) {
    if (b->res_ok)
        B_HIP(batch_residual_sell(resid_dev(b), b->d_res_val, b->d_b, b->d_x, b->d_res, b->d_part,
                                  b->d_sysout, b->stream));
    else
        B_HIP(batch_residual(b->bd, b->d_val, b->d_b, b->d_x, b->d_res, b->d_part, b->d_sysout,
                             b->stream));
    return NPBCG_OK;
}

std::string shape_key(const npbcg_batch *b) {
    char buf[256];
    // "batch4": records of the tuner with window items for layout 3 (older
    // records are not reused)
    std::snprintf(buf, sizeof buf, "batch4:sys=%d:rows=%d:nnz=%lld:nrhs=%d:mode=%d", b->nsys,
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
    B_HIP(batch_spmm<T>(b->bd, v, 1, b->stream));  // warm
    B_HIP(hipEventRecord(b->ev0, b->stream));
    // alternating item order, as in the solve's CG steps
    for (int i = 0; i < reps; i++) B_HIP(batch_spmm<T>(b->bd, v, i & 1, b->stream));
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
    if (b->res_owns)
        for (void *p : {(void *)b->d_res_val, (void *)b->d_res_col, (void *)b->d_res_off,
                        (void *)b->d_res_row})
            (void)hipFree(p);
    for (void *p : {(void *)b->d_res_items, (void *)b->d_res_sys, (void *)b->d_res_part,
                    (void *)b->d_res_block_item, (void *)b->d_res_item_sl})
        (void)hipFree(p);
    for (void *p : {b->d_sell_val, (void *)b->d_sell_col, (void *)b->d_sell_off,
                    (void *)b->d_sell_row, (void *)b->d_block_item, (void *)b->d_item_sl,
                    (void *)b->d_items, (void *)b->d_sysinfo, (void *)b->d_sys64, (void *)b->d_seg,
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
    npbcg_detail::TraceRange trace("npbcg_batch_create");
    if (!out || nsys < 1 || !n || !rowstr || !colidx || !a)
        return set_error(NPBCG_ERR_INVALID, "npbcg_batch_create: invalid argument");
    *out = nullptr;
    npbcg_batch_opts o;
    npbcg_batch_default_opts(&o);
    if (opts) o = *opts;
    if ((o.nrhs != 1 && o.nrhs != 2 && o.nrhs != 4) ||
        (o.mode != NPBCG_SOLVE_FP64 && o.mode != NPBCG_SOLVE_MIXED) || o.layout < 0 ||
        o.layout > 3)
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
    TRY(hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, b->device));
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
    // SELL copy in the inner type (layout 3)
    std::vector<float> sval32;
    std::vector<double> sval64;
    std::vector<uint16_t> scol;
    std::vector<int> soff, srow;
    b->sell_ok = o.mode == NPBCG_SOLVE_MIXED
                     ? build_batch_sell<float>(b, b->cbw, rp, col, val, &b->sell_h, &sval32, &scol,
                                               &soff, &srow)
                     : build_batch_sell<double>(b, b->cbw, rp, col, val, &b->sell_h, &sval64,
                                                &scol, &soff, &srow);
    // FP64 copy of the true residual (MIXED: a copy of its own)
    std::vector<double> rval;
    std::vector<uint16_t> rcol;
    std::vector<int> roff, rrow;
    if (o.mode == NPBCG_SOLVE_MIXED) {
        b->res_ok = build_batch_sell<double>(b, res_width(b, b->device), rp, col, val, &b->res_h,
                                             &rval, &rcol, &roff, &rrow);
        b->res_owns = b->res_ok;
    } else {
        b->res_ok = b->sell_ok;
        b->res_h = b->sell_h;
    }
    if (o.layout == 3 && !b->sell_ok)
        return fail(set_error(NPBCG_ERR_CAPACITY,
                              "npbcg_batch_create: SELL copy exceeds 32-bit offsets"));
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
    TRY(hipMalloc(&b->d_flags, sizeof(int) * nsys * b->nr));
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
    if (b->sell_ok) {
        const size_t stored = scol.size();
        const size_t vbytes = o.mode == NPBCG_SOLVE_MIXED ? sizeof(float) * stored
                                                          : sizeof(double) * stored;
        TRY(hipMalloc(&b->d_sell_val, std::max<size_t>(vbytes, 8)));
        TRY(hipMalloc(&b->d_sell_col, std::max<size_t>(sizeof(uint16_t) * stored, 8)));
        TRY(hipMalloc(&b->d_sell_off, sizeof(int) * soff.size()));
        TRY(hipMalloc(&b->d_sell_row, sizeof(int) * std::max<size_t>(srow.size(), 1)));
        const void *vsrc = o.mode == NPBCG_SOLVE_MIXED ? (const void *)sval32.data()
                                                       : (const void *)sval64.data();
        TRY(hipMemcpyAsync(b->d_sell_val, vsrc, vbytes, hipMemcpyHostToDevice, b->stream));
        TRY(hipMemcpyAsync(b->d_sell_col, scol.data(), sizeof(uint16_t) * stored,
                           hipMemcpyHostToDevice, b->stream));
        TRY(hipMemcpyAsync(b->d_sell_off, soff.data(), sizeof(int) * soff.size(),
                           hipMemcpyHostToDevice, b->stream));
        TRY(hipMemcpyAsync(b->d_sell_row, srow.data(), sizeof(int) * srow.size(),
                           hipMemcpyHostToDevice, b->stream));
    }
    if (b->res_ok) {
        if (b->res_owns) {
            const size_t stored = rcol.size();
            TRY(hipMalloc(&b->d_res_val, std::max<size_t>(sizeof(double) * stored, 8)));
            TRY(hipMalloc(&b->d_res_col, std::max<size_t>(sizeof(uint16_t) * stored, 8)));
            TRY(hipMalloc(&b->d_res_off, sizeof(int) * roff.size()));
            TRY(hipMalloc(&b->d_res_row, sizeof(int) * std::max<size_t>(rrow.size(), 1)));
            TRY(hipMemcpyAsync(b->d_res_val, rval.data(), sizeof(double) * stored,
                               hipMemcpyHostToDevice, b->stream));
            TRY(hipMemcpyAsync(b->d_res_col, rcol.data(), sizeof(uint16_t) * stored,
                               hipMemcpyHostToDevice, b->stream));
            TRY(hipMemcpyAsync(b->d_res_off, roff.data(), sizeof(int) * roff.size(),
                               hipMemcpyHostToDevice, b->stream));
            TRY(hipMemcpyAsync(b->d_res_row, rrow.data(), sizeof(int) * rrow.size(),
                               hipMemcpyHostToDevice, b->stream));
        } else {  // FP64 mode: the inner SELL copy
            b->d_res_val = static_cast<double *>(b->d_sell_val);
            b->d_res_col = b->d_sell_col;
            b->d_res_off = b->d_sell_off;
            b->d_res_row = b->d_sell_row;
        }
        std::vector<int> items, sys_item, item_sl;
        make_sell_items(b->res_h, b->n, &items, &sys_item, &item_sl);
        b->res_items = sys_item[nsys];
        std::vector<int> sys(3 * (size_t)nsys + 2);
        for (int s = 0; s <= nsys; s++) {
            sys[s] = sys_item[s];
            sys[nsys + 1 + s] = b->row_off[s];
        }
        for (int s = 0; s < nsys; s++) sys[2 * nsys + 2 + s] = b->res_h.nb[s];
        std::vector<int> order;
        sell_wg_lists(b, b->res_h, items, sys_item, item_sl, 8, &order);
        b->res_item_blocks = (int)order.size() - b->res_items - 1;
        TRY(hipMalloc(&b->d_res_items, sizeof(int) * std::max<size_t>(items.size(), 4)));
        TRY(hipMalloc(&b->d_res_sys, sizeof(int) * sys.size()));
        TRY(hipMalloc(&b->d_res_part, sizeof(long long) * nsys));
        TRY(hipMalloc(&b->d_res_block_item, sizeof(int) * std::max<size_t>(order.size(), 1)));
        TRY(hipMalloc(&b->d_res_item_sl, sizeof(int) * std::max<size_t>(item_sl.size(), 1)));
        TRY(hipMemcpyAsync(b->d_res_item_sl, item_sl.data(), sizeof(int) * item_sl.size(),
                           hipMemcpyHostToDevice, b->stream));
        TRY(hipMemcpyAsync(b->d_res_items, items.data(), sizeof(int) * items.size(),
                           hipMemcpyHostToDevice, b->stream));
        TRY(hipMemcpyAsync(b->d_res_sys, sys.data(), sizeof(int) * sys.size(),
                           hipMemcpyHostToDevice, b->stream));
        TRY(hipMemcpyAsync(b->d_res_part, b->res_h.part.data(), sizeof(long long) * nsys,
                           hipMemcpyHostToDevice, b->stream));
        TRY(hipMemcpyAsync(b->d_res_block_item, order.data(), sizeof(int) * order.size(),
                           hipMemcpyHostToDevice, b->stream));
        const hipError_t e3 =
            batch_prepare_resid((int)sell_lds_bytes(b->res_h.W, b->nr, sizeof(double)));
        if (e3 != hipSuccess)
            return fail(set_error(NPBCG_ERR_DEVICE, std::string("batch_prepare_resid: ") +
                                                        hipGetErrorString(e3)));
    }
    TRY(hipMemcpyAsync(b->d_rowptr, rp.data(), sizeof(int) * (rows + 1), hipMemcpyHostToDevice,
                       b->stream));
    TRY(hipMemcpyAsync(b->d_col, col.data(), sizeof(int) * nnz, hipMemcpyHostToDevice, b->stream));
    TRY(hipMemcpyAsync(b->d_val, val.data(), sizeof(double) * nnz, hipMemcpyHostToDevice, b->stream));
    if (b->d_val32)
        TRY(hipMemcpyAsync(b->d_val32, val32.data(), sizeof(float) * nnz, hipMemcpyHostToDevice,
                           b->stream));
    TRY(hipMemsetAsync(b->d_x, 0, sizeof(double) * m, b->stream));
    TRY(hipMemsetAsync(b->d_inner, 0, inner_elem * 4 * m, b->stream));
    TRY(hipMemsetAsync(b->d_flags, 0, sizeof(int) * nsys * b->nr, b->stream));
    TRY(hipMemsetAsync(b->d_sc, 0, sizeof(BatchScalars) * nsys * b->nr, b->stream));
    TRY(hipStreamSynchronize(b->stream));  // host staging goes out of scope
#undef TRY
    b->bd.nsys = nsys;
    b->bd.nr = b->nr;
    b->bd.rows = b->rows;
    b->bd.rowptr = b->d_rowptr;
    b->bd.col = b->d_col;
    b->bd.done = b->d_flags;
    b->bd.sc = b->d_sc;
    b->bd.cbw = b->cbw;
    b->bd.item = b->d_items;
    b->bd.sys_item = b->d_sysinfo;
    b->bd.sys_row = b->d_sysinfo + nsys + 1;
    b->bd.sys_nb = b->d_sysinfo + 2 * nsys + 2;
    b->bd.sys_seg = b->d_sys64;
    b->bd.sys_part = b->d_sys64 + nsys;
    b->bd.seg = b->d_seg;
    b->bd.sell_off = b->d_sell_off;
    b->bd.sell_row = b->d_sell_row;
    b->bd.sell_col = b->d_sell_col;
    {
        const int lds = (int)sell_lds_bytes(b->cbw, b->nr, o.mode == NPBCG_SOLVE_MIXED ? 4 : 8);
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
    const int layout = o.layout != 0 ? o.layout : b->sell_ok ? 3 : 2;
    if (layout == 3 && o.rows_per_tile == 0) rt = kSellWindow;
    int rc = apply_shape(b, lpr, rt, layout);
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
    npbcg_detail::TraceRange trace("npbcg_batch_tune");
    if (!b) return set_error(NPBCG_ERR_INVALID, "npbcg_batch_tune: NULL");
    B_HIP(hipSetDevice(b->device));
    const std::string key = shape_key(b);
    int lpr = 0, rt = 0, layout = 2;
    double best = 0.0;
    if (!cache_lookup(cache_path, key, &lpr, &rt, &layout, &best)) {
        // All solves marked running so every tile does its work; vectors hold
        // whatever they hold (timing only); state is reset afterwards.
        B_HIP(hipMemsetAsync(b->d_flags, 0, sizeof(int) * b->nsys * b->nr, b->stream));
        best = 1e30;
        int dl = 0, dr = 0;
        default_shape(b, &dl, &dr);
        // Layout 2 (column blocks in CSR order) is not timed: layout 3 beat
        // it on every batch measured.
        for (int cy : {1, 3})
            for (int cl : {4, 8, 16, 32, 64})
                for (int cr : {64, 128, 256, 512, 1024, 2048, 4096}) {
                    // layout 3: one lane per row, one sorting window per
                    // item; lanes_per_row stays at its default (it shapes
                    // only the row-gather residual kernel)
                    if (cy == 3 && (!b->sell_ok || cl != dl || cr != kSellWindow)) continue;
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
    npbcg_detail::TraceRange trace("npbcg_batch_solve");
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
    B_HIP(batch_residual_zero(b->bd, b->d_b, b->d_res, b->d_part, b->d_sysout, b->stream));
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
        int rc = mixed ? run_inner<float>(b, b->bd.maxit) : run_inner<double>(b, b->bd.maxit);
        if (rc != NPBCG_OK) return rc;
        rc = true_residual(b);
        if (rc != NPBCG_OK) return rc;
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
    B_HIP(hipMemsetAsync(b->d_flags, 0, sizeof(int) * b->nsys * b->nr, b->stream));
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
