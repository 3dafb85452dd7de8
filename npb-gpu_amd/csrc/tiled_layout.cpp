// Host construction of the column-blocked SELL copy (tiled_layout.h).
#include "tiled_layout.h"

#include <algorithm>
#include <thread>

namespace npbcg_detail {

namespace {

template <class F>
void parallel_items(int count, F &&body) {
    unsigned hc = std::thread::hardware_concurrency();
    int nth = (int)std::min<unsigned>(hc ? hc : 1, 16u);
    nth = std::min(nth, count);
    if (nth <= 1) {
        for (int i = 0; i < count; i++) body(i);
        return;
    }
    std::vector<std::thread> pool;
    for (int w = 0; w < nth; w++)
        pool.emplace_back([&, w] {
            for (int i = w; i < count; i += nth) body(i);
        });
    for (auto &th : pool) th.join();
}

// Rows of range rr ordered by their segment length in block b (longest
// first, ties by row) within windows of `window` rows, with the segment
// bounds.  The window keeps the rows of one slice, whose block sums are
// stored as 8-byte partials, within a few thousand rows of each other, so
// the partial reads of the combining kernels (row by row, through row_slot)
// stay in L2-sized neighbourhoods; a wider window pads fewer entries (rows of
// like length share a slice).  It is a multiple of the slice height, so a
// slice never spans two windows.
struct RowSeg {
    int row, lo, hi;
};

void block_segments(int n, const int *rowstr, const int *colidx, int W, int R, int b,
                    int rr, int window, std::vector<RowSeg> *out) {
    out->clear();
    const int r0 = rr * R, r1 = std::min(n, r0 + R);
    const int c0 = b * W, c1 = c0 + W;
    for (int j = r0; j < r1; j++) {
        const int *first = colidx + rowstr[j], *last = colidx + rowstr[j + 1];
        const int lo = (int)(std::lower_bound(first, last, c0) - colidx);
        const int hi = (int)(std::lower_bound(first, last, c1) - colidx);
        out->push_back({j, lo, hi});
    }
    for (size_t w0 = 0; w0 < out->size(); w0 += window) {
        const size_t w1 = std::min(out->size(), w0 + (size_t)window);
        std::stable_sort(out->begin() + w0, out->begin() + w1,
                         [](const RowSeg &x, const RowSeg &y) {
                             return (x.hi - x.lo) > (y.hi - y.lo);
                         });
    }
}

// Position of each slice of a workgroup (slices in sorted order, widths
// given).  The kernels give position k of round rd = k / waves to wavefront
// (rd even) k % waves or (rd odd) waves - 1 - k % waves.  With waves > 0 the
// slices are placed so that the wavefronts' total widths are close: widest
// first, each to the wavefront with the least width so far among those with
// a position left (every wavefront keeps the number of positions the rounds
// give it).  waves = 0 keeps the sorted order.
void place_slices(const std::vector<int> &width, int waves, std::vector<int> *pos) {
    const int nsl = (int)width.size();
    pos->resize(nsl);
    if (waves <= 1 || nsl <= waves) {
        for (int i = 0; i < nsl; i++) (*pos)[i] = i;
        return;
    }
    std::vector<std::vector<int>> slots(waves);  // positions of each wavefront, in order
    for (int k = 0; k < nsl; k++) {
        const int rd = k / waves, i = k % waves;
        slots[(rd & 1) ? waves - 1 - i : i].push_back(k);
    }
    std::vector<int> order(nsl);
    for (int i = 0; i < nsl; i++) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int x, int y) { return width[x] > width[y]; });
    std::vector<long long> load(waves, 0);
    std::vector<int> used(waves, 0);
    for (int i : order) {
        int best = -1;
        for (int w = 0; w < waves; w++)
            if (used[w] < (int)slots[w].size() && (best < 0 || load[w] < load[best])) best = w;
        (*pos)[i] = slots[best][used[best]++];
        load[best] += width[i] + 2;  // + the per-slice start
    }
}

template <class V>
bool build_impl(int n, const int *rowstr, const int *colidx, const double *a, int W, int R,
                int window, int waves, SellBlocks<V> *out) {
    if (n < 1 || W < 1 || W > 65536 || R < 1 || window < kSliceRows || window % kSliceRows)
        return false;
    SellBlocks<V> &m = *out;
    m.n = n;
    m.W = W;
    m.R = R;
    m.nblk = (n + W - 1) / W;
    m.nrange = (n + R - 1) / R;
    m.nnz = rowstr[n];
    const int nwg = m.nblk * m.nrange;

    // pass 1: slices and stored entries per workgroup
    std::vector<long long> wg_entries(nwg, 0);
    std::vector<int> wg_slices(nwg, 0);
    parallel_items(nwg, [&](int wg) {
        std::vector<RowSeg> segs;
        const int b = wg / m.nrange, rr = wg % m.nrange;
        block_segments(n, rowstr, colidx, W, R, b, rr, window, &segs);
        long long e = 0;
        int s = 0;
        for (size_t i = 0; i < segs.size(); i += kSliceRows, s++)
            e += (long long)kSliceRows * (segs[i].hi - segs[i].lo);
        wg_entries[wg] = e;
        wg_slices[wg] = s;
    });
    m.wg_slice.assign(nwg + 1, 0);
    std::vector<long long> wg_base(nwg + 1, 0);
    for (int wg = 0; wg < nwg; wg++) {
        m.wg_slice[wg + 1] = m.wg_slice[wg] + wg_slices[wg];
        wg_base[wg + 1] = wg_base[wg] + wg_entries[wg];
    }
    m.stored = wg_base[nwg];
    if (m.stored > (long long)INT32_MAX) return false;
    const int nslices = m.wg_slice[nwg];
    m.slice_off.assign(nslices + 1, 0);
    if ((long long)nslices * kSliceRows > (long long)INT32_MAX) return false;
    m.slice_row.assign((size_t)nslices * kSliceRows, -1);
    m.row_slot.assign((size_t)m.nblk * n, -1);
    m.val.assign((size_t)m.stored, V(0));
    m.col.assign((size_t)m.stored, 0);

    // pass 2: fill
    parallel_items(nwg, [&](int wg) {
        std::vector<RowSeg> segs;
        const int b = wg / m.nrange, rr = wg % m.nrange;
        block_segments(n, rowstr, colidx, W, R, b, rr, window, &segs);
        long long off = wg_base[wg];
        const int nsl = m.wg_slice[wg + 1] - m.wg_slice[wg];
        std::vector<int> width(nsl), pos, at_pos(nsl);
        for (int i = 0; i < nsl; i++)
            width[i] = segs[(size_t)i * kSliceRows].hi - segs[(size_t)i * kSliceRows].lo;
        place_slices(width, waves, &pos);
        for (int i = 0; i < nsl; i++) at_pos[pos[i]] = i;
        for (int k = 0; k < nsl; k++) {
            const int s = m.wg_slice[wg] + k;
            const size_t i0 = (size_t)at_pos[k] * kSliceRows;
            const int width_k = width[at_pos[k]];
            m.slice_off[s] = (int)off;
            for (int l = 0; l < kSliceRows && i0 + l < segs.size(); l++) {
                const RowSeg &rs = segs[i0 + l];
                m.slice_row[(size_t)s * kSliceRows + l] = rs.row;
                m.row_slot[(size_t)b * n + rs.row] = s * kSliceRows + l;
                for (int j = 0; j < rs.hi - rs.lo; j++) {
                    const size_t at = (size_t)off + (size_t)j * kSliceRows + l;
                    m.val[at] = (V)a[rs.lo + j];
                    m.col[at] = (uint16_t)(colidx[rs.lo + j] - b * W);
                }
            }
            off += (long long)kSliceRows * width_k;
        }
        if (wg == nwg - 1) m.slice_off[nslices] = (int)off;
    });
    if (nslices == 0) m.slice_off[0] = 0;
    return true;
}

}  // namespace

bool build_sell_blocks(int n, const int *rowstr, const int *colidx, const double *a, int W,
                       int R, SellBlocks<double> *out, int window, int waves) {
    return build_impl(n, rowstr, colidx, a, W, R, window, waves, out);
}

bool build_sell_blocks(int n, const int *rowstr, const int *colidx, const double *a, int W,
                       int R, SellBlocks<float> *out, int window, int waves) {
    return build_impl(n, rowstr, colidx, a, W, R, window, waves, out);
}

}  // namespace npbcg_detail
