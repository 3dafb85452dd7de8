// Host check of the column-blocked SELL layout (npb-gpu_amd/csrc/tiled_layout.h)
// on generated CSR matrices, for the invariants the SpMV kernels rely on:
//   * every row has exactly one slot in each column block, slice_row names it
//     back, and rows of a workgroup stay inside its row range;
//   * each slice stores its lanes' block segments in CSR order, padded with
//     (0, 0), and slice_off advances by 64 * width;
//   * a workgroup's slices are a permutation of the sorted cut (the
//     wavefront placement only moves them): the same multiset of widths and
//     the same row sets, and the same count per wavefront under the kernels'
//     round order;
//   * y = A x through the layout (each row's block sum in entry order, blocks
//     added in order, as the kernels and the combining kernels form it) equals
//     the same sums taken from the CSR directly, bit for bit.
// Prints one line per case: "n window waves stored slices ok"; exits nonzero
// on the first violation.
//
// Usage: sell_layout_check <n> <row_len_max> <W> <R> <window> <waves> <seed>
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <random>
#include <set>
#include <vector>

#include "tiled_layout.h"

using npbcg_detail::SellBlocks;
using npbcg_detail::kSliceRows;

static int fail(const char *what, long long a, long long b) {
    std::printf("FAIL %s (%lld, %lld)\n", what, a, b);
    return 1;
}

int main(int argc, char **argv) {
    if (argc != 8) {
        std::fprintf(stderr, "usage: %s n row_len_max W R window waves seed\n", argv[0]);
        return 2;
    }
    const int n = std::atoi(argv[1]), lmax = std::atoi(argv[2]), W = std::atoi(argv[3]);
    const int R = std::atoi(argv[4]), window = std::atoi(argv[5]), waves = std::atoi(argv[6]);
    std::mt19937_64 rng(std::strtoull(argv[7], nullptr, 10));

    // CSR with ascending columns, ragged rows (0 .. lmax entries, some empty)
    std::vector<int> rowstr(n + 1, 0), colidx;
    std::vector<double> a;
    for (int j = 0; j < n; j++) {
        const int len = (int)(rng() % (unsigned)(lmax + 1));
        std::set<int> cols;
        while ((int)cols.size() < std::min(len, n)) cols.insert((int)(rng() % (unsigned)n));
        for (int c : cols) {
            colidx.push_back(c);
            a.push_back(std::ldexp((double)(rng() % 2000001) - 1e6, -20));
        }
        rowstr[j + 1] = (int)colidx.size();
    }
    std::vector<double> x(n);
    for (auto &v : x) v = std::ldexp((double)(rng() % 2000001) - 1e6, -19);

    SellBlocks<double> m;
    if (!npbcg_detail::build_sell_blocks(n, rowstr.data(), colidx.data(), a.data(), W, R, &m,
                                         window, waves))
        return fail("build", 0, 0);
    // the same cut without placement, for the permutation check
    SellBlocks<double> m0;
    if (!npbcg_detail::build_sell_blocks(n, rowstr.data(), colidx.data(), a.data(), W, R, &m0,
                                         window, 0))
        return fail("build (sorted order)", 0, 0);

    const int nwg = m.nblk * m.nrange;
    if ((int)m.wg_slice.size() != nwg + 1) return fail("wg_slice size", m.wg_slice.size(), nwg + 1);
    const int nsl_total = m.wg_slice[nwg];
    if (m.slice_off[nsl_total] != m.stored) return fail("slice_off end", m.slice_off[nsl_total], m.stored);
    if (m.stored < m.nnz) return fail("stored < nnz", m.stored, m.nnz);
    if (m0.stored != m.stored || m0.wg_slice != m.wg_slice)
        return fail("placement changed the cut", m0.stored, m.stored);

    std::vector<int> seen((size_t)m.nblk * n, 0);
    for (int wg = 0; wg < nwg; wg++) {
        const int b = wg / m.nrange, rr = wg % m.nrange;
        const int r0 = rr * R, r1 = std::min(n, r0 + R), c0 = b * W;
        const int s0 = m.wg_slice[wg], nsl = m.wg_slice[wg + 1] - s0;
        std::multiset<int> w_now, w_sorted;
        std::multiset<std::vector<int>> rows_now, rows_sorted;
        for (int k = 0; k < nsl; k++) {
            const int s = s0 + k;
            const long long off = m.slice_off[s];
            const long long span = m.slice_off[s + 1] - off;
            if (span < 0 || span % kSliceRows) return fail("slice span", s, span);
            const int width = (int)(span / kSliceRows);
            w_now.insert(width);
            w_sorted.insert((m0.slice_off[s + 1] - m0.slice_off[s]) / kSliceRows);
            std::vector<int> rs, rs0;
            for (int l = 0; l < kSliceRows; l++) {
                rs.push_back(m.slice_row[(size_t)s * kSliceRows + l]);
                rs0.push_back(m0.slice_row[(size_t)s * kSliceRows + l]);
                const int row = m.slice_row[(size_t)s * kSliceRows + l];
                if (row < 0) {
                    for (int jj = 0; jj < width; jj++) {
                        const size_t at = (size_t)off + (size_t)jj * kSliceRows + l;
                        if (m.val[at] != 0.0 || m.col[at] != 0) return fail("padding lane", s, l);
                    }
                    continue;
                }
                if (row < r0 || row >= r1) return fail("row outside range", row, wg);
                if (m.row_slot[(size_t)b * n + row] != s * kSliceRows + l)
                    return fail("row_slot", row, m.row_slot[(size_t)b * n + row]);
                seen[(size_t)b * n + row]++;
                // the row's segment of block b, in CSR order, then padding
                int jj = 0;
                for (int e = rowstr[row]; e < rowstr[row + 1]; e++) {
                    if (colidx[e] < c0 || colidx[e] >= c0 + W) continue;
                    if (jj >= width) return fail("segment wider than slice", row, width);
                    const size_t at = (size_t)off + (size_t)jj * kSliceRows + l;
                    if (m.val[at] != a[e] || (int)m.col[at] + c0 != colidx[e])
                        return fail("entry", row, jj);
                    jj++;
                }
                for (; jj < width; jj++) {
                    const size_t at = (size_t)off + (size_t)jj * kSliceRows + l;
                    if (m.val[at] != 0.0 || m.col[at] != 0) return fail("row padding", row, jj);
                }
            }
            std::sort(rs.begin(), rs.end());
            std::sort(rs0.begin(), rs0.end());
            rows_now.insert(rs);
            rows_sorted.insert(rs0);
        }
        if (w_now != w_sorted || rows_now != rows_sorted)
            return fail("placement is not a permutation of the workgroup's slices", wg, nsl);
        (void)r1;
    }
    for (size_t i = 0; i < seen.size(); i++)
        if (seen[i] != 1) return fail("slots per (block, row)", (long long)i, seen[i]);

    // y through the layout against the CSR, same summation order
    std::vector<double> part((size_t)nsl_total * kSliceRows, 0.0);
    for (int wg = 0; wg < nwg; wg++) {
        const int b = wg / m.nrange, c0 = b * W;
        for (int s = m.wg_slice[wg]; s < m.wg_slice[wg + 1]; s++) {
            const int width = (m.slice_off[s + 1] - m.slice_off[s]) / kSliceRows;
            for (int l = 0; l < kSliceRows; l++) {
                double acc = 0.0;
                for (int jj = 0; jj < width; jj++) {
                    const size_t at = (size_t)m.slice_off[s] + (size_t)jj * kSliceRows + l;
                    acc = std::fma(m.val[at], x[c0 + m.col[at]], acc);
                }
                part[(size_t)s * kSliceRows + l] = acc;
            }
        }
    }
    for (int j = 0; j < n; j++) {
        double y_layout = part[m.row_slot[j]], y_csr = 0.0;
        for (int b = 1; b < m.nblk; b++) y_layout += part[m.row_slot[(size_t)b * n + j]];
        for (int b = 0; b < m.nblk; b++) {
            double acc = 0.0;
            for (int e = rowstr[j]; e < rowstr[j + 1]; e++)
                if (colidx[e] >= b * W && colidx[e] < (b + 1) * W) acc = std::fma(a[e], x[colidx[e]], acc);
            y_csr = b == 0 ? acc : y_csr + acc;
        }
        if (std::memcmp(&y_layout, &y_csr, sizeof(double)) != 0) return fail("y bits", j, 0);
    }

    // wavefront loads under the kernels' round order (snake), and the
    // per-wavefront slice counts, which placement must keep
    long long worst_now = 0, worst_sorted = 0;
    for (int wg = 0; wg < nwg && waves > 1; wg++) {
        const int s0 = m.wg_slice[wg], nsl = m.wg_slice[wg + 1] - s0;
        std::vector<long long> load(waves, 0), load0(waves, 0);
        std::vector<int> cnt(waves, 0), cnt0(waves, 0);
        for (int k = 0; k < nsl; k++) {
            const int rd = k / waves, i = k % waves, w = (rd & 1) ? waves - 1 - i : i;
            load[w] += (m.slice_off[s0 + k + 1] - m.slice_off[s0 + k]) / kSliceRows;
            load0[w] += (m0.slice_off[s0 + k + 1] - m0.slice_off[s0 + k]) / kSliceRows;
            cnt[w]++;
            cnt0[w]++;
        }
        if (cnt != cnt0) return fail("slices per wavefront", wg, 0);
        worst_now += *std::max_element(load.begin(), load.end());
        worst_sorted += *std::max_element(load0.begin(), load0.end());
    }
    std::printf("n %d window %d waves %d stored %lld nnz %lld slices %d worst_wave_width %lld sorted %lld ok\n",
                n, window, waves, m.stored, m.nnz, nsl_total, worst_now, worst_sorted);
    return 0;
}
