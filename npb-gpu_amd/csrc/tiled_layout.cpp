// Host construction of the column-blocked copy used by the LDS-staged SpMV.
#include "tiled_layout.h"

#include <algorithm>
#include <thread>

namespace npbcg_detail {

namespace {

template <class F>
void for_tiles(int ntile, F &&body) {
    unsigned hc = std::thread::hardware_concurrency();
    int nth = (int)std::min<unsigned>(hc ? hc : 1, 16u);
    nth = std::min(nth, ntile);
    if (nth <= 1) {
        for (int t = 0; t < ntile; t++) body(t);
        return;
    }
    std::vector<std::thread> pool;
    for (int w = 0; w < nth; w++)
        pool.emplace_back([&, w] {
            for (int t = w; t < ntile; t += nth) body(t);
        });
    for (auto &th : pool) th.join();
}

template <class V>
bool build_impl(int n, const int *rowstr, const int *colidx, const double *a, int T, int W,
                TiledMatrix<V> *out) {
    if (n < 1 || T < 1 || W < 1 || W > 65536) return false;
    TiledMatrix<V> &m = *out;
    m.n = n;
    m.T = T;
    m.W = W;
    m.nblk = (n + W - 1) / W;
    m.ntile = (n + T - 1) / T;
    m.nnz = rowstr[n];
    const int nblk = m.nblk, ntile = m.ntile;
    const size_t segs = (size_t)ntile * nblk * (T + 1);
    if (segs > (size_t)INT32_MAX) return false;

    std::vector<long long> cnt((size_t)ntile * nblk, 0);
    for_tiles(ntile, [&](int t) {
        for (int i = 0; i < T; i++) {
            const int row = t * T + i;
            if (row >= n) break;
            for (int k = rowstr[row]; k < rowstr[row + 1]; k++)
                cnt[(size_t)t * nblk + colidx[k] / W]++;
        }
    });
    std::vector<long long> base(cnt.size() + 1, 0);
    for (size_t i = 0; i < cnt.size(); i++) base[i + 1] = base[i] + cnt[i];
    if (base.back() != m.nnz || m.nnz > (long long)INT32_MAX) return false;

    m.seg.assign(segs, 0);
    m.val.resize((size_t)m.nnz);
    m.col.resize((size_t)m.nnz);
    for_tiles(ntile, [&](int t) {
        std::vector<int> cursor(T);
        for (int i = 0; i < T; i++) {
            const int row = t * T + i;
            cursor[i] = row < n ? rowstr[row] : 0;
        }
        for (int b = 0; b < nblk; b++) {
            long long off = base[(size_t)t * nblk + b];
            int *sg = &m.seg[((size_t)t * nblk + b) * (T + 1)];
            const int lim = (b + 1) * W;  // b*W <= col < lim
            for (int i = 0; i < T; i++) {
                sg[i] = (int)off;
                const int row = t * T + i;
                if (row >= n) continue;
                int k = cursor[i];
                const int e = rowstr[row + 1];
                while (k < e && colidx[k] < lim) {
                    m.val[off] = (V)a[k];
                    m.col[off] = (uint16_t)(colidx[k] - b * W);
                    off++;
                    k++;
                }
                cursor[i] = k;
            }
            sg[T] = (int)off;
        }
    });
    return true;
}

}  // namespace

bool build_tiled(int n, const int *rowstr, const int *colidx, const double *a, int T, int W,
                 TiledMatrix<double> *out) {
    return build_impl(n, rowstr, colidx, a, T, W, out);
}

bool build_tiled(int n, const int *rowstr, const int *colidx, const double *a, int T, int W,
                 TiledMatrix<float> *out) {
    return build_impl(n, rowstr, colidx, a, T, W, out);
}

}  // namespace npbcg_detail
