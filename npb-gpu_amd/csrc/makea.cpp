// Host-side NPB CG matrix generator for npbcg (product code).
//
// Produces exactly the CSR the reference assembles (CUDA/CG/cg.cu makea
// :1303-1368, sparse :1611-1785, sprnvc :1798-1833, vecset :1841-1857, RNG set
// up at :406-408), but not by the reference's method.  The reference inserts
// every (row, column, value) triple into a pre-sized row with shifting, which is
// O(triples * row length) and serial.  Here:
//   1. the sparse vectors are drawn serially (the LCG stream with rejections is
//      inherently sequential, and cheap);
//   2. each vector's outer-product triples are scattered into row buckets in
//      vector order, in parallel over vector ranges (per-thread row cursors keep
//      the vector order inside each row);
//   3. every row is stably sorted by column and equal columns are folded left to
//      right from 0.0 — the same additions, in the same order, as the
//      reference's `a[k] = a[k] + va` on a slot that starts at 0.0.
// Bit-exactness therefore rests on: identical LCG integers (exact arithmetic),
// identical per-triple products (no FMA contraction: built -ffp-contract=off),
// identical scale sequence size_{i+1} = size_i * ratio with ratio from libm
// pow, and identical fold order per (row, column).
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <thread>
#include <vector>

#include "npbcg.h"
#include "npbcg_internal.h"

namespace {

// 46-bit multiplicative LCG x <- a*x mod 2^46 (Bailey's NPB generator), in
// exact 128-bit integer arithmetic.  The returned double 2^-46 * x is exact.
struct Lcg46 {
    uint64_t x;
    uint64_t a;
    double next() {
        unsigned __int128 prod = (unsigned __int128)a * x;
        x = (uint64_t)(prod & ((((unsigned __int128)1) << 46) - 1));
        return std::ldexp((double)x, -46);
    }
};

thread_local int t_gen_threads = 0;  // per-thread override (0: none)

int gen_threads() {
    if (t_gen_threads > 0)
        return t_gen_threads;
    const char *env = std::getenv("NPBCG_GEN_THREADS");
    if (env && std::atoi(env) > 0)
        return std::atoi(env);
    unsigned hc = std::thread::hardware_concurrency();
    if (hc == 0)
        hc = 1;
    return (int)std::min(hc, 16u);
}

template <class F>
void parallel_for(int nthreads, long count, F &&body) {
    if (nthreads <= 1 || count < 4096) {
        body(0, 0L, count);
        return;
    }
    std::vector<std::thread> pool;
    long chunk = (count + nthreads - 1) / nthreads;
    for (int t = 0; t < nthreads; t++) {
        long lo = t * chunk, hi = std::min(count, lo + chunk);
        if (lo >= hi)
            break;
        pool.emplace_back([&, t, lo, hi] { body(t, lo, hi); });
    }
    for (auto &th : pool)
        th.join();
}

}  // namespace

namespace npbcg_detail {

// Worker threads of the generator used by npbcg_makea calls from this thread.
void set_generator_threads(int threads) { t_gen_threads = threads; }

}  // namespace npbcg_detail

extern "C" void npbcg_set_generator_threads(int threads) {
    npbcg_detail::set_generator_threads(threads > 0 ? threads : 0);
}

extern "C" long npbcg_makea(int n, int nonzer, double rcond, double shift,
                            double seed, long nz_cap, int *rowstr, int *colidx,
                            double *a, double *ratio_out) {
    using namespace npbcg_detail;
    if (n < 1 || nonzer < 0 || nonzer > n || nz_cap < 1 || !rowstr || !colidx ||
        !a)
        return set_error(NPBCG_ERR_INVALID, "npbcg_makea: invalid argument");
    if (!(seed > 0.0) || seed >= 70368744177664.0 || seed != std::floor(seed))
        return set_error(NPBCG_ERR_INVALID,
                         "npbcg_makea: seed must be an integer in (0, 2^46)");
    const int w = nonzer + 1;

    // ---- 1. sparse vectors (serial LCG stream) -----------------------------
    Lcg46 rng{(uint64_t)seed, 1220703125ull};  // a = 5^13
    (void)rng.next();                          // the discarded draw of :408
    int nn1 = 1;
    do {
        nn1 *= 2;
    } while (nn1 < n);

    std::vector<int> vlen(n);
    std::vector<int> vcol((size_t)n * w);
    std::vector<double> vval((size_t)n * w);
    for (int i = 0; i < n; i++) {
        int *col = &vcol[(size_t)i * w];
        double *val = &vval[(size_t)i * w];
        int cnt = 0;
        while (cnt < nonzer) {
            double v = rng.next();
            double u = rng.next();
            int pos = (int)(nn1 * u);  // 0-based; reference uses pos+1 <= n
            if (pos >= n)
                continue;
            bool seen = false;
            for (int t = 0; t < cnt; t++)
                if (col[t] == pos) {
                    seen = true;
                    break;
                }
            if (seen)
                continue;
            col[cnt] = pos;
            val[cnt] = v;
            cnt++;
        }
        // the vector's own index carries 0.5 (vecset)
        bool present = false;
        for (int t = 0; t < cnt; t++)
            if (col[t] == i) {
                val[t] = 0.5;
                present = true;
            }
        if (!present) {
            col[cnt] = i;
            val[cnt] = 0.5;
            cnt++;
        }
        vlen[i] = cnt;
    }

    // ---- 2. row buckets of outer-product triples, in vector order ----------
    const int nth = gen_threads();
    std::vector<long> row_triples(n + 1, 0);  // triples landing in each row
    for (int i = 0; i < n; i++)
        for (int s = 0; s < vlen[i]; s++)
            row_triples[vcol[(size_t)i * w + s]] += vlen[i];
    long total = 0;
    for (int j = 0; j < n; j++) {
        long c = row_triples[j];
        row_triples[j] = total;
        total += c;
    }
    row_triples[n] = total;
    // The reference sizes rows by triple count and rejects nza > nz
    // (cg.cu:1663-1674) with nza = triples - 1.
    if (total - 1 > nz_cap)
        return set_error(NPBCG_ERR_CAPACITY,
                         "npbcg_makea: space for matrix elements exceeded");
    if (total > (long)INT32_MAX)
        return set_error(NPBCG_ERR_CAPACITY,
                         "npbcg_makea: more than 2^31-1 triples (int32 rowstr)");

    // scale of vector i is the running product 1*ratio*ratio*... (:1755)
    const double ratio = std::pow(rcond, 1.0 / (double)n);
    if (ratio_out)
        *ratio_out = ratio;
    std::vector<double> vscale(n);
    {
        double size = 1.0;
        for (int i = 0; i < n; i++) {
            vscale[i] = size;
            size = size * ratio;
        }
    }

    std::vector<int> tcol((size_t)total);
    std::vector<double> tval((size_t)total);
    {
        // Per-thread row cursors: thread t owns vectors [lo_t, hi_t); its
        // triples for row j start after those of threads < t.
        int T = (n >= 4096) ? nth : 1;
        long chunk = ((long)n + T - 1) / T;
        std::vector<std::vector<int>> counts(T, std::vector<int>(n, 0));
        parallel_for(T, (long)T, [&](int, long t0, long t1) {
            for (long t = t0; t < t1; t++) {
                long lo = t * chunk, hi = std::min((long)n, lo + chunk);
                std::vector<int> &c = counts[t];
                for (long i = lo; i < hi; i++)
                    for (int s = 0; s < vlen[i]; s++)
                        c[vcol[(size_t)i * w + s]] += vlen[i];
            }
        });
        // exclusive scan over threads for each row
        parallel_for(nth, (long)n, [&](int, long j0, long j1) {
            for (long j = j0; j < j1; j++) {
                long base = row_triples[j];
                for (int t = 0; t < T; t++) {
                    int c = counts[t][j];
                    counts[t][j] = (int)(base - row_triples[j]);
                    base += c;
                }
            }
        });
        parallel_for(T, (long)T, [&](int, long t0, long t1) {
            for (long t = t0; t < t1; t++) {
                long lo = t * chunk, hi = std::min((long)n, lo + chunk);
                std::vector<int> &cur = counts[t];
                for (long i = lo; i < hi; i++) {
                    const int *col = &vcol[(size_t)i * w];
                    const double *val = &vval[(size_t)i * w];
                    const int len = vlen[i];
                    for (int s = 0; s < len; s++) {
                        const int j = col[s];
                        const double scale = vscale[i] * val[s];
                        long at = row_triples[j] + cur[j];
                        cur[j] += len;
                        for (int q = 0; q < len; q++) {
                            double va = val[q] * scale;
                            if (col[q] == j && j == (int)i)
                                va = va + rcond - shift;
                            tcol[at + q] = col[q];
                            tval[at + q] = va;
                        }
                    }
                }
            }
        });
    }

    // ---- 3. per-row stable sort by column and left fold of duplicates -------
    std::vector<int> distinct(n, 0);
    parallel_for(nth, (long)n, [&](int, long j0, long j1) {
        std::vector<int> order;
        std::vector<int> kcol;
        std::vector<double> kval;
        for (long j = j0; j < j1; j++) {
            long b = row_triples[j], e = row_triples[j + 1];
            int len = (int)(e - b);
            order.resize(len);
            for (int k = 0; k < len; k++)
                order[k] = k;
            const int *c = &tcol[b];
            std::stable_sort(order.begin(), order.end(),
                             [c](int x, int y) { return c[x] < c[y]; });
            kcol.clear();
            kval.clear();
            for (int k = 0; k < len; k++) {
                int col = c[order[k]];
                double v = tval[b + order[k]];
                if (kcol.empty() || kcol.back() != col) {
                    kcol.push_back(col);
                    kval.push_back(0.0 + v);
                } else {
                    kval.back() = kval.back() + v;
                }
            }
            // write back in place (distinct <= len), compacted later
            for (size_t k = 0; k < kcol.size(); k++) {
                tcol[b + k] = kcol[k];
                tval[b + k] = kval[k];
            }
            distinct[j] = (int)kcol.size();
        }
    });

    // ---- 4. compaction into the caller's CSR -------------------------------
    long nnz = 0;
    rowstr[0] = 0;
    for (int j = 0; j < n; j++) {
        nnz += distinct[j];
        rowstr[j + 1] = (int)nnz;
    }
    if (nnz > nz_cap)
        return set_error(NPBCG_ERR_CAPACITY,
                         "npbcg_makea: nz_cap smaller than the assembled nnz");
    parallel_for(nth, (long)n, [&](int, long j0, long j1) {
        for (long j = j0; j < j1; j++) {
            long src = row_triples[j];
            long dst = rowstr[j];
            for (int k = 0; k < distinct[j]; k++) {
                colidx[dst + k] = tcol[src + k];
                a[dst + k] = tval[src + k];
            }
        }
    });
    clear_error();
    return nnz;
}
