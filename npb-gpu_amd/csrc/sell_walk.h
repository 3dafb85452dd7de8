// One wavefront's walk over its slices of the column-blocked SELL SpMV
// (layout in tiled_layout.h), shared by the kernels that stream the matrix
// inside a larger kernel.
//
// Workgroup wg owns column block b = wg / nrange of row range wg % nrange and
// has nsl slices of 64 rows.  Wavefront w takes slice rd * waves + (w or
// waves-1-w) in round rd ("snake" order; the layout places the slices so
// that the wavefronts' total widths are close) and walks its rounds forward
// or, with rev, backward: successive SpMVs alternate so each starts on the slices the
// previous one streamed last, which the memory-side cache still holds.  Lane
// l accumulates its own row down the slice in entry order, so neither the
// order nor the direction changes any sum.
//
// The entry stream is software-pipelined in chunks of UNR entries
// (sell_chunk): the next chunk's loads, or the next slice's first chunk and
// lane rows, go out before the current chunk's LDS gathers.  begin() loads
// the first chunk, so a caller that stages the vector block afterwards
// overlaps that memory latency with the fill.
#ifndef NPBCG_SELL_WALK_H
#define NPBCG_SELL_WALK_H

#include "cg_device.h"

namespace npbcg_dev {

template <typename T, int BLOCK, int UNR>
struct SellWalk {
    static constexpr int kWaves = BLOCK / kWave;
    const int *slice_off;
    const int *slice_row;
    const T *val;
    const uint16_t *col;
    int s0, nsl, wave, lane, nrw, rev;
    int cur, cbase, cwidth, nrow;
    T na[UNR];
    int nc[UNR];

    // Slice (index in the workgroup) of round rd for this wave, or -1; the
    // index grows with rd, so after the first -1 there is none.
    __device__ __forceinline__ int slice_of(int rd) const {
        const int k = rd * kWaves + ((rd & 1) ? kWaves - 1 - wave : wave);
        return k < nsl ? k : -1;
    }

    __device__ __forceinline__ int slice_at(int i) const {
        return i < nrw ? slice_of(rev ? nrw - 1 - i : i) : -1;
    }

    // Locates this wave's slices of workgroup wg and loads the first chunk.
    // wave and the slice bounds are wave-uniform, and readfirstlane says so
    // (the compiler cannot see it when the index arrays come from memory
    // rather than from restrict kernel arguments), so every branch on them is
    // a scalar branch and the wavefront stays converged around the
    // cross-lane reductions of the callers.
    __device__ __forceinline__ void begin(const int *wg_slice, const int *so, const int *sr,
                                          const T *v, const uint16_t *c, int wg, int reverse) {
        slice_off = so;
        slice_row = sr;
        val = v;
        col = c;
        int t = threadIdx.x;
        asm volatile("" : "+v"(t));  // not hoisted out of a caller's loop
        lane = t & (kWave - 1);
        wave = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
        s0 = __builtin_amdgcn_readfirstlane(wg_slice[wg]);
        nsl = __builtin_amdgcn_readfirstlane(wg_slice[wg + 1]) - s0;
        rev = reverse;
        nrw = 0;
        while (slice_of(nrw) >= 0) nrw++;
        cur = slice_at(0);
        cbase = 0;
        cwidth = 0;
        nrow = -1;
        if (cur >= 0) {
            cbase = __builtin_amdgcn_readfirstlane(slice_off[s0 + cur]);
            cwidth = (__builtin_amdgcn_readfirstlane(slice_off[s0 + cur + 1]) - cbase) / kWave;
            nrow = slice_row[(size_t)(s0 + cur) * kWave + lane];
            if (cwidth > 0) sell_chunk<T, UNR>(val, col, cbase, cwidth, 0, lane, na, nc);
        }
    }

    // Streams every slice against the staged block vblk.  start(row) runs
    // when a slice begins (row = this lane's row, -1 for padding) and returns
    // a value handed back to finish(k, row, acc, value) with the lane's row
    // sum acc once the slice is done (k = slice index in the workgroup).
    template <class Start, class Finish>
    __device__ __forceinline__ void run(const T *vblk, Start start, Finish finish) const {
        // the walk's state in locals (not members), so nothing of it has an
        // address and every value stays in registers
        const int *const so = slice_off;
        const int *const sr = slice_row;
        const T *const vp = val;
        const uint16_t *const cp = col;
        const int ln = lane, sb = s0;
        int c_cur = cur, c_base = cbase, c_width = cwidth, c_row = nrow;
        T pa[UNR];
        int pc[UNR];
#pragma unroll
        for (int u = 0; u < UNR; u++) {
            pa[u] = na[u];
            pc[u] = nc[u];
        }
        for (int i = 0; c_cur >= 0; i++) {
            const int k = c_cur, width = c_width, base = c_base, row = c_row;
            const int nxt = slice_at(i + 1);
            int xbase = 0, xwidth = 0;
            if (nxt >= 0) {
                xbase = __builtin_amdgcn_readfirstlane(so[sb + nxt]);
                xwidth = (__builtin_amdgcn_readfirstlane(so[sb + nxt + 1]) - xbase) / kWave;
            }
            const auto pr = start(row);
            T acc = T(0);
            const int nch = (width + UNR - 1) / UNR;
            for (int ch = 0; ch < nch; ch++) {
                T av[UNR];
                int c[UNR];
#pragma unroll
                for (int u = 0; u < UNR; u++) {
                    av[u] = pa[u];
                    c[u] = pc[u];
                }
                sell_trim<T, UNR>(width - ch * UNR, av, c);
                if (ch + 1 < nch) {
                    sell_chunk<T, UNR>(vp, cp, base, width, (ch + 1) * UNR, ln, pa, pc);
                } else if (nxt >= 0) {
                    c_row = sr[(size_t)(sb + nxt) * kWave + ln];
                    if (xwidth > 0) sell_chunk<T, UNR>(vp, cp, xbase, xwidth, 0, ln, pa, pc);
                }
#pragma unroll
                for (int u = 0; u < UNR; u++) acc = fma(av[u], vblk[c[u]], acc);
            }
            if (nch == 0 && nxt >= 0) {
                c_row = sr[(size_t)(sb + nxt) * kWave + ln];
                if (xwidth > 0) sell_chunk<T, UNR>(vp, cp, xbase, xwidth, 0, ln, pa, pc);
            }
            finish(k, row, acc, pr);
            c_cur = nxt;
            c_base = xbase;
            c_width = xwidth;
        }
    }
};

}  // namespace npbcg_dev

#endif
