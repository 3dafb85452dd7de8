// Batched CG over independent systems with NR right-hand sides each
// (batch_kernels.hip): device layout and launchers.
#ifndef NPBCG_BATCH_KERNELS_H
#define NPBCG_BATCH_KERNELS_H

#include <hip/hip_runtime.h>

#include <cstdint>

namespace npbcg_dev {

// Rows of a SELL sorting window (layout 3): a system's rows are sorted by
// their segment length inside windows of this many rows.
constexpr int kSellWindow = 2048;

// Per (system, right-hand side) scalars of the inner CG.
struct BatchScalars {
    double rtol2;     // inner stop level relative to the start r.r (set by the host)
    double tol2;      // rtol2 * r0.r0, set by the first step
    double rho;       // r.r of the last step
    int steps;        // inner CG steps, all sweeps
    int sweep_steps;  // inner CG steps of this sweep
};

// The concatenated batch (device pointers) and its launch shape.
struct BatchDev {
    int nsys = 0;
    int nr = 1;          // right-hand sides per system (1, 2 or 4)
    int lpr = 16;        // lanes per row of the SpMM (4 .. 64)
    int tiles = 0;       // workgroups; tiles never cross systems
    int rows = 0;        // total rows
    int maxit = 0;       // inner steps per sweep
    int blocks = 0;                 // workgroups of the gathering kernels (SpMM, residual)
    const int *block_tile = nullptr;  // [blocks] tile of each of those workgroups, -1: none.
                                      // Workgroup b runs on XCD b % 8, so the tiles of one
                                      // system are placed on one XCD and its vector stays
                                      // in that XCD's L2.
    const int *tile_sys = nullptr;  // [tiles]
    const int *tile_row = nullptr;  // [tiles + 1] first row of each tile
    const int *sys_tile = nullptr;  // [nsys + 1] first tile of each system
    const int *rowptr = nullptr;    // [rows + 1]
    const int *col = nullptr;       // global column indices
    // Column-blocked SpMM (cb = 1): system s's columns are cut into blocks of
    // cbw columns whose vector entries are staged in LDS; one workgroup per
    // item (system, column block, row range) writes the block's row partials,
    // which the update kernel adds up (q is never stored).
    int cb = 0;
    int cbw = 0;                    // column block width
    int items = 0;
    const int *item = nullptr;      // [items][4]: system, block, first row, end row (global)
    const int *sys_item = nullptr;  // [nsys + 1] first item of each system
    const int *sys_row = nullptr;   // [nsys + 1] first global row of each system
    const int *sys_nb = nullptr;    // [nsys] column blocks of each system
    const long long *sys_seg = nullptr;   // [nsys] start of the system's segment table
    const long long *sys_part = nullptr;  // [nsys] start of the system's row partials
    const int *seg = nullptr;       // per row, nb + 1 entry offsets: block b's entries
                                    // are [seg[b], seg[b+1]) (columns sorted in a row)
    // SELL SpMM (sell = 1, also cb = 1): a column-blocked SELL copy of every
    // system (blocks of cbw columns, rows sorted by segment length inside
    // kSellWindow-row windows, 64-row slices stored column-major, 16-bit
    // columns inside the block).  Item = one window: (system, first local row,
    // rows, offset into item_sl), where item_sl holds the window's slice range
    // [first, end) in each of the system's blocks.  The SpMM writes q (not
    // block partials) and one p.q partial per item.
    int sell = 0;
    const int *item_sl = nullptr;
    int item_blocks = 0;               // workgroups of the SELL SpMM (one per CU)
    const int *block_item = nullptr;   // [item_blocks + 1] list offsets, then the lists:
                                       // workgroup g runs items list[first[g] .. first[g+1])
    const int *sell_off = nullptr;     // [slices + 1] first entry of each slice
    const int *sell_row = nullptr;     // [slices * 64] global row of each lane, -1: padding
    const uint16_t *sell_col = nullptr;  // [stored] column inside the block
    int *done = nullptr;            // [nsys * nr] right-hand side finished
    BatchScalars *sc = nullptr;     // [nsys * nr]
};

// Inner-solve vectors ([rows][nr]) and partial sums ([tiles][nr]) in type T.
template <typename T>
struct BatchVecs {
    const T *val = nullptr;   // matrix values in type T
    const T *sval = nullptr;  // SELL copy's values in type T (layout 3)
    T *d = nullptr, *r = nullptr, *p = nullptr, *q = nullptr;
    T *rows_part = nullptr;   // column-blocked SpMM: per-block row partials
    double *part_init = nullptr, *part_pq = nullptr, *part_rr[2] = {nullptr, nullptr};
};

// Step `step` of an inner solve of every unfinished system: SpMM + p.q,
// x/r update + r.r, p update + stop test.
template <typename T>
hipError_t batch_cg_step(const BatchDev &bd, const BatchVecs<T> &v, int step, int first,
                         hipStream_t st);
// Dynamic LDS of the SELL SpMM: the vector block (cbw rows of nr values)
// and the row sums of one sorting window.
inline size_t sell_lds_bytes(int cbw, int nr, size_t elem) {
    return (size_t)(cbw + kSellWindow) * nr * elem;
}
// Allows `bytes` of dynamic LDS for the column-blocked and SELL SpMM kernels
// of type T.
template <typename T>
hipError_t batch_prepare_cb(int bytes);
// Allows `bytes` of dynamic LDS for the FP64 SELL residual kernels.
hipError_t batch_prepare_resid(int bytes);
// Workgroup size of the SELL SpMM: one 1024-thread workgroup per CU, which
// holds a column block of up to 128 KB in LDS.
constexpr int kSellBlock = 1024;
// q = A p and p.q partials only (the SpMM the autotuner times); rev: item
// order of the SELL SpMM (0 forward, 1 backward within each XCD).
template <typename T>
hipError_t batch_spmm(const BatchDev &bd, const BatchVecs<T> &v, int rev, hipStream_t st);
// r = p = (T) res, d = 0 for the unfinished right-hand sides.
template <typename T>
hipError_t batch_init(const BatchDev &bd, const BatchVecs<T> &v, const double *res,
                      hipStream_t st);
template <typename T>
hipError_t batch_refine(const BatchDev &bd, double *x, const T *d, hipStream_t st);
// res = b - A x (FP64) and |res|^2 per (system, rhs) into sys_out[nsys * nr].
hipError_t batch_residual(const BatchDev &bd, const double *val, const double *b, const double *x,
                          double *res, double *part, double *sys_out, hipStream_t st);
// The same from an FP64 SELL copy described by bd (its window items and
// column blocks): one SELL SpMM writes res and the |res|^2 partials of its
// items, then the per-system sums.
hipError_t batch_residual_sell(const BatchDev &bd, const double *sval, const double *b,
                               const double *x, double *res, double *part, double *sys_out,
                               hipStream_t st);
// Residual of x = 0 (res = b).
hipError_t batch_residual_zero(const BatchDev &bd, const double *b, double *res, double *part,
                               double *sys_out, hipStream_t st);
hipError_t batch_narrow(const double *in, float *out, long long count, hipStream_t st);

}  // namespace npbcg_dev

#endif
