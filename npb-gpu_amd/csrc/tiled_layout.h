// Column-blocked SELL copy of a CSR matrix for the LDS-staged SpMV.
//
// Columns are cut into nblk blocks of W columns and rows into nrange ranges
// of R rows.  Workgroup wg = b * nrange + rr owns the part of row range rr that
// falls into column block b.  Inside a workgroup the rows are ordered by their
// segment length in block b (longest first, ties by row) within sort windows
// (build_sell_blocks), and cut into slices of 64 rows; slice s stores its
// entries column-major: entry j of the slice's lane l sits at
// slice_off[s] + 64*j + l, and rows shorter than the slice
// width are padded with (value 0, column 0).  So one wavefront streams a slice
// with fully coalesced loads, one row per lane, and no cross-lane reduction.
// Column indices are 16-bit offsets inside the block (W <= 65536).
//   wg_slice[wg] .. wg_slice[wg+1]   slices of workgroup wg
//   slice_off[s] .. slice_off[s+1]   entries of slice s (multiple of 64)
//   slice_row[64*s + l]              row (global) of lane l, -1 for padding
//   row_slot[b*n + j]                slot 64*s + l of row j in column block b
// Within a row the entries keep their CSR (ascending column) order.  Every
// row has one slot in each column block (rows with no entry there sit in a
// slice with value 0), so the FP64 kernels store a row's block sum at its
// slot, 64 consecutive sums per slice, and the block sums of row j are found
// through row_slot.
#ifndef NPBCG_TILED_LAYOUT_H
#define NPBCG_TILED_LAYOUT_H

#include <cstdint>
#include <vector>

namespace npbcg_detail {

constexpr int kSliceRows = 64;

template <class V>
struct SellBlocks {
    int n = 0, W = 0, nblk = 0, R = 0, nrange = 0;
    long long nnz = 0;     // entries of the CSR
    long long stored = 0;  // entries stored, padding included
    std::vector<int> wg_slice;   // nblk * nrange + 1
    std::vector<int> slice_off;  // slices + 1
    std::vector<int> slice_row;  // slices * 64
    std::vector<int> row_slot;   // nblk * n
    std::vector<V> val;          // stored
    std::vector<uint16_t> col;   // stored
};

// Rows are sorted by segment length within windows of `window` rows (a
// multiple of 64; kSortWindow unless the caller widens it).  With waves > 0
// (wavefronts per workgroup) a workgroup's slices are placed so that the
// wavefronts of the kernels' round order get close total widths; with 0 they
// stay in sorted order (widest first).
constexpr int kSortWindow = 2048;

// Builds the copy; the CSR must have ascending columns within each row.
// Returns false if W > 65536, the window is not a positive multiple of 64
// or the sizes overflow 32-bit offsets.
bool build_sell_blocks(int n, const int *rowstr, const int *colidx, const double *a,
                       int W, int R, SellBlocks<double> *out, int window = kSortWindow,
                       int waves = 0);
bool build_sell_blocks(int n, const int *rowstr, const int *colidx, const double *a,
                       int W, int R, SellBlocks<float> *out, int window = kSortWindow,
                       int waves = 0);

}  // namespace npbcg_detail

#endif
