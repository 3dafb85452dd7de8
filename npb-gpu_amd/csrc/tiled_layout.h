// Column-blocked ("tiled") copy of a CSR matrix for the LDS-staged SpMV.
//
// Rows are cut into tiles of T rows and columns into blocks of W columns.
// Storage order is tile -> block -> row -> column, so the part of a tile that
// falls into one column block is contiguous.  Column indices are stored as
// 16-bit offsets inside their block (W <= 65536).  seg[(t*nblk + b)*(T+1) + i]
// is the offset of row t*T+i's segment in block b; entry T ends the last row.
// Rows past n in the last tile have empty segments.  Within a row the
// entries keep their CSR (ascending column) order.
#ifndef NPBCG_TILED_LAYOUT_H
#define NPBCG_TILED_LAYOUT_H

#include <cstdint>
#include <vector>

namespace npbcg_detail {

template <class V>
struct TiledMatrix {
    int n = 0, T = 0, W = 0, nblk = 0, ntile = 0;
    long long nnz = 0;
    std::vector<int> seg;        // ntile * nblk * (T+1)
    std::vector<V> val;          // nnz
    std::vector<uint16_t> col;   // nnz
};

// Builds the tiled copy.  Returns false if W > 65536 or the sizes overflow.
bool build_tiled(int n, const int *rowstr, const int *colidx, const double *a,
                 int T, int W, TiledMatrix<double> *out);
bool build_tiled(int n, const int *rowstr, const int *colidx, const double *a,
                 int T, int W, TiledMatrix<float> *out);

}  // namespace npbcg_detail

#endif
