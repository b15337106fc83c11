// Internal interface of the R^k contraction kernels.
#pragma once

#include <cstdint>
#include <vector>

#include "ax_common.h"

namespace ax {

// One output block: R = D[rows] * Y^T (m x m, upper triangle, packed at
// out_base + out_off).  For the split path the operands are read from their
// int8 slice planes (Ds/Ys, plane stride *_plane elements, row pitch *_ld)
// and per-(row, 64-wide K-tile) exponents (De/Ye, row pitch *_eld).
struct RkSeg {
    const double *D;
    int64_t ld_d;
    const int32_t *rows;  // D row of output row i (nullptr = identity)
    const double *Y;
    int64_t ld_y;
    int64_t n;
    int64_t out_off;
    int m;
    int pad0;
    const int8_t *Ds;
    int64_t ds_ld, ds_plane;
    const int32_t *De;
    int64_t de_ld;
    const int8_t *Ys;
    int64_t ys_ld, ys_plane;
    const int32_t *Ye;
    int64_t ye_ld;
};

struct RkWork {
    int seg;
    int ti;
    int tj;
    int pad;
};

constexpr int kSplitSlices = 7;   // int8 slices per operand (7 bits each)
constexpr int kSplitKTile = 64;   // K-tile that shares one exponent

int gemm_tile();
int64_t rk_gemm_partial_doubles(int n_work, int n_split);
void upper_tiles(int m, int seg, std::vector<RkWork> &out);

// Launches the contraction (and the deterministic split reduction).
ax_status rk_gemm_launch(const RkSeg *segs_dev, const RkWork *work_dev, int n_work, int n_split,
                         int precision, double *partial, double *out_base, hipStream_t stream);
ax_status rk_gemm_contract_launch(const RkSeg *segs_dev, const RkWork *work_dev, int n_work,
                                 int n_split, int precision, double *partial, hipStream_t stream);
ax_status rk_gemm_reduce_launch(const RkSeg *segs_dev, const RkWork *work_dev, int n_work,
                               int n_split, const double *partial, double *out_base,
                               hipStream_t stream);
ax_status rk_gemm_split_launch(const RkSeg *segs_dev, const RkWork *work_dev, int n_work,
                               int n_split, double *partial, hipStream_t stream);

// Slices n_rows rows of src (row r at src + rows[r]*ld, or r*ld) into
// kSplitSlices int8 planes dst[s*plane + r*dst_ld + x] (x < npad, zero
// padded beyond n) and exponents e[r*e_ld + t] per 64-wide K-tile.
ax_status slice_rows_launch(const double *src, int64_t ld, const int32_t *rows, int n_rows,
                            int64_t n, int64_t npad, int8_t *dst, int64_t dst_ld,
                            int64_t dst_plane, int32_t *e, int64_t e_ld, hipStream_t stream);

}  // namespace ax
