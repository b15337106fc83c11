// Split-precision R^k contraction: error-free int8 slicing on the int8 MFMA.
//
// Every operand row is cut into 64-wide K-tiles.  Each (row, tile) gets a
// power-of-two exponent e with max|v| < 2^e, and the scaled values v*2^-e in
// (-1, 1) are written as kSplitSlices signed 7-bit slices c_s (truncation),
//     v = 2^e * sum_{s=1..S} c_s 2^{-7s} + r,   |r| < 2^{e-7S}.
// A slice product summed over one K-tile is an exact int32 sum on
// v_mfma_i32_16x16x64_i8 (|c| <= 127, 64 terms).  Per output element and
// K-tile the levels L = s+s' <= S+1 are combined in FP64 (Horner in 2^-7),
// scaled by 2^{eD+eY} and accumulated in FP64.  The dropped levels and the
// truncation remainder bound the error of each term by about
// (S+2) 2^{-7S} |local max D| |local max Y| — relative to the magnitudes inside
// the 64-point tile, not to the row maximum, so the dynamic range of r^k
// across the radial grid does not cost accuracy.
#include "ax_common.h"
#include "rk_gemm.h"

namespace {

typedef int v4i __attribute__((ext_vector_type(4)));
typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int S = ax::kSplitSlices;
constexpr int KT = ax::kSplitKTile;
constexpr int BM = 64;

__global__ __launch_bounds__(256) void slice_rows_kernel(const double *__restrict__ src,
                                                         int64_t ld,
                                                         const int32_t *__restrict__ rows,
                                                         int n_rows, int64_t n, int64_t npad,
                                                         int8_t *__restrict__ dst, int64_t dst_ld,
                                                         int64_t dst_plane,
                                                         int32_t *__restrict__ e, int64_t e_ld) {
    const int lane = threadIdx.x & 63;
    const int64_t n_tiles = npad / KT;
    const int64_t total = (int64_t)n_rows * n_tiles;
    for (int64_t item = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); item < total;
         item += (int64_t)gridDim.x * 4) {
        const int r = static_cast<int>(item / n_tiles);
        const int64_t t = item - (int64_t)r * n_tiles;
        const int64_t x = t * KT + lane;
        const int64_t rr = rows ? (int64_t)rows[r] : (int64_t)r;
        const double v = x < n ? src[rr * ld + x] : 0.0;
        double mx = fabs(v);
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
        const int ex = mx > 0.0 ? ilogb(mx) + 1 : 0;
        double rho = ldexp(v, 7 - ex);  // v * 2^-e * 2^7, |rho| < 128
        int8_t *d = dst + (int64_t)r * dst_ld + x;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const double c = trunc(rho);
            d[(int64_t)s * dst_plane] = static_cast<int8_t>(c);
            rho = (rho - c) * 128.0;
        }
        if (lane == 0) e[(int64_t)r * e_ld + t] = ex;
    }
}

__device__ inline v4i load16(const int8_t *p) { return *reinterpret_cast<const v4i *>(p); }

__global__ __launch_bounds__(256) void rk_gemm_split_kernel(const ax::RkSeg *__restrict__ segs,
                                                            const ax::RkWork *__restrict__ work,
                                                            int n_work, int n_split,
                                                            double *__restrict__ partial) {
    const int e = blockIdx.x;
    const int split = blockIdx.y;
    if (e >= n_work) return;
    const ax::RkWork wk = work[e];
    const ax::RkSeg sg = segs[wk.seg];
    const int64_t n_tiles = (sg.n + KT - 1) / KT;
    const int64_t per = (n_tiles + n_split - 1) / n_split;
    const int64_t tb = (int64_t)split * per;
    const int64_t te = tb + per < n_tiles ? tb + per : n_tiles;

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    const int fr = lane & 15, fg = lane >> 4;

    // operand rows of this lane (A: output rows, B: output columns)
    int64_t arow[2], brow[2];
    bool aok[2], bok[2];
    int64_t erow[2][4];  // exponent rows of the 4 output rows this lane holds
    bool eok[2][4];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const int i = wk.ti * BM + wr + a * 16 + fr;
        aok[a] = i < sg.m;
        arow[a] = aok[a] ? (sg.rows ? (int64_t)sg.rows[i] : (int64_t)i) : 0;
        const int j = wk.tj * BM + wc + a * 16 + fr;
        bok[a] = j < sg.m;
        brow[a] = bok[a] ? (int64_t)j : 0;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int io = wk.ti * BM + wr + a * 16 + 4 * fg + r;
            eok[a][r] = io < sg.m;
            erow[a][r] = eok[a][r] ? (sg.rows ? (int64_t)sg.rows[io] : (int64_t)io) : 0;
        }
    }

    d4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};

    const v4i zero = {0, 0, 0, 0};
    for (int64_t t = tb; t < te; ++t) {
        const int64_t xo = t * KT + fg * 16;
        v4i av[2][S], bv[2][S];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int s = 0; s < S; ++s) {
                av[a][s] = aok[a] ? load16(sg.Ds + s * sg.ds_plane + arow[a] * sg.ds_ld + xo) : zero;
                bv[a][s] = bok[a] ? load16(sg.Ys + s * sg.ys_plane + brow[a] * sg.ys_ld + xo) : zero;
            }
        int ea[2][4], eb[2];
#pragma unroll
        for (int a = 0; a < 2; ++a) {
#pragma unroll
            for (int r = 0; r < 4; ++r) ea[a][r] = eok[a][r] ? sg.De[erow[a][r] * sg.de_ld + t] : 0;
            eb[a] = bok[a] ? sg.Ye[brow[a] * sg.ye_ld + t] : 0;
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                // levels L = 2..S+1, highest first for Horner
                double lvl[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int L = S + 1; L >= 2; --L) {
                    v4i c = zero;
#pragma unroll
                    for (int s = 1; s <= S; ++s) {
                        const int s2 = L - s;
                        if (s2 < 1 || s2 > S) continue;
                        c = __builtin_amdgcn_mfma_i32_16x16x64_i8(av[a][s - 1], bv[b][s2 - 1], c,
                                                                  0, 0, 0);
                    }
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        lvl[r] = (L == S + 1) ? static_cast<double>(c[r])
                                              : fma(lvl[r], 0x1p-7, static_cast<double>(c[r]));
                }
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[a][b][r] += ldexp(lvl[r], ea[a][r] + eb[b] - 14);
            }
    }

    double *dst = partial + ((int64_t)e * n_split + split) * (BM * BM);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = wr + a * 16 + 4 * fg + r;
                const int j = wc + b * 16 + fr;
                dst[i * BM + j] = acc[a][b][r];
            }
}

}  // namespace

namespace ax {

ax_status slice_rows_launch(const double *src, int64_t ld, const int32_t *rows, int n_rows,
                            int64_t n, int64_t npad, int8_t *dst, int64_t dst_ld,
                            int64_t dst_plane, int32_t *e, int64_t e_ld, hipStream_t stream) {
    if (n_rows <= 0) return AX_OK;
    if (npad % KT != 0 || npad < n || dst_ld < npad || (dst_ld % 16) != 0) {
        set_error("slice_rows: bad padding (n=%lld npad=%lld ld=%lld)", (long long)n,
                  (long long)npad, (long long)dst_ld);
        return AX_ERR_VALUE;
    }
    const int64_t items = (int64_t)n_rows * (npad / KT);
    const int grid = static_cast<int>((items + 3) / 4 < 65535 ? (items + 3) / 4 : 65535);
    hipLaunchKernelGGL(slice_rows_kernel, dim3(grid), dim3(256), 0, stream, src, ld, rows, n_rows, n,
                       npad, dst, dst_ld, dst_plane, e, e_ld);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

ax_status rk_gemm_split_launch(const RkSeg *segs_dev, const RkWork *work_dev, int n_work,
                               int n_split, double *partial, hipStream_t stream) {
    hipLaunchKernelGGL(rk_gemm_split_kernel, dim3(n_work, n_split), dim3(256), 0, stream, segs_dev,
                       work_dev, n_work, n_split, partial);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

}  // namespace ax
