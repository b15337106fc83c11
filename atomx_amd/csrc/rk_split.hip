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

constexpr int KT = ax::kSplitKTile;
constexpr int BM = 64;

__global__ __launch_bounds__(256) void slice_rows_kernel(const double *__restrict__ src,
                                                         int64_t ld,
                                                         const int32_t *__restrict__ rows,
                                                         int n_rows, int64_t n, int64_t npad,
                                                         int8_t *__restrict__ dst, int64_t dst_ld,
                                                         int64_t dst_plane,
                                                         int32_t *__restrict__ e, int64_t e_ld) {
    const int64_t n_tiles = npad / KT;
    const int64_t total = (int64_t)n_rows * n_tiles;
    for (int64_t item = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); item < total;
         item += (int64_t)gridDim.x * 4) {
        const int r = static_cast<int>(item / n_tiles);
        const int64_t t = item - (int64_t)r * n_tiles;
        const int64_t rr = rows ? (int64_t)rows[r] : (int64_t)r;
        ax::slice_tile(src + rr * ld, n, t, dst + (int64_t)r * dst_ld, dst_plane,
                       e + (int64_t)r * e_ld + t, ax::kSplitSlices);
    }
}

// Batched slicing of the rows that no producing kernel sliced (Y^k rows of
// grids whose scan does not slice in place).  Four waves per block, one
// 64-wide K-tile per wave; the plan lookup is per wave.
__global__ __launch_bounds__(256) void slice_batch_kernel(const ax::BatchOf<ax::SliceArg> A) {
    const int64_t total = A.end[A.n - 1];
    for (int64_t item = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6); item < total;
         item += (int64_t)gridDim.x * 4) {
        int64_t local;
        const int p = ax::batch_find(A, item, local);
        const ax::SliceArg &g = A.a[p];
        const int64_t n_tiles = g.npad / KT;
        const int64_t r = local / n_tiles;
        const int64_t t = local - r * n_tiles;
        ax::slice_tile(g.src + r * g.n, g.n, t, g.dst + r * g.npad, g.plane, g.e + r * n_tiles + t,
                       g.slices);
    }
}

// Global-address-space loads: they count on the vector memory counter only,
// so waiting on LDS reads never waits on the prefetch of the next K-tile.
typedef const v4i __attribute__((address_space(1))) gv4i;
typedef const int32_t __attribute__((address_space(1))) gint;
__device__ inline v4i load16(const int8_t *p) { return *(gv4i *)p; }

// LDS tile of one operand for one K-tile: [S][64 rows][ROWB bytes] int8
// (64 data bytes + 16 pad per row) and the 64 row exponents.
constexpr int ROWB = KT + 16;
constexpr int PLANEB = BM * ROWB;

// Global -> register staging: thread t moves S chunks of 16 bytes, chunk
// (s, row = t/4, q = t%4) of the K-tile.
// Base of this thread's staged row (thread t -> row t/4, 16-byte chunk t%4);
// a row past the segment's last row reads row 0 and is zeroed when staged.
__device__ inline const int8_t *stage_base(const int8_t *__restrict__ base, int64_t ld,
                                           const int32_t *__restrict__ rows, int row0, int m,
                                           bool &valid) {
    const int r = threadIdx.x >> 2, q = threadIdx.x & 3;
    const int gr = row0 + r;
    valid = gr < m;
    const int g = valid ? gr : 0;
    const int64_t rr = rows ? (int64_t)rows[g] : (int64_t)g;
    return base + rr * ld + q * 16;
}

template <int S>
__device__ inline void stage_load(const int8_t *__restrict__ p, int64_t plane, int64_t t,
                                  v4i (&reg)[S]) {
#pragma unroll
    for (int s2 = 0; s2 < S; ++s2) reg[s2] = load16(p + t * KT + s2 * plane);
}

template <int S>
__device__ inline void stage_store(int8_t *__restrict__ lds, const v4i (&reg)[S], bool valid) {
    const int r = threadIdx.x >> 2, q = threadIdx.x & 3;
    const v4i zero = {0, 0, 0, 0};
#pragma unroll
    for (int s2 = 0; s2 < S; ++s2)
        *reinterpret_cast<v4i *>(lds + s2 * PLANEB + r * ROWB + q * 16) = valid ? reg[s2] : zero;
}

// One job (a K-tile range of a tile) of the split contraction.
template <int S>
__device__ inline void split_job(const ax::BatchOf<ax::GemmArg> &A, const ax::JobInfo &J,
                                 int8_t *__restrict__ As, int8_t *__restrict__ Bs, int *eAs,
                                 int *eBs) {
    const ax::GemmArg &g = A.a[J.p];
    const ax::RkSeg sg = g.segs[J.seg];
    const int64_t tb = J.kb, te = J.ke;  // K-tiles [tb, te)
    const int row0 = J.ti * BM, col0 = J.tj * BM;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    const int fr = lane & 15, fg = lane >> 4;

    // exponent row of output-row / output-column slot `tid` (tid < 64 / < 128)
    int64_t erow = 0;
    bool eok = false;
    if (tid < BM) {
        const int i = row0 + tid;
        eok = i < sg.m;
        erow = eok ? (sg.rows ? (int64_t)sg.rows[i] : (int64_t)i) : 0;
    } else if (tid < 2 * BM) {
        const int j = col0 + tid - BM;
        eok = j < sg.m;
        erow = eok ? (int64_t)j : 0;
    }

    d4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};

    // K-tile 0 (nodes 0..63) holds the cell that reaches r = 0, where the
    // powers of r inside one tile span more binary orders than the slices
    // carry: it is contracted in FP64, from the FP64 densities and the FP64
    // first tile of the Y^k rows, through the LDS region of the slices.
    int64_t t0 = tb;
    if (tb == 0 && te > 0 && sg.Y0 != nullptr) {
        constexpr int LD = KT + 1;  // padded row (LDS banks)
        double *Dt = reinterpret_cast<double *>(As);
        double *Yt = Dt + BM * LD;
        const int64_t nk = sg.n < KT ? sg.n : KT;
        for (int e = tid; e < BM * KT; e += 256) {
            const int r = e / KT, x = e - (e / KT) * KT;
            const int i = row0 + r, j = col0 + r;
            double dvl = 0.0, yvl = 0.0;
            if (x < nk && i < sg.m) {
                const int64_t dr = sg.d_rows ? (int64_t)sg.d_rows[i] : (int64_t)i;
                dvl = sg.D[dr * sg.ld_d + x];
            }
            if (x < nk && j < sg.m) yvl = sg.Y0[(int64_t)j * sg.y0_ld + x];
            Dt[r * LD + x] = dvl;
            Yt[r * LD + x] = yvl;
        }
        __syncthreads();
        for (int x = 0; x < KT; ++x) {
            double dv[2][4], yv[2];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int r = 0; r < 4; ++r) dv[a][r] = Dt[(wr + a * 16 + 4 * fg + r) * LD + x];
#pragma unroll
            for (int b = 0; b < 2; ++b) yv[b] = Yt[(wc + b * 16 + fr) * LD + x];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
#pragma unroll
                    for (int r = 0; r < 4; ++r) acc[a][b][r] = fma(dv[a][r], yv[b], acc[a][b][r]);
        }
        t0 = 1;  // the slice loop's first barrier frees the LDS again
    }

    bool va_ok, vb_ok;
    const int8_t *pa = stage_base(sg.Ds, sg.ds_ld, sg.rows, row0, sg.m, va_ok);
    const int8_t *pb = stage_base(sg.Ys, sg.ys_ld, nullptr, col0, sg.m, vb_ok);
    // exponent row of this thread's slot (rows for tid < 64, columns for < 128)
    gint *ep = tid < BM ? (gint *)(sg.De + erow * sg.de_ld) : (gint *)(sg.Ye + erow * sg.ye_ld);
    const bool has_e = tid < 2 * BM;
    v4i ra[S], rb[S];
    int rexp = 0;
    if (t0 < te) {
        stage_load<S>(pa, sg.ds_plane, t0, ra);
        stage_load<S>(pb, sg.ys_plane, t0, rb);
        if (has_e) rexp = ep[t0];
    }
    const v4i zero = {0, 0, 0, 0};
    for (int64_t t = t0; t < te; ++t) {
        __syncthreads();  // previous tile's LDS reads are complete
        stage_store<S>(As, ra, va_ok);
        stage_store<S>(Bs, rb, vb_ok);
        if (tid < BM) eAs[tid] = eok ? rexp : 0;
        else if (tid < 2 * BM) eBs[tid - BM] = eok ? rexp : 0;
        __syncthreads();
        if (t + 1 < te) {  // prefetch the next K-tile while this one is multiplied
            stage_load<S>(pa, sg.ds_plane, t + 1, ra);
            stage_load<S>(pb, sg.ys_plane, t + 1, rb);
            if (has_e) rexp = ep[t + 1];
        }
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            v4i av[S];
#pragma unroll
            for (int s2 = 0; s2 < S; ++s2)
                av[s2] = *reinterpret_cast<const v4i *>(As + s2 * PLANEB + (wr + a * 16 + fr) * ROWB + fg * 16);
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                v4i bv[S];
#pragma unroll
                for (int s2 = 0; s2 < S; ++s2)
                    bv[s2] = *reinterpret_cast<const v4i *>(Bs + s2 * PLANEB + (wc + b * 16 + fr) * ROWB + fg * 16);
                double lvl[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
                for (int L = S + 1; L >= 2; --L) {
                    v4i c = zero;
#pragma unroll
                    for (int s1 = 1; s1 <= S; ++s1) {
                        const int s2 = L - s1;
                        if (s2 < 1 || s2 > S) continue;
                        c = __builtin_amdgcn_mfma_i32_16x16x64_i8(av[s1 - 1], bv[s2 - 1], c, 0, 0, 0);
                    }
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        lvl[r] = (L == S + 1) ? static_cast<double>(c[r])
                                              : fma(lvl[r], 0x1p-7, static_cast<double>(c[r]));
                }
                const int eb = eBs[wc + b * 16 + fr];
#pragma unroll
                for (int r = 0; r < 4; ++r)
                    acc[a][b][r] += ldexp(lvl[r], eAs[wr + a * 16 + 4 * fg + r] + eb - 14);
            }
        }
    }

    if (J.count == 1) {  // a single chunk: the packed output directly
        double *out = g.out + sg.out_off;
        const int64_t m = sg.m;
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int64_t i = row0 + wr + a * 16 + 4 * fg + r;
                    const int64_t j = col0 + wc + b * 16 + fr;
                    if (i < m && j < m && j >= i) out[ax::packed_index(i, j, m)] = acc[a][b][r];
                }
        return;
    }
    double *dst = g.partial + (int64_t)J.slot * (BM * BM);
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

template <int S>
__global__ __launch_bounds__(256) void rk_gemm_split_kernel(const ax::BatchOf<ax::GemmArg> A) {
    extern __shared__ __attribute__((aligned(16))) int8_t lds_split[];
    int8_t *As = lds_split;                // [S * PLANEB]
    int8_t *Bs = lds_split + S * PLANEB;   // [S * PLANEB]
    __shared__ int eAs[BM], eBs[BM];
    // this work-group's jobs, one after the other (every K-tile step starts
    // with a barrier, so the LDS tiles are free when the next job stages)
    const ax::GroupInfo G = ax::gemm_group(A);
    for (int j = G.j0; j < G.j1; ++j) split_job<S>(A, ax::gemm_job(A, G.p, j), As, Bs, eAs, eBs);
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

template <int S>
ax_status launch_split(const BatchOf<GemmArg> &A, dim3 grid, hipStream_t stream) {
    // the slice tiles, or the two FP64 tiles of K-tile 0 (64 x 65 doubles each)
    const size_t tile0 = (size_t)2 * BM * (KT + 1) * sizeof(double);
    const size_t slices = (size_t)2 * S * PLANEB;
    const size_t lds = slices > tile0 ? slices : tile0;
    if (lds > 64 * 1024)
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(rk_gemm_split_kernel<S>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    hipLaunchKernelGGL(rk_gemm_split_kernel<S>, grid, dim3(256), lds, stream, A);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

ax_status rk_gemm_split_launch(const BatchOf<GemmArg> &A, int slices, dim3 grid,
                               hipStream_t stream) {
    if (slices == kSplitSlices) return launch_split<kSplitSlices>(A, grid, stream);
    if (slices == kSplitSlicesFine) return launch_split<kSplitSlicesFine>(A, grid, stream);
    set_error("split contraction: unsupported slice count %d", slices);
    return AX_ERR_VALUE;
}

ax_status slice_batch_launch(const PlanHost *const *plans, int n, hipStream_t stream) {
    if (n <= 0) return AX_OK;
    if (n > kMaxBatch) {
        set_error("slice_batch_launch: %d plans exceed the batch limit %d", n, kMaxBatch);
        return AX_ERR_VALUE;
    }
    BatchOf<SliceArg> A{};
    A.n = n;
    int64_t total = 0;
    for (int i = 0; i < n; ++i) {
        const PlanHost *h = plans[i];
        const bool todo = h->precision == AX_PREC_SPLIT && h->Y != nullptr;
        if (todo) total += (int64_t)h->n_yrows * (h->npad / KT);
        A.end[i] = total;
        const int64_t yr = h->n_yrows > 0 ? h->n_yrows : 1;
        A.a[i] = SliceArg{h->Y, h->n, h->npad, yr * h->npad, h->Ys, h->Ye, h->slices, 0};
    }
    if (total == 0) return AX_OK;
    const int grid = static_cast<int>((total + 3) / 4 < 65535 ? (total + 3) / 4 : 65535);
    hipLaunchKernelGGL(slice_batch_kernel, dim3(grid), dim3(256), 0, stream, A);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

}  // namespace ax
