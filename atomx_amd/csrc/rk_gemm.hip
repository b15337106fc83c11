// R^k contraction R[i][j] = sum_x D[row_i][x] * Y[j][x] (numerics/bspline/
// slater.py:129-130 for a whole pair block), upper triangle only.
//
// Work decomposition: a list of segments (one per rank k, per request, ...),
// each an m x m symmetric-structured output; every upper 64x64 tile of every
// segment is one work entry; K is split into `n_split` contiguous chunks.
// Partial tiles are written to a workspace and summed over the splits in a
// fixed order by rk_reduce_kernel, so a result depends only on its own
// segment's data (never on what else is in the batch).
//
// FP64 path: v_mfma_f64_16x16x4_f64.  256 threads = 4 waves in a 2x2 layout,
// each wave owns a 32x32 sub-tile = 2x2 MFMA tiles; K is staged through LDS in
// chunks of 16 in k-major layout, double buffered (35 KB of LDS: four
// work-groups per CU).  A 16x16 block that lies wholly below the diagonal of
// its segment or past its last row/column is not stored, and a wave whose
// four blocks are all such blocks skips its MFMAs (the lower-left wave of a
// diagonal tile, waves past the segment edge).
#include "ax_common.h"
#include "rk_gemm.h"

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int BM = 64;
constexpr int BK = 16;
constexpr int LDS_LD = BM + 4;  // padded row (in doubles) of the k-major LDS tiles

constexpr int KPT = BK * BM / 256;  // k values per thread in one staged tile (4)

// Global-address-space view of device memory: loads through it are
// global_load (vector memory counter only), so waiting on LDS reads never
// waits on the prefetch of the next K stage.
typedef const double __attribute__((address_space(1))) gdouble;

// Row pointer of this thread's staged row (thread t -> row t/4); rows past
// the segment's last row read its row 0 and are zeroed when staged to LDS.
__device__ inline gdouble *stage_row(const double *__restrict__ base, int64_t ld,
                                     const int32_t *__restrict__ rows, int row0, int m,
                                     bool &valid) {
    const int gr = row0 + (threadIdx.x >> 2);
    valid = gr < m;
    const int r = valid ? gr : 0;
    const int64_t rr = rows ? (int64_t)rows[r] : (int64_t)r;
    return (gdouble *)(base + rr * ld);
}

// Loads this thread's KPT k values [k0 + kq, k0 + kq + KPT) of its row (zero
// past kend).  The full-tile test is uniform over the work-group.
__device__ inline void load_stage(gdouble *__restrict__ p, int64_t k0, int64_t kend,
                                  double (&v)[KPT]) {
    const int kq = (threadIdx.x & 3) * KPT;
    if (k0 + BK <= kend) {
#pragma unroll
        for (int q = 0; q < KPT; ++q) v[q] = p[k0 + kq + q];
    } else {
#pragma unroll
        for (int q = 0; q < KPT; ++q) {
            const int64_t kx = k0 + kq + q;
            v[q] = kx < kend ? p[kx] : 0.0;
        }
    }
}

// Stages the loaded values k-major into LDS (zeros for a row past the end).
__device__ inline void store_tile_lds(double *__restrict__ s, const double (&v)[KPT],
                                      bool valid) {
    const int t = threadIdx.x;
    const int r = t >> 2;
    const int kq = (t & 3) * KPT;
#pragma unroll
    for (int q = 0; q < KPT; ++q) s[(kq + q) * LDS_LD + r] = valid ? v[q] : 0.0;
}

__global__ __launch_bounds__(256) void rk_gemm_f64_kernel(const ax::BatchOf<ax::GemmArg> A) {
    __shared__ double As[2][BK * LDS_LD];
    __shared__ double Bs[2][BK * LDS_LD];

    // one 1-D grid of tiles x max_split blocks in the launch's block order
    int split;
    const int64_t tile = ax::gemm_block(A, split);
    int64_t e;
    const int p = ax::batch_find(A, tile, e);
    const ax::GemmArg &g = A.a[p];
    const int n_split = g.n_split;
    if (split >= n_split) return;
    const ax::RkWork wk = g.work[e];
    const ax::RkSeg sg = g.segs[wk.seg];
    const int64_t chunk = ((sg.n + n_split - 1) / n_split + BK - 1) / BK * BK;
    const int64_t kbeg = (int64_t)split * chunk;
    const int64_t kend = kbeg + chunk < sg.n ? kbeg + chunk : sg.n;
    const int row0 = wk.ti * BM, col0 = wk.tj * BM;

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    const int fr = lane & 15, fk = lane >> 4;

    // wave-uniform mask of the 16x16 blocks that hold upper-triangle
    // elements of the segment (i <= j < m); bit 2a+b for block (a, b)
    int bits = 0;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int i0 = row0 + wr + 16 * a, j0 = col0 + wc + 16 * b;
            if (i0 < sg.m && j0 < sg.m && i0 <= j0 + 15) bits |= 1 << (2 * a + b);
        }
    const int msk = __builtin_amdgcn_readfirstlane(bits);

    d4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};

    bool va_ok, vb_ok;
    gdouble *pa = stage_row(sg.D, sg.ld_d, sg.rows, row0, sg.m, va_ok);
    gdouble *pb = stage_row(sg.Y, sg.ld_y, nullptr, col0, sg.m, vb_ok);
    double va[KPT], vb[KPT];
    int buf = 0;
    if (kbeg < kend) {
        load_stage(pa, kbeg, kend, va);
        load_stage(pb, kbeg, kend, vb);
        store_tile_lds(As[0], va, va_ok);
        store_tile_lds(Bs[0], vb, vb_ok);
    }
    __syncthreads();
    for (int64_t k0 = kbeg; k0 < kend; k0 += BK) {
        const int64_t kn = k0 + BK;
        if (kn < kend) {
            load_stage(pa, kn, kend, va);
            load_stage(pb, kn, kend, vb);
        }
        if (msk) {
            const double *as = As[buf];
            const double *bs = Bs[buf];
            double a0[BK / 4], a1[BK / 4], b0[BK / 4], b1[BK / 4];
#pragma unroll
            for (int kk = 0; kk < BK / 4; ++kk) {
                const int kr = (kk * 4 + fk) * LDS_LD;
                a0[kk] = as[kr + wr + fr];
                a1[kk] = as[kr + wr + 16 + fr];
                b0[kk] = bs[kr + wc + fr];
                b1[kk] = bs[kr + wc + 16 + fr];
            }
#pragma unroll
            for (int kk = 0; kk < BK / 4; ++kk) {
                acc[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[kk], b0[kk], acc[0][0], 0, 0, 0);
                acc[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0[kk], b1[kk], acc[0][1], 0, 0, 0);
                acc[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[kk], b0[kk], acc[1][0], 0, 0, 0);
                acc[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1[kk], b1[kk], acc[1][1], 0, 0, 0);
            }
        }
        if (kn < kend) {
            store_tile_lds(As[buf ^ 1], va, va_ok);
            store_tile_lds(Bs[buf ^ 1], vb, vb_ok);
        }
        __syncthreads();
        buf ^= 1;
    }

    // C layout of the 16x16x4 f64 MFMA: result register r of lane l holds
    // row (l/16) + 4r, column l%16 (one-row groups, unlike the 4-row groups
    // of the i32/f32 16x16 MFMAs).  Skipped blocks are never read back.
    double *dst = g.partial + (e * n_split + split) * (BM * BM);
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            if (!(msk & (1 << (2 * a + b)))) continue;
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = wr + a * 16 + fk + 4 * r;
                const int j = wc + b * 16 + fr;
                dst[i * BM + j] = acc[a][b][r];
            }
        }
}

// Sums the split partials of each tile in split order and writes the packed
// upper triangle of the segment.
__global__ __launch_bounds__(256) void rk_reduce_kernel(const ax::BatchOf<ax::GemmArg> A) {
    int64_t e;
    const int p = ax::batch_find(A, blockIdx.x, e);
    const ax::GemmArg &g = A.a[p];
    const ax::RkWork wk = g.work[e];
    const ax::RkSeg sg = g.segs[wk.seg];
    const int n_split = g.n_split;
    const int64_t m = sg.m;
    const double *src = g.partial + e * n_split * (BM * BM);
    double *out = g.out + sg.out_off;
    const int per = (BM * BM) / gridDim.y;
    for (int t = blockIdx.y * per + threadIdx.x; t < (blockIdx.y + 1) * per; t += blockDim.x) {
        const int li = t / BM, lj = t % BM;
        const int64_t i = (int64_t)wk.ti * BM + li;
        const int64_t j = (int64_t)wk.tj * BM + lj;
        if (i >= m || j >= m || j < i) continue;
        double s = src[t];
        for (int sp = 1; sp < n_split; ++sp) s += src[(int64_t)sp * BM * BM + t];
        out[i * m - i * (i - 1) / 2 + (j - i)] = s;
    }
}

}  // namespace

namespace ax {

int gemm_tile() { return BM; }

int64_t rk_gemm_partial_doubles(int n_work, int n_split) {
    return (int64_t)n_work * n_split * BM * BM;
}

namespace {

// Fills the per-plan prefix of contraction tiles; returns the total tile
// count (the largest K split of the batch goes to A.max_split).
int64_t fill_work(BatchOf<GemmArg> &A, const PlanHost *const *plans, int n,
                  double *const *out) {
    A.n = n;
    A.max_split = 1;
    A.order = plans[0]->order;
    int64_t total = 0;
    for (int i = 0; i < n; ++i) {
        const PlanHost *h = plans[i];
        total += h->n_work;
        A.end[i] = total;
        A.a[i] = GemmArg{h->segs, h->work, h->partial, out ? out[i] : nullptr, h->n_split, 0};
        if (h->n_split > A.max_split) A.max_split = h->n_split;
    }
    return total;
}

}  // namespace

ax_status contract_batch_launch(const PlanHost *const *plans, int n, hipStream_t stream) {
    if (n <= 0) return AX_OK;
    if (n > kMaxBatch) {
        set_error("contract_batch_launch: %d plans exceed the batch limit %d", n, kMaxBatch);
        return AX_ERR_VALUE;
    }
    for (int i = 0; i < n; ++i) {
        const PlanHost *h = plans[i];
        const bool ok = h->precision == AX_PREC_FP64 ||
                        (h->precision == AX_PREC_SPLIT &&
                         (h->slices == kSplitSlices || h->slices == kSplitSlicesFine));
        if (!ok) {
            set_error("contract_batch_launch: plan %d has precision %d with %d slices", i,
                      h->precision, h->slices);
            return AX_ERR_VALUE;
        }
    }
    // one launch per contraction kind (FP64, split with 7 or 8 slices)
    const int kinds[3][2] = {{AX_PREC_FP64, 0},
                             {AX_PREC_SPLIT, kSplitSlices},
                             {AX_PREC_SPLIT, kSplitSlicesFine}};
    for (const auto &kd : kinds) {
        const PlanHost *sub[kMaxBatch];
        int ns = 0;
        for (int i = 0; i < n; ++i)
            if (plans[i]->precision == kd[0] && (kd[0] == AX_PREC_FP64 || plans[i]->slices == kd[1]))
                sub[ns++] = plans[i];
        if (ns == 0) continue;
        BatchOf<GemmArg> A{};
        const int64_t total = fill_work(A, sub, ns, nullptr);
        if (total == 0) continue;
        if (total * A.max_split > 0x7fffffff) {
            set_error("contract_batch_launch: %lld tiles exceed one launch", (long long)total);
            return AX_ERR_VALUE;
        }
        const dim3 grid(static_cast<unsigned>(total * A.max_split));
        if (kd[0] == AX_PREC_FP64) {
            hipLaunchKernelGGL(rk_gemm_f64_kernel, grid, dim3(256), 0, stream, A);
            AX_CHECK_LAUNCH();
        } else {
            ax_status st = rk_gemm_split_launch(A, kd[1], grid, stream);
            if (st != AX_OK) return st;
        }
    }
    return AX_OK;
}

ax_status reduce_batch_launch(const PlanHost *const *plans, int n, double *const *out,
                              hipStream_t stream) {
    if (n <= 0) return AX_OK;
    if (n > kMaxBatch) {
        set_error("reduce_batch_launch: %d plans exceed the batch limit %d", n, kMaxBatch);
        return AX_ERR_VALUE;
    }
    BatchOf<GemmArg> A{};
    const int64_t total = fill_work(A, plans, n, out);
    if (total == 0) return AX_OK;
    hipLaunchKernelGGL(rk_reduce_kernel, dim3(static_cast<unsigned>(total), 16), dim3(256), 0,
                       stream, A);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

void upper_tiles(int m, int seg, std::vector<RkWork> &out) {
    const int T = (m + BM - 1) / BM;
    for (int ti = 0; ti < T; ++ti)
        for (int tj = ti; tj < T; ++tj) out.push_back(RkWork{seg, ti, tj, 0});
}

}  // namespace ax
