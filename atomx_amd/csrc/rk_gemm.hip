// R^k contraction R[i][j] = sum_x D[row_i][x] * Y[j][x] (numerics/bspline/
// slater.py:129-130 for a whole pair block), upper triangle only.
//
// Work decomposition: a list of segments (one per rank k, per request, ...),
// each an m x m symmetric-structured output; every upper 64x64 tile of every
// segment is one work entry, and its K range is cut into contiguous pieces
// (one job = one piece of one tile).  A work-group runs a range of jobs: one
// job for the static and dynamic orders, a packed list of whole K-slice
// pieces for the balanced order (make_balanced_jobs).  A tile of one piece
// is written to the packed output by its work-group; the pieces of the
// others are written to a partial workspace and summed in K order by
// rk_reduce_kernel, so a result depends only on its own segment's data and
// the schedule (never on what else is in the batch).
// (Summing a tile's chunks in the contraction, by the work-group that
// finishes last, needs an agent-scope release per work-group — an L2
// write-back — and measured 30-70 us slower per launch than the separate
// reduction.)
//
// FP64 path: v_mfma_f64_16x16x4_f64, K split across the four waves of a
// work-group (below).
#include <algorithm>
#include <numeric>

#include "ax_common.h"
#include "rk_gemm.h"

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int BM = 64;
// Global-address-space view of device memory: loads through it are
// global_load (vector memory counter only), never flat loads that share a
// counter with LDS operations.
typedef const double __attribute__((address_space(1))) gdouble;

// ---------------------------------------------------------------------------
// FP64 contraction.  Every wave of the work-group computes the
// whole 64x64 tile (4x4 blocks of 16x16) over its own quarter of the block's
// K chunk, with operands loaded straight from global memory into registers
// (no LDS and no barrier in the K loop); the four partial tiles are summed
// through LDS at the end.  Blocks wholly below the diagonal or past the
// segment's last row or column are skipped at 16x16 granularity, uniformly
// for all four waves, so the waves of a work-group carry equal work.
//
// K order inside a stage of 8: MFMA step s (0, 1) takes, in lane group
// g = lane / 16, node k0 + 2g + s, so each lane reads two consecutive nodes
// of each of its 8 rows per stage (A and B use the same assignment, so the
// products pair up).
// ---------------------------------------------------------------------------
constexpr int KW = 8;         // nodes per register stage
constexpr int KW_ALIGN = 32;  // job K ranges start on multiples of 4 waves x KW nodes

struct KwOperands {
    double a[4][2], b[4][2];
};

__device__ inline void kw_load(gdouble *const (&pa)[4], gdouble *const (&pb)[4], int64_t k,
                               KwOperands &o) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        o.a[i][0] = pa[i][k];
        o.a[i][1] = pa[i][k + 1];
        o.b[i][0] = pb[i][k];
        o.b[i][1] = pb[i][k + 1];
    }
}

// Marks the loaded operands as used here (an empty asm that reads and writes
// them), after the MFMAs that overlap their latency.
__device__ inline void kw_pin(KwOperands &o) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int q = 0; q < 2; ++q) asm volatile("" : "+v"(o.a[i][q]), "+v"(o.b[i][q]));
}

// MODE 0: all 16 blocks; 1: blocks a <= b (diagonal tile); 2: blocks set in
// the wave-uniform mask `msk` (bit 4a + b).
template <int MODE>
__device__ inline void kw_mfma(const KwOperands &o, d4 (&acc)[4][4], int msk) {
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                if (MODE == 1 && a > b) continue;
                if (MODE == 2 && !(msk & (1 << (4 * a + b)))) continue;
                acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(o.a[a][s], o.b[b][s], acc[a][b],
                                                                 0, 0, 0);
            }
}

template <int MODE>
__device__ inline void kw_loop(gdouble *const (&pa)[4], gdouble *const (&pb)[4], int64_t kb,
                               int64_t ke, d4 (&acc)[4][4], int msk) {
    const int64_t g2 = (threadIdx.x & 63) >> 4 << 1;
    const int64_t nfull = ke > kb ? (ke - kb) / KW : 0;
    if (nfull > 0) {
        KwOperands cur, nxt;
        kw_load(pa, pb, kb + g2, cur);
        for (int64_t s = 0; s < nfull; ++s) {
            // prefetch the next stage (a valid stage of this range when none is left)
            const int64_t kn = s + 1 < nfull ? kb + (s + 1) * KW : kb;
            kw_load(pa, pb, kn + g2, nxt);
            kw_mfma<MODE>(cur, acc, msk);
            // keeps the prefetch in this iteration, behind all of its MFMAs:
            // without the pin the loads are sunk to their use in the next
            // iteration, without the barrier the wait is hoisted among the
            // MFMAs, and either way their latency is exposed
            __builtin_amdgcn_sched_barrier(0);
            kw_pin(nxt);
            cur = nxt;
        }
    }
    const int64_t kt = kb + nfull * KW;
    if (kt < ke) {  // uniform: a last partial stage, zero past ke
        KwOperands t;
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int64_t kx = kt + g2 + q;
            const int64_t kc = kx < ke ? kx : ke - 1;
            const bool in = kx < ke;
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const double va = pa[i][kc], vb = pb[i][kc];
                t.a[i][q] = in ? va : 0.0;
                t.b[i][q] = in ? vb : 0.0;
            }
        }
        kw_mfma<MODE>(t, acc, msk);
    }
}

// One job (a K chunk of a tile) of the FP64 contraction by one work-group.
__device__ inline void kw_job(const ax::BatchOf<ax::GemmArg> &A, const ax::JobInfo &J,
                              double (&park)[2][16 * 4 * 64]) {
    const ax::GemmArg &g = A.a[J.p];
    const ax::RkSeg sg = g.segs[J.seg];
    // the job's K range in quarters of whole 8-node stages, one per wave
    const int64_t kbeg = J.kb, kend = J.ke > J.kb ? J.ke : J.kb;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
    const int64_t wch = ((kend - kbeg + 3) / 4 + KW - 1) / KW * KW;
    const int64_t kb0 = kbeg + wave * wch;
    const int64_t kb = kb0 < kend ? kb0 : kend;
    const int64_t ke = kb0 + wch < kend ? kb0 + wch : kend;
    const int row0 = J.ti * BM, col0 = J.tj * BM;
    const int m = sg.m;

    // active 16x16 blocks: rows and columns inside the segment, a <= b on
    // the diagonal tile
    const int rb = min(4, (m - row0 + 15) >> 4), cb = min(4, (m - col0 + 15) >> 4);
    const bool diag = J.ti == J.tj;
    int bits = 0;
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b)
            if (a < rb && b < cb && (!diag || a <= b)) bits |= 1 << (4 * a + b);
    const int msk = __builtin_amdgcn_readfirstlane(bits);
    const int mode = __builtin_amdgcn_readfirstlane(
        rb == 4 && cb == 4 ? (diag ? 1 : 0) : 2);

    const int fr = lane & 15;
    gdouble *pa[4];
    gdouble *pb[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const int ra = min(row0 + 16 * i + fr, m - 1);
        const int64_t rr = sg.rows ? (int64_t)sg.rows[ra] : (int64_t)ra;
        pa[i] = (gdouble *)(sg.D + rr * sg.ld_d);
        const int cc = min(col0 + 16 * i + fr, m - 1);
        pb[i] = (gdouble *)(sg.Y + (int64_t)cc * sg.ld_y);
    }

    d4 acc[4][4];
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
        for (int b = 0; b < 4; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};
    if (mode == 0)
        kw_loop<0>(pa, pb, kb, ke, acc, msk);
    else if (mode == 1)
        kw_loop<1>(pa, pb, kb, ke, acc, msk);
    else
        kw_loop<2>(pa, pb, kb, ke, acc, msk);

    // sum the four waves' tiles in a fixed order: (w0 + w2) + (w1 + w3)
#define KW_FOR_ACTIVE(BODY)                                   \
    _Pragma("unroll") for (int a = 0; a < 4; ++a)             \
    _Pragma("unroll") for (int b = 0; b < 4; ++b)             \
    if (msk & (1 << (4 * a + b)))                             \
    _Pragma("unroll") for (int r = 0; r < 4; ++r) { BODY; }
    const int slot = lane;
    if (wave >= 2) {
        double *dst = park[wave - 2];
        KW_FOR_ACTIVE(dst[((4 * a + b) * 4 + r) * 64 + slot] = acc[a][b][r])
    }
    __syncthreads();
    if (wave < 2) {
        const double *src = park[wave];
        KW_FOR_ACTIVE(acc[a][b][r] += src[((4 * a + b) * 4 + r) * 64 + slot])
    }
    __syncthreads();
    if (wave == 1) {
        double *dst = park[0];
        KW_FOR_ACTIVE(dst[((4 * a + b) * 4 + r) * 64 + slot] = acc[a][b][r])
    }
    __syncthreads();
    if (wave == 0) {
        const double *src = park[0];
        const int fk = lane >> 4;
        if (J.count == 1) {  // a single chunk: the packed output directly
            double *out = g.out + sg.out_off;
            KW_FOR_ACTIVE(const int64_t i = row0 + 16 * a + fk + 4 * r;
                          const int64_t j = col0 + 16 * b + fr;
                          if (i < m && j < m && j >= i) out[ax::packed_index(i, j, m)] =
                              acc[a][b][r] + src[((4 * a + b) * 4 + r) * 64 + slot])
        } else {
            double *out = g.partial + (int64_t)J.slot * (BM * BM);
            KW_FOR_ACTIVE(out[(16 * a + fk + 4 * r) * BM + 16 * b + fr] =
                              acc[a][b][r] + src[((4 * a + b) * 4 + r) * 64 + slot])
        }
    }
#undef KW_FOR_ACTIVE
}

// The jobs of one work-group, one after the other.
__device__ inline void kw_group(const ax::BatchOf<ax::GemmArg> &A, const ax::GroupInfo &G,
                                double (&park)[2][16 * 4 * 64]) {
    for (int j = G.j0; j < G.j1; ++j) {
        if (j > G.j0) __syncthreads();  // the partial-sum area is reused
        kw_job(A, ax::gemm_job(A, G.p, j), park);
    }
}

__global__ __launch_bounds__(256, 2) void rk_gemm_f64_kernel(const ax::BatchOf<ax::GemmArg> A) {
    __shared__ double park[2][16 * 4 * 64];
    if (A.order != ax::kOrderDynamic) {
        kw_group(A, ax::gemm_group(A), park);
        return;
    }
    // dynamic order: every resident work-group takes the next job from the
    // ticket until none is left (the jobs are listed heaviest first); the
    // last work-group out resets the ticket and the exit count
    __shared__ int ticket;
    const int64_t total = A.end[A.n - 1];
    for (;;) {
        if (threadIdx.x == 0) ticket = atomicAdd(&A.counter[0], 1);
        __syncthreads();
        const int jb = ticket;
        __syncthreads();
        if (jb >= total) break;
        kw_group(A, ax::gemm_group(A, jb), park);
        __syncthreads();  // the partial-sum area is reused by the next job
    }
    if (threadIdx.x == 0 && atomicAdd(&A.counter[1], 1) == static_cast<int>(gridDim.x) - 1) {
        atomicExch(&A.counter[0], 0);
        atomicExch(&A.counter[1], 0);
    }
}

// Sums the split partials of each tile in split order and writes the packed
// upper triangle of the segment (tiles of one chunk were written by the
// contraction).
__global__ __launch_bounds__(256) void rk_reduce_kernel(const ax::BatchOf<ax::GemmArg> A) {
    int64_t e;
    const int p = ax::batch_find(A, blockIdx.x, e);
    const ax::GemmArg &g = A.a[p];
    const ax::RkWork wk = g.work[e];
    const ax::RkSeg sg = g.segs[wk.seg];
    const ax::RkTileJobs tj = g.tile_jobs[e];
    if (tj.count == 1) return;
    const int64_t m = sg.m;
    const double *src = g.partial + (int64_t)tj.first * (BM * BM);
    double *out = g.out + sg.out_off;
    const int per = (BM * BM) / gridDim.y;
    for (int t = blockIdx.y * per + threadIdx.x; t < (blockIdx.y + 1) * per; t += blockDim.x) {
        const int li = t / BM, lj = t % BM;
        const int64_t i = (int64_t)wk.ti * BM + li;
        const int64_t j = (int64_t)wk.tj * BM + lj;
        if (i >= m || j >= m || j < i) continue;
        double s = src[t];
        for (int sp = 1; sp < tj.count; ++sp) s += src[(int64_t)sp * BM * BM + t];
        out[i * m - i * (i - 1) / 2 + (j - i)] = s;
    }
}

}  // namespace

namespace ax {

int gemm_tile() { return BM; }

// Resident work-groups of a dynamic-order FP64 contraction: two per CU on the
// 256 CUs of an MI355X (64 KB of LDS and 232 VGPRs each).
constexpr int64_t kDynamicGroups = 512;

int64_t rk_gemm_partial_doubles(int n_work, int n_split) {
    return (int64_t)n_work * n_split * BM * BM;
}

namespace {

// Fills the per-plan prefix of contraction work-groups (or, for the
// reduction, of tiles); returns the total.
int64_t fill_work(BatchOf<GemmArg> &A, const PlanHost *const *plans, int n,
                  double *const *out, bool jobs) {
    A.n = n;
    A.order = plans[0]->order;
    int64_t total = 0;
    for (int i = 0; i < n; ++i) {
        const PlanHost *h = plans[i];
        total += jobs ? h->n_groups : h->n_work;
        A.end[i] = total;
        A.a[i] = GemmArg{h->segs, h->work, h->jobs, h->groups, h->tile_jobs, h->partial,
                         out ? out[i] : nullptr};
    }
    return total;
}

}  // namespace

ax_status contract_batch_launch(const PlanHost *const *plans, int n, double *const *out,
                                hipStream_t stream) {
    if (n <= 0) return AX_OK;
    if (n > kMaxBatch) {
        set_error("contract_batch_launch: %d plans exceed the batch limit %d", n, kMaxBatch);
        return AX_ERR_VALUE;
    }
    for (int i = 0; i < n; ++i) {
        const PlanHost *h = plans[i];
        if (!(out && out[i])) {
            set_error("contract_batch_launch: null output for plan %d", i);
            return AX_ERR_VALUE;
        }
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
        double *sub_out[kMaxBatch];
        int ns = 0;
        for (int i = 0; i < n; ++i)
            if (plans[i]->precision == kd[0] &&
                (kd[0] == AX_PREC_FP64 || plans[i]->slices == kd[1])) {
                sub_out[ns] = out[i];
                sub[ns++] = plans[i];
            }
        if (ns == 0) continue;
        BatchOf<GemmArg> A{};
        const int64_t total = fill_work(A, sub, ns, sub_out, true);
        if (total == 0) continue;
        if (total > 0x7fffffff) {
            set_error("contract_batch_launch: %lld work-groups exceed one launch",
                      (long long)total);
            return AX_ERR_VALUE;
        }
        dim3 grid(static_cast<unsigned>(total));
        if (kd[0] == AX_PREC_FP64) {
            if (A.order == kOrderDynamic) {
                // as many work-groups as fit at once (two per CU), each taking jobs
                A.counter = sub[0]->counter;
                grid = dim3(static_cast<unsigned>(std::min<int64_t>(total, kDynamicGroups)));
            }
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
    // plans contracted in one K piece per tile have nothing to reduce
    const PlanHost *sub[kMaxBatch];
    double *sub_out[kMaxBatch];
    int ns = 0;
    for (int i = 0; i < n; ++i)
        if (plans[i]->needs_reduce) {
            sub_out[ns] = out[i];
            sub[ns++] = plans[i];
        }
    if (ns == 0) return AX_OK;
    BatchOf<GemmArg> A{};
    const int64_t total = fill_work(A, sub, ns, sub_out, false);
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

namespace {

// 16x16 blocks of a tile inside the triangle of its segment (the FP64
// kernel's MFMA work per K unit).
int tile_blocks(const RkWork &w, int m) {
    const int rb = std::min(4, (m - w.ti * BM + 15) / 16);
    const int cb = std::min(4, (m - w.tj * BM + 15) / 16);
    int n = 0;
    for (int a = 0; a < rb; ++a)
        for (int b = 0; b < cb; ++b) n += (w.ti != w.tj || a <= b) ? 1 : 0;
    return n;
}

RkJob make_job(const RkWork &w, int split, int count, int slot, int64_t kb, int64_t ke) {
    return RkJob{w.seg, w.ti | (w.tj << 16), split | (count << 16), slot, static_cast<int>(kb),
                 static_cast<int>(ke)};
}

}  // namespace

KSpace k_space(int64_t n, bool fp64) {
    return fp64 ? KSpace{n, KW_ALIGN, true}
                : KSpace{(n + kSplitKTile - 1) / kSplitKTile, 1, false};
}

JobTables make_jobs(const std::vector<RkWork> &work, const std::vector<int> &seg_m, int S,
                    const KSpace &ks, bool all) {
    JobTables t;
    const int nw = static_cast<int>(work.size());
    // every tile in S chunks of equal length (weighting the chunk count by
    // the tile's blocks inside the triangle was measured no faster)
    const int64_t per = ((ks.k_len + S - 1) / S + ks.k_align - 1) / ks.k_align * ks.k_align;
    t.tiles.resize(nw);
    for (int e = 0; e < nw; ++e) t.tiles[e] = RkTileJobs{e * S, S};
    auto job = [&](int e, int sp) {
        const int64_t kb = std::min<int64_t>(sp * per, ks.k_len);
        const int64_t ke = std::min<int64_t>(kb + per, ks.k_len);
        return make_job(work[e], sp, S, e * S + sp, kb, ke);
    };
    t.tile_fastest.reserve((size_t)nw * S);
    if (all) t.split_fastest.reserve((size_t)nw * S);
    for (int sp = 0; sp < S; ++sp)
        for (int e = 0; e < nw; ++e) t.tile_fastest.push_back(job(e, sp));
    if (!all) return t;
    for (int e = 0; e < nw; ++e)
        for (int sp = 0; sp < S; ++sp) t.split_fastest.push_back(job(e, sp));
    // heaviest first: tiles by their 16x16 blocks inside the triangle (the
    // MFMA work of each of their jobs), ties in tile-fastest order
    std::vector<int> blocks(nw);
    for (int e = 0; e < nw; ++e) blocks[e] = tile_blocks(work[e], seg_m[work[e].seg]);
    std::vector<int> ord(nw);
    for (int e = 0; e < nw; ++e) ord[e] = e;
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) { return blocks[x] > blocks[y]; });
    t.heaviest.reserve((size_t)nw * S);
    for (int e : ord)
        for (int sp = 0; sp < S; ++sp) t.heaviest.push_back(job(e, sp));
    return t;
}

int balanced_groups(int S) { return 64 * S; }

BalancedJobs make_balanced_jobs(const std::vector<RkWork> &work, const std::vector<int> &seg_m,
                                int n_groups, const KSpace &ks) {
    BalancedJobs b;
    const int nw = static_cast<int>(work.size());
    const int64_t U = (ks.k_len + ks.k_align - 1) / ks.k_align;  // K units per tile
    if (nw == 0 || U == 0) {
        b.groups.assign(1, 0);
        b.tiles.assign(nw, RkTileJobs{0, 0});
        return b;
    }
    std::vector<int64_t> wt(nw);  // cost of one K unit of each tile
    for (int e = 0; e < nw; ++e) wt[e] = ks.fp64 ? tile_blocks(work[e], seg_m[work[e].seg]) : 16;
    // K slices, one per XCD (each XCD reads its own K slice of every operand
    // row)
    const int X = static_cast<int>(std::min<int64_t>(8, U));
    const int gx = std::max(1, n_groups / X);
    struct Piece {
        int e;
        int64_t u0, u1;  // K units of tile e
    };
    std::vector<Piece> pieces;
    b.groups.push_back(0);
    // per slice: whole tile-slices, heaviest first, each to the least
    // loaded work-group of the slice (ties to the lowest index)
    std::vector<int> ord(nw);
    for (int e = 0; e < nw; ++e) ord[e] = e;
    std::stable_sort(ord.begin(), ord.end(), [&](int p, int q) { return wt[p] > wt[q]; });
    for (int x = 0; x < X; ++x) {
        const int64_t s0 = U * x / X, s1 = U * (x + 1) / X;
        const int G = std::min(gx, nw);
        std::vector<int64_t> load(G, 0);
        std::vector<std::vector<int>> bins(G);
        for (int e : ord) {
            const int g = static_cast<int>(std::min_element(load.begin(), load.end()) - load.begin());
            load[g] += wt[e];
            bins[g].push_back(e);
        }
        for (int g = 0; g < gx; ++g) {
            if (g < G)
                for (int e : bins[g]) pieces.push_back(Piece{e, s0, s1});
            b.groups.push_back(static_cast<int>(pieces.size()));
        }
    }
    // slots tile-major, a tile's pieces in K order (one per slice)
    std::vector<int> count(nw, 0), first(nw, 0), seen(nw, 0);
    for (const Piece &pc : pieces) ++count[pc.e];
    for (int t = 1; t < nw; ++t) first[t] = first[t - 1] + count[t - 1];
    b.tiles.resize(nw);
    for (int t = 0; t < nw; ++t) b.tiles[t] = RkTileJobs{first[t], count[t]};
    b.jobs.reserve(pieces.size());
    for (const Piece &pc : pieces) {
        const int sp = seen[pc.e]++;
        const int64_t kb = std::min<int64_t>(pc.u0 * ks.k_align, ks.k_len);
        const int64_t ke = std::min<int64_t>(pc.u1 * ks.k_align, ks.k_len);
        b.jobs.push_back(make_job(work[pc.e], sp, count[pc.e], first[pc.e] + sp, kb, ke));
    }
    return b;
}

}  // namespace ax
