// Batched R^k tensor over many orbitals on one grid (the all-quartet
// Slater-integral workload), and the single-block ax_rk_gemm entry point.
//
// Pipeline of one run (all on one stream, capturable in a HIP graph):
//   1. pair_density   D[p] = w * (P_a P_b + Q_a Q_b)  for all pairs a <= b
//   2. yk_batch       Y^k[slot(p,k)] for every pair valid at rank k
//   3. (split) slice  int8 planes + exponents of D and of every Y^k row
//   4. rk_gemm        packed upper triangle of D[V_k] Y_k^T for every k,
//                     all ranks in one launch (one segment per rank)
#include <algorithm>
#include <cstring>
#include <vector>

#include "ax_common.h"
#include "rk_gemm.h"

struct ax_rk_tensor {
    int nv = 0, ng = 0, n_orb = 0, kmax = 0, precision = 0;
    int64_t n = 0, npad = 0;
    int n_pairs = 0, n_yrows = 0, n_items = 0, n_work = 0, n_split = 1;
    std::vector<int> m_k;
    std::vector<int64_t> out_off;
    int64_t out_size = 0;
    int64_t footprint = 0;
    // device buffers
    double *r = nullptr, *w = nullptr, *rk = nullptr, *rkm = nullptr, *mo = nullptr, *mi = nullptr;
    int32_t *pa = nullptr, *pb = nullptr, *rows_all = nullptr;
    int32_t *item_row = nullptr, *item_k = nullptr, *item_out = nullptr;
    double *D = nullptr, *Y = nullptr, *partial = nullptr;
    ax::RkSeg *segs = nullptr;
    ax::RkWork *work = nullptr;
    int8_t *Ds = nullptr, *Ys = nullptr;
    int32_t *De = nullptr, *Ye = nullptr;
    std::vector<void *> allocs;
};

namespace {

template <typename T>
ax_status dev_alloc(ax_rk_tensor *p, T **ptr, size_t count) {
    void *q = nullptr;
    if (count == 0) count = 1;
    hipError_t e = hipMalloc(&q, count * sizeof(T));
    if (e != hipSuccess) {
        ax::set_error("hipMalloc of %zu bytes failed: %s", count * sizeof(T), hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? AX_ERR_MEMORY : AX_ERR_HIP;
    }
    p->allocs.push_back(q);
    p->footprint += static_cast<int64_t>(count * sizeof(T));
    *ptr = static_cast<T *>(q);
    return AX_OK;
}

template <typename T>
ax_status upload(T *dst, const T *src, size_t count) {
    if (count == 0) return AX_OK;
    AX_CHECK_HIP(hipMemcpy(dst, src, count * sizeof(T), hipMemcpyHostToDevice));
    return AX_OK;
}

void free_plan(ax_rk_tensor *p) {
    for (void *q : p->allocs) (void)hipFree(q);
    p->allocs.clear();
}

#define AX_TRY(expr)                   \
    do {                               \
        ax_status st_ = (expr);        \
        if (st_ != AX_OK) {            \
            free_plan(plan);           \
            delete plan;               \
            return st_;                \
        }                              \
    } while (0)

int choose_split(int n_work, int64_t n) {
    // enough work-groups for 256 CUs (about four per CU), K chunks >= 256
    int s = 1;
    while ((int64_t)n_work * s < 1024 && n / (2 * s) >= 256 && s < 16) s *= 2;
    return s;
}

}  // namespace

extern "C" {

ax_status ax_rk_tensor_create(const double *r_host, const double *w_host, int nv, int ng,
                              int n_orb, const int32_t *l_host, int kmax, int precision,
                              ax_rk_tensor **out_plan) {
    AX_REQUIRE(out_plan != nullptr, "ax_rk_tensor_create: null plan pointer");
    AX_REQUIRE(nv >= 1 && ng >= 1 && ng <= 16, "ax_rk_tensor_create: bad grid nv=%d ng=%d", nv, ng);
    AX_REQUIRE(n_orb >= 1 && kmax >= 0 && kmax <= 64, "ax_rk_tensor_create: bad n_orb/kmax");
    AX_REQUIRE(precision == AX_PREC_FP64 || precision == AX_PREC_SPLIT,
               "ax_rk_tensor_create: unknown precision %d", precision);
    for (int a = 0; a < n_orb; ++a)
        AX_REQUIRE(l_host[a] >= 0, "ax_rk_tensor_create: negative l for orbital %d", a);
    ax_rk_tensor *plan = new ax_rk_tensor();
    plan->nv = nv;
    plan->ng = ng;
    plan->n_orb = n_orb;
    plan->kmax = kmax;
    plan->precision = precision;
    plan->n = (int64_t)nv * ng;
    plan->npad = (plan->n + ax::kSplitKTile - 1) / ax::kSplitKTile * ax::kSplitKTile;
    const int64_t n = plan->n;

    // pairs a <= b in lexicographic order
    std::vector<int32_t> pa, pb;
    for (int a = 0; a < n_orb; ++a)
        for (int b = a; b < n_orb; ++b) {
            pa.push_back(a);
            pb.push_back(b);
        }
    plan->n_pairs = static_cast<int>(pa.size());

    // valid pairs per rank, Y^k slots, scan items
    std::vector<int32_t> rows_all, it_row, it_k, it_out;
    std::vector<int64_t> row_off(kmax + 1), y_off(kmax + 1);
    plan->m_k.assign(kmax + 1, 0);
    plan->out_off.assign(kmax + 1, 0);
    int64_t yrows = 0, outsz = 0;
    for (int k = 0; k <= kmax; ++k) {
        row_off[k] = static_cast<int64_t>(rows_all.size());
        y_off[k] = yrows;
        int m = 0;
        for (int p = 0; p < plan->n_pairs; ++p) {
            const int la = l_host[pa[p]], lb = l_host[pb[p]];
            const int lo = la > lb ? la - lb : lb - la;
            if (k < lo || k > la + lb || ((la + lb + k) & 1)) continue;
            rows_all.push_back(p);
            it_row.push_back(p);
            it_k.push_back(k);
            it_out.push_back(static_cast<int32_t>(yrows + m));
            ++m;
        }
        plan->m_k[k] = m;
        plan->out_off[k] = outsz;
        outsz += (int64_t)m * (m + 1) / 2;
        yrows += m;
    }
    plan->out_size = outsz;
    plan->n_yrows = static_cast<int>(yrows);
    plan->n_items = static_cast<int>(it_row.size());

    // grid tables
    AX_TRY(dev_alloc(plan, &plan->r, n));
    AX_TRY(dev_alloc(plan, &plan->w, n));
    AX_TRY(dev_alloc(plan, &plan->rk, n * (kmax + 1)));
    AX_TRY(dev_alloc(plan, &plan->rkm, n * (kmax + 1)));
    AX_TRY(dev_alloc(plan, &plan->mo, ng * ng));
    AX_TRY(dev_alloc(plan, &plan->mi, ng * ng));
    AX_TRY(upload(plan->r, r_host, n));
    AX_TRY(upload(plan->w, w_host, n));
    {
        std::vector<double> mo(ng * ng), mi(ng * ng);
        AX_TRY(ax_yk_partial_matrices(ng, mo.data(), mi.data()));
        AX_TRY(upload(plan->mo, mo.data(), mo.size()));
        AX_TRY(upload(plan->mi, mi.data(), mi.size()));
    }
    AX_TRY(ax_grid_rpow(plan->r, n, kmax, plan->rk, plan->rkm, nullptr));

    AX_TRY(dev_alloc(plan, &plan->pa, pa.size()));
    AX_TRY(dev_alloc(plan, &plan->pb, pb.size()));
    AX_TRY(upload(plan->pa, pa.data(), pa.size()));
    AX_TRY(upload(plan->pb, pb.data(), pb.size()));
    AX_TRY(dev_alloc(plan, &plan->rows_all, rows_all.size()));
    AX_TRY(upload(plan->rows_all, rows_all.data(), rows_all.size()));
    AX_TRY(dev_alloc(plan, &plan->item_row, it_row.size()));
    AX_TRY(dev_alloc(plan, &plan->item_k, it_k.size()));
    AX_TRY(dev_alloc(plan, &plan->item_out, it_out.size()));
    AX_TRY(upload(plan->item_row, it_row.data(), it_row.size()));
    AX_TRY(upload(plan->item_k, it_k.data(), it_k.size()));
    AX_TRY(upload(plan->item_out, it_out.data(), it_out.size()));
    AX_TRY(dev_alloc(plan, &plan->D, (size_t)plan->n_pairs * n));
    AX_TRY(dev_alloc(plan, &plan->Y, (size_t)(yrows > 0 ? yrows : 1) * n));

    if (precision == AX_PREC_SPLIT) {
        const int64_t tiles = plan->npad / ax::kSplitKTile;
        AX_TRY(dev_alloc(plan, &plan->Ds, (size_t)ax::kSplitSlices * plan->n_pairs * plan->npad));
        AX_TRY(dev_alloc(plan, &plan->Ys, (size_t)ax::kSplitSlices * (yrows > 0 ? yrows : 1) * plan->npad));
        AX_TRY(dev_alloc(plan, &plan->De, (size_t)plan->n_pairs * tiles));
        AX_TRY(dev_alloc(plan, &plan->Ye, (size_t)(yrows > 0 ? yrows : 1) * tiles));
    }

    // contraction segments (one per rank) and the tile work list
    std::vector<ax::RkSeg> segs;
    std::vector<ax::RkWork> work;
    const int64_t tiles = plan->npad / ax::kSplitKTile;
    for (int k = 0; k <= kmax; ++k) {
        if (plan->m_k[k] == 0) continue;
        ax::RkSeg s{};
        s.D = plan->D;
        s.ld_d = n;
        s.rows = plan->rows_all + row_off[k];
        s.Y = plan->Y + y_off[k] * n;
        s.ld_y = n;
        s.n = n;
        s.out_off = plan->out_off[k];
        s.m = plan->m_k[k];
        if (precision == AX_PREC_SPLIT) {
            s.Ds = plan->Ds;
            s.ds_ld = plan->npad;
            s.ds_plane = (int64_t)plan->n_pairs * plan->npad;
            s.De = plan->De;
            s.de_ld = tiles;
            s.Ys = plan->Ys + y_off[k] * plan->npad;
            s.ys_ld = plan->npad;
            s.ys_plane = (int64_t)(yrows > 0 ? yrows : 1) * plan->npad;
            s.Ye = plan->Ye + y_off[k] * tiles;
            s.ye_ld = tiles;
        }
        ax::upper_tiles(s.m, static_cast<int>(segs.size()), work);
        segs.push_back(s);
    }
    plan->n_work = static_cast<int>(work.size());
    plan->n_split = choose_split(plan->n_work, n);
    AX_TRY(dev_alloc(plan, &plan->segs, segs.size()));
    AX_TRY(dev_alloc(plan, &plan->work, work.size()));
    AX_TRY(upload(plan->segs, segs.data(), segs.size()));
    AX_TRY(upload(plan->work, work.data(), work.size()));
    AX_TRY(dev_alloc(plan, &plan->partial,
                     (size_t)ax::rk_gemm_partial_doubles(plan->n_work, plan->n_split)));
    AX_CHECK_HIP(hipDeviceSynchronize());
    *out_plan = plan;
    return AX_OK;
}

ax_status ax_rk_tensor_destroy(ax_rk_tensor *plan) {
    if (!plan) return AX_OK;
    (void)hipDeviceSynchronize();
    free_plan(plan);
    delete plan;
    return AX_OK;
}

int64_t ax_rk_tensor_size(const ax_rk_tensor *plan) { return plan ? plan->out_size : -1; }

int ax_rk_tensor_pairs(const ax_rk_tensor *plan, int k) {
    if (!plan || k < 0 || k > plan->kmax) return -1;
    return plan->m_k[k];
}

int64_t ax_rk_tensor_offset(const ax_rk_tensor *plan, int k) {
    if (!plan || k < 0 || k > plan->kmax) return -1;
    return plan->out_off[k];
}

int64_t ax_rk_tensor_footprint(const ax_rk_tensor *plan) { return plan ? plan->footprint : -1; }

ax_status ax_rk_tensor_run_stage(ax_rk_tensor *plan, int stage, const double *P,
                                 const double *Q, double *out, void *stream) {
    AX_REQUIRE(plan != nullptr, "ax_rk_tensor_run_stage: null plan");
    hipStream_t s = ax::as_stream(stream);
    const int64_t n = plan->n;
    switch (stage) {
        case AX_STAGE_DENSITY:
            AX_REQUIRE(P != nullptr, "ax_rk_tensor_run_stage: null orbitals");
            return ax_pair_density(P, Q, n, n, plan->n_pairs, plan->pa, plan->pb, plan->w, plan->D,
                                   n, stream);
        case AX_STAGE_SCAN:
            return ax_yk_batch(plan->D, n, nullptr, plan->rk, plan->rkm, n, plan->nv, plan->ng,
                               plan->kmax, plan->mo, plan->mi, plan->n_items, plan->item_row,
                               plan->item_k, plan->item_out, plan->Y, n, stream);
        case AX_STAGE_SLICE: {
            if (plan->precision != AX_PREC_SPLIT) return AX_OK;
            const int64_t tiles = plan->npad / ax::kSplitKTile;
            ax_status st = ax::slice_rows_launch(plan->D, n, nullptr, plan->n_pairs, n, plan->npad,
                                                 plan->Ds, plan->npad,
                                                 (int64_t)plan->n_pairs * plan->npad, plan->De,
                                                 tiles, s);
            if (st != AX_OK) return st;
            return ax::slice_rows_launch(
                plan->Y, n, nullptr, plan->n_yrows, n, plan->npad, plan->Ys, plan->npad,
                (int64_t)(plan->n_yrows > 0 ? plan->n_yrows : 1) * plan->npad, plan->Ye, tiles, s);
        }
        case AX_STAGE_CONTRACT:
            return ax::rk_gemm_contract_launch(plan->segs, plan->work, plan->n_work, plan->n_split,
                                               plan->precision, plan->partial, s);
        case AX_STAGE_REDUCE:
            AX_REQUIRE(out != nullptr, "ax_rk_tensor_run_stage: null output");
            return ax::rk_gemm_reduce_launch(plan->segs, plan->work, plan->n_work, plan->n_split,
                                             plan->partial, out, s);
        default:
            ax::set_error("ax_rk_tensor_run_stage: unknown stage %d", stage);
            return AX_ERR_VALUE;
    }
}

ax_status ax_rk_tensor_run(ax_rk_tensor *plan, const double *P, const double *Q, double *out,
                           void *stream) {
    AX_REQUIRE(plan && P && out, "ax_rk_tensor_run: null argument");
    for (int stage = AX_STAGE_DENSITY; stage <= AX_STAGE_REDUCE; ++stage) {
        ax_status st = ax_rk_tensor_run_stage(plan, stage, P, Q, out, stream);
        if (st != AX_OK) return st;
    }
    return AX_OK;
}

// ---------------------------------------------------------------------------
// Single block: packed upper triangle of D[rows] Y^T.
// ---------------------------------------------------------------------------
int64_t ax_rk_gemm_workspace_bytes(int m, int64_t n, int precision) {
    if (m <= 0 || n <= 0) return 0;
    std::vector<ax::RkWork> work;
    ax::upper_tiles(m, 0, work);
    const int nw = static_cast<int>(work.size());
    const int split = choose_split(nw, n);
    int64_t bytes = 256 + (int64_t)sizeof(ax::RkSeg) + (int64_t)nw * sizeof(ax::RkWork);
    bytes = (bytes + 255) / 256 * 256;
    bytes += ax::rk_gemm_partial_doubles(nw, split) * 8;
    if (precision == AX_PREC_SPLIT) {
        const int64_t npad = (n + ax::kSplitKTile - 1) / ax::kSplitKTile * ax::kSplitKTile;
        const int64_t tiles = npad / ax::kSplitKTile;
        bytes += 2 * ((int64_t)ax::kSplitSlices * m * npad + 256);
        bytes += 2 * ((int64_t)m * tiles * 4 + 256);
    }
    return bytes;
}

ax_status ax_rk_gemm(const double *D, int64_t ld_d, const int32_t *rows, const double *Y,
                     int64_t ld_y, int m, int64_t n, int precision, void *workspace, double *out,
                     void *stream) {
    AX_REQUIRE(m >= 0 && n > 0 && ld_d >= n && ld_y >= n, "ax_rk_gemm: bad sizes");
    AX_REQUIRE(precision == AX_PREC_FP64 || precision == AX_PREC_SPLIT,
               "ax_rk_gemm: unknown precision %d", precision);
    if (m == 0) return AX_OK;
    AX_REQUIRE(workspace != nullptr, "ax_rk_gemm: workspace required");
    hipStream_t s = ax::as_stream(stream);
    std::vector<ax::RkWork> work;
    ax::upper_tiles(m, 0, work);
    const int nw = static_cast<int>(work.size());
    const int split = choose_split(nw, n);
    char *ws = static_cast<char *>(workspace);
    ax::RkSeg *segs_dev = reinterpret_cast<ax::RkSeg *>(ws);
    ax::RkWork *work_dev = reinterpret_cast<ax::RkWork *>(ws + 256);
    int64_t off = 256 + (int64_t)nw * sizeof(ax::RkWork);
    off = (off + 255) / 256 * 256;
    double *partial = reinterpret_cast<double *>(ws + off);
    off += ax::rk_gemm_partial_doubles(nw, split) * 8;
    ax::RkSeg sg{};
    sg.D = D;
    sg.ld_d = ld_d;
    sg.rows = rows;
    sg.Y = Y;
    sg.ld_y = ld_y;
    sg.n = n;
    sg.out_off = 0;
    sg.m = m;
    if (precision == AX_PREC_SPLIT) {
        const int64_t npad = (n + ax::kSplitKTile - 1) / ax::kSplitKTile * ax::kSplitKTile;
        const int64_t tiles = npad / ax::kSplitKTile;
        int8_t *Ds = reinterpret_cast<int8_t *>(ws + off);
        off += (int64_t)ax::kSplitSlices * m * npad + 256;
        off = (off + 255) / 256 * 256;
        int8_t *Ys = reinterpret_cast<int8_t *>(ws + off);
        off += (int64_t)ax::kSplitSlices * m * npad + 256;
        off = (off + 255) / 256 * 256;
        int32_t *De = reinterpret_cast<int32_t *>(ws + off);
        off += (int64_t)m * tiles * 4 + 256;
        off = (off + 255) / 256 * 256;
        int32_t *Ye = reinterpret_cast<int32_t *>(ws + off);
        // D rows are gathered while slicing, so the sliced operand is identity-indexed
        ax_status st = ax::slice_rows_launch(D, ld_d, rows, m, n, npad, Ds, npad, (int64_t)m * npad,
                                             De, tiles, s);
        if (st != AX_OK) return st;
        st = ax::slice_rows_launch(Y, ld_y, nullptr, m, n, npad, Ys, npad, (int64_t)m * npad, Ye,
                                   tiles, s);
        if (st != AX_OK) return st;
        sg.rows = nullptr;
        sg.Ds = Ds;
        sg.ds_ld = npad;
        sg.ds_plane = (int64_t)m * npad;
        sg.De = De;
        sg.de_ld = tiles;
        sg.Ys = Ys;
        sg.ys_ld = npad;
        sg.ys_plane = (int64_t)m * npad;
        sg.Ye = Ye;
        sg.ye_ld = tiles;
    }
    AX_CHECK_HIP(hipMemcpyAsync(segs_dev, &sg, sizeof(sg), hipMemcpyHostToDevice, s));
    AX_CHECK_HIP(hipMemcpyAsync(work_dev, work.data(), work.size() * sizeof(ax::RkWork),
                                hipMemcpyHostToDevice, s));
    ax_status st = ax::rk_gemm_launch(segs_dev, work_dev, nw, split, precision, partial, out, s);
    if (st != AX_OK) return st;
    // the host-side descriptors above must outlive the asynchronous copies
    AX_CHECK_HIP(hipStreamSynchronize(s));
    return AX_OK;
}

}  // extern "C"
