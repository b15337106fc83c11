// Batched R^k tensor over many orbitals on one grid (the all-quartet
// Slater-integral workload), batches of such plans, and the single-block
// ax_rk_gemm entry point.
//
// Pipeline of one run (all on one stream, capturable in a HIP graph):
//   1. pair_density   D[p] = w * (P_a P_b + Q_a Q_b)  for all pairs a <= b
//   2. yk scan        Y^k[slot(p,k)] for every pair valid at rank k
//   3. (split) slice  int8 planes + exponents of D and of every Y^k row
//   4. contraction    packed upper triangle of D[V_k] Y_k^T for every k,
//                     all ranks in one launch (one segment per rank); tiles
//                     of one K chunk are written to the output directly
//   5. reduction      K-chunk partials summed in a fixed order (no launch
//                     when every tile is one chunk)
// A batch of plans (different grids, orbital counts, k ranges) runs the same
// five launches once for all of them; every kernel looks its plan up from a
// prefix table in its arguments.  A plan's tiles, K splits and summation
// order do not depend on the batch, so its results are bitwise the same
// alone and in any batch.
#include <algorithm>
#include <map>
#include <mutex>
#include <cmath>
#include <cstring>
#include <vector>

#include "ax_common.h"
#include "rk_gemm.h"

struct ax_rk_tensor {
    int nv = 0, ng = 0, n_orb = 0, kmax = 0, precision = 0;
    int64_t n = 0, npad = 0;
    int n_pairs = 0, n_yrows = 0, n_items = 0, n_work = 0, n_split = 1, split_cap = 1;
    int slices = 0;  // int8 slices per operand (split precision)
    std::vector<int> m_k;
    std::vector<int64_t> out_off;
    int64_t out_size = 0;
    int64_t footprint = 0;
    ax::PlanHost host{};
    // job lists of the K-split levels S = 1..split_cap (index S - 1; null
    // until built), on the device: tile-fastest, split-fastest, heaviest
    // first (one job per work-group: `identity` groups) with their per-tile
    // slot ranges, and the balanced order's jobs, work-groups and slot
    // ranges.  Creation builds the default level's; set_schedule and tune
    // build others when first asked (ensure_tables), from the host copies
    // of the tile list below.
    std::vector<const ax::RkJob *> jobs_tf, jobs_sf, jobs_hv, jobs_bal;
    std::vector<const ax::RkTileJobs *> tile_jobs, tile_jobs_bal;
    std::vector<const int *> groups_bal;
    std::vector<int> n_jobs, n_groups_bal, reduce_bal;
    const int *identity = nullptr;
    std::vector<ax::RkWork> work_host;
    std::vector<int> seg_m_host;
    ax::KSpace ks{};
    int64_t partial_slots = 0;  // 64x64 partial tiles the workspace holds
    struct DoneMark *done = nullptr;  // recorded after the plan's last launch set
    // recorded after the set-up upload; shared (reference-counted) by the
    // plans created by one ax_rk_tensor_create_many call
    struct DoneMark *ready = nullptr;
    bool ready_pending = true;  // launches must still wait on `ready`
    bool captured = false;      // launched into a stream capture at least once
    std::vector<void *> allocs;
};

// Completion mark of a launch set: one event recorded once per set, shared
// (reference-counted) by every plan of the set.
struct DoneMark {
    hipEvent_t ev = nullptr;
    int refs = 0;
};

namespace {

std::mutex g_mark_mu;
std::vector<DoneMark *> g_mark_free;

// A mark with no references (a recycled one or a new one); nullptr on failure.
DoneMark *mark_acquire() {
    {
        std::lock_guard<std::mutex> lk(g_mark_mu);
        if (!g_mark_free.empty()) {
            DoneMark *m = g_mark_free.back();
            g_mark_free.pop_back();
            return m;
        }
    }
    DoneMark *m = new DoneMark();
    if (hipEventCreateWithFlags(&m->ev, hipEventDisableTiming) != hipSuccess) {
        delete m;
        return nullptr;
    }
    return m;
}

void mark_release(DoneMark *m) {
    if (!m) return;
    std::lock_guard<std::mutex> lk(g_mark_mu);
    if (--m->refs == 0) g_mark_free.push_back(m);
}

int64_t align256(int64_t x) { return (x + 255) / 256 * 256; }

// Layout of a plan's single device block: every array at a 256-byte aligned
// offset.  `add` reserves and returns the offset.
struct Layout {
    int64_t size = 0;
    int64_t add(int64_t bytes) {
        const int64_t off = size;
        size = align256(size + (bytes > 0 ? bytes : 1));
        return off;
    }
};

// Points the launch descriptors at the job list of the plan's schedule.
void apply_schedule(ax_rk_tensor *p) {
    const int S = p->host.n_split;
    const int o = p->host.order;
    if (o == ax::kOrderBalanced) {
        p->host.jobs = p->jobs_bal[S - 1];
        p->host.groups = p->groups_bal[S - 1];
        p->host.tile_jobs = p->tile_jobs_bal[S - 1];
        p->host.n_groups = p->n_groups_bal[S - 1];
        p->host.needs_reduce = p->reduce_bal[S - 1];
        return;
    }
    p->host.jobs = o == ax::kOrderXcdSplit   ? p->jobs_sf[S - 1]
                   : o == ax::kOrderDynamic ? p->jobs_hv[S - 1]
                                            : p->jobs_tf[S - 1];
    p->host.groups = p->identity;
    p->host.tile_jobs = p->tile_jobs[S - 1];
    p->host.n_groups = p->n_jobs[S - 1];
    p->host.needs_reduce = S > 1;
}

// Builds and uploads the job tables of K-split level S if the plan does not
// hold them yet, growing the partial workspace when the level needs more
// slots (the old workspace stays allocated until the plan is destroyed:
// launches already queued or captured keep using it).  Synchronous.
ax_status ensure_tables(ax_rk_tensor *p, int S) {
    if (S < 1 || S > p->split_cap) {
        ax::set_error("K split %d outside [1, %d]", S, p->split_cap);
        return AX_ERR_VALUE;
    }
    if (p->jobs_tf[S - 1] && p->jobs_sf[S - 1] && p->jobs_hv[S - 1] && p->jobs_bal[S - 1])
        return AX_OK;  // a new plan holds its default level's tile-fastest list only
    const ax::JobTables t = ax::make_jobs(p->work_host, p->seg_m_host, S, p->ks);
    const ax::BalancedJobs b =
        ax::make_balanced_jobs(p->work_host, p->seg_m_host, ax::balanced_groups(S), p->ks);
    const int64_t slots = std::max<int64_t>((int64_t)p->n_work * S, (int64_t)b.jobs.size());
    Layout L;
    const int64_t o_tf = L.add((int64_t)t.tile_fastest.size() * sizeof(ax::RkJob));
    const int64_t o_sf = L.add((int64_t)t.split_fastest.size() * sizeof(ax::RkJob));
    const int64_t o_hv = L.add((int64_t)t.heaviest.size() * sizeof(ax::RkJob));
    const int64_t o_tj = L.add((int64_t)t.tiles.size() * sizeof(ax::RkTileJobs));
    const int64_t o_jb = L.add((int64_t)b.jobs.size() * sizeof(ax::RkJob));
    const int64_t o_gb = L.add((int64_t)b.groups.size() * sizeof(int));
    const int64_t o_tjb = L.add((int64_t)b.tiles.size() * sizeof(ax::RkTileJobs));
    const int64_t head = L.size;
    const bool grow = slots > p->partial_slots;
    const int64_t o_part = grow ? L.add(slots * ax::gemm_tile() * ax::gemm_tile() * 8) : -1;
    std::vector<char> img(static_cast<size_t>(head), 0);
    auto put = [&](int64_t off, const void *src, size_t bytes) {
        if (bytes) std::memcpy(img.data() + off, src, bytes);
    };
    put(o_tf, t.tile_fastest.data(), t.tile_fastest.size() * sizeof(ax::RkJob));
    put(o_sf, t.split_fastest.data(), t.split_fastest.size() * sizeof(ax::RkJob));
    put(o_hv, t.heaviest.data(), t.heaviest.size() * sizeof(ax::RkJob));
    put(o_tj, t.tiles.data(), t.tiles.size() * sizeof(ax::RkTileJobs));
    put(o_jb, b.jobs.data(), b.jobs.size() * sizeof(ax::RkJob));
    put(o_gb, b.groups.data(), b.groups.size() * sizeof(int));
    put(o_tjb, b.tiles.data(), b.tiles.size() * sizeof(ax::RkTileJobs));
    void *blk = nullptr;
    hipError_t e = ax::pool_alloc(&blk, static_cast<size_t>(L.size));
    if (e != hipSuccess) {
        ax::set_error("device allocation of %lld bytes failed: %s", (long long)L.size,
                      hipGetErrorString(e));
        return e == hipErrorOutOfMemory ? AX_ERR_MEMORY : AX_ERR_HIP;
    }
    hipStream_t ss = ax::setup_stream();
    e = hipMemcpyAsync(blk, img.data(), static_cast<size_t>(head), hipMemcpyHostToDevice, ss);
    if (e == hipSuccess) e = hipStreamSynchronize(ss);
    if (e != hipSuccess) {
        ax::pool_free(blk);
        ax::set_error("job table upload failed: %s", hipGetErrorString(e));
        return AX_ERR_HIP;
    }
    p->allocs.push_back(blk);
    p->footprint += static_cast<int64_t>(ax::pool_class(static_cast<size_t>(L.size)));
    char *base = static_cast<char *>(blk);
    p->jobs_tf[S - 1] = reinterpret_cast<const ax::RkJob *>(base + o_tf);
    p->jobs_sf[S - 1] = reinterpret_cast<const ax::RkJob *>(base + o_sf);
    p->jobs_hv[S - 1] = reinterpret_cast<const ax::RkJob *>(base + o_hv);
    p->tile_jobs[S - 1] = reinterpret_cast<const ax::RkTileJobs *>(base + o_tj);
    p->n_jobs[S - 1] = static_cast<int>(t.tile_fastest.size());
    p->jobs_bal[S - 1] = reinterpret_cast<const ax::RkJob *>(base + o_jb);
    p->groups_bal[S - 1] = reinterpret_cast<const int *>(base + o_gb);
    p->tile_jobs_bal[S - 1] = reinterpret_cast<const ax::RkTileJobs *>(base + o_tjb);
    p->n_groups_bal[S - 1] = static_cast<int>(b.groups.size()) - 1;
    p->reduce_bal[S - 1] = b.jobs.size() > b.tiles.size() ? 1 : 0;
    if (grow) {
        p->host.partial = reinterpret_cast<double *>(base + o_part);
        p->partial_slots = slots;
    }
    return AX_OK;
}

void free_plan(ax_rk_tensor *p) {
    for (void *q : p->allocs) ax::pool_free(q);
    p->allocs.clear();
    mark_release(p->done);
    mark_release(p->ready);
    p->done = nullptr;
    p->ready = nullptr;
}

// Makes stream s wait for the set-up upload of every plan (a device-side
// dependency; under stream capture, where an outside event cannot be waited
// on, the host waits instead).
ax_status await_ready(ax_rk_tensor *const *plans, int n, hipStream_t s) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    AX_CHECK_HIP(hipStreamIsCapturing(s, &cs));
    for (int i = 0; i < n; ++i) {
        ax_rk_tensor *p = plans[i];
        if (!p->ready_pending) continue;
        if (hipEventQuery(p->ready->ev) == hipSuccess) {
            p->ready_pending = false;
            continue;
        }
        if (cs != hipStreamCaptureStatusNone) {
            AX_CHECK_HIP(hipEventSynchronize(p->ready->ev));
            p->ready_pending = false;
        } else {
            AX_CHECK_HIP(hipStreamWaitEvent(s, p->ready->ev, 0));
        }
    }
    return AX_OK;
}

// Partial-interval matrices of each n_gauss, computed once.
ax_status partial_matrices_cached(int ng, const double **mo, const double **mi) {
    static std::mutex mu;
    static std::map<int, std::pair<std::vector<double>, std::vector<double>>> cache;
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(ng);
    if (it == cache.end()) {
        std::vector<double> a(ng * ng), b(ng * ng);
        ax_status st = ax_yk_partial_matrices(ng, a.data(), b.data());
        if (st != AX_OK) return st;
        it = cache.emplace(ng, std::make_pair(std::move(a), std::move(b))).first;
    }
    *mo = it->second.first.data();
    *mi = it->second.second.data();
    return AX_OK;
}

// Int8 slices per operand for a split-precision plan, or 0 when the plan's
// contraction has to run in FP64.  The slices of one 64-node K-tile share
// one exponent, so a term's error is bounded relative to the largest |D| and
// |Y^k| inside its tile, and across a tile r^-(k+1) and r^k change by
// (r_last / r_first)^(k+1) = 2^bits.  The bound scales as 2^(bits - 7 S):
// 7 slices up to 5 bits, 8 slices up to 12 bits (coarse exponential grids at
// high k), FP64 beyond.  The first cell reaches r = 0 and is left out (the
// Gauss weights there are of the order of its width).  Without the grid
// (footprint estimates) 8 slices are assumed.
int choose_slices(const double *r, int nv, int ng, int kmax) {
    if (!r) return ax::kSplitSlicesFine;
    const int64_t n = (int64_t)nv * ng;
    double worst = 1.0;
    for (int64_t x0 = 0; x0 < n; x0 += ax::kSplitKTile) {
        const int64_t x1 = std::min(n, x0 + ax::kSplitKTile) - 1;
        const int64_t start = std::max<int64_t>(x0, ng);
        if (start > x1 || r[start] <= 0.0) continue;
        worst = std::max(worst, r[x1] / r[start]);
    }
    const double bits = (kmax + 1) * std::log2(worst);
    if (bits <= 5.0) return ax::kSplitSlices;
    if (bits <= 12.0) return ax::kSplitSlicesFine;
    return 0;
}

int choose_split(int n_work, int64_t n) {
    // about one round of FP64 contraction work-groups (two per CU on 256
    // CUs), K chunks of at least 128 nodes, at most 16 chunks
    int64_t s = 512 / std::max(n_work, 1);
    s = std::min<int64_t>(s, n / 128);
    s = std::min<int64_t>(s, 16);
    return static_cast<int>(std::max<int64_t>(s, 1));
}

// Records one completion mark after a launch sequence and hands it to every
// plan of the sequence (skipped under stream capture, where destroy
// synchronises the device).
ax_status mark_done(ax_rk_tensor *const *plans, int n, hipStream_t s) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    AX_CHECK_HIP(hipStreamIsCapturing(s, &cs));
    if (cs != hipStreamCaptureStatusNone) {
        for (int i = 0; i < n; ++i) plans[i]->captured = true;
        return AX_OK;
    }
    if (n <= 0) return AX_OK;
    DoneMark *m = mark_acquire();
    if (!m) {
        ax::set_error("completion event creation failed");
        return AX_ERR_HIP;
    }
    const hipError_t e = hipEventRecord(m->ev, s);
    if (e != hipSuccess) {
        std::lock_guard<std::mutex> lk(g_mark_mu);
        g_mark_free.push_back(m);
        ax::set_error("hipEventRecord: %s", hipGetErrorString(e));
        return AX_ERR_HIP;
    }
    {
        std::lock_guard<std::mutex> lk(g_mark_mu);
        m->refs = n;
    }
    for (int i = 0; i < n; ++i) {
        mark_release(plans[i]->done);
        plans[i]->done = m;
    }
    return AX_OK;
}

ax_status run_stage_batch(ax_rk_tensor *const *plans, int n, int stage, const double *const *P,
                          const double *const *Q, double *const *out, hipStream_t s) {
    const ax::PlanHost *h[ax::kMaxBatch];
    for (int i = 0; i < n; ++i) h[i] = &plans[i]->host;
    switch (stage) {
        case AX_STAGE_DENSITY:
            return ax::density_batch_launch(h, n, P, Q, s);
        case AX_STAGE_SCAN:
            return ax::scan_batch_launch(h, n, s);
        case AX_STAGE_SLICE:
            return ax::slice_batch_launch(h, n, s);
        case AX_STAGE_CONTRACT:
            return ax::contract_batch_launch(h, n, out, s);
        case AX_STAGE_REDUCE:
            for (int i = 0; i < n; ++i)
                AX_REQUIRE(out[i] != nullptr, "R^k tensor run: null output for plan %d", i);
            return ax::reduce_batch_launch(h, n, out, s);
        default:
            ax::set_error("R^k tensor run: unknown stage %d", stage);
            return AX_ERR_VALUE;
    }
}

}  // namespace

namespace {

// Explicit pair list of a pair-product plan (every pair valid at every rank);
// null for the all-pairs plan with the angular rule.
struct PairList {
    int n_pairs;
    const int32_t *pa, *pb;
};

// A built plan whose device block still has to receive its set-up image
// (grid, index lists, segments, job tables) and its r-power tables.
struct Setup {
    ax_rk_tensor *plan = nullptr;
    void *block = nullptr;
    std::vector<char> image;  // the first image.size() bytes of the block
    const double *r = nullptr;
    double *rk = nullptr, *rkm = nullptr;
};

// Set-up of up to kMaxBatch plans in one launch: each plan's uploaded image
// is copied from the device mirror of the staging buffer into its block, 16
// bytes per item, and its r^k and r^-(k+1) tables (the values of
// rpow_kernel, numerics/yk.py:113-117) are formed from the r values of that
// same mirror, so neither part waits for the other.
struct SetupArg {
    const uint4 *src;  // the plan's image in the device mirror
    uint4 *dst;        // the plan's block
    const double *r;   // its r values in the mirror
    double *rk, *rkm;
    int64_t n16;  // 16-byte units of the image (images are 256-byte multiples)
    int64_t n;    // nodes
    int64_t kmax;
};
__global__ void setup_batch_kernel(const ax::BatchOf<SetupArg> A) {
    const int64_t total = A.end[A.n - 1];
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        int64_t local;
        const int p = ax::batch_find(A, t, local);
        const SetupArg &g = A.a[p];
        if (local < g.n16) {
            g.dst[local] = g.src[local];
            continue;
        }
        local -= g.n16;
        const int k = static_cast<int>(local / g.n);
        const double rv = g.r[local - (int64_t)k * g.n];
        g.rk[local] = pow(rv, static_cast<double>(k));
        g.rkm[local] = pow(rv, static_cast<double>(-(k + 1)));
    }
}

// Uploads the set-up images of n built plans on `ss`: one pinned staging
// buffer moved with one copy into its device mirror, then launches of up to
// kMaxBatch plans that scatter the images into the plans' blocks and form
// the r-power tables, and one `ready` mark that every plan shares.  On
// failure every plan is freed.
ax_status upload_plans(Setup *su, int n, hipStream_t ss) {
    auto fail_all = [&](ax_status st) {
        for (int i = 0; i < n; ++i) {
            if (!su[i].plan) continue;
            free_plan(su[i].plan);
            delete su[i].plan;
            su[i].plan = nullptr;
        }
        return st;
    };
    std::vector<int64_t> off(static_cast<size_t>(n) + 1, 0);
    for (int i = 0; i < n; ++i) off[i + 1] = off[i] + align256((int64_t)su[i].image.size());
    void *stage = nullptr, *mirror = nullptr;
    int slot = -1;
    hipError_t e = ax::staging_acquire(static_cast<size_t>(off[n]), ss, &stage, &slot, &mirror);
    if (e != hipSuccess) {
        ax::set_error("set-up staging of %lld bytes failed: %s", (long long)off[n],
                      hipGetErrorString(e));
        return fail_all(e == hipErrorOutOfMemory ? AX_ERR_MEMORY : AX_ERR_HIP);
    }
    char *img = static_cast<char *>(stage);
    for (int i = 0; i < n; ++i)
        if (!su[i].image.empty()) std::memcpy(img + off[i], su[i].image.data(), su[i].image.size());
    e = hipMemcpyAsync(mirror, stage, static_cast<size_t>(off[n]), hipMemcpyHostToDevice, ss);
    const char *dimg = static_cast<const char *>(mirror);
    for (int i0 = 0; i0 < n && e == hipSuccess; i0 += ax::kMaxBatch) {
        ax::BatchOf<SetupArg> A{};
        A.n = std::min(ax::kMaxBatch, n - i0);
        int64_t total = 0;
        for (int j = 0; j < A.n; ++j) {
            const Setup &u = su[i0 + j];
            const ax_rk_tensor *p = u.plan;
            const int64_t o_r = reinterpret_cast<const char *>(u.r) -
                                static_cast<const char *>(u.block);
            SetupArg &g = A.a[j];
            g.src = reinterpret_cast<const uint4 *>(dimg + off[i0 + j]);
            g.dst = static_cast<uint4 *>(u.block);
            g.r = reinterpret_cast<const double *>(dimg + off[i0 + j] + o_r);
            g.rk = u.rk;
            g.rkm = u.rkm;
            g.n16 = static_cast<int64_t>(u.image.size()) / 16;
            g.n = p->n;
            g.kmax = p->kmax;
            total += g.n16 + p->n * (p->kmax + 1);
            A.end[j] = total;
        }
        hipLaunchKernelGGL(setup_batch_kernel, dim3(ax::grid_for(total, 256, 4096)), dim3(256), 0,
                           ss, A);
        e = hipGetLastError();
    }
    const hipError_t er = ax::staging_release(slot);
    if (e == hipSuccess) e = er;
    DoneMark *m = e == hipSuccess ? mark_acquire() : nullptr;
    if (e == hipSuccess && !m) e = hipErrorOutOfMemory;
    if (e == hipSuccess) e = hipEventRecord(m->ev, ss);
    if (e != hipSuccess) {
        if (m) {
            std::lock_guard<std::mutex> lk(g_mark_mu);
            g_mark_free.push_back(m);
        }
        ax::set_error("R^k tensor set-up failed: %s", hipGetErrorString(e));
        return fail_all(AX_ERR_HIP);
    }
    {
        std::lock_guard<std::mutex> lk(g_mark_mu);
        m->refs = n;
    }
    for (int i = 0; i < n; ++i) su[i].plan->ready = m;
    return AX_OK;
}

// Builds a plan (uploading it, or with setup_out handing it back for a
// batched upload), or (estimate != nullptr) only computes the device bytes
// it would hold, without touching the device.

ax_status plan_create(const double *r_host, const double *w_host, int nv, int ng, int n_orb,
                      const int32_t *l_host, int kmax, int precision, ax_rk_tensor **out_plan,
                      int64_t *estimate, const PairList *explicit_pairs = nullptr,
                      Setup *setup_out = nullptr) {
    AX_REQUIRE(out_plan != nullptr || estimate != nullptr || setup_out != nullptr,
               "ax_rk_tensor_create: null plan pointer");
    AX_REQUIRE((l_host || explicit_pairs) && (estimate || (r_host && w_host)),
               "ax_rk_tensor_create: null grid or l values");
    AX_REQUIRE(nv >= 1 && ng >= 1 && ng <= 16, "ax_rk_tensor_create: bad grid nv=%d ng=%d", nv, ng);
    AX_REQUIRE(n_orb >= 1 && kmax >= 0 && kmax <= 64, "ax_rk_tensor_create: bad n_orb/kmax");
    AX_REQUIRE(precision == AX_PREC_FP64 || precision == AX_PREC_SPLIT,
               "ax_rk_tensor_create: unknown precision %d", precision);
    if (!explicit_pairs)
        for (int a = 0; a < n_orb; ++a)
            AX_REQUIRE(l_host[a] >= 0, "ax_rk_tensor_create: negative l for orbital %d", a);
    if (precision == AX_PREC_SPLIT && choose_slices(r_host, nv, ng, kmax) == 0)
        precision = AX_PREC_FP64;  // the grid is too coarse for the split error bound
    const int64_t n = (int64_t)nv * ng;
    const int64_t npad = (n + ax::kSplitKTile - 1) / ax::kSplitKTile * ax::kSplitKTile;
    const int64_t tiles = npad / ax::kSplitKTile;

    // pairs a <= b in lexicographic order, or the caller's list
    std::vector<int32_t> pa, pb;
    if (explicit_pairs) {
        AX_REQUIRE(explicit_pairs->n_pairs >= 1 && explicit_pairs->pa && explicit_pairs->pb,
                   "ax_rk_tensor_create_pairs: empty pair list");
        for (int p = 0; p < explicit_pairs->n_pairs; ++p) {
            const int a = explicit_pairs->pa[p], b = explicit_pairs->pb[p];
            AX_REQUIRE(a >= 0 && a < n_orb && b >= 0 && b < n_orb,
                       "ax_rk_tensor_create_pairs: pair %d = (%d, %d) outside [0, %d)", p, a, b,
                       n_orb);
            pa.push_back(a);
            pb.push_back(b);
        }
    } else {
        for (int a = 0; a < n_orb; ++a)
            for (int b = a; b < n_orb; ++b) {
                pa.push_back(a);
                pb.push_back(b);
            }
    }
    const int n_pairs = static_cast<int>(pa.size());

    // valid pairs per rank, Y^k slots, scan items: pair p is valid at the
    // ranks lo[p] <= k <= hi[p] with k = lo[p] (mod 2) (every rank for an
    // explicit pair list).  The scan items are the rank-major list of valid
    // (pair, k): item rows = rows_all, item_out = its position.
    std::vector<int32_t> plo(n_pairs, 0), phi(n_pairs, kmax);
    int64_t n_valid = 0;
    for (int p = 0; p < n_pairs; ++p) {
        if (!explicit_pairs) {
            const int la = l_host[pa[p]], lb = l_host[pb[p]];
            plo[p] = la > lb ? la - lb : lb - la;
            phi[p] = la + lb;
            const int top = std::min(phi[p], kmax);
            n_valid += top >= plo[p] ? (top - plo[p]) / 2 + 1 : 0;
        } else {
            n_valid += kmax + 1;
        }
    }
    const int step = explicit_pairs ? 1 : 2;
    std::vector<int32_t> rows_all(static_cast<size_t>(n_valid) + 1), it_k(n_valid), it_out(n_valid);
    std::vector<int64_t> row_off(kmax + 1), y_off(kmax + 1), out_off(kmax + 1);
    std::vector<int> m_k(kmax + 1, 0);
    int64_t yrows = 0, outsz = 0, idx = 0;
    for (int k = 0; k <= kmax; ++k) {
        row_off[k] = idx;
        y_off[k] = yrows;
        const int64_t start = idx;
        for (int p = 0; p < n_pairs; ++p) {
            rows_all[idx] = p;  // kept when valid (one spare slot at the end)
            const int d = k - plo[p];
            idx += d >= 0 && k <= phi[p] && d % step == 0;
        }
        const int m = static_cast<int>(idx - start);
        std::fill(it_k.begin() + start, it_k.begin() + idx, k);
        for (int64_t i = start; i < idx; ++i) it_out[i] = static_cast<int32_t>(i);
        m_k[k] = m;
        out_off[k] = outsz;
        outsz += (int64_t)m * (m + 1) / 2;
        yrows += m;
    }
    rows_all.resize(static_cast<size_t>(n_valid));
    const std::vector<int32_t> &it_row = rows_all;
    AX_REQUIRE(yrows < 0x7fffffff, "ax_rk_tensor_create: too many Y^k rows");
    const int64_t yr = yrows > 0 ? yrows : 1;

    // contraction segments (one per rank) and the tile work list
    std::vector<ax::RkWork> work;
    std::vector<int> seg_k;
    for (int k = 0; k <= kmax; ++k) {
        if (m_k[k] == 0) continue;
        ax::upper_tiles(m_k[k], static_cast<int>(seg_k.size()), work);
        seg_k.push_back(k);
    }
    const int n_work = static_cast<int>(work.size());
    const int n_split = choose_split(n_work, n);
    const int split_cap = std::min(16, 4 * n_split);  // partials sized for tuning up to here
    AX_REQUIRE(n_work == 0 || work.back().tj < 0x8000, "ax_rk_tensor_create: too many tiles");
    std::vector<int> seg_m;
    for (int k : seg_k) seg_m.push_back(m_k[k]);
    const ax::KSpace ks = ax::k_space(n, precision == AX_PREC_FP64);
    // the default schedule's job list only (tile-fastest at the default K
    // split); the other orders and levels are built when a schedule or the
    // tuner asks for them (ensure_tables): a one-shot served plan never does
    const ax::JobTables jobs0 = ax::make_jobs(work, seg_m, n_split, ks, false);
    const int64_t slots = (int64_t)n_work * n_split;
    std::vector<int> ident((size_t)n_work * split_cap + 1);
    for (size_t i = 0; i < ident.size(); ++i) ident[i] = static_cast<int>(i);

    // one device block for everything the plan owns
    Layout L;
    const int64_t o_r = L.add(n * 8), o_w = L.add(n * 8);
    const int64_t o_mo = L.add((int64_t)ng * ng * 8), o_mi = L.add((int64_t)ng * ng * 8);
    const int64_t o_pa = L.add(n_pairs * 4), o_pb = L.add(n_pairs * 4);
    const int64_t o_rows = L.add((int64_t)rows_all.size() * 4);
    const int64_t o_irow = L.add((int64_t)it_row.size() * 4);
    const int64_t o_ik = L.add((int64_t)it_k.size() * 4);
    const int64_t o_iout = L.add((int64_t)it_out.size() * 4);
    const int64_t o_segs = L.add((int64_t)seg_k.size() * sizeof(ax::RkSeg));
    const int64_t o_work = L.add((int64_t)n_work * sizeof(ax::RkWork));
    const int64_t o_jtf = L.add((int64_t)jobs0.tile_fastest.size() * sizeof(ax::RkJob));
    const int64_t o_tj = L.add((int64_t)jobs0.tiles.size() * sizeof(ax::RkTileJobs));
    const int64_t o_ident = L.add((int64_t)ident.size() * sizeof(int));
    const int64_t o_ctr = L.add(2 * sizeof(int));  // dynamic-order ticket (zero)
    const int64_t head = L.size;  // everything above is uploaded from the host
    // r^k and r^-(k+1) tables: formed on the device from the uploaded grid
    const int64_t o_rk = L.add(n * (kmax + 1) * 8), o_rkm = L.add(n * (kmax + 1) * 8);
    // split precision on 8-node cells: the scan writes Y^k slices only
    const bool y_fp64 = !(precision == AX_PREC_SPLIT && ax::scan_slices_fused(ng, nv));
    const int64_t o_D = L.add((int64_t)n_pairs * n * 8);
    const int64_t o_Y = y_fp64 ? L.add(yr * n * 8) : -1;
    const int64_t o_part = L.add(slots * ax::gemm_tile() * ax::gemm_tile() * 8);
    int64_t o_Ds = 0, o_Ys = 0, o_De = 0, o_Ye = 0;
    const int slices = precision == AX_PREC_SPLIT ? choose_slices(r_host, nv, ng, kmax) : 0;
    if (precision == AX_PREC_SPLIT) {
        o_Ds = L.add((int64_t)slices * n_pairs * npad);
        o_Ys = L.add((int64_t)slices * yr * npad);
        o_De = L.add((int64_t)n_pairs * tiles * 4);
        o_Ye = L.add(yr * tiles * 4);
    }
    // FP64 first K-tile of the Y^k rows, when the scan keeps no FP64 rows
    const int64_t o_Y0 = (precision == AX_PREC_SPLIT && !y_fp64)
                             ? L.add(yr * ax::kSplitKTile * 8) : -1;
    if (estimate) {
        *estimate = static_cast<int64_t>(ax::pool_class(static_cast<size_t>(L.size)));
        return AX_OK;
    }

    ax_rk_tensor *plan = new ax_rk_tensor();
    plan->nv = nv;
    plan->ng = ng;
    plan->n_orb = n_orb;
    plan->kmax = kmax;
    plan->precision = precision;
    plan->n = n;
    plan->npad = npad;
    plan->n_pairs = n_pairs;
    plan->n_yrows = static_cast<int>(yrows);
    plan->n_items = static_cast<int>(it_row.size());
    plan->n_work = n_work;
    plan->n_split = n_split;
    plan->split_cap = split_cap;
    plan->slices = slices;
    plan->m_k = m_k;
    plan->out_off = out_off;
    plan->out_size = outsz;
    plan->work_host = work;
    plan->seg_m_host = seg_m;
    plan->ks = ks;
    plan->partial_slots = slots;
    for (auto *v : {&plan->jobs_tf, &plan->jobs_sf, &plan->jobs_hv, &plan->jobs_bal})
        v->assign(split_cap, nullptr);
    plan->tile_jobs.assign(split_cap, nullptr);
    plan->tile_jobs_bal.assign(split_cap, nullptr);
    plan->groups_bal.assign(split_cap, nullptr);
    plan->n_jobs.assign(split_cap, 0);
    plan->n_groups_bal.assign(split_cap, 0);
    plan->reduce_bal.assign(split_cap, 0);

    void *blk = nullptr;
    hipError_t he = ax::pool_alloc(&blk, static_cast<size_t>(L.size));
    if (he != hipSuccess) {
        delete plan;
        ax::set_error("device allocation of %lld bytes failed: %s", (long long)L.size,
                      hipGetErrorString(he));
        return he == hipErrorOutOfMemory ? AX_ERR_MEMORY : AX_ERR_HIP;
    }
    plan->allocs.push_back(blk);
    plan->footprint = static_cast<int64_t>(ax::pool_class(static_cast<size_t>(L.size)));
    char *base = static_cast<char *>(blk);
    auto dptr = [&](int64_t off) { return static_cast<void *>(base + off); };

    // host image of the uploaded head of the block
    Setup su;
    su.image.assign(static_cast<size_t>(head), 0);
    char *img = su.image.data();
    auto put = [&](int64_t off, const void *src, size_t bytes) {
        if (bytes) std::memcpy(img + off, src, bytes);
    };
    put(o_r, r_host, n * 8);
    put(o_w, w_host, n * 8);
    {
        const double *mo = nullptr, *mi = nullptr;
        ax_status st = partial_matrices_cached(ng, &mo, &mi);
        if (st != AX_OK) {
            free_plan(plan);
            delete plan;
            return st;
        }
        put(o_mo, mo, (size_t)ng * ng * 8);
        put(o_mi, mi, (size_t)ng * ng * 8);
    }
    put(o_pa, pa.data(), pa.size() * 4);
    put(o_pb, pb.data(), pb.size() * 4);
    put(o_rows, rows_all.data(), rows_all.size() * 4);
    put(o_irow, it_row.data(), it_row.size() * 4);
    put(o_ik, it_k.data(), it_k.size() * 4);
    put(o_iout, it_out.data(), it_out.size() * 4);

    double *D = static_cast<double *>(dptr(o_D));
    double *Y = y_fp64 ? static_cast<double *>(dptr(o_Y)) : nullptr;
    int8_t *Ds = precision == AX_PREC_SPLIT ? static_cast<int8_t *>(dptr(o_Ds)) : nullptr;
    int8_t *Ys = precision == AX_PREC_SPLIT ? static_cast<int8_t *>(dptr(o_Ys)) : nullptr;
    int32_t *De = precision == AX_PREC_SPLIT ? static_cast<int32_t *>(dptr(o_De)) : nullptr;
    int32_t *Ye = precision == AX_PREC_SPLIT ? static_cast<int32_t *>(dptr(o_Ye)) : nullptr;
    const int32_t *rows_dev = static_cast<const int32_t *>(dptr(o_rows));
    // FP64 first K-tile of the Y^k rows for the split contraction: the scan's
    // own copy, or the FP64 rows when they exist
    double *Y0 = o_Y0 >= 0 ? static_cast<double *>(dptr(o_Y0)) : nullptr;
    const int64_t y0_ld = Y0 ? ax::kSplitKTile : n;
    double *Y0base = precision == AX_PREC_SPLIT ? (Y0 ? Y0 : Y) : nullptr;
    {
        std::vector<ax::RkSeg> segs;
        for (int k : seg_k) {
            ax::RkSeg sg{};
            sg.D = D;
            sg.ld_d = n;
            sg.rows = rows_dev + row_off[k];
            sg.Y = Y ? Y + y_off[k] * n : nullptr;
            sg.ld_y = n;
            sg.n = n;
            sg.out_off = out_off[k];
            sg.m = m_k[k];
            if (precision == AX_PREC_SPLIT) {
                sg.Ds = Ds;
                sg.ds_ld = npad;
                sg.ds_plane = (int64_t)n_pairs * npad;
                sg.De = De;
                sg.de_ld = tiles;
                sg.Ys = Ys + y_off[k] * npad;
                sg.ys_ld = npad;
                sg.ys_plane = yr * npad;
                sg.Ye = Ye + y_off[k] * tiles;
                sg.ye_ld = tiles;
                sg.Y0 = Y0base ? Y0base + y_off[k] * y0_ld : nullptr;
                sg.y0_ld = y0_ld;
                sg.d_rows = sg.rows;
            }
            segs.push_back(sg);
        }
        put(o_segs, segs.data(), segs.size() * sizeof(ax::RkSeg));
        put(o_work, work.data(), work.size() * sizeof(ax::RkWork));
        const int d = n_split - 1;  // the default level's tables
        put(o_jtf, jobs0.tile_fastest.data(), jobs0.tile_fastest.size() * sizeof(ax::RkJob));
        put(o_tj, jobs0.tiles.data(), jobs0.tiles.size() * sizeof(ax::RkTileJobs));
        plan->jobs_tf[d] = static_cast<const ax::RkJob *>(dptr(o_jtf));
        plan->tile_jobs[d] = static_cast<const ax::RkTileJobs *>(dptr(o_tj));
        plan->n_jobs[d] = static_cast<int>(jobs0.tile_fastest.size());
        put(o_ident, ident.data(), ident.size() * sizeof(int));
        plan->identity = static_cast<const int *>(dptr(o_ident));
    }
    ax::PlanHost &h = plan->host;
    h.n = n;
    h.npad = npad;
    h.nv = nv;
    h.ng = ng;
    h.kmax = kmax;
    h.n_pairs = n_pairs;
    h.n_items = plan->n_items;
    h.n_yrows = plan->n_yrows;
    h.n_work = n_work;
    h.n_split = n_split;
    h.precision = precision;
    h.order = ax::kOrderLinear;
    h.slices = slices;
    h.mo = static_cast<const double *>(dptr(o_mo));
    h.mi = static_cast<const double *>(dptr(o_mi));
    h.w = static_cast<const double *>(dptr(o_w));
    h.rk = static_cast<const double *>(dptr(o_rk));
    h.rkm = static_cast<const double *>(dptr(o_rkm));
    h.pa = static_cast<const int32_t *>(dptr(o_pa));
    h.pb = static_cast<const int32_t *>(dptr(o_pb));
    h.item_row = static_cast<const int32_t *>(dptr(o_irow));
    h.item_k = static_cast<const int32_t *>(dptr(o_ik));
    h.item_out = static_cast<const int32_t *>(dptr(o_iout));
    h.D = D;
    h.Y = Y;
    h.Y0 = Y0;
    h.Ds = Ds;
    h.Ys = Ys;
    h.De = De;
    h.Ye = Ye;
    h.segs = static_cast<const ax::RkSeg *>(dptr(o_segs));
    h.work = static_cast<const ax::RkWork *>(dptr(o_work));
    h.partial = static_cast<double *>(dptr(o_part));
    h.counter = static_cast<int *>(dptr(o_ctr));
    apply_schedule(plan);

    su.plan = plan;
    su.block = blk;
    su.r = static_cast<const double *>(dptr(o_r));
    su.rk = static_cast<double *>(dptr(o_rk));
    su.rkm = static_cast<double *>(dptr(o_rkm));
    if (setup_out) {
        *setup_out = std::move(su);
        return AX_OK;
    }
    const ax_status st = upload_plans(&su, 1, ax::setup_stream());
    if (st == AX_OK) *out_plan = plan;
    return st;
}

}  // namespace

extern "C" {

ax_status ax_rk_tensor_create(const double *r_host, const double *w_host, int nv, int ng,
                              int n_orb, const int32_t *l_host, int kmax, int precision,
                              ax_rk_tensor **out_plan) {
    AX_REQUIRE(out_plan != nullptr, "ax_rk_tensor_create: null plan pointer");
    return plan_create(r_host, w_host, nv, ng, n_orb, l_host, kmax, precision, out_plan, nullptr);
}

ax_status ax_rk_tensor_create_many(int n_plans, const double *const *r_host,
                                   const double *const *w_host, const int32_t *nv,
                                   const int32_t *ng, const int32_t *n_orb,
                                   const int32_t *const *l_host, const int32_t *kmax,
                                   const int32_t *precision, void *stream,
                                   ax_rk_tensor **out_plans) {
    AX_REQUIRE(n_plans >= 0 && (n_plans == 0 || (r_host && w_host && nv && ng && n_orb && l_host &&
                                                   kmax && precision && out_plans)),
               "ax_rk_tensor_create_many: null argument");
    std::vector<Setup> su(static_cast<size_t>(n_plans));
    for (int i = 0; i < n_plans; ++i) out_plans[i] = nullptr;
    for (int i = 0; i < n_plans; ++i) {
        const ax_status st = plan_create(r_host[i], w_host[i], nv[i], ng[i], n_orb[i], l_host[i],
                                         kmax[i], precision[i], nullptr, nullptr, nullptr, &su[i]);
        if (st != AX_OK) {
            for (int j = 0; j < i; ++j) {
                free_plan(su[j].plan);
                delete su[j].plan;
            }
            return st;
        }
    }
    if (n_plans == 0) return AX_OK;
    const ax_status st = upload_plans(
        su.data(), n_plans, stream ? static_cast<hipStream_t>(stream) : ax::setup_stream());
    if (st != AX_OK) return st;
    for (int i = 0; i < n_plans; ++i) out_plans[i] = su[i].plan;
    return AX_OK;
}

ax_status ax_rk_tensor_describe(const ax_rk_tensor *plan, int64_t *size, int32_t *slices,
                                int32_t *pairs, int64_t *offsets) {
    AX_REQUIRE(plan && size && slices && pairs && offsets, "ax_rk_tensor_describe: null argument");
    *size = plan->out_size;
    *slices = plan->slices;
    for (int k = 0; k <= plan->kmax; ++k) {
        pairs[k] = plan->m_k[k];
        offsets[k] = plan->out_off[k];
    }
    return AX_OK;
}

ax_status ax_rk_tensor_create_pairs(const double *r_host, const double *w_host, int nv, int ng,
                                    int n_fn, int n_pairs, const int32_t *pa_host,
                                    const int32_t *pb_host, int kmax, int precision,
                                    ax_rk_tensor **out_plan) {
    AX_REQUIRE(out_plan != nullptr, "ax_rk_tensor_create_pairs: null plan pointer");
    const PairList pl{n_pairs, pa_host, pb_host};
    return plan_create(r_host, w_host, nv, ng, n_fn, nullptr, kmax, precision, out_plan, nullptr,
                       &pl);
}

int64_t ax_rk_tensor_estimate(int nv, int ng, int n_orb, const int32_t *l_host, int kmax,
                              int precision) {
    int64_t bytes = -1;
    if (plan_create(nullptr, nullptr, nv, ng, n_orb, l_host, kmax, precision, nullptr, &bytes) !=
        AX_OK)
        return -1;
    return bytes;
}

ax_status ax_rk_tensor_destroy(ax_rk_tensor *plan) {
    if (!plan) return AX_OK;
    // the blocks return to the cache: wait until neither the set-up upload
    // nor a launch can still use them
    if (plan->ready) (void)hipEventSynchronize(plan->ready->ev);
    if (plan->captured)
        (void)hipDeviceSynchronize();
    else if (plan->done)
        (void)hipEventSynchronize(plan->done->ev);
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

int ax_rk_tensor_slices(const ax_rk_tensor *plan) { return plan ? plan->slices : -1; }

ax_status ax_rk_tensor_run_stage(ax_rk_tensor *plan, int stage, const double *P,
                                 const double *Q, double *out, void *stream) {
    AX_REQUIRE(plan != nullptr, "ax_rk_tensor_run_stage: null plan");
    AX_REQUIRE(stage != AX_STAGE_DENSITY || P != nullptr, "ax_rk_tensor_run_stage: null orbitals");
    hipStream_t s = ax::as_stream(stream);
    ax_status st = await_ready(&plan, 1, s);
    if (st != AX_OK) return st;
    st = run_stage_batch(&plan, 1, stage, &P, &Q, &out, s);
    if (st != AX_OK) return st;
    return mark_done(&plan, 1, s);
}

ax_status ax_rk_tensor_run(ax_rk_tensor *plan, const double *P, const double *Q, double *out,
                           void *stream) {
    AX_REQUIRE(plan && P && out, "ax_rk_tensor_run: null argument");
    return ax_rk_tensor_run_batch(&plan, 1, &P, &Q, &out, stream);
}

ax_status ax_rk_tensor_run_batch(ax_rk_tensor *const *plans, int n_plans, const double *const *P,
                                 const double *const *Q, double *const *out, void *stream) {
    AX_REQUIRE(n_plans >= 0, "ax_rk_tensor_run_batch: negative plan count");
    if (n_plans == 0) return AX_OK;
    AX_REQUIRE(plans && P && out, "ax_rk_tensor_run_batch: null argument");
    for (int i = 0; i < n_plans; ++i)
        AX_REQUIRE(plans[i] && P[i] && out[i], "ax_rk_tensor_run_batch: null entry %d", i);
    hipStream_t s = ax::as_stream(stream);
    ax_status wst = await_ready(plans, n_plans, s);
    if (wst != AX_OK) return wst;
    // launch sets of up to 16 plans; a plan listed again starts a new set, so
    // its buffers (densities, Y^k, partials, tile counts) serve one listing
    // at a time, in list order
    for (int b0 = 0, nb = 0; b0 < n_plans; b0 += nb) {
        nb = 0;
        while (b0 + nb < n_plans && nb < ax::kMaxBatch &&
               std::find(plans + b0, plans + b0 + nb, plans[b0 + nb]) == plans + b0 + nb)
            ++nb;
        const double *q[ax::kMaxBatch];
        for (int i = 0; i < nb; ++i) q[i] = Q ? Q[b0 + i] : nullptr;
        for (int stage = AX_STAGE_DENSITY; stage <= AX_STAGE_REDUCE; ++stage) {
            ax_status st = run_stage_batch(plans + b0, nb, stage, P + b0, q, out + b0, s);
            if (st != AX_OK) return st;
        }
        ax_status st = mark_done(plans + b0, nb, s);
        if (st != AX_OK) return st;
    }
    return AX_OK;
}

// ---------------------------------------------------------------------------
// Contraction schedule: block order and K split, set directly or chosen by
// timing every candidate on the plan's own data.  The split fixes the
// summation order of the reduction, so a plan's results are deterministic
// for a given schedule (and the same alone and in a batch).
// ---------------------------------------------------------------------------
ax_status ax_rk_tensor_set_schedule(ax_rk_tensor *plan, int order, int n_split) {
    AX_REQUIRE(plan != nullptr, "ax_rk_tensor_set_schedule: null plan");
    AX_REQUIRE(order >= 0 && order < ax::kOrderCount, "ax_rk_tensor_set_schedule: bad order %d",
               order);
    AX_REQUIRE(n_split >= 1 && n_split <= plan->split_cap,
               "ax_rk_tensor_set_schedule: split %d outside [1, %d]", n_split, plan->split_cap);
    if (plan->done) AX_CHECK_HIP(hipEventSynchronize(plan->done->ev));
    const ax_status st = ensure_tables(plan, n_split);
    if (st != AX_OK) return st;
    plan->host.order = order;
    plan->host.n_split = n_split;
    plan->n_split = n_split;
    apply_schedule(plan);
    return AX_OK;
}

ax_status ax_rk_tensor_get_schedule(const ax_rk_tensor *plan, int *order_host,
                                    int *n_split_host) {
    AX_REQUIRE(plan && order_host && n_split_host, "ax_rk_tensor_get_schedule: null argument");
    *order_host = plan->host.order;
    *n_split_host = plan->host.n_split;
    return AX_OK;
}

ax_status ax_rk_tensor_tune(ax_rk_tensor *plan, const double *P, const double *Q, double *out,
                            int reps, void *stream, float *best_ms_host) {
    AX_REQUIRE(plan && P && out, "ax_rk_tensor_tune: null argument");
    AX_REQUIRE(reps >= 1 && reps <= 100, "ax_rk_tensor_tune: reps must be in [1, 100]");
    hipStream_t s = ax::as_stream(stream);
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    AX_CHECK_HIP(hipStreamIsCapturing(s, &cs));
    AX_REQUIRE(cs == hipStreamCaptureStatusNone, "ax_rk_tensor_tune: stream is capturing");
    ax_status st = ax_rk_tensor_run(plan, P, Q, out, stream);
    for (int split = 1; split <= plan->split_cap && st == AX_OK; ++split)
        st = ensure_tables(plan, split);  // every level's job tables, once
    if (st != AX_OK) return st;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    AX_CHECK_HIP(hipEventCreate(&e0));
    AX_CHECK_HIP(hipEventCreate(&e1));
    const int save_order = plan->host.order, save_split = plan->host.n_split;
    float best = -1.0f;
    int best_order = save_order, best_split = save_split;
    for (int split = 1; split <= plan->split_cap && st == AX_OK; ++split) {
        for (int order = 0; order < ax::kOrderCount && st == AX_OK; ++order) {
            plan->host.order = order;
            plan->host.n_split = split;
            apply_schedule(plan);
            for (int r = 0; r <= reps && st == AX_OK; ++r) {  // r = 0 warms up
                if (r == 1 && hipEventRecord(e0, s) != hipSuccess) st = AX_ERR_HIP;
                if (st == AX_OK)
                    st = run_stage_batch(&plan, 1, AX_STAGE_CONTRACT, &P, &Q, &out, s);
                if (st == AX_OK) st = run_stage_batch(&plan, 1, AX_STAGE_REDUCE, &P, &Q, &out, s);
            }
            float ms = 0.0f;
            if (st == AX_OK && (hipEventRecord(e1, s) != hipSuccess ||
                                hipEventSynchronize(e1) != hipSuccess ||
                                hipEventElapsedTime(&ms, e0, e1) != hipSuccess)) {
                ax::set_error("ax_rk_tensor_tune: event timing failed");
                st = AX_ERR_HIP;
            }
            if (st == AX_OK && (best < 0.0f || ms < best)) {
                best = ms;
                best_order = order;
                best_split = split;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    plan->host.order = st == AX_OK ? best_order : save_order;
    plan->host.n_split = st == AX_OK ? best_split : save_split;
    plan->n_split = plan->host.n_split;
    apply_schedule(plan);
    if (st != AX_OK) return st;
    if (best_ms_host) *best_ms_host = best / reps;
    // leave `out` holding the result of the chosen schedule
    st = run_stage_batch(&plan, 1, AX_STAGE_CONTRACT, &P, &Q, &out, s);
    if (st == AX_OK) st = run_stage_batch(&plan, 1, AX_STAGE_REDUCE, &P, &Q, &out, s);
    if (st == AX_OK) st = mark_done(&plan, 1, s);
    return st;
}

// ---------------------------------------------------------------------------
// Single block: packed upper triangle of D[rows] Y^T.
// ---------------------------------------------------------------------------
namespace {

struct GemmWorkspace {
    int64_t o_seg, o_work, o_jobs, o_grp, o_tj, o_part, o_Ds, o_Ys, o_De, o_Ye, size;
};

GemmWorkspace gemm_workspace(int m, int64_t n, int precision, int n_work, int n_split) {
    Layout L;
    GemmWorkspace g{};
    g.o_seg = L.add(sizeof(ax::RkSeg));
    g.o_work = L.add((int64_t)n_work * sizeof(ax::RkWork));
    g.o_jobs = L.add((int64_t)n_work * n_split * sizeof(ax::RkJob));
    g.o_grp = L.add(((int64_t)n_work * n_split + 1) * sizeof(int));
    g.o_tj = L.add((int64_t)n_work * sizeof(ax::RkTileJobs));
    g.o_part = L.add(ax::rk_gemm_partial_doubles(n_work, n_split) * 8);
    if (precision == AX_PREC_SPLIT) {
        const int64_t npad = (n + ax::kSplitKTile - 1) / ax::kSplitKTile * ax::kSplitKTile;
        const int64_t tiles = npad / ax::kSplitKTile;
        g.o_Ds = L.add((int64_t)ax::kSplitSlices * m * npad);
        g.o_Ys = L.add((int64_t)ax::kSplitSlices * m * npad);
        g.o_De = L.add((int64_t)m * tiles * 4);
        g.o_Ye = L.add((int64_t)m * tiles * 4);
    }
    g.size = L.size;
    return g;
}

}  // namespace

int64_t ax_rk_gemm_workspace_bytes(int m, int64_t n, int precision) {
    if (m <= 0 || n <= 0) return 0;
    std::vector<ax::RkWork> work;
    ax::upper_tiles(m, 0, work);
    const int nw = static_cast<int>(work.size());
    return gemm_workspace(m, n, precision, nw, choose_split(nw, n)).size;
}

}  // extern "C"

namespace {

// The contraction of one block into a workspace of at least
// gemm_workspace(...).size bytes (from the caller, or the library's static
// scratch; ws_bytes < 0: not checked).  Every workspace byte the kernels
// read is written earlier in the same call, so the workspace needs no
// clearing.
ax_status rk_gemm_into(const double *D, int64_t ld_d, const int32_t *rows, const double *Y,
                       int64_t ld_y, int m, int64_t n, int precision, char *ws,
                       int64_t ws_bytes, const char *who, double *out, hipStream_t s) {
    std::vector<ax::RkWork> work;
    ax::upper_tiles(m, 0, work);
    const int nw = static_cast<int>(work.size());
    const int split = choose_split(nw, n);
    const GemmWorkspace g = gemm_workspace(m, n, precision, nw, split);
    AX_REQUIRE_CODE(ws_bytes < 0 || g.size <= ws_bytes, AX_ERR_MEMORY,
                    "%s: m = %d, n = %lld needs %lld bytes of scratch, the maximum is %lld",
                    who, m, static_cast<long long>(n), static_cast<long long>(g.size),
                    static_cast<long long>(ws_bytes));

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
        int8_t *Ds = reinterpret_cast<int8_t *>(ws + g.o_Ds);
        int8_t *Ys = reinterpret_cast<int8_t *>(ws + g.o_Ys);
        int32_t *De = reinterpret_cast<int32_t *>(ws + g.o_De);
        int32_t *Ye = reinterpret_cast<int32_t *>(ws + g.o_Ye);
        // D rows are gathered while slicing, so the sliced operand is identity-indexed
        ax_status st = ax::slice_rows_launch(D, ld_d, rows, m, n, npad, Ds, npad, (int64_t)m * npad,
                                             De, tiles, s);
        if (st != AX_OK) return st;
        st = ax::slice_rows_launch(Y, ld_y, nullptr, m, n, npad, Ys, npad, (int64_t)m * npad, Ye,
                                   tiles, s);
        if (st != AX_OK) return st;
        sg.rows = nullptr;
        // K-tile 0 is contracted in FP64 from the caller's operands
        sg.Y0 = Y;
        sg.y0_ld = ld_y;
        sg.d_rows = rows;
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
    AX_CHECK_HIP(hipMemcpyAsync(ws + g.o_seg, &sg, sizeof(sg), hipMemcpyHostToDevice, s));
    const ax::JobTables jt = ax::make_jobs(work, std::vector<int>{m}, split,
                                           ax::k_space(n, precision == AX_PREC_FP64));
    AX_CHECK_HIP(hipMemcpyAsync(ws + g.o_jobs, jt.tile_fastest.data(),
                                jt.tile_fastest.size() * sizeof(ax::RkJob), hipMemcpyHostToDevice,
                                s));
    std::vector<int> ident(jt.tile_fastest.size() + 1);  // one job per work-group
    for (size_t i = 0; i < ident.size(); ++i) ident[i] = static_cast<int>(i);
    AX_CHECK_HIP(hipMemcpyAsync(ws + g.o_grp, ident.data(), ident.size() * sizeof(int),
                                hipMemcpyHostToDevice, s));
    AX_CHECK_HIP(hipMemcpyAsync(ws + g.o_tj, jt.tiles.data(),
                                jt.tiles.size() * sizeof(ax::RkTileJobs), hipMemcpyHostToDevice,
                                s));
    AX_CHECK_HIP(hipMemcpyAsync(ws + g.o_work, work.data(), work.size() * sizeof(ax::RkWork),
                                hipMemcpyHostToDevice, s));
    ax::PlanHost h{};
    h.n = n;
    h.n_work = nw;
    h.n_split = split;
    h.precision = precision;
    h.order = ax::kOrderLinear;
    h.slices = ax::kSplitSlices;
    h.segs = reinterpret_cast<const ax::RkSeg *>(ws + g.o_seg);
    h.work = reinterpret_cast<const ax::RkWork *>(ws + g.o_work);
    h.jobs = reinterpret_cast<const ax::RkJob *>(ws + g.o_jobs);
    h.groups = reinterpret_cast<const int *>(ws + g.o_grp);
    h.tile_jobs = reinterpret_cast<const ax::RkTileJobs *>(ws + g.o_tj);
    h.n_groups = static_cast<int>(jt.tile_fastest.size());
    h.needs_reduce = split > 1;
    h.partial = reinterpret_cast<double *>(ws + g.o_part);
    const ax::PlanHost *hp = &h;
    ax_status st = ax::contract_batch_launch(&hp, 1, &out, s);
    if (st != AX_OK) return st;
    st = ax::reduce_batch_launch(&hp, 1, &out, s);
    if (st != AX_OK) return st;
    // the host-side descriptors above must outlive the asynchronous copies
    AX_CHECK_HIP(hipStreamSynchronize(s));
    return AX_OK;
}

ax_status check_gemm_args(const char *who, int m, int64_t n, int64_t ld_d, int64_t ld_y,
                          int precision) {
    AX_REQUIRE(m >= 0 && m < 0x8000 * 64 && n > 0 && ld_d >= n && ld_y >= n, "%s: bad sizes",
               who);
    AX_REQUIRE(precision == AX_PREC_FP64 || precision == AX_PREC_SPLIT,
               "%s: unknown precision %d", who, precision);
    return AX_OK;
}

// Scratch of ax_rk_gemm_noalloc: one copy per device, held until the process
// ends.  Host code uses the per-device address from hipGetSymbolAddress,
// never the array's name.
static __device__ __align__(256) char scratch[AX_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
char *g_scratch_ptr[kMaxDevices] = {};

ax_status scratch_for_current_device(char **ptr) {
    int dev = -1;
    AX_CHECK_HIP(hipGetDevice(&dev));
    AX_REQUIRE(dev >= 0 && dev < kMaxDevices, "ax_rk_gemm_noalloc: device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[dev] == nullptr) {
        void *p = nullptr;
        AX_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        AX_REQUIRE_CODE(p != nullptr, AX_ERR_HIP,
                        "ax_rk_gemm_noalloc: no scratch address on device %d", dev);
        g_scratch_ptr[dev] = static_cast<char *>(p);
    }
    *ptr = g_scratch_ptr[dev];
    return AX_OK;
}

}  // namespace

extern "C" {

ax_status ax_rk_gemm(const double *D, int64_t ld_d, const int32_t *rows, const double *Y,
                     int64_t ld_y, int m, int64_t n, int precision, void *workspace, double *out,
                     void *stream) {
    ax_status st = check_gemm_args("ax_rk_gemm", m, n, ld_d, ld_y, precision);
    if (st != AX_OK || m == 0) return st;
    AX_REQUIRE(workspace != nullptr && out != nullptr, "ax_rk_gemm: workspace and output required");
    return rk_gemm_into(D, ld_d, rows, Y, ld_y, m, n, precision, static_cast<char *>(workspace),
                        -1, "ax_rk_gemm", out, ax::as_stream(stream));
}

ax_status ax_rk_gemm_noalloc(const double *D, int64_t ld_d, const int32_t *rows, const double *Y,
                             int64_t ld_y, int m, int64_t n, int precision, double *out,
                             void *stream) {
    ax_status st = check_gemm_args("ax_rk_gemm_noalloc", m, n, ld_d, ld_y, precision);
    if (st != AX_OK || m == 0) return st;
    AX_REQUIRE(out != nullptr, "ax_rk_gemm_noalloc: output required");
    char *ws = nullptr;
    st = scratch_for_current_device(&ws);
    if (st != AX_OK) return st;
    return rk_gemm_into(D, ld_d, rows, Y, ld_y, m, n, precision, ws, AX_NOALLOC_SCRATCH_BYTES,
                        "ax_rk_gemm_noalloc", out, ax::as_stream(stream));
}

}  // extern "C"
