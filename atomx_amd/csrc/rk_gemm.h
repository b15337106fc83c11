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
    // split precision: K-tile 0 (nodes 0..63, from r = 0, where r^k spans
    // more binary orders inside the tile than the slices carry) is
    // contracted in FP64 from D rows d_rows[i] (nullptr = identity) and the
    // FP64 first K-tile of the Y^k rows, Y0 + j*y0_ld (nullptr: no FP64 tile)
    const double *Y0;
    int64_t y0_ld;
    const int32_t *d_rows;
};

struct RkWork {
    int seg;
    int ti;
    int tj;
    int pad;
};

// One contraction job: tile (ti, tj) of segment `seg` over the K range
// [kb, ke) (nodes for the FP64 kernel, 64-node K-tiles for the split kernel),
// piece `split` of the tile's `count` pieces in K order, written to partial
// slot `slot` (or, for a tile of one piece, to the packed output).
struct RkJob {
    int seg;
    int tile;   // ti | tj << 16
    int split;  // split | count << 16
    int slot;
    int kb, ke;
};
// The partial slots [first, first + count) of one tile, summed in order.
struct RkTileJobs {
    int first;
    int count;
};

constexpr int kSplitSlices = 7;      // int8 slices per operand (7 bits each)
constexpr int kSplitSlicesFine = 8;  // for grids whose K-tiles span a wide radius range
constexpr int kSplitKTile = 64;   // K-tile that shares one exponent
constexpr int kMaxBatch = 16;     // plans per batched launch

// Per-plan arguments of the batched kernels, passed by value in the launch
// table (kernel arguments), so a work-group reaches its plan's arrays without
// a chain of dependent loads.
struct DensArg {             // weighted pair densities (+ their slices)
    const double *P, *Q;     // orbital samples (n_orb, ld); Q may be null
    int64_t ld, n;
    const int32_t *pa, *pb;
    const double *w;
    double *D;               // (n_pairs, n)
    int8_t *Ds;              // split precision: slice planes of D, else null
    int32_t *De;
    int64_t npad, ds_plane;
    int slices, chunks;      // chunks: 512-node blocks of one pair row
};
struct ScanArg {             // Y^k items of one plan
    const double *D;
    double *Y;               // (n_yrows, n) FP64 rows, or null when sliced here
    const double *rk, *rkm;
    const int32_t *item_row, *item_k, *item_out;
    int64_t n;
    int nv, kmax;
    int8_t *Ys;              // slices written by the scan (n_gauss = 8, split)
    int32_t *Ye;
    int64_t npad, ys_plane;
    int slices, pad;
    double *Y0;              // FP64 first K-tile of the sliced rows (kSplitKTile per row)
};
struct SliceArg {            // rows sliced by a separate pass
    const double *src;
    int64_t n, npad, plane;
    int8_t *dst;
    int32_t *e;
    int slices, pad;
};
struct GemmArg {             // contraction work-groups and split reduction
    const RkSeg *segs;
    const RkWork *work;      // tiles (reduction)
    const RkJob *jobs;       // contraction jobs, in the schedule's order
    const int *groups;       // work-group w runs jobs [groups[w], groups[w + 1])
    const RkTileJobs *tile_jobs;
    double *partial;         // one 64*64 tile per job slot
    double *out;             // packed R^k output: a tile contracted in one K
                             // chunk is written here by the contraction, the
                             // others by rk_reduce_kernel
};

// Launch table: end[] is the exclusive prefix of the launch's items per plan
// (pairs, scan items, slice tiles, contraction work-groups or tiles).
template <typename T>
struct BatchOf {
    int n;
    int order;      // contraction launches: block order (kOrder*)
    int pad0, pad1;
    int *counter;   // dynamic contraction order: job ticket, finished work-groups
    int64_t end[kMaxBatch];
    T a[kMaxBatch];
};

// Block orders of a contraction launch over its jobs (results do not depend
// on the order, only the L2 locality does):
enum {
    kOrderLinear = 0,    // block b -> job b of the tile-fastest list
    kOrderXcd = 1,       // same, after the XCD-contiguous renumbering
    kOrderXcdSplit = 2,  // XCD-contiguous over the split-fastest list
    kOrderDynamic = 3,   // FP64: resident work-groups take jobs heaviest first from a ticket
    kOrderBalanced = 4,  // 64 x S work-groups: per XCD one K slice, its tiles
                         // heaviest first to the least-loaded work-group
    kOrderCount = 5
};

// Work-group renumbering for L2 locality: the dispatcher deals consecutive
// work-groups round-robin to the 8 XCDs, each with its own L2.  Mapping
// hardware block b of nb to this logical index gives every XCD one
// contiguous range of logical blocks (a bijection on [0, nb) for any nb, so
// correctness never depends on the dispatch order).
__device__ inline int64_t xcd_block(int64_t b, int64_t nb) {
    constexpr int64_t X = 8;
    const int64_t xcd = b % X, idx = b / X, q = nb / X, r = nb % X;
    return xcd < r ? xcd * (q + 1) + idx : r * (q + 1) + (xcd - r) * q + idx;
}

// Plan of item `it` (uniform across the caller's wave) and the item's index
// inside that plan.
template <typename T>
__device__ inline int batch_find(const BatchOf<T> &B, int64_t it, int64_t &local) {
    int p = 0;
    while (p + 1 < B.n && it >= B.end[p]) ++p;
    local = it - (p ? B.end[p - 1] : 0);
    return p;
}

// The work-group of a static contraction launch under its block order:
// plan p of the batch and its jobs [j0, j1).  gb >= 0 names the work-group
// directly (dynamic order).
struct GroupInfo {
    int p, j0, j1;
};
__device__ inline GroupInfo gemm_group(const BatchOf<GemmArg> &A, int64_t gb = -1) {
    const int64_t nb = gridDim.x, b = blockIdx.x;
    const int64_t lb = gb >= 0 ? gb : A.order == kOrderLinear ? b : xcd_block(b, nb);
    int64_t local;
    GroupInfo G;
    G.p = batch_find(A, lb, local);
    const int *grp = A.a[G.p].groups;
    G.j0 = grp[local];
    G.j1 = grp[local + 1];
    return G;
}

struct JobInfo {
    int p;  // plan in the batch
    int seg, ti, tj, split, count, slot, kb, ke;
};
__device__ inline JobInfo gemm_job(const BatchOf<GemmArg> &A, int p, int jb) {
    const RkJob j = A.a[p].jobs[jb];
    JobInfo J;
    J.p = p;
    J.seg = j.seg;
    J.ti = j.tile & 0xffff;
    J.tj = static_cast<int>(static_cast<unsigned>(j.tile) >> 16);
    J.split = j.split & 0xffff;
    J.count = static_cast<int>(static_cast<unsigned>(j.split) >> 16);
    J.slot = j.slot;
    J.kb = j.kb;
    J.ke = j.ke;
    return J;
}

// Split-precision slicing of one value per lane of a wave whose 64 lanes
// hold one 64-wide K-tile (lane l = element t*64 + l): power-of-two exponent
// e with max|v| < 2^e over the tile, then `slices` signed 7-bit slices
// of v*2^-e by truncation, plane s at dst_row[s*plane + x]; lane 0 writes e.
// Every lane of the wave must call it (it reduces across the wave).
__device__ inline void slice_value(double v, int64_t x, int8_t *__restrict__ dst_row,
                                   int64_t plane, int32_t *__restrict__ e_slot, int slices) {
    double mx = fabs(v);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
    const int ex = mx > 0.0 ? ilogb(mx) + 1 : 0;
    double rho = ldexp(v, 7 - ex);  // v * 2^-e * 2^7, |rho| < 128
    int8_t *d = dst_row + x;
    for (int s = 0; s < slices; ++s) {
        const double c = trunc(rho);
        d[(int64_t)s * plane] = static_cast<int8_t>(c);
        rho = (rho - c) * 128.0;
    }
    if ((threadIdx.x & 63) == 0) *e_slot = ex;
}

// One wave slices K-tile t of a row of n values (zero beyond n).
__device__ inline void slice_tile(const double *__restrict__ src_row, int64_t n, int64_t t,
                                  int8_t *__restrict__ dst_row, int64_t dst_plane,
                                  int32_t *__restrict__ e_slot, int slices) {
    const int64_t x = t * kSplitKTile + (threadIdx.x & 63);
    slice_value(x < n ? src_row[x] : 0.0, x, dst_row, dst_plane, e_slot, slices);
}

// Packed upper-triangle index of (i, j), i <= j < m.
__device__ inline int64_t packed_index(int64_t i, int64_t j, int64_t m) {
    return i * m - i * (i - 1) / 2 + (j - i);
}

int gemm_tile();
int64_t rk_gemm_partial_doubles(int n_work, int n_split);
void upper_tiles(int m, int seg, std::vector<RkWork> &out);

// K extent of a contraction's jobs: k_len units (nodes for the FP64 kernel,
// 64-node K-tiles for the split kernel), chunk boundaries on multiples of
// k_align units.
struct KSpace {
    int64_t k_len;
    int k_align;
    bool fp64;  // FP64 kernel: a tile's cost is its 16x16 blocks inside the triangle
};
KSpace k_space(int64_t n, bool fp64);

// Jobs of a contraction cut into S K chunks per tile (one job per
// work-group), listed tile-fastest, split-fastest and heaviest first (most
// 16x16 blocks inside the triangle of segment rows seg_m); the slots (and the
// reduction order) are tile-major and do not depend on the list.
struct JobTables {
    std::vector<RkJob> tile_fastest, split_fastest, heaviest;
    std::vector<RkTileJobs> tiles;
};
// all = false: the tile-fastest list and the slot ranges only (the default
// order's tables; the others are built when a schedule asks for them).
JobTables make_jobs(const std::vector<RkWork> &work, const std::vector<int> &seg_m, int S,
                    const KSpace &ks, bool all = true);

// Balanced jobs: the K range is cut into one slice per XCD (at most 8);
// each slice's tiles, heaviest first (cost = 16x16 blocks inside the
// triangle for the FP64 kernel), go whole to the least-loaded of the
// slice's n_groups / slices work-groups, listed slice by slice (so the
// XCD-contiguous block order gives each XCD one slice and dispatches its
// heavy tiles first).  A tile's pieces in K order are its slots.  groups
// has one entry per work-group plus one; work-groups beyond the slice's
// tile count stay empty.  (Cutting the slices into equal-weight ranges of
// tile pieces instead measured 1.6x slower: many short pieces per
// work-group.)
struct BalancedJobs {
    std::vector<RkJob> jobs;
    std::vector<int> groups;
    std::vector<RkTileJobs> tiles;
};
BalancedJobs make_balanced_jobs(const std::vector<RkWork> &work, const std::vector<int> &seg_m,
                                int n_groups, const KSpace &ks);
// Work-groups of the balanced order at level S.
int balanced_groups(int S);

// Host copy of a plan: sizes and the device pointers the launchers need.
struct PlanHost {
    int64_t n, npad;
    int nv, ng, kmax, n_pairs, n_items, n_yrows, n_work, n_split, precision;
    int order;              // contraction block order (kOrder*)
    int slices;             // split precision: int8 slices per operand
    const double *mo, *mi;  // partial matrices of ng (device)
    const double *w, *rk, *rkm;
    const int32_t *pa, *pb, *item_row, *item_k, *item_out;
    double *D, *Y;          // Y null when the scan writes slices directly
    double *Y0;             // then the FP64 first K-tile of each Y^k row (split)
    int8_t *Ds, *Ys;
    int32_t *De, *Ye;
    const RkSeg *segs;
    const RkWork *work;
    const RkJob *jobs;      // the schedule's job list (order and K-split level)
    const int *groups;      // jobs of each work-group (GemmArg::groups)
    const RkTileJobs *tile_jobs;
    int n_groups;           // work-groups of a static launch
    int needs_reduce;       // some tile has more than one K piece
    int *counter;           // two ints, zero between launches (dynamic order)
    double *partial;
};
// Whether the scan of a split-precision plan on n_gauss nodes per cell writes
// the int8 slices of Y^k itself (one wave covers exactly one 64-node K-tile).
bool scan_slices_fused(int ng, int nv);

ax_status density_batch_launch(const PlanHost *const *plans, int n, const double *const *P,
                               const double *const *Q, hipStream_t stream);
ax_status scan_batch_launch(const PlanHost *const *plans, int n, hipStream_t stream);
ax_status slice_batch_launch(const PlanHost *const *plans, int n, hipStream_t stream);
// out: the plans' packed outputs (tiles contracted in one K chunk are
// written by the contraction itself).
ax_status contract_batch_launch(const PlanHost *const *plans, int n, double *const *out,
                                hipStream_t stream);
ax_status reduce_batch_launch(const PlanHost *const *plans, int n, double *const *out,
                              hipStream_t stream);
ax_status rk_gemm_split_launch(const BatchOf<GemmArg> &A, int slices, dim3 grid,
                               hipStream_t stream);

// Slices n_rows rows of src (row r at src + rows[r]*ld, or r*ld) into
// kSplitSlices int8 planes dst[s*plane + r*dst_ld + x] (x < npad, zero
// padded beyond n) and exponents e[r*e_ld + t] per 64-wide K-tile.
ax_status slice_rows_launch(const double *src, int64_t ld, const int32_t *rows, int n_rows,
                            int64_t n, int64_t npad, int8_t *dst, int64_t dst_ld,
                            int64_t dst_plane, int32_t *e, int64_t e_ld, hipStream_t stream);

}  // namespace ax
