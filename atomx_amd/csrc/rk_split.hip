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
typedef const double __attribute__((address_space(1))) gdouble;
__device__ inline v4i load16(const int8_t *p) { return *(gv4i *)p; }

// ---------------------------------------------------------------------------
// Contraction of one job (a K-tile range of a 64x64 output tile).
//
// LDS holds two stages (double buffer), each the S slice planes of the
// tile's 64 D rows and 64 Y^k rows for one K-tile: plane = 64 rows x 64
// bytes, the four 16-byte chunks of row r stored at chunk position
// q ^ ((r >> 1) & 3), so the 16 rows one MFMA operand read touches spread
// over all banks without padding.  Per K-tile: the staged registers of
// K-tile t go to stage t % 2, one barrier, the loads of K-tile t + 1 are
// issued, and the waves contract K-tile t from LDS (the barrier of step t
// also guarantees that every wave has finished reading stage (t + 1) % 2 in
// step t - 1).  (Loading two K-tiles ahead, in two register sets, measured
// no faster.)
//
// Eight waves own the tile's output blocks, a 32x16 strip (2x1 blocks of
// 16x16) each (one work-group per CU: two waves per SIMD); a wave reads its
// row blocks' and column block's slices once per K-tile and reuses them
// across its blocks, whose MFMA chains it interleaves when all are active.
// Blocks wholly below the diagonal or past the segment's last row or
// column are skipped.  (Four waves of 32x32 quadrants hold one wave per
// SIMD and measured slower: 60-70 against 50 us on the benchmark tensor.)
//
// Per block and K-tile the slice products of level L = s + s' (L <= S + 1)
// are summed exactly on v_mfma_i32_16x16x64_i8, two adjacent levels in one
// int32 accumulator (c_lo * 2^7 + c_hi: at most (S - 1) * 64 * 127^2 * 2^7 +
// S * 64 * 127^2 < 2^31 for S <= 8), converted to FP64 and combined by
// Horner in 2^-14, scaled by 2^(eD + eY) and accumulated in FP64.
//
// K-tile 0 (nodes 0..63, the cell from r = 0, where the powers of r inside
// one tile span more binary orders than the slices carry) is contracted in
// FP64 on v_mfma_f64_16x16x4_f64 from the FP64 densities and the FP64 first
// tile of the Y^k rows, with operands loaded straight into registers.  That
// MFMA leaves in lane (g, c) the rows g + 4r of its A operand; feeding A's
// row 4(m % 4) + m / 4 as MFMA row m puts output row 4g + r there, the row
// the int8 MFMA leaves in the same accumulator slot, so both accumulate
// into the same registers.
// ---------------------------------------------------------------------------
constexpr int PLANE = BM * KT;  // bytes per slice plane of one operand

__device__ inline int swz(int row, int q) { return row * KT + ((q ^ ((row >> 1) & 3)) << 4); }

template <int S, int WAVES>
struct SplitCfg {
    static constexpr int kThreads = 64 * WAVES;
    static constexpr int kPer = 2 * S * 256 / kThreads;  // staged 16-byte chunks per thread
    static constexpr int kStage = 2 * S * PLANE;         // slice bytes per stage
    static constexpr int BR = 2;                         // row blocks per wave
    static constexpr int BC = WAVES == 4 ? 2 : 1;        // column blocks per wave
    static constexpr int kLds = 2 * kStage + 2 * 2 * BM * (int)sizeof(int);
};

// Sum of the slice products of level L into c (s1 + s2 = L, 1 <= s <= S).
template <int S, int L>
__device__ inline v4i level_sum(const v4i (&av)[S], const v4i (&bv)[S], v4i c) {
#pragma unroll
    for (int s1 = 1; s1 <= S; ++s1) {
        const int s2 = L - s1;
        if (s2 < 1 || s2 > S) continue;
        c = __builtin_amdgcn_mfma_i32_16x16x64_i8(av[s1 - 1], bv[s2 - 1], c, 0, 0, 0);
    }
    return c;
}

// Levels [2, HI] of one block, paired from the top: (HI - 1, HI), then
// (HI - 3, HI - 2), ...; a pair is one int32 sum c_lo * 2^7 + c_hi in units
// of 2^{-7 (hi - 2)}, and Horner steps between pairs are 2^-14 (also down to
// a lone level 2, which follows the pair (3, 4)).  Returns in lvl
// sum_L c_L 2^{-7 (L - LOW)}: LOW = 2 for an odd count of levels (level 2
// alone last), 3 otherwise (pairs down to (2, 3)).
template <int S, int HI>
struct Levels {
    static __device__ inline void run(const v4i (&av)[S], const v4i (&bv)[S], double (&lvl)[4]) {
        const v4i zero = {0, 0, 0, 0};
        if constexpr (HI >= 3) {
            v4i c = level_sum<S, HI - 1>(av, bv, zero);
            c = c * 128;
            c = level_sum<S, HI>(av, bv, c);
            if constexpr (HI == S + 1) {
#pragma unroll
                for (int r = 0; r < 4; ++r) lvl[r] = static_cast<double>(c[r]);
            } else {
#pragma unroll
                for (int r = 0; r < 4; ++r) lvl[r] = fma(lvl[r], 0x1p-14, static_cast<double>(c[r]));
            }
            Levels<S, HI - 2>::run(av, bv, lvl);
        } else if constexpr (HI == 2) {
            const v4i c = level_sum<S, 2>(av, bv, zero);
            if constexpr (HI == S + 1) {
#pragma unroll
                for (int r = 0; r < 4; ++r) lvl[r] = static_cast<double>(c[r]);
            } else {
#pragma unroll
                for (int r = 0; r < 4; ++r) lvl[r] = fma(lvl[r], 0x1p-14, static_cast<double>(c[r]));
            }
        }
    }
};
// The same for all BR x BC blocks of a wave at once: every MFMA step is
// issued for every block before the next step, so the blocks' dependent
// chains (MFMA -> shift -> MFMA -> convert) interleave.
template <int S, int L, int BR, int BC>
__device__ inline void level_sum_all(const v4i (&av)[BR][S], const v4i (&bv)[BC][S],
                                     v4i (&c)[BR][BC]) {
#pragma unroll
    for (int s1 = 1; s1 <= S; ++s1) {
        const int s2 = L - s1;
        if (s2 < 1 || s2 > S) continue;
#pragma unroll
        for (int a = 0; a < BR; ++a)
#pragma unroll
            for (int b = 0; b < BC; ++b)
                c[a][b] = __builtin_amdgcn_mfma_i32_16x16x64_i8(av[a][s1 - 1], bv[b][s2 - 1],
                                                               c[a][b], 0, 0, 0);
    }
}

template <int S, int HI, int BR, int BC>
struct LevelsAll {
    static __device__ inline void run(const v4i (&av)[BR][S], const v4i (&bv)[BC][S],
                                      double (&lvl)[BR][BC][4]) {
        if constexpr (HI >= 2) {
            v4i c[BR][BC];
#pragma unroll
            for (int a = 0; a < BR; ++a)
#pragma unroll
                for (int b = 0; b < BC; ++b) c[a][b] = v4i{0, 0, 0, 0};
            if constexpr (HI >= 3) {
                level_sum_all<S, HI - 1, BR, BC>(av, bv, c);
#pragma unroll
                for (int a = 0; a < BR; ++a)
#pragma unroll
                    for (int b = 0; b < BC; ++b) c[a][b] = c[a][b] * 128;
            }
            level_sum_all<S, HI, BR, BC>(av, bv, c);
            constexpr double step = 0x1p-14;  // between pairs, and down to a lone level 2
#pragma unroll
            for (int a = 0; a < BR; ++a)
#pragma unroll
                for (int b = 0; b < BC; ++b)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        lvl[a][b][r] = HI == S + 1 ? static_cast<double>(c[a][b][r])
                                                   : fma(lvl[a][b][r], step,
                                                         static_cast<double>(c[a][b][r]));
            if constexpr (HI >= 4) LevelsAll<S, HI - 2, BR, BC>::run(av, bv, lvl);
        }
    }
};

// binary exponent of the unit of lvl relative to 2^(eD + eY - 14): S levels
// (2 .. S + 1); odd S ends on the lone level 2, even S on the pair (2, 3)
template <int S>
constexpr int levels_shift() { return S % 2 ? 0 : -7; }

template <int S, int WAVES>
__device__ inline void split_job(const ax::BatchOf<ax::GemmArg> &A, const ax::JobInfo &J,
                                 int8_t *__restrict__ lds, int *__restrict__ elds) {
    using C = SplitCfg<S, WAVES>;
    const ax::GemmArg &g = A.a[J.p];
    const ax::RkSeg sg = g.segs[J.seg];
    const int64_t tb = J.kb, te = J.ke;  // K-tiles [tb, te)
    const int row0 = J.ti * BM, col0 = J.tj * BM;
    const int m = sg.m;

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    const int wr = WAVES == 4 ? (wave >> 1) * 32 : (wave >> 2) * 32;
    const int wc = WAVES == 4 ? (wave & 1) * 32 : (wave & 3) * 16;
    const int fr = lane & 15, fg = lane >> 4;

    // active blocks (wave-uniform): inside the segment and not wholly below
    // the diagonal of a diagonal tile
    bool act[C::BR][C::BC];
    bool any = false, all = true;
#pragma unroll
    for (int a = 0; a < C::BR; ++a)
#pragma unroll
        for (int b = 0; b < C::BC; ++b) {
            const int i0 = row0 + wr + 16 * a, j0 = col0 + wc + 16 * b;
            act[a][b] = i0 < m && j0 < m && j0 + 15 >= i0;
            any = any || act[a][b];
            all = all && act[a][b];
        }

    d4 acc[C::BR][C::BC];
#pragma unroll
    for (int a = 0; a < C::BR; ++a)
#pragma unroll
        for (int b = 0; b < C::BC; ++b) acc[a][b] = d4{0.0, 0.0, 0.0, 0.0};

    int64_t t0 = tb;
    if (tb == 0 && te > 0 && sg.Y0 != nullptr) {
        if (any) {
            const int64_t nk = sg.n < KT ? sg.n : KT;
            const int prow = 4 * (fr & 3) + (fr >> 2);  // A row fed as MFMA row fr
            gdouble *pa[C::BR];
            gdouble *pb[C::BC];
            bool oka[C::BR], okb[C::BC];
#pragma unroll
            for (int a = 0; a < C::BR; ++a) {
                const int i = row0 + wr + 16 * a + prow;
                oka[a] = i < m;
                const int64_t dr = oka[a] ? (sg.d_rows ? (int64_t)sg.d_rows[i] : (int64_t)i) : 0;
                pa[a] = (gdouble *)(sg.D + dr * sg.ld_d);
            }
#pragma unroll
            for (int b = 0; b < C::BC; ++b) {
                const int j = col0 + wc + 16 * b + fr;
                okb[b] = j < m;
                pb[b] = (gdouble *)(sg.Y0 + (okb[b] ? (int64_t)j : 0) * sg.y0_ld);
            }
#pragma unroll 4
            for (int kk = 0; kk < KT / 4; ++kk) {
                const int64_t x = 4 * kk + fg;
                double av0[C::BR], bv0[C::BC];
#pragma unroll
                for (int a = 0; a < C::BR; ++a) av0[a] = (oka[a] && x < nk) ? pa[a][x] : 0.0;
#pragma unroll
                for (int b = 0; b < C::BC; ++b) bv0[b] = (okb[b] && x < nk) ? pb[b][x] : 0.0;
#pragma unroll
                for (int a = 0; a < C::BR; ++a)
#pragma unroll
                    for (int b = 0; b < C::BC; ++b)
                        if (act[a][b])
                            acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av0[a], bv0[b],
                                                                             acc[a][b], 0, 0, 0);
            }
        }
        t0 = 1;
    }

    // staging: chunk c = tid + i * kThreads of the K-tile, c = (op*S + s)*256
    // + row*4 + q: this thread's row and chunk are fixed, its planes vary
    const int srow = (tid & 255) >> 2, sq = tid & 3;
    bool va, vb;
    const int8_t *sa, *sb;
    {
        const int i = row0 + srow, j = col0 + srow;
        va = i < m;
        vb = j < m;
        const int64_t ra = va ? (sg.rows ? (int64_t)sg.rows[i] : (int64_t)i) : 0;
        sa = sg.Ds + ra * sg.ds_ld + sq * 16;
        sb = sg.Ys + (vb ? (int64_t)j : 0) * sg.ys_ld + sq * 16;
    }
    const int half = tid >> 8;  // 0 for 4 waves; 0 / 1 for 8 waves
    // exponent slot: threads 0..63 the tile's D rows, 64..127 its Y^k rows
    int64_t erow = 0;
    bool eok = false;
    if (tid < BM) {
        const int i = row0 + tid;
        eok = i < m;
        erow = eok ? (sg.rows ? (int64_t)sg.rows[i] : (int64_t)i) : 0;
    } else if (tid < 2 * BM) {
        const int j = col0 + tid - BM;
        eok = j < m;
        erow = eok ? (int64_t)j : 0;
    }
    gint *ep = tid < BM ? (gint *)(sg.De + erow * sg.de_ld) : (gint *)(sg.Ye + erow * sg.ye_ld);
    const bool has_e = tid < 2 * BM;

    auto load_tile = [&](int64_t t, v4i (&reg)[C::kPer], int &rexp) {
#pragma unroll
        for (int i = 0; i < C::kPer; ++i) {
            const int idx = i * (C::kThreads / 256) + half;
            const int op = idx / S, s2 = idx - (idx / S) * S;
            reg[i] = op == 0 ? load16(sa + s2 * sg.ds_plane + t * KT)
                             : load16(sb + s2 * sg.ys_plane + t * KT);
        }
        if (has_e) rexp = ep[t];
    };
    auto store_tile = [&](int stage, const v4i (&reg)[C::kPer], int rexp) {
        int8_t *base = lds + stage * C::kStage;
        const v4i zero = {0, 0, 0, 0};
#pragma unroll
        for (int i = 0; i < C::kPer; ++i) {
            const int idx = i * (C::kThreads / 256) + half;
            const int op = idx / S;
            const bool ok = op == 0 ? va : vb;
            *reinterpret_cast<v4i *>(base + idx * PLANE + swz(srow, sq)) = ok ? reg[i] : zero;
        }
        if (tid < 2 * BM) elds[stage * 2 * BM + tid] = eok ? rexp : 0;
    };
    auto contract = [&](int stage) {
        const int8_t *As = lds + stage * C::kStage;
        const int8_t *Bs = As + S * PLANE;
        const int *eA = elds + stage * 2 * BM;
        const int *eB = eA + BM;
        v4i av[C::BR][S], bv[C::BC][S];
#pragma unroll
        for (int a = 0; a < C::BR; ++a)
#pragma unroll
            for (int s2 = 0; s2 < S; ++s2)
                av[a][s2] = *reinterpret_cast<const v4i *>(As + s2 * PLANE + swz(wr + 16 * a + fr, fg));
#pragma unroll
        for (int b = 0; b < C::BC; ++b)
#pragma unroll
            for (int s2 = 0; s2 < S; ++s2)
                bv[b][s2] = *reinterpret_cast<const v4i *>(Bs + s2 * PLANE + swz(wc + 16 * b + fr, fg));
        int ea[C::BR][4], eb[C::BC];
#pragma unroll
        for (int a = 0; a < C::BR; ++a)
#pragma unroll
            for (int r = 0; r < 4; ++r) ea[a][r] = eA[wr + 16 * a + 4 * fg + r];
#pragma unroll
        for (int b = 0; b < C::BC; ++b) eb[b] = eB[wc + 16 * b + fr] - 14 + levels_shift<S>();
        if (all) {  // every block of the wave: interleaved chains
            double lvl[C::BR][C::BC][4];
            LevelsAll<S, S + 1, C::BR, C::BC>::run(av, bv, lvl);
#pragma unroll
            for (int a = 0; a < C::BR; ++a)
#pragma unroll
                for (int b = 0; b < C::BC; ++b)
#pragma unroll
                    for (int r = 0; r < 4; ++r)
                        acc[a][b][r] += ldexp(lvl[a][b][r], ea[a][r] + eb[b]);
            return;
        }
#pragma unroll
        for (int a = 0; a < C::BR; ++a) {
#pragma unroll
            for (int b = 0; b < C::BC; ++b) {
                if (!act[a][b]) continue;
                double lvl[4] = {0.0, 0.0, 0.0, 0.0};
                Levels<S, S + 1>::run(av[a], bv[b], lvl);
#pragma unroll
                for (int r = 0; r < 4; ++r) acc[a][b][r] += ldexp(lvl[r], ea[a][r] + eb[b]);
            }
        }
    };

    const int64_t nt = te - t0;
    v4i reg[C::kPer];
    int rexp = 0;
    if (nt > 0) load_tile(t0, reg, rexp);
    for (int64_t it = 0; it < nt; ++it) {
        const int stage = static_cast<int>(it & 1);
        store_tile(stage, reg, rexp);
        __syncthreads();
        if (it + 1 < nt) load_tile(t0 + it + 1, reg, rexp);
        if (any) contract(stage);
    }
    __syncthreads();  // the next job of this work-group restages both buffers

    if (J.count == 1) {  // a single chunk: the packed output directly
        double *out = g.out + sg.out_off;
#pragma unroll
        for (int a = 0; a < C::BR; ++a)
#pragma unroll
            for (int b = 0; b < C::BC; ++b)
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
    for (int a = 0; a < C::BR; ++a)
#pragma unroll
        for (int b = 0; b < C::BC; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = wr + a * 16 + 4 * fg + r;
                const int j = wc + b * 16 + fr;
                dst[i * BM + j] = acc[a][b][r];
            }
}

template <int S, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void rk_gemm_split_kernel(const ax::BatchOf<ax::GemmArg> A) {
    extern __shared__ __attribute__((aligned(16))) int8_t lds_split[];
    int *elds = reinterpret_cast<int *>(lds_split + 2 * SplitCfg<S, WAVES>::kStage);
    const ax::GroupInfo G = ax::gemm_group(A);
    for (int j = G.j0; j < G.j1; ++j)
        split_job<S, WAVES>(A, ax::gemm_job(A, G.p, j), lds_split, elds);
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

template <int S, int WAVES>
ax_status launch_split(const BatchOf<GemmArg> &A, dim3 grid, hipStream_t stream) {
    const int lds = SplitCfg<S, WAVES>::kLds;
    static bool attr = false;  // once per instantiation (the size is fixed)
    if (!attr) {
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(rk_gemm_split_kernel<S, WAVES>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, lds);
        attr = true;
    }
    hipLaunchKernelGGL((rk_gemm_split_kernel<S, WAVES>), grid, dim3(64 * WAVES), lds, stream, A);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

ax_status rk_gemm_split_launch(const BatchOf<GemmArg> &A, int slices, dim3 grid,
                               hipStream_t stream) {
    if (slices == kSplitSlices) return launch_split<kSplitSlices, 8>(A, grid, stream);
    if (slices == kSplitSlicesFine) return launch_split<kSplitSlicesFine, 8>(A, grid, stream);
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
