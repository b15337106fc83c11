// (Y) Poisson discretisation on the Gauss grid: pair densities, r-power
// tables, the cell-blocked Y^k scan, quartet contraction, B-spline orbital
// evaluation and the Y^k-form exchange matrix.
//
// Data layout in HBM: every radial function is one contiguous row of n =
// nv*ng FP64 samples (cell-major, node-minor), rows of a batch stacked with a
// leading dimension.  Kernels map 64-lane wavefronts onto whole cells
// (floor(64/ng) cells per wave) so that every global load of a wave is one
// contiguous run of doubles.
#include "ax_common.h"
#include "rk_gemm.h"

namespace {

constexpr int kThreads = 256;
constexpr int kMaxNg = 16;

// ---------------------------------------------------------------------------
// r^k and r^-(k+1) tables (numerics/yk.py:113-117)
// ---------------------------------------------------------------------------
__global__ void rpow_kernel(const double *__restrict__ r, int64_t n, int kmax,
                            double *__restrict__ rk, double *__restrict__ rkm) {
    int64_t total = n * (kmax + 1);
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        int k = static_cast<int>(t / n);
        int64_t x = t - (int64_t)k * n;
        double rv = r[x];
        rk[t] = pow(rv, static_cast<double>(k));
        rkm[t] = pow(rv, static_cast<double>(-(k + 1)));
    }
}

// ---------------------------------------------------------------------------
// Pair densities rho_p = P_a P_b + Q_a Q_b  (optionally times w)
// ---------------------------------------------------------------------------
__global__ void pair_density_kernel(const double *__restrict__ P, const double *__restrict__ Q,
                                    int64_t ld_pq, int64_t n, int n_pairs,
                                    const int32_t *__restrict__ pa,
                                    const int32_t *__restrict__ pb,
                                    const double *__restrict__ w, double *__restrict__ out,
                                    int64_t ld_out) {
    for (int p = blockIdx.y; p < n_pairs; p += gridDim.y) {
        const double *Pa = P + (int64_t)pa[p] * ld_pq;
        const double *Pb = P + (int64_t)pb[p] * ld_pq;
        const double *Qa = Q ? Q + (int64_t)pa[p] * ld_pq : nullptr;
        const double *Qb = Q ? Q + (int64_t)pb[p] * ld_pq : nullptr;
        double *o = out + (int64_t)p * ld_out;
        for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < n;
             x += (int64_t)gridDim.x * blockDim.x) {
            double v = Pa[x] * Pb[x];
            if (Q) v += Qa[x] * Qb[x];
            if (w) v = w[x] * v;
            o[x] = v;
        }
    }
}

// Batched weighted pair densities D_p = w (P_a P_b + Q_a Q_b) of several plans:
// blockIdx.y walks the concatenated pair list, blockIdx.x the grid points.
// For split-precision plans the same pass writes the int8 slices of D: every
// wave holds one 64-aligned run of x, i.e. one K-tile, so the loop runs to
// the padded length (wave-uniform) with zeros beyond n.
// Slices two consecutive values per lane into the int8 planes of one
// 64-node K-tile held by 32 lanes (half a wave): same arithmetic as
// slice_value (tile exponent from the tile maximum, truncated 7-bit digits).
__device__ inline void slice_pair(double v0, double v1, int64_t x, int8_t *__restrict__ dst_row,
                                  int64_t plane, int32_t *__restrict__ e_slot, int slices) {
    double mx = fmax(fabs(v0), fabs(v1));
#pragma unroll
    for (int off = 16; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
    const int ex = mx > 0.0 ? ilogb(mx) + 1 : 0;
    double r0 = ldexp(v0, 7 - ex), r1 = ldexp(v1, 7 - ex);
    int8_t *d = dst_row + x;
    for (int s = 0; s < slices; ++s) {
        const double c0 = trunc(r0), c1 = trunc(r1);
        const uint16_t word = (uint16_t)(uint8_t)static_cast<int8_t>(c0) |
                              (uint16_t)((uint16_t)(uint8_t)static_cast<int8_t>(c1) << 8);
        *reinterpret_cast<uint16_t *>(d + (int64_t)s * plane) = word;
        r0 = (r0 - c0) * 128.0;
        r1 = (r1 - c1) * 128.0;
    }
    if ((threadIdx.x & 31) == 0) *e_slot = ex;
}

// One work-group per (plan, pair, 512-node chunk) of the batch, taken from
// the prefix of each plan's pairs x chunks, so a plan's rows launch only
// their own chunks (a short row beside a long one in the batch launches no
// empty work-groups).  Each thread forms two consecutive nodes; rows of even
// length with 16-byte aligned bases are read and written as double2.  Split
// plans also slice their rows (npad is a multiple of 64 and a chunk starts
// at a multiple of 512, so every half-wave holds one whole K-tile or lies
// wholly past npad).
__global__ __launch_bounds__(256) void pair_density_batch_kernel(
    const ax::BatchOf<ax::DensArg> A) {
    typedef double d2 __attribute__((ext_vector_type(2)));
    const int64_t total = A.end[A.n - 1];
    for (int64_t gb = blockIdx.x; gb < total; gb += gridDim.x) {
        int64_t local;
        const int i = ax::batch_find(A, gb, local);
        const ax::DensArg &g = A.a[i];
        const int64_t p = local / g.chunks;
        const int64_t x = (local - p * g.chunks) * 512 + (int64_t)threadIdx.x * 2;
        const int64_t n = g.n, ld = g.ld;
        const int64_t lim = g.Ds ? g.npad : n;
        if (x >= lim) continue;  // uniform over each half-wave (see above)
        const int64_t a = g.pa[p], b = g.pb[p];
        const double *Pa = g.P + a * ld;
        const double *Pb = g.P + b * ld;
        const double *Qa = g.Q ? g.Q + a * ld : nullptr;
        const double *Qb = g.Q ? g.Q + b * ld : nullptr;
        double *o = g.D + p * n;
        const bool vec = (n % 2 == 0) && (ld % 2 == 0) &&
                         ((reinterpret_cast<uintptr_t>(g.P) | reinterpret_cast<uintptr_t>(g.w) |
                           reinterpret_cast<uintptr_t>(g.D) |
                           (g.Q ? reinterpret_cast<uintptr_t>(g.Q) : 0)) % 16 == 0);
        double v0 = 0.0, v1 = 0.0;
        if (vec) {
            const int64_t xc = x < n ? x : n - 2;
            const d2 pa = *reinterpret_cast<const d2 *>(Pa + xc);
            const d2 pb = *reinterpret_cast<const d2 *>(Pb + xc);
            const d2 w = *reinterpret_cast<const d2 *>(g.w + xc);
            d2 t = pa * pb;
            if (Qa) {
                const d2 qa = *reinterpret_cast<const d2 *>(Qa + xc);
                const d2 qb = *reinterpret_cast<const d2 *>(Qb + xc);
                t = t + qa * qb;
            }
            t = w * t;
            if (x < n) {
                *reinterpret_cast<d2 *>(o + x) = t;
                v0 = t.x;
                v1 = t.y;
            }
        } else {
#pragma unroll
            for (int u = 0; u < 2; ++u) {
                const int64_t xu = x + u;
                if (xu < n) {
                    double t = Pa[xu] * Pb[xu];
                    if (Qa) t += Qa[xu] * Qb[xu];
                    t = g.w[xu] * t;
                    o[xu] = t;
                    (u ? v1 : v0) = t;
                }
            }
        }
        if (g.Ds)
            slice_pair(v0, v1, x, g.Ds + p * g.npad, g.ds_plane,
                       g.De + p * (g.npad / ax::kSplitKTile) + x / ax::kSplitKTile, g.slices);
    }
}

// ---------------------------------------------------------------------------
// Y^k scan (numerics/yk.py:87-154).
//
// One work-group per (density row, k) item, grid-strided over items.
//   pass 1: per-cell full sums S_out[c] = sum_j d_j r_j^k and
//           S_in[c] = sum_j d_j r_j^-(k+1)  (d = w*rho), into LDS;
//   scan:   exclusive prefix of S_out (outward) and exclusive suffix of S_in
//           (inward) over cells: segmented sequential sums per thread, then a
//           Hillis-Steele scan of the 256 segment totals;
//   pass 2: I_out[i] = base_out[c] + sum_j mo[i,j] g_out_j,
//           I_in[i]  = base_in[c]  + sum_j mi[i,j] g_in_j,
//           Y[i] = r_i^-(k+1) I_out[i] + r_i^k I_in[i].
// The row is read twice (the second read hits L2) and Y written once.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void yk_batch_kernel(
    const double *__restrict__ rho, int64_t ld_rho, const double *__restrict__ w,
    const double *__restrict__ rk_tab, const double *__restrict__ rkm_tab, int64_t n, int nv,
    int ng, int kmax, const double *__restrict__ mo, const double *__restrict__ mi, int n_items,
    const int32_t *__restrict__ row, const int32_t *__restrict__ kk,
    const int32_t *__restrict__ out_row, double *__restrict__ y, int64_t ld_y) {
    extern __shared__ double cell_sums[];  // [2*nv]: S_out then S_in
    __shared__ double mo_s[kMaxNg * kMaxNg], mi_s[kMaxNg * kMaxNg];
    __shared__ double g_out[kThreads], g_in[kThreads];
    __shared__ double tot_a[kThreads], tot_b[kThreads];

    double *s_out = cell_sums;
    double *s_in = cell_sums + nv;
    const int tid = threadIdx.x;
    for (int t = tid; t < ng * ng; t += kThreads) {
        mo_s[t] = mo[t];
        mi_s[t] = mi[t];
    }
    const int lane = tid & 63, wave = tid >> 6;
    const int cpw = 64 / ng;           // cells per wave
    const int cpb = cpw * (kThreads / 64);  // cells per block pass
    const int cell_in_wave = lane / ng, node = lane - cell_in_wave * ng;
    const bool lane_used = lane < cpw * ng;
    const int lane_cell_base = tid - node;  // LDS slot of node 0 of this lane's cell
    const int seg = (nv + kThreads - 1) / kThreads;
    __syncthreads();

    for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
        const double *src = rho + (int64_t)row[it] * ld_rho;
        const int k = kk[it];
        if (k < 0 || k > kmax) {
            // rank outside the r-power tables: poison the row instead of reading past them
            double *bad = y + (int64_t)out_row[it] * ld_y;
            for (int64_t x = tid; x < n; x += kThreads) bad[x] = __builtin_nan("");
            continue;
        }
        const double *rkp = rk_tab + (int64_t)k * n;
        const double *rkmp = rkm_tab + (int64_t)k * n;

        // pass 1: cell sums
        for (int c0 = 0; c0 < nv; c0 += cpb) {
            const int cell = c0 + wave * cpw + cell_in_wave;
            double go = 0.0, gi = 0.0;
            if (lane_used && cell < nv) {
                const int64_t x = (int64_t)cell * ng + node;
                double d = src[x];
                if (w) d = w[x] * d;
                go = d * rkp[x];
                gi = d * rkmp[x];
            }
            g_out[tid] = go;
            g_in[tid] = gi;
            __syncthreads();
            if (lane_used && node == 0 && cell < nv) {
                double so = 0.0, si = 0.0;
                for (int j = 0; j < ng; ++j) {
                    so += g_out[tid + j];
                    si += g_in[tid + j];
                }
                s_out[cell] = so;
                s_in[cell] = si;
            }
            __syncthreads();
        }

        // scan of cell sums: per-thread contiguous segments
        const int c_lo = tid * seg;
        const int c_hi = min(c_lo + seg, nv);
        double a = 0.0, b = 0.0;
        for (int c = c_lo; c < c_hi; ++c) {
            a += s_out[c];
            b += s_in[c];
        }
        tot_a[tid] = a;
        tot_b[kThreads - 1 - tid] = b;  // reversed: suffix becomes prefix
        __syncthreads();
        for (int off = 1; off < kThreads; off <<= 1) {
            double va = (tid >= off) ? tot_a[tid - off] : 0.0;
            double vb = (tid >= off) ? tot_b[tid - off] : 0.0;
            __syncthreads();
            tot_a[tid] += va;
            tot_b[tid] += vb;
            __syncthreads();
        }
        // exclusive bases read from the neighbouring inclusive sums (no
        // subtraction of nearly equal partial sums)
        double base_o = (tid == 0) ? 0.0 : tot_a[tid - 1];
        double base_i = (tid == kThreads - 1) ? 0.0 : tot_b[kThreads - 2 - tid];
        __syncthreads();
        for (int c = c_lo; c < c_hi; ++c) {
            double s = s_out[c];
            s_out[c] = base_o;
            base_o += s;
        }
        for (int c = c_hi - 1; c >= c_lo; --c) {
            double s = s_in[c];
            s_in[c] = base_i;
            base_i += s;
        }
        __syncthreads();

        // pass 2: in-cell partials and Y
        double *dst = y + (int64_t)out_row[it] * ld_y;
        for (int c0 = 0; c0 < nv; c0 += cpb) {
            const int cell = c0 + wave * cpw + cell_in_wave;
            const bool live = lane_used && cell < nv;
            const int64_t x = (int64_t)cell * ng + node;
            double rkv = 0.0, rkmv = 0.0;
            double go = 0.0, gi = 0.0;
            if (live) {
                double d = src[x];
                if (w) d = w[x] * d;
                rkv = rkp[x];
                rkmv = rkmp[x];
                go = d * rkv;
                gi = d * rkmv;
            }
            g_out[tid] = go;
            g_in[tid] = gi;
            __syncthreads();
            if (live) {
                double po = 0.0, pi = 0.0;
                for (int j = 0; j < ng; ++j) {
                    po += mo_s[node * ng + j] * g_out[lane_cell_base + j];
                    pi += mi_s[node * ng + j] * g_in[lane_cell_base + j];
                }
                const double iout = s_out[cell] + po;
                const double iin = s_in[cell] + pi;
                dst[x] = rkmv * iout + rkv * iin;
            }
            __syncthreads();
        }
    }
}

// ---------------------------------------------------------------------------
// LDS-resident Y^k scan (the production variant of the same arithmetic).
//   1. one coalesced pass loads the row and writes g_out = d r^k and
//      g_in = d r^-(k+1) to LDS, each cell padded to NG+1 doubles;
//   2. thread t owns cells t, t+T, t+2T, ... (round-major): per-cell sums and
//      the exclusive prefix (outward) / suffix (inward) bases over cells via
//      wave shuffle scans and the per-wave totals; bases go to LDS;
//   3. lanes map onto (cell, node) with NG fixed per lane, so each lane keeps
//      its rows of mo/mi in registers; the cell's g values are LDS broadcasts
//      and Y is stored with contiguous, coalesced wave stores.
// One global load round trip and three barriers per (row, k) item.
// ---------------------------------------------------------------------------
constexpr int kScanThreads = 512;
constexpr int kScanWaves = kScanThreads / 64;
constexpr int kMaxRounds = 4;  // nv <= kScanThreads * kMaxRounds

// LDS-resident scan of one (density row, k) item by a whole work-group:
// g_out/g_in of the row in LDS (cell pitch NG+1 against bank conflicts),
// per-cell sums, wave-shuffle prefix/suffix scans over cells, then one lane
// per (cell, node) applies the partial matrices and writes Y coalesced.
template <int NG>
struct ScanLds {
    double (*wt_o)[kScanWaves];
    double (*wt_i)[kScanWaves];
    double mo_r[NG], mi_r[NG];
};

// SLICE (NG = 8 only): instead of the FP64 row, write its split-precision
// slices (planes sdst[s*splane + x], exponents edst[tile]); the eight cells of
// one wave are exactly one 64-node K-tile.
template <int NG, bool SLICE = false>
__device__ inline void scan_item(double *__restrict__ sm, ScanLds<NG> &L,
                                 const double *__restrict__ src, const double *__restrict__ w,
                                 const double *__restrict__ rkp, const double *__restrict__ rkmp,
                                 int64_t n, int nv, double *__restrict__ dst,
                                 int8_t *__restrict__ sdst = nullptr, int64_t splane = 0,
                                 int32_t *__restrict__ edst = nullptr, int slices = 0,
                                 double *__restrict__ y0 = nullptr) {
    static_assert(!SLICE || NG * (64 / NG) == 64, "in-scan slicing needs whole K-tiles per wave");
    constexpr int P = NG + 1;
    double *gos = sm;                          // [nv * P] g_out
    double *gis = sm + (int64_t)nv * P;        // [nv * P] g_in
    double *bo = sm + (int64_t)2 * nv * P;     // [nv] outward base per cell
    double *bi = bo + nv;                      // [nv] inward base per cell
    auto wt_o = L.wt_o;
    auto wt_i = L.wt_i;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    constexpr int cpw = 64 / NG;
    const int ciw = lane / NG, node = lane - (lane / NG) * NG;
    const bool lane_used = lane < cpw * NG;
    const double *mo_r = L.mo_r;
    const double *mi_r = L.mi_r;
    const int rounds = (nv + kScanThreads - 1) / kScanThreads;
    __syncthreads();  // the previous item is done with the LDS
    for (int64_t x = tid; x < n; x += kScanThreads) {
        double d = src[x];
        if (w) d = w[x] * d;
        const int64_t px = x + x / NG;
        gos[px] = d * rkp[x];
        gis[px] = d * rkmp[x];
    }
    __syncthreads();
    double so[kMaxRounds], si[kMaxRounds], ex_o[kMaxRounds], ex_i[kMaxRounds];
#pragma unroll
    for (int r = 0; r < kMaxRounds; ++r) {
        so[r] = si[r] = ex_o[r] = ex_i[r] = 0.0;
        if (r < rounds) {
            const int c = r * kScanThreads + tid;
            if (c < nv) {
                double sa = 0.0, sb = 0.0;
#pragma unroll
                for (int j = 0; j < NG; ++j) {
                    sa += gos[(int64_t)c * P + j];
                    sb += gis[(int64_t)c * P + j];
                }
                so[r] = sa;
                si[r] = sb;
            }
            double pa = so[r], sb2 = si[r];
#pragma unroll
            for (int off = 1; off < 64; off <<= 1) {
                const double va = __shfl_up(pa, off, 64);
                const double vb = __shfl_down(sb2, off, 64);
                if (lane >= off) pa += va;
                if (lane + off < 64) sb2 += vb;
            }
            if (lane == 63) wt_o[r][wave] = pa;
            if (lane == 0) wt_i[r][wave] = sb2;
            const double ea = __shfl_up(pa, 1, 64);
            const double eb = __shfl_down(sb2, 1, 64);
            ex_o[r] = lane == 0 ? 0.0 : ea;
            ex_i[r] = lane == 63 ? 0.0 : eb;
        }
    }
    __syncthreads();
    double carry = 0.0;
#pragma unroll
    for (int r = 0; r < kMaxRounds; ++r) {
        if (r < rounds) {
            double before = carry;
            for (int v = 0; v < wave; ++v) before += wt_o[r][v];
            for (int v = 0; v < kScanWaves; ++v) carry += wt_o[r][v];
            const int c = r * kScanThreads + tid;
            if (c < nv) bo[c] = ex_o[r] + before;
        }
    }
    carry = 0.0;
#pragma unroll
    for (int r = kMaxRounds - 1; r >= 0; --r) {
        if (r < rounds) {
            double after = carry;
            for (int v = kScanWaves - 1; v > wave; --v) after += wt_i[r][v];
            for (int v = 0; v < kScanWaves; ++v) carry += wt_i[r][v];
            const int c = r * kScanThreads + tid;
            if (c < nv) bi[c] = ex_i[r] + after;
        }
    }
    __syncthreads();
    for (int c0 = 0; c0 < nv; c0 += cpw * kScanWaves) {
        const int c = c0 + wave * cpw + ciw;
        if constexpr (SLICE) {
            if (c0 + wave * cpw < nv) {  // wave-uniform: the wave's tile exists
                double v = 0.0;
                if (c < nv) {
                    const double *a = gos + (int64_t)c * P;
                    const double *b = gis + (int64_t)c * P;
                    double po = 0.0, pi = 0.0;
#pragma unroll
                    for (int j = 0; j < NG; ++j) {
                        po += mo_r[j] * a[j];
                        pi += mi_r[j] * b[j];
                    }
                    const int64_t x = (int64_t)c * NG + node;
                    v = rkmp[x] * (bo[c] + po) + rkp[x] * (bi[c] + pi);
                }
                const int64_t x = (int64_t)c * NG + node;
                ax::slice_value(v, x, sdst, splane, edst + x / ax::kSplitKTile, slices);
                // the first K-tile also in FP64 (it is contracted in FP64)
                if (y0 != nullptr && x < ax::kSplitKTile) y0[x] = v;
            }
        } else if (lane_used && c < nv) {
            const double *a = gos + (int64_t)c * P;
            const double *b = gis + (int64_t)c * P;
            double po = 0.0, pi = 0.0;
#pragma unroll
            for (int j = 0; j < NG; ++j) {
                po += mo_r[j] * a[j];
                pi += mi_r[j] * b[j];
            }
            const int64_t x = (int64_t)c * NG + node;
            dst[x] = rkmp[x] * (bo[c] + po) + rkp[x] * (bi[c] + pi);
        }
    }
}

template <int NG>
__device__ inline void scan_load_partials(ScanLds<NG> &L, const double *__restrict__ mo,
                                          const double *__restrict__ mi) {
    const int lane = threadIdx.x & 63;
    const int node = lane - (lane / NG) * NG;
#pragma unroll
    for (int j = 0; j < NG; ++j) {
        L.mo_r[j] = mo[node * NG + j];
        L.mi_r[j] = mi[node * NG + j];
    }
}

template <int NG>
__global__ __launch_bounds__(kScanThreads) void yk_scan_lds_kernel(
    const double *__restrict__ rho, int64_t ld_rho, const double *__restrict__ w,
    const double *__restrict__ rk_tab, const double *__restrict__ rkm_tab, int64_t n, int nv,
    int kmax, const double *__restrict__ mo, const double *__restrict__ mi, int n_items,
    const int32_t *__restrict__ row, const int32_t *__restrict__ kk,
    const int32_t *__restrict__ out_row, double *__restrict__ y, int64_t ld_y) {
    extern __shared__ double sm[];
    __shared__ double wt_o[kMaxRounds][kScanWaves], wt_i[kMaxRounds][kScanWaves];
    ScanLds<NG> L;
    L.wt_o = wt_o;
    L.wt_i = wt_i;
    scan_load_partials<NG>(L, mo, mi);
    for (int it = blockIdx.x; it < n_items; it += gridDim.x) {
        const int k = kk[it];
        double *dst = y + (int64_t)out_row[it] * ld_y;
        if (k < 0 || k > kmax) {
            for (int64_t x = threadIdx.x; x < n; x += kScanThreads) dst[x] = __builtin_nan("");
            continue;
        }
        scan_item<NG>(sm, L, rho + (int64_t)row[it] * ld_rho, w, rk_tab + (int64_t)k * n,
                      rkm_tab + (int64_t)k * n, n, nv, dst);
    }
}

// Batched scan over the items of several plans (same NG; the partial
// matrices depend only on NG).  Plan densities are already weighted.
template <int NG, bool SLICE>
__global__ __launch_bounds__(kScanThreads) void yk_scan_batch_kernel(
    const ax::BatchOf<ax::ScanArg> A, const double *__restrict__ mo,
    const double *__restrict__ mi) {
    extern __shared__ double sm[];
    __shared__ double wt_o[kMaxRounds][kScanWaves], wt_i[kMaxRounds][kScanWaves];
    ScanLds<NG> L;
    L.wt_o = wt_o;
    L.wt_i = wt_i;
    scan_load_partials<NG>(L, mo, mi);
    const int64_t total = A.end[A.n - 1];
    for (int64_t it = blockIdx.x; it < total; it += gridDim.x) {
        int64_t local;
        const int p = ax::batch_find(A, it, local);
        const ax::ScanArg &g = A.a[p];
        const int k = g.item_k[local];
        const int64_t n = g.n;
        const int64_t orow = g.item_out[local];
        const double *src = g.D + (int64_t)g.item_row[local] * n;
        if constexpr (SLICE)
            scan_item<NG, true>(sm, L, src, nullptr, g.rk + (int64_t)k * n,
                                g.rkm + (int64_t)k * n, n, g.nv, nullptr, g.Ys + orow * g.npad,
                                g.ys_plane, g.Ye + orow * (g.npad / ax::kSplitKTile), g.slices,
                                g.Y0 ? g.Y0 + orow * ax::kSplitKTile : nullptr);
        else
            scan_item<NG>(sm, L, src, nullptr, g.rk + (int64_t)k * n, g.rkm + (int64_t)k * n, n,
                          g.nv, g.Y + orow * n);
    }
}

template <int NG>
void launch_scan_lds(size_t lds, int grid, hipStream_t s, const double *rho, int64_t ld_rho,
                     const double *w, const double *rk, const double *rkm, int64_t n, int nv,
                     int kmax, const double *mo, const double *mi, int n_items, const int32_t *row,
                     const int32_t *kk, const int32_t *out_row, double *y, int64_t ld_y) {
    if (lds > 64 * 1024)
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(yk_scan_lds_kernel<NG>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    hipLaunchKernelGGL(yk_scan_lds_kernel<NG>, dim3(grid), dim3(kScanThreads), lds, s, rho, ld_rho,
                       w, rk, rkm, n, nv, kmax, mo, mi, n_items, row, kk, out_row, y, ld_y);
}

// ---------------------------------------------------------------------------
// R[t] = sum_x D[p_t, x] * Y[q_t, x]  (numerics/bspline/slater.py:129-130)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(kThreads) void rk_dot_kernel(
    const double *__restrict__ D, int64_t ld_d, const double *__restrict__ Y, int64_t ld_y,
    int64_t n, int n_quart, const int32_t *__restrict__ p, const int32_t *__restrict__ q,
    double *__restrict__ out) {
    __shared__ double red[kThreads];
    for (int t = blockIdx.x; t < n_quart; t += gridDim.x) {
        const double *a = D + (int64_t)p[t] * ld_d;
        const double *b = Y + (int64_t)q[t] * ld_y;
        double s = 0.0;
        for (int64_t x = threadIdx.x; x < n; x += kThreads) s += a[x] * b[x];
        red[threadIdx.x] = s;
        __syncthreads();
        for (int off = kThreads / 2; off > 0; off >>= 1) {
            if (threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[t] = red[0];
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Orbital samples f(x) = sum_s c[iv+s] B_s(x)  (numerics/bspline/evaluate.py:71-97)
// ---------------------------------------------------------------------------
__global__ void bspline_eval_kernel(const double *__restrict__ coeffs, int64_t ld_c, int n_orb,
                                    const double *__restrict__ bsp, int nv, int ng, int ks,
                                    int ns, double *__restrict__ out, int64_t ld_out) {
    const int64_t n = (int64_t)nv * ng;
    for (int o = blockIdx.y; o < n_orb; o += gridDim.y) {
        const double *c = coeffs + (int64_t)o * ld_c;
        for (int64_t x = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; x < n;
             x += (int64_t)gridDim.x * blockDim.x) {
            const int iv = static_cast<int>(x / ng);
            const double *b = bsp + x * ks;
            double s = 0.0;
            for (int sidx = 0; sidx < ks; ++sidx) {
                const int gi = iv + sidx;
                const double cv = gi < ns ? c[gi] : 0.0;
                s += cv * b[sidx];
            }
            out[(int64_t)o * ld_out + x] = s;
        }
    }
}

// ---------------------------------------------------------------------------
// Y^k-form exchange matrix (numerics/bspline/slater_yk.py:35-122).
// One work-group per column n_idx: the product density fb*B_n is nonzero on
// at most ks cells, so the node-wise cumulative integrals are constant
// outside that support; the column is assembled over all cells.
// ---------------------------------------------------------------------------
// <B1_gi | w | B2_gj> (numerics/bspline/integrals.py:18-56) for n_w weight
// rows, one thread per matrix element, gathering the cells that hold both
// B-splines in ascending order (deterministic, no atomics).
__global__ void assemble_integral_kernel(int nv, int ng, int ks1, int ks2, int mm, int n_w,
                                         const double *__restrict__ b1,
                                         const double *__restrict__ b2,
                                         const double *__restrict__ wts,
                                         double *__restrict__ mats) {
    const int64_t per = (int64_t)mm * mm, total = per * n_w;
    for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < total;
         u += (int64_t)gridDim.x * blockDim.x) {
        const int64_t wi = u / per, t = u - wi * per;
        const double *wt = wts + wi * (int64_t)nv * ng;
        const int gi = static_cast<int>(t / mm), gj = static_cast<int>(t - (int64_t)gi * mm);
        int lo = gi - ks1 + 1;
        if (gj - ks2 + 1 > lo) lo = gj - ks2 + 1;
        if (lo < 0) lo = 0;
        int hi = gi < gj ? gi : gj;
        if (hi > nv - 1) hi = nv - 1;
        double acc = 0.0;
        for (int iv = lo; iv <= hi; ++iv) {
            const int i = gi - iv, j = gj - iv;
            double s = 0.0;
            for (int m = 0; m < ng; ++m) {
                const int64_t c = (int64_t)iv * ng + m;
                s += wt[c] * b1[c * ks1 + i] * b2[c * ks2 + j];
            }
            acc += s;
        }
        mats[u] = acc;
    }
}

// M = (M + M^T) / 2 in place for n_mat square matrices (one thread per i < j).
__global__ void symmetrize_kernel(double *__restrict__ mats, int mm, int n_mat) {
    const int64_t per = (int64_t)mm * mm, total = per * n_mat;
    for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < total;
         u += (int64_t)gridDim.x * blockDim.x) {
        const int64_t t = u % per;
        const int i = static_cast<int>(t / mm), j = static_cast<int>(t - (int64_t)i * mm);
        if (j <= i) continue;
        double *M = mats + (u - t);
        const double v = 0.5 * (M[(int64_t)i * mm + j] + M[(int64_t)j * mm + i]);
        M[(int64_t)i * mm + j] = v;
        M[(int64_t)j * mm + i] = v;
    }
}

// Fast form of the exchange matrix, in two passes per (orbital, rank) item.
// Y^k of the density f B_n (supported on the cells of B_n) is, outside that
// support, r^-(k+1) T_out[n] above it and r^k T_in[n] below it.  So when the
// supports of B_m and B_n are disjoint, K[m][n] = A^-[m] T_out[n] (m above)
// or A^+[m] T_in[n] (m below) with the moments A^-/A^+ of w f B_m r^-(k+1) /
// w f B_m r^k; only the band |m - n| < ks needs the in-support profiles of
// the scan, evaluated exactly as the reference does (iv ascending, then the
// Gauss nodes).
//
// Pass 1, one thread per (item, n): in-support inclusive prefix / suffix
// (reference order), the totals and the two moments of B_n.
__global__ void exchange_profiles_kernel(const double *__restrict__ fb, int64_t ld_fb,
                                         const double *__restrict__ bsp,
                                         const double *__restrict__ w,
                                         const double *__restrict__ rk_tab,
                                         const double *__restrict__ rkm_tab, int kmax_tab,
                                         int nv, int ng, int ks, int ns, int n_items,
                                         const int32_t *__restrict__ item_b,
                                         const int32_t *__restrict__ item_k,
                                         double *__restrict__ prof) {
    const int L = ks * ng;
    const int64_t n_all = (int64_t)nv * ng;
    const int64_t total = (int64_t)n_items * ns;
    for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < total;
         u += (int64_t)gridDim.x * blockDim.x) {
        const int it = static_cast<int>(u / ns), nidx = static_cast<int>(u - (int64_t)it * ns);
        const int64_t kb = item_b ? item_b[it] : it, kk = item_k ? item_k[it] : kmax_tab;
        const double *f = fb + kb * ld_fb;
        const double *rk = rk_tab + kk * n_all;
        const double *rkm = rkm_tab + kk * n_all;
        double *pr = prof + u * (2 * L + 4);  // [L] out-prefix, [L] in-suffix, Tout, Tin, A-, A+
        int lo = nidx - ks + 1;
        if (lo < 0) lo = 0;
        int hi = nidx + 1;
        if (hi > nv) hi = nv;
        const int64_t x_lo = (int64_t)lo * ng;
        const int len = hi > lo ? (hi - lo) * ng : 0;
        double s = 0.0, am = 0.0, ap = 0.0;
        for (int t = 0; t < len; ++t) {
            const int64_t x = x_lo + t;
            const int iv = static_cast<int>(x / ng);
            const double rho = f[x] * bsp[x * ks + (nidx - iv)];
            s += w[x] * rho * rk[x];
            pr[t] = s;
        }
        pr[2 * L] = s;
        s = 0.0;
        for (int t = len - 1; t >= 0; --t) {
            const int64_t x = x_lo + t;
            const int iv = static_cast<int>(x / ng);
            const double rho = f[x] * bsp[x * ks + (nidx - iv)];
            s += w[x] * rho * rkm[x];
            pr[L + t] = s;
        }
        pr[2 * L + 1] = s;
        // moments of B_n as the left factor: cells iv ascending, nodes inside
        for (int sidx = ks - 1; sidx >= 0; --sidx) {
            const int iv = nidx - sidx;
            if (iv < 0 || iv >= nv) continue;
            double a1 = 0.0, a2 = 0.0;
            for (int g = 0; g < ng; ++g) {
                const int64_t x = (int64_t)iv * ng + g;
                const double v = (w[x] * f[x]) * bsp[x * ks + sidx];
                a1 += v * rkm[x];
                a2 += v * rk[x];
            }
            am += a1;
            ap += a2;
        }
        pr[2 * L + 2] = am;
        pr[2 * L + 3] = ap;
    }
}

// Pass 2, one thread per (item, m, n).
__global__ void exchange_assemble_kernel(const double *__restrict__ fb, int64_t ld_fb,
                                         const double *__restrict__ bsp,
                                         const double *__restrict__ w,
                                         const double *__restrict__ rk_tab,
                                         const double *__restrict__ rkm_tab, int kmax_tab,
                                         int nv, int ng, int ks, int ns, int n_items,
                                         const int32_t *__restrict__ item_b,
                                         const int32_t *__restrict__ item_k,
                                         const double *__restrict__ prof,
                                         double *__restrict__ K) {
    const int L = ks * ng;
    const int64_t n_all = (int64_t)nv * ng;
    const int64_t per = (int64_t)ns * ns, total = per * n_items;
    for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < total;
         u += (int64_t)gridDim.x * blockDim.x) {
        const int it = static_cast<int>(u / per);
        const int64_t e = u - (int64_t)it * per;
        const int m = static_cast<int>(e / ns), nidx = static_cast<int>(e - (int64_t)m * ns);
        const double *pn = prof + ((int64_t)it * ns + nidx) * (2 * L + 4);
        double val;
        if (m - nidx >= ks) {
            val = prof[((int64_t)it * ns + m) * (2 * L + 4) + 2 * L + 2] * pn[2 * L];
        } else if (nidx - m >= ks) {
            val = prof[((int64_t)it * ns + m) * (2 * L + 4) + 2 * L + 3] * pn[2 * L + 1];
        } else {
            const int64_t kb = item_b ? item_b[it] : it, kk = item_k ? item_k[it] : kmax_tab;
            const double *f = fb + kb * ld_fb;
            const double *rk = rk_tab + kk * n_all;
            const double *rkm = rkm_tab + kk * n_all;
            int lo = nidx - ks + 1;
            if (lo < 0) lo = 0;
            int hi = nidx + 1;
            if (hi > nv) hi = nv;
            const int64_t x_lo = (int64_t)lo * ng, x_hi = (int64_t)hi * ng;
            const double tot_out = pn[2 * L], tot_in = pn[2 * L + 1];
            double acc_m = 0.0;
            for (int sidx = ks - 1; sidx >= 0; --sidx) {
                const int iv = m - sidx;
                if (iv < 0 || iv >= nv) continue;
                double acc = 0.0;
                for (int g = 0; g < ng; ++g) {
                    const int64_t x = (int64_t)iv * ng + g;
                    double io, ii;
                    if (x < x_lo) {
                        io = 0.0;
                        ii = tot_in;
                    } else if (x >= x_hi) {
                        io = tot_out;
                        ii = 0.0;
                    } else {
                        io = pn[x - x_lo];
                        ii = pn[L + (x - x_lo)];
                    }
                    const double yk = rkm[x] * io + rk[x] * ii;
                    acc += (w[x] * f[x]) * bsp[x * ks + sidx] * yk;
                }
                acc_m += acc;
            }
            val = acc_m;
        }
        K[(int64_t)it * per + (int64_t)m * ns + nidx] = val;
    }
}


// ---------------------------------------------------------------------------
// "direct" O(N^2) double quadrature with the coincident-cell cusp split
// (numerics/bspline/slater.py:51-115).  Disjoint cells: product rule with the
// bare kernel; coincident cells: the triangle rule r = a + h x, s = a + h x y
// with densities from their cell interpolants (tables outer/inner built on
// the host).  Partial sums per block/cell, then a fixed-order final sum.
// ---------------------------------------------------------------------------
__device__ inline double block_sum(double v, double *red) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
        __syncthreads();
    }
    double r = red[0];
    __syncthreads();
    return r;
}

__global__ __launch_bounds__(kThreads) void direct_disjoint_kernel(
    const double *__restrict__ rho13, const double *__restrict__ rho24,
    const double *__restrict__ r, const double *__restrict__ w, int64_t n, int ng, int k,
    double *__restrict__ partial) {
    __shared__ double red[kThreads];
    double acc = 0.0;
    for (int64_t i = blockIdx.x; i < n; i += gridDim.x) {
        const int64_t j0 = (i / ng + 1) * ng;
        const double ri = r[i], wi = w[i], a13 = rho13[i], a24 = rho24[i];
        for (int64_t j = j0 + threadIdx.x; j < n; j += kThreads) {
            const double kern = pow(ri / r[j], static_cast<double>(k)) / r[j];
            acc += wi * w[j] * kern * (a13 * rho24[j] + rho13[j] * a24);
        }
    }
    const double s = block_sum(acc, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(kThreads) void direct_diag_kernel(
    const double *__restrict__ rho13, const double *__restrict__ rho24,
    const double *__restrict__ gr, const double *__restrict__ grw, int nv, int ng, int k,
    int order, const double *__restrict__ xq, const double *__restrict__ wq,
    const double *__restrict__ outer, const double *__restrict__ inner, double node0,
    double *__restrict__ partial) {
    __shared__ double red[kThreads];
    __shared__ double fo[64], go[64];
    for (int c = blockIdx.x; c < nv; c += gridDim.x) {
        const double *f = rho13 + (int64_t)c * ng;
        const double *g = rho24 + (int64_t)c * ng;
        double h = 0.0;
        for (int m = 0; m < ng; ++m) h += grw[(int64_t)c * ng + m];
        const double a = gr[(int64_t)c * ng] - h * (1.0 + node0) / 2.0;
        for (int x = threadIdx.x; x < order; x += blockDim.x) {
            double sf = 0.0, sg = 0.0;
            for (int j = 0; j < ng; ++j) {
                sf += f[j] * outer[x * ng + j];
                sg += g[j] * outer[x * ng + j];
            }
            fo[x] = sf;
            go[x] = sg;
        }
        __syncthreads();
        double acc = 0.0;
        for (int t = threadIdx.x; t < order * order; t += blockDim.x) {
            const int x = t / order, y = t - (t / order) * order;
            const double *in = inner + (int64_t)t * ng;
            double fi = 0.0, gi = 0.0;
            for (int j = 0; j < ng; ++j) {
                fi += f[j] * in[j];
                gi += g[j] * in[j];
            }
            const double rr = a + h * xq[x];
            const double ss = a + h * xq[x] * xq[y];
            const double kj = pow(ss / rr, static_cast<double>(k)) / rr * (h * h) * xq[x];
            const double dens = fo[x] * gi + go[x] * fi;
            acc += dens * kj * wq[x] * wq[y];
        }
        const double s = block_sum(acc, red);
        if (threadIdx.x == 0) partial[c] = s;
    }
}

__global__ __launch_bounds__(kThreads) void sum_partials_kernel(const double *__restrict__ a,
                                                                int na,
                                                                const double *__restrict__ b,
                                                                int nb, double *__restrict__ out) {
    __shared__ double red[kThreads];
    double s = 0.0;
    for (int t = threadIdx.x; t < na; t += kThreads) s += a[t];
    const double sa = block_sum(s, red);
    s = 0.0;
    for (int t = threadIdx.x; t < nb; t += kThreads) s += b[t];
    const double sb = block_sum(s, red);
    if (threadIdx.x == 0) out[0] = sa + sb;
}

// ---------------------------------------------------------------------------
// Batched scan, one 256-thread work-group per (row, k) item, rows of up to
// 64 * 4 * MAXQ cells.  Cells are taken in passes of 64; pass P belongs to
// wave P % 4.  Global memory is always read and written node-per-lane
// (lane l holds node P*64*NG + r*64 + l, fully coalesced, every load of the
// item issued before any arithmetic); a per-wave LDS region transposes the
// products g = d r^k, d r^-(k+1) to cell-per-lane for the cell sums, the
// in-wave Hillis-Steele scans and the in-cell partial integrals, and the
// two cumulative integrals back to node-per-lane for the output (or, with
// SLICE, for the int8 slices: for a fixed r the 64 lanes hold one K-tile).
// 8-node cells skip the transposes on the way in (each lane reads its cell's
// 64 contiguous bytes); their FP64 output goes through a per-wave LDS stage
// so that every store is 1 KB contiguous (cell-per-lane stores, 16 bytes
// per lane at a 64-byte stride, measured 1.6 us slower on 1389 rows of 3200
// nodes).  The per-pass wave totals are combined through LDS in a fixed
// order after one barrier.  Same discretisation as scan_item
// (numerics/yk.py:87-154).
// ---------------------------------------------------------------------------
constexpr int kWgScanThreads = 256;
constexpr int kWgScanWaves = kWgScanThreads / 64;

template <int NG, bool SLICE, int MAXQ>
__global__ __launch_bounds__(kWgScanThreads) void yk_scan_wg_kernel(
    const ax::BatchOf<ax::ScanArg> A, const double *__restrict__ mo,
    const double *__restrict__ mi) {
    constexpr int MAXP = MAXQ * kWgScanWaves;
    constexpr int CS = NG + 1;  // padded cell stride of the transpose region
    __shared__ double tot_o[MAXP], tot_i[MAXP];
    // transpose regions: only 7- and 9-node cells go through them; 8-node
    // cells are read and written (or sliced) cell-per-lane
    constexpr int TW = NG != 8 ? kWgScanWaves : 1;
    __shared__ double tr[TW][2][64 * CS];
    // FP64 output of 8-node cells: a per-wave stage (cell stride padded to
    // 10 doubles) that turns the cell-per-lane results into contiguous stores
    constexpr int YS = NG == 8 ? 10 : 1;
    __shared__ __attribute__((aligned(16))) double ystage[NG == 8 ? kWgScanWaves : 1][64 * YS];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int64_t local;
    const int p = ax::batch_find(A, blockIdx.x, local);
    const ax::ScanArg &g = A.a[p];
    const int k = g.item_k[local];
    const int64_t n = g.n;
    const int nv = g.nv;
    const int passes = (nv + 63) / 64;
    const double *src = g.D + (int64_t)g.item_row[local] * n;
    const double *rkp = g.rk + (int64_t)k * n;
    const double *rkmp = g.rkm + (int64_t)k * n;
    double *ybuf = ystage[NG == 8 ? wave : 0];
    double *to = tr[TW > 1 ? wave : 0][0];
    double *ti = tr[TW > 1 ? wave : 0][1];

    // node-per-lane products of every pass of this wave; the loads use a
    // clamped index and are all issued before the first product (a load
    // under a per-lane branch would be waited for inside the branch)
    double no[MAXQ][NG], ni[MAXQ][NG];
    double go[MAXQ][NG], gi[MAXQ][NG];
    if constexpr (NG == 8) {
        // 8-node cells: the lane's cell is read directly (64 contiguous
        // bytes per row), no transpose
        typedef double dd2 __attribute__((ext_vector_type(2)));
        dd2 d2[MAXQ][NG / 2], a2[MAXQ][NG / 2], b2[MAXQ][NG / 2];
#pragma unroll
        for (int q = 0; q < MAXQ; ++q) {
            const int64_t c0 = (int64_t)(q * kWgScanWaves + wave) * 64 * NG + (int64_t)lane * NG;
            const int64_t cc = c0 + NG <= n ? c0 : 0;
#pragma unroll
            for (int h = 0; h < NG / 2; ++h) {
                d2[q][h] = *reinterpret_cast<const dd2 *>(src + cc + 2 * h);
                a2[q][h] = *reinterpret_cast<const dd2 *>(rkp + cc + 2 * h);
                b2[q][h] = *reinterpret_cast<const dd2 *>(rkmp + cc + 2 * h);
            }
        }
#pragma unroll
        for (int q = 0; q < MAXQ; ++q) {
            const int64_t c0 = (int64_t)(q * kWgScanWaves + wave) * 64 * NG + (int64_t)lane * NG;
            const bool in = c0 + NG <= n;
#pragma unroll
            for (int h = 0; h < NG / 2; ++h) {
                go[q][2 * h] = in ? d2[q][h].x * a2[q][h].x : 0.0;
                go[q][2 * h + 1] = in ? d2[q][h].y * a2[q][h].y : 0.0;
                gi[q][2 * h] = in ? d2[q][h].x * b2[q][h].x : 0.0;
                gi[q][2 * h + 1] = in ? d2[q][h].y * b2[q][h].y : 0.0;
            }
        }
    } else {
        double dv[MAXQ][NG], av[MAXQ][NG], bv[MAXQ][NG];
#pragma unroll
        for (int q = 0; q < MAXQ; ++q) {
            const int64_t base = (int64_t)(q * kWgScanWaves + wave) * 64 * NG;
#pragma unroll
            for (int r = 0; r < NG; ++r) {
                const int64_t x = base + r * 64 + lane;
                const int64_t xc = x < n ? x : n - 1;
                dv[q][r] = src[xc];
                av[q][r] = rkp[xc];
                bv[q][r] = rkmp[xc];
            }
        }
#pragma unroll
        for (int q = 0; q < MAXQ; ++q) {
            const int64_t base = (int64_t)(q * kWgScanWaves + wave) * 64 * NG;
#pragma unroll
            for (int r = 0; r < NG; ++r) {
                const bool in = base + r * 64 + lane < n;
                no[q][r] = in ? dv[q][r] * av[q][r] : 0.0;
                ni[q][r] = in ? dv[q][r] * bv[q][r] : 0.0;
            }
        }
    }
    // cell-per-lane values, cell sums and in-wave scans
    double ex_o[MAXQ], ex_i[MAXQ];
#pragma unroll
    for (int q = 0; q < MAXQ; ++q) {
        const int P = q * kWgScanWaves + wave;
        if constexpr (NG != 8) {
#pragma unroll
            for (int r = 0; r < NG; ++r) {
                const int xl = r * 64 + lane;  // node of the pass, cell xl / NG
                to[(xl / NG) * CS + xl % NG] = no[q][r];
                ti[(xl / NG) * CS + xl % NG] = ni[q][r];
            }
#pragma unroll
            for (int j = 0; j < NG; ++j) {
                go[q][j] = to[lane * CS + j];
                gi[q][j] = ti[lane * CS + j];
            }
        }
        double so = 0.0, si = 0.0;
#pragma unroll
        for (int j = 0; j < NG; ++j) {
            so += go[q][j];
            si += gi[q][j];
        }
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const double va = __shfl_up(so, off, 64);
            const double vb = __shfl_down(si, off, 64);
            if (lane >= off) so += va;
            if (lane + off < 64) si += vb;
        }
        const double ea = __shfl_up(so, 1, 64);
        const double eb = __shfl_down(si, 1, 64);
        ex_o[q] = lane == 0 ? 0.0 : ea;
        ex_i[q] = lane == 63 ? 0.0 : eb;
        const double tso = __shfl(so, 63, 64), tsi = __shfl(si, 0, 64);
        if (lane == 0) {
            tot_o[P] = P < passes ? tso : 0.0;
            tot_i[P] = P < passes ? tsi : 0.0;
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXQ; ++q) {
        const int P = q * kWgScanWaves + wave;
        if (P >= passes) break;  // wave-uniform
        double co = 0.0, ci = 0.0;
        for (int v = 0; v < P; ++v) co += tot_o[v];
        for (int v = passes - 1; v > P; --v) ci += tot_i[v];
        const double bo = co + ex_o[q], bi = ci + ex_i[q];
        const int64_t base = (int64_t)P * 64 * NG;
        if constexpr (!SLICE && NG == 8) {
            // the lane's cell: its 8 consecutive nodes are read and written
            // by this lane (64 contiguous bytes per row)
            typedef double dd2 __attribute__((ext_vector_type(2)));
            const int64_t c0 = base + (int64_t)lane * NG;
            if (c0 + NG <= n) {
                dd2 a2[NG / 2], b2[NG / 2];
#pragma unroll
                for (int h = 0; h < NG / 2; ++h) {
                    a2[h] = *reinterpret_cast<const dd2 *>(rkp + c0 + 2 * h);
                    b2[h] = *reinterpret_cast<const dd2 *>(rkmp + c0 + 2 * h);
                }
                dd2 y2[NG / 2];
#pragma unroll
                for (int m = 0; m < NG; ++m) {
                    double po = 0.0, pi = 0.0;
#pragma unroll
                    for (int j = 0; j < NG; ++j) {
                        po += mo[m * NG + j] * go[q][j];
                        pi += mi[m * NG + j] * gi[q][j];
                    }
                    const double ra = (m & 1) ? a2[m >> 1].y : a2[m >> 1].x;
                    const double rb = (m & 1) ? b2[m >> 1].y : b2[m >> 1].x;
                    const double y = rb * (bo + po) + ra * (bi + pi);
                    if (m & 1) y2[m >> 1].y = y; else y2[m >> 1].x = y;
                }
                // through the wave's LDS stage: written cell-per-lane, read
                // back as node pairs so that each store is 1 KB contiguous
#pragma unroll
                for (int h = 0; h < NG / 2; ++h)
                    *reinterpret_cast<dd2 *>(ybuf + lane * YS + 2 * h) = y2[h];
            }
            __builtin_amdgcn_wave_barrier();
            {
                double *yo = g.Y + (int64_t)g.item_out[local] * n + base;
#pragma unroll
                for (int i = 0; i < NG / 2; ++i) {
                    const int xl = 128 * i + 2 * lane;  // node pair of the pass's 512 nodes
                    const dd2 v = *reinterpret_cast<const dd2 *>(ybuf + (xl / NG) * YS + xl % NG);
                    if (base + xl + 1 < n) *reinterpret_cast<dd2 *>(yo + xl) = v;
                }
            }
            __builtin_amdgcn_wave_barrier();
            continue;
        }
        if constexpr (SLICE && NG == 8) {
            // the lane's cell, sliced in place: eight lanes hold one 64-node
            // K-tile (their exponent is the tile's), and each slice plane
            // gets the cell's 8 bytes in one store (the bytes of
            // ax::slice_value, node by node)
            typedef double dd2 __attribute__((ext_vector_type(2)));
            const int64_t c0 = base + (int64_t)lane * NG;
            double y[NG];
            if (c0 + NG <= n) {
                dd2 a2[NG / 2], b2[NG / 2];
#pragma unroll
                for (int h = 0; h < NG / 2; ++h) {
                    a2[h] = *reinterpret_cast<const dd2 *>(rkp + c0 + 2 * h);
                    b2[h] = *reinterpret_cast<const dd2 *>(rkmp + c0 + 2 * h);
                }
#pragma unroll
                for (int m = 0; m < NG; ++m) {
                    double po = 0.0, pi = 0.0;
#pragma unroll
                    for (int j = 0; j < NG; ++j) {
                        po += mo[m * NG + j] * go[q][j];
                        pi += mi[m * NG + j] * gi[q][j];
                    }
                    const double ra = (m & 1) ? a2[m >> 1].y : a2[m >> 1].x;
                    const double rb = (m & 1) ? b2[m >> 1].y : b2[m >> 1].x;
                    y[m] = rb * (bo + po) + ra * (bi + pi);
                }
            } else {
#pragma unroll
                for (int m = 0; m < NG; ++m) y[m] = 0.0;
            }
            if (g.Y0 != nullptr && c0 < ax::kSplitKTile) {  // cells of K-tile 0, also in FP64
                double *y0 = g.Y0 + g.item_out[local] * ax::kSplitKTile + c0;
#pragma unroll
                for (int m = 0; m < NG; ++m) y0[m] = y[m];
            }
            double mx = 0.0;
#pragma unroll
            for (int m = 0; m < NG; ++m) mx = fmax(mx, fabs(y[m]));
#pragma unroll
            for (int off = 1; off < 8; off <<= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
            if (c0 < g.npad) {  // the same for the tile's eight lanes
                const int ex = mx > 0.0 ? ilogb(mx) + 1 : 0;
                const int64_t row = g.item_out[local];
                double rho[NG];
#pragma unroll
                for (int m = 0; m < NG; ++m) rho[m] = ldexp(y[m], 7 - ex);
                int8_t *dst = g.Ys + row * g.npad + c0;
                for (int sl = 0; sl < g.slices; ++sl) {
                    uint64_t word = 0;
#pragma unroll
                    for (int m = 0; m < NG; ++m) {
                        const double c = trunc(rho[m]);
                        word |= (uint64_t)(uint8_t)static_cast<int8_t>(c) << (8 * m);
                        rho[m] = (rho[m] - c) * 128.0;
                    }
                    *reinterpret_cast<uint64_t *>(dst + (int64_t)sl * g.ys_plane) = word;
                }
                if ((lane & 7) == 0) g.Ye[row * (g.npad / ax::kSplitKTile) + c0 / ax::kSplitKTile] = ex;
            }
            continue;
        }
        // cumulative integrals of the lane's cell, back to node-per-lane
#pragma unroll
        for (int m = 0; m < NG; ++m) {
            double po = 0.0, pi = 0.0;
#pragma unroll
            for (int j = 0; j < NG; ++j) {
                po += mo[m * NG + j] * go[q][j];
                pi += mi[m * NG + j] * gi[q][j];
            }
            to[lane * CS + m] = bo + po;
            ti[lane * CS + m] = bi + pi;
        }
        double ra[NG], rb[NG];
#pragma unroll
        for (int r = 0; r < NG; ++r) {
            const int64_t x = base + r * 64 + lane;
            const int64_t xc = x < n ? x : n - 1;
            ra[r] = rkp[xc];
            rb[r] = rkmp[xc];
        }
#pragma unroll
        for (int r = 0; r < NG; ++r) {
            const int xl = r * 64 + lane;
            const int64_t x = base + xl;
            double y = rb[r] * to[(xl / NG) * CS + xl % NG] + ra[r] * ti[(xl / NG) * CS + xl % NG];
            if (x >= n) y = 0.0;
            if constexpr (SLICE) {
                static_assert(!SLICE || NG == 8, "in-scan slicing needs 8-node cells");
                const int64_t row = g.item_out[local];
                if (base + r * 64 < g.npad)  // wave-uniform: the K-tile exists
                    ax::slice_value(y, x, g.Ys + row * g.npad, g.ys_plane,
                                    g.Ye + row * (g.npad / ax::kSplitKTile) + x / ax::kSplitKTile,
                                    g.slices);
            } else if (x < n) {
                g.Y[(int64_t)g.item_out[local] * n + x] = y;
            }
        }
    }
}

// Largest row (cells) the register-resident scan takes.
constexpr int kWgScanMaxCells = 64 * kWgScanWaves * 4;

template <int NG, bool SLICE>
void launch_scan_wg(int nv_max, int64_t items, hipStream_t s, const ax::BatchOf<ax::ScanArg> &A,
                    const double *mo, const double *mi) {
    const dim3 grid(static_cast<unsigned>(items));
    if (nv_max <= 64 * kWgScanWaves)
        hipLaunchKernelGGL((yk_scan_wg_kernel<NG, SLICE, 1>), grid, dim3(kWgScanThreads), 0, s, A,
                           mo, mi);
    else if (nv_max <= 2 * 64 * kWgScanWaves)
        hipLaunchKernelGGL((yk_scan_wg_kernel<NG, SLICE, 2>), grid, dim3(kWgScanThreads), 0, s, A,
                           mo, mi);
    else
        hipLaunchKernelGGL((yk_scan_wg_kernel<NG, SLICE, 4>), grid, dim3(kWgScanThreads), 0, s, A,
                           mo, mi);
}

template <int NG, bool SLICE>
void launch_scan_batch(size_t lds, int grid, hipStream_t s, const ax::BatchOf<ax::ScanArg> &A,
                       const double *mo, const double *mi) {
    if (lds > 64 * 1024)
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(yk_scan_batch_kernel<NG, SLICE>),
                                  hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    hipLaunchKernelGGL((yk_scan_batch_kernel<NG, SLICE>), dim3(grid), dim3(kScanThreads), lds, s, A,
                       mo, mi);
}

int lds_capacity() {
    static int cap = -1;
    if (cap < 0) {
        int dev = 0, v = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess)
            cap = v;
        else
            cap = 65536;
    }
    return cap;
}

size_t scan_lds_bytes(int nv, int ng) {
    return ((size_t)2 * nv * (ng + 1) + 2 * (size_t)nv) * sizeof(double);
}

// Whether the LDS-resident scan serves a grid of nv cells of ng nodes.
bool scan_lds_fits(int nv, int ng) {
    if (ng != 7 && ng != 8 && ng != 9) return false;
    const size_t static_lds = 2 * kMaxRounds * kScanWaves * sizeof(double);
    return nv <= kScanThreads * kMaxRounds &&
           scan_lds_bytes(nv, ng) + static_lds <= (size_t)lds_capacity();
}

}  // namespace

namespace ax {

bool scan_slices_fused(int ng, int nv) {
    return ng == 8 && (nv <= kWgScanMaxCells || scan_lds_fits(nv, ng));
}

ax_status density_batch_launch(const PlanHost *const *plans, int n, const double *const *P,
                               const double *const *Q, hipStream_t stream) {
    if (n <= 0) return AX_OK;
    if (n > kMaxBatch) {
        set_error("density_batch_launch: %d plans exceed the batch limit %d", n, kMaxBatch);
        return AX_ERR_VALUE;
    }
    BatchOf<DensArg> A{};
    A.n = n;
    int64_t total = 0;  // work-groups: pairs x 512-node chunks, plan after plan
    for (int i = 0; i < n; ++i) {
        const PlanHost *h = plans[i];
        if (!P[i]) {
            set_error("density_batch_launch: null orbital samples for plan %d", i);
            return AX_ERR_VALUE;
        }
        const bool sl = h->precision == AX_PREC_SPLIT;
        const int64_t len = sl ? h->npad : h->n;
        const int chunks = static_cast<int>((len + 511) / 512);
        total += (int64_t)h->n_pairs * chunks;
        A.end[i] = total;
        A.a[i] = DensArg{P[i], Q ? Q[i] : nullptr, h->n, h->n, h->pa, h->pb, h->w, h->D,
                         sl ? h->Ds : nullptr, sl ? h->De : nullptr, h->npad,
                         (int64_t)h->n_pairs * h->npad, h->slices, chunks};
    }
    if (total == 0) return AX_OK;
    const dim3 grid(static_cast<unsigned>(total < (1 << 22) ? total : (1 << 22)));
    hipLaunchKernelGGL(pair_density_batch_kernel, grid, dim3(256), 0, stream, A);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

ax_status scan_batch_launch(const PlanHost *const *plans, int n, hipStream_t stream) {
    if (n <= 0) return AX_OK;
    if (n > kMaxBatch) {
        set_error("scan_batch_launch: %d plans exceed the batch limit %d", n, kMaxBatch);
        return AX_ERR_VALUE;
    }
    // rows of up to kWgScanMaxCells cells: one register-resident scan launch
    // per (n_gauss, in-scan slicing, passes per wave), so that a short row
    // never carries the registers and loads of the batch's longest (the
    // passes per wave do not change a row's sums); longer rows: the LDS scan,
    // and rows the LDS scan does not hold go through the generic kernel one
    // plan at a time
    auto q_class = [](int nv) {
        return nv <= 64 * kWgScanWaves ? 1 : nv <= 128 * kWgScanWaves ? 2 : 4;
    };
    for (int c = 0; c < 9; ++c) {
        const int ng = 7 + c / 3, qc = 1 << (c % 3);
        for (int sl = 0; sl < (ng == 8 ? 2 : 1); ++sl) {
            BatchOf<ScanArg> A{};
            int64_t total = 0;
            int nv_max = 0;
            const double *mo = nullptr, *mi = nullptr;
            for (int i = 0; i < n; ++i) {
                const PlanHost *h = plans[i];
                if (h->ng != ng || h->nv > kWgScanMaxCells || h->n_items == 0) continue;
                if (q_class(h->nv) != qc) continue;
                const bool fused = h->precision == AX_PREC_SPLIT && h->Y == nullptr;
                if (fused != (sl == 1)) continue;
                total += h->n_items;
                A.end[A.n] = total;
                const int64_t yr = h->n_yrows > 0 ? h->n_yrows : 1;
                A.a[A.n] = ScanArg{h->D, h->Y, h->rk, h->rkm, h->item_row, h->item_k,
                                   h->item_out, h->n, h->nv, h->kmax, h->Ys, h->Ye, h->npad,
                                   yr * h->npad, h->slices, 0, h->Y0};
                ++A.n;
                if (h->nv > nv_max) nv_max = h->nv;
                mo = h->mo;
                mi = h->mi;
            }
            if (A.n == 0) continue;
            if (total > 2147483647LL) {
                set_error("scan_batch_launch: %lld items exceed one launch", (long long)total);
                return AX_ERR_VALUE;
            }
            if (ng == 7) launch_scan_wg<7, false>(nv_max, total, stream, A, mo, mi);
            if (ng == 8 && sl == 0) launch_scan_wg<8, false>(nv_max, total, stream, A, mo, mi);
            if (ng == 8 && sl == 1) launch_scan_wg<8, true>(nv_max, total, stream, A, mo, mi);
            if (ng == 9) launch_scan_wg<9, false>(nv_max, total, stream, A, mo, mi);
            AX_CHECK_LAUNCH();
        }
    }
    for (int ng : {7, 8, 9}) {
        for (int sl = 0; sl < (ng == 8 ? 2 : 1); ++sl) {
            BatchOf<ScanArg> A{};
            int64_t total = 0;
            int nv_max = 0;
            const double *mo = nullptr, *mi = nullptr;
            for (int i = 0; i < n; ++i) {
                const PlanHost *h = plans[i];
                if (h->ng != ng || h->nv <= kWgScanMaxCells || !scan_lds_fits(h->nv, h->ng) ||
                    h->n_items == 0)
                    continue;
                const bool fused = h->precision == AX_PREC_SPLIT && h->Y == nullptr;
                if (fused != (sl == 1)) continue;
                total += h->n_items;
                A.end[A.n] = total;
                const int64_t yr = h->n_yrows > 0 ? h->n_yrows : 1;
                A.a[A.n] = ScanArg{h->D, h->Y, h->rk, h->rkm, h->item_row, h->item_k,
                                   h->item_out, h->n, h->nv, h->kmax, h->Ys, h->Ye, h->npad,
                                   yr * h->npad, h->slices, 0, h->Y0};
                ++A.n;
                if (h->nv > nv_max) nv_max = h->nv;
                mo = h->mo;
                mi = h->mi;
            }
            if (A.n == 0) continue;
            const int grid = total < 8192 ? static_cast<int>(total) : 8192;
            const size_t lds = scan_lds_bytes(nv_max, ng);
            if (ng == 7) launch_scan_batch<7, false>(lds, grid, stream, A, mo, mi);
            if (ng == 8 && sl == 0) launch_scan_batch<8, false>(lds, grid, stream, A, mo, mi);
            if (ng == 8 && sl == 1) launch_scan_batch<8, true>(lds, grid, stream, A, mo, mi);
            if (ng == 9) launch_scan_batch<9, false>(lds, grid, stream, A, mo, mi);
            AX_CHECK_LAUNCH();
        }
    }
    for (int i = 0; i < n; ++i) {
        const PlanHost *h = plans[i];
        const bool wg = (h->ng == 7 || h->ng == 8 || h->ng == 9) && h->nv <= kWgScanMaxCells;
        if (wg || scan_lds_fits(h->nv, h->ng) || h->n_items == 0) continue;
        ax_status st = ax_yk_batch(h->D, h->n, nullptr, h->rk, h->rkm, h->n, h->nv, h->ng,
                                   h->kmax, h->mo, h->mi, h->n_items, h->item_row, h->item_k,
                                   h->item_out, h->Y, h->n, stream);
        if (st != AX_OK) return st;
    }
    return AX_OK;
}

}  // namespace ax

namespace {
}  // namespace

extern "C" {

ax_status ax_grid_rpow(const double *r, int64_t n, int kmax, double *rk, double *rkm,
                       void *stream) {
    AX_REQUIRE(n > 0 && kmax >= 0, "ax_grid_rpow: bad sizes n=%lld kmax=%d", (long long)n, kmax);
    hipLaunchKernelGGL(rpow_kernel, dim3(ax::grid_for(n * (kmax + 1), 256, 4096)), dim3(256), 0,
                       ax::as_stream(stream), r, n, kmax, rk, rkm);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

ax_status ax_pair_density(const double *P, const double *Q, int64_t ld_pq, int64_t n,
                          int n_pairs, const int32_t *pa, const int32_t *pb, const double *w,
                          double *out, int64_t ld_out, void *stream) {
    AX_REQUIRE(n > 0 && n_pairs >= 0 && ld_pq >= n && ld_out >= n,
               "ax_pair_density: bad sizes");
    if (n_pairs == 0) return AX_OK;
    dim3 grid(ax::grid_for(n, 256, 64), n_pairs < 65535 ? n_pairs : 65535);
    hipLaunchKernelGGL(pair_density_kernel, grid, dim3(256), 0, ax::as_stream(stream), P, Q,
                       ld_pq, n, n_pairs, pa, pb, w, out, ld_out);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

ax_status ax_yk_batch(const double *rho, int64_t ld_rho, const double *w, const double *rk,
                      const double *rkm, int64_t n, int nv, int ng, int kmax, const double *mo,
                      const double *mi, int n_items, const int32_t *row, const int32_t *kk,
                      const int32_t *out_row, double *y, int64_t ld_y, void *stream) {
    AX_REQUIRE(ng >= 1 && ng <= kMaxNg, "ax_yk_batch: n_gauss must be in [1, %d], got %d",
               kMaxNg, ng);
    AX_REQUIRE(nv >= 1 && n == (int64_t)nv * ng, "ax_yk_batch: n (%lld) != nv*ng (%d*%d)",
               (long long)n, nv, ng);
    AX_REQUIRE(ld_rho >= n && ld_y >= n, "ax_yk_batch: leading dimension below n");
    AX_REQUIRE((size_t)2 * nv * sizeof(double) <= 48 * 1024,
               "ax_yk_batch: nv=%d exceeds the LDS cell-sum capacity", nv);
    if (n_items <= 0) return AX_OK;
    if (scan_lds_fits(nv, ng)) {
        const size_t lds = scan_lds_bytes(nv, ng);
        const int grid = n_items < 8192 ? n_items : 8192;
        hipStream_t s = ax::as_stream(stream);
        if (ng == 7)
            launch_scan_lds<7>(lds, grid, s, rho, ld_rho, w, rk, rkm, n, nv, kmax, mo, mi, n_items,
                               row, kk, out_row, y, ld_y);
        else if (ng == 8)
            launch_scan_lds<8>(lds, grid, s, rho, ld_rho, w, rk, rkm, n, nv, kmax, mo, mi, n_items,
                               row, kk, out_row, y, ld_y);
        else
            launch_scan_lds<9>(lds, grid, s, rho, ld_rho, w, rk, rkm, n, nv, kmax, mo, mi, n_items,
                               row, kk, out_row, y, ld_y);
        AX_CHECK_LAUNCH();
        return AX_OK;
    }
    const size_t shmem = (size_t)2 * nv * sizeof(double);
    const int grid = n_items < 4096 ? n_items : 4096;
    hipLaunchKernelGGL(yk_batch_kernel, dim3(grid), dim3(kThreads), shmem, ax::as_stream(stream),
                       rho, ld_rho, w, rk, rkm, n, nv, ng, kmax, mo, mi, n_items, row, kk, out_row,
                       y, ld_y);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

ax_status ax_rk_dot(const double *D, int64_t ld_d, const double *Y, int64_t ld_y, int64_t n,
                    int n_quart, const int32_t *p, const int32_t *q, double *out, void *stream) {
    AX_REQUIRE(n > 0 && n_quart >= 0, "ax_rk_dot: bad sizes");
    if (n_quart == 0) return AX_OK;
    const int grid = n_quart < 8192 ? n_quart : 8192;
    hipLaunchKernelGGL(rk_dot_kernel, dim3(grid), dim3(kThreads), 0, ax::as_stream(stream), D,
                       ld_d, Y, ld_y, n, n_quart, p, q, out);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

ax_status ax_bspline_eval(const double *coeffs, int64_t ld_c, int n_orb, const double *bsp,
                          int nv, int ng, int ks, int ns, double *out, int64_t ld_out,
                          void *stream) {
    AX_REQUIRE(nv > 0 && ng > 0 && ks > 0 && ns > 0 && n_orb >= 0 && ld_c >= ns &&
                   ld_out >= (int64_t)nv * ng,
               "ax_bspline_eval: bad sizes");
    if (n_orb == 0) return AX_OK;
    dim3 grid(ax::grid_for((int64_t)nv * ng, 256, 64), n_orb < 65535 ? n_orb : 65535);
    hipLaunchKernelGGL(bspline_eval_kernel, grid, dim3(256), 0, ax::as_stream(stream), coeffs,
                       ld_c, n_orb, bsp, nv, ng, ks, ns, out, ld_out);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

ax_status ax_assemble_integral(int nv, int ng, int ks1, int ks2, int mm, const double *bsp1,
                               const double *bsp2, const double *weights, int n_w, double *mat,
                               void *stream) {
    AX_REQUIRE(nv > 0 && ng > 0 && ks1 > 0 && ks2 > 0 && mm > 0 && n_w >= 0,
               "ax_assemble_integral: bad sizes (nv=%d ng=%d ks=%d,%d mm=%d)", nv, ng, ks1, ks2, mm);
    if (n_w == 0) return AX_OK;
    AX_REQUIRE(bsp1 && bsp2 && weights && mat, "ax_assemble_integral: null argument");
    const int64_t total = (int64_t)mm * mm * n_w;
    hipLaunchKernelGGL(assemble_integral_kernel, dim3(ax::grid_for(total, 256, 16384)), dim3(256),
                       0, ax::as_stream(stream), nv, ng, ks1, ks2, mm, n_w, bsp1, bsp2, weights,
                       mat);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

ax_status ax_symmetrize(double *mat, int mm, int n_mat, void *stream) {
    AX_REQUIRE(mm > 0 && n_mat >= 0, "ax_symmetrize: bad sizes");
    if (n_mat == 0) return AX_OK;
    const int64_t total = (int64_t)mm * mm * n_mat;
    hipLaunchKernelGGL(symmetrize_kernel, dim3(ax::grid_for(total, 256, 16384)), dim3(256), 0,
                       ax::as_stream(stream), mat, mm, n_mat);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

ax_status ax_exchange_matrix_yk(const double *fb, const double *bsp, const double *r,
                                const double *w, int nv, int ng, int ks, int ns, int k,
                                double *K, void *stream) {
    AX_REQUIRE(nv > 0 && ng > 0 && ks > 0 && ns > 0 && k >= 0, "ax_exchange_matrix_yk: bad sizes");
    const int64_t n = (int64_t)nv * ng;
    hipStream_t s = ax::as_stream(stream);
    // r^j and r^-(j+1) for j = 0..k; the item reads the rank-k rows
    double *tab = nullptr;
    AX_CHECK_HIP(hipMallocAsync(&tab, sizeof(double) * 2 * n * (k + 1), s));
    double *rk_all = tab, *rkm_all = tab + n * (k + 1);
    hipLaunchKernelGGL(rpow_kernel, dim3(ax::grid_for(n * (k + 1), 256, 4096)), dim3(256), 0, s, r,
                       n, k, rk_all, rkm_all);
    AX_CHECK_LAUNCH();
    ax_status st = ax_exchange_matrix_yk_batch(fb, n, bsp, w, rk_all, rkm_all, k, nv, ng, ks, ns, 1,
                                               nullptr, nullptr, K, stream);
    AX_CHECK_HIP(hipFreeAsync(tab, s));
    return st;
}

ax_status ax_exchange_matrix_yk_batch(const double *fb, int64_t ld_fb, const double *bsp,
                                      const double *w, const double *rk_tab,
                                      const double *rkm_tab, int kmax_tab, int nv, int ng, int ks,
                                      int ns, int n_items, const int32_t *item_b,
                                      const int32_t *item_k, double *K, void *stream) {
    AX_REQUIRE(nv > 0 && ng > 0 && ks > 0 && ns > 0 && kmax_tab >= 0 && n_items >= 0 &&
                   ld_fb >= (int64_t)nv * ng,
               "ax_exchange_matrix_yk_batch: bad sizes");
    AX_REQUIRE(ns <= nv + ks - 1, "ax_exchange_matrix_yk_batch: ns=%d exceeds nv+ks-1", ns);
    if (n_items == 0) return AX_OK;
    hipStream_t s = ax::as_stream(stream);
    const int L = ks * ng;
    double *prof = nullptr;
    const size_t bytes = sizeof(double) * (size_t)n_items * ns * (2 * L + 4);
    AX_CHECK_HIP(hipMallocAsync(&prof, bytes, s));
    const int64_t t1 = (int64_t)n_items * ns;
    hipLaunchKernelGGL(exchange_profiles_kernel, dim3(ax::grid_for(t1, 64, 65535)), dim3(64), 0, s,
                       fb, ld_fb, bsp, w, rk_tab, rkm_tab, kmax_tab, nv, ng, ks, ns, n_items,
                       item_b, item_k, prof);
    AX_CHECK_LAUNCH();
    const int64_t t2 = (int64_t)n_items * ns * ns;
    hipLaunchKernelGGL(exchange_assemble_kernel, dim3(ax::grid_for(t2, 256, 65535)), dim3(256), 0,
                       s, fb, ld_fb, bsp, w, rk_tab, rkm_tab, kmax_tab, nv, ng, ks, ns, n_items,
                       item_b, item_k, prof, K);
    AX_CHECK_LAUNCH();
    AX_CHECK_HIP(hipFreeAsync(prof, s));
    return AX_OK;
}

ax_status ax_slater_rk_direct(const double *rho13, const double *rho24, const double *gr,
                              const double *grw, int nv, int ng, int k, int order,
                              const double *xq, const double *wq, const double *outer,
                              const double *inner, double node0, double *workspace, double *out,
                              void *stream) {
    AX_REQUIRE(nv > 0 && ng > 0 && k >= 0 && order > 0 && order <= 64,
               "ax_slater_rk_direct: bad sizes (nv=%d ng=%d k=%d order=%d)", nv, ng, k, order);
    const int64_t n = (int64_t)nv * ng;
    const int g1 = n < 1024 ? static_cast<int>(n) : 1024;
    hipLaunchKernelGGL(direct_disjoint_kernel, dim3(g1), dim3(kThreads), 0, ax::as_stream(stream),
                       rho13, rho24, gr, grw, n, ng, k, workspace);
    AX_CHECK_LAUNCH();
    hipLaunchKernelGGL(direct_diag_kernel, dim3(nv < 4096 ? nv : 4096), dim3(kThreads), 0,
                       ax::as_stream(stream), rho13, rho24, gr, grw, nv, ng, k, order, xq, wq,
                       outer, inner, node0, workspace + 1024);
    AX_CHECK_LAUNCH();
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(kThreads), 0, ax::as_stream(stream),
                       workspace, g1, workspace + 1024, nv, out);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

}  // extern "C"
