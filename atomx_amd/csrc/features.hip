// Auxiliary-node form of the cell Coulomb quadrature and the pair-density
// matrix action built on it.
//
// Reference: numerics/bspline/coulomb_features.py:30-61 (_apply: cell
// moments, completed-cell prefix/suffix, symmetric same-cell edges),
// :281-294 (sample), numerics/bspline/pair_contraction.py:84-148 (moments,
// cumulative moments, edge batches).
//
// Layouts: a set of sampled columns is node-major (n_nodes, n_col) FP64, or
// complex FP64 as interleaved (re, im) pairs.  The Coulomb action of one
// rank is given by per-node slots into the (cell, outer node) weight tables
// and by the same-cell edges gathered by destination (CSR: for every node
// the (source node, weight) of each directed edge that ends there), so the
// action is a gather with no atomics and a fixed summation order.

#include <hip/hip_runtime.h>

#include <type_traits>

#include "ax_common.h"

namespace {

struct Cplx {
    double re, im;
};

__device__ inline double zero_of(double) { return 0.0; }
__device__ inline Cplx zero_of(Cplx) { return Cplx{0.0, 0.0}; }
__device__ inline double mul(double a, double b) { return a * b; }
__device__ inline Cplx mul(Cplx a, Cplx b) {
    return Cplx{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ inline double scale(double w, double a) { return w * a; }
__device__ inline Cplx scale(double w, Cplx a) { return Cplx{w * a.re, w * a.im}; }
__device__ inline double add(double a, double b) { return a + b; }
__device__ inline Cplx add(Cplx a, Cplx b) { return Cplx{a.re + b.re, a.im + b.im}; }

// ---------------------------------------------------------------------------
// Orbital samples at the auxiliary nodes of one component:
// out[node_base + cell*width + j, col] = sum_s coef[col, cell + s] * tab[cell, j, s]
// (coefficients beyond n_coef are zero, as evaluate.py:71-97).
// ---------------------------------------------------------------------------
template <typename T>
__global__ void feat_sample_kernel(const double *__restrict__ tab, int nv, int width, int order,
                                   int n_coef, const T *__restrict__ coef, int64_t ld_coef,
                                   int n_col, T *__restrict__ out, int64_t ld_out,
                                   int64_t node_base) {
    const int64_t total = (int64_t)nv * width * n_col;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int col = (int)(t % n_col);
        const int64_t node = t / n_col;
        const int cell = (int)(node / width);
        const double *b = tab + node * order;
        const T *c = coef + col * ld_coef;
        T acc = zero_of(T{});
        for (int s = 0; s < order; ++s) {
            const int i = cell + s;
            if (i < n_coef) acc = add(acc, scale(b[s], c[i]));
        }
        out[(node_base + node) * ld_out + col] = acc;
    }
}

// ---------------------------------------------------------------------------
// Coulomb action, step 1: per column, the rising/falling moments of every
// cell (both components' outer nodes) and their exclusive prefix (rising,
// cells before) and suffix (falling, cells after).  One work-group per
// column; each thread owns a contiguous run of cells.
// ---------------------------------------------------------------------------
constexpr int kScanThreads = 256;

__global__ void __launch_bounds__(kScanThreads)
feat_moments_kernel(const double *__restrict__ rho, int64_t ld, int n_col,
                    const int *__restrict__ outer, const double *__restrict__ rising,
                    const double *__restrict__ falling, int nv, int ngo,
                    double *__restrict__ prefix, double *__restrict__ suffix) {
    __shared__ double s_rise[kScanThreads], s_fall[kScanThreads];
    const int col = blockIdx.x;
    if (col >= n_col) return;
    const int tid = threadIdx.x;
    const int per = (nv + kScanThreads - 1) / kScanThreads;
    const int c0 = min(nv, tid * per), c1 = min(nv, c0 + per);
    // per-cell moments (kept in the output slots until the scan) and the
    // run totals of this thread's cells
    double rt = 0.0, ft = 0.0;
    for (int c = c0; c < c1; ++c) {
        double r = 0.0, f = 0.0;
        for (int comp = 0; comp < 2; ++comp)
            for (int g = 0; g < ngo; ++g) {
                const double v = rho[(int64_t)outer[(comp * nv + c) * ngo + g] * ld + col];
                r += rising[c * ngo + g] * v;
                f += falling[c * ngo + g] * v;
            }
        prefix[(int64_t)c * n_col + col] = r;
        suffix[(int64_t)c * n_col + col] = f;
        rt += r;
        ft += f;
    }
    s_rise[tid] = rt;
    s_fall[tid] = ft;
    __syncthreads();
    // exclusive prefix of the rising run totals, exclusive suffix of the
    // falling ones (sequential over the 256 runs: fixed order)
    if (tid == 0) {
        double acc = 0.0;
        for (int i = 0; i < kScanThreads; ++i) {
            const double v = s_rise[i];
            s_rise[i] = acc;
            acc += v;
        }
    } else if (tid == 64) {
        double acc = 0.0;
        for (int i = kScanThreads - 1; i >= 0; --i) {
            const double v = s_fall[i];
            s_fall[i] = acc;
            acc += v;
        }
    }
    __syncthreads();
    // each thread rereads only the slots it wrote itself
    double pre = s_rise[tid];
    for (int c = c0; c < c1; ++c) {
        const double r = prefix[(int64_t)c * n_col + col];
        prefix[(int64_t)c * n_col + col] = pre;
        pre += r;
    }
    double suf = s_fall[tid];
    for (int c = c1 - 1; c >= c0; --c) {
        const double f = suffix[(int64_t)c * n_col + col];
        suffix[(int64_t)c * n_col + col] = suf;
        suf += f;
    }
}

// Coulomb action, step 2: out[i, j] = (outer node: falling * prefix + rising
// * suffix of its cell) + sum over the edges ending at i of w * rho[src, j].
__global__ void feat_apply_kernel(const double *__restrict__ rho, int64_t ld, int n_col,
                                  int64_t n_nodes, const int *__restrict__ slot,
                                  const double *__restrict__ rising,
                                  const double *__restrict__ falling, int ngo,
                                  const double *__restrict__ prefix,
                                  const double *__restrict__ suffix,
                                  const int64_t *__restrict__ row_ptr,
                                  const int *__restrict__ src, const double *__restrict__ w,
                                  double *__restrict__ out, int64_t ld_out) {
    const int64_t total = n_nodes * n_col;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int col = (int)(t % n_col);
        const int64_t i = t / n_col;
        double acc = 0.0;
        const int s = slot[i];
        if (s >= 0) {
            const int c = s / ngo;
            acc = falling[s] * prefix[(int64_t)c * n_col + col];
            acc += rising[s] * suffix[(int64_t)c * n_col + col];
        }
        for (int64_t e = row_ptr[i]; e < row_ptr[i + 1]; ++e)
            acc += w[e] * rho[(int64_t)src[e] * ld + col];
        out[i * ld_out + col] = acc;
    }
}

// ---------------------------------------------------------------------------
// Pair contraction.
// moments: out[(c*d1 + i)*d2 + j] = sum over the 2*ngo outer nodes o of cell
// c of w[c, g] * Y[o, i] * X[o, j]  (X: (N, d2), Y: (N, d1), node-major).
// ---------------------------------------------------------------------------
template <typename T>
__global__ void cpc_moments_kernel(const T *__restrict__ X, int64_t ldx, int d2,
                                   const T *__restrict__ Y, int64_t ldy, int d1,
                                   const int *__restrict__ outer, const double *__restrict__ wt,
                                   int nv, int ngo, T *__restrict__ out) {
    const int64_t total = (int64_t)nv * d1 * d2;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(t % d2);
        const int64_t ci = t / d2;
        const int i = (int)(ci % d1);
        const int c = (int)(ci / d1);
        T acc = zero_of(T{});
        for (int comp = 0; comp < 2; ++comp)
            for (int g = 0; g < ngo; ++g) {
                const int64_t o = outer[(comp * nv + c) * ngo + g];
                acc = add(acc, scale(wt[c * ngo + g], mul(Y[o * ldy + i], X[o * ldx + j])));
            }
        out[t] = acc;
    }
}

// prefix[c] = sum_{c' < c} rise[c'], suffix[c] = sum_{c' > c} fall[c'] for
// (nv, n_el) stacks (sequential over cells).
template <typename T>
__global__ void cpc_scan_kernel(const T *__restrict__ rise, const T *__restrict__ fall, int nv,
                                int64_t n_el, T *__restrict__ prefix, T *__restrict__ suffix) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n_el;
         e += (int64_t)gridDim.x * blockDim.x) {
        T acc = zero_of(T{});
        for (int c = 0; c < nv; ++c) {
            prefix[c * n_el + e] = acc;
            acc = add(acc, rise[c * n_el + e]);
        }
        acc = zero_of(T{});
        for (int c = nv - 1; c >= 0; --c) {
            suffix[c * n_el + e] = acc;
            acc = add(acc, fall[c * n_el + e]);
        }
    }
}

// Z[c][r][n] = sum_s rho[r, s] * S[c][s][n]  (batched over cells).
template <typename T>
__global__ void cpc_transform_kernel(const T *__restrict__ rho, int nr, int ns_,
                                     const T *__restrict__ S, int nv, int nn,
                                     T *__restrict__ Z) {
    const int64_t total = (int64_t)nv * nr * nn;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int n = (int)(t % nn);
        const int64_t cr = t / nn;
        const int r = (int)(cr % nr);
        const int c = (int)(cr / nr);
        T acc = zero_of(T{});
        for (int s = 0; s < ns_; ++s)
            acc = add(acc, mul(rho[r * ns_ + s], S[((int64_t)c * ns_ + s) * nn + n]));
        Z[t] = acc;
    }
}

// BP[i, s] = sum_r B[i, r] rho[r, s].
template <typename T>
__global__ void cpc_bp_kernel(const T *__restrict__ B, int64_t ldb, int nr,
                              const T *__restrict__ rho, int ns_, int64_t n_nodes,
                              T *__restrict__ BP) {
    const int64_t total = n_nodes * ns_;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int s = (int)(t % ns_);
        const int64_t i = t / ns_;
        T acc = zero_of(T{});
        for (int r = 0; r < nr; ++r) acc = add(acc, mul(B[i * ldb + r], rho[r * ns_ + s]));
        BP[t] = acc;
    }
}

// coef[e] = w[e] * sum_s BP[src_e, s] * C[dst_e, s]  (one directed edge each).
template <typename T>
__global__ void cpc_edge_coef_kernel(const T *__restrict__ BP, int ns_,
                                     const T *__restrict__ C, int64_t ldc,
                                     const int *__restrict__ src, const int *__restrict__ dst,
                                     const double *__restrict__ w, int64_t n_edges,
                                     T *__restrict__ coef) {
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n_edges;
         e += (int64_t)gridDim.x * blockDim.x) {
        const T *a = BP + (int64_t)src[e] * ns_;
        const T *b = C + (int64_t)dst[e] * ldc;
        T acc = zero_of(T{});
        for (int s = 0; s < ns_; ++s) acc = add(acc, mul(a[s], b[s]));
        coef[e] = scale(w[e], acc);
    }
}

// Gathered, weighted product, split over K:
// part[split][i][j] = sum_{k in chunk} coef[k] * L[li(k), i] * R[ri(k), j]
// with li/ri the identity when the index vector is null and coef 1 when null.
template <typename T>
__global__ void cpc_gemm_kernel(int64_t K, int chunk, const T *__restrict__ L, int64_t ldl,
                                int m, const int *__restrict__ li, const T *__restrict__ R,
                                int64_t ldr, int n, const int *__restrict__ ri,
                                const T *__restrict__ coef, T *__restrict__ part) {
    const int split = blockIdx.y;
    const int64_t k0 = (int64_t)split * chunk;
    const int64_t k1 = min(K, k0 + chunk);
    const int64_t mn = (int64_t)m * n;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < mn;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int j = (int)(t % n);
        const int i = (int)(t / n);
        T acc = zero_of(T{});
        for (int64_t k = k0; k < k1; ++k) {
            const int64_t a = li ? li[k] : k;
            const int64_t b = ri ? ri[k] : k;
            T p = mul(L[a * ldl + i], R[b * ldr + j]);
            if (coef) p = mul(coef[k], p);
            acc = add(acc, p);
        }
        part[(int64_t)split * mn + t] = acc;
    }
}

// The same product for real operands on the FP64 MFMA: one 64x64 output tile
// and one K chunk per work-group.  K is staged through LDS 16 rows at a time
// (row li[k] of L scaled by coef[k], row ri[k] of R, columns of the tile,
// zero outside the matrix or past the chunk; rows padded by 8 doubles), and
// each wave accumulates a 32x32 quadrant as 2x2 blocks of
// v_mfma_f64_16x16x4_f64: lane (f, g) feeds L^T row f and R column f at K
// row g of each group of 4 and holds output rows g + 4 r, column f.
constexpr int kCpcTile = 64;
constexpr int kCpcK = 16;
constexpr int kCpcPad = kCpcTile + 8;

typedef double cpc_d4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void cpc_gemm_mfma_kernel(
    int64_t K, int chunk, const double *__restrict__ L, int64_t ldl, int m,
    const int *__restrict__ li, const double *__restrict__ R, int64_t ldr, int n,
    const int *__restrict__ ri, const double *__restrict__ coef, double *__restrict__ part) {
    __shared__ double Ls[kCpcK][kCpcPad], Rs[kCpcK][kCpcPad];
    const int tiles_n = (n + kCpcTile - 1) / kCpcTile;
    const int m0 = (int)(blockIdx.x / tiles_n) * kCpcTile;
    const int n0 = (int)(blockIdx.x % tiles_n) * kCpcTile;
    const int split = blockIdx.y;
    const int64_t k0 = (int64_t)split * chunk;
    const int64_t k1 = min(K, k0 + chunk);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = (wave >> 1) * 32, wc = (wave & 1) * 32;
    const int fr = lane & 15, fg = lane >> 4;
    cpc_d4 acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b) acc[a][b] = cpc_d4{0.0, 0.0, 0.0, 0.0};
    for (int64_t kb = k0; kb < k1; kb += kCpcK) {
#pragma unroll
        for (int q = 0; q < kCpcK * kCpcTile / 256; ++q) {
            const int e = tid + 256 * q, kk = e / kCpcTile, c = e % kCpcTile;
            const int64_t k = kb + kk;
            double lv = 0.0, rv = 0.0;
            if (k < k1) {
                const int64_t ra = li ? (int64_t)li[k] : k, rb = ri ? (int64_t)ri[k] : k;
                if (m0 + c < m) lv = L[ra * ldl + m0 + c] * (coef ? coef[k] : 1.0);
                if (n0 + c < n) rv = R[rb * ldr + n0 + c];
            }
            Ls[kk][c] = lv;
            Rs[kk][c] = rv;
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < kCpcK / 4; ++s) {
            const int kk = 4 * s + fg;
            double av[2], bv[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) av[a] = Ls[kk][wr + 16 * a + fr];
#pragma unroll
            for (int b = 0; b < 2; ++b) bv[b] = Rs[kk][wc + 16 * b + fr];
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < 2; ++b)
                    acc[a][b] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[a], bv[b], acc[a][b], 0, 0,
                                                                     0);
        }
        __syncthreads();
    }
    double *dst = part + (int64_t)split * m * n;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < 2; ++b)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int i = m0 + wr + 16 * a + fg + 4 * r, j = n0 + wc + 16 * b + fr;
                if (i < m && j < n) dst[(int64_t)i * n + j] = acc[a][b][r];
            }
}

// cpc_gemm_kernel for complex operands, the MFMA form for real ones.
template <typename T>
void launch_cpc_gemm(int64_t K, int chunk, int n_split, const T *L, int64_t ldl, int m,
                     const int *li, const T *R, int64_t ldr, int n, const int *ri, const T *coef,
                     T *part, hipStream_t st) {
    if constexpr (std::is_same<T, double>::value) {
        const int tiles = ((m + kCpcTile - 1) / kCpcTile) * ((n + kCpcTile - 1) / kCpcTile);
        hipLaunchKernelGGL(cpc_gemm_mfma_kernel, dim3(tiles, n_split), dim3(256), 0, st, K, chunk,
                           L, ldl, m, li, R, ldr, n, ri, coef, part);
    } else {
        const dim3 gmn(ax::grid_for((int64_t)m * n, 256, 64));
        hipLaunchKernelGGL(cpc_gemm_kernel<T>, dim3(gmn.x, n_split), dim3(256), 0, st, K,
                           chunk, L, ldl, m, li, R, ldr, n, ri, coef, part);
    }
}

// out[t] (+)= sum over splits in order.
template <typename T>
__global__ void cpc_reduce_kernel(const T *__restrict__ part, int n_split, int64_t mn,
                                  int accumulate, T *__restrict__ out) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < mn;
         t += (int64_t)gridDim.x * blockDim.x) {
        T acc = accumulate ? out[t] : zero_of(T{});
        for (int s = 0; s < n_split; ++s) acc = add(acc, part[(int64_t)s * mn + t]);
        out[t] = acc;
    }
}

// out[i, p] = S[i, ia[p]] * S[i, ib[p]]  (pair densities of sampled columns,
// c-product).
template <typename T>
__global__ void feat_products_kernel(const T *__restrict__ S, int64_t lds, int64_t n_nodes,
                                     int n_pairs, const int *__restrict__ ia,
                                     const int *__restrict__ ib, T *__restrict__ out,
                                     int64_t ld_out) {
    const int64_t total = n_nodes * n_pairs;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int p = (int)(t % n_pairs);
        const int64_t i = t / n_pairs;
        out[i * ld_out + p] = mul(S[i * lds + ia[p]], S[i * lds + ib[p]]);
    }
}

// out[p] = sum_i X[i, ix[p]] * Y[i, iy[p]]: one work-group per pair, fixed
// reduction order.
constexpr int kDotThreads = 256;

template <typename T>
__global__ void __launch_bounds__(kDotThreads)
feat_dots_kernel(const T *__restrict__ X, int64_t ldx, const T *__restrict__ Y, int64_t ldy,
                 int64_t n_nodes, const int *__restrict__ ix, const int *__restrict__ iy,
                 T *__restrict__ out) {
    __shared__ T red[kDotThreads];
    const int p = blockIdx.x;
    const int a = ix[p], b = iy[p];
    T acc = zero_of(T{});
    for (int64_t i = threadIdx.x; i < n_nodes; i += kDotThreads)
        acc = add(acc, mul(X[i * ldx + a], Y[i * ldy + b]));
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int h = kDotThreads / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] = add(red[threadIdx.x], red[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[p] = red[0];
}

constexpr int kThreads = 256;

template <typename T>
int launch_products(const void *S, int64_t lds, int64_t n_nodes, int n_pairs, const int *ia,
                    const int *ib, void *out, int64_t ld_out, hipStream_t st) {
    hipLaunchKernelGGL(feat_products_kernel<T>, dim3(ax::grid_for(n_nodes * n_pairs, kThreads)),
                       dim3(kThreads), 0, st, static_cast<const T *>(S), lds, n_nodes, n_pairs,
                       ia, ib, static_cast<T *>(out), ld_out);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

template <typename T>
int launch_dots(const void *X, int64_t ldx, const void *Y, int64_t ldy, int64_t n_nodes,
                int n_pairs, const int *ix, const int *iy, void *out, hipStream_t st) {
    hipLaunchKernelGGL(feat_dots_kernel<T>, dim3(n_pairs), dim3(kDotThreads), 0, st,
                       static_cast<const T *>(X), ldx, static_cast<const T *>(Y), ldy, n_nodes,
                       ix, iy, static_cast<T *>(out));
    AX_CHECK_LAUNCH();
    return AX_OK;
}

template <typename T>
int launch_sample(const double *tab, int nv, int width, int order, int n_coef, const void *coef,
                  int64_t ld_coef, int n_col, void *out, int64_t ld_out, int64_t node_base,
                  hipStream_t st) {
    const int64_t work = (int64_t)nv * width * n_col;
    hipLaunchKernelGGL(feat_sample_kernel<T>, dim3(ax::grid_for(work, kThreads)), dim3(kThreads),
                       0, st, tab, nv, width, order, n_coef, static_cast<const T *>(coef),
                       ld_coef, n_col, static_cast<T *>(out), ld_out, node_base);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

}  // namespace

extern "C" int ax_feat_sample(const double *tab, int nv, int width, int order, int n_coef,
                              const void *coef, int64_t ld_coef, int n_col, int is_complex,
                              void *out, int64_t ld_out, int64_t node_base, void *stream) {
    AX_REQUIRE(tab && coef && out, "ax_feat_sample: null pointer");
    AX_REQUIRE(nv > 0 && width > 0 && order > 0 && n_coef > 0 && n_col > 0,
               "ax_feat_sample: invalid sizes");
    AX_REQUIRE(ld_coef >= n_coef && ld_out >= n_col && node_base >= 0,
               "ax_feat_sample: invalid leading dimensions");
    hipStream_t st = ax::as_stream(stream);
    return is_complex ? launch_sample<Cplx>(tab, nv, width, order, n_coef, coef, ld_coef, n_col,
                                            out, ld_out, node_base, st)
                      : launch_sample<double>(tab, nv, width, order, n_coef, coef, ld_coef,
                                              n_col, out, ld_out, node_base, st);
}

extern "C" int ax_feat_products(const void *S, int64_t lds, int64_t n_nodes, int n_pairs,
                                const int *ia, const int *ib, int is_complex, void *out,
                                int64_t ld_out, void *stream) {
    AX_REQUIRE(S && ia && ib && out, "ax_feat_products: null pointer");
    AX_REQUIRE(n_nodes > 0 && n_pairs > 0 && ld_out >= n_pairs,
               "ax_feat_products: invalid sizes");
    hipStream_t st = ax::as_stream(stream);
    return is_complex ? launch_products<Cplx>(S, lds, n_nodes, n_pairs, ia, ib, out, ld_out, st)
                      : launch_products<double>(S, lds, n_nodes, n_pairs, ia, ib, out, ld_out,
                                                st);
}

extern "C" int ax_feat_dots(const void *X, int64_t ldx, const void *Y, int64_t ldy,
                            int64_t n_nodes, int n_pairs, const int *ix, const int *iy,
                            int is_complex, void *out, void *stream) {
    AX_REQUIRE(X && Y && ix && iy && out, "ax_feat_dots: null pointer");
    AX_REQUIRE(n_nodes > 0 && n_pairs > 0, "ax_feat_dots: invalid sizes");
    hipStream_t st = ax::as_stream(stream);
    return is_complex ? launch_dots<Cplx>(X, ldx, Y, ldy, n_nodes, n_pairs, ix, iy, out, st)
                      : launch_dots<double>(X, ldx, Y, ldy, n_nodes, n_pairs, ix, iy, out, st);
}

extern "C" int ax_feat_apply(const double *rho, int64_t ld, int n_col, int64_t n_nodes,
                             const int *outer, const int *slot, const double *rising,
                             const double *falling, int nv, int ngo, const int64_t *row_ptr,
                             const int *src, const double *w, double *work, double *out,
                             int64_t ld_out, void *stream) {
    AX_REQUIRE(rho && outer && slot && rising && falling && row_ptr && work && out,
               "ax_feat_apply: null pointer");
    AX_REQUIRE(n_col > 0 && n_nodes > 0 && nv > 0 && ngo > 0 && ld >= n_col && ld_out >= n_col,
               "ax_feat_apply: invalid sizes");
    AX_REQUIRE(out != rho, "ax_feat_apply: output must not alias the density");
    hipStream_t st = ax::as_stream(stream);
    double *prefix = work;
    double *suffix = work + (int64_t)nv * n_col;
    hipLaunchKernelGGL(feat_moments_kernel, dim3(n_col), dim3(kScanThreads), 0, st, rho, ld,
                       n_col, outer, rising, falling, nv, ngo, prefix, suffix);
    AX_CHECK_LAUNCH();
    const int64_t work_items = n_nodes * n_col;
    hipLaunchKernelGGL(feat_apply_kernel, dim3(ax::grid_for(work_items, kThreads)),
                       dim3(kThreads), 0, st, rho, ld, n_col, n_nodes, slot, rising, falling,
                       ngo, prefix, suffix, row_ptr, src, w, out, ld_out);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

extern "C" int64_t ax_feat_apply_workspace(int nv, int n_col) {
    if (nv <= 0 || n_col <= 0) return -1;
    return (int64_t)2 * nv * n_col * (int64_t)sizeof(double);
}

namespace {

template <typename T>
int cpc_moments(const void *X, int64_t ldx, int d2, const void *Y, int64_t ldy, int d1,
                const int *outer, const double *wr, const double *wf, int nv, int ngo,
                void *rise, void *fall, hipStream_t st) {
    const int64_t work = (int64_t)nv * d1 * d2;
    const dim3 grid(ax::grid_for(work, kThreads));
    hipLaunchKernelGGL(cpc_moments_kernel<T>, grid, dim3(kThreads), 0, st,
                       static_cast<const T *>(X), ldx, d2, static_cast<const T *>(Y), ldy, d1,
                       outer, wr, nv, ngo, static_cast<T *>(rise));
    hipLaunchKernelGGL(cpc_moments_kernel<T>, grid, dim3(kThreads), 0, st,
                       static_cast<const T *>(X), ldx, d2, static_cast<const T *>(Y), ldy, d1,
                       outer, wf, nv, ngo, static_cast<T *>(fall));
    AX_CHECK_LAUNCH();
    return AX_OK;
}

template <typename T>
int cpc_scan(const void *rise, const void *fall, int nv, int64_t n_el, void *prefix,
             void *suffix, hipStream_t st) {
    hipLaunchKernelGGL(cpc_scan_kernel<T>, dim3(ax::grid_for(n_el, kThreads)), dim3(kThreads), 0,
                       st, static_cast<const T *>(rise), static_cast<const T *>(fall), nv, n_el,
                       static_cast<T *>(prefix), static_cast<T *>(suffix));
    AX_CHECK_LAUNCH();
    return AX_OK;
}

template <typename T>
int cpc_apply(const void *rho, int nr, int ns_, const void *A, int64_t lda, int m,
              const void *B, int64_t ldb, const void *C, int64_t ldc, const void *D,
              int64_t ldd, int n, int64_t n_nodes, int nv, const void *left_stack,
              const void *sfx, const void *pfx, const int *src, const int *dst, const double *w,
              int64_t n_edges, int n_split, void *work, void *out, hipStream_t st) {
    const T *rho_ = static_cast<const T *>(rho);
    T *Z = static_cast<T *>(work);                       // (2, nv, nr, n)
    T *BP = Z + (int64_t)2 * nv * nr * n;                // (n_nodes, ns)
    T *coef = BP + n_nodes * ns_;                        // (n_edges)
    T *part = coef + (n_edges > 0 ? n_edges : 1);        // (n_split, m, n)
    T *out_ = static_cast<T *>(out);
    const int64_t zn = (int64_t)nv * nr * n;
    const dim3 gz(ax::grid_for(zn, kThreads));
    hipLaunchKernelGGL(cpc_transform_kernel<T>, gz, dim3(kThreads), 0, st, rho_, nr, ns_,
                       static_cast<const T *>(sfx), nv, n, Z);
    hipLaunchKernelGGL(cpc_transform_kernel<T>, gz, dim3(kThreads), 0, st, rho_, nr, ns_,
                       static_cast<const T *>(pfx), nv, n, Z + zn);
    AX_CHECK_LAUNCH();
    const int64_t mn = (int64_t)m * n;
    // cross-cell part: [L_rise; L_fall] (2 nv nr, m) against [Z_sfx; Z_pfx]
    {
        const int64_t K = 2 * zn / n;
        const int chunk = (int)((K + n_split - 1) / n_split);
        launch_cpc_gemm<T>(K, chunk, n_split, static_cast<const T *>(left_stack), (int64_t)m, m,
                           nullptr, Z, (int64_t)n, n, nullptr, nullptr, part, st);
        hipLaunchKernelGGL(cpc_reduce_kernel<T>, dim3(ax::grid_for(mn, kThreads)),
                           dim3(kThreads), 0, st, part, n_split, mn, 0, out_);
        AX_CHECK_LAUNCH();
    }
    if (n_edges > 0) {
        hipLaunchKernelGGL(cpc_bp_kernel<T>, dim3(ax::grid_for(n_nodes * ns_, kThreads)),
                           dim3(kThreads), 0, st, static_cast<const T *>(B), ldb, nr, rho_, ns_,
                           n_nodes, BP);
        hipLaunchKernelGGL(cpc_edge_coef_kernel<T>, dim3(ax::grid_for(n_edges, kThreads)),
                           dim3(kThreads), 0, st, BP, ns_, static_cast<const T *>(C), ldc, src,
                           dst, w, n_edges, coef);
        const int chunk = (int)((n_edges + n_split - 1) / n_split);
        launch_cpc_gemm<T>(n_edges, chunk, n_split, static_cast<const T *>(A), lda, m, src,
                           static_cast<const T *>(D), ldd, n, dst, coef, part, st);
        hipLaunchKernelGGL(cpc_reduce_kernel<T>, dim3(ax::grid_for(mn, kThreads)),
                           dim3(kThreads), 0, st, part, n_split, mn, 1, out_);
        AX_CHECK_LAUNCH();
    }
    return AX_OK;
}

}  // namespace

extern "C" int ax_cpc_moments(const void *X, int64_t ldx, int d2, const void *Y, int64_t ldy,
                              int d1, const int *outer, const double *wr, const double *wf,
                              int nv, int ngo, int is_complex, void *rise, void *fall,
                              void *stream) {
    AX_REQUIRE(X && Y && outer && wr && wf && rise && fall, "ax_cpc_moments: null pointer");
    AX_REQUIRE(d1 > 0 && d2 > 0 && nv > 0 && ngo > 0 && ldx >= d2 && ldy >= d1,
               "ax_cpc_moments: invalid sizes");
    hipStream_t st = ax::as_stream(stream);
    return is_complex
               ? cpc_moments<Cplx>(X, ldx, d2, Y, ldy, d1, outer, wr, wf, nv, ngo, rise, fall, st)
               : cpc_moments<double>(X, ldx, d2, Y, ldy, d1, outer, wr, wf, nv, ngo, rise, fall,
                                     st);
}

extern "C" int ax_cpc_scan(const void *rise, const void *fall, int nv, int64_t n_el,
                           int is_complex, void *prefix, void *suffix, void *stream) {
    AX_REQUIRE(rise && fall && prefix && suffix, "ax_cpc_scan: null pointer");
    AX_REQUIRE(nv > 0 && n_el > 0, "ax_cpc_scan: invalid sizes");
    hipStream_t st = ax::as_stream(stream);
    return is_complex ? cpc_scan<Cplx>(rise, fall, nv, n_el, prefix, suffix, st)
                      : cpc_scan<double>(rise, fall, nv, n_el, prefix, suffix, st);
}

extern "C" int64_t ax_cpc_workspace_bytes(int nv, int nr, int ns_, int m, int n,
                                          int64_t n_nodes, int64_t n_edges, int n_split,
                                          int is_complex) {
    if (nv <= 0 || nr <= 0 || ns_ <= 0 || m <= 0 || n <= 0 || n_nodes <= 0 || n_edges < 0 ||
        n_split <= 0)
        return -1;
    const int64_t el = (int64_t)2 * nv * nr * n + n_nodes * ns_ + (n_edges > 0 ? n_edges : 1) +
                       (int64_t)n_split * m * n;
    return el * (is_complex ? 16 : 8);
}

extern "C" int ax_cpc_apply(const void *rho, int nr, int ns_, const void *A, int64_t lda, int m,
                            const void *B, int64_t ldb, const void *C, int64_t ldc,
                            const void *D, int64_t ldd, int n, int64_t n_nodes, int nv,
                            const void *left_stack, const void *sfx, const void *pfx,
                            const int *src, const int *dst, const double *w, int64_t n_edges,
                            int n_split, int is_complex, void *work, void *out, void *stream) {
    AX_REQUIRE(rho && A && B && C && D && left_stack && sfx && pfx && work && out,
               "ax_cpc_apply: null pointer");
    AX_REQUIRE(nr > 0 && ns_ > 0 && m > 0 && n > 0 && n_nodes > 0 && nv > 0 && n_split > 0 &&
                   n_split <= 4096 && n_edges >= 0,
               "ax_cpc_apply: invalid sizes");
    AX_REQUIRE(n_edges == 0 || (src && dst && w), "ax_cpc_apply: null edge arrays");
    AX_REQUIRE(lda >= m && ldb >= nr && ldc >= ns_ && ldd >= n,
               "ax_cpc_apply: invalid leading dimensions");
    hipStream_t st = ax::as_stream(stream);
    return is_complex ? cpc_apply<Cplx>(rho, nr, ns_, A, lda, m, B, ldb, C, ldc, D, ldd, n,
                                        n_nodes, nv, left_stack, sfx, pfx, src, dst, w, n_edges,
                                        n_split, work, out, st)
                      : cpc_apply<double>(rho, nr, ns_, A, lda, m, B, ldb, C, ldc, D, ldd, n,
                                          n_nodes, nv, left_stack, sfx, pfx, src, dst, w,
                                          n_edges, n_split, work, out, st);
}
