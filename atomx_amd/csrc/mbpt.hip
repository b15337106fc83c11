// Second-order MBPT pair sums on a closed-shell reference.
//
// Inputs prepared on the device by the caller:
//   * for every (occupied pair, rank k) item, the Slater table
//     T[r][s] = R^k(a r, b s) over all N virtual orbitals (a GEMM product of
//     the bilinear form M^k_ab with the virtual coefficient columns);
//   * per pair type (kappa_a, kappa_b) and virtual channel (kappa_r,
//     kappa_s), the J-coupling coefficients of methods/mbpt.py:752-790:
//       cd[J][k] = (-1)^((j_b + j_r + J)) {j_a j_b J; j_s j_r k} C^k_ar C^k_bs
//       ce[J][k] = (-1)^((j_r + j_s - J)) (-1)^((j_b + j_s + J))
//                  {j_a j_b J; j_r j_s k} C^k_as C^k_br
//     (zero where the reference drops the rank), the first 2J and the J
//     count (0: the channel contributes nothing).
// Per virtual pair r <= s (virtuals ordered by kappa channel, so r <= s is
// the reference's unordered channel sum with r <= s inside one kappa):
//   amp_J = sum_k cd[J][k] T^k[r][s] - ce[J][k] T^k[s][r]
//   E_rs  = -sum_J (2J+1) amp_J^2 / (n_ab n_rs (e_r + e_s - e_a - e_b)),
// skipped where |e_r + e_s - e_a - e_b| < 1e-10; n = 2 for a repeated
// orbital.  Work-group partial sums are reduced per pair in a fixed order.
#include "ax_common.h"

namespace {

constexpr int kMaxJ = 8;   // J values of one pair type: (2 j_a + 2 j_b) / 2 + 1 <= 8 for j <= 7/2
constexpr int kTile = 16;  // 16 x 16 (r, s) tile per work-group

// Stage 2 of the bilinear R^k forms (numerics/bspline/convol.py:1900-1978),
// the gather over the right band pair in the reference's j-ascending order:
//   M[m][n] = sum_{j<n} T[m][j][n-j] wb[j] + sum_{jp} T[m][n][jp] wb[n+jp],
// j + jp < n_right.  Job descriptors (int64 x 5): T, wb, M addresses, ksr,
// n_right; M has row stride ld.
__global__ void bilinear_batch_kernel(const int64_t *__restrict__ jobs, int ns, int nm, int ld) {
    const int64_t *job = jobs + 5 * (int64_t)blockIdx.y;
    const double *T = reinterpret_cast<const double *>(job[0]);
    const double *wb = reinterpret_cast<const double *>(job[1]);
    double *M = reinterpret_cast<double *>(job[2]);
    const int ksr = static_cast<int>(job[3]), n_right = static_cast<int>(job[4]);
    const int64_t total = (int64_t)nm * ns;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int m = static_cast<int>(t / ns), n = static_cast<int>(t - (int64_t)m * ns);
        const double *Tm = T + (int64_t)m * ns * ksr;
        double s = 0.0;
        if (n < n_right) {
            for (int jp = ksr - 1; jp >= 1; --jp) {
                const int j = n - jp;
                if (j < 0) continue;
                s += Tm[(int64_t)j * ksr + jp] * wb[j];
            }
            for (int jp = 0; jp < ksr; ++jp) {
                if (n + jp >= n_right) break;
                s += Tm[(int64_t)n * ksr + jp] * wb[n + jp];
            }
        }
        M[(int64_t)m * ld + n] = s;
    }
}

struct PairSum {
    int n_virt, n_kappa, n_rank, n_pair, max_j, n_tiles, tiles_per_row;
    const int32_t *vkappa;
    const double *veps;
    const double *tables;      // entry (item, r, s) at item * item_stride + r * row_stride + s
    int64_t item_stride, row_stride;
    const int32_t *pair_item;  // n_pair x n_rank: table item, -1 if absent
    const double *pair_eps, *pair_norm;
    const int32_t *pair_type;
    const int32_t *chan_nj, *chan_tjlo;
    const double *cd, *ce;     // [chan][max_j][n_rank]
    double *partial;           // n_pair x n_tiles
    double *jsum;              // optional n_pair x N x N (r <= s entries written)
};

__device__ inline void tile_of(int idx, int nt, int &tr, int &ts) {
    tr = 0;
    while (idx >= nt - tr) {
        idx -= nt - tr;
        ++tr;
    }
    ts = tr + idx;
}

__global__ __launch_bounds__(256) void pair_sum_kernel(PairSum A) {
    const int pair = blockIdx.y;
    int tr, ts;
    tile_of(blockIdx.x, A.tiles_per_row, tr, ts);
    const int r = tr * kTile + static_cast<int>(threadIdx.x >> 4);
    const int s = ts * kTile + static_cast<int>(threadIdx.x & 15);
    const int N = A.n_virt;
    double contrib = 0.0;
    if (r < N && s < N && r <= s) {
        const int chan = (A.pair_type[pair] * A.n_kappa + A.vkappa[r]) * A.n_kappa + A.vkappa[s];
        const int nj = A.chan_nj[chan];
        const double de = A.veps[r] + A.veps[s] - A.pair_eps[pair];
        if (nj > 0 && fabs(de) >= 1e-10) {
            double amp[kMaxJ];
#pragma unroll
            for (int j = 0; j < kMaxJ; ++j) amp[j] = 0.0;
            const double *cdc = A.cd + (int64_t)chan * A.max_j * A.n_rank;
            const double *cec = A.ce + (int64_t)chan * A.max_j * A.n_rank;
            const int32_t *items = A.pair_item + (int64_t)pair * A.n_rank;
            for (int k = 0; k < A.n_rank; ++k) {
                const int item = items[k];
                if (item < 0) continue;
                const double *Tk = A.tables + (int64_t)item * A.item_stride;
                const double td = Tk[(int64_t)r * A.row_stride + s];
                const double te = Tk[(int64_t)s * A.row_stride + r];
#pragma unroll
                for (int j = 0; j < kMaxJ; ++j)
                    if (j < nj)
                        amp[j] += cdc[j * A.n_rank + k] * td - cec[j * A.n_rank + k] * te;
            }
            double js = 0.0;
            const int tj0 = A.chan_tjlo[chan];
#pragma unroll
            for (int j = 0; j < kMaxJ; ++j)
                if (j < nj) js += static_cast<double>(tj0 + 2 * j + 1) * amp[j] * amp[j];
            const double nrs = r == s ? 2.0 : 1.0;
            contrib = -js / (A.pair_norm[pair] * nrs * de);
            if (A.jsum) A.jsum[((int64_t)pair * N + r) * N + s] = js;
        }
    }
    // fixed-order work-group sum: wave shuffles, then the four wave sums
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) contrib += __shfl_down(contrib, off, 64);
    __shared__ double wave_sum[4];
    if ((threadIdx.x & 63) == 0) wave_sum[threadIdx.x >> 6] = contrib;
    __syncthreads();
    if (threadIdx.x == 0)
        A.partial[(int64_t)pair * A.n_tiles + blockIdx.x] =
            ((wave_sum[0] + wave_sum[1]) + wave_sum[2]) + wave_sum[3];
}

// One wave per pair: lane l sums tiles l, l + 64, ... in order, then a fixed
// butterfly combines the lanes (the same order on every run).
__global__ __launch_bounds__(64) void pair_reduce_kernel(const double *__restrict__ partial,
                                                         int n_tiles, double *__restrict__ out) {
    const int p = blockIdx.x;
    const double *row = partial + (int64_t)p * n_tiles;
    double s = 0.0;
    for (int t = threadIdx.x; t < n_tiles; t += 64) s += row[t];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (threadIdx.x == 0) out[p] = s;
}

}  // namespace

extern "C" {

int64_t ax_mbpt2_workspace_doubles(int n_virt, int n_pair) {
    if (n_virt < 1 || n_pair < 1) return 0;
    const int64_t nt = (n_virt + kTile - 1) / kTile;
    return nt * (nt + 1) / 2 * n_pair;
}

ax_status ax_bilinear_stage2_batch(int ns, int nm, int ld, int n_jobs, const int64_t *jobs,
                                   void *stream) {
    AX_REQUIRE(ns > 0 && nm >= 0 && nm <= ns && ld >= ns && n_jobs >= 0,
               "ax_bilinear_stage2_batch: bad sizes");
    AX_REQUIRE(n_jobs <= 65535, "ax_bilinear_stage2_batch: %d jobs (at most 65535)", n_jobs);
    if (n_jobs == 0 || nm == 0) return AX_OK;
    AX_REQUIRE(jobs, "ax_bilinear_stage2_batch: null job table");
    const int gx = ax::grid_for((int64_t)nm * ns, 256, 64);
    hipLaunchKernelGGL(bilinear_batch_kernel, dim3(gx, n_jobs), dim3(256), 0,
                       ax::as_stream(stream), jobs, ns, nm, ld);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

ax_status ax_mbpt2_pair_sum(int n_virt, int n_kappa, int n_rank, int n_pair, int max_j,
                            const int32_t *vkappa, const double *veps, const double *tables,
                            int64_t item_stride, int64_t row_stride, const int32_t *pair_item,
                            const double *pair_eps,
                            const double *pair_norm, const int32_t *pair_type,
                            const int32_t *chan_nj, const int32_t *chan_tjlo, const double *cd,
                            const double *ce, double *pair_energy, double *jsum,
                            double *workspace, void *stream) {
    AX_REQUIRE(n_virt >= 1 && n_kappa >= 1 && n_rank >= 1 && n_pair >= 0,
               "ax_mbpt2_pair_sum: bad sizes");
    AX_REQUIRE(max_j >= 1 && max_j <= kMaxJ, "ax_mbpt2_pair_sum: %d J values (at most %d)", max_j,
               kMaxJ);
    if (n_pair == 0) return AX_OK;
    AX_REQUIRE(n_pair <= 65535, "ax_mbpt2_pair_sum: %d pairs in one call (at most 65535)", n_pair);
    AX_REQUIRE(vkappa && veps && tables && pair_item && pair_eps && pair_norm && pair_type &&
                   chan_nj && chan_tjlo && cd && ce && pair_energy && workspace,
               "ax_mbpt2_pair_sum: null argument");
    const int nt = (n_virt + kTile - 1) / kTile;
    const int64_t tiles = (int64_t)nt * (nt + 1) / 2;
    AX_REQUIRE(tiles <= 0x7fffffff, "ax_mbpt2_pair_sum: %d virtuals", n_virt);
    AX_REQUIRE(item_stride >= 0 && row_stride >= n_virt,
               "ax_mbpt2_pair_sum: table strides %lld, %lld", (long long)item_stride,
               (long long)row_stride);
    PairSum A{n_virt, n_kappa, n_rank, n_pair, max_j, static_cast<int>(tiles), nt,
              vkappa, veps, tables, item_stride, row_stride, pair_item, pair_eps, pair_norm, pair_type,
              chan_nj, chan_tjlo, cd, ce, workspace, jsum};
    hipStream_t st = ax::as_stream(stream);
    hipLaunchKernelGGL(pair_sum_kernel, dim3(static_cast<unsigned>(tiles), n_pair), dim3(256), 0,
                       st, A);
    AX_CHECK_LAUNCH();
    hipLaunchKernelGGL(pair_reduce_kernel, dim3(n_pair), dim3(64), 0, st, workspace,
                       static_cast<int>(tiles), pair_energy);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

}  // extern "C"
