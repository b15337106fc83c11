// Signed antisymmetrised two-electron integrals of atomic spinors |n kappa m>
// for tensor correlation engines (CC / FSCC export).
//
// Direct Coulomb integral of spinors a, b, c, d (M conserved):
//   <ab|cd> = (-1)^((m_a - m_c)) sum_k T_k[a][c] T_k[b][d] R^k(ac, bd),
//   T_k[a][c] = (-1)^(j_a - m_a) (j_a k j_c; -m_a, m_a - m_c, m_c) <kappa_a||C^k||kappa_c>
// (the angular factor of angular/csf_hamiltonian.py:874-920 split into one
// factor per electron: with M conserved the second 3j, of (b, d) with
// projection -(m_a - m_c) = m_b - m_d, has the same form as the first),
// and R^k(ac, bd) the Slater integral of the target pair (a, c) against the
// source pair (b, d) from a radial table indexed by unordered radial pairs.
// The exported value is g = (x + y) / 2 with x = <pq|rs> - <pq|sr> and
// y = <rs|pq> - <rs|qp> (physics/csf/spinor_integrals.py:389-399): for real
// radial functions x = y up to rounding, and max |x - y| is returned for
// the caller's Hermiticity gate.
//
// Tables are small (n^2 (kmax+1) angular factors, n_pair^2 (kmax+1) radial
// values) and stay in L2; the kernels are write-bound on the packed output.
#include "ax_common.h"

namespace {

struct EriTables {
    int n, kmax, n_rad, n_rpair;
    const int32_t *rad;      // radial orbital of spinor p
    const int32_t *m2;       // 2 m_p
    const int32_t *pair_id;  // n_rad x n_rad -> unordered radial pair
    const double *T;         // (kmax+1) x n x n
    const double *R;         // (kmax+1) x n_rpair x n_rpair: R^k(target pair, source pair)
};

__device__ inline double direct(const EriTables &E, int a, int b, int c, int d) {
    const int q = (E.m2[a] - E.m2[c]) / 2;  // 2m are odd: the difference is even
    const double sign = (q & 1) ? -1.0 : 1.0;
    const int64_t tp = E.pair_id[E.rad[a] * E.n_rad + E.rad[c]];
    const int64_t sp = E.pair_id[E.rad[b] * E.n_rad + E.rad[d]];
    const int64_t nn = (int64_t)E.n * E.n, np2 = (int64_t)E.n_rpair * E.n_rpair;
    const double *tac = E.T + (int64_t)a * E.n + c;
    const double *tbd = E.T + (int64_t)b * E.n + d;
    const double *rk = E.R + tp * E.n_rpair + sp;
    double v = 0.0;
    for (int k = 0; k <= E.kmax; ++k) v += tac[k * nn] * tbd[k * nn] * rk[k * np2];
    return sign * v;
}

__device__ inline double signed_eri(const EriTables &E, int p, int q, int r, int s, double &err) {
    if (E.m2[p] + E.m2[q] != E.m2[r] + E.m2[s]) return 0.0;
    const double x = direct(E, p, q, r, s) - direct(E, p, q, s, r);
    const double y = direct(E, r, s, p, q) - direct(E, r, s, q, p);
    err = fmax(err, fabs(x - y));
    return 0.5 * (x + y);
}

// Largest |x - y| of a wave into *err_out (non-negative doubles order as
// their bit patterns).
__device__ inline void wave_max_to(double err, unsigned long long *err_out) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) err = fmax(err, __shfl_xor(err, off, 64));
    if ((threadIdx.x & 63) == 0 && err > 0.0)
        atomicMax(err_out, static_cast<unsigned long long>(__double_as_longlong(err)));
}

// One wave per (p, q) block of the packed layout: entries (r, s), p <= r < s,
// row r after row, s across the lanes (consecutive outputs).
__global__ __launch_bounds__(64) void eri_packed_kernel(EriTables E, int n_packed,
                                                        const int32_t *__restrict__ bp,
                                                        const int32_t *__restrict__ bq,
                                                        const int64_t *__restrict__ boff,
                                                        int64_t base, int nblk,
                                                        double *__restrict__ out,
                                                        unsigned long long *err_out) {
    double err = 0.0;
    for (int b = blockIdx.x; b < nblk; b += gridDim.x) {
        const int p = bp[b], q = bq[b];
        const int m = n_packed - p;
        double *o = out + (boff[b] - base);
        int64_t start = 0;
        for (int rx = 0; rx < m - 1; ++rx) {
            for (int sx = rx + 1 + static_cast<int>(threadIdx.x); sx < m; sx += 64)
                o[start + (sx - rx - 1)] = signed_eri(E, p, q, p + rx, p + sx, err);
            start += m - 1 - rx;
        }
    }
    wave_max_to(err, err_out);
}

__global__ __launch_bounds__(256) void eri_list_kernel(EriTables E, const int32_t *__restrict__ quad,
                                                       int64_t count, double *__restrict__ out,
                                                       unsigned long long *err_out) {
    double err = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count;
         i += (int64_t)gridDim.x * blockDim.x) {
        const int32_t *t = quad + 4 * i;
        out[i] = signed_eri(E, t[0], t[1], t[2], t[3], err);
    }
    wave_max_to(err, err_out);
}

ax_status check_tables(const char *who, int n, int kmax, int n_rad, int n_rpair,
                       const int32_t *rad, const int32_t *m2, const int32_t *pair_id,
                       const double *T, const double *R, const double *max_err) {
    AX_REQUIRE(n >= 1 && kmax >= 0 && n_rad >= 1 && n_rpair >= 1, "%s: bad sizes", who);
    AX_REQUIRE(rad && m2 && pair_id && T && R && max_err, "%s: null table", who);
    return AX_OK;
}

}  // namespace

extern "C" {

ax_status ax_spinor_eri_packed(int n, int kmax, int n_rad, int n_rpair, const int32_t *rad,
                               const int32_t *m2, const int32_t *pair_id, const double *T,
                               const double *R, int n_packed, int nblk, const int32_t *blk_p,
                               const int32_t *blk_q, const int64_t *blk_off, int64_t base,
                               double *out, double *max_err, void *stream) {
    ax_status st = check_tables("ax_spinor_eri_packed", n, kmax, n_rad, n_rpair, rad, m2,
                                pair_id, T, R, max_err);
    if (st != AX_OK) return st;
    AX_REQUIRE(nblk >= 0, "ax_spinor_eri_packed: negative block count");
    AX_REQUIRE(n_packed >= 2 && n_packed <= n, "ax_spinor_eri_packed: %d packed spinors of %d",
               n_packed, n);
    if (nblk == 0) return AX_OK;
    AX_REQUIRE(blk_p && blk_q && blk_off && out, "ax_spinor_eri_packed: null block table");
    const EriTables E{n, kmax, n_rad, n_rpair, rad, m2, pair_id, T, R};
    const int grid = nblk < 65535 ? nblk : 65535;
    hipLaunchKernelGGL(eri_packed_kernel, dim3(grid), dim3(64), 0, ax::as_stream(stream), E,
                       n_packed, blk_p, blk_q, blk_off, base, nblk, out,
                       reinterpret_cast<unsigned long long *>(max_err));
    AX_CHECK_LAUNCH();
    return AX_OK;
}

ax_status ax_spinor_eri_list(int n, int kmax, int n_rad, int n_rpair, const int32_t *rad,
                             const int32_t *m2, const int32_t *pair_id, const double *T,
                             const double *R, int64_t count, const int32_t *quartets, double *out,
                             double *max_err, void *stream) {
    ax_status st = check_tables("ax_spinor_eri_list", n, kmax, n_rad, n_rpair, rad, m2, pair_id,
                                T, R, max_err);
    if (st != AX_OK) return st;
    AX_REQUIRE(count >= 0, "ax_spinor_eri_list: negative count");
    if (count == 0) return AX_OK;
    AX_REQUIRE(quartets && out, "ax_spinor_eri_list: null quartets or output");
    const EriTables E{n, kmax, n_rad, n_rpair, rad, m2, pair_id, T, R};
    hipLaunchKernelGGL(eri_list_kernel, dim3(ax::grid_for(count, 256, 16384)), dim3(256), 0,
                       ax::as_stream(stream), E, quartets, count, out,
                       reinterpret_cast<unsigned long long *>(max_err));
    AX_CHECK_LAUNCH();
    return AX_OK;
}

}  // extern "C"
