/*
 * atomx_slater.h — C ABI of the MI355X Slater-integral engine.
 *
 * The reference (zezhong-zhang/atomx) has no FFI on this path: its boundary is
 * a set of Python functions (numerics/yk.py, numerics/bspline/slater.py,
 * slater_yk.py, convol.py, methods/dirac_fock.py).  Each entry point below
 * names the reference function whose arithmetic it replaces; the Python
 * package atomx_amd re-exports those names and signatures on top of this ABI.
 *
 * Conventions
 *   - Every array argument is a DEVICE pointer unless its name ends in _host.
 *   - Arrays are row-major FP64 (int32 for index lists); leading dimensions
 *     are given in elements.
 *   - `stream` is a hipStream_t (NULL = default stream).  Calls are
 *     asynchronous on that stream unless documented otherwise.
 *   - Every function returns an ax_status; on failure ax_last_error() returns
 *     a message describing the failure (thread-local).
 *   - Status codes map to the reference's exceptions: AX_ERR_VALUE ->
 *     ValueError, AX_ERR_TYPE -> TypeError, AX_ERR_MEMORY -> MemoryError,
 *     AX_ERR_NONFINITE -> FloatingPointError, AX_ERR_HIP -> RuntimeError.
 */
#ifndef ATOMX_SLATER_H
#define ATOMX_SLATER_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int ax_status;
enum {
    AX_OK = 0,
    AX_ERR_VALUE = 1,
    AX_ERR_TYPE = 2,
    AX_ERR_MEMORY = 3,
    AX_ERR_NONFINITE = 4,
    AX_ERR_HIP = 5
};

/* Contraction precision of the R^k GEMM. */
enum {
    AX_PREC_FP64 = 0,  /* FP64 MFMA (v_mfma_f64_16x16x4_f64), FP64 accumulate        */
    AX_PREC_SPLIT = 1  /* error-free int8 slices on MFMA, exact int32 partial sums,
                          FP64 reduction per K-tile (Ozaki-type split precision)     */
};

/* ---- library ----------------------------------------------------------- */
const char *ax_last_error(void);
int ax_version(void);
/* Split-precision configuration: int8 slices per operand, K-tile sharing one
 * exponent, and the number of slice products per K-tile (S(S+1)/2). */
int ax_split_slices(void);
int ax_split_ktile(void);
int ax_split_products(void);
/* Number of visible devices; fails with AX_ERR_HIP when none is present. */
ax_status ax_device_count(int *count_host);

/* ---- (Y) Poisson Y^k discretisation ----------------------------------- */

/* numerics/yk.py:65-84 yk_partial_matrices: the interpolatory partial
 * interval matrices (ng x ng) on the host; mi = 1 - mo. */
ax_status ax_yk_partial_matrices(int ng, double *mo_host, double *mi_host);

/* numerics/yk.py:113-117: r^k and r^-(k+1) tables for k = 0..kmax,
 * out (kmax+1, n) each. */
ax_status ax_grid_rpow(const double *r, int64_t n, int kmax, double *rk, double *rkm,
                       void *stream);

/* numerics/bspline/slater_yk.py:184-194 and numerics/yk.py:93-94: pair
 * densities rho_p = P_a P_b + Q_a Q_b (times w when w != NULL) for pairs
 * p = (pa[p], pb[p]); P, Q are (n_orb, ld_pq) orbital samples; Q may be NULL. */
ax_status ax_pair_density(const double *P, const double *Q, int64_t ld_pq, int64_t n,
                          int n_pairs, const int32_t *pa, const int32_t *pb, const double *w,
                          double *out, int64_t ld_out, void *stream);

/* numerics/yk.py:87-154 yk_flat, batched: for every item t computes
 *   Y^k[rho_{row[t]}] with k = kk[t] into y row out_row[t].
 * If w != NULL the input rows are rho (weighted inside); if w == NULL they
 * are already the weighted densities w*rho.  rk/rkm are the ax_grid_rpow
 * tables (kmax+1, n); mo/mi (ng, ng) device copies of the partial matrices. */
ax_status ax_yk_batch(const double *rho, int64_t ld_rho, const double *w, const double *rk,
                      const double *rkm, int64_t n, int nv, int ng, int kmax,
                      const double *mo, const double *mi, int n_items, const int32_t *row,
                      const int32_t *kk, const int32_t *out_row, double *y, int64_t ld_y,
                      void *stream);

/* numerics/bspline/slater.py:129-130: R[t] = sum_x D[p[t], x] * Y[q[t], x]
 * for quartet lists (D = w*rho13 rows, Y = Y^k rows). */
ax_status ax_rk_dot(const double *D, int64_t ld_d, const double *Y, int64_t ld_y, int64_t n,
                    int n_quart, const int32_t *p, const int32_t *q, double *out, void *stream);

/* numerics/bspline/evaluate.py:71-97 eval_orbital_on_grid for n_orb
 * coefficient vectors (n_orb, ld_c) of length ns: out (n_orb, nv*ng). */
ax_status ax_bspline_eval(const double *coeffs, int64_t ld_c, int n_orb, const double *bsp,
                          int nv, int ng, int ks, int ns, double *out, int64_t ld_out,
                          void *stream);

/* numerics/bspline/slater_yk.py:35-122 _exchange_matrix_kernel: K (ns, ns)
 * for orbital samples fb (nv*ng), node-wise cumulative Y^k scans. */
ax_status ax_exchange_matrix_yk(const double *fb, const double *bsp, const double *r,
                                const double *w, int nv, int ng, int ks, int ns, int k,
                                double *K, void *stream);

/* numerics/bspline/slater.py:51-115 slater_rk_direct: O(N^2) double
 * quadrature, disjoint cells by the product rule, coincident cells split
 * along r = s (triangle rule of order `order` with nodes xq/weights wq on
 * [0,1], interpolation tables outer (order, ng) and inner (order, order, ng),
 * node0 = first Gauss node on [-1,1]).  workspace >= 1024 + nv doubles;
 * out: one double. */
ax_status ax_slater_rk_direct(const double *rho13, const double *rho24, const double *gr,
                              const double *grw, int nv, int ng, int k, int order,
                              const double *xq, const double *wq, const double *outer,
                              const double *inner, double node0, double *workspace, double *out,
                              void *stream);

/* ---- R^k GEMM (all pair-pairs of one rank) ------------------------------ */

/* numerics/bspline/slater.py:130 for a whole pair block: the upper triangle
 * of R = D[rows] * Y^T written packed: out[i*m - i(i-1)/2 + (j-i)], i <= j,
 * where D row rows[i] pairs with Y row i.  precision = AX_PREC_*.
 * `workspace` must hold ax_rk_gemm_workspace_bytes(m, n, precision) bytes. */
int64_t ax_rk_gemm_workspace_bytes(int m, int64_t n, int precision);
ax_status ax_rk_gemm(const double *D, int64_t ld_d, const int32_t *rows, const double *Y,
                     int64_t ld_y, int m, int64_t n, int precision, void *workspace,
                     double *out, void *stream);

/* ---- batched R^k tensor over many orbitals (the synthetic workload) ------ */

/* All R^k(ab;cd), k = 0..kmax, over quartets whose pairs satisfy the
 * triangle and parity rules |l_a-l_b| <= k <= l_a+l_b, l_a+l_b+k even,
 * counted once per unordered pair-pair (numerics/bspline/slater.py:118-130
 * per integral).  The plan owns the device workspaces; run() takes orbital
 * samples P, Q (n_orb, n) and writes out (ax_rk_tensor_size values):
 * for each k the packed upper triangle over the valid pairs of that k, in
 * pair order (a<=b lexicographic).  All launches go to `stream`. */
typedef struct ax_rk_tensor ax_rk_tensor;
ax_status ax_rk_tensor_create(const double *r_host, const double *w_host, int nv, int ng,
                              int n_orb, const int32_t *l_host, int kmax, int precision,
                              ax_rk_tensor **plan);
ax_status ax_rk_tensor_destroy(ax_rk_tensor *plan);
int64_t ax_rk_tensor_size(const ax_rk_tensor *plan);
/* number of valid pairs of rank k (rows of that rank's packed triangle) */
int ax_rk_tensor_pairs(const ax_rk_tensor *plan, int k);
/* offset of rank k's packed triangle inside the output */
int64_t ax_rk_tensor_offset(const ax_rk_tensor *plan, int k);
/* device bytes held by the plan (admission accounting) */
int64_t ax_rk_tensor_footprint(const ax_rk_tensor *plan);
ax_status ax_rk_tensor_run(ax_rk_tensor *plan, const double *P, const double *Q, double *out,
                           void *stream);
/* One stage of run(), for per-kernel timing: the stages in order are
 * density (needs P, Q), Y^k scan, slicing (split precision only), the
 * contraction (split partials) and the fixed-order reduction (needs out).
 * Running all five in order equals run(). */
enum {
    AX_STAGE_DENSITY = 0,
    AX_STAGE_SCAN = 1,
    AX_STAGE_SLICE = 2,
    AX_STAGE_CONTRACT = 3,
    AX_STAGE_REDUCE = 4
};
ax_status ax_rk_tensor_run_stage(ax_rk_tensor *plan, int stage, const double *P,
                                 const double *Q, double *out, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* ATOMX_SLATER_H */
