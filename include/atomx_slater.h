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

/* numerics/bspline/integrals.py:18-56 _assemble_integral, for n_w weight
 * rows at once:
 * mat[w][gi][gj] = sum_iv sum_m weights[w][iv][m] bsp1[iv][m][gi-iv] bsp2[iv][m][gj-iv]
 * over the cells iv that hold both B-splines (iv ascending, as the
 * reference accumulates), for gi, gj < mm; mat (n_w, mm, mm) is overwritten
 * and holds zeros outside the band.  bsp1 (nv, ng, ks1), bsp2 (nv, ng, ks2),
 * weights (n_w, nv, ng).  With weights = grw * Y^0[rho] this is the direct
 * potential of methods/dirac_fock.py:954-965. */
ax_status ax_assemble_integral(int nv, int ng, int ks1, int ks2, int mm, const double *bsp1,
                               const double *bsp2, const double *weights, int n_w, double *mat,
                               void *stream);
/* methods/dirac_fock.py:957,962: M = (M + M^T)/2 in place, n_mat (mm, mm). */
ax_status ax_symmetrize(double *mat, int mm, int n_mat, void *stream);

/* numerics/bspline/slater_yk.py:35-122 _exchange_matrix_kernel: K (ns, ns)
 * for orbital samples fb (nv*ng), node-wise cumulative Y^k scans. */
ax_status ax_exchange_matrix_yk(const double *fb, const double *bsp, const double *r,
                                const double *w, int nv, int ng, int ks, int ns, int k,
                                double *K, void *stream);

/* ax_exchange_matrix_yk for n_items (orbital, rank) pairs at once:
 * K[t] (ns, ns) from row item_b[t] of fb (rows of ld_fb samples) and rank
 * item_k[t] <= kmax_tab of the ax_grid_rpow tables rk_tab / rkm_tab
 * (kmax_tab+1, nv*ng).  item_b, item_k are device arrays (NULL item_b: row t;
 * NULL item_k: rank kmax_tab).  The caller checks item_k[t] <= kmax_tab.
 * Entries whose B-spline supports are disjoint use the factorised form
 * K[m][n] = A^-(m) T_out(n) or A^+(m) T_in(n) of the same integral; the band
 * |m - n| < ks follows the reference's per-node sum. */
ax_status ax_exchange_matrix_yk_batch(const double *fb, int64_t ld_fb, const double *bsp,
                                      const double *w, const double *rk_tab,
                                      const double *rkm_tab, int kmax_tab, int nv, int ng, int ks,
                                      int ns, int n_items, const int32_t *item_b,
                                      const int32_t *item_k, double *K, void *stream);

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

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): ax_rk_gemm with its workspace taken from a static
 * device array of AX_NOALLOC_SCRATCH_BYTES that the library holds on every
 * device it runs on until the process ends.  The call allocates nothing;
 * it fails with AX_ERR_MEMORY when ax_rk_gemm_workspace_bytes(m, n,
 * precision) exceeds the array.  All calls on one device share that
 * device's array: calls that overlap on one device (other threads or
 * streams) race and must be serialised by the caller.  Like ax_rk_gemm it
 * synchronises `stream` before returning. */
#define AX_NOALLOC_SCRATCH_BYTES (32ll << 20)
ax_status ax_rk_gemm_noalloc(const double *D, int64_t ld_d, const int32_t *rows,
                             const double *Y, int64_t ld_y, int m, int64_t n, int precision,
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
/* n_plans plans in one call, plan i as ax_rk_tensor_create(r_host[i],
 * w_host[i], nv[i], ng[i], n_orb[i], l_host[i], kmax[i], precision[i]): the
 * set-up images go up with one copy through one pinned staging buffer, one
 * launch per 16 plans scatters them into the plans' blocks and forms their
 * r-power tables, and the plans share one set-up completion event (serving:
 * every one-shot plan of a launch set).  That set-up is queued on `stream`
 * (NULL here means the calling thread's own set-up stream, not the default
 * stream); launches of the plans on any stream wait for it on the device.
 * Passing the stream the plans will run on keeps the set-up off other
 * streams' hardware queues; that stream must stay alive until each plan has
 * run once or been destroyed.  Each plan is destroyed on its own.  On failure
 * no plan is left and every out_plans[i] is NULL.  No reference
 * counterpart: it batches the set-up of plans that
 * numerics/bspline/slater.py:118-130 would evaluate one by one. */
ax_status ax_rk_tensor_create_many(int n_plans, const double *const *r_host,
                                   const double *const *w_host, const int32_t *nv,
                                   const int32_t *ng, const int32_t *n_orb,
                                   const int32_t *const *l_host, const int32_t *kmax,
                                   const int32_t *precision, void *stream,
                                   ax_rk_tensor **out_plans);
/* Output size, slice count (0: FP64 contraction) and, for k = 0..kmax, the
 * valid pairs and the packed offset of rank k, in one call (the values of
 * ax_rk_tensor_size / _slices / _pairs / _offset). */
ax_status ax_rk_tensor_describe(const ax_rk_tensor *plan, int64_t *size, int32_t *slices,
                                int32_t *pairs, int64_t *offsets);
/* ---- native serving loop (continuous batching of R^k requests) -------- */

/* A server owns one scheduler thread on the calling thread's device.  It
 * admits submitted all-pairs R^k requests in arrival order into launch sets
 * of at most max_batch plans while one of max_inflight slots is free and the
 * in-flight device bytes (plans, staged inputs and outputs) stay within
 * budget_bytes; each set's plans come from one ax_rk_tensor_create_many
 * call, its orbitals are packed into pinned staging by worker threads and
 * moved with one copy, its plans run as one ax_rk_tensor_run_batch and its
 * results come back with one copy into a pinned buffer of the set.  No
 * reference counterpart: the reference evaluates integrals one call at a
 * time (numerics/bspline/slater.py:118-130); this is the serving loop of the
 * same values.  destroy() finishes every queued and in-flight request; the
 * values of tickets not released by then stay valid until the process ends. */
typedef struct ax_server ax_server;
ax_status ax_server_create(int64_t budget_bytes, int max_inflight, int max_batch,
                           ax_server **server);
ax_status ax_server_destroy(ax_server *server);
/* Queues the request of ax_rk_tensor_create(r_host, w_host, nv, ng, n_orb,
 * l_host, kmax, precision) run on P_host, Q_host (n_orb x nv*ng FP64 host
 * arrays, Q_host may be NULL).  r_host, w_host, P_host and Q_host are read
 * when the request is launched and must stay valid until its ticket is
 * complete; l_host is copied.  AX_ERR_MEMORY when the request alone needs
 * more than the budget.  state != 0 names a resident state (one atom's
 * requests, SCF iteration after SCF iteration): its first request fixes the
 * shape (grid values, l, kmax, precision, Q given or not) and its plan is
 * kept for the later ones, which must have that shape (AX_ERR_VALUE
 * otherwise); one request of a state runs at a time, requests of other
 * states may go ahead of a waiting one, and idle states give up their plans
 * (least recently used first) when a request does not fit the budget. */
ax_status ax_server_submit(ax_server *server, const double *r_host, const double *w_host, int nv,
                           int ng, int n_orb, const int32_t *l_host, int kmax, int precision,
                           const double *P_host, const double *Q_host, int64_t state,
                           int64_t *ticket);
ax_status ax_server_poll(ax_server *server, int64_t ticket, int32_t *done);
ax_status ax_server_wait(ax_server *server, int64_t ticket);
/* A complete ticket's packed R^k values (the layout of ax_rk_tensor_run:
 * per k the packed upper triangle over the valid pairs) in server-owned
 * pinned memory, valid until ax_server_release; its status is the request's
 * (a failed request returns its error here).  slices: int8 slices the
 * contraction used (0: FP64). */
ax_status ax_server_result(ax_server *server, int64_t ticket, const double **values,
                           int64_t *size, int32_t *slices);
ax_status ax_server_release(ax_server *server, int64_t ticket);
/* Counters since creation, the first n of: launch sets, plans run, peak
 * device bytes (in flight plus resident), host nanoseconds of the scheduler
 * thread in plan set-up, waiting for the packing, allocations, and queueing
 * copies and launches; plans made; and now: resident states holding a plan,
 * their bytes. */
ax_status ax_server_stats(ax_server *server, int64_t *out, int n);
/* Frees a resident state once no request of it is queued or in flight (at
 * once if none is); a request naming the key before that keeps the state
 * (the drop is cancelled), one after it starts a new state.  Unknown keys
 * are ignored. */
ax_status ax_server_drop_state(ax_server *server, int64_t state);
/* Pair-product tensor (the CI / CCSD B-spline product workload): n_fn
 * functions sampled on the grid (P rows; Q may be NULL), an explicit list of
 * n_pairs pairs (pa_host[p], pb_host[p]), every pair at every rank
 * k = 0..kmax (no angular rule).  Output as ax_rk_tensor_create: per k the
 * packed upper triangle over all pairs in list order.  R^k(B_i B_j; B_k B_l)
 * of numerics/bspline/convol.py:323-380 (_assemble_rkb_sym) in the (Y)
 * discretisation. */
ax_status ax_rk_tensor_create_pairs(const double *r_host, const double *w_host, int nv, int ng,
                                    int n_fn, int n_pairs, const int32_t *pa_host,
                                    const int32_t *pb_host, int kmax, int precision,
                                    ax_rk_tensor **plan);
/* Device bytes ax_rk_tensor_create would hold for this shape (admission by
 * HBM footprint before anything is allocated); -1 on invalid arguments. */
int64_t ax_rk_tensor_estimate(int nv, int ng, int n_orb, const int32_t *l_host, int kmax,
                              int precision);
int64_t ax_rk_tensor_size(const ax_rk_tensor *plan);
/* number of valid pairs of rank k (rows of that rank's packed triangle) */
int ax_rk_tensor_pairs(const ax_rk_tensor *plan, int k);
/* offset of rank k's packed triangle inside the output */
int64_t ax_rk_tensor_offset(const ax_rk_tensor *plan, int k);
/* device bytes held by the plan now (admission accounting): creation
 * holds the default K split's job tables and partial workspace; a schedule
 * of another split (set_schedule, tune) adds that split's tables, and a
 * larger workspace when it needs one */
int64_t ax_rk_tensor_footprint(const ax_rk_tensor *plan);
/* int8 slices per operand of a split-precision plan, from the radius ratio
 * rho one 64-node K-tile spans (cells after the first) and
 * bits = (kmax+1) log2(rho): 7 up to 5 bits, 8 up to 12 bits; beyond that
 * (and for FP64 plans) 0: the plan contracts in FP64. */
int ax_rk_tensor_slices(const ax_rk_tensor *plan);
ax_status ax_rk_tensor_run(ax_rk_tensor *plan, const double *P, const double *Q, double *out,
                           void *stream);
/* Several plans (any grids, orbital counts, k ranges and precision modes)
 * as one set of launches: density, Y^k scan (one per n_gauss value),
 * slicing, contraction (one per contraction kind) and reduction each run
 * once for the whole batch, in chunks of 16 plans; a plan listed more than
 * once starts a new chunk at each repeat, so its listings run one after
 * the other.  P, Q and out are host arrays of n_plans device
 * pointers (Q may be NULL, or hold NULL entries, for no small component).
 * Each plan's results are bitwise those of its own ax_rk_tensor_run: tiles,
 * K splits and summation order are fixed per plan. */
ax_status ax_rk_tensor_run_batch(ax_rk_tensor *const *plans, int n_plans,
                                 const double *const *P, const double *const *Q,
                                 double *const *out, void *stream);
/* One stage of run(), for per-kernel timing: the stages in order are
 * density (needs P, Q), Y^k scan, slicing (split precision only), the
 * contraction (needs out: K-chunk partials, and the packed values of tiles
 * contracted in one chunk) and the fixed-order reduction of the partials
 * (needs out; nothing to launch at K split 1).
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

/* Contraction schedule of a plan: block order and level S (1 .. min(16,
 * 4x the default K split)).  Orders 0-3 cut every tile's K range into S
 * equal chunks, one work-group each: 0 = tile-major, 1 = work-groups dealt
 * to the 8 XCDs in contiguous ranges, 2 = as 1 with the K split fastest,
 * 3 = (FP64) resident work-groups take the chunks heaviest tile first from
 * a ticket.  Order 4 runs 64 x S work-groups: the K range is cut into one
 * slice per XCD (at most 8), and each slice's tiles, heaviest first (16x16
 * blocks inside the triangle), go whole to the least-loaded of that XCD's
 * 8 x S work-groups.  The schedule fixes how each
 * tile's K range is cut and the order its pieces are summed, so results
 * are deterministic per schedule.  A split's job tables are built and
 * uploaded the first time a schedule of that split is set (synchronous).
 * tune() runs the plan once on P, Q, then times every (order, split) with
 * `reps` timed repetitions of the contraction and reduction, keeps the
 * fastest, and leaves its result in `out`; best_ms_host (may be NULL)
 * receives the chosen schedule's time per contraction + reduction.
 * No counterpart in the reference (its contraction is one numpy dot per
 * integral).  tune() synchronises `stream` and must not be captured. */
ax_status ax_rk_tensor_set_schedule(ax_rk_tensor *plan, int order, int n_split);
ax_status ax_rk_tensor_get_schedule(const ax_rk_tensor *plan, int *order_host,
                                    int *n_split_host);
ax_status ax_rk_tensor_tune(ax_rk_tensor *plan, const double *P, const double *Q, double *out,
                            int reps, void *stream, float *best_ms_host);

/* ---- device workspace cache -------------------------------------------------
 * Plan workspaces come from a caching allocator (size-classed free lists per
 * device), so creating and destroying plans per request costs no hipMalloc
 * once the cache is warm.  No counterpart in the reference (its workspaces
 * are numpy arrays).  cached_bytes: device bytes held in free lists;
 * trim: waits for the device and returns every cached block to HIP. */
int64_t ax_pool_cached_bytes(void);
ax_status ax_pool_trim(void);

/* ---- (C) DBSR cell algorithm ------------------------------------------------
 * Bank layout: bankT[ip][jp][i][j] = R^k(B_i B_{i+ip}; B_j B_{j+jp}), i.e. the
 * reference's transposed bank rkbT (convol.py:1572-1591), (ksl, ksr, ns, ns).
 * Contractions take nb densities that share one bank (leading dims ld_*). */

/* numerics/bspline/convol.py:252-291 _gauss_legendre_01 (host, n nodes on [0,1]) */
ax_status ax_gauss_legendre_01(int n, double *x_host, double *w_host);

/* convol.py:57-90 _compute_moments_sym (sym=1, packed upper triangle, ks1 == ks2)
 * and :93-128 _compute_moments_nonsym (sym=0): out (kk, nv). */
ax_status ax_cell_moments(const double *bsp1, const double *bsp2, const double *gr,
                          const double *grw, int nv, int ng, int ks1, int ks2, int k, int sym,
                          double *out, void *stream);

/* convol.py:135-249 _triang_sym (left == right basis) and :387-529
 * _triang_pqpq: diagonal-cell integrals out (kkl, kkr, nv); n_sub sub-nodes
 * x01/w01 from ax_gauss_legendre_01 (n_sub = max(ksl,ksr) + k/2 for k >= 0). */
ax_status ax_cell_triang(const double *bsp_l, const double *bsp_r, const double *gr,
                         const double *grw, const double *t_l, const double *t_r, int nv, int ng,
                         int ksl, int ksr, int k, int n_sub, const double *x01, const double *w01,
                         double *out, void *stream);

/* convol.py:323-380 _assemble_rkb_sym / :532-572 _assemble_rkb_pqpq (mode 0:
 * the bank bankT) and :1076-1122 _assemble_near_rkb (mode 1: the overlap
 * strip near (ksl, ksr, ns, ksl+ksr-1)), from r^k moments lo_*, r^-(k+1)
 * moments hi_* and diagonal cells. */
ax_status ax_cell_pair_integrals(const double *lo_l, const double *hi_r, const double *lo_r,
                                 const double *hi_l, const double *diag, int nv, int ksl, int ksr,
                                 int ns, int mode, double *out, void *stream);

/* convol.py:1061-1073 _global_pair_moments: out (ks, ns). */
ax_status ax_cell_global_moments(const double *local, int nv, int ks, int ns, double *out,
                                 void *stream);

/* convol.py:688-753 _density_sym for np coefficient-row pairs (A[pa], A[pb]),
 * zero-padded beyond n1 / n2: out (np, ns, ks). */
ax_status ax_density_sym(const double *A, int64_t lda, const int32_t *pa, const int32_t *pb,
                         int np, int n1, int n2, int ns, int ks, double *out, void *stream);

/* convol.py:646-685 _accumulate_direct_density_batch as a row combination:
 * out[t] = sum_j coef[t, j] dens[j] (rows of L values). */
ax_status ax_combine_rows(const double *dens, int64_t L, int n_src, const double *coef, int nt,
                          double *out, void *stream);

/* convol.py:688-709 _accumulate_exchange_density_batch for nt targets:
 * out[t][i][k] = sum_j (coef[t,j] A[j][i]) B[j][k]. */
ax_status ax_exchange_density(const double *A, const double *B, int64_t ld, int n_src, int ns,
                              const double *coef, int nt, double *out, void *stream);

/* convol.py:788-849 / :971-997 _convol_case1[_mixed|_t]: out (nb, ns, ksl). */
ax_status ax_convol_case1(const double *bankT, int ns, int ksl, int ksr, const double *d,
                          int64_t ld_d, int nb, double *out, int64_t ld_o, void *stream);

/* convol.py:852-874 / :1310-1345 _convol_case2[_mixed]: left densities
 * (nb, n_left, ksl) -> out (nb, ns, ksr). */
ax_status ax_convol_case2(const double *bankT, int ns, int ksl, int ksr, int n_left,
                          const double *d, int64_t ld_d, int nb, double *out, int64_t ld_o,
                          void *stream);

/* convol.py:877-948 / :1000-1058 _convol_case4[_t]: full densities
 * (nb, ns, ns) -> out (nb, ns, ns); symmetric = symmetric_output. */
ax_status ax_convol_case4(const double *bankT, int ns, int ksl, int ksr, const double *d,
                          int64_t ld_d, int nb, int symmetric, double *out, int64_t ld_o,
                          void *stream);
/* case 4 accumulated into nb matrices at out_base + dst_off[b] (device int64
 * element offsets): out += the contraction (the "hfm_XX += Kk" step of
 * methods/dirac_fock.py:1293-1301 fused into the contraction). */
ax_status ax_convol_case4_acc(const double *bankT, int ns, int ksl, int ksr, const double *d,
                              int64_t ld_d, int nb, int symmetric, double *out_base,
                              const int64_t *dst_off, void *stream);

/* convol.py:1125-1162 _convol_case1_semiseparable_t; workspace holds
 * ax_convol_semi_workspace(nb, ns, ksr) doubles. */
int64_t ax_convol_semi_workspace(int nb, int ns, int ksr);
ax_status ax_convol_case1_semi(const double *near, const double *rl, const double *fl,
                               const double *rr, const double *fr, int ns, int ksl, int ksr,
                               const double *d, int64_t ld_d, int nb, double *workspace,
                               double *out, int64_t ld_o, void *stream);

/* convol.py:1187-1222 _convol_case4_semiseparable_t (symmetric = 1) and
 * :1250-1303 _convol_case4_semiseparable_mixed_t (symmetric = 0). */
ax_status ax_convol_case4_semi(const double *near, const double *rl, const double *fl,
                               const double *rr, const double *fr, int ns, int ksl, int ksr,
                               const double *d, int64_t ld_d, int nb, int symmetric, double *out,
                               int64_t ld_o, void *stream);

/* convol.py:1348-1356 _sum_amb_banded for nb pairs (row strides sa, sb). */
ax_status ax_sum_amb_banded(const double *a, int64_t lda, int sa, const double *b, int64_t ldb,
                            int sb, int nb, int ns, int ks, double *out, void *stream);

/* convol.py:1348-1356 _sum_amb_banded for nq operand pairs given by device
 * pointer tables a[q], b[q] (each (ns, ks) row-major); out[q] is
 * overwritten, or accumulated into when accumulate != 0 (the component sum
 * of physics/csf/matrix.py:793-797). */
ax_status ax_sum_amb_ptr(const double *const *a, const double *const *b, int nq, int ns, int ks,
                         int accumulate, double *out, void *stream);

/* convol.py:814-825 _banded_to_full_add for nb matrices. */
ax_status ax_banded_to_full_add(double *V, int64_t ldv, const double *vb, int64_t ldb, int nb,
                                int ns, int ks, double coeff, void *stream);
/* The same into nb matrices at V_base + dst_off[b] (dst_off: device int64
 * element offsets), touching only the band. */
ax_status ax_banded_to_full_add_at(double *V_base, const int64_t *dst_off, const double *vb,
                                   int64_t ldb, int nb, int ns, int ks, double coeff,
                                   void *stream);

/* y += alpha x (the "hfm_XX += Kk" accumulations of methods/dirac_fock.py:1293-1301) */
ax_status ax_axpy(double *y, const double *x, int64_t n, double alpha, void *stream);

/* y[r][x] = a[r][x] w[x]: weighted densities (numerics/bspline/slater.py:130) */
ax_status ax_mul_rows(double *y, const double *a, const double *w, int nrows, int64_t n,
                      void *stream);

/* convol.py:1933-1942 (_rkb_bilinear stage 2): M (nm, ns) from the case-2
 * contractions T (nm, ns, ksr) of the symmetrised unit densities and wb. */
ax_status ax_bilinear_stage2(const double *T, int ns, int ksr, int n_right, const double *wb,
                             int nm, double *M, void *stream);

/* methods/dirac_fock.py:1303-1317: symmetrise the four Fock blocks and divide
 * by the target's occupation divisor (skip_pq: PQ = QP^T). */
ax_status ax_fock_finalize(double *pp, double *qq, double *pq, double *qp, int ns, double norm,
                           int skip_pq, void *stream);

/* ---- batched Fock build of many atoms or ions ----------------------------
 * methods/dirac_fock.py:1128-1319 (_build_hf_matrix_dbsr) for every target
 * of every ion in one plan: five launches per run (direct densities,
 * semiseparable chains, direct contractions, exchange + assembly per output
 * tile, symmetrisation and occupation divisor), capturable in a HIP graph.
 * Per ion, in order, ion_meta_host holds 7 ints
 *   ns, ksp, ksq, n_orb, n_targets, kmax, skip_pq_exchange
 * and the other host arrays hold, ion after ion:
 *   targets_host       (n_targets) orbital index of each target
 *   direct_host        (n_targets, n_orb, kmax+1) angular direct weights
 *   exchange_host      (n_targets, n_orb, kmax+1) angular exchange weights
 *                      (both zero at ranks a target does not take)
 *   divisor_host       (n_targets) occupation divisors (no division <= 1e-15)
 *   P[i], Q[i]         device coefficient stacks (n_orb, ns), zero past nsp / nsq
 *   H[i]               device output (n_targets, 4, ns, ns): PP, QQ, PQ, QP
 *   banks_host         (4, kmax+1) device banks bankT of blocks pppp, pqpq,
 *                      qpqp, qqqq (NULL where unused)
 *   factors_host       (4, kmax+1, 5) device semiseparable factors (near,
 *                      rising_l, falling_l, rising_r, falling_r) used for the
 *                      direct term when ns >= semi_direct_ns (NULL otherwise).
 * max_group (1..8, 0 = 8) bounds the targets one assembly work-group holds
 * (balanced groups): fewer raise occupancy, more read each bank less often.
 * The plan keeps the device pointers: a run reads P, Q as they are at that
 * moment and overwrites H.  ns >= 320 (the semiseparable exchange) and
 * orders above 12 are refused with AX_ERR_VALUE. */
typedef struct ax_fock_plan ax_fock_plan;
ax_status ax_fock_plan_create(int n_ions, const int32_t *ion_meta_host,
                              const int32_t *targets_host, const double *direct_host,
                              const double *exchange_host, const double *divisor_host,
                              const double *const *P, const double *const *Q, double *const *H,
                              const double *const *banks_host,
                              const double *const *factors_host, int semi_direct_ns,
                              int max_group, ax_fock_plan **plan);
ax_status ax_fock_plan_run(ax_fock_plan *plan, void *stream);
/* device bytes the plan holds; (target, source, k, block) contractions with
 * a nonzero weight, direct and exchange counted separately */
int64_t ax_fock_plan_bytes(const ax_fock_plan *plan);
int64_t ax_fock_plan_rk_equivalents(const ax_fock_plan *plan);
/* waits for the device, then frees the plan */
ax_status ax_fock_plan_destroy(ax_fock_plan *plan);

/* ---- auxiliary-node Coulomb action and pair contraction ------------------
 * Sampled columns are node-major (n_nodes, ld) FP64, or complex FP64 as
 * interleaved (re, im) when is_complex (c-products, no conjugation). */

/* numerics/bspline/coulomb_features.py:281-294 (sample) for n_col
 * coefficient vectors (n_col, ld_coef) of one component: rows node_base +
 * cell*width + j of out get sum_s coef[cell+s] tab[cell][j][s] (s < order,
 * cell+s < n_coef). */
ax_status ax_feat_sample(const double *tab, int nv, int width, int order, int n_coef,
                         const void *coef, int64_t ld_coef, int n_col, int is_complex, void *out,
                         int64_t ld_out, int64_t node_base, void *stream);

/* coulomb_features.py:30-61 (_apply) for n_col real columns: outer
 * (2, nv, ngo) node indices, slot[node] = cell*ngo + g for outer nodes and
 * -1 otherwise, rising/falling (nv, ngo) weights, same-cell edges by
 * destination (row_ptr (n_nodes+1) int64, src, w).  work holds
 * ax_feat_apply_workspace(nv, n_col) bytes. */
ax_status ax_feat_apply(const double *rho, int64_t ld, int n_col, int64_t n_nodes,
                        const int *outer, const int *slot, const double *rising,
                        const double *falling, int nv, int ngo, const int64_t *row_ptr,
                        const int *src, const double *w, double *work, double *out,
                        int64_t ld_out, void *stream);
int64_t ax_feat_apply_workspace(int nv, int n_col);

/* coulomb_features.py:332-337 (integral) pieces: column products
 * out[i][p] = S[i][ia[p]] S[i][ib[p]] and column dots
 * out[p] = sum_i X[i][ix[p]] Y[i][iy[p]]. */
ax_status ax_feat_products(const void *S, int64_t lds, int64_t n_nodes, int n_pairs,
                           const int *ia, const int *ib, int is_complex, void *out,
                           int64_t ld_out, void *stream);
ax_status ax_feat_dots(const void *X, int64_t ldx, const void *Y, int64_t ldy, int64_t n_nodes,
                       int n_pairs, const int *ix, const int *iy, int is_complex, void *out,
                       void *stream);

/* numerics/bspline/pair_contraction.py:84-99 (moment): per cell c,
 * rise/fall[c][i][j] = sum over both components' outer nodes o of c of
 * w[c][g] Y[o][i] X[o][j] with w = wr / wf. */
ax_status ax_cpc_moments(const void *X, int64_t ldx, int d2, const void *Y, int64_t ldy, int d1,
                         const int *outer, const double *wr, const double *wf, int nv, int ngo,
                         int is_complex, void *rise, void *fall, void *stream);
/* pair_contraction.py:100-105: prefix[c] = sum_{c'<c} rise[c'],
 * suffix[c] = sum_{c'>c} fall[c'] over (nv, n_el) stacks. */
ax_status ax_cpc_scan(const void *rise, const void *fall, int nv, int64_t n_el, int is_complex,
                      void *prefix, void *suffix, void *stream);
/* pair_contraction.py:121-148 (__call__): out (m, n) = sum_c L_rise[c] rho
 * sfx[c]^T + L_fall[c] rho pfx[c]^T (left_stack = [L_rise; L_fall] as
 * (2, nv, nr, m), sfx/pfx (nv, ns, n)) + the same-cell directed edges
 * (src, dst, w): w (B[src] rho . C[dst]) A[src]^T D[dst].  work holds
 * ax_cpc_workspace_bytes(...) bytes. */
int64_t ax_cpc_workspace_bytes(int nv, int nr, int ns, int m, int n, int64_t n_nodes,
                               int64_t n_edges, int n_split, int is_complex);
ax_status ax_cpc_apply(const void *rho, int nr, int ns, const void *A, int64_t lda, int m,
                       const void *B, int64_t ldb, const void *C, int64_t ldc, const void *D,
                       int64_t ldd, int n, int64_t n_nodes, int nv, const void *left_stack,
                       const void *sfx, const void *pfx, const int *src, const int *dst,
                       const double *w, int64_t n_edges, int n_split, int is_complex, void *work,
                       void *out, void *stream);

/* ---- spinor integrals for correlation engines ---------------------------
 * physics/csf/spinor_integrals.py:258-466 (build_spinor_integrals): the
 * signed antisymmetrised integrals g(p,q,r,s) of spinors |n kappa m>, with
 * the angular factor of angular/csf_hamiltonian.py:874-920 split per
 * electron: <ab|cd> = (-1)^((m_a - m_c)) sum_k T[k][a][c] T[k][b][d]
 * R[k][pair(rad a, rad c)][pair(rad b, rad d)] when m_a + m_b = m_c + m_d
 * (else 0), T[k][a][c] = (-1)^(j_a - m_a) (j_a k j_c; -m_a, m_a - m_c, m_c)
 * <kappa_a||C^k||kappa_c>; g = (x + y) / 2, x = <pq|rs> - <pq|sr>,
 * y = <rs|pq> - <rs|qp>, and max_err (one double, zeroed by the caller)
 * receives the largest |x - y| (the Hermiticity gate of :389-399).
 * Device pointers: rad, m2 (n, int32: radial orbital, 2m), pair_id (n_rad x
 * n_rad, int32, symmetric), T ((kmax+1) n n), R ((kmax+1) n_rpair n_rpair,
 * rows the target pair). */
/* The packed layout of spinor_integrals.py:20-50 (packed_size) over the
 * first n_packed spinors (2 <= n_packed <= n; the tables may hold further,
 * frozen, spinors): for the (p, q) blocks b = 0..nblk-1 (p < q, block
 * offsets blk_off), every (r, s) with p <= r < s < n_packed at
 * out[blk_off[b] - base + m r' + s' - (r'+1)(r'+2)/2], m = n_packed - p,
 * r' = r - p, s' = s - p. */
ax_status ax_spinor_eri_packed(int n, int kmax, int n_rad, int n_rpair, const int32_t *rad,
                               const int32_t *m2, const int32_t *pair_id, const double *T,
                               const double *R, int n_packed, int nblk, const int32_t *blk_p,
                               const int32_t *blk_q, const int64_t *blk_off, int64_t base,
                               double *out, double *max_err, void *stream);
/* g of an explicit list of quartets (4 int32 spinor indices each): the
 * Fock, frozen-core and reference-energy contractions of :401-432. */
ax_status ax_spinor_eri_list(int n, int kmax, int n_rad, int n_rpair, const int32_t *rad,
                             const int32_t *m2, const int32_t *pair_id, const double *T,
                             const double *R, int64_t count, const int32_t *quartets, double *out,
                             double *max_err, void *stream);

/* ---- second-order MBPT on a closed-shell reference (methods/mbpt.py) ---- */
/* Batched stage 2 of the bilinear R^k forms (numerics/bspline/convol.py:
 * 1900-1978, build_slater_rk_bilinear's gather over the right band pair,
 * as ax_bilinear_stage2): job i (5 int64 each: stage-1 contraction T
 * (nm x ns x ksr) address, right coefficient row wb address, output
 * address, ksr, n_right) writes an nm x ns block with row stride ld, so the
 * four component blocks land in one 2ns x 2ns form. */
ax_status ax_bilinear_stage2_batch(int ns, int nm, int ld, int n_jobs, const int64_t *jobs,
                                   void *stream);
/* Workspace (doubles) of ax_mbpt2_pair_sum for n_virt virtuals and n_pair
 * pairs. */
int64_t ax_mbpt2_workspace_doubles(int n_virt, int n_pair);
/* The J-coupled MBPT(2) pair sums of _mbpt2_pair_contribution
 * (methods/mbpt.py:602-790): for every occupied pair p and virtual pair
 * r <= s (virtuals ordered by kappa channel, so r <= s is the reference's
 * unordered channel sum), amp_J = sum_k cd[c][J][k] T^k[r][s] -
 * ce[c][J][k] T^k[s][r] with T^k the Slater table R^k(a r, b s) of item
 * pair_item[p][k] (-1 absent; entry (item, r, s) at tables[item *
 * item_stride + r * row_stride + s]) and c the channel of (pair_type[p],
 * vkappa[r], vkappa[s]) (chan_nj[c] J values from 2J = chan_tjlo[c], step 2;
 * 0 = no contribution); E_p = -sum (2J+1) amp_J^2 / (pair_norm[p] n_rs
 * (veps[r] + veps[s] - pair_eps[p])), n_rs = 2 for r = s, skipped where the
 * denominator is below 1e-10 in magnitude.  pair_energy[p] is written
 * (fixed-order sums); jsum (nullable, n_pair x N x N) receives sum_J
 * (2J+1) amp_J^2 of every counted r <= s entry. */
ax_status ax_mbpt2_pair_sum(int n_virt, int n_kappa, int n_rank, int n_pair, int max_j,
                            const int32_t *vkappa, const double *veps, const double *tables,
                            int64_t item_stride, int64_t row_stride, const int32_t *pair_item,
                            const double *pair_eps,
                            const double *pair_norm, const int32_t *pair_type,
                            const int32_t *chan_nj, const int32_t *chan_tjlo, const double *cd,
                            const double *ce, double *pair_energy, double *jsum,
                            double *workspace, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* ATOMX_SLATER_H */
