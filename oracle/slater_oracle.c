/*
 * CPU FP64 oracle for the Slater-integral path.  TEST INFRASTRUCTURE ONLY:
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this
 * library as the checker.  The product path (atomx_amd) never links or calls it.
 *
 * Every function restates one reference routine (paths relative to
 * src/atomx/ of zezhong-zhang/atomx) in plain C, with the same loop order
 * where the reference fixes a summation order.  Build with
 * -O2 -ffp-contract=off -fno-fast-math so that no FMA contraction or
 * reassociation changes the arithmetic.
 *
 * Array layouts (all row-major, FP64):
 *   gr, grw         (nv, ng)
 *   bsp             (nv, ng, ks)
 *   moments         (kk, nv)            kk = ks(ks+1)/2, or ks1*ks2 (nonsym)
 *   diagonal cells  (kk1, kk2, nv)
 *   rkb bank        (ns, ns, ksl, ksr)
 *   banded density  (ns, ks)
 *   full density    (ns, ns)
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define IDX2(a, b, nb) ((long)(a) * (nb) + (b))

/* ---------------------------------------------------------------------- */
/* Quadrature tables                                                       */
/* ---------------------------------------------------------------------- */

/* Gauss-Legendre nodes/weights on [-1, 1] by Newton iteration on P_n in
 * long double (the reference calls numpy.polynomial.legendre.leggauss,
 * numerics/bspline/basis.py:195-208; both give the nodes to <=1 ulp). */
void orc_gauss_legendre(int n, double *x, double *w) {
    for (int i = 0; i < n; ++i) {
        long double xi = cosl(3.14159265358979323846264338327950288L * (i + 0.75L) / (n + 0.5L));
        long double dp = 1.0L;
        for (int it = 0; it < 100; ++it) {
            long double p0 = 1.0L, p1 = xi;
            for (int j = 2; j <= n; ++j) {
                long double p2 = ((2 * j - 1) * xi * p1 - (j - 1) * p0) / j;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (xi * p1 - p0) / (xi * xi - 1.0L);
            long double dx = p1 / dp;
            xi -= dx;
            if (fabsl(dx) < 1e-19L) break;
        }
        {
            long double p0 = 1.0L, p1 = xi;
            for (int j = 2; j <= n; ++j) {
                long double p2 = ((2 * j - 1) * xi * p1 - (j - 1) * p0) / j;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (xi * p1 - p0) / (xi * xi - 1.0L);
        }
        /* ascending order like leggauss */
        x[n - 1 - i] = (double)xi;
        w[n - 1 - i] = (double)(2.0L / ((1.0L - xi * xi) * dp * dp));
    }
    if (n == 1) { x[0] = 0.0; w[0] = 2.0; }
}

/* numerics/bspline/convol.py:252-291 (_gauss_legendre_01): the same Newton
 * iteration in double, mapped to [0, 1].  Restated literally because the
 * diagonal-cell sub-quadrature of the reference uses exactly these nodes. */
void orc_gauss_legendre_01(int n, double *x, double *w) {
    if (n == 1) { x[0] = 0.5; w[0] = 1.0; return; }
    for (int i = 0; i < n; ++i) {
        double xi = cos(M_PI * (i + 0.75) / (n + 0.5));
        double dp = 0.0, p0, p1;
        for (int it = 0; it < 50; ++it) {
            p0 = 1.0;
            p1 = xi;
            for (int j = 2; j <= n; ++j) {
                double p2 = ((2 * j - 1) * xi * p1 - (j - 1) * p0) / j;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (xi * p1 - p0) / (xi * xi - 1.0);
            double dx = p1 / dp;
            xi -= dx;
            if (fabs(dx) < 1e-15) break;
        }
        x[i] = 0.5 * (xi + 1.0);
        w[i] = 1.0 / ((1.0 - xi * xi) * dp * dp);
    }
}

/* numerics/yk.py:65-84 (yk_partial_matrices): the interpolatory partial
 * interval matrices in the Legendre basis.
 *   mo[i,j] = sum_l partial[i,l] (2l+1)/2 P_l(x_j),
 *   partial[i,0] = x_i + 1, partial[i,l] = (P_{l+1}(x_i) - P_{l-1}(x_i))/(2l+1),
 *   mi = 1 - mo.  Nodes x are passed in (leggauss nodes). */
void orc_yk_partial_matrices_from_nodes(int ng, const double *xi, double *mo, double *mi) {
    double *vals = (double *)malloc(sizeof(double) * ng * (ng + 1));
    double *partial = (double *)malloc(sizeof(double) * ng * ng);
    for (int i = 0; i < ng; ++i) {
        double x = xi[i];
        vals[IDX2(i, 0, ng + 1)] = 1.0;
        if (ng >= 1) vals[IDX2(i, 1, ng + 1)] = x;
        for (int l = 2; l <= ng; ++l)
            vals[IDX2(i, l, ng + 1)] = ((2 * l - 1) * x * vals[IDX2(i, l - 1, ng + 1)]
                                        - (l - 1) * vals[IDX2(i, l - 2, ng + 1)]) / l;
    }
    for (int i = 0; i < ng; ++i) {
        partial[IDX2(i, 0, ng)] = xi[i] + 1.0;
        for (int l = 1; l < ng; ++l)
            partial[IDX2(i, l, ng)] = (vals[IDX2(i, l + 1, ng + 1)] - vals[IDX2(i, l - 1, ng + 1)]) / (2 * l + 1);
    }
    for (int i = 0; i < ng; ++i)
        for (int j = 0; j < ng; ++j) {
            double s = 0.0;
            for (int l = 0; l < ng; ++l)
                s += partial[IDX2(i, l, ng)] * ((2 * l + 1) / 2.0) * vals[IDX2(j, l, ng + 1)];
            mo[IDX2(i, j, ng)] = s;
            mi[IDX2(i, j, ng)] = 1.0 - s;
        }
    free(vals);
    free(partial);
}

void orc_yk_partial_matrices(int ng, double *mo, double *mi) {
    double *x = (double *)malloc(sizeof(double) * ng), *w = (double *)malloc(sizeof(double) * ng);
    orc_gauss_legendre(ng, x, w);
    orc_yk_partial_matrices_from_nodes(ng, x, mo, mi);
    free(x);
    free(w);
}

/* ---------------------------------------------------------------------- */
/* (Y) discretisation: Poisson Y^k on the Gauss grid                        */
/* ---------------------------------------------------------------------- */

/* numerics/yk.py:87-154 (yk_flat), same loop and summation order. */
void orc_yk_flat(const double *rho, const double *r, const double *w, long n, int k,
                 int nv, int ng, const double *mo, const double *mi, double *yk) {
    double *rk = (double *)malloc(sizeof(double) * n);
    double *rkm = (double *)malloc(sizeof(double) * n);
    double *iout = (double *)malloc(sizeof(double) * n);
    double *iin = (double *)malloc(sizeof(double) * n);
    for (long i = 0; i < n; ++i) {
        rk[i] = pow(r[i], (double)k);
        rkm[i] = pow(r[i], (double)(-(k + 1)));
    }
    double base = 0.0;
    for (int m = 0; m < nv; ++m) {
        long off = (long)m * ng;
        for (int i = 0; i < ng; ++i) {
            double s = base;
            for (int j = 0; j < ng; ++j) s += w[off + j] * rho[off + j] * rk[off + j] * mo[IDX2(i, j, ng)];
            iout[off + i] = s;
        }
        for (int j = 0; j < ng; ++j) base += w[off + j] * rho[off + j] * rk[off + j];
    }
    base = 0.0;
    for (int m = nv - 1; m >= 0; --m) {
        long off = (long)m * ng;
        for (int i = 0; i < ng; ++i) {
            double s = base;
            for (int j = 0; j < ng; ++j) s += w[off + j] * rho[off + j] * rkm[off + j] * mi[IDX2(i, j, ng)];
            iin[off + i] = s;
        }
        for (int j = 0; j < ng; ++j) base += w[off + j] * rho[off + j] * rkm[off + j];
    }
    for (long i = 0; i < n; ++i) yk[i] = rkm[i] * iout[i] + rk[i] * iin[i];
    free(rk);
    free(rkm);
    free(iout);
    free(iin);
}

/* numerics/bspline/slater.py:118-130 (slater_rk_poisson): sum(grw*rho13*Y^k[rho24]). */
double orc_slater_rk_poisson(const double *rho13, const double *rho24, const double *gr,
                             const double *grw, int nv, int ng, int k, const double *mo,
                             const double *mi) {
    long n = (long)nv * ng;
    double *yk = (double *)malloc(sizeof(double) * n);
    orc_yk_flat(rho24, gr, grw, n, k, nv, ng, mo, mi, yk);
    double s = 0.0;
    for (long i = 0; i < n; ++i) s += grw[i] * rho13[i] * yk[i];
    free(yk);
    return s;
}

/* numerics/bspline/slater.py:51-60 (_disjoint_cell_integral). */
double orc_disjoint_cell_integral(const double *rho13, const double *rho24, const double *r,
                                  const double *w, long n, int k, int ng) {
    double total = 0.0;
    for (long i = 0; i < n; ++i) {
        for (long j = (i / ng + 1) * ng; j < n; ++j) {
            double kernel = pow(r[i] / r[j], (double)k) / r[j];
            total += w[i] * w[j] * kernel * (rho13[i] * rho24[j] + rho13[j] * rho24[i]);
        }
    }
    return total;
}

/* Batched Y^k for the synthetic-orbital all-quartet workload: for each
 * (density row, k) computes Y^k, then every R^k(p,q) = sum_x grw rho_p Y^k[rho_q]
 * for the requested quartets.  This is the CPU FP64 baseline of the R^k
 * tensor (same arithmetic as orc_yk_flat + slater_rk_poisson, with Y^k
 * computed once per (pair, k) and reused), OpenMP over quartets.
 *   rho     (n_rho, n) pair densities
 *   qp, qq, qk (n_quart) quartet left pair, right pair, rank
 *   returns out (n_quart)                                              */
void orc_rk_quartets(const double *rho, long n_rho, const double *r, const double *w, int nv,
                     int ng, const double *mo, const double *mi, const int *qp, const int *qq,
                     const int *qk, long n_quart, int kmax, int nthreads, double *out) {
    long n = (long)nv * ng;
    /* which (q, k) Y^k rows are needed */
    char *need = (char *)calloc((size_t)n_rho * (kmax + 1), 1);
    for (long t = 0; t < n_quart; ++t) need[(long)qq[t] * (kmax + 1) + qk[t]] = 1;
    long *slot = (long *)malloc(sizeof(long) * n_rho * (kmax + 1));
    long nslots = 0;
    for (long t = 0; t < n_rho * (kmax + 1); ++t) slot[t] = need[t] ? nslots++ : -1;
    double *Y = (double *)malloc(sizeof(double) * (nslots > 0 ? nslots : 1) * n);
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4)
#endif
    for (long t = 0; t < n_rho * (kmax + 1); ++t) {
        if (slot[t] < 0) continue;
        long q = t / (kmax + 1);
        int k = (int)(t % (kmax + 1));
        orc_yk_flat(rho + q * n, r, w, n, k, nv, ng, mo, mi, Y + slot[t] * n);
    }
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (long t = 0; t < n_quart; ++t) {
        const double *a = rho + (long)qp[t] * n;
        const double *y = Y + slot[(long)qq[t] * (kmax + 1) + qk[t]] * n;
        double s = 0.0;
        for (long i = 0; i < n; ++i) s += w[i] * a[i] * y[i];
        out[t] = s;
    }
    free(need);
    free(slot);
    free(Y);
}

/* numerics/bspline/slater_yk.py:35-122 (_exchange_matrix_kernel): ns full
 * cumulative scans (node-wise, no partial-interval rule) plus banded assembly. */
void orc_exchange_matrix_kernel(const double *fb, const double *bsp, const double *gr,
                                const double *grw, int nv, int ng, int ks, int ns, int k,
                                double *K) {
    long n = (long)nv * ng;
    double *rk = (double *)malloc(sizeof(double) * n), *rkm = (double *)malloc(sizeof(double) * n);
    double *integ = (double *)malloc(sizeof(double) * n * ks);
    double *rho = (double *)malloc(sizeof(double) * n);
    double *iout = (double *)malloc(sizeof(double) * n), *iin = (double *)malloc(sizeof(double) * n);
    double *yk = (double *)malloc(sizeof(double) * n);
    memset(K, 0, sizeof(double) * ns * ns);
    for (long i = 0; i < n; ++i) {
        rk[i] = pow(gr[i], (double)k);
        rkm[i] = pow(gr[i], (double)(-(k + 1)));
    }
    for (int iv = 0; iv < nv; ++iv)
        for (int g = 0; g < ng; ++g) {
            long idx = (long)iv * ng + g;
            double wfb = grw[idx] * fb[idx];
            for (int s = 0; s < ks; ++s) integ[idx * ks + s] = wfb * bsp[idx * ks + s];
        }
    for (int nidx = 0; nidx < ns; ++nidx) {
        for (long i = 0; i < n; ++i) rho[i] = 0.0;
        int lo = nidx - ks + 1;
        if (lo < 0) lo = 0;
        int hi = nidx + 1;
        if (hi > nv) hi = nv;
        for (int iv = lo; iv < hi; ++iv) {
            int ls = nidx - iv;
            for (int g = 0; g < ng; ++g) {
                long idx = (long)iv * ng + g;
                rho[idx] = fb[idx] * bsp[idx * ks + ls];
            }
        }
        double s = 0.0;
        for (long i = 0; i < n; ++i) { s += grw[i] * rho[i] * rk[i]; iout[i] = s; }
        s = 0.0;
        for (long i = n - 1; i >= 0; --i) { s += grw[i] * rho[i] * rkm[i]; iin[i] = s; }
        for (long i = 0; i < n; ++i) yk[i] = rkm[i] * iout[i] + rk[i] * iin[i];
        for (int iv = 0; iv < nv; ++iv)
            for (int sidx = 0; sidx < ks; ++sidx) {
                int m = iv + sidx;
                if (m < ns) {
                    double acc = 0.0;
                    for (int g = 0; g < ng; ++g) acc += integ[((long)iv * ng + g) * ks + sidx] * yk[(long)iv * ng + g];
                    K[IDX2(m, nidx, ns)] += acc;
                }
            }
    }
    free(rk); free(rkm); free(integ); free(rho); free(iout); free(iin); free(yk);
}

/* ---------------------------------------------------------------------- */
/* (C) discretisation: DBSR_HF cell algorithm                               */
/* ---------------------------------------------------------------------- */

static double rank_weight(double wgt, double r, int k) {
    if (k > 0) return wgt * pow(r, (double)k);
    if (k < 0) return wgt / pow(r, (double)(-k));
    return wgt;
}

/* numerics/bspline/convol.py:57-90 (_compute_moments_sym). */
void orc_moments_sym(const double *bsp, const double *gr, const double *grw, int nv, int ng,
                     int ks, int k, double *rkm) {
    int kk = ks * (ks + 1) / 2;
    double gw[64];
    memset(rkm, 0, sizeof(double) * kk * nv);
    for (int iv = 0; iv < nv; ++iv) {
        for (int m = 0; m < ng; ++m) gw[m] = rank_weight(grw[IDX2(iv, m, ng)], gr[IDX2(iv, m, ng)], k);
        int ii = 0;
        for (int i = 0; i < ks; ++i)
            for (int j = i; j < ks; ++j) {
                double s = 0.0;
                for (int m = 0; m < ng; ++m)
                    s += bsp[((long)iv * ng + m) * ks + i] * bsp[((long)iv * ng + m) * ks + j] * gw[m];
                rkm[IDX2(ii, iv, nv)] = s;
                ++ii;
            }
    }
}

/* numerics/bspline/convol.py:93-128 (_compute_moments_nonsym). */
void orc_moments_nonsym(const double *bsp1, const double *bsp2, const double *gr,
                        const double *grw, int nv, int ng, int ks1, int ks2, int k, double *rkm) {
    double gw[64];
    for (int iv = 0; iv < nv; ++iv) {
        for (int m = 0; m < ng; ++m) gw[m] = rank_weight(grw[IDX2(iv, m, ng)], gr[IDX2(iv, m, ng)], k);
        int ii = 0;
        for (int i = 0; i < ks1; ++i)
            for (int j = 0; j < ks2; ++j) {
                double s = 0.0;
                for (int m = 0; m < ng; ++m)
                    s += bsp1[((long)iv * ng + m) * ks1 + i] * bsp2[((long)iv * ng + m) * ks2 + j] * gw[m];
                rkm[IDX2(ii, iv, nv)] = s;
                ++ii;
            }
    }
}

/* numerics/bspline/convol.py:294-316 (_bsplvb_inplace). */
void orc_bsplvb(const double *t, int ks, double x, int left, double *b) {
    double dr[32], dl[32];
    b[0] = 1.0;
    for (int i = 1; i < ks; ++i) b[i] = 0.0;
    for (int j = 1; j < ks; ++j) {
        dr[j] = t[left + j] - x;
        dl[j] = x - t[left + 1 - j];
        double saved = 0.0;
        for (int r = 0; r < j; ++r) {
            double term = b[r] / (dr[r + 1] + dl[j - r]);
            b[r] = saved + dr[r + 1] * term;
            saved = dl[j - r] * term;
        }
        b[j] = saved;
    }
}

/* Inner sub-quadrature INTP[j,jp,m] = int_{xbase}^{gr[iv,m]} B_j B_jp r^k
 * (convol.py:187-224 and :423-448), packed upper triangle (kk, ng). */
static void triang_inner(const double *t, int ks, int iv, const double *gr_row, int ng, int k,
                         int n_sub, const double *x01, const double *w01, double *intp) {
    int left = ks - 1 + iv;
    double xbase = t[left];
    double bsub[32 * 32];
    double gwrk[64];
    for (int m = 0; m < ng; ++m) {
        double span = gr_row[m] - xbase;
        int jj = 0;
        if (span < 1e-30) {
            for (int j = 0; j < ks; ++j)
                for (int jp = j; jp < ks; ++jp) intp[IDX2(jj++, m, ng)] = 0.0;
            continue;
        }
        for (int nn = 0; nn < n_sub; ++nn) {
            double gx = xbase + span * x01[nn];
            orc_bsplvb(t, ks, gx, left, bsub + nn * ks);
            double wgt = span * w01[nn];
            gwrk[nn] = rank_weight(wgt, gx, k);
        }
        for (int j = 0; j < ks; ++j)
            for (int jp = j; jp < ks; ++jp) {
                double s = 0.0;
                for (int nn = 0; nn < n_sub; ++nn) s += bsub[nn * ks + j] * bsub[nn * ks + jp] * gwrk[nn];
                intp[IDX2(jj++, m, ng)] = s;
            }
    }
}

/* numerics/bspline/convol.py:135-249 (_triang_sym). rkd (kk, kk, nv). */
void orc_triang_sym(const double *bsp, const double *gr, const double *grw, const double *t,
                    int nv, int ng, int ks, int k, double *rkd) {
    int kk = ks * (ks + 1) / 2;
    int n_sub = ks + (k >= 0 ? k / 2 : 0);
    double x01[64], w01[64], ow[64], bw[64];
    orc_gauss_legendre_01(n_sub, x01, w01);
    double *intp = (double *)malloc(sizeof(double) * kk * ng);
    double *a = (double *)malloc(sizeof(double) * kk * kk);
    for (int iv = 0; iv < nv; ++iv) {
        triang_inner(t, ks, iv, gr + (long)iv * ng, ng, k, n_sub, x01, w01, intp);
        for (int m = 0; m < ng; ++m) ow[m] = grw[IDX2(iv, m, ng)] / pow(gr[IDX2(iv, m, ng)], (double)(k + 1));
        int ii = 0;
        for (int i = 0; i < ks; ++i)
            for (int ip = i; ip < ks; ++ip) {
                for (int m = 0; m < ng; ++m)
                    bw[m] = bsp[((long)iv * ng + m) * ks + i] * bsp[((long)iv * ng + m) * ks + ip] * ow[m];
                for (int jj = 0; jj < kk; ++jj) {
                    double s = 0.0;
                    for (int m = 0; m < ng; ++m) s += bw[m] * intp[IDX2(jj, m, ng)];
                    a[IDX2(ii, jj, kk)] = s;
                }
                ++ii;
            }
        for (int ii2 = 0; ii2 < kk; ++ii2)
            for (int jj = 0; jj < kk; ++jj)
                rkd[((long)ii2 * kk + jj) * nv + iv] = a[IDX2(ii2, jj, kk)] + a[IDX2(jj, ii2, kk)];
    }
    free(intp);
    free(a);
}

/* numerics/bspline/convol.py:387-529 (_triang_pqpq).  Left basis (bsp_p,
 * ksp, tp) is the "P-pair", right basis (bsp_q, ksq, tq).  rkd (kkp, kkq, nv). */
void orc_triang_pqpq(const double *bsp_p, const double *bsp_q, const double *gr,
                     const double *grw, const double *tp, const double *tq, int nv, int ng,
                     int ksp, int ksq, int k, double *rkd) {
    int kkp = ksp * (ksp + 1) / 2, kkq = ksq * (ksq + 1) / 2;
    int n_sub = (ksp > ksq ? ksp : ksq) + (k >= 0 ? k / 2 : 0);
    double x01[64], w01[64], ow[64], bw[64];
    orc_gauss_legendre_01(n_sub, x01, w01);
    double *intq = (double *)malloc(sizeof(double) * kkq * ng);
    double *intp = (double *)malloc(sizeof(double) * kkp * ng);
    double *a = (double *)malloc(sizeof(double) * kkp * kkq);
    double *at = (double *)malloc(sizeof(double) * kkp * kkq);
    for (int iv = 0; iv < nv; ++iv) {
        triang_inner(tq, ksq, iv, gr + (long)iv * ng, ng, k, n_sub, x01, w01, intq);
        for (int m = 0; m < ng; ++m) ow[m] = grw[IDX2(iv, m, ng)] / pow(gr[IDX2(iv, m, ng)], (double)(k + 1));
        int ii = 0;
        for (int i = 0; i < ksp; ++i)
            for (int ip = i; ip < ksp; ++ip) {
                for (int m = 0; m < ng; ++m)
                    bw[m] = bsp_p[((long)iv * ng + m) * ksp + i] * bsp_p[((long)iv * ng + m) * ksp + ip] * ow[m];
                for (int jj = 0; jj < kkq; ++jj) {
                    double s = 0.0;
                    for (int m = 0; m < ng; ++m) s += bw[m] * intq[IDX2(jj, m, ng)];
                    a[IDX2(ii, jj, kkq)] = s;
                }
                ++ii;
            }
        triang_inner(tp, ksp, iv, gr + (long)iv * ng, ng, k, n_sub, x01, w01, intp);
        int jj = 0;
        for (int j = 0; j < ksq; ++j)
            for (int jp = j; jp < ksq; ++jp) {
                for (int m = 0; m < ng; ++m)
                    bw[m] = bsp_q[((long)iv * ng + m) * ksq + j] * bsp_q[((long)iv * ng + m) * ksq + jp] * ow[m];
                for (int ii2 = 0; ii2 < kkp; ++ii2) {
                    double s = 0.0;
                    for (int m = 0; m < ng; ++m) s += bw[m] * intp[IDX2(ii2, m, ng)];
                    at[IDX2(ii2, jj, kkq)] = s;
                }
                ++jj;
            }
        for (int ii2 = 0; ii2 < kkp; ++ii2)
            for (int jj2 = 0; jj2 < kkq; ++jj2)
                rkd[((long)ii2 * kkq + jj2) * nv + iv] = a[IDX2(ii2, jj2, kkq)] + at[IDX2(ii2, jj2, kkq)];
    }
    free(intq); free(intp); free(a); free(at);
}

/* numerics/bspline/convol.py:323-380 (_assemble_rkb_sym) and :532-572
 * (_assemble_rkb_pqpq), written once for left order ksl, right order ksr:
 *   iv < jv: c = lo_l[ii, iv] * hi_r[jj, jv]
 *   iv > jv: c = lo_r[jj, jv] * hi_l[ii, iv]
 *   iv = jv: c = diag[ii, jj, iv]
 * For the symmetric type lo_l = lo_r = rkd1 (r^k), hi_l = hi_r = rkd2. */
void orc_assemble_rkb(const double *lo_l, const double *hi_r, const double *lo_r,
                      const double *hi_l, const double *diag, int nv, int ksl, int ksr, int ns,
                      double *rkb) {
    int kkr = ksr * (ksr + 1) / 2;
    memset(rkb, 0, sizeof(double) * (long)ns * ns * ksl * ksr);
    for (int jv = 0; jv < nv; ++jv) {
        int jj = 0;
        for (int jh = 0; jh < ksr; ++jh) {
            int j = jv + jh;
            for (int jhp = jh; jhp < ksr; ++jhp) {
                int jp = jhp - jh;
                for (int iv = 0; iv < nv; ++iv) {
                    int ii = 0;
                    for (int ih = 0; ih < ksl; ++ih) {
                        int i = iv + ih;
                        for (int ihp = ih; ihp < ksl; ++ihp) {
                            int ip = ihp - ih;
                            double c;
                            if (iv < jv) c = lo_l[IDX2(ii, iv, nv)] * hi_r[IDX2(jj, jv, nv)];
                            else if (iv > jv) c = lo_r[IDX2(jj, jv, nv)] * hi_l[IDX2(ii, iv, nv)];
                            else c = diag[((long)ii * kkr + jj) * nv + iv];
                            if (i < ns && j < ns)
                                rkb[(((long)i * ns + j) * ksl + ip) * ksr + jp] += c;
                            ++ii;
                        }
                    }
                }
                ++jj;
            }
        }
    }
}

/* numerics/bspline/convol.py:688-753 (_density_sym with zero padding). */
void orc_density_sym(const double *p1, int n1, const double *p2, int n2, int ns, int ks, double *d) {
    memset(d, 0, sizeof(double) * ns * ks);
    int lim = n1 < ns ? n1 : ns;
    for (int i = 0; i < lim; ++i) {
        if (i < n2) d[IDX2(i, 0, ks)] = p1[i] * p2[i];
        for (int s = 1; s < ks; ++s) {
            int ip = i + s;
            if (ip >= ns) break;
            double v = 0.0;
            if (ip < n2) v += p1[i] * p2[ip];
            if (ip < n1 && i < n2) v += p1[ip] * p2[i];
            d[IDX2(i, s, ks)] = v;
        }
    }
}

/* numerics/bspline/convol.py:756-771 (_density_full). */
void orc_density_full(const double *p1, int n1, const double *p2, int n2, int ns, double *d) {
    memset(d, 0, sizeof(double) * ns * ns);
    int a = n1 < ns ? n1 : ns, b = n2 < ns ? n2 : ns;
    for (int i = 0; i < a; ++i)
        for (int j = 0; j < b; ++j) d[IDX2(i, j, ns)] = p1[i] * p2[j];
}

/* numerics/bspline/convol.py:788-811 and :829-849 (_convol_case1[_mixed]):
 * a[i, ip] = sum_{jp, j} rkb[i, j, ip, jp] * d[j, jp]. */
void orc_convol_case1(const double *rkb, const double *d, int ns, int ksl, int ksr, double *a) {
    memset(a, 0, sizeof(double) * ns * ksl);
    for (long idx = 0; idx < (long)ns * ksl; ++idx) {
        int i = (int)(idx / ksl), ip = (int)(idx % ksl);
        if (i + ip >= ns) continue;
        double s = 0.0;
        for (int jp = 0; jp < ksr; ++jp)
            for (int j = 0; j < ns; ++j) {
                if (j + jp >= ns) continue;
                s += rkb[(((long)i * ns + j) * ksl + ip) * ksr + jp] * d[IDX2(j, jp, ksr)];
            }
        a[IDX2(i, ip, ksl)] = s;
    }
}

/* numerics/bspline/convol.py:852-874 (_convol_case2) and :1310-1345
 * (_convol_case2_mixed): a[j, jp] = sum_{ip, i < n_left} rkb[i,j,ip,jp] d[i,ip]. */
void orc_convol_case2(const double *rkb, const double *d, int n_left, int ns, int ksl, int ksr,
                      double *a) {
    memset(a, 0, sizeof(double) * ns * ksr);
    for (long idx = 0; idx < (long)ns * ksr; ++idx) {
        int j = (int)(idx / ksr), jp = (int)(idx % ksr);
        if (j + jp >= ns) continue;
        double s = 0.0;
        for (int ip = 0; ip < ksl; ++ip)
            for (int i = 0; i < n_left; ++i) {
                if (i + ip >= n_left) continue;
                s += rkb[(((long)i * ns + j) * ksl + ip) * ksr + jp] * d[IDX2(i, ip, ksl)];
            }
        a[IDX2(j, jp, ksr)] = s;
    }
}

/* numerics/bspline/convol.py:877-948 (_convol_case4), gather formulation,
 * four index families in the reference order. */
void orc_convol_case4(const double *rkb, const double *d, int ns, int ks, int ksr, double *a) {
#define RKB(i, j, ip, jp) rkb[(((long)(i) * ns + (j)) * ks + (ip)) * ksr + (jp)]
    for (int row = 0; row < ns; ++row)
        for (int col = 0; col < ns; ++col) {
            double s = 0.0;
            for (int ip = 0; ip < ks; ++ip) {
                int i = row - ip;
                if (i < 0) continue;
                for (int jp = 0; jp < ksr; ++jp) {
                    int jj = col + jp;
                    if (jj >= ns) continue;
                    s += RKB(i, col, ip, jp) * d[IDX2(i, jj, ns)];
                }
            }
            for (int ip = 1; ip < ks; ++ip) {
                int ii = row + ip;
                if (ii >= ns) continue;
                for (int jp = 0; jp < ksr; ++jp) {
                    int jj = col + jp;
                    if (jj >= ns) continue;
                    s += RKB(row, col, ip, jp) * d[IDX2(ii, jj, ns)];
                }
            }
            for (int jp = 1; jp < ksr; ++jp) {
                int jv = col - jp;
                if (jv < 0) continue;
                for (int ip = 0; ip < ks; ++ip) {
                    int i = row - ip;
                    if (i < 0) continue;
                    s += RKB(i, jv, ip, jp) * d[IDX2(i, jv, ns)];
                }
            }
            for (int ip = 1; ip < ks; ++ip)
                for (int jp = 1; jp < ksr; ++jp) {
                    int jv = col - jp;
                    if (jv < 0) continue;
                    int ii = row + ip;
                    if (ii >= ns) continue;
                    s += RKB(row, jv, ip, jp) * d[IDX2(ii, jv, ns)];
                }
            a[IDX2(row, col, ns)] = s;
        }
#undef RKB
}

/* numerics/bspline/convol.py:1348-1356 (_sum_amb_banded); a and b have
 * row strides lda/ldb (their band widths). */
double orc_sum_amb_banded(const double *a, int lda, const double *b, int ldb, int ns, int ks) {
    double s = 0.0;
    for (int ip = 0; ip < ks; ++ip)
        for (int i = 0; i < ns; ++i)
            if (i + ip < ns) s += a[IDX2(i, ip, lda)] * b[IDX2(i, ip, ldb)];
    return s;
}

/* numerics/bspline/convol.py:814-825 (_banded_to_full_add). */
void orc_banded_to_full_add(double *V, const double *vb, int ns, int ks, double coeff) {
    for (int s = 0; s < ks; ++s)
        for (int i = 0; i < ns - s; ++i) {
            double val = coeff * vb[IDX2(i, s, ks)];
            V[IDX2(i, i + s, ns)] += val;
            if (s > 0) V[IDX2(i + s, i, ns)] += val;
        }
}

/* numerics/bspline/integrals.py:18-56 (_assemble_integral): <B1_i|w|B2_j>
 * on an (mm, mm) matrix, accumulated cell by cell in the reference's order. */
void orc_assemble_integral(int nv, int ng, int ks1, int ks2, int mm, const double *bsp1,
                           const double *bsp2, const double *weights, double *mat) {
    memset(mat, 0, sizeof(double) * (size_t)mm * mm);
    for (int iv = 0; iv < nv; ++iv)
        for (int i = 0; i < ks1; ++i) {
            const int gi = iv + i;
            if (gi >= mm) continue;
            for (int j = 0; j < ks2; ++j) {
                const int gj = iv + j;
                if (gj >= mm) continue;
                double s = 0.0;
                for (int m = 0; m < ng; ++m)
                    s += weights[IDX2(iv, m, ng)] * bsp1[((long)iv * ng + m) * ks1 + i] *
                         bsp2[((long)iv * ng + m) * ks2 + j];
                mat[IDX2(gi, gj, mm)] += s;
            }
        }
}

/* ---------------------------------------------------------------------- */
/* Semiseparable (scan) form of the cell algorithm                          */
/* ---------------------------------------------------------------------- */

/* numerics/bspline/convol.py:1061-1073 (_global_pair_moments): (ks, ns). */
void orc_global_pair_moments(const double *local, int nv, int ks, int ns, double *mom) {
    memset(mom, 0, sizeof(double) * ks * ns);
    for (int iv = 0; iv < nv; ++iv) {
        int lp = 0;
        for (int first = 0; first < ks; ++first) {
            int i = iv + first;
            for (int second = first; second < ks; ++second) {
                int off = second - first;
                if (i + off < ns) mom[IDX2(off, i, ns)] += local[IDX2(lp, iv, nv)];
                ++lp;
            }
        }
    }
}

/* numerics/bspline/convol.py:1076-1122 (_assemble_near_rkb):
 * near (ksl, ksr, ns, ksl+ksr-1). */
void orc_assemble_near_rkb(const double *rising_left, const double *falling_right,
                           const double *rising_right, const double *falling_left,
                           const double *diag, int nv, int ksl, int ksr, int ns, double *near) {
    int W = ksl + ksr - 1;
    int kkr = ksr * (ksr + 1) / 2;
    memset(near, 0, sizeof(double) * (long)ksl * ksr * ns * W);
    for (int jv = 0; jv < nv; ++jv) {
        int rl = 0;
        for (int jh = 0; jh < ksr; ++jh) {
            int j = jv + jh;
            for (int jhp = jh; jhp < ksr; ++jhp) {
                int jp = jhp - jh;
                if (j < ns) {
                    int ivs = j - ksr + jp - ksl + 2;
                    if (ivs < 0) ivs = 0;
                    int ive = j + ksl;
                    if (ive > nv) ive = nv;
                    for (int iv = ivs; iv < ive; ++iv) {
                        int ll = 0;
                        for (int ih = 0; ih < ksl; ++ih) {
                            int i = iv + ih;
                            for (int ihp = ih; ihp < ksl; ++ihp) {
                                int ip = ihp - ih;
                                int first_j = i + ip - ksl + 1;
                                int last_j = i + ksr - jp - 1;
                                if (i < ns && first_j <= j && j <= last_j) {
                                    double v;
                                    if (iv < jv) v = rising_left[IDX2(ll, iv, nv)] * falling_right[IDX2(rl, jv, nv)];
                                    else if (iv > jv) v = rising_right[IDX2(rl, jv, nv)] * falling_left[IDX2(ll, iv, nv)];
                                    else v = diag[((long)ll * kkr + rl) * nv + iv];
                                    near[(((long)ip * ksr + jp) * ns + i) * W + (j - first_j)] += v;
                                }
                                ++ll;
                            }
                        }
                    }
                }
                ++rl;
            }
        }
    }
}

/* numerics/bspline/convol.py:1125-1162 (_convol_case1_semiseparable_t).
 * rising/falling_{left,right} are global pair moments (ks, ns). */
void orc_convol_case1_semiseparable(const double *near, const double *rising_left,
                                    const double *falling_left, const double *rising_right,
                                    const double *falling_right, const double *density, int ns,
                                    int ksl, int ksr, double *result) {
    int W = ksl + ksr - 1;
    double *prefix = (double *)calloc((size_t)ksr * (ns + 1), sizeof(double));
    double *suffix = (double *)calloc((size_t)ksr * (ns + 1), sizeof(double));
    for (int jp = 0; jp < ksr; ++jp) {
        int end = ns - jp;
        for (int j = 0; j < end; ++j)
            prefix[IDX2(jp, j + 1, ns + 1)] = prefix[IDX2(jp, j, ns + 1)] + rising_right[IDX2(jp, j, ns)] * density[IDX2(j, jp, ksr)];
        for (int j = end - 1; j >= 0; --j)
            suffix[IDX2(jp, j, ns + 1)] = suffix[IDX2(jp, j + 1, ns + 1)] + falling_right[IDX2(jp, j, ns)] * density[IDX2(j, jp, ksr)];
    }
    memset(result, 0, sizeof(double) * ns * ksl);
    for (int ip = 0; ip < ksl; ++ip)
        for (int i = 0; i < ns - ip; ++i) {
            int first_near = i + ip - ksl + 1;
            double value = 0.0;
            for (int jp = 0; jp < ksr; ++jp) {
                int end = ns - jp;
                int fn0 = first_near > 0 ? first_near : 0;
                int before = end < fn0 ? end : fn0;
                int aft = i + ksr - jp;
                if (aft < 0) aft = 0;
                int after = end < aft ? end : aft;
                value += falling_left[IDX2(ip, i, ns)] * prefix[IDX2(jp, before, ns + 1)];
                value += rising_left[IDX2(ip, i, ns)] * suffix[IDX2(jp, after, ns + 1)];
                for (int j = fn0; j < after; ++j)
                    value += near[(((long)ip * ksr + jp) * ns + i) * W + (j - first_near)] * density[IDX2(j, jp, ksr)];
            }
            result[IDX2(i, ip, ksl)] = value;
        }
    free(prefix);
    free(suffix);
}

/* numerics/bspline/convol.py:1225-1247 (_add_semiseparable_mixed_row);
 * the symmetric variant (:1165-1184) is the same with left == right. */
static void semi_mixed_row(double *buf, const double *drow, const double *near,
                           const double *rl, const double *fl, const double *rr, const double *fr,
                           int i, int ip, int jp, int ksl, int ksr, int ns, int j_stop,
                           int out_shift, int d_shift) {
    int W = ksl + ksr - 1;
    int first_near = i + ip - ksl + 1;
    int last_near = i + ksr - jp - 1;
    int e1 = j_stop < first_near ? j_stop : first_near;
    double inside = fl[IDX2(ip, i, ns)];
    for (int j = 0; j < e1; ++j) buf[j + out_shift] += inside * rr[IDX2(jp, j, ns)] * drow[j + d_shift];
    int s2 = first_near > 0 ? first_near : 0;
    int e2 = j_stop < last_near + 1 ? j_stop : last_near + 1;
    const double *nr = near + (((long)ip * ksr + jp) * ns + i) * W;
    for (int j = s2; j < e2; ++j) buf[j + out_shift] += nr[j - first_near] * drow[j + d_shift];
    double outside = rl[IDX2(ip, i, ns)];
    int s3 = last_near + 1 > 0 ? last_near + 1 : 0;
    for (int j = s3; j < j_stop; ++j) buf[j + out_shift] += outside * fr[IDX2(jp, j, ns)] * drow[j + d_shift];
}

/* numerics/bspline/convol.py:1187-1222 (_convol_case4_semiseparable_t, symmetric,
 * lower triangle mirrored) when symmetric != 0, else :1250-1303
 * (_convol_case4_semiseparable_mixed_t, full).  Note the symmetric kernel
 * evaluates (falling*rising)*density in the order inside*rising*density,
 * which the shared row routine reproduces. */
void orc_convol_case4_semiseparable(const double *near, const double *rl, const double *fl,
                                    const double *rr, const double *fr, const double *density,
                                    int ns, int ksl, int ksr, int symmetric, double *result) {
    memset(result, 0, sizeof(double) * ns * ns);
    for (int row = 0; row < ns; ++row) {
        double *buf = result + (long)row * ns;
        int limit = symmetric ? row + 1 : ns;
        for (int ip = 0; ip < ksl; ++ip) {
            int i = row - ip;
            if (i < 0) continue;
            for (int jp = 0; jp < ksr; ++jp) {
                int js = (ns - jp) < limit ? (ns - jp) : limit;
                semi_mixed_row(buf, density + (long)i * ns, near, rl, fl, rr, fr, i, ip, jp, ksl, ksr, ns, js, 0, jp);
            }
        }
        for (int ip = 1; ip < ksl; ++ip) {
            int other = row + ip;
            if (other >= ns) continue;
            for (int jp = 0; jp < ksr; ++jp) {
                int js = (ns - jp) < limit ? (ns - jp) : limit;
                semi_mixed_row(buf, density + (long)other * ns, near, rl, fl, rr, fr, row, ip, jp, ksl, ksr, ns, js, 0, jp);
            }
        }
        for (int jp = 1; jp < ksr; ++jp)
            for (int ip = 0; ip < ksl; ++ip) {
                int i = row - ip;
                if (i < 0) continue;
                int js = symmetric ? limit - jp : ns - jp;
                semi_mixed_row(buf, density + (long)i * ns, near, rl, fl, rr, fr, i, ip, jp, ksl, ksr, ns, js, jp, 0);
            }
        for (int ip = 1; ip < ksl; ++ip) {
            int other = row + ip;
            if (other >= ns) continue;
            for (int jp = 1; jp < ksr; ++jp) {
                int js = symmetric ? limit - jp : ns - jp;
                semi_mixed_row(buf, density + (long)other * ns, near, rl, fl, rr, fr, row, ip, jp, ksl, ksr, ns, js, jp, 0);
            }
        }
        if (symmetric)
            for (int col = 0; col < row; ++col) result[IDX2(col, row, ns)] = buf[col];
    }
}

/* numerics/bspline/convol.py:1000-1058 (_convol_case4_t) with
 * symmetric_output: the four families in the transposed-bank order
 * (bit-identical to _convol_case4 without symmetric_output). */
void orc_convol_case4_t(const double *rkb, const double *d, int ns, int ks, int ksr,
                        int symmetric_output, double *a) {
#define RKB(i, j, ip, jp) rkb[(((long)(i) * ns + (j)) * ks + (ip)) * ksr + (jp)]
    memset(a, 0, sizeof(double) * ns * ns);
    for (int row = 0; row < ns; ++row) {
        double *buf = a + (long)row * ns;
        int col_limit = symmetric_output ? row + 1 : ns;
        for (int ip = 0; ip < ks; ++ip) {
            int i = row - ip;
            if (i < 0) continue;
            for (int jp = 0; jp < ksr; ++jp) {
                int lim = (ns - jp) < col_limit ? (ns - jp) : col_limit;
                for (int col = 0; col < lim; ++col) buf[col] += RKB(i, col, ip, jp) * d[IDX2(i, col + jp, ns)];
            }
        }
        for (int ip = 1; ip < ks; ++ip) {
            int ii = row + ip;
            if (ii >= ns) continue;
            for (int jp = 0; jp < ksr; ++jp) {
                int lim = (ns - jp) < col_limit ? (ns - jp) : col_limit;
                for (int col = 0; col < lim; ++col) buf[col] += RKB(row, col, ip, jp) * d[IDX2(ii, col + jp, ns)];
            }
        }
        for (int jp = 1; jp < ksr; ++jp)
            for (int ip = 0; ip < ks; ++ip) {
                int i = row - ip;
                if (i < 0) continue;
                for (int col = jp; col < col_limit; ++col) buf[col] += RKB(i, col - jp, ip, jp) * d[IDX2(i, col - jp, ns)];
            }
        for (int ip = 1; ip < ks; ++ip) {
            int ii = row + ip;
            if (ii >= ns) continue;
            for (int jp = 1; jp < ksr; ++jp)
                for (int col = jp; col < col_limit; ++col) buf[col] += RKB(row, col - jp, ip, jp) * d[IDX2(ii, col - jp, ns)];
        }
        if (symmetric_output)
            for (int col = 0; col < row; ++col) a[IDX2(col, row, ns)] = buf[col];
    }
#undef RKB
}

/* numerics/bspline/convol.py:971-997 (_convol_case1_t) — same terms as
 * case 1, jp outer / j inner (the reference compiles it with fastmath, so
 * only agreement to ~5e-15 relative is meaningful). */
void orc_convol_case1_t(const double *rkb, const double *d, int ns, int ksl, int ksr, double *a) {
    orc_convol_case1(rkb, d, ns, ksl, ksr, a);
}
