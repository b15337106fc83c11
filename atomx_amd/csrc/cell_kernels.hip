// (C) DBSR cell algorithm on the GPU: cell moments, diagonal-cell double
// integrals, the pair x pair R^k bank (or only its overlap strip), global
// pair moments, banded/full densities and the direct (case 1), left-pair
// (case 2) and exchange (case 4) contractions in bank and semiseparable form.
//
// Bank layout in HBM: bankT[ip][jp][i][j] = R^k(B_i B_{i+ip}; B_j B_{j+jp})
// (the reference's transposed layout rkbT, convol.py:1572-1591), so for a
// fixed band pair (ip, jp) the bank is an ns x ns matrix with j contiguous.
// Every contraction is batched over nb densities that share one bank.
#include "ax_common.h"

namespace {

constexpr int kT = 256;
constexpr int kMaxKs = 16;

__device__ inline double rank_weight(double wgt, double r, int k) {
    if (k > 0) return wgt * pow(r, static_cast<double>(k));
    if (k < 0) return wgt / pow(r, static_cast<double>(-k));
    return wgt;
}

// packed upper-triangle index of local pair (h, hp), hp >= h, order ks
__device__ inline int packed(int h, int hp, int ks) { return h * ks - h * (h - 1) / 2 + (hp - h); }

// Cox-de Boor values of the ks nonzero B-splines at x, t[left] <= x < t[left+1]
// (convol.py:294-316; same operation order).
__device__ inline void bsplvb(const double *__restrict__ t, int ks, double x, int left, double *b) {
    double dr[kMaxKs], dl[kMaxKs];
    b[0] = 1.0;
    for (int i = 1; i < ks; ++i) b[i] = 0.0;
    for (int j = 1; j < ks; ++j) {
        dr[j] = t[left + j] - x;
        dl[j] = x - t[left + 1 - j];
        double saved = 0.0;
        for (int r = 0; r < j; ++r) {
            const double term = b[r] / (dr[r + 1] + dl[j - r]);
            b[r] = saved + dr[r + 1] * term;
            saved = dl[j - r] * term;
        }
        b[j] = saved;
    }
}

// ---- (a10) cell moments  rkm[ii, iv] = sum_m B_i B_j grw r^k -----------------
__global__ void moments_kernel(const double *__restrict__ bsp1, const double *__restrict__ bsp2,
                               const double *__restrict__ gr, const double *__restrict__ grw,
                               int nv, int ng, int ks1, int ks2, int k, int sym,
                               double *__restrict__ out) {
    const int kk = sym ? ks1 * (ks1 + 1) / 2 : ks1 * ks2;
    const int64_t total = (int64_t)kk * nv;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int ii = static_cast<int>(t / nv);
        const int iv = static_cast<int>(t - (int64_t)ii * nv);
        int i, j;
        if (sym) {
            i = 0;
            int rem = ii;
            while (rem >= ks1 - i) {
                rem -= ks1 - i;
                ++i;
            }
            j = i + rem;
        } else {
            i = ii / ks2;
            j = ii - i * ks2;
        }
        double s = 0.0;
        for (int m = 0; m < ng; ++m) {
            const int64_t x = (int64_t)iv * ng + m;
            const double gw = rank_weight(grw[x], gr[x], k);
            s += bsp1[x * ks1 + i] * bsp2[x * ks2 + j] * gw;
        }
        out[t] = s;
    }
}

// ---- (a11) diagonal cells ----------------------------------------------------
// rkd[ii, jj, iv] = a[ii, jj] + aT[ii, jj],
//   a  = sum_m (B^l_i B^l_ip grw / r^{k+1})(m) INTP^r[jj](m)
//   aT = sum_m (B^r_j B^r_jp grw / r^{k+1})(m) INTP^l[ii](m)
// with INTP[jj](m) the n_sub-point Gauss integral of B_j B_jp r^k from the
// cell start to node m (convol.py:135-249 symmetric, :387-529 mixed).
__device__ void triang_inner(const double *__restrict__ t, int ks, int iv, int ng, int k,
                             int n_sub, const double *__restrict__ x01,
                             const double *__restrict__ w01, const double *__restrict__ gr,
                             double *bsub, double *gw, double *intp) {
    const int left = ks - 1 + iv;
    const double xbase = t[left];
    for (int p = threadIdx.x; p < ng * n_sub; p += blockDim.x) {
        const int m = p / n_sub, nn = p - (p / n_sub) * n_sub;
        const double span = gr[(int64_t)iv * ng + m] - xbase;
        if (span < 1e-30) {
            gw[p] = 0.0;
            for (int q = 0; q < ks; ++q) bsub[p * ks + q] = 0.0;
            continue;
        }
        const double gx = xbase + span * x01[nn];
        bsplvb(t, ks, gx, left, bsub + p * ks);
        gw[p] = rank_weight(span * w01[nn], gx, k);
    }
    __syncthreads();
    const int kk = ks * (ks + 1) / 2;
    for (int p = threadIdx.x; p < kk * ng; p += blockDim.x) {
        const int jj = p / ng, m = p - (p / ng) * ng;
        int j = 0, rem = jj;
        while (rem >= ks - j) {
            rem -= ks - j;
            ++j;
        }
        const int jp = j + rem;
        const double span = gr[(int64_t)iv * ng + m] - xbase;
        double s = 0.0;
        if (!(span < 1e-30)) {
            for (int nn = 0; nn < n_sub; ++nn) {
                const int q = m * n_sub + nn;
                s += bsub[q * ks + j] * bsub[q * ks + jp] * gw[q];
            }
        }
        intp[jj * ng + m] = s;
    }
    __syncthreads();
}

__global__ __launch_bounds__(kT) void triang_kernel(
    const double *__restrict__ bsp_l, const double *__restrict__ bsp_r,
    const double *__restrict__ gr, const double *__restrict__ grw, const double *__restrict__ t_l,
    const double *__restrict__ t_r, int nv, int ng, int ksl, int ksr, int k, int n_sub,
    const double *__restrict__ x01, const double *__restrict__ w01, double *__restrict__ out) {
    extern __shared__ double sm[];
    const int kkl = ksl * (ksl + 1) / 2, kkr = ksr * (ksr + 1) / 2;
    const int ksm = ksl > ksr ? ksl : ksr;
    double *bsub = sm;                              // [ng*n_sub*ksm]
    double *gw = bsub + ng * n_sub * ksm;           // [ng*n_sub]
    double *intp_l = gw + ng * n_sub;               // [kkl*ng]
    double *intp_r = intp_l + kkl * ng;             // [kkr*ng]
    double *bw_l = intp_r + kkr * ng;               // [kkl*ng]
    double *bw_r = bw_l + kkl * ng;                 // [kkr*ng]
    for (int iv = blockIdx.x; iv < nv; iv += gridDim.x) {
        triang_inner(t_r, ksr, iv, ng, k, n_sub, x01, w01, gr, bsub, gw, intp_r);
        triang_inner(t_l, ksl, iv, ng, k, n_sub, x01, w01, gr, bsub, gw, intp_l);
        for (int p = threadIdx.x; p < (kkl + kkr) * ng; p += blockDim.x) {
            const bool left = p < kkl * ng;
            const int q = left ? p : p - kkl * ng;
            const int ks = left ? ksl : ksr;
            const double *bsp = left ? bsp_l : bsp_r;
            const int ii = q / ng, m = q - (q / ng) * ng;
            int i = 0, rem = ii;
            while (rem >= ks - i) {
                rem -= ks - i;
                ++i;
            }
            const int ip = i + rem;
            const int64_t x = (int64_t)iv * ng + m;
            const double ow = grw[x] / pow(gr[x], static_cast<double>(k + 1));
            (left ? bw_l : bw_r)[q] = bsp[x * ks + i] * bsp[x * ks + ip] * ow;
        }
        __syncthreads();
        for (int p = threadIdx.x; p < kkl * kkr; p += blockDim.x) {
            const int ii = p / kkr, jj = p - (p / kkr) * kkr;
            double a = 0.0, at = 0.0;
            for (int m = 0; m < ng; ++m) a += bw_l[ii * ng + m] * intp_r[jj * ng + m];
            for (int m = 0; m < ng; ++m) at += bw_r[jj * ng + m] * intp_l[ii * ng + m];
            out[((int64_t)ii * kkr + jj) * nv + iv] = a + at;
        }
        __syncthreads();
    }
}

// ---- (a12, a17) pair x pair integrals from cell moments -------------------------
// R(i,ip; j,jp) = sum_{jv} sum_{iv} c(iv, jv) over the cells of the two
// pairs, jv then iv ascending (the accumulation order of convol.py:323-380):
//   iv < jv: lo_l[ii, iv] hi_r[jj, jv];  iv > jv: lo_r[jj, jv] hi_l[ii, iv];
//   iv = jv: diag[ii, jj, iv].
// mode 0 writes the whole bank bankT[ip][jp][i][j]; mode 1 only the overlap
// strip near[ip][jp][i][j - first_j], first_j = i+ip-ksl+1 .. i+ksr-jp-1
// (convol.py:1076-1122).
__device__ inline double pair_integral(const double *__restrict__ lo_l,
                                       const double *__restrict__ hi_r,
                                       const double *__restrict__ lo_r,
                                       const double *__restrict__ hi_l,
                                       const double *__restrict__ diag, int nv, int ksl, int ksr,
                                       int i, int ip, int j, int jp) {
    const int kkr = ksr * (ksr + 1) / 2;
    int jv0 = j + jp - ksr + 1;
    if (jv0 < 0) jv0 = 0;
    const int jv1 = j < nv - 1 ? j : nv - 1;
    int iv0 = i + ip - ksl + 1;
    if (iv0 < 0) iv0 = 0;
    const int iv1 = i < nv - 1 ? i : nv - 1;
    double s = 0.0;
    for (int jv = jv0; jv <= jv1; ++jv) {
        const int jh = j - jv;
        const int jj = packed(jh, jh + jp, ksr);
        for (int iv = iv0; iv <= iv1; ++iv) {
            const int ih = i - iv;
            const int ii = packed(ih, ih + ip, ksl);
            double c;
            if (iv < jv)
                c = lo_l[(int64_t)ii * nv + iv] * hi_r[(int64_t)jj * nv + jv];
            else if (iv > jv)
                c = lo_r[(int64_t)jj * nv + jv] * hi_l[(int64_t)ii * nv + iv];
            else
                c = diag[((int64_t)ii * kkr + jj) * nv + iv];
            s += c;
        }
    }
    return s;
}

__global__ void pair_integrals_kernel(const double *__restrict__ lo_l,
                                      const double *__restrict__ hi_r,
                                      const double *__restrict__ lo_r,
                                      const double *__restrict__ hi_l,
                                      const double *__restrict__ diag, int nv, int ksl, int ksr,
                                      int ns, int mode, double *__restrict__ out) {
    const int W = mode == 0 ? ns : ksl + ksr - 1;
    const int64_t total = (int64_t)ksl * ksr * ns * W;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        int64_t rem = t;
        const int c = static_cast<int>(rem % W);
        rem /= W;
        const int i = static_cast<int>(rem % ns);
        rem /= ns;
        const int jp = static_cast<int>(rem % ksr);
        const int ip = static_cast<int>(rem / ksr);
        int j;
        if (mode == 0) {
            j = c;
        } else {
            j = i + ip - ksl + 1 + c;
            const int last_j = i + ksr - jp - 1;
            if (j < 0 || j >= ns || j > last_j) {
                out[t] = 0.0;
                continue;
            }
        }
        out[t] = pair_integral(lo_l, hi_r, lo_r, hi_l, diag, nv, ksl, ksr, i, ip, j, jp);
    }
}

// ---- global pair moments mom[off][i] (convol.py:1061-1073) ---------------------
__global__ void global_moments_kernel(const double *__restrict__ local, int nv, int ks, int ns,
                                      double *__restrict__ mom) {
    const int64_t total = (int64_t)ks * ns;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int off = static_cast<int>(t / ns);
        const int i = static_cast<int>(t - (int64_t)off * ns);
        double s = 0.0;
        if (i + off < ns) {
            int iv0 = i + off - ks + 1;
            if (iv0 < 0) iv0 = 0;
            const int iv1 = i < nv - 1 ? i : nv - 1;
            for (int iv = iv0; iv <= iv1; ++iv) {
                const int ih = i - iv;
                s += local[(int64_t)packed(ih, ih + off, ks) * nv + iv];
            }
        }
        mom[t] = s;
    }
}

// ---- (a13) densities -------------------------------------------------------------
// banded symmetric density d[i, s] of coefficient rows p1 = A[a], p2 = A[b]
// with zero padding beyond n1 / n2 (convol.py:688-753).
__global__ void density_sym_kernel(const double *__restrict__ A, int64_t lda,
                                   const int32_t *__restrict__ pa, const int32_t *__restrict__ pb,
                                   int np, int n1, int n2, int ns, int ks,
                                   double *__restrict__ out) {
    const int64_t per = (int64_t)ns * ks;
    const int64_t total = per * np;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int p = static_cast<int>(t / per);
        const int64_t e = t - (int64_t)p * per;
        const int i = static_cast<int>(e / ks);
        const int s = static_cast<int>(e - (int64_t)i * ks);
        const double *p1 = A + (int64_t)pa[p] * lda;
        const double *p2 = A + (int64_t)pb[p] * lda;
        const int lim = n1 < ns ? n1 : ns;
        double v = 0.0;
        if (i < lim) {
            if (s == 0) {
                if (i < n2) v = p1[i] * p2[i];
            } else {
                const int ip = i + s;
                if (ip < ns) {
                    double q = 0.0;
                    if (ip < n2) q += p1[i] * p2[ip];
                    if (ip < n1 && i < n2) q += p1[ip] * p2[i];
                    v = q;
                }
            }
        }
        out[t] = v;
    }
}

// out[t][e] = sum_j coef[t][j] dens[j][e] (rows of length L, j ascending):
// work-group row blockIdx.y takes kCombineT targets and each thread one
// column e of them, so the coefficients are work-group uniform (scalar
// loads, and a uniform branch skips a zero one) and each density value is
// loaded once per kCombineT targets, four j ahead of their FMAs; every
// output sums over j in order, skipping zero coefficients (the values of one
// target at a time, bit for bit).
constexpr int kCombineT = 8;
__global__ void combine_rows_kernel(const double *__restrict__ dens, int64_t L, int n_src,
                                    const double *__restrict__ coef, int nt,
                                    double *__restrict__ out) {
    const int t0 = blockIdx.y * kCombineT;
    const int tn = min(kCombineT, nt - t0);
    const double *cr = coef + (int64_t)t0 * n_src;
    for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < L;
         e += (int64_t)gridDim.x * blockDim.x) {
        double s[kCombineT];
#pragma unroll
        for (int u = 0; u < kCombineT; ++u) s[u] = 0.0;
        int j = 0;
        for (; j + 4 <= n_src; j += 4) {
            double d[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) d[q] = dens[(int64_t)(j + q) * L + e];
#pragma unroll
            for (int q = 0; q < 4; ++q)
#pragma unroll
                for (int u = 0; u < kCombineT; ++u) {
                    if (u >= tn) break;
                    const double c = cr[(int64_t)u * n_src + j + q];
                    if (c != 0.0) s[u] += c * d[q];
                }
        }
        for (; j < n_src; ++j) {
            const double d = dens[(int64_t)j * L + e];
#pragma unroll
            for (int u = 0; u < kCombineT; ++u) {
                if (u >= tn) break;
                const double c = cr[(int64_t)u * n_src + j];
                if (c != 0.0) s[u] += c * d;
            }
        }
#pragma unroll
        for (int u = 0; u < kCombineT; ++u)
            if (u < tn) out[(int64_t)(t0 + u) * L + e] = s[u];
    }
}

// exchange density xx[t][i][k] = sum_j (b[t,j] A[j][i]) B[j][k]
// (convol.py:688-709, same operation order).
__global__ void exchange_density_kernel(const double *__restrict__ A, const double *__restrict__ B,
                                        int64_t ld, int n_src, int ns,
                                        const double *__restrict__ coef, int nt,
                                        double *__restrict__ out) {
    const int64_t per = (int64_t)ns * ns;
    const int64_t total = per * nt;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int tt = static_cast<int>(t / per);
        const int64_t e = t - (int64_t)tt * per;
        const int i = static_cast<int>(e / ns);
        const int kx = static_cast<int>(e - (int64_t)i * ns);
        double s = 0.0;
        for (int j = 0; j < n_src; ++j) {
            const double c = coef[(int64_t)tt * n_src + j];
            if (c == 0.0) continue;
            s += (c * A[(int64_t)j * ld + i]) * B[(int64_t)j * ld + kx];
        }
        out[t] = s;
    }
}

// ---- (a14) case 1: a[i, ip] = sum_jp sum_j bankT[ip][jp][i][j] d[j, jp] ------------
constexpr int kMaxBatch = 8;

__global__ __launch_bounds__(kT) void case1_kernel(const double *__restrict__ bank, int ns, int ksl,
                                                   int ksr, const double *__restrict__ d,
                                                   int64_t ld_d, int nb, double *__restrict__ out,
                                                   int64_t ld_o) {
    __shared__ double red[kMaxBatch][kT / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // grid.y: groups of kMaxBatch densities
    const int b0 = blockIdx.y * kMaxBatch;
    d += (int64_t)b0 * ld_d;
    out += (int64_t)b0 * ld_o;
    nb = nb - b0 < kMaxBatch ? nb - b0 : kMaxBatch;
    for (int r = blockIdx.x; r < ns * ksl; r += gridDim.x) {
        const int ip = r / ns, i = r - (r / ns) * ns;
        double acc[kMaxBatch];
#pragma unroll
        for (int b = 0; b < kMaxBatch; ++b) acc[b] = 0.0;
        if (i + ip < ns) {
            for (int jp = 0; jp < ksr; ++jp) {
                const double *row = bank + (((int64_t)ip * ksr + jp) * ns + i) * ns;
                for (int j = threadIdx.x; j < ns - jp; j += kT) {
                    const double bv = row[j];
#pragma unroll
                    for (int b = 0; b < kMaxBatch; ++b)
                        if (b < nb) acc[b] += bv * d[(int64_t)b * ld_d + (int64_t)j * ksr + jp];
                }
            }
        }
#pragma unroll
        for (int b = 0; b < kMaxBatch; ++b) {
            double v = acc[b];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            if (lane == 0) red[b][wave] = v;
        }
        __syncthreads();
        if (threadIdx.x < nb) {
            double v = 0.0;
            for (int q = 0; q < kT / 64; ++q) v += red[threadIdx.x][q];
            out[(int64_t)threadIdx.x * ld_o + (int64_t)i * ksl + ip] = (i + ip < ns) ? v : 0.0;
        }
        __syncthreads();
    }
}

// ---- (a15) case 2: a[j, jp] = sum_ip sum_{i<n_left, i+ip<n_left} bankT d_l[i, ip] --
// One work-group per (jp, tile of 64 j, tile of 32 densities): chunks of 32
// i of one ip are staged in LDS (the bank rows [i][j] and the densities
// [b][i]) by all 256 threads at once, then each thread (j = lane, densities
// 8 w .. 8 w + 7 of its wave w) accumulates its outputs from LDS.  Every
// output is one sequential sum in the reference's (ip, i) order
// (convol.py:852-874, :1310-1345), so results do not depend on the tiling.
constexpr int kC2J = 64, kC2B = 32, kC2I = 32;

__global__ __launch_bounds__(256) void case2_kernel(const double *__restrict__ bank, int ns,
                                                    int ksl, int ksr, int n_left,
                                                    const double *__restrict__ d, int64_t ld_d,
                                                    int nb, double *__restrict__ out,
                                                    int64_t ld_o) {
    __shared__ double sB[kC2I][kC2J];
    __shared__ double sD[kC2B][kC2I + 1];
    const int jp = blockIdx.x;
    const int j0 = blockIdx.y * kC2J;
    const int b0 = blockIdx.z * kC2B;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int j = j0 + lane;
    double acc[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] = 0.0;
    for (int ip = 0; ip < ksl; ++ip) {
        const int n_i = n_left - ip;  // i < n_left - ip
        const double *brow = bank + ((int64_t)ip * ksr + jp) * ns * ns;
        for (int i0 = 0; i0 < n_i; i0 += kC2I) {
            const int cnt = n_i - i0 < kC2I ? n_i - i0 : kC2I;
            __syncthreads();
            for (int e = threadIdx.x; e < kC2I * kC2J; e += 256) {
                const int ii = e / kC2J, jj = e - (e / kC2J) * kC2J;
                sB[ii][jj] = (ii < cnt && j0 + jj < ns) ? brow[(int64_t)(i0 + ii) * ns + j0 + jj]
                                                        : 0.0;
            }
            for (int e = threadIdx.x; e < kC2B * kC2I; e += 256) {
                const int bb = e / kC2I, ii = e - (e / kC2I) * kC2I;
                sD[bb][ii] = (ii < cnt && b0 + bb < nb)
                                 ? d[(int64_t)(b0 + bb) * ld_d + (int64_t)(i0 + ii) * ksl + ip]
                                 : 0.0;
            }
            __syncthreads();
            for (int ii = 0; ii < cnt; ++ii) {
                const double bv = sB[ii][lane];
#pragma unroll
                for (int u = 0; u < 8; ++u) acc[u] += bv * sD[w * 8 + u][ii];
            }
        }
    }
    if (j < ns) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int b = b0 + w * 8 + u;
            if (b < nb) out[(int64_t)b * ld_o + (int64_t)j * ksr + jp] = j + jp < ns ? acc[u] : 0.0;
        }
    }
}

// ---- (a16) case 4 exchange: four gather families in the order of
// _convol_case4_t (convol.py:1000-1058), one thread per (density b, output
// element); `symmetric` computes the lower triangle col <= row and mirrors
// it on store.  With dst_off the result is added to out + dst_off[b] (one
// addition per element, as the accumulation of a separate result would do);
// without, out + b*ld_o is overwritten.
__device__ inline int tri_row(int64_t e) {
    int r = static_cast<int>((sqrt(8.0 * static_cast<double>(e) + 1.0) - 1.0) * 0.5);
    while ((int64_t)(r + 1) * (r + 2) / 2 <= e) ++r;
    while ((int64_t)r * (r + 1) / 2 > e) --r;
    return r;
}

__global__ void case4_bt_kernel(const double *__restrict__ bank, int ns, int ksl, int ksr,
                                const double *__restrict__ d, int64_t ld_d, int nb, int symmetric,
                                double *__restrict__ out, int64_t ld_o,
                                const int64_t *__restrict__ dst_off) {
    const int64_t per = symmetric ? (int64_t)ns * (ns + 1) / 2 : (int64_t)ns * ns;
    const int64_t total = per * nb;
    for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < total;
         u += (int64_t)gridDim.x * blockDim.x) {
        const int b = static_cast<int>(u / per);
        const int64_t e = u - (int64_t)b * per;
        int row, col;
        if (symmetric) {
            row = tri_row(e);
            col = static_cast<int>(e - (int64_t)row * (row + 1) / 2);
        } else {
            row = static_cast<int>(e / ns);
            col = static_cast<int>(e - (int64_t)row * ns);
        }
        const double *db = d + (int64_t)b * ld_d;
#define BANK(ip_, jp_, i_, j_) bank[((((int64_t)(ip_)) * ksr + (jp_)) * ns + (i_)) * ns + (j_)]
        double acc = 0.0;
        for (int ip = 0; ip < ksl; ++ip) {  // family 1
            const int i = row - ip;
            if (i < 0) continue;
            for (int jp = 0; jp < ksr; ++jp) {
                if (col >= ns - jp) continue;
                acc += BANK(ip, jp, i, col) * db[(int64_t)i * ns + col + jp];
            }
        }
        for (int ip = 1; ip < ksl; ++ip) {  // family 2
            const int ii = row + ip;
            if (ii >= ns) continue;
            for (int jp = 0; jp < ksr; ++jp) {
                if (col >= ns - jp) continue;
                acc += BANK(ip, jp, row, col) * db[(int64_t)ii * ns + col + jp];
            }
        }
        for (int jp = 1; jp < ksr; ++jp) {  // family 3
            if (col < jp) continue;
            for (int ip = 0; ip < ksl; ++ip) {
                const int i = row - ip;
                if (i < 0) continue;
                acc += BANK(ip, jp, i, col - jp) * db[(int64_t)i * ns + col - jp];
            }
        }
        for (int ip = 1; ip < ksl; ++ip) {  // family 4
            const int ii = row + ip;
            if (ii >= ns) continue;
            for (int jp = 1; jp < ksr; ++jp) {
                if (col < jp) continue;
                acc += BANK(ip, jp, row, col - jp) * db[(int64_t)ii * ns + col - jp];
            }
        }
#undef BANK
        double *o = dst_off ? out + dst_off[b] : out + (int64_t)b * ld_o;
        if (dst_off) {
            o[(int64_t)row * ns + col] += acc;
            if (symmetric && col < row) o[(int64_t)col * ns + row] += acc;
        } else {
            o[(int64_t)row * ns + col] = acc;
            if (symmetric && col < row) o[(int64_t)col * ns + row] = acc;
        }
    }
}

// banded -> full add into out + dst_off[b], one thread per band entry (lo, s)
__global__ void banded_to_full_at_kernel(double *__restrict__ V, const int64_t *__restrict__ dst_off,
                                         const double *__restrict__ vb, int64_t ldb, int nb,
                                         int ns, int ks, double coeff) {
    const int64_t per = (int64_t)ns * ks;
    const int64_t total = per * nb;
    for (int64_t u = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; u < total;
         u += (int64_t)gridDim.x * blockDim.x) {
        const int b = static_cast<int>(u / per);
        const int64_t e = u - (int64_t)b * per;
        const int lo = static_cast<int>(e / ks), sft = static_cast<int>(e - (int64_t)lo * ks);
        if (lo + sft >= ns) continue;
        const double val = coeff * vb[(int64_t)b * ldb + e];
        double *Vb = V + dst_off[b];
        Vb[(int64_t)lo * ns + lo + sft] += val;
        if (sft > 0) Vb[(int64_t)(lo + sft) * ns + lo] += val;
    }
}

// ---- (a17) semiseparable direct contraction (convol.py:1125-1162) --------------
// pass 1: per (b, jp) sequential prefix/suffix over j into scratch
// One wave per (b, jp) chain: the chain's factor rows and density column are
// staged in LDS by all lanes, then lane 0 runs the prefix and suffix sums
// in the reference's sequential order.
__global__ __launch_bounds__(64) void semi_scan_kernel(const double *__restrict__ rising_r,
                                                       const double *__restrict__ falling_r,
                                                       int ns, int ksr,
                                                       const double *__restrict__ d, int64_t ld_d,
                                                       int nb, double *__restrict__ prefix,
                                                       double *__restrict__ suffix) {
    extern __shared__ double sh[];
    double *r_s = sh, *f_s = sh + ns, *d_s = sh + 2 * ns;
    const int total = nb * ksr;
    for (int t = blockIdx.x; t < total; t += gridDim.x) {
        const int b = t / ksr, jp = t - (t / ksr) * ksr;
        const double *db = d + (int64_t)b * ld_d;
        for (int j = threadIdx.x; j < ns; j += blockDim.x) {
            r_s[j] = rising_r[(int64_t)jp * ns + j];
            f_s[j] = falling_r[(int64_t)jp * ns + j];
            d_s[j] = db[(int64_t)j * ksr + jp];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double *pre = prefix + ((int64_t)b * ksr + jp) * (ns + 1);
            double *suf = suffix + ((int64_t)b * ksr + jp) * (ns + 1);
            const int end = ns - jp;
            double s = 0.0;
            pre[0] = 0.0;
            for (int j = 0; j < end; ++j) {
                s = s + r_s[j] * d_s[j];
                pre[j + 1] = s;
            }
            for (int j = end + 1; j <= ns; ++j) pre[j] = s;
            s = 0.0;
            for (int j = ns; j >= end; --j) suf[j] = 0.0;
            for (int j = end - 1; j >= 0; --j) {
                s = s + f_s[j] * d_s[j];
                suf[j] = s;
            }
        }
        __syncthreads();
    }
}

// pass 2: result[i, ip] per (b, ip, i)
__global__ void semi_case1_kernel(const double *__restrict__ near, const double *__restrict__ rl,
                                  const double *__restrict__ fl, int ns, int ksl, int ksr,
                                  const double *__restrict__ d, int64_t ld_d, int nb,
                                  const double *__restrict__ prefix,
                                  const double *__restrict__ suffix, double *__restrict__ out,
                                  int64_t ld_o) {
    const int W = ksl + ksr - 1;
    const int64_t per = (int64_t)ns * ksl;
    const int64_t total = per * nb;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int b = static_cast<int>(t / per);
        const int64_t e = t - (int64_t)b * per;
        const int i = static_cast<int>(e / ksl);
        const int ip = static_cast<int>(e - (int64_t)i * ksl);
        double value = 0.0;
        if (i < ns - ip) {
            const int first_near = i + ip - ksl + 1;
            const double *db = d + (int64_t)b * ld_d;
            for (int jp = 0; jp < ksr; ++jp) {
                const int end = ns - jp;
                const int fn0 = first_near > 0 ? first_near : 0;
                const int before = end < fn0 ? end : fn0;
                int aft = i + ksr - jp;
                if (aft < 0) aft = 0;
                const int after = end < aft ? end : aft;
                const double *pre = prefix + ((int64_t)b * ksr + jp) * (ns + 1);
                const double *suf = suffix + ((int64_t)b * ksr + jp) * (ns + 1);
                value += fl[(int64_t)ip * ns + i] * pre[before];
                value += rl[(int64_t)ip * ns + i] * suf[after];
                const double *nr = near + (((int64_t)ip * ksr + jp) * ns + i) * W;
                for (int j = fn0; j < after; ++j)
                    value += nr[j - first_near] * db[(int64_t)j * ksr + jp];
            }
        }
        out[(int64_t)b * ld_o + (int64_t)i * ksl + ip] = value;
    }
}

// ---- (a17) semiseparable exchange (convol.py:1165-1303), per output element -----
__device__ inline double semi_term(const double *__restrict__ drow, const double *__restrict__ near,
                                   const double *__restrict__ rl, const double *__restrict__ fl,
                                   const double *__restrict__ rr, const double *__restrict__ fr,
                                   int i, int ip, int jp, int ksl, int ksr, int ns, int j,
                                   int d_shift) {
    const int W = ksl + ksr - 1;
    const int first_near = i + ip - ksl + 1;
    const int last_near = i + ksr - jp - 1;
    if (j < first_near) return fl[(int64_t)ip * ns + i] * rr[(int64_t)jp * ns + j] * drow[j + d_shift];
    if (j <= last_near)
        return near[(((int64_t)ip * ksr + jp) * ns + i) * W + (j - first_near)] * drow[j + d_shift];
    return rl[(int64_t)ip * ns + i] * fr[(int64_t)jp * ns + j] * drow[j + d_shift];
}

__global__ void semi_case4_kernel(const double *__restrict__ near, const double *__restrict__ rl,
                                  const double *__restrict__ fl, const double *__restrict__ rr,
                                  const double *__restrict__ fr, int ns, int ksl, int ksr,
                                  const double *__restrict__ d, int64_t ld_d, int nb,
                                  int symmetric, double *__restrict__ out, int64_t ld_o) {
    const int64_t per = (int64_t)ns * ns;
    const int64_t total = per * nb;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int b = static_cast<int>(t / per);
        const int64_t e = t - (int64_t)b * per;
        const int row = static_cast<int>(e / ns);
        const int col = static_cast<int>(e - (int64_t)row * ns);
        if (symmetric && col > row) continue;
        const int limit = symmetric ? row + 1 : ns;
        const double *db = d + (int64_t)b * ld_d;
        double acc = 0.0;
        // family 1: rows i = row - ip, output shift 0, density shift jp
        for (int ip = 0; ip < ksl; ++ip) {
            const int i = row - ip;
            if (i < 0) continue;
            for (int jp = 0; jp < ksr; ++jp) {
                const int js = (ns - jp) < limit ? (ns - jp) : limit;
                const int j = col;
                if (j < js) acc += semi_term(db + (int64_t)i * ns, near, rl, fl, rr, fr, i, ip, jp, ksl, ksr, ns, j, jp);
            }
        }
        // family 2: density row row+ip, pair row `row`
        for (int ip = 1; ip < ksl; ++ip) {
            const int other = row + ip;
            if (other >= ns) continue;
            for (int jp = 0; jp < ksr; ++jp) {
                const int js = (ns - jp) < limit ? (ns - jp) : limit;
                const int j = col;
                if (j < js) acc += semi_term(db + (int64_t)other * ns, near, rl, fl, rr, fr, row, ip, jp, ksl, ksr, ns, j, jp);
            }
        }
        // family 3: output shift jp
        for (int jp = 1; jp < ksr; ++jp) {
            const int j = col - jp;
            if (j < 0) continue;
            const int js = symmetric ? limit - jp : ns - jp;
            if (j >= js) continue;
            for (int ip = 0; ip < ksl; ++ip) {
                const int i = row - ip;
                if (i < 0) continue;
                acc += semi_term(db + (int64_t)i * ns, near, rl, fl, rr, fr, i, ip, jp, ksl, ksr, ns, j, 0);
            }
        }
        // family 4
        for (int ip = 1; ip < ksl; ++ip) {
            const int other = row + ip;
            if (other >= ns) continue;
            for (int jp = 1; jp < ksr; ++jp) {
                const int j = col - jp;
                if (j < 0) continue;
                const int js = symmetric ? limit - jp : ns - jp;
                if (j >= js) continue;
                acc += semi_term(db + (int64_t)other * ns, near, rl, fl, rr, fr, row, ip, jp, ksl, ksr, ns, j, 0);
            }
        }
        out[(int64_t)b * ld_o + (int64_t)row * ns + col] = acc;
        if (symmetric && col < row) out[(int64_t)b * ld_o + (int64_t)col * ns + row] = acc;
    }
}

// ---- (a18) trace contraction, banded -> full add, Fock finalisation -------------
__global__ void sum_amb_kernel(const double *__restrict__ a, int64_t lda, int sa,
                               const double *__restrict__ b, int64_t ldb, int sb, int nb, int ns,
                               int ks, double *__restrict__ out) {
    // one block per batch element; a[i*sa + ip], b[i*sb + ip]
    __shared__ double red[kT];
    for (int q = blockIdx.x; q < nb; q += gridDim.x) {
        double s = 0.0;
        for (int e = threadIdx.x; e < ns * ks; e += kT) {
            const int ip = e / ns, i = e - (e / ns) * ns;
            if (i + ip < ns)
                s += a[(int64_t)q * lda + (int64_t)i * sa + ip] * b[(int64_t)q * ldb + (int64_t)i * sb + ip];
        }
        red[threadIdx.x] = s;
        __syncthreads();
        for (int off = kT / 2; off > 0; off >>= 1) {
            if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[q] = red[0];
        __syncthreads();
    }
}

// out[q] (+)= sum_{i, ip < ks, i+ip < ns} a_q[i][ip] b_q[i][ip] for banded
// operands addressed through device pointer tables (operands shared by many
// q are stored once).
__global__ void sum_amb_ptr_kernel(const double *const *__restrict__ a,
                                   const double *const *__restrict__ b, int nq, int ns, int ks,
                                   int accumulate, double *__restrict__ out) {
    __shared__ double red[kT];
    for (int q = blockIdx.x; q < nq; q += gridDim.x) {
        const double *x = a[q];
        const double *y = b[q];
        double s = 0.0;
        for (int e = threadIdx.x; e < ns * ks; e += kT) {
            const int ip = e / ns, i = e - (e / ns) * ns;
            if (i + ip < ns) s += x[(int64_t)i * ks + ip] * y[(int64_t)i * ks + ip];
        }
        red[threadIdx.x] = s;
        __syncthreads();
        for (int off = kT / 2; off > 0; off >>= 1) {
            if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
            __syncthreads();
        }
        if (threadIdx.x == 0) out[q] = accumulate ? out[q] + red[0] : red[0];
        __syncthreads();
    }
}

__global__ void banded_to_full_kernel(double *__restrict__ V, int64_t ldv,
                                      const double *__restrict__ vb, int64_t ldb, int nb, int ns,
                                      int ks, double coeff) {
    // V[b][i][i+s] += coeff vb[b][i][s]  and  V[b][i+s][i] += (s > 0)
    const int64_t per = (int64_t)ns * ns;
    const int64_t total = per * nb;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int b = static_cast<int>(t / per);
        const int64_t e = t - (int64_t)b * per;
        const int r = static_cast<int>(e / ns), c = static_cast<int>(e - (int64_t)r * ns);
        const int lo = r < c ? r : c, s = r < c ? c - r : r - c;
        if (s >= ks) continue;
        V[(int64_t)b * ldv + e] += coeff * vb[(int64_t)b * ldb + (int64_t)lo * ks + s];
    }
}

// hfm symmetrisation and the 1/q_i division of methods/dirac_fock.py:1303-1317
__global__ void fock_finalize_kernel(double *__restrict__ pp, double *__restrict__ qq,
                                     double *__restrict__ pq, double *__restrict__ qp, int ns,
                                     double norm, int skip_pq) {
    const int64_t total = (int64_t)ns * ns;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int r = static_cast<int>(t / ns), c = static_cast<int>(t - (int64_t)r * ns);
        if (c < r) continue;  // each thread owns the pair (r, c), (c, r)
        const int64_t a = t, b = (int64_t)c * ns + r;
        const double inv = norm > 1e-15 ? norm : 1.0;
        const double ppv = 0.5 * (pp[a] + pp[b]);
        const double qqv = 0.5 * (qq[a] + qq[b]);
        pp[a] = ppv / inv;
        pp[b] = ppv / inv;
        qq[a] = qqv / inv;
        qq[b] = qqv / inv;
        double pq_rc, pq_cr;
        if (skip_pq) {
            pq_rc = qp[b];
            pq_cr = qp[a];
        } else {
            pq_rc = 0.5 * (pq[a] + qp[b]);
            pq_cr = 0.5 * (pq[b] + qp[a]);
        }
        pq[a] = pq_rc / inv;
        pq[b] = pq_cr / inv;
        qp[a] = pq_cr / inv;
        qp[b] = pq_rc / inv;
    }
}

__global__ void axpy_kernel(double *__restrict__ y, const double *__restrict__ x, int64_t n,
                            double alpha) {
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
         t += (int64_t)gridDim.x * blockDim.x)
        y[t] += alpha * x[t];
}

// y[r][x] = a[r][x] * w[x]
__global__ void mul_rows_kernel(double *__restrict__ y, const double *__restrict__ a,
                                const double *__restrict__ w, int nrows, int64_t n) {
    const int64_t total = (int64_t)nrows * n;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x)
        y[t] = a[t] * w[t % n];
}

// Second stage of the bilinear form (convol.py:1933-1942): from
// T[m][j][jp] (the case-2 contraction of sym(wa, e_m)),
//   M[m][n] = sum_{j<n} T[m][j][n-j] wb[j]  +  sum_{jp} T[m][n][jp] wb[n+jp],
// over j + jp < n_right, in the reference's j-ascending accumulation order.
__global__ void bilinear_stage2_kernel(const double *__restrict__ T, int ns, int ksr, int n_right,
                                       const double *__restrict__ wb, int nm,
                                       double *__restrict__ M) {
    const int64_t total = (int64_t)nm * ns;
    for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
         t += (int64_t)gridDim.x * blockDim.x) {
        const int m = static_cast<int>(t / ns), n = static_cast<int>(t - (int64_t)(t / ns) * ns);
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
        M[t] = s;
    }
}

inline int blocks(int64_t work) { return ax::grid_for(work, kT, 8192); }

}  // namespace

namespace ax {

// Gauss-Legendre nodes on [0, 1] by the Newton iteration of
// numerics/bspline/convol.py:252-291 (the nodes of the diagonal-cell
// sub-quadrature are these, not leggauss's).
void gauss_legendre_01(int n, double *x, double *w) {
    if (n == 1) {
        x[0] = 0.5;
        w[0] = 1.0;
        return;
    }
    for (int i = 0; i < n; ++i) {
        double xi = cos(M_PI * (i + 0.75) / (n + 0.5));
        double dp = 0.0;
        for (int it = 0; it < 50; ++it) {
            double p0 = 1.0, p1 = xi;
            for (int j = 2; j <= n; ++j) {
                const double p2 = ((2 * j - 1) * xi * p1 - (j - 1) * p0) / j;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (xi * p1 - p0) / (xi * xi - 1.0);
            const double dx = p1 / dp;
            xi -= dx;
            if (fabs(dx) < 1e-15) break;
        }
        x[i] = 0.5 * (xi + 1.0);
        w[i] = 1.0 / ((1.0 - xi * xi) * dp * dp);
    }
}

}  // namespace ax

extern "C" {

ax_status ax_gauss_legendre_01(int n, double *x_host, double *w_host) {
    AX_REQUIRE(n >= 1 && n <= 128, "ax_gauss_legendre_01: n must be in [1, 128]");
    ax::gauss_legendre_01(n, x_host, w_host);
    return AX_OK;
}

ax_status ax_cell_moments(const double *bsp1, const double *bsp2, const double *gr,
                          const double *grw, int nv, int ng, int ks1, int ks2, int k, int sym,
                          double *out, void *stream) {
    AX_REQUIRE(nv > 0 && ng > 0 && ks1 > 0 && ks2 > 0, "ax_cell_moments: bad sizes");
    AX_REQUIRE(!sym || ks1 == ks2, "ax_cell_moments: symmetric moments need ks1 == ks2");
    const int kk = sym ? ks1 * (ks1 + 1) / 2 : ks1 * ks2;
    hipLaunchKernelGGL(moments_kernel, dim3(blocks((int64_t)kk * nv)), dim3(kT), 0,
                       ax::as_stream(stream), bsp1, bsp2, gr, grw, nv, ng, ks1, ks2, k, sym, out);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

ax_status ax_cell_triang(const double *bsp_l, const double *bsp_r, const double *gr,
                         const double *grw, const double *t_l, const double *t_r, int nv, int ng,
                         int ksl, int ksr, int k, int n_sub, const double *x01, const double *w01,
                         double *out, void *stream) {
    AX_REQUIRE(nv > 0 && ng > 0 && ksl > 0 && ksr > 0 && ksl <= kMaxKs && ksr <= kMaxKs &&
                   n_sub > 0,
               "ax_cell_triang: bad sizes (ks must be <= %d)", kMaxKs);
    const int kkl = ksl * (ksl + 1) / 2, kkr = ksr * (ksr + 1) / 2;
    const int ksm = ksl > ksr ? ksl : ksr;
    const size_t lds = sizeof(double) * ((size_t)ng * n_sub * ksm + (size_t)ng * n_sub +
                                         2 * (size_t)(kkl + kkr) * ng);
    AX_REQUIRE(lds <= 60 * 1024, "ax_cell_triang: ng*n_sub*ks too large for LDS");
    hipLaunchKernelGGL(triang_kernel, dim3(nv < 4096 ? nv : 4096), dim3(kT), lds,
                       ax::as_stream(stream), bsp_l, bsp_r, gr, grw, t_l, t_r, nv, ng, ksl, ksr,
                       k, n_sub, x01, w01, out);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

ax_status ax_cell_pair_integrals(const double *lo_l, const double *hi_r, const double *lo_r,
                                 const double *hi_l, const double *diag, int nv, int ksl, int ksr,
                                 int ns, int mode, double *out, void *stream) {
    AX_REQUIRE(nv > 0 && ksl > 0 && ksr > 0 && ns > 0 && (mode == 0 || mode == 1),
               "ax_cell_pair_integrals: bad arguments");
    const int W = mode == 0 ? ns : ksl + ksr - 1;
    hipLaunchKernelGGL(pair_integrals_kernel, dim3(blocks((int64_t)ksl * ksr * ns * W)), dim3(kT),
                       0, ax::as_stream(stream), lo_l, hi_r, lo_r, hi_l, diag, nv, ksl, ksr, ns,
                       mode, out);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

ax_status ax_cell_global_moments(const double *local, int nv, int ks, int ns, double *out,
                                 void *stream) {
    AX_REQUIRE(nv > 0 && ks > 0 && ns > 0, "ax_cell_global_moments: bad sizes");
    hipLaunchKernelGGL(global_moments_kernel, dim3(blocks((int64_t)ks * ns)), dim3(kT), 0,
                       ax::as_stream(stream), local, nv, ks, ns, out);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

ax_status ax_density_sym(const double *A, int64_t lda, const int32_t *pa, const int32_t *pb,
                         int np, int n1, int n2, int ns, int ks, double *out, void *stream) {
    AX_REQUIRE(np >= 0 && ns > 0 && ks > 0 && n1 >= 0 && n2 >= 0 && lda >= (n1 > n2 ? n1 : n2),
               "ax_density_sym: bad sizes");
    if (np == 0) return AX_OK;
    hipLaunchKernelGGL(density_sym_kernel, dim3(blocks((int64_t)np * ns * ks)), dim3(kT), 0,
                       ax::as_stream(stream), A, lda, pa, pb, np, n1, n2, ns, ks, out);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

ax_status ax_combine_rows(const double *dens, int64_t L, int n_src, const double *coef, int nt,
                          double *out, void *stream) {
    AX_REQUIRE(L > 0 && n_src >= 0 && nt >= 0, "ax_combine_rows: bad sizes");
    if (nt == 0) return AX_OK;
    const int groups = (nt + kCombineT - 1) / kCombineT;
    AX_REQUIRE(groups <= 65535, "ax_combine_rows: %d targets exceed one launch", nt);
    hipLaunchKernelGGL(combine_rows_kernel, dim3(blocks(L), groups), dim3(kT), 0,
                       ax::as_stream(stream), dens, L, n_src, coef, nt, out);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

ax_status ax_exchange_density(const double *A, const double *B, int64_t ld, int n_src, int ns,
                              const double *coef, int nt, double *out, void *stream) {
    AX_REQUIRE(ns > 0 && ld >= ns && n_src >= 0 && nt >= 0, "ax_exchange_density: bad sizes");
    if (nt == 0) return AX_OK;
    hipLaunchKernelGGL(exchange_density_kernel, dim3(blocks((int64_t)ns * ns * nt)), dim3(kT), 0,
                       ax::as_stream(stream), A, B, ld, n_src, ns, coef, nt, out);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

ax_status ax_convol_case1(const double *bankT, int ns, int ksl, int ksr, const double *d,
                          int64_t ld_d, int nb, double *out, int64_t ld_o, void *stream) {
    AX_REQUIRE(ns > 0 && ksl > 0 && ksr > 0 && nb >= 0 && ld_d >= (int64_t)ns * ksr &&
                   ld_o >= (int64_t)ns * ksl,
               "ax_convol_case1: bad sizes");
    // rows on grid.x, groups of kMaxBatch densities on grid.y (65535 groups per launch)
    constexpr int64_t kPerLaunch = (int64_t)65535 * kMaxBatch;
    for (int64_t b0 = 0; b0 < nb; b0 += kPerLaunch) {
        const int cnt = static_cast<int>(nb - b0 < kPerLaunch ? nb - b0 : kPerLaunch);
        hipLaunchKernelGGL(case1_kernel, dim3(ns * ksl < 8192 ? ns * ksl : 8192,
                                              (cnt + kMaxBatch - 1) / kMaxBatch),
                           dim3(kT), 0, ax::as_stream(stream), bankT, ns, ksl, ksr,
                           d + b0 * ld_d, ld_d, cnt, out + b0 * ld_o, ld_o);
        AX_CHECK_LAUNCH();
    }
    return AX_OK;
}

ax_status ax_convol_case2(const double *bankT, int ns, int ksl, int ksr, int n_left,
                          const double *d, int64_t ld_d, int nb, double *out, int64_t ld_o,
                          void *stream) {
    AX_REQUIRE(ns > 0 && ksl > 0 && ksr > 0 && nb >= 0 && n_left >= 0 && n_left <= ns &&
                   ld_d >= (int64_t)n_left * ksl && ld_o >= (int64_t)ns * ksr,
               "ax_convol_case2: bad sizes");
    // grid (jp, j tiles, density tiles); one launch per 65535 density tiles
    const int jt = (ns + kC2J - 1) / kC2J;
    constexpr int64_t kPerLaunch = (int64_t)65535 * kC2B;
    for (int64_t b0 = 0; b0 < nb; b0 += kPerLaunch) {
        const int cnt = static_cast<int>(nb - b0 < kPerLaunch ? nb - b0 : kPerLaunch);
        hipLaunchKernelGGL(case2_kernel, dim3(ksr, jt, (cnt + kC2B - 1) / kC2B), dim3(256), 0,
                           ax::as_stream(stream), bankT, ns, ksl, ksr, n_left, d + b0 * ld_d,
                           ld_d, cnt, out + b0 * ld_o, ld_o);
        AX_CHECK_LAUNCH();
    }
    return AX_OK;
}

ax_status ax_convol_case4(const double *bankT, int ns, int ksl, int ksr, const double *d,
                          int64_t ld_d, int nb, int symmetric, double *out, int64_t ld_o,
                          void *stream) {
    AX_REQUIRE(ns > 0 && ksl > 0 && ksr > 0 && nb >= 0 && ld_d >= (int64_t)ns * ns &&
                   ld_o >= (int64_t)ns * ns,
               "ax_convol_case4: bad sizes");
    if (nb == 0) return AX_OK;
    const int64_t per = symmetric ? (int64_t)ns * (ns + 1) / 2 : (int64_t)ns * ns;
    hipLaunchKernelGGL(case4_bt_kernel, dim3(blocks(per * nb)), dim3(kT), 0, ax::as_stream(stream),
                       bankT, ns, ksl, ksr, d, ld_d, nb, symmetric, out, ld_o, nullptr);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

ax_status ax_convol_case4_acc(const double *bankT, int ns, int ksl, int ksr, const double *d,
                              int64_t ld_d, int nb, int symmetric, double *out_base,
                              const int64_t *dst_off, void *stream) {
    AX_REQUIRE(ns > 0 && ksl > 0 && ksr > 0 && nb >= 0 && ld_d >= (int64_t)ns * ns,
               "ax_convol_case4_acc: bad sizes");
    if (nb == 0) return AX_OK;
    AX_REQUIRE(out_base && dst_off, "ax_convol_case4_acc: null destination");
    const int64_t per = symmetric ? (int64_t)ns * (ns + 1) / 2 : (int64_t)ns * ns;
    hipLaunchKernelGGL(case4_bt_kernel, dim3(blocks(per * nb)), dim3(kT), 0, ax::as_stream(stream),
                       bankT, ns, ksl, ksr, d, ld_d, nb, symmetric, out_base, (int64_t)0, dst_off);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

ax_status ax_banded_to_full_add_at(double *V_base, const int64_t *dst_off, const double *vb,
                                   int64_t ldb, int nb, int ns, int ks, double coeff,
                                   void *stream) {
    AX_REQUIRE(nb >= 0 && ns > 0 && ks > 0 && ldb >= (int64_t)ns * ks,
               "ax_banded_to_full_add_at: bad sizes");
    if (nb == 0) return AX_OK;
    AX_REQUIRE(V_base && dst_off, "ax_banded_to_full_add_at: null destination");
    hipLaunchKernelGGL(banded_to_full_at_kernel, dim3(blocks((int64_t)nb * ns * ks)), dim3(kT), 0,
                       ax::as_stream(stream), V_base, dst_off, vb, ldb, nb, ns, ks, coeff);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

int64_t ax_convol_semi_workspace(int nb, int ns, int ksr) {
    return (int64_t)2 * nb * ksr * (ns + 1);
}

ax_status ax_convol_case1_semi(const double *near, const double *rl, const double *fl,
                               const double *rr, const double *fr, int ns, int ksl, int ksr,
                               const double *d, int64_t ld_d, int nb, double *workspace,
                               double *out, int64_t ld_o, void *stream) {
    AX_REQUIRE(ns > 0 && ksl > 0 && ksr > 0 && nb >= 0 && ld_d >= (int64_t)ns * ksr &&
                   ld_o >= (int64_t)ns * ksl,
               "ax_convol_case1_semi: bad sizes");
    if (nb == 0) return AX_OK;
    double *prefix = workspace;
    double *suffix = workspace + (int64_t)nb * ksr * (ns + 1);
    AX_REQUIRE((size_t)3 * ns * sizeof(double) <= 64 * 1024,
               "ax_convol_case1_semi: ns=%d exceeds the LDS staging of the scan", ns);
    hipLaunchKernelGGL(semi_scan_kernel, dim3(nb * ksr < 65535 ? nb * ksr : 65535), dim3(64),
                       (size_t)3 * ns * sizeof(double), ax::as_stream(stream), rr, fr, ns, ksr, d,
                       ld_d, nb, prefix, suffix);
    AX_CHECK_LAUNCH();
    hipLaunchKernelGGL(semi_case1_kernel, dim3(blocks((int64_t)nb * ns * ksl)), dim3(kT), 0,
                       ax::as_stream(stream), near, rl, fl, ns, ksl, ksr, d, ld_d, nb, prefix,
                       suffix, out, ld_o);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

ax_status ax_convol_case4_semi(const double *near, const double *rl, const double *fl,
                               const double *rr, const double *fr, int ns, int ksl, int ksr,
                               const double *d, int64_t ld_d, int nb, int symmetric, double *out,
                               int64_t ld_o, void *stream) {
    AX_REQUIRE(ns > 0 && ksl > 0 && ksr > 0 && nb >= 0 && ld_d >= (int64_t)ns * ns &&
                   ld_o >= (int64_t)ns * ns,
               "ax_convol_case4_semi: bad sizes");
    if (nb == 0) return AX_OK;
    hipLaunchKernelGGL(semi_case4_kernel, dim3(blocks((int64_t)nb * ns * ns)), dim3(kT), 0,
                       ax::as_stream(stream), near, rl, fl, rr, fr, ns, ksl, ksr, d, ld_d, nb,
                       symmetric, out, ld_o);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

ax_status ax_sum_amb_banded(const double *a, int64_t lda, int sa, const double *b, int64_t ldb,
                            int sb, int nb, int ns, int ks, double *out, void *stream) {
    AX_REQUIRE(nb >= 0 && ns > 0 && ks > 0 && sa >= ks && sb >= ks, "ax_sum_amb_banded: bad sizes");
    if (nb == 0) return AX_OK;
    hipLaunchKernelGGL(sum_amb_kernel, dim3(nb < 4096 ? nb : 4096), dim3(kT), 0,
                       ax::as_stream(stream), a, lda, sa, b, ldb, sb, nb, ns, ks, out);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

ax_status ax_sum_amb_ptr(const double *const *a, const double *const *b, int nq, int ns,
                         int ks, int accumulate, double *out, void *stream) {
    AX_REQUIRE(nq >= 0 && ns > 0 && ks > 0, "ax_sum_amb_ptr: bad sizes");
    if (nq == 0) return AX_OK;
    AX_REQUIRE(a && b && out, "ax_sum_amb_ptr: null pointer");
    hipLaunchKernelGGL(sum_amb_ptr_kernel, dim3(nq < 4096 ? nq : 4096), dim3(kT), 0,
                       ax::as_stream(stream), a, b, nq, ns, ks, accumulate, out);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

ax_status ax_banded_to_full_add(double *V, int64_t ldv, const double *vb, int64_t ldb, int nb,
                                int ns, int ks, double coeff, void *stream) {
    AX_REQUIRE(nb >= 0 && ns > 0 && ks > 0 && ldv >= (int64_t)ns * ns && ldb >= (int64_t)ns * ks,
               "ax_banded_to_full_add: bad sizes");
    if (nb == 0) return AX_OK;
    hipLaunchKernelGGL(banded_to_full_kernel, dim3(blocks((int64_t)nb * ns * ns)), dim3(kT), 0,
                       ax::as_stream(stream), V, ldv, vb, ldb, nb, ns, ks, coeff);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

ax_status ax_axpy(double *y, const double *x, int64_t n, double alpha, void *stream) {
    AX_REQUIRE(n >= 0, "ax_axpy: bad size");
    if (n == 0) return AX_OK;
    hipLaunchKernelGGL(axpy_kernel, dim3(blocks(n)), dim3(kT), 0, ax::as_stream(stream), y, x, n,
                       alpha);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

ax_status ax_mul_rows(double *y, const double *a, const double *w, int nrows, int64_t n,
                      void *stream) {
    AX_REQUIRE(nrows >= 0 && n > 0, "ax_mul_rows: bad sizes");
    if (nrows == 0) return AX_OK;
    hipLaunchKernelGGL(mul_rows_kernel, dim3(blocks((int64_t)nrows * n)), dim3(kT), 0,
                       ax::as_stream(stream), y, a, w, nrows, n);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

ax_status ax_bilinear_stage2(const double *T, int ns, int ksr, int n_right, const double *wb,
                             int nm, double *M, void *stream) {
    AX_REQUIRE(ns > 0 && ksr > 0 && n_right >= 0 && n_right <= ns && nm >= 0,
               "ax_bilinear_stage2: bad sizes");
    if (nm == 0) return AX_OK;
    hipLaunchKernelGGL(bilinear_stage2_kernel, dim3(blocks((int64_t)nm * ns)), dim3(kT), 0,
                       ax::as_stream(stream), T, ns, ksr, n_right, wb, nm, M);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

ax_status ax_fock_finalize(double *pp, double *qq, double *pq, double *qp, int ns, double norm,
                           int skip_pq, void *stream) {
    AX_REQUIRE(ns > 0, "ax_fock_finalize: bad size");
    hipLaunchKernelGGL(fock_finalize_kernel, dim3(blocks((int64_t)ns * ns)), dim3(kT), 0,
                       ax::as_stream(stream), pp, qq, pq, qp, ns, norm, skip_pq);
    AX_CHECK_LAUNCH();
    return AX_OK;
}

}  // extern "C"
