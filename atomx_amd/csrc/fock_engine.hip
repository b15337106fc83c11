// Batched Dirac-Fock two-electron operator of many atoms or ions (the
// per-iteration Fock build of methods/dirac_fock.py:1128-1319,
// _build_hf_matrix_dbsr, for every target of every ion) as five launches
// driven by job tables built once per plan:
//
//   A  direct densities: per (ion, block, k) the banded ket densities of all
//      targets, sum_j a(t, j, k) d_j with the target's own term doubled
//      (convol.py:646-685);
//   B  semiseparable chains: prefix / suffix sums of rising and falling
//      right factors times those densities (convol.py:1137-1146), one wave
//      per (target, band offset);
//   C  direct contractions: bank rows against the densities (case 1,
//      convol.py:971-997) or the semiseparable form (convol.py:1125-1162),
//      giving the banded potentials v (ns, ks_bra) of every target;
//   D  assembly: one work-group per (ion, quadrant, 16 x 16 output tile,
//      group of up to 8 targets) runs the quadrant's contributions in the
//      reference's order (block outer, k inner, direct before exchange):
//      direct band values, and exchange (case 4, convol.py:1000-1058) with
//      the exchange densities sum_j b(t, j, k) A_j B_j^T of the group formed
//      in LDS from the coefficient stacks and the bank read straight from
//      HBM, each bank window once for every target of the group;
//   E  symmetrisation and division by the target's occupation divisor
//      (dirac_fock.py:1303-1317).
//
// Layouts: coefficient stacks P, Q (n_orb, ns) FP64 (zero past nsp / nsq);
// banks bankT[ip][jp][i][j] (ksl, ksr, ns, ns); output per ion (n_targets,
// 4, ns, ns) in the order PP, QQ, PQ, QP.  Symmetric quadrants (PP, QQ) are
// computed on their lower triangle and mirrored, as the reference's
// symmetric_output.  A run is five kernel launches and nothing else, so it
// can be captured in a HIP graph.
#include <algorithm>
#include <cstring>
#include <vector>

#include "ax_common.h"

namespace {

constexpr int kMaxG = 8;     // targets per assembly work item (register accumulators)
constexpr int kTile = 16;    // assembly output tile edge
constexpr int kMaxKs = 12;   // largest B-spline order the engine takes
constexpr int kThreads = 256;

struct DdJob {               // A: direct densities of one (ion, block, k)
    const double *src;       // ket-component coefficient stack (n_orb, ns)
    const double *coef;      // (nt, n_orb), the target's own term doubled
    double *dd;              // (nt, ns, ks)
    int ns, ks, n_orb, nt;
};

struct ChainJob {            // B: chains of one semiseparable (ion, block, k)
    const double *rr, *fr;   // right rising / falling factors (ksr, ns)
    const double *dd;        // (nt, ns, ksr)
    double *pre, *suf;       // (nt, ksr, ns + 1)
    int ns, ksr, nt, pad;
};

struct DirJob {              // C: direct contraction of one (ion, block, k)
    const double *bank;      // mode 0: bankT (ksl, ksr, ns, ns)
    const double *nr, *rl, *fl;  // mode 1: near (ksl, ksr, ns, ksl+ksr-1), left factors (ksl, ns)
    const double *pre, *suf; // mode 1: chains of B
    const double *dd;        // (nt, ns, ksr)
    double *v;               // (nt, ns, ksl)
    int mode, ns, ksl, ksr, nt, pad;
};

struct AsmEntry {            // D: one contribution to a quadrant
    const double *v;         // kind 0: direct band potentials (nt, ns, ksl)
    const double *bank;      // kind 1: bankT (ksl, ksr, ns, ns)
    const double *coef;      // kind 1: exchange coefficients (nt, n_orb)
    const double *A, *B;     // kind 1: bra / ket coefficient stacks (n_orb, ns)
    int kind, ksl, ksr, pad;
};

struct AsmJob {              // D: one output tile of one quadrant for a target group
    double *H;               // the ion's output (nt, 4, ns, ns)
    int ns, n_orb, quadrant, symmetric;
    int r0, c0, t0, tg;      // tile origin (elements), targets [t0, t0 + tg)
    int e0, e1, pad0, pad1;  // entries [e0, e1)
};

struct FinJob {              // E: one ion
    double *H;
    const double *divisor;   // (nt)
    int ns, nt, skip, pad;
};

// ---- A: direct densities ---------------------------------------------------------
__global__ __launch_bounds__(kThreads) void fock_dd_kernel(const DdJob *__restrict__ jobs) {
    const DdJob J = jobs[blockIdx.y];
    const int64_t per = (int64_t)J.ns * J.ks;
    const int64_t total = per * J.nt;
    for (int64_t e = blockIdx.x * (int64_t)kThreads + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * kThreads) {
        const int t = static_cast<int>(e / per);
        const int64_t rem = e - (int64_t)t * per;
        const int i = static_cast<int>(rem / J.ks), s = static_cast<int>(rem - (rem / J.ks) * J.ks);
        double acc = 0.0;
        if (i + s < J.ns) {
            const double *c = J.coef + (int64_t)t * J.n_orb;
            for (int j = 0; j < J.n_orb; ++j) {
                const double cj = c[j];
                if (cj == 0.0) continue;
                const double *p = J.src + (int64_t)j * J.ns;
                const double pi = p[i];
                if (s == 0) {
                    acc += cj * pi * pi;
                } else {
                    const double ps = p[i + s];
                    acc += cj * (pi * ps + ps * pi);
                }
            }
        }
        J.dd[e] = acc;
    }
}

// ---- B: semiseparable chains, one wave per (target, band offset) ------------------
// pre[j] = sum_{j' < j} rr[jp][j'] d[j'][jp] and suf[j] = sum_{j' >= j}
// fr[jp][j'] d[j'][jp], over j' < ns - jp (pre[0] = suf[ns] = 0).  Each lane
// sums a contiguous segment; the wave scans the segment sums.
__global__ __launch_bounds__(64) void fock_chain_kernel(const ChainJob *__restrict__ jobs) {
    const ChainJob J = jobs[blockIdx.y];
    const int chain = blockIdx.x;
    if (chain >= J.nt * J.ksr) return;  // uniform over the work-group
    const int t = chain / J.ksr, jp = chain - (chain / J.ksr) * J.ksr;
    const int lane = threadIdx.x;
    const int ns = J.ns, end = ns - jp;
    const double *rr = J.rr + (int64_t)jp * ns;
    const double *fr = J.fr + (int64_t)jp * ns;
    const double *d = J.dd + (int64_t)t * ns * J.ksr + jp;
    double *pre = J.pre + (int64_t)chain * (ns + 1);
    double *suf = J.suf + (int64_t)chain * (ns + 1);
    const int seg = (ns + 63) / 64;
    const int lo = lane * seg, hi = min(lo + seg, ns);
    double so = 0.0, si = 0.0;
    for (int j = lo; j < hi; ++j) {
        if (j < end) {
            const double dj = d[(int64_t)j * J.ksr];
            so += rr[j] * dj;
            si += fr[j] * dj;
        }
    }
    double incl_o = so, incl_i = si;  // inclusive scans: left-to-right and right-to-left
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const double a = __shfl_up(incl_o, off, 64);
        const double b = __shfl_down(incl_i, off, 64);
        if (lane >= off) incl_o += a;
        if (lane + off < 64) incl_i += b;
    }
    const double left = __shfl_up(incl_o, 1, 64), right = __shfl_down(incl_i, 1, 64);
    double run = lane == 0 ? 0.0 : left;  // sum of the segments to the left
    if (lane == 0) pre[0] = 0.0;
    for (int j = lo; j < hi; ++j) {
        if (j < end) run += rr[j] * d[(int64_t)j * J.ksr];
        pre[j + 1] = run;
    }
    run = lane == 63 ? 0.0 : right;  // sum of the segments to the right
    for (int j = hi - 1; j >= lo; --j) {
        if (j < end) run += fr[j] * d[(int64_t)j * J.ksr];
        suf[j] = run;
    }
    if (lane == 63) suf[ns] = 0.0;
}

// ---- C: direct contractions ---------------------------------------------------------
__global__ __launch_bounds__(kThreads) void fock_direct_kernel(const DirJob *__restrict__ jobs) {
    const DirJob J = jobs[blockIdx.y];
    const int ns = J.ns, ksl = J.ksl, ksr = J.ksr;
    if (J.mode == 0) {
        // bank rows (ip, i), one wave each, lanes across j; targets in groups of 8
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        for (int r = blockIdx.x * 4 + wave; r < ns * ksl; r += gridDim.x * 4) {
            const int ip = r / ns, i = r - (r / ns) * ns;
            for (int tb = 0; tb < J.nt; tb += kMaxG) {
                double acc[kMaxG];
#pragma unroll
                for (int g = 0; g < kMaxG; ++g) acc[g] = 0.0;
                if (i + ip < ns) {
                    for (int jp = 0; jp < ksr; ++jp) {
                        const double *row = J.bank + (((int64_t)ip * ksr + jp) * ns + i) * ns;
                        for (int j = lane; j < ns - jp; j += 64) {
                            const double bv = row[j];
#pragma unroll
                            for (int g = 0; g < kMaxG; ++g)
                                if (tb + g < J.nt)
                                    acc[g] += bv * J.dd[((int64_t)(tb + g) * ns + j) * ksr + jp];
                        }
                    }
                }
#pragma unroll
                for (int g = 0; g < kMaxG; ++g) {
                    double x = acc[g];
                    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
                    if (lane == 0 && tb + g < J.nt)
                        J.v[((int64_t)(tb + g) * ns + i) * ksl + ip] = (i + ip < ns) ? x : 0.0;
                }
            }
        }
        return;
    }
    // semiseparable: one thread per (target, i, ip)
    const int W = ksl + ksr - 1;
    const int64_t per = (int64_t)ns * ksl;
    const int64_t total = per * J.nt;
    for (int64_t e = blockIdx.x * (int64_t)kThreads + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * kThreads) {
        const int t = static_cast<int>(e / per);
        const int64_t rem = e - (int64_t)t * per;
        const int i = static_cast<int>(rem / ksl), ip = static_cast<int>(rem - (rem / ksl) * ksl);
        double value = 0.0;
        if (i < ns - ip) {
            const int first_near = i + ip - ksl + 1;
            const double *d = J.dd + (int64_t)t * ns * ksr;
            for (int jp = 0; jp < ksr; ++jp) {
                const int end = ns - jp;
                const int fn0 = first_near > 0 ? first_near : 0;
                const int before = end < fn0 ? end : fn0;
                const int aft = i + ksr - jp > 0 ? i + ksr - jp : 0;
                const int after = end < aft ? end : aft;
                const double *pre = J.pre + ((int64_t)t * ksr + jp) * (ns + 1);
                const double *suf = J.suf + ((int64_t)t * ksr + jp) * (ns + 1);
                value += J.fl[(int64_t)ip * ns + i] * pre[before];
                value += J.rl[(int64_t)ip * ns + i] * suf[after];
                const double *nr = J.nr + (((int64_t)ip * ksr + jp) * ns + i) * W;
                for (int j = fn0; j < after; ++j) value += nr[j - first_near] * d[(int64_t)j * ksr + jp];
            }
        }
        J.v[((int64_t)t * ns + i) * ksl + ip] = value;
    }
}

// ---- D: assembly of the quadrants ----------------------------------------------------
// LDS (doubles): the exchange densities of the target group on the tile plus
// a halo of ks - 1 rows / columns on each side (xx[g][lr][lc], planes of
// xx_plane), then the rows of the bra / ket coefficient stacks over the
// halo (sa[s][lr], sb[s][lc]) and the group's coefficients (sc[g][s]).
struct AsmLds {
    int xx_plane, n_orb_max, xr_max, xc_max, g_max, pad;
};

__device__ inline void bank_values(const double *__restrict__ bank, int ns, int ksr, int q,
                                   int row, int col, double &b1, double &b2, double &b3,
                                   double &b4) {
    const int ip = q / ksr, jp = q - (q / ksr) * ksr;
    const double *base = bank + ((int64_t)ip * ksr + jp) * ns * ns;
    const bool up = row >= ip, down = ip >= 1 && row + ip < ns;
    const bool right = col + jp < ns, left = jp >= 1 && col >= jp;
    // the four index families of the exchange contraction
    b1 = (up && right) ? base[(int64_t)(row - ip) * ns + col] : 0.0;
    b2 = (down && right) ? base[(int64_t)row * ns + col] : 0.0;
    b3 = (up && left) ? base[(int64_t)(row - ip) * ns + col - jp] : 0.0;
    b4 = (down && left) ? base[(int64_t)row * ns + col - jp] : 0.0;
}

__global__ __launch_bounds__(kThreads) void fock_assemble_kernel(
    const AsmJob *__restrict__ jobs, const AsmEntry *__restrict__ entries, const AsmLds geo) {
    extern __shared__ double lds[];
    const AsmJob J = jobs[blockIdx.x];
    const int ns = J.ns, n_orb = J.n_orb, tg = J.tg, t0 = J.t0;
    double *sa = lds + geo.g_max * geo.xx_plane;
    double *sb = sa + geo.n_orb_max * geo.xr_max;
    double *sc = sb + geo.n_orb_max * geo.xc_max;
    const int tid = threadIdx.x;
    const int row = J.r0 + tid / kTile, col = J.c0 + tid % kTile;
    const bool active = row < ns && col < ns && (!J.symmetric || col <= row);
    double acc[kMaxG];
#pragma unroll
    for (int g = 0; g < kMaxG; ++g) acc[g] = 0.0;

    for (int e = J.e0; e < J.e1; ++e) {
        const AsmEntry E = entries[e];
        if (E.kind == 0) {  // direct band value of v[lo][row - col] (lower triangle)
            const int s = row - col;
            if (active && s >= 0 && s < E.ksl) {
#pragma unroll
                for (int g = 0; g < kMaxG; ++g)
                    if (g < tg) acc[g] += E.v[((int64_t)(t0 + g) * ns + col) * E.ksl + s];
            }
            continue;
        }
        const int ksl = E.ksl, ksr = E.ksr;
        const int xr = kTile + 2 * (ksl - 1), xc = kTile + 2 * (ksr - 1);
        const int plane = xr * xc;
        const int rb = J.r0 - (ksl - 1), cb = J.c0 - (ksr - 1);
        __syncthreads();  // the previous entry's LDS reads are done
        for (int idx = tid; idx < n_orb * xr; idx += kThreads) {
            const int sx = idx / xr, i = idx - (idx / xr) * xr;
            const int gi = rb + i;
            sa[sx * geo.xr_max + i] = (gi >= 0 && gi < ns) ? E.A[(int64_t)sx * ns + gi] : 0.0;
        }
        for (int idx = tid; idx < n_orb * xc; idx += kThreads) {
            const int sx = idx / xc, j = idx - (idx / xc) * xc;
            const int gj = cb + j;
            sb[sx * geo.xc_max + j] = (gj >= 0 && gj < ns) ? E.B[(int64_t)sx * ns + gj] : 0.0;
        }
        for (int idx = tid; idx < tg * n_orb; idx += kThreads) {
            const int g = idx / n_orb, sx = idx - (idx / n_orb) * n_orb;
            sc[g * geo.n_orb_max + sx] = E.coef[(int64_t)(t0 + g) * n_orb + sx];
        }
        __syncthreads();
        // exchange densities of the group: sum_s (b(t, s) A_s[i]) B_s[j]
        for (int idx = tid; idx < tg * plane; idx += kThreads) {
            const int g = idx / plane, rem = idx - (idx / plane) * plane;
            const int i = rem / xc, jx = rem - (rem / xc) * xc;
            double val = 0.0;
            for (int sx = 0; sx < n_orb; ++sx) {
                const double cs = sc[g * geo.n_orb_max + sx];
                if (cs == 0.0) continue;
                val += (cs * sa[sx * geo.xr_max + i]) * sb[sx * geo.xc_max + jx];
            }
            lds[g * geo.xx_plane + i * xc + jx] = val;
        }
        __syncthreads();
        if (!active) continue;
        double part[kMaxG];
#pragma unroll
        for (int g = 0; g < kMaxG; ++g) part[g] = 0.0;
        const int lr = row - J.r0 + (ksl - 1), lc = col - J.c0 + (ksr - 1);
        const int nq = ksl * ksr;
        double b1, b2, b3, b4;
        bank_values(E.bank, ns, ksr, 0, row, col, b1, b2, b3, b4);
        for (int q = 0; q < nq; ++q) {
            // the next band pair's bank values are in flight during this one's FMAs
            double n1 = 0.0, n2 = 0.0, n3 = 0.0, n4 = 0.0;
            if (q + 1 < nq) bank_values(E.bank, ns, ksr, q + 1, row, col, n1, n2, n3, n4);
            const int ip = q / ksr, jp = q - (q / ksr) * ksr;
            const int o1 = (lr - ip) * xc + lc + jp, o2 = (lr + ip) * xc + lc + jp;
            const int o3 = (lr - ip) * xc + lc - jp, o4 = (lr + ip) * xc + lc - jp;
#pragma unroll
            for (int g = 0; g < kMaxG; ++g) {
                if (g < tg) {
                    const double *x = lds + g * geo.xx_plane;
                    part[g] += b1 * x[o1] + b2 * x[o2] + b3 * x[o3] + b4 * x[o4];
                }
            }
            b1 = n1;
            b2 = n2;
            b3 = n3;
            b4 = n4;
        }
#pragma unroll
        for (int g = 0; g < kMaxG; ++g) acc[g] += part[g];
    }
    if (active) {
#pragma unroll
        for (int g = 0; g < kMaxG; ++g)
            if (g < tg)
                J.H[(((int64_t)(t0 + g) * 4 + J.quadrant) * ns + row) * ns + col] = acc[g];
    }
}

// ---- E: symmetrisation and occupation divisor --------------------------------------------
__global__ __launch_bounds__(kThreads) void fock_finalize_kernel(const FinJob *__restrict__ jobs) {
    const FinJob J = jobs[blockIdx.y];
    const int ns = J.ns;
    const int64_t per = (int64_t)ns * ns;
    const int64_t total = per * J.nt;
    for (int64_t e = blockIdx.x * (int64_t)kThreads + threadIdx.x; e < total;
         e += (int64_t)gridDim.x * kThreads) {
        const int t = static_cast<int>(e / per);
        const int64_t rem = e - (int64_t)t * per;
        const int r = static_cast<int>(rem / ns), c = static_cast<int>(rem - (rem / ns) * ns);
        const double norm = J.divisor[t];
        const bool divide = norm > 1e-15;
        double *pp = J.H + ((int64_t)t * 4 + 0) * per;
        double *qq = J.H + ((int64_t)t * 4 + 1) * per;
        double *pq = J.H + ((int64_t)t * 4 + 2) * per;
        double *qp = J.H + ((int64_t)t * 4 + 3) * per;
        if (r >= c) {  // symmetric quadrants: the lower triangle holds the values
            double a = pp[(int64_t)r * ns + c], b = qq[(int64_t)r * ns + c];
            if (divide) {
                a /= norm;
                b /= norm;
            }
            pp[(int64_t)r * ns + c] = a;
            pp[(int64_t)c * ns + r] = a;
            qq[(int64_t)r * ns + c] = b;
            qq[(int64_t)c * ns + r] = b;
        }
        // PQ(r, c) and QP(c, r) belong to this thread alone
        const double x = qp[(int64_t)c * ns + r];
        double f = J.skip ? x : 0.5 * (pq[(int64_t)r * ns + c] + x);
        if (divide) f /= norm;
        pq[(int64_t)r * ns + c] = f;
        qp[(int64_t)c * ns + r] = f;
    }
}

size_t align_up(size_t x) { return (x + 255) & ~static_cast<size_t>(255); }

}  // namespace

struct ax_fock_plan {
    void *block = nullptr;   // one device block: tables, coefficients, work buffers
    const DdJob *dd = nullptr;
    const ChainJob *chains = nullptr;
    const DirJob *dir = nullptr;
    const AsmJob *asmj = nullptr;
    const AsmEntry *entries = nullptr;
    const FinJob *fin = nullptr;
    int n_dd = 0, n_chains = 0, n_dir = 0, n_asm = 0, n_fin = 0;
    int max_chain = 0;
    size_t asm_lds = 0;
    AsmLds geo{};
    int dir_blocks = 1;
    int64_t bytes = 0;
    int64_t rk_equivalents = 0;
};

extern "C" {

ax_status ax_fock_plan_create(int n_ions, const int32_t *ion_meta_host,
                              const int32_t *targets_host, const double *direct_host,
                              const double *exchange_host, const double *divisor_host,
                              const double *const *P, const double *const *Q, double *const *H,
                              const double *const *banks_host,
                              const double *const *factors_host, int semi_direct_ns,
                              int max_group, ax_fock_plan **plan_out) {
    AX_REQUIRE(n_ions >= 1 && ion_meta_host && targets_host && direct_host && exchange_host &&
                   divisor_host && P && Q && H && banks_host && factors_host && plan_out,
               "ax_fock_plan_create: missing argument");
    *plan_out = nullptr;
    if (max_group <= 0) max_group = kMaxG;
    AX_REQUIRE(max_group <= kMaxG, "ax_fock_plan_create: target groups of at most %d, got %d",
               kMaxG, max_group);
    // host images of every table, with offsets into one device block
    struct Buf {
        size_t off, bytes;
    };
    size_t cursor = 0;
    auto reserve = [&](size_t bytes) {
        Buf b{cursor, bytes};
        cursor += align_up(bytes > 0 ? bytes : 1);
        return b;
    };
    // Host images of every table.  Coefficient pointers hold an index into
    // coef_img and work-buffer pointers the byte offset + 1 inside the work
    // area until the device block exists; both are fixed up before upload.
    std::vector<double> coef_img;
    std::vector<DdJob> dd;
    std::vector<ChainJob> chains;
    std::vector<DirJob> dir;
    std::vector<AsmEntry> entries;
    std::vector<AsmJob> asmj;
    std::vector<FinJob> fin;

    const int32_t *meta = ion_meta_host;
    const int32_t *tptr = targets_host;
    const double *dptr = direct_host, *xptr = exchange_host, *nptr = divisor_host;
    const double *const *bptr = banks_host;
    const double *const *fptr = factors_host;
    int64_t rk_eq = 0;
    int xr_max = 1, xc_max = 1, n_orb_max = 1, g_max = 1;
    // a device-pointer placeholder: the byte offset + 1 (0 stays "none")
    auto ph = [](size_t off) { return reinterpret_cast<double *>(off + 1); };

    for (int ion = 0; ion < n_ions; ++ion) {
        const int ns = meta[0], ksp = meta[1], ksq = meta[2], n_orb = meta[3], nt = meta[4],
                  kmax = meta[5], skip = meta[6];
        meta += 7;
        AX_REQUIRE(ns >= 1 && n_orb >= 1 && nt >= 0 && kmax >= 0, "ax_fock_plan_create: ion %d "
                   "has bad sizes (ns %d, n_orb %d, targets %d, kmax %d)", ion, ns, n_orb, nt, kmax);
        AX_REQUIRE(ksp >= 1 && ksq >= 1 && ksp <= kMaxKs && ksq <= kMaxKs,
                   "ax_fock_plan_create: ion %d: B-spline orders %d / %d outside 1..%d", ion,
                   ksp, ksq, kMaxKs);
        AX_REQUIRE(ns < 320, "ax_fock_plan_create: ion %d: ns = %d takes the semiseparable "
                   "exchange (ns >= 320), which the batched engine does not run", ion, ns);
        AX_REQUIRE(P[ion] && Q[ion] && (nt == 0 || H[ion]),
                   "ax_fock_plan_create: ion %d: missing device buffers", ion);
        const int32_t *targets = tptr;
        tptr += nt;
        for (int t = 0; t < nt; ++t)
            AX_REQUIRE(targets[t] >= 0 && targets[t] < n_orb,
                       "ax_fock_plan_create: ion %d: target %d out of range", ion, targets[t]);
        const int K = kmax + 1;
        const double *dcoef = dptr, *xcoef = xptr;  // (nt, n_orb, K)
        dptr += (size_t)nt * n_orb * K;
        xptr += (size_t)nt * n_orb * K;
        const double *divisor = nptr;
        nptr += nt;
        const double *const *banks = bptr;      // (4, K)
        const double *const *factors = fptr;    // (4, K, 5)
        bptr += 4 * K;
        fptr += 4 * K * 5;
        if (nt == 0) continue;

        const bool semi = ns >= semi_direct_ns;
        // quadrant entry lists per target group: PP, QQ, PQ, QP
        // targets in balanced groups of at most max_group
        const int n_groups = (nt + max_group - 1) / max_group;
        const int g_size = (nt + n_groups - 1) / n_groups;
        std::vector<std::vector<std::vector<AsmEntry>>> qe(
            4, std::vector<std::vector<AsmEntry>>(n_groups));
        const int blocks[4][2] = {{0, 0}, {0, 1}, {1, 0}, {1, 1}};  // (bra, ket): 0 P, 1 Q
        for (int b = 0; b < 4; ++b) {
            const int ii = blocks[b][0], jj = blocks[b][1];
            const int ks_bra = ii ? ksq : ksp, ks_ket = jj ? ksq : ksp;
            const double *ket = jj ? Q[ion] : P[ion];
            const double *bra = ii ? Q[ion] : P[ion];
            const int exch_q = b == 0 ? 0 : b == 1 ? 2 : b == 2 ? 3 : 1;
            const int dir_q = ii ? 1 : 0;
            for (int k = 0; k < K; ++k) {
                bool any_d = false, any_x = false;
                std::vector<bool> gd(n_groups, false), gx(n_groups, false);
                for (int t = 0; t < nt; ++t)
                    for (int j = 0; j < n_orb; ++j) {
                        const double a = dcoef[((size_t)t * n_orb + j) * K + k];
                        const double x = xcoef[((size_t)t * n_orb + j) * K + k];
                        if (a != 0.0) {
                            any_d = gd[t / g_size] = true;
                            ++rk_eq;
                        }
                        if (x != 0.0 && !(skip && b == 1)) {
                            any_x = gx[t / g_size] = true;
                            ++rk_eq;
                        }
                    }
                if (any_d) {
                    // direct densities with the target's own term doubled
                    const size_t c0 = coef_img.size();
                    for (int t = 0; t < nt; ++t)
                        for (int j = 0; j < n_orb; ++j) {
                            double a = dcoef[((size_t)t * n_orb + j) * K + k];
                            if (j == targets[t]) a = a + a;
                            coef_img.push_back(a);
                        }
                    const Buf dd_buf = reserve((size_t)nt * ns * ks_ket * 8);
                    const Buf v_buf = reserve((size_t)nt * ns * ks_bra * 8);
                    DdJob A{};
                    A.src = ket;
                    A.coef = reinterpret_cast<const double *>(c0);  // coefficient index, fixed below
                    A.dd = ph(dd_buf.off);
                    A.ns = ns;
                    A.ks = ks_ket;
                    A.n_orb = n_orb;
                    A.nt = nt;
                    dd.push_back(A);
                    DirJob C{};
                    C.dd = ph(dd_buf.off);
                    C.v = ph(v_buf.off);
                    C.ns = ns;
                    C.ksl = ks_bra;
                    C.ksr = ks_ket;
                    C.nt = nt;
                    if (semi) {
                        const double *const *f = factors + ((size_t)b * K + k) * 5;
                        AX_REQUIRE(f[0] && f[1] && f[2] && f[3] && f[4],
                                   "ax_fock_plan_create: ion %d: missing semiseparable factors "
                                   "for block %d, k %d", ion, b, k);
                        const Buf pre = reserve((size_t)nt * ks_ket * (ns + 1) * 8);
                        const Buf suf = reserve((size_t)nt * ks_ket * (ns + 1) * 8);
                        ChainJob B{};
                        B.rr = f[3];
                        B.fr = f[4];
                        B.dd = ph(dd_buf.off);
                        B.pre = ph(pre.off);
                        B.suf = ph(suf.off);
                        B.ns = ns;
                        B.ksr = ks_ket;
                        B.nt = nt;
                        chains.push_back(B);
                        C.mode = 1;
                        C.nr = f[0];
                        C.rl = f[1];
                        C.fl = f[2];
                        C.pre = ph(pre.off);
                        C.suf = ph(suf.off);
                    } else {
                        const double *bank = banks[(size_t)b * K + k];
                        AX_REQUIRE(bank, "ax_fock_plan_create: ion %d: missing bank for block %d, "
                                   "k %d", ion, b, k);
                        C.mode = 0;
                        C.bank = bank;
                    }
                    dir.push_back(C);
                    AsmEntry E{};
                    E.kind = 0;
                    E.v = ph(v_buf.off);
                    E.ksl = ks_bra;
                    E.ksr = ks_ket;
                    for (int g = 0; g < n_groups; ++g)
                        if (gd[g]) qe[dir_q][g].push_back(E);
                }
                if (any_x) {
                    const double *bank = banks[(size_t)b * K + k];
                    AX_REQUIRE(bank, "ax_fock_plan_create: ion %d: missing bank for block %d, k %d",
                               ion, b, k);
                    const size_t c0 = coef_img.size();
                    for (int t = 0; t < nt; ++t)
                        for (int j = 0; j < n_orb; ++j)
                            coef_img.push_back(xcoef[((size_t)t * n_orb + j) * K + k]);
                    AsmEntry E{};
                    E.kind = 1;
                    E.bank = bank;
                    E.coef = reinterpret_cast<const double *>(c0);
                    E.A = bra;
                    E.B = ket;
                    E.ksl = ks_bra;
                    E.ksr = ks_ket;
                    xr_max = std::max(xr_max, kTile + 2 * (ks_bra - 1));
                    xc_max = std::max(xc_max, kTile + 2 * (ks_ket - 1));
                    n_orb_max = std::max(n_orb_max, n_orb);
                    g_max = std::max(g_max, g_size);
                    for (int g = 0; g < n_groups; ++g)
                        if (gx[g]) qe[exch_q][g].push_back(E);
                }
            }
        }
        // assembly jobs: every tile of every quadrant (PQ only without skip)
        const int nT = (ns + kTile - 1) / kTile;
        for (int q = 0; q < 4; ++q) {
            if (q == 2 && skip) continue;
            const bool symmetric = q < 2;
            for (int g = 0; g < n_groups; ++g) {
                const int e0 = static_cast<int>(entries.size());
                // entries carry coefficient indices until the block exists
                for (const AsmEntry &E : qe[q][g]) entries.push_back(E);
                const int e1 = static_cast<int>(entries.size());
                for (int tr = 0; tr < nT; ++tr)
                    for (int tc = 0; tc < nT; ++tc) {
                        if (symmetric && tc > tr) continue;
                        AsmJob J{};
                        J.H = H[ion];
                        J.ns = ns;
                        J.n_orb = n_orb;
                        J.quadrant = q;
                        J.symmetric = symmetric ? 1 : 0;
                        J.r0 = tr * kTile;
                        J.c0 = tc * kTile;
                        J.t0 = g * g_size;
                        J.tg = std::min(g_size, nt - g * g_size);
                        J.e0 = e0;
                        J.e1 = e1;
                        asmj.push_back(J);
                    }
            }
        }
        const size_t d0 = coef_img.size();
        for (int t = 0; t < nt; ++t) coef_img.push_back(divisor[t]);
        FinJob F{};
        F.H = H[ion];
        F.divisor = reinterpret_cast<const double *>(d0);
        F.ns = ns;
        F.nt = nt;
        F.skip = skip;
        fin.push_back(F);
    }
    AsmLds geo{xr_max * xc_max, n_orb_max, xr_max, xc_max, g_max, 0};
    // g_max planes of the exchange densities, the two coefficient strips and
    // the group's coefficients
    const size_t max_lds = sizeof(double) * ((size_t)g_max * geo.xx_plane +
                                             (size_t)n_orb_max * (xr_max + xc_max) +
                                             (size_t)g_max * n_orb_max);
    AX_REQUIRE(max_lds <= 160 * 1024, "ax_fock_plan_create: LDS need %zu exceeds the CU", max_lds);

    // layout of the device block: coefficients, job tables, work buffers
    const size_t work_bytes = cursor;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off += align_up(bytes > 0 ? bytes : 1);
        return o;
    };
    const size_t o_coef = take(coef_img.size() * sizeof(double));
    const size_t o_dd = take(dd.size() * sizeof(DdJob));
    const size_t o_ch = take(chains.size() * sizeof(ChainJob));
    const size_t o_dir = take(dir.size() * sizeof(DirJob));
    const size_t o_ent = take(entries.size() * sizeof(AsmEntry));
    const size_t o_asm = take(asmj.size() * sizeof(AsmJob));
    const size_t o_fin = take(fin.size() * sizeof(FinJob));
    const size_t o_work = take(work_bytes);
    void *block = nullptr;
    AX_CHECK_HIP(ax::pool_alloc(&block, off));
    char *base = static_cast<char *>(block);
    const double *coef_dev = reinterpret_cast<const double *>(base + o_coef);
    auto work_ptr = [&](const double *p) -> double * {  // placeholder -> device pointer
        return p ? reinterpret_cast<double *>(base + o_work + (reinterpret_cast<size_t>(p) - 1))
                 : nullptr;
    };
    auto coef_ptr = [&](const double *p) {
        return coef_dev + reinterpret_cast<size_t>(p);
    };
    for (DdJob &A : dd) {
        A.coef = coef_ptr(A.coef);
        A.dd = work_ptr(A.dd);
    }
    for (ChainJob &B : chains) {
        B.dd = work_ptr(B.dd);
        B.pre = work_ptr(B.pre);
        B.suf = work_ptr(B.suf);
    }
    for (DirJob &C : dir) {
        C.dd = work_ptr(C.dd);
        C.v = work_ptr(C.v);
        if (C.mode == 1) {
            C.pre = work_ptr(C.pre);
            C.suf = work_ptr(C.suf);
        }
    }
    for (AsmEntry &E : entries) {
        if (E.kind == 0) E.v = work_ptr(E.v);
        else E.coef = coef_ptr(E.coef);
    }
    for (FinJob &F : fin) F.divisor = coef_ptr(F.divisor);

    ax_status st = AX_OK;
    auto put = [&](size_t o, const void *src, size_t bytes) {
        if (st == AX_OK && bytes > 0 && hipMemcpy(base + o, src, bytes, hipMemcpyHostToDevice) !=
                                            hipSuccess) {
            ax::set_error("ax_fock_plan_create: table upload failed");
            st = AX_ERR_HIP;
        }
    };
    put(o_coef, coef_img.data(), coef_img.size() * sizeof(double));
    put(o_dd, dd.data(), dd.size() * sizeof(DdJob));
    put(o_ch, chains.data(), chains.size() * sizeof(ChainJob));
    put(o_dir, dir.data(), dir.size() * sizeof(DirJob));
    put(o_ent, entries.data(), entries.size() * sizeof(AsmEntry));
    put(o_asm, asmj.data(), asmj.size() * sizeof(AsmJob));
    put(o_fin, fin.data(), fin.size() * sizeof(FinJob));
    if (st != AX_OK) {
        ax::pool_free(block);
        return st;
    }
    ax_fock_plan *plan = new ax_fock_plan();
    plan->block = block;
    plan->dd = reinterpret_cast<const DdJob *>(base + o_dd);
    plan->chains = reinterpret_cast<const ChainJob *>(base + o_ch);
    plan->dir = reinterpret_cast<const DirJob *>(base + o_dir);
    plan->entries = reinterpret_cast<const AsmEntry *>(base + o_ent);
    plan->asmj = reinterpret_cast<const AsmJob *>(base + o_asm);
    plan->fin = reinterpret_cast<const FinJob *>(base + o_fin);
    plan->n_dd = static_cast<int>(dd.size());
    plan->n_chains = static_cast<int>(chains.size());
    plan->n_dir = static_cast<int>(dir.size());
    plan->n_asm = static_cast<int>(asmj.size());
    plan->n_fin = static_cast<int>(fin.size());
    for (const ChainJob &B : chains) plan->max_chain = std::max(plan->max_chain, B.nt * B.ksr);
    plan->asm_lds = max_lds;
    plan->geo = geo;
    for (const DirJob &C : dir) {
        // one bank row per wave (4 per work-group), or one element per thread
        const int64_t units = C.mode == 0 ? ((int64_t)C.ns * C.ksl + 3) / 4
                                          : ((int64_t)C.ns * C.ksl * C.nt + kThreads - 1) / kThreads;
        plan->dir_blocks = static_cast<int>(std::min<int64_t>(std::max<int64_t>(
            plan->dir_blocks, units), 1024));
    }
    plan->bytes = static_cast<int64_t>(off);
    plan->rk_equivalents = rk_eq;
    if (max_lds > 64 * 1024)
        (void)hipFuncSetAttribute(reinterpret_cast<const void *>(fock_assemble_kernel),
                                  hipFuncAttributeMaxDynamicSharedMemorySize,
                                  static_cast<int>(max_lds));
    *plan_out = plan;
    return AX_OK;
}

ax_status ax_fock_plan_run(ax_fock_plan *plan, void *stream) {
    AX_REQUIRE(plan != nullptr, "ax_fock_plan_run: null plan");
    hipStream_t s = ax::as_stream(stream);
    AX_REQUIRE(plan->n_dd < 65536 && plan->n_chains < 65536 && plan->n_dir < 65536 &&
                   plan->n_fin < 65536,
               "ax_fock_plan_run: job tables exceed the grid limit");
    if (plan->n_dd > 0) {
        hipLaunchKernelGGL(fock_dd_kernel, dim3(8, plan->n_dd), dim3(kThreads), 0, s, plan->dd);
        AX_CHECK_LAUNCH();
    }
    if (plan->n_chains > 0) {
        hipLaunchKernelGGL(fock_chain_kernel, dim3(plan->max_chain, plan->n_chains), dim3(64), 0,
                           s, plan->chains);
        AX_CHECK_LAUNCH();
    }
    if (plan->n_dir > 0) {
        // enough waves per job that each bank row (or semiseparable element)
        // chain is short: the kernel is latency-bound, not bandwidth-bound
        hipLaunchKernelGGL(fock_direct_kernel, dim3(plan->dir_blocks, plan->n_dir), dim3(kThreads),
                           0, s, plan->dir);
        AX_CHECK_LAUNCH();
    }
    if (plan->n_asm > 0) {
        hipLaunchKernelGGL(fock_assemble_kernel, dim3(plan->n_asm), dim3(kThreads),
                           plan->asm_lds, s, plan->asmj, plan->entries, plan->geo);
        AX_CHECK_LAUNCH();
    }
    if (plan->n_fin > 0) {
        hipLaunchKernelGGL(fock_finalize_kernel, dim3(32, plan->n_fin), dim3(kThreads), 0, s,
                           plan->fin);
        AX_CHECK_LAUNCH();
    }
    return AX_OK;
}

int64_t ax_fock_plan_bytes(const ax_fock_plan *plan) { return plan ? plan->bytes : -1; }

int64_t ax_fock_plan_rk_equivalents(const ax_fock_plan *plan) {
    return plan ? plan->rk_equivalents : -1;
}

ax_status ax_fock_plan_destroy(ax_fock_plan *plan) {
    if (plan == nullptr) return AX_OK;
    // the caller has finished every run on the plan's streams
    AX_CHECK_HIP(hipDeviceSynchronize());
    ax::pool_free(plan->block);
    delete plan;
    return AX_OK;
}

}  // extern "C"
