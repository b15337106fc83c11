// Library-level entry points: error reporting, version, host-side tables.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "ax_common.h"
#include "rk_gemm.h"

namespace ax {

static thread_local char g_error[1024] = "";

void set_error(const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_error, sizeof(g_error), fmt, ap);
    va_end(ap);
}

void gauss_legendre(int n, double *x, double *w) {
    if (n == 1) {
        x[0] = 0.0;
        w[0] = 2.0;
        return;
    }
    const long double pi = 3.14159265358979323846264338327950288L;
    for (int i = 0; i < n; ++i) {
        long double xi = cosl(pi * (i + 0.75L) / (n + 0.5L));
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
        long double p0 = 1.0L, p1 = xi;
        for (int j = 2; j <= n; ++j) {
            long double p2 = ((2 * j - 1) * xi * p1 - (j - 1) * p0) / j;
            p0 = p1;
            p1 = p2;
        }
        dp = n * (xi * p1 - p0) / (xi * xi - 1.0L);
        x[n - 1 - i] = static_cast<double>(xi);
        w[n - 1 - i] = static_cast<double>(2.0L / ((1.0L - xi * xi) * dp * dp));
    }
}

}  // namespace ax

extern "C" {

const char *ax_last_error(void) { return ax::g_error; }

int ax_version(void) { return 100; }

int ax_split_slices(void) { return ax::kSplitSlices; }
int ax_split_ktile(void) { return ax::kSplitKTile; }
int ax_split_products(void) { return ax::kSplitSlices * (ax::kSplitSlices + 1) / 2; }

ax_status ax_device_count(int *count_host) {
    int c = 0;
    hipError_t e = hipGetDeviceCount(&c);
    if (e != hipSuccess || c <= 0) {
        ax::set_error("no HIP device available (%s)", hipGetErrorString(e));
        if (count_host) *count_host = 0;
        return AX_ERR_HIP;
    }
    *count_host = c;
    return AX_OK;
}

// Interpolatory partial-interval matrices in the Legendre basis:
//   mo[i][j] = sum_l partial[i][l] (2l+1)/2 P_l(x_j),
//   partial[i][0] = x_i + 1,  partial[i][l] = (P_{l+1}(x_i) - P_{l-1}(x_i)) / (2l+1),
// i.e. the exact integral from -1 to x_i of the Lagrange polynomial L_j
// divided by its Gauss weight; mi = 1 - mo.
ax_status ax_yk_partial_matrices(int ng, double *mo_host, double *mi_host) {
    AX_REQUIRE(ng >= 1 && ng <= 64, "n_gauss must be in [1, 64], got %d", ng);
    std::vector<double> x(ng), w(ng);
    ax::gauss_legendre(ng, x.data(), w.data());
    std::vector<long double> vals(static_cast<size_t>(ng) * (ng + 1));
    for (int i = 0; i < ng; ++i) {
        long double xv = x[i];
        vals[i * (ng + 1) + 0] = 1.0L;
        vals[i * (ng + 1) + 1] = xv;
        for (int l = 2; l <= ng; ++l)
            vals[i * (ng + 1) + l] = ((2 * l - 1) * xv * vals[i * (ng + 1) + l - 1] -
                                      (l - 1) * vals[i * (ng + 1) + l - 2]) / l;
    }
    for (int i = 0; i < ng; ++i)
        for (int j = 0; j < ng; ++j) {
            long double s = 0.0L;
            for (int l = 0; l < ng; ++l) {
                long double part = (l == 0) ? (static_cast<long double>(x[i]) + 1.0L)
                                            : (vals[i * (ng + 1) + l + 1] - vals[i * (ng + 1) + l - 1]) / (2 * l + 1);
                s += part * ((2 * l + 1) / 2.0L) * vals[j * (ng + 1) + l];
            }
            mo_host[i * ng + j] = static_cast<double>(s);
            mi_host[i * ng + j] = static_cast<double>(1.0L - s);
        }
    return AX_OK;
}

}  // extern "C"
