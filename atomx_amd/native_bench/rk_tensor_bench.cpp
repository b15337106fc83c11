// Native driver of the all-quartet R^k tensor for hardware-counter runs: no
// Python and no PyTorch in the process, only the HIP runtime and
// libatomx_slater.so.  Reads a workload file written by write_workload.py,
// builds one plan, fixes (or tunes) its contraction schedule, runs
// `warmup` + `steps` steps, prints the per-step time, and ends with one
// device-to-device copy of `copy_mib` MiB that calibrates byte counters.
//
// usage: rk_tensor_bench WORKLOAD fp64|split ORDER SPLITS STEPS [COPY_MIB]
//        ORDER in 0..4 (ax_rk_tensor_set_schedule) and SPLITS a
//        comma-separated list of levels, each timed in turn; an entry O:S
//        runs order O at level S instead.  ORDER = -1 tunes.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <string_view>
#include <vector>

#include "atomx_slater.h"

namespace {

int fail(const char *what) {
    std::fprintf(stderr, "rk_tensor_bench: %s: %s\n", what, ax_last_error());
    return 1;
}

int hip_fail(const char *what, hipError_t e) {
    std::fprintf(stderr, "rk_tensor_bench: %s: %s\n", what, hipGetErrorString(e));
    return 1;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 6) {
        std::fprintf(stderr,
                     "usage: %s WORKLOAD fp64|split ORDER SPLIT STEPS [COPY_MIB]\n"
                     "  ORDER -1: the tuned schedule; ORDER -2: every schedule the tuner\n"
                     "  considers (orders 0-4, K split 1 .. the plan's cap), SPLIT ignored\n",
                     argv[0]);
        return 2;
    }
    const char *path = argv[1];
    const int precision = std::string_view(argv[2]) == "split" ? AX_PREC_SPLIT : AX_PREC_FP64;
    const int order = std::atoi(argv[3]);
    std::vector<int> splits, orders;
    for (const char *c = argv[4]; *c;) {
        const char *colon = c;
        while (*colon && *colon != ',' && *colon != ':') ++colon;
        if (*colon == ':') {
            orders.push_back(std::atoi(c));
            c = colon + 1;
        } else {
            orders.push_back(order);
        }
        splits.push_back(std::atoi(c));
        while (*c && *c != ',') ++c;
        if (*c == ',') ++c;
    }
    const int steps = std::atoi(argv[5]);
    const int copy_mib = argc > 6 ? std::atoi(argv[6]) : 0;

    // header: nv, ng, n_orb, kmax (int32); then r, w (nv*ng doubles each),
    // P, Q (n_orb*nv*ng doubles each), l (n_orb int32)
    FILE *f = std::fopen(path, "rb");
    if (!f) {
        std::fprintf(stderr, "rk_tensor_bench: cannot open %s\n", path);
        return 1;
    }
    int32_t hdr[4];
    if (std::fread(hdr, sizeof(int32_t), 4, f) != 4) return 1;
    const int nv = hdr[0], ng = hdr[1], n_orb = hdr[2], kmax = hdr[3];
    const size_t n = static_cast<size_t>(nv) * ng;
    if (nv <= 0 || ng <= 0 || n_orb <= 0 || kmax < 0 || n > (1u << 26)) {
        std::fprintf(stderr, "rk_tensor_bench: bad header\n");
        return 1;
    }
    std::vector<double> r(n), w(n), P(n * n_orb), Q(n * n_orb);
    std::vector<int32_t> l(n_orb);
    bool ok = std::fread(r.data(), 8, n, f) == n && std::fread(w.data(), 8, n, f) == n &&
              std::fread(P.data(), 8, P.size(), f) == P.size() &&
              std::fread(Q.data(), 8, Q.size(), f) == Q.size() &&
              std::fread(l.data(), 4, l.size(), f) == l.size();
    std::fclose(f);
    if (!ok) {
        std::fprintf(stderr, "rk_tensor_bench: short workload file\n");
        return 1;
    }

    double *dP = nullptr, *dQ = nullptr, *dout = nullptr;
    hipError_t e = hipMalloc(&dP, P.size() * 8);
    if (e == hipSuccess) e = hipMalloc(&dQ, Q.size() * 8);
    if (e != hipSuccess) return hip_fail("hipMalloc", e);
    e = hipMemcpy(dP, P.data(), P.size() * 8, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(dQ, Q.data(), Q.size() * 8, hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail("hipMemcpy", e);

    ax_rk_tensor *plan = nullptr;
    if (ax_rk_tensor_create(r.data(), w.data(), nv, ng, n_orb, l.data(), kmax, precision, &plan))
        return fail("create");
    const int64_t size = ax_rk_tensor_size(plan);
    e = hipMalloc(&dout, static_cast<size_t>(size) * 8);
    if (e != hipSuccess) return hip_fail("hipMalloc", e);
    hipStream_t s = nullptr;
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return 1;
    if (order == -2) {  // every (order, split) that ax_rk_tensor_tune times
        int o0 = 0, s0 = 0;
        if (ax_rk_tensor_get_schedule(plan, &o0, &s0)) return fail("get_schedule");
        const int cap = std::min(16, 4 * s0);
        orders.clear();
        splits.clear();
        for (int sp = 1; sp <= cap; ++sp)
            for (int o = 0; o < 5; ++o) {
                orders.push_back(o);
                splits.push_back(sp);
            }
    }
    const size_t n_sched = order == -1 ? 1 : splits.size();
    for (size_t si = 0; si < n_sched; ++si) {
        if (order == -1) {
            float tms = 0.0f;
            if (ax_rk_tensor_tune(plan, dP, dQ, dout, 3, s, &tms)) return fail("tune");
        } else if (ax_rk_tensor_set_schedule(plan, orders[si], splits[si])) {
            return fail("set_schedule");
        }
        int o = 0, sp = 0;
        if (ax_rk_tensor_get_schedule(plan, &o, &sp)) return fail("get_schedule");
        for (int i = 0; i < 3; ++i)
            if (ax_rk_tensor_run(plan, dP, dQ, dout, s)) return fail("run");
        e = hipEventRecord(e0, s);
        for (int i = 0; i < steps && e == hipSuccess; ++i)
            if (ax_rk_tensor_run(plan, dP, dQ, dout, s)) return fail("run");
        if (e == hipSuccess) e = hipEventRecord(e1, s);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.0f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) return hip_fail("timing", e);
        std::printf("workload %s: %d orbitals, %d x %d nodes, kmax %d, %s, order %d, K split %d: "
                    "%lld R^k, %.2f us per step\n",
                    path, n_orb, nv, ng, kmax, precision == AX_PREC_SPLIT ? "split" : "fp64", o,
                    sp, static_cast<long long>(size), 1e3 * ms / (steps > 0 ? steps : 1));
        std::fflush(stdout);
    }

    if (copy_mib > 0) {
        const size_t bytes = static_cast<size_t>(copy_mib) << 20;
        void *a = nullptr, *b = nullptr;
        e = hipMalloc(&a, bytes);
        if (e == hipSuccess) e = hipMalloc(&b, bytes);
        if (e == hipSuccess) e = hipMemset(a, 0, bytes);
        if (e == hipSuccess) e = hipMemcpyAsync(b, a, bytes, hipMemcpyDeviceToDevice, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) return hip_fail("copy calibration", e);
        std::printf("copy calibration: %d MiB read + %d MiB written\n", copy_mib, copy_mib);
        (void)hipFree(a);
        (void)hipFree(b);
    }
    if (ax_rk_tensor_destroy(plan)) return fail("destroy");
    (void)hipFree(dP);
    (void)hipFree(dQ);
    (void)hipFree(dout);
    std::printf("rk_tensor_bench: done\n");
    std::fflush(stdout);
    return 0;
}
