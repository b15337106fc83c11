// Invariants of the contraction job tables (host code of the library), over
// random plans: every tile's K range is covered once, contiguously, by its
// pieces in slot order; slots are unique and tile-major; work-group job
// ranges are ordered and cover the job list.  Prints "bad N" and exits
// non-zero on any violation.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "rk_gemm.h"

using namespace ax;

namespace {

int bad = 0;

void fail(const char *what, int trial) {
    if (bad++ < 10) std::printf("trial %d: %s\n", trial, what);
}

void check_tiles(const std::vector<RkWork> &work, const std::vector<RkJob> &jobs,
                 const std::vector<RkTileJobs> &tiles, const KSpace &ks, int trial) {
    const int nw = static_cast<int>(work.size());
    std::vector<std::vector<const RkJob *>> per(nw);
    std::vector<int> seen(jobs.size(), 0);
    for (const RkJob &j : jobs) {
        const int ti = j.tile & 0xffff, tj = j.tile >> 16;
        int e = -1;
        for (int q = 0; q < nw; ++q)
            if (work[q].seg == j.seg && work[q].ti == ti && work[q].tj == tj) e = q;
        if (e < 0) { fail("job of no tile", trial); continue; }
        if (j.slot < 0 || j.slot >= static_cast<int>(jobs.size()) || seen[j.slot]++)
            fail("slot out of range or repeated", trial);
        per[e].push_back(&j);
    }
    for (int e = 0; e < nw; ++e) {
        const int cnt = static_cast<int>(per[e].size());
        if (cnt == 0 || tiles[e].count != cnt) { fail("tile piece count", trial); continue; }
        std::vector<const RkJob *> ord(cnt, nullptr);
        for (const RkJob *j : per[e]) {
            const int sp = j->split & 0xffff;
            if (sp >= cnt || (j->split >> 16) != cnt || j->slot != tiles[e].first + sp || ord[sp])
                fail("piece index or slot", trial);
            else
                ord[sp] = j;
        }
        int64_t k = 0;
        for (int i = 0; i < cnt; ++i) {
            if (!ord[i]) continue;
            if (ord[i]->kb != std::min<int64_t>(k, ks.k_len)) fail("K gap or overlap", trial);
            if (ord[i]->ke < ord[i]->kb) fail("K range reversed", trial);
            if (ord[i]->kb < ks.k_len && ord[i]->kb % ks.k_align) fail("K start unaligned", trial);
            k = std::max<int64_t>(k, ord[i]->ke);
        }
        if (k != ks.k_len) fail("K range not covered", trial);
    }
}

}  // namespace

int main() {
    for (int trial = 0; trial < 300; ++trial) {
        std::srand(trial + 1);
        std::vector<RkWork> work;
        std::vector<int> seg_m;
        const int nseg = 1 + std::rand() % 5;
        for (int s = 0; s < nseg; ++s) {
            seg_m.push_back(1 + std::rand() % 260);
            upper_tiles(seg_m.back(), s, work);
        }
        const int64_t n = 1 + std::rand() % 4000;
        const KSpace ks = k_space(n, std::rand() % 2 == 0);
        const int S = 1 + std::rand() % 16;
        const JobTables t = make_jobs(work, seg_m, S, ks);
        check_tiles(work, t.tile_fastest, t.tiles, ks, trial);
        check_tiles(work, t.split_fastest, t.tiles, ks, trial);
        check_tiles(work, t.heaviest, t.tiles, ks, trial);
        const BalancedJobs b = make_balanced_jobs(work, seg_m, balanced_groups(S), ks);
        check_tiles(work, b.jobs, b.tiles, ks, trial);
        if (b.groups.empty() || b.groups.front() != 0 ||
            b.groups.back() != static_cast<int>(b.jobs.size()))
            fail("work-group ranges do not cover the jobs", trial);
        for (size_t g = 1; g < b.groups.size(); ++g)
            if (b.groups[g] < b.groups[g - 1]) fail("work-group ranges out of order", trial);
    }
    std::printf("bad %d\n", bad);
    return bad != 0;
}
