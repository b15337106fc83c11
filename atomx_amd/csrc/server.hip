// Native serving loop for all-quartet R^k requests (continuous batching).
//
// One scheduler thread per server owns the device work.  Callers submit
// requests (grid, angular momenta, kmax, precision and host orbital samples
// that stay valid until the request completes) from any thread and get a
// ticket.  The scheduler admits queued requests in arrival order into a
// launch set while an in-flight slot is free, the set holds fewer than
// max_batch requests and the in-flight device bytes (plans plus staged
// inputs and outputs) stay within the budget; it then
//   * starts packing the set's orbitals into the slot's pinned staging on
//     the pack threads,
//   * creates the set's plans with one ax_rk_tensor_create_many call (a
//     failing set is retried plan by plan, so a bad request fails alone),
//     except for requests that name a resident state holding a plan,
//   * waits for the packing, and on the slot's stream copies the inputs to
//     the device with one copy, runs every plan as one batch and copies all
//     results into a pinned buffer of the set with one copy, then records
//     the set's completion event.
// Completed sets hand their results to the tickets as views of that buffer
// (no host copy); the buffer returns to the server's pinned cache when every
// ticket of the set has been released.  Several sets are in flight on
// separate streams.  Each plan's results are bitwise those of the same plan
// run alone (the batch does not change a plan's tiles, splits or order).
// A request may name a resident state (one atom's requests, SCF iteration
// after SCF iteration): the state keeps the plan its first request made and
// later requests of the same shape reuse it; a state runs one request at a
// time (later requests of other states go ahead of a waiting one), its plan
// counts against the budget while it is held, and idle states give their
// plans up, least recently used first, when a request does not fit.
// The request surface is that of the reference's per-integral entry points
// (numerics/bspline/slater.py:118-130) taken over all quartets at once.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "ax_common.h"

namespace {

struct Request {
    int64_t ticket = 0;
    const double *r = nullptr, *w = nullptr, *P = nullptr, *Q = nullptr;
    int nv = 0, ng = 0, n_orb = 0, kmax = 0, precision = 0;
    std::vector<int32_t> l;
    int64_t n_in = 0;   // staged input values (P, and Q when given)
    int64_t bytes = 0;  // device bytes while in flight: plan + staged inputs and outputs
    int64_t plan_bytes = 0;  // the plan's part of `bytes`
    int64_t state = 0;       // resident state key (0: none)
    struct State *st = nullptr;  // that state, set at admission
};

// A resident state: the plan of the requests that name one key (the next
// SCF iteration of an atom), kept between them, and the shape it was made
// for.  At most one request of a state is admitted at a time.
struct State {
    ax_rk_tensor *plan = nullptr;
    int nv = 0, ng = 0, n_orb = 0, kmax = 0, precision = 0;
    bool has_q = false;
    std::vector<int32_t> l;
    std::vector<double> r, w;  // the grid the plan was made for
    int64_t plan_bytes = 0;    // charged to the resident bytes while reserved
    bool charged = false;      // plan_bytes reserved (plan made or being made)
    bool busy = false;         // a request of it is admitted and not complete
    bool drop = false;         // free it once idle with nothing queued
    int queued = 0;            // its requests still in the queue
    int64_t last = 0;          // use stamp (least recently used goes first)
};

struct LaunchSet {
    int slot = -1;
    std::vector<ax_rk_tensor *> plans;
    std::vector<State *> owner;    // per plan: its resident state (null: freed at completion)
    std::vector<int64_t> tickets;  // one per plan
    std::vector<int64_t> offset;   // result offset of each plan in `results`
    hipEvent_t done = nullptr;
    double *results = nullptr;  // pinned, from the server's cache
    size_t result_bytes = 0;
    double *staging = nullptr;  // pinned input staging, from the same cache
    size_t staging_bytes = 0;
    int refs = 0;  // completed tickets not yet released
    int64_t bytes = 0;
};

struct Ticket {
    bool done = false;
    ax_status status = AX_OK;
    std::string error;
    int slices = 0;
    int64_t size = 0;
    const double *values = nullptr;
    LaunchSet *set = nullptr;
};

struct Slot {
    hipStream_t stream = nullptr;
    void *dev = nullptr;     // device staging (inputs, then outputs)
    size_t dev_bytes = 0;
    bool busy = false;
};

// Persistent threads that copy the segments of one launch set into its
// staging in equal byte ranges of their concatenation.
class PackPool {
  public:
    struct Seg {
        char *dst;
        const char *src;
        size_t bytes;
    };
    explicit PackPool(int n) {
        for (int i = 0; i < n; ++i) th_.emplace_back([this] { loop(); });
    }
    ~PackPool() {
        {
            std::lock_guard<std::mutex> lk(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto &t : th_) t.join();
    }
    // Starts copying `segs` (kept alive by the caller until wait()).
    void start(const std::vector<Seg> *segs) {
        std::lock_guard<std::mutex> lk(m_);
        segs_ = segs;
        total_ = 0;
        for (const Seg &s : *segs) total_ += s.bytes;
        parts_ = th_.empty() ? 0 : static_cast<int>(th_.size());
        if (total_ < (size_t(4) << 20)) parts_ = std::min(parts_, 1);  // small sets: one thread
        next_ = 0;
        left_ = parts_;
        ++gen_;
        cv_.notify_all();
    }
    void wait() {
        std::unique_lock<std::mutex> lk(m_);
        done_cv_.wait(lk, [this] { return left_ == 0; });
        segs_ = nullptr;
    }

  private:
    void copy_range(size_t lo, size_t hi) {
        size_t pos = 0;
        for (const Seg &s : *segs_) {
            const size_t a = std::max(lo, pos), b = std::min(hi, pos + s.bytes);
            if (a < b) std::memcpy(s.dst + (a - pos), s.src + (a - pos), b - a);
            pos += s.bytes;
            if (pos >= hi) break;
        }
    }
    void loop() {
        uint64_t seen = 0;
        std::unique_lock<std::mutex> lk(m_);
        for (;;) {
            cv_.wait(lk, [&] { return stop_ || (gen_ != seen && next_ < parts_); });
            if (stop_) return;
            seen = gen_;
            while (next_ < parts_) {
                const int p = next_++;
                const size_t lo = total_ * p / parts_, hi = total_ * (p + 1) / parts_;
                lk.unlock();
                copy_range(lo, hi);
                lk.lock();
                if (--left_ == 0) done_cv_.notify_all();
            }
        }
    }
    std::vector<std::thread> th_;
    std::mutex m_;
    std::condition_variable cv_, done_cv_;
    const std::vector<Seg> *segs_ = nullptr;
    size_t total_ = 0;
    int parts_ = 0, next_ = 0, left_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
};

size_t pinned_class(size_t n) {
    size_t c = size_t(1) << 16;
    while (c < n) c <<= 1;
    return c;
}

}  // namespace

struct ax_server {
    int device = 0;
    int64_t budget = 0;
    int max_inflight = 4, max_batch = 16;
    std::mutex mu;
    std::condition_variable work_cv, done_cv;
    std::deque<Request> queue;
    std::unordered_map<int64_t, Ticket> tickets;
    std::vector<LaunchSet *> inflight;
    std::vector<Slot> slots;
    std::multimap<size_t, double *> pinned_free;  // result buffers by size class
    std::unordered_map<int64_t, std::unique_ptr<State>> states;
    int64_t state_bytes = 0;  // plan bytes of the resident states
    int64_t use_clock = 0, plans_made = 0;
    int64_t next_ticket = 1, inflight_bytes = 0;
    int preparing = 0;  // launch sets a scheduler thread is preparing
    int64_t launches = 0, plans_created = 0, peak_bytes = 0;
    // host nanoseconds of the scheduler thread: plan set-up, waiting for the
    // packing, pinned and device allocations, queueing copies and launches
    int64_t ns_setup = 0, ns_pack_wait = 0, ns_alloc = 0, ns_queue = 0;
    bool stop = false;
    // scheduler threads, each with its own pack threads (two measured no
    // faster than one: the device work of a saturated stream bounds it)
    static constexpr int kWorkers = 1;
    PackPool *pack[kWorkers] = {nullptr};
    std::thread worker[kWorkers];

    void run(int id);
    bool complete_locked();
    bool evict_locked(int64_t need);
    void free_state_plan_locked(State *st);
    bool launch(std::vector<Request> &batch, int slot, int64_t bytes, PackPool *pk);
    void fail_ticket_locked(int64_t t, ax_status st, const std::string &msg);
    double *pinned_acquire(size_t bytes, size_t *got);
    void pinned_release_locked(double *p, size_t bytes);
};

void ax_server::fail_ticket_locked(int64_t t, ax_status st, const std::string &msg) {
    Ticket &k = tickets[t];
    k.done = true;
    k.status = st;
    k.error = msg;
}

// A pinned buffer of at least `bytes` (the smallest cached one that is large
// enough, else a new one of the size class); *got = its size.
static int64_t now_ns() {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

double *ax_server::pinned_acquire(size_t bytes, size_t *got) {
    const size_t cls = pinned_class(bytes);
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = pinned_free.lower_bound(cls);
        if (it != pinned_free.end()) {
            double *p = it->second;
            *got = it->first;
            pinned_free.erase(it);
            return p;
        }
    }
    void *p = nullptr;
    const hipError_t e = hipHostMalloc(&p, cls, hipHostMallocDefault);
    if (e != hipSuccess) return nullptr;
    *got = cls;
    return static_cast<double *>(p);
}

// Returns a buffer of `bytes` (its size from pinned_acquire) to the cache.
void ax_server::pinned_release_locked(double *p, size_t bytes) {
    if (p) pinned_free.insert({bytes, p});
}

// Hands the results of finished sets to their tickets and frees their slots
// and plans.  Called with `mu` held; returns whether a set finished.
bool ax_server::complete_locked() {
    bool any = false;
    for (size_t i = 0; i < inflight.size();) {
        LaunchSet *s = inflight[i];
        const hipError_t q = hipEventQuery(s->done);
        if (q == hipErrorNotReady) {
            ++i;
            continue;
        }
        for (size_t j = 0; j < s->plans.size(); ++j) {
            Ticket &k = tickets[s->tickets[j]];
            if (q == hipSuccess) {
                k.values = s->results + s->offset[j];
                k.slices = ax_rk_tensor_slices(s->plans[j]);
                k.size = ax_rk_tensor_size(s->plans[j]);
                k.set = s;
                ++s->refs;
            } else {
                k.status = AX_ERR_HIP;
                k.error = std::string("launch set failed: ") + hipGetErrorString(q);
            }
            k.done = true;
            State *st = s->owner[j];
            if (!st) {
                (void)ax_rk_tensor_destroy(s->plans[j]);
                continue;
            }
            st->busy = false;
            st->last = ++use_clock;
            if (q != hipSuccess || (st->drop && st->queued == 0)) free_state_plan_locked(st);
        }
        (void)hipEventDestroy(s->done);
        pinned_release_locked(s->staging, s->staging_bytes);
        s->staging = nullptr;
        slots[s->slot].busy = false;
        inflight_bytes -= s->bytes;
        if (s->refs == 0) {
            pinned_release_locked(s->results, s->result_bytes);
            delete s;
        }
        inflight[i] = inflight.back();
        inflight.pop_back();
        any = true;
    }
    return any;
}

// Frees a state's plan and its charge (the state keeps its shape; a state
// marked for dropping with nothing queued goes away).
void ax_server::free_state_plan_locked(State *st) {
    if (st->plan) (void)ax_rk_tensor_destroy(st->plan);
    st->plan = nullptr;
    if (st->charged) state_bytes -= st->plan_bytes;
    st->charged = false;
    if (st->drop && st->queued == 0 && !st->busy)
        for (auto it = states.begin(); it != states.end(); ++it)
            if (it->second.get() == st) {
                states.erase(it);
                break;
            }
}

// Frees the plans of idle resident states, least recently used first, until
// `need` more bytes fit the budget; returns whether they do.
bool ax_server::evict_locked(int64_t need) {
    while (inflight_bytes + state_bytes + need > budget) {
        State *lru = nullptr;
        for (auto &kv : states) {
            State *st = kv.second.get();
            // busy first: the set holding a busy state sets its plan unlocked
            if (!st->busy && st->charged && st->plan && (!lru || st->last < lru->last)) lru = st;
        }
        if (!lru) return false;
        free_state_plan_locked(lru);
    }
    return true;
}

// Prepares and queues one launch set on slot `slot_id` (called without `mu`).
// Returns whether the set went in flight; it then holds the `bytes` reserved
// for it at admission until it completes.
bool ax_server::launch(std::vector<Request> &batch, int slot_id, int64_t bytes, PackPool *pk) {
    Slot &slot = slots[slot_id];
    const int n = static_cast<int>(batch.size());
    auto fail_all = [&](ax_status st, const std::string &msg) {
        std::lock_guard<std::mutex> lk(mu);
        for (const Request &r : batch) fail_ticket_locked(r.ticket, st, msg);
        slot.busy = false;
        done_cv.notify_all();
    };
    // staging for the inputs (sizes known from the requests), packed while
    // the plans are made
    int64_t n_in = 0;
    for (const Request &r : batch) n_in += r.n_in;
    const size_t in_bytes = static_cast<size_t>(n_in) * 8;
    size_t staging_bytes = 0;
    const int64_t t_stage = now_ns();
    double *staging = pinned_acquire(in_bytes, &staging_bytes);
    const int64_t ns_stage = now_ns() - t_stage;
    if (!staging) {
        fail_all(AX_ERR_MEMORY, "pinned input staging allocation failed");
        return false;
    }
    std::vector<PackPool::Seg> segs;
    std::vector<int64_t> in_off(n);
    {
        int64_t off = 0;
        for (int i = 0; i < n; ++i) {
            const Request &r = batch[i];
            const int64_t m = (int64_t)r.n_orb * r.nv * r.ng;
            in_off[i] = off;
            segs.push_back({reinterpret_cast<char *>(staging + off),
                            reinterpret_cast<const char *>(r.P), static_cast<size_t>(m) * 8});
            off += m;
            if (r.Q) {
                segs.push_back({reinterpret_cast<char *>(staging + off),
                                reinterpret_cast<const char *>(r.Q), static_cast<size_t>(m) * 8});
                off += m;
            }
        }
    }
    pk->start(&segs);

    // plans: a resident state's own (this set holds the state busy), the
    // others made by one call for the set, plan by plan when that fails
    const int64_t t_setup = now_ns();
    std::vector<ax_rk_tensor *> plans(n, nullptr);
    std::vector<std::string> errs(n);
    std::vector<ax_status> sts(n, AX_OK);
    std::vector<int> make;
    for (int i = 0; i < n; ++i) {
        if (batch[i].st && batch[i].st->plan)
            plans[i] = batch[i].st->plan;
        else
            make.push_back(i);
    }
    const int nm = static_cast<int>(make.size());
    std::vector<const double *> rr(nm), ww(nm);
    std::vector<const int32_t *> ll(nm);
    std::vector<int32_t> nv(nm), ng(nm), no(nm), km(nm), pr(nm);
    std::vector<ax_rk_tensor *> made(nm, nullptr);
    for (int j = 0; j < nm; ++j) {
        const Request &q = batch[make[j]];
        rr[j] = q.r;
        ww[j] = q.w;
        ll[j] = q.l.data();
        nv[j] = q.nv;
        ng[j] = q.ng;
        no[j] = q.n_orb;
        km[j] = q.kmax;
        pr[j] = q.precision;
    }
    if (nm > 0 && ax_rk_tensor_create_many(nm, rr.data(), ww.data(), nv.data(), ng.data(),
                                           no.data(), ll.data(), km.data(), pr.data(), slot.stream,
                                           made.data()) != AX_OK) {
        for (int j = 0; j < nm; ++j) {
            const int i = make[j];
            sts[i] = ax_rk_tensor_create(rr[j], ww[j], nv[j], ng[j], no[j], ll[j], km[j], pr[j],
                                         &made[j]);
            if (sts[i] != AX_OK) {
                errs[i] = ax_last_error();
                made[j] = nullptr;
            }
        }
    }
    int made_ok = 0;
    for (int j = 0; j < nm; ++j) {
        plans[make[j]] = made[j];
        if (batch[make[j]].st) batch[make[j]].st->plan = made[j];  // null if it failed
        made_ok += made[j] != nullptr;
    }
    const int64_t t_pack = now_ns();
    pk->wait();
    const int64_t t_alloc = now_ns();

    LaunchSet *s = new LaunchSet();
    s->slot = slot_id;
    s->staging = staging;
    s->staging_bytes = staging_bytes;
    int64_t n_out = 0;
    std::vector<const double *> pp, qq;
    std::vector<double *> oo;
    std::vector<int> keep;
    for (int i = 0; i < n; ++i) {
        if (!plans[i]) continue;
        keep.push_back(i);
        s->offset.push_back(n_out);
        n_out += ax_rk_tensor_size(plans[i]);
    }
    const size_t dev_need = static_cast<size_t>(n_in + n_out) * 8;
    hipError_t e = hipSuccess;
    if (!keep.empty() && slot.dev_bytes < dev_need) {
        if (slot.dev) ax::pool_free(slot.dev);
        slot.dev = nullptr;
        slot.dev_bytes = 0;
        const size_t grow = dev_need + dev_need / 2;  // headroom: sets vary in size
        e = ax::pool_alloc(&slot.dev, grow);
        if (e == hipSuccess) slot.dev_bytes = ax::pool_class(grow);
    }
    if (!keep.empty() && e == hipSuccess) {
        s->results = pinned_acquire(static_cast<size_t>(n_out) * 8, &s->result_bytes);
        if (!s->results) e = hipErrorOutOfMemory;
    }
    double *dev = static_cast<double *>(slot.dev);
    const int64_t t_queue = now_ns();
    ax_status run = AX_OK;
    if (!keep.empty() && e == hipSuccess) {
        for (size_t j = 0; j < keep.size(); ++j) {
            const int i = keep[j];
            const int64_t m = (int64_t)batch[i].n_orb * batch[i].nv * batch[i].ng;
            pp.push_back(dev + in_off[i]);
            qq.push_back(batch[i].Q ? dev + in_off[i] + m : nullptr);
            oo.push_back(dev + n_in + s->offset[j]);
            s->plans.push_back(plans[i]);
            s->owner.push_back(batch[i].st);
            s->tickets.push_back(batch[i].ticket);
        }
        e = hipMemcpyAsync(dev, staging, in_bytes, hipMemcpyHostToDevice, slot.stream);
        if (e == hipSuccess)
            run = ax_rk_tensor_run_batch(s->plans.data(), static_cast<int>(s->plans.size()),
                                         pp.data(), qq.data(), oo.data(), slot.stream);
        if (e == hipSuccess && run == AX_OK)
            e = hipMemcpyAsync(s->results, dev + n_in, static_cast<size_t>(n_out) * 8,
                               hipMemcpyDeviceToHost, slot.stream);
        if (e == hipSuccess && run == AX_OK)
            e = hipEventCreateWithFlags(&s->done, hipEventDisableTiming);
        if (e == hipSuccess && run == AX_OK) e = hipEventRecord(s->done, slot.stream);
    }
    const int64_t t_end = now_ns();
    std::lock_guard<std::mutex> lk(mu);
    ns_setup += t_pack - t_setup;
    ns_pack_wait += t_alloc - t_pack;
    ns_alloc += ns_stage + (t_queue - t_alloc);  // input staging, device and result buffers
    ns_queue += t_end - t_queue;
    plans_made += made_ok;
    // a state whose plan could not be made, or whose set failed, is idle
    // again and gives back its charge (its next request makes a new plan)
    auto release_state = [&](State *st) {
        if (!st) return;
        st->plan = nullptr;
        st->busy = false;
        st->last = ++use_clock;
        free_state_plan_locked(st);
    };
    for (int i = 0; i < n; ++i)
        if (!plans[i]) {
            fail_ticket_locked(batch[i].ticket, sts[i], errs[i]);
            release_state(batch[i].st);
        }
    if (keep.empty() || e != hipSuccess || run != AX_OK) {
        const std::string msg = run != AX_OK ? std::string(ax_last_error())
                                             : std::string("launch set failed: ") +
                                                   hipGetErrorString(e);
        if (e != hipSuccess || run != AX_OK) (void)hipStreamSynchronize(slot.stream);
        for (size_t j = 0; j < s->plans.size(); ++j) {
            fail_ticket_locked(s->tickets[j], run != AX_OK ? run : AX_ERR_HIP, msg);
            (void)ax_rk_tensor_destroy(s->plans[j]);
            release_state(s->owner[j]);
        }
        for (int i : keep)
            if (std::find(s->plans.begin(), s->plans.end(), plans[i]) == s->plans.end()) {
                fail_ticket_locked(batch[i].ticket, AX_ERR_MEMORY, "staging allocation failed");
                (void)ax_rk_tensor_destroy(plans[i]);
                release_state(batch[i].st);
            }
        if (s->done) (void)hipEventDestroy(s->done);
        pinned_release_locked(s->results, s->result_bytes);
        pinned_release_locked(s->staging, s->staging_bytes);
        delete s;
        slot.busy = false;
        done_cv.notify_all();
        return false;
    }
    s->bytes = bytes;
    inflight.push_back(s);
    ++launches;
    plans_created += static_cast<int64_t>(s->plans.size());
    return true;
}

void ax_server::run(int id) {
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        if (complete_locked()) done_cv.notify_all();
        if (stop && queue.empty() && inflight.empty()) return;
        // admission: arrival order, while a slot is free, the set has room
        // and the in-flight bytes fit the budget
        int slot = -1;
        for (int i = 0; i < max_inflight; ++i)
            if (!slots[i].busy) {
                slot = i;
                break;
            }
        std::vector<Request> batch;
        if (slot >= 0) {
            // a request whose state is busy (its previous request is in flight
            // or in this set) waits and later ones go ahead, so a state's own
            // requests keep their order; a resident plan is charged to the
            // state's bytes, not to the set's, and idle states lose their
            // plans (least recently used first) when a request does not fit
            int64_t bytes = 0;
            for (auto it = queue.begin(); it != queue.end() && (int)batch.size() < max_batch;) {
                State *st = it->state ? states.at(it->state).get() : nullptr;
                if (st && st->busy) {
                    ++it;
                    continue;
                }
                const bool new_plan = st && !st->charged;
                const int64_t need = st ? it->bytes - it->plan_bytes : it->bytes;
                const int64_t charge = need + (new_plan ? st->plan_bytes : 0);
                if (inflight_bytes + state_bytes + bytes + charge > budget &&
                    !evict_locked(bytes + charge))
                    break;
                if (st) {
                    st->busy = true;
                    --st->queued;
                    if (new_plan) {
                        st->charged = true;
                        state_bytes += st->plan_bytes;
                    }
                }
                bytes += need;
                it->st = st;
                batch.push_back(std::move(*it));
                it = queue.erase(it);
            }
            if (!batch.empty()) {
                slots[slot].busy = true;
                inflight_bytes += bytes;  // reserved while the set is prepared
                peak_bytes = std::max(peak_bytes, inflight_bytes + state_bytes);
                ++preparing;
                lk.unlock();
                const bool queued = launch(batch, slot, bytes, pack[id]);
                lk.lock();
                --preparing;
                if (!queued) inflight_bytes -= bytes;
                work_cv.notify_all();  // a slot or the budget may have been waited for
                continue;
            }
        }
        // nothing to admit now: sleep until new work, or poll the sets in
        // flight (and the other thread's preparation) every 50 us
        if (inflight.empty() && preparing == 0 && (queue.empty() || slot >= 0))
            work_cv.wait(lk, [&] { return stop || !queue.empty(); });
        else
            work_cv.wait_for(lk, std::chrono::microseconds(50));
    }
}

extern "C" {

ax_status ax_server_create(int64_t budget_bytes, int max_inflight, int max_batch,
                           ax_server **out) {
    AX_REQUIRE(out != nullptr, "ax_server_create: null server pointer");
    AX_REQUIRE(budget_bytes > 0, "ax_server_create: budget must be positive");
    AX_REQUIRE(max_inflight >= 1 && max_inflight <= 16,
               "ax_server_create: max_inflight must be in [1, 16]");
    AX_REQUIRE(max_batch >= 1 && max_batch <= 16, "ax_server_create: max_batch must be in [1, 16]");
    ax_server *s = new ax_server();
    AX_CHECK_HIP(hipGetDevice(&s->device));
    s->budget = budget_bytes;
    s->max_inflight = max_inflight;
    s->max_batch = max_batch;
    s->slots.resize(max_inflight);
    for (Slot &sl : s->slots) {
        const hipError_t e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            for (Slot &t : s->slots)
                if (t.stream) (void)hipStreamDestroy(t.stream);
            delete s;
            ax::set_error("ax_server_create: stream creation failed: %s", hipGetErrorString(e));
            return AX_ERR_HIP;
        }
    }
    const unsigned hw = std::thread::hardware_concurrency();
    const int per = static_cast<int>(std::max(1u, std::min(8u, hw / (2 * ax_server::kWorkers))));
    for (int i = 0; i < ax_server::kWorkers; ++i) s->pack[i] = new PackPool(per);
    for (int i = 0; i < ax_server::kWorkers; ++i) s->worker[i] = std::thread([s, i] { s->run(i); });
    *out = s;
    return AX_OK;
}

ax_status ax_server_destroy(ax_server *s) {
    if (!s) return AX_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->work_cv.notify_all();
    for (auto &w : s->worker) w.join();  // drain the queue and every in-flight set
    for (PackPool *p : s->pack) delete p;
    for (auto &kv : s->states)
        if (kv.second->plan) (void)ax_rk_tensor_destroy(kv.second->plan);
    s->states.clear();
    for (Slot &sl : s->slots) {
        if (sl.dev) ax::pool_free(sl.dev);
        if (sl.stream) {
            // plan set-up staging bound to this stream goes with it
            (void)hipStreamSynchronize(sl.stream);
            ax::staging_forget_stream(sl.stream);
            (void)hipStreamDestroy(sl.stream);
        }
    }
    // results not released yet stay valid: their pinned buffers are left to
    // the process (a caller may still hold views of them)
    for (auto &kv : s->tickets)
        if (kv.second.set && --kv.second.set->refs == 0) delete kv.second.set;
    for (auto &kv : s->pinned_free) (void)hipHostFree(kv.second);
    delete s;
    return AX_OK;
}

ax_status ax_server_submit(ax_server *s, const double *r_host, const double *w_host, int nv,
                           int ng, int n_orb, const int32_t *l_host, int kmax, int precision,
                           const double *P_host, const double *Q_host, int64_t state,
                           int64_t *ticket) {
    AX_REQUIRE(s && r_host && w_host && l_host && P_host && ticket,
               "ax_server_submit: null argument");
    // a request of an existing state of the same scalars has its plan size
    // (its full shape is checked below); others are estimated
    int64_t plan_bytes = -1;
    if (state != 0) {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->states.find(state);
        if (it != s->states.end()) {
            const State &st = *it->second;
            if (st.nv == nv && st.ng == ng && st.n_orb == n_orb && st.kmax == kmax &&
                st.precision == precision)
                plan_bytes = st.plan_bytes;
        }
    }
    if (plan_bytes < 0) plan_bytes = ax_rk_tensor_estimate(nv, ng, n_orb, l_host, kmax, precision);
    AX_REQUIRE(plan_bytes >= 0, "ax_server_submit: invalid request (%s)", ax_last_error());
    const int64_t n = (int64_t)nv * ng;
    // R^k count of the request (the plan's output size), by angular class
    std::map<int, int64_t> cls;
    for (int a = 0; a < n_orb; ++a) ++cls[l_host[a]];
    std::vector<int64_t> m(kmax + 1, 0);
    for (auto i = cls.begin(); i != cls.end(); ++i)
        for (auto j = i; j != cls.end(); ++j) {
            const int64_t c = i == j ? i->second * (i->second + 1) / 2 : i->second * j->second;
            for (int k = j->first - i->first; k <= std::min(i->first + j->first, kmax); k += 2)
                m[k] += c;
        }
    int64_t n_rk = 0;
    for (int64_t x : m) n_rk += x * (x + 1) / 2;
    Request q;
    q.r = r_host;
    q.w = w_host;
    q.P = P_host;
    q.Q = Q_host;
    q.nv = nv;
    q.ng = ng;
    q.n_orb = n_orb;
    q.kmax = kmax;
    q.precision = precision;
    q.l.assign(l_host, l_host + n_orb);
    q.n_in = (int64_t)n_orb * n * (Q_host ? 2 : 1);
    q.bytes = plan_bytes + 8 * (q.n_in + n_rk);
    q.plan_bytes = plan_bytes;
    q.state = state;
    std::lock_guard<std::mutex> lk(s->mu);
    AX_REQUIRE_CODE(!s->stop, AX_ERR_VALUE, "ax_server_submit: the server is closing");
    AX_REQUIRE_CODE(q.bytes <= s->budget, AX_ERR_MEMORY,
                    "ax_server_submit: request needs %lld device bytes, above the budget %lld",
                    (long long)q.bytes, (long long)s->budget);
    if (state != 0) {
        // the state's shape is that of its first request; later ones must
        // match it (grid values included), since they reuse its plan
        auto &slot = s->states[state];
        if (!slot) {
            slot.reset(new State());
            State &st = *slot;
            st.nv = nv;
            st.ng = ng;
            st.n_orb = n_orb;
            st.kmax = kmax;
            st.precision = precision;
            st.has_q = Q_host != nullptr;
            st.l = q.l;
            st.r.assign(r_host, r_host + n);
            st.w.assign(w_host, w_host + n);
            st.plan_bytes = plan_bytes;
        } else {
            const State &st = *slot;
            const bool same = st.nv == nv && st.ng == ng && st.n_orb == n_orb &&
                              st.kmax == kmax && st.precision == precision &&
                              st.has_q == (Q_host != nullptr) && st.l == q.l &&
                              std::memcmp(st.r.data(), r_host, n * sizeof(double)) == 0 &&
                              std::memcmp(st.w.data(), w_host, n * sizeof(double)) == 0;
            AX_REQUIRE(same, "ax_server_submit: state %lld was created with another shape",
                       (long long)state);
        }
        slot->drop = false;
        ++slot->queued;
    }
    q.ticket = s->next_ticket++;
    s->tickets[q.ticket] = Ticket{};
    s->tickets[q.ticket].size = n_rk;
    *ticket = q.ticket;
    s->queue.push_back(std::move(q));
    s->work_cv.notify_one();
    return AX_OK;
}

ax_status ax_server_poll(ax_server *s, int64_t ticket, int32_t *done) {
    AX_REQUIRE(s && done, "ax_server_poll: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    AX_REQUIRE(it != s->tickets.end(), "ax_server_poll: unknown ticket %lld", (long long)ticket);
    *done = it->second.done ? 1 : 0;
    return AX_OK;
}

ax_status ax_server_wait(ax_server *s, int64_t ticket) {
    AX_REQUIRE(s != nullptr, "ax_server_wait: null server");
    std::unique_lock<std::mutex> lk(s->mu);
    AX_REQUIRE(s->tickets.count(ticket) != 0, "ax_server_wait: unknown ticket %lld",
               (long long)ticket);
    s->done_cv.wait(lk, [&] {  // found again each time: submits may rehash the table
        const auto it = s->tickets.find(ticket);
        return it == s->tickets.end() || it->second.done;
    });
    return AX_OK;
}

ax_status ax_server_result(ax_server *s, int64_t ticket, const double **values, int64_t *size,
                           int32_t *slices) {
    AX_REQUIRE(s && values && size && slices, "ax_server_result: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    AX_REQUIRE(it != s->tickets.end(), "ax_server_result: unknown ticket %lld", (long long)ticket);
    const Ticket &k = it->second;
    AX_REQUIRE(k.done, "ax_server_result: ticket %lld is not complete", (long long)ticket);
    if (k.status != AX_OK) {
        ax::set_error("%s", k.error.c_str());
        return k.status;
    }
    *values = k.values;
    *size = k.size;
    *slices = k.slices;
    return AX_OK;
}

ax_status ax_server_release(ax_server *s, int64_t ticket) {
    AX_REQUIRE(s != nullptr, "ax_server_release: null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    AX_REQUIRE(it != s->tickets.end(), "ax_server_release: unknown ticket %lld", (long long)ticket);
    AX_REQUIRE(it->second.done, "ax_server_release: ticket %lld is not complete",
               (long long)ticket);
    LaunchSet *ls = it->second.set;
    if (ls && --ls->refs == 0) {
        s->pinned_release_locked(ls->results, ls->result_bytes);
        delete ls;
    }
    s->tickets.erase(it);
    return AX_OK;
}

ax_status ax_server_stats(ax_server *s, int64_t *out, int n) {
    AX_REQUIRE(s && out && n >= 0, "ax_server_stats: null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    int64_t resident = 0;
    for (auto &kv : s->states)
        if (kv.second->plan) ++resident;
    const int64_t v[10] = {s->launches,     s->plans_created, s->peak_bytes, s->ns_setup,
                           s->ns_pack_wait, s->ns_alloc,      s->ns_queue,   s->plans_made,
                           resident,        s->state_bytes};
    for (int i = 0; i < n && i < 10; ++i) out[i] = v[i];
    return AX_OK;
}

ax_status ax_server_drop_state(ax_server *s, int64_t state) {
    AX_REQUIRE(s != nullptr, "ax_server_drop_state: null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->states.find(state);
    if (it == s->states.end()) return AX_OK;
    State *st = it->second.get();
    st->drop = true;
    if (!st->busy && st->queued == 0) s->free_state_plan_locked(st);  // erases it
    return AX_OK;
}

}  // extern "C"
