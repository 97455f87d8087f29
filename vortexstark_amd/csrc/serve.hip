// Proof server: continuous batching of Fibonacci proof jobs of mixed sizes on
// one GPU, admitted against an HBM budget (Part 5 of the C ABI).
//
// A job is n_cols Fibonacci columns of 2^log_n rows; each column becomes one
// KRAK v1 proof, bit-identical to proving it alone (every proof of a batch
// owns its arena, so nothing is shared between proofs but the twiddles).
//
// Scheduling, all on the caller's thread (vx_server_poll is one step):
//  * every size class (log_n) keeps a FIFO of pending proofs and up to
//    MAX_PROVERS_PER_CLASS batched provers, each on its own stream;
//  * an idle prover takes up to its capacity of pending proofs at once, so
//    requests that arrive while a batch runs join the next batch of their
//    class, and batches of different classes overlap on the device;
//  * a prover is created when its class has work and its arena fits the HBM
//    budget (or ahead of traffic by vx_server_prewarm); idle provers of
//    classes without work are released (least recently used first) to make
//    room;
//  * finished proofs are written by the device into one region of a pinned
//    host ring, where they stay until the job is released; a batch waits while
//    the ring has no room for it;
//  * a finished batch is noticed through an event query on its stream;
//  * a failure is confined to the jobs it touches: their status turns
//    negative and every other job still completes (vx_server_wait_all returns
//    once all jobs are finished, -1 if a job failed since the previous wait).
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "common.hpp"
#include "result_ring.hpp"

namespace {

constexpr uint32_t MAX_PROVERS_PER_CLASS = 2;
constexpr uint32_t M31_P = 0x7FFFFFFFu;
constexpr uint64_t DEFAULT_RING_BYTES = 1ull << 30;

double now_s() {
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// proofs per batch: about 2^26 LDE rows per launch (a batch then spends most
// of its time in stage kernels of thousands of waves, and the fixed cost of a
// batch's narrow tree levels and FRI tail is shared by many proofs), at least
// 2 and at most 1024
uint32_t class_capacity(uint32_t log_n) {
    if (log_n >= 23) return 2;
    return std::min<uint32_t>(1024, std::max<uint32_t>(2, 1u << (24 - log_n)));
}

struct Job {
    uint32_t log_n = 0, n_cols = 0, remaining = 0;
    double arrival = 0, done = -1;
    int32_t error = 0;
    bool released = false;
    std::vector<uint8_t> inject;  // per column: report this proof failed (vx_server_inject_fault)
    uint64_t proof_size = 0;
    std::vector<int64_t> region;  // per column: ring region holding its proof
    std::vector<uint8_t*> where;  // per column: address of its proof in the ring
};

struct Item {
    int64_t job;
    uint32_t col;
    uint32_t a, b;
};

struct Slot {
    vx_prover* prover = nullptr;
    uint32_t capacity = 0;
    uint64_t bytes = 0;
    uint64_t proof_size = 0;
    hipEvent_t done = nullptr;
    uint32_t* h_failed = nullptr;  // pinned: capacity words
    int64_t region = -1;
    std::vector<Item> running;
    double last_used = 0;
    bool big = false;  // the running batch counts against the server's big-batch cap
    bool busy() const { return !running.empty(); }
};

struct Class {
    std::deque<Item> pending;
    std::vector<std::unique_ptr<Slot>> slots;
};

}  // namespace

namespace vx {
// Raises the transcript failure flag of proof `index` of the prover's last
// launch sequence, on the prover's stream (defined with the prover).
bool prover_flag_failed(vx_prover* p, uint32_t index);
// The event recorded after the heavy part (LDE, trace, quotient and circle-fold
// trees) of the prover's latest launch sequence (defined with the prover).
hipEvent_t prover_heavy_done(vx_prover* p);
}  // namespace vx

struct vx_server {
    uint64_t budget = 0, used = 0;
    uint32_t n_q = 80, log_last = 3;
    bool graphs = false;  // batches replay captured graphs
    bool copy_kernel = false;  // proofs reach the ring through a kernel instead of the DMA engines
    // Launch order within a scheduling step. Smallest first (the default): a
    // small class's batch is short and latency-bound, so it is dispatched ahead
    // of the long batches that share its hardware queue instead of waiting for
    // them to finish.
    bool smallest_first = true;
    // Batches of at least 2^big_log_rows rows (proofs x rows) in flight at once;
    // 0: no limit. Large batches are VALU-bound on their own; run side by side
    // they slow each other down more than they fill each other's idle time.
    // The same threshold marks the batches that are staggered (below).
    uint32_t big_cap = 0, big_log_rows = 23, big_running = 0;
    bool is_big(const Class& c, const Slot& s, uint32_t log_n) const {
        const uint64_t count = std::min<uint64_t>(s.capacity, c.pending.size());
        return big_cap && (count << log_n) >= (1ull << big_log_rows);
    }
    bool may_launch(const Class& c, const Slot& s, uint32_t log_n) const {
        return !is_big(c, s, log_n) || big_running < big_cap;
    }
    // Staggered big batches (off by default): the stream of a batch of at least
    // 2^big_log_rows rows first waits, on the device, for the heavy part of the
    // big batch launched before it. Their VALU-bound stage kernels then run one
    // batch after another at full speed, and consecutive big batches overlap
    // only where one of them is latency- or PCIe-bound (its FRI line folds,
    // queries, decommitment and proof copy), instead of all of them reaching
    // those phases together at the end of a pass. Waits follow launch order,
    // so they cannot form a cycle. This pays when a whole job list is submitted
    // at once (the big batches would otherwise finish together); under a
    // stream of arrivals, where big batches start at different times anyway,
    // the waits only delay them.
    bool stagger = false;
    hipEvent_t last_heavy = nullptr;  // owned by the prover of the latest big batch
    vx_prover* last_heavy_prover = nullptr;
    vx::ResultRing ring;
    std::map<uint32_t, Class> classes;
    std::vector<Job> jobs;
    uint64_t completed = 0, batches = 0, proofs_done = 0, peak_used = 0;
    uint64_t failed_jobs = 0;
    // First job failure since the last wait: reported to the waiting caller
    // through the thread's error slot, which the server keeps clear meanwhile
    // (a recorded error would make every later launch look failed).
    int32_t fail_code = 0;
    std::string fail_msg;
    uint64_t failed_reported = 0;
    void note_failure(int32_t code, const char* what) {
        std::string cause;
        const int e = vx::take_error(cause);
        if (!fail_code) {
            fail_code = code;
            fail_msg = what;
            if (e && !cause.empty()) fail_msg += " (" + cause + ")";
        }
    }
    // 0 when no job failed since the previous call, else -1 with the first
    // failure recorded as the thread's error
    int32_t report_failures() {
        if (failed_jobs == failed_reported) return 0;
        failed_reported = failed_jobs;
        vx::put_error(fail_code ? fail_code : -1, fail_msg.empty() ? "vx_server: a job failed" : fail_msg);
        fail_code = 0;
        fail_msg.clear();
        return -1;
    }

    ~vx_server() {
        for (auto& kv : classes)
            for (auto& s : kv.second.slots) release(*s);
        if (ring.base) (void)hipHostFree(ring.base);
    }
    void release(Slot& s) {
        if (s.prover) {
            if (s.prover == last_heavy_prover) {
                last_heavy = nullptr;
                last_heavy_prover = nullptr;
            }
            (void)hipStreamSynchronize((hipStream_t)vx_prover_stream(s.prover));
            vx_prover_destroy(s.prover);
            used -= s.bytes;
        }
        if (s.done) (void)hipEventDestroy(s.done);
        if (s.h_failed) (void)hipHostFree(s.h_failed);
        s = Slot();
    }
    bool make_slot(uint32_t log_n, Slot& s) {
        const uint32_t cap = class_capacity(log_n);
        vx_prover* p = vx_prover_create_batch(log_n, n_q, log_last, cap);
        if (!p) return false;
        s.prover = p;
        s.capacity = cap;
        if (graphs) vx_prover_set_graph(p, 1);
        if (copy_kernel) vx_prover_set_copy_kernel(p, 1);
        s.bytes = vx_prover_device_bytes(p);
        s.proof_size = vx_prover_proof_size(p);
        used += s.bytes;
        peak_used = std::max(peak_used, used);
        bool ok = vx::check(hipEventCreateWithFlags(&s.done, hipEventDisableTiming), "hipEventCreate") &&
                  vx::check(hipHostMalloc((void**)&s.h_failed, 4ull * cap), "hipHostMalloc");
        if (!ok) release(s);
        return ok;
    }
    // device bytes of a prover of this class, known before creating it
    uint64_t estimate_bytes(uint32_t log_n) const {
        return vx_prover_estimate_bytes(log_n, n_q, log_last, class_capacity(log_n));
    }
    // Releases the idle prover used least recently in a class other than
    // for_class that has no pending work; false when there is none.
    bool release_idle(uint32_t for_class) {
        Slot* victim = nullptr;
        uint32_t victim_class = 0;
        for (auto& kv : classes) {
            if (kv.first == for_class || !kv.second.pending.empty()) continue;
            for (auto& s : kv.second.slots)
                if (s->prover && !s->busy() && (!victim || s->last_used < victim->last_used)) {
                    victim = s.get();
                    victim_class = kv.first;
                }
        }
        if (!victim) return false;
        release(*victim);
        auto& v = classes[victim_class].slots;
        v.erase(std::remove_if(v.begin(), v.end(), [](const std::unique_ptr<Slot>& x) { return !x->prover; }),
                v.end());
        return true;
    }
    bool reserve_room(uint64_t need, uint32_t for_class) {
        while (used + need > budget)
            if (!release_idle(for_class)) return false;
        return true;
    }
    // Launches up to capacity pending proofs of the class on an idle prover.
    // Returns false when the result ring has no room for even one proof.
    bool launch(Class& c, Slot& s, uint32_t log_n) {
        const bool big = is_big(c, s, log_n);
        uint32_t count = (uint32_t)std::min<size_t>(s.capacity, c.pending.size());
        int64_t region = -1;
        while (count && (region = ring.alloc(count * s.proof_size, count)) < 0) count /= 2;
        if (region < 0) return false;
        std::vector<uint32_t> ab(2 * count);
        s.running.clear();
        s.region = region;
        s.big = big;
        if (big) big_running++;
        for (uint32_t i = 0; i < count; i++) {
            s.running.push_back(c.pending.front());
            c.pending.pop_front();
            ab[2 * i] = s.running[i].a;
            ab[2 * i + 1] = s.running[i].b;
        }
        hipStream_t st = (hipStream_t)vx_prover_stream(s.prover);
        const bool heavy = stagger && ((uint64_t)count << log_n) >= (1ull << big_log_rows);
        bool ok = true;
        if (heavy && last_heavy) ok = vx::check(hipStreamWaitEvent(st, last_heavy, 0), "hipStreamWaitEvent");
        ok = ok && vx_prove_fibonacci_batch(s.prover, ab.data(), count) == 0;
        if (heavy && ok) {
            last_heavy = vx::prover_heavy_done(s.prover);
            last_heavy_prover = s.prover;
        }
        // an injected fault takes the path of a device-side failure: the proof's
        // transcript flag is raised after its proving and read back with the proofs
        for (uint32_t i = 0; ok && i < count; i++) {
            const Job& j = jobs[s.running[i].job];
            if (j.inject[s.running[i].col]) ok = vx::prover_flag_failed(s.prover, i);
        }
        // proofs (into the ring) and transcript flags, on the same stream
        ok = ok && vx_prover_enqueue_copy(s.prover, count, ring.at(region), s.h_failed, st) == 0;
        ok = ok && vx::check(hipEventRecord(s.done, st), "hipEventRecord");
        s.last_used = now_s();
        batches++;
        if (!ok) {
            // work already queued may still write the ring region: let it drain first
            (void)hipStreamSynchronize(st);
            note_failure(-1, "vx_server: a batch could not be launched");
            fail_running(s, -1);
        }
        return true;
    }
    void end_batch(Slot& s) {
        if (s.big) big_running--;
        s.big = false;
    }
    void fail_running(Slot& s, int32_t code) {
        end_batch(s);
        for (const Item& it : s.running) {
            Job& j = jobs[it.job];
            j.error = code;
            ring.release(s.region);
            if (--j.remaining == 0) finish(j);
        }
        s.running.clear();
        s.region = -1;
    }
    void finish(Job& j) {
        j.done = now_s();
        completed++;
        if (j.error) failed_jobs++;
    }
    void retire(Slot& s) {
        end_batch(s);
        uint8_t* base = ring.at(s.region);
        for (uint32_t i = 0; i < s.running.size(); i++) {
            const Item& it = s.running[i];
            Job& j = jobs[it.job];
            j.region[it.col] = s.region;
            j.where[it.col] = base + (uint64_t)i * s.proof_size;
            if (s.h_failed[i]) {
                j.error = -3;
                note_failure(-3, "vx_server: a proof's device-side failure flag was raised");
            }
            proofs_done++;
            if (--j.remaining == 0) finish(j);
        }
        s.running.clear();
        s.region = -1;
        s.last_used = now_s();
    }
    bool any_busy() const {
        for (auto& kv : classes)
            for (auto& s : kv.second.slots)
                if (s->busy()) return true;
        return false;
    }
    void fail_pending(Class& c, int32_t code) {
        for (const Item& it : c.pending) {
            Job& j = jobs[it.job];
            j.error = code;
            if (--j.remaining == 0) finish(j);
        }
        c.pending.clear();
    }
    // One scheduler step; returns the number of batches launched or retired.
    // An error the caller had recorded before the step is set aside during it
    // and restored afterwards.
    int32_t poll() {
        std::string held_msg;
        const int held = vx::take_error(held_msg);
        const int32_t acted = poll_step();
        if (held) {
            std::string ignore;
            (void)vx::take_error(ignore);
            vx::put_error(held, held_msg);
        }
        return acted;
    }
    int32_t poll_step() {
        int32_t acted = 0;
        for (auto& kv : classes) {
            for (auto& s : kv.second.slots) {
                if (!s->busy()) continue;
                hipError_t q = hipEventQuery(s->done);
                if (q == hipSuccess) {
                    retire(*s);
                    acted++;
                } else if (q != hipErrorNotReady) {
                    vx::check(q, "vx_server: batch failed");
                    note_failure((int32_t)q, "vx_server: batch failed");
                    fail_running(*s, (int32_t)q);
                    acted++;
                }
            }
        }
        bool ring_full = false;
        // classes in the server's order (smallest or largest proofs first)
        std::vector<std::pair<const uint32_t, Class>*> order;
        for (auto& kv : classes) order.push_back(&kv);
        if (!smallest_first) std::reverse(order.begin(), order.end());
        for (auto* it : order) {
            auto& kv = *it;
            Class& c = kv.second;
            if (c.pending.empty()) continue;
            bool held = false;  // a big batch of this class waits for the big-batch cap
            for (auto& s : c.slots)
                if (!s->busy() && !c.pending.empty()) {
                    if (!may_launch(c, *s, kv.first)) {
                        held = true;
                        continue;
                    }
                    if (launch(c, *s, kv.first))
                        acted++;
                    else
                        ring_full = true;
                }
            // more proofs waiting than the idle provers take: add a prover if the budget allows
            if (!c.pending.empty() && !ring_full && !held && c.slots.size() < MAX_PROVERS_PER_CLASS) {
                const uint64_t need = estimate_bytes(kv.first);
                if (reserve_room(need, kv.first)) {
                    auto s = std::make_unique<Slot>();
                    if (make_slot(kv.first, *s)) {
                        c.slots.push_back(std::move(s));
                        if (!may_launch(c, *c.slots.back(), kv.first))
                            ;  // made ahead: it takes the class's next batch once the cap allows
                        else if (launch(c, *c.slots.back(), kv.first))
                            acted++;
                        else
                            ring_full = true;
                    } else if (c.slots.empty() && !any_busy() && !release_idle(kv.first)) {
                        // the prover cannot be made and nothing running or idle can
                        // free memory for it: waiting cannot help
                        note_failure(-6, "vx_server: could not create a prover for this size class");
                        fail_pending(c, -6);
                        acted++;
                    } else {
                        // out of device memory now: retry once a batch finishes or
                        // an idle prover has been released
                        (void)vx_get_last_error();
                        acted += c.slots.empty() && !any_busy() ? 1 : 0;  // released one: poll again at once
                    }
                } else if (c.slots.empty() && !any_busy()) {
                    // nothing will free memory: this class can never run under the budget
                    note_failure(-4, "vx_server: a proof of this size does not fit the HBM budget");
                    fail_pending(c, -4);
                    acted++;
                }
            }
        }
        if (ring_full && !acted && !any_busy()) {
            // nothing runs and no finished job has been released: waiting cannot help
            note_failure(-5, "vx_server: result ring full; release finished jobs");
            for (auto& kv : classes) fail_pending(kv.second, -5);
        }
        return acted;
    }
};

extern "C" {

vx_server* vx_server_create(uint64_t budget_bytes, uint64_t result_bytes, uint32_t n_queries,
                            uint32_t log_last_layer) {
    // the smallest trace a last layer of 2^log_last_layer allows must make a
    // prover: otherwise no job could ever be served
    const uint32_t log_last = log_last_layer ? log_last_layer : 3;
    if (vx_prover_estimate_bytes(std::max<uint32_t>(2, log_last - 1), n_queries, log_last_layer, 1) == 0) {
        vx::set_error(-2, "vx_server_create: unsupported n_queries or log_last_layer");
        return nullptr;
    }
    size_t free_b = 0, total_b = 0;
    if (!vx::check(hipMemGetInfo(&free_b, &total_b), "hipMemGetInfo")) return nullptr;
    vx_server* s = new vx_server();
    const uint64_t avail = (uint64_t)(0.9 * (double)free_b);
    s->budget = budget_bytes ? std::min<uint64_t>(budget_bytes, avail) : avail;
    s->n_q = n_queries ? n_queries : 80;
    s->log_last = log_last_layer ? log_last_layer : 3;
    s->ring.size = result_bytes ? result_bytes : DEFAULT_RING_BYTES;
    if (!vx::check(hipHostMalloc((void**)&s->ring.base, s->ring.size), "vx_server: pinned result ring")) {
        delete s;
        return nullptr;
    }
    return s;
}

void vx_server_destroy(vx_server* s) { delete s; }

int64_t vx_server_submit(vx_server* s, uint32_t log_n, uint32_t n_cols, const uint32_t* ab) {
    if (!s || !ab || n_cols == 0 || s->estimate_bytes(log_n) == 0) {
        vx::set_error(-2, "vx_server_submit: bad job");
        return -1;
    }
    const int64_t id = (int64_t)s->jobs.size();
    Job j;
    j.log_n = log_n;
    j.n_cols = n_cols;
    j.remaining = n_cols;
    j.arrival = now_s();
    j.proof_size = vx_krak_proof_size(log_n, s->n_q, s->log_last);
    j.region.assign(n_cols, -1);
    j.where.assign(n_cols, nullptr);
    j.inject.assign(n_cols, 0);
    s->jobs.push_back(std::move(j));
    Class& c = s->classes[log_n];
    for (uint32_t k = 0; k < n_cols; k++) c.pending.push_back(Item{id, k, ab[2 * k] % M31_P, ab[2 * k + 1] % M31_P});
    return id;
}

int32_t vx_server_poll(vx_server* s) { return s ? s->poll() : -1; }

int32_t vx_server_prewarm(vx_server* s, uint32_t log_n, uint32_t n_provers) {
    if (!s || s->estimate_bytes(log_n) == 0) {
        vx::set_error(-2, "vx_server_prewarm: bad size class");
        return -1;
    }
    Class& c = s->classes[log_n];
    const uint32_t want = std::min<uint32_t>(n_provers, MAX_PROVERS_PER_CLASS);
    int32_t made = 0;
    while (c.slots.size() < want) {
        if (!s->reserve_room(s->estimate_bytes(log_n), log_n)) break;  // the budget is full
        auto sl = std::make_unique<Slot>();
        if (!s->make_slot(log_n, *sl)) return -1;
        sl->last_used = now_s();
        c.slots.push_back(std::move(sl));
        made++;
    }
    return made;
}

int32_t vx_server_set_order(vx_server* s, int32_t smallest_first) {
    if (!s) return -1;
    s->smallest_first = smallest_first != 0;
    return 0;
}

int32_t vx_server_set_big_batches(vx_server* s, uint32_t max_in_flight, uint32_t log_rows) {
    if (!s || log_rows > 40) {
        vx::set_error(-2, "vx_server_set_big_batches: bad arguments");
        return -1;
    }
    s->big_cap = max_in_flight;
    s->big_log_rows = log_rows;
    return 0;
}

int32_t vx_server_set_stagger(vx_server* s, int32_t enable) {
    if (!s) {
        vx::set_error(-2, "vx_server_set_stagger: null server");
        return -1;
    }
    s->stagger = enable != 0;
    if (!s->stagger) {
        s->last_heavy = nullptr;
        s->last_heavy_prover = nullptr;
    }
    return 0;
}

int32_t vx_server_set_copy_kernel(vx_server* s, int32_t enable) {
    if (!s) {
        vx::set_error(-2, "vx_server_set_copy_kernel: null server");
        return -1;
    }
    s->copy_kernel = enable != 0;
    for (auto& kv : s->classes)
        for (auto& sl : kv.second.slots)
            if (sl->prover && vx_prover_set_copy_kernel(sl->prover, enable)) return -1;
    return 0;
}

int32_t vx_server_set_graphs(vx_server* s, int32_t enable) {
    if (!s) return -1;
    s->graphs = enable != 0;
    for (auto& kv : s->classes)
        for (auto& sl : kv.second.slots)
            if (sl->prover && vx_prover_set_graph(sl->prover, enable)) return -1;
    return 0;
}

uint64_t vx_server_completed(const vx_server* s) { return s ? s->completed : 0; }
uint64_t vx_server_jobs(const vx_server* s) { return s ? s->jobs.size() : 0; }

int32_t vx_server_wait_all(vx_server* s) {
    if (!s) return -1;
    // every pass either acts or has a batch in flight: poll fails the jobs of a
    // class that can make no progress, so this loop ends
    while (s->completed < s->jobs.size())
        if (s->poll() == 0) std::this_thread::sleep_for(std::chrono::microseconds(20));
    return s->report_failures();
}

int32_t vx_server_inject_fault(vx_server* s, int64_t job, uint32_t col) {
    if (!s || job < 0 || job >= (int64_t)s->jobs.size() || col >= s->jobs[job].n_cols) {
        vx::set_error(-2, "vx_server_inject_fault: no such job column");
        return -1;
    }
    Job& j = s->jobs[job];
    for (const auto& kv : s->classes)
        for (const auto& sl : kv.second.slots)
            for (const Item& it : sl->running)
                if (it.job == job && it.col == col) {
                    vx::set_error(-2, "vx_server_inject_fault: that proof is already on the device");
                    return -1;
                }
    if (j.done >= 0) {
        vx::set_error(-2, "vx_server_inject_fault: the job has finished");
        return -1;
    }
    j.inject[col] = 1;
    return 0;
}

void vx_server_release_job(vx_server* s, int64_t job) {
    if (!s || job < 0 || job >= (int64_t)s->jobs.size()) return;
    Job& j = s->jobs[job];
    if (j.done < 0 || j.released) return;
    j.released = true;
    for (int64_t r : j.region)
        if (r >= 0) s->ring.release(r);
    j.where.assign(j.n_cols, nullptr);
}

int32_t vx_server_job_status(const vx_server* s, int64_t job) {
    if (!s || job < 0 || job >= (int64_t)s->jobs.size()) return -2;
    const Job& j = s->jobs[job];
    if (j.error) return j.error;
    return j.done >= 0 ? 1 : 0;
}

double vx_server_job_latency_ms(const vx_server* s, int64_t job) {
    if (!s || job < 0 || job >= (int64_t)s->jobs.size() || s->jobs[job].done < 0) return -1.0;
    return 1e3 * (s->jobs[job].done - s->jobs[job].arrival);
}

uint64_t vx_server_proof_copy(const vx_server* s, int64_t job, uint32_t col, void* out, uint64_t cap) {
    if (!s || job < 0 || job >= (int64_t)s->jobs.size()) return 0;
    const Job& j = s->jobs[job];
    if (j.done < 0 || j.error || j.released || col >= j.n_cols || cap < j.proof_size || !out || !j.where[col])
        return 0;
    std::memcpy(out, j.where[col], j.proof_size);
    return j.proof_size;
}

int32_t vx_server_run_trace(vx_server* s, uint32_t n_jobs, const double* arrival_s, const uint32_t* log_n,
                            const uint32_t* n_cols, const uint32_t* ab, double* latency_ms) {
    if (!s || (n_jobs && (!arrival_s || !log_n || !n_cols || !ab))) {
        vx::set_error(-2, "vx_server_run_trace: bad arguments");
        return -1;
    }
    const double t0 = now_s();
    const int64_t first = (int64_t)s->jobs.size();
    uint32_t next = 0;
    uint64_t ab_pos = 0;
    while (next < n_jobs || s->completed < s->jobs.size()) {
        const double t = now_s() - t0;
        while (next < n_jobs && arrival_s[next] <= t) {
            if (vx_server_submit(s, log_n[next], n_cols[next], ab + 2 * ab_pos) < 0) return -1;
            s->jobs.back().arrival = t0 + arrival_s[next];
            ab_pos += n_cols[next];
            next++;
        }
        if (s->poll() == 0) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
    if (latency_ms)
        for (uint32_t i = 0; i < n_jobs; i++) latency_ms[i] = vx_server_job_latency_ms(s, first + i);
    return s->report_failures();
}

void vx_server_stats(const vx_server* s, uint64_t out[6]) {
    if (!s || !out) return;
    out[0] = s->batches;
    out[1] = s->proofs_done;
    out[2] = s->completed;
    out[3] = s->used;
    out[4] = s->peak_used;
    out[5] = s->budget;
}

}  // extern "C"
