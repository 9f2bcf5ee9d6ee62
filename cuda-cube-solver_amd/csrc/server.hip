// Serving loop: continuous batching of cube solves on a resident table.
//
// One persistent search kernel runs for the life of the server.  Searches
// live in device "slots" (SearchJob); the number of slots is set by the
// device-memory budget for per-search state.  A host pump thread moves
// submitted cubes into free slots through a mailbox in pinned host memory
// (the kernel's idle lanes admit them: root bounds, first round, ring
// entry) and collects results from a pinned completion ring that the kernel
// fills when a finished cube's last in-flight work item has drained, which
// is also when its slot is free again.  Cubes that arrive while others are
// searched start at once if a slot is free; all of them share the GPU.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <vector>

#include "capi_internal.hpp"
#include "kernels.hpp"

using namespace ccs;

namespace {

struct Pending {
    Cubie cube;
    uint64_t tag;
    double t_submit;
};

struct Admitted {
    Cubie cube;
    uint64_t tag;
    double t_submit;
    double t_admit;
};

}  // namespace

struct ccs_server {
    ccs_tables* t = nullptr;
    int n_slots = 0;
    hipStream_t stream = nullptr;
    SearchJob* d_jobs = nullptr;
    QueueCtl* d_q = nullptr;
    unsigned long long* d_ring = nullptr;
    unsigned long long* d_stats = nullptr;
    SplitRec* d_splits = nullptr;
    ServeRec* mb = nullptr;                 // pinned
    unsigned long long* mb_tail = nullptr;  // pinned
    ServeDone* done = nullptr;              // pinned
    unsigned int* host_stop = nullptr;      // pinned
    unsigned long long mb_size = 0, done_size = 0;
    unsigned long long mb_next = 0, done_head = 0;
    std::vector<int> free_slots;
    std::vector<Admitted> slot_info;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Pending> pending;
    std::deque<ccs_served_t> results;
    std::atomic<bool> quit{false};
    std::thread pump;
    std::string error;

    // Frees whatever was set up; the serving kernel must have ended (or never started).
    ~ccs_server() {
        if (stream) (void)hipStreamDestroy(stream);
        for (void* p : {(void*)d_jobs, (void*)d_q, (void*)d_ring, (void*)d_stats, (void*)d_splits})
            if (p) (void)hipFree(p);
        for (void* p : {(void*)mb, (void*)mb_tail, (void*)done, (void*)host_stop})
            if (p) (void)hipHostFree(p);
    }
};

namespace {

void pump_loop(ccs_server* s) {
    while (!s->quit.load()) {
        bool busy = false;
        // results: the kernel publishes them in order of completion-record index
        for (;;) {
            ServeDone& d = s->done[s->done_head % s->done_size];
            if (__atomic_load_n(&d.seq, __ATOMIC_ACQUIRE) != s->done_head + 1) break;
            busy = true;
            int slot = (int)d.slot;
            ccs_served_t r;
            memset(&r, 0, sizeof r);
            r.tag = d.tag;
            r.length = d.length;
            r.nodes = d.nodes;
            Cubie c{};
            bool ok = slot >= 0 && slot < s->n_slots;
            if (ok) {
                const Admitted& a = s->slot_info[slot];
                c = a.cube;
                r.queued_ms = a.t_admit - a.t_submit;
                r.latency_ms = now_ms() - a.t_submit;
            }
            for (int l = 0; l < r.length && l < 24; ++l) {
                r.moves[l] = d.moves[l];
                c = then(c, move(d.moves[l]));
            }
            Cubie sv = solved();
            if (ok && r.length >= 0 && memcmp(&c, &sv, sizeof c) != 0) r.length = -2;  // failed verification
            ++s->done_head;
            std::lock_guard<std::mutex> g(s->mu);
            s->results.push_back(r);
            if (ok) s->free_slots.push_back(slot);
            s->cv.notify_all();
        }
        // admissions: pending cubes into free slots
        for (;;) {
            Pending p;
            int slot;
            {
                std::lock_guard<std::mutex> g(s->mu);
                // at most one unconsumed mailbox record per slot, and mb_size > n_slots
                if (s->pending.empty() || s->free_slots.empty()) break;
                p = s->pending.front();
                s->pending.pop_front();
                slot = s->free_slots.back();
                s->free_slots.pop_back();
            }
            busy = true;
            Cubie f[6];
            axis_frames(p.cube, f);
            ServeRec& rec = s->mb[s->mb_next % s->mb_size];
            for (int j = 0; j < 6; ++j) {
                Packed pk = pack(f[j]);
                rec.frame[j] = Frame{pk.w[0], pk.w[1], pk.w[2], pk.w[3], pk.w[4]};
            }
            rec.tag = p.tag;
            rec.slot = (uint32_t)slot;
            s->slot_info[slot] = Admitted{p.cube, p.tag, p.t_submit, now_ms()};
            __atomic_store_n(&rec.seq, s->mb_next + 1, __ATOMIC_RELEASE);
            ++s->mb_next;
            __atomic_store_n(s->mb_tail, s->mb_next, __ATOMIC_RELEASE);
        }
        if (!busy) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
}

}  // namespace

extern "C" {

int ccs_server_start(ccs_tables_t* t, const ccs_server_opts_t* opts, ccs_server_t** out) {
    if (!t || !out) return fail(CCS_EINVAL, "null argument");
    if (t->format < 0 || !t->d_table[t->format]) return fail(CCS_EINVAL, "no solving table resident");
    ccs_server_opts_t o;
    memset(&o, 0, sizeof o);
    if (opts) o = *opts;
    if (o.prefix_len == 0) o.prefix_len = 4;
    if (o.prefix_len < 4 || o.prefix_len > MAX_PLEN) return fail(CCS_EINVAL, "prefix_len must be 4..7");
    // Longer prefixes sooner than a batch uses them: an item runs on one lane, and
    // a lane of a shared GPU covers ~50k nodes/s, so a round's longest item sets its
    // latency (p99 at 50% load: 0.80 s against 1.05 s with the batch's 5e7, 3e9).
    if (o.deeper_at[0] == 0 && o.deeper_at[1] == 0) o.deeper_at[0] = 1e6, o.deeper_at[1] = 1e8;
    // Admission by device memory: each search holds one SearchJob and four ring
    // entries.  (Servers keep no item keys: the banded round order, which they
    // need, cost more in visits than it saved at serving's cube counts.)
    for (int p = 0; p <= std::min(o.prefix_len + 2, MAX_PLEN); ++p) {
        int rc = ensure_prefixes(t, p);
        if (rc) return rc;
    }
    const uint64_t per_search = sizeof(SearchJob) + 4 * sizeof(unsigned long long);
    uint64_t budget = o.state_budget_bytes ? o.state_budget_bytes : (64ull << 20);
    int n_slots = (int)std::min<uint64_t>(budget / per_search, 1u << 20);
    if (o.max_inflight > 0) n_slots = std::min(n_slots, o.max_inflight);
    if (n_slots < 1) return fail(CCS_ENOMEM, "state budget holds no search");
    std::unique_ptr<ccs_server> s(new ccs_server());
    s->t = t;
    s->n_slots = n_slots;
    s->mb_size = (unsigned long long)n_slots + 64;
    s->done_size = 2ull * n_slots + 64;
    const unsigned long long ring_size = 4ull * n_slots + 1024;
    HIP_TRY(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
    HIP_TRY(hipMalloc(&s->d_jobs, sizeof(SearchJob) * (size_t)n_slots));
    HIP_TRY(hipMalloc(&s->d_q, sizeof(QueueCtl)));
    HIP_TRY(hipMalloc(&s->d_ring, sizeof(unsigned long long) * ring_size));
    HIP_TRY(hipMalloc(&s->d_stats, STAT_WORDS * sizeof(unsigned long long)));
    HIP_TRY(hipMalloc(&s->d_splits, sizeof(SplitRec) * SPLIT_RING));
    HIP_TRY(hipMemset(s->d_splits, 0, sizeof(SplitRec) * SPLIT_RING));
    const unsigned flags = hipHostMallocCoherent | hipHostMallocMapped;
    HIP_TRY(hipHostMalloc(&s->mb, sizeof(ServeRec) * s->mb_size, flags));
    HIP_TRY(hipHostMalloc(&s->mb_tail, 64, flags));
    HIP_TRY(hipHostMalloc(&s->done, sizeof(ServeDone) * s->done_size, flags));
    HIP_TRY(hipHostMalloc(&s->host_stop, 64, flags));
    memset(s->mb, 0, sizeof(ServeRec) * s->mb_size);
    memset(s->done, 0, sizeof(ServeDone) * s->done_size);
    *s->mb_tail = 0;
    *s->host_stop = 0;
    // Every slot starts retired: found, no round, no item in flight.
    std::vector<SearchJob> jobs(n_slots);
    memset(jobs.data(), 0, sizeof(SearchJob) * jobs.size());
    for (auto& J : jobs) {
        J.found = 1;
        J.status = 1;
        J.ticket = 1ull << 58;
    }
    HIP_TRY(hipMemcpy(s->d_jobs, jobs.data(), sizeof(SearchJob) * jobs.size(), hipMemcpyHostToDevice));
    std::vector<unsigned long long> ring(ring_size, ~0ull);
    HIP_TRY(hipMemcpy(s->d_ring, ring.data(), sizeof(unsigned long long) * ring_size, hipMemcpyHostToDevice));
    QueueCtl q{};
    q.n_jobs = 0xFFFFFFFFu;
    HIP_TRY(hipMemcpy(s->d_q, &q, sizeof q, hipMemcpyHostToDevice));
    HIP_TRY(hipMemset(s->d_stats, 0, STAT_WORDS * sizeof(unsigned long long)));
    s->free_slots.reserve(n_slots);
    for (int i = n_slots - 1; i >= 0; --i) s->free_slots.push_back(i);
    s->slot_info.resize(n_slots);

    SearchArgs A;
    memset(&A, 0, sizeof A);
    A.lut = t->d_lut;
    A.class_lut = t->d_class_lut;
    A.classidx = t->d_classidx;
    const int SF = search_format(t, true);  // the search's half-block copy of a block table, made once
    A.table = search_table(t, SF);
    A.jobs = s->d_jobs;
    A.q = s->d_q;
    A.ring = s->d_ring;
    A.ring_size = ring_size;
    for (int p = 0; p <= MAX_PLEN; ++p) {
        A.pre_tab[p] = t->d_prefix[p];
        A.pre_n[p] = t->prefix_n[p] ? t->prefix_n[p] : 0xFFFFFFFFu;
    }
    A.prefix_len = o.prefix_len;
    A.max_bound = STACK_DEPTH;
    A.deeper0 = (unsigned long long)o.deeper_at[0];
    A.deeper1 = (unsigned long long)o.deeper_at[1];
    A.poll_mask = 63;  // found flag refreshed every 64 picks (read one pick later)
    A.chunk_nodes = o.chunk_nodes > 0 ? (unsigned long long)o.chunk_nodes : SERVE_CHUNK_NODES;
    // paired reads by default: under load the server holds as many cubes as a
    // large batch (served rate at 80% load 265 against 260 cubes/s in turn)
    A.read_mode = o.read_mode == 2 ? 1 : 0;
    A.idle_limit = ~0ull;  // idling is normal while serving
    {
        int rate_khz = 0, dev = 0;
        HIP_TRY(hipGetDevice(&dev));
        HIP_TRY(hipDeviceGetAttribute(&rate_khz, hipDeviceAttributeWallClockRate, dev));
        const unsigned long long per_us = (unsigned long long)std::max(rate_khz, 1000) / 1000ull;
        A.mb_poll_ticks = 2ull * per_us;   // one read of the host-side mailbox tail per 2 us
        A.age_ticks = 50000ull * per_us;   // one more fetch per turn for every 50 ms a cube has been in
        A.backoff_ticks = 20ull * per_us;
    }
    A.splits = s->d_splits;
    A.serving = 1;
    A.mb = s->mb;
    A.mb_tail = s->mb_tail;
    A.mb_size = s->mb_size;
    A.done = s->done;
    A.done_size = s->done_size;
    A.host_stop = s->host_stop;
    A.stats = s->d_stats;
    int grid = 0, per_cu = o.blocks_per_cu;
    if (per_cu <= 0) HIP_TRY(search_occupancy(t->K, SF, 2, SEARCH_WG, &per_cu));
    grid = std::max(per_cu, 1) * device_cus();
    A.n_lanes = (unsigned long long)grid * SEARCH_WG;
    HIP_TRY(launch_search(t->K, SF, 2, A, grid, SEARCH_WG, s->stream));
    s->pump = std::thread(pump_loop, s.get());
    *out = s.release();
    return CCS_OK;
}

int ccs_server_submit(ccs_server_t* s, const uint8_t* cube, uint64_t tag) {
    if (!s || !cube) return fail(CCS_EINVAL, "null argument");
    const Cubie& c = *reinterpret_cast<const Cubie*>(cube);
    if (!is_legal(c)) return fail(CCS_EINVAL, "not a legal cube");
    Cubie sv = solved();
    std::lock_guard<std::mutex> g(s->mu);
    if (memcmp(&c, &sv, sizeof c) == 0) {  // nothing to search
        ccs_served_t r;
        memset(&r, 0, sizeof r);
        r.tag = tag;
        s->results.push_back(r);
        s->cv.notify_all();
        return CCS_OK;
    }
    s->pending.push_back(Pending{c, tag, now_ms()});
    return CCS_OK;
}

int ccs_server_poll(ccs_server_t* s, ccs_served_t* out, int max, int timeout_ms) {
    if (!s || (max > 0 && !out)) return fail(CCS_EINVAL, "bad arguments");
    std::unique_lock<std::mutex> g(s->mu);
    if (s->results.empty() && timeout_ms > 0)
        s->cv.wait_for(g, std::chrono::milliseconds(timeout_ms), [&] { return !s->results.empty(); });
    int n = 0;
    while (n < max && !s->results.empty()) {
        out[n++] = s->results.front();
        s->results.pop_front();
    }
    return n;
}

int ccs_server_stop(ccs_server_t* s) {
    if (!s) return fail(CCS_EINVAL, "null argument");
    __atomic_store_n(s->host_stop, 1u, __ATOMIC_RELEASE);
    s->quit.store(true);
    if (s->pump.joinable()) s->pump.join();
    hipError_t e = hipStreamSynchronize(s->stream);
    delete s;  // frees the device and pinned memory
    if (e != hipSuccess) return fail(CCS_EDEVICE, std::string("serving kernel: ") + hipGetErrorString(e));
    return CCS_OK;
}

}  // extern "C"
