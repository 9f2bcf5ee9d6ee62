// ccs_solve: IDA* over a resident table.
//
// mode 0 (plain) keeps the reference's structure (solver.py:254-320): the host
// runs the rounds, launching one persistent grid per prefix-length group per
// round and synchronising in between.  mode 1 (persistent batched) hands
// every cube's rounds to the device: one launch, each cube advances its own
// bound as soon as its previous round drains, cubes of very different
// difficulty share the GPU without host round trips.
//
// ccs_solve_noalloc runs the persistent form with its device state in a
// static device array instead of hipMalloc'ed buffers; both share every
// kernel and every step below, templated on where device memory comes from.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <vector>

#include "capi_internal.hpp"
#include "kernels.hpp"

using namespace ccs;

namespace ccs {

namespace {

__global__ void read_clock_kernel(unsigned long long* out) { *out = wall_clock64(); }

// Fill n words with v (used instead of hipMemsetAsync, which a caller's graph
// capture must not record).
__global__ void fill_kernel(uint32_t* p, unsigned long long n, uint32_t v) {
    for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; i < n;
         i += (unsigned long long)gridDim.x * blockDim.x)
        p[i] = v;
}

hipError_t zero_async(void* p, size_t bytes, hipStream_t s) {
    const unsigned long long words = bytes / 4;  // every buffer cleared here is a whole number of words
    if (!words) return hipSuccess;
    const unsigned blocks = (unsigned)std::min<unsigned long long>((words + 255) / 256, 4096ull);
    hipLaunchKernelGGL(fill_kernel, dim3(blocks), dim3(256), 0, s, static_cast<uint32_t*>(p), words, 0u);
    return hipGetLastError();
}

// Device memory from a fixed region, 256-byte aligned; nothing is allocated.
struct Arena {
    static constexpr bool holds_keys = false;  // the keys of a banded order need 1.24 MB per cube
    uint8_t* base = nullptr;
    size_t cap = 0, used = 0;
    template <typename T>
    hipError_t alloc(T** p, size_t bytes) {
        const size_t at = (used + 255) & ~size_t(255);
        if (at + bytes > cap) {
            *p = nullptr;
            return hipErrorOutOfMemory;
        }
        *p = reinterpret_cast<T*>(base + at);
        used = at + bytes;
        return hipSuccess;
    }
    hipError_t event(hipEvent_t*) { return hipErrorNotSupported; }  // events are objects to create: none here
};

void fill_opts(const ccs_solve_opts_t* in, ccs_solve_opts_t& o) {
    memset(&o, 0, sizeof o);
    o.mode = 0;
    o.prefix_len = 4;
    o.deeper_at[0] = 5e7;
    o.deeper_at[1] = 3e9;
    if (in) o = *in;
    if (o.prefix_len == 0) o.prefix_len = 4;
    if (o.block == 0) o.block = SEARCH_WG;
    if (o.max_bound == 0) o.max_bound = MAX_BOUND;
    if (o.poll_interval == 0) o.poll_interval = 64;
    if (o.probes_per_step == 0) o.probes_per_step = 2;
    if (o.deeper_at[0] == 0 && o.deeper_at[1] == 0) o.deeper_at[0] = 5e7, o.deeper_at[1] = 3e9;
}

// Persistent batches from this size on read a step's two probes paired
// (read_mode 1): most of their steps probe twice.  Smaller ones, mostly lone
// and hard cubes whose steps are mostly a child's first, one-probe step, read
// in turn (2).  The plain form reads paired: its rounds keep under one busy
// wave per SIMD, where a step waits out its reads' latency, and paired reads
// wait once (45.8 against 42.7 cubes/s, one workload cube per call).
constexpr int PAIRED_READS_MIN_CUBES = 1024;

int check_opts(const ccs_solve_opts_t& o) {
    if (o.mode != 0 && o.mode != 1) return fail(CCS_EINVAL, "mode must be 0 (plain) or 1 (persistent)");
    if (o.prefix_len < 4 || o.prefix_len > MAX_PLEN) return fail(CCS_EINVAL, "prefix_len must be 4..7");
    if (o.max_bound < 1 || o.max_bound > MAX_BOUND) return fail(CCS_EINVAL, "max_bound must be 1..24");
    if (o.block < 64 || o.block > SEARCH_WG || o.block % 64)
        return fail(CCS_EINVAL, "block must be a multiple of 64 up to " + std::to_string(SEARCH_WG));
    if (o.probes_per_step != 1 && o.probes_per_step != 2) return fail(CCS_EINVAL, "probes_per_step must be 1 or 2");
    if (o.read_mode < 0 || o.read_mode > 2) return fail(CCS_EINVAL, "read_mode must be 0, 1 or 2");
    if (o.poll_interval & (o.poll_interval - 1)) return fail(CCS_EINVAL, "poll_interval must be a power of two");
    return CCS_OK;
}

unsigned long long to_u64(double x) { return x >= 1.8e19 ? ~0ull : (unsigned long long)x; }

int wall_clock_khz(int* khz) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(khz, hipDeviceAttributeWallClockRate, dev));
    return CCS_OK;
}

}  // namespace

// Grid of a persistent search launch.
int search_grid(ccs_tables* t, bool queue, int block, int blocks_per_cu, int* grid) {
    int per_cu = blocks_per_cu;
    if (per_cu <= 0) {
        HIP_TRY(search_occupancy(t->K, search_format(t, false), queue ? 1 : 0, block, &per_cu));
        if (per_cu <= 0) return fail(CCS_EDEVICE, "search kernel does not fit on a CU");
    }
    *grid = per_cu * device_cus();
    return CCS_OK;
}

template <typename Mem>
int root_setup(ccs_tables* t, const uint8_t* cubes, int n, std::vector<SearchJob>& jobs, std::vector<int>& solved0,
               Mem& mem) {
    jobs.assign(n, SearchJob{});
    std::vector<Packed> frames(6ull * n);
    for (int i = 0; i < n; ++i) {
        const Cubie& c = *reinterpret_cast<const Cubie*>(cubes + 40ull * i);
        if (!is_legal(c)) return fail(CCS_EINVAL, "cube " + std::to_string(i) + " is not a legal cube");
        Cubie f[6];
        axis_frames(c, f);
        for (int j = 0; j < 6; ++j) frames[6ull * i + j] = pack(f[j]);
        Cubie s = solved();
        solved0[i] = memcmp(&c, &s, sizeof c) == 0;
    }
    Frame* d_frames = nullptr;
    uint32_t* d_lb = nullptr;
    HIP_TRY(mem.alloc(&d_frames, sizeof(Frame) * 6ull * n));
    HIP_TRY(mem.alloc(&d_lb, 4ull * 6 * n));
    HIP_TRY(hipMemcpyAsync(d_frames, frames.data(), sizeof(Frame) * 6ull * n, hipMemcpyHostToDevice, t->stream));
    HIP_TRY(launch_probe(t->K, t->format, t->d_lut, t->d_classidx, t->d_table[t->format], d_frames, 6 * n, -1, nullptr,
                         d_lb, nullptr, t->stream));
    std::vector<uint32_t> lb(6ull * n);
    HIP_TRY(hipMemcpyAsync(lb.data(), d_lb, 4ull * 6 * n, hipMemcpyDeviceToHost, t->stream));
    HIP_TRY(hipStreamSynchronize(t->stream));
    for (int i = 0; i < n; ++i) {
        SearchJob& J = jobs[i];
        memset(&J, 0, sizeof J);
        uint32_t b = 0;
        for (int j = 0; j < 6; ++j) {
            const Packed& p = frames[6ull * i + j];
            J.frame[j] = Frame{p.w[0], p.w[1], p.w[2], p.w[3], p.w[4]};
            J.h[j] = lb[6ull * i + j];
            b = std::max(b, J.h[j]);
        }
        J.bound = b;  // the largest frame bound (solver.py:261)
    }
    return CCS_OK;
}

// The persistent batched form on n cubes whose jobs root_setup prepared.
// With events, the kernel time comes from HIP events; otherwise from the
// device wall clock read before and after the launch.
template <typename Mem>
int solve_persistent(ccs_tables* t, const ccs_solve_opts_t& o, SearchArgs& A, int grid, int n,
                     std::vector<SearchJob>& jobs, const std::vector<int>& solved0, Mem& mem, bool events,
                     double t_start, ccs_result_t* out, std::vector<double>& done_ms, double& kernel_ms,
                     uint32_t& launches) {
    // A cube whose root bound is already above the limit is given up without
    // a round, as the plain form does (its first round would search at that
    // bound and could return a solution longer than the limit).
    std::vector<int> live;
    for (int i = 0; i < n; ++i) {
        if (solved0[i]) {
            done_ms[i] = 0;
            continue;
        }
        if ((int)jobs[i].bound > A.max_bound) continue;
        live.push_back(i);
    }
    const int nl = (int)live.size();
    if (!nl) return CCS_OK;
    std::vector<SearchJob> lj(nl);
    for (int k = 0; k < nl; ++k) {
        lj[k] = jobs[live[k]];
        uint32_t b = lj[k].bound;
        uint32_t plen = std::min(std::min(b, (uint32_t)o.prefix_len), (uint32_t)MAX_PLEN);
        // first round: chunks of up to 64 items, and at least max(256, 1.5 x the
        // cube's share of the lanes) chunks per round, as the device's later rounds
        const unsigned long long share = A.n_lanes / (unsigned long long)nl;
        const unsigned long long want = std::max<unsigned long long>(256ull, share + share / 2);
        unsigned long long clog = 0;
        while (clog < 6 && (unsigned long long)(t->prefix_n[plen] >> (clog + 1)) >= want) ++clog;
        lj[k].ticket = ((unsigned long long)b << 49) | ((unsigned long long)plen << 54) | (clog << 59);
    }
    const unsigned long long ring_size = 2ull * nl + 1024;
    std::vector<unsigned long long> ring(ring_size, ~0ull);
    // ring entries: slot << 37 | round bound << 32 | cube (the kernel's ring_entry)
    for (int k = 0; k < nl; ++k)
        ring[k] = ((unsigned long long)k << 37) | ((unsigned long long)lj[k].bound << 32) | (unsigned)k;
    QueueCtl q{};
    q.head = 0;
    q.tail = (unsigned long long)nl;
    q.n_jobs = (unsigned)nl;
    QueueCtl* d_q = nullptr;
    unsigned long long *d_ring = nullptr, *d_clock = nullptr;
    SplitRec* d_splits = nullptr;
    HIP_TRY(mem.alloc(&d_q, sizeof(QueueCtl)));
    HIP_TRY(mem.alloc(&d_ring, sizeof(unsigned long long) * ring_size));
    HIP_TRY(mem.alloc(&d_clock, 2 * sizeof(unsigned long long)));
    HIP_TRY(mem.alloc(&d_splits, sizeof(SplitRec) * SPLIT_RING));
    // item keys (two buffers per cube and prefix length) and key histograms
    size_t key_total = 0;
    // (rows of whole 8-byte words: the search reads eight keys at a time)
    for (int p = o.prefix_len; p <= KEY_MAX_PLEN; ++p) key_total += 2ull * nl * ((t->prefix_n[p] + 7ull) / 8ull * 2ull);
    if (Mem::holds_keys && nl >= BAND_MIN_CUBES && key_total && key_total * 4 <= KEY_MAX_BYTES) {
        uint32_t *d_keys = nullptr, *d_hist = nullptr;
        HIP_TRY(mem.alloc(&d_keys, key_total * 4));
        HIP_TRY(mem.alloc(&d_hist, (size_t)nl * KEY_BINS * 4));
        size_t off = 0;
        for (int p = o.prefix_len; p <= KEY_MAX_PLEN; ++p) {
            A.keys[p] = d_keys + off;
            A.key_words[p] = (t->prefix_n[p] + 7ull) / 8ull * 2ull;
            off += 2ull * nl * A.key_words[p];
        }
        A.key_hist = d_hist;
        HIP_TRY(zero_async(d_keys, key_total * 4, t->stream));
        HIP_TRY(zero_async(d_hist, (size_t)nl * KEY_BINS * 4, t->stream));
    }
    HIP_TRY(zero_async(d_splits, sizeof(SplitRec) * SPLIT_RING, t->stream));
    HIP_TRY(hipMemcpyAsync(A.jobs, lj.data(), sizeof(SearchJob) * (size_t)nl, hipMemcpyHostToDevice, t->stream));
    HIP_TRY(hipMemcpyAsync(d_ring, ring.data(), sizeof(unsigned long long) * ring_size, hipMemcpyHostToDevice,
                           t->stream));
    HIP_TRY(hipMemcpyAsync(d_q, &q, sizeof q, hipMemcpyHostToDevice, t->stream));
    A.q = d_q;
    A.ring = d_ring;
    A.ring_size = ring_size;
    int rate_khz = 0;
    int rc = wall_clock_khz(&rate_khz);
    if (rc) return rc;
    rate_khz = std::max(rate_khz, 1);
    A.idle_limit = (unsigned long long)rate_khz * 60000ull;  // 60 s
    A.backoff_ticks = (unsigned long long)rate_khz / 50ull;  // 20 us
    A.splits = d_splits;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    if (events) {
        HIP_TRY(mem.event(&ev0));
        HIP_TRY(mem.event(&ev1));
    }
    hipLaunchKernelGGL(read_clock_kernel, dim3(1), dim3(1), 0, t->stream, d_clock);
    if (events) HIP_TRY(hipEventRecord(ev0, t->stream));
    HIP_TRY(launch_search(t->K, search_format(t, false), 1, A, grid, o.block, t->stream));
    ++launches;
    if (events) HIP_TRY(hipEventRecord(ev1, t->stream));
    hipLaunchKernelGGL(read_clock_kernel, dim3(1), dim3(1), 0, t->stream, d_clock + 1);
    HIP_TRY(hipMemcpyAsync(lj.data(), A.jobs, sizeof(SearchJob) * (size_t)nl, hipMemcpyDeviceToHost, t->stream));
    unsigned long long clk[2] = {0, 0};
    HIP_TRY(hipMemcpyAsync(clk, d_clock, sizeof clk, hipMemcpyDeviceToHost, t->stream));
    HIP_TRY(hipMemcpyAsync(&q, d_q, sizeof q, hipMemcpyDeviceToHost, t->stream));
    HIP_TRY(hipStreamSynchronize(t->stream));
    if (q.stop == 2) return fail(CCS_EFAULT, "persistent search stalled: no work for 60 s with cubes unfinished");
    float ms = (float)((double)(clk[1] - clk[0]) / (double)rate_khz);
    if (events) HIP_TRY(hipEventElapsedTime(&ms, ev0, ev1));
    kernel_ms += ms;
    double launch_ms = now_ms() - t_start - ms;  // host time before the kernel started
    for (int k = 0; k < nl; ++k) {
        SearchJob& J = lj[k];
        int i = live[k];
        ccs_result_t& r = out[i];
        r.n_rounds = std::min(J.n_rounds, 24u);
        for (uint32_t j = 0; j < r.n_rounds; ++j) {
            r.round_bound[j] = J.round_bound[j];
            r.round_nodes[j] = J.round_nodes[j];
            r.nodes += J.round_nodes[j];
        }
        if (J.found && (r.n_rounds == 0 || r.round_bound[r.n_rounds - 1] != J.length) && r.n_rounds < 24) {
            // the solving round ends when the solution is claimed
            r.round_bound[r.n_rounds] = J.length;
            r.round_nodes[r.n_rounds] = J.nodes;
            r.nodes += J.nodes;
            ++r.n_rounds;
        }
        jobs[i].found = J.found;
        jobs[i].length = J.length;
        memcpy(jobs[i].solution, J.solution, sizeof J.solution);
        if (J.status) done_ms[i] = launch_ms + (double)(J.t_final - clk[0]) / (double)rate_khz;
    }
    return CCS_OK;
}

// The diagnostics words of a launch into the caller's statistics.
void copy_diag(const unsigned long long* hs, ccs_solve_stats_t* stats) {
    static_assert(DIAG_WORDS <= (int)(sizeof stats->diag / sizeof stats->diag[0]), "diag words");
#ifdef CCS_SEARCH_STATS
    stats->diag_built = 1;
#endif
    for (int i = 0; i < DIAG_WORDS; ++i) stats->diag[i] = hs[STAT_DIAG + i];
}

// Verify every solution (solver.py:317-319) and fill the results.
int finish_results(const uint8_t* cubes, int n, const std::vector<SearchJob>& jobs, const std::vector<int>& solved0,
                   const std::vector<double>& done_ms, ccs_result_t* out) {
    int bad = -1;
    for (int i = 0; i < n; ++i) {
        ccs_result_t& r = out[i];
        r.seconds = done_ms[i] / 1e3;
        if (solved0[i]) {
            r.length = 0;
            continue;
        }
        if (!jobs[i].found) continue;
        r.length = (int)jobs[i].length;
        Cubie c = *reinterpret_cast<const Cubie*>(cubes + 40ull * i);
        for (int l = 0; l < r.length; ++l) {
            r.moves[l] = jobs[i].solution[l];
            c = then(c, move(r.moves[l]));
        }
        Cubie s = solved();
        if (memcmp(&c, &s, sizeof c) != 0 && bad < 0) bad = i;
    }
    if (bad >= 0) return fail(CCS_EFAULT, "solution of cube " + std::to_string(bad) + " does not solve it");
    for (int i = 0; i < n; ++i)
        if (out[i].length < 0) return fail(CCS_EFAULT, "cube " + std::to_string(i) + " not solved within max_bound");
    return CCS_OK;
}

void fill_args(ccs_tables* t, const ccs_solve_opts_t& o, int grid, SearchJob* d_jobs, unsigned long long* d_stats,
               SearchArgs& A) {
    memset(&A, 0, sizeof A);
    A.lut = t->d_lut;
    A.class_lut = t->d_class_lut;
    A.classidx = t->d_classidx;
    A.table = search_table(t, search_format(t, false));
    A.jobs = d_jobs;
    A.prefix_len = o.prefix_len;
    A.max_bound = std::min(o.max_bound, STACK_DEPTH);
    A.deeper0 = to_u64(o.deeper_at[0]);
    A.deeper1 = to_u64(o.deeper_at[1]);
    A.poll_mask = (unsigned)o.poll_interval - 1u;
    A.one_probe = o.probes_per_step == 1;
    A.read_mode = o.read_mode == 1 ? 0 : 1;
    A.chunk_nodes = o.chunk_nodes > 0 ? (unsigned long long)o.chunk_nodes : DEFAULT_CHUNK_NODES;
    A.n_lanes = (unsigned long long)grid * (unsigned long long)o.block;
    A.stats = d_stats;
    for (int p = 0; p <= MAX_PLEN; ++p) {
        A.pre_tab[p] = t->d_prefix[p];
        A.pre_n[p] = t->prefix_n[p] ? t->prefix_n[p] : 0xFFFFFFFFu;
    }
}

// ---- the allocation-free variant's device state -------------------------
//
// One array per device (each GPU the library runs on holds its own copy until
// the process ends).  In host code `scratch` names a host-side stand-in, not
// device memory: its device address comes from hipGetSymbolAddress, once per
// device with that device current, and only that address is ever used.
constexpr size_t NOALLOC_SCRATCH_BYTES = 1u << 20;
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

int scratch_address(uint8_t** out) {
    static std::mutex mu;
    static std::array<uint8_t*, 64> addr{};
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= (int)addr.size()) return fail(CCS_EINVAL, "device index out of range");
    std::lock_guard<std::mutex> g(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) return fail(CCS_EDEVICE, "no device address for the scratch array");
        addr[dev] = static_cast<uint8_t*>(p);
    }
    *out = addr[dev];
    return CCS_OK;
}

// Device bytes the allocation-free variant needs for n cubes, as the Arena
// lays them out (256-byte aligned pieces; batches this small keep no keys).
size_t noalloc_bytes(int n) {
    auto up = [](size_t b) { return (b + 255) & ~size_t(255); };
    return up(sizeof(Frame) * 6ull * n) + up(4ull * 6 * n) + up(sizeof(SearchJob) * (size_t)n) +
           up(STAT_WORDS * sizeof(unsigned long long)) + up(sizeof(QueueCtl)) + up(8ull * (2ull * n + 1024)) + up(16) +
           up(sizeof(SplitRec) * SPLIT_RING);
}

}  // namespace ccs

extern "C" int ccs_solve(ccs_tables_t* t, const uint8_t* cubes, int n, const ccs_solve_opts_t* opts_in,
                         ccs_result_t* out, ccs_solve_stats_t* stats) {
    if (!t || n < 0 || (n && (!cubes || !out))) return fail(CCS_EINVAL, "bad arguments");
    if (t->format < 0 || !t->d_table[t->format]) return fail(CCS_EINVAL, "no solving table resident");
    ccs_solve_opts_t o;
    fill_opts(opts_in, o);
    int rc = check_opts(o);
    if (rc) return rc;
    if (o.read_mode == 0) o.read_mode = o.mode == 0 || n >= PAIRED_READS_MIN_CUBES ? 1 : 2;
    double t_start = now_ms();
    if (stats) memset(stats, 0, sizeof *stats);
    if (n == 0) return CCS_OK;
    search_format(t, true);  // the search's half-block copy of a block table, made once
    for (int i = 0; i < n; ++i) {
        memset(&out[i], 0, sizeof out[i]);
        out[i].length = -1;
    }
    DeviceBuffers buf;
    std::vector<SearchJob> jobs;
    std::vector<int> solved0(n, 0);
    rc = root_setup(t, cubes, n, jobs, solved0, buf);
    if (rc) return rc;
    const bool queue = o.mode == 1;
    int grid = 0;
    rc = search_grid(t, queue, o.block, o.blocks_per_cu, &grid);
    if (rc) return rc;
    for (int p = 0; p <= MAX_PLEN; ++p) {
        if (p > o.prefix_len + 2) break;
        if ((rc = ensure_prefixes(t, p))) return rc;
    }

    SearchJob* d_jobs = nullptr;
    unsigned long long* d_stats = nullptr;
    HIP_TRY(buf.alloc(&d_jobs, sizeof(SearchJob) * (size_t)n));
    HIP_TRY(buf.alloc(&d_stats, STAT_WORDS * sizeof(unsigned long long)));
    HIP_TRY(zero_async(d_stats, STAT_WORDS * sizeof(unsigned long long), t->stream));
    double kernel_ms = 0;
    uint32_t launches = 0;
    SearchArgs A;
    fill_args(t, o, grid, d_jobs, d_stats, A);

    std::vector<double> done_ms(n, -1.0);
    if (!queue) {
        // ---- plain: host-driven rounds (solver.py:270-316) ----
        hipEvent_t ev0, ev1;
        HIP_TRY(buf.event(&ev0));
        HIP_TRY(buf.event(&ev1));
        const unsigned long long d0 = A.deeper0, d1 = A.deeper1;
        int* d_active = nullptr;
        unsigned long long* d_counter = nullptr;
        HIP_TRY(buf.alloc(&d_active, sizeof(int) * (size_t)n));
        HIP_TRY(buf.alloc(&d_counter, sizeof(unsigned long long)));
        HIP_TRY(hipMemcpyAsync(d_jobs, jobs.data(), sizeof(SearchJob) * (size_t)n, hipMemcpyHostToDevice, t->stream));
        std::vector<unsigned long long> last(n, 0);
        std::vector<int> finished(n, 0);
        for (int i = 0; i < n; ++i)
            if (solved0[i]) finished[i] = 1, done_ms[i] = now_ms() - t_start;
        for (;;) {
            std::map<int, std::vector<int>> groups;
            for (int i = 0; i < n; ++i) {
                if (finished[i]) continue;
                if ((int)jobs[i].bound > std::min(o.max_bound, STACK_DEPTH)) {
                    finished[i] = 1;
                    continue;
                }
                int plen = o.prefix_len + (last[i] >= d0) + (last[i] >= d1);
                plen = std::min(std::min(plen, (int)jobs[i].bound), MAX_PLEN);
                groups[plen].push_back(i);
            }
            if (groups.empty()) break;
            std::vector<int> ids;
            HIP_TRY(hipEventRecord(ev0, t->stream));
            for (auto& kv : groups) {
                if ((rc = ensure_prefixes(t, kv.first))) return rc;
                ids.insert(ids.end(), kv.second.begin(), kv.second.end());
            }
            HIP_TRY(hipMemcpyAsync(d_active, ids.data(), sizeof(int) * ids.size(), hipMemcpyHostToDevice, t->stream));
            size_t off = 0;
            for (auto& kv : groups) {
                int plen = kv.first;
                HIP_TRY(zero_async(d_counter, sizeof(unsigned long long), t->stream));
                A.active = d_active + off;
                A.n_active = (int)kv.second.size();
                A.plen = plen;
                A.prefixes = t->d_prefix[plen];
                A.n_prefix = t->prefix_n[plen];
                A.counter = d_counter;
                HIP_TRY(launch_search(t->K, search_format(t, false), 0, A, grid, o.block, t->stream));
                ++launches;
                off += kv.second.size();
            }
            HIP_TRY(hipEventRecord(ev1, t->stream));
            HIP_TRY(hipMemcpyAsync(jobs.data(), d_jobs, sizeof(SearchJob) * (size_t)n, hipMemcpyDeviceToHost, t->stream));
            HIP_TRY(hipStreamSynchronize(t->stream));
            float ms = 0;
            HIP_TRY(hipEventElapsedTime(&ms, ev0, ev1));
            kernel_ms += ms;
            double now = now_ms() - t_start;
            for (int i : ids) {
                SearchJob& J = jobs[i];
                ccs_result_t& r = out[i];
                if (r.n_rounds < 24) {
                    r.round_bound[r.n_rounds] = J.bound;
                    r.round_nodes[r.n_rounds] = J.nodes;
                    ++r.n_rounds;
                }
                r.nodes += J.nodes;
                last[i] = J.nodes;
                J.nodes = 0;
                if (J.found) {
                    finished[i] = 1;
                    done_ms[i] = now;
                } else {
                    J.bound += 1;
                }
            }
            HIP_TRY(hipMemcpyAsync(d_jobs, jobs.data(), sizeof(SearchJob) * (size_t)n, hipMemcpyHostToDevice, t->stream));
        }
        HIP_TRY(hipStreamSynchronize(t->stream));
    } else {
        // ---- persistent batched: all rounds on the device ----
        rc = solve_persistent(t, o, A, grid, n, jobs, solved0, buf, true, t_start, out, done_ms, kernel_ms, launches);
        if (rc) return rc;
    }
    if (stats) {
        unsigned long long hs[STAT_WORDS];
        HIP_TRY(hipMemcpy(hs, d_stats, sizeof hs, hipMemcpyDeviceToHost));
        stats->probes = hs[0];
        copy_diag(hs, stats);
        stats->nodes = 0;  // children evaluated plus prefix moves replayed, every round of every cube
        for (int i = 0; i < n; ++i) stats->nodes += out[i].nodes;
        stats->kernel_ms = kernel_ms;
        stats->launches = launches;
        stats->wall_ms = now_ms() - t_start;
    }
    return finish_results(cubes, n, jobs, solved0, done_ms, out);
}

extern "C" int ccs_solve_noalloc(ccs_tables_t* t, const uint8_t* cubes, int n, const ccs_solve_opts_t* opts_in,
                                 ccs_result_t* out, ccs_solve_stats_t* stats) {
    if (!t || n < 0 || (n && (!cubes || !out))) return fail(CCS_EINVAL, "bad arguments");
    if (t->format < 0 || !t->d_table[t->format]) return fail(CCS_EINVAL, "no solving table resident");
    if (n > CCS_NOALLOC_MAX_CUBES)
        return fail(CCS_EINVAL, std::to_string(n) + " cubes: the allocation-free solve takes at most " +
                                    std::to_string(CCS_NOALLOC_MAX_CUBES));
    ccs_solve_opts_t o;
    fill_opts(opts_in, o);
    o.mode = 1;  // the persistent form only: all rounds in one launch
    int rc = check_opts(o);
    if (rc) return rc;
    if (o.read_mode == 0) o.read_mode = n >= PAIRED_READS_MIN_CUBES ? 1 : 2;
    for (int p = 0; p <= std::min(o.prefix_len + 2, MAX_PLEN); ++p)
        if (!t->prefix_n[p])
            return fail(CCS_EINVAL, "prefix table " + std::to_string(p) +
                                        " is not resident (tables keep lengths 0..6; prefix_len above 4 needs a "
                                        "ccs_solve with that prefix_len first)");
    double t_start = now_ms();
    if (stats) memset(stats, 0, sizeof *stats);
    if (n == 0) return CCS_OK;
    for (int i = 0; i < n; ++i) {
        memset(&out[i], 0, sizeof out[i]);
        out[i].length = -1;
    }
    Arena mem;
    if ((rc = scratch_address(&mem.base))) return rc;
    mem.cap = NOALLOC_SCRATCH_BYTES;
    if (noalloc_bytes(n) > mem.cap) return fail(CCS_EINVAL, "scratch array too small for this many cubes");
    std::vector<SearchJob> jobs;
    std::vector<int> solved0(n, 0);
    rc = root_setup(t, cubes, n, jobs, solved0, mem);
    if (rc) return rc;
    int grid = 0;
    rc = search_grid(t, true, o.block, o.blocks_per_cu, &grid);
    if (rc) return rc;
    SearchJob* d_jobs = nullptr;
    unsigned long long* d_stats = nullptr;
    HIP_TRY(mem.alloc(&d_jobs, sizeof(SearchJob) * (size_t)n));
    HIP_TRY(mem.alloc(&d_stats, STAT_WORDS * sizeof(unsigned long long)));
    HIP_TRY(zero_async(d_stats, STAT_WORDS * sizeof(unsigned long long), t->stream));
    SearchArgs A;
    fill_args(t, o, grid, d_jobs, d_stats, A);
    double kernel_ms = 0;
    uint32_t launches = 0;
    std::vector<double> done_ms(n, -1.0);
    rc = solve_persistent(t, o, A, grid, n, jobs, solved0, mem, false, t_start, out, done_ms, kernel_ms, launches);
    if (rc) return rc;
    if (stats) {
        unsigned long long hs[STAT_WORDS];
        HIP_TRY(hipMemcpy(hs, d_stats, sizeof hs, hipMemcpyDeviceToHost));
        stats->probes = hs[0];
        copy_diag(hs, stats);
        for (int i = 0; i < n; ++i) stats->nodes += out[i].nodes;
        stats->kernel_ms = kernel_ms;
        stats->launches = launches;
        stats->wall_ms = now_ms() - t_start;
    }
    return finish_results(cubes, n, jobs, solved0, done_ms, out);
}
