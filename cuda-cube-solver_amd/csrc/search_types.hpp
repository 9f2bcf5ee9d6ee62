// Structures shared by the host driver and the kernels.
#pragma once
#include <cstdint>

#include "device_cube.hpp"

namespace ccs {

constexpr int MAX_BOUND = 24;   // longest searched sequence (the reference's MAX_DEPTH, kernels.cu:700)
// Depth of the per-lane search stack.  Every legal position is within 20 face
// turns, so no search of a legal cube goes deeper; bounds above it are refused.
constexpr int STACK_DEPTH = 20;
// Persistent form: nodes a lane aims to take per fetch (a 1,024-cube d4t sweep
// of 16K..1M peaks at 256K; at least 4,096 chunks per round keep one cube wide).
constexpr unsigned long long DEFAULT_CHUNK_NODES = 262144;
// Serving: one lane covers ~40k nodes/s, so small chunks keep any one cube
// from waiting seconds on a single lane.
constexpr unsigned long long SERVE_CHUNK_NODES = 16384;
constexpr int MAX_PLEN = 7;     // longest work-item prefix
constexpr int STAT_WORDS = 40;  // device statistics words of a search launch
constexpr int STAT_DIAG = 8;    // first diagnostics word (DIAG_WORDS of them)
constexpr int DIAG_WORDS = 30;
// Persistent batches of at least BAND_MIN_CUBES cubes: item keys are kept for
// rounds with prefixes of at most KEY_MAX_PLEN moves, and a long round (at
// least BAND_MIN_NODES_PER_ITEM nodes per item in the round before) whose
// previous round had the same prefix length visits its items in N_BANDS bands
// of rising key, while at least BAND_MIN_LIVE cubes are unfinished.  (With
// banded chunks of at most 8 to 32 visits, by the cube's share of the lanes:
// 4,096 cubes 3-4% faster than in move order, 256 to 1,024 cubes 12-19%;
// 48 to 96 cubes 2-8% slower, and 320 hard cubes on d4 (length 18 or more,
// and flip-heavy) 4% slower unless the last 64 go in move order.)
constexpr int KEY_MAX_PLEN = 5;
constexpr int BAND_MIN_CUBES = 256;
constexpr unsigned BAND_MIN_LIVE = 64;
// ... and whose keys fit in this many bytes (1.24 MB per cube: about 13,500 cubes)
constexpr unsigned long long KEY_MAX_BYTES = 16ull << 30;
constexpr uint32_t N_BANDS = 16;  // a multiple of 4, at most 16
constexpr uint32_t KEY_BINS = 128;
constexpr unsigned long long BAND_MIN_NODES_PER_ITEM = 32;

// One cube being searched, resident in HBM (the role of the reference's
// CubeJob, kernels.cu:702-710, plus the per-cube round state of the
// persistent form).
struct alignas(64) SearchJob {
    Frame frame[6];              // 0..2: R^k g R^-k; 3..5: their inverses
    uint32_t h[6];               // root lower bounds per frame
    uint32_t bound;              // current bound (plain form: set by the host)
    uint32_t found;              // 1 once a lane has claimed a solution
    uint32_t length;
    uint8_t solution[MAX_BOUND];
    uint32_t status;             // persistent form: 0 searching, 1 solved, 2 gave up
    uint32_t n_rounds;
    unsigned long long nodes;    // nodes of the current round
    unsigned long long ticket;   // persistent form: round ticket (see kernels.hip)
    unsigned long long t_final;  // device clock when the cube became final
    unsigned long long tag;      // serving: the caller's tag
    unsigned long long t_admit;  // serving: device clock at admission
    uint32_t round_bound[MAX_BOUND];
    unsigned long long round_nodes[MAX_BOUND];
    uint32_t band_thr[4];        // persistent form: item-key bounds of the round's bands (N_BANDS - 1 bytes)
    uint32_t credit;             // serving: fetches the cube keeps the ring's head for, on its turn
};

// Control block of the persistent form.
struct QueueCtl {
    unsigned long long head;
    unsigned long long tail;
    unsigned int n_final;
    unsigned int n_jobs;
    unsigned int stop;           // 1: stop requested; 2: no work for idle_limit while cubes were unfinished
    unsigned int pad;
    unsigned long long mb_head;  // serving: mailbox records admitted
    unsigned long long done_next;// serving: completion records claimed
    unsigned long long mb_seen;  // serving: the mailbox tail as last read from host memory
    unsigned long long mb_polled;// serving: device clock of that read
    // donated subtrees (SplitRec ring): records claimed, records reserved,
    // records reserved and not yet claimed, and whether lanes are out of work
    unsigned long long split_head;
    unsigned long long split_tail;
    unsigned int split_pend;
    unsigned int split_want;
};

// A subtree donated by a lane to lanes without work once no cube has items
// left to hand out: the node reached by `path` (move j in byte j, j < k),
// searched from the move after p on; byte k of `path` holds p + 1.
struct SplitRec {
    unsigned long long path;
    uint32_t cube;
    uint32_t info;             // round bound | k << 8
    unsigned long long seq;    // ring index + 1 once the record is written
    unsigned long long pad;
};
// Ring of donated subtrees (a power of two) and the most reserved and not yet
// claimed at once; half the ring, so a writer never reaches a record that a
// reader may still be reading.
constexpr unsigned int SPLIT_RING = 16384;
constexpr unsigned int SPLIT_PEND_MAX = SPLIT_RING / 2;

// Serving mailbox record (host -> device, in pinned host memory); seq is
// written last, as record index + 1.
struct ServeRec {
    Frame frame[6];
    unsigned long long tag;
    uint32_t slot;
    uint32_t pad;
    unsigned long long seq;
};

// Serving completion record (device -> host, pinned host memory).
struct ServeDone {
    unsigned long long tag;
    uint32_t slot;
    int32_t length;
    uint8_t moves[MAX_BOUND];
    unsigned long long nodes;
    unsigned long long t_admit;
    unsigned long long t_final;
    unsigned long long seq;
};

struct SearchArgs {
    const CubeLUT* lut;
    const ClassLUT* class_lut;  // corner classes for the probes (copied to LDS)
    const uint32_t* classidx;   // the same classes as one array (admission's root bounds)
    const uint8_t* table;
    SearchJob* jobs;
    // plain form: one round of the listed cubes
    const int* active;
    int n_active;
    int plen;
    const unsigned long long* prefixes;  // one item per word, move j in byte j
    unsigned long long n_prefix;
    unsigned long long* counter;
    // persistent form
    QueueCtl* q;
    unsigned long long* ring;
    unsigned long long ring_size;
    const unsigned long long* pre_tab[MAX_PLEN + 1];
    unsigned int pre_n[MAX_PLEN + 1];
    int prefix_len;
    int max_bound;
    unsigned long long deeper0, deeper1;
    unsigned int poll_mask;      // found flag polled every poll_mask + 1 picks
    int one_probe;               // 1: one table probe per step (the top remaining frame); 0: a forward and an inverse
    int read_mode;               // a step's two table reads: 0 paired (in flight together), 1 one after the other
    unsigned long long idle_limit;  // persistent form: device-clock ticks a lane may wait for work
    unsigned long long backoff_ticks; // persistent form: idle ticks after which a lane polls less often
    unsigned long long chunk_nodes; // persistent form: target nodes per fetched chunk of items
    unsigned long long n_lanes;     // persistent form: lanes of the launch (grid * workgroup)
    // serving
    int serving;
    const ServeRec* mb;
    const unsigned long long* mb_tail;   // pinned host: records published
    unsigned long long mb_size;
    ServeDone* done;                     // pinned host ring
    unsigned long long done_size;
    const unsigned int* host_stop;       // pinned host: 1 = stop
    unsigned long long mb_poll_ticks;    // device-clock ticks between two reads of the mailbox tail
    unsigned long long age_ticks;        // serving: a cube's turn at the head grows by one fetch per this age
    SplitRec* splits;                    // persistent form: SPLIT_RING donated subtrees
    // persistent batches: per cube and prefix length, two buffers of one key
    // byte per item (1 + the sum of the six bounds at the prefix's end; 0
    // unknown), four to a word; null where not kept.  key_hist: per cube,
    // KEY_BINS counters of a sample of the round's keys.
    uint32_t* keys[MAX_PLEN + 1];
    unsigned long long key_words[MAX_PLEN + 1];
    uint32_t* key_hist;
    // statistics (STAT_WORDS words): [0] probes; from STAT_DIAG on, the
    // diagnostics of a -DCCS_SEARCH_STATS build (ccs_solve_stats_t::diag)
    unsigned long long* stats;
};

}  // namespace ccs
