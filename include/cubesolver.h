/* libcubesolver: MI355X (gfx950) optimal Rubik's-cube solver, half-turn metric.
 *
 * C ABI of the solver's hot path: the BFS that builds the coset pruning
 * tables and the IDA* search over them.  Plain pointers and sizes only.
 *
 * The reference (ozimchuk/cuda-cube-solver) crosses its host/GPU boundary in
 * Python through CuPy RawModule launches; each entry point below names the
 * reference interface it replaces.  Paths are relative to the reference
 * checkout.
 *
 * Conventions
 *  - Every call returns CCS_OK (0) or a negative code; ccs_last_error() gives
 *    the message of the calling thread's last failure.
 *  - A cube is 40 bytes: cp[8], co[8], ep[12], eo[12] in the reference's
 *    Kociemba cubie convention (src/cuda_cube_solver/cubie.py:1-51).
 *  - Host buffers belong to the caller; device memory belongs to the handles.
 *  - A tables handle is used by one thread at a time; a server accepts
 *    submissions from any thread.
 */
#ifndef CUBESOLVER_H
#define CUBESOLVER_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CCS_OK 0
#define CCS_ENOENT (-2)   /* table file missing  (solver.py:82-88 TableNotFound) */
#define CCS_EIO (-5)      /* file size or read error (table.py:36-37 ValueError) */
#define CCS_ENOMEM (-12)  /* device memory budget / allocation (kinds.py:184-195, NoTableFits) */
#define CCS_EFAULT (-14)  /* self-check failed (table.py:183-189, solver.py:317-319 AssertionError) */
#define CCS_EINVAL (-22)  /* bad name or argument (kinds.py:203-206 ValueError) */
#define CCS_EBUSY (-16)   /* server queue full */
#define CCS_EDEVICE (-100) /* HIP runtime error */

#define CCS_FORMAT_BLOCK 0 /* prune-ud-<kind>-blk.bin: u8 minimum + 124 two-bit offsets per 32 B */
#define CCS_FORMAT_B3 1    /* prune-ud-<kind>-b3.bin: header + 155 base-3 digits per 32 B */
#define CCS_TABLE_DEPTHS 2 /* ccs_tables_read/bytes: the BFS build table, an exact depth per entry in
                              a nibble (8 entries per u32, entry j at bits 4 (j mod 8)); resident only
                              if the build was asked to keep it */
#define CCS_BUILD_KEEP_DEPTHS 4 /* OR'ed into ccs_tables_build's format: keep the build table resident */

typedef struct ccs_tables ccs_tables_t;
typedef struct ccs_server ccs_server_t;

/* Library version (major * 10000 + minor * 100 + patch). */
int ccs_version(void);

/* Message of the calling thread's last failed call ("" if none). */
const char* ccs_last_error(void);

typedef struct {
    uint64_t free_bytes;
    uint64_t total_bytes;
    int32_t compute_units;
    int32_t device;
    char name[64];
    char arch[32];
} ccs_device_info_t;

/* Free/total HBM and CU count of the current device.
 * Replaces cp.cuda.runtime.memGetInfo() in kinds.pick (kinds.py:207-211) and
 * the MultiProcessorCount query of GpuSolver (solver.py:197). */
int ccs_device_info(ccs_device_info_t* out);

typedef struct {
    uint64_t n_edge;
    uint64_t n_entries;
    int32_t n_classes;  /* corner classes (9,930) */
    int32_t nec;        /* E-content classes of a group kind, 0 for 304 */
} ccs_kind_info_t;

/* Sizes of a table kind ("304", "s4", "d4f", "a4t", "d4t", "d4", "a4"); no GPU needed.
 * Replaces kinds.get(...).n_edge / Preset.entries (kinds.py:110-115, 141-143). */
int ccs_kind_info(const char* kind, ccs_kind_info_t* out);

/* Load a solving table file into HBM (chunked, pinned, overlapped with the
 * copy) and set up the device constants of its kind.
 * Replaces DeviceTables + load_device + GpuSolver.__init__ table loading
 * (device.py:125-141, table.py:20-38, solver.py:118-224). */
int ccs_tables_open(const char* path, const char* kind, int format, ccs_tables_t** out);

/* Build a kind's tables on the GPU by BFS and keep the solving table(s)
 * resident: format CCS_FORMAT_BLOCK, CCS_FORMAT_B3, or 2 for both.
 * depth_counts (capacity *n_depths on input) receives the entries per depth,
 * *n_depths the number of depths.  log (may be NULL) gets one line per depth.
 * Runs the reference's self-checks (histogram equals the depth counts, no
 * entry left unvisited).
 * Replaces table.generate + histogram + build_preset's checks
 * (table.py:82-137, 170-189; kernels.cu seed_solved/expand_forward/
 * expand_backward/count_values/build_b3 at :532-624, :1839-1869). */
typedef void (*ccs_log_fn)(const char* line, void* user);
int ccs_tables_build(const char* kind, int format, ccs_log_fn log, void* user, uint64_t* depth_counts,
                     int* n_depths, ccs_tables_t** out);

/* Write a resident solving table to path (written to path.tmp, then renamed).
 * Replaces save_device / write_b3 and the .tmp rename (table.py:20-25, 66-79, 190-198). */
int ccs_tables_save(ccs_tables_t* t, int format, const char* path);

/* Choose which resident solving table (CCS_FORMAT_BLOCK or CCS_FORMAT_B3)
 * probes and searches read; a table built with both formats keeps both.
 * Replaces picking the table file by format in GpuSolver (solver.py:164-169). */
int ccs_tables_use(ccs_tables_t* t, int format);

/* Bytes of a resident solving table, or of the kept build table with
 * CCS_TABLE_DEPTHS (0 if not resident). */
uint64_t ccs_tables_bytes(ccs_tables_t* t, int format);

/* Copy n bytes at offset of a resident solving table (or, with
 * CCS_TABLE_DEPTHS, of the kept build table) to host memory.  The reference
 * reads its tables back with CuPy .get() (table.py:20-25). */
int ccs_tables_read(ccs_tables_t* t, int format, uint64_t offset, uint64_t n, void* host);

void ccs_tables_close(ccs_tables_t* t);

/* Table index and lower bound of n cubes after an optional move
 * (move >= 0: cube then move; move <= -2: move (-2 - move) then cube).
 * state_out (n x 40 bytes, may be NULL) receives the moved cube.
 * Replaces the probe_index / probe_block kernels (kernels.cu:626-642, 737-746)
 * used by GpuSolver.root_h and the reference's GPU tests. */
int ccs_probe(ccs_tables_t* t, const uint8_t* cubes, int n, int move, uint64_t* idx_out, uint32_t* lb_out,
              uint8_t* state_out);

/* table_index(unrank(idx)) for n indices (the identity when both agree).
 * Replaces the probe_unrank test kernel (kernels.cu:1168-1178). */
int ccs_unrank_check(ccs_tables_t* t, const uint64_t* idx, int n, uint64_t* out);

typedef struct {
    int32_t mode;           /* 0: plain (per-round launches, the reference's form); 1: persistent batched */
    int32_t prefix_len;     /* work-item prefix length, >= 4 (solver.py:119, default 4) */
    double deeper_at[2];    /* longer prefixes after rounds of this many nodes (solver.py:123: 5e7, 3e9) */
    int32_t block;          /* lanes per workgroup, a multiple of 64 up to 768 (0: 768); one workgroup runs per CU */
    int32_t blocks_per_cu;  /* persistent grid per CU (0: from occupancy) */
    int32_t max_bound;      /* give up above this bound (0: 24); searches stop at 20, the longest optimal solution of any legal cube */
    int32_t poll_interval;  /* picks between polls of the found flag, power of two (0: 64) */
    int32_t chunk_nodes;    /* persistent form: target nodes per chunk of items a lane takes (0: 262144) */
    int32_t probes_per_step; /* table probes a lane keeps in flight per step, 1 or 2 (0: 2): a child's
                               first step probes its top frame alone, later steps one forward and one
                               inverse frame (2) or the top remaining frame (1); nodes are the same */
    int32_t read_mode;       /* a step's two table reads: 1 paired (issued together; a lane with one
                               probe reads its block twice), 2 in turn (each read waited for before the
                               next), 0: paired in the plain form and for persistent batches of at
                               least 1,024 cubes, else in turn */
} ccs_solve_opts_t;

typedef struct {
    int32_t length;               /* optimal length, -1 if not solved */
    uint8_t moves[24];            /* 0..17 = U U2 U' R R2 R' F F2 F' D D2 D' L L2 L' B B2 B' */
    uint32_t n_rounds;
    uint32_t round_bound[24];
    uint64_t round_nodes[24];     /* nodes per round (children evaluated + prefix moves replayed) */
    uint64_t nodes;
    double seconds;               /* from the start of the call to this cube's result */
} ccs_result_t;

typedef struct {
    uint64_t probes;              /* table lookups in the search kernels */
    uint64_t nodes;
    double kernel_ms;             /* time inside search kernels (HIP events) */
    uint32_t launches;
    double wall_ms;
    /* Diagnostics, filled by a library built with -DCCS_SEARCH_STATS
     * (libcubesolver_diag.so; zero otherwise), the counterpart of the reference's
     * LOOKUP_STATS / LANE_STATS / PHASE_STATS builds (kernels.cu:1485-1505, 1773-1800):
     *  [0..5]   probes per frame (0..2 forward, 3..5 inverse)
     *  [6..11]  probes that pruned their child, per frame
     *  [12]     steps (summed over waves)   [13] frame updates (lanes)
     *  [14..22] steps by the number of lanes that probe: 0, 1-8, 9-16, ..., 57-64
     *  [23..26] shader-clock cycles per wave in a step's parts: the pass (bookkeeping,
     *           fetches, picks), the frame updates, the probes' addresses, the wait
     *           for the table reads
     *  [27]     pass cycles of the passes in which a lane of the wave started an item
     *  [28]     steps in which a lane of the wave started an item
     *  [29]     iterations of the pass loop (summed over waves; a wave's pass
     *           runs until each of its lanes has a probe or a frame update) */
    int32_t diag_built;
    uint64_t diag[32];
} ccs_solve_stats_t;

/* Solve n cubes optimally; every returned sequence is verified to solve its
 * cube (CCS_EFAULT otherwise).  stats may be NULL.
 * Replaces GpuSolver.solve (solver.py:254-320) and the search_step kernel
 * (kernels.cu:1447-1801); mode 1 keeps every cube's IDA* rounds on the
 * device in one persistent launch. */
int ccs_solve(ccs_tables_t* t, const uint8_t* cubes, int n, const ccs_solve_opts_t* opts, ccs_result_t* out,
              ccs_solve_stats_t* stats);

/* Most cubes ccs_solve_noalloc takes in one call. */
#define CCS_NOALLOC_MAX_CUBES 256

/* ccs_solve in the persistent batched form (mode 1, whatever opts->mode
 * says) with no device allocation: its device state (jobs, round queue, ring,
 * donated-subtree records, counters; about 0.7 MB at 256 cubes) lives in one
 * static 1 MiB device array per GPU, and the prefix tables it reads are the
 * ones every tables handle keeps (lengths 0..6, i.e. prefix_len 4).  No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  n above CCS_NOALLOC_MAX_CUBES is refused (CCS_EINVAL).
 * Calls that overlap on one GPU share that array and race: serialise them.
 * Kernel time in stats comes from the device clock instead of HIP events. */
int ccs_solve_noalloc(ccs_tables_t* t, const uint8_t* cubes, int n, const ccs_solve_opts_t* opts, ccs_result_t* out,
                      ccs_solve_stats_t* stats);

/* ---- serving: continuous batching on a resident table ------------------ */

typedef struct {
    uint64_t state_budget_bytes; /* device memory for per-search state (0: 64 MiB); a search holds its job
                                    and ring entries (608 B), so the budget sets how many cubes are
                                    searched at once */
    int32_t max_inflight;        /* cubes searched at once (0: from the budget) */
    int32_t prefix_len;          /* as in ccs_solve_opts_t */
    double deeper_at[2];         /* as for ccs_solve, but 0, 0 means 1e6, 1e8: items of long rounds get
                                    shorter sooner than in ccs_solve, so that no single item holds a
                                    cube back for long on a shared GPU */
    int32_t blocks_per_cu;
    int32_t chunk_nodes;         /* target nodes per chunk of items a lane takes (0: 16384); a
                                    lane covers ~40k nodes/s, so this bounds how long one chunk
                                    holds a cube back */
    int32_t read_mode;           /* as in ccs_solve_opts_t: 1 paired, 2 in turn (0: paired) */
} ccs_server_opts_t;

typedef struct {
    uint64_t tag;
    int32_t length;
    uint8_t moves[24];
    uint64_t nodes;
    double queued_ms;            /* submission -> admission to the device */
    double latency_ms;           /* submission -> result */
} ccs_served_t;

/* Start a serving loop on a resident table: a persistent kernel keeps the
 * searches of admitted cubes on the device and takes new cubes as slots and
 * memory free up.  No counterpart in the reference (its CLI solves fixed
 * batches, __main__.py:128-132). */
int ccs_server_start(ccs_tables_t* t, const ccs_server_opts_t* opts, ccs_server_t** out);
/* Queue one cube (thread-safe). */
int ccs_server_submit(ccs_server_t* s, const uint8_t* cube, uint64_t tag);
/* Collect up to max finished results, waiting up to timeout_ms for the first. */
int ccs_server_poll(ccs_server_t* s, ccs_served_t* out, int max, int timeout_ms);
/* Stop the loop (pending cubes are dropped) and free its device state. */
int ccs_server_stop(ccs_server_t* s);

/* ---- diagnostics ------------------------------------------------------- */

/* Random-probe micro-benchmark on the resident solving table: n_probes
 * random 32-byte blocks read as a probe of the block layout reads them (two
 * 4-byte reads: mode 0; mode 1 with nontemporal loads), one random 16-byte
 * half-block in one load as the search reads its half-block copy (mode 3),
 * or one streaming pass over the table (mode 2).  The measured rate is the
 * memory system's ceiling for the search's probes.
 * Counterpart of the reference's bench/gpu_gather.py. */
int ccs_gather_bench(ccs_tables_t* t, int mode, uint64_t n_probes, int blocks_per_cu, double* ms, uint64_t* checksum);

/* Table index of n cubes computed on the host with the same LDS lookup-table
 * arithmetic the kernels use (no GPU; for checking that arithmetic). */
int ccs_index_host(const char* kind, const uint8_t* cubes, int n, uint64_t* idx_out);

#ifdef __cplusplus
}
#endif
#endif
