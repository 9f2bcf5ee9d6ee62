/* CPU oracle for the optimal-cube-solver hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C restatement of the reference's algorithm, used by tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker.
 * Nothing in the product links or calls this code.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference checkout).
 */
#ifndef CUBE_ORACLE_H
#define CUBE_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Cubie state, Kociemba convention (src/cuda_cube_solver/cubie.py:1-51). */
typedef struct {
    uint8_t cp[8], co[8], ep[12], eo[12];
} ocube;

/* Table kind: 304 = EdgeKind(3,4) (coords.py:37-76), else a GroupKind
 * (kinds.py:56-96) with E content classes. */
typedef struct {
    int is304;
    int nec;
    uint64_t n_edge;
    uint64_t n_entries;
    uint16_t gen_class[384];
    uint16_t gen_rep[2][384];
} okind;

int or_init(void);
void or_move(int m, ocube* out);
void or_compose(const ocube* a, const ocube* b, ocube* out);
void or_inverse(const ocube* a, ocube* out);
void or_sym_matrix(int sym48, int32_t out[9]);
void or_sym_conjugate(int sym48, const ocube* in, ocube* out);
int or_ud_sym(int k); /* index into the 48 of the k-th UD symmetry */
void or_axis_frames(const ocube* in, ocube out[6]);
int or_axis_move_map(int m);
uint32_t or_corner_raw(const ocube* c);
const uint32_t* or_classidx(void);
const uint32_t* or_class_rep(void);
const uint16_t* or_class_stab(void);
int or_n_classes(void);
int or_kind_init(okind* k, const char* name);
int or_group_kind(okind* k, const int* gens, int n_gens);
uint64_t or_table_index(const okind* k, const ocube* c);
void or_unrank(const okind* k, uint64_t idx, ocube* out);

/* BFS (kernels.cu:389-624, table.py:82-129). Writes the block table (n_blocks * 32 B)
 * and the mod-3 table ((n_entries+15)/16 words); returns the number of depths or < 0. */
int or_bfs(const okind* k, uint32_t* mod3, uint32_t* blk, uint64_t* depth_counts, int max_depths,
           int n_threads, int verbose);
void or_histogram(const uint32_t* mod3, uint64_t n_entries, uint64_t counts[4]);
/* b3 encoding (kernels.cu:1825-1869). out: n_blocks * 8 words. */
void or_build_b3(const uint32_t* mod3, const uint32_t* blk, uint64_t b0, uint64_t nb, uint64_t n_entries,
                 uint32_t* out);

/* Lower bound from a solving table: fmt 0 = block, 1 = b3 (kernels.cu:408-434). */
uint32_t or_lookup(const uint8_t* table, int fmt, uint64_t idx);

/* IDA* with the reference's step-mode pruning (kernels.cu:1447-1801, solver.py:254-320). */
typedef struct {
    int prefix_len;           /* 4 */
    double deeper_at[2];      /* 5e7, 3e9 */
    int max_bound;            /* give up above this bound (-1 = 24) */
} osolve_opts;

typedef struct {
    int length;               /* -1 if not solved */
    uint8_t moves[24];
    int n_rounds;
    uint32_t round_bound[24];
    uint64_t round_nodes[24];
    uint64_t nodes;
    uint32_t root_h[6];
} osolve_result;

int or_solve(const okind* k, const uint8_t* table, int fmt, const ocube* cube, const osolve_opts* opts,
             osolve_result* res);
int or_solve_many(const okind* k, const uint8_t* table, int fmt, const ocube* cubes, int n,
                  const osolve_opts* opts, osolve_result* res, int n_threads, double* seconds);
/* Stop searches after this many seconds of wall time (0 = no limit). */
void or_set_deadline(double seconds_from_now);
/* Canonical move sequences of a length (solver.py:33-48). */
int64_t or_canonical_prefixes(int length, uint8_t* out, int64_t cap);

#ifdef __cplusplus
}
#endif
#endif
