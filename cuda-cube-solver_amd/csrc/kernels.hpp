// Launch entry points of the kernels (kernels.hip); K is the table kind (0 for
// 304, else the number of E-content classes), F the solving-table format
// (0 block, 1 base 3).
#pragma once
#include <hip/hip_runtime.h>

#include "search_types.hpp"

// Lanes per search workgroup: one workgroup of 12 waves per CU holds one copy
// of the lookup tables plus every lane's stack in LDS.
#ifndef SEARCH_WG
#define SEARCH_WG 768
#endif
// Waves per SIMD the search kernel is compiled for (caps it at 512 / N VGPRs).
#ifndef SEARCH_WAVES_PER_SIMD
#define SEARCH_WAVES_PER_SIMD 3
#endif

namespace ccs {

size_t search_smem_bytes(int block);
// form: 0 plain (one round of the listed cubes), 1 persistent batch, 2 persistent serving
hipError_t search_occupancy(int K, int F, int form, int block, int* blocks_per_cu);
hipError_t launch_search(int K, int F, int form, const SearchArgs& args, int grid, int block, hipStream_t s);
hipError_t launch_probe(int K, int F, const CubeLUT* lut, const uint32_t* classidx, const uint8_t* table,
                        const Frame* in, int n, int move, unsigned long long* idx_out, uint32_t* lb_out,
                        Frame* state_out, hipStream_t s);
hipError_t launch_unrank_check(int K, const CubeLUT* lut, const uint32_t* classidx, const uint32_t* class_rep,
                               const unsigned long long* idx, int n, unsigned long long* out, hipStream_t s);
hipError_t launch_bfs_seed(int K, const CubeLUT* lut, uint32_t* tbl, const uint32_t* classidx,
                           const uint16_t* class_stab, unsigned long long* counter, hipStream_t s);
hipError_t launch_bfs_layer(int K, bool forward, int grid, const CubeLUT* lut, uint32_t* tbl, const uint32_t* classidx,
                            const uint32_t* class_rep, const uint16_t* class_stab, unsigned long long n_words,
                            unsigned long long n_entries, uint32_t depth, unsigned long long* counter, hipStream_t s);
hipError_t launch_histogram(int grid, const uint32_t* tbl, unsigned long long n_words, unsigned long long n_entries,
                            unsigned long long* counts, hipStream_t s);
// format 0 writes the block table (out; none if null) and every block's
// minimum (mins, one byte per 124-entry block); format 1 writes the base-3
// table, reading those minima for the reference's saturated-offset values.
hipError_t launch_encode(int format, const uint32_t* tbl, uint8_t* mins, unsigned long long n_entries,
                         unsigned long long b0, unsigned long long nb, uint32_t* out, unsigned int* bad, hipStream_t s);

// The block table in the search's half-block layout (FMT_HALVES); *bad counts
// blocks whose minimum exceeds 15.
hipError_t launch_halves(int grid, const uint8_t* blk, uint8_t* out, unsigned long long n_blocks, unsigned int* bad,
                         hipStream_t s);

// mode 0: random probes; 1: random probes with nontemporal loads; 2: stream the table once.
hipError_t launch_gather(int mode, int grid, const uint32_t* blk, unsigned long long n_blocks,
                         unsigned long long n_probes, unsigned long long* sum, hipStream_t s);

}  // namespace ccs
