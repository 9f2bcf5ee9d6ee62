// CDNA4 (gfx950) kernels of the optimal cube solver: probes, the BFS that
// builds the pruning tables, the table encoders and the IDA* search.
// All integer work on the vector ALUs; lookup tables live in LDS.
#include <hip/hip_runtime.h>

#include <type_traits>

#include "encode.hpp"
#include "kernels.hpp"

namespace ccs {

namespace {

__device__ __forceinline__ void load_lut(CubeLUT* dst, const CubeLUT* __restrict__ src) {
    const uint4* s = reinterpret_cast<const uint4*>(src);
    uint4* d = reinterpret_cast<uint4*>(dst);
    for (uint32_t i = threadIdx.x; i < sizeof(CubeLUT) / 16; i += blockDim.x) d[i] = s[i];
    __syncthreads();
}

__device__ __forceinline__ uint32_t lane_id() { return __lane_id(); }

__device__ __forceinline__ unsigned long long lanemask_lt() {
    uint32_t l = lane_id();
    return l == 0 ? 0ull : (~0ull >> (64 - l));
}

__device__ __forceinline__ uint32_t first_lane_u32(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ unsigned long long first_lane_u64(unsigned long long v) {
    uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return ((unsigned long long)hi << 32) | lo;
}

// Cross-lane state of the persistent form is read and written with relaxed
// agent-scope atomics: on gfx950 they go past the per-XCD L2 (sc1), so they
// are coherent across the eight XCDs without the L2 invalidate/writeback
// that agent-scope acquire/release fences cost.  Ordering between two such
// accesses of one lane comes from waiting for the first to complete.
template <typename T>
__device__ __forceinline__ T ld_c(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <typename T>
__device__ __forceinline__ void st_c(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Wait until this wave's outstanding memory operations have completed.
__device__ __forceinline__ void complete() { __builtin_amdgcn_s_waitcnt(0); }

// ---------------------------------------------------------------------------
// probes

template <int K, int F>
__global__ void __launch_bounds__(256) probe_kernel(const CubeLUT* __restrict__ gL, const uint32_t* __restrict__ classidx,
                                                    const uint8_t* __restrict__ table, const Frame* __restrict__ in,
                                                    int n, int move, unsigned long long* __restrict__ idx_out,
                                                    uint32_t* __restrict__ lb_out, Frame* __restrict__ state_out) {
    __shared__ CubeLUT L;
    load_lut(&L, gL);
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Frame f = in[i];
    if (move >= 0) apply_move(L, f, move);
    if (move <= -2) apply_left(L, f, -2 - move);
    uint64_t idx = table_index<K>(L, classidx, f);
    if (idx_out) idx_out[i] = idx;
    if (lb_out && table) lb_out[i] = table_lb<F>(table, idx);
    if (state_out) state_out[i] = f;
}

template <int K>
__global__ void __launch_bounds__(256) unrank_kernel(const CubeLUT* __restrict__ gL, const uint32_t* __restrict__ classidx,
                                                     const uint32_t* __restrict__ class_rep,
                                                     const unsigned long long* __restrict__ idx, int n,
                                                     unsigned long long* __restrict__ out) {
    __shared__ CubeLUT L;
    load_lut(&L, gL);
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Frame f = unrank<K>(L, class_rep, idx[i]);
    out[i] = table_index<K>(L, classidx, f);
}

// ---------------------------------------------------------------------------
// BFS over the coset space.  The build table holds one exact depth per entry
// in a nibble (8 entries per u32, 0xF = unvisited); the solving formats are
// encoded from it afterwards.  The reference keeps depth mod 3 in 2 bits and
// updates its block table with atomics during the search (kernels.cu:389-611);
// exact nibbles cost 2x the bytes (15 GB for d4t, small next to 288 GB of
// HBM) and turn both encoders into streaming passes.

constexpr uint32_t BFS_WG = 256;

__device__ __forceinline__ uint32_t nib_get(const uint32_t* tbl, uint64_t idx) {
    return (tbl[idx >> 3] >> ((uint32_t)(idx & 7u) * 4u)) & 15u;
}

// Claim entry idx at `depth` if it is unvisited; 1 if this call set it.
__device__ __forceinline__ uint32_t claim_entry(uint32_t* tbl, uint64_t idx, uint32_t depth) {
    uint32_t* wp = tbl + (idx >> 3);
    const uint32_t sh = (uint32_t)(idx & 7u) * 4u;
    if (((*wp >> sh) & 15u) != 15u) return 0;
    const uint32_t old = atomicAnd(wp, ~((15u ^ depth) << sh));
    return ((old >> sh) & 15u) == 15u;
}

// Claim the canonical index of nb and its stabiliser images at `depth`
// (kernels.cu:507-528 claims the same set).  Returns the number newly set.
template <int K>
__device__ uint32_t claim_images(const CubeLUT& L, uint32_t* tbl, const uint32_t* __restrict__ classidx,
                                 const uint16_t* __restrict__ class_stab, const Frame& nb, uint32_t depth) {
    const uint32_t packed = classidx[corner_raw(L, nb)];
    const uint32_t cls = packed >> 4, sym = packed & 15u;
    const uint64_t base = (uint64_t)cls * KindT<K>::N_EDGE;
    uint32_t claimed = 0;
    for (uint32_t stab = class_stab[cls]; stab; stab &= stab - 1u)
        claimed += claim_entry(tbl, base + edge_coord<K>(L, nb, L.s_mul[ctz(stab)][sym]), depth);
    return claimed;
}

// Memory order among the lanes of one wave for its LDS lists (the LDS runs a
// wave's instructions in order; the fences keep the compiler from moving the
// accesses across the barrier).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int K>
__global__ void bfs_seed_kernel(const CubeLUT* __restrict__ gL, uint32_t* tbl, const uint32_t* __restrict__ classidx,
                                const uint16_t* __restrict__ class_stab, unsigned long long* counter) {
    __shared__ CubeLUT L;
    load_lut(&L, gL);
    if (threadIdx.x) return;
    atomicAdd(counter, (unsigned long long)claim_images<K>(L, tbl, classidx, class_stab, solved_frame(), 0u));
}

// BFS layers.  Each wave scans 64 consecutive words (512 entries) at a time,
// grid-strided, and compacts the entries to process into an LDS list, so that
// lanes work on entries rather than on words.

// Top-down layer: expand the entries at `depth` and claim their neighbours at
// depth + 1.  The list carries over from chunk to chunk and is worked off 64
// entries at a time, so sparse layers keep every lane busy.  A neighbour whose
// corner class has a nontrivial stabiliser has several table entries
// (kernels.cu:507-528 claims them all); its first is claimed at once and the
// rest are queued in LDS and claimed 64 neighbours at a time, so the few lanes
// with extra images do not hold up the wave at every move.
constexpr uint32_t FWD_LIST = 576;   // < 64 carried plus one chunk's 512
constexpr uint32_t FWD_QUEUE = 128;  // < 64 carried plus one move's 64
constexpr uint32_t FWD_TASK = 7;     // queued neighbour: frame (5 words), class, sym | images << 8
constexpr uint32_t FWD_WG = 512;     // eight waves share the LUT copy: 4 waves per SIMD fit the LDS

template <int K>
__global__ void __launch_bounds__(FWD_WG) bfs_forward_kernel(const CubeLUT* __restrict__ gL, uint32_t* tbl,
                                                            const uint32_t* __restrict__ classidx,
                                                            const uint32_t* __restrict__ class_rep,
                                                            const uint16_t* __restrict__ class_stab,
                                                            unsigned long long n_words, unsigned long long n_entries,
                                                            uint32_t depth, unsigned long long* counter) {
    __shared__ CubeLUT L;
    __shared__ uint32_t lst[FWD_WG / 64][FWD_LIST];
    __shared__ uint32_t xq[FWD_WG / 64][FWD_QUEUE * FWD_TASK];
    __shared__ uint32_t cnt[FWD_WG / 64], xcnt[FWD_WG / 64];
    load_lut(&L, gL);
    const uint32_t wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const uint32_t next = depth + 1u;
    const uint32_t pat = depth * 0x11111111u;
    unsigned long long found = 0;
    const unsigned long long n_chunks = (n_words + 63) / 64;
    const unsigned long long stride = (unsigned long long)gridDim.x * (FWD_WG / 64);
    const unsigned long long c0 = (unsigned long long)blockIdx.x * (FWD_WG / 64) + wave;
    if (lane == 0) cnt[wave] = xcnt[wave] = 0;
    wave_sync();
    // claim the remaining images of up to 64 queued neighbours, from the top of the queue
    auto drain_queue = [&](uint32_t take) {
        const uint32_t q = xcnt[wave];
        if (lane < take) {
            const uint32_t* t = &xq[wave][(q - take + lane) * FWD_TASK];
            const Frame nb{t[0], t[1], t[2], t[3], t[4]};
            const uint64_t base = (uint64_t)t[5] * KindT<K>::N_EDGE;
            const uint32_t sym = t[6] & 15u;
            for (uint32_t st = t[6] >> 8; st; st &= st - 1u)
                found += claim_entry(tbl, base + edge_coord<K>(L, nb, L.s_mul[ctz(st)][sym]), next);
        }
        wave_sync();
        if (lane == 0) xcnt[wave] = q - take;
        wave_sync();
    };
    for (uint32_t k = 0;; ++k) {
        const unsigned long long chunk = c0 + (unsigned long long)k * stride;
        const bool have = chunk < n_chunks;  // uniform across the wave
        if (have) {
            const unsigned long long w = chunk * 64 + lane;
            uint32_t hits = 0;
            if (w < n_words) {
                const uint32_t x = tbl[w] ^ pat;
                hits = ~(x | (x >> 1) | (x >> 2) | (x >> 3)) & 0x11111111u;
                const unsigned long long first = w * 8u;
                if (first + 8u > n_entries) hits &= (1u << (4u * (uint32_t)(n_entries - first))) - 1u;
            }
            // entry = (chunk number k within this wave, word lane, nibble); k < 2^23
            uint32_t off = atomicAdd(&cnt[wave], popc(hits));
            for (uint32_t h = hits; h; h &= h - 1) lst[wave][off++] = (k << 9) | (lane << 3) | (ctz(h) >> 2);
        }
        wave_sync();
        uint32_t total = cnt[wave];
        const bool last = !have || chunk + stride >= n_chunks;
        while (total >= 64u || (last && total > 0u)) {
            const uint32_t take = min(total, 64u);
            const bool act = lane < take;
            Frame s{};
            if (act) {
                const uint32_t e = lst[wave][total - take + lane];
                const unsigned long long ch = c0 + (unsigned long long)(e >> 9) * stride;
                s = unrank<K>(L, class_rep, (ch * 64u + ((e >> 3) & 63u)) * 8u + (e & 7u));
            }
#pragma unroll 1
            for (int m = 0; m < 18; ++m) {
                if (act) {
                    Frame nb = s;
                    apply_move(L, nb, m);
                    const uint32_t packed = classidx[corner_raw(L, nb)];
                    const uint32_t cls = packed >> 4, sym = packed & 15u;
                    const uint32_t stab = class_stab[cls];
                    found += claim_entry(tbl, (uint64_t)cls * KindT<K>::N_EDGE +
                                                  edge_coord<K>(L, nb, L.s_mul[ctz(stab)][sym]), next);
                    const uint32_t rest = stab & (stab - 1u);
                    if (rest) {
                        uint32_t* t = &xq[wave][atomicAdd(&xcnt[wave], 1u) * FWD_TASK];
                        t[0] = nb.c0, t[1] = nb.c1, t[2] = nb.e0, t[3] = nb.e1, t[4] = nb.e2;
                        t[5] = cls, t[6] = sym | (rest << 8);
                    }
                }
                wave_sync();
                if (xcnt[wave] >= 64u) drain_queue(64u);
            }
            total -= take;
        }
        if (lane == 0) cnt[wave] = total;
        wave_sync();
        if (last) break;
    }
    drain_queue(xcnt[wave]);
    // one atomic per wave
    for (int o = 32; o > 0; o >>= 1) found += __shfl_xor(found, o);
    if (lane == 0 && found) atomicAdd(counter, found);
}

// Bottom-up layer: an unvisited entry with a neighbour at `depth` is at
// depth + 1.  A lane stops at an entry's first such neighbour, so entries take
// 1 to 18 moves; lanes that finish take the next entry of the chunk's list
// once a quarter of the wave is idle, instead of every lane waiting for the
// wave's slowest entry.  Each lane then writes back its word of the chunk.
template <int K>
__global__ void __launch_bounds__(BFS_WG) bfs_backward_kernel(const CubeLUT* __restrict__ gL, uint32_t* tbl,
                                                             const uint32_t* __restrict__ classidx,
                                                             const uint32_t* __restrict__ class_rep,
                                                             unsigned long long n_words, unsigned long long n_entries,
                                                             uint32_t depth, unsigned long long* counter) {
    __shared__ CubeLUT L;
    __shared__ uint16_t lst[BFS_WG / 64][512];
    __shared__ uint32_t cnt[BFS_WG / 64];
    __shared__ uint32_t clr[BFS_WG / 64][64];
    load_lut(&L, gL);
    const uint32_t wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const uint32_t next = depth + 1u;
    unsigned long long found = 0;
    const unsigned long long n_chunks = (n_words + 63) / 64;
    const unsigned long long stride = (unsigned long long)gridDim.x * (BFS_WG / 64);
    for (unsigned long long chunk = (unsigned long long)blockIdx.x * (BFS_WG / 64) + wave; chunk < n_chunks;
         chunk += stride) {
        const unsigned long long w = chunk * 64 + lane;
        uint32_t word = 0, hits = 0;
        if (w < n_words) {
            word = tbl[w];
            hits = word & (word >> 1) & (word >> 2) & (word >> 3) & 0x11111111u;  // unvisited (15)
            const unsigned long long first = w * 8u;
            if (first + 8u > n_entries) hits &= (1u << (4u * (uint32_t)(n_entries - first))) - 1u;
        }
        if (lane == 0) cnt[wave] = 0;
        clr[wave][lane] = 0;
        wave_sync();
        uint32_t off = atomicAdd(&cnt[wave], popc(hits));
        for (uint32_t h = hits; h; h &= h - 1) lst[wave][off++] = (uint16_t)((lane << 3) | (ctz(h) >> 2));
        wave_sync();
        const uint32_t total = cnt[wave];
        uint32_t handed = 0, e = 0;
        int m = 0;
        bool act = false;
        Frame s{};
        for (;;) {
            const bool need = !act;
            const unsigned long long nm = __ballot(need);
            const uint32_t pos = handed + __popcll(nm & lanemask_lt());
            if (need && pos < total) {
                e = lst[wave][pos];
                s = unrank<K>(L, class_rep, (chunk * 64 + (e >> 3)) * 8u + (e & 7u));
                m = 0;
                act = true;
            }
            handed = min(total, handed + (uint32_t)__popcll(nm));
            if (!__any(act)) break;
            const uint32_t keep = handed < total ? 48u : 0u;  // refill below this many busy lanes
            do {
                if (act) {
                    Frame nb = s;
                    apply_move(L, nb, m);
                    if (nib_get(tbl, table_index<K>(L, classidx, nb)) == depth) {
                        atomicOr(&clr[wave][e >> 3], (15u ^ next) << ((e & 7u) * 4u));
                        ++found;
                        act = false;
                    } else if (++m == 18) {
                        act = false;
                    }
                }
            } while ((uint32_t)__popcll(__ballot(act)) > keep);
        }
        wave_sync();
        if (w < n_words) {
            const uint32_t c = clr[wave][lane];
            if (c) tbl[w] = word & ~c;  // this lane owns word w during a bottom-up layer
        }
        wave_sync();
    }
    for (int o = 32; o > 0; o >>= 1) found += __shfl_xor(found, o);
    if (lane == 0 && found) atomicAdd(counter, found);
}

// Histogram of the nibble values over the valid entries (the reference's
// count_values self-check, kernels.cu:614-624).  Per word and value k, the
// nibbles equal to k are marked by a zero test on word ^ (k * 0x11111111) and
// counted with one popcount.
__global__ void __launch_bounds__(256) histogram_kernel(const uint32_t* __restrict__ tbl, unsigned long long n_words,
                                                        unsigned long long n_entries,
                                                        unsigned long long* __restrict__ counts) {
    uint32_t local[16];
#pragma unroll
    for (int v = 0; v < 16; ++v) local[v] = 0;
    for (unsigned long long w = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; w < n_words;
         w += (unsigned long long)gridDim.x * blockDim.x) {
        const uint32_t word = tbl[w];
        uint32_t keep = 0x11111111u;  // one bit per valid nibble
        if (w * 8u + 8u > n_entries) keep &= (1u << (4u * (uint32_t)(n_entries - w * 8u))) - 1u;
#pragma unroll
        for (uint32_t k = 0; k < 16; ++k) {
            const uint32_t z = word ^ (k * 0x11111111u);
            local[k] += popc(~(z | (z >> 1) | (z >> 2) | (z >> 3)) & keep);
        }
    }
#pragma unroll
    for (int v = 0; v < 16; ++v) {
        unsigned long long s = local[v];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o);
        if (lane_id() == 0 && s) atomicAdd(&counts[v], s);
    }
}

// Encoders.  One wave encodes 64 consecutive output blocks.  The wave first
// copies its span of the build table to LDS with 16-byte loads (the span of 64
// blocks starts on a 16-byte boundary because the host launches from block
// numbers that are multiples of 64), then each lane encodes one block from
// registers (encode.hpp).  The table's last block, if partial, takes the
// per-entry path below.

// Copy words [W, W + 256 * R) of tbl (zero past n_words) to LDS, 16 bytes a lane.
template <int R>
__device__ __forceinline__ void stage_span(const uint32_t* __restrict__ tbl, unsigned long long n_words,
                                           unsigned long long W, uint32_t* st, uint32_t lane) {
#pragma unroll
    for (int r = 0; r < R; ++r) {
        const unsigned long long wi = W + (unsigned long long)(r * 64 + (int)lane) * 4u;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (wi + 3u < n_words) {
            v = *reinterpret_cast<const uint4*>(tbl + wi);
        } else {
            if (wi < n_words) v.x = tbl[wi];
            if (wi + 1u < n_words) v.y = tbl[wi + 1u];
            if (wi + 2u < n_words) v.z = tbl[wi + 2u];
        }
        reinterpret_cast<uint4*>(st)[r * 64 + (int)lane] = v;
    }
}

// Block format, one block of fewer than 124 entries (the table's last).
__device__ void encode_block_partial(const uint32_t* __restrict__ tbl, unsigned long long first, uint32_t n,
                                     uint32_t* __restrict__ out, uint8_t* __restrict__ mn_out,
                                     unsigned int* __restrict__ bad) {
    uint32_t m = 15;
    for (uint32_t j = 0; j < n; ++j) m = min(m, nib_get(tbl, first + j));
    uint32_t w[8] = {m, 0, 0, 0, 0, 0, 0, 0};
    uint32_t unset = m == 15u;
    for (uint32_t j = 0; j < n; ++j) {
        uint32_t v = nib_get(tbl, first + j);
        unset |= v == 15u;
        uint32_t pos = 8u + 2u * j;
        w[pos >> 5] |= min(v - m, 3u) << (pos & 31u);
    }
    if (unset) atomicOr(bad, 1u);
    *mn_out = (uint8_t)m;
    if (!out) return;
#pragma unroll
    for (int k = 0; k < 8; ++k) out[k] = w[k];
}

// Block format (kernels.cu:396-406, 485-503 define it): per 124 entries, u8
// minimum then 2-bit offsets min(v - min, 3) at bit 8 + 2j.  Blocks b0 .. b0 + nb - 1.
// Each block's minimum also goes to mins[block] (the base-3 encoder reads them);
// with out == nullptr only the minima are written.
__global__ void __launch_bounds__(256) encode_block_kernel(const uint32_t* __restrict__ tbl, unsigned long long n_words,
                                                           unsigned long long n_entries, unsigned long long b0,
                                                           unsigned long long nb, uint32_t* __restrict__ out,
                                                           uint8_t* __restrict__ mins, unsigned int* __restrict__ bad) {
    __shared__ uint32_t stage[4][1024];  // 64 blocks x 15.5 words, plus the word after
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const unsigned long long i0 = ((unsigned long long)blockIdx.x * 4u + wave) * 64u;  // wave's first block - b0
    const unsigned long long gb0 = b0 + i0;
    if (i0 < nb) stage_span<4>(tbl, n_words, gb0 / 2u * 31u, stage[wave], lane);  // word (gb0 * 124) / 8
    __syncthreads();
    const unsigned long long i = i0 + lane;
    if (i >= nb) return;
    const unsigned long long first = (gb0 + lane) * BLK_ENTRIES;
    uint32_t* o = out ? out + i * 8u : nullptr;
    if (n_entries - first < BLK_ENTRIES) {
        encode_block_partial(tbl, first, (uint32_t)(n_entries - first), o, mins + gb0 + lane, bad);
        return;
    }
    const uint32_t* st = stage[wave] + ((lane * 31u) >> 1);  // block's first word within the span
    const uint32_t sh = (lane & 1u) * 16u;                   // its first nibble: 0 or 4
    uint32_t a[16], w[8];
#pragma unroll
    for (int k = 0; k < 16; ++k) a[k] = funnel(st[k + 1], st[k], sh);
    if (encode_block_words(a, w)) atomicOr(bad, 1u);
    mins[gb0 + lane] = (uint8_t)w[0];
    if (!o) return;
    reinterpret_cast<uint4*>(o)[0] = make_uint4(w[0], w[1], w[2], w[3]);
    reinterpret_cast<uint4*>(o)[1] = make_uint4(w[4], w[5], w[6], w[7]);
}

// Base-3 format (kernels.cu:1839-1869 define it): per 155 entries, header
// m | c << 6 with the cut c in 0..3 maximising the sum of encoded bounds
// (lowest c on ties), then 31 bytes of five digits, little-endian base 3.
__global__ void __launch_bounds__(256) encode_b3_kernel(const uint32_t* __restrict__ tbl, unsigned long long n_words,
                                                        const uint8_t* __restrict__ mins, unsigned long long n_entries,
                                                        unsigned long long b0, unsigned long long nb,
                                                        uint32_t* __restrict__ out) {
    __shared__ uint32_t stage[4][1280];  // 64 blocks x 19.375 words, plus the word after
    const uint32_t wave = threadIdx.x >> 6, lane = threadIdx.x & 63u;
    const unsigned long long i0 = ((unsigned long long)blockIdx.x * 4u + wave) * 64u;
    const unsigned long long gb0 = b0 + i0;
    if (i0 < nb) stage_span<5>(tbl, n_words, gb0 / 8u * 155u, stage[wave], lane);  // word (gb0 * 155) / 8
    __syncthreads();
    const unsigned long long i = i0 + lane;
    if (i >= nb) return;
    const unsigned long long first = (gb0 + lane) * B3_ENTRIES;
    const uint32_t n = (uint32_t)min((unsigned long long)B3_ENTRIES, n_entries - first);
    uint32_t* o = out + i * 8u;
    // the block's entries meet at most three 124-entry blocks of the block format
    const unsigned long long bl = first / BLK_ENTRIES;
    const uint32_t j1 = (uint32_t)((bl + 1u) * BLK_ENTRIES - first), j2 = j1 + BLK_ENTRIES;
    const uint32_t mn0 = mins[bl];
    const uint32_t mn1 = j1 < n ? mins[bl + 1u] : 0u;
    const uint32_t mn2 = j2 < n ? mins[bl + 2u] : 0u;
    auto put = [&](uint32_t k, uint32_t word) __attribute__((always_inline)) { o[k] = word; };
    if (n == B3_ENTRIES) {
        const uint32_t off = lane * B3_ENTRIES;  // block's first nibble within the span
        const uint32_t* st = stage[wave];
        auto nib = [&](uint32_t j) __attribute__((always_inline)) { return (st[(off + j) >> 3] >> (((off + j) & 7u) * 4u)) & 15u; };
        encode_b3_block(nib, B3_ENTRIES, mn0, mn1, mn2, j1, j2, put);
    } else {  // the table's last block: read the build table directly
        auto nib = [&](uint32_t j) __attribute__((always_inline)) { return nib_get(tbl, first + j); };
        encode_b3_block(nib, n, mn0, mn1, mn2, j1, j2, put);
    }
}

// ---------------------------------------------------------------------------
// IDA* search: the reference's step machine (kernels.cu:1447-1801) on wave64.
// Each outer iteration every lane either probes its current child m (its next
// forward and next inverse frame, two table reads in flight) or idles; control
// work (push, pick the next canonical child, pop, fetch a work item, replay a
// prefix move) runs between probes, one pass per step in the common cases.
// The per-lane DFS stack (parent bounds, moves) lives in LDS, lane-major.
//
// h and child values are packed 6 x 4 bits in one register.  Table values
// never exceed 15 (the deepest coset table reaches 14), and a value clamped to
// 15 would still be a lower bound; a stack level is then one u32: the bounds
// in bits 0..23 and the move in bits 24..28.

__device__ __forceinline__ uint32_t hget(uint32_t hp, int j) { return (hp >> (4 * j)) & 15u; }
__device__ __forceinline__ uint32_t hset(uint32_t hp, int j, uint32_t v) {
    return (hp & ~(15u << (4 * j))) | (min(v, 15u) << (4 * j));
}
constexpr uint32_t STK_H = 0xFFFFFFu;  // stack word: packed bounds
__device__ __forceinline__ int stk_move(uint32_t w) { return (int)(w >> 24); }
// Axes whose first move is excluded: an inverse frame at the budget
// (kernels.cu:750-763).
__device__ __forceinline__ uint32_t excl_axes(uint32_t hp, uint32_t budget) {
    return (uint32_t)(hget(hp, 3) == budget) | ((uint32_t)(hget(hp, 4) == budget) << 1) |
           ((uint32_t)(hget(hp, 5) == budget) << 2);
}
__device__ __forceinline__ uint32_t axis_mask(uint32_t axes) {
    return ((axes & 1u) ? 0x00E07u : 0u) | ((axes & 2u) ? 0x07038u : 0u) | ((axes & 4u) ? 0x381C0u : 0u);
}
// All three forward (or inverse) frames at the budget rule out every last
// (first) move (kernels.cu:765-774).
__device__ __forceinline__ bool all_tight(uint32_t cp, uint32_t budget) {
    if (budget == 0 || budget > 15u) return false;  // no value reaches 16
    uint32_t b3 = budget * 0x111u;
    return (cp & 0xFFFu) == b3 || ((cp >> 12) & 0xFFFu) == b3;
}
// Frame probe order (kernels.cu:1398-1410 orders all six frames: at the budget
// edge, one below, the rest).  Here a step probes one forward and one inverse
// frame together, so the order is kept per group: forward frame j at bit
// 4 * class + j, inverse frame 3 + j at bit 16 + 4 * class + j (class 0 = at
// the edge, 1 = one below, 2 = the rest); the lowest set bit goes first.  The
// order decides only which probes a pruned child costs, never the nodes.
__device__ __forceinline__ uint32_t frame_order(uint32_t hp, uint32_t edge) {
    uint32_t o = 0;
#pragma unroll
    for (int j = 0; j < 6; ++j) {
        uint32_t h = hget(hp, j);
        uint32_t sh = h >= edge ? 0u : (h + 1u == edge ? 4u : 8u);
        o |= 1u << (sh + (uint32_t)(j % 3) + (j < 3 ? 0u : 16u));
    }
    return o;
}
// Axis (U/D = 0, R/L = 1, F/B = 2) of HTM move m < 18: face m / 3, mod 3.
__device__ __forceinline__ uint32_t move_axis(int m) {
    uint32_t face = ((uint32_t)m * 11u) >> 5;
    return face >= 3u ? face - 3u : face;
}
template <int I, int N, typename Fn>
__device__ __forceinline__ void static_for(Fn&& fn) {
    if constexpr (I < N) {
        fn(std::integral_constant<int, I>{});
        static_for<I + 1, N>(fn);
    }
}
__device__ __forceinline__ Frame load_frame(const Frame* p) {
    const uint32_t* w = reinterpret_cast<const uint32_t*>(p);
    return Frame{w[0], w[1], w[2], w[3], w[4]};
}
// Ticket of the persistent form: one 64-bit word per cube packs the round
// (bound, prefix length), the items handed out and the items in flight, so a
// single atomic both takes items and keeps the round open.
constexpr unsigned long long T_NEXT_MASK = (1ull << 28) - 1;
constexpr unsigned long long T_INF_ONE = 1ull << 28;
constexpr unsigned long long T_INF_MASK = ((1ull << 21) - 1) << 28;
constexpr int T_BOUND_SHIFT = 49;
constexpr int T_PLEN_SHIFT = 54;
constexpr unsigned long long T_CLOSING = 1ull << 57;
constexpr unsigned long long T_RETIRED = 1ull << 58;  // serving: the cube's slot is free again
constexpr int T_CHUNK_SHIFT = 59;
constexpr unsigned long long T_BANDED = 1ull << 62;  // the round visits its items in N_BANDS key bands
// Serving: a fetching wave admits new cubes while fewer rounds than this are queued.
constexpr unsigned long long ADMIT_BELOW = 256;
// Serving: while at most this many cubes are in the server, lanes out of work
// ask for donated subtrees and chunks aim at 64x the serving chunk size (a
// lone hard cube then gets the whole GPU, as in a batch).
constexpr uint32_t SERVE_DONATE_LIVE = 4;
// Serving: a cube's turn at the head of the ring lasts 1 + age / age_ticks
// fetches, at most this many, so that cubes that have waited longest get a
// larger share of the lanes (the round robin's plain equal shares made every
// cube's latency grow with the number of cubes in the server).
constexpr uint32_t SERVE_MAX_CREDIT = 16;
__device__ __forceinline__ uint32_t serve_credit(const SearchArgs& A, const SearchJob* J) {
    const unsigned long long age = (unsigned long long)wall_clock64() - ld_c(&J->t_admit);
    const unsigned long long c = 1ull + age / (A.age_ticks ? A.age_ticks : 1ull);
    return (uint32_t)(c < SERVE_MAX_CREDIT ? c : SERVE_MAX_CREDIT);
}
// Ring entry: (slot mod 2^27) << 37 | round bound << 32 | cube.  The slot part
// tells a reader that the entry at its head has been published; the bound
// tells a stale entry (of a round that has since closed) from the cube's live one.
constexpr unsigned long long RING_SEQ_MASK = (1ull << 27) - 1;
__device__ __forceinline__ unsigned long long ring_entry(unsigned long long slot, uint32_t bound, uint32_t cube) {
    return ((slot & RING_SEQ_MASK) << 37) | ((unsigned long long)bound << 32) | cube;
}
__device__ __forceinline__ void ring_push(const SearchArgs& A, uint32_t bound, uint32_t cube) {
    unsigned long long slot = atomicAdd(&A.q->tail, 1ull);
    st_c(&A.ring[slot % A.ring_size], ring_entry(slot, bound, cube));
}
// Chunk layout of the persistent form.  Chunk q of a round with 2^lg items per
// chunk and Q = ceil(n / 2^lg) chunks takes blocks of ITEM_BLOCK consecutive
// items spaced Q blocks apart (blocks q, q + Q, q + 2Q, ...).  The lanes that
// hold a cube's chunks then sweep its items roughly in order, so the round that
// finds the solution stops near where a sequential search would; with
// contiguous chunks every chunk was the same fraction done when the solution
// turned up, half the round on average.  Consecutive items within a block
// still share their leading prefix moves.
constexpr uint32_t ITEM_BLOCK = 8;
constexpr uint32_t CL_BANDED = 1u << 16;  // chunk_left: the round is banded (visits, not items)
constexpr uint32_t CL_KEYS = 1u << 17;    // chunk_left: record the items' keys
__device__ __forceinline__ uint32_t next_item(uint32_t item, uint32_t left, uint32_t lg, uint32_t n) {
    const uint32_t blk = min(ITEM_BLOCK, 1u << lg);
    if (left & (blk - 1u)) return item + 1u;  // within a block
    const uint32_t q = (n + (1u << lg) - 1u) >> lg;
    return item + (q - 1u) * blk + 1u;        // first item of the chunk's next block
}

__device__ __forceinline__ uint32_t t_next(unsigned long long t) { return (uint32_t)(t & T_NEXT_MASK); }
__device__ __forceinline__ uint32_t t_inflight(unsigned long long t) { return (uint32_t)((t & T_INF_MASK) >> 28); }
__device__ __forceinline__ uint32_t t_bound(unsigned long long t) { return (uint32_t)(t >> T_BOUND_SHIFT) & 31u; }
__device__ __forceinline__ uint32_t t_plen(unsigned long long t) { return (uint32_t)(t >> T_PLEN_SHIFT) & 7u; }
__device__ __forceinline__ uint32_t t_chunk(unsigned long long t) { return (uint32_t)(t >> T_CHUNK_SHIFT) & 7u; }
__device__ __forceinline__ unsigned long long make_ticket(unsigned long long keep, uint32_t bound, uint32_t plen,
                                                          uint32_t chunk_log, bool banded = false) {
    return (keep & T_INF_MASK) | ((unsigned long long)bound << T_BOUND_SHIFT) |
           ((unsigned long long)plen << T_PLEN_SHIFT) | ((unsigned long long)chunk_log << T_CHUNK_SHIFT) |
           (banded ? T_BANDED : 0ull);
}
// Fetch slots of a round: its items; when it is banded, each band's items
// rounded up to whole chunks, once per band.  Visit v of a banded round is item
// v mod n of band v / n (an item is searched in the visit of its band); the
// chunks of a band are laid out over its items as a plain round's are.
__device__ __forceinline__ uint32_t t_visits(const SearchArgs& A, unsigned long long t) {
    const uint32_t n = A.pre_n[t_plen(t)], lg = t_chunk(t);
    return (t & T_BANDED) ? N_BANDS * (((n + (1u << lg) - 1u) >> lg) << lg) : n;
}
// Whether an item key (0: unknown, searched in the last band) is in band b,
// whose keys are those above lo and up to hi (the cube's band bounds b - 1
// and b: N_BANDS - 1 non-decreasing bytes, four to a word).
__device__ __forceinline__ bool in_band(uint32_t key, uint32_t b, uint32_t lo, uint32_t hi) {
    return key == 0u ? b == N_BANDS - 1u : (lo < key && key <= hi);
}
// Items per fetch for a round: about chunk_nodes nodes per chunk, estimated from
// the previous round's nodes and the usual growth of ~13x per bound, but at
// least 1.5 chunks per lane of this cube's share of the launch's lanes (and
// never fewer than 256), so that the last cubes of a batch still fill the GPU.
// Finer chunks cost more in fetches than they save at a round's end: with at
// least 4096 chunks per round, a 4,096-cube batch fetched 4-move rounds four
// items at a time and ran at 415 against 564 cubes/s (256 cubes: 192 against
// 242); with 16 chunks per lane, serving fell from 140 to 87 cubes/s at 150
// QPS.
__device__ __forceinline__ uint32_t chunk_log_for(const SearchArgs& A, unsigned long long last_nodes,
                                                  uint32_t n_items, bool banded = false) {
    unsigned long long per_item = (last_nodes * 13u) / (n_items ? n_items : 1u) + 1u;
    // cubes admitted and not yet finished (u32 differences stay right across wraparound)
    unsigned long long live = A.serving ? (unsigned long long)((uint32_t)ld_c(&A.q->mb_head) - ld_c(&A.q->n_final))
                                        : (unsigned long long)(A.q->n_jobs - ld_c(&A.q->n_final));
    const unsigned long long share = A.n_lanes / (live ? live : 1ull);  // the cube's lanes
    const unsigned long long want = share + share / 2 < 256ull ? 256ull : share + share / 2;
    const unsigned long long target = A.serving && live <= SERVE_DONATE_LIVE ? A.chunk_nodes << 6 : A.chunk_nodes;
    // A banded round's chunks hold at most 32, 16 or 8 visits (about 2, 1 or
    // 0.5 of their band's items).  Its first bands hold the heavy items; in
    // chunks of 64 visits they left a cube's lanes waiting on a few long
    // chunks at every round's end (lanes without work: 20 of 64 per step on
    // 512 cubes, against 9 in move order).  The more lanes a cube has, the
    // more that costs: 32 visits while a cube has at most 64 lanes, 16 up to
    // 128, 8 beyond.  Rounds in move order take at most 128 items per chunk
    // while a cube has at most 512 lanes, 32 up to 2,048, 16 beyond (4,096
    // cubes: 575-591 against 555-564 cubes/s with 64 items; small batches
    // and the last cubes of a large one: 0-3% faster with 32 or 16).
    const uint32_t cap = !banded ? (share <= 512ull ? 7u : share <= 2048ull ? 5u : 4u)
                                 : share <= 64ull ? 5u : share <= 128ull ? 4u : 3u;
    uint32_t lg = 0;
    while (lg < cap && (per_item << (lg + 1)) <= target && (unsigned long long)(n_items >> (lg + 1)) >= want) ++lg;
    return lg;
}

// A cube's band bounds for its next round from the key histogram of the round
// that ends (if `want`: the round was long and the prefix length stays), and
// the histogram cleared.  Out of line: it runs once per round, and its
// working set stays out of the search loop's registers.
__device__ __attribute__((noinline)) bool band_bounds(uint32_t* hist, SearchJob* J, bool want) {
    bool banded = false;
    if (want) {
        uint32_t total = 0;
        for (uint32_t v = 1; v < KEY_BINS; ++v) total += ld_c(&hist[v]);
        if (total >= 64u) {
            uint32_t acc = 0, b = 0, word = 0;
            for (uint32_t v = 1; v < KEY_BINS && b < N_BANDS - 1; ++v) {
                acc += ld_c(&hist[v]);
                while (b < N_BANDS - 1 && (unsigned long long)acc * N_BANDS >= (unsigned long long)total * (b + 1)) {
                    word |= v << (8u * (b & 3u));
                    if ((++b & 3u) == 0u) {
                        st_c(&J->band_thr[(b >> 2) - 1u], word);
                        word = 0;
                    }
                }
            }
            for (; b < N_BANDS; ++b) {  // bounds never reached, and the pad byte
                word |= 255u << (8u * (b & 3u));
                if (((b + 1u) & 3u) == 0u) {
                    st_c(&J->band_thr[b >> 2], word);
                    word = 0;
                }
            }
            banded = true;
        }
    }
    for (uint32_t v = 0; v < KEY_BINS; ++v) st_c(&hist[v], 0u);
    return banded;
}

__device__ void close_round(const SearchArgs& A, uint32_t c, unsigned long long tk) {
    SearchJob* J = &A.jobs[c];
    uint32_t bound = t_bound(tk);
    unsigned long long nodes = ld_c(&J->nodes);
    uint32_t r = ld_c(&J->n_rounds);
    if (r < (uint32_t)MAX_BOUND) {
        st_c(&J->round_bound[r], bound);
        st_c(&J->round_nodes[r], nodes);
        st_c(&J->n_rounds, r + 1);
    }
    if (ld_c(&J->found)) return;  // the finder already made the cube final
    if ((int)bound + 1 > A.max_bound) {
        st_c(&J->status, 2u);
        st_c(&J->t_final, (unsigned long long)wall_clock64());
        complete();
        atomicAdd(&A.q->n_final, 1u);
        return;
    }
    uint32_t nb = bound + 1;
    uint32_t p = (uint32_t)A.prefix_len + (nodes >= A.deeper0) + (nodes >= A.deeper1);
    uint32_t plen = min(min(nb, p), (uint32_t)MAX_PLEN);
    // Item order of the next round.  The key of an item (the bounds at its
    // prefix's end) does not depend on the round's bound, so the keys this
    // round recorded order the next one when the prefix length stays: items
    // with lower bounds at their prefix's end hold a solution more often, and
    // going through them first cuts the final round's nodes.  Bands of equal
    // sampled counts stand in for a sort; only long rounds are banded, where
    // the visits of other bands' items cost little.
    bool banded = false;
    if (A.key_hist)
        banded = band_bounds(A.key_hist + (size_t)c * KEY_BINS, J,
                             plen == t_plen(tk) && A.keys[plen] && nodes >= BAND_MIN_NODES_PER_ITEM * A.pre_n[plen] &&
                                 A.q->n_jobs - ld_c(&A.q->n_final) >= BAND_MIN_LIVE);
    uint32_t clog = chunk_log_for(A, nodes, A.pre_n[plen] * (banded ? N_BANDS : 1u), banded);
    st_c(&J->nodes, 0ull);
    st_c(&J->bound, nb);
    if (A.serving) st_c(&J->credit, serve_credit(A, J));
    complete();
    unsigned long long cur = ld_c(&J->ticket);
    for (;;) {  // keep the in-flight count of late fetchers, reset the item cursor
        unsigned long long prev = atomicCAS(&J->ticket, cur, make_ticket(cur, nb, plen, clog, banded));
        if (prev == cur) break;
        cur = prev;
    }
    ring_push(A, nb, c);
    st_c(&A.q->split_want, 0u);  // new work in the ring: no more donations needed
}

__device__ __forceinline__ bool round_drained(const SearchArgs& A, unsigned long long t) {
    return t_inflight(t) == 0 && !(t & (T_CLOSING | T_RETIRED)) && t_next(t) >= t_visits(A, t);
}

template <typename T>
__device__ __forceinline__ T sys_load(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}
template <typename T>
__device__ __forceinline__ void sys_store(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

// Serving: post the result of a final cube whose last in-flight item has
// drained; its slot can then take a new cube.
__device__ void post_done(const SearchArgs& A, uint32_t c) {
    SearchJob* J = &A.jobs[c];
    unsigned long long idx = atomicAdd(&A.q->done_next, 1ull);
    ServeDone* d = &A.done[idx % A.done_size];
    uint32_t found = ld_c(&J->found);
    uint32_t len = ld_c(&J->length);
    int32_t length = found ? (int32_t)len : -1;
    unsigned long long nodes = 0;
    uint32_t nr = ld_c(&J->n_rounds);
    for (uint32_t r = 0; r < nr && r < (uint32_t)MAX_BOUND; ++r) nodes += ld_c(&J->round_nodes[r]);
    if (found && (nr == 0 || ld_c(&J->round_bound[nr - 1]) != len)) nodes += ld_c(&J->nodes);
    sys_store(&d->tag, ld_c(&J->tag));
    sys_store(&d->slot, c);
    sys_store(&d->length, length);
    const uint32_t* sol = reinterpret_cast<const uint32_t*>(J->solution);
    uint32_t* dm = reinterpret_cast<uint32_t*>(d->moves);
    for (int w = 0; w < MAX_BOUND / 4; ++w) sys_store(&dm[w], ld_c(&sol[w]));
    sys_store(&d->nodes, nodes);
    sys_store(&d->t_admit, ld_c(&J->t_admit));
    sys_store(&d->t_final, ld_c(&J->t_final));
    __hip_atomic_store(&d->seq, idx + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

__device__ void try_retire(const SearchArgs& A, uint32_t c) {
    SearchJob* J = &A.jobs[c];
    if (!ld_c(&J->found) && ld_c(&J->status) != 2u) return;
    unsigned long long t = ld_c(&J->ticket);
    while (t_inflight(t) == 0 && !(t & T_RETIRED)) {
        unsigned long long prev = atomicCAS(&J->ticket, t, t | T_RETIRED);
        if (prev == t) {
            post_done(A, c);
            return;
        }
        t = prev;
    }
}

// Give back n in-flight items of cube c; the release that drains a fully
// handed-out round closes it (exactly one CAS wins).
__device__ void release_items(const SearchArgs& A, uint32_t c, uint32_t n) {
    SearchJob* J = &A.jobs[c];
    unsigned long long now = atomicAdd(&J->ticket, 0ull - (unsigned long long)n * T_INF_ONE) -
                             (unsigned long long)n * T_INF_ONE;
    while (round_drained(A, now)) {
        unsigned long long prev = atomicCAS(&J->ticket, now, now | T_CLOSING);
        if (prev == now) {
            close_round(A, c, now);
            break;
        }
        now = prev;
    }
    if (A.serving) try_retire(A, c);
}

__device__ __forceinline__ void report_solution(const SearchArgs& A, SearchJob* J, bool queue, int g,
                                                const uint32_t* stk, uint32_t stride) {
    if (atomicCAS(&J->found, 0u, 1u) != 0u) return;
    st_c(&J->length, (uint32_t)g);
    uint32_t* sol = reinterpret_cast<uint32_t*>(J->solution);
    for (int w = 0; w < MAX_BOUND / 4; ++w) {
        uint32_t word = 0;
        for (int b = 0; b < 4; ++b)
            if (4 * w + b < g) word |= (uint32_t)stk_move(stk[(4 * w + b) * stride]) << (8 * b);
        st_c(&sol[w], word);
    }
    if (queue) {
        st_c(&J->status, 1u);
        st_c(&J->t_final, (unsigned long long)wall_clock64());
        complete();
        atomicAdd(&A.q->n_final, 1u);
    }
}

// Serving: take one mailbox record into its slot: root bounds of the six
// frames, a fresh round ticket (keeping the in-flight count of late
// fetchers), and a ring entry.  Returns false when the mailbox is empty.
//
// The tail lives in host memory, so reading it crosses the bus: idle waves
// poll here, and the whole grid reads it at most once per mb_poll_ticks,
// keeping the last value read in mb_seen.  The read is relaxed: the record
// itself is read with system-scope loads that bypass the caches, and an
// acquire at system scope would invalidate the XCD's L2 on every poll.
template <int K, int F>
__device__ bool admit_one(const SearchArgs& A, const CubeLUT& L) {
    unsigned long long h = ld_c(&A.q->mb_head);
    unsigned long long tail = ld_c(&A.q->mb_seen);
    if (h >= tail) {
        const unsigned long long now = wall_clock64(), last = ld_c(&A.q->mb_polled);
        if (now - last < A.mb_poll_ticks || atomicCAS(&A.q->mb_polled, last, now) != last) return false;
        tail = sys_load(A.mb_tail);
        atomicMax(&A.q->mb_seen, tail);
        if (h >= tail) return false;
    }
    if (atomicCAS(&A.q->mb_head, h, h + 1) != h) return true;  // another wave took it
    const ServeRec* r = &A.mb[h % A.mb_size];
    uint32_t c = sys_load(&r->slot);
    SearchJob* J = &A.jobs[c];
    uint32_t bound = 0;
    for (int j = 0; j < 6; ++j) {
        const uint32_t* w = reinterpret_cast<const uint32_t*>(&r->frame[j]);
        Frame f{sys_load(&w[0]), sys_load(&w[1]), sys_load(&w[2]), sys_load(&w[3]), sys_load(&w[4])};
        uint32_t* dst = reinterpret_cast<uint32_t*>(&J->frame[j]);
        st_c(&dst[0], f.c0);
        st_c(&dst[1], f.c1);
        st_c(&dst[2], f.e0);
        st_c(&dst[3], f.e1);
        st_c(&dst[4], f.e2);
        uint32_t v = frame_bound<K, F>(L, A.classidx, A.table, f);
        st_c(&J->h[j], v);
        bound = max(bound, v);
    }
    st_c(&J->tag, sys_load(&r->tag));
    st_c(&J->bound, bound);
    st_c(&J->length, 0u);
    st_c(&J->n_rounds, 0u);
    st_c(&J->t_admit, (unsigned long long)wall_clock64());
    st_c(&J->credit, 1u);
    st_c(&J->nodes, 0ull);
    st_c(&J->status, 0u);
    st_c(&J->found, 0u);
    uint32_t plen = min(min(bound, (uint32_t)A.prefix_len), (uint32_t)MAX_PLEN);
    complete();
    unsigned long long cur = ld_c(&J->ticket);
    for (;;) {
        unsigned long long prev = atomicCAS(&J->ticket, cur, make_ticket(cur, bound, plen, chunk_log_for(A, 0, A.pre_n[plen])));
        if (prev == cur) break;
        cur = prev;
    }
    ring_push(A, bound, c);
    st_c(&A.q->split_want, 0u);
    return true;
}

// A lane finished its chunk of items: add its nodes to the round, wait for
// that add, then give back its in-flight share of the ticket (the closer
// reads the round's nodes).
__device__ __forceinline__ void end_chunk(const SearchArgs& A, uint32_t c, SearchJob* J,
                                          unsigned long long& chunk_nodes) {
    if (chunk_nodes) atomicAdd(&J->nodes, chunk_nodes);
    chunk_nodes = 0;
    complete();
    release_items(A, c, 1);
}

// Donated subtrees (batches).  Once no cube has items left to hand out and
// lanes have waited a while for work, they raise split_want; every
// SPLIT_CHECK passes each wave looks at it, and a lane that has spent
// SPLIT_MIN_NODES nodes on its item (so the item is a long one) hands the
// untried moves of its shallowest level, if they are at least SPLIT_MIN_LEFT
// moves from the bound, to a SplitRec: that level is then done for it, so its
// item root moves below it.  Donating short items as well costs more in
// replays (and their probes) than it gains.  The taker replays the path (its
// moves are not counted again: node counts stay those of an undivided round)
// and searches the node from the move after the donor's one.  A record holds
// one in-flight share of the round until its taker releases it, so a round
// still closes only when all of its work has drained.
constexpr uint32_t SPLIT_CHECK = 512;  // passes between two looks of a wave (a power of two)
constexpr int SPLIT_MIN_LEFT = 7;
constexpr unsigned long long SPLIT_MIN_NODES = 2048;
constexpr int SPLIT_MAX_K = 7;         // deepest donated node: its path and skip byte fill one u64

//
// Returns the lane's new item root, or -1 if it donated nothing; a root below
// the current depth g means that the node's untried moves went too (the
// caller clears them).  Out of line, like take_split.
__device__ __attribute__((noinline)) int donate(QueueCtl* q, SplitRec* splits, const CubeLUT& L, SearchJob* J, uint32_t c,
                                                uint32_t bound, int g, int m, uint32_t todo, int root, int phase,
                                                uint32_t found_seen, unsigned long long nodes, const uint32_t* stk,
                                                uint32_t nt) {
    int k = -1;
    uint32_t p = 0;
    if (phase == 3 && !found_seen && nodes >= SPLIT_MIN_NODES) {
        int d = root;
        for (; d < g && d <= SPLIT_MAX_K && (int)bound - d >= SPLIT_MIN_LEFT; ++d) {
            const uint32_t pd = (uint32_t)stk_move(stk[d * nt]);
            const int ld = d > 0 ? stk_move(stk[(d - 1) * nt]) : -1;
            const uint32_t rem = L.amask[ld + 1] & ~axis_mask(excl_axes(stk[d * nt] & STK_H, bound - (uint32_t)d)) &
                                 ~((2u << pd) - 1u);
            if (rem) {
                k = d;
                p = pd;
                break;
            }
        }
        // every level above is done: the current node's untried moves (the one
        // under evaluation, m, stays with the donor)
        if (k < 0 && d == g && m >= 0 && todo && g <= SPLIT_MAX_K && (int)bound - g >= SPLIT_MIN_LEFT) {
            k = g;
            p = (uint32_t)m;
        }
    }
    const unsigned long long cm = __ballot(k >= 0);
    if (!cm) return -1;
    const uint32_t n = (uint32_t)__popcll(cm);
    const int leader = (int)__builtin_ctzll(cm);
    uint32_t granted = 0, blo = 0, bhi = 0;
    if ((int)lane_id() == leader) {
        const uint32_t old = atomicAdd(&q->split_pend, n);
        granted = old < SPLIT_PEND_MAX ? min(n, SPLIT_PEND_MAX - old) : 0u;
        if (granted < n) atomicSub(&q->split_pend, n - granted);
        if (granted) {
            const unsigned long long b = atomicAdd(&q->split_tail, (unsigned long long)granted);
            blo = (uint32_t)b;
            bhi = (uint32_t)(b >> 32);
        }
    }
    granted = __builtin_amdgcn_readlane(granted, leader);
    blo = __builtin_amdgcn_readlane(blo, leader);
    bhi = __builtin_amdgcn_readlane(bhi, leader);
    const uint32_t rank = (uint32_t)__popcll(cm & lanemask_lt());
    if (k < 0 || rank >= granted) return -1;
    const unsigned long long slot = (((unsigned long long)bhi << 32) | blo) + rank;
    SplitRec* R = &splits[slot & (SPLIT_RING - 1u)];
    unsigned long long path = (unsigned long long)(p + 1u) << (8 * k);
    for (int j = 0; j < k; ++j) path |= (unsigned long long)stk_move(stk[j * nt]) << (8 * j);
    atomicAdd(&J->ticket, T_INF_ONE);  // the record's share of the round, released by its taker
    st_c(&R->path, path);
    st_c(&R->cube, c);
    st_c(&R->info, bound | ((uint32_t)k << 8));
    complete();
    st_c(&R->seq, slot + 1ull);
    // k < g: the levels down to k are done; k == g: the node keeps only m, and
    // the item ends once m is done (or pruned)
    return k < g ? k + 1 : g + 1;
}

// A fetching lane's donated subtree, if one is published: each lane reads the
// record it would take before the wave claims the run of published ones with
// one CAS on the head.  While the head stays put no writer can reach those
// records (at most SPLIT_PEND_MAX are reserved past it).  Called by the lanes
// that fetch, together; out of line, so that this rarely used code does not
// add to the search loop's registers.
struct SplitTake {
    unsigned long long path;
    uint32_t cube, info;
    bool ok;
};
__device__ __attribute__((noinline)) SplitTake take_split(QueueCtl* q, const SplitRec* splits) {
    SplitTake r{0ull, 0u, 0u, false};
    const unsigned long long smask = __ballot(1);
    const uint32_t srank = (uint32_t)__popcll(smask & lanemask_lt());
    unsigned long long sh = 0, stl = 0;
    if (srank == 0) {
        sh = ld_c(&q->split_head);
        stl = ld_c(&q->split_tail);
    }
    sh = first_lane_u64(sh);
    stl = first_lane_u64(stl);
    if (sh >= stl) return r;
    const unsigned long long sl = sh + srank;
    const SplitRec* R = &splits[sl & (SPLIT_RING - 1u)];
    const bool ok = sl < stl && ld_c(&R->seq) == sl + 1ull;
    if (ok) {
        r.path = ld_c(&R->path);
        r.cube = ld_c(&R->cube);
        r.info = ld_c(&R->info);
    }
    const unsigned long long bad = __ballot(!ok);
    const uint32_t n = bad ? (uint32_t)__popcll(smask & ((1ull << __builtin_ctzll(bad)) - 1ull))
                           : (uint32_t)__popcll(smask);
    uint32_t won = 0;
    if (srank == 0 && n) won = atomicCAS(&q->split_head, sh, sh + n) == sh;
    won = first_lane_u32(won);
    if (won && srank == 0) atomicSub(&q->split_pend, n);
    r.ok = won && srank < n;
    return r;
}

#ifdef CCS_SEARCH_STATS
// Shader clock (s_memtime; its read waits for the wave's outstanding LDS and
// scalar loads), low 32 bits: differences are taken modulo 2^32.
constexpr uint32_t CYCLES_MASK = 0xFFFFFFFFu;
__device__ __forceinline__ uint32_t shader_cycles() { return (uint32_t)clock64(); }
#endif

__device__ __forceinline__ Frame load_frame_c(const Frame* p) {
    const uint32_t* w = reinterpret_cast<const uint32_t*>(p);
    return Frame{ld_c(&w[0]), ld_c(&w[1]), ld_c(&w[2]), ld_c(&w[3]), ld_c(&w[4])};
}

// QUEUE: the persistent form (cubes and rounds on the device); SERVE: its
// serving variant (admission from the mailbox, host stop flag), compiled
// separately so that the batch kernel carries none of that code.
template <int K, int F, bool QUEUE, bool SERVE>
__global__ void __launch_bounds__(SEARCH_WG, SEARCH_WAVES_PER_SIMD) search_kernel(SearchArgs A) {
    extern __shared__ __align__(16) unsigned char smem[];
    // LDS: CubeLUT, ClassLUT, then the lanes' stacks (one u32 per level), lane-major
    CubeLUT& L = *reinterpret_cast<CubeLUT*>(smem);
    ClassLUT& CL = *reinterpret_cast<ClassLUT*>(smem + sizeof(CubeLUT));
    const uint32_t nt = blockDim.x;  // lanes of the workgroup (a multiple of 64, at most SEARCH_WG)
    const uint32_t tid = threadIdx.x;
    unsigned char* stk = smem + sizeof(CubeLUT) + sizeof(ClassLUT);
    uint32_t* stw = reinterpret_cast<uint32_t*>(stk) + tid;  // this lane's stack words, STACK_DEPTH of them
    {
        const uint4* s = reinterpret_cast<const uint4*>(A.class_lut);
        uint4* d = reinterpret_cast<uint4*>(&CL);
        for (uint32_t i = tid; i < sizeof(ClassLUT) / 16; i += nt) d[i] = s[i];
    }
    load_lut(&L, A.lut);  // ends with a workgroup barrier
    const uint8_t* __restrict__ table = A.table;

    // The node's state (axis frame 0) and its inverse (frame 3).  Frame k and
    // 3 + k are R^k g R^-k of these: a probe conjugates the child's corners
    // (ConjRec) and the edge coordinate conjugates its edges by R^k with the
    // class symmetry in one table, so a push or pop moves two frames, not six.
    Frame f0, f3;
    uint32_t hp = 0, cp = 0;
    uint32_t c = 0xFFFFFFFFu;  // cube of the current item
    SearchJob* J = nullptr;
    uint32_t bound = 0, todo = 0, ord = 0, idle = 0, plen = 0;
    int g = 0, tgt = 0, m = -1;
    // Depth of the item's root: the search below it ends when the lane is back
    // at (or above) it.  The prefix length for a plain item; deeper once the
    // lane has donated the rest of the levels above.
    int root = 0;
    uint32_t wtick = 0;  // persistent form: passes of this wave
    // Child m under evaluation: forward / inverse frames still to probe (bits
    // in frame_order's layout), whether a probe pruned it, whether it is done.
    uint32_t remF = 0, remI = 0;
    bool pruned = false, cdone = false, first = false;
    // Phase of the lane's item: 0 fetch one, 1 undo down to the prefix shared
    // with the previous item, 2 replay the prefix, 3 search below it, 4 done.
    int phase = 0;
    unsigned long long nodes = 0, idle_since = 0, chunk_nodes = 0;
    uint32_t probes = 0;  // added to the launch's counter as it nears 2^31
    unsigned long long pbits = 0;  // the item's prefix, move j in byte j
    // Persistent form: the next item's prefix and the cube's found flag, both
    // read one item ahead so that no item start waits on a global load.
    unsigned long long pnext = 0;
    uint32_t found_seen = 0;
    // Persistent form: the current chunk's round data, read once when the chunk
    // is taken (per-item reads of the kernel arguments indexed by the lane's
    // prefix length, and of the cube's band bounds, each cost a memory round
    // trip at every item start): the prefix table and its item count, the key
    // rows the round reads and writes, and its band's key bounds (lo | hi << 8).
    const unsigned long long* ptab = nullptr;
    uint32_t pn = 0, band_lh = 0;
    const unsigned long long* krd = nullptr;
    uint32_t* kwr = nullptr;
    // persistent form: the lane's chunk: items (visits) left in the low byte,
    // log2 chunk size << 8, CL_BANDED, CL_KEYS
    uint32_t chunk_left = 0, item_cur = 0;
    // The lane's frames stay at depth `hold` on the previous item's prefix
    // (-1: nothing held); `hfail` is the depth of that item's failing prefix
    // move (-1: none).  The next item of the same cube and round undoes back
    // to the prefix it shares and replays only the rest.
    int hold = -1, hfail = -1;
    uint32_t atick = 0;  // serving: fetches since the mailbox was last checked
    bool done = false;
#ifdef CCS_SEARCH_STATS
    // Diagnostics (a build with -DCCS_SEARCH_STATS only; see SearchArgs::stats):
    // per lane, probes and prunes per frame and frame updates; per wave, steps,
    // a histogram of the lanes that probe in a step, and the shader-clock cycles
    // of a step's parts.
    unsigned long long d_look[6] = {}, d_prune[6] = {}, d_upd = 0;
    unsigned long long d_steps = 0, d_hist[9] = {}, d_cyc[4] = {}, d_istart_cyc = 0, d_istart_steps = 0;
    unsigned long long d_iters = 0;  // iterations of the pass loop (the wave's)
    uint32_t d_t0 = 0, d_t1 = 0;
    bool d_start = false;  // the lane started an item in this step's pass
#endif

    for (;;) {
#ifdef CCS_SEARCH_STATS
        d_t0 = shader_cycles();
        d_start = false;
#endif
        if (QUEUE) {
            // Donation: every SPLIT_CHECK passes the wave looks at whether lanes are
            // out of work (one load, a wait of about a microsecond in some
            // milliseconds of passes); if so, each lane deep in a subtree hands the
            // untried moves of its shallowest level to them (a SplitRec) and keeps
            // the rest.
            wtick = __builtin_amdgcn_readfirstlane(wtick + 1u);
            if ((wtick & (SPLIT_CHECK - 1u)) == 0u && __builtin_amdgcn_readfirstlane(ld_c(&A.q->split_want))) {
                const int nr = donate(A.q, A.splits, L, J, c, bound, g, m, todo, root, phase, found_seen, nodes, stw, nt);
                if (nr > g) todo = 0;
                if (nr >= 0) root = nr;
            }
        }
        bool eval = false;
        // The pass's frame update, applied after the pass for every lane at once
        // (push: move mv; pop: undo move mv): -1 none.  `leaf`: the push reached
        // the bound, so the frames are checked for the solved cube after it.
        int mv = -1;
        uint32_t msel = 0;
        bool leaf = false;
        while (!done) {
#ifdef CCS_SEARCH_STATS
            d_iters += __builtin_amdgcn_readfirstlane(1u);
#endif
            // One pass resolves the lane's evaluated child, then finishes, fetches,
            // undoes or backs up as needed and picks the next child, so that the
            // common sequences (push then pick, discard then pick, pop then pick)
            // cost the wave a single pass.  Once a frame update is pending, only
            // work that does not read the frames (a pick, the next prefix move)
            // may follow in the same pass.
            if (mv >= 0) break;
            if (m >= 0) {
                if (!cdone) {  // frames of child m still to probe
                    eval = true;
                    break;
                }
                if (!pruned && !all_tight(cp, bound - (uint32_t)g - 1u)) {  // survived: push
                    mv = m;
                    msel = 0;
                    stw[g * nt] = hp | ((uint32_t)m << 24);
                    hp = cp;
                    ++g;
                    if ((uint32_t)g == bound) {
                        leaf = true;
                        todo = 0;
                    } else {
                        todo = L.amask[m + 1] & ~axis_mask(excl_axes(hp, bound - (uint32_t)g));
                        ord = 0;  // recomputed before its next use
                    }
                    if (phase == 2 && g == root) {
                        // the prefix is in place; a donated subtree resumes after the
                        // donor's move, whose successor is in the byte past its path
                        phase = 3;
                        todo &= ~((1u << (uint32_t)((pbits >> (8 * g)) & 0xFFu)) - 1u);
                        if (QUEUE && !SERVE && (chunk_left & CL_KEYS)) {  // the item's key, for the next round's order
                            const uint32_t pi = (chunk_left & CL_BANDED) ? item_cur % pn : item_cur;
                            uint32_t key = 1u;
#pragma unroll
                            for (int j = 0; j < 6; ++j) key += hget(hp, j);
                            // this round's buffer: the round reads only the other one, so an
                            // item's key never changes under the round's own visits
                            atomicOr(kwr + (pi >> 2), key << (8u * (pi & 3u)));
                            if ((pi & 7u) == 0u) atomicAdd(&A.key_hist[(size_t)c * KEY_BINS + key], 1u);
                        }
                    }
                } else if (phase == 2) {  // a prefix move was pruned: the item ends
                    hfail = g;
                    phase = 4;
                }
                m = -1;
                if (mv >= 0 && !(phase == 2 || (phase == 3 && todo && !found_seen))) break;
            }
            bool pop = phase == 1;
            if (phase == 3 && todo == 0) {  // node exhausted: the item is complete, or back up
                if (g <= root || found_seen) phase = 4;
                else pop = true;
            }
            if (phase == 4) {
                // ---- the item is finished: account its nodes ----
                hold = g <= root ? g : -1;  // deeper only when found ended the search
                if (probes >= 0x80000000u && A.stats) {
                    atomicAdd(&A.stats[0], (unsigned long long)probes);
                    probes = 0;
                }
                if (QUEUE) {
                    chunk_nodes += nodes;
                    if (!(chunk_left & 0xFFu) || found_seen) {
                        chunk_left = 0;
                        end_chunk(A, c, J, chunk_nodes);
                    }
                } else {
                    atomicAdd(&J->nodes, nodes);
                }
                phase = 0;
            }
            if (phase == 0) {
                // ---- fetch a work item: (cube, bound, prefix length, prefix) ----
                uint32_t item;
                bool fresh;
                bool spl = false;  // a donated subtree (its path in spath)
                unsigned long long spath = 0;
                if (!QUEUE) {
                    unsigned long long mask = __ballot(1);
                    uint32_t rank = __popcll(mask & lanemask_lt());
                    unsigned long long w0 = 0;
                    if (rank == 0) w0 = atomicAdd(A.counter, (unsigned long long)__popcll(mask));
                    unsigned long long w = first_lane_u64(w0) + rank;
                    if (w >= (unsigned long long)A.n_active * A.n_prefix) {
                        done = true;
                        break;
                    }
                    uint32_t cw = (uint32_t)A.active[w / A.n_prefix];
                    item = (uint32_t)(w % A.n_prefix);
                    if (cw != c) hold = -1;  // one launch is one round: same cube, same bound
                    fresh = false;
                    c = cw;
                    J = &A.jobs[c];
                    if (ld_c(&J->found)) continue;
                    found_seen = 0;
                    bound = J->bound;
                    plen = (uint32_t)A.plen;
                } else if (chunk_left & 0xFFu) {
                    // next item of this lane's chunk: same cube and round, no atomics
                    const uint32_t bbase = (!SERVE && (chunk_left & CL_BANDED)) ? (item_cur / pn) * pn : 0u;
                    const uint32_t w = next_item(item_cur - bbase, chunk_left & 0xFFu, (chunk_left >> 8) & 7u, pn);
                    item = bbase + w;
                    --chunk_left;
                    fresh = false;
                    if (w >= pn) {  // the rest of the chunk lies past the round's (band's) last item
                        chunk_left = 0;
                        end_chunk(A, c, J, chunk_nodes);
                        continue;
                    }
                    item_cur = item;
                } else {
                    if (idle && (--idle)) break;  // back off between empty fetches
                    {
                        unsigned long long mask = __ballot(1);
                        uint32_t cnt = __popcll(mask), rank = __popcll(mask & lanemask_lt());
                        uint32_t cc = 0xFFFFFFFFu, clog = 0;
                        unsigned long long old = 0;
                        // serving: a stop request ends the launch within a few fetches even
                        // while cubes remain.  The fetch count is kept wave-uniform (each
                        // fetch takes its leader's count + 1) and the flag, in host memory,
                        // is read on every 4th fetch.
                        if (SERVE) atick = first_lane_u32(atick) + 1u;
                        const bool stopping = SERVE && rank == 0 && (atick & 3u) == 0u && sys_load(A.host_stop);
                        if (rank == 0 && !stopping) {
                            for (int tries = 0; tries < 32; ++tries) {
                                unsigned long long h = ld_c(&A.q->head), t = ld_c(&A.q->tail);
                                // serving: take new cubes when the ring is empty, and while
                                // fewer than ADMIT_BELOW rounds are queued on every 16th fetch
                                // (the mailbox is in host memory; its reads cross the bus)
                                if (SERVE && (h >= t || (t - h < ADMIT_BELOW && (atick & 15u) == 1u && tries == 0)) &&
                                    admit_one<K, F>(A, L))
                                    continue;
                                if (h >= t) break;
                                unsigned long long e = ld_c(&A.ring[h % A.ring_size]);
                                if ((e >> 37) != (h & RING_SEQ_MASK)) break;  // entry still being published
                                uint32_t ce = (uint32_t)e, eb = (uint32_t)(e >> 32) & 31u;
                                SearchJob* Je = &A.jobs[ce];
                                unsigned long long tk = ld_c(&Je->ticket);
                                if ((tk & (T_CLOSING | T_RETIRED)) || t_bound(tk) != eb ||
                                    t_next(tk) >= t_visits(A, tk) || ld_c(&Je->found)) {
                                    atomicCAS(&A.q->head, h, h + 1);  // stale, finished or fully handed out
                                    continue;
                                }
                                clog = t_chunk(tk);
                                old = atomicAdd(&Je->ticket, (unsigned long long)cnt * (T_INF_ONE + (1ull << clog)));
                                // Round robin: the cube moves behind the others, so each live cube
                                // gets about its share of the lanes and its round is searched
                                // roughly in item order (a final round then stops near where a
                                // sequential search would).  Only the wave that advances the head
                                // re-queues the entry, and only while items remain.  A server's
                                // cube keeps the head for its turn's credit of fetches.
                                bool move_on = true;
                                if (SERVE) move_on = (int)atomicSub(&Je->credit, 1u) <= 1;
                                if (move_on && atomicCAS(&A.q->head, h, h + 1) == h &&
                                    t_next(old) + ((unsigned long long)cnt << clog) < t_visits(A, old)) {
                                    if (SERVE) st_c(&Je->credit, serve_credit(A, Je));
                                    ring_push(A, eb, ce);
                                }
                                cc = ce;
                                break;
                            }
                        }
                        cc = first_lane_u32(cc);
                        old = first_lane_u64(old);
                        clog = first_lane_u32(clog);
                        // no round has items left: a donated subtree, if one is published
                        const SplitTake st = cc == 0xFFFFFFFFu ? take_split(A.q, A.splits) : SplitTake{0ull, 0u, 0u, false};
                        if (cc == 0xFFFFFFFFu && !st.ok) {
                            if (ld_c(&A.q->n_final) >= A.q->n_jobs || ld_c(&A.q->stop) ||
                                (SERVE && sys_load(A.host_stop))) {
                                done = true;
                            } else {
                                // out of work: once the wait has lasted, ask for donations and
                                // back off longer, so that idle waves poll the queue words far
                                // less.  A server asks only while few cubes are in it: under
                                // load, donated work would wait behind new rounds in the ring
                                // and hold its own round open (13% of the served rate at 80%
                                // load), while a lone hard cube needs the whole GPU.
                                unsigned long long now = wall_clock64();
                                const bool waited = idle_since && now - idle_since > A.backoff_ticks;
                                const bool few = !SERVE || (uint32_t)ld_c(&A.q->mb_head) - ld_c(&A.q->n_final) <=
                                                               SERVE_DONATE_LIVE;
                                if (few && waited && rank == 0 && !ld_c(&A.q->split_want))
                                    st_c(&A.q->split_want, 1u);
                                idle = waited ? 1024 : 64;
                                // watchdog: a lane that finds no work for idle_limit ends the launch
                                if (!idle_since) idle_since = now;
                                else if (now - idle_since > A.idle_limit) atomicMax(&A.q->stop, 2u);
                            }
                            break;
                        }
                        idle_since = 0;
                        if (st.ok) {
                            spl = true;
                            spath = st.path;
                            c = st.cube;
                            J = &A.jobs[c];
                            bound = st.info & 31u;
                            plen = (st.info >> 8) & 31u;  // the path: replayed like a prefix
                            found_seen = 0;
                            item = 0;
                            item_cur = 0;
                            fresh = true;
                            hold = -1;
                            chunk_left = 0;
                            chunk_nodes = 0;
                        } else {
                            const uint32_t pl = t_plen(old);
                            item = t_next(old) + (rank << clog);
                            bool valid = !(old & (T_CLOSING | T_RETIRED)) && item < t_visits(A, old);
                            unsigned long long inval = __ballot(!valid);
                            if (rank == 0 && inval) release_items(A, cc, __popcll(inval));
                            if (!valid) continue;
                            c = cc;
                            J = &A.jobs[c];
                            bound = t_bound(old);
                            plen = pl;
                            ptab = A.pre_tab[pl];
                            pn = A.pre_n[pl];
                            if (!SERVE && A.keys[pl]) {
                                uint32_t* krow0 = A.keys[pl] + (size_t)c * 2u * A.key_words[pl];
                                kwr = krow0 + (bound & 1u) * A.key_words[pl];
                                krd = reinterpret_cast<const unsigned long long*>(krow0 + ((bound - 1u) & 1u) * A.key_words[pl]);
                            }
                            if (!SERVE && (old & T_BANDED)) {  // chunk ci of the round is chunk ci mod Q of band ci / Q
                                const uint32_t nq = (pn + (1u << clog) - 1u) >> clog;
                                const uint32_t ci = item >> clog, band = ci / nq;
                                item = band * pn + (ci - band * nq) * min(ITEM_BLOCK, 1u << clog);
                                // the band's key bounds, fixed for the round (close_round set them)
                                const uint32_t lo =
                                    band ? (ld_c(&J->band_thr[(band - 1u) >> 2]) >> (8u * ((band - 1u) & 3u))) & 0xFFu : 0u;
                                const uint32_t hi =
                                    band < N_BANDS - 1u ? (ld_c(&J->band_thr[band >> 2]) >> (8u * (band & 3u))) & 0xFFu : 255u;
                                band_lh = lo | hi << 8;
                            } else {
                                item = (item >> clog) * min(ITEM_BLOCK, 1u << clog);  // chunk's first item
                            }
                            item_cur = item;
                            fresh = true;  // a chunk may follow a reuse of the slot: start from the root
                            hold = -1;
                            chunk_left = ((1u << clog) - 1u) | (clog << 8) | ((old & T_BANDED) ? CL_BANDED : 0u) |
                                         (!SERVE && A.keys[pl] ? CL_KEYS : 0u);
                            chunk_nodes = 0;
                            if (ld_c(&J->found)) {
                                chunk_left = 0;
                                end_chunk(A, c, J, chunk_nodes);
                                continue;
                            }
                            found_seen = 0;  // refreshed by the picks' polls
                        }
                    }
                }
#ifdef CCS_SEARCH_STATS
                d_start = true;
#endif
                // ---- start the item: its prefix is replayed move by move through
                // the probe step below, every frame (kernels.cu:1552-1614) ----
                unsigned long long nb;
                if (QUEUE) {
                    const bool banded = !SERVE && !spl && (chunk_left & CL_BANDED);
                    uint32_t pitem = item;
                    if (banded) {
                        // visit `item` is item pitem of band `band`: searched only there,
                        // by the keys of the previous round (its buffer: bound - 1)
                        const uint32_t band = item / pn;
                        pitem = item - band * pn;
                        const unsigned long long kw = ld_c(&krd[pitem >> 3]);  // keys of 8 items
                        const uint32_t lo = band_lh & 0xFFu, hi = band_lh >> 8;
                        bool member = in_band((uint32_t)(kw >> (8u * (pitem & 7u))) & 0xFFu, band, lo, hi);
                        // the chunk's next visits whose keys came with the same load (the
                        // rest of its block of ITEM_BLOCK consecutive items)
                        const uint32_t blk = min(ITEM_BLOCK, 1u << ((chunk_left >> 8) & 7u));
                        while (!member && (chunk_left & 0xFFu & (blk - 1u)) && ((pitem + 1u) & 7u) && pitem + 1u < pn) {
                            ++item;
                            ++pitem;
                            --chunk_left;
                            item_cur = item;
                            member = in_band((uint32_t)(kw >> (8u * (pitem & 7u))) & 0xFFu, band, lo, hi);
                        }
                        if (!member) {
                            if (!(chunk_left & 0xFFu)) {
                                chunk_left = 0;
                                end_chunk(A, c, J, chunk_nodes);
                            }
                            // nothing searched and nothing held changes; the pass ends
                            // here, so that a run of other bands' visits does not
                            // hold the whole wave
                            break;
                        }
                    }
                    nb = spl ? spath : (fresh || banded) ? ptab[pitem] : pnext;
                    if ((chunk_left & 0xFFu) && !banded) {  // the next item's prefix, used when it starts
                        const uint32_t nx = next_item(item, chunk_left & 0xFFu, (chunk_left >> 8) & 7u, pn);
                        if (nx < pn) pnext = ptab[nx];
                    }
                } else {
                    nb = A.prefixes[item];
                }
                fresh = fresh || hold < 0;
                unsigned long long x = nb ^ pbits;
                int shared = fresh ? 0 : min((int)plen, x ? (int)(__builtin_ctzll(x) >> 3) : 8);
                pbits = nb;
                if (!fresh && hfail >= 0 && shared > hfail) {
                    // the previous item's failing move is on this prefix too: same
                    // nodes as a full replay (moves up to the failing one), no probes
                    nodes = (unsigned long long)hfail + 1u;
                    phase = 4;
                    continue;
                }
                hfail = -1;
                if (fresh) {
                    hp = 0;
                    if (SERVE) {  // the slot may have held another cube: read past the L2
                        f0 = load_frame_c(&J->frame[0]);
                        f3 = load_frame_c(&J->frame[3]);
#pragma unroll
                        for (int j = 0; j < 6; ++j) hp |= min(ld_c(&J->h[j]), 15u) << (4 * j);
                    } else {
                        f0 = load_frame(&J->frame[0]);
                        f3 = load_frame(&J->frame[3]);
#pragma unroll
                        for (int j = 0; j < 6; ++j) hp |= min(J->h[j], 15u) << (4 * j);
                    }
                    g = 0;
                    ord = 0;
                }
                // moves shared with the held state count as replayed (node counts are
                // those of a replay from the root); a donated subtree's path was
                // counted by its donor
                nodes = spl ? 0ull - plen : (unsigned long long)shared;
                tgt = shared;  // undo down to the shared depth first
                todo = 0;
                root = (int)plen;
                phase = g > tgt ? 1 : (g < root ? 2 : 3);
                if (phase == 3) todo = L.amask[0] & ~axis_mask(excl_axes(hp, bound));  // empty prefix
                pop = phase == 1;
            }
            if (pop) {
                --g;
                const uint32_t w = stw[g * nt];
                const int pm = stk_move(w);
                mv = pm;
                msel = 1;
                hp = w & STK_H;
                const int last = g > 0 ? stk_move(stw[(g - 1) * nt]) : -1;
                todo = L.amask[last + 1] & ~axis_mask(excl_axes(hp, bound - (uint32_t)g)) & ~((2u << pm) - 1u);
                ord = 0;
                if (phase == 1 && g == tgt) phase = 2;
            }
            // The frame order of the lane's node, after a push, a pop or a new
            // item's root (which mark it stale with 0; an actual order has one
            // bit per frame): computed here once for all three, so that a wave
            // whose lanes did different ones runs it once.
            if (ord == 0u) ord = frame_order(hp, bound - (uint32_t)g);
            if (phase == 2) {
                // ---- next prefix move, evaluated like a child ----
                m = (int)((pbits >> (8 * g)) & 0xFFu);
                ++nodes;
                const uint32_t axis = move_axis(m);
                if ((excl_axes(hp, bound - (uint32_t)g) >> axis) & 1u) {  // the move is not applied
                    hfail = g;
                    m = -1;
                    phase = 4;
                    continue;
                }
                cp = hp;  // the inherited inverse frame keeps the parent's value
                remF = ord & 0x777u;
                remI = (ord >> 16) & ~(0x111u << axis);
                pruned = cdone = false;
                first = true;
                eval = true;
                break;
            }
            if (phase == 3 && todo) {
                if (found_seen) {  // another lane solved the cube
                    phase = 4;
                    continue;
                }
                // refresh the flag every poll_mask + 1 picks of the lane's chunk (its items
                // start with no read of it); it is read from the next pick on, so the
                // load's latency hides behind this step's probes
                if ((((QUEUE ? chunk_nodes : 0ull) + nodes) & A.poll_mask) == 0) found_seen = ld_c(&J->found);
                // ---- pick the next canonical child ----
                m = (int)ctz(todo);
                todo &= todo - 1;
                ++nodes;
                cp = hp;  // the inherited inverse frame keeps the parent's value
                remF = ord & 0x777u;
                remI = (ord >> 16) & ~(0x111u << move_axis(m));
                pruned = cdone = false;
                first = true;
                eval = true;
                break;
            }
        }
        if (__all(done)) break;
        if (QUEUE && !__any(eval || mv >= 0)) {  // the whole wave is waiting for work
            __builtin_amdgcn_s_sleep(2);
            continue;
        }
#ifdef CCS_SEARCH_STATS
        d_t1 = shader_cycles();
        d_cyc[0] += (d_t1 - d_t0) & CYCLES_MASK;  // the pass: bookkeeping, fetches, picks
        if (__any(d_start)) {  // passes in which a lane of the wave started an item
            d_istart_cyc += (d_t1 - d_t0) & CYCLES_MASK;
            ++d_istart_steps;
        }
        d_t0 = d_t1;
        ++d_steps;
        ++d_hist[((uint32_t)__popcll(__ballot(eval)) + 7u) >> 3];
        d_upd += mv >= 0;
#endif
        if (mv >= 0) {
            right_mul(f0, L.rr[msel][mv]);
            left_mul(f3, L.lr[msel][mv]);
            if (leaf && is_solved(f0)) report_solution(A, J, QUEUE, g, stw, nt);
        }
#ifdef CCS_SEARCH_STATS
        d_t1 = shader_cycles();
        d_cyc[1] += (d_t1 - d_t0) & CYCLES_MASK;  // the frame updates
        d_t0 = d_t1;
#endif
        if (eval) {
            // ---- one step of child m: its next forward and next inverse frame,
            // both table reads in flight together ----
            // raised wave priority from here until the table addresses are known: a
            // wave computing a probe goes ahead of waves doing bookkeeping, so more
            // table misses overlap
            __builtin_amdgcn_s_setprio(2);
            bool hasF = remF != 0u, hasI = remI != 0u;
            // a child's first step probes only its top frame, where most are pruned
            // (with one probe per step, every step probes only the top remaining frame)
            if (first || A.one_probe) {
                if ((ctz(remF | 0x10000u) >> 2) <= (ctz(remI | 0x10000u) >> 2)) hasI = false;
                else hasF = false;
                first = false;
            }
            const uint32_t kf = ctz(remF | 0x10000u) & 3u, ki = ctz(remI | 0x10000u) & 3u;
            if (hasF) remF &= remF - 1u;
            if (hasI) remI &= remI - 1u;
            Frame tf = f0, ti = f3;
            right_mul(tf, L.rr[0][m]);
            left_mul(ti, L.lr[0][m]);
            conj_corners(tf, L.cj[kf]);
            conj_corners(ti, L.cj[ki]);
            const uint32_t pkF = corner_class(CL, tf), pkI = corner_class(CL, ti);
            uint32_t jF, jI;
            const uint8_t* bF = probe_block<K, F>(L, pkF, kf, table, tf, jF);
            const uint8_t* bI = probe_block<K, F>(L, pkI, ki, table, ti, jI);
            __builtin_amdgcn_s_setprio(0);
#ifdef CCS_SEARCH_STATS
            d_t1 = shader_cycles();
            d_cyc[2] += (d_t1 - d_t0) & CYCLES_MASK;  // the probes' addresses
            d_t0 = d_t1;
#endif
            // The step's table reads.  Paired (A.read_mode 0): both issued with no
            // branch around them, so that they are in flight together; a lane
            // without a probe of one kind reads its other probe's block again.
            // In turn (1): each read in a branch of its own that waits for it;
            // lanes read only what they probe.  Paired reads are faster in large
            // batches, where most lanes probe twice; in turn, when most steps are
            // a child's first, one-probe step (lone and hard cubes).
            uint32_t vf = 0u, vi = 0u;
            if (A.read_mode == 0) {
                const uint32_t tF = block_value<F>(L, hasF ? bF : bI, hasF ? jF : jI);
                const uint32_t tI = block_value<F>(L, hasI ? bI : bF, hasI ? jI : jF);
                vf = hasF ? tF : 0u;
                vi = hasI ? tI : 0u;
            } else {
                if (hasF) vf = block_value<F>(L, bF, jF);
                if (hasI) vi = block_value<F>(L, bI, jI);
            }
            probes += (uint32_t)hasF + (uint32_t)hasI;
            const uint32_t pf = hget(hp, (int)kf);
            if (pf) vf = max(vf, pf - 1u);  // forward frames: at least the parent's value - 1
            bool pr = false;
            if (hasF) {
                cp = hset(cp, (int)kf, vf);
                pr = (uint32_t)g + 1u + vf > bound;
            }
            if (hasI) {
                cp = hset(cp, 3 + (int)ki, vi);
                pr = pr || (uint32_t)g + 1u + vi > bound;
            }
            pruned = pr;
            cdone = pr || (remF | remI) == 0u;
#ifdef CCS_SEARCH_STATS
            static_for<0, 3>([&](auto j) {
                d_look[j] += hasF && kf == (uint32_t)j;
                d_look[3 + j] += hasI && ki == (uint32_t)j;
                d_prune[j] += hasF && kf == (uint32_t)j && (uint32_t)g + 1u + vf > bound;
                d_prune[3 + j] += hasI && ki == (uint32_t)j && (uint32_t)g + 1u + vi > bound;
            });
            d_t1 = shader_cycles();
            d_cyc[3] += (d_t1 - d_t0) & CYCLES_MASK;  // waiting for the table reads, and the values
#endif
        }
    }
    if (A.stats) {
        unsigned long long p = probes;
        for (int o = 32; o > 0; o >>= 1) p += __shfl_xor(p, o);
        if (lane_id() == 0) atomicAdd(&A.stats[0], p);
#ifdef CCS_SEARCH_STATS
        // lanes' counters summed over the wave; the wave's (all lanes equal) once
        auto wave_sum = [&](unsigned long long v) {
            for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
            return v;
        };
        for (int i = 0; i < 6; ++i) {
            const unsigned long long a = wave_sum(d_look[i]), b = wave_sum(d_prune[i]);
            if (lane_id() == 0) atomicAdd(&A.stats[STAT_DIAG + i], a), atomicAdd(&A.stats[STAT_DIAG + 6 + i], b);
        }
        const unsigned long long u = wave_sum(d_upd);
        if (lane_id() == 0) {
            atomicAdd(&A.stats[STAT_DIAG + 12], d_steps);
            atomicAdd(&A.stats[STAT_DIAG + 13], u);
            for (int i = 0; i < 9; ++i) atomicAdd(&A.stats[STAT_DIAG + 14 + i], d_hist[i]);
            for (int i = 0; i < 4; ++i) atomicAdd(&A.stats[STAT_DIAG + 23 + i], d_cyc[i]);
            atomicAdd(&A.stats[STAT_DIAG + 27], d_istart_cyc);
            atomicAdd(&A.stats[STAT_DIAG + 28], d_istart_steps);
            atomicAdd(&A.stats[STAT_DIAG + 29], d_iters);
        }
#endif
    }
}

// ---------------------------------------------------------------------------
// Gather micro-benchmark: random 32-byte table blocks read the way a probe of
// the block layout reads them (block minimum word + one offset word; mode 1
// the same with nontemporal loads), or one 16-byte half of a random block in
// a single load, as the search reads its half-block layout (mode 3), to
// measure the random-probe ceiling of the memory system; mode 2 streams the
// table for a calibration of the fetch counters.  Four independent probes per
// thread per iteration keep several misses in flight.
__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

template <int MODE>
__global__ void __launch_bounds__(256) gather_kernel(const uint32_t* __restrict__ blk, unsigned long long n_blocks,
                                                     unsigned long long n_probes, unsigned long long* sum) {
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    unsigned long long tid = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x;
    uint32_t acc = 0;
    if (MODE == 2) {  // stream: every word once
        const uint4* p = reinterpret_cast<const uint4*>(blk);
        unsigned long long n4 = n_blocks * 2;
        for (unsigned long long i = tid; i < n4; i += stride) {
            uint4 v = p[i];
            acc += v.x ^ v.y ^ v.z ^ v.w;
        }
    } else {
        for (unsigned long long i = tid * 4; i < n_probes; i += stride * 4) {
            uint32_t v[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) {
                unsigned long long h = mix64(i + k + 0x9E3779B97F4A7C15ull);
                const uint32_t* b = blk + (h % n_blocks) * 8u;
                uint32_t w = (uint32_t)(h >> 61);
                if (MODE == 3) {
                    const uint4 q = reinterpret_cast<const uint4*>(b)[w & 1u];
                    v[k] = q.x ^ q.y ^ q.z ^ q.w;
                } else if (MODE == 1) {
                    v[k] = __builtin_nontemporal_load(b) + __builtin_nontemporal_load(b + w);
                } else {
                    v[k] = b[0] + b[w];
                }
            }
            acc += v[0] ^ v[1] ^ v[2] ^ v[3];
        }
    }
    atomicAdd(sum, (unsigned long long)acc);
}

// The block format laid out for the search (FMT_HALVES, device_cube.hpp): half
// 0 of a block is its bits 4..131, half 1 its bits 128..255, the low 4 bits of
// each replaced by the block minimum.  Blocks whose minimum does not fit in 4
// bits are counted in *bad (the layout cannot hold them).
__global__ void __launch_bounds__(256) halves_kernel(const uint4* __restrict__ blk, uint4* __restrict__ out,
                                                      unsigned long long n_blocks, unsigned int* bad) {
    const unsigned long long stride = (unsigned long long)gridDim.x * blockDim.x;
    for (unsigned long long b = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; b < n_blocks; b += stride) {
        const uint4 lo = blk[2 * b], hi = blk[2 * b + 1];
        const uint32_t m = lo.x & 0xFFu;
        if (m > 15u) atomicAdd(bad, 1u);
        out[2 * b] = make_uint4((((lo.x >> 4) | (lo.y << 28)) & ~15u) | (m & 15u), (lo.y >> 4) | (lo.z << 28),
                                (lo.z >> 4) | (lo.w << 28), (lo.w >> 4) | (hi.x << 28));
        out[2 * b + 1] = make_uint4((hi.x & ~15u) | (m & 15u), hi.y, hi.z, hi.w);
    }
}

template <int K, int F>
void* search_fn_f(int form) {
    return form == 2 ? (void*)search_kernel<K, F, true, true>
         : form == 1 ? (void*)search_kernel<K, F, true, false> : (void*)search_kernel<K, F, false, false>;
}
template <int K>
void* search_fn(int F, int form) {
    return F == FMT_HALVES ? search_fn_f<K, FMT_HALVES>(form) : F == 0 ? search_fn_f<K, 0>(form) : search_fn_f<K, 1>(form);
}

#define CCS_KIND_SWITCH(K, CALL)                  \
    switch (K) {                                  \
        case 0: CALL(0); break;                   \
        case 8: CALL(8); break;                   \
        case 12: CALL(12); break;                 \
        case 16: CALL(16); break;                 \
        case 24: CALL(24); break;                 \
        default: return hipErrorInvalidValue;     \
    }

}  // namespace

size_t search_smem_bytes(int block) {
    return sizeof(CubeLUT) + sizeof(ClassLUT) + (size_t)block * STACK_DEPTH * 4;
}

hipError_t search_occupancy(int K, int F, int form, int block, int* blocks_per_cu) {
    void* fn = nullptr;
#define CALL(KK) fn = search_fn<KK>(F, form)
    CCS_KIND_SWITCH(K, CALL)
#undef CALL
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, fn, block, search_smem_bytes(block));
}

hipError_t launch_search(int K, int F, int form, const SearchArgs& args, int grid, int block, hipStream_t s) {
    size_t smem = search_smem_bytes(block);
    void* fn = nullptr;
#define CALL(KK) fn = search_fn<KK>(F, form)
    CCS_KIND_SWITCH(K, CALL)
#undef CALL
    void* params[] = {const_cast<SearchArgs*>(&args)};
    return hipLaunchKernel(fn, dim3(grid), dim3(block), params, smem, s);
}

hipError_t launch_probe(int K, int F, const CubeLUT* lut, const uint32_t* classidx, const uint8_t* table,
                        const Frame* in, int n, int move, unsigned long long* idx_out, uint32_t* lb_out,
                        Frame* state_out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    dim3 grid((n + 255) / 256), block(256);
#define CALL(KK)                                                                                                 \
    if (F == 0) hipLaunchKernelGGL((probe_kernel<KK, 0>), grid, block, 0, s, lut, classidx, table, in, n, move, \
                                   idx_out, lb_out, state_out);                                                  \
    else hipLaunchKernelGGL((probe_kernel<KK, 1>), grid, block, 0, s, lut, classidx, table, in, n, move, idx_out, \
                            lb_out, state_out);
    CCS_KIND_SWITCH(K, CALL)
#undef CALL
    return hipGetLastError();
}

hipError_t launch_unrank_check(int K, const CubeLUT* lut, const uint32_t* classidx, const uint32_t* class_rep,
                               const unsigned long long* idx, int n, unsigned long long* out, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    dim3 grid((n + 255) / 256), block(256);
#define CALL(KK) hipLaunchKernelGGL((unrank_kernel<KK>), grid, block, 0, s, lut, classidx, class_rep, idx, n, out);
    CCS_KIND_SWITCH(K, CALL)
#undef CALL
    return hipGetLastError();
}

hipError_t launch_bfs_seed(int K, const CubeLUT* lut, uint32_t* tbl, const uint32_t* classidx,
                           const uint16_t* class_stab, unsigned long long* counter, hipStream_t s) {
#define CALL(KK) hipLaunchKernelGGL((bfs_seed_kernel<KK>), dim3(1), dim3(64), 0, s, lut, tbl, classidx, class_stab, counter);
    CCS_KIND_SWITCH(K, CALL)
#undef CALL
    return hipGetLastError();
}

hipError_t launch_bfs_layer(int K, bool forward, int grid, const CubeLUT* lut, uint32_t* tbl, const uint32_t* classidx,
                            const uint32_t* class_rep, const uint16_t* class_stab, unsigned long long n_words,
                            unsigned long long n_entries, uint32_t depth, unsigned long long* counter, hipStream_t s) {
    const int fgrid = max(1, grid * (int)BFS_WG / (int)FWD_WG);  // the same number of waves
#define CALL(KK)                                                                                                    \
    if (forward) hipLaunchKernelGGL((bfs_forward_kernel<KK>), dim3(fgrid), dim3(FWD_WG), 0, s, lut, tbl, classidx,   \
                                    class_rep, class_stab, n_words, n_entries, depth, counter);                     \
    else hipLaunchKernelGGL((bfs_backward_kernel<KK>), dim3(grid), dim3(BFS_WG), 0, s, lut, tbl, classidx,          \
                            class_rep, n_words, n_entries, depth, counter);
    CCS_KIND_SWITCH(K, CALL)
#undef CALL
    return hipGetLastError();
}

hipError_t launch_histogram(int grid, const uint32_t* tbl, unsigned long long n_words, unsigned long long n_entries,
                            unsigned long long* counts, hipStream_t s) {
    hipLaunchKernelGGL(histogram_kernel, dim3(grid), dim3(256), 0, s, tbl, n_words, n_entries, counts);
    return hipGetLastError();
}

hipError_t launch_halves(int grid, const uint8_t* blk, uint8_t* out, unsigned long long n_blocks, unsigned int* bad,
                         hipStream_t s) {
    hipLaunchKernelGGL(halves_kernel, dim3(grid), dim3(256), 0, s, reinterpret_cast<const uint4*>(blk),
                       reinterpret_cast<uint4*>(out), n_blocks, bad);
    return hipGetLastError();
}

hipError_t launch_gather(int mode, int grid, const uint32_t* blk, unsigned long long n_blocks,
                         unsigned long long n_probes, unsigned long long* sum, hipStream_t s) {
    if (mode == 0) hipLaunchKernelGGL(gather_kernel<0>, dim3(grid), dim3(256), 0, s, blk, n_blocks, n_probes, sum);
    else if (mode == 1) hipLaunchKernelGGL(gather_kernel<1>, dim3(grid), dim3(256), 0, s, blk, n_blocks, n_probes, sum);
    else if (mode == 3) hipLaunchKernelGGL(gather_kernel<3>, dim3(grid), dim3(256), 0, s, blk, n_blocks, n_probes, sum);
    else hipLaunchKernelGGL(gather_kernel<2>, dim3(grid), dim3(256), 0, s, blk, n_blocks, n_probes, sum);
    return hipGetLastError();
}

hipError_t launch_encode(int format, const uint32_t* tbl, uint8_t* mins, unsigned long long n_entries,
                         unsigned long long b0, unsigned long long nb, uint32_t* out, unsigned int* bad, hipStream_t s) {
    if (nb == 0) return hipSuccess;
    if (b0 % 64u) return hipErrorInvalidValue;  // a wave's span must start on a 16-byte boundary
    const unsigned long long n_words = (n_entries + 7u) / 8u;
    dim3 grid((unsigned)((nb + 255) / 256)), block(256);
    if (format == 0)
        hipLaunchKernelGGL(encode_block_kernel, grid, block, 0, s, tbl, n_words, n_entries, b0, nb, out, mins, bad);
    else if (out) hipLaunchKernelGGL(encode_b3_kernel, grid, block, 0, s, tbl, n_words, mins, n_entries, b0, nb, out);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

}  // namespace ccs
