// Encoders of the solving formats from the BFS build table (exact depth per
// entry, one nibble each, 8 per u32), one output block of 32 bytes at a time.
// The kernels stage a wave's span of the build table in LDS and hand each lane
// its block's nibbles realigned so that entry j sits in word j / 8 at nibble
// j % 8; the functions below then work on registers only.  They are
// __host__ __device__ so that the host can check them against a direct
// per-entry encoding.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#define CCS_EHD __host__ __device__ __forceinline__

namespace ccs {

constexpr uint32_t BLK_ENTRIES = 124;  // block format: entries per 32-byte block
constexpr uint32_t B3_ENTRIES = 155;   // base-3 format: entries per 32-byte block

// Bits 32 * k + sh .. of the 64-bit value hi:lo (v_alignbit_b32), sh < 32.
CCS_EHD uint32_t funnel(uint32_t hi, uint32_t lo, uint32_t sh) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_alignbit(hi, lo, sh);
#else
    return (uint32_t)((((uint64_t)hi << 32) | lo) >> sh);
#endif
}

// Entry j (compile-time after unrolling) of realigned words a.
template <typename W>
CCS_EHD uint32_t nib(const W& a, uint32_t j) {
    return (a[j >> 3] >> ((j & 7u) * 4u)) & 15u;
}

// Block format (kernels.cu:396-406, 485-503 define it): u8 minimum, then the
// 2-bit offsets min(v - min, 3) at bit 8 + 2j.  a holds 124 entries (16 words).
// Returns true if an entry is still unvisited (15), which the caller reports.
CCS_EHD bool encode_block_words(const uint32_t (&a)[16], uint32_t (&w)[8]) {
    uint32_t m = 15, unset = 0;
#pragma unroll
    for (uint32_t j = 0; j < BLK_ENTRIES; ++j) {
        const uint32_t v = nib(a, j);
        m = v < m ? v : m;
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) w[k] = 0;
    w[0] = m;
#pragma unroll
    for (uint32_t j = 0; j < BLK_ENTRIES; ++j) {
        const uint32_t v = nib(a, j);
        unset |= v == 15u;
        const uint32_t d = v - m, pos = 8u + 2u * j;
        w[pos >> 5] |= (d < 3u ? d : 3u) << (pos & 31u);
    }
    return unset != 0;
}

// The value the base-3 encoder sees for an entry of exact depth d whose
// 124-entry block has minimum mn (kernels.cu:1825-1835): the block-table bound,
// and when its 2-bit offset saturates, the first value at or above mn + 3 with
// the entry's depth mod 3.
CCS_EHD uint32_t b3_value(uint32_t d, uint32_t mn) {
    const uint32_t x = d - mn;
    return x < 3u ? d : mn + 3u + (x - 3u) % 3u;
}

// Base-3 format (kernels.cu:1839-1869 define it): header m | c << 6 with the cut
// c in 0..3 maximising the sum of encoded bounds (the lowest c on ties), then
// 31 bytes of five digits, little-endian base 3.  The block has n <= 155
// entries (fewer only at the table's end; missing entries get digit 0); nib(j)
// gives the exact depth of entry j < n, which lies in the 124-entry block with
// minimum mn0 below j1, mn1 below j2, else mn2.  The loops stay rolled: 155
// entries in registers would spill.
// put(k, word) receives output word k = 0..7 once it is complete.
template <typename Nib, typename Put>
CCS_EHD void encode_b3_block(const Nib& nib, uint32_t n, uint32_t mn0, uint32_t mn1, uint32_t mn2, uint32_t j1,
                             uint32_t j2, const Put& put) {
    auto val = [=](uint32_t j) __attribute__((always_inline)) {  // by value: a choice among references spills
        // the block minimum by masks: a conditional chain becomes a stack array
        const uint32_t s1 = 0u - (uint32_t)(j >= j1), s2 = 0u - (uint32_t)(j >= j2);
        const uint32_t mn = (mn0 & ~s1) | (mn1 & s1 & ~s2) | (mn2 & s2);
        return b3_value(nib(j), mn);
    };
    uint32_t m = 255;
    for (uint32_t j = 0; j < n; ++j) {
        const uint32_t v = val(j);
        m = v < m ? v : m;
    }
    // counts of v - m = 0, 1, 2, 3, 4 and >= 5, one byte each (at most 155)
    uint64_t cnt = 0;
    for (uint32_t j = 0; j < n; ++j) {
        const uint32_t x = val(j) - m;
        cnt += 1ull << (8u * (x < 5u ? x : 5u));
    }
    // encoded bound of an entry with v - m = x under cut c: m if x <= c,
    // m + c + 1 if x == c + 1, else m + c + 2
    uint32_t best = 0, best_score = 0;
#pragma unroll
    for (uint32_t c = 0; c < 4; ++c) {
        uint32_t score = 0;
#pragma unroll
        for (uint32_t x = 0; x < 6; ++x) {
            const uint32_t n = (uint32_t)(cnt >> (8u * x)) & 0xFFu;
            score += n * (x <= c ? m : (x == c + 1u ? m + c + 1u : m + c + 2u));
        }
        if (score > best_score) best_score = score, best = c;
    }
    const uint32_t cut = m + best;
    uint32_t acc = m | (best << 6);  // the output word being filled; byte 0 is the header
    for (uint32_t q = 0; q < 31; ++q) {
        uint32_t x = 0;
#pragma unroll
        for (int k = 4; k >= 0; --k) {
            const uint32_t j = q * 5u + (uint32_t)k;
            uint32_t d = 0;
            if (j < n) {
                const uint32_t v = val(j);
                d = v <= cut ? 0u : (v == cut + 1u ? 1u : 2u);
            }
            x = x * 3u + d;
        }
        const uint32_t byte = 1u + q;
        acc |= x << ((byte & 3u) * 8u);
        if ((byte & 3u) == 3u) {  // bytes 3, 7, ..., 31 end words 0 .. 7
            put(byte >> 2, acc);
            acc = 0;
        }
    }
}

}  // namespace ccs
