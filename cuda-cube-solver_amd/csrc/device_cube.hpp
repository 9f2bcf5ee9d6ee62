// Packed-cube arithmetic for CDNA4: moves, conjugation, table index, unrank
// and table lookups, all integer VALU + LDS work.
//
// Functions are __host__ __device__ so that the library's diagnostic entry
// ccs_index_host can evaluate exactly this LUT logic on the CPU; kernels are
// the only product users.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>

#include "lut.hpp"

#define CCS_HD __host__ __device__ __forceinline__

namespace ccs {

struct Frame {
    uint32_t c0, c1, e0, e1, e2;
};

constexpr uint32_t SOLVED_C0 = 0x03020100u, SOLVED_C1 = 0x07060504u;
constexpr uint32_t SOLVED_E0 = 0x03020100u, SOLVED_E1 = 0x07060504u, SOLVED_E2 = 0x0B0A0908u;

CCS_HD Frame solved_frame() { return Frame{SOLVED_C0, SOLVED_C1, SOLVED_E0, SOLVED_E1, SOLVED_E2}; }
CCS_HD bool is_solved(const Frame& f) {
    return ((f.c0 ^ SOLVED_C0) | (f.c1 ^ SOLVED_C1) | (f.e0 ^ SOLVED_E0) | (f.e1 ^ SOLVED_E1) | (f.e2 ^ SOLVED_E2)) == 0;
}

// v_perm_b32: bytes of {s0:s1} picked by selector bytes (see lut.hpp).
CCS_HD uint32_t vperm(uint32_t s0, uint32_t s1, uint32_t sel) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_perm(s0, s1, sel);
#else
    uint64_t data = ((uint64_t)s0 << 32) | s1;
    uint32_t r = 0;
    for (int i = 0; i < 4; ++i) {
        uint32_t s = (sel >> (8 * i)) & 0xFFu, b;
        if (s >= 13) b = 0xFFu;
        else if (s == 12) b = 0;
        else if (s >= 8) b = ((data >> (16 * (s - 8) + 15)) & 1u) ? 0xFFu : 0u;
        else b = (uint32_t)(data >> (8 * s)) & 0xFFu;
        r |= b << (8 * i);
    }
    return r;
#endif
}

CCS_HD uint32_t popc(uint32_t x) { return (uint32_t)__builtin_popcount(x); }
CCS_HD uint32_t ctz(uint32_t x) { return (uint32_t)__builtin_ctz(x); }
CCS_HD uint32_t shl(uint32_t x, uint32_t s) { return s >= 32u ? 0u : x << s; }
CCS_HD uint64_t shl64(uint64_t x, uint32_t s) { return s >= 64u ? 0u : x << s; }

// Sum of the four byte products of a and b, plus c (v_dot4_u32_u8, full rate;
// 32-bit multiplies are quarter rate).
CCS_HD uint32_t dot4(uint32_t a, uint32_t b, uint32_t c) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_udot4(a, b, c, false);
#else
    for (int i = 0; i < 4; ++i) c += ((a >> (8 * i)) & 0xFFu) * ((b >> (8 * i)) & 0xFFu);
    return c;
#endif
}
// Product of two values below 2^24 (v_mul_u32_u24, full rate).
CCS_HD uint32_t mul24(uint32_t a, uint32_t b) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __umul24(a, b);
#else
    return a * b;
#endif
}

// Bit `bit` of each byte -> 4-bit mask (byte 0 -> bit 0).
CCS_HD uint32_t gather4(uint32_t v, int bit) { return dot4((v >> bit) & 0x01010101u, 0x08040201u, 0u); }

// Corner twist stored in bits 3..5 of each byte may reach 4 after an add:
// reduce mod 3 by taking 0x18 off the bytes whose twist is 3 or 4 (s marks
// them with 0x20 per byte, and 0x20 - 0x08 = 0x18).
CCS_HD uint32_t twist_mod3(uint32_t x) {
    uint32_t s = (x + 0x08080808u) & 0x20202020u;
    return x - s + (s >> 2);
}

// Relabel edge cubies of a word through a 12-byte table (rel0..rel2): new
// byte = rel[cubie] with the existing flip bit xored in.  Table bytes are
// below 0x80, so selectors 8..11 (sign fill) read as zero: cubie c selects
// c from rel0:rel1 (8..11 -> zero) and (c & 11) ^ 8 from rel2 (0..7 -> 8..11
// -> zero, 8..11 -> 0..3).
CCS_HD uint32_t relabel_edges(uint32_t v, uint32_t r0, uint32_t r1, uint32_t r2) {
    uint32_t x = vperm(r1, r0, v & 0x0F0F0F0Fu);                  // cubies 0..7
    uint32_t y = vperm(r2, r2, (v & 0x0B0B0B0Bu) ^ 0x08080808u);  // cubies 8..11
    return (x | y) ^ (v & 0x10101010u);
}

// state then move (right multiplication; kernels.cu:84-101 computes the same).
CCS_HD void right_mul(Frame& f, const RightRec& r) {
    uint32_t a = twist_mod3(vperm(f.c1, f.c0, r.csel[0]) + r.cadd[0]);
    uint32_t b = twist_mod3(vperm(f.c1, f.c0, r.csel[1]) + r.cadd[1]);
    uint32_t e0 = (vperm(f.e1, f.e0, r.elo[0]) | vperm(f.e2, f.e2, r.ehi[0])) ^ r.eflip[0];
    uint32_t e1 = (vperm(f.e1, f.e0, r.elo[1]) | vperm(f.e2, f.e2, r.ehi[1])) ^ r.eflip[1];
    uint32_t e2 = (vperm(f.e1, f.e0, r.elo[2]) | vperm(f.e2, f.e2, r.ehi[2])) ^ r.eflip[2];
    f = Frame{a, b, e0, e1, e2};
}

// move then state (left multiplication; used by the inverse frames).
CCS_HD void left_mul(Frame& f, const LeftRec& r) {
    uint32_t a = twist_mod3(vperm(r.lc[1], r.lc[0], f.c0 & 0x07070707u) + (f.c0 & 0x18181818u));
    uint32_t b = twist_mod3(vperm(r.lc[1], r.lc[0], f.c1 & 0x07070707u) + (f.c1 & 0x18181818u));
    uint32_t e0 = relabel_edges(f.e0, r.le[0], r.le[1], r.le[2]);
    uint32_t e1 = relabel_edges(f.e1, r.le[0], r.le[1], r.le[2]);
    uint32_t e2 = relabel_edges(f.e2, r.le[0], r.le[1], r.le[2]);
    f = Frame{a, b, e0, e1, e2};
}

// Corners of R^k g R^-k (lut.hpp ConjRec); the edges are left as they are; the
// edge coordinate conjugates them through the composite tables instead.
CCS_HD void conj_corners(Frame& f, const ConjRec& r) {
    const uint32_t a = twist_mod3(vperm(f.c1, f.c0, r.csel[0]) + r.cadd[0]);
    const uint32_t b = twist_mod3(vperm(f.c1, f.c0, r.csel[1]) + r.cadd[1]);
    f.c0 = twist_mod3(vperm(r.lc[1], r.lc[0], a & 0x07070707u) + (a & 0x18181818u));
    f.c1 = twist_mod3(vperm(r.lc[1], r.lc[0], b & 0x07070707u) + (b & 0x18181818u));
}

CCS_HD void apply_move(const CubeLUT& L, Frame& f, int m) { right_mul(f, L.rr[0][m]); }
CCS_HD void apply_left(const CubeLUT& L, Frame& f, int m) { left_mul(f, L.lr[1][m]); }

// Byte i (0..15) of a 16-byte table held in four words, by v_perm.
CCS_HD uint32_t byte16(uint32_t w0, uint32_t w1, uint32_t w2, uint32_t w3, uint32_t i) {
    const uint32_t lo = vperm(w1, w0, i & 7u), hi = vperm(w3, w2, i & 7u);
    return ((i & 8u) ? hi : lo) & 0xFFu;
}
// Lexicographic rank (itertools.combinations order) of a 4-of-8 position mask.
// The positions below 4 are a prefix of the sorted tuple, so the tuples with
// one low part are contiguous and rank = TL[low nibble] + TH[high nibble],
// TH being the rank of the high part among subsets of its size.  Equal to the
// ccomb_rank table on all 70 masks; no LDS round trip before the classidx read.
CCS_HD uint32_t comb8_rank(uint32_t m) {
    return byte16(0x09331F45u, 0x0127133Du, 0x052D1941u, 0x00230F37u, m & 15u) +
           byte16(0x00010000u, 0x00030102u, 0x01040203u, 0x00030205u, m >> 4);
}

// twist (3^7) * 70 + rank of the D-corner positions (coords.py:112-118).
CCS_HD void corner_parts(const Frame& f, uint32_t& twist, uint32_t& pos) {
    uint32_t olo = (f.c0 >> 3) & 0x03030303u, ohi = (f.c1 >> 3) & 0x00030303u;
    // sum of co[i] * 3^i over corners 0..6 (corner 7 is fixed by the others)
    twist = dot4(olo, 0x1B090301u, mul24(dot4(ohi, 0x00090301u, 0u), 81u));
    uint32_t dmask = dot4((f.c0 >> 2) & 0x01010101u, 0x08040201u, dot4((f.c1 >> 2) & 0x01010101u, 0x80402010u, 0u));
    pos = comb8_rank(dmask);
}
CCS_HD uint32_t corner_raw(const CubeLUT& L, const Frame& f) {
    uint32_t twist, pos;
    corner_parts(f, twist, pos);
    return mul24(twist, 70u) + pos;
}

// classidx[twist * 70 + pos] (class << 4 | sym) from the ClassLUT tables
// (lut.hpp): one read of the twist's entry, then the placement and, for
// several-symmetry twists, the symmetry, then the rank within a symmetric
// orbit.  Indices stay in their tables for every lane, also where the value
// is not used.
CCS_HD uint32_t corner_class(const ClassLUT& C, uint32_t twist, uint32_t pos) {
    const uint32_t x = C.tw[twist][0], y = C.tw[twist][1];
    const uint32_t s0 = x & 15u, mi = x >> 4, mo = y >> 16;
    const uint32_t q = C.ps[s0][pos];
    const uint32_t ms = C.ms[mi ? mi - 1u : 0u][pos];
    const uint32_t rk = C.rk[mo ? mo - 1u : 0u][q];
    return (((y & 0xFFFFu) + (mo ? rk : q)) << 4) | (mi ? ms : s0);
}
CCS_HD uint32_t corner_class(const ClassLUT& C, const Frame& f) {
    uint32_t twist, pos;
    corner_parts(f, twist, pos);
    return corner_class(C, twist, pos);
}

// K = 0: the 304 kind; K > 0: a group kind with K E-content classes.
template <int K>
struct KindT {
    static constexpr uint64_t N_EDGE = K == 0 ? 495ull * 70ull * 16ull : 495ull * 256ull * (uint64_t)K;
};

// Edge coordinate of M_s g M_s^-1 (coords.py:121-154 for 304; kernels.cu:175-186
// for the group kinds), from the packed edges; s indexes the edge conjugation
// tables (below 16: a UD symmetry; 16 k + s: that symmetry after R^k).
template <int K>
CCS_HD uint32_t edge_coord(const CubeLUT& L, const Frame& f, uint32_t s) {
    uint32_t e[3];
#pragma unroll
    for (int w = 0; w < 3; ++w) {
        uint32_t x = (vperm(f.e1, f.e0, L.cs_lo[s][w]) | vperm(f.e2, f.e2, L.cs_hi[s][w])) ^ L.cs_flip[s][w];
        e[w] = relabel_edges(x, L.cs_rel[s][0], L.cs_rel[s][1], L.cs_rel[s][2]);
    }
    uint32_t m0 = gather4(e[0], 3), m1 = gather4(e[1], 3), m2 = gather4(e[2], 3);
    uint32_t emask = m0 | (m1 << 4) | (m2 << 8);
    uint32_t lo6 = emask & 63u;
    uint32_t slice = (uint32_t)L.slice_lo[lo6] + L.slice_hi[popc(lo6)][emask >> 6];
    uint32_t n0 = popc(m0), n1 = popc(m1);
    // E bytes in position order, compacted into one word
    uint32_t ew = vperm(e[0], e[0], L.csel[m0]) | shl(vperm(e[1], e[1], L.csel[m1]), 8u * n0) |
                  shl(vperm(e[2], e[2], L.csel[m2]), 8u * (n0 + n1));
    uint32_t eflip = gather4(ew, 4);
    // non-E bytes in position order: 8 bytes in two words
    uint32_t k0 = 4u - n0, k1 = 4u - n1;
    uint64_t nw = (uint64_t)vperm(e[0], e[0], L.csel[m0 ^ 15u]) |
                  shl64(vperm(e[1], e[1], L.csel[m1 ^ 15u]), 8u * k0) |
                  shl64(vperm(e[2], e[2], L.csel[m2 ^ 15u]), 8u * (k0 + k1));
    uint32_t nlo = (uint32_t)nw, nhi = (uint32_t)(nw >> 32);
    if constexpr (K == 0) {
        // U-layer edges are cubies 0..3 (bits 2 and 3 clear)
        uint32_t umask = gather4(~(nlo | (nlo >> 1)), 2) | (gather4(~(nhi | (nhi >> 1)), 2) << 4);
        return (mul24(slice, 70u) + comb8_rank(umask)) * 16u + eflip;
    } else {
        uint32_t lab = ew & 0x03030303u;
        uint32_t key = (lab | (lab >> 6) | (lab >> 12) | (lab >> 18)) & 0xFFu;
        uint32_t nflip = gather4(nlo, 4) | (gather4(nhi, 4) << 4);
        uint32_t c = L.key_class[key * 16u + eflip];
        return mul24(slice * 256u + nflip, (uint32_t)K) + c;
    }
}

// Table index: corner class * N_EDGE + edge coordinate after the class
// symmetry (coords.py:347-352).
template <int K>
CCS_HD uint64_t table_index(const CubeLUT& L, const uint32_t* __restrict__ classidx, const Frame& f) {
    uint32_t packed = classidx[corner_raw(L, f)];
    return (uint64_t)(packed >> 4) * KindT<K>::N_EDGE + edge_coord<K>(L, f, packed & 15u);
}

// The same index with the corner class taken from the ClassLUT tables, as the
// search computes it.
template <int K>
CCS_HD uint64_t table_index_c(const CubeLUT& L, const ClassLUT& C, const Frame& f) {
    const uint32_t packed = corner_class(C, f);
    return (uint64_t)(packed >> 4) * KindT<K>::N_EDGE + edge_coord<K>(L, f, packed & 15u);
}

// Lower bound from the solving table (kernels.cu:396-434 defines the formats):
// F = 0: block (u8 minimum + 124 two-bit offsets per 32 B), F = 1: base 3
// (header m | c << 6, 155 digits).
template <int F>
CCS_HD uint32_t table_lb(const uint8_t* __restrict__ tbl, uint64_t idx) {
    if constexpr (F == 0) {
        uint64_t b = idx / 124u;
        uint32_t pos = 8u + 2u * (uint32_t)(idx - b * 124u);
        const uint32_t* p = reinterpret_cast<const uint32_t*>(tbl + b * 32u);
        return (p[0] & 0xFFu) + ((p[pos >> 5] >> (pos & 31u)) & 3u);
    } else {
        uint64_t b = idx / 155u;
        uint32_t j = (uint32_t)(idx - b * 155u);
        const uint8_t* p = tbl + b * 32u;
        uint32_t hdr = p[0], byte = p[1u + j / 5u], k = j % 5u;
        uint32_t d = (byte / (k == 0 ? 1u : k == 1 ? 3u : k == 2 ? 9u : k == 3 ? 27u : 81u)) % 3u;
        uint32_t m = hdr & 63u;
        return d == 0u ? m : m + (hdr >> 6) + d;
    }
}

// Lower bound of a frame from the solving table, with the block number and
// the entry within the block computed in 32-bit arithmetic: with
// N_EDGE = Q * D + R, index / D = class * Q + (class * R + edge) / D, and the
// bracket stays below 2^23 (class * N_EDGE itself reaches 2^35).  Split in two
// so that a caller can put several probes' table reads in flight before it
// uses any: probe_block (from the frame's classidx word) and block_value.
// `frame` k: f holds R^k g R^-k's corners and g's edges (the edges are
// conjugated by R^k with the class symmetry, in one table).
//
// F = 2 (FMT_HALVES) is the block format laid out for the search: each 32-byte
// block as two self-contained 16-byte halves, entries 0..61 and 62..123, each
// half the block minimum in its low 4 bits and the entries' 2-bit offsets from
// bit 4 on.  A probe then reads one 16-byte half with a single load; the
// values are the block format's, bit for bit.  probe_block returns the half
// and the entry within it.
constexpr int FMT_HALVES = 2;
constexpr uint32_t HALF_ENTRIES = 62;
template <int K, int F>
CCS_HD const uint8_t* probe_block(const CubeLUT& L, uint32_t packed, uint32_t frame, const uint8_t* __restrict__ tbl,
                                  const Frame& f, uint32_t& j) {
    constexpr uint32_t D = F == 1 ? 155u : 124u;
    constexpr uint32_t Q = (uint32_t)(KindT<K>::N_EDGE / D), R = (uint32_t)(KindT<K>::N_EDGE % D);
    static_assert(Q < (1u << 24), "class * Q must stay a 24-bit product");
    const uint32_t cls = packed >> 4;
    const uint32_t x = mul24(cls, R) + edge_coord<K>(L, f, (packed & 15u) + 16u * frame);
    const uint32_t q = x / D;
    j = x - mul24(q, D);
    const uint8_t* p = tbl + (uint64_t)(mul24(cls, Q) + q) * 32u;
    if constexpr (F == FMT_HALVES) {
        const uint32_t h = j >= HALF_ENTRIES ? 1u : 0u;
        j -= h * HALF_ENTRIES;
        p += 16u * h;
    }
    return p;
}
template <int F>
CCS_HD uint32_t block_value(const CubeLUT& L, const uint8_t* p, uint32_t j) {
    if constexpr (F == FMT_HALVES) {
        const uint4 w = *reinterpret_cast<const uint4*>(p);
        const uint32_t pos = 4u + 2u * j, k = pos >> 5;
        const uint32_t word = k == 0u ? w.x : k == 1u ? w.y : k == 2u ? w.z : w.w;
        return (w.x & 15u) + ((word >> (pos & 31u)) & 3u);
    } else if constexpr (F == 0) {
        const uint32_t* w = reinterpret_cast<const uint32_t*>(p);
        const uint32_t pos = 8u + 2u * j;
        return (w[0] & 0xFFu) + ((w[pos >> 5] >> (pos & 31u)) & 3u);
    } else {
        const uint32_t hdr = p[0], byte = p[1u + j / 5u];
        const uint32_t d = (L.b3digits[byte] >> (2u * (j % 5u))) & 3u;
        const uint32_t m = hdr & 63u;
        return d == 0u ? m : m + (hdr >> 6) + d;
    }
}
template <int K, int F>
CCS_HD uint32_t frame_bound(const CubeLUT& L, const uint32_t* __restrict__ classidx,
                            const uint8_t* __restrict__ tbl, const Frame& f) {
    uint32_t j;
    const uint8_t* p = probe_block<K, F>(L, classidx[corner_raw(L, f)], 0u, tbl, f, j);
    return block_value<F>(L, p, j);
}

// Representative frame of a table entry (the BFS expands it); any member of
// the coset works since moves act on cosets.
template <int K>
CCS_HD Frame unrank(const CubeLUT& L, const uint32_t* __restrict__ class_rep, uint64_t idx) {
    uint64_t cls = idx / KindT<K>::N_EDGE;
    uint32_t ec = (uint32_t)(idx - cls * KindT<K>::N_EDGE);
    uint32_t raw = class_rep[cls];
    uint32_t twist = raw / 70u, dmask = L.ccomb_unrank[raw - twist * 70u];
    uint32_t c[2] = {0, 0}, osum = 0, nu = 0, nd = 0;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        uint32_t o;
        if (i < 7) {
            o = twist % 3u;
            twist /= 3u;
            osum += o;
        } else {
            o = (3u - osum % 3u) % 3u;
        }
        uint32_t cp = ((dmask >> i) & 1u) ? 4u + nd++ : nu++;
        c[i >> 2] |= (cp | (o << 3)) << ((i & 3) * 8);
    }
    uint32_t emask, eord, umask, eflips, nflips;
    if constexpr (K == 0) {
        uint32_t eo = ec % 16u, rest = ec / 16u;
        emask = L.slice_unrank[rest / 70u];
        umask = L.ccomb_unrank[rest % 70u];
        eord = 0xE4u;
        eflips = eo;
        nflips = 0;
    } else {
        uint32_t gc = ec % (uint32_t)K, grest = ec / (uint32_t)K;
        nflips = grest % 256u;
        emask = L.slice_unrank[grest / 256u];
        uint32_t content = L.gen_rep[popc(nflips) & 1u][gc];
        eord = L.eorder_unrank[content >> 4];
        eflips = content & 15u;
        umask = 0x0Fu;
    }
    uint32_t e[3] = {0, 0, 0}, ne = 0, no = 0, nue = 0, nde = 0;
#pragma unroll
    for (int i = 0; i < 12; ++i) {
        uint32_t b, fl;
        if ((emask >> i) & 1u) {
            b = 8u | ((eord >> (2 * ne)) & 3u);
            fl = (eflips >> ne) & 1u;
            ++ne;
        } else {
            b = ((umask >> no) & 1u) ? nue++ : 4u + nde++;
            fl = (nflips >> no) & 1u;
            ++no;
        }
        e[i >> 2] |= (b | (fl << 4)) << ((i & 3) * 8);
    }
    return Frame{c[0], c[1], e[0], e[1], e[2]};
}

}  // namespace ccs
