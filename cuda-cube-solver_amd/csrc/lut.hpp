// Constant lookup tables shared by every kernel, built on the host and copied
// into LDS by each workgroup.  Layout is private to this library (the
// reference's `Tables` blob, kernels.cu:44-77, is laid out for CUDA's
// __byte_perm nibble selectors; these are selectors for CDNA's v_perm_b32).
//
// Packed cube ("frame"): corners in two u32 (byte = cp | co << 3), edges in
// three u32 (byte = ep | eo << 4), as in the reference (kernels.cu:79-82).
//
// v_perm_b32 (__builtin_amdgcn_perm(S0, S1, sel)): result byte i is chosen by
// selector byte i: 0-3 = byte of S1, 4-7 = byte of S0, 8-11 = sign fill of
// bytes 1/3/5/7, 12 = 0x00, >= 13 = 0xFF.  Every table byte fed through a
// selector here is < 0x80, so selectors 8-11 yield zero as well.
#pragma once
#include <cstdint>

namespace ccs {

constexpr uint32_t SEL_ZERO = 0x0Cu;

// Move records.  The search's state takes a move as right multiplication by
// it; the state's inverse takes it as left multiplication by the move's
// inverse.  (Axis frames 1 and 2 are reached by conjugation, ConjRec below.)
struct alignas(16) RightRec {
    uint32_t csel[2];   // corner word w: source corner index per byte (perm over c1:c0)
    uint32_t cadd[2];   // co << 3 added per byte
    uint32_t elo[3];    // edge word w: source index < 8 (perm over e1:e0), else SEL_ZERO
    uint32_t ehi[3];    // source index - 8 for sources >= 8 (perm over e2:e2), else SEL_ZERO
    uint32_t eflip[3];  // eo << 4 xored per byte
    uint32_t pad[3];
};
struct alignas(16) LeftRec {
    uint32_t lc[2];     // corners: move.cp[c] | move.co[c] << 3 for c = 0..7
    uint32_t le[3];     // edges:   move.ep[c] | move.eo[c] << 4 for c = 0..11
    uint32_t pad[3];
};

// Corners of R^k g R^-k, R the URF-diagonal rotation that turns axis frame 0
// into frame k: the position part (source corner and twist added per target
// byte, as in RightRec), then the cubie relabel with its twist (as in
// LeftRec).  A rotation turns every corner's stickers cyclically, so both
// twist parts are additions mod 3.
struct alignas(16) ConjRec {
    uint32_t csel[2];
    uint32_t cadd[2];
    uint32_t lc[2];
    uint32_t pad[2];
};
// Edge conjugation tables: symmetry index 16 k + s conjugates by the UD
// symmetry s after the axis rotation R^k (k = 0: the UD symmetry alone).
constexpr int N_ESYM = 48;

struct alignas(16) CubeLUT {
    // rr[0][m] applies move m to the state (right multiplication by m), rr[1][m]
    // undoes it; lr[0][m] applies it to the inverse (left multiplication by
    // m^-1), lr[1][m] undoes that (left multiplication by m).
    RightRec rr[2][18];
    LeftRec lr[2][18];
    ConjRec cj[3];  // corner conjugation by R^k (cj[0]: identity)
    // edge conjugation by symmetry t = 16 k + s (M g M^-1, M = UD symmetry s after
    // R^k): position part, then cubie relabel
    uint32_t cs_lo[N_ESYM][3];
    uint32_t cs_hi[N_ESYM][3];
    uint32_t cs_flip[N_ESYM][3];  // flip of the source position, << 4 per byte
    uint32_t cs_rel[N_ESYM][3];   // new byte for each cubie c: perm(c) | flip(c) << 4
    uint32_t csel[16];        // selector compacting the bytes picked by a 4-bit mask to the low end
    uint16_t slice_lo[64];    // 12C4 lexicographic rank = slice_lo[low 6 bits] + slice_hi[|low|][high 6 bits]
    uint8_t slice_hi[5][64];
    uint8_t ccomb_rank[256];  // 8C4 lexicographic rank
    uint8_t eorder_rank[256]; // 4 two-bit labels (first in low bits) -> permutation rank
    uint8_t gen_class[384];   // group kinds: E content (order rank * 16 + flips) -> class
    uint8_t key_class[256 * 16];  // the same by order key (four 2-bit labels) * 16 + flips: one lookup
    uint8_t s_mul[16][16];    // index of M_t M_s among the UD symmetries
    uint8_t ccomb_unrank[72];
    uint16_t slice_unrank[496];
    uint8_t eorder_unrank[24];
    uint16_t gen_rep[2][48];  // class -> E content, by flip parity of the non-E edges
    uint16_t b3digits[256];   // base-3 byte (< 243) -> its five digits, two bits each
    uint32_t amask[19];       // canonical successors of the last move (index last + 1), 18-bit mask
    uint32_t pad1;
};

static_assert(sizeof(CubeLUT) % 16 == 0, "LUT copied in 16-byte units");

// Corner class and class symmetry (the classidx word, class << 4 | sym) of a
// raw corner coordinate twist * 70 + pos, from small tables instead of the
// 153,090-entry classidx array, so that the search finds them in LDS.
//
// Under a UD symmetry s the twist and the D-corner placement transform
// separately: conj(twist * 70 + pos, s) = tw_s(twist) * 70 + ps[s][pos].  The
// orbit minimum therefore takes the smallest twist tmin first; M(twist), the
// symmetries that reach it, is a coset of the stabiliser of tmin, and the
// class symmetry is the first symmetry of M(twist) that also minimises the
// placement.  Classes are numbered in order of their representatives' raw
// coordinate, so a class is base[tmin] plus its placement's rank among tmin's
// representatives, and that rank depends only on q = ps[s0][pos], s0 being the
// first symmetry of M(twist).  The host checks these tables against classidx
// for every raw coordinate.
constexpr uint32_t N_TWIST = 2187;
constexpr uint32_t N_DPOS = 70;
constexpr uint32_t N_MULTI_TWIST = 363;  // twists reaching their minimum through several symmetries
constexpr uint32_t N_MULTI_ORBIT = 54;   // minimal twists with a nontrivial stabiliser

struct alignas(16) ClassLUT {
    // per twist: x = s0 | mi << 4 (mi: 1 + row of ms, 0 if s0 is the only symmetry);
    // y = base[tmin] | mo << 16 (mo: 1 + row of rk, 0 if tmin's stabiliser is trivial)
    uint32_t tw[N_TWIST][2];
    uint8_t ms[N_MULTI_TWIST][N_DPOS];  // class symmetry of (twist, pos) for several-symmetry twists
    uint8_t ps[16][N_DPOS];             // D-corner placement rank after symmetry s
    uint8_t rk[N_MULTI_ORBIT][N_DPOS];  // rank of the representative among tmin's, by q
    uint8_t pad[2];
};
static_assert(sizeof(ClassLUT) % 16 == 0, "LUT copied in 16-byte units");

}  // namespace ccs
