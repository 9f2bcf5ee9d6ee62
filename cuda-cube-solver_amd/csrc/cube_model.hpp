// Host-side cube model, symmetries, coordinates and table kinds.
//
// Cubie convention follows the reference (src/cuda_cube_solver/cubie.py:1-51):
// cp[position] = cubie at that position; "a then b" gives cp = a.cp[b.cp[i]].
#pragma once
#include <array>
#include <cstdint>
#include <string>
#include <vector>

#include "lut.hpp"

namespace ccs {

struct Cubie {
    uint8_t cp[8], co[8], ep[12], eo[12];  // 40 bytes, the C-ABI cube layout
};
static_assert(sizeof(Cubie) == 40, "C-ABI cube layout");

struct Packed {
    uint32_t w[5];  // c0, c1, e0, e1, e2
};

Cubie solved();
Cubie then(const Cubie& a, const Cubie& b);  // b applied after a (cubie.py:97-105)
Cubie inverse(const Cubie& a);
bool is_legal(const Cubie& a);
const Cubie& move(int m);                    // HTM moves U,U2,U',R,..,B' (cubie.py:115-124)
Packed pack(const Cubie& c);
Cubie unpack(const Packed& p);

// Cubie-level action of a signed-permutation symmetry matrix (a cube symmetry).
struct SymAction {
    int mat[9];
    uint8_t cpos[8], csig[8][3];    // corner position map and sticker map per position
    uint8_t epos[12], esig[12][2];  // edge position map and sticker map per position
};
SymAction sym_from_matrix(const int m[9]);
Cubie conjugate(const SymAction& s, const Cubie& c);  // M g M^-1
SymAction compose_sym(const SymAction& a, const SymAction& b);  // conjugation by a after b

// The 16 UD-preserving symmetries in the reference's order
// (geometry.py:96-116) and the URF-diagonal rotation (geometry.py:119-124).
const std::array<SymAction, 16>& ud_syms();
const SymAction& axis_rotation();
int axis_move(int m);  // index of R m R^-1
void axis_frames(const Cubie& c, Cubie out[6]);  // solver.py:58-64

// Corner classes under the 16 UD symmetries (coords.py:327-340).
struct CornerClasses {
    std::vector<uint32_t> classidx;  // raw -> class << 4 | sym
    std::vector<uint32_t> rep;       // class -> representative raw
    std::vector<uint16_t> stab;      // class -> stabiliser mask of the representative
    ClassLUT class_lut;              // the same classes from small tables (lut.hpp)
    int class_lut_rows[2] = {0, 0};  // several-symmetry twists and orbits found while building it
};
const CornerClasses& corner_classes();
uint32_t corner_raw(const Cubie& c);

// Table kinds: "304" (EdgeKind(3,4), coords.py:37-76) or a GroupKind
// (kinds.py:56-107) numbered by E-content class.
struct Kind {
    std::string name;
    bool is304 = false;
    int nec = 0;
    uint64_t n_edge = 0;
    uint64_t n_entries = 0;
    uint8_t gen_class[384] = {};
    uint16_t gen_rep[2][48] = {};
};
// Returns false (and sets err) for an unknown name or a group that the UD
// symmetries do not normalise.
bool make_kind(const std::string& name, Kind& out, std::string& err);
bool make_group_kind(const std::string& name, const std::vector<std::array<int, 8>>& gens, Kind& out,
                     std::string& err);

// Host reference of the device index (used for probes on the host side and
// for tests of the library itself).
uint64_t table_index(const Kind& k, const Cubie& c);

// Canonical move sequences of a length, row-major in move order (solver.py:33-48).
bool allowed_after(int last, int m);
std::vector<uint8_t> canonical_prefixes(int length);

// Build the LDS lookup tables for a kind.
void build_lut(const Kind& k, CubeLUT& lut);

}  // namespace ccs
