// Host-side cube model, symmetries, corner classes, table kinds and the LDS
// lookup tables.  Moves and symmetries are derived here from the sticker
// geometry (face normals and signed-permutation matrices) and reproduce the
// reference's numbering: moves of cubie.py:57-124, symmetry order of
// geometry.py:96-116, corner classes of coords.py:327-340, group-class
// numbering of kinds.py:56-96.
#include "cube_model.hpp"

#include <algorithm>
#include <cstring>
#include <mutex>

namespace ccs {

namespace {

// Face normals in U, R, F, D, L, B order (x = R, y = U, z = F).
constexpr int FACE_N[6][3] = {{0, 1, 0}, {1, 0, 0}, {0, 0, 1}, {0, -1, 0}, {-1, 0, 0}, {0, 0, -1}};
// Cubie positions by their faces, sticker k = face k (cubie.py:12-14).
constexpr int CORNER_FACES[8][3] = {{0, 1, 2}, {0, 2, 4}, {0, 4, 5}, {0, 5, 1},
                                    {3, 2, 1}, {3, 4, 2}, {3, 5, 4}, {3, 1, 5}};
constexpr int EDGE_FACES[12][2] = {{0, 1}, {0, 2}, {0, 4}, {0, 5}, {3, 1}, {3, 2},
                                   {3, 4}, {3, 5}, {2, 1}, {2, 4}, {5, 4}, {5, 1}};

struct V3 {
    int x, y, z;
    bool operator==(const V3& o) const { return x == o.x && y == o.y && z == o.z; }
};

V3 normal(int f) { return {FACE_N[f][0], FACE_N[f][1], FACE_N[f][2]}; }
V3 add(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
V3 mul(const int* m, V3 v) {
    return {m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z,
            m[6] * v.x + m[7] * v.y + m[8] * v.z};
}
int face_of(V3 n) {
    for (int f = 0; f < 6; ++f)
        if (normal(f) == n) return f;
    return -1;
}
V3 corner_center(int p) {
    return add(add(normal(CORNER_FACES[p][0]), normal(CORNER_FACES[p][1])), normal(CORNER_FACES[p][2]));
}
V3 edge_center(int p) { return add(normal(EDGE_FACES[p][0]), normal(EDGE_FACES[p][1])); }
int corner_at(V3 c) {
    for (int p = 0; p < 8; ++p)
        if (corner_center(p) == c) return p;
    return -1;
}
int edge_at(V3 c) {
    for (int p = 0; p < 12; ++p)
        if (edge_center(p) == c) return p;
    return -1;
}

// Cubie action of any linear map of the cube (a symmetry, or a layer turn
// restricted to the cubies with center . axis > 0).
struct Action {
    uint8_t cpos[8], csig[8][3], epos[12], esig[12][2];
};

Action action_of(const int* m) {
    Action a{};
    for (int p = 0; p < 8; ++p) {
        int q = corner_at(mul(m, corner_center(p)));
        a.cpos[p] = (uint8_t)q;
        for (int k = 0; k < 3; ++k) {
            int f = face_of(mul(m, normal(CORNER_FACES[p][k])));
            for (int j = 0; j < 3; ++j)
                if (CORNER_FACES[q][j] == f) a.csig[p][k] = (uint8_t)j;
        }
    }
    for (int p = 0; p < 12; ++p) {
        int q = edge_at(mul(m, edge_center(p)));
        a.epos[p] = (uint8_t)q;
        for (int k = 0; k < 2; ++k) {
            int f = face_of(mul(m, normal(EDGE_FACES[p][k])));
            for (int j = 0; j < 2; ++j)
                if (EDGE_FACES[q][j] == f) a.esig[p][k] = (uint8_t)j;
        }
    }
    return a;
}

// A clockwise quarter turn of face f (viewed from outside): rotation by -90
// degrees about the face normal, v -> n (n.v) - n x v.
Cubie quarter_turn(int f) {
    V3 n = normal(f);
    int m[9];
    V3 basis[3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int col = 0; col < 3; ++col) {
        V3 v = basis[col];
        int d = n.x * v.x + n.y * v.y + n.z * v.z;
        V3 cr = {n.y * v.z - n.z * v.y, n.z * v.x - n.x * v.z, n.x * v.y - n.y * v.x};
        V3 r = {n.x * d - cr.x, n.y * d - cr.y, n.z * d - cr.z};
        m[0 + col] = r.x, m[3 + col] = r.y, m[6 + col] = r.z;
    }
    Action a = action_of(m);
    Cubie c = solved();
    for (int p = 0; p < 8; ++p) {
        V3 ctr = corner_center(p);
        if (ctr.x * n.x + ctr.y * n.y + ctr.z * n.z <= 0) continue;
        int q = a.cpos[p];
        c.cp[q] = (uint8_t)p;                // position q receives the cubie from p
        c.co[q] = (uint8_t)(a.csig[p][0] % 3);  // its sticker 0 now sits at slot csig
    }
    for (int p = 0; p < 12; ++p) {
        V3 ctr = edge_center(p);
        if (ctr.x * n.x + ctr.y * n.y + ctr.z * n.z <= 0) continue;
        int q = a.epos[p];
        c.ep[q] = (uint8_t)p;
        c.eo[q] = (uint8_t)(a.esig[p][0] & 1);
    }
    return c;
}

struct Statics {
    Cubie moves[18];
    std::array<SymAction, 16> ud;
    SymAction rot;
    int axis_map[18];
    uint8_t s_mul[16][16];
    CornerClasses cc;
    uint16_t ccomb_rank[256], ccomb_unrank[70], slice_rank[4096], slice_unrank[495];
    uint8_t eorder_rank[256], eorder_unrank[24];
    int perms4[24][4];
};

Statics* g_statics = nullptr;
std::once_flag g_once;

void comb(int bits, uint16_t* rank, uint16_t* unrank) {
    // lexicographic order of the sorted 4-subsets (itertools.combinations order)
    for (int i = 0; i < (1 << bits); ++i) rank[i] = 0xFFFF;
    int r = 0;
    for (int a = 0; a < bits; ++a)
        for (int b = a + 1; b < bits; ++b)
            for (int c = b + 1; c < bits; ++c)
                for (int d = c + 1; d < bits; ++d) {
                    uint16_t mask = (uint16_t)((1 << a) | (1 << b) | (1 << c) | (1 << d));
                    rank[mask] = (uint16_t)r;
                    unrank[r++] = mask;
                }
}

Cubie corners_from_raw(uint32_t raw) {
    Cubie c = solved();
    uint32_t twist = raw / 70u, dmask = g_statics->ccomb_unrank[raw % 70u], sum = 0;
    for (int i = 0; i < 7; ++i) {
        c.co[i] = (uint8_t)(twist % 3u);
        sum += c.co[i];
        twist /= 3u;
    }
    c.co[7] = (uint8_t)((3u - sum % 3u) % 3u);
    int nu = 0, nd = 0;
    for (int i = 0; i < 8; ++i) c.cp[i] = (uint8_t)(((dmask >> i) & 1u) ? 4 + nd++ : nu++);
    return c;
}

void init_statics() {
    Statics* S = new Statics();
    g_statics = S;
    for (int f = 0; f < 6; ++f) {
        Cubie q = quarter_turn(f), p = q;
        for (int t = 0; t < 3; ++t) {
            S->moves[3 * f + t] = p;
            p = then(p, q);
        }
    }
    // Symmetry order of geometry.py:96-116: permutations of the axes in
    // lexicographic order, sign patterns in itertools.product((1,-1)) order;
    // keep those that fix the UD axis.
    const int P3[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    int nud = 0;
    for (auto& perm : P3)
        for (int sg = 0; sg < 8; ++sg) {
            int m[9] = {0};
            for (int row = 0; row < 3; ++row) m[3 * row + perm[row]] = ((sg >> (2 - row)) & 1) ? -1 : 1;
            if (m[4] == 0) continue;
            S->ud[nud++] = sym_from_matrix(m);
        }
    const int R[9] = {0, 0, 1, 1, 0, 0, 0, 1, 0};
    S->rot = sym_from_matrix(R);
    for (int m = 0; m < 18; ++m) {
        Cubie c = conjugate(S->rot, S->moves[m]);
        S->axis_map[m] = -1;
        for (int j = 0; j < 18; ++j)
            if (!memcmp(&c, &S->moves[j], sizeof c)) S->axis_map[m] = j;
    }
    for (int t = 0; t < 16; ++t)
        for (int s = 0; s < 16; ++s) {
            int pr[9];
            const int *a = S->ud[t].mat, *b = S->ud[s].mat;
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) pr[3 * r + c] = a[3 * r] * b[c] + a[3 * r + 1] * b[3 + c] + a[3 * r + 2] * b[6 + c];
            for (int u = 0; u < 16; ++u)
                if (!memcmp(S->ud[u].mat, pr, sizeof pr)) S->s_mul[t][s] = (uint8_t)u;
        }
    comb(8, S->ccomb_rank, S->ccomb_unrank);
    comb(12, S->slice_rank, S->slice_unrank);
    memset(S->eorder_rank, 0xFF, sizeof S->eorder_rank);
    int r = 0;
    for (int a = 0; a < 4; ++a)
        for (int b = 0; b < 4; ++b)
            for (int c = 0; c < 4; ++c)
                for (int d = 0; d < 4; ++d) {
                    if (a == b || a == c || a == d || b == c || b == d || c == d) continue;
                    S->perms4[r][0] = a, S->perms4[r][1] = b, S->perms4[r][2] = c, S->perms4[r][3] = d;
                    S->eorder_unrank[r] = (uint8_t)(a | b << 2 | c << 4 | d << 6);
                    S->eorder_rank[S->eorder_unrank[r]] = (uint8_t)r;
                    ++r;
                }
    // Corner classes: orbit minimum of the raw coordinate over the 16 UD
    // symmetries; the class symmetry is the first one reaching it.
    const uint32_t N_RAW = 153090;
    CornerClasses& cc = S->cc;
    cc.classidx.assign(N_RAW, 0);
    std::vector<uint32_t> omin(N_RAW);
    std::vector<uint8_t> osym(N_RAW);
    for (uint32_t raw = 0; raw < N_RAW; ++raw) {
        Cubie c = corners_from_raw(raw);
        uint32_t best = 0xFFFFFFFFu;
        int bs = 0;
        for (int s = 0; s < 16; ++s) {
            uint32_t v = corner_raw(conjugate(S->ud[s], c));
            if (v < best) best = v, bs = s;
        }
        omin[raw] = best;
        osym[raw] = (uint8_t)bs;
    }
    std::vector<int32_t> cls_of(N_RAW, -1);
    for (uint32_t raw = 0; raw < N_RAW; ++raw)
        if (omin[raw] == raw) {
            cls_of[raw] = (int32_t)cc.rep.size();
            cc.rep.push_back(raw);
        }
    for (uint32_t raw = 0; raw < N_RAW; ++raw) cc.classidx[raw] = ((uint32_t)cls_of[omin[raw]] << 4) | osym[raw];
    cc.stab.resize(cc.rep.size());
    for (size_t k = 0; k < cc.rep.size(); ++k) {
        Cubie c = corners_from_raw(cc.rep[k]);
        uint16_t st = 0;
        for (int s = 0; s < 16; ++s)
            if (corner_raw(conjugate(S->ud[s], c)) == cc.rep[k]) st |= (uint16_t)(1u << s);
        cc.stab[k] = st;
    }
    // The class tables (lut.hpp, ClassLUT): twist and placement transform
    // separately under each symmetry; the entries are read off classidx.
    ClassLUT& C = cc.class_lut;
    memset(&C, 0, sizeof C);
    std::vector<uint32_t> tw(16 * N_TWIST);
    for (int s = 0; s < 16; ++s) {
        for (uint32_t t = 0; t < N_TWIST; ++t) tw[s * N_TWIST + t] = corner_raw(conjugate(S->ud[s], corners_from_raw(t * N_DPOS))) / N_DPOS;
        for (uint32_t p = 0; p < N_DPOS; ++p)
            C.ps[s][p] = (uint8_t)(corner_raw(conjugate(S->ud[s], corners_from_raw(p))) % N_DPOS);
    }
    std::vector<uint32_t> orbit_row(N_TWIST, 0);
    uint32_t n_mt = 0, n_mo = 0;
    for (uint32_t t = 0; t < N_TWIST; ++t) {
        uint32_t tmin = 0xFFFFFFFFu, s0 = 0, nm = 0;
        for (int s = 0; s < 16; ++s) tmin = std::min(tmin, tw[s * N_TWIST + t]);
        for (int s = 15; s >= 0; --s)
            if (tw[s * N_TWIST + t] == tmin) s0 = (uint32_t)s, ++nm;
        uint32_t base = 0xFFFFFFFFu;
        for (uint32_t p = 0; p < N_DPOS; ++p) base = std::min(base, cc.classidx[tmin * N_DPOS + p] >> 4);
        uint32_t mi = 0, mo = 0;
        if (nm > 1) {
            if (n_mt < N_MULTI_TWIST) {
                mi = ++n_mt;
                for (uint32_t p = 0; p < N_DPOS; ++p) C.ms[mi - 1][p] = (uint8_t)(cc.classidx[t * N_DPOS + p] & 15u);
            } else {
                ++n_mt;  // more than the table holds: the check below fails
            }
            if (!orbit_row[tmin]) orbit_row[tmin] = ++n_mo;
            mo = orbit_row[tmin] <= N_MULTI_ORBIT ? orbit_row[tmin] : 0;
            if (mo)
                for (uint32_t p = 0; p < N_DPOS; ++p)
                    C.rk[mo - 1][C.ps[s0][p]] = (uint8_t)((cc.classidx[t * N_DPOS + p] >> 4) - base);
        }
        C.tw[t][0] = s0 | (mi << 4);
        C.tw[t][1] = base | (mo << 16);
    }
    cc.class_lut_rows[0] = (int)n_mt;
    cc.class_lut_rows[1] = (int)n_mo;
}

Statics& st() {
    std::call_once(g_once, init_statics);
    return *g_statics;
}

}  // namespace

Cubie solved() {
    Cubie c;
    for (int i = 0; i < 8; ++i) c.cp[i] = (uint8_t)i, c.co[i] = 0;
    for (int i = 0; i < 12; ++i) c.ep[i] = (uint8_t)i, c.eo[i] = 0;
    return c;
}

Cubie then(const Cubie& a, const Cubie& b) {
    Cubie r;
    for (int i = 0; i < 8; ++i) {
        r.cp[i] = a.cp[b.cp[i]];
        r.co[i] = (uint8_t)((a.co[b.cp[i]] + b.co[i]) % 3);
    }
    for (int i = 0; i < 12; ++i) {
        r.ep[i] = a.ep[b.ep[i]];
        r.eo[i] = (uint8_t)(a.eo[b.ep[i]] ^ b.eo[i]);
    }
    return r;
}

Cubie inverse(const Cubie& a) {
    Cubie r;
    for (int i = 0; i < 8; ++i) r.cp[a.cp[i]] = (uint8_t)i;
    for (int i = 0; i < 12; ++i) r.ep[a.ep[i]] = (uint8_t)i;
    for (int i = 0; i < 8; ++i) r.co[i] = (uint8_t)((3 - a.co[r.cp[i]]) % 3);
    for (int i = 0; i < 12; ++i) r.eo[i] = a.eo[r.ep[i]];
    return r;
}

static int parity(const uint8_t* p, int n) {
    int inv = 0;
    for (int i = 0; i < n; ++i)
        for (int j = i + 1; j < n; ++j) inv += p[i] > p[j];
    return inv & 1;
}

bool is_legal(const Cubie& a) {
    int seen_c = 0, seen_e = 0, tw = 0, fl = 0;
    for (int i = 0; i < 8; ++i) {
        if (a.cp[i] > 7 || a.co[i] > 2) return false;
        seen_c |= 1 << a.cp[i];
        tw += a.co[i];
    }
    for (int i = 0; i < 12; ++i) {
        if (a.ep[i] > 11 || a.eo[i] > 1) return false;
        seen_e |= 1 << a.ep[i];
        fl += a.eo[i];
    }
    return seen_c == 0xFF && seen_e == 0xFFF && tw % 3 == 0 && fl % 2 == 0 && parity(a.cp, 8) == parity(a.ep, 12);
}

const Cubie& move(int m) { return st().moves[m]; }

Packed pack(const Cubie& c) {
    Packed p{};
    for (int i = 0; i < 8; ++i) p.w[i >> 2] |= (uint32_t)(c.cp[i] | c.co[i] << 3) << (8 * (i & 3));
    for (int i = 0; i < 12; ++i) p.w[2 + (i >> 2)] |= (uint32_t)(c.ep[i] | c.eo[i] << 4) << (8 * (i & 3));
    return p;
}

Cubie unpack(const Packed& p) {
    Cubie c;
    for (int i = 0; i < 8; ++i) {
        uint32_t b = (p.w[i >> 2] >> (8 * (i & 3))) & 0xFFu;
        c.cp[i] = (uint8_t)(b & 7u), c.co[i] = (uint8_t)(b >> 3);
    }
    for (int i = 0; i < 12; ++i) {
        uint32_t b = (p.w[2 + (i >> 2)] >> (8 * (i & 3))) & 0xFFu;
        c.ep[i] = (uint8_t)(b & 15u), c.eo[i] = (uint8_t)(b >> 4);
    }
    return c;
}

SymAction sym_from_matrix(const int m[9]) {
    Action a = action_of(m);
    SymAction s;
    memcpy(s.mat, m, sizeof s.mat);
    memcpy(s.cpos, a.cpos, sizeof s.cpos);
    memcpy(s.csig, a.csig, sizeof s.csig);
    memcpy(s.epos, a.epos, sizeof s.epos);
    memcpy(s.esig, a.esig, sizeof s.esig);
    return s;
}

// M g M^-1 at cubie level: the cubie c at position i moves to position M(i)
// and becomes cubie M(c); its orientation follows the sticker maps of both.
Cubie conjugate(const SymAction& s, const Cubie& c) {
    Cubie r;
    for (int i = 0; i < 8; ++i) {
        int cub = c.cp[i], q = s.cpos[i];
        r.cp[q] = s.cpos[cub];
        r.co[q] = (uint8_t)((s.csig[i][c.co[i] % 3] - s.csig[cub][0] + 3) % 3);
    }
    for (int i = 0; i < 12; ++i) {
        int cub = c.ep[i], q = s.epos[i];
        r.ep[q] = s.epos[cub];
        r.eo[q] = (uint8_t)(s.esig[i][c.eo[i] & 1] ^ s.esig[cub][0]);
    }
    return r;
}

const std::array<SymAction, 16>& ud_syms() { return st().ud; }
const SymAction& axis_rotation() { return st().rot; }
int axis_move(int m) { return st().axis_map[m]; }

void axis_frames(const Cubie& c, Cubie out[6]) {
    out[0] = c;
    out[1] = conjugate(st().rot, out[0]);
    out[2] = conjugate(st().rot, out[1]);
    for (int j = 0; j < 3; ++j) out[3 + j] = inverse(out[j]);
}

const CornerClasses& corner_classes() { return st().cc; }

uint32_t corner_raw(const Cubie& c) {
    Statics& S = *g_statics;  // valid once statics exist (called during their construction)
    uint32_t twist = 0, w = 1, dmask = 0;
    for (int i = 0; i < 7; ++i) twist += c.co[i] * w, w *= 3;
    for (int i = 0; i < 8; ++i) dmask |= (uint32_t)(c.cp[i] >> 2 & 1) << i;
    return twist * 70u + S.ccomb_rank[dmask];
}

// ---- table kinds -------------------------------------------------------

namespace {
struct B4 {
    int p[4], f[4];  // relabelling: label x -> p[x], its flip ^= f[x]
};
B4 b4_after(const B4& a, const B4& b) {  // a after b (kinds.py:28-32)
    B4 r;
    for (int x = 0; x < 4; ++x) r.p[x] = a.p[b.p[x]], r.f[x] = b.f[x] ^ a.f[b.p[x]];
    return r;
}
B4 b4_inv(const B4& a) {
    B4 r;
    for (int z = 0; z < 4; ++z) r.p[a.p[z]] = z;
    for (int z = 0; z < 4; ++z) r.f[z] = a.f[r.p[z]];
    return r;
}
int b4_code(const B4& a) {
    return ((a.p[0] * 4 + a.p[1]) * 4 + a.p[2]) * 4 + a.p[3] + 256 * (a.f[0] | a.f[1] << 1 | a.f[2] << 2 | a.f[3] << 3);
}
}  // namespace

bool make_group_kind(const std::string& name, const std::vector<std::array<int, 8>>& gens, Kind& out,
                     std::string& err) {
    Statics& S = st();
    std::vector<B4> group{B4{{0, 1, 2, 3}, {0, 0, 0, 0}}};
    std::vector<char> in(256 * 16, 0);
    in[b4_code(group[0])] = 1;
    for (size_t i = 0; i < group.size(); ++i)
        for (auto& g : gens) {
            B4 ge{{g[0], g[1], g[2], g[3]}, {g[4], g[5], g[6], g[7]}};
            B4 y = b4_after(ge, group[i]);
            if (!in[b4_code(y)]) in[b4_code(y)] = 1, group.push_back(y);
        }
    // The relabellings of the E edges induced by the 16 UD symmetries must
    // normalise the group, or symmetry reduction would be wrong.
    for (auto& s : S.ud) {
        B4 e;
        for (int x = 0; x < 4; ++x) e.p[x] = s.epos[8 + x] - 8, e.f[x] = s.esig[8 + x][0];
        for (auto& x : group)
            if (!in[b4_code(b4_after(b4_after(e, x), b4_inv(e)))]) {
                err = name + ": group not normalised by the UD symmetries; symmetry reduction would be wrong";
                return false;
            }
    }
    // E contents c = perm rank * 16 + flips by position; classes numbered in
    // order of their smallest content.
    int cls[384], nrep = 0, reps[384];
    for (int c = 0; c < 384; ++c) cls[c] = -1;
    for (int c = 0; c < 384; ++c) {
        if (cls[c] >= 0) continue;
        const int* q = S.perms4[c / 16];
        for (auto& k : group) {
            int key = 0, flips = 0;
            for (int i = 0; i < 4; ++i) {
                key |= k.p[q[i]] << (2 * i);
                flips |= (((c >> i) & 1) ^ k.f[q[i]]) << i;
            }
            cls[S.eorder_rank[key] * 16 + flips] = nrep;
        }
        reps[nrep++] = c;
    }
    bool even = true;
    for (auto& k : group)
        if ((k.f[0] + k.f[1] + k.f[2] + k.f[3]) & 1) even = false;
    Kind k;
    k.name = name;
    k.is304 = false;
    if (even) {
        // flip parity of the E edges follows from the non-E flips: number
        // the classes within each parity
        int count[2] = {0, 0}, index[384];
        for (int r = 0; r < nrep; ++r) {
            int p = __builtin_popcount(reps[r] & 15) & 1;
            index[r] = count[p];
            if (count[p] < 48) k.gen_rep[p][count[p]] = (uint16_t)reps[r];
            ++count[p];
        }
        k.nec = nrep / 2;
        for (int c = 0; c < 384; ++c) k.gen_class[c] = (uint8_t)index[cls[c]];
    } else {
        k.nec = nrep;
        for (int c = 0; c < 384; ++c) {
            k.gen_class[c] = (uint8_t)cls[c];
            if (cls[c] < 48) k.gen_rep[__builtin_popcount(c & 15) & 1][cls[c]] = (uint16_t)c;
        }
    }
    if (k.nec > 48) {
        err = name + ": more than 48 E-content classes are not supported";
        return false;
    }
    k.n_edge = 495ull * 256ull * (uint64_t)k.nec;
    k.n_entries = (uint64_t)S.cc.rep.size() * k.n_edge;
    out = k;
    return true;
}

bool make_kind(const std::string& name, Kind& out, std::string& err) {
    Statics& S = st();
    if (name == "304") {
        out = Kind();
        out.name = name;
        out.is304 = true;
        out.n_edge = 495ull * 70ull * 16ull;
        out.n_entries = (uint64_t)S.cc.rep.size() * out.n_edge;
        return true;
    }
    // kinds.py:100-107
    static const struct {
        const char* name;
        std::vector<std::array<int, 8>> gens;
    } G[] = {
        {"s4", {{0, 1, 3, 2, 0, 0, 0, 0}, {0, 2, 1, 3, 0, 0, 0, 0}, {1, 0, 2, 3, 0, 0, 0, 0}}},
        {"d4f", {{0, 1, 2, 3, 1, 1, 1, 1}, {0, 3, 2, 1, 0, 0, 0, 0}, {1, 0, 3, 2, 0, 0, 0, 0}}},
        {"a4t", {{0, 2, 3, 1, 0, 1, 1, 0}, {1, 0, 3, 2, 1, 1, 1, 1}}},
        {"d4t", {{0, 3, 2, 1, 0, 0, 0, 0}, {1, 0, 3, 2, 1, 1, 1, 1}}},
        {"d4", {{1, 2, 3, 0, 0, 0, 0, 0}, {0, 3, 2, 1, 0, 0, 0, 0}}},
        {"a4", {{0, 2, 3, 1, 0, 0, 0, 0}, {1, 0, 3, 2, 0, 0, 0, 0}}},
    };
    for (auto& g : G)
        if (name == g.name) return make_group_kind(name, g.gens, out, err);
    err = "unknown table kind '" + name + "'";
    return false;
}

uint64_t table_index(const Kind& k, const Cubie& c) {
    Statics& S = st();
    uint32_t packed = S.cc.classidx[corner_raw(c)];
    Cubie e = conjugate(S.ud[packed & 15u], c);
    uint32_t emask = 0, eorder = 0, umask = 0, eflip = 0, nflip = 0;
    int ne = 0, no = 0;
    for (int i = 0; i < 12; ++i) {
        if (e.ep[i] >= 8) {
            emask |= 1u << i;
            eorder |= (uint32_t)(e.ep[i] & 3) << (2 * ne);
            eflip |= (uint32_t)e.eo[i] << ne;
            ++ne;
        } else {
            umask |= (uint32_t)(e.ep[i] < 4) << no;
            nflip |= (uint32_t)e.eo[i] << no;
            ++no;
        }
    }
    uint64_t cls = packed >> 4, ec;
    if (k.is304)
        ec = ((uint64_t)S.slice_rank[emask] * 70u + S.ccomb_rank[umask]) * 16u + eflip;
    else
        ec = ((uint64_t)S.slice_rank[emask] * 256u + nflip) * (uint64_t)k.nec +
             k.gen_class[S.eorder_rank[eorder] * 16u + eflip];
    return cls * k.n_edge + ec;
}

bool allowed_after(int last, int m) {
    if (last < 0) return true;
    int lf = last / 3, mf = m / 3;
    return mf != lf && !(lf >= 3 && mf == lf - 3);
}

std::vector<uint8_t> canonical_prefixes(int length) {
    std::vector<uint8_t> out;
    if (length <= 0) return out;
    int seq[16];
    int d = 0;
    seq[0] = -1;
    for (;;) {
        int last = d ? seq[d - 1] : -1, m = seq[d] + 1;
        while (m < 18 && !allowed_after(last, m)) ++m;
        if (m >= 18) {
            if (d == 0) break;
            --d;
            continue;
        }
        seq[d] = m;
        if (d + 1 < length) {
            seq[++d] = -1;
            continue;
        }
        for (int j = 0; j < length; ++j) out.push_back((uint8_t)seq[j]);
    }
    return out;
}

// M_a M_b g M_b^-1 M_a^-1 as one action: conjugation by a after conjugation by b.
SymAction compose_sym(const SymAction& a, const SymAction& b) {
    SymAction t{};
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            t.mat[3 * r + c] = a.mat[3 * r] * b.mat[c] + a.mat[3 * r + 1] * b.mat[3 + c] + a.mat[3 * r + 2] * b.mat[6 + c];
    for (int i = 0; i < 8; ++i) {
        t.cpos[i] = a.cpos[b.cpos[i]];
        for (int o = 0; o < 3; ++o) t.csig[i][o] = a.csig[b.cpos[i]][b.csig[i][o]];
    }
    for (int i = 0; i < 12; ++i) {
        t.epos[i] = a.epos[b.epos[i]];
        for (int o = 0; o < 2; ++o) t.esig[i][o] = a.esig[b.epos[i]][b.esig[i][o]];
    }
    return t;
}

void build_lut(const Kind& k, CubeLUT& L) {
    Statics& S = st();
    memset(&L, 0, sizeof L);
    auto put = [](uint32_t& w, int j, uint32_t b) { w |= (b & 0xFFu) << (8 * j); };
    auto right_rec = [&](RightRec& r, int m) {
        const Cubie& mv = S.moves[m];
        for (int w = 0; w < 2; ++w)
            for (int j = 0; j < 4; ++j) {
                put(r.csel[w], j, mv.cp[4 * w + j]);
                put(r.cadd[w], j, (uint32_t)mv.co[4 * w + j] << 3);
            }
        for (int w = 0; w < 3; ++w)
            for (int j = 0; j < 4; ++j) {
                uint32_t src = mv.ep[4 * w + j];
                put(r.elo[w], j, src < 8 ? src : SEL_ZERO);
                put(r.ehi[w], j, src >= 8 ? src - 8 : SEL_ZERO);
                put(r.eflip[w], j, (uint32_t)mv.eo[4 * w + j] << 4);
            }
    };
    auto left_rec = [&](LeftRec& r, int m) {
        const Cubie& mv = S.moves[m];
        for (int w = 0; w < 2; ++w)
            for (int j = 0; j < 4; ++j) put(r.lc[w], j, mv.cp[4 * w + j] | mv.co[4 * w + j] << 3);
        for (int w = 0; w < 3; ++w)
            for (int j = 0; j < 4; ++j) put(r.le[w], j, mv.ep[4 * w + j] | mv.eo[4 * w + j] << 4);
    };
    auto inverse = [](int m) { return 3 * (m / 3) + 2 - m % 3; };
    for (int m = 0; m < 18; ++m) {
        right_rec(L.rr[0][m], m);
        right_rec(L.rr[1][m], inverse(m));
        left_rec(L.lr[0][m], inverse(m));
        left_rec(L.lr[1][m], m);
    }
    // R^k, k = 0..2, and the corner part of conjugation by it (lut.hpp ConjRec)
    static const int ID[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1};
    const SymAction rk[3] = {sym_from_matrix(ID), S.rot, compose_sym(S.rot, S.rot)};
    for (int kk = 0; kk < 3; ++kk) {
        const SymAction& y = rk[kk];
        ConjRec& r = L.cj[kk];
        for (int i = 0; i < 8; ++i) {
            const int q = y.cpos[i];
            put(r.csel[q >> 2], q & 3, (uint32_t)i);
            put(r.cadd[q >> 2], q & 3, (uint32_t)y.csig[i][0] << 3);
            put(r.lc[i >> 2], i & 3, (uint32_t)y.cpos[i] | (uint32_t)((3 - y.csig[i][0]) % 3) << 3);
        }
    }
    for (int t = 0; t < N_ESYM; ++t) {
        const SymAction y = compose_sym(S.ud[t % 16], rk[t / 16]);
        const int s = t;
        uint8_t pinv[12];
        for (int i = 0; i < 12; ++i) pinv[y.epos[i]] = (uint8_t)i;
        for (int w = 0; w < 3; ++w)
            for (int j = 0; j < 4; ++j) {
                int n = 4 * w + j, src = pinv[n];
                put(L.cs_lo[s][w], j, src < 8 ? (uint32_t)src : SEL_ZERO);
                put(L.cs_hi[s][w], j, src >= 8 ? (uint32_t)src - 8 : SEL_ZERO);
                put(L.cs_flip[s][w], j, (uint32_t)y.esig[src][0] << 4);
                put(L.cs_rel[s][w], j, y.epos[n] | (uint32_t)y.esig[n][0] << 4);
            }
    }
    for (int m = 0; m < 16; ++m) {
        uint32_t sel = 0x0C0C0C0Cu;
        int o = 0;
        for (int b = 0; b < 4; ++b)
            if (m >> b & 1) {
                sel &= ~(0xFFu << (8 * o));
                sel |= (uint32_t)b << (8 * o);
                ++o;
            }
        L.csel[m] = sel;
    }
    // 12C4 rank split at bit 6 (see lut.hpp).
    for (int lo = 0; lo < 64; ++lo) {
        int a = __builtin_popcount(lo);
        if (a > 4) continue;
        uint32_t low_hi = 0;  // the lowest (4 - a) positions of 6..11
        for (int i = 0; i < 4 - a; ++i) low_hi |= 1u << (6 + i);
        L.slice_lo[lo] = S.slice_rank[lo | low_hi];
        for (int hi = 0; hi < 64; ++hi)
            if (__builtin_popcount(hi) == 4 - a)
                L.slice_hi[a][hi] = (uint8_t)(S.slice_rank[lo | hi << 6] - L.slice_lo[lo]);
    }
    for (int i = 0; i < 256; ++i) {
        L.ccomb_rank[i] = S.ccomb_rank[i] == 0xFFFF ? 0xFF : (uint8_t)S.ccomb_rank[i];
        L.eorder_rank[i] = S.eorder_rank[i];
    }
    for (int i = 0; i < 70; ++i) L.ccomb_unrank[i] = (uint8_t)S.ccomb_unrank[i];
    for (int i = 0; i < 495; ++i) L.slice_unrank[i] = S.slice_unrank[i];
    for (int i = 0; i < 24; ++i) L.eorder_unrank[i] = S.eorder_unrank[i];
    memcpy(L.gen_class, k.gen_class, sizeof L.gen_class);
    for (int key = 0; key < 256; ++key)
        if (S.eorder_rank[key] != 0xFF)
            for (int fl = 0; fl < 16; ++fl) L.key_class[key * 16 + fl] = k.gen_class[S.eorder_rank[key] * 16 + fl];
    memcpy(L.gen_rep, k.gen_rep, sizeof L.gen_rep);
    memcpy(L.s_mul, S.s_mul, sizeof L.s_mul);
    for (int b = 0; b < 243; ++b)
        for (int k = 0, v = b; k < 5; ++k, v /= 3) L.b3digits[b] |= (uint16_t)((v % 3) << (2 * k));
    for (int last = -1; last < 18; ++last) {
        uint32_t mask = 0;
        for (int m = 0; m < 18; ++m)
            if (allowed_after(last, m)) mask |= 1u << m;
        L.amask[last + 1] = mask;
    }
}

}  // namespace ccs
