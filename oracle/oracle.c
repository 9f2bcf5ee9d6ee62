/* CPU oracle -- TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates, in plain C, the reference's host cube model and geometry
 * (cubie.py, geometry.py), index numbering (coords.py, kinds.py), the
 * mod-3 + block-table BFS and the b3 encoder (kernels.cu, table.py), and the
 * step-mode IDA* pruning (kernels.cu search_step, solver.py GpuSolver.solve).
 * The product never links this file.
 */
#include "oracle.h"

#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ---- cube model: cubie.py:57-124 ------------------------------------- */

static const uint8_t QUARTER[6][40] = {
    /* U */ {3, 0, 1, 2, 4, 5, 6, 7, 0, 0, 0, 0, 0, 0, 0, 0, 3, 0, 1, 2, 4, 5, 6, 7, 8, 9, 10, 11, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    /* R */ {4, 1, 2, 0, 7, 5, 6, 3, 2, 0, 0, 1, 1, 0, 0, 2, 8, 1, 2, 3, 11, 5, 6, 7, 4, 9, 10, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    /* F */ {1, 5, 2, 3, 0, 4, 6, 7, 1, 2, 0, 0, 2, 1, 0, 0, 0, 9, 2, 3, 4, 8, 6, 7, 1, 5, 10, 11, 0, 1, 0, 0, 0, 1, 0, 0, 1, 1, 0, 0},
    /* D */ {0, 1, 2, 3, 5, 6, 7, 4, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 2, 3, 5, 6, 7, 4, 8, 9, 10, 11, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    /* L */ {0, 2, 6, 3, 4, 1, 5, 7, 0, 1, 2, 0, 0, 2, 1, 0, 0, 1, 10, 3, 4, 5, 9, 7, 8, 2, 6, 11, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0},
    /* B */ {0, 1, 3, 7, 4, 5, 2, 6, 0, 0, 1, 2, 0, 0, 2, 1, 0, 1, 2, 11, 4, 5, 6, 10, 8, 9, 3, 7, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 1, 1},
};

static ocube MOVES[18];
static const ocube SOLVED = {{0, 1, 2, 3, 4, 5, 6, 7}, {0}, {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11}, {0}};

void or_compose(const ocube* a, const ocube* b, ocube* out) {
    ocube r;
    for (int i = 0; i < 8; ++i) {
        r.cp[i] = a->cp[b->cp[i]];
        r.co[i] = (uint8_t)((a->co[b->cp[i]] + b->co[i]) % 3);
    }
    for (int i = 0; i < 12; ++i) {
        r.ep[i] = a->ep[b->ep[i]];
        r.eo[i] = (uint8_t)((a->eo[b->ep[i]] + b->eo[i]) & 1);
    }
    *out = r;
}

void or_inverse(const ocube* a, ocube* out) {
    ocube r;
    uint8_t cinv[8], einv[12];
    for (int i = 0; i < 8; ++i) cinv[a->cp[i]] = (uint8_t)i;
    for (int i = 0; i < 12; ++i) einv[a->ep[i]] = (uint8_t)i;
    for (int i = 0; i < 8; ++i) {
        r.cp[i] = cinv[i];
        r.co[i] = (uint8_t)((3 - a->co[cinv[i]]) % 3);
    }
    for (int i = 0; i < 12; ++i) {
        r.ep[i] = einv[i];
        r.eo[i] = a->eo[einv[i]];
    }
    *out = r;
}

void or_move(int m, ocube* out) { *out = MOVES[m]; }

static int is_solved(const ocube* c) { return memcmp(c, &SOLVED, sizeof(ocube)) == 0; }

/* ---- sticker geometry and the 48 symmetries: geometry.py:18-124 -------- */

static const char* CORNER_NAMES[8] = {"URF", "UFL", "ULB", "UBR", "DFR", "DLF", "DBL", "DRB"};
static const char* EDGE_NAMES[12] = {"UR", "UF", "UL", "UB", "DR", "DF", "DL", "DB", "FR", "FL", "BL", "BR"};

static void face_normal(char f, int v[3]) {
    v[0] = v[1] = v[2] = 0;
    switch (f) {
        case 'U': v[1] = 1; break;
        case 'R': v[0] = 1; break;
        case 'F': v[2] = 1; break;
        case 'D': v[1] = -1; break;
        case 'L': v[0] = -1; break;
        case 'B': v[2] = -1; break;
    }
}

typedef struct { int c[3], n[3]; } slot_t;
static slot_t CSLOT[24], ESLOT[24];

typedef struct {
    int m[9];
    int det;
    uint8_t corner[24], edge[24], corner_inv[24], edge_inv[24];
} osym;

static osym SYMS[48];
static int UD[16];
static osym ROT;

static void mat_apply(const int* m, const int* v, int* out) {
    for (int r = 0; r < 3; ++r) out[r] = m[3 * r] * v[0] + m[3 * r + 1] * v[1] + m[3 * r + 2] * v[2];
}

static int find_slot(const slot_t* slots, const int c[3], const int n[3]) {
    for (int i = 0; i < 24; ++i)
        if (!memcmp(slots[i].c, c, sizeof(int) * 3) && !memcmp(slots[i].n, n, sizeof(int) * 3)) return i;
    return -1;
}

static void make_sym(osym* s, const int* m) {
    memcpy(s->m, m, sizeof(int) * 9);
    s->det = m[0] * (m[4] * m[8] - m[5] * m[7]) - m[1] * (m[3] * m[8] - m[5] * m[6]) + m[2] * (m[3] * m[7] - m[4] * m[6]);
    for (int j = 0; j < 24; ++j) {
        int c[3], n[3];
        mat_apply(m, CSLOT[j].c, c);
        mat_apply(m, CSLOT[j].n, n);
        s->corner[j] = (uint8_t)find_slot(CSLOT, c, n);
        mat_apply(m, ESLOT[j].c, c);
        mat_apply(m, ESLOT[j].n, n);
        s->edge[j] = (uint8_t)find_slot(ESLOT, c, n);
    }
    for (int j = 0; j < 24; ++j) {
        s->corner_inv[s->corner[j]] = (uint8_t)j;
        s->edge_inv[s->edge[j]] = (uint8_t)j;
    }
}

/* geometry.py:84-93 (conjugate), :36-63 (cubie <-> slots). */
static void sym_conj(const osym* s, const ocube* in, ocube* out) {
    uint8_t cs[24], es[24], nc[24], ne[24];
    for (int i = 0; i < 8; ++i)
        for (int k = 0; k < 3; ++k) cs[3 * i + (k + in->co[i]) % 3] = (uint8_t)(3 * in->cp[i] + k);
    for (int i = 0; i < 12; ++i)
        for (int k = 0; k < 2; ++k) es[2 * i + (k + in->eo[i]) % 2] = (uint8_t)(2 * in->ep[i] + k);
    for (int j = 0; j < 24; ++j) {
        nc[j] = s->corner[cs[s->corner_inv[j]]];
        ne[j] = s->edge[es[s->edge_inv[j]]];
    }
    for (int i = 0; i < 8; ++i) {
        out->cp[i] = nc[3 * i] / 3;
        out->co[i] = (uint8_t)((3 - nc[3 * i] % 3) % 3);
    }
    for (int i = 0; i < 12; ++i) {
        out->ep[i] = ne[2 * i] / 2;
        out->eo[i] = ne[2 * i] % 2;
    }
}

void or_sym_matrix(int k, int32_t out[9]) {
    for (int i = 0; i < 9; ++i) out[i] = SYMS[k].m[i];
}
void or_sym_conjugate(int k, const ocube* in, ocube* out) { sym_conj(&SYMS[k], in, out); }
int or_ud_sym(int k) { return UD[k]; }

/* solver.py:58-64: the three axis frames and their inverses. */
void or_axis_frames(const ocube* in, ocube out[6]) {
    out[0] = *in;
    sym_conj(&ROT, &out[0], &out[1]);
    sym_conj(&ROT, &out[1], &out[2]);
    for (int j = 0; j < 3; ++j) or_inverse(&out[j], &out[3 + j]);
}

static int AXIS_MAP[18];
int or_axis_move_map(int m) { return AXIS_MAP[m]; }

/* ---- coordinates: coords.py:82-118, 157-168, 227-340 ------------------- */

static uint16_t CCOMB_RANK[256], CCOMB_UNRANK[70], SLICE_RANK[4096], SLICE_UNRANK[495];
static uint8_t EORDER_RANK[256], EORDER_UNRANK[24];
static int PERMS4[24][4];

static void comb_tables(int bits, uint16_t* rank, uint16_t* unrank) {
    /* itertools.combinations(range(bits), 4) in lexicographic order. */
    int r = 0;
    for (int i = 0; i < (1 << bits); ++i) rank[i] = 0xFFFF;
    for (int a = 0; a < bits; ++a)
        for (int b = a + 1; b < bits; ++b)
            for (int c = b + 1; c < bits; ++c)
                for (int d = c + 1; d < bits; ++d) {
                    int mask = (1 << a) | (1 << b) | (1 << c) | (1 << d);
                    rank[mask] = (uint16_t)r;
                    unrank[r++] = (uint16_t)mask;
                }
}

uint32_t or_corner_raw(const ocube* c) {
    uint32_t twist = 0, w = 1, dmask = 0;
    for (int i = 0; i < 7; ++i) {
        twist += c->co[i] * w;
        w *= 3;
    }
    for (int i = 0; i < 8; ++i) dmask |= (uint32_t)((c->cp[i] >> 2) & 1) << i;
    return twist * 70u + CCOMB_RANK[dmask];
}

static void corners_from_raw(uint32_t raw, ocube* c) {
    uint32_t twist = raw / 70u, ccomb = raw % 70u, dmask = CCOMB_UNRANK[ccomb], sum = 0;
    for (int i = 0; i < 7; ++i) {
        c->co[i] = (uint8_t)(twist % 3u);
        sum += c->co[i];
        twist /= 3u;
    }
    c->co[7] = (uint8_t)((3 - sum % 3) % 3);
    int nu = 0, nd = 0;
    for (int i = 0; i < 8; ++i) c->cp[i] = (uint8_t)(((dmask >> i) & 1) ? 4 + nd++ : nu++);
}

#define N_RAW 153090
static uint32_t* CLASSIDX;
static uint32_t CLASS_REP[9930];
static uint16_t CLASS_STAB[9930];
static int N_CLASSES;

/* Cubie-level edge action of a UD symmetry, taken from its sticker slots
 * (coords.py:227-231, 243-250). */
static uint8_t EPERM[16][12], EPERM_INV[16][12], EFLIP[16][12];

static void edge_conj_fast(int s, const ocube* in, uint8_t ep[12], uint8_t eo[12]) {
    for (int i = 0; i < 12; ++i) {
        int src = EPERM_INV[s][i];
        int cub = in->ep[src];
        ep[i] = EPERM[s][cub];
        eo[i] = (uint8_t)(in->eo[src] ^ EFLIP[s][src] ^ EFLIP[s][cub]);
    }
}

static int g_init = 0;

static void corner_classes(void) {
    /* coords.py:327-340: orbit minimum over the 16 UD symmetries; sym = the
     * first symmetry (in ud_symmetries order) reaching it. */
    CLASSIDX = (uint32_t*)malloc(sizeof(uint32_t) * N_RAW);
    uint32_t* omin = (uint32_t*)malloc(sizeof(uint32_t) * N_RAW);
    uint8_t* osym_ = (uint8_t*)malloc(N_RAW);
    int32_t* cls_of = (int32_t*)malloc(sizeof(int32_t) * N_RAW);
    ocube base = SOLVED, img;
    for (uint32_t raw = 0; raw < N_RAW; ++raw) {
        corners_from_raw(raw, &base);
        uint32_t best = 0xFFFFFFFFu;
        int bs = 0;
        for (int s = 0; s < 16; ++s) {
            sym_conj(&SYMS[UD[s]], &base, &img);
            uint32_t r = or_corner_raw(&img);
            if (r < best) {
                best = r;
                bs = s;
            }
        }
        omin[raw] = best;
        osym_[raw] = (uint8_t)bs;
    }
    for (uint32_t r = 0; r < N_RAW; ++r) cls_of[r] = -1;
    int n = 0;
    for (uint32_t r = 0; r < N_RAW; ++r)
        if (omin[r] == r) {
            cls_of[r] = n;
            CLASS_REP[n++] = r;
        }
    N_CLASSES = n;
    for (uint32_t r = 0; r < N_RAW; ++r) CLASSIDX[r] = ((uint32_t)cls_of[omin[r]] << 4) | osym_[r];
    for (int c = 0; c < n; ++c) {
        corners_from_raw(CLASS_REP[c], &base);
        uint16_t st = 0;
        for (int s = 0; s < 16; ++s) {
            sym_conj(&SYMS[UD[s]], &base, &img);
            if (or_corner_raw(&img) == CLASS_REP[c]) st |= (uint16_t)(1u << s);
        }
        CLASS_STAB[c] = st;
    }
    free(omin);
    free(osym_);
    free(cls_of);
}

const uint32_t* or_classidx(void) { return CLASSIDX; }
const uint32_t* or_class_rep(void) { return CLASS_REP; }
const uint16_t* or_class_stab(void) { return CLASS_STAB; }
int or_n_classes(void) { return N_CLASSES; }

static uint8_t S_MUL[16][16];

int or_init(void) {
    if (g_init) return N_CLASSES;
    for (int f = 0; f < 6; ++f) {
        ocube q;
        memcpy(q.cp, QUARTER[f], 8);
        memcpy(q.co, QUARTER[f] + 8, 8);
        memcpy(q.ep, QUARTER[f] + 16, 12);
        memcpy(q.eo, QUARTER[f] + 28, 12);
        ocube p = q;
        for (int t = 0; t < 3; ++t) {
            MOVES[3 * f + t] = p;
            or_compose(&p, &q, &p);
        }
    }
    for (int i = 0; i < 8; ++i) {
        int c[3] = {0, 0, 0}, v[3];
        for (int k = 0; k < 3; ++k) {
            face_normal(CORNER_NAMES[i][k], v);
            for (int a = 0; a < 3; ++a) c[a] += v[a];
        }
        for (int k = 0; k < 3; ++k) {
            memcpy(CSLOT[3 * i + k].c, c, sizeof c);
            face_normal(CORNER_NAMES[i][k], CSLOT[3 * i + k].n);
        }
    }
    for (int i = 0; i < 12; ++i) {
        int c[3] = {0, 0, 0}, v[3];
        for (int k = 0; k < 2; ++k) {
            face_normal(EDGE_NAMES[i][k], v);
            for (int a = 0; a < 3; ++a) c[a] += v[a];
        }
        for (int k = 0; k < 2; ++k) {
            memcpy(ESLOT[2 * i + k].c, c, sizeof c);
            face_normal(EDGE_NAMES[i][k], ESLOT[2 * i + k].n);
        }
    }
    /* geometry.py:96-107: permutations(range(3)) x product((1,-1), repeat=3),
     * identity first (it already is). */
    static const int P3[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};
    int n = 0, nud = 0;
    for (int p = 0; p < 6; ++p)
        for (int sg = 0; sg < 8; ++sg) {
            int m[9] = {0};
            for (int row = 0; row < 3; ++row) m[3 * row + P3[p][row]] = ((sg >> (2 - row)) & 1) ? -1 : 1;
            make_sym(&SYMS[n], m);
            if (m[4] == 1 || m[4] == -1) UD[nud++] = n;
            ++n;
        }
    static const int RM[9] = {0, 0, 1, 1, 0, 0, 0, 1, 0};
    make_sym(&ROT, RM);
    for (int m = 0; m < 18; ++m) {
        ocube c;
        sym_conj(&ROT, &MOVES[m], &c);
        AXIS_MAP[m] = -1;
        for (int j = 0; j < 18; ++j)
            if (!memcmp(&c, &MOVES[j], sizeof c)) AXIS_MAP[m] = j;
    }
    for (int s = 0; s < 16; ++s) {
        const osym* y = &SYMS[UD[s]];
        for (int i = 0; i < 12; ++i) {
            EPERM[s][i] = y->edge[2 * i] / 2;
            EFLIP[s][i] = y->edge[2 * i] % 2;
        }
        for (int i = 0; i < 12; ++i) EPERM_INV[s][EPERM[s][i]] = (uint8_t)i;
    }
    for (int t = 0; t < 16; ++t)
        for (int s = 0; s < 16; ++s) {
            const int* a = SYMS[UD[t]].m;
            const int* b = SYMS[UD[s]].m;
            int pr[9];
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) pr[3 * r + c] = a[3 * r] * b[c] + a[3 * r + 1] * b[3 + c] + a[3 * r + 2] * b[6 + c];
            for (int u = 0; u < 16; ++u)
                if (!memcmp(SYMS[UD[u]].m, pr, sizeof pr)) S_MUL[t][s] = (uint8_t)u;
        }
    comb_tables(8, CCOMB_RANK, CCOMB_UNRANK);
    comb_tables(12, SLICE_RANK, SLICE_UNRANK);
    {
        int r = 0;
        for (int a = 0; a < 4; ++a)
            for (int b = 0; b < 4; ++b)
                for (int c = 0; c < 4; ++c)
                    for (int d = 0; d < 4; ++d) {
                        if (a == b || a == c || a == d || b == c || b == d || c == d) continue;
                        int key = a | (b << 2) | (c << 4) | (d << 6);
                        PERMS4[r][0] = a, PERMS4[r][1] = b, PERMS4[r][2] = c, PERMS4[r][3] = d;
                        EORDER_UNRANK[r] = (uint8_t)key;
                        ++r;
                    }
        memset(EORDER_RANK, 0xFF, sizeof EORDER_RANK);
        for (int i = 0; i < 24; ++i) EORDER_RANK[EORDER_UNRANK[i]] = (uint8_t)i;
    }
    corner_classes();
    g_init = 1;
    return N_CLASSES;
}

/* ---- table kinds: coords.py:37-76, kinds.py:23-107 --------------------- */

typedef struct { int p[4], f[4]; } b4el;

static b4el b4mul(b4el a, b4el b) { /* kinds.py:28-32: a after b */
    b4el r;
    for (int x = 0; x < 4; ++x) {
        r.p[x] = a.p[b.p[x]];
        r.f[x] = b.f[x] ^ a.f[b.p[x]];
    }
    return r;
}
static b4el b4inv(b4el a) {
    b4el r;
    for (int z = 0; z < 4; ++z) r.p[a.p[z]] = z;
    for (int z = 0; z < 4; ++z) r.f[z] = a.f[r.p[z]];
    return r;
}
static int b4id(b4el a) {
    int pr = -1;
    for (int i = 0; i < 24; ++i)
        if (!memcmp(PERMS4[i], a.p, sizeof a.p)) pr = i;
    return pr * 16 + a.f[0] + 2 * a.f[1] + 4 * a.f[2] + 8 * a.f[3];
}

/* gens: n_gens x 8 ints (perm[4], flips[4]). Returns 0, or -1 if not normalised. */
int or_group_kind(okind* k, const int* gens, int n_gens) {
    or_init();
    static b4el grp[384];
    int in[384];
    memset(in, 0, sizeof in);
    int n = 1;
    b4el id = {{0, 1, 2, 3}, {0, 0, 0, 0}};
    grp[0] = id;
    in[b4id(id)] = 1;
    for (int i = 0; i < n; ++i)
        for (int g = 0; g < n_gens; ++g) {
            b4el ge;
            memcpy(ge.p, gens + 8 * g, sizeof ge.p);
            memcpy(ge.f, gens + 8 * g + 4, sizeof ge.f);
            b4el y = b4mul(ge, grp[i]);
            int yi = b4id(y);
            if (!in[yi]) {
                in[yi] = 1;
                grp[n++] = y;
            }
        }
    /* normalised by the relabellings the 16 UD symmetries induce (kinds.py:41-45, 62-63) */
    for (int s = 0; s < 16; ++s) {
        b4el e;
        for (int x = 0; x < 4; ++x) {
            e.p[x] = EPERM[s][8 + x] - 8;
            e.f[x] = EFLIP[s][8 + x];
        }
        for (int i = 0; i < n; ++i)
            if (!in[b4id(b4mul(b4mul(e, grp[i]), b4inv(e)))]) return -1;
    }
    int cls[384], reps[384], nrep = 0;
    for (int c = 0; c < 384; ++c) cls[c] = -1;
    for (int c = 0; c < 384; ++c) {
        if (cls[c] >= 0) continue;
        const int* q = PERMS4[c / 16];
        for (int i = 0; i < n; ++i) {
            int q2[4], g2 = 0;
            for (int j = 0; j < 4; ++j) {
                q2[j] = grp[i].p[q[j]];
                g2 |= (((c >> j) & 1) ^ grp[i].f[q[j]]) << j;
            }
            int pr = -1;
            for (int r = 0; r < 24; ++r)
                if (!memcmp(PERMS4[r], q2, sizeof q2)) pr = r;
            cls[pr * 16 + g2] = nrep;
        }
        reps[nrep++] = c;
    }
    int even = 1;
    for (int i = 0; i < n; ++i)
        if ((grp[i].f[0] + grp[i].f[1] + grp[i].f[2] + grp[i].f[3]) & 1) even = 0;
    memset(k->gen_class, 0, sizeof k->gen_class);
    memset(k->gen_rep, 0, sizeof k->gen_rep);
    if (even) {
        int count[2] = {0, 0}, index[384];
        for (int r = 0; r < nrep; ++r) {
            int p = __builtin_popcount(reps[r] & 15) & 1;
            index[r] = count[p];
            k->gen_rep[p][count[p]] = (uint16_t)reps[r];
            ++count[p];
        }
        k->nec = nrep / 2;
        for (int c = 0; c < 384; ++c) k->gen_class[c] = (uint16_t)index[cls[c]];
    } else {
        k->nec = nrep;
        for (int c = 0; c < 384; ++c) {
            k->gen_class[c] = (uint16_t)cls[c];
            k->gen_rep[__builtin_popcount(c & 15) & 1][cls[c]] = (uint16_t)c;
        }
    }
    k->is304 = 0;
    k->n_edge = 495ull * 256ull * (uint64_t)k->nec;
    k->n_entries = (uint64_t)N_CLASSES * k->n_edge;
    return 0;
}

int or_kind_init(okind* k, const char* name) {
    or_init();
    memset(k, 0, sizeof *k);
    if (!strcmp(name, "304")) {
        k->is304 = 1;
        k->nec = 0;
        k->n_edge = 495ull * 70ull * 16ull;
        k->n_entries = (uint64_t)N_CLASSES * k->n_edge;
        return 0;
    }
    /* kinds.py:100-107 */
    static const struct { const char* name; int n; int g[3][8]; } G[] = {
        {"s4", 3, {{0, 1, 3, 2, 0, 0, 0, 0}, {0, 2, 1, 3, 0, 0, 0, 0}, {1, 0, 2, 3, 0, 0, 0, 0}}},
        {"d4f", 3, {{0, 1, 2, 3, 1, 1, 1, 1}, {0, 3, 2, 1, 0, 0, 0, 0}, {1, 0, 3, 2, 0, 0, 0, 0}}},
        {"a4t", 2, {{0, 2, 3, 1, 0, 1, 1, 0}, {1, 0, 3, 2, 1, 1, 1, 1}}},
        {"d4t", 2, {{0, 3, 2, 1, 0, 0, 0, 0}, {1, 0, 3, 2, 1, 1, 1, 1}}},
        {"d4", 2, {{1, 2, 3, 0, 0, 0, 0, 0}, {0, 3, 2, 1, 0, 0, 0, 0}}},
        {"a4", 2, {{0, 2, 3, 1, 0, 0, 0, 0}, {1, 0, 3, 2, 0, 0, 0, 0}}},
    };
    for (unsigned i = 0; i < sizeof G / sizeof G[0]; ++i)
        if (!strcmp(G[i].name, name)) return or_group_kind(k, &G[i].g[0][0], G[i].n);
    return -2;
}

/* Edge coordinate after the corner symmetry: coords.py:121-154 (EdgeKind 3/4)
 * and kernels.cu:175-186 (EP 9 / EO 1). */
static uint64_t edge_coord(const okind* k, const uint8_t ep[12], const uint8_t eo[12]) {
    uint32_t emask = 0, eorder = 0, umask = 0, eo_e = 0, eo_n = 0;
    int ecount = 0, ocount = 0;
    for (int i = 0; i < 12; ++i) {
        if (ep[i] >= 8) {
            emask |= 1u << i;
            eorder |= (uint32_t)(ep[i] & 3) << (2 * ecount);
            eo_e |= (uint32_t)eo[i] << ecount;
            ++ecount;
        } else {
            umask |= (uint32_t)(ep[i] < 4) << ocount;
            eo_n |= (uint32_t)eo[i] << ocount;
            ++ocount;
        }
    }
    if (k->is304) return ((uint64_t)SLICE_RANK[emask] * 70u + CCOMB_RANK[umask]) * 16u + eo_e;
    uint32_t content = (uint32_t)EORDER_RANK[eorder] * 16u + eo_e;
    return ((uint64_t)SLICE_RANK[emask] * 256u + eo_n) * (uint64_t)k->nec + k->gen_class[content];
}

/* coords.py:347-352 */
uint64_t or_table_index(const okind* k, const ocube* c) {
    uint32_t packed = CLASSIDX[or_corner_raw(c)];
    uint8_t ep[12], eo[12];
    edge_conj_fast(packed & 15u, c, ep, eo);
    return (uint64_t)(packed >> 4) * k->n_edge + edge_coord(k, ep, eo);
}

/* kernels.cu:294-379 (EP 3/EO 4 and EP 9/EO 1 branches); coords.py:157-204. */
void or_unrank(const okind* k, uint64_t idx, ocube* out) {
    uint32_t cls = (uint32_t)(idx / k->n_edge);
    uint64_t ec = idx - (uint64_t)cls * k->n_edge;
    corners_from_raw(CLASS_REP[cls], out);
    uint32_t emask, eord, umask, eflips, nflips;
    if (k->is304) {
        uint32_t eo = (uint32_t)(ec % 16u), rest = (uint32_t)(ec / 16u);
        uint32_t ep = rest % 70u;
        emask = SLICE_UNRANK[rest / 70u];
        eord = 0xE4u;
        umask = CCOMB_UNRANK[ep];
        eflips = eo;
        nflips = 0;
    } else {
        uint32_t gc = (uint32_t)(ec % (uint64_t)k->nec), grest = (uint32_t)(ec / (uint64_t)k->nec);
        nflips = grest % 256u;
        emask = SLICE_UNRANK[grest / 256u];
        uint32_t content = k->gen_rep[__builtin_popcount(nflips) & 1][gc];
        eord = EORDER_UNRANK[content >> 4];
        eflips = content & 15u;
        umask = 0x0Fu;
    }
    int ne = 0, no = 0, nu = 0, nd = 0;
    for (int i = 0; i < 12; ++i) {
        if ((emask >> i) & 1u) {
            out->ep[i] = (uint8_t)(8u | ((eord >> (2 * ne)) & 3u));
            out->eo[i] = (uint8_t)((eflips >> ne) & 1u);
            ++ne;
        } else {
            out->ep[i] = (uint8_t)(((umask >> no) & 1u) ? nu++ : 4 + nd++);
            out->eo[i] = (uint8_t)((nflips >> no) & 1u);
            ++no;
        }
    }
}

/* ---- BFS: kernels.cu:389-624, table.py:82-129 -------------------------- */

#define BLOCK_ENTRIES 124ull
#define B3_ENTRIES 155ull

static inline uint32_t mod3_get(const uint32_t* t, uint64_t idx) {
    uint32_t w = __atomic_load_n(&t[idx >> 4], __ATOMIC_RELAXED);
    return (w >> ((idx & 15u) * 2u)) & 3u;
}

static inline int mod3_claim(uint32_t* t, uint64_t idx, uint32_t value) {
    uint32_t shift = (uint32_t)(idx & 15u) * 2u;
    uint32_t clear = (3u ^ value) << shift;
    uint32_t old = __atomic_fetch_and(&t[idx >> 4], ~clear, __ATOMIC_RELAXED);
    return ((old >> shift) & 3u) == 3u;
}

static void block_set(uint32_t* blk, uint64_t idx, uint32_t depth) {
    uint64_t b = idx / BLOCK_ENTRIES;
    uint32_t pos = 8u + 2u * (uint32_t)(idx - b * BLOCK_ENTRIES);
    uint32_t* p = blk + b * 8u;
    uint32_t w0 = __atomic_load_n(p, __ATOMIC_RELAXED);
    while ((w0 & 0xFFu) == 0xFFu) {
        uint32_t want = (w0 & ~0xFFu) | depth;
        if (__atomic_compare_exchange_n(p, &w0, want, 0, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
            w0 = want;
            break;
        }
    }
    uint32_t v = depth - (w0 & 0xFFu);
    if (v > 3u) v = 3u;
    if (v) __atomic_fetch_or(&p[pos >> 5], v << (pos & 31u), __ATOMIC_RELAXED);
}

static uint64_t claim_with_stab(const okind* k, uint32_t* mod3, uint32_t* blk, const ocube* n, uint32_t depth) {
    uint32_t value = depth % 3u;
    uint32_t packed = CLASSIDX[or_corner_raw(n)];
    uint32_t cls = packed >> 4, sym = packed & 15u;
    uint32_t stab = CLASS_STAB[cls];
    uint64_t base = (uint64_t)cls * k->n_edge, claimed = 0;
    while (stab) {
        uint32_t t = (uint32_t)__builtin_ctz(stab);
        stab &= stab - 1;
        uint8_t ep[12], eo[12];
        edge_conj_fast(S_MUL[t][sym], n, ep, eo);
        uint64_t idx = base + edge_coord(k, ep, eo);
        if (mod3_get(mod3, idx) == 3u && mod3_claim(mod3, idx, value)) {
            ++claimed;
            block_set(blk, idx, depth);
        }
    }
    return claimed;
}

int or_bfs(const okind* k, uint32_t* mod3, uint32_t* blk, uint64_t* counts, int max_depths, int n_threads,
           int verbose) {
    or_init();
    uint64_t n_entries = k->n_entries, n_words = (n_entries + 15) / 16;
    uint64_t n_blocks = (n_entries + BLOCK_ENTRIES - 1) / BLOCK_ENTRIES;
    if (n_threads > 0) omp_set_num_threads(n_threads);
    memset(mod3, 0xFF, n_words * 4);
    memset(blk, 0, n_blocks * 32);
    for (uint64_t b = 0; b < n_blocks; ++b) blk[b * 8] = 0xFFu;
    counts[0] = claim_with_stab(k, mod3, blk, &SOLVED, 0);
    uint64_t remaining = n_entries - counts[0];
    int nd = 1;
    uint32_t depth = 0;
    while (remaining) {
        if (nd >= max_depths) return -3;
        int fwd = counts[nd - 1] < remaining / 4;
        uint64_t found = 0;
        uint32_t cur = depth % 3u, next = (depth + 1u) % 3u;
        double t0 = omp_get_wtime();
        if (fwd) {
            uint32_t pattern = cur * 0x55555555u;
#pragma omp parallel for schedule(dynamic, 8192) reduction(+ : found)
            for (uint64_t w = 0; w < n_words; ++w) {
                uint32_t word = mod3[w];
                uint32_t x = word ^ pattern;
                uint32_t hit = ~(x | (x >> 1)) & 0x55555555u;
                while (hit) {
                    uint32_t bit = (uint32_t)__builtin_ctz(hit);
                    hit &= hit - 1;
                    uint64_t idx = w * 16u + bit / 2u;
                    if (idx >= n_entries) break;
                    ocube s = SOLVED, nb;
                    or_unrank(k, idx, &s);
                    for (int m = 0; m < 18; ++m) {
                        or_compose(&s, &MOVES[m], &nb);
                        found += claim_with_stab(k, mod3, blk, &nb, depth + 1u);
                    }
                }
            }
        } else {
#pragma omp parallel for schedule(dynamic, 8192) reduction(+ : found)
            for (uint64_t w = 0; w < n_words; ++w) {
                uint32_t word = mod3[w];
                uint32_t hit = word & (word >> 1) & 0x55555555u;
                uint32_t clear = 0;
                while (hit) {
                    uint32_t bit = (uint32_t)__builtin_ctz(hit);
                    hit &= hit - 1;
                    uint64_t idx = w * 16u + bit / 2u;
                    if (idx >= n_entries) break;
                    ocube s = SOLVED, nb;
                    or_unrank(k, idx, &s);
                    for (int m = 0; m < 18; ++m) {
                        or_compose(&s, &MOVES[m], &nb);
                        if (mod3_get(mod3, or_table_index(k, &nb)) == cur) {
                            clear |= (3u ^ next) << bit;
                            block_set(blk, idx, depth + 1u);
                            ++found;
                            break;
                        }
                    }
                }
                if (clear) __atomic_fetch_and(&mod3[w], ~clear, __ATOMIC_RELAXED);
            }
        }
        if (verbose)
            fprintf(stderr, "oracle bfs depth %2u %s found %llu (%.1fs)\n", depth + 1, fwd ? "fwd" : "bwd",
                    (unsigned long long)found, omp_get_wtime() - t0);
        if (!found) return -4;
        counts[nd++] = found;
        remaining -= found;
        ++depth;
    }
    return nd;
}

void or_histogram(const uint32_t* mod3, uint64_t n_entries, uint64_t counts[4]) {
    uint64_t c0 = 0, c1 = 0, c2 = 0, c3 = 0;
    uint64_t n_words = (n_entries + 15) / 16;
#pragma omp parallel for reduction(+ : c0, c1, c2, c3)
    for (uint64_t w = 0; w < n_words; ++w) {
        uint32_t word = mod3[w];
        uint32_t valid = 16;
        if (w * 16u + 16u > n_entries) valid = (uint32_t)(n_entries - w * 16u);
        for (uint32_t j = 0; j < valid; ++j) {
            uint32_t v = (word >> (2 * j)) & 3u;
            c0 += v == 0, c1 += v == 1, c2 += v == 2, c3 += v == 3;
        }
    }
    counts[0] = c0, counts[1] = c1, counts[2] = c2, counts[3] = c3;
}

/* kernels.cu:1825-1835 */
static uint32_t exact_value(const uint32_t* mod3, const uint32_t* blk, uint64_t x) {
    uint64_t b = x / BLOCK_ENTRIES;
    uint32_t pos = 8u + 2u * (uint32_t)(x - b * BLOCK_ENTRIES);
    const uint32_t* p = blk + b * 8u;
    uint32_t mn = p[0] & 0xFFu;
    uint32_t v = mn + ((p[pos >> 5] >> (pos & 31u)) & 3u);
    if (v - mn == 3u) {
        uint32_t m3 = mod3_get(mod3, x);
        while (v % 3u != m3) ++v;
    }
    return v;
}

/* kernels.cu:1839-1869 */
void or_build_b3(const uint32_t* mod3, const uint32_t* blk, uint64_t b0, uint64_t nb, uint64_t n_entries,
                 uint32_t* out) {
#pragma omp parallel for schedule(static, 4096)
    for (uint64_t i = 0; i < nb; ++i) {
        uint64_t first = (b0 + i) * B3_ENTRIES;
        uint8_t v[155];
        uint64_t left = n_entries - first;
        uint32_t n = left < B3_ENTRIES ? (uint32_t)left : (uint32_t)B3_ENTRIES, m = 255;
        for (uint32_t j = 0; j < n; ++j) {
            v[j] = (uint8_t)exact_value(mod3, blk, first + j);
            if (v[j] < m) m = v[j];
        }
        uint32_t best = 0, best_score = 0;
        for (uint32_t c = 0; c < 4; ++c) {
            uint32_t score = 0;
            for (uint32_t j = 0; j < n; ++j) score += v[j] <= m + c ? m : (v[j] == m + c + 1 ? m + c + 1 : m + c + 2);
            if (score > best_score) best_score = score, best = c;
        }
        uint8_t bytes[32];
        bytes[0] = (uint8_t)(m | (best << 6));
        for (uint32_t q = 0; q < 31; ++q) {
            uint32_t x = 0;
            for (int kk = 4; kk >= 0; --kk) {
                uint32_t j = q * 5u + (uint32_t)kk, d = 0;
                if (j < n) d = v[j] <= m + best ? 0u : (v[j] == m + best + 1 ? 1u : 2u);
                x = x * 3u + d;
            }
            bytes[1 + q] = (uint8_t)x;
        }
        memcpy(out + i * 8, bytes, 32);
    }
}

/* ---- lookups: kernels.cu:408-434 ---------------------------------------- */

uint32_t or_lookup(const uint8_t* table, int fmt, uint64_t idx) {
    if (fmt == 1) {
        uint64_t b = idx / B3_ENTRIES;
        uint32_t j = (uint32_t)(idx - b * B3_ENTRIES);
        const uint8_t* p = table + b * 32u;
        uint32_t hdr = p[0], byte = p[1u + j / 5u], k = j % 5u;
        static const uint32_t POW3[5] = {1, 3, 9, 27, 81};
        uint32_t d = (byte / POW3[k]) % 3u, m = hdr & 63u;
        return d == 0u ? m : m + (hdr >> 6) + d;
    }
    uint64_t b = idx / BLOCK_ENTRIES;
    uint32_t pos = 8u + 2u * (uint32_t)(idx - b * BLOCK_ENTRIES);
    const uint32_t* p = (const uint32_t*)(table + b * 32u);
    return (p[0] & 0xFFu) + ((p[pos >> 5] >> (pos & 31u)) & 3u);
}

/* ---- IDA*: kernels.cu:1447-1801 (step mode, STEP_ALL_BLOCK, no extra
 * tables), solver.py:33-64, 254-320 ------------------------------------- */

static int allowed(int last, int m) { /* solver.py:33-37 */
    if (last < 0) return 1;
    int lf = last / 3, mf = m / 3;
    return mf != lf && !(lf >= 3 && mf == lf - 3);
}

static int inverse_move(int m) { return 3 * (m / 3) + 2 - m % 3; }

/* Wall-clock deadline for bounded timing runs (0 = none): a search past it
 * stops and reports the nodes it visited. */
static double g_deadline = 0.0;
void or_set_deadline(double seconds_from_now) { g_deadline = seconds_from_now > 0 ? omp_get_wtime() + seconds_from_now : 0.0; }

typedef struct {
    const okind* k;
    const uint8_t* table;
    int fmt;
    int expired;
    int frame_move[3][18];
    uint32_t bound;
    uint64_t nodes;
    int found;
    uint8_t path[24];
    uint8_t solution[24];
} sctx;

static uint32_t lb(const sctx* S, const ocube* c) { return or_lookup(S->table, S->fmt, or_table_index(S->k, c)); }

/* frame_do: kernels.cu:777-781 */
static void frame_do(const sctx* S, ocube* f, int j, int m) {
    if (j < 3) {
        or_compose(f, &MOVES[S->frame_move[j][m]], f);
    } else {
        or_compose(&MOVES[inverse_move(S->frame_move[j - 3][m])], f, f);
    }
}

static uint32_t excluded_axes(const uint32_t* h, uint32_t budget) { /* kernels.cu:756-763 */
    return (uint32_t)(h[3] == budget) | ((uint32_t)(h[4] == budget) << 1) | ((uint32_t)(h[5] == budget) << 2);
}

static int all_axes_tight(const uint32_t* c, uint32_t budget) { /* kernels.cu:766-774 */
    if (budget == 0) return 0;
    if (c[0] == budget && c[1] == budget && c[2] == budget) return 1;
    if (c[3] == budget && c[4] == budget && c[5] == budget) return 1;
    return 0;
}

/* One node of the DFS at depth g with frames f and values h. */
static void dfs(sctx* S, ocube* f, const uint32_t* h, uint32_t g, int last) {
    uint32_t bound = S->bound;
    uint32_t excl = excluded_axes(h, bound - g);
    if (g_deadline > 0 && (S->nodes & 1023u) == 0 && omp_get_wtime() > g_deadline) S->expired = 1;
    for (int m = 0; m < 18 && !S->found && !S->expired; ++m) {
        if (!allowed(last, m)) continue;
        if ((excl >> ((m / 3) % 3)) & 1u) continue;
        S->nodes++;
        int inherit = 3 + (m / 3) % 3;
        uint32_t c[6];
        ocube t[6];
        int pruned = 0;
        for (int j = 0; j < 6; ++j) {
            t[j] = f[j];
            if (j == inherit) {
                c[j] = h[j];
                continue;
            }
            frame_do(S, &t[j], j, m);
            uint32_t v = lb(S, &t[j]);
            if (j < 3 && h[j] && h[j] - 1u > v) v = h[j] - 1u;
            c[j] = v;
            if (g + 1 + v > bound) {
                pruned = 1;
                break;
            }
        }
        if (pruned || all_axes_tight(c, bound - g - 1)) continue;
        frame_do(S, &t[inherit], inherit, m);
        S->path[g] = (uint8_t)m;
        if (g + 1 == bound) {
            if (is_solved(&t[0])) {
                S->found = 1;
                memcpy(S->solution, S->path, bound);
            }
        } else {
            dfs(S, t, c, g + 1, m);
        }
    }
}

/* Enumerate canonical prefixes of length plen in row-major order and run the
 * item at each (kernels.cu:1552-1621). */
static void run_items(sctx* S, const ocube* root_f, const uint32_t* root_h, int plen) {
    int seq[24];
    int depth = 0;
    seq[0] = -1;
    /* iterative lexicographic enumeration */
    for (;;) {
        if (S->found || S->expired) return;
        /* advance seq[depth] to the next allowed move */
        int last = depth ? seq[depth - 1] : -1;
        int m = seq[depth] + 1;
        while (m < 18 && !allowed(last, m)) ++m;
        if (m >= 18) {
            if (depth == 0) return;
            --depth;
            continue;
        }
        seq[depth] = m;
        if (depth + 1 < plen) {
            ++depth;
            seq[depth] = -1;
            continue;
        }
        /* full prefix (or plen == 0 handled below) */
        ocube f[6];
        uint32_t h[6];
        memcpy(f, root_f, sizeof f);
        memcpy(h, root_h, sizeof h);
        uint32_t bound = S->bound;
        int ok = 1;
        uint32_t g = 0;
        for (; g < (uint32_t)plen && ok; ++g) {
            int pm = seq[g];
            S->nodes++;
            if ((excluded_axes(h, bound - g) >> ((pm / 3) % 3)) & 1u) {
                ok = 0;
                break;
            }
            for (int j = 0; j < 6; ++j) {
                frame_do(S, &f[j], j, pm);
                if (j != 3 + (pm / 3) % 3) {
                    uint32_t v = lb(S, &f[j]);
                    h[j] = (j < 3 && h[j]) ? (v > h[j] - 1u ? v : h[j] - 1u) : v;
                }
                if (g + 1 + h[j] > bound) ok = 0;
            }
            if (ok && all_axes_tight(h, bound - g - 1)) ok = 0;
            S->path[g] = (uint8_t)pm;
        }
        if (ok) {
            if (g == bound) {
                if (is_solved(&f[0])) {
                    S->found = 1;
                    memcpy(S->solution, S->path, bound);
                }
            } else {
                dfs(S, f, h, g, seq[plen - 1]);
            }
        }
    }
}

static void run_round(sctx* S, const ocube* root_f, const uint32_t* root_h, int plen) {
    if (plen == 0) {
        ocube f[6];
        memcpy(f, root_f, sizeof f);
        if (S->bound == 0) {
            if (is_solved(&f[0])) S->found = 1;
        } else {
            dfs(S, f, root_h, 0, -1);
        }
        return;
    }
    run_items(S, root_f, root_h, plen);
}

int or_solve(const okind* k, const uint8_t* table, int fmt, const ocube* cube, const osolve_opts* opts,
             osolve_result* res) {
    or_init();
    memset(res, 0, sizeof *res);
    res->length = -1;
    sctx S;
    memset(&S, 0, sizeof S);
    S.k = k;
    S.table = table;
    S.fmt = fmt;
    for (int m = 0; m < 18; ++m) {
        S.frame_move[0][m] = m;
        S.frame_move[1][m] = AXIS_MAP[m];
        S.frame_move[2][m] = AXIS_MAP[AXIS_MAP[m]];
    }
    ocube f[6];
    or_axis_frames(cube, f);
    uint32_t h[6], bound = 0;
    for (int j = 0; j < 6; ++j) {
        h[j] = lb(&S, &f[j]);
        res->root_h[j] = h[j];
        if (h[j] > bound) bound = h[j];
    }
    if (is_solved(cube)) {
        res->length = 0;
        return 0;
    }
    int max_bound = opts->max_bound > 0 ? opts->max_bound : 24;
    uint64_t last = 0;
    while ((int)bound <= max_bound) {
        int plen = opts->prefix_len + (last >= opts->deeper_at[0]) + (last >= opts->deeper_at[1]);
        if ((uint32_t)plen > bound) plen = (int)bound;
        S.bound = bound;
        S.nodes = 0;
        run_round(&S, f, h, plen);
        res->round_bound[res->n_rounds] = bound;
        res->round_nodes[res->n_rounds] = S.nodes;
        res->n_rounds++;
        res->nodes += S.nodes;
        last = S.nodes;
        if (S.expired) return -2;
        if (S.found) {
            res->length = (int)bound;
            memcpy(res->moves, S.solution, bound);
            return 0;
        }
        ++bound;
    }
    return -1;
}

int or_solve_many(const okind* k, const uint8_t* table, int fmt, const ocube* cubes, int n,
                  const osolve_opts* opts, osolve_result* res, int n_threads, double* seconds) {
    or_init();
    if (n_threads > 0) omp_set_num_threads(n_threads);
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) reduction(+ : bad)
    for (int i = 0; i < n; ++i) {
        double t0 = omp_get_wtime();
        if (or_solve(k, table, fmt, &cubes[i], opts, &res[i])) ++bad;
        if (seconds) seconds[i] = omp_get_wtime() - t0;
    }
    return bad;
}

int64_t or_canonical_prefixes(int length, uint8_t* out, int64_t cap) {
    int64_t n = 0;
    int seq[24];
    if (length == 0) return 1;
    int depth = 0;
    seq[0] = -1;
    for (;;) {
        int last = depth ? seq[depth - 1] : -1;
        int m = seq[depth] + 1;
        while (m < 18 && !allowed(last, m)) ++m;
        if (m >= 18) {
            if (depth == 0) return n;
            --depth;
            continue;
        }
        seq[depth] = m;
        if (depth + 1 < length) {
            ++depth;
            seq[depth] = -1;
            continue;
        }
        if (out && n < cap)
            for (int j = 0; j < length; ++j) out[n * length + j] = (uint8_t)seq[j];
        ++n;
    }
}
