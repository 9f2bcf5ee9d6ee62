// C ABI of libcubesolver: table load/build/save, probes, solve, serving.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/cubesolver.h"
#include "cube_model.hpp"
#include "capi_internal.hpp"
#include "kernels.hpp"

using namespace ccs;

namespace ccs {
thread_local std::string g_last_error;
}  // namespace ccs

namespace ccs {

// The ClassLUT tables must give classidx for every raw corner coordinate;
// checked once per process, before any table is set up.
static int class_lut_check() {
    static int rc = -1;
    static std::once_flag once;
    std::call_once(once, [] {
        const CornerClasses& cc = corner_classes();
        uint32_t bad = 0;
        for (uint32_t t = 0; t < N_TWIST; ++t)
            for (uint32_t p = 0; p < N_DPOS; ++p) bad += corner_class(cc.class_lut, t, p) != cc.classidx[t * N_DPOS + p];
        rc = bad == 0 && cc.class_lut_rows[0] == (int)N_MULTI_TWIST && cc.class_lut_rows[1] == (int)N_MULTI_ORBIT;
    });
    return rc;
}

// The search keeps a cube's state and its inverse only, and reaches axis
// frame k through the LUT's corner conjugation records and composite edge
// tables: for sampled cubes, every frame's table index that way must equal the
// index of the frame built explicitly (axis_frames).
template <int K>
static bool frame_index_check_k(const CubeLUT& L) {
    const ClassLUT& C = corner_classes().class_lut;
    uint32_t rng = 12345u;
    for (int n = 0; n < 256; ++n) {
        Cubie c = solved();
        for (int d = 0; d < 25; ++d) {
            rng = rng * 1664525u + 1013904223u;
            c = then(c, move((int)((rng >> 8) % 18u)));
        }
        Cubie fr[6];
        axis_frames(c, fr);
        for (int j = 0; j < 6; ++j) {
            const Packed want = pack(fr[j]), base = pack(j < 3 ? fr[0] : fr[3]);
            const Frame fw{want.w[0], want.w[1], want.w[2], want.w[3], want.w[4]};
            Frame fb{base.w[0], base.w[1], base.w[2], base.w[3], base.w[4]};
            conj_corners(fb, L.cj[j % 3]);
            const uint32_t pk = corner_class(C, fb);
            const uint64_t got = (uint64_t)(pk >> 4) * KindT<K>::N_EDGE + edge_coord<K>(L, fb, (pk & 15u) + 16u * (j % 3));
            if (got != table_index_c<K>(L, C, fw)) return false;
        }
    }
    return true;
}
static bool frame_index_check(const Kind& k, const CubeLUT& L) {
    switch (kind_K(k)) {
        case 0: return frame_index_check_k<0>(L);
        case 8: return frame_index_check_k<8>(L);
        case 12: return frame_index_check_k<12>(L);
        case 16: return frame_index_check_k<16>(L);
        case 24: return frame_index_check_k<24>(L);
        default: return false;
    }
}

int search_format(ccs_tables* t, bool make) {
    if (t->format != CCS_FORMAT_BLOCK) return t->format;
    if (t->halves_state == 0 && make) {
        t->halves_state = -1;
        const uint64_t nb = t->table_bytes[CCS_FORMAT_BLOCK] / 32;
        unsigned int* d_bad = nullptr;
        unsigned int bad = 1;
        if (hipMalloc(&t->d_halves, nb * 32) == hipSuccess && hipMalloc(&d_bad, 4) == hipSuccess &&
            hipMemsetAsync(d_bad, 0, 4, t->stream) == hipSuccess &&
            launch_halves(device_cus() * 16, t->d_table[CCS_FORMAT_BLOCK], t->d_halves, nb, d_bad, t->stream) ==
                hipSuccess &&
            hipMemcpyAsync(&bad, d_bad, 4, hipMemcpyDeviceToHost, t->stream) == hipSuccess &&
            hipStreamSynchronize(t->stream) == hipSuccess && bad == 0)
            t->halves_state = 1;
        (void)hipFree(d_bad);
        if (t->halves_state != 1) {
            (void)hipFree(t->d_halves);
            t->d_halves = nullptr;
            (void)hipGetLastError();  // an allocation that failed leaves nothing behind
        }
    }
    return t->halves_state == 1 ? FMT_HALVES : CCS_FORMAT_BLOCK;
}

const uint8_t* search_table(const ccs_tables* t, int F) {
    return F == FMT_HALVES ? t->d_halves : t->d_table[t->format];
}

int tables_setup(ccs_tables* t, const std::string& name) {
    std::string err;
    if (!class_lut_check()) return fail(CCS_EFAULT, "corner class tables disagree with classidx");
    if (!make_kind(name, t->kind, err)) return fail(CCS_EINVAL, err);
    t->K = kind_K(t->kind);
    if (t->K != 0 && t->K != 8 && t->K != 12 && t->K != 16 && t->K != 24)
        return fail(CCS_EINVAL, name + ": no kernel instance for " + std::to_string(t->K) + " E-content classes");
    build_lut(t->kind, t->lut);
    if (!frame_index_check(t->kind, t->lut)) return fail(CCS_EFAULT, "axis-frame conjugation tables give wrong indices");
    const CornerClasses& cc = corner_classes();
    HIP_TRY(hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking));
    HIP_TRY(hipMalloc(&t->d_lut, sizeof(CubeLUT)));
    HIP_TRY(hipMalloc(&t->d_classidx, cc.classidx.size() * 4));
    HIP_TRY(hipMalloc(&t->d_class_rep, cc.rep.size() * 4));
    HIP_TRY(hipMalloc(&t->d_class_stab, cc.stab.size() * 2));
    HIP_TRY(hipMalloc(&t->d_class_lut, sizeof(ClassLUT)));
    HIP_TRY(hipMemcpy(t->d_class_lut, &cc.class_lut, sizeof(ClassLUT), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(t->d_lut, &t->lut, sizeof(CubeLUT), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(t->d_classidx, cc.classidx.data(), cc.classidx.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(t->d_class_rep, cc.rep.data(), cc.rep.size() * 4, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(t->d_class_stab, cc.stab.data(), cc.stab.size() * 2, hipMemcpyHostToDevice));
    // The work-item prefixes of the default prefix length 4 and its deeper rounds
    // (lengths 0..6, 62 MB): resident with the table, so that a solve needs no
    // allocation for them.
    for (int p = 0; p <= 6; ++p) {
        int rc = ensure_prefixes(t, p);
        if (rc) return rc;
    }
    return CCS_OK;
}

int ensure_prefixes(ccs_tables* t, int plen) {
    if (plen < 0 || plen > MAX_PLEN) return fail(CCS_EINVAL, "prefix length out of range");
    if (t->prefix_n[plen]) return CCS_OK;
    // one item per 64-bit word, move j in byte j (zero past the prefix)
    std::vector<unsigned long long> words(1, 0ull);
    if (plen > 0) {
        std::vector<uint8_t> pre = canonical_prefixes(plen);
        const size_t n = pre.size() / plen;
        words.assign(n, 0ull);
        for (size_t i = 0; i < n; ++i)
            for (int j = 0; j < plen; ++j) words[i] |= (unsigned long long)pre[i * plen + j] << (8 * j);
    }
    HIP_TRY(hipMalloc(&t->d_prefix[plen], words.size() * 8));
    HIP_TRY(hipMemcpy(t->d_prefix[plen], words.data(), words.size() * 8, hipMemcpyHostToDevice));
    t->prefix_n[plen] = (uint32_t)words.size();
    return CCS_OK;
}

}  // namespace ccs


extern "C" {

int ccs_version(void) { return 100; }

const char* ccs_last_error(void) { return g_last_error.c_str(); }

int ccs_device_info(ccs_device_info_t* out) {
    if (!out) return fail(CCS_EINVAL, "null output");
    memset(out, 0, sizeof *out);
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    hipDeviceProp_t p;
    HIP_TRY(hipGetDeviceProperties(&p, dev));
    size_t fr = 0, tot = 0;
    HIP_TRY(hipMemGetInfo(&fr, &tot));
    out->free_bytes = fr;
    out->total_bytes = tot;
    out->compute_units = p.multiProcessorCount;
    out->device = dev;
    snprintf(out->name, sizeof out->name, "%s", p.name);
    snprintf(out->arch, sizeof out->arch, "%s", p.gcnArchName);
    return CCS_OK;
}

int ccs_kind_info(const char* kind, ccs_kind_info_t* out) {
    if (!kind || !out) return fail(CCS_EINVAL, "null argument");
    Kind k;
    std::string err;
    if (!make_kind(kind, k, err)) return fail(CCS_EINVAL, err);
    out->n_edge = k.n_edge;
    out->n_entries = k.n_entries;
    out->n_classes = (int32_t)corner_classes().rep.size();
    out->nec = k.nec;
    return CCS_OK;
}

int ccs_tables_open(const char* path, const char* kind, int format, ccs_tables_t** out) {
    if (!path || !kind || !out) return fail(CCS_EINVAL, "null argument");
    if (format != CCS_FORMAT_BLOCK && format != CCS_FORMAT_B3) return fail(CCS_EINVAL, "unknown table format");
    *out = nullptr;
    std::unique_ptr<ccs_tables> t(new ccs_tables());
    int rc = tables_setup(t.get(), kind);
    if (rc) return rc;
    FILE* fh = fopen(path, "rb");
    if (!fh) return fail(CCS_ENOENT, std::string(path) + " not found");
    fseeko(fh, 0, SEEK_END);
    uint64_t size = (uint64_t)ftello(fh);
    fseeko(fh, 0, SEEK_SET);
    uint64_t want = format_blocks(t->kind, format) * 32u;
    if (size != want) {
        fclose(fh);
        return fail(CCS_EIO, std::string(path) + ": size " + std::to_string(size) + ", expected " + std::to_string(want));
    }
    hipError_t e = hipMalloc(&t->d_table[format], want);
    if (e != hipSuccess) {
        fclose(fh);
        return fail(CCS_ENOMEM, "table allocation: " + std::string(hipGetErrorString(e)));
    }
    t->table_bytes[format] = want;
    t->format = format;
    // Read into two pinned buffers in turn while the previous chunk copies.
    const uint64_t CH = 64ull << 20;
    void* pin[2] = {nullptr, nullptr};
    hipEvent_t ev[2] = {nullptr, nullptr};
    int result = CCS_OK;
    for (int i = 0; i < 2 && result == CCS_OK; ++i) {
        if (hipHostMalloc(&pin[i], CH, hipHostMallocDefault) != hipSuccess || hipEventCreate(&ev[i]) != hipSuccess)
            result = fail(CCS_ENOMEM, "pinned staging buffer");
    }
    for (uint64_t off = 0, i = 0; off < want && result == CCS_OK; off += CH, ++i) {
        int b = (int)(i & 1);
        if (i >= 2 && hipEventSynchronize(ev[b]) != hipSuccess) {  // the buffer's previous upload
            result = fail(CCS_EDEVICE, "table upload failed");
            break;
        }
        uint64_t n = std::min(CH, want - off);
        if (fread(pin[b], 1, n, fh) != n) {
            result = fail(CCS_EIO, std::string(path) + ": short read");
            break;
        }
        if (hipMemcpyAsync(t->d_table[format] + off, pin[b], n, hipMemcpyHostToDevice, t->stream) != hipSuccess ||
            hipEventRecord(ev[b], t->stream) != hipSuccess)
            result = fail(CCS_EDEVICE, "table upload failed");
    }
    if (hipStreamSynchronize(t->stream) != hipSuccess && result == CCS_OK)
        result = fail(CCS_EDEVICE, "table upload failed");
    for (int i = 0; i < 2; ++i) {
        if (pin[i]) (void)hipHostFree(pin[i]);
        if (ev[i]) (void)hipEventDestroy(ev[i]);
    }
    fclose(fh);
    if (result != CCS_OK) return result;
    *out = t.release();
    return CCS_OK;
}

int ccs_tables_build(const char* kind, int format, ccs_log_fn log, void* user, uint64_t* depth_counts,
                     int* n_depths, ccs_tables_t** out) {
    if (!kind || !out) return fail(CCS_EINVAL, "null argument");
    const bool keep_depths = (format & CCS_BUILD_KEEP_DEPTHS) != 0;
    format &= ~CCS_BUILD_KEEP_DEPTHS;
    if (format < 0 || format > 2) return fail(CCS_EINVAL, "unknown table format");
    *out = nullptr;
    std::unique_ptr<ccs_tables> t(new ccs_tables());
    int rc = tables_setup(t.get(), kind);
    if (rc) return rc;
    const uint64_t n_entries = t->kind.n_entries, n_words = (n_entries + 7) / 8;
    DeviceBuffers buf;  // build-only memory, freed on every return
    uint32_t* tbl = nullptr;
    unsigned long long* d_cnt = nullptr;
    unsigned int* d_bad = nullptr;
    auto logf = [&](const std::string& s) {
        if (log) log(s.c_str(), user);
    };
    if (buf.alloc(&tbl, n_words * 4) != hipSuccess)
        return fail(CCS_ENOMEM, "build table allocation (" + std::to_string(n_words * 4) + " bytes)");
    HIP_TRY(buf.alloc(&d_cnt, 48 * sizeof(unsigned long long)));
    HIP_TRY(buf.alloc(&d_bad, sizeof(unsigned int)));
    HIP_TRY(hipMemsetAsync(tbl, 0xFF, n_words * 4, t->stream));
    HIP_TRY(hipMemsetAsync(d_cnt, 0, 48 * sizeof(unsigned long long), t->stream));
    HIP_TRY(hipMemsetAsync(d_bad, 0, sizeof(unsigned int), t->stream));
    HIP_TRY(launch_bfs_seed(t->K, t->d_lut, tbl, t->d_classidx, t->d_class_stab, d_cnt, t->stream));
    std::vector<uint64_t> counts;
    unsigned long long c0 = 0;
    HIP_TRY(hipMemcpyAsync(&c0, d_cnt, 8, hipMemcpyDeviceToHost, t->stream));
    HIP_TRY(hipStreamSynchronize(t->stream));
    counts.push_back(c0);
    uint64_t remaining = n_entries - c0;
    const int grid = device_cus() * 16;
    uint32_t depth = 0;
    double t_all = now_ms();
    while (remaining) {
        if (counts.size() >= 16) return fail(CCS_EFAULT, "BFS did not finish within 15 depths");
        // top-down while the frontier is small next to what is left (table.py:112)
        bool fwd = counts.back() < remaining / 4;
        unsigned long long* slot = d_cnt + 1 + depth;
        double t0 = now_ms();
        HIP_TRY(launch_bfs_layer(t->K, fwd, grid, t->d_lut, tbl, t->d_classidx, t->d_class_rep, t->d_class_stab,
                                 n_words, n_entries, depth, slot, t->stream));
        unsigned long long found = 0;
        HIP_TRY(hipMemcpyAsync(&found, slot, 8, hipMemcpyDeviceToHost, t->stream));
        HIP_TRY(hipStreamSynchronize(t->stream));
        char line[160];
        snprintf(line, sizeof line, "depth %2u %s found %14llu (%7.3fs)", depth + 1, fwd ? "fwd" : "bwd", found,
                 (now_ms() - t0) / 1e3);
        logf(line);
        if (found == 0) return fail(CCS_EFAULT, std::to_string(remaining) + " entries unreachable");
        counts.push_back(found);
        remaining -= found;
        ++depth;
    }
    // Self-check: the histogram of exact depths equals the per-depth counts.
    unsigned long long* d_hist = d_cnt + 32;
    HIP_TRY(hipMemsetAsync(d_hist, 0, 16 * 8, t->stream));
    HIP_TRY(launch_histogram(grid, tbl, n_words, n_entries, d_hist, t->stream));
    unsigned long long hist[16];
    HIP_TRY(hipMemcpyAsync(hist, d_hist, sizeof hist, hipMemcpyDeviceToHost, t->stream));
    HIP_TRY(hipStreamSynchronize(t->stream));
    for (int v = 0; v < 16; ++v) {
        uint64_t want = (size_t)v < counts.size() ? counts[v] : 0;
        if (hist[v] != want)
            return fail(CCS_EFAULT, "histogram of depth " + std::to_string(v) + " is " + std::to_string(hist[v]) +
                                        ", depth counts say " + std::to_string(want));
    }
    // Encoders.  The block pass also writes each 124-entry block's minimum to a
    // byte array, which the base-3 encoder reads for the reference's saturated-
    // offset values; a base-3-only build runs the pass for the minima alone and
    // never holds the block table.
    const uint64_t nb124 = format_blocks(t->kind, CCS_FORMAT_BLOCK);
    uint8_t* d_min = nullptr;
    HIP_TRY(buf.alloc(&d_min, nb124));
    const uint64_t CH = 1ull << 26;  // blocks per launch
    for (int fmt = 0; fmt < 2; ++fmt) {
        const bool keep = format == 2 || format == fmt;
        if (fmt == 1 && !keep) continue;
        const uint64_t nb = format_blocks(t->kind, fmt);
        if (keep) {
            if (hipMalloc(&t->d_table[fmt], nb * 32) != hipSuccess) {
                t->d_table[fmt] = nullptr;
                return fail(CCS_ENOMEM, "solving table allocation");
            }
            t->table_bytes[fmt] = nb * 32;
        }
        const double te = now_ms();
        uint32_t* dst = reinterpret_cast<uint32_t*>(t->d_table[fmt]);
        for (uint64_t b0 = 0; b0 < nb; b0 += CH)
            HIP_TRY(launch_encode(fmt, tbl, d_min, n_entries, b0, std::min(CH, nb - b0), dst ? dst + b0 * 8 : nullptr,
                                  d_bad, t->stream));
        HIP_TRY(hipStreamSynchronize(t->stream));
        char line[160];
        if (keep)
            snprintf(line, sizeof line, "encoded %s table, %llu bytes (%7.3fs)", fmt ? "b3" : "block",
                     (unsigned long long)(nb * 32), (now_ms() - te) / 1e3);
        else
            snprintf(line, sizeof line, "block minima for the b3 encoder, %llu bytes (%7.3fs)",
                     (unsigned long long)nb124, (now_ms() - te) / 1e3);
        logf(line);
    }
    unsigned int bad = 0;
    HIP_TRY(hipMemcpyAsync(&bad, d_bad, 4, hipMemcpyDeviceToHost, t->stream));
    HIP_TRY(hipStreamSynchronize(t->stream));
    if (bad) return fail(CCS_EFAULT, "block table has an unset block minimum");
    t->format = format == 2 ? CCS_FORMAT_BLOCK : format;  // both: probes and searches use the block table
    if (keep_depths) {  // the handle takes the build table
        buf.release(tbl);
        t->d_depths = tbl;
        t->depth_bytes = n_words * 4;
    }
    char line[160];
    snprintf(line, sizeof line, "built %s in %.1fs", kind, (now_ms() - t_all) / 1e3);
    logf(line);
    if (depth_counts && n_depths) {
        int cap = *n_depths;
        for (int i = 0; i < (int)counts.size() && i < cap; ++i) depth_counts[i] = counts[i];
        *n_depths = (int)counts.size();
    } else if (n_depths) {
        *n_depths = (int)counts.size();
    }
    *out = t.release();
    return CCS_OK;
}

int ccs_tables_use(ccs_tables_t* t, int format) {
    if (!t || format < 0 || format > 1) return fail(CCS_EINVAL, "bad arguments");
    if (!t->d_table[format]) return fail(CCS_EINVAL, "that table format is not resident");
    t->format = format;
    return CCS_OK;
}

uint64_t ccs_tables_bytes(ccs_tables_t* t, int format) {
    if (!t) return 0;
    if (format == CCS_TABLE_DEPTHS) return t->depth_bytes;
    if (format < 0 || format > 1) return 0;
    return t->table_bytes[format];
}

int ccs_tables_read(ccs_tables_t* t, int format, uint64_t offset, uint64_t n, void* host) {
    if (!t || (n && !host)) return fail(CCS_EINVAL, "bad arguments");
    const uint8_t* src = nullptr;
    uint64_t size = 0;
    if (format == CCS_TABLE_DEPTHS) {
        src = reinterpret_cast<const uint8_t*>(t->d_depths);
        size = t->depth_bytes;
    } else if (format == 0 || format == 1) {
        src = t->d_table[format];
        size = t->table_bytes[format];
    }
    if (!src) return fail(CCS_EINVAL, "table not resident");
    if (offset > size || n > size - offset) return fail(CCS_EINVAL, "read past the end of the table");
    HIP_TRY(hipMemcpy(host, src + offset, n, hipMemcpyDeviceToHost));
    return CCS_OK;
}

int ccs_tables_save(ccs_tables_t* t, int format, const char* path) {
    if (!t || !path) return fail(CCS_EINVAL, "null argument");
    if (format < 0 || format > 1 || !t->d_table[format]) return fail(CCS_EINVAL, "table not resident");
    std::string tmp = std::string(path) + ".tmp";
    FILE* fh = fopen(tmp.c_str(), "wb");
    if (!fh) return fail(CCS_EIO, "cannot write " + tmp);
    const uint64_t CH = 64ull << 20;
    void* pin = nullptr;
    if (hipHostMalloc(&pin, CH, hipHostMallocDefault) != hipSuccess) {
        fclose(fh);
        return fail(CCS_ENOMEM, "pinned staging buffer");
    }
    int result = CCS_OK;
    for (uint64_t off = 0; off < t->table_bytes[format]; off += CH) {
        uint64_t n = std::min(CH, t->table_bytes[format] - off);
        if (hipMemcpy(pin, t->d_table[format] + off, n, hipMemcpyDeviceToHost) != hipSuccess) {
            result = fail(CCS_EDEVICE, "table download failed");
            break;
        }
        if (fwrite(pin, 1, n, fh) != n) {
            result = fail(CCS_EIO, "short write to " + tmp);
            break;
        }
    }
    (void)hipHostFree(pin);
    if (fclose(fh) != 0 && result == CCS_OK) result = fail(CCS_EIO, "close " + tmp);
    if (result != CCS_OK) {
        remove(tmp.c_str());
        return result;
    }
    if (rename(tmp.c_str(), path) != 0) return fail(CCS_EIO, "rename " + tmp);
    return CCS_OK;
}

void ccs_tables_close(ccs_tables_t* t) { delete t; }

int ccs_probe(ccs_tables_t* t, const uint8_t* cubes, int n, int move, uint64_t* idx_out, uint32_t* lb_out,
              uint8_t* state_out) {
    if (!t || (!cubes && n > 0) || n < 0) return fail(CCS_EINVAL, "bad arguments");
    if (move > 17 || move < -19) return fail(CCS_EINVAL, "bad move");
    if (n == 0) return CCS_OK;
    std::vector<Packed> in(n);
    for (int i = 0; i < n; ++i) in[i] = pack(*reinterpret_cast<const Cubie*>(cubes + 40 * (size_t)i));
    DeviceBuffers buf;
    Frame* d_in = nullptr;
    unsigned long long* d_idx = nullptr;
    uint32_t* d_lb = nullptr;
    Frame* d_state = nullptr;
    HIP_TRY(buf.alloc(&d_in, sizeof(Frame) * n));
    HIP_TRY(buf.alloc(&d_idx, 8ull * n));
    HIP_TRY(buf.alloc(&d_lb, 4ull * n));
    HIP_TRY(buf.alloc(&d_state, sizeof(Frame) * n));
    HIP_TRY(hipMemcpy(d_in, in.data(), sizeof(Frame) * n, hipMemcpyHostToDevice));
    int fmt = t->format < 0 ? 0 : t->format;
    HIP_TRY(launch_probe(t->K, fmt, t->d_lut, t->d_classidx, t->format < 0 ? nullptr : t->d_table[fmt], d_in, n, move,
                         d_idx, t->format < 0 ? nullptr : d_lb, d_state, t->stream));
    HIP_TRY(hipStreamSynchronize(t->stream));
    if (idx_out) HIP_TRY(hipMemcpy(idx_out, d_idx, 8ull * n, hipMemcpyDeviceToHost));
    if (lb_out && t->format >= 0) HIP_TRY(hipMemcpy(lb_out, d_lb, 4ull * n, hipMemcpyDeviceToHost));
    if (state_out) {
        std::vector<Packed> st(n);
        HIP_TRY(hipMemcpy(st.data(), d_state, sizeof(Frame) * n, hipMemcpyDeviceToHost));
        for (int i = 0; i < n; ++i) *reinterpret_cast<Cubie*>(state_out + 40 * (size_t)i) = unpack(st[i]);
    }
    return CCS_OK;
}

int ccs_unrank_check(ccs_tables_t* t, const uint64_t* idx, int n, uint64_t* out) {
    if (!t || n < 0 || (n && (!idx || !out))) return fail(CCS_EINVAL, "bad arguments");
    for (int i = 0; i < n; ++i)
        if (idx[i] >= t->kind.n_entries) return fail(CCS_EINVAL, "index out of range");
    if (n == 0) return CCS_OK;
    DeviceBuffers buf;
    unsigned long long *d_idx = nullptr, *d_out = nullptr;
    HIP_TRY(buf.alloc(&d_idx, 8ull * n));
    HIP_TRY(buf.alloc(&d_out, 8ull * n));
    HIP_TRY(hipMemcpy(d_idx, idx, 8ull * n, hipMemcpyHostToDevice));
    HIP_TRY(launch_unrank_check(t->K, t->d_lut, t->d_classidx, t->d_class_rep, d_idx, n, d_out, t->stream));
    HIP_TRY(hipStreamSynchronize(t->stream));
    HIP_TRY(hipMemcpy(out, d_out, 8ull * n, hipMemcpyDeviceToHost));
    return CCS_OK;
}

int ccs_gather_bench(ccs_tables_t* t, int mode, uint64_t n_probes, int blocks_per_cu, double* ms,
                     uint64_t* checksum) {
    if (!t || mode < 0 || mode > 3 || !ms) return fail(CCS_EINVAL, "bad arguments");
    int fmt = t->format < 0 ? 0 : t->format;
    if (!t->d_table[fmt]) return fail(CCS_EINVAL, "no table resident");
    DeviceBuffers buf;
    unsigned long long* d_sum = nullptr;
    HIP_TRY(buf.alloc(&d_sum, 8));
    HIP_TRY(hipMemsetAsync(d_sum, 0, 8, t->stream));
    hipEvent_t e0, e1;
    HIP_TRY(buf.event(&e0));
    HIP_TRY(buf.event(&e1));
    int grid = device_cus() * (blocks_per_cu > 0 ? blocks_per_cu : 8);
    HIP_TRY(hipEventRecord(e0, t->stream));
    HIP_TRY(launch_gather(mode, grid, reinterpret_cast<const uint32_t*>(t->d_table[fmt]), t->table_bytes[fmt] / 32,
                          n_probes, d_sum, t->stream));
    HIP_TRY(hipEventRecord(e1, t->stream));
    HIP_TRY(hipEventSynchronize(e1));
    float f = 0;
    HIP_TRY(hipEventElapsedTime(&f, e0, e1));
    *ms = f;
    unsigned long long sum = 0;
    HIP_TRY(hipMemcpy(&sum, d_sum, 8, hipMemcpyDeviceToHost));
    if (checksum) *checksum = sum;
    return CCS_OK;
}

int ccs_index_host(const char* kind, const uint8_t* cubes, int n, uint64_t* idx_out) {
    if (!kind || n < 0 || (n && (!cubes || !idx_out))) return fail(CCS_EINVAL, "bad arguments");
    Kind k;
    std::string err;
    if (!make_kind(kind, k, err)) return fail(CCS_EINVAL, err);
    if (!class_lut_check()) return fail(CCS_EFAULT, "corner class tables disagree with classidx");
    std::unique_ptr<CubeLUT> L(new CubeLUT());
    build_lut(k, *L);
    if (!frame_index_check(k, *L)) return fail(CCS_EFAULT, "axis-frame conjugation tables give wrong indices");
    const ClassLUT& C = corner_classes().class_lut;  // the search's path to the corner class
    for (int i = 0; i < n; ++i) {
        Packed p = pack(*reinterpret_cast<const Cubie*>(cubes + 40 * (size_t)i));
        Frame f{p.w[0], p.w[1], p.w[2], p.w[3], p.w[4]};
        switch (kind_K(k)) {
            case 0: idx_out[i] = table_index_c<0>(*L, C, f); break;
            case 8: idx_out[i] = table_index_c<8>(*L, C, f); break;
            case 12: idx_out[i] = table_index_c<12>(*L, C, f); break;
            case 16: idx_out[i] = table_index_c<16>(*L, C, f); break;
            case 24: idx_out[i] = table_index_c<24>(*L, C, f); break;
            default: return fail(CCS_EINVAL, "unsupported kind");
        }
    }
    return CCS_OK;
}

}  // extern "C"
