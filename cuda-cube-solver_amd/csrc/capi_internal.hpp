// Internals shared by the C-ABI translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <chrono>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/cubesolver.h"
#include "cube_model.hpp"
#include "search_types.hpp"

namespace ccs {

extern thread_local std::string g_last_error;

inline int fail(int code, const std::string& msg) {
    g_last_error = msg;
    return code;
}

#define HIP_TRY(expr)                                                                                   \
    do {                                                                                                \
        hipError_t e_ = (expr);                                                                         \
        if (e_ != hipSuccess)                                                                           \
            return fail(CCS_EDEVICE, std::string(#expr) + ": " + hipGetErrorString(e_));                \
    } while (0)

inline double now_ms() {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

inline int device_cus() {
    static int cus = 0;
    if (!cus) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return 0;
        hipDeviceProp_t p;
        if (hipGetDeviceProperties(&p, dev) != hipSuccess) return 0;
        cus = p.multiProcessorCount;
    }
    return cus;
}

// Device allocations and events owned by one call: freed on every return path.
struct DeviceBuffers {
    static constexpr bool holds_keys = true;  // room for the item keys of a banded order
    std::vector<void*> ptrs;
    std::vector<hipEvent_t> events;
    template <typename T>
    hipError_t alloc(T** p, size_t bytes) {
        *p = nullptr;
        hipError_t e = hipMalloc((void**)p, bytes);
        if (e == hipSuccess) ptrs.push_back(*p);
        return e;
    }
    hipError_t event(hipEvent_t* ev) {
        hipError_t e = hipEventCreate(ev);
        if (e == hipSuccess) events.push_back(*ev);
        return e;
    }
    // Stop owning p (its ownership moves elsewhere).
    void release(void* p) {
        for (auto& q : ptrs)
            if (q == p) q = nullptr;
    }
    DeviceBuffers() = default;
    DeviceBuffers(const DeviceBuffers&) = delete;
    DeviceBuffers& operator=(const DeviceBuffers&) = delete;
    ~DeviceBuffers() {
        for (void* p : ptrs)
            if (p) (void)hipFree(p);
        for (hipEvent_t e : events) (void)hipEventDestroy(e);
    }
};

inline int kind_K(const Kind& k) { return k.is304 ? 0 : k.nec; }

inline uint64_t format_blocks(const Kind& k, int format) {
    uint64_t per = format == CCS_FORMAT_BLOCK ? 124u : 155u;
    return (k.n_entries + per - 1) / per;
}

}  // namespace ccs

// A kind's device constants plus its resident solving table(s).
struct ccs_tables {
    ccs::Kind kind;
    int K = 0;
    int format = -1;  // solving format used by probes and searches
    ccs::CubeLUT lut{};
    ccs::CubeLUT* d_lut = nullptr;
    uint32_t* d_classidx = nullptr;
    uint32_t* d_class_rep = nullptr;
    uint16_t* d_class_stab = nullptr;
    ccs::ClassLUT* d_class_lut = nullptr;  // corner classes from small tables, for the search's LDS
    uint8_t* d_table[2] = {nullptr, nullptr};
    uint64_t table_bytes[2] = {0, 0};
    // The block table in the search's half-block layout (FMT_HALVES): made at the
    // first solve or server start on a block table; halves_state 1 made, -1 not
    // possible (no memory, or a block minimum above 15: searches then read the
    // block table itself), 0 not tried yet.
    uint8_t* d_halves = nullptr;
    int halves_state = 0;
    uint32_t* d_depths = nullptr;  // the BFS build table, kept only on request (CCS_BUILD_KEEP_DEPTHS)
    uint64_t depth_bytes = 0;
    unsigned long long* d_prefix[ccs::MAX_PLEN + 1] = {};  // packed prefixes, 8 B per item
    uint32_t prefix_n[ccs::MAX_PLEN + 1] = {};
    hipStream_t stream = nullptr;

    // Frees everything; errors are ignored here (nothing is left to report them to).
    ~ccs_tables() {
        if (stream) (void)hipStreamSynchronize(stream);
        (void)hipFree(d_lut);
        (void)hipFree(d_classidx);
        (void)hipFree(d_class_rep);
        (void)hipFree(d_class_stab);
        (void)hipFree(d_class_lut);
        for (auto* p : d_table) (void)hipFree(p);
        (void)hipFree(d_halves);
        (void)hipFree(d_depths);
        for (auto* p : d_prefix) (void)hipFree(p);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace ccs {
int tables_setup(ccs_tables* t, const std::string& name);
// Table format of the searches (FMT_HALVES for a block table whose half-block
// copy exists; made here first if `make` and not tried yet), and that table.
int search_format(ccs_tables* t, bool make);
const uint8_t* search_table(const ccs_tables* t, int F);
int ensure_prefixes(ccs_tables* t, int plen);
}  // namespace ccs
