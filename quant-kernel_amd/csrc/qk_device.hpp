// qk_device.hpp - the process-wide device context shared by the MC engine
// and the surrogate: HIP device, one default-priority stream, timing events,
// grow-only device / pinned buffers and the lock that serialises device work.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdlib>
#include <mutex>
#include <string>
#include <vector>

#include "qk_kernels.hpp"

namespace qk {

#define QK_HIP(call)                                                                   \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            if (err) *err = std::string("HIP error: ") + hipGetErrorString(e_) + " (" + \
                            #call + ")";                                               \
            return QK_ERR_RUNTIME_INIT;                                                \
        }                                                                              \
    } while (0)

// Grow-only device buffer.
struct DevBuf {
    void* p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap) return hipSuccess;
        if (p) {
            hipError_t e = hipFree(p);
            if (e != hipSuccess) return e;
            p = nullptr;
            cap = 0;
        }
        size_t want = std::max<size_t>(bytes, 4096);
        want = (want + 4095) & ~size_t(4095);
        hipError_t e = hipMalloc(&p, want);
        if (e != hipSuccess) {
            p = nullptr;
            return e;
        }
        cap = want;
        return hipSuccess;
    }
};

// Grow-only pinned host buffer.
struct PinBuf {
    void* p = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap) return hipSuccess;
        if (p) {
            hipError_t e = hipHostFree(p);
            if (e != hipSuccess) return e;
            p = nullptr;
            cap = 0;
        }
        size_t want = std::max<size_t>(bytes, 4096);
        want = (want + 4095) & ~size_t(4095);
        hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
        if (e != hipSuccess) {
            p = nullptr;
            return e;
        }
        cap = want;
        return hipSuccess;
    }
};

inline int env_int(const char* name, int dflt) {
    const char* v = std::getenv(name);
    if (!v || !*v) return dflt;
    return std::atoi(v);
}

struct Device {
    int dev = 0;
    int cus = 0;
    std::string name;
    hipStream_t stream = nullptr;
    hipEvent_t ev[4] = {};
    hipEvent_t ev_sur[3] = {};  // surrogate pricing phases
    std::mutex mu;

    DevBuf in, recs, ntiles, tile_start, counters, partials, out, stream_buf, stream_off, scan;
    DevBuf lsmc_path, lsmc_cf, lsmc_recs, lsmc_part, lsmc_beta;
    PinBuf pin_in, pin_out;

    double ms_total = -1.0, ms_sim = -1.0;

    SimShape shape_step{8, 0};
    SimShape shape_term{24, 0};

    // instantiated launch graphs of the host-memory MC path, keyed by
    // everything their nodes depend on (sizes, shapes and buffer addresses)
    struct Graph {
        std::vector<uint64_t> key;
        hipGraphExec_t exec = nullptr;
    };
    DevBuf gl;               // Gauss-Legendre nodes / weights of the Heston CF pricer
    bool gl_ready = false;
    std::vector<Graph> graphs;
    bool graphs_ok = true;        // capture and instantiation work on this runtime
    bool graphs_enabled = true;   // caller switch (qk_mc_set_graphs)
    int64_t graph_launches = 0;   // replays + first launches, for diagnostics

    // Device address of this device's copy of the allocation-free entry
    // point's fixed scratch array (hipGetSymbolAddress at set-up), or null
    // with noalloc_err saying why.
    char* noalloc_scratch = nullptr;
    std::string noalloc_err;
};

// Returns the context (creating it on first use) or null with *err set.
Device* get_device(std::string* err);

}  // namespace qk
