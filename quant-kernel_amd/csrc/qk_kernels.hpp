// qk_kernels.hpp - host-side interface of the gfx950 Monte Carlo kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "qk_common.hpp"

namespace qk {

// Device-resident SoA inputs of one batch.  Arrays a method does not use may
// be null.  `method` holds Method values, or QK_MODEL_* ids when
// method_is_model is set; when `method` is null every element uses
// uniform_method.
struct BatchIn {
    const int32_t* method;
    const double *spot, *strike, *t, *vol, *r, *q;
    const double *v0, *kappa, *theta, *sigma, *rho;
    const int32_t *ot, *paths, *steps;
    const uint64_t* seed;
    const double* aux;       // optional per-element shift / weight clip
    int32_t uniform_method;
    int32_t method_is_model;
    const uint8_t* active;   // optional: elements with active[i] == 0 are skipped (NaN)
    int32_t paths_override;  // > 0: every element uses this path count
};

// Path precision / normal source of a simulation launch.
enum SimMode : int32_t {
    SIM_PHILOX_F32 = 0,  // production: Philox + Box-Muller, FP32 paths
    SIM_STREAM_F64 = 1,  // validation: host normal stream, FP64 paths
    SIM_STREAM_F32 = 2,  // host normal stream, FP32 paths
};

// Launch geometry (autotuned per family).
struct SimShape {
    int32_t lane_units;  // stepwise: paths per lane per tile; terminal: Philox groups per lane
    int32_t grid;        // persistent workgroups
};

constexpr int kBlock = 256;  // threads per workgroup (4 wave64)

inline int32_t tile_units_stepwise(const SimShape& s) { return kBlock * s.lane_units; }
// terminal tiles hold kBlock * lane_units groups of term_group_units(method) units
inline int32_t tile_groups_terminal(const SimShape& s) { return kBlock * s.lane_units; }

// counters[0] = next tile, counters[1] = tiles to run, counters[2] = overflow
// (1: more tiles than `capacity`, 2: more than 2^31-1; no tile runs and the
// finalizer writes NaN), counters[3] = tiles needed (clamped to 2^31-1)
// n: *n_dev when n_dev is set, else n_host (then stored to n_out if set)
hipError_t launch_prepare(hipStream_t st, const BatchIn& in, const int32_t* n_dev, int32_t n_host,
                          int32_t* n_out, const int64_t* stream_off, int32_t tg_terminal,
                          int32_t tu_stepwise, OptionRec* recs, int32_t* ntiles);
// scratch: (n + 1)/1024 + 1 int64 block sums
hipError_t launch_scan(hipStream_t st, int32_t n, const int32_t* n_dev, const int32_t* ntiles,
                       OptionRec* recs, int32_t* tile_start, int32_t* counters, void* scratch,
                       int32_t capacity);
// pair_only: a terminal batch whose every option is standard MC or
// antithetic variates (a kernel without the other estimators' branches).
hipError_t launch_simulate(hipStream_t st, int32_t family, int32_t mode, const SimShape& shape,
                           const OptionRec* recs, const int32_t* tile_start,
                           const int32_t* n_dev, int32_t* counters, Partial* partials,
                           const void* stream, bool pair_only);
// host_flags (nullable): device address of pinned host memory that receives
// counters[0..3]
hipError_t launch_finalize(hipStream_t st, const int32_t* n_dev, const OptionRec* recs,
                           const Partial* partials, const int32_t* counters, double* out_price,
                           double* out_se, int32_t* host_flags);

// Writes the angle table of the four-pair Philox layout into device memory
// (the current device); called once per device at set-up.
hipError_t upload_trig_table();

hipError_t launch_philox_probe8(hipStream_t st, uint64_t seed, uint32_t c0, uint32_t c1,
                                uint32_t c3, int32_t count, float* normals);
hipError_t launch_philox_probe(hipStream_t st, uint64_t seed, uint32_t c0, uint32_t c1,
                               uint32_t c3, int32_t count, uint32_t* words, float* normals);

}  // namespace qk
