// qk_engine.hpp - host engine behind the C ABI: device context, buffers,
// validation-mode normal streams and the launch sequence of one MC batch.
#pragma once

#include <stdint.h>

#include <string>

#include "qk_kernels.hpp"

namespace qk {

// Host-memory SoA batch (length n).  Null arrays are unused by the method.
struct HostBatch {
    const int32_t* model_ids = nullptr;  // QK_MODEL_* per element (mixed batches)
    int32_t uniform_method = M_HESTON;   // Method when model_ids is null
    const double *spot = nullptr, *strike = nullptr, *t = nullptr, *vol = nullptr, *r = nullptr,
                 *q = nullptr;
    const double *v0 = nullptr, *kappa = nullptr, *theta = nullptr, *sigma = nullptr,
                 *rho = nullptr;
    const int32_t *ot = nullptr, *paths = nullptr, *steps = nullptr;
    const uint64_t* seed = nullptr;
    const double* aux = nullptr;  // shift (importance sampling) / weight clip (LR deltas) / MLMC level
    // validation modes: element i continues element i-1's host stream (the
    // levels of one MLMC option draw from one generator, level by level)
    const uint8_t* stream_continues = nullptr;
};

// Runs one batch on the GPU.  Returns a QK_* code; on failure *err names the
// cause.  out_se may be null.
int32_t run_mc_batch(const HostBatch& hb, int32_t n, double* out_price, double* out_se,
                     std::string* err);

// Same with device-resident inputs/outputs (QK_RNG_PHILOX only).
int32_t run_mc_batch_device(const BatchIn& in, int32_t n, double* out_price_dev,
                            double* out_se_dev, std::string* err);

// Limits of the allocation-free variant's fixed device scratch.
constexpr int32_t kNoAllocMaxOptions = 65535;
constexpr int32_t kNoAllocMaxTiles = 1 << 19;  // 2^19 tiles x 64-byte partials = 32 MiB

// run_mc_batch_device without any device or pinned allocation: its work
// buffers live in a fixed device array of the library (about 44.5 MiB per
// device).  Batches of more than kNoAllocMaxOptions options, or that need
// more than kNoAllocMaxTiles path tiles, are rejected.  Launches directly
// (no launch graph).
int32_t run_mc_batch_device_noalloc(const BatchIn& in, int32_t n, double* out_price_dev,
                                    double* out_se_dev, std::string* err);

struct Device;

// The device MC pipeline on device-resident inputs; the caller holds d->mu
// and has sized the per-option buffers (ensure_mc_capacity).  Synchronises
// once to size the tile partials; leaves the finalize queued on d->stream.
int32_t mc_device_locked(Device* d, const BatchIn& in, int32_t n, const int32_t* n_dev,
                         double* out_price_dev, double* out_se_dev, std::string* err);
int32_t ensure_mc_capacity(Device* d, int32_t n, std::string* err);
// Packs a host batch into the device staging buffer (one H2D copy, queued).
// With bytes_out set, the copy is not queued and its size is returned
// (d->pin_in -> d->in), for callers that capture it in a graph.
int32_t stage_host_batch(Device* d, const HostBatch& hb, int32_t n, BatchIn* bi,
                         const int32_t** n_dev, std::string* err, size_t* bytes_out = nullptr);

// Process-wide RNG mode (QK_RNG_*).
void set_rng_mode(int32_t mode);
int32_t get_rng_mode();

// The validation modes' normal stream for one seed (mt19937_64 +
// std::normal_distribution<double>, as mc_engine.h:15-23).
void host_normal_stream(uint64_t seed, int64_t count, double* out);

// Device Philox words and FP32 normals of blocks (c0+i, c1, 0, c3), i < count.
// The eight four-pair normals of each of `count` blocks (c0 + i, c1, 0, c3).
int32_t philox_probe8(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c3, int32_t count,
                      float* normals, std::string* err);
int32_t philox_probe(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c3, int32_t count,
                     uint32_t* words, float* normals, std::string* err);

// Tile shape of a family (FAM_TERMINAL / FAM_STEPWISE): paths (stepwise) or
// Philox groups (terminal) per lane, and persistent workgroups per CU.
int32_t set_shape(int32_t family, int32_t lane_units, int32_t grid_per_cu, std::string* err);
int32_t get_shape(int32_t family, int32_t* lane_units, int32_t* grid_per_cu, std::string* err);
int32_t autotune(int32_t family, int32_t* best_lane, int32_t* best_grid, double* best_ms,
                 std::string* err);

// Graph switch (enable < 0 leaves it) and counters: cached graphs (-1 when
// capture is unsupported) and graph launches so far.
int32_t set_graphs(int32_t enable, int32_t* cached, int64_t* launches, std::string* err);

// Event timings of the last completed batch.
bool last_timing(double* ms_total, double* ms_simulate);

// Device name / CU count; false if no usable device.
bool device_info(std::string* name, int32_t* cus, std::string* err);

}  // namespace qk
