// qk_lsmc.hpp - Longstaff-Schwartz (American) batch pricing on gfx950.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace qk {

constexpr int kLsmcBlock = 256;    // threads per workgroup
// paths per lane per tile: 32 measured best with the sweep's 16-byte
// accesses (19.6 ms for 1024 options x 100k paths x 50 dates, against 20.0,
// 20.3, 20.7 and 26.0 ms at 48, 64, 16 and 8)
constexpr int kLsmcPerLane = 32;
constexpr int kLsmcTile = kLsmcBlock * kLsmcPerLane;

// One option of an LSMC launch (host-prepared).
struct alignas(16) LsmcRec {
    int32_t out_index;  // position in the caller's output arrays
    int32_t ot, paths, steps;
    uint32_t seed_lo, seed_hi;
    int64_t path_off;    // first element of this option's [date][path] matrix
    int64_t cf_off;      // first element of this option's cashflow vector
    int64_t stream_off;  // first draw of this option's host stream (validation modes)
    double spot, strike;
    double disc;   // e^{-r dt}
    double drift;  // (r - q - vol^2/2) dt
    double vs;     // vol sqrt(dt)
};

// Host entry: validates like longstaff_schwartz.cpp:50-55, resolves the
// degenerate branch, then simulates on the GPU.  basis is QK_LSMC_BASIS_*.
int32_t run_lsmc_batch(const double* spot, const double* strike, const double* t,
                       const double* vol, const double* r, const double* q, const int32_t* ot,
                       const int32_t* paths, const int32_t* steps, const uint64_t* seed,
                       int32_t basis, int32_t n, double* out_price, double* out_se,
                       std::string* err);

// Doubles per (option, tile) partial record of a basis.
int lsmc_record_doubles(int32_t basis);

// Launches forward, (solve, sweep) x (max_steps - 1) and final for one chunk
// (qk_lsmc.hip).  mode is a SimMode, basis QK_LSMC_BASIS_*; part holds
// count * tiles_max records, beta count * 5 doubles.
hipError_t launch_lsmc(hipStream_t st, int32_t mode, int32_t basis, const LsmcRec* recs,
                       int32_t count, int32_t tiles_max, int32_t max_steps, void* pathbuf,
                       void* cfbuf, const void* stream, double* part, double* beta,
                       double* out_price, double* out_se);

}  // namespace qk
