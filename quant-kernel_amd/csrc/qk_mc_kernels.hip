// qk_mc_kernels.hip - gfx950 kernels of the Monte Carlo batch-pricing path.
//
// Pipeline per batch (all on one stream):
//   prepare  : validate every element and build its OptionRec (degenerate
//              branches resolved here; the rest get a tile count);
//   scan     : one workgroup turns tile counts into tile offsets and resets
//              the persistent tile counter;
//   simulate : persistent workgroups pull tiles from an atomic counter; a tile
//              is kBlock lanes x lane_units paths of one option; each lane
//              simulates whole paths (one path per lane at a time) and the
//              tile's centred payoff sums are reduced over the wave64 with
//              shuffles, then over the 4 waves in LDS, in FP64, in a fixed
//              order;
//   finalize : per option, a compensated (Neumaier) FP64 sum of its tiles'
//              partials, discounting and the standard error.
// Results do not depend on which workgroup ran which tile, so repeated runs
// are bitwise identical.
//
// Reference loops restated: mc_engine.h:32-53 (antithetic terminal),
// mc_engine.h:84-103 (stepwise), mc_engine.h:113-136 (2-D stepwise),
// local_vol_monte_carlo.cpp:20-48; step models model_concepts.h:93-266.
#include <cmath>

#include "qk_common.hpp"
#include "qk_kernels.hpp"
#include "qk_rng.hpp"

namespace qk {

// ----------------------------------------------------------------------------
// precision helpers
// ----------------------------------------------------------------------------
__device__ __forceinline__ float q_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double q_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float q_exp2(float x) { return __builtin_amdgcn_exp2f(x); }
__device__ __forceinline__ double q_exp2(double x) { return exp2(x); }
__device__ __forceinline__ float q_log2(float x) { return __builtin_amdgcn_logf(x); }
__device__ __forceinline__ double q_log2(double x) { return log2(x); }
__device__ __forceinline__ float q_fma(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double q_fma(double a, double b, double c) { return __builtin_fma(a, b, c); }
// Inputs are validated finite, so IEEE maxNum and the reference's
// std::max agree on every value that can occur here.
__device__ __forceinline__ float q_max(float a, float b) { return __builtin_fmaxf(a, b); }
__device__ __forceinline__ double q_max(double a, double b) { return __builtin_fmax(a, b); }
__device__ __forceinline__ float q_min(float a, float b) { return __builtin_fminf(a, b); }
__device__ __forceinline__ double q_min(double a, double b) { return __builtin_fmin(a, b); }

// Likelihood-ratio score z / den clipped to [-clip, clip]
// (likelihood_ratio.cpp:35-36, heston_lr_delta.cpp:88-91).
template <typename R>
__device__ __forceinline__ R lr_score(R z, R den, R inv_den, R clip) {
    // FP64 keeps the reference's division; FP32 multiplies by the reciprocal
    R sc;
    if constexpr (sizeof(R) == 8) sc = z / den;
    else sc = z * inv_den;
    return q_max(-clip, q_min(clip, sc));
}

template <typename R>
__device__ __forceinline__ R payoff(R s, R k, int32_t ot) {
    const R d = ot == QK_CALL ? s - k : k - s;
    return d > R(0) ? d : R(0);
}

// ----------------------------------------------------------------------------
// normal sources
// ----------------------------------------------------------------------------
template <typename R>
struct Quad {
    R z[4];
};
template <typename R, int N>
struct Draws {
    R z[N];
};

// Production: Philox blocks keyed by the option seed (wave-uniform key).
struct PhiloxSrc {
    uint32_t k0, k1;
    uint32_t tag = 0;               // fourth counter word of path blocks (the MLMC level)
    const float2* trig = nullptr;   // LDS angle table of the four-pair layout (terminal draws)
    __device__ __forceinline__ Quad<float> path_block(uint32_t p, uint32_t b, int32_t) const {
        const Normal4 n = philox_normals(p, b, tag, k0, k1);
        return Quad<float>{{n.z0, n.z1, n.z2, n.z3}};
    }
    // uniforms k = 4g .. 4g+3: (w + 0.5) 2^-32 from block (g, 0, 0, 2)
    __device__ __forceinline__ void term_uniforms(uint32_t g, double u[4]) const {
        const U4 w = philox4x32_10(g, 0u, 2u, k0, k1);
        u[0] = (double(w.x) + 0.5) * 0x1p-32;
        u[1] = (double(w.y) + 0.5) * 0x1p-32;
        u[2] = (double(w.z) + 0.5) * 0x1p-32;
        u[3] = (double(w.w) + 0.5) * 0x1p-32;
    }
    // normals k = 8b .. 8b+7 of a terminal estimator: the four pairs of
    // block (b, 0, 0, 1)
    __device__ __forceinline__ Draws<float, 8> term_normals8(uint32_t b) const {
        const Normal8 n = philox_normals8(philox4x32_10(b, 0u, 1u, k0, k1), trig);
        return Draws<float, 8>{{n.z[0], n.z[1], n.z[2], n.z[3], n.z[4], n.z[5], n.z[6], n.z[7]}};
    }
    __device__ __forceinline__ float term_draw(uint32_t k) const {
        const Draws<float, 8> q = term_normals8(k / 8u);
        return q.z[k % 8u];
    }
};

// Validation: the host's per-option stream laid out draw-major
// (draw d of path p at off + d*paths + p) so a wave's loads coalesce.
template <typename R>
struct StreamSrc {
    const R* base;
    int64_t paths;
    int64_t ndraws;  // draws per path
    __device__ __forceinline__ Quad<R> path_block(uint32_t p, uint32_t b, int32_t) const {
        Quad<R> q;
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const int64_t d = int64_t(b) * 4 + c;
            q.z[c] = d < ndraws ? base[d * paths + p] : R(0);
        }
        return q;
    }
    // draws k = N g .. N g + N-1
    template <int N>
    __device__ __forceinline__ Draws<R, N> term_draws(uint32_t g) const {
        Draws<R, N> q;
#pragma unroll
        for (int c = 0; c < N; ++c) {
            const int64_t k = int64_t(g) * N + c;
            q.z[c] = k < ndraws ? base[k] : R(0);
        }
        return q;
    }
    __device__ __forceinline__ R term_draw(uint32_t k) const { return base[k]; }
};

// ----------------------------------------------------------------------------
// angle table of the four-pair Philox layout: entry a = (cos, sin) of
// 2 pi a / 512, FP64 values rounded to FP32 on the host (upload_trig_table)
// ----------------------------------------------------------------------------
__device__ __align__(16) float2 g_trig[kTrigEntries];

// ----------------------------------------------------------------------------
// prepare / scan
// ----------------------------------------------------------------------------
__device__ __forceinline__ int32_t model_to_method(int32_t m) {
    switch (m) {
        case QK_MODEL_GBM_EULER: return M_EULER;
        case QK_MODEL_GBM_MILSTEIN: return M_MILSTEIN;
        case QK_MODEL_HESTON_EULER: return M_HESTON;
        case QK_MODEL_LOCAL_VOL: return M_LOCAL_VOL;
        default: return -1;
    }
}

// The batch size comes from n_dev when it is set (a count produced on the
// device), else from n_host; with n_out set, prepare stores it there for the
// later kernels of the batch, so a host-sized batch needs no copy of it.
__global__ __launch_bounds__(256) void k_prepare(BatchIn in, const int32_t* __restrict__ n_dev,
                                                 int32_t n_host, int32_t* __restrict__ n_out,
                                                 const int64_t* __restrict__ stream_off,
                                                 int32_t tg_terminal, int32_t tu_stepwise,
                                                 OptionRec* __restrict__ recs,
                                                 int32_t* __restrict__ ntiles) {
    const int32_t n = n_dev ? *n_dev : n_host;
    if (n_out && blockIdx.x == 0 && threadIdx.x == 0) *n_out = n;
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        RawOption o;
        int32_t m = in.method ? in.method[i] : in.uniform_method;
        if (in.method && in.method_is_model) m = model_to_method(m);
        o.method = m;
        o.spot = in.spot[i];
        o.strike = in.strike[i];
        o.t = in.t[i];
        o.vol = in.vol ? in.vol[i] : 0.0;
        o.r = in.r[i];
        o.q = in.q[i];
        o.v0 = in.v0 ? in.v0[i] : 0.0;
        o.kappa = in.kappa ? in.kappa[i] : 0.0;
        o.theta = in.theta ? in.theta[i] : 0.0;
        o.sigma = in.sigma ? in.sigma[i] : 0.0;
        o.rho = in.rho ? in.rho[i] : 0.0;
        o.ot = in.ot[i];
        o.paths = in.paths_override > 0 ? in.paths_override : in.paths[i];
        o.steps = in.steps ? in.steps[i] : 1;
        o.seed = in.seed[i];
        o.aux = in.aux ? in.aux[i] : 0.0;
        OptionRec rec;
        if (m < 0 || (in.active && !in.active[i])) {
            prepare_option(o, 1, rec);  // fills defaults
            rec.family = FAM_NONE;
            rec.closed = qnan_d();
            rec.ntiles = 0;
        } else {
            prepare_option(o, method_is_terminal(m) ? tg_terminal * term_group_units(m) : tu_stepwise,
                           rec);
        }
        if (rec.family == FAM_NONE) rec.ntiles = 0;
        if (stream_off) rec.stream_off = stream_off[i];
        recs[i] = rec;
        ntiles[i] = rec.ntiles;
    }
}

// Exclusive scan of the tile counts in three passes: per-1024-element block
// totals, one workgroup scanning the block totals (and resetting the tile
// counter), then per-block offsets.  Sums are 64-bit; a total beyond 2^31-1
// sets the overflow flag and schedules no tile.
constexpr int kScanBlock = 1024;

__device__ __forceinline__ long long block_inclusive_scan(long long v, long long* sh) {
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int off = 1; off < kScanBlock; off <<= 1) {
        const long long add = tid >= off ? sh[tid - off] : 0;
        __syncthreads();
        sh[tid] += add;
        __syncthreads();
    }
    return sh[tid];
}

__global__ __launch_bounds__(kScanBlock) void k_scan_blocks(const int32_t* __restrict__ n_dev,
                                                            const int32_t* __restrict__ ntiles,
                                                            long long* __restrict__ block_sum) {
    __shared__ long long sh[kScanBlock];
    const int32_t n = *n_dev;
    const int32_t i = blockIdx.x * kScanBlock + threadIdx.x;
    const long long v = i < n ? ntiles[i] : 0;
    const long long inc = block_inclusive_scan(v, sh);
    if (threadIdx.x == kScanBlock - 1) block_sum[blockIdx.x] = inc;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_top(int32_t nblocks,
                                                         long long* __restrict__ block_sum,
                                                         int32_t* __restrict__ counters,
                                                         int32_t capacity) {
    __shared__ long long sh[kScanBlock];
    long long carry = 0;
    for (int32_t base = 0; base < nblocks; base += kScanBlock) {
        const int32_t i = base + threadIdx.x;
        const long long v = i < nblocks ? block_sum[i] : 0;
        const long long inc = block_inclusive_scan(v, sh);
        if (i < nblocks) block_sum[i] = carry + inc - v;  // exclusive block offset
        carry += sh[kScanBlock - 1];
        __syncthreads();
    }
    // counters: [0] tile counter, [1] tiles to run, [2] 0 ok / 1 more tiles than
    // the partials hold / 2 more than 2^31-1, [3] tiles needed (clamped)
    if (threadIdx.x == 0) {
        const bool too_big = carry > 0x7fffffffLL;
        const bool over = too_big || carry > capacity;
        counters[0] = 0;
        counters[1] = over ? 0 : int32_t(carry);
        counters[2] = too_big ? 2 : (over ? 1 : 0);
        counters[3] = too_big ? 0x7fffffff : int32_t(carry);
    }
}

__global__ __launch_bounds__(kScanBlock) void k_scan_apply(const int32_t* __restrict__ n_dev,
                                                           const int32_t* __restrict__ ntiles,
                                                           const long long* __restrict__ block_off,
                                                           const int32_t* __restrict__ counters,
                                                           OptionRec* __restrict__ recs,
                                                           int32_t* __restrict__ tile_start) {
    __shared__ long long sh[kScanBlock];
    const int32_t n = *n_dev;
    const int32_t i = blockIdx.x * kScanBlock + threadIdx.x;
    const long long v = i < n ? ntiles[i] : 0;
    const long long inc = block_inclusive_scan(v, sh);
    const bool overflow = counters[2] != 0;
    if (i < n) {
        const int32_t off = overflow ? 0 : int32_t(block_off[blockIdx.x] + inc - v);
        tile_start[i] = off;
        recs[i].tile_first = off;
    }
    if (i == n) tile_start[n] = counters[1];
}

// One-pass scan for batches of at most kScanSmallBlocks blocks: each block
// sums the tile counts before its own range, and of the whole batch, itself
// (at most 64 counts per thread), which costs less than the two launches of
// the three-pass scan it replaces.  Offsets, tile_start[n] and the counters
// are the three-pass scan's.
constexpr int kScanSmallBlocks = 64;

__global__ __launch_bounds__(kScanBlock) void k_scan_small(const int32_t* __restrict__ n_dev,
                                                           const int32_t* __restrict__ ntiles,
                                                           int32_t* __restrict__ counters,
                                                           int32_t capacity,
                                                           OptionRec* __restrict__ recs,
                                                           int32_t* __restrict__ tile_start) {
    __shared__ long long sh[kScanBlock];
    const int32_t n = *n_dev;
    const int32_t base = blockIdx.x * kScanBlock;
    long long before = 0, all = 0;
    for (int32_t k = threadIdx.x; k < n; k += kScanBlock) {
        const long long v = ntiles[k];
        all += v;
        if (k < base) before += v;
    }
    block_inclusive_scan(before, sh);
    before = sh[kScanBlock - 1];
    __syncthreads();
    block_inclusive_scan(all, sh);
    all = sh[kScanBlock - 1];
    __syncthreads();
    const bool too_big = all > 0x7fffffffLL;
    const bool over = too_big || all > capacity;
    const int32_t i = base + threadIdx.x;
    const long long v = i < n ? ntiles[i] : 0;
    const long long inc = block_inclusive_scan(v, sh);
    if (i < n) {
        const int32_t off = over ? 0 : int32_t(before + inc - v);
        tile_start[i] = off;
        recs[i].tile_first = off;
    }
    if (i == n) tile_start[n] = over ? 0 : int32_t(all);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        counters[0] = 0;
        counters[1] = over ? 0 : int32_t(all);
        counters[2] = too_big ? 2 : (over ? 1 : 0);
        counters[3] = too_big ? 0x7fffffff : int32_t(all);
    }
}

// ----------------------------------------------------------------------------
// block reduction of (s1, s2, extra) in FP64, fixed order
// ----------------------------------------------------------------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ void block_store_partial(double s1, double s2, double extra,
                                                    Partial* out, double (*red)[5]) {
    s1 = wave_sum(s1);
    s2 = wave_sum(s2);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) {
        red[wave][0] = s1;
        red[wave][1] = s2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a = 0.0, b = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) {
            a += red[w][0];
            b += red[w][1];
        }
        Partial p;
        p.s1 = a;
        p.s2 = b;
        p.extra = extra;
        p.x1 = p.x2 = p.xy = 0.0;
        p.pad[0] = p.pad[1] = 0.0;
        *out = p;
    }
    __syncthreads();
}

// Control variates: the payoff sums and the control's sums, same order.
__device__ __forceinline__ void block_store_partial_cv(double s1, double s2, double x1, double x2,
                                                       double xy, Partial* out,
                                                       double (*red)[5]) {
    double v[5] = {s1, s2, x1, x2, xy};
#pragma unroll
    for (int i = 0; i < 5; ++i) v[i] = wave_sum(v[i]);
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < 5; ++i) red[wave][i] = v[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double a[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        for (int w = 0; w < kBlock / 64; ++w)
            for (int i = 0; i < 5; ++i) a[i] += red[w][i];
        Partial p;
        p.s1 = a[0];
        p.s2 = a[1];
        p.extra = 0.0;
        p.x1 = a[2];
        p.x2 = a[3];
        p.xy = a[4];
        p.pad[0] = p.pad[1] = 0.0;
        *out = p;
    }
    __syncthreads();
}

// Largest option index whose first tile is <= tile (tile_start is
// non-decreasing; empty options share their successor's start).
__device__ __forceinline__ int32_t find_option(const int32_t* __restrict__ tile_start, int32_t n,
                                               int32_t tile) {
    int32_t lo = 0, hi = n;
    while (hi - lo > 1) {
        const int32_t mid = (lo + hi) >> 1;
        if (tile_start[mid] <= tile) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Persistent workgroups pull tiles from counters[0].  Every workgroup takes
// one index past the last tile it runs, so the counter can pass 2^31 - 1
// when the tile count is close to it: callers compare tiles as unsigned, and
// a wrapped index reads as past the end.  The scan of each launch resets it.
__device__ __forceinline__ int32_t next_tile(int32_t* counters, int32_t* slot) {
    if (threadIdx.x == 0) *slot = atomicAdd(&counters[0], 1);
    __syncthreads();
    return __builtin_amdgcn_readfirstlane(*slot);
}

// ----------------------------------------------------------------------------
// path loops
// ----------------------------------------------------------------------------

// GBM Euler / local vol (Milstein when MIL): s <- max(1e-12, s*(c0 + c1 z + c2 z^2)),
// the factored form of model_concepts.h:136-158 / 257-266.
template <typename R, bool MIL, class Src>
__device__ __forceinline__ R gbm_path(const Src& src, uint32_t p, int32_t steps, R s, R c0, R c1,
                                      R c2) {
    const R floor_s = R(1e-12);
    const int32_t nb = steps >> 2;
    for (int32_t b = 0; b < nb; ++b) {
        const Quad<R> z = src.path_block(p, uint32_t(b), steps);
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            const R f = MIL ? q_fma(q_fma(c2, z.z[c], c1), z.z[c], c0) : q_fma(c1, z.z[c], c0);
            s = q_max(floor_s, s * f);
        }
    }
    const int32_t rem = steps & 3;
    if (rem) {
        const Quad<R> z = src.path_block(p, uint32_t(nb), steps);
        for (int c = 0; c < rem; ++c) {
            const R f = MIL ? q_fma(q_fma(c2, z.z[c], c1), z.z[c], c0) : q_fma(c1, z.z[c], c0);
            s = q_max(floor_s, s * f);
        }
    }
    return s;
}

// One MLMC correction level (multilevel_monte_carlo.cpp:37-61): per coarse
// step two fine Euler steps on (z1, z2) and one coarse step of twice the
// length on (z1 + z2)/sqrt 2, both in the factored form s (c + c' z) with
// the 1e-12 floor.  steps (fine) is even.
template <typename R, class Src>
__device__ __forceinline__ void mlmc_pair_path(const Src& src, uint32_t p, int32_t steps, R s0,
                                               R c0, R c1, R c2, R c3, R& sf, R& sc) {
    const R fl = R(1e-12), rs2 = R(0.70710678118654752440);
    R f = s0, c = s0;
    auto coarse_step = [&](R z1, R z2) {
        f = q_max(fl, f * q_fma(c1, z1, c0));
        f = q_max(fl, f * q_fma(c1, z2, c0));
        c = q_max(fl, c * q_fma(c3, (z1 + z2) * rs2, c2));
    };
    const int32_t nb = steps >> 2;  // one block of 4 draws = two coarse steps
    for (int32_t b = 0; b < nb; ++b) {
        const Quad<R> z = src.path_block(p, uint32_t(b), steps);
        coarse_step(z.z[0], z.z[1]);
        coarse_step(z.z[2], z.z[3]);
    }
    if (steps & 3) {
        const Quad<R> z = src.path_block(p, uint32_t(nb), steps);
        coarse_step(z.z[0], z.z[1]);
    }
    sf = f;
    sc = c;
}

// Heston full-truncation Euler in x = log2(S/spot) (model_concepts.h:205-226):
//   x <- max(x + (r-q-v/2)dt log2e + sqrt(v) sqrt(dt) log2e z1, log2(1e-12/spot))
//   v <- max(0, v + kappa(theta - v)dt + sigma sqrt(v) sqrt(dt) (rho z1 + rho' z2))
// v >= 0 holds after every step and v0 >= 0 is validated, so max(v,0) = v.
template <typename R>
struct HestonC {
    R a, h, c, kdt, kth, sg, rho, rc, xf, v0;
};

template <typename R>
__device__ __forceinline__ void heston_step(const HestonC<R>& k, R& x, R& v, R z1, R z2) {
    const R sv = q_sqrt(v);
    const R drift = q_fma(-k.h, v, k.a);
    x = q_fma(k.c * sv, z1, x + drift);
    x = q_max(x, k.xf);
    const R mix = q_fma(k.rho, z1, k.rc * z2);
    const R vn = q_fma(k.sg * sv, mix, q_fma(R(1) - k.kdt, v, k.kth));
    v = q_max(vn, R(0));
}

template <typename R, class Src>
__device__ __forceinline__ R heston_path(const Src& src, uint32_t p, int32_t steps,
                                         const HestonC<R>& k) {
    R x = R(0), v = k.v0;
    const int32_t nb = steps >> 1;
    for (int32_t b = 0; b < nb; ++b) {
        const Quad<R> z = src.path_block(p, uint32_t(b), steps);
        heston_step(k, x, v, z.z[0], z.z[1]);
        heston_step(k, x, v, z.z[2], z.z[3]);
    }
    if (steps & 1) {
        const Quad<R> z = src.path_block(p, uint32_t(nb), steps);
        heston_step(k, x, v, z.z[0], z.z[1]);
    }
    return x;
}

// Production Heston path: the scheme of heston_path on the Philox stream,
// rearranged to issue fewer instructions.  With Box-Muller radius
// r = sqrt(2 ln 2) * sqrt(-log2 u1) and angle theta,
//   z1               = r cos(theta)
//   rho z1 + rho' z2 = r (rho cos(theta) + rho' sin(theta))
// and the log-spot is kept in units of c sqrt(2 ln 2) (c = sqrt(dt) log2e),
// so per step: y <- max(y + (a - h v)/cK + sqrt(v) r' cos(theta), floor/cK),
// v <- max(0, (1 - kappa dt) v + kappa theta dt + sqrt(v) r' (A cos(theta) +
// B sin(theta))), with r' = sqrt(-log2 u1), A = sigma sqrt(dt) sqrt(2 ln 2) rho
// and B the same with rho' = sqrt(1 - rho^2).
//
// Heston stream (four-pair layout): Philox block (p, j/4, 0, 0) drives
// steps 4(j/4) .. 4(j/4)+3 of path p; step j takes word s = j%4 alone:
//   u1 = (2m + 1) 2^-24, m = low 23 bits of w_s (exact in FP32, in (0,1));
//   theta = 2 pi a / 512, a = top 9 bits of w_s, read as (cos, A cos + B sin)
//   from the LDS table.
// The radius is capped at sqrt(48 ln 2) = 5.77 sigma.  The angle takes 512
// equally spaced values: E[cos^n theta] and E[sin^n theta] equal their
// values under a uniform angle for every n < 512, so every moment of
// (z1, z2) below order 512 is that of the Gaussian pair the 23-bit radius
// gives.  One block per four steps (instead of two with four full
// Box-Muller normals) quarters the Philox work per step; the table replaces
// the two cosines of a step, whose transcendental issue cost three FMAs each.
struct HestonPx {
    float a, h, omk, kth, A, B, yf, v0, ck;
};

// max(x, f) for finite x, f as v_med3_f32(x, f, FLT_MAX): unlike fmaxf it
// needs no canonicalisation of the loop-invariant floor inside the loop.
__device__ __forceinline__ float floor_max(float x, float f) {
    return __builtin_amdgcn_fmed3f(x, f, 0x1.fffffep+127f);
}

// sqrt(v) r' = sqrt(-v log2 u1): one square root per step instead of two.
// cm = (cos(theta), A cos(theta) + B sin(theta)) from the Heston table.
__device__ __forceinline__ void heston_px_step(const HestonPx& k, float& y, float& v, float u1,
                                               float2 cm) {
    const float lg = __builtin_amdgcn_logf(u1);  // < 0
    const float t = __builtin_amdgcn_sqrtf(-(v * lg));
    y = floor_max(__builtin_fmaf(t, cm.x, y + __builtin_fmaf(-k.h, v, k.a)), k.yf);
    v = __builtin_fmaxf(__builtin_fmaf(t, cm.y, __builtin_fmaf(k.omk, v, k.kth)), 0.0f);
}

// The stepwise kernel's LDS angle table (512 entries, the angles of the
// four-pair layout) holds (cos, sin) for GBM tiles and, for Heston tiles,
// (cos, fma(B, sin, A cos)) with the tile option's A and B: the variance
// step's correlated term costs one lookup instead of a multiply and an FMA
// per step (the entry is rounded exactly as that per-step expression would
// be).  Rebuilt from the device table (global memory) only when a tile
// needs other contents than the last one; every thread of the workgroup
// calls it with the same arguments.
struct TrigTableState {
    int32_t kind;  // 0: (cos, sin); 1: Heston (cos, m) for (A, B)
    float A, B;
};

__device__ __forceinline__ void retable(float2* s_trig, TrigTableState* st, bool heston,
                                        float A, float B) {
    const int32_t want = heston ? 1 : 0;
    if (st->kind == want && (!heston || (st->A == A && st->B == B))) return;
    __syncthreads();  // every thread has read st and finished with the old contents
    for (int i = threadIdx.x; i < kTrigEntries; i += kBlock) {
        const float2 c = g_trig[i];  // angle 2 pi i / 512
        s_trig[i] = heston ? make_float2(c.x, __builtin_fmaf(B, c.y, A * c.x)) : c;
    }
    if (threadIdx.x == 0) {
        st->kind = want;
        st->A = A;
        st->B = B;
    }
    __syncthreads();
}

__device__ __forceinline__ float heston_path_px(uint32_t p, int32_t steps, const HestonPx& k,
                                                uint32_t k0, uint32_t k1,
                                                const float2* trig) {
    float y = 0.0f, v = k.v0;
    const int32_t nb = steps / 4, rem = steps - 4 * nb;
    for (int32_t b = 0; b < nb; ++b) {
        const U4 w = philox4x32_10(p, uint32_t(b), 0u, k0, k1);
        const float2 c0 = trig_at(trig, quad_offset(w.x)), c1 = trig_at(trig, quad_offset(w.y)),
                     c2 = trig_at(trig, quad_offset(w.z)), c3 = trig_at(trig, quad_offset(w.w));
        heston_px_step(k, y, v, u23(w.x), c0);
        heston_px_step(k, y, v, u23(w.y), c1);
        heston_px_step(k, y, v, u23(w.z), c2);
        heston_px_step(k, y, v, u23(w.w), c3);
    }
    if (rem) {
        const U4 w = philox4x32_10(p, uint32_t(nb), 0u, k0, k1);
        heston_px_step(k, y, v, u23(w.x), trig_at(trig, quad_offset(w.x)));
        if (rem > 1) heston_px_step(k, y, v, u23(w.y), trig_at(trig, quad_offset(w.y)));
        if (rem > 2) heston_px_step(k, y, v, u23(w.z), trig_at(trig, quad_offset(w.z)));
    }
    return y * k.ck;  // log2(S_T / spot)
}

// Production GBM Euler / Milstein / local-vol path: the scheme of gbm_path on
// the Heston stream's pairs, step j taking the cosine (even j) or sine (odd
// j) of the pair of Heston step j/2, so one Philox block drives eight steps.
// m = -2 ln 2 c1^2 for Euler and local vol folds the volatility into the
// radius (c1 z = sqrt(m log2 u1) cos theta); Milstein keeps z (m = -2 ln 2).
template <bool MIL>
__device__ __forceinline__ void gbm_px_pair(float& s, float u1, float2 cs, float c0, float c1,
                                            float c2, float m, bool both) {
    const float fl = 1e-12f;
    const float rad = __builtin_amdgcn_sqrtf(__builtin_amdgcn_logf(u1) * m);
    if constexpr (MIL) {
        const float z = rad * cs.x;
        s = q_max(fl, s * q_fma(q_fma(c2, z, c1), z, c0));
    } else {
        s = q_max(fl, s * q_fma(rad, cs.x, c0));
    }
    if (both) {
        if constexpr (MIL) {
            const float z = rad * cs.y;
            s = q_max(fl, s * q_fma(q_fma(c2, z, c1), z, c0));
        } else {
            s = q_max(fl, s * q_fma(rad, cs.y, c0));
        }
    }
}

template <bool MIL>
__device__ __forceinline__ float gbm_path_px(uint32_t p, int32_t steps, float s, float c0, float c1,
                                             float c2, uint32_t k0, uint32_t k1,
                                             const float2* trig) {
    const float m = MIL ? -1.3862943611198906f : -1.3862943611198906f * c1 * c1;
    const int32_t nb = steps / 8, rem = steps - 8 * nb;
    for (int32_t b = 0; b < nb; ++b) {
        const U4 w = philox4x32_10(p, uint32_t(b), 0u, k0, k1);
        const float2 a0 = trig_at(trig, quad_offset(w.x)), a1 = trig_at(trig, quad_offset(w.y)),
                     a2 = trig_at(trig, quad_offset(w.z)), a3 = trig_at(trig, quad_offset(w.w));
        gbm_px_pair<MIL>(s, u23(w.x), a0, c0, c1, c2, m, true);
        gbm_px_pair<MIL>(s, u23(w.y), a1, c0, c1, c2, m, true);
        gbm_px_pair<MIL>(s, u23(w.z), a2, c0, c1, c2, m, true);
        gbm_px_pair<MIL>(s, u23(w.w), a3, c0, c1, c2, m, true);
    }
    if (rem) {
        const U4 w = philox4x32_10(p, uint32_t(nb), 0u, k0, k1);
        gbm_px_pair<MIL>(s, u23(w.x), trig_at(trig, quad_offset(w.x)), c0, c1, c2, m, rem > 1);
        if (rem > 2)
            gbm_px_pair<MIL>(s, u23(w.y), trig_at(trig, quad_offset(w.y)), c0, c1, c2, m, rem > 3);
        if (rem > 4)
            gbm_px_pair<MIL>(s, u23(w.z), trig_at(trig, quad_offset(w.z)), c0, c1, c2, m, rem > 5);
        if (rem > 6)
            gbm_px_pair<MIL>(s, u23(w.w), trig_at(trig, quad_offset(w.w)), c0, c1, c2, m, false);
    }
    return s;
}

// Antithetic Heston pair of heston_lr_delta.cpp:67-96: the forward path and
// the path driven by the same draws negated; z1f is the forward path's first
// z1 (the likelihood-ratio score's draw).
template <typename R, class Src>
__device__ __forceinline__ void heston_pair(const Src& src, uint32_t p, int32_t steps,
                                            const HestonC<R>& k, R& xf, R& xa, R& z1f) {
    R x = R(0), v = k.v0, y = R(0), w = k.v0;
    const int32_t nb = steps >> 1;
    for (int32_t b = 0; b < nb; ++b) {
        const Quad<R> z = src.path_block(p, uint32_t(b), steps);
        if (b == 0) z1f = z.z[0];
        heston_step(k, x, v, z.z[0], z.z[1]);
        heston_step(k, y, w, -z.z[0], -z.z[1]);
        heston_step(k, x, v, z.z[2], z.z[3]);
        heston_step(k, y, w, -z.z[2], -z.z[3]);
    }
    if (steps & 1) {
        const Quad<R> z = src.path_block(p, uint32_t(nb), steps);
        if (nb == 0) z1f = z.z[0];
        heston_step(k, x, v, z.z[0], z.z[1]);
        heston_step(k, y, w, -z.z[0], -z.z[1]);
    }
    xf = x;
    xa = y;
}

// Production form of heston_pair: one Box-Muller evaluation drives both
// paths (negating both normals is negating the radius).  cs is a Heston
// table entry (cos, A cos + B sin).  lg and c1 return log2 u1 and
// cos(theta) of the step (z1 = sqrt(2 ln 2) sqrt(-lg) c1).
__device__ __forceinline__ void heston_px_pair_step(const HestonPx& k, float& y, float& v,
                                                    float& ya, float& va, float u1, float2 cs,
                                                    float& lg, float& c1) {
    lg = __builtin_amdgcn_logf(u1);
    c1 = cs.x;
    const float m = cs.y;  // A cos + B sin from the Heston table
    const float t = __builtin_amdgcn_sqrtf(-(v * lg));
    y = floor_max(__builtin_fmaf(t, c1, y + __builtin_fmaf(-k.h, v, k.a)), k.yf);
    v = __builtin_fmaxf(__builtin_fmaf(t, m, __builtin_fmaf(k.omk, v, k.kth)), 0.0f);
    const float ta = __builtin_amdgcn_sqrtf(-(va * lg));
    ya = floor_max(__builtin_fmaf(-ta, c1, ya + __builtin_fmaf(-k.h, va, k.a)), k.yf);
    va = __builtin_fmaxf(__builtin_fmaf(-ta, m, __builtin_fmaf(k.omk, va, k.kth)), 0.0f);
}

__device__ __forceinline__ void heston_pair_px(uint32_t p, int32_t steps, const HestonPx& k,
                                               uint32_t k0, uint32_t k1, const float2* trig,
                                               float& xf, float& xa, float& z1f) {
    float y = 0.0f, v = k.v0, ya = 0.0f, va = k.v0, lg = 0.0f, c1 = 0.0f, lg0 = 0.0f, c10 = 0.0f;
    const int32_t nb = steps / 4, rem = steps - 4 * nb;
    for (int32_t b = 0; b < nb; ++b) {
        const U4 w = philox4x32_10(p, uint32_t(b), 0u, k0, k1);
        heston_px_pair_step(k, y, v, ya, va, u23(w.x), trig_at(trig, quad_offset(w.x)), lg, c1);
        if (b == 0) {
            lg0 = lg;
            c10 = c1;
        }
        heston_px_pair_step(k, y, v, ya, va, u23(w.y), trig_at(trig, quad_offset(w.y)), lg, c1);
        heston_px_pair_step(k, y, v, ya, va, u23(w.z), trig_at(trig, quad_offset(w.z)), lg, c1);
        heston_px_pair_step(k, y, v, ya, va, u23(w.w), trig_at(trig, quad_offset(w.w)), lg, c1);
    }
    if (rem) {
        const U4 w = philox4x32_10(p, uint32_t(nb), 0u, k0, k1);
        heston_px_pair_step(k, y, v, ya, va, u23(w.x), trig_at(trig, quad_offset(w.x)), lg, c1);
        if (nb == 0) {
            lg0 = lg;
            c10 = c1;
        }
        if (rem > 1)
            heston_px_pair_step(k, y, v, ya, va, u23(w.y), trig_at(trig, quad_offset(w.y)), lg, c1);
        if (rem > 2)
            heston_px_pair_step(k, y, v, ya, va, u23(w.z), trig_at(trig, quad_offset(w.z)), lg, c1);
    }
    xf = y * k.ck;
    xa = ya * k.ck;
    z1f = __builtin_amdgcn_sqrtf(-lg0) * c10 * 1.1774100225154746910f;  // sqrt(2 ln 2)
}

// Heston constants of a record in both forms.
template <typename R>
__device__ __forceinline__ HestonC<R> heston_consts(const OptionRec& rec) {
    HestonC<R> k;
    k.a = R(rec.c[0]);
    k.h = R(rec.c[1]);
    k.c = R(rec.c[2]);
    k.kdt = R(rec.c[3]);
    k.kth = R(rec.c[4]);
    k.sg = R(rec.c[5]);
    k.rho = R(rec.c[6]);
    k.rc = R(rec.c[7]);
    k.xf = R(log2(1e-12 / rec.spot));
    k.v0 = R(rec.v0);
    return k;
}

__device__ __forceinline__ HestonPx heston_px_consts(const OptionRec& rec) {
    HestonPx kp;
    const double sq = 1.1774100225154746910;  // sqrt(2 ln 2)
    const double ck = rec.c[2] * sq;          // sqrt(dt) log2e sqrt(2 ln 2)
    kp.a = float(rec.c[0] / ck);
    kp.h = float(rec.c[1] / ck);
    kp.omk = float(1.0 - rec.c[3]);
    kp.kth = float(rec.c[4]);
    kp.A = float(rec.c[5] * sq * rec.c[6]);  // sigma sqrt(dt) sqrt(2 ln 2) rho
    kp.B = float(rec.c[5] * sq * rec.c[7]);  // ... rho'
    kp.yf = float(log2(1e-12 / rec.spot) / ck);
    kp.v0 = float(rec.v0);
    kp.ck = float(ck);
    return kp;
}

// ----------------------------------------------------------------------------
// stepwise simulation (Euler, Milstein, local vol, Heston)
// ----------------------------------------------------------------------------
template <typename R, int MODE>
__global__ __launch_bounds__(kBlock) void k_sim_stepwise(const OptionRec* __restrict__ recs,
                                                         const int32_t* __restrict__ tile_start,
                                                         const int32_t* __restrict__ n_dev,
                                                         int32_t* __restrict__ counters,
                                                         Partial* __restrict__ partials,
                                                         const R* __restrict__ stream,
                                                         int32_t lane_paths) {
    __shared__ int32_t s_tile;
    __shared__ double s_red[kBlock / 64][5];
    __shared__ __align__(16) float2 s_trig[MODE == SIM_PHILOX_F32 ? kTrigEntries : 1];
    __shared__ TrigTableState s_tts;
    const float2* trig = nullptr;
    if constexpr (MODE == SIM_PHILOX_F32) {
        if (threadIdx.x == 0) s_tts = TrigTableState{-1, 0.0f, 0.0f};  // filled by the first tile
        __syncthreads();
        trig = s_trig;
    }
    const int32_t n = *n_dev;
    const int32_t total = __builtin_amdgcn_readfirstlane(counters[1]);
    for (;;) {
        const int32_t tile = next_tile(counters, &s_tile);
        if (uint32_t(tile) >= uint32_t(total)) break;
        const int32_t opt = find_option(tile_start, n, tile);
        const OptionRec& rec = recs[opt];
        const int32_t method = rec.method, steps = rec.steps, paths = rec.paths, ot = rec.ot;
        const int64_t first = int64_t(tile - rec.tile_first) * kBlock * lane_paths;
        const bool heston_tile = method == M_HESTON || method == M_HESTON_LR;
        HestonPx kp{};
        if constexpr (MODE == SIM_PHILOX_F32) {
            if (heston_tile) kp = heston_px_consts(rec);
            retable(s_trig, &s_tts, heston_tile, kp.A, kp.B);
        }
        const R spot = R(rec.spot), strike = R(rec.strike);
        const double shift = rec.shift;

        StreamSrc<R> ss{stream + rec.stream_off, paths,
                        int64_t(steps) * (method == M_HESTON ? 2 : 1)};

        double s1 = 0.0, s2 = 0.0, extra = 0.0;
        if (method == M_HESTON_LR) {
            // units are antithetic pairs; an odd path count adds one forward
            // path (Philox path index / stream unit = pairs) on lane 0 of the
            // option's last tile
            const HestonC<R> k = heston_consts<R>(rec);
            const int32_t pairs = rec.units;
            StreamSrc<R> sp{stream + rec.stream_off, int64_t(pairs) + (paths & 1),
                            int64_t(steps) * 2};
            const R den = R(rec.aux[2]), inv_den = R(1.0 / rec.aux[2]), clip = R(rec.aux[0]);
            const bool no_score = rec.aux[2] == 0.0;  // sqrt(v0) <= 1e-15: score 0
            const bool last = tile - rec.tile_first == rec.ntiles - 1;
            for (int32_t j = 0; j <= lane_paths; ++j) {
                int64_t p = first + threadIdx.x + int64_t(j) * kBlock;
                if (j == lane_paths) {  // the odd path
                    if (!((paths & 1) && last && threadIdx.x == 0)) break;
                    p = pairs;
                } else if (p >= pairs) {
                    continue;
                }
                R xf, xa, z1;
                if constexpr (MODE == SIM_PHILOX_F32)
                    heston_pair_px(uint32_t(p), steps, kp, rec.seed_lo, rec.seed_hi, trig, xf, xa,
                                   z1);
                else heston_pair<R>(sp, uint32_t(p), steps, k, xf, xa, z1);
                const R sc = no_score ? R(0) : lr_score(z1, den, inv_den, clip);
                const R pf = payoff(spot * q_exp2(xf), strike, ot);
                if (p == pairs) {
                    extra = double(pf * sc);
                    break;
                }
                const R pa = payoff(spot * q_exp2(xa), strike, ot);
                const double f = double(pf * sc) - double(pa * sc);
                s1 += f;
                s2 = __builtin_fma(f, f, s2);
            }
        } else if (method == M_MLMC_LEVEL) {
            // level 0: plain Euler paths; level l >= 1: fine - coarse.  Philox
            // draws of level l come from blocks (p, d/4, 0, l).
            const int32_t level = int32_t(rec.aux[0]);
            const R c0 = R(rec.c[0]), c1 = R(rec.c[1]), c2 = R(rec.c[2]), c3 = R(rec.c[3]);
            PhiloxSrc pl{rec.seed_lo, rec.seed_hi, uint32_t(level)};
            for (int32_t j = 0; j < lane_paths; ++j) {
                const int64_t p = first + threadIdx.x + int64_t(j) * kBlock;
                if (p >= paths) break;
                double f;
                if (level == 0) {
                    R sp;
                    if constexpr (MODE == SIM_PHILOX_F32)
                        sp = gbm_path<R, false>(pl, uint32_t(p), steps, spot, c0, c1, R(0));
                    else sp = gbm_path<R, false>(ss, uint32_t(p), steps, spot, c0, c1, R(0));
                    f = double(payoff(sp, strike, ot)) - shift;
                } else {
                    R sf, sc;
                    if constexpr (MODE == SIM_PHILOX_F32)
                        mlmc_pair_path<R>(pl, uint32_t(p), steps, spot, c0, c1, c2, c3, sf, sc);
                    else mlmc_pair_path<R>(ss, uint32_t(p), steps, spot, c0, c1, c2, c3, sf, sc);
                    f = double(payoff(sf, strike, ot)) - double(payoff(sc, strike, ot));
                }
                s1 += f;
                s2 = __builtin_fma(f, f, s2);
            }
        } else if (method == M_HESTON) {
            const HestonC<R> k = heston_consts<R>(rec);
            for (int32_t j = 0; j < lane_paths; ++j) {
                const int64_t p = first + threadIdx.x + int64_t(j) * kBlock;
                if (p >= paths) break;
                R x;
                if constexpr (MODE == SIM_PHILOX_F32)
                    x = heston_path_px(uint32_t(p), steps, kp, rec.seed_lo, rec.seed_hi, trig);
                else x = heston_path<R>(ss, uint32_t(p), steps, k);
                const double f = double(payoff(spot * q_exp2(x), strike, ot)) - shift;
                s1 += f;
                s2 = __builtin_fma(f, f, s2);
            }
        } else {
            const R c0 = R(rec.c[0]), c1 = R(rec.c[1]), c2 = R(rec.c[2]);
            const bool mil = method == M_MILSTEIN;
            for (int32_t j = 0; j < lane_paths; ++j) {
                const int64_t p = first + threadIdx.x + int64_t(j) * kBlock;
                if (p >= paths) break;
                R s;
                if constexpr (MODE == SIM_PHILOX_F32) {
                    s = mil ? gbm_path_px<true>(uint32_t(p), steps, spot, c0, c1, c2, rec.seed_lo,
                                                rec.seed_hi, trig)
                            : gbm_path_px<false>(uint32_t(p), steps, spot, c0, c1, c2, rec.seed_lo,
                                                 rec.seed_hi, trig);
                } else {
                    s = mil ? gbm_path<R, true>(ss, uint32_t(p), steps, spot, c0, c1, c2)
                            : gbm_path<R, false>(ss, uint32_t(p), steps, spot, c0, c1, c2);
                }
                const double f = double(payoff(s, strike, ot)) - shift;
                s1 += f;
                s2 = __builtin_fma(f, f, s2);
            }
        }
        // extra is only non-zero on thread 0, which writes the partial
        block_store_partial(s1, s2, extra, partials + tile, s_red);
    }
}

// ----------------------------------------------------------------------------
// terminal GBM (mc_engine.h:32-75): S_T = spot * 2^(A + B z),
// A = (r-q-vol^2/2)T log2e, B = vol sqrt(T) log2e.  One unit per draw:
//   TK_PAIR  standard MC / antithetic variates: (pay(z) + pay(-z)) / 2
//   TK_LR    likelihood-ratio delta: (pay(z) - pay(-z)) clip(z / den) / 2
//            (likelihood_ratio.cpp:30-37; the mirror path's score is -score)
//   TK_IS    importance sampling: pay(z + h) e^{-h(z+h) + h^2/2}
//            (importance_sampling.cpp:26-34)
//   TK_PW    pathwise delta: 1[S_T > K] S_T / spot (call), -1[S_T < K] S_T / spot
//            (put) (pathwise_derivative.cpp:29-35)
// ----------------------------------------------------------------------------
enum TermKind : int32_t {
    TK_PAIR = 0, TK_LR = 1, TK_IS = 2, TK_PW = 3,
    TK_CV = 4,      // control variates: payoff and the control S_T - E[S_T]
    TK_STRAT = 5,   // stratified: u = (i + U_i) / n
    TK_SOBOL = 6,   // u = Sobol point i + 1
    TK_HALTON = 7,  // u = base-2 Halton point i + 1
};

// Acklam's inverse normal CDF (P. J. Acklam, 2003; relative error < 1.15e-9)
// in FP64, as monte_carlo_methods/common/internal_util.h:43-83 evaluates it.
__device__ double acklam(double p) {
    if (!(p > 0.0 && p < 1.0)) return qnan_d();
    const double lo = 0.02425, hi = 1.0 - lo;
    if (p < lo || p > hi) {
        const double qq = sqrt(-2.0 * log(p < lo ? p : 1.0 - p));
        const double v = (((((-7.784894002430293e-03 * qq + -3.223964580411365e-01) * qq +
                             -2.400758277161838e+00) * qq + -2.549732539343734e+00) * qq +
                           4.374664141464968e+00) * qq + 2.938163982698783e+00) /
                         ((((7.784695709041462e-03 * qq + 3.224671290700398e-01) * qq +
                            2.445134137142996e+00) * qq + 3.754408661907416e+00) * qq + 1.0);
        return p < lo ? v : -v;
    }
    const double qq = p - 0.5, rr = qq * qq;
    return (((((-3.969683028665376e+01 * rr + 2.209460984245205e+02) * rr +
               -2.759285104469687e+02) * rr + 1.383577518672690e+02) * rr +
             -3.066479806614716e+01) * rr + 2.506628277459239e+00) * qq /
           (((((-5.447609879822406e+01 * rr + 1.615858368580409e+02) * rr +
               -1.556989798598866e+02) * rr + 6.680131188771972e+01) * rr +
             -1.328068155288572e+01) * rr + 1.0);
}

// internal_util.h:96-122: the Sobol direction number of bit j is 1 << (31-j),
// so point i is the bit reverse of gray(i); base-2 Halton is the bit reverse.
__device__ __forceinline__ double sobol_point(uint32_t i) {
    return (double(__builtin_bitreverse32(i ^ (i >> 1))) + 0.5) * 0x1p-32;
}
__device__ __forceinline__ double halton2_point(uint32_t i) {
    return double(__builtin_bitreverse32(i)) * 0x1p-32;
}
__device__ __forceinline__ double clamp_u(double u) { return fmin(fmax(u, 1e-12), 1.0 - 1e-12); }

template <typename R>
struct TermC {
    R spot, strike, A, B;
    R Ah;         // TK_IS: A + B h
    R hn, hc;     // TK_IS: weight exponent -h log2e z - h^2/2 log2e
    R den, inv_den, clip;  // TK_LR
    R fwd;        // TK_CV: E[S_T]
    R hs, hk;     // FP32 pair units: 0.5 sg spot, -0.5 sg strike (sg = +1 call, -1 put)
    double n;     // TK_STRAT: strata (= paths)
    int32_t ot;
};

template <typename R, int KIND>
__device__ __forceinline__ double term_unit(const TermC<R>& k, R z) {
    if constexpr (KIND == TK_PAIR) {
        const R e = k.B * z;
        const R up = payoff(k.spot * q_exp2(k.A + e), k.strike, k.ot);
        const R dn = payoff(k.spot * q_exp2(k.A - e), k.strike, k.ot);
        return 0.5 * (double(up) + double(dn));
    } else if constexpr (KIND == TK_LR) {
        const R e = k.B * z;
        const R up = payoff(k.spot * q_exp2(k.A + e), k.strike, k.ot);
        const R dn = payoff(k.spot * q_exp2(k.A - e), k.strike, k.ot);
        const R sc = lr_score(z, k.den, k.inv_den, k.clip);
        return 0.5 * (double(up * sc) - double(dn * sc));
    } else if constexpr (KIND == TK_IS) {
        const R st = k.spot * q_exp2(q_fma(k.B, z, k.Ah));
        const R w = q_exp2(q_fma(k.hn, z, k.hc));
        return double(payoff(st, k.strike, k.ot) * w);
    } else {
        const R e = q_exp2(q_fma(k.B, z, k.A));
        const R st = k.spot * e;
        if (k.ot == QK_CALL) return st > k.strike ? double(e) : 0.0;
        return st < k.strike ? -double(e) : 0.0;
    }
}

// The same units in FP32 for the Philox path's group sums: the pair mean and
// the LR difference are formed in FP32 instead of FP64.  The pair mean is
//   max(hs 2^(A + B z) + hk, 0) + max(hs 2^(A - B z) + hk, 0)
// with hs = sg spot / 2 and hk = -sg strike / 2: the call/put sign and the
// halving (exact in binary) sit in the constants, so each path costs one
// exp2, one FMA and one max, with no per-path select.
template <int KIND>
__device__ __forceinline__ float term_unit_f32(const TermC<float>& k, float z) {
    if constexpr (KIND == TK_PAIR) {
        const float up = __builtin_fmaf(k.hs, q_exp2(__builtin_fmaf(k.B, z, k.A)), k.hk);
        const float dn = __builtin_fmaf(k.hs, q_exp2(__builtin_fmaf(-k.B, z, k.A)), k.hk);
        return (up > 0.0f ? up : 0.0f) + (dn > 0.0f ? dn : 0.0f);
    } else if constexpr (KIND == TK_LR) {
        const float e = k.B * z;
        const float up = payoff(k.spot * q_exp2(k.A + e), k.strike, k.ot);
        const float dn = payoff(k.spot * q_exp2(k.A - e), k.strike, k.ot);
        return 0.5f * (up - dn) * lr_score(z, k.den, k.inv_den, k.clip);
    } else if constexpr (KIND == TK_IS) {
        const float st = k.spot * q_exp2(q_fma(k.B, z, k.Ah));
        return payoff(st, k.strike, k.ot) * q_exp2(q_fma(k.hn, z, k.hc));
    } else {
        const float e = q_exp2(q_fma(k.B, z, k.A));
        const float st = k.spot * e;
        if (k.ot == QK_CALL) return st > k.strike ? e : 0.0f;
        return st < k.strike ? -e : 0.0f;
    }
}

// One tile's units: lane handles groups g0 + lane + j*kBlock of GU units:
// eight (one Philox block of four (radius, angle) pairs) for the estimators
// that draw normals, four for the uniform-driven ones.  x1, x2, xy
// accumulate the control (TK_CV only).
//
// FP32 paths (TK_PAIR, TK_LR, TK_IS, TK_PW) sum each group's eight centred
// units and their squares in FP32 and add the two group sums to the lane's
// FP64 sums: two FP64 additions per eight units instead of five FP64
// operations per unit.  The centring uses float(shift); the lane adds
// count * (float(shift) - shift) at the end, so s1 stays the sum of
// unit - shift.  The FP32 group sums round at 2^-24 of a few units, far below
// any standard error the sums produce.
template <typename R, int MODE, int KIND, class PS, class SS>
__device__ __forceinline__ void term_tile(const TermC<R>& k, const PS& ph, const SS& ss,
                                          int64_t g0, int32_t lane_groups, int32_t units,
                                          double shift, double& s1, double& s2, double& x1,
                                          double& x2, double& xy) {
    constexpr bool from_u = KIND == TK_STRAT || KIND == TK_SOBOL || KIND == TK_HALTON;
    constexpr int GU = from_u ? 4 : 8;  // units per group (term_group_units)
    constexpr bool f32_sums = sizeof(R) == 4 && !from_u && KIND != TK_CV;
    const float sh = float(shift);
    int32_t cnt = 0;
    for (int32_t j = 0; j < lane_groups; ++j) {
        const int64_t g = g0 + threadIdx.x + int64_t(j) * kBlock;
        if (g * GU >= units) break;
        Draws<R, GU> z;
        double uu[4] = {0.0, 0.0, 0.0, 0.0};
        if constexpr (KIND == TK_STRAT) {
            if constexpr (MODE == SIM_PHILOX_F32) {
                ph.term_uniforms(uint32_t(g), uu);
            } else {
                const Draws<R, 4> q = ss.template term_draws<4>(uint32_t(g));  // uniforms
#pragma unroll
                for (int c = 0; c < 4; ++c) uu[c] = double(q.z[c]);
            }
        } else if constexpr (!from_u) {
            if constexpr (MODE == SIM_PHILOX_F32) z = ph.term_normals8(uint32_t(g));
            else z = ss.template term_draws<8>(uint32_t(g));
        }
        if constexpr (f32_sums) {
            const int32_t left = int32_t(units - g * GU);  // >= 1 here
            float t1 = 0.0f, t2 = 0.0f;
            auto add = [&](int c) {
                const float f = term_unit_f32<KIND>(k, z.z[c]) - sh;
                t1 += f;
                t2 = __builtin_fmaf(f, f, t2);
            };
            if (left >= GU) {  // every group but an option's last
#pragma unroll
                for (int c = 0; c < GU; ++c) add(c);
                cnt += GU;
            } else {
#pragma unroll
                for (int c = 0; c < GU - 1; ++c)
                    if (c < left) add(c);
                cnt += left;
            }
            s1 += double(t1);
            s2 += double(t2);
            continue;
        }
#pragma unroll
        for (int c = 0; c < GU; ++c) {
            const int64_t unit = g * GU + c;
            if (unit < units) {
                double f;
                if constexpr (from_u) {
                    double u;
                    if constexpr (KIND == TK_STRAT) u = (double(unit) + uu[c]) / k.n;
                    else if constexpr (KIND == TK_SOBOL) u = sobol_point(uint32_t(unit + 1));
                    else u = halton2_point(uint32_t(unit + 1));
                    const R zr = R(acklam(clamp_u(u)));
                    f = double(payoff(k.spot * q_exp2(q_fma(k.B, zr, k.A)), k.strike, k.ot)) - shift;
                } else if constexpr (KIND == TK_CV) {
                    const R st = k.spot * q_exp2(q_fma(k.B, z.z[c], k.A));
                    f = double(payoff(st, k.strike, k.ot)) - shift;
                    const double x = double(st - k.fwd);
                    x1 += x;
                    x2 = __builtin_fma(x, x, x2);
                    xy = __builtin_fma(x, f, xy);
                } else {
                    f = term_unit<R, KIND>(k, z.z[c]) - shift;
                }
                s1 += f;
                s2 = __builtin_fma(f, f, s2);
            }
        }
    }
    if constexpr (f32_sums) s1 = __builtin_fma(double(cnt), double(sh) - shift, s1);
}

// PAIR: a batch of standard / antithetic options only (the estimator of
// mc_engine.h:32-53), compiled without the other estimators' branches: 49
// instead of 113 VGPRs, so twice the waves per SIMD.
template <typename R, int MODE, bool PAIR>
__global__ __launch_bounds__(kBlock) void k_sim_terminal(const OptionRec* __restrict__ recs,
                                                         const int32_t* __restrict__ tile_start,
                                                         const int32_t* __restrict__ n_dev,
                                                         int32_t* __restrict__ counters,
                                                         Partial* __restrict__ partials,
                                                         const R* __restrict__ stream,
                                                         int32_t lane_groups) {
    __shared__ int32_t s_tile;
    __shared__ double s_red[kBlock / 64][5];
    const float2* trig = nullptr;
    if constexpr (MODE == SIM_PHILOX_F32) {
        __shared__ __align__(16) float2 s_trig[kTrigEntries];
        stage_trig_table(g_trig, s_trig);
        trig = s_trig;
    }
    const int32_t n = *n_dev;
    const int32_t total = __builtin_amdgcn_readfirstlane(counters[1]);
    for (;;) {
        const int32_t tile = next_tile(counters, &s_tile);
        if (uint32_t(tile) >= uint32_t(total)) break;
        const int32_t opt = find_option(tile_start, n, tile);
        const OptionRec& rec = recs[opt];
        const int32_t units = rec.units, paths = rec.paths, method = rec.method;
        const int32_t local_tile = tile - rec.tile_first;
        const int64_t g0 = int64_t(local_tile) * kBlock * lane_groups;  // first group
        TermC<R> k;
        k.spot = R(rec.spot);
        k.strike = R(rec.strike);
        k.ot = rec.ot;
        k.A = R(rec.c[0] * kLog2e);
        k.B = R(rec.c[1] * kLog2e);
        const double h = rec.aux[0];
        k.Ah = R((rec.c[0] + rec.c[1] * h) * kLog2e);
        k.hn = R(-h * kLog2e);
        k.hc = R(-0.5 * h * h * kLog2e);
        k.den = R(rec.aux[2]);
        k.inv_den = R(1.0 / rec.aux[2]);
        k.clip = R(rec.aux[0]);
        k.fwd = R(rec.aux[0]);
        const double sg = rec.ot == QK_CALL ? 0.5 : -0.5;
        k.hs = R(sg * rec.spot);
        k.hk = R(-sg * rec.strike);
        k.n = double(paths);
        const double shift = rec.shift;
        const bool anti = method_is_antithetic_terminal(method);
        PhiloxSrc ph{rec.seed_lo, rec.seed_hi, 0u, trig};
        StreamSrc<R> ss{stream + rec.stream_off, 1, int64_t(units) + (anti ? (paths & 1) : 0)};

        double s1 = 0.0, s2 = 0.0, x1 = 0.0, x2 = 0.0, xy = 0.0;
#define QK_TERM_TILE(KIND) \
    term_tile<R, MODE, KIND>(k, ph, ss, g0, lane_groups, units, shift, s1, s2, x1, x2, xy)
        if constexpr (PAIR) {
            QK_TERM_TILE(TK_PAIR);
        } else {
            switch (method) {
                case M_LR: QK_TERM_TILE(TK_LR); break;
                case M_IMPORTANCE: QK_TERM_TILE(TK_IS); break;
                case M_PATHWISE: QK_TERM_TILE(TK_PW); break;
                case M_CONTROL: QK_TERM_TILE(TK_CV); break;
                case M_STRATIFIED: QK_TERM_TILE(TK_STRAT); break;
                case M_SOBOL: QK_TERM_TILE(TK_SOBOL); break;
                case M_HALTON: QK_TERM_TILE(TK_HALTON); break;
                default: QK_TERM_TILE(TK_PAIR); break;
            }
        }
#undef QK_TERM_TILE
        // antithetic estimators: an odd path count adds one path with a fresh draw
        double extra = 0.0;
        if (anti && (paths & 1) && local_tile == rec.ntiles - 1 && threadIdx.x == 0) {
            R z;
            if constexpr (MODE == SIM_PHILOX_F32) z = ph.term_draw(uint32_t(units));
            else z = ss.term_draw(uint32_t(units));
            const R pay = payoff(k.spot * q_exp2(k.A + k.B * z), k.strike, k.ot);
            extra = method == M_LR ? double(pay * lr_score(z, k.den, k.inv_den, k.clip))
                                   : double(pay);
        }
        if (method == M_CONTROL)
            block_store_partial_cv(s1, s2, x1, x2, xy, partials + tile, s_red);
        else
            block_store_partial(s1, s2, extra, partials + tile, s_red);
    }
}

// ----------------------------------------------------------------------------
// finalize
// ----------------------------------------------------------------------------
struct Neumaier {  // compensated FP64 sum (Neumaier 1974)
    double s = 0.0, c = 0.0;
    __device__ __forceinline__ void add(double x) {
        const double t = s + x;
        c += fabs(s) >= fabs(x) ? (s - t) + x : (x - t) + s;
        s = t;
    }
    __device__ __forceinline__ double sum() const { return s + c; }
};

// host_flags, when set, is pinned host memory (its device address): the
// scan's counters[0..3] are stored there, so the host reads the overflow flag
// after the stream completes without a copy of its own.
__global__ __launch_bounds__(256) void k_finalize(const int32_t* __restrict__ n_dev,
                                                  const OptionRec* __restrict__ recs,
                                                  const Partial* __restrict__ partials,
                                                  const int32_t* __restrict__ counters,
                                                  double* __restrict__ out_price,
                                                  double* __restrict__ out_se,
                                                  int32_t* __restrict__ host_flags) {
    const int32_t n = *n_dev;
    const bool overflow = counters[2] != 0;  // no tile ran: partials were not written
    if (host_flags && blockIdx.x == 0 && threadIdx.x < 4) host_flags[threadIdx.x] = counters[threadIdx.x];
    for (int32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const OptionRec& rec = recs[i];
        if (rec.family == FAM_NONE) {
            out_price[i] = rec.closed;
            out_se[i] = rec.closed == rec.closed ? 0.0 : qnan_d();
            continue;
        }
        if (overflow) {
            out_price[i] = qnan_d();
            out_se[i] = qnan_d();
            continue;
        }
        // Neumaier compensated sums over the option's tiles, in tile order.
        const bool cv = rec.method == M_CONTROL;
        Neumaier a, b, xa, xb, xc;
        double extra = 0.0;
        for (int32_t t = 0; t < rec.ntiles; ++t) {
            const Partial& p = partials[rec.tile_first + t];
            a.add(p.s1);
            b.add(p.s2);
            extra += p.extra;
            if (cv) {
                xa.add(p.x1);
                xb.add(p.x2);
                xc.add(p.xy);
            }
        }
        const double s1 = a.sum(), s2 = b.sum();
        if (cv) {
            // control_variates.cpp:42-53 on centred sums: y = payoff - shift,
            // x = S_T - E[S_T]; the reference's x and y carry the discount,
            // so its var_x <= 1e-12 test is disc^2 var_x here
            const double n = double(rec.units);
            const double my = s1 / n, mx = xa.sum() / n;
            const double var_x = xb.sum() / n - mx * mx;
            const double var_y = s2 / n - my * my;
            double est = rec.shift + my, res = var_y;
            if (rec.disc * rec.disc * var_x > kEps) {
                const double cov = xc.sum() / n - mx * my;
                const double beta = cov / var_x;
                est -= beta * mx;
                res = var_y - cov * beta;
            }
            out_price[i] = rec.disc * est;
            out_se[i] = rec.disc * sqrt(fmax(res, 0.0) / (n - 1.0));
            continue;
        }
        const double units = double(rec.units);
        const double mean_d = s1 / units;
        double var = units > 1.0 ? (s2 - s1 * mean_d) / (units - 1.0) : 0.0;
        var = var > 0.0 ? var : 0.0;
        double price;
        double se = rec.disc * sqrt(var / units);
        const int32_t m = rec.method;
        if (rec.family == FAM_TERMINAL) {
            // units are pairs (sum counts both paths) or single paths
            const double sum = method_is_antithetic_terminal(m)
                                   ? 2.0 * (rec.shift * units + s1) + extra
                                   : rec.shift * units + s1;
            price = rec.disc * (sum / double(rec.paths));
            if (m == M_PATHWISE || m == M_LR) {
                // adjoint_greeks/common/internal_util.h:47-50
                const double qf = rec.aux[1];
                price = fmin(qf, fmax(-qf, price));
            }
        } else if (m == M_LOCAL_VOL) {
            const double sum = rec.shift * units + s1;
            price = rec.disc * sum / double(rec.paths);  // local_vol_monte_carlo.cpp:47
        } else if (m == M_HESTON_LR) {
            // heston_lr_delta.cpp:131: disc * sum / paths, units are pair sums
            const double sum = rec.shift * units + s1 + extra;
            price = rec.disc * sum / double(rec.paths);
            se *= units / double(rec.paths);
        } else {
            price = rec.disc * (rec.shift + mean_d);
        }
        out_price[i] = price;
        out_se[i] = method_is_qmc(m) ? qnan_d() : se;  // no sampling error for point sets
    }
}

// ----------------------------------------------------------------------------
// stream self-check: raw Philox words and FP32 normals of blocks
// (c0 + i, c1, 0, c3), i < count
// ----------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_philox_probe(uint32_t k0, uint32_t k1, uint32_t c0,
                                                      uint32_t c1, uint32_t c3, int32_t count,
                                                      uint32_t* __restrict__ words,
                                                      float* __restrict__ normals) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const U4 w = philox4x32_10(c0 + uint32_t(i), c1, c3, k0, k1);
    words[4 * i + 0] = w.x;
    words[4 * i + 1] = w.y;
    words[4 * i + 2] = w.z;
    words[4 * i + 3] = w.w;
    float a, b, c, d;
    box_muller_f32(w.x, w.y, a, b);
    box_muller_f32(w.z, w.w, c, d);
    normals[4 * i + 0] = a;
    normals[4 * i + 1] = b;
    normals[4 * i + 2] = c;
    normals[4 * i + 3] = d;
}

// The eight normals of the four-pair layout of blocks (c0 + i, c1, 0, c3),
// read through the device's angle table as the path kernels read them.
__global__ __launch_bounds__(256) void k_philox_probe8(uint32_t k0, uint32_t k1, uint32_t c0,
                                                       uint32_t c1, uint32_t c3, int32_t count,
                                                       float* __restrict__ normals) {
    const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const Normal8 n = philox_normals8(philox4x32_10(c0 + uint32_t(i), c1, c3, k0, k1), g_trig);
#pragma unroll
    for (int k = 0; k < 8; ++k) normals[8 * i + k] = n.z[k];
}

hipError_t launch_philox_probe8(hipStream_t st, uint64_t seed, uint32_t c0, uint32_t c1,
                                uint32_t c3, int32_t count, float* normals) {
    hipLaunchKernelGGL(k_philox_probe8, dim3((count + 255) / 256), dim3(256), 0, st,
                       uint32_t(seed), uint32_t(seed >> 32), c0, c1, c3, count, normals);
    return hipGetLastError();
}

hipError_t launch_philox_probe(hipStream_t st, uint64_t seed, uint32_t c0, uint32_t c1,
                               uint32_t c3, int32_t count, uint32_t* words, float* normals) {
    hipLaunchKernelGGL(k_philox_probe, dim3((count + 255) / 256), dim3(256), 0, st,
                       uint32_t(seed), uint32_t(seed >> 32), c0, c1, c3, count, words, normals);
    return hipGetLastError();
}

// ----------------------------------------------------------------------------
// host launchers
// ----------------------------------------------------------------------------
hipError_t upload_trig_table() {
    // (cos, sin) of 2 pi a / 512: a quarter turn in FP64 (the C library's
    // cos / sin of an exact multiple of 2 pi / 512, rounded once to FP32),
    // the other three quarters by exact symmetry, so sign and zero patterns
    // are exact (entry 128 is (0, 1)).
    static float2 host[kTrigEntries];
    constexpr int Q = kTrigEntries / 4;
    const double step = 6.283185307179586476925286766559 / kTrigEntries;
    for (int a = 0; a < Q; ++a) {
        const float c = float(std::cos(step * a)), s = float(std::sin(step * a));
        host[a] = make_float2(c, s);
        host[a + Q] = make_float2(-s, c);
        host[a + 2 * Q] = make_float2(-c, -s);
        host[a + 3 * Q] = make_float2(s, -c);
    }
    return hipMemcpyToSymbol(HIP_SYMBOL(g_trig), host, sizeof(host), 0, hipMemcpyHostToDevice);
}

hipError_t launch_prepare(hipStream_t st, const BatchIn& in, const int32_t* n_dev, int32_t n_host,
                          int32_t* n_out, const int64_t* stream_off, int32_t tg_terminal,
                          int32_t tu_stepwise, OptionRec* recs, int32_t* ntiles) {
    hipLaunchKernelGGL(k_prepare, dim3(1024), dim3(256), 0, st, in, n_dev, n_host, n_out,
                       stream_off, tg_terminal, tu_stepwise, recs, ntiles);
    return hipGetLastError();
}

hipError_t launch_scan(hipStream_t st, int32_t n, const int32_t* n_dev, const int32_t* ntiles,
                       OptionRec* recs, int32_t* tile_start, int32_t* counters, void* scratch,
                       int32_t capacity) {
    // n + 1 elements so that the thread of index n writes tile_start[n]
    const int32_t nb = (n + 1 + kScanBlock - 1) / kScanBlock;
    if (nb <= kScanSmallBlocks) {
        hipLaunchKernelGGL(k_scan_small, dim3(nb), dim3(kScanBlock), 0, st, n_dev, ntiles, counters,
                           capacity, recs, tile_start);
        return hipGetLastError();
    }
    long long* bs = static_cast<long long*>(scratch);
    hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(kScanBlock), 0, st, n_dev, ntiles, bs);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(kScanBlock), 0, st, nb, bs, counters, capacity);
    hipLaunchKernelGGL(k_scan_apply, dim3(nb), dim3(kScanBlock), 0, st, n_dev, ntiles, bs, counters,
                       recs, tile_start);
    return hipGetLastError();
}

hipError_t launch_simulate(hipStream_t st, int32_t family, int32_t mode, const SimShape& shape,
                           const OptionRec* recs, const int32_t* tile_start,
                           const int32_t* n_dev, int32_t* counters, Partial* partials,
                           const void* stream, bool pair_only) {
    const dim3 grid(shape.grid), block(kBlock);
    if (family == FAM_TERMINAL) {
        switch (mode) {
            case SIM_PHILOX_F32:
                if (pair_only)
                    hipLaunchKernelGGL((k_sim_terminal<float, SIM_PHILOX_F32, true>), grid, block, 0,
                                       st, recs, tile_start, n_dev, counters, partials,
                                       static_cast<const float*>(stream), shape.lane_units);
                else
                    hipLaunchKernelGGL((k_sim_terminal<float, SIM_PHILOX_F32, false>), grid, block, 0,
                                       st, recs, tile_start, n_dev, counters, partials,
                                       static_cast<const float*>(stream), shape.lane_units);
                break;
            case SIM_STREAM_F64:
                hipLaunchKernelGGL((k_sim_terminal<double, SIM_STREAM_F64, false>), grid, block, 0, st,
                                   recs, tile_start, n_dev, counters, partials,
                                   static_cast<const double*>(stream), shape.lane_units);
                break;
            default:
                hipLaunchKernelGGL((k_sim_terminal<float, SIM_STREAM_F32, false>), grid, block, 0, st,
                                   recs, tile_start, n_dev, counters, partials,
                                   static_cast<const float*>(stream), shape.lane_units);
                break;
        }
    } else {
        switch (mode) {
            case SIM_PHILOX_F32:
                hipLaunchKernelGGL((k_sim_stepwise<float, SIM_PHILOX_F32>), grid, block, 0, st,
                                   recs, tile_start, n_dev, counters, partials,
                                   static_cast<const float*>(stream), shape.lane_units);
                break;
            case SIM_STREAM_F64:
                hipLaunchKernelGGL((k_sim_stepwise<double, SIM_STREAM_F64>), grid, block, 0, st,
                                   recs, tile_start, n_dev, counters, partials,
                                   static_cast<const double*>(stream), shape.lane_units);
                break;
            default:
                hipLaunchKernelGGL((k_sim_stepwise<float, SIM_STREAM_F32>), grid, block, 0, st,
                                   recs, tile_start, n_dev, counters, partials,
                                   static_cast<const float*>(stream), shape.lane_units);
                break;
        }
    }
    return hipGetLastError();
}

hipError_t launch_finalize(hipStream_t st, const int32_t* n_dev, const OptionRec* recs,
                           const Partial* partials, const int32_t* counters, double* out_price,
                           double* out_se, int32_t* host_flags) {
    hipLaunchKernelGGL(k_finalize, dim3(256), dim3(256), 0, st, n_dev, recs, partials, counters,
                       out_price, out_se, host_flags);
    return hipGetLastError();
}

}  // namespace qk
