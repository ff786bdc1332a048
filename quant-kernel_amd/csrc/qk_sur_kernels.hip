// qk_sur_kernels.hip - per-option kernels of surrogate pricing with Monte
// Carlo fallback: feature building with the training-box test, surrogate
// price (calls clamped to their no-arbitrage bounds; puts by put-call parity,
// which the log-Euler scheme satisfies in expectation), the per-query spot
// check, the batch judge and the final merge.
#include "qk_common.hpp"
#include "qk_kernels.hpp"
#include "qk_mlp.hpp"
#include "qk_surrogate.hpp"

namespace qk {

using mlp::kBox;
using mlp::kIn;
using mlp::kInPad;

// Judge region of an in-box query: one bit each for the upper halves of the
// box in log K/S, T, sigma (vol of vol) and v0, the features the price
// depends on most.
__device__ __forceinline__ int8_t judge_region(const double* f) {
    int r = 0;
    const int dims[4] = {0, 1, 5, 7};
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const int c = dims[b];
        r |= (f[c] > 0.5 * (kBox.lo[c] + kBox.hi[c]) ? 1 : 0) << b;
    }
    return int8_t(r);
}

// Per query: validation (the Heston MC rules), the box test, features of
// in-box queries, and pass A of the MC: the spot checks (spot_paths paths,
// every spot_every-th query answered by the network) together with the
// out-of-box queries (their own paths).
__global__ __launch_bounds__(256) void k_sur_prepare(BatchIn in, int n, int Mpad, int spot_every,
                                                     int spot_paths, __bf16* __restrict__ X,
                                                     uint8_t* __restrict__ act_a,
                                                     int32_t* __restrict__ paths_a,
                                                     uint8_t* __restrict__ act_spot,
                                                     uint8_t* __restrict__ act_b,
                                                     int8_t* __restrict__ status,
                                                     int8_t* __restrict__ region,
                                                     double* __restrict__ closed) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < Mpad; i += gridDim.x * blockDim.x) {
        __bf16* row = X + size_t(i) * kInPad;
        for (int c = 0; c < kInPad; ++c) row[c] = __bf16(0.0f);
        if (i >= n) continue;
        RawOption o;
        o.method = M_HESTON;
        o.spot = in.spot[i];
        o.strike = in.strike[i];
        o.t = in.t[i];
        o.vol = 0.0;
        o.r = in.r[i];
        o.q = in.q[i];
        o.v0 = in.v0[i];
        o.kappa = in.kappa[i];
        o.theta = in.theta[i];
        o.sigma = in.sigma[i];
        o.rho = in.rho[i];
        o.ot = in.ot[i];
        o.paths = in.paths[i];
        o.steps = in.steps[i];
        o.seed = in.seed[i];
        o.aux = 0.0;
        OptionRec rec;
        act_a[i] = 0;
        act_spot[i] = 0;
        act_b[i] = 0;
        paths_a[i] = o.paths;
        region[i] = -1;
        closed[i] = 0.0;
        if (prepare_option(o, 1, rec) == FAM_NONE) {
            status[i] = SUR_CLOSED;
            closed[i] = rec.closed;
            continue;
        }
        const double f[kIn] = {log(o.strike / o.spot), o.t, o.r, o.kappa, o.theta, o.sigma, o.rho,
                               o.v0};
        bool inside = o.q == 0.0 && o.steps == kSurSteps;
        for (int c = 0; c < kIn; ++c) inside = inside && f[c] >= kBox.lo[c] && f[c] <= kBox.hi[c];
        if (!inside) {
            status[i] = SUR_FALLBACK_BOX;
            act_a[i] = 1;
            continue;
        }
        status[i] = SUR_NETWORK;
        region[i] = judge_region(f);
        for (int c = 0; c < kIn; ++c) {
            const double mid = 0.5 * (kBox.lo[c] + kBox.hi[c]);
            const double half = 0.5 * (kBox.hi[c] - kBox.lo[c]);
            row[c] = __bf16(float((f[c] - mid) / half));
        }
        if (spot_every > 0 && i % spot_every == 0) {
            act_spot[i] = 1;
            act_a[i] = 1;
            paths_a[i] = spot_paths;
        }
    }
}

// Network price: call = S (control + network), clamped to the call's
// no-arbitrage bounds [max(S - K e^{-rT}, 0), S] (q = 0 inside the box),
// which the MC expectation of the log-Euler scheme also satisfies; puts by
// put-call parity from the clamped call.
__global__ __launch_bounds__(256) void k_sur_price(BatchIn in, int n, const float* __restrict__ y,
                                                   const int8_t* __restrict__ status,
                                                   double* __restrict__ sur_price) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (status[i] != SUR_NETWORK) {
            sur_price[i] = 0.0;
            continue;
        }
        const double s = in.spot[i], k = in.strike[i], t = in.t[i];
        const double f[kIn] = {log(k / s), t, in.r[i], in.kappa[i], in.theta[i], in.sigma[i],
                               in.rho[i], in.v0[i]};
        const double dk = k * exp(-in.r[i] * t);
        const double call = fmin(s, fmax(fmax(s - dk, 0.0), s * (mlp::heston_control(f) + double(y[i]))));
        sur_price[i] = in.ot[i] == QK_CALL ? call : call - s + dk;
    }
}

// Per-query spot check: the network answer fails when it differs from the
// check's MC price by more than k_se MC standard errors plus tol (half the
// accuracy target, per unit spot).  The tolerance keeps checks whose MC saw
// few or no in-the-money paths (SE ~ 0) from rejecting answers that are
// within the target anyway.
__global__ __launch_bounds__(256) void k_sur_compare(BatchIn in, int n, double k_se, double tol,
                                                     const uint8_t* __restrict__ act_spot,
                                                     const double* __restrict__ sur_price,
                                                     const double* __restrict__ mc_price,
                                                     const double* __restrict__ mc_se,
                                                     uint8_t* __restrict__ act_b,
                                                     int8_t* __restrict__ status) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (!act_spot[i]) continue;
        const double diff = fabs(sur_price[i] - mc_price[i]);
        if (!(diff <= k_se * mc_se[i] + tol * in.spot[i])) {
            act_b[i] = 1;
            status[i] = SUR_FALLBACK_CHECK;
        }
    }
}

// ---- batch judge ------------------------------------------------------------
// Over the spot checks of a judge region, e_i = ((sur_i - mc_i) / S_i)^2 -
// (SE_i / S_i)^2 has mean equal to the network's mean squared error per unit
// spot (the MC estimate is unbiased and independent of the network).  Each
// block sums (count, sum e, sum e^2) per region over a fixed range of
// queries, in a fixed order; one block then sums the blocks in order and
// rejects a region (or, over all checks, the whole batch) when its mean minus
// k standard errors exceeds target^2 with at least kSurJudgeMinChecks checks.
constexpr int kJudgeChunk = 4096;  // queries per block of the partial sums

__global__ __launch_bounds__(256) void k_sur_judge_partial(BatchIn in, int n,
                                                           const uint8_t* __restrict__ act_spot,
                                                           const int8_t* __restrict__ region,
                                                           const double* __restrict__ sur_price,
                                                           const double* __restrict__ mc_price,
                                                           const double* __restrict__ mc_se,
                                                           double* __restrict__ part) {
    __shared__ double red[4][3];
    int cnt[kSurRegions];
    double s1[kSurRegions], s2[kSurRegions];
#pragma unroll
    for (int r = 0; r < kSurRegions; ++r) {
        cnt[r] = 0;
        s1[r] = 0.0;
        s2[r] = 0.0;
    }
    const int base = blockIdx.x * kJudgeChunk;
    for (int j = 0; j < kJudgeChunk / 256; ++j) {
        const int i = base + j * 256 + int(threadIdx.x);
        if (i >= n || !act_spot[i]) continue;
        const double s = in.spot[i];
        const double d = (sur_price[i] - mc_price[i]) / s, v = mc_se[i] / s;
        const double e = d * d - v * v;
        const int rg = region[i];
#pragma unroll
        for (int r = 0; r < kSurRegions; ++r) {
            if (r == rg) {
                ++cnt[r];
                s1[r] += e;
                s2[r] = fma(e, e, s2[r]);
            }
        }
    }
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int r = 0; r < kSurRegions; ++r) {
        double a = double(cnt[r]), b = s1[r], c = s2[r];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            a += __shfl_xor(a, off, 64);
            b += __shfl_xor(b, off, 64);
            c += __shfl_xor(c, off, 64);
        }
        if (lane == 0) {
            red[wave][0] = a;
            red[wave][1] = b;
            red[wave][2] = c;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            double* out = part + (size_t(blockIdx.x) * kSurRegions + r) * 3;
            out[0] = red[0][0] + red[1][0] + red[2][0] + red[3][0];
            out[1] = red[0][1] + red[1][1] + red[2][1] + red[3][1];
            out[2] = red[0][2] + red[1][2] + red[2][2] + red[3][2];
        }
        __syncthreads();
    }
}

// stats row r (r < kSurRegions: a region; r == kSurRegions: all checks):
// {checks, mean e (implied MSE / S^2), its standard error, rejected 0/1}
__global__ __launch_bounds__(64) void k_sur_judge_final(int nblocks, const double* __restrict__ part,
                                                        double target, double k_sig,
                                                        double* __restrict__ stats) {
    const int r = threadIdx.x;
    if (r > kSurRegions) return;
    double a = 0.0, b = 0.0, c = 0.0;
    for (int blk = 0; blk < nblocks; ++blk) {
        for (int q = 0; q < kSurRegions; ++q) {
            if (r != q && r != kSurRegions) continue;
            const double* p = part + (size_t(blk) * kSurRegions + q) * 3;
            a += p[0];
            b += p[1];
            c += p[2];
        }
    }
    double mean = 0.0, se = 0.0;
    if (a > 0.0) mean = b / a;
    if (a > 1.0) se = sqrt(fmax(c - b * mean, 0.0) / (a - 1.0) / a);
    const bool reject = a >= double(kSurJudgeMinChecks) && mean - k_sig * se > target * target;
    stats[r * 4 + 0] = a;
    stats[r * 4 + 1] = mean;
    stats[r * 4 + 2] = se;
    stats[r * 4 + 3] = reject ? 1.0 : 0.0;
}

// Network answers of a rejected region (or of a rejected batch) go to MC.
__global__ __launch_bounds__(256) void k_sur_judge_apply(int n, const int8_t* __restrict__ region,
                                                         const double* __restrict__ stats,
                                                         uint8_t* __restrict__ act_b,
                                                         int8_t* __restrict__ status) {
    const bool all = stats[kSurRegions * 4 + 3] != 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (status[i] != SUR_NETWORK) continue;
        if (all || stats[region[i] * 4 + 3] != 0.0) {
            act_b[i] = 1;
            status[i] = SUR_FALLBACK_JUDGE;
        }
    }
}

__global__ __launch_bounds__(256) void k_sur_combine(int n, const int8_t* __restrict__ status,
                                                     const double* __restrict__ sur_price,
                                                     const double* __restrict__ a_price,
                                                     const double* __restrict__ b_price,
                                                     const double* __restrict__ closed,
                                                     double* __restrict__ out_price,
                                                     int32_t* __restrict__ out_source) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int s = status[i];
        out_price[i] = s == SUR_NETWORK        ? sur_price[i]
                       : s == SUR_CLOSED       ? closed[i]
                       : s == SUR_FALLBACK_BOX ? a_price[i]
                                               : b_price[i];
        out_source[i] = s;
    }
}

static int grid_of(int n) {
    const int g = (n + 255) / 256;
    return g < 2048 ? (g ? g : 1) : 2048;
}

int sur_judge_blocks(int n) { return (n + kJudgeChunk - 1) / kJudgeChunk; }

hipError_t sur_prepare(hipStream_t st, const BatchIn& in, int n, int Mpad, int spot_every,
                       int spot_paths, void* X, uint8_t* act_a, int32_t* paths_a,
                       uint8_t* act_spot, uint8_t* act_b, int8_t* status, int8_t* region,
                       double* closed) {
    hipLaunchKernelGGL(k_sur_prepare, dim3(grid_of(Mpad)), dim3(256), 0, st, in, n, Mpad,
                       spot_every, spot_paths, static_cast<__bf16*>(X), act_a, paths_a, act_spot,
                       act_b, status, region, closed);
    return hipGetLastError();
}

hipError_t sur_price(hipStream_t st, const BatchIn& in, int n, const float* y, const int8_t* status,
                     double* sur) {
    hipLaunchKernelGGL(k_sur_price, dim3(grid_of(n)), dim3(256), 0, st, in, n, y, status, sur);
    return hipGetLastError();
}

hipError_t sur_compare(hipStream_t st, const BatchIn& in, int n, double k_se, double tol,
                       const uint8_t* act_spot, const double* sur, const double* mc,
                       const double* mc_se, uint8_t* act_b, int8_t* status) {
    hipLaunchKernelGGL(k_sur_compare, dim3(grid_of(n)), dim3(256), 0, st, in, n, k_se, tol,
                       act_spot, sur, mc, mc_se, act_b, status);
    return hipGetLastError();
}

hipError_t sur_judge(hipStream_t st, const BatchIn& in, int n, const uint8_t* act_spot,
                     const int8_t* region, const double* sur, const double* mc,
                     const double* mc_se, double target, double k_sig, double* part,
                     double* stats, uint8_t* act_b, int8_t* status) {
    const int nb = sur_judge_blocks(n);
    hipLaunchKernelGGL(k_sur_judge_partial, dim3(nb), dim3(256), 0, st, in, n, act_spot, region, sur,
                       mc, mc_se, part);
    hipLaunchKernelGGL(k_sur_judge_final, dim3(1), dim3(64), 0, st, nb, part, target, k_sig, stats);
    hipLaunchKernelGGL(k_sur_judge_apply, dim3(grid_of(n)), dim3(256), 0, st, n, region, stats,
                       act_b, status);
    return hipGetLastError();
}

hipError_t sur_combine(hipStream_t st, int n, const int8_t* status, const double* sur,
                       const double* a, const double* b, const double* closed, double* out,
                       int32_t* src) {
    hipLaunchKernelGGL(k_sur_combine, dim3(grid_of(n)), dim3(256), 0, st, n, status, sur, a, b,
                       closed, out, src);
    return hipGetLastError();
}

}  // namespace qk
