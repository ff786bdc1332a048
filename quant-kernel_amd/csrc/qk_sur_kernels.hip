// qk_sur_kernels.hip - per-option kernels of surrogate pricing with Monte
// Carlo fallback: feature building with the training-box test, surrogate
// price (calls; puts by put-call parity, which the log-Euler scheme satisfies
// in expectation), spot-check comparison and the final merge.
#include "qk_common.hpp"
#include "qk_kernels.hpp"
#include "qk_mlp.hpp"
#include "qk_surrogate.hpp"

namespace qk {

using mlp::kBox;
using mlp::kIn;
using mlp::kInPad;

__global__ __launch_bounds__(256) void k_sur_prepare(BatchIn in, int n, int Mpad, int spot_every,
                                                     __bf16* __restrict__ X,
                                                     uint8_t* __restrict__ act_fb,
                                                     uint8_t* __restrict__ act_spot,
                                                     int8_t* __restrict__ status,
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
        act_fb[i] = 0;
        act_spot[i] = 0;
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
            act_fb[i] = 1;
            continue;
        }
        status[i] = SUR_NETWORK;
        for (int c = 0; c < kIn; ++c) {
            const double mid = 0.5 * (kBox.lo[c] + kBox.hi[c]);
            const double half = 0.5 * (kBox.hi[c] - kBox.lo[c]);
            row[c] = __bf16(float((f[c] - mid) / half));
        }
        act_spot[i] = (spot_every > 0 && i % spot_every == 0) ? 1 : 0;
    }
}

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
        const double call = s * (mlp::heston_control(f) + double(y[i]));
        sur_price[i] = in.ot[i] == QK_CALL ? call : call - s + k * exp(-in.r[i] * t);
    }
}

__global__ __launch_bounds__(256) void k_sur_compare(BatchIn in, int n, double k_se,
                                                     const uint8_t* __restrict__ act_spot,
                                                     const double* __restrict__ sur_price,
                                                     const double* __restrict__ mc_price,
                                                     const double* __restrict__ mc_se,
                                                     uint8_t* __restrict__ act_fb,
                                                     int8_t* __restrict__ status) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (!act_spot[i]) continue;
        const double diff = fabs(sur_price[i] - mc_price[i]);
        if (!(diff <= k_se * mc_se[i] + 1e-9 * in.spot[i])) {
            act_fb[i] = 1;
            status[i] = SUR_FALLBACK_CHECK;
        }
    }
}

__global__ __launch_bounds__(256) void k_sur_combine(int n, const int8_t* __restrict__ status,
                                                     const double* __restrict__ sur_price,
                                                     const double* __restrict__ fb_price,
                                                     const double* __restrict__ closed,
                                                     double* __restrict__ out_price,
                                                     int32_t* __restrict__ out_source) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const int s = status[i];
        out_price[i] = s == SUR_NETWORK ? sur_price[i] : s == SUR_CLOSED ? closed[i] : fb_price[i];
        out_source[i] = s;
    }
}

static int grid_of(int n) {
    const int g = (n + 255) / 256;
    return g < 2048 ? (g ? g : 1) : 2048;
}

hipError_t sur_prepare(hipStream_t st, const BatchIn& in, int n, int Mpad, int spot_every, void* X,
                       uint8_t* act_fb, uint8_t* act_spot, int8_t* status, double* closed) {
    hipLaunchKernelGGL(k_sur_prepare, dim3(grid_of(Mpad)), dim3(256), 0, st, in, n, Mpad,
                       spot_every, static_cast<__bf16*>(X), act_fb, act_spot, status, closed);
    return hipGetLastError();
}

hipError_t sur_price(hipStream_t st, const BatchIn& in, int n, const float* y, const int8_t* status,
                     double* sur) {
    hipLaunchKernelGGL(k_sur_price, dim3(grid_of(n)), dim3(256), 0, st, in, n, y, status, sur);
    return hipGetLastError();
}

hipError_t sur_compare(hipStream_t st, const BatchIn& in, int n, double k_se,
                       const uint8_t* act_spot, const double* sur, const double* mc,
                       const double* mc_se, uint8_t* act_fb, int8_t* status) {
    hipLaunchKernelGGL(k_sur_compare, dim3(grid_of(n)), dim3(256), 0, st, in, n, k_se, act_spot,
                       sur, mc, mc_se, act_fb, status);
    return hipGetLastError();
}

hipError_t sur_combine(hipStream_t st, int n, const int8_t* status, const double* sur,
                       const double* fb, const double* closed, double* out, int32_t* src) {
    hipLaunchKernelGGL(k_sur_combine, dim3(grid_of(n)), dim3(256), 0, st, n, status, sur, fb,
                       closed, out, src);
    return hipGetLastError();
}

}  // namespace qk
