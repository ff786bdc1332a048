// qk_cf.hip - Heston semi-analytic prices on gfx950 (characteristic function,
// Gil-Pelaez P1/P2 integrals over Gauss-Legendre panels), FP64 throughout.
//
// Restates closed_form_semi_analytical/heston/heston.cpp:7-53 and the
// integration of closed_form_semi_analytical/common/internal_util.h:80-142:
// the Albrecher et al. (2007) form of the Heston (1993) characteristic
// function, P2 and P1 (the latter normalised by phi(-i)) as sums over
// panels of 64 Gauss-Legendre nodes on [1e-8, integration_limit], and the
// put by parity.  One wavefront prices one option: lane L evaluates node
// (L/2, side L%2) of every panel, and the wave reduces the two sums in a
// fixed order, so results are deterministic.
#include "qk_cf.hpp"
#include "qk_common.hpp"

namespace qk {

namespace {

struct Cx {
    double re, im;
};

__device__ __forceinline__ Cx cadd(Cx a, Cx b) { return {a.re + b.re, a.im + b.im}; }
__device__ __forceinline__ Cx csub(Cx a, Cx b) { return {a.re - b.re, a.im - b.im}; }
__device__ __forceinline__ Cx cmul(Cx a, Cx b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__device__ __forceinline__ Cx cscale(Cx a, double s) { return {a.re * s, a.im * s}; }

// Smith's division (scaled by the larger denominator component), with one
// reciprocal of the scaled denominator in place of two divisions.
__device__ __forceinline__ Cx cdiv(Cx a, Cx b) {
    if (fabs(b.re) >= fabs(b.im)) {
        const double r = b.im / b.re, inv = 1.0 / (b.re + b.im * r);
        return {(a.re + a.im * r) * inv, (a.im - a.re * r) * inv};
    }
    const double r = b.re / b.im, inv = 1.0 / (b.re * r + b.im);
    return {(a.re * r + a.im) * inv, (a.im * r - a.re) * inv};
}

// Principal square root (branch cut on the negative real axis, the sign of
// the imaginary part kept as for a signed zero).
__device__ __forceinline__ Cx csqrt(Cx z) {
    const double m = hypot(z.re, z.im);
    if (z.re >= 0.0) {
        const double s = sqrt(0.5 * (m + z.re));
        return {s, s > 0.0 ? z.im / (2.0 * s) : z.im};
    }
    const double s = sqrt(0.5 * (m - z.re));
    return {fabs(z.im) / (2.0 * s), copysign(s, z.im)};
}

__device__ __forceinline__ Cx cexp(Cx z) {
    const double m = exp(z.re);
    double sn, cs;
    sincos(z.im, &sn, &cs);
    return {m * cs, m * sn};
}

// Principal logarithm: arg in (-pi, pi].
__device__ __forceinline__ Cx clog(Cx z) { return {log(hypot(z.re, z.im)), atan2(z.im, z.re)}; }

struct HestonCf {
    double v0, kappa, rho_sigma, sigma2, inv_sigma2, kt_inv_s2, drift, t;
};

// heston.cpp:33-44: phi(u) = E[exp(i u ln S_T)], times exp(-i rot) for a real
// rot folded into the final exponential (the integrands' e^{-i u ln K}).
__device__ Cx heston_phi(const HestonCf& h, Cx u, double rot) {
    const Cx iu = {-u.im, u.re};
    const Cx beta = {h.kappa - h.rho_sigma * iu.re, -h.rho_sigma * iu.im};
    const Cx d = csqrt(cadd(cmul(beta, beta), cscale(cadd(iu, cmul(u, u)), h.sigma2)));
    const Cx bmd = csub(beta, d);
    const Cx g = cdiv(bmd, cadd(beta, d));
    const Cx e = cexp(cscale(d, -h.t));
    const Cx one_m_ge = {1.0 - (g.re * e.re - g.im * e.im), -(g.re * e.im + g.im * e.re)};
    const Cx lg = clog(cdiv(one_m_ge, {1.0 - g.re, -g.im}));
    const Cx C = cadd(cscale(iu, h.drift),
                      cscale(csub(cscale(bmd, h.t), cscale(lg, 2.0)), h.kt_inv_s2));
    const Cx D = cmul(cscale(bmd, h.inv_sigma2), cdiv({1.0 - e.re, -e.im}, one_m_ge));
    const Cx x = cadd(C, cscale(D, h.v0));
    return cexp({x.re, x.im - rot});
}

__device__ __forceinline__ double wave_sum64(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

constexpr int kCfBlock = 256;  // 4 options per workgroup

__global__ __launch_bounds__(kCfBlock) void k_heston_cf(HestonCfIn in, int32_t n,
                                                       const double* __restrict__ gl,
                                                       double* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int node = lane >> 1, side = lane & 1;
    const double xi = gl[node], wi = gl[32 + node];  // nodes and weights on [0, 1]
    for (int32_t i = blockIdx.x * (kCfBlock / 64) + (threadIdx.x >> 6); i < n;
         i += gridDim.x * (kCfBlock / 64)) {
        const double s = in.spot[i], k = in.strike[i], t = in.t[i], r = in.r[i], q = in.q[i];
        const double v0 = in.v0[i], ka = in.kappa[i], th = in.theta[i], sg = in.sigma[i],
                     rho = in.rho[i], lim = in.limit[i];
        const int32_t ot = in.ot[i], steps = in.steps[i];
        double price;
        // heston.cpp:10-24
        if ((ot != QK_CALL && ot != QK_PUT) || !finite_d(s) || !finite_d(k) || !finite_d(t) ||
            !finite_d(r) || !finite_d(q) || !finite_d(v0) || !finite_d(ka) || !finite_d(th) ||
            !finite_d(sg) || !finite_d(rho) || s <= 0.0 || k <= 0.0 || t < 0.0 || v0 < 0.0 ||
            ka <= 0.0 || th < 0.0 || sg <= 0.0 || rho <= -1.0 || rho >= 1.0) {
            price = qnan_d();
        } else if (t <= kEps) {
            price = intrinsic(s, k, ot);
        } else if (steps < 16 || lim <= 1e-6) {
            price = qnan_d();
        } else {
            HestonCf h;
            h.v0 = v0;
            h.kappa = ka;
            h.rho_sigma = rho * sg;
            h.sigma2 = sg * sg;
            h.inv_sigma2 = 1.0 / h.sigma2;
            h.kt_inv_s2 = ka * th * h.inv_sigma2;
            h.drift = log(s) + (r - q) * t;
            h.t = t;
            const double log_k = log(k);
            // internal_util.h:83-95: phi(-i) normalises P1; panels of 64 nodes
            const Cx phi_mi = heston_phi(h, {0.0, -1.0}, 0.0);
            const double phi_abs = hypot(phi_mi.re, phi_mi.im);
            const bool p1_ok = phi_abs >= 1e-14;
            const double phi_abs2 = phi_mi.re * phi_mi.re + phi_mi.im * phi_mi.im;
            const int32_t panels = steps > 0 ? (steps + 63) / 64 : 4;
            const double a = 1e-8, width = (lim - a) / double(panels);
            double s1 = 0.0, s2 = 0.0;
            for (int32_t p = 0; p < panels; ++p) {
                const double pa = a + double(p) * width, pb = pa + width;
                const double mid = 0.5 * (pa + pb), half = 0.5 * (pb - pa);
                const double u = side == 0 ? mid - half * xi : mid + half * xi;
                const double inv_u = 1.0 / u, rot = u * log_k;
                // e^{-i u ln K} phi(u) and e^{-i u ln K} phi(u - i)
                const Cx cf2 = heston_phi(h, {u, 0.0}, rot);
                s2 += half * wi * (cf2.im * inv_u);
                if (p1_ok) {
                    const Cx n1 = heston_phi(h, {u, -1.0}, rot);
                    const double d_re = n1.im * inv_u, d_im = -n1.re * inv_u;
                    s1 += half * wi * ((d_re * phi_mi.re + d_im * phi_mi.im) / phi_abs2);
                }
            }
            s1 = wave_sum64(s1);
            s2 = wave_sum64(s2);
            const double kPi = 3.1415926535897932384626433832795;
            const double p2 = fmin(1.0, fmax(0.0, 0.5 + s2 / kPi));
            const double p1 = p1_ok ? fmin(1.0, fmax(0.0, 0.5 + s1 / kPi)) : qnan_d();
            if (!finite_d(p1) || !finite_d(p2)) {
                price = qnan_d();
            } else {
                // heston.cpp:49-52 and internal_util.h:160-167
                const double call = s * exp(-q * t) * p1 - k * exp(-r * t) * p2;
                price = ot == QK_CALL ? call : call - s * exp(-q * t) + k * exp(-r * t);
            }
        }
        if (lane == 0) out[i] = price;
    }
}

}  // namespace

hipError_t launch_heston_cf(hipStream_t st, const HestonCfIn& in, int32_t n, const double* gl,
                            double* out, int32_t grid) {
    hipLaunchKernelGGL(k_heston_cf, dim3(grid), dim3(kCfBlock), 0, st, in, n, gl, out);
    return hipGetLastError();
}

}  // namespace qk
