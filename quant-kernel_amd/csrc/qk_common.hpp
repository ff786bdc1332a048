// qk_common.hpp - definitions shared by the host side and the gfx950 kernels
// of the Monte Carlo batch-pricing library: per-option records, input
// classification (the reference's validation and degenerate branches) and
// the per-model constants the path kernels consume.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "quantkernel/qk_abi.h"

namespace qk {

#define QK_HD __host__ __device__ __forceinline__

constexpr double kEps = 1e-12;           // mc internal_util.h:13
constexpr double kLog2e = 1.4426950408889634073599246810019;

// Estimator family of a prepared option (what kernel simulates it).
enum Family : int32_t {
    FAM_NONE = 0,      // resolved without simulation (NaN / intrinsic / closed form)
    FAM_TERMINAL = 1,  // GBM terminal value, antithetic pairs (mc_engine.h:32-53)
    FAM_STEPWISE = 2,  // Euler / Milstein / local vol / Heston (mc_engine.h:84-136)
};

// Which reference entry point's rules apply to an element.
enum Method : int32_t {
    M_STANDARD = 0,    // standard_monte_carlo.cpp:11-33 (has the vol <= eps branch)
    M_ANTITHETIC = 1,  // antithetic_variates.cpp:11-29
    M_EULER = 2,       // euler_maruyama.cpp:11-28
    M_MILSTEIN = 3,    // milstein.cpp:11-28
    M_HESTON = 4,      // heston_monte_carlo.cpp:11-36
    M_LOCAL_VOL = 5,   // local_vol_monte_carlo.cpp:57-73
    M_IMPORTANCE = 6,  // importance_sampling.cpp:11-36 (estimate_terminal, shifted draws)
    M_PATHWISE = 7,    // pathwise_derivative.cpp:11-40 (estimate_terminal)
    M_LR = 8,          // likelihood_ratio.cpp:12-42 (estimate_terminal_antithetic)
    M_HESTON_LR = 9,   // heston_lr_delta.cpp:41-132 (antithetic Heston pairs)
    M_CONTROL = 10,    // control_variates.cpp:11-54 (control e^{-rT} S_T)
    M_STRATIFIED = 11, // stratified_sampling.cpp:11-35 (one uniform per stratum)
    M_SOBOL = 12,      // quasi_monte_carlo.cpp:10-32 (1-D Sobol points)
    M_HALTON = 13,     // quasi_monte_carlo.cpp:34-56 (base-2 Halton points)
    M_MLMC_LEVEL = 14, // one level of multilevel_monte_carlo.cpp:10-72 (aux = level)
};

// Terminal methods simulate S_T in one draw; the antithetic ones use a pair
// (z, -z) per unit.
QK_HD bool method_is_terminal(int32_t m) {
    return m == M_STANDARD || m == M_ANTITHETIC || m == M_IMPORTANCE || m == M_PATHWISE ||
           m == M_LR || m == M_CONTROL || m == M_STRATIFIED || m == M_SOBOL || m == M_HALTON;
}
// Units per lane group of a terminal estimator: eight for those that draw
// normals (one Philox block gives four (radius, angle) pairs), four for the
// uniform-driven ones (stratified, Sobol, Halton).
QK_HD int32_t term_group_units(int32_t m) {
    return (m == M_STRATIFIED || m == M_SOBOL || m == M_HALTON) ? 4 : 8;
}
// Deterministic point sets: no stream, and no sampling-error estimate.
QK_HD bool method_is_qmc(int32_t m) { return m == M_SOBOL || m == M_HALTON; }
QK_HD bool method_is_antithetic_terminal(int32_t m) {
    return m == M_STANDARD || m == M_ANTITHETIC || m == M_LR;
}

QK_HD double qnan_d() { return __builtin_nan(""); }

QK_HD bool finite_d(double x) {          // math_util.h:17-22
    union { double d; uint64_t u; } b;
    b.d = x;
    return (b.u & 0x7FF0000000000000ULL) != 0x7FF0000000000000ULL;
}

QK_HD double intrinsic(double x, double k, int32_t ot) {   // option_util.h:27-31
    if (ot == QK_CALL) return x - k > 0.0 ? x - k : 0.0;
    if (ot == QK_PUT) return k - x > 0.0 ? k - x : 0.0;
    return qnan_d();
}

QK_HD bool common_ok(double s, double k, double t, double vol, double r, double q, int32_t ot) {
    // option_util.h:35-43
    if (ot != QK_CALL && ot != QK_PUT) return false;
    if (!finite_d(s) || !finite_d(k) || !finite_d(t) || !finite_d(vol) || !finite_d(r) ||
        !finite_d(q))
        return false;
    return s > 0.0 && k > 0.0 && t >= 0.0 && vol >= 0.0;
}

// Raw per-option inputs of any method (unused fields are ignored).
struct RawOption {
    int32_t method;
    double spot, strike, t, vol, r, q;
    double v0, kappa, theta, sigma, rho;
    int32_t ot, paths, steps;
    uint64_t seed;
    double aux;  // importance sampling: shift; LR deltas: weight clip; MLMC: level
};

// Prepared option: everything a path kernel and the finaliser need.
// 192 bytes, 16-byte aligned, written by the prepare step.
struct alignas(16) OptionRec {
    int32_t family;      // Family
    int32_t method;      // Method
    int32_t ot;          // QK_CALL / QK_PUT
    int32_t units;       // sample units: paths (stepwise) or pairs (terminal)
    int32_t paths;
    int32_t steps;       // stepwise: steps; terminal: 1
    uint32_t seed_lo, seed_hi;
    int64_t stream_off;  // offset of this option's host normal stream (validation modes)
    int32_t ntiles;
    int32_t tile_first;  // filled by the scan
    // FP64 constants (kernels convert to the path precision once per tile)
    double spot, strike;
    double disc;         // e^{-rT}
    double shift;        // payoff of the deterministic forward path; centring for the sums
    double c[8];         // model constants, see prepare_option()
    double closed;       // result when family == FAM_NONE
    double v0;           // Heston initial variance
    // Greeks and importance sampling:
    //  aux[0]: importance-sampling shift h, or the LR weight clip
    //  aux[1]: e^{-qT}, the delta clamp (pathwise and LR delta)
    //  aux[2]: LR score denominator: vol sqrt(T) spot (GBM) or
    //          sqrt(v0) sqrt(dt) spot (Heston; 0 when sqrt(v0) <= 1e-15: score 0)
    //  control variates: aux[0] = E[S_T] = S e^{(r-q)T} (the control's mean,
    //          undiscounted); MLMC level: aux[0] = level
    double aux[4];
};
static_assert(sizeof(OptionRec) == 192, "OptionRec layout");

// Model constants c[] by method (path kernels work on log2(S/spot) for
// Heston and on S for the GBM schemes):
//  terminal: c0 = (r-q-vol^2/2) T, c1 = vol sqrt(T)              (model_concepts.h:93-102)
//  Euler/LV: c0 = 1 + (r-q) dt, c1 = vol sqrt(dt)                 (model_concepts.h:136-143, 257-266)
//  Milstein: c0 = 1 + (r-q) dt - vol^2 dt/2, c1 = vol sqrt(dt), c2 = vol^2 dt/2   (model_concepts.h:150-158)
//  Heston:   c0 = (r-q) dt log2e, c1 = dt/2 log2e, c2 = sqrt(dt) log2e,
//            c3 = kappa dt, c4 = kappa theta dt, c5 = sigma sqrt(dt),
//            c6 = rho, c7 = sqrt(max(1-rho^2,0))                  (model_concepts.h:205-226)
//  The Heston spot floor 1e-12 becomes log2(1e-12/spot), computed in-kernel.

// Classify one element and fill its record.  Returns the family.  Shared by
// the host (validation-mode stream sizing) and the prepare kernel so both
// make the same decision.
QK_HD int32_t prepare_option(const RawOption& in, int32_t tile_units, OptionRec& rec) {
    rec.family = FAM_NONE;
    rec.method = in.method;
    rec.ot = in.ot;
    rec.units = 0;
    rec.paths = in.paths;
    rec.steps = in.steps;
    rec.seed_lo = uint32_t(in.seed);
    rec.seed_hi = uint32_t(in.seed >> 32);
    rec.stream_off = 0;
    rec.ntiles = 0;
    rec.tile_first = 0;
    rec.spot = in.spot;
    rec.strike = in.strike;
    rec.disc = 0.0;
    rec.shift = 0.0;
    for (int i = 0; i < 8; ++i) rec.c[i] = 0.0;
    rec.closed = qnan_d();
    rec.v0 = 0.0;
    for (int i = 0; i < 4; ++i) rec.aux[i] = 0.0;

    const double s = in.spot, k = in.strike, t = in.t, r = in.r, q = in.q;
    const int32_t ot = in.ot;

    if (in.method == M_HESTON || in.method == M_HESTON_LR) {
        const bool lr = in.method == M_HESTON_LR;
        // heston_monte_carlo.cpp:18-24: no finiteness test on the inputs.
        // Non-finite Heston parameters give NaN here (the reference's result
        // for them is an artefact of std::max's NaN ordering).
        if ((ot != QK_CALL && ot != QK_PUT) || !(s > 0.0) || !(k > 0.0) || !(t >= 0.0) ||
            !(in.v0 >= 0.0) || !(in.kappa >= 0.0) || !(in.theta >= 0.0) ||
            !(in.sigma >= 0.0) || !(in.rho >= -1.0) || !(in.rho <= 1.0) || in.paths <= 1 ||
            in.steps <= 0)
            return FAM_NONE;
        // heston_lr_delta.cpp:53: weight_clip <= 0 is invalid; +inf means no clipping
        if (lr && !(in.aux > 0.0)) return FAM_NONE;
        if (!finite_d(s) || !finite_d(k) || !finite_d(t) || !finite_d(r) || !finite_d(q) ||
            !finite_d(in.v0) || !finite_d(in.kappa) || !finite_d(in.theta) ||
            !finite_d(in.sigma))
            return FAM_NONE;
        if (t <= kEps) {
            if (lr)  // heston_lr_delta.cpp:57-61: the intrinsic delta
                rec.closed = ot == QK_CALL ? (s > k ? 1.0 : 0.0) : (s < k ? -1.0 : 0.0);
            else
                rec.closed = intrinsic(s, k, ot);
            return FAM_NONE;
        }
        const double dt = t / double(in.steps);
        const double sdt = sqrt(dt);
        rec.disc = exp(-r * t);
        rec.c[0] = (r - q) * dt * kLog2e;
        rec.c[1] = 0.5 * dt * kLog2e;
        rec.c[2] = sdt * kLog2e;
        rec.c[3] = in.kappa * dt;
        rec.c[4] = in.kappa * in.theta * dt;
        rec.c[5] = in.sigma * sdt;
        rec.c[6] = in.rho;
        const double one_m = 1.0 - in.rho * in.rho;
        rec.c[7] = sqrt(one_m > 0.0 ? one_m : 0.0);
        rec.v0 = in.v0;
        rec.family = FAM_STEPWISE;
        if (lr) {
            // unit = one antithetic pair; an odd last path is simulated apart
            const double sv0 = sqrt(in.v0);
            rec.aux[0] = in.aux;
            rec.aux[2] = sv0 > 1e-15 ? sv0 * sdt * s : 0.0;
            rec.units = in.paths / 2;
        } else {
            rec.shift = intrinsic(s * exp((r - q) * t), k, ot);
            rec.units = in.paths;
        }
    } else if (in.method == M_PATHWISE || in.method == M_LR) {
        // pathwise_derivative.cpp:15-22, likelihood_ratio.cpp:15-23
        if (!common_ok(s, k, t, in.vol, r, q, ot) || in.paths < 256) return FAM_NONE;
        if (in.method == M_LR && (!finite_d(in.aux) || in.aux <= 0.0)) return FAM_NONE;
        if (t <= kEps || in.vol <= kEps) {
            // adjoint_greeks/common/internal_util.h:22-31
            const double fwd = s * exp((r - q) * t);
            const double qf = exp(-q * t);
            if (fabs(fwd - k) < kEps * k) rec.closed = ot == QK_CALL ? 0.5 * qf : -0.5 * qf;
            else if (ot == QK_CALL) rec.closed = fwd > k ? qf : 0.0;
            else rec.closed = fwd < k ? -qf : 0.0;
            return FAM_NONE;
        }
        rec.disc = exp(-r * t);
        rec.c[0] = (r - q - 0.5 * in.vol * in.vol) * t;
        rec.c[1] = in.vol * sqrt(t);
        rec.aux[1] = exp(-q * t);
        rec.family = FAM_TERMINAL;
        rec.steps = 1;
        if (in.method == M_LR) {
            rec.aux[0] = in.aux;
            rec.aux[2] = in.vol * sqrt(t) * s;
            rec.units = in.paths / 2;
        } else {
            rec.units = in.paths;
        }
    } else if (in.method == M_CONTROL || in.method == M_STRATIFIED || in.method == M_SOBOL ||
               in.method == M_HALTON) {
        // control_variates.cpp:14-19 (paths > 2), stratified_sampling.cpp:14-19,
        // quasi_monte_carlo.cpp:13-18 (paths > 1); no vol <= eps branch
        const int32_t min_paths = in.method == M_CONTROL ? 3 : 2;
        if (!common_ok(s, k, t, in.vol, r, q, ot) || in.paths < min_paths) return FAM_NONE;
        if (t <= kEps) {
            rec.closed = intrinsic(s, k, ot);
            return FAM_NONE;
        }
        rec.disc = exp(-r * t);
        rec.c[0] = (r - q - 0.5 * in.vol * in.vol) * t;
        rec.c[1] = in.vol * sqrt(t);
        const double fwd = s * exp((r - q) * t);
        rec.shift = intrinsic(fwd, k, ot);
        rec.aux[0] = fwd;
        rec.family = FAM_TERMINAL;
        rec.units = in.paths;
        rec.steps = 1;
    } else if (in.method == M_MLMC_LEVEL) {
        // one level of multilevel_monte_carlo.cpp:24-66: paths = max(16,
        // base_paths >> level) and steps = base_steps << level are set by the
        // host, which also validates base_paths, levels and base_steps
        if (!common_ok(s, k, t, in.vol, r, q, ot) || in.paths <= 1 || in.steps <= 0)
            return FAM_NONE;
        const bool fine_only = !(in.aux > 0.0);  // level 0
        if (t <= kEps) {
            rec.closed = fine_only ? intrinsic(s, k, ot) : 0.0;  // the levels sum to intrinsic
            return FAM_NONE;
        }
        const double dt = t / double(in.steps);
        rec.disc = exp(-r * t);
        rec.c[0] = 1.0 + (r - q) * dt;
        rec.c[1] = in.vol * sqrt(dt);
        rec.c[2] = 1.0 + (r - q) * (2.0 * dt);
        rec.c[3] = in.vol * sqrt(2.0 * dt);
        rec.aux[0] = in.aux;
        rec.shift = fine_only ? intrinsic(s * exp((r - q) * t), k, ot) : 0.0;
        rec.family = FAM_STEPWISE;
        rec.units = in.paths;
    } else if (in.method == M_IMPORTANCE) {
        // importance_sampling.cpp:15-20
        if (!common_ok(s, k, t, in.vol, r, q, ot) || in.paths <= 1 || !finite_d(in.aux))
            return FAM_NONE;
        if (t <= kEps) {
            rec.closed = intrinsic(s, k, ot);
            return FAM_NONE;
        }
        rec.disc = exp(-r * t);
        rec.c[0] = (r - q - 0.5 * in.vol * in.vol) * t;
        rec.c[1] = in.vol * sqrt(t);
        rec.aux[0] = in.aux;
        rec.shift = intrinsic(s * exp((r - q) * t), k, ot);
        rec.family = FAM_TERMINAL;
        rec.units = in.paths;
        rec.steps = 1;
    } else if (in.method == M_STANDARD || in.method == M_ANTITHETIC) {
        // standard_monte_carlo.cpp:14-23, antithetic_variates.cpp:14-19
        if (!common_ok(s, k, t, in.vol, r, q, ot) || in.paths <= 1) return FAM_NONE;
        if (t <= kEps) {
            rec.closed = intrinsic(s, k, ot);
            return FAM_NONE;
        }
        if (in.method == M_STANDARD && in.vol <= kEps) {
            const double fwd = s * exp((r - q) * t);
            rec.closed = exp(-r * t) * intrinsic(fwd, k, ot);
            return FAM_NONE;
        }
        rec.disc = exp(-r * t);
        rec.c[0] = (r - q - 0.5 * in.vol * in.vol) * t;
        rec.c[1] = in.vol * sqrt(t > 0.0 ? t : 0.0);
        rec.shift = intrinsic(s * exp((r - q) * t), k, ot);
        rec.family = FAM_TERMINAL;
        rec.units = in.paths / 2;  // pairs; an odd path is handled by the last pair's lane
        rec.steps = 1;
    } else {
        // euler_maruyama.cpp:15-19, milstein.cpp:15-19, local_vol_monte_carlo.cpp:61-66
        if (!common_ok(s, k, t, in.vol, r, q, ot) || in.paths <= 1 || in.steps <= 0)
            return FAM_NONE;
        if (t <= kEps) {
            rec.closed = intrinsic(s, k, ot);
            return FAM_NONE;
        }
        const double dt = t / double(in.steps);
        const double sdt = sqrt(dt);
        const double mu = r - q;
        rec.disc = exp(-r * t);
        if (in.method == M_MILSTEIN) {
            const double h = 0.5 * in.vol * in.vol * dt;
            rec.c[0] = 1.0 + mu * dt - h;
            rec.c[1] = in.vol * sdt;
            rec.c[2] = h;
        } else {
            rec.c[0] = 1.0 + mu * dt;
            rec.c[1] = in.vol * sdt;
        }
        rec.shift = intrinsic(s * exp(mu * t), k, ot);
        rec.family = FAM_STEPWISE;
        rec.units = in.paths;
    }
    rec.ntiles = (rec.units + tile_units - 1) / tile_units;
    if (rec.units == 0) rec.ntiles = 1;  // every simulated option owns at least one tile
    return rec.family;
}

// Validation-mode stream geometry: units that draw from the host stream
// (paths, pairs, plus an odd path) and draws per unit.
QK_HD int64_t stream_units(const OptionRec& rec) {
    if (rec.family == FAM_TERMINAL) {
        if (method_is_qmc(rec.method)) return 0;
        return method_is_antithetic_terminal(rec.method)
                   ? int64_t(rec.paths / 2) + (rec.paths & 1)
                   : int64_t(rec.paths);
    }
    if (rec.family == FAM_STEPWISE)
        return rec.method == M_HESTON_LR ? int64_t(rec.paths / 2) + (rec.paths & 1)
                                         : int64_t(rec.paths);
    return 0;
}
QK_HD int64_t stream_draws_per_unit(const OptionRec& rec) {
    if (rec.family == FAM_STEPWISE)
        return int64_t(rec.steps) * (rec.method == M_HESTON || rec.method == M_HESTON_LR ? 2 : 1);
    return rec.family == FAM_TERMINAL ? 1 : 0;
}
// Number of host-stream normals an option consumes in the validation modes.
QK_HD int64_t stream_normals(const OptionRec& rec) {
    return stream_units(rec) * stream_draws_per_unit(rec);
}

// Per-tile partial sums of centred payoff units (64 bytes: one HBM line).
struct alignas(64) Partial {
    double s1;     // sum of (unit - shift)
    double s2;     // sum of (unit - shift)^2
    double extra;  // payoff of an odd extra path (0 otherwise)
    double x1;     // control variates: sum of (S_T - E[S_T])
    double x2;     //   sum of (S_T - E[S_T])^2
    double xy;     //   sum of (S_T - E[S_T]) (unit - shift)
    double pad[2];
};
static_assert(sizeof(Partial) == 64, "Partial layout");

}  // namespace qk
