// qk_oracle.cpp - CPU oracle of QuantKernel's Monte Carlo batch-pricing path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: only the test suite,
// the smoke check and the CPU-baseline leg of the benchmark may load it.  The
// product library never links or calls it.
//
// Contents (each function names the reference file:line it restates; all
// paths are relative to the reference repository root):
//   1. The reference MC path in FP64 with its own normal stream
//      (std::mt19937_64 + std::normal_distribution<double> from libstdc++;
//      GCC 11.4 headers in this image, /usr/include/c++/11/bits/random.tcc:
//      1811-1829, Marsaglia polar method).  Same draw order, floors,
//      degenerate branches, summation order and discount placement, so it is
//      bit-identical to the reference when both are built the same way.
//      Pinned by the reference goldens (cpp/tests/test_monte_carlo.cpp:57-73)
//      and, where no golden exists, by a build of the reference itself.
//   2. The same estimators in FP64 fed by the Philox4x32-10 + Box-Muller
//      stream that the GPU kernels use (the "FP64 MC oracle" for the GPU's
//      production mode).  Philox is pinned by the Random123 known-answer
//      vectors (Salmon et al., SC'11), Box-Muller by distribution tests.
//   3. Closed forms used as independent checks: Black-Scholes-Merton (erfc
//      and the reference's A&S 7.1.26 variant), the Heston characteristic-
//      function price (Heston 1993, Albrecher et al. 2007 form), and a CRR
//      binomial tree (European / American).
//   4. A host-side Philox MLP helper is NOT here: the surrogate's CPU checks
//      live in the Python tests (numpy).
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <random>
#include <vector>

#define ORC_EXPORT extern "C" __attribute__((visibility("default")))

namespace {

constexpr double kEps = 1e-12;   // internal_util.h:13

double qnan() {                  // math_util.h:12-15 (quiet NaN bit pattern)
    const uint64_t bits = 0x7FF8000000000000ULL;
    double d;
    std::memcpy(&d, &bits, sizeof d);
    return d;
}

bool finite_bits(double x) {     // math_util.h:17-22
    uint64_t b;
    std::memcpy(&b, &x, sizeof b);
    return (b & 0x7FF0000000000000ULL) != 0x7FF0000000000000ULL;
}

bool type_ok(int32_t ot) { return ot == 0 || ot == 1; }   // option_util.h:21-23

double intrinsic(double x, double k, int32_t ot) {        // option_util.h:27-31
    if (ot == 0) return std::max(0.0, x - k);
    if (ot == 1) return std::max(0.0, k - x);
    return qnan();
}

bool common_ok(double s, double k, double t, double vol, double r, double q, int32_t ot) {
    // option_util.h:35-43
    if (!type_ok(ot)) return false;
    if (!finite_bits(s) || !finite_bits(k) || !finite_bits(t) || !finite_bits(vol) ||
        !finite_bits(r) || !finite_bits(q))
        return false;
    return s > 0.0 && k > 0.0 && t >= 0.0 && vol >= 0.0;
}

bool counts_ok(int32_t paths, int32_t steps) { return paths > 1 && steps > 0; }  // internal_util.h:24-26

// mc_engine.h:15-23: one stateful N(0,1) stream per option.
struct MtStream {
    std::mt19937_64 eng;
    std::normal_distribution<double> nd{0.0, 1.0};
    explicit MtStream(uint64_t s) : eng(s) {}
    double operator()() { return nd(eng); }
};

// model_concepts.h:93-102 (also internal_util.h:32-36).
inline double gbm_terminal(double spot, double t, double vol, double r, double q, double z) {
    const double mu_t = (r - q - 0.5 * vol * vol) * t;
    const double shock = vol * std::sqrt(std::max(t, 0.0)) * z;
    return spot * std::exp(mu_t + shock);
}

// ---------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel random numbers: as
// easy as 1, 2, 3", SC'11).  Round: (hi,lo)=M*x on words 0 and 2, then
// out = {hi2^x1^k0, lo2, hi0^x3^k1, lo0}; the key gains the Weyl constants
// between rounds.
// ---------------------------------------------------------------------------
void philox10(const uint32_t in[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t x0 = in[0], x1 = in[1], x2 = in[2], x3 = in[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int round = 0; round < 10; ++round) {
        const uint64_t pa = uint64_t(0xD2511F53u) * x0;
        const uint64_t pb = uint64_t(0xCD9E8D57u) * x2;
        const uint32_t n0 = uint32_t(pb >> 32) ^ x1 ^ k0;
        const uint32_t n2 = uint32_t(pa >> 32) ^ x3 ^ k1;
        x1 = uint32_t(pb);
        x3 = uint32_t(pa);
        x0 = n0;
        x2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = x0; out[1] = x1; out[2] = x2; out[3] = x3;
}

// Box-Muller (1958) on two 32-bit words, with exactly the uniforms the GPU
// forms: u1 = fmaf(float(a), 2^-32, 2^-33) in FP32 (in (0, 1]) and the angle
// fraction f = (b & 0x7FFFFF) / 2^23; (z0, z1) = sqrt(-2 ln u1) *
// (cos 2 pi f, sin 2 pi f), evaluated here in FP64.
void box_muller(uint32_t a, uint32_t b, double* z0, double* z1) {
    const double u1 = static_cast<double>(std::fmaf(static_cast<float>(a), 0x1p-32f, 0x1p-33f));
    const double f = static_cast<double>(b & 0x007FFFFFu) * (1.0 / 8388608.0);
    const double rad = std::sqrt(-2.0 * std::log(u1));
    const double ang = 6.283185307179586476925286766559 * f;
    *z0 = rad * std::cos(ang);
    *z1 = rad * std::sin(ang);
}

void philox_normals4(uint32_t c0, uint32_t c1, uint32_t c3, uint64_t seed, double z[4]) {
    const uint32_t in[4] = {c0, c1, 0u, c3};
    const uint32_t key[2] = {uint32_t(seed), uint32_t(seed >> 32)};
    uint32_t w[4];
    philox10(in, key, w);
    box_muller(w[0], w[1], &z[0], &z[1]);
    box_muller(w[2], w[3], &z[2], &z[3]);
}

// Draw d of path p of an LSMC option: block (p, d/4), tag 0, component d%4.
double philox_path_normal(uint64_t seed, uint32_t p, uint32_t d) {
    double z[4];
    philox_normals4(p, d >> 2, 0u, seed, z);
    return z[d & 3u];
}

// Four-pair layout of one Philox block (every production stream): slot s is word
// s alone, u1 = (2m + 1) 2^-24 with m its low 23 bits and a 9-bit angle
// index a = its top 9 bits; (z1, z2) = sqrt(-2 ln u1) (cos 2 pi a / 512,
// sin 2 pi a / 512) in FP64 (the GPU reads FP32 roundings of the same
// cosine and sine from a table).
void philox_quad_pair(const uint32_t in[4], uint64_t seed, uint32_t s, double* z1, double* z2) {
    const uint32_t key[2] = {uint32_t(seed), uint32_t(seed >> 32)};
    uint32_t w[4];
    philox10(in, key, w);
    const double u1 = (2.0 * double(w[s] & 0x007FFFFFu) + 1.0) * 0x1p-24;
    const double f = double(w[s] >> 23) * (1.0 / 512.0);
    const double rad = std::sqrt(-2.0 * std::log(u1));
    const double ang = 6.283185307179586476925286766559 * f;
    *z1 = rad * std::cos(ang);
    *z2 = rad * std::sin(ang);
}

// Heston (production) stream: step j of path p uses slot j%4 of block
// (p, j/4, 0, 0) in the four-pair layout.
void philox_heston_normals(uint64_t seed, uint32_t p, uint32_t j, double* z1, double* z2) {
    const uint32_t in[4] = {p, j / 4u, 0u, 0u};
    philox_quad_pair(in, seed, j % 4u, z1, z2);
}

// GBM / local-vol stepwise (production) stream: step j of path p takes
// component j%2 of the Heston pair of step j/2, i.e. slot (j%8)/2 of block
// (p, j/8, 0, 0), the cosine for even j and the sine for odd j: eight
// normals per Philox block.
double philox_gbm_normal(uint64_t seed, uint32_t p, uint32_t j) {
    double z1, z2;
    philox_heston_normals(seed, p, j >> 1, &z1, &z2);
    return (j & 1u) ? z2 : z1;
}

// Terminal estimators: draw k is component k%8 of block (k/8, 0, 0, 1) in
// the four-pair layout: slot (k%8)/2, the cosine for even k, the sine for
// odd k.
double philox_terminal_normal(uint64_t seed, uint32_t k) {
    const uint32_t in[4] = {k / 8u, 0u, 0u, 1u};
    double z1, z2;
    philox_quad_pair(in, seed, (k % 8u) / 2u, &z1, &z2);
    return (k & 1u) ? z2 : z1;
}

struct Moments {  // running sums of payoff units
    double s1 = 0.0, s2 = 0.0;
    void add(double x) { s1 += x; s2 += x * x; }
};

double se_of(const Moments& m, double units, double disc) {
    if (units < 2.0) return 0.0;
    const double mean = m.s1 / units;
    double var = (m.s2 - m.s1 * mean) / (units - 1.0);
    if (var < 0.0) var = 0.0;
    return disc * std::sqrt(var / units);
}

// --------------------------- Gauss-Legendre nodes --------------------------
// 64-point rule computed by Newton iteration on P_64 (no tables).
struct GL64 {
    double x[32], w[32];
    GL64() {
        const int n = 64;
        for (int i = 0; i < 32; ++i) {
            double z = std::cos(M_PI * (i + 0.75) / (n + 0.5));
            double dp = 0.0;
            for (int it = 0; it < 100; ++it) {
                double p0 = 1.0, p1 = z;
                for (int k = 2; k <= n; ++k) {
                    const double pk = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                    p0 = p1;
                    p1 = pk;
                }
                dp = n * (z * p1 - p0) / (z * z - 1.0);
                const double dz = p1 / dp;
                z -= dz;
                if (std::fabs(dz) < 1e-16) break;
            }
            // recompute derivative at the converged node
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; ++k) {
                const double pk = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = pk;
            }
            dp = n * (z * p1 - p0) / (z * z - 1.0);
            x[31 - i] = z;
            w[31 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
        }
    }
};
const GL64& gl64() {
    static const GL64 g;
    return g;
}

}  // namespace

// ===========================================================================
// 1. Reference path, mt19937_64 stream, FP64.
// ===========================================================================

// The first `count` normals of the per-option stream (mc_engine.h:15-23).
ORC_EXPORT void orc_mt_normals(uint64_t seed, int64_t count, double* out) {
    MtStream g(seed);
    for (int64_t i = 0; i < count; ++i) out[i] = g();
}

// standard_monte_carlo.cpp:11-33 with mc_engine.h:32-53 (antithetic loop).
// antithetic_variates.cpp:11-29 is the same without the vol <= eps branch.
static double terminal_antithetic_price(double spot, double strike, double t, double vol,
                                        double r, double q, int32_t ot, int32_t paths,
                                        uint64_t seed, bool vol_branch) {
    if (!common_ok(spot, strike, t, vol, r, q, ot) || paths <= 1) return qnan();
    if (t <= kEps) return intrinsic(spot, strike, ot);
    if (vol_branch && vol <= kEps) {
        const double fwd = spot * std::exp((r - q) * t);
        return std::exp(-r * t) * intrinsic(fwd, strike, ot);
    }
    const double disc = std::exp(-r * t);
    MtStream g(seed);
    double acc = 0.0;
    const int32_t pairs = paths / 2;
    for (int32_t i = 0; i < pairs; ++i) {
        const double z = g();
        const double up = gbm_terminal(spot, t, vol, r, q, z);
        const double dn = gbm_terminal(spot, t, vol, r, q, -z);
        acc += intrinsic(up, strike, ot);
        acc += intrinsic(dn, strike, ot);
    }
    if (paths & 1) {
        const double z = g();
        acc += intrinsic(gbm_terminal(spot, t, vol, r, q, z), strike, ot);
    }
    const double mean = acc / static_cast<double>(paths);
    return disc * mean;
}

ORC_EXPORT double orc_standard_mc(double spot, double strike, double t, double vol, double r,
                                  double q, int32_t ot, int32_t paths, uint64_t seed) {
    return terminal_antithetic_price(spot, strike, t, vol, r, q, ot, paths, seed, true);
}

ORC_EXPORT double orc_antithetic(double spot, double strike, double t, double vol, double r,
                                 double q, int32_t ot, int32_t paths, uint64_t seed) {
    return terminal_antithetic_price(spot, strike, t, vol, r, q, ot, paths, seed, false);
}

// euler_maruyama.cpp:11-28 / milstein.cpp:11-28 with mc_engine.h:84-103 and
// model_concepts.h:136-158.  scheme 0 = Euler, 1 = Milstein.
static double stepwise_price(double spot, double strike, double t, double vol, double r,
                             double q, int32_t ot, int32_t paths, int32_t steps, uint64_t seed,
                             int scheme) {
    if (!common_ok(spot, strike, t, vol, r, q, ot) || !counts_ok(paths, steps)) return qnan();
    if (t <= kEps) return intrinsic(spot, strike, ot);
    const double disc = std::exp(-r * t);
    const double mu = r - q;
    MtStream g(seed);
    const double dt = t / static_cast<double>(steps);
    const double sdt = std::sqrt(dt);
    double acc = 0.0;
    for (int32_t i = 0; i < paths; ++i) {
        double s = spot;
        for (int32_t j = 0; j < steps; ++j) {
            const double dw = sdt * g();
            double nxt;
            if (scheme == 0)
                nxt = s + mu * s * dt + vol * s * dw;
            else
                nxt = s + mu * s * dt + vol * s * dw + 0.5 * vol * vol * s * (dw * dw - dt);
            s = std::max(1e-12, nxt);
        }
        acc += intrinsic(s, strike, ot);
    }
    const double mean = acc / static_cast<double>(paths);
    return disc * mean;
}

ORC_EXPORT double orc_euler(double spot, double strike, double t, double vol, double r, double q,
                            int32_t ot, int32_t paths, int32_t steps, uint64_t seed) {
    return stepwise_price(spot, strike, t, vol, r, q, ot, paths, steps, seed, 0);
}

ORC_EXPORT double orc_milstein(double spot, double strike, double t, double vol, double r,
                               double q, int32_t ot, int32_t paths, int32_t steps, uint64_t seed) {
    return stepwise_price(spot, strike, t, vol, r, q, ot, paths, steps, seed, 1);
}

// local_vol_monte_carlo.cpp:20-48, 57-73 with model_concepts.h:240-266.
// Own loop with a time counter; result is disc*sum/paths (not disc*mean).
ORC_EXPORT double orc_local_vol(double spot, double strike, double t, double vol, double r,
                                double q, int32_t ot, int32_t paths, int32_t steps, uint64_t seed) {
    if (!common_ok(spot, strike, t, vol, r, q, ot) || !counts_ok(paths, steps)) return qnan();
    if (t <= kEps) return intrinsic(spot, strike, ot);
    const double disc = std::exp(-r * t);
    const double dt = t / static_cast<double>(steps);
    const double sdt = std::sqrt(dt);
    MtStream g(seed);
    double acc = 0.0;
    for (int32_t i = 0; i < paths; ++i) {
        double s = spot;
        double now = 0.0;
        for (int32_t j = 0; j < steps; ++j) {
            const double dw = sdt * g();
            const double sig = vol;  // constant sigma(s, now)
            s = std::max(1e-12, s + (r - q) * s * dt + sig * s * dw);
            now += dt;
        }
        (void)now;
        acc += intrinsic(s, strike, ot);
    }
    return disc * acc / static_cast<double>(paths);
}

// heston_monte_carlo.cpp:11-36 with mc_engine.h:113-136 and
// model_concepts.h:205-226 (log-Euler spot, full-truncation Euler variance).
ORC_EXPORT double orc_heston(double spot, double strike, double t, double r, double q, double v0,
                             double kappa, double theta, double sigma, double rho, int32_t ot,
                             int32_t paths, int32_t steps, uint64_t seed) {
    if (!type_ok(ot) || spot <= 0.0 || strike <= 0.0 || t < 0.0 || v0 < 0.0 || kappa < 0.0 ||
        theta < 0.0 || sigma < 0.0 || rho < -1.0 || rho > 1.0 || !counts_ok(paths, steps))
        return qnan();
    if (t <= kEps) return intrinsic(spot, strike, ot);
    const double disc = std::exp(-r * t);
    MtStream g(seed);
    const double rc = std::sqrt(std::max(1.0 - rho * rho, 0.0));
    const double dt = t / static_cast<double>(steps);
    const double sdt = std::sqrt(dt);
    double acc = 0.0;
    for (int32_t i = 0; i < paths; ++i) {
        double s = spot, v = v0;
        for (int32_t j = 0; j < steps; ++j) {
            const double z1 = g();
            const double z2 = g();
            const double vp = std::max(v, 0.0);
            const double sv = std::sqrt(vp);
            const double w1 = sdt * z1;
            const double w2 = sdt * (rho * z1 + rc * z2);
            const double s_new = s * std::exp((r - q - 0.5 * vp) * dt + sv * w1);
            const double v_new = std::max(0.0, v + kappa * (theta - vp) * dt + sigma * sv * w2);
            s = std::max(s_new, 1e-12);
            v = v_new;
        }
        acc += intrinsic(s, strike, ot);
    }
    const double mean = acc / static_cast<double>(paths);
    return disc * mean;
}

// longstaff_schwartz.cpp:15-42: Gauss-Jordan with partial pivoting on an
// augmented n x (n+1) system; a pivot below 1e-14 in magnitude gives beta = 0.
static void gauss_jordan(int n, double m[4][5], double beta[4]) {
    for (int col = 0; col < n; ++col) {
        int piv = col;
        for (int row = col + 1; row < n; ++row)
            if (std::fabs(m[row][col]) > std::fabs(m[piv][col])) piv = row;
        if (std::fabs(m[piv][col]) < 1e-14) {
            for (int i = 0; i < n; ++i) beta[i] = 0.0;
            return;
        }
        if (piv != col)
            for (int j = 0; j <= n; ++j) std::swap(m[piv][j], m[col][j]);
        const double d = m[col][col];
        for (int j = col; j <= n; ++j) m[col][j] /= d;
        for (int row = 0; row < n; ++row) {
            if (row == col) continue;
            const double f = m[row][col];
            for (int j = col; j <= n; ++j) m[row][j] -= f * m[col][j];
        }
    }
    for (int i = 0; i < n; ++i) beta[i] = m[i][n];
}

// Backward induction shared by both normal sources.  paths_mat is
// [path][date] with date 0 = spot.  basis 0: {1,S,S^2} (reference,
// longstaff_schwartz.cpp:84-125); basis 1: {1,x,x^2,x^3}, x = S/K.
static double lsmc_backward(std::vector<double>& mat, int32_t paths, int32_t steps, double strike,
                            int32_t ot, double disc, int basis, double* se_out) {
    const size_t w = size_t(steps) + 1;
    std::vector<double> cf(static_cast<size_t>(paths));
    for (int32_t i = 0; i < paths; ++i) cf[i] = intrinsic(mat[i * w + steps], strike, ot);
    std::vector<int32_t> itm;
    itm.reserve(size_t(paths));
    const int nb = basis == 0 ? 3 : 4;
    for (int32_t j = steps - 1; j >= 1; --j) {
        itm.clear();
        for (int32_t i = 0; i < paths; ++i) {
            cf[i] *= disc;
            if (intrinsic(mat[i * w + j], strike, ot) > 0.0) itm.push_back(i);
        }
        if (itm.size() < 3u) continue;
        double a[4][4] = {{0}}, b[4] = {0};
        for (int32_t idx : itm) {
            const double s = mat[idx * w + j];
            const double y = cf[idx];
            double x[4];
            if (basis == 0) {
                x[0] = 1.0; x[1] = s; x[2] = s * s; x[3] = 0.0;
            } else {
                const double u = s / strike;
                x[0] = 1.0; x[1] = u; x[2] = u * u; x[3] = u * u * u;
            }
            for (int ri = 0; ri < nb; ++ri) {
                b[ri] += x[ri] * y;
                for (int ci = 0; ci < nb; ++ci) a[ri][ci] += x[ri] * x[ci];
            }
        }
        double m[4][5];
        for (int i = 0; i < nb; ++i) {
            for (int k = 0; k < nb; ++k) m[i][k] = a[i][k];
            m[i][nb] = b[i];
        }
        double beta[4] = {0, 0, 0, 0};
        gauss_jordan(nb, m, beta);
        for (int32_t idx : itm) {
            const double s = mat[idx * w + j];
            const double ex = intrinsic(s, strike, ot);
            double cont;
            if (basis == 0) {
                cont = beta[0] + beta[1] * s + beta[2] * s * s;
            } else {
                const double u = s / strike;
                cont = beta[0] + beta[1] * u + beta[2] * u * u + beta[3] * u * u * u;
            }
            if (ex > cont) cf[idx] = ex;
        }
    }
    double acc = 0.0;
    Moments mo;
    for (int32_t i = 0; i < paths; ++i) {
        acc += cf[i];
        mo.add(cf[i]);
    }
    if (se_out) *se_out = se_of(mo, double(paths), disc);
    const double mean_cf = acc / static_cast<double>(paths);
    return mean_cf * disc;
}

// longstaff_schwartz.cpp:46-130 (basis 0) and the {1,x,x^2,x^3} variant.
ORC_EXPORT double orc_lsmc(double spot, double strike, double t, double vol, double r, double q,
                           int32_t ot, int32_t paths, int32_t steps, uint64_t seed, int32_t basis) {
    if (!common_ok(spot, strike, t, vol, r, q, ot) || !counts_ok(paths, steps)) return qnan();
    if (t <= kEps) return intrinsic(spot, strike, ot);
    const double dt = t / static_cast<double>(steps);
    const double disc = std::exp(-r * dt);
    const double mu_dt = (r - q - 0.5 * vol * vol) * dt;
    const double vs = vol * std::sqrt(dt);
    std::mt19937_64 eng(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    const size_t w = size_t(steps) + 1;
    std::vector<double> mat(size_t(paths) * w, spot);
    for (int32_t i = 0; i < paths; ++i)
        for (int32_t j = 1; j <= steps; ++j) {
            const double z = nd(eng);
            mat[i * w + j] = mat[i * w + j - 1] * std::exp(mu_dt + vs * z);
        }
    return lsmc_backward(mat, paths, steps, strike, ot, disc, basis, nullptr);
}

// ===========================================================================
// 2. The same estimators on the GPU's Philox stream, FP64 arithmetic.
//    Each returns the price and writes its standard error.
// ===========================================================================

ORC_EXPORT void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    philox10(ctr, key, out);
}

ORC_EXPORT void orc_box_muller(uint32_t a, uint32_t b, double* z0, double* z1) {
    box_muller(a, b, z0, z1);
}

ORC_EXPORT double orc_philox_gbm_normal(uint64_t seed, uint32_t path, uint32_t step) {
    return philox_gbm_normal(seed, path, step);
}

ORC_EXPORT double orc_philox_path_normal(uint64_t seed, uint32_t path, uint32_t draw) {
    return philox_path_normal(seed, path, draw);
}

ORC_EXPORT void orc_philox_heston_normals(uint64_t seed, uint32_t path, uint32_t step, double* z1,
                                          double* z2) {
    philox_heston_normals(seed, path, step, z1, z2);
}

ORC_EXPORT double orc_philox_terminal_normal(uint64_t seed, uint32_t k) {
    return philox_terminal_normal(seed, k);
}

// Antithetic terminal GBM (both standard MC and antithetic variates use it);
// vol_branch selects standard MC's vol <= eps closed form.
ORC_EXPORT double orc_px_terminal(double spot, double strike, double t, double vol, double r,
                                  double q, int32_t ot, int32_t paths, uint64_t seed,
                                  int32_t vol_branch, double* se) {
    *se = qnan();
    if (!common_ok(spot, strike, t, vol, r, q, ot) || paths <= 1) return qnan();
    *se = 0.0;
    if (t <= kEps) return intrinsic(spot, strike, ot);
    if (vol_branch && vol <= kEps) {
        const double fwd = spot * std::exp((r - q) * t);
        return std::exp(-r * t) * intrinsic(fwd, strike, ot);
    }
    const double disc = std::exp(-r * t);
    const int32_t pairs = paths / 2;
    double acc = 0.0;
    Moments mo;
    for (int32_t i = 0; i < pairs; ++i) {
        const double z = philox_terminal_normal(seed, uint32_t(i));
        const double a = intrinsic(gbm_terminal(spot, t, vol, r, q, z), strike, ot);
        const double b = intrinsic(gbm_terminal(spot, t, vol, r, q, -z), strike, ot);
        acc += a + b;
        mo.add(0.5 * (a + b));
    }
    if (paths & 1) {
        const double z = philox_terminal_normal(seed, uint32_t(pairs));
        acc += intrinsic(gbm_terminal(spot, t, vol, r, q, z), strike, ot);
    }
    *se = se_of(mo, double(pairs), disc);
    return disc * (acc / double(paths));
}

// Stepwise 1-D models: model 0 Euler, 1 Milstein, 3 constant local vol, on the
// eight-normals-per-block stream of philox_gbm_normal.
ORC_EXPORT double orc_px_stepwise(int32_t model, double spot, double strike, double t, double vol,
                                  double r, double q, int32_t ot, int32_t paths, int32_t steps,
                                  uint64_t seed, double* se) {
    *se = qnan();
    if (!common_ok(spot, strike, t, vol, r, q, ot) || !counts_ok(paths, steps)) return qnan();
    *se = 0.0;
    if (t <= kEps) return intrinsic(spot, strike, ot);
    const double disc = std::exp(-r * t);
    const double mu = r - q;
    const double dt = t / double(steps);
    const double sdt = std::sqrt(dt);
    Moments mo;
    for (int32_t i = 0; i < paths; ++i) {
        double s = spot;
        for (int32_t j = 0; j < steps; ++j) {
            const double dw = sdt * philox_gbm_normal(seed, uint32_t(i), uint32_t(j));
            double nxt = s + mu * s * dt + vol * s * dw;
            if (model == 1) nxt += 0.5 * vol * vol * s * (dw * dw - dt);
            s = std::max(1e-12, nxt);
        }
        mo.add(intrinsic(s, strike, ot));
    }
    *se = se_of(mo, double(paths), disc);
    return disc * (mo.s1 / double(paths));
}

ORC_EXPORT double orc_px_heston(double spot, double strike, double t, double r, double q,
                                double v0, double kappa, double theta, double sigma, double rho,
                                int32_t ot, int32_t paths, int32_t steps, uint64_t seed,
                                double* se) {
    *se = qnan();
    if (!type_ok(ot) || spot <= 0.0 || strike <= 0.0 || t < 0.0 || v0 < 0.0 || kappa < 0.0 ||
        theta < 0.0 || sigma < 0.0 || rho < -1.0 || rho > 1.0 || !counts_ok(paths, steps))
        return qnan();
    *se = 0.0;
    if (t <= kEps) return intrinsic(spot, strike, ot);
    const double disc = std::exp(-r * t);
    const double rc = std::sqrt(std::max(1.0 - rho * rho, 0.0));
    const double dt = t / double(steps);
    const double sdt = std::sqrt(dt);
    Moments mo;
    for (int32_t i = 0; i < paths; ++i) {
        double s = spot, v = v0;
        for (int32_t j = 0; j < steps; ++j) {
            double z1, z2;
            philox_heston_normals(seed, uint32_t(i), uint32_t(j), &z1, &z2);
            const double vp = std::max(v, 0.0);
            const double sv = std::sqrt(vp);
            const double w1 = sdt * z1;
            const double w2 = sdt * (rho * z1 + rc * z2);
            const double s_new = s * std::exp((r - q - 0.5 * vp) * dt + sv * w1);
            v = std::max(0.0, v + kappa * (theta - vp) * dt + sigma * sv * w2);
            s = std::max(s_new, 1e-12);
        }
        mo.add(intrinsic(s, strike, ot));
    }
    *se = se_of(mo, double(paths), disc);
    return disc * (mo.s1 / double(paths));
}

ORC_EXPORT double orc_px_lsmc(double spot, double strike, double t, double vol, double r, double q,
                              int32_t ot, int32_t paths, int32_t steps, uint64_t seed,
                              int32_t basis, double* se) {
    *se = qnan();
    if (!common_ok(spot, strike, t, vol, r, q, ot) || !counts_ok(paths, steps)) return qnan();
    *se = 0.0;
    if (t <= kEps) return intrinsic(spot, strike, ot);
    const double dt = t / double(steps);
    const double disc = std::exp(-r * dt);
    const double mu_dt = (r - q - 0.5 * vol * vol) * dt;
    const double vs = vol * std::sqrt(dt);
    const size_t w = size_t(steps) + 1;
    std::vector<double> mat(size_t(paths) * w, spot);
    for (int32_t i = 0; i < paths; ++i)
        for (int32_t j = 1; j <= steps; ++j) {
            const double z = philox_path_normal(seed, uint32_t(i), uint32_t(j - 1));
            mat[i * w + j] = mat[i * w + j - 1] * std::exp(mu_dt + vs * z);
        }
    return lsmc_backward(mat, paths, steps, strike, ot, disc, basis, se);
}

// ===========================================================================
// 3. Closed forms and the binomial tree.
// ===========================================================================

// Black-Scholes-Merton with N(x) = erfc(-x/sqrt 2)/2 (exact to rounding).
ORC_EXPORT double orc_bsm(double spot, double strike, double t, double vol, double r, double q,
                          int32_t ot) {
    if (!common_ok(spot, strike, t, vol, r, q, ot)) return qnan();
    if (t <= kEps) return intrinsic(spot, strike, ot);
    if (vol <= kEps) return std::exp(-r * t) * intrinsic(spot * std::exp((r - q) * t), strike, ot);
    const double st = vol * std::sqrt(t);
    const double d1 = (std::log(spot / strike) + (r - q + 0.5 * vol * vol) * t) / st;
    const double d2 = d1 - st;
    auto N = [](double x) { return 0.5 * std::erfc(-x * M_SQRT1_2); };
    const double df = std::exp(-r * t), qf = std::exp(-q * t);
    if (ot == 0) return spot * qf * N(d1) - strike * df * N(d2);
    return strike * df * N(-d2) - spot * qf * N(-d1);
}

// The reference's BSM (black_scholes_merton.cpp:7-32) uses the A&S 7.1.26
// approximation of N (math_util.h:27-46, max error 1.5e-7); kept so tests can
// reproduce the reference's own "MC near BSM" check exactly.
ORC_EXPORT double orc_bsm_as(double spot, double strike, double t, double vol, double r,
                             double q, int32_t ot) {
    if (!common_ok(spot, strike, t, vol, r, q, ot)) return qnan();
    if (t <= kEps) return intrinsic(spot, strike, ot);
    if (vol <= kEps) return std::exp(-r * t) * intrinsic(spot * std::exp((r - q) * t), strike, ot);
    auto N = [](double x) {
        // Abramowitz & Stegun 7.1.26 for erf, published coefficients.
        const double p = 0.3275911;
        const double c[5] = {0.254829592, -0.284496736, 1.421413741, -1.453152027, 1.061405429};
        const double ax = std::fabs(x) * M_SQRT1_2;
        const double k = 1.0 / (1.0 + p * ax);
        const double poly = k * (c[0] + k * (c[1] + k * (c[2] + k * (c[3] + k * c[4]))));
        const double erf_ax = 1.0 - poly * std::exp(-ax * ax);
        return 0.5 * (1.0 + (x < 0.0 ? -erf_ax : erf_ax));
    };
    const double st = vol * std::sqrt(t);
    const double d1 = (std::log(spot / strike) + (r - q + 0.5 * vol * vol) * t) / st;
    const double d2 = d1 - st;
    const double df = std::exp(-r * t), qf = std::exp(-q * t);
    if (ot == 0) return spot * qf * N(d1) - strike * df * N(d2);
    return strike * df * N(-d2) - spot * qf * N(-d1);
}

// Heston (1993) price via P1/P2 Fourier inversion with the Albrecher et al.
// (2007) "little trap" characteristic function, integrated over
// [1e-8, limit] with 64-point Gauss-Legendre panels (as heston.cpp:7-53 and
// closed_form common/internal_util.h:80-142 do).  Requires kappa > 0,
// sigma > 0, |rho| < 1.
ORC_EXPORT double orc_heston_cf(double spot, double strike, double t, double r, double q,
                                double v0, double kappa, double theta, double sigma, double rho,
                                int32_t ot, int32_t integration_steps, double limit) {
    // heston.cpp:10-24: type, finiteness of every input, ranges, then the
    // t <= eps branch, then the integration controls
    if (!type_ok(ot)) return qnan();
    for (double x : {spot, strike, t, r, q, v0, kappa, theta, sigma, rho})
        if (!finite_bits(x)) return qnan();
    if (spot <= 0.0 || strike <= 0.0 || t < 0.0 || v0 < 0.0 || kappa <= 0.0 || theta < 0.0 ||
        sigma <= 0.0 || rho <= -1.0 || rho >= 1.0)
        return qnan();
    if (t <= kEps) return intrinsic(spot, strike, ot);
    if (integration_steps < 16 || limit <= 1e-6) return qnan();
    using cd = std::complex<double>;
    const cd I(0.0, 1.0);
    const double s2 = sigma * sigma;
    const double x0 = std::log(spot) + (r - q) * t;
    // phi(u) = E[exp(i u ln S_T)]
    auto phi = [&](cd u) -> cd {
        const cd iu = I * u;
        const cd b = kappa - rho * sigma * iu;
        const cd d = std::sqrt(b * b + s2 * (iu + u * u));
        const cd g = (b - d) / (b + d);
        const cd e = std::exp(-d * t);
        const cd C = iu * x0 + (kappa * theta / s2) *
                                   ((b - d) * t - 2.0 * std::log((1.0 - g * e) / (1.0 - g)));
        const cd D = ((b - d) / s2) * ((1.0 - e) / (1.0 - g * e));
        return std::exp(C + D * v0);
    };
    const double lk = std::log(strike);
    const cd phi_mi = phi(cd(0.0, -1.0));
    const int panels = std::max(1, (std::max(integration_steps, 1) + 63) / 64);
    const double a = 1e-8, bnd = limit;
    const double width = (bnd - a) / panels;
    const GL64& g = gl64();
    double i1 = 0.0, i2 = 0.0;
    for (int p = 0; p < panels; ++p) {
        const double mid = a + (p + 0.5) * width, half = 0.5 * width;
        for (int k = 0; k < 32; ++k)
            for (int sgn = -1; sgn <= 1; sgn += 2) {
                const double u = mid + sgn * half * g.x[k];
                const cd ek = std::exp(-I * u * lk);
                // P2 integrand Re[e^{-iu lnK} phi(u) / (iu)]
                i2 += half * g.w[k] * std::real(ek * phi(cd(u, 0.0)) / (I * u));
                // P1 integrand Re[e^{-iu lnK} phi(u - i) / (iu phi(-i))]
                i1 += half * g.w[k] * std::real(ek * phi(cd(u, -1.0)) / (I * u * phi_mi));
            }
    }
    auto clamp01 = [](double x) { return std::min(1.0, std::max(0.0, x)); };
    const double P1 = clamp01(0.5 + i1 / M_PI);
    const double P2 = clamp01(0.5 + i2 / M_PI);
    const double call = spot * std::exp(-q * t) * P1 - strike * std::exp(-r * t) * P2;
    if (ot == 0) return call;
    return call - spot * std::exp(-q * t) + strike * std::exp(-r * t);
}

// Cox-Ross-Rubinstein binomial tree (textbook): u = e^{vol sqrt dt},
// d = 1/u, p = (e^{(r-q)dt} - d)/(u - d); American exercise checked at every
// node when american != 0.
ORC_EXPORT double orc_crr(double spot, double strike, double t, double vol, double r, double q,
                          int32_t ot, int32_t steps, int32_t american) {
    if (!common_ok(spot, strike, t, vol, r, q, ot) || steps < 1) return qnan();
    if (t <= kEps) return intrinsic(spot, strike, ot);
    const double dt = t / steps;
    const double u = std::exp(vol * std::sqrt(dt)), d = 1.0 / u;
    const double p = (std::exp((r - q) * dt) - d) / (u - d);
    const double disc = std::exp(-r * dt);
    std::vector<double> v(size_t(steps) + 1);
    for (int i = 0; i <= steps; ++i)
        v[i] = intrinsic(spot * std::pow(u, 2.0 * i - steps), strike, ot);
    for (int n = steps - 1; n >= 0; --n)
        for (int i = 0; i <= n; ++i) {
            double c = disc * (p * v[i + 1] + (1.0 - p) * v[i]);
            if (american) c = std::max(c, intrinsic(spot * std::pow(u, 2.0 * i - n), strike, ot));
            v[i] = c;
        }
    return v[0];
}

// ===========================================================================
// Batch wrappers (OpenMP over options, as the reference's qk_api.cpp batches
// do) used as the CPU baseline when no build of the reference is available.
// ===========================================================================
ORC_EXPORT void orc_heston_batch(const double* spot, const double* strike, const double* t,
                                 const double* r, const double* q, const double* v0,
                                 const double* kappa, const double* theta, const double* sigma,
                                 const double* rho, const int32_t* ot, const int32_t* paths,
                                 const int32_t* steps, const uint64_t* seed, int32_t n,
                                 double* out) {
#pragma omp parallel for schedule(dynamic)
    for (int32_t i = 0; i < n; ++i)
        out[i] = orc_heston(spot[i], strike[i], t[i], r[i], q[i], v0[i], kappa[i], theta[i],
                            sigma[i], rho[i], ot[i], paths[i], steps[i], seed[i]);
}

ORC_EXPORT void orc_standard_mc_batch(const double* spot, const double* strike, const double* t,
                                      const double* vol, const double* r, const double* q,
                                      const int32_t* ot, const int32_t* paths,
                                      const uint64_t* seed, int32_t n, double* out) {
#pragma omp parallel for schedule(dynamic)
    for (int32_t i = 0; i < n; ++i)
        out[i] = orc_standard_mc(spot[i], strike[i], t[i], vol[i], r[i], q[i], ot[i], paths[i],
                                 seed[i]);
}

// ===========================================================================
// 4. Callers of estimate_terminal (mc_engine.h:62-75) and the Greeks on the
//    MC engine: importance sampling, pathwise and likelihood-ratio delta
//    (GBM), Heston likelihood-ratio delta.  Each exists twice: on the
//    reference's mt19937_64 stream (bit-identical to the reference when both
//    are built the same way) and on the GPU's Philox stream with the
//    per-estimator-unit standard error.
// ===========================================================================
namespace {

// adjoint_greeks/common/internal_util.h:22-31
double deterministic_delta(double spot, double strike, double t, double r, double q, int32_t ot) {
    const double forward = spot * std::exp((r - q) * t);
    const double qf = std::exp(-q * t);
    if (std::fabs(forward - strike) < kEps * strike) return ot == 0 ? 0.5 * qf : -0.5 * qf;
    if (ot == 0) return forward > strike ? qf : 0.0;
    return forward < strike ? -qf : 0.0;
}

// adjoint_greeks/common/internal_util.h:47-50
double clamp_delta(double delta, double t, double q) {
    const double qf = std::exp(-q * t);
    return std::min(qf, std::max(-qf, delta));
}

// importance_sampling.cpp:26-31: the accumulated value of one path.
double is_unit(double spot, double strike, double t, double vol, double r, double q, int32_t ot,
               double shift, double z) {
    const double st = gbm_terminal(spot, t, vol, r, q, z + shift);
    const double y = z + shift;
    const double w = std::exp(-shift * y + 0.5 * shift * shift);
    return intrinsic(st, strike, ot) * w;
}

// pathwise_derivative.cpp:29-35 with model_concepts.h:108-111.
double pw_unit(double spot, double strike, double t, double vol, double r, double q, int32_t ot,
               double z) {
    const double st = gbm_terminal(spot, t, vol, r, q, z);
    if (ot == 0) return st > strike ? st / spot : 0.0;
    return st < strike ? -(st / spot) : 0.0;
}

// likelihood_ratio.cpp:30-37 with model_concepts.h:121-128.
double lr_unit(double spot, double strike, double t, double vol, double r, double q, int32_t ot,
               double wclip, double z) {
    const double st = gbm_terminal(spot, t, vol, r, q, z);
    const double pay = ot == 0 ? std::max(st - strike, 0.0) : std::max(strike - st, 0.0);
    const double sqrt_t = std::sqrt(t);
    double score = z / (vol * sqrt_t * spot);
    score = std::max(-wclip, std::min(wclip, score));
    return pay * score;
}

// Heston path of heston_lr_delta.cpp:74-82 (model_concepts.h:205-226); the
// draws come from `draw(j, &z1, &z2)`; sign = -1 gives the antithetic path.
template <class Draw>
double heston_lr_path(double spot, double r, double q, double v0, double kappa, double theta,
                      double sigma, double rho, int32_t steps, double dt, double sdt, double sign,
                      Draw&& draw, double* z1_first) {
    const double rc = std::sqrt(std::max(1.0 - rho * rho, 0.0));
    double s = spot, v = v0;
    for (int32_t j = 0; j < steps; ++j) {
        double z1, z2;
        draw(j, &z1, &z2);
        z1 *= sign;
        z2 *= sign;
        if (j == 0) *z1_first = z1;
        const double vp = std::max(v, 0.0);
        const double sv = std::sqrt(vp);
        const double w1 = sdt * z1;
        const double w2 = sdt * (rho * z1 + rc * z2);
        const double s_new = s * std::exp((r - q - 0.5 * vp) * dt + sv * w1);
        v = std::max(0.0, v + kappa * (theta - vp) * dt + sigma * sv * w2);
        s = std::max(s_new, 1e-12);
    }
    return s;
}

bool heston_lr_ok(double spot, double strike, double t, double v0, double kappa, double theta,
                  double sigma, double rho, int32_t ot, int32_t paths, int32_t steps,
                  double wclip) {
    // heston_lr_delta.cpp:47-54
    return !(!type_ok(ot) || spot <= 0.0 || strike <= 0.0 || t < 0.0 || v0 < 0.0 ||
             kappa < 0.0 || theta < 0.0 || sigma < 0.0 || rho < -1.0 || rho > 1.0 ||
             !counts_ok(paths, steps) || wclip <= 0.0);
}

double heston_lr_intrinsic_delta(double spot, double strike, int32_t ot) {
    // heston_lr_delta.cpp:57-61
    if (ot == 0) return spot > strike ? 1.0 : 0.0;
    return spot < strike ? -1.0 : 0.0;
}

}  // namespace

// importance_sampling.cpp:11-36 (estimate_terminal, one draw per path).
ORC_EXPORT double orc_importance_sampling(double spot, double strike, double t, double vol,
                                          double r, double q, int32_t ot, int32_t paths,
                                          double shift, uint64_t seed) {
    if (!common_ok(spot, strike, t, vol, r, q, ot) || paths <= 1 || !finite_bits(shift))
        return qnan();
    if (t <= kEps) return intrinsic(spot, strike, ot);
    const double disc = std::exp(-r * t);
    MtStream g(seed);
    double sum = 0.0;
    for (int32_t i = 0; i < paths; ++i) sum += is_unit(spot, strike, t, vol, r, q, ot, shift, g());
    return disc * (sum / double(paths));
}

// pathwise_derivative.cpp:11-40.
ORC_EXPORT double orc_pathwise_delta(double spot, double strike, double t, double vol, double r,
                                     double q, int32_t ot, int32_t paths, uint64_t seed) {
    if (!common_ok(spot, strike, t, vol, r, q, ot) || paths < 256) return qnan();
    if (t <= kEps || vol <= kEps) return deterministic_delta(spot, strike, t, r, q, ot);
    const double disc = std::exp(-r * t);
    MtStream g(seed);
    double sum = 0.0;
    for (int32_t i = 0; i < paths; ++i) sum += pw_unit(spot, strike, t, vol, r, q, ot, g());
    return clamp_delta(disc * (sum / double(paths)), t, q);
}

// likelihood_ratio.cpp:12-42 (estimate_terminal_antithetic, mc_engine.h:32-53).
ORC_EXPORT double orc_lr_delta(double spot, double strike, double t, double vol, double r,
                               double q, int32_t ot, int32_t paths, uint64_t seed, double wclip) {
    if (!common_ok(spot, strike, t, vol, r, q, ot) || paths < 256 || !finite_bits(wclip) ||
        wclip <= 0.0)
        return qnan();
    if (t <= kEps || vol <= kEps) return deterministic_delta(spot, strike, t, r, q, ot);
    const double disc = std::exp(-r * t);
    MtStream g(seed);
    double sum = 0.0;
    const int32_t pairs = paths / 2;
    for (int32_t i = 0; i < pairs; ++i) {
        const double z = g();
        sum += lr_unit(spot, strike, t, vol, r, q, ot, wclip, z);
        sum += lr_unit(spot, strike, t, vol, r, q, ot, wclip, -z);
    }
    if (paths & 1) sum += lr_unit(spot, strike, t, vol, r, q, ot, wclip, g());
    return clamp_delta(disc * (sum / double(paths)), t, q);
}

// heston_lr_delta.cpp:41-132: antithetic pairs replay the pair's draws
// negated; an odd path takes fresh draws; result disc * sum / paths.
ORC_EXPORT double orc_heston_lr_delta(double spot, double strike, double t, double r, double q,
                                      double v0, double kappa, double theta, double sigma,
                                      double rho, int32_t ot, int32_t paths, int32_t steps,
                                      uint64_t seed, double wclip) {
    if (!heston_lr_ok(spot, strike, t, v0, kappa, theta, sigma, rho, ot, paths, steps, wclip))
        return qnan();
    if (t <= kEps) return heston_lr_intrinsic_delta(spot, strike, ot);
    const double disc = std::exp(-r * t);
    const double dt = t / double(steps);
    const double sdt = std::sqrt(dt);
    const double sqrt_v0 = std::sqrt(std::max(v0, 0.0));
    MtStream g(seed);
    std::vector<double> zs(size_t(2) * steps);
    auto from_buf = [&](int32_t j, double* z1, double* z2) {
        *z1 = zs[2 * j];
        *z2 = zs[2 * j + 1];
    };
    auto score_of = [&](double z1) {
        const double sc = sqrt_v0 > 1e-15 ? z1 / (sqrt_v0 * sdt * spot) : 0.0;
        return std::max(-wclip, std::min(wclip, sc));
    };
    double sum = 0.0, z1f = 0.0, z1a = 0.0;
    const int32_t pairs = paths / 2;
    for (int32_t i = 0; i < pairs; ++i) {
        for (auto& z : zs) z = g();
        const double sf = heston_lr_path(spot, r, q, v0, kappa, theta, sigma, rho, steps, dt, sdt,
                                         1.0, from_buf, &z1f);
        const double sa = heston_lr_path(spot, r, q, v0, kappa, theta, sigma, rho, steps, dt, sdt,
                                         -1.0, from_buf, &z1a);
        const double score = score_of(z1f);
        sum += intrinsic(sf, strike, ot) * score + intrinsic(sa, strike, ot) * (-score);
    }
    if (paths & 1) {
        for (auto& z : zs) z = g();
        const double s = heston_lr_path(spot, r, q, v0, kappa, theta, sigma, rho, steps, dt, sdt,
                                        1.0, from_buf, &z1f);
        sum += intrinsic(s, strike, ot) * score_of(z1f);
    }
    return disc * sum / double(paths);
}

// ---- the same four on the Philox stream (FP64), with standard errors ----
// Units: importance sampling and pathwise delta one path each (terminal
// draw k for path k); LR delta one antithetic pair (draw k for pair k, the
// odd path draw `pairs`); Heston LR delta one pair (Philox path index k,
// draws 2j and 2j+1 at step j; the odd path is path index `pairs`).
ORC_EXPORT double orc_px_importance_sampling(double spot, double strike, double t, double vol,
                                             double r, double q, int32_t ot, int32_t paths,
                                             double shift, uint64_t seed, double* se) {
    *se = qnan();
    if (!common_ok(spot, strike, t, vol, r, q, ot) || paths <= 1 || !finite_bits(shift))
        return qnan();
    *se = 0.0;
    if (t <= kEps) return intrinsic(spot, strike, ot);
    const double disc = std::exp(-r * t);
    Moments mo;
    for (int32_t i = 0; i < paths; ++i)
        mo.add(is_unit(spot, strike, t, vol, r, q, ot, shift,
                       philox_terminal_normal(seed, uint32_t(i))));
    *se = se_of(mo, double(paths), disc);
    return disc * (mo.s1 / double(paths));
}

ORC_EXPORT double orc_px_pathwise_delta(double spot, double strike, double t, double vol, double r,
                                        double q, int32_t ot, int32_t paths, uint64_t seed,
                                        double* se) {
    *se = qnan();
    if (!common_ok(spot, strike, t, vol, r, q, ot) || paths < 256) return qnan();
    *se = 0.0;
    if (t <= kEps || vol <= kEps) return deterministic_delta(spot, strike, t, r, q, ot);
    const double disc = std::exp(-r * t);
    Moments mo;
    for (int32_t i = 0; i < paths; ++i)
        mo.add(pw_unit(spot, strike, t, vol, r, q, ot, philox_terminal_normal(seed, uint32_t(i))));
    *se = se_of(mo, double(paths), disc);
    return clamp_delta(disc * (mo.s1 / double(paths)), t, q);
}

ORC_EXPORT double orc_px_lr_delta(double spot, double strike, double t, double vol, double r,
                                  double q, int32_t ot, int32_t paths, uint64_t seed, double wclip,
                                  double* se) {
    *se = qnan();
    if (!common_ok(spot, strike, t, vol, r, q, ot) || paths < 256 || !finite_bits(wclip) ||
        wclip <= 0.0)
        return qnan();
    *se = 0.0;
    if (t <= kEps || vol <= kEps) return deterministic_delta(spot, strike, t, r, q, ot);
    const double disc = std::exp(-r * t);
    const int32_t pairs = paths / 2;
    double sum = 0.0;
    Moments mo;
    for (int32_t i = 0; i < pairs; ++i) {
        const double z = philox_terminal_normal(seed, uint32_t(i));
        const double a = lr_unit(spot, strike, t, vol, r, q, ot, wclip, z);
        const double b = lr_unit(spot, strike, t, vol, r, q, ot, wclip, -z);
        sum += a + b;
        mo.add(0.5 * (a + b));
    }
    if (paths & 1)
        sum += lr_unit(spot, strike, t, vol, r, q, ot, wclip,
                       philox_terminal_normal(seed, uint32_t(pairs)));
    *se = se_of(mo, double(pairs), disc);  // pair average as the unit, as for orc_px_terminal
    return clamp_delta(disc * (sum / double(paths)), t, q);
}

ORC_EXPORT double orc_px_heston_lr_delta(double spot, double strike, double t, double r, double q,
                                         double v0, double kappa, double theta, double sigma,
                                         double rho, int32_t ot, int32_t paths, int32_t steps,
                                         uint64_t seed, double wclip, double* se) {
    *se = qnan();
    if (!heston_lr_ok(spot, strike, t, v0, kappa, theta, sigma, rho, ot, paths, steps, wclip))
        return qnan();
    *se = 0.0;
    if (t <= kEps) return heston_lr_intrinsic_delta(spot, strike, ot);
    const double disc = std::exp(-r * t);
    const double dt = t / double(steps);
    const double sdt = std::sqrt(dt);
    const double sqrt_v0 = std::sqrt(std::max(v0, 0.0));
    auto score_of = [&](double z1) {
        const double sc = sqrt_v0 > 1e-15 ? z1 / (sqrt_v0 * sdt * spot) : 0.0;
        return std::max(-wclip, std::min(wclip, sc));
    };
    const int32_t pairs = paths / 2;
    double sum = 0.0, z1f = 0.0, z1a = 0.0;
    Moments mo;
    for (int32_t i = 0; i <= pairs; ++i) {
        if (i == pairs && !(paths & 1)) break;
        auto draw = [&](int32_t j, double* z1, double* z2) {
            philox_heston_normals(seed, uint32_t(i), uint32_t(j), z1, z2);
        };
        const double sf = heston_lr_path(spot, r, q, v0, kappa, theta, sigma, rho, steps, dt, sdt,
                                         1.0, draw, &z1f);
        const double score = score_of(z1f);
        if (i == pairs) {
            sum += intrinsic(sf, strike, ot) * score;
            break;
        }
        const double sa = heston_lr_path(spot, r, q, v0, kappa, theta, sigma, rho, steps, dt, sdt,
                                         -1.0, draw, &z1a);
        const double u = intrinsic(sf, strike, ot) * score - intrinsic(sa, strike, ot) * score;
        sum += u;
        mo.add(u);
    }
    // unit = pair sum; the estimator is sum / paths = (pairs / paths) mean(unit)
    *se = se_of(mo, double(pairs), disc) * pairs / double(paths);
    return disc * sum / double(paths);
}

// ===========================================================================
// 5. Remaining variance-reduction estimators of the MC family: control
//    variates, stratified sampling, Sobol / Halton QMC and multilevel MC
//    (monte_carlo_methods/{control_variates,stratified_sampling,
//    quasi_monte_carlo,multilevel_monte_carlo}), on the reference's stream and
//    on the GPU's Philox stream.
// ===========================================================================
namespace {

// Acklam's rational approximation of the inverse normal CDF (P. J. Acklam,
// 2003; relative error < 1.15e-9), as monte_carlo_methods/common/
// internal_util.h:43-83 evaluates it (same coefficients, branches and
// Horner order).
double acklam(double p) {
    if (p <= 0.0 || p >= 1.0) return qnan();
    const double a[6] = {-3.969683028665376e+01, 2.209460984245205e+02, -2.759285104469687e+02,
                         1.383577518672690e+02, -3.066479806614716e+01, 2.506628277459239e+00};
    const double b[5] = {-5.447609879822406e+01, 1.615858368580409e+02, -1.556989798598866e+02,
                         6.680131188771972e+01, -1.328068155288572e+01};
    const double c[6] = {-7.784894002430293e-03, -3.223964580411365e-01, -2.400758277161838e+00,
                         -2.549732539343734e+00, 4.374664141464968e+00, 2.938163982698783e+00};
    const double d[4] = {7.784695709041462e-03, 3.224671290700398e-01, 2.445134137142996e+00,
                         3.754408661907416e+00};
    const double lo = 0.02425, hi = 1.0 - lo;
    auto tail = [&](double qq) {
        return (((((c[0] * qq + c[1]) * qq + c[2]) * qq + c[3]) * qq + c[4]) * qq + c[5]) /
               ((((d[0] * qq + d[1]) * qq + d[2]) * qq + d[3]) * qq + 1.0);
    };
    if (p < lo) return tail(std::sqrt(-2.0 * std::log(p)));
    if (p > hi) return -tail(std::sqrt(-2.0 * std::log(1.0 - p)));
    const double qq = p - 0.5, rr = qq * qq;
    return (((((a[0] * rr + a[1]) * rr + a[2]) * rr + a[3]) * rr + a[4]) * rr + a[5]) * qq /
           (((((b[0] * rr + b[1]) * rr + b[2]) * rr + b[3]) * rr + b[4]) * rr + 1.0);
}

uint32_t bit_reverse(uint32_t x) {
    uint32_t y = 0;
    for (int i = 0; i < 32; ++i) y |= ((x >> i) & 1u) << (31 - i);
    return y;
}

// internal_util.h:96-110: direction number of bit j is 1 << (31 - j), so the
// Gray-code XOR sum is the bit reverse of gray(i).
double sobol_point(uint32_t i) {
    if (i == 0) return 0.0;
    return (double(bit_reverse(i ^ (i >> 1))) + 0.5) / 4294967296.0;
}

// internal_util.h:112-122 with base 2 (van der Corput): exact binary fractions.
double halton2_point(uint32_t i) { return double(bit_reverse(i)) / 4294967296.0; }

double clamp_u(double u) { return std::min(std::max(u, 1e-12), 1.0 - 1e-12); }

// mcm internal_util.h:32-36: (spot, t, r, q, vol, z) argument order.
double gbm_terminal_mcm(double spot, double t, double r, double q, double vol, double z) {
    const double drift = (r - q - 0.5 * vol * vol) * t;
    const double diffusion = vol * std::sqrt(std::max(t, 0.0)) * z;
    return spot * std::exp(drift + diffusion);
}

// Uniform k of the Philox uniform stream: word k%4 of block (k/4, 0, 0, 2),
// (w + 0.5) 2^-32.
double philox_uniform(uint64_t seed, uint32_t k) {
    const uint32_t in[4] = {k >> 2, 0u, 0u, 2u};
    const uint32_t key[2] = {uint32_t(seed), uint32_t(seed >> 32)};
    uint32_t w[4];
    philox10(in, key, w);
    return (double(w[k & 3u]) + 0.5) * 0x1p-32;
}

// MLMC draw d of path p at level l: component d%4 of block (p, d/4, 0, l).
double philox_level_normal(uint64_t seed, uint32_t p, uint32_t d, uint32_t level) {
    double z[4];
    philox_normals4(p, d >> 2, level, seed, z);
    return z[d & 3u];
}

// The control-variate combination of control_variates.cpp:42-53 from
// population moments; returns the estimate and the residual standard error.
double cv_combine(double sum_y, double sum_x, double sum_xx, double sum_xy, double sum_yy,
                  int32_t paths, double expected_x, double* se) {
    const double n = double(paths);
    const double mean_y = sum_y / n;
    const double mean_x = sum_x / n;
    const double var_x = (sum_xx / n) - mean_x * mean_x;
    const double var_y = (sum_yy / n) - mean_y * mean_y;
    if (var_x <= kEps) {
        if (se) *se = std::sqrt(std::max(var_y, 0.0) / (n - 1.0));
        return mean_y;
    }
    const double cov_xy = (sum_xy / n) - mean_x * mean_y;
    const double beta = cov_xy / var_x;
    if (se) *se = std::sqrt(std::max(var_y - cov_xy * beta, 0.0) / (n - 1.0));
    return mean_y - beta * (mean_x - expected_x);
}

}  // namespace

// control_variates.cpp:11-54 (one draw per path; control x = e^{-rT} S_T with
// known mean S e^{-qT}).
ORC_EXPORT double orc_control_variates(double spot, double strike, double t, double vol, double r,
                                       double q, int32_t ot, int32_t paths, uint64_t seed) {
    if (!common_ok(spot, strike, t, vol, r, q, ot) || paths <= 2) return qnan();
    if (t <= kEps) return intrinsic(spot, strike, ot);
    MtStream g(seed);
    const double disc = std::exp(-r * t);
    const double expected_x = spot * std::exp(-q * t);
    double sy = 0.0, sx = 0.0, sxx = 0.0, sxy = 0.0;
    for (int32_t i = 0; i < paths; ++i) {
        const double st = gbm_terminal_mcm(spot, t, r, q, vol, g());
        const double x = disc * st;
        const double y = disc * intrinsic(st, strike, ot);
        sy += y;
        sx += x;
        sxx += x * x;
        sxy += x * y;
    }
    return cv_combine(sy, sx, sxx, sxy, 0.0, paths, expected_x, nullptr);
}

// stratified_sampling.cpp:11-35: u_i = (i + U_i)/n, U_i from
// std::uniform_real_distribution<double>(0, 1) on the option's mt19937_64.
ORC_EXPORT double orc_stratified(double spot, double strike, double t, double vol, double r,
                                 double q, int32_t ot, int32_t paths, uint64_t seed) {
    if (!common_ok(spot, strike, t, vol, r, q, ot) || paths <= 1) return qnan();
    if (t <= kEps) return intrinsic(spot, strike, ot);
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> unif(0.0, 1.0);
    const double disc = std::exp(-r * t);
    double sum = 0.0;
    for (int32_t i = 0; i < paths; ++i) {
        const double u = clamp_u((double(i) + unif(rng)) / double(paths));
        sum += intrinsic(gbm_terminal_mcm(spot, t, r, q, vol, acklam(u)), strike, ot);
    }
    return disc * (sum / double(paths));
}

// The uniforms of the reference's stratified stream (for the product's
// validation-stream check).
ORC_EXPORT void orc_mt_uniforms(uint64_t seed, int64_t count, double* out) {
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> unif(0.0, 1.0);
    for (int64_t i = 0; i < count; ++i) out[i] = unif(rng);
}

ORC_EXPORT double orc_acklam(double p) { return acklam(p); }
ORC_EXPORT double orc_sobol_point(uint32_t i) { return sobol_point(i); }
ORC_EXPORT double orc_halton2_point(uint32_t i) { return halton2_point(i); }

// quasi_monte_carlo.cpp:10-56: points i = 1..paths, no seed.  kind 0 Sobol,
// 1 Halton (base 2).
ORC_EXPORT double orc_qmc(int32_t kind, double spot, double strike, double t, double vol, double r,
                          double q, int32_t ot, int32_t paths) {
    if (!common_ok(spot, strike, t, vol, r, q, ot) || paths <= 1) return qnan();
    if (t <= kEps) return intrinsic(spot, strike, ot);
    const double disc = std::exp(-r * t);
    double sum = 0.0;
    for (int32_t i = 1; i <= paths; ++i) {
        const double u0 = kind == 0 ? sobol_point(uint32_t(i)) : halton2_point(uint32_t(i));
        const double z = acklam(clamp_u(u0));
        sum += intrinsic(gbm_terminal_mcm(spot, t, r, q, vol, z), strike, ot);
    }
    return disc * (sum / double(paths));
}

// multilevel_monte_carlo.cpp:10-72.  Level l: max(16, base_paths >> l) paths
// of base_steps << l Euler steps; levels >= 1 add fine - coarse with the
// coarse path driven by (z1 + z2)/sqrt 2 over two fine steps.  `draw(l, p, d)`
// returns draw d of path p at level l.
template <class Draw>
static double mlmc_impl(double spot, double strike, double t, double vol, double r, double q,
                        int32_t ot, int32_t base_paths, int32_t levels, int32_t base_steps,
                        Draw&& draw, double* se) {
    double estimate = 0.0, var_sum = 0.0;
    for (int32_t level = 0; level < levels; ++level) {
        const int32_t n_level = std::max(16, base_paths >> level);
        const int32_t fine_steps = base_steps << level;
        const double dt_fine = t / double(fine_steps);
        const double sqrt_dt_fine = std::sqrt(dt_fine);
        const double disc = std::exp(-r * t);
        double level_sum = 0.0;
        Moments mo;
        for (int32_t i = 0; i < n_level; ++i) {
            double unit;
            if (level == 0) {
                double s = spot;
                for (int32_t j = 0; j < fine_steps; ++j) {
                    const double z = draw(level, i, j);
                    s += (r - q) * s * dt_fine + vol * s * sqrt_dt_fine * z;
                    s = std::max(1e-12, s);
                }
                unit = intrinsic(s, strike, ot);
            } else {
                double s_fine = spot, s_coarse = spot;
                for (int32_t cs = 0; cs < fine_steps / 2; ++cs) {
                    const double z1 = draw(level, i, 2 * cs);
                    const double z2 = draw(level, i, 2 * cs + 1);
                    s_fine += (r - q) * s_fine * dt_fine + vol * s_fine * sqrt_dt_fine * z1;
                    s_fine = std::max(1e-12, s_fine);
                    s_fine += (r - q) * s_fine * dt_fine + vol * s_fine * sqrt_dt_fine * z2;
                    s_fine = std::max(1e-12, s_fine);
                    const double zc = (z1 + z2) * M_SQRT1_2;
                    const double dt_coarse = 2.0 * dt_fine;
                    s_coarse += (r - q) * s_coarse * dt_coarse +
                                vol * s_coarse * std::sqrt(dt_coarse) * zc;
                    s_coarse = std::max(1e-12, s_coarse);
                }
                unit = intrinsic(s_fine, strike, ot) - intrinsic(s_coarse, strike, ot);
            }
            level_sum += unit;
            mo.add(unit);
        }
        estimate += disc * (level_sum / double(n_level));
        const double sl = se_of(mo, double(n_level), disc);
        var_sum += sl * sl;
    }
    if (se) *se = std::sqrt(var_sum);
    return estimate;
}

static bool mlmc_ok(double spot, double strike, double t, double vol, double r, double q,
                    int32_t ot, int32_t base_paths, int32_t levels, int32_t base_steps) {
    return common_ok(spot, strike, t, vol, r, q, ot) && base_paths > 1 && levels > 0 &&
           base_steps > 0;
}

ORC_EXPORT double orc_mlmc(double spot, double strike, double t, double vol, double r, double q,
                           int32_t ot, int32_t base_paths, int32_t levels, int32_t base_steps,
                           uint64_t seed) {
    if (!mlmc_ok(spot, strike, t, vol, r, q, ot, base_paths, levels, base_steps)) return qnan();
    if (t <= kEps) return intrinsic(spot, strike, ot);
    MtStream g(seed);  // one stream, consumed level by level, path by path
    return mlmc_impl(spot, strike, t, vol, r, q, ot, base_paths, levels, base_steps,
                     [&](int32_t, int32_t, int32_t) { return g(); }, nullptr);
}

// ---- Philox-stream versions with standard errors ----
ORC_EXPORT double orc_px_control_variates(double spot, double strike, double t, double vol,
                                          double r, double q, int32_t ot, int32_t paths,
                                          uint64_t seed, double* se) {
    *se = qnan();
    if (!common_ok(spot, strike, t, vol, r, q, ot) || paths <= 2) return qnan();
    *se = 0.0;
    if (t <= kEps) return intrinsic(spot, strike, ot);
    const double disc = std::exp(-r * t);
    const double expected_x = spot * std::exp(-q * t);
    double sy = 0.0, sx = 0.0, sxx = 0.0, sxy = 0.0, syy = 0.0;
    for (int32_t i = 0; i < paths; ++i) {
        const double st =
            gbm_terminal_mcm(spot, t, r, q, vol, philox_terminal_normal(seed, uint32_t(i)));
        const double x = disc * st;
        const double y = disc * intrinsic(st, strike, ot);
        sy += y;
        sx += x;
        sxx += x * x;
        sxy += x * y;
        syy += y * y;
    }
    return cv_combine(sy, sx, sxx, sxy, syy, paths, expected_x, se);
}

ORC_EXPORT double orc_px_stratified(double spot, double strike, double t, double vol, double r,
                                    double q, int32_t ot, int32_t paths, uint64_t seed,
                                    double* se) {
    *se = qnan();
    if (!common_ok(spot, strike, t, vol, r, q, ot) || paths <= 1) return qnan();
    *se = 0.0;
    if (t <= kEps) return intrinsic(spot, strike, ot);
    const double disc = std::exp(-r * t);
    Moments mo;
    for (int32_t i = 0; i < paths; ++i) {
        const double u =
            clamp_u((double(i) + philox_uniform(seed, uint32_t(i))) / double(paths));
        mo.add(intrinsic(gbm_terminal_mcm(spot, t, r, q, vol, acklam(u)), strike, ot));
    }
    *se = se_of(mo, double(paths), disc);
    return disc * (mo.s1 / double(paths));
}

ORC_EXPORT double orc_px_mlmc(double spot, double strike, double t, double vol, double r,
                              double q, int32_t ot, int32_t base_paths, int32_t levels,
                              int32_t base_steps, uint64_t seed, double* se) {
    *se = qnan();
    if (!mlmc_ok(spot, strike, t, vol, r, q, ot, base_paths, levels, base_steps)) return qnan();
    *se = 0.0;
    if (t <= kEps) return intrinsic(spot, strike, ot);
    return mlmc_impl(spot, strike, t, vol, r, q, ot, base_paths, levels, base_steps,
                     [&](int32_t l, int32_t p, int32_t d) {
                         return philox_level_normal(seed, uint32_t(p), uint32_t(d), uint32_t(l));
                     },
                     se);
}
