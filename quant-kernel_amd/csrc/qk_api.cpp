// qk_api.cpp - flat C ABI of the MI355X Monte Carlo batch-pricing library.
//
// Error conventions follow the reference's qk_api.cpp:19-117:
//   - a thread-local 256-byte message, strncpy-truncated, sticky on success,
//     cleared only by qk_clear_last_error;
//   - batch calls check every pointer in declaration order first
//     ("null pointer: <param>" -> QK_ERR_NULL_PTR), then n <= 0
//     ("bad batch size: <n>" -> QK_ERR_BAD_SIZE);
//   - invalid elements give NaN with QK_OK.
// Scalar entry points are batches of one, so scalar == batch bitwise.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>
#include <exception>
#include <new>

#include "qk_cf.hpp"
#include "qk_engine.hpp"
#include "qk_lsmc.hpp"
#include "qk_mlp.hpp"
#include "qk_surrogate.hpp"
#include "quantkernel/qk_api.h"

namespace {

thread_local char tl_msg[256] = "";

void set_msg(const char* m) {
    std::strncpy(tl_msg, m, sizeof(tl_msg) - 1);
    tl_msg[sizeof(tl_msg) - 1] = '\0';
}

void set_msg(const std::string& m) { set_msg(m.c_str()); }

// Runs one engine call; an exception (a failed host allocation, above all)
// becomes QK_ERR_RUNTIME_INIT with a message instead of crossing the C ABI.
template <class F>
int32_t guard(std::string& err, F&& f) {
    try {
        return f();
    } catch (const std::bad_alloc&) {
        err = "host memory allocation failed";
    } catch (const std::exception& e) {
        err = std::string("internal error: ") + e.what();
    } catch (...) {
        err = "internal error";
    }
    return QK_ERR_RUNTIME_INIT;
}

bool check_ptrs(std::initializer_list<std::pair<const void*, const char*>> ptrs) {
    for (const auto& p : ptrs) {
        if (!p.first) {
            char buf[256];
            std::snprintf(buf, sizeof buf, "null pointer: %s", p.second);
            set_msg(buf);
            return false;
        }
    }
    return true;
}

bool check_n(int32_t n) {
    if (n <= 0) {
        char buf[256];
        std::snprintf(buf, sizeof buf, "bad batch size: %d", n);
        set_msg(buf);
        return false;
    }
    return true;
}

#define QK_P(x) std::pair<const void*, const char*>(static_cast<const void*>(x), #x)

const QKPluginAPI k_api = {QK_ABI_MAJOR, QK_ABI_MINOR,
                           "quantkernel.mi355x.gfx950.mc_batch.v1"};

int32_t run(const qk::HostBatch& hb, int32_t n, double* out, double* se) {
    std::string err;
    const int32_t rc = guard(err, [&] {
        return qk::run_mc_batch(hb, n, out, se, &err);
    });
    if (rc != QK_OK) set_msg(err);
    return rc;
}

// GBM families sharing (spot, strike, t, vol, r, q, ot, paths[, steps], seed).
int32_t gbm_batch(int32_t method, const double* spot, const double* strike, const double* t,
                  const double* vol, const double* r, const double* q, const int32_t* ot,
                  const int32_t* paths, const int32_t* steps, const uint64_t* seed, int32_t n,
                  double* out, double* se) {
    qk::HostBatch hb;
    hb.uniform_method = method;
    hb.spot = spot;
    hb.strike = strike;
    hb.t = t;
    hb.vol = vol;
    hb.r = r;
    hb.q = q;
    hb.ot = ot;
    hb.paths = paths;
    hb.steps = steps;
    hb.seed = seed;
    return run(hb, n, out, se);
}

int32_t heston_batch(const double* spot, const double* strike, const double* t, const double* r,
                     const double* q, const double* v0, const double* kappa, const double* theta,
                     const double* sigma, const double* rho, const int32_t* ot,
                     const int32_t* paths, const int32_t* steps, const uint64_t* seed, int32_t n,
                     double* out, double* se) {
    qk::HostBatch hb;
    hb.uniform_method = qk::M_HESTON;
    hb.spot = spot;
    hb.strike = strike;
    hb.t = t;
    hb.r = r;
    hb.q = q;
    hb.v0 = v0;
    hb.kappa = kappa;
    hb.theta = theta;
    hb.sigma = sigma;
    hb.rho = rho;
    hb.ot = ot;
    hb.paths = paths;
    hb.steps = steps;
    hb.seed = seed;
    return run(hb, n, out, se);
}

// GBM terminal estimators with a per-element parameter (shift / weight clip).
int32_t gbm_aux_batch(int32_t method, const double* spot, const double* strike, const double* t,
                      const double* vol, const double* r, const double* q, const int32_t* ot,
                      const int32_t* paths, const double* aux, const uint64_t* seed, int32_t n,
                      double* out, double* se) {
    qk::HostBatch hb;
    hb.uniform_method = method;
    hb.spot = spot;
    hb.strike = strike;
    hb.t = t;
    hb.vol = vol;
    hb.r = r;
    hb.q = q;
    hb.ot = ot;
    hb.paths = paths;
    hb.seed = seed;
    hb.aux = aux;
    return run(hb, n, out, se);
}

int32_t heston_lr_batch(const double* spot, const double* strike, const double* t,
                        const double* r, const double* q, const double* v0, const double* kappa,
                        const double* theta, const double* sigma, const double* rho,
                        const int32_t* ot, const int32_t* paths, const int32_t* steps,
                        const uint64_t* seed, const double* wclip, int32_t n, double* out,
                        double* se) {
    qk::HostBatch hb;
    hb.uniform_method = qk::M_HESTON_LR;
    hb.spot = spot;
    hb.strike = strike;
    hb.t = t;
    hb.r = r;
    hb.q = q;
    hb.v0 = v0;
    hb.kappa = kappa;
    hb.theta = theta;
    hb.sigma = sigma;
    hb.rho = rho;
    hb.ot = ot;
    hb.paths = paths;
    hb.steps = steps;
    hb.seed = seed;
    hb.aux = wclip;
    return run(hb, n, out, se);
}

// Quasi-MC has no seed in the reference's signature; the engine takes one
// per element and ignores it for point sets.  The seed vector is allocated
// inside guard so a failed allocation becomes an error code.
int32_t qmc_batch(int32_t method, const double* spot, const double* strike, const double* t,
                  const double* vol, const double* r, const double* q, const int32_t* ot,
                  const int32_t* paths, int32_t n, double* out, double* se) {
    std::vector<uint64_t> seed;
    std::string err;
    const int32_t rc = guard(err, [&] {
        seed.assign(static_cast<size_t>(n), 0);
        return QK_OK;
    });
    if (rc != QK_OK) {
        set_msg(err);
        return rc;
    }
    return gbm_batch(method, spot, strike, t, vol, r, q, ot, paths, nullptr, seed.data(), n, out,
                     se);
}

// Multilevel MC (multilevel_monte_carlo.cpp:10-72): each option becomes one
// element per level (paths max(16, base_paths >> l), steps base_steps << l,
// aux = l) in one batch; the levels' discounted means are then summed in
// level order, and their standard errors in quadrature.  base_paths <= 1,
// levels <= 0 or base_steps <= 0 give NaN (multilevel_monte_carlo.cpp:14-17);
// so do level counts whose finest step count overflows int32 (the reference
// shifts past 2^31 there).
int32_t mlmc_batch_impl(const double* spot, const double* strike, const double* t,
                        const double* vol, const double* r, const double* q, const int32_t* ot,
                        const int32_t* base_paths, const int32_t* levels,
                        const int32_t* base_steps, const uint64_t* seed, int32_t n, double* out,
                        double* se) {
    std::vector<int64_t> first(static_cast<size_t>(n));
    std::vector<int32_t> count(static_cast<size_t>(n));
    int64_t total = 0;
    for (int32_t i = 0; i < n; ++i) {
        const int32_t L = levels[i];
        const bool ok = base_paths[i] > 1 && L > 0 && base_steps[i] > 0 && L <= 31 &&
                        (int64_t(base_steps[i]) << (L - 1)) <= 0x7fffffffLL;
        count[i] = ok ? L : 0;
        first[i] = total;
        total += count[i];
    }
    if (total > 0x7fffffffLL) {
        set_msg("multilevel batch too large: more than 2^31-1 levels in total");
        return QK_ERR_INVALID_INPUT;
    }
    const size_t m = static_cast<size_t>(total);
    std::vector<double> es(m), ek(m), et(m), ev(m), er(m), eq(m), ea(m), sp(m), ss(m);
    std::vector<int32_t> eot(m), ep(m), est(m);
    std::vector<uint64_t> eseed(m);
    std::vector<uint8_t> cont(m);
    for (int32_t i = 0; i < n; ++i)
        for (int32_t l = 0; l < count[i]; ++l) {
            const size_t j = static_cast<size_t>(first[i] + l);
            es[j] = spot[i];
            ek[j] = strike[i];
            et[j] = t[i];
            ev[j] = vol[i];
            er[j] = r[i];
            eq[j] = q[i];
            eot[j] = ot[i];
            ep[j] = std::max(16, base_paths[i] >> l);
            est[j] = base_steps[i] << l;
            eseed[j] = seed[i];
            ea[j] = double(l);
            cont[j] = l > 0;
        }
    if (total > 0) {
        qk::HostBatch hb;
        hb.uniform_method = qk::M_MLMC_LEVEL;
        hb.spot = es.data();
        hb.strike = ek.data();
        hb.t = et.data();
        hb.vol = ev.data();
        hb.r = er.data();
        hb.q = eq.data();
        hb.ot = eot.data();
        hb.paths = ep.data();
        hb.steps = est.data();
        hb.seed = eseed.data();
        hb.aux = ea.data();
        hb.stream_continues = cont.data();
        const int32_t rc = run(hb, int32_t(total), sp.data(), ss.data());
        if (rc != QK_OK) return rc;
    }
    for (int32_t i = 0; i < n; ++i) {
        if (count[i] == 0) {
            out[i] = qk::qnan_d();
            if (se) se[i] = qk::qnan_d();
            continue;
        }
        double est = 0.0, var = 0.0;
        for (int32_t l = 0; l < count[i]; ++l) {
            const size_t j = static_cast<size_t>(first[i] + l);
            est += sp[j];
            var += ss[j] * ss[j];
        }
        out[i] = est;
        if (se) se[i] = std::sqrt(var);
    }
    return QK_OK;
}

// mlmc_batch_impl builds its per-level expansion in host vectors.
int32_t mlmc_batch(const double* spot, const double* strike, const double* t, const double* vol,
                   const double* r, const double* q, const int32_t* ot, const int32_t* base_paths,
                   const int32_t* levels, const int32_t* base_steps, const uint64_t* seed,
                   int32_t n, double* out, double* se) {
    std::string err;
    const int32_t rc = guard(err, [&] {
        return mlmc_batch_impl(spot, strike, t, vol, r, q, ot, base_paths, levels, base_steps,
                               seed, n, out, se);
    });
    if (!err.empty()) set_msg(err);
    return rc;
}

double scalar_of(int32_t rc, double v) { return rc == QK_OK ? v : qk::qnan_d(); }

}  // namespace

namespace qk {
// The thread-local last error, for entry points defined in other files.
void set_last_error(const char* msg) { set_msg(msg); }
}  // namespace qk

namespace {

// Shared by the two device-resident mixed entry points: pointer and size
// checks, then the BatchIn of the device arrays.
bool mixed_device_args(const int32_t* model, const double* spot, const double* strike,
                       const double* t, const double* vol, const double* r, const double* q,
                       const double* v0, const double* kappa, const double* theta,
                       const double* sigma, const double* rho, const int32_t* option_type,
                       const int32_t* paths, const int32_t* steps, const uint64_t* seed,
                       int32_t n, double* out_prices, double* out_stderr, qk::BatchIn& bi,
                       int32_t& rc) {
    if (!check_ptrs({QK_P(model), QK_P(spot), QK_P(strike), QK_P(t), QK_P(vol), QK_P(r), QK_P(q),
                     QK_P(v0), QK_P(kappa), QK_P(theta), QK_P(sigma), QK_P(rho),
                     QK_P(option_type), QK_P(paths), QK_P(steps), QK_P(seed), QK_P(out_prices),
                     QK_P(out_stderr)})) {
        rc = QK_ERR_NULL_PTR;
        return false;
    }
    if (!check_n(n)) {
        rc = QK_ERR_BAD_SIZE;
        return false;
    }
    bi = qk::BatchIn{};
    bi.method = model;
    bi.method_is_model = 1;
    bi.spot = spot;
    bi.strike = strike;
    bi.t = t;
    bi.vol = vol;
    bi.r = r;
    bi.q = q;
    bi.v0 = v0;
    bi.kappa = kappa;
    bi.theta = theta;
    bi.sigma = sigma;
    bi.rho = rho;
    bi.ot = option_type;
    bi.paths = paths;
    bi.steps = steps;
    bi.seed = seed;
    return true;
}

}  // namespace

extern "C" {

// ---- ABI negotiation and errors (reference qk_api.cpp:94-117) -------------
void qk_abi_version(int32_t* major, int32_t* minor) {
    if (major) *major = QK_ABI_MAJOR;
    if (minor) *minor = QK_ABI_MINOR;
}

int32_t qk_plugin_get_api(int32_t host_abi_major, int32_t host_abi_minor,
                          const QKPluginAPI** out_api) {
    if (!out_api) return QK_ERR_NULL_PTR;
    if (host_abi_major != QK_ABI_MAJOR || host_abi_minor > QK_ABI_MINOR) {
        *out_api = nullptr;
        return QK_ERR_ABI_MISMATCH;
    }
    *out_api = &k_api;
    return QK_OK;
}

const char* qk_get_last_error(void) { return tl_msg; }

void qk_clear_last_error(void) { tl_msg[0] = '\0'; }

// ---- standard MC (reference qk_api.cpp:837-867) ---------------------------
int32_t qk_mcm_standard_monte_carlo_price_batch_se(const double* spot, const double* strike,
                                                   const double* t, const double* vol,
                                                   const double* r, const double* q,
                                                   const int32_t* option_type,
                                                   const int32_t* paths, const uint64_t* seed,
                                                   int32_t n, double* out_prices,
                                                   double* out_stderr) {
    if (!check_ptrs({QK_P(spot), QK_P(strike), QK_P(t), QK_P(vol), QK_P(r), QK_P(q),
                     QK_P(option_type), QK_P(paths), QK_P(seed), QK_P(out_prices),
                     QK_P(out_stderr)}))
        return QK_ERR_NULL_PTR;
    if (!check_n(n)) return QK_ERR_BAD_SIZE;
    return gbm_batch(qk::M_STANDARD, spot, strike, t, vol, r, q, option_type, paths, nullptr, seed,
                     n, out_prices, out_stderr);
}

int32_t qk_mcm_standard_monte_carlo_price_batch(const double* spot, const double* strike,
                                                const double* t, const double* vol,
                                                const double* r, const double* q,
                                                const int32_t* option_type, const int32_t* paths,
                                                const uint64_t* seed, int32_t n,
                                                double* out_prices) {
    if (!check_ptrs({QK_P(spot), QK_P(strike), QK_P(t), QK_P(vol), QK_P(r), QK_P(q),
                     QK_P(option_type), QK_P(paths), QK_P(seed), QK_P(out_prices)}))
        return QK_ERR_NULL_PTR;
    if (!check_n(n)) return QK_ERR_BAD_SIZE;
    return gbm_batch(qk::M_STANDARD, spot, strike, t, vol, r, q, option_type, paths, nullptr, seed,
                     n, out_prices, nullptr);
}

double qk_mcm_standard_monte_carlo_price(double spot, double strike, double t, double vol,
                                         double r, double q, int32_t option_type, int32_t paths,
                                         uint64_t seed) {
    double out = 0.0;
    const int32_t rc = gbm_batch(qk::M_STANDARD, &spot, &strike, &t, &vol, &r, &q, &option_type,
                                 &paths, nullptr, &seed, 1, &out, nullptr);
    return scalar_of(rc, out);
}

// ---- antithetic variates (reference qk_api.cpp:1091-1107) -----------------
int32_t qk_mcm_antithetic_variates_price_batch(const double* spot, const double* strike,
                                               const double* t, const double* vol,
                                               const double* r, const double* q,
                                               const int32_t* option_type, const int32_t* paths,
                                               const uint64_t* seed, int32_t n,
                                               double* out_prices) {
    if (!check_ptrs({QK_P(spot), QK_P(strike), QK_P(t), QK_P(vol), QK_P(r), QK_P(q),
                     QK_P(option_type), QK_P(paths), QK_P(seed), QK_P(out_prices)}))
        return QK_ERR_NULL_PTR;
    if (!check_n(n)) return QK_ERR_BAD_SIZE;
    return gbm_batch(qk::M_ANTITHETIC, spot, strike, t, vol, r, q, option_type, paths, nullptr,
                     seed, n, out_prices, nullptr);
}

double qk_mcm_antithetic_variates_price(double spot, double strike, double t, double vol,
                                        double r, double q, int32_t option_type, int32_t paths,
                                        uint64_t seed) {
    double out = 0.0;
    const int32_t rc = gbm_batch(qk::M_ANTITHETIC, &spot, &strike, &t, &vol, &r, &q, &option_type,
                                 &paths, nullptr, &seed, 1, &out, nullptr);
    return scalar_of(rc, out);
}

// ---- stepwise GBM families (reference qk_api.cpp:950-984, 1210-1227) ------
#define QK_STEPWISE_ENTRY(NAME, METHOD)                                                          \
    int32_t NAME##_batch(const double* spot, const double* strike, const double* t,              \
                         const double* vol, const double* r, const double* q,                    \
                         const int32_t* option_type, const int32_t* paths,                       \
                         const int32_t* steps, const uint64_t* seed, int32_t n,                  \
                         double* out_prices) {                                                   \
        if (!check_ptrs({QK_P(spot), QK_P(strike), QK_P(t), QK_P(vol), QK_P(r), QK_P(q),         \
                         QK_P(option_type), QK_P(paths), QK_P(steps), QK_P(seed),                \
                         QK_P(out_prices)}))                                                     \
            return QK_ERR_NULL_PTR;                                                              \
        if (!check_n(n)) return QK_ERR_BAD_SIZE;                                                 \
        return gbm_batch(METHOD, spot, strike, t, vol, r, q, option_type, paths, steps, seed, n, \
                         out_prices, nullptr);                                                   \
    }                                                                                            \
    int32_t NAME##_batch_se(const double* spot, const double* strike, const double* t,           \
                            const double* vol, const double* r, const double* q,                 \
                            const int32_t* option_type, const int32_t* paths,                    \
                            const int32_t* steps, const uint64_t* seed, int32_t n,               \
                            double* out_prices, double* out_stderr) {                            \
        if (!check_ptrs({QK_P(spot), QK_P(strike), QK_P(t), QK_P(vol), QK_P(r), QK_P(q),         \
                         QK_P(option_type), QK_P(paths), QK_P(steps), QK_P(seed),                \
                         QK_P(out_prices), QK_P(out_stderr)}))                                   \
            return QK_ERR_NULL_PTR;                                                              \
        if (!check_n(n)) return QK_ERR_BAD_SIZE;                                                 \
        return gbm_batch(METHOD, spot, strike, t, vol, r, q, option_type, paths, steps, seed, n, \
                         out_prices, out_stderr);                                                \
    }                                                                                            \
    double NAME(double spot, double strike, double t, double vol, double r, double q,            \
                int32_t option_type, int32_t paths, int32_t steps, uint64_t seed) {              \
        double out = 0.0;                                                                        \
        const int32_t rc = gbm_batch(METHOD, &spot, &strike, &t, &vol, &r, &q, &option_type,     \
                                     &paths, &steps, &seed, 1, &out, nullptr);                   \
        return scalar_of(rc, out);                                                               \
    }

QK_STEPWISE_ENTRY(qk_mcm_euler_maruyama_price, qk::M_EULER)
QK_STEPWISE_ENTRY(qk_mcm_milstein_price, qk::M_MILSTEIN)
QK_STEPWISE_ENTRY(qk_mcm_local_vol_monte_carlo_price, qk::M_LOCAL_VOL)

// ---- Heston (reference qk_api.cpp:1127-1161) ------------------------------
int32_t qk_mcm_heston_monte_carlo_price_batch(const double* spot, const double* strike,
                                              const double* t, const double* r, const double* q,
                                              const double* v0, const double* kappa,
                                              const double* theta, const double* sigma,
                                              const double* rho, const int32_t* option_type,
                                              const int32_t* paths, const int32_t* steps,
                                              const uint64_t* seed, int32_t n,
                                              double* out_prices) {
    if (!check_ptrs({QK_P(spot), QK_P(strike), QK_P(t), QK_P(r), QK_P(q), QK_P(v0), QK_P(kappa),
                     QK_P(theta), QK_P(sigma), QK_P(rho), QK_P(option_type), QK_P(paths),
                     QK_P(steps), QK_P(seed), QK_P(out_prices)}))
        return QK_ERR_NULL_PTR;
    if (!check_n(n)) return QK_ERR_BAD_SIZE;
    return heston_batch(spot, strike, t, r, q, v0, kappa, theta, sigma, rho, option_type, paths,
                        steps, seed, n, out_prices, nullptr);
}

int32_t qk_mcm_heston_monte_carlo_price_batch_se(const double* spot, const double* strike,
                                                 const double* t, const double* r,
                                                 const double* q, const double* v0,
                                                 const double* kappa, const double* theta,
                                                 const double* sigma, const double* rho,
                                                 const int32_t* option_type, const int32_t* paths,
                                                 const int32_t* steps, const uint64_t* seed,
                                                 int32_t n, double* out_prices,
                                                 double* out_stderr) {
    if (!check_ptrs({QK_P(spot), QK_P(strike), QK_P(t), QK_P(r), QK_P(q), QK_P(v0), QK_P(kappa),
                     QK_P(theta), QK_P(sigma), QK_P(rho), QK_P(option_type), QK_P(paths),
                     QK_P(steps), QK_P(seed), QK_P(out_prices), QK_P(out_stderr)}))
        return QK_ERR_NULL_PTR;
    if (!check_n(n)) return QK_ERR_BAD_SIZE;
    return heston_batch(spot, strike, t, r, q, v0, kappa, theta, sigma, rho, option_type, paths,
                        steps, seed, n, out_prices, out_stderr);
}

double qk_mcm_heston_monte_carlo_price(double spot, double strike, double t, double r, double q,
                                       double v0, double kappa, double theta, double sigma,
                                       double rho, int32_t option_type, int32_t paths,
                                       int32_t steps, uint64_t seed) {
    double out = 0.0;
    const int32_t rc = heston_batch(&spot, &strike, &t, &r, &q, &v0, &kappa, &theta, &sigma, &rho,
                                    &option_type, &paths, &steps, &seed, 1, &out, nullptr);
    return scalar_of(rc, out);
}

// ---- importance sampling (reference qk_api.cpp:918-925, 1055-1071) --------
int32_t qk_mcm_importance_sampling_price_batch_se(const double* spot, const double* strike,
                                                  const double* t, const double* vol,
                                                  const double* r, const double* q,
                                                  const int32_t* option_type,
                                                  const int32_t* paths, const double* shift,
                                                  const uint64_t* seed, int32_t n,
                                                  double* out_prices, double* out_stderr) {
    if (!check_ptrs({QK_P(spot), QK_P(strike), QK_P(t), QK_P(vol), QK_P(r), QK_P(q),
                     QK_P(option_type), QK_P(paths), QK_P(shift), QK_P(seed), QK_P(out_prices),
                     QK_P(out_stderr)}))
        return QK_ERR_NULL_PTR;
    if (!check_n(n)) return QK_ERR_BAD_SIZE;
    return gbm_aux_batch(qk::M_IMPORTANCE, spot, strike, t, vol, r, q, option_type, paths, shift,
                         seed, n, out_prices, out_stderr);
}

int32_t qk_mcm_importance_sampling_price_batch(const double* spot, const double* strike,
                                               const double* t, const double* vol,
                                               const double* r, const double* q,
                                               const int32_t* option_type, const int32_t* paths,
                                               const double* shift, const uint64_t* seed,
                                               int32_t n, double* out_prices) {
    if (!check_ptrs({QK_P(spot), QK_P(strike), QK_P(t), QK_P(vol), QK_P(r), QK_P(q),
                     QK_P(option_type), QK_P(paths), QK_P(shift), QK_P(seed),
                     QK_P(out_prices)}))
        return QK_ERR_NULL_PTR;
    if (!check_n(n)) return QK_ERR_BAD_SIZE;
    return gbm_aux_batch(qk::M_IMPORTANCE, spot, strike, t, vol, r, q, option_type, paths, shift,
                         seed, n, out_prices, nullptr);
}

double qk_mcm_importance_sampling_price(double spot, double strike, double t, double vol,
                                        double r, double q, int32_t option_type, int32_t paths,
                                        double shift, uint64_t seed) {
    double out = 0.0;
    const int32_t rc = gbm_aux_batch(qk::M_IMPORTANCE, &spot, &strike, &t, &vol, &r, &q,
                                     &option_type, &paths, &shift, &seed, 1, &out, nullptr);
    return scalar_of(rc, out);
}

// ---- Heston likelihood-ratio delta (reference qk_api.cpp:1163-1199) -------
int32_t qk_mcm_heston_lr_delta_batch_se(const double* spot, const double* strike, const double* t,
                                        const double* r, const double* q, const double* v0,
                                        const double* kappa, const double* theta,
                                        const double* sigma, const double* rho,
                                        const int32_t* option_type, const int32_t* paths,
                                        const int32_t* steps, const uint64_t* seed,
                                        const double* weight_clip, int32_t n,
                                        double* out_prices, double* out_stderr) {
    if (!check_ptrs({QK_P(spot), QK_P(strike), QK_P(t), QK_P(r), QK_P(q), QK_P(v0), QK_P(kappa),
                     QK_P(theta), QK_P(sigma), QK_P(rho), QK_P(option_type), QK_P(paths),
                     QK_P(steps), QK_P(seed), QK_P(weight_clip), QK_P(out_prices),
                     QK_P(out_stderr)}))
        return QK_ERR_NULL_PTR;
    if (!check_n(n)) return QK_ERR_BAD_SIZE;
    return heston_lr_batch(spot, strike, t, r, q, v0, kappa, theta, sigma, rho, option_type, paths,
                           steps, seed, weight_clip, n, out_prices, out_stderr);
}

int32_t qk_mcm_heston_lr_delta_batch(const double* spot, const double* strike, const double* t,
                                     const double* r, const double* q, const double* v0,
                                     const double* kappa, const double* theta,
                                     const double* sigma, const double* rho,
                                     const int32_t* option_type, const int32_t* paths,
                                     const int32_t* steps, const uint64_t* seed,
                                     const double* weight_clip, int32_t n, double* out_prices) {
    if (!check_ptrs({QK_P(spot), QK_P(strike), QK_P(t), QK_P(r), QK_P(q), QK_P(v0), QK_P(kappa),
                     QK_P(theta), QK_P(sigma), QK_P(rho), QK_P(option_type), QK_P(paths),
                     QK_P(steps), QK_P(seed), QK_P(weight_clip), QK_P(out_prices)}))
        return QK_ERR_NULL_PTR;
    if (!check_n(n)) return QK_ERR_BAD_SIZE;
    return heston_lr_batch(spot, strike, t, r, q, v0, kappa, theta, sigma, rho, option_type, paths,
                           steps, seed, weight_clip, n, out_prices, nullptr);
}

double qk_mcm_heston_lr_delta(double spot, double strike, double t, double r, double q,
                              double v0, double kappa, double theta, double sigma, double rho,
                              int32_t option_type, int32_t paths, int32_t steps, uint64_t seed,
                              double weight_clip) {
    double out = 0.0;
    const int32_t rc = heston_lr_batch(&spot, &strike, &t, &r, &q, &v0, &kappa, &theta, &sigma,
                                       &rho, &option_type, &paths, &steps, &seed, &weight_clip, 1,
                                       &out, nullptr);
    return scalar_of(rc, out);
}

// ---- pathwise / likelihood-ratio delta (reference qk_api.cpp:1650-1667,
//      1873-1907) ---------------------------------------------------------------
int32_t qk_agm_pathwise_derivative_delta_batch_se(const double* spot, const double* strike,
                                                  const double* t, const double* vol,
                                                  const double* r, const double* q,
                                                  const int32_t* option_type,
                                                  const int32_t* paths, const uint64_t* seed,
                                                  int32_t n, double* out_prices,
                                                  double* out_stderr) {
    if (!check_ptrs({QK_P(spot), QK_P(strike), QK_P(t), QK_P(vol), QK_P(r), QK_P(q),
                     QK_P(option_type), QK_P(paths), QK_P(seed), QK_P(out_prices),
                     QK_P(out_stderr)}))
        return QK_ERR_NULL_PTR;
    if (!check_n(n)) return QK_ERR_BAD_SIZE;
    return gbm_aux_batch(qk::M_PATHWISE, spot, strike, t, vol, r, q, option_type, paths, nullptr,
                         seed, n, out_prices, out_stderr);
}

int32_t qk_agm_pathwise_derivative_delta_batch(const double* spot, const double* strike,
                                               const double* t, const double* vol,
                                               const double* r, const double* q,
                                               const int32_t* option_type, const int32_t* paths,
                                               const uint64_t* seed, int32_t n,
                                               double* out_prices) {
    if (!check_ptrs({QK_P(spot), QK_P(strike), QK_P(t), QK_P(vol), QK_P(r), QK_P(q),
                     QK_P(option_type), QK_P(paths), QK_P(seed), QK_P(out_prices)}))
        return QK_ERR_NULL_PTR;
    if (!check_n(n)) return QK_ERR_BAD_SIZE;
    return gbm_aux_batch(qk::M_PATHWISE, spot, strike, t, vol, r, q, option_type, paths, nullptr,
                         seed, n, out_prices, nullptr);
}

double qk_agm_pathwise_derivative_delta(double spot, double strike, double t, double vol,
                                        double r, double q, int32_t option_type, int32_t paths,
                                        uint64_t seed) {
    double out = 0.0;
    const int32_t rc = gbm_aux_batch(qk::M_PATHWISE, &spot, &strike, &t, &vol, &r, &q,
                                     &option_type, &paths, nullptr, &seed, 1, &out, nullptr);
    return scalar_of(rc, out);
}

int32_t qk_agm_likelihood_ratio_delta_batch_se(const double* spot, const double* strike,
                                               const double* t, const double* vol,
                                               const double* r, const double* q,
                                               const int32_t* option_type, const int32_t* paths,
                                               const uint64_t* seed, const double* weight_clip,
                                               int32_t n, double* out_prices,
                                               double* out_stderr) {
    if (!check_ptrs({QK_P(spot), QK_P(strike), QK_P(t), QK_P(vol), QK_P(r), QK_P(q),
                     QK_P(option_type), QK_P(paths), QK_P(seed), QK_P(weight_clip),
                     QK_P(out_prices), QK_P(out_stderr)}))
        return QK_ERR_NULL_PTR;
    if (!check_n(n)) return QK_ERR_BAD_SIZE;
    return gbm_aux_batch(qk::M_LR, spot, strike, t, vol, r, q, option_type, paths, weight_clip,
                         seed, n, out_prices, out_stderr);
}

int32_t qk_agm_likelihood_ratio_delta_batch(const double* spot, const double* strike,
                                            const double* t, const double* vol, const double* r,
                                            const double* q, const int32_t* option_type,
                                            const int32_t* paths, const uint64_t* seed,
                                            const double* weight_clip, int32_t n,
                                            double* out_prices) {
    if (!check_ptrs({QK_P(spot), QK_P(strike), QK_P(t), QK_P(vol), QK_P(r), QK_P(q),
                     QK_P(option_type), QK_P(paths), QK_P(seed), QK_P(weight_clip),
                     QK_P(out_prices)}))
        return QK_ERR_NULL_PTR;
    if (!check_n(n)) return QK_ERR_BAD_SIZE;
    return gbm_aux_batch(qk::M_LR, spot, strike, t, vol, r, q, option_type, paths, weight_clip,
                         seed, n, out_prices, nullptr);
}

double qk_agm_likelihood_ratio_delta(double spot, double strike, double t, double vol, double r,
                                     double q, int32_t option_type, int32_t paths, uint64_t seed,
                                     double weight_clip) {
    double out = 0.0;
    const int32_t rc = gbm_aux_batch(qk::M_LR, &spot, &strike, &t, &vol, &r, &q, &option_type,
                                     &paths, &weight_clip, &seed, 1, &out, nullptr);
    return scalar_of(rc, out);
}

// ---- control variates (reference qk_api.cpp:926-934, 1073-1089) ----
int32_t qk_mcm_control_variates_price_batch_se(const double* spot, const double* strike, const double* t,
        const double* vol, const double* r, const double* q, const int32_t* option_type,
        const int32_t* paths, const uint64_t* seed, int32_t n, double* out_prices,
        double* out_stderr) {
    if (!check_ptrs({QK_P(spot), QK_P(strike), QK_P(t), QK_P(vol), QK_P(r), QK_P(q),
                     QK_P(option_type), QK_P(paths), QK_P(seed), QK_P(out_prices),
                     QK_P(out_stderr)}))
        return QK_ERR_NULL_PTR;
    if (!check_n(n)) return QK_ERR_BAD_SIZE;
    return gbm_batch(qk::M_CONTROL, spot, strike, t, vol, r, q, option_type, paths, nullptr, seed, n,
                     out_prices, out_stderr);
}

int32_t qk_mcm_control_variates_price_batch(const double* spot, const double* strike, const double* t,
        const double* vol, const double* r, const double* q, const int32_t* option_type,
        const int32_t* paths, const uint64_t* seed, int32_t n, double* out_prices) {
    if (!check_ptrs({QK_P(spot), QK_P(strike), QK_P(t), QK_P(vol), QK_P(r), QK_P(q),
                     QK_P(option_type), QK_P(paths), QK_P(seed), QK_P(out_prices)}))
        return QK_ERR_NULL_PTR;
    if (!check_n(n)) return QK_ERR_BAD_SIZE;
    return gbm_batch(qk::M_CONTROL, spot, strike, t, vol, r, q, option_type, paths, nullptr, seed, n,
                     out_prices, nullptr);
}

double qk_mcm_control_variates_price(double spot, double strike, double t, double vol, double r, double q,
        int32_t option_type, int32_t paths, uint64_t seed) {
    double out = 0.0;
    const int32_t rc = gbm_batch(qk::M_CONTROL, &spot, &strike, &t, &vol, &r, &q, &option_type, &paths,
                                 nullptr, &seed, 1, &out, nullptr);
    return scalar_of(rc, out);
}

// ---- stratified sampling (reference qk_api.cpp:942-950, 1109-1125) ----
int32_t qk_mcm_stratified_sampling_price_batch_se(const double* spot, const double* strike, const double* t,
        const double* vol, const double* r, const double* q, const int32_t* option_type,
        const int32_t* paths, const uint64_t* seed, int32_t n, double* out_prices,
        double* out_stderr) {
    if (!check_ptrs({QK_P(spot), QK_P(strike), QK_P(t), QK_P(vol), QK_P(r), QK_P(q),
                     QK_P(option_type), QK_P(paths), QK_P(seed), QK_P(out_prices),
                     QK_P(out_stderr)}))
        return QK_ERR_NULL_PTR;
    if (!check_n(n)) return QK_ERR_BAD_SIZE;
    return gbm_batch(qk::M_STRATIFIED, spot, strike, t, vol, r, q, option_type, paths, nullptr, seed, n,
                     out_prices, out_stderr);
}

int32_t qk_mcm_stratified_sampling_price_batch(const double* spot, const double* strike, const double* t,
        const double* vol, const double* r, const double* q, const int32_t* option_type,
        const int32_t* paths, const uint64_t* seed, int32_t n, double* out_prices) {
    if (!check_ptrs({QK_P(spot), QK_P(strike), QK_P(t), QK_P(vol), QK_P(r), QK_P(q),
                     QK_P(option_type), QK_P(paths), QK_P(seed), QK_P(out_prices)}))
        return QK_ERR_NULL_PTR;
    if (!check_n(n)) return QK_ERR_BAD_SIZE;
    return gbm_batch(qk::M_STRATIFIED, spot, strike, t, vol, r, q, option_type, paths, nullptr, seed, n,
                     out_prices, nullptr);
}

double qk_mcm_stratified_sampling_price(double spot, double strike, double t, double vol, double r, double q,
        int32_t option_type, int32_t paths, uint64_t seed) {
    double out = 0.0;
    const int32_t rc = gbm_batch(qk::M_STRATIFIED, &spot, &strike, &t, &vol, &r, &q, &option_type, &paths,
                                 nullptr, &seed, 1, &out, nullptr);
    return scalar_of(rc, out);
}

// ---- quasi-MC, Sobol (reference qk_api.cpp:893-899, 1004-1019) ----
int32_t qk_mcm_quasi_monte_carlo_sobol_price_batch_se(const double* spot, const double* strike, const double* t,
        const double* vol, const double* r, const double* q, const int32_t* option_type,
        const int32_t* paths, int32_t n, double* out_prices, double* out_stderr) {
    if (!check_ptrs({QK_P(spot), QK_P(strike), QK_P(t), QK_P(vol), QK_P(r), QK_P(q),
                     QK_P(option_type), QK_P(paths), QK_P(out_prices), QK_P(out_stderr)}))
        return QK_ERR_NULL_PTR;
    if (!check_n(n)) return QK_ERR_BAD_SIZE;
    return qmc_batch(qk::M_SOBOL, spot, strike, t, vol, r, q, option_type, paths, n, out_prices,
                     out_stderr);
}

int32_t qk_mcm_quasi_monte_carlo_sobol_price_batch(const double* spot, const double* strike, const double* t,
        const double* vol, const double* r, const double* q, const int32_t* option_type,
        const int32_t* paths, int32_t n, double* out_prices) {
    if (!check_ptrs({QK_P(spot), QK_P(strike), QK_P(t), QK_P(vol), QK_P(r), QK_P(q),
                     QK_P(option_type), QK_P(paths), QK_P(out_prices)}))
        return QK_ERR_NULL_PTR;
    if (!check_n(n)) return QK_ERR_BAD_SIZE;
    return qmc_batch(qk::M_SOBOL, spot, strike, t, vol, r, q, option_type, paths, n, out_prices,
                     nullptr);
}

double qk_mcm_quasi_monte_carlo_sobol_price(double spot, double strike, double t, double vol, double r, double q,
        int32_t option_type, int32_t paths) {
    double out = 0.0;
    const int32_t rc = qmc_batch(qk::M_SOBOL, &spot, &strike, &t, &vol, &r, &q, &option_type, &paths,
                                 1, &out, nullptr);
    return scalar_of(rc, out);
}

// ---- quasi-MC, Halton (reference qk_api.cpp:901-907, 1020-1035) ----
int32_t qk_mcm_quasi_monte_carlo_halton_price_batch_se(const double* spot, const double* strike, const double* t,
        const double* vol, const double* r, const double* q, const int32_t* option_type,
        const int32_t* paths, int32_t n, double* out_prices, double* out_stderr) {
    if (!check_ptrs({QK_P(spot), QK_P(strike), QK_P(t), QK_P(vol), QK_P(r), QK_P(q),
                     QK_P(option_type), QK_P(paths), QK_P(out_prices), QK_P(out_stderr)}))
        return QK_ERR_NULL_PTR;
    if (!check_n(n)) return QK_ERR_BAD_SIZE;
    return qmc_batch(qk::M_HALTON, spot, strike, t, vol, r, q, option_type, paths, n, out_prices,
                     out_stderr);
}

int32_t qk_mcm_quasi_monte_carlo_halton_price_batch(const double* spot, const double* strike, const double* t,
        const double* vol, const double* r, const double* q, const int32_t* option_type,
        const int32_t* paths, int32_t n, double* out_prices) {
    if (!check_ptrs({QK_P(spot), QK_P(strike), QK_P(t), QK_P(vol), QK_P(r), QK_P(q),
                     QK_P(option_type), QK_P(paths), QK_P(out_prices)}))
        return QK_ERR_NULL_PTR;
    if (!check_n(n)) return QK_ERR_BAD_SIZE;
    return qmc_batch(qk::M_HALTON, spot, strike, t, vol, r, q, option_type, paths, n, out_prices,
                     nullptr);
}

double qk_mcm_quasi_monte_carlo_halton_price(double spot, double strike, double t, double vol, double r, double q,
        int32_t option_type, int32_t paths) {
    double out = 0.0;
    const int32_t rc = qmc_batch(qk::M_HALTON, &spot, &strike, &t, &vol, &r, &q, &option_type, &paths,
                                 1, &out, nullptr);
    return scalar_of(rc, out);
}

// ---- multilevel MC (reference qk_api.cpp:909-916, 1036-1053) ----
int32_t qk_mcm_multilevel_monte_carlo_price_batch_se(const double* spot, const double* strike,
        const double* t, const double* vol, const double* r, const double* q,
        const int32_t* option_type, const int32_t* base_paths, const int32_t* levels,
        const int32_t* base_steps, const uint64_t* seed, int32_t n, double* out_prices,
        double* out_stderr) {
    if (!check_ptrs({QK_P(spot), QK_P(strike), QK_P(t), QK_P(vol), QK_P(r), QK_P(q),
                     QK_P(option_type), QK_P(base_paths), QK_P(levels), QK_P(base_steps),
                     QK_P(seed), QK_P(out_prices), QK_P(out_stderr)}))
        return QK_ERR_NULL_PTR;
    if (!check_n(n)) return QK_ERR_BAD_SIZE;
    return mlmc_batch(spot, strike, t, vol, r, q, option_type, base_paths, levels, base_steps,
                      seed, n, out_prices, out_stderr);
}

int32_t qk_mcm_multilevel_monte_carlo_price_batch(const double* spot, const double* strike,
        const double* t, const double* vol, const double* r, const double* q,
        const int32_t* option_type, const int32_t* base_paths, const int32_t* levels,
        const int32_t* base_steps, const uint64_t* seed, int32_t n, double* out_prices) {
    if (!check_ptrs({QK_P(spot), QK_P(strike), QK_P(t), QK_P(vol), QK_P(r), QK_P(q),
                     QK_P(option_type), QK_P(base_paths), QK_P(levels), QK_P(base_steps),
                     QK_P(seed), QK_P(out_prices)}))
        return QK_ERR_NULL_PTR;
    if (!check_n(n)) return QK_ERR_BAD_SIZE;
    return mlmc_batch(spot, strike, t, vol, r, q, option_type, base_paths, levels, base_steps,
                      seed, n, out_prices, nullptr);
}

double qk_mcm_multilevel_monte_carlo_price(double spot, double strike, double t, double vol,
                                           double r, double q, int32_t option_type,
                                           int32_t base_paths, int32_t levels, int32_t base_steps,
                                           uint64_t seed) {
    double out = 0.0;
    const int32_t rc = mlmc_batch(&spot, &strike, &t, &vol, &r, &q, &option_type, &base_paths,
                                  &levels, &base_steps, &seed, 1, &out, nullptr);
    return scalar_of(rc, out);
}

// ---- Heston characteristic-function prices (reference qk_api.cpp:195-207,
//      260-281) ------------------------------------------------------------------
int32_t qk_cf_heston_price_cf_batch(const double* spot, const double* strike, const double* t,
                                    const double* r, const double* q, const double* v0,
                                    const double* kappa, const double* theta,
                                    const double* sigma, const double* rho,
                                    const int32_t* option_type, const int32_t* integration_steps,
                                    const double* integration_limit, int32_t n,
                                    double* out_prices) {
    if (!check_ptrs({QK_P(spot), QK_P(strike), QK_P(t), QK_P(r), QK_P(q), QK_P(v0), QK_P(kappa),
                     QK_P(theta), QK_P(sigma), QK_P(rho), QK_P(option_type),
                     QK_P(integration_steps), QK_P(integration_limit), QK_P(out_prices)}))
        return QK_ERR_NULL_PTR;
    if (!check_n(n)) return QK_ERR_BAD_SIZE;
    std::string err;
    const int32_t rc = guard(err, [&] {
        return qk::run_heston_cf_batch(spot, strike, t, r, q, v0, kappa, theta, sigma, rho,
                                       option_type, integration_steps, integration_limit, n,
                                       out_prices, &err);
    });
    if (rc != QK_OK) set_msg(err);
    return rc;
}

int32_t qk_cf_heston_price_cf_batch_device(const double* spot, const double* strike,
                                           const double* t, const double* r, const double* q,
                                           const double* v0, const double* kappa,
                                           const double* theta, const double* sigma,
                                           const double* rho, const int32_t* option_type,
                                           const int32_t* integration_steps,
                                           const double* integration_limit, int32_t n,
                                           double* out_prices) {
    if (!check_ptrs({QK_P(spot), QK_P(strike), QK_P(t), QK_P(r), QK_P(q), QK_P(v0), QK_P(kappa),
                     QK_P(theta), QK_P(sigma), QK_P(rho), QK_P(option_type),
                     QK_P(integration_steps), QK_P(integration_limit), QK_P(out_prices)}))
        return QK_ERR_NULL_PTR;
    if (!check_n(n)) return QK_ERR_BAD_SIZE;
    const qk::HestonCfIn in{spot, strike, t,     r,     q,   v0, kappa, theta, sigma, rho,
                            integration_limit, option_type, integration_steps};
    std::string err;
    const int32_t rc = guard(err, [&] {
        return qk::run_heston_cf_batch_device(in, n, out_prices, &err);
    });
    if (rc != QK_OK) set_msg(err);
    return rc;
}

double qk_cf_heston_price_cf(double spot, double strike, double t, double r, double q, double v0,
                             double kappa, double theta, double sigma, double rho,
                             int32_t option_type, int32_t integration_steps,
                             double integration_limit) {
    double out = 0.0;
    std::string err;
    const int32_t rc = guard(err, [&] {
        return qk::run_heston_cf_batch(&spot, &strike, &t, &r, &q, &v0, &kappa, &theta, &sigma,
                                       &rho, &option_type, &integration_steps, &integration_limit,
                                       1, &out, &err);
    });
    if (rc != QK_OK) set_msg(err);
    return scalar_of(rc, out);
}

// ---- mixed-model batches (ABI 2.13) ---------------------------------------
int32_t qk_mcm_mixed_price_batch(const int32_t* model, const double* spot, const double* strike,
                                 const double* t, const double* vol, const double* r,
                                 const double* q, const double* v0, const double* kappa,
                                 const double* theta, const double* sigma, const double* rho,
                                 const int32_t* option_type, const int32_t* paths,
                                 const int32_t* steps, const uint64_t* seed, int32_t n,
                                 double* out_prices, double* out_stderr) {
    if (!check_ptrs({QK_P(model), QK_P(spot), QK_P(strike), QK_P(t), QK_P(vol), QK_P(r), QK_P(q),
                     QK_P(v0), QK_P(kappa), QK_P(theta), QK_P(sigma), QK_P(rho),
                     QK_P(option_type), QK_P(paths), QK_P(steps), QK_P(seed), QK_P(out_prices),
                     QK_P(out_stderr)}))
        return QK_ERR_NULL_PTR;
    if (!check_n(n)) return QK_ERR_BAD_SIZE;
    qk::HostBatch hb;
    hb.model_ids = model;
    hb.spot = spot;
    hb.strike = strike;
    hb.t = t;
    hb.vol = vol;
    hb.r = r;
    hb.q = q;
    hb.v0 = v0;
    hb.kappa = kappa;
    hb.theta = theta;
    hb.sigma = sigma;
    hb.rho = rho;
    hb.ot = option_type;
    hb.paths = paths;
    hb.steps = steps;
    hb.seed = seed;
    return run(hb, n, out_prices, out_stderr);
}

int32_t qk_mcm_mixed_price_batch_device(const int32_t* model, const double* spot,
                                        const double* strike, const double* t, const double* vol,
                                        const double* r, const double* q, const double* v0,
                                        const double* kappa, const double* theta,
                                        const double* sigma, const double* rho,
                                        const int32_t* option_type, const int32_t* paths,
                                        const int32_t* steps, const uint64_t* seed, int32_t n,
                                        double* out_prices, double* out_stderr) {
    qk::BatchIn bi;
    int32_t rc = QK_OK;
    if (!mixed_device_args(model, spot, strike, t, vol, r, q, v0, kappa, theta, sigma, rho,
                           option_type, paths, steps, seed, n, out_prices, out_stderr, bi, rc))
        return rc;
    std::string err;
    rc = guard(err, [&] { return qk::run_mc_batch_device(bi, n, out_prices, out_stderr, &err); });
    if (rc != QK_OK) set_msg(err);
    return rc;
}

int32_t qk_mcm_mixed_price_batch_device_noalloc(const int32_t* model, const double* spot,
                                                const double* strike, const double* t,
                                                const double* vol, const double* r,
                                                const double* q, const double* v0,
                                                const double* kappa, const double* theta,
                                                const double* sigma, const double* rho,
                                                const int32_t* option_type, const int32_t* paths,
                                                const int32_t* steps, const uint64_t* seed,
                                                int32_t n, double* out_prices,
                                                double* out_stderr) {
    qk::BatchIn bi;
    int32_t rc = QK_OK;
    if (!mixed_device_args(model, spot, strike, t, vol, r, q, v0, kappa, theta, sigma, rho,
                           option_type, paths, steps, seed, n, out_prices, out_stderr, bi, rc))
        return rc;
    std::string err;
    rc = guard(err, [&] {
        return qk::run_mc_batch_device_noalloc(bi, n, out_prices, out_stderr, &err);
    });
    if (rc != QK_OK) set_msg(err);
    return rc;
}

int32_t qk_mcm_standard_monte_carlo_price_batch_device(const double* spot, const double* strike,
                                                       const double* t, const double* vol,
                                                       const double* r, const double* q,
                                                       const int32_t* option_type,
                                                       const int32_t* paths, const uint64_t* seed,
                                                       int32_t n, double* out_prices,
                                                       double* out_stderr) {
    if (!check_ptrs({QK_P(spot), QK_P(strike), QK_P(t), QK_P(vol), QK_P(r), QK_P(q),
                     QK_P(option_type), QK_P(paths), QK_P(seed), QK_P(out_prices),
                     QK_P(out_stderr)}))
        return QK_ERR_NULL_PTR;
    if (!check_n(n)) return QK_ERR_BAD_SIZE;
    qk::BatchIn bi{};
    bi.uniform_method = qk::M_STANDARD;
    bi.spot = spot;
    bi.strike = strike;
    bi.t = t;
    bi.vol = vol;
    bi.r = r;
    bi.q = q;
    bi.ot = option_type;
    bi.paths = paths;
    bi.seed = seed;
    std::string err;
    const int32_t rc = guard(err, [&] {
        return qk::run_mc_batch_device(bi, n, out_prices, out_stderr, &err);
    });
    if (rc != QK_OK) set_msg(err);
    return rc;
}

// ---- RNG mode, timing, device info (ABI 2.13) -----------------------------
int32_t qk_mc_set_rng_mode(int32_t mode) {
    if (mode != QK_RNG_PHILOX && mode != QK_RNG_MT19937_FP64 && mode != QK_RNG_MT19937_FP32) {
        char buf[256];
        std::snprintf(buf, sizeof buf, "unknown rng mode: %d", mode);
        set_msg(buf);
        return QK_ERR_INVALID_INPUT;
    }
    qk::set_rng_mode(mode);
    return QK_OK;
}

int32_t qk_mc_get_rng_mode(void) { return qk::get_rng_mode(); }

int32_t qk_mc_last_timing(double* ms_total, double* ms_simulate) {
    if (!qk::last_timing(ms_total, ms_simulate)) {
        set_msg("no MC batch has completed on the device yet");
        return QK_ERR_RUNTIME_INIT;
    }
    return QK_OK;
}

int32_t qk_device_info(char* name_buf, int32_t buf_len, int32_t* compute_units) {
    std::string name, err;
    int32_t cus = 0;
    if (!qk::device_info(&name, &cus, &err)) {
        set_msg(err);
        return QK_ERR_RUNTIME_INIT;
    }
    if (name_buf && buf_len > 0) {
        std::strncpy(name_buf, name.c_str(), size_t(buf_len) - 1);
        name_buf[buf_len - 1] = '\0';
    }
    if (compute_units) *compute_units = cus;
    return QK_OK;
}

int32_t qk_mc_philox_probe(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c3, int32_t count,
                           uint32_t* out_words, float* out_normals) {
    if (!check_ptrs({QK_P(out_words), QK_P(out_normals)})) return QK_ERR_NULL_PTR;
    if (!check_n(count)) return QK_ERR_BAD_SIZE;
    std::string err;
    const int32_t rc = guard(err, [&] {
        return qk::philox_probe(seed, c0, c1, c3, count, out_words, out_normals, &err);
    });
    if (rc != QK_OK) set_msg(err);
    return rc;
}

int32_t qk_mc_philox_probe8(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c3, int32_t count,
                            float* out_normals) {
    if (!check_ptrs({QK_P(out_normals)})) return QK_ERR_NULL_PTR;
    if (!check_n(count)) return QK_ERR_BAD_SIZE;
    std::string err;
    const int32_t rc = guard(err, [&] {
        return qk::philox_probe8(seed, c0, c1, c3, count, out_normals, &err);
    });
    if (rc != QK_OK) set_msg(err);
    return rc;
}

int32_t qk_mc_graphs(int32_t enable, int32_t* out_cached, int64_t* out_launches) {
    std::string err;
    const int32_t rc = guard(err, [&] {
        return qk::set_graphs(enable, out_cached, out_launches, &err);
    });
    if (rc != QK_OK) set_msg(err);
    return rc;
}

int32_t qk_mc_set_shape(int32_t family, int32_t lane_units, int32_t grid_per_cu) {
    std::string err;
    const int32_t rc = guard(err, [&] {
        return qk::set_shape(family, lane_units, grid_per_cu, &err);
    });
    if (rc != QK_OK) set_msg(err);
    return rc;
}

int32_t qk_mc_get_shape(int32_t family, int32_t* lane_units, int32_t* grid_per_cu) {
    if (!check_ptrs({QK_P(lane_units), QK_P(grid_per_cu)})) return QK_ERR_NULL_PTR;
    std::string err;
    const int32_t rc = guard(err, [&] {
        return qk::get_shape(family, lane_units, grid_per_cu, &err);
    });
    if (rc != QK_OK) set_msg(err);
    return rc;
}

int32_t qk_mc_autotune(int32_t family, int32_t* out_lane_units, int32_t* out_grid_per_cu,
                       double* out_ms) {
    std::string err;
    const int32_t rc = guard(err, [&] {
        return qk::autotune(family, out_lane_units, out_grid_per_cu, out_ms, &err);
    });
    if (rc != QK_OK) set_msg(err);
    return rc;
}

int32_t qk_mc_validation_stream(uint64_t seed, int64_t count, double* out) {
    if (!check_ptrs({QK_P(out)})) return QK_ERR_NULL_PTR;
    if (count < 0) {
        set_msg("bad stream length");
        return QK_ERR_BAD_SIZE;
    }
    qk::host_normal_stream(seed, count, out);
    return QK_OK;
}

}  // extern "C"

// ---- Longstaff-Schwartz (reference qk_api.cpp:986-1002, 875-880) ----------
extern "C" {

int32_t qk_mcm_longstaff_schwartz_price_batch_ex(const double* spot, const double* strike,
                                                 const double* t, const double* vol,
                                                 const double* r, const double* q,
                                                 const int32_t* option_type, const int32_t* paths,
                                                 const int32_t* steps, const uint64_t* seed,
                                                 int32_t basis, int32_t n, double* out_prices,
                                                 double* out_stderr) {
    if (!check_ptrs({QK_P(spot), QK_P(strike), QK_P(t), QK_P(vol), QK_P(r), QK_P(q),
                     QK_P(option_type), QK_P(paths), QK_P(steps), QK_P(seed), QK_P(out_prices),
                     QK_P(out_stderr)}))
        return QK_ERR_NULL_PTR;
    if (!check_n(n)) return QK_ERR_BAD_SIZE;
    if (basis != QK_LSMC_BASIS_REFERENCE && basis != QK_LSMC_BASIS_CUBIC) {
        char buf[256];
        std::snprintf(buf, sizeof buf, "unknown LSMC basis: %d", basis);
        set_msg(buf);
        return QK_ERR_INVALID_INPUT;
    }
    std::string err;
    const int32_t rc = guard(err, [&] {
        return qk::run_lsmc_batch(spot, strike, t, vol, r, q, option_type, paths, steps, seed,
                                  basis, n, out_prices, out_stderr, &err);
    });
    if (rc != QK_OK) set_msg(err);
    return rc;
}

int32_t qk_mcm_longstaff_schwartz_price_batch(const double* spot, const double* strike,
                                              const double* t, const double* vol,
                                              const double* r, const double* q,
                                              const int32_t* option_type, const int32_t* paths,
                                              const int32_t* steps, const uint64_t* seed,
                                              int32_t n, double* out_prices) {
    if (!check_ptrs({QK_P(spot), QK_P(strike), QK_P(t), QK_P(vol), QK_P(r), QK_P(q),
                     QK_P(option_type), QK_P(paths), QK_P(steps), QK_P(seed), QK_P(out_prices)}))
        return QK_ERR_NULL_PTR;
    if (!check_n(n)) return QK_ERR_BAD_SIZE;
    std::string err;
    const int32_t rc = guard(err, [&] {
        return qk::run_lsmc_batch(spot, strike, t, vol, r, q, option_type, paths, steps, seed,
                                  QK_LSMC_BASIS_REFERENCE, n, out_prices, nullptr, &err);
    });
    if (rc != QK_OK) set_msg(err);
    return rc;
}

double qk_mcm_longstaff_schwartz_price(double spot, double strike, double t, double vol,
                                       double r, double q, int32_t option_type, int32_t paths,
                                       int32_t steps, uint64_t seed) {
    double out = 0.0;
    std::string err;
    const int32_t rc = guard(err, [&] {
        return qk::run_lsmc_batch(&spot, &strike, &t, &vol, &r, &q, &option_type, &paths, &steps,
                                  &seed, QK_LSMC_BASIS_REFERENCE, 1, &out, nullptr, &err);
    });
    if (rc != QK_OK) set_msg(err);
    return scalar_of(rc, out);
}

}  // extern "C"

// ---- Heston surrogate (ABI 2.13) ------------------------------------------
extern "C" {

#define QK_SUR_RC(expr)                                         \
    do {                                                        \
        std::string err_;                                       \
        const int32_t rc_ = guard(err_, [&] { return (expr); }); \
        if (rc_ != QK_OK) set_msg(err_);                        \
        return rc_;                                             \
    } while (0)

int32_t qk_sur_param_count(void) { return qk::mlp::kParams; }

int32_t qk_sur_init(uint64_t seed) { QK_SUR_RC(qk::sur_init(seed, &err_)); }

int32_t qk_sur_train(const double* features, const double* targets, int32_t n, int32_t epochs,
                     int32_t batch, double lr, double lr_final, uint64_t seed,
                     double* out_epoch_loss) {
    if (!check_ptrs({QK_P(features), QK_P(targets)})) return QK_ERR_NULL_PTR;
    if (!check_n(n)) return QK_ERR_BAD_SIZE;
    QK_SUR_RC(qk::sur_train(features, targets, n, epochs, batch, lr, lr_final, seed,
                            out_epoch_loss, &err_));
}

int32_t qk_sur_predict(const double* features, int32_t n, double* out) {
    if (!check_ptrs({QK_P(features), QK_P(out)})) return QK_ERR_NULL_PTR;
    if (!check_n(n)) return QK_ERR_BAD_SIZE;
    QK_SUR_RC(qk::sur_predict(features, n, out, &err_));
}

int32_t qk_sur_gradients(const double* features, const double* targets, int32_t n,
                         float* out_grads, double* out_loss) {
    if (!check_ptrs({QK_P(features), QK_P(targets), QK_P(out_grads), QK_P(out_loss)}))
        return QK_ERR_NULL_PTR;
    if (!check_n(n)) return QK_ERR_BAD_SIZE;
    QK_SUR_RC(qk::sur_gradients(features, targets, n, out_grads, out_loss, &err_));
}

int32_t qk_sur_get_params(float* out) {
    if (!check_ptrs({QK_P(out)})) return QK_ERR_NULL_PTR;
    QK_SUR_RC(qk::sur_get_params(out, &err_));
}

int32_t qk_sur_set_params(const float* params) {
    if (!check_ptrs({QK_P(params)})) return QK_ERR_NULL_PTR;
    QK_SUR_RC(qk::sur_set_params(params, &err_));
}

int32_t qk_sur_save(const char* path) {
    if (!check_ptrs({QK_P(path)})) return QK_ERR_NULL_PTR;
    QK_SUR_RC(qk::sur_save(path, &err_));
}

int32_t qk_sur_load(const char* path) {
    if (!check_ptrs({QK_P(path)})) return QK_ERR_NULL_PTR;
    QK_SUR_RC(qk::sur_load(path, &err_));
}

int32_t qk_sur_configure(int32_t spot_every, int32_t spot_paths, double spot_k_se) {
    if (spot_every < 0 || spot_paths < 2 || !(spot_k_se > 0.0)) {
        set_msg("surrogate spot check needs spot_every >= 0, spot_paths >= 2, spot_k_se > 0");
        return QK_ERR_INVALID_INPUT;
    }
    qk::sur_configure(spot_every, spot_paths, spot_k_se);
    return QK_OK;
}

static int32_t sur_price_common(bool device, const double* spot, const double* strike,
                                const double* t, const double* r, const double* q,
                                const double* v0, const double* kappa, const double* theta,
                                const double* sigma, const double* rho,
                                const int32_t* option_type, const int32_t* paths,
                                const int32_t* steps, const uint64_t* seed, int32_t n,
                                double* out_prices, int32_t* out_source) {
    if (!check_ptrs({QK_P(spot), QK_P(strike), QK_P(t), QK_P(r), QK_P(q), QK_P(v0), QK_P(kappa),
                     QK_P(theta), QK_P(sigma), QK_P(rho), QK_P(option_type), QK_P(paths),
                     QK_P(steps), QK_P(seed), QK_P(out_prices), QK_P(out_source)}))
        return QK_ERR_NULL_PTR;
    if (!check_n(n)) return QK_ERR_BAD_SIZE;
    if (qk::get_rng_mode() != QK_RNG_PHILOX) {
        set_msg("surrogate pricing uses the Philox MC fallback: set QK_RNG_PHILOX");
        return QK_ERR_INVALID_INPUT;
    }
    qk::BatchIn bi{};
    bi.spot = spot;
    bi.strike = strike;
    bi.t = t;
    bi.r = r;
    bi.q = q;
    bi.v0 = v0;
    bi.kappa = kappa;
    bi.theta = theta;
    bi.sigma = sigma;
    bi.rho = rho;
    bi.ot = option_type;
    bi.paths = paths;
    bi.steps = steps;
    bi.seed = seed;
    QK_SUR_RC(qk::sur_heston_price(bi, n, device, out_prices, out_source, &err_));
}

int32_t qk_sur_heston_price_batch(const double* spot, const double* strike, const double* t,
                                  const double* r, const double* q, const double* v0,
                                  const double* kappa, const double* theta, const double* sigma,
                                  const double* rho, const int32_t* option_type,
                                  const int32_t* paths, const int32_t* steps,
                                  const uint64_t* seed, int32_t n, double* out_prices,
                                  int32_t* out_source) {
    return sur_price_common(false, spot, strike, t, r, q, v0, kappa, theta, sigma, rho,
                            option_type, paths, steps, seed, n, out_prices, out_source);
}

int32_t qk_sur_heston_price_batch_device(const double* spot, const double* strike,
                                         const double* t, const double* r, const double* q,
                                         const double* v0, const double* kappa,
                                         const double* theta, const double* sigma,
                                         const double* rho, const int32_t* option_type,
                                         const int32_t* paths, const int32_t* steps,
                                         const uint64_t* seed, int32_t n, double* out_prices,
                                         int32_t* out_source) {
    return sur_price_common(true, spot, strike, t, r, q, v0, kappa, theta, sigma, rho,
                            option_type, paths, steps, seed, n, out_prices, out_source);
}

int32_t qk_sur_set_target(double target_rms_over_spot) {
    QK_SUR_RC(qk::sur_set_target(target_rms_over_spot, &err_));
}

int32_t qk_sur_last_judge(double* out_rows, int32_t max_rows) {
    if (!check_ptrs({QK_P(out_rows)})) return QK_ERR_NULL_PTR;
    std::string err;
    const int32_t rc = guard(err, [&] { return qk::sur_last_judge(out_rows, max_rows, &err); });
    if (rc < 0) set_msg(err);
    return rc;
}

int32_t qk_sur_last_timing(double* ms_network, double* ms_monte_carlo) {
    if (!qk::sur_last_timing(ms_network, ms_monte_carlo)) {
        set_msg("no surrogate pricing call has completed yet");
        return QK_ERR_RUNTIME_INIT;
    }
    return QK_OK;
}

}  // extern "C"
