/*
 * qk_api.h - flat C ABI of the MI355X Monte Carlo batch-pricing library.
 *
 * Every declaration marked "reference" has the exact name, argument order,
 * types and return convention of the reference's declaration at the cited
 * line of cpp/include/quantkernel/qk_api.h, so the reference's callers
 * (python/quantkernel/_loader.py:152-199, engine.py:163-194, C tests) bind to
 * this library unchanged.  Semantics kept from the reference
 * (cpp/src/qk_api.cpp:19-117, 837-1227):
 *   - batch calls null-check every pointer in declaration order first
 *     ("null pointer: <param>", QK_ERR_NULL_PTR), then n <= 0
 *     ("bad batch size: <n>", QK_ERR_BAD_SIZE);
 *   - an invalid element writes a quiet NaN and the call still returns QK_OK;
 *   - the last-error string is thread-local, 256 bytes, sticky on success and
 *     cleared only by qk_clear_last_error.
 * What changes: the paths are simulated on a gfx950 GPU.  A missing or
 * unusable device makes a batch call return QK_ERR_RUNTIME_INIT (there is no
 * CPU fallback) and a scalar call return NaN with the error string set.
 *
 * All pointers of the non-"_device" entry points are caller-owned host memory
 * of length n.  Calls are synchronous and may come from several host threads
 * (device work is serialised internally).
 */
#ifndef QK_API_H
#define QK_API_H

#include "qk_abi.h"

#define QK_EXPORT __attribute__((visibility("default")))

#ifdef __cplusplus
extern "C" {
#endif

/* ---- ABI negotiation and errors (reference qk_api.h:16-21) ------------- */
QK_EXPORT void        qk_abi_version(int32_t* major, int32_t* minor);
QK_EXPORT int32_t     qk_plugin_get_api(int32_t host_abi_major, int32_t host_abi_minor,
                                        const QKPluginAPI** out_api);
QK_EXPORT const char* qk_get_last_error(void);
QK_EXPORT void        qk_clear_last_error(void);

/* ---- GBM terminal, antithetic pairs (reference qk_api.h:259-268) -------- */
QK_EXPORT double  qk_mcm_standard_monte_carlo_price(double spot, double strike, double t, double vol,
                                                    double r, double q, int32_t option_type,
                                                    int32_t paths, uint64_t seed);
QK_EXPORT int32_t qk_mcm_standard_monte_carlo_price_batch(const double* spot, const double* strike,
                                                          const double* t, const double* vol,
                                                          const double* r, const double* q,
                                                          const int32_t* option_type,
                                                          const int32_t* paths,
                                                          const uint64_t* seed,
                                                          int32_t n, double* out_prices);

/* ---- antithetic variates (reference qk_api.h:294-296, 347-352) ---------- */
QK_EXPORT double  qk_mcm_antithetic_variates_price(double spot, double strike, double t, double vol,
                                                   double r, double q, int32_t option_type,
                                                   int32_t paths, uint64_t seed);
QK_EXPORT int32_t qk_mcm_antithetic_variates_price_batch(const double* spot, const double* strike,
                                                         const double* t, const double* vol,
                                                         const double* r, const double* q,
                                                         const int32_t* option_type, const int32_t* paths,
                                                         const uint64_t* seed,
                                                         int32_t n, double* out_prices);

/* ---- GBM Euler-Maruyama (reference qk_api.h:269-271, 300-305) ----------- */
QK_EXPORT double  qk_mcm_euler_maruyama_price(double spot, double strike, double t, double vol,
                                              double r, double q, int32_t option_type,
                                              int32_t paths, int32_t steps, uint64_t seed);
QK_EXPORT int32_t qk_mcm_euler_maruyama_price_batch(const double* spot, const double* strike,
                                                    const double* t, const double* vol,
                                                    const double* r, const double* q,
                                                    const int32_t* option_type, const int32_t* paths,
                                                    const int32_t* steps, const uint64_t* seed,
                                                    int32_t n, double* out_prices);

/* ---- GBM Milstein (reference qk_api.h:272-274, 306-311) ----------------- */
QK_EXPORT double  qk_mcm_milstein_price(double spot, double strike, double t, double vol,
                                        double r, double q, int32_t option_type,
                                        int32_t paths, int32_t steps, uint64_t seed);
QK_EXPORT int32_t qk_mcm_milstein_price_batch(const double* spot, const double* strike,
                                              const double* t, const double* vol,
                                              const double* r, const double* q,
                                              const int32_t* option_type, const int32_t* paths,
                                              const int32_t* steps, const uint64_t* seed,
                                              int32_t n, double* out_prices);

/* ---- Longstaff-Schwartz American (reference qk_api.h:275-277, 312-317) -- */
QK_EXPORT double  qk_mcm_longstaff_schwartz_price(double spot, double strike, double t, double vol,
                                                  double r, double q, int32_t option_type,
                                                  int32_t paths, int32_t steps, uint64_t seed);
QK_EXPORT int32_t qk_mcm_longstaff_schwartz_price_batch(const double* spot, const double* strike,
                                                        const double* t, const double* vol,
                                                        const double* r, const double* q,
                                                        const int32_t* option_type, const int32_t* paths,
                                                        const int32_t* steps, const uint64_t* seed,
                                                        int32_t n, double* out_prices);

/* ---- Heston full-truncation Euler (reference qk_api.h:359-374) ---------- */
QK_EXPORT double  qk_mcm_heston_monte_carlo_price(double spot, double strike, double t,
                                                  double r, double q,
                                                  double v0, double kappa, double theta,
                                                  double sigma, double rho,
                                                  int32_t option_type,
                                                  int32_t paths, int32_t steps, uint64_t seed);
QK_EXPORT int32_t qk_mcm_heston_monte_carlo_price_batch(const double* spot, const double* strike,
                                                        const double* t,
                                                        const double* r, const double* q,
                                                        const double* v0, const double* kappa,
                                                        const double* theta, const double* sigma,
                                                        const double* rho,
                                                        const int32_t* option_type,
                                                        const int32_t* paths, const int32_t* steps,
                                                        const uint64_t* seed,
                                                        int32_t n, double* out_prices);

/* ---- constant local vol, Euler (reference qk_api.h:393-403) ------------- */
QK_EXPORT double  qk_mcm_local_vol_monte_carlo_price(double spot, double strike, double t,
                                                     double vol, double r, double q,
                                                     int32_t option_type,
                                                     int32_t paths, int32_t steps, uint64_t seed);
QK_EXPORT int32_t qk_mcm_local_vol_monte_carlo_price_batch(const double* spot, const double* strike,
                                                           const double* t, const double* vol,
                                                           const double* r, const double* q,
                                                           const int32_t* option_type,
                                                           const int32_t* paths, const int32_t* steps,
                                                           const uint64_t* seed,
                                                           int32_t n, double* out_prices);

/* ---- control variates, stratified sampling, quasi-MC, multilevel MC
 *      (reference qk_api.h:278-299, 318-358) ------------------------------- */
QK_EXPORT double  qk_mcm_quasi_monte_carlo_sobol_price(double spot, double strike, double t, double vol,
                                                       double r, double q, int32_t option_type,
                                                       int32_t paths);
QK_EXPORT double  qk_mcm_quasi_monte_carlo_halton_price(double spot, double strike, double t, double vol,
                                                        double r, double q, int32_t option_type,
                                                        int32_t paths);
QK_EXPORT double  qk_mcm_multilevel_monte_carlo_price(double spot, double strike, double t, double vol,
                                                      double r, double q, int32_t option_type,
                                                      int32_t base_paths, int32_t levels, int32_t base_steps,
                                                      uint64_t seed);
QK_EXPORT double  qk_mcm_control_variates_price(double spot, double strike, double t, double vol,
                                                double r, double q, int32_t option_type,
                                                int32_t paths, uint64_t seed);
QK_EXPORT double  qk_mcm_stratified_sampling_price(double spot, double strike, double t, double vol,
                                                   double r, double q, int32_t option_type,
                                                   int32_t paths, uint64_t seed);
QK_EXPORT int32_t qk_mcm_quasi_monte_carlo_sobol_price_batch(const double* spot, const double* strike,
                                                             const double* t, const double* vol,
                                                             const double* r, const double* q,
                                                             const int32_t* option_type, const int32_t* paths,
                                                             int32_t n, double* out_prices);
QK_EXPORT int32_t qk_mcm_quasi_monte_carlo_halton_price_batch(const double* spot, const double* strike,
                                                              const double* t, const double* vol,
                                                              const double* r, const double* q,
                                                              const int32_t* option_type, const int32_t* paths,
                                                              int32_t n, double* out_prices);
QK_EXPORT int32_t qk_mcm_multilevel_monte_carlo_price_batch(const double* spot, const double* strike,
                                                            const double* t, const double* vol,
                                                            const double* r, const double* q,
                                                            const int32_t* option_type, const int32_t* base_paths,
                                                            const int32_t* levels, const int32_t* base_steps,
                                                            const uint64_t* seed,
                                                            int32_t n, double* out_prices);
QK_EXPORT int32_t qk_mcm_control_variates_price_batch(const double* spot, const double* strike,
                                                      const double* t, const double* vol,
                                                      const double* r, const double* q,
                                                      const int32_t* option_type, const int32_t* paths,
                                                      const uint64_t* seed,
                                                      int32_t n, double* out_prices);
QK_EXPORT int32_t qk_mcm_stratified_sampling_price_batch(const double* spot, const double* strike,
                                                         const double* t, const double* vol,
                                                         const double* r, const double* q,
                                                         const int32_t* option_type, const int32_t* paths,
                                                         const uint64_t* seed,
                                                         int32_t n, double* out_prices);

/* ---- importance sampling (reference qk_api.h:288-290, 335-340) ---------- */
QK_EXPORT double  qk_mcm_importance_sampling_price(double spot, double strike, double t, double vol,
                                                   double r, double q, int32_t option_type,
                                                   int32_t paths, double shift, uint64_t seed);
QK_EXPORT int32_t qk_mcm_importance_sampling_price_batch(const double* spot, const double* strike,
                                                         const double* t, const double* vol,
                                                         const double* r, const double* q,
                                                         const int32_t* option_type, const int32_t* paths,
                                                         const double* shift, const uint64_t* seed,
                                                         int32_t n, double* out_prices);

/* ---- Heston likelihood-ratio delta (reference qk_api.h:375-392) --------- */
QK_EXPORT double  qk_mcm_heston_lr_delta(double spot, double strike, double t,
                                         double r, double q,
                                         double v0, double kappa, double theta,
                                         double sigma, double rho,
                                         int32_t option_type,
                                         int32_t paths, int32_t steps, uint64_t seed,
                                         double weight_clip);
QK_EXPORT int32_t qk_mcm_heston_lr_delta_batch(const double* spot, const double* strike,
                                               const double* t,
                                               const double* r, const double* q,
                                               const double* v0, const double* kappa,
                                               const double* theta, const double* sigma,
                                               const double* rho,
                                               const int32_t* option_type,
                                               const int32_t* paths, const int32_t* steps,
                                               const uint64_t* seed,
                                               const double* weight_clip,
                                               int32_t n, double* out_prices);

/* ---- pathwise and likelihood-ratio delta (reference qk_api.h:609-614,
 *      618-630) ------------------------------------------------------------ */
QK_EXPORT double  qk_agm_pathwise_derivative_delta(double spot, double strike, double t, double vol,
                                                   double r, double q, int32_t option_type,
                                                   int32_t paths, uint64_t seed);
QK_EXPORT double  qk_agm_likelihood_ratio_delta(double spot, double strike, double t, double vol,
                                                double r, double q, int32_t option_type,
                                                int32_t paths, uint64_t seed, double weight_clip);
QK_EXPORT int32_t qk_agm_pathwise_derivative_delta_batch(const double* spot, const double* strike,
                                                         const double* t, const double* vol,
                                                         const double* r, const double* q,
                                                         const int32_t* option_type,
                                                         const int32_t* paths, const uint64_t* seed,
                                                         int32_t n, double* out_prices);
QK_EXPORT int32_t qk_agm_likelihood_ratio_delta_batch(const double* spot, const double* strike,
                                                      const double* t, const double* vol,
                                                      const double* r, const double* q,
                                                      const int32_t* option_type,
                                                      const int32_t* paths, const uint64_t* seed,
                                                      const double* weight_clip,
                                                      int32_t n, double* out_prices);

/* ---- Heston characteristic-function price (reference qk_api.h:53-55,
 *      71-76): the closed-form check of the Heston MC path, on the GPU ---- */
QK_EXPORT double  qk_cf_heston_price_cf(double spot, double strike, double t, double r, double q,
                                        double v0, double kappa, double theta, double sigma, double rho,
                                        int32_t option_type, int32_t integration_steps, double integration_limit);
QK_EXPORT int32_t qk_cf_heston_price_cf_batch(const double* spot, const double* strike,
                                              const double* t, const double* r, const double* q,
                                              const double* v0, const double* kappa, const double* theta,
                                              const double* sigma, const double* rho,
                                              const int32_t* option_type, const int32_t* integration_steps,
                                              const double* integration_limit,
                                              int32_t n, double* out_prices);

/* ======================================================================== *
 * ABI 2.13 additions.  No reference counterpart: new names only.
 * ======================================================================== */

/* Process-wide normal-stream mode (QK_RNG_*) used by every MC entry point.
 * Returns QK_ERR_INVALID_INPUT for an unknown mode.  Default QK_RNG_PHILOX. */
QK_EXPORT int32_t qk_mc_set_rng_mode(int32_t mode);
QK_EXPORT int32_t qk_mc_get_rng_mode(void);

/* Batch entry points that also return the standard error of each price
 * (discounted sample standard deviation / sqrt(paths); for antithetic
 * estimators the pair average is the sample unit).  Same argument order as the
 * reference batch plus out_stderr; out_stderr is 0 for closed-form branches
 * and NaN for invalid elements. */
QK_EXPORT int32_t qk_mcm_standard_monte_carlo_price_batch_se(const double* spot, const double* strike,
        const double* t, const double* vol, const double* r, const double* q,
        const int32_t* option_type, const int32_t* paths, const uint64_t* seed,
        int32_t n, double* out_prices, double* out_stderr);
QK_EXPORT int32_t qk_mcm_euler_maruyama_price_batch_se(const double* spot, const double* strike,
        const double* t, const double* vol, const double* r, const double* q,
        const int32_t* option_type, const int32_t* paths, const int32_t* steps, const uint64_t* seed,
        int32_t n, double* out_prices, double* out_stderr);
QK_EXPORT int32_t qk_mcm_milstein_price_batch_se(const double* spot, const double* strike,
        const double* t, const double* vol, const double* r, const double* q,
        const int32_t* option_type, const int32_t* paths, const int32_t* steps, const uint64_t* seed,
        int32_t n, double* out_prices, double* out_stderr);
QK_EXPORT int32_t qk_mcm_local_vol_monte_carlo_price_batch_se(const double* spot, const double* strike,
        const double* t, const double* vol, const double* r, const double* q,
        const int32_t* option_type, const int32_t* paths, const int32_t* steps, const uint64_t* seed,
        int32_t n, double* out_prices, double* out_stderr);
QK_EXPORT int32_t qk_mcm_heston_monte_carlo_price_batch_se(const double* spot, const double* strike,
        const double* t, const double* r, const double* q,
        const double* v0, const double* kappa, const double* theta, const double* sigma,
        const double* rho, const int32_t* option_type, const int32_t* paths, const int32_t* steps,
        const uint64_t* seed, int32_t n, double* out_prices, double* out_stderr);

/* Standard errors of importance sampling (per path), of the pathwise delta
 * (per path), of the likelihood-ratio delta (per antithetic pair) and of the
 * Heston likelihood-ratio delta (of sum / paths, with the antithetic pair as
 * the sample unit).  Same argument order as the batch entry points above plus
 * out_stderr. */
QK_EXPORT int32_t qk_mcm_importance_sampling_price_batch_se(const double* spot, const double* strike,
        const double* t, const double* vol, const double* r, const double* q,
        const int32_t* option_type, const int32_t* paths, const double* shift, const uint64_t* seed,
        int32_t n, double* out_prices, double* out_stderr);
QK_EXPORT int32_t qk_mcm_heston_lr_delta_batch_se(const double* spot, const double* strike,
        const double* t, const double* r, const double* q,
        const double* v0, const double* kappa, const double* theta, const double* sigma,
        const double* rho, const int32_t* option_type, const int32_t* paths, const int32_t* steps,
        const uint64_t* seed, const double* weight_clip, int32_t n, double* out_prices,
        double* out_stderr);
QK_EXPORT int32_t qk_agm_pathwise_derivative_delta_batch_se(const double* spot, const double* strike,
        const double* t, const double* vol, const double* r, const double* q,
        const int32_t* option_type, const int32_t* paths, const uint64_t* seed,
        int32_t n, double* out_prices, double* out_stderr);
QK_EXPORT int32_t qk_agm_likelihood_ratio_delta_batch_se(const double* spot, const double* strike,
        const double* t, const double* vol, const double* r, const double* q,
        const int32_t* option_type, const int32_t* paths, const uint64_t* seed,
        const double* weight_clip, int32_t n, double* out_prices, double* out_stderr);

/* Standard errors of control variates (residual of the control-variate
 * fit, per path), stratified sampling (per path, the iid formula, which
 * overstates a stratified estimator's error), multilevel MC (the levels'
 * errors in quadrature) and quasi-MC (NaN: a point set has no sampling
 * error). */
QK_EXPORT int32_t qk_mcm_control_variates_price_batch_se(const double* spot, const double* strike,
        const double* t, const double* vol, const double* r, const double* q,
        const int32_t* option_type, const int32_t* paths, const uint64_t* seed,
        int32_t n, double* out_prices, double* out_stderr);
QK_EXPORT int32_t qk_mcm_stratified_sampling_price_batch_se(const double* spot, const double* strike,
        const double* t, const double* vol, const double* r, const double* q,
        const int32_t* option_type, const int32_t* paths, const uint64_t* seed,
        int32_t n, double* out_prices, double* out_stderr);
QK_EXPORT int32_t qk_mcm_quasi_monte_carlo_sobol_price_batch_se(const double* spot, const double* strike,
        const double* t, const double* vol, const double* r, const double* q,
        const int32_t* option_type, const int32_t* paths, int32_t n, double* out_prices,
        double* out_stderr);
QK_EXPORT int32_t qk_mcm_quasi_monte_carlo_halton_price_batch_se(const double* spot, const double* strike,
        const double* t, const double* vol, const double* r, const double* q,
        const int32_t* option_type, const int32_t* paths, int32_t n, double* out_prices,
        double* out_stderr);
QK_EXPORT int32_t qk_mcm_multilevel_monte_carlo_price_batch_se(const double* spot, const double* strike,
        const double* t, const double* vol, const double* r, const double* q,
        const int32_t* option_type, const int32_t* base_paths, const int32_t* levels,
        const int32_t* base_steps, const uint64_t* seed, int32_t n, double* out_prices,
        double* out_stderr);

/* Mixed-model uneven batch: one persistent launch over options of different
 * models (QK_MODEL_*), path counts and step counts.  vol is used by the GBM
 * and local-vol models, v0..rho by Heston; unused arrays may hold anything
 * but must not be null.  Per element the result equals the single-model batch
 * entry point's result for the same inputs. */
QK_EXPORT int32_t qk_mcm_mixed_price_batch(const int32_t* model,
        const double* spot, const double* strike, const double* t, const double* vol,
        const double* r, const double* q,
        const double* v0, const double* kappa, const double* theta, const double* sigma,
        const double* rho, const int32_t* option_type, const int32_t* paths, const int32_t* steps,
        const uint64_t* seed, int32_t n, double* out_prices, double* out_stderr);

/* Same as qk_mcm_mixed_price_batch, but every pointer is device memory of
 * the current HIP device (inputs already resident in HBM, outputs left there).
 * QK_RNG_PHILOX only; returns QK_ERR_INVALID_INPUT in another RNG mode.  The
 * library works on its own stream, which does not wait for the caller's
 * streams: inputs must be complete when the call is made.  The call returns
 * when the outputs are written. */
QK_EXPORT int32_t qk_mcm_mixed_price_batch_device(const int32_t* model,
        const double* spot, const double* strike, const double* t, const double* vol,
        const double* r, const double* q,
        const double* v0, const double* kappa, const double* theta, const double* sigma,
        const double* rho, const int32_t* option_type, const int32_t* paths, const int32_t* steps,
        const uint64_t* seed, int32_t n, double* out_prices, double* out_stderr);

/* Allocation-free variant of qk_mcm_mixed_price_batch_device, with the same
 * arguments and results.  It has no counterpart in the reference: it is for
 * callers that cannot allocate device memory during the call (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  It
 * calls no hipMalloc, hipFree, hipHostMalloc or graph instantiation: its work
 * buffers live in one fixed device array of the library (about 44.5 MiB,
 * held by every device the library runs on until the process ends).  Limits:
 * at most 65535 options (larger n returns QK_ERR_BAD_SIZE before any device
 * work) and at most 2^19 path tiles of the current tile shape (a tile is 256
 * lanes x the stepwise or terminal lane_units of qk_mc_set_shape; a batch
 * needing more runs no tile, its outputs are NaN and the call returns
 * QK_ERR_INVALID_INPUT).  All calls in a process share that array; the
 * library's device lock serialises them, and each call finishes its device
 * work before it returns.  The device context (stream, events, the array's
 * device address) is set up by the library's first call, which may be this
 * one. */
QK_EXPORT int32_t qk_mcm_mixed_price_batch_device_noalloc(const int32_t* model,
        const double* spot, const double* strike, const double* t, const double* vol,
        const double* r, const double* q,
        const double* v0, const double* kappa, const double* theta, const double* sigma,
        const double* rho, const int32_t* option_type, const int32_t* paths, const int32_t* steps,
        const uint64_t* seed, int32_t n, double* out_prices, double* out_stderr);

/* Standard MC (antithetic terminal estimator) on device-resident inputs and
 * outputs, with standard errors; QK_RNG_PHILOX only.  Equal to
 * qk_mcm_standard_monte_carlo_price_batch_se for the same inputs.
 * Device-resident calls run without a host synchronisation inside the batch,
 * so each distinct (size, buffers) shape is captured once as a launch graph
 * and replayed (see qk_mc_graphs). */
QK_EXPORT int32_t qk_mcm_standard_monte_carlo_price_batch_device(const double* spot,
        const double* strike, const double* t, const double* vol, const double* r, const double* q,
        const int32_t* option_type, const int32_t* paths, const uint64_t* seed, int32_t n,
        double* out_prices, double* out_stderr);

/* Heston characteristic-function prices on device-resident inputs and
 * output (pointers of the current HIP device); equal to
 * qk_cf_heston_price_cf_batch for the same inputs. */
QK_EXPORT int32_t qk_cf_heston_price_cf_batch_device(const double* spot, const double* strike,
        const double* t, const double* r, const double* q, const double* v0, const double* kappa,
        const double* theta, const double* sigma, const double* rho, const int32_t* option_type,
        const int32_t* integration_steps, const double* integration_limit, int32_t n,
        double* out_prices);

/* Longstaff-Schwartz with a selectable regression basis (QK_LSMC_BASIS_*)
 * and standard-error output.  basis QK_LSMC_BASIS_REFERENCE reproduces
 * qk_mcm_longstaff_schwartz_price_batch. */
QK_EXPORT int32_t qk_mcm_longstaff_schwartz_price_batch_ex(const double* spot, const double* strike,
        const double* t, const double* vol, const double* r, const double* q,
        const int32_t* option_type, const int32_t* paths, const int32_t* steps,
        const uint64_t* seed, int32_t basis, int32_t n, double* out_prices, double* out_stderr);

/* Device timings of the last MC call made by this process, from HIP events on
 * the library's stream: total device time of the call and the time of its
 * path-simulation kernel alone, in milliseconds.  A host-memory QK_RNG_PHILOX
 * call that ran as a launch graph (qk_mc_graphs) is timed as a whole (copies
 * included) and reports ms_simulate = NaN.  Returns QK_ERR_RUNTIME_INIT if no
 * call has completed yet. */
QK_EXPORT int32_t qk_mc_last_timing(double* ms_total, double* ms_simulate);

/* The normal stream the QK_RNG_MT19937_* modes feed to the GPU for one
 * option seed: the first `count` draws of std::mt19937_64(seed) through
 * std::normal_distribution<double> (reference mc_engine.h:15-23).  Host
 * only; needs no device. */
QK_EXPORT int32_t qk_mc_validation_stream(uint64_t seed, int64_t count, double* out);

/* Self-check of the device normal generator: the Philox4x32-10 words and
 * the FP32 Box-Muller normals of blocks (c0 + i, c1, 0, c3), i < count, keyed
 * by seed (4 words and 4 normals per block). */
QK_EXPORT int32_t qk_mc_philox_probe(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c3,
                                     int32_t count, uint32_t* out_words, float* out_normals);

/* Self-check of the production normal layout: the eight FP32 normals of
 * blocks (c0 + i, c1, 0, c3), i < count, keyed by seed, in the four-pair
 * layout the path and terminal kernels draw from (word s gives a 23-bit
 * radius and a 9-bit angle read from the device's 512-entry table;
 * normals 2s and 2s+1 are its cosine and sine components).  No reference
 * counterpart (an addition of this ABI, like qk_mc_philox_probe). */
QK_EXPORT int32_t qk_mc_philox_probe8(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c3,
                                      int32_t count, float* out_normals);

/* Launch graphs of the QK_RNG_PHILOX paths (on by default): each distinct
 * batch shape of the host-memory entry points is captured once (H2D copy,
 * prepare, scan, simulate, finalize, D2H copy) and replayed, and likewise each
 * (size, buffers) shape of the device-resident entry points.  enable: 1 on, 0 off, -1 unchanged.
 * out_cached: graphs held (-1 if the runtime cannot capture); out_launches:
 * graph launches so far.  Either output may be null.  QK_NO_GRAPH=1 turns
 * graphs off for the process. */
QK_EXPORT int32_t qk_mc_graphs(int32_t enable, int32_t* out_cached, int64_t* out_launches);

/* Tile shapes of the persistent simulation kernels.  family 1 = terminal
 * (lane_units = Philox groups of 4 antithetic pairs per lane), 2 = stepwise
 * (lane_units = paths per lane); grid_per_cu = persistent workgroups per CU.
 * Results are bitwise reproducible for a fixed shape; another shape changes
 * the summation order (differences ~1e-16 relative).  Defaults: 8 and 8. */
QK_EXPORT int32_t qk_mc_set_shape(int32_t family, int32_t lane_units, int32_t grid_per_cu);
QK_EXPORT int32_t qk_mc_get_shape(int32_t family, int32_t* lane_units, int32_t* grid_per_cu);
/* Autotuner: times the candidate shapes of a family (lane_units in
 * {2,4,8,16,32} x grid_per_cu in {4,8,16}) on a representative synthetic batch
 * (C2-like uneven stepwise mix, or 4096 x 100k terminal), applies the fastest
 * and appends it to $QK_AUTOTUNE_CACHE (default ~/.cache/
 * quantkernel_mi355x_autotune.txt); with QK_AUTOTUNE=1 the cached shapes are
 * applied when the device is opened. */
QK_EXPORT int32_t qk_mc_autotune(int32_t family, int32_t* out_lane_units, int32_t* out_grid_per_cu,
                                 double* out_ms);

/* Name of the device in use and its compute-unit count; QK_ERR_RUNTIME_INIT
 * if there is none. */
QK_EXPORT int32_t qk_device_info(char* name_buf, int32_t buf_len, int32_t* compute_units);

/* ---- Heston surrogate (ABI 2.13; no reference counterpart) -------------
 * An MLP (8 -> 4 x 256 GELU -> 1; BF16 MFMA, FP32 accumulate) maps
 * (log K/S, T, r, kappa, theta, sigma, rho, v0) to the price/S of a European
 * call under the Heston full-truncation Euler scheme with 256 steps.  Features
 * are passed raw (n x 8, row-major); the library maps its training box
 * log K/S in [ln .7, ln 1.3], T in [.05, 2], r in [0, .08], kappa in [.5, 5],
 * theta in [.01, .16], sigma in [.1, 1], rho in [-.9, -.1], v0 in [.01, .16]
 * to [-1, 1].  The model is a closed-form control (the Black-Scholes call
 * per unit spot at the expected average variance
 * theta + (v0 - theta)(1 - e^{-kappa T})/(kappa T)) plus the network:
 * qk_sur_train takes full prices/S and fits the network to the difference,
 * qk_sur_predict returns control + network, and qk_sur_gradients returns the
 * network's own MSE gradients against the given network targets.  Weights
 * live on the device; one model per process. */
QK_EXPORT int32_t qk_sur_param_count(void);
QK_EXPORT int32_t qk_sur_init(uint64_t seed);
QK_EXPORT int32_t qk_sur_train(const double* features, const double* targets, int32_t n,
                               int32_t epochs, int32_t batch, double lr, double lr_final,
                               uint64_t seed, double* out_epoch_loss);
QK_EXPORT int32_t qk_sur_predict(const double* features, int32_t n, double* out);
QK_EXPORT int32_t qk_sur_gradients(const double* features, const double* targets, int32_t n,
                                   float* out_grads, double* out_loss);
QK_EXPORT int32_t qk_sur_get_params(float* out);
QK_EXPORT int32_t qk_sur_set_params(const float* params);
QK_EXPORT int32_t qk_sur_save(const char* path);
QK_EXPORT int32_t qk_sur_load(const char* path);
/* Spot check: every spot_every-th query answered by the network is also
 * priced by plain MC with spot_paths paths; a difference above
 * spot_k_se * SE + target/2 * spot sends it to the full MC fallback.
 * spot_every = 0 disables the checks and the batch judge. */
QK_EXPORT int32_t qk_sur_configure(int32_t spot_every, int32_t spot_paths, double spot_k_se);
/* Accuracy target of the served answers, RMS |price - MC expectation| / spot
 * (default 1e-3; in (0, 1)).  Besides the per-query tolerance above, it sets
 * the batch judge: over the spot checks of each of 16 regions of the box
 * (halves in log K/S, T, sigma and v0), and over all checks, the mean of
 * ((network - MC)/S)^2 - (SE/S)^2 estimates the network's mean squared
 * error; a region (or the whole batch) with at least 32 checks whose
 * estimate minus 3 of its standard errors exceeds target^2 has all its
 * network answers replaced by plain MC (source 4). */
QK_EXPORT int32_t qk_sur_set_target(double target_rms_over_spot);
/* The batch judge of the last pricing call with spot checks: up to max_rows
 * rows of 4 doubles {checks, implied MSE / spot^2, its standard error,
 * rejected (0/1)}, rows 0-15 the regions (bit 0: log K/S in the upper half
 * of the box, bit 1: T, bit 2: sigma, bit 3: v0), row 16 all checks.
 * Returns the number of rows written (17 at most), or a negative QK_ERR_*. */
QK_EXPORT int32_t qk_sur_last_judge(double* out_rows, int32_t max_rows);
/* Heston prices with the same arguments as qk_mcm_heston_monte_carlo_price_batch:
 * queries inside the box (q == 0, steps == 256) are answered by the network
 * (calls clamped to their no-arbitrage bounds, puts by put-call parity), the
 * others by plain MC with their own paths, steps and seed; out_source[i] is
 * 0 network, 1 MC (outside the box), 2 MC (spot check failed),
 * 3 invalid / t <= eps (NaN / intrinsic), 4 MC (region rejected by the
 * batch judge).  Every MC answer equals qk_mcm_heston_monte_carlo_price_batch's
 * for the same query. */
QK_EXPORT int32_t qk_sur_heston_price_batch(const double* spot, const double* strike,
        const double* t, const double* r, const double* q, const double* v0,
        const double* kappa, const double* theta, const double* sigma, const double* rho,
        const int32_t* option_type, const int32_t* paths, const int32_t* steps,
        const uint64_t* seed, int32_t n, double* out_prices, int32_t* out_source);
/* Same with device-resident inputs and outputs. */
QK_EXPORT int32_t qk_sur_heston_price_batch_device(const double* spot, const double* strike,
        const double* t, const double* r, const double* q, const double* v0,
        const double* kappa, const double* theta, const double* sigma, const double* rho,
        const int32_t* option_type, const int32_t* paths, const int32_t* steps,
        const uint64_t* seed, int32_t n, double* out_prices, int32_t* out_source);
/* Device time of the last surrogate call: network phase and MC phase (ms). */
QK_EXPORT int32_t qk_sur_last_timing(double* ms_network, double* ms_monte_carlo);

#ifdef __cplusplus
}
#endif

#endif /* QK_API_H */
