// qk_cf.hpp - Heston characteristic-function pricer (qk_cf.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

namespace qk {

// Device pointers of one batch (length n each).
struct HestonCfIn {
    const double *spot, *strike, *t, *r, *q, *v0, *kappa, *theta, *sigma, *rho, *limit;
    const int32_t *ot, *steps;
};

// gl: 32 Gauss-Legendre nodes in (0, 1) ascending, then their 32 weights.
hipError_t launch_heston_cf(hipStream_t st, const HestonCfIn& in, int32_t n, const double* gl,
                            double* out, int32_t grid);

// Host entry: closed_form_semi_analytical/heston/heston.cpp:7-53 per element.
int32_t run_heston_cf_batch(const double* spot, const double* strike, const double* t,
                            const double* r, const double* q, const double* v0,
                            const double* kappa, const double* theta, const double* sigma,
                            const double* rho, const int32_t* ot, const int32_t* steps,
                            const double* limit, int32_t n, double* out, std::string* err);

// Same on device-resident inputs and output (pointers of the current device).
int32_t run_heston_cf_batch_device(const HestonCfIn& in, int32_t n, double* out_dev,
                                   std::string* err);

}  // namespace qk
