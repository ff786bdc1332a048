// qk_surrogate.hpp - host side of the Heston surrogate: training, inference,
// weights file, and pricing with Monte Carlo fallback.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "qk_kernels.hpp"

namespace qk {

constexpr int kSurSteps = 256;  // time steps of the MC labels the network was trained on

// Per-option source of a surrogate-priced result.
enum SurSource : int8_t {
    SUR_NETWORK = 0,         // answered by the network
    SUR_FALLBACK_BOX = 1,    // outside the training box: plain MC
    SUR_FALLBACK_CHECK = 2,  // its spot-check MC disagreed: plain MC
    SUR_CLOSED = 3,          // invalid (NaN) or t <= eps (intrinsic), as the MC entry point
    SUR_FALLBACK_JUDGE = 4,  // the batch judge rejected its region (or the batch): plain MC
};

// Batch judge: regions of the training box (halves in log K/S, T, sigma and
// v0), the fewest spot checks a region (or the batch) needs before it is
// judged, and the standard errors of margin.
constexpr int kSurRegions = 16;
constexpr int kSurJudgeMinChecks = 32;
constexpr double kSurJudgeSigmas = 3.0;

int32_t sur_init(uint64_t seed, std::string* err);
int32_t sur_train(const double* feat, const double* target, int32_t n, int32_t epochs,
                  int32_t batch, double lr, double lr_final, uint64_t seed, double* epoch_loss,
                  std::string* err);
int32_t sur_predict(const double* feat, int32_t n, double* out, std::string* err);
int32_t sur_gradients(const double* feat, const double* target, int32_t n, float* grads,
                      double* loss, std::string* err);
int32_t sur_get_params(float* out, std::string* err);
int32_t sur_set_params(const float* in, std::string* err);
int32_t sur_save(const char* path, std::string* err);
int32_t sur_load(const char* path, std::string* err);
void sur_configure(int32_t spot_every, int32_t spot_paths, double spot_k_se);
// Accuracy target of the served answers (RMS |price - MC expectation| / spot):
// the per-query spot check allows half of it on top of k SE, and the batch
// judge rejects a region whose implied RMS error exceeds it.
int32_t sur_set_target(double target, std::string* err);
// Judge statistics of the last pricing call: kSurRegions + 1 rows (the last
// over all checks) of {checks, implied MSE / S^2, its SE, rejected}.
int32_t sur_last_judge(double* rows, int32_t max_rows, std::string* err);
// Pricing with fallback; inputs and outputs in host memory (device == false)
// or device memory (device == true).
int32_t sur_heston_price(const BatchIn& in, int32_t n, bool device, double* out_price,
                         int32_t* out_source, std::string* err);
bool sur_last_timing(double* ms_network, double* ms_mc);

// kernels (qk_sur_kernels.hip)
int sur_judge_blocks(int n);
hipError_t sur_prepare(hipStream_t st, const BatchIn& in, int n, int Mpad, int spot_every,
                       int spot_paths, void* X, uint8_t* act_a, int32_t* paths_a,
                       uint8_t* act_spot, uint8_t* act_b, int8_t* status, int8_t* region,
                       double* closed);
hipError_t sur_price(hipStream_t st, const BatchIn& in, int n, const float* y, const int8_t* status,
                     double* sur);
hipError_t sur_compare(hipStream_t st, const BatchIn& in, int n, double k_se, double tol,
                       const uint8_t* act_spot, const double* sur, const double* mc,
                       const double* mc_se, uint8_t* act_b, int8_t* status);
hipError_t sur_judge(hipStream_t st, const BatchIn& in, int n, const uint8_t* act_spot,
                     const int8_t* region, const double* sur, const double* mc,
                     const double* mc_se, double target, double k_sig, double* part,
                     double* stats, uint8_t* act_b, int8_t* status);
hipError_t sur_combine(hipStream_t st, int n, const int8_t* status, const double* sur,
                       const double* a, const double* b, const double* closed, double* out,
                       int32_t* src);

}  // namespace qk
