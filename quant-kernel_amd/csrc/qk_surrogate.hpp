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
    SUR_FALLBACK_CHECK = 2,  // spot-check MC disagreed: plain MC
    SUR_CLOSED = 3,          // invalid (NaN) or t <= eps (intrinsic), as the MC entry point
};

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
// Pricing with fallback; inputs and outputs in host memory (device == false)
// or device memory (device == true).
int32_t sur_heston_price(const BatchIn& in, int32_t n, bool device, double* out_price,
                         int32_t* out_source, std::string* err);
bool sur_last_timing(double* ms_network, double* ms_mc);

// kernels (qk_sur_kernels.hip)
hipError_t sur_prepare(hipStream_t st, const BatchIn& in, int n, int Mpad, int spot_every, void* X,
                       uint8_t* act_fb, uint8_t* act_spot, int8_t* status, double* closed);
hipError_t sur_price(hipStream_t st, const BatchIn& in, int n, const float* y, const int8_t* status,
                     double* sur);
hipError_t sur_compare(hipStream_t st, const BatchIn& in, int n, double k_se,
                       const uint8_t* act_spot, const double* sur, const double* mc,
                       const double* mc_se, uint8_t* act_fb, int8_t* status);
hipError_t sur_combine(hipStream_t st, int n, const int8_t* status, const double* sur,
                       const double* fb, const double* closed, double* out, int32_t* src);

}  // namespace qk
