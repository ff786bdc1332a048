// qk_mlp.hpp - the Heston price surrogate: an 8 -> 4x256 (GELU) -> 1 MLP.
//
// Hidden layers run on MFMA (v_mfma_f32_32x32x16_bf16): BF16 weights and
// activations, FP32 accumulation.  The output layer is an FP32 dot product
// on the FP32 activations of the last hidden layer.  Master weights, Adam
// moments and gradients are FP32.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qk {
namespace mlp {

constexpr int kIn = 8;      // log(K/S), T, r, kappa, theta, sigma(vol of vol), rho, v0
constexpr int kInPad = 32;  // first GEMM's K, zero padded
constexpr int kH = 256;     // hidden width
constexpr int kHidden = 4;  // hidden layers

// FP32 master parameters, layer l = 1..4 weights stored transposed
// (Wt[out][in]); w5 is the output row vector.
constexpr int kW1 = 0;
constexpr int kB1 = kW1 + kH * kInPad;
constexpr int kW2 = kB1 + kH;
constexpr int kB2 = kW2 + kH * kH;
constexpr int kW3 = kB2 + kH;
constexpr int kB3 = kW3 + kH * kH;
constexpr int kW4 = kB3 + kH;
constexpr int kB4 = kW4 + kH * kH;
constexpr int kW5 = kB4 + kH;
constexpr int kB5 = kW5 + kH;
constexpr int kParams = kB5 + 4;  // padded to a multiple of 4
constexpr int kTrueParams = kIn * kH + 3 * kH * kH + 4 * kH + kH + 1;  // 199937

constexpr int kWOff[kHidden] = {kW1, kW2, kW3, kW4};
constexpr int kBOff[kHidden] = {kB1, kB2, kB3, kB4};
constexpr int kK[kHidden] = {kInPad, kH, kH, kH};

// BF16 copies: Wt_l for the forward GEMMs, W_l ([in][out]) for the
// backward-data GEMMs of layers 2..4.
constexpr int kBfWt[kHidden] = {0, kH * kInPad, kH * kInPad + kH * kH, kH * kInPad + 2 * kH * kH};
constexpr int kBfW[kHidden] = {-1, kH * kInPad + 3 * kH * kH, kH * kInPad + 4 * kH * kH,
                               kH * kInPad + 5 * kH * kH};
constexpr int kBfCount = kH * kInPad + 6 * kH * kH;

// Training box of the features (the Heston C1 parameter ranges).
struct Box {
    double lo[kIn], hi[kIn];
};
constexpr Box kBox = {{-0.35667494393873245, 0.05, 0.0, 0.5, 0.01, 0.1, -0.9, 0.01},
                      {0.26236426446749106, 2.0, 0.08, 5.0, 0.16, 1.0, -0.1, 0.16}};

// Closed-form control the network corrects: the Black-Scholes call price per
// unit spot (q = 0) at the expected average variance of the Heston variance
// process over [0, T], theta + (v0 - theta)(1 - e^{-kappa T})/(kappa T).
// The model's output is control + network.
__host__ __device__ inline double heston_control(const double* f) {
    const double lk = f[0], t = f[1], r = f[2], kappa = f[3], theta = f[4], v0 = f[7];
    const double kt = kappa * t;
    const double w = kt > 1e-8 ? (1.0 - exp(-kt)) / kt : 1.0 - 0.5 * kt;
    const double var = theta + (v0 - theta) * w;
    const double sd = sqrt(var * t);
    const double d1 = (-lk + (r + 0.5 * var) * t) / sd;
    const double d2 = d1 - sd;
    return 0.5 * erfc(-d1 * 0.70710678118654752440) -
           exp(lk - r * t) * 0.5 * erfc(-d2 * 0.70710678118654752440);
}

// ---- kernels' launchers (qk_mlp.hip) -----------------------------------
// C[m][n] = sum_k A[m][k] B[n][k]; bf16 inputs (K-contiguous rows), FP32 out.
// M, N multiples of 32, K a multiple of 16, lda/ldb multiples of 8.
hipError_t gemm_nt(hipStream_t st, const void* A, int lda, const void* B, int ldb, float* C, int ldc,
                   int M, int N, int K);
// y[m] = network(X[m]) for m < M, all layers in one launch: X is [M][kInPad]
// bf16 (rows >= M are neither read nor written), wbf the BF16 forward
// weights, params the FP32 biases, w5 and b5.
hipError_t infer(hipStream_t st, const void* X, const void* wbf, const float* params, float* y,
                 int M);
// Zb = Z + b; H = gelu(Zb) (bf16, [M][N]); optional Zb (fp32), Ht (bf16 [N][ldt]), Hf (fp32).
hipError_t act_forward(hipStream_t st, const float* Z, const float* bias, int M, int N, float* Zb,
                       void* H, void* Ht, int ldt, float* Hf);
// y[m] = b5 + sum_k Hf[m][k] w5[k]; with target: dy = 2 (y - t)/M_valid on valid rows, 0 else;
// sq[m] = (y - t)^2 on valid rows.
hipError_t out_forward(hipStream_t st, const float* Hf, const float* params, int M, int M_valid,
                       const float* target, float* y, float* dy, float* sq);
// dZ = (dH or dy (x) w5) * gelu'(Zb): fp32 dZf, bf16 dZ [M][N] and dZt [N][M].
hipError_t act_backward(hipStream_t st, const float* dH, const float* dy, const float* w5,
                        const float* Zb, int M, int N, float* dZf, void* dZ, void* dZt);
// out[n] (+)= sum_m X[m][n] (* w[m]) in a fixed order; part needs ceil(M/128)*N floats.
hipError_t colsum(hipStream_t st, const float* X, int M, int N, const float* w, float* part,
                  float* out);
hipError_t adam(hipStream_t st, float* p, const float* g, float* m, float* v, int count, float lr,
                float b1, float b2, float eps, float bc1, float bc2);
hipError_t refresh_bf16(hipStream_t st, const float* p, void* wbf);
// X/Xt rows from normalised FP32 features feat[src][8] (src = perm[start + m]).
hipError_t gather(hipStream_t st, const float* feat, const float* tgt, const int32_t* perm,
                  int start, int M_valid, int M, void* X, void* Xt, float* t);
hipError_t zero_f32(hipStream_t st, float* p, int count);

}  // namespace mlp
}  // namespace qk
