// qk_rng.hpp - on-device normal streams for the gfx950 path kernels.
//
// Production stream: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) keyed
// by the option's 64-bit seed, followed by Box-Muller (1958) in FP32 on the
// hardware transcendental units.
//
// Four normals per block: draw d of path p of an LSMC or MLMC path is
// component d%4 of block (p, d/4, 0, level).  Words (w0,w1) give components
// 0,1 and (w2,w3) components 2,3:
//   u1 = fma(float(w_even), 2^-32, 2^-33)           (FP32, in (0, 1])
//   f  = (w_odd & 0x7FFFFF) / 2^23                   (exact, in [0, 1))
//   (z_even, z_odd) = sqrt(-2 ln u1) * (cos 2 pi f, sin 2 pi f)
//
// Eight normals per block (philox_normals8), the four-pair layout: slot s
// is word w_s alone,
//   u1 = (2m + 1) 2^-24, m = the low 23 bits of w_s  (exact in FP32)
//   theta_s = 2 pi a_s / 512, a_s = the top 9 bits of w_s,
// (z_2s, z_2s+1) = sqrt(-2 ln u1) (cos theta_s, sin theta_s), with the
// cosine and sine read from a 512-entry table that each workgroup stages in
// LDS.  The radius is capped at sqrt(48 ln 2) = 5.77 sigma; 512 equally
// spaced angles give E[cos^n sin^k] of a uniform angle exactly for
// n + k < 512.  Draw k of a terminal estimator is component k%8 of block
// (k/8, 0, 0, 1) (cosine of slot (k%8)/2 for even k, sine for odd k).
// Heston paths take step j from slot j%4 of block (p, j/4, 0, 0) (both
// components), stepwise GBM paths step j from component j%2 of the pair of
// Heston step j/2.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qk {

struct U4 {
    uint32_t x, y, z, w;
};

constexpr uint32_t kPhiloxM0 = 0xD2511F53u, kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u, kPhiloxW1 = 0xBB67AE85u;

// a ^ b ^ s in one gfx950 v_bitop3_b32 (truth table 0x96); s is wave-uniform.
__device__ __forceinline__ uint32_t xor3_vvs(uint32_t a, uint32_t b, uint32_t s) {
    uint32_t r;
    asm("v_bitop3_b32 %0, %1, %2, %3 bitop3:0x96" : "=v"(r) : "v"(a), "v"(b), "s"(s));
    return r;
}

__device__ __forceinline__ void mulhilo(uint32_t m, uint32_t x, uint32_t& hi, uint32_t& lo) {
    const uint64_t p = uint64_t(m) * x;
    hi = uint32_t(p >> 32);
    lo = uint32_t(p);
}

// Philox4x32-10 of the counter (c0, c1, 0, c3) where c0 varies per lane and
// c1, c3, k0, k1 are wave-uniform.  Rounds 1-3 are left to the compiler so
// the uniform parts stay on the scalar unit; rounds 4-10 use the 3-input xor.
__device__ __forceinline__ U4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c3, uint32_t k0,
                                            uint32_t k1) {
    uint32_t x0 = c0, x1 = c1, x2 = 0u, x3 = c3;
#pragma unroll
    for (int round = 0; round < 3; ++round) {
        uint32_t h0, l0, h2, l2;
        mulhilo(kPhiloxM0, x0, h0, l0);
        mulhilo(kPhiloxM1, x2, h2, l2);
        const uint32_t n0 = h2 ^ x1 ^ k0;
        const uint32_t n2 = h0 ^ x3 ^ k1;
        x0 = n0;
        x1 = l2;
        x2 = n2;
        x3 = l0;
        k0 += kPhiloxW0;
        k1 += kPhiloxW1;
    }
#pragma unroll
    for (int round = 3; round < 10; ++round) {
        uint32_t h0, l0, h2, l2;
        mulhilo(kPhiloxM0, x0, h0, l0);
        mulhilo(kPhiloxM1, x2, h2, l2);
        const uint32_t n0 = xor3_vvs(h2, x1, k0);
        const uint32_t n2 = xor3_vvs(h0, x3, k1);
        x0 = n0;
        x1 = l2;
        x2 = n2;
        x3 = l0;
        k0 += kPhiloxW0;
        k1 += kPhiloxW1;
    }
    return U4{x0, x1, x2, x3};
}

// Box-Muller in FP32: v_log_f32 (log2), v_sqrt_f32, and v_sin/v_cos, which
// take their argument in revolutions: 1 + f in [1, 2) has the same sine and
// cosine as f and is built with one and-or.
__device__ __forceinline__ void box_muller_f32(uint32_t a, uint32_t b, float& z0, float& z1) {
    const float u1 = __builtin_fmaf(float(a), 0x1p-32f, 0x1p-33f);
    const float rev = __builtin_bit_cast(float, (b & 0x007FFFFFu) | 0x3F800000u);
    const float rad = __builtin_amdgcn_sqrtf(__builtin_amdgcn_logf(u1) * -1.3862943611198906f);
    z0 = rad * __builtin_amdgcn_cosf(rev);
    z1 = rad * __builtin_amdgcn_sinf(rev);
}

struct Normal4 {
    float z0, z1, z2, z3;
};

// Four-pair layout of the production streams: slot s of a Philox block is
// word s alone.  u1 = (2m + 1) 2^-24 from its low 23 bits m (u23: exact in
// FP32 and in (0, 1)), and theta = 2 pi a / 512 from its top 9 bits a, read
// as (cos, sin) from a 512-entry table; quad_offset is the entry's byte
// offset (8a).
__device__ __forceinline__ float u23(uint32_t w) {
    return __builtin_bit_cast(float, (w & 0x007FFFFFu) | 0x3F800000u) - 0x1.fffffep-1f;
}

constexpr int kTrigEntries = 512;  // angle table: (cos, sin) of 2 pi a / 512
__device__ __forceinline__ uint32_t quad_offset(uint32_t w) { return (w >> 20) & 0xFF8u; }
__device__ __forceinline__ float2 trig_at(const float2* trig, uint32_t byte_off) {
    return *reinterpret_cast<const float2*>(reinterpret_cast<const char*>(trig) + byte_off);
}

// Copies the device's angle table (global memory) into a workgroup's LDS
// copy; every thread of the workgroup takes part, then a barrier.
__device__ __forceinline__ void stage_trig_table(const float2* __restrict__ src, float2* dst) {
    const float4* s4 = reinterpret_cast<const float4*>(src);
    float4* d4 = reinterpret_cast<float4*>(dst);
    for (int i = threadIdx.x; i < kTrigEntries / 2; i += blockDim.x) d4[i] = s4[i];
    __syncthreads();
}

// The eight normals of one block: (r_s cos theta_s, r_s sin theta_s) of
// slots s = 0 .. 3, r_s = sqrt(-2 ln u1_s).
struct Normal8 {
    float z[8];
};

__device__ __forceinline__ Normal8 philox_normals8(const U4& w, const float2* trig) {
    const uint32_t ws[4] = {w.x, w.y, w.z, w.w};
    Normal8 n;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const float2 cs = trig_at(trig, quad_offset(ws[s]));
        const float rad =
            __builtin_amdgcn_sqrtf(__builtin_amdgcn_logf(u23(ws[s])) * -1.3862943611198906f);
        n.z[2 * s] = rad * cs.x;
        n.z[2 * s + 1] = rad * cs.y;
    }
    return n;
}

__device__ __forceinline__ Normal4 philox_normals(uint32_t c0, uint32_t c1, uint32_t c3,
                                                  uint32_t k0, uint32_t k1) {
    const U4 w = philox4x32_10(c0, c1, c3, k0, k1);
    Normal4 n;
    box_muller_f32(w.x, w.y, n.z0, n.z1);
    box_muller_f32(w.z, w.w, n.z2, n.z3);
    return n;
}

}  // namespace qk
