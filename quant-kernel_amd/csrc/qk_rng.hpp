// qk_rng.hpp - on-device normal streams for the gfx950 path kernels.
//
// Production stream: Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) keyed
// by the option's 64-bit seed, followed by Box-Muller (1958) in FP32 on the
// hardware transcendental units.  Draw d of path p of an option is component
// d%4 of block (p, d/4, 0, 0); draw k of a terminal (antithetic) estimator is
// component k%4 of block (k/4, 0, 0, 1).  Components: words (0,1) give
// (r cos, r sin) and words (2,3) the next pair, with
// u1 = (w>>8 + 1/2)/2^24, u2 = (w>>8)/2^24 and r = sqrt(-2 ln u1).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace qk {

struct U4 {
    uint32_t x, y, z, w;
};

// One Philox4x32-10 block.  The key words are wave-uniform in every caller,
// so the Weyl key schedule runs on the scalar unit.
__device__ __forceinline__ U4 philox4x32_10(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        const uint64_t pa = uint64_t(0xD2511F53u) * c.x;
        const uint64_t pb = uint64_t(0xCD9E8D57u) * c.z;
        const uint32_t nx = uint32_t(pb >> 32) ^ c.y ^ k0;
        const uint32_t nz = uint32_t(pa >> 32) ^ c.w ^ k1;
        c.y = uint32_t(pb);
        c.w = uint32_t(pa);
        c.x = nx;
        c.z = nz;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c;
}

// Box-Muller in FP32: v_log_f32 (log2), v_sqrt_f32, and v_sin/v_cos, which
// take their argument in revolutions, so u2 is used as is.
__device__ __forceinline__ void box_muller_f32(uint32_t a, uint32_t b, float& z0, float& z1) {
    const float u1 = __builtin_fmaf(float(a >> 8), 0x1p-24f, 0x1p-25f);
    const float u2 = float(b >> 8) * 0x1p-24f;
    const float rad = __builtin_amdgcn_sqrtf(__builtin_amdgcn_logf(u1) * -1.3862943611198906f);
    z0 = rad * __builtin_amdgcn_cosf(u2);
    z1 = rad * __builtin_amdgcn_sinf(u2);
}

struct Normal4 {
    float z0, z1, z2, z3;
};

__device__ __forceinline__ Normal4 philox_normals(uint32_t c0, uint32_t c1, uint32_t c3,
                                                  uint32_t k0, uint32_t k1) {
    const U4 w = philox4x32_10(U4{c0, c1, 0u, c3}, k0, k1);
    Normal4 n;
    box_muller_f32(w.x, w.y, n.z0, n.z1);
    box_muller_f32(w.z, w.w, n.z2, n.z3);
    return n;
}

}  // namespace qk
