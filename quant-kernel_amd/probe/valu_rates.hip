// valu_rates.hip - throughput of single wave64 VALU instruction kinds on the
// GPU it runs on (a measurement tool; not part of the library).
//
// Each kernel runs ITER iterations of CH independent dependency chains of one
// operation (or a fixed mix) per thread, over 16 workgroups of 256 threads
// per compute unit.  The time per operation group and wave on one SIMD is
// kernel time x SIMDs / (waves x ITER x CH); the ratios between kinds are the
// cost model used to read the Monte Carlo path loops (FP32 FMA = 1).
//
// Build and run:
//   hipcc --offload-arch=gfx950 -Xclang -target-feature -Xclang -packed-fp32-ops -O3 \
//       -o /tmp/valu_rates quant-kernel_amd/probe/valu_rates.hip && /tmp/valu_rates
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <utility>

#define CHECK(x)                                                                   \
    do {                                                                           \
        hipError_t e_ = (x);                                                       \
        if (e_ != hipSuccess) {                                                    \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));           \
            std::exit(1);                                                          \
        }                                                                          \
    } while (0)

constexpr int kIter = 4096;
constexpr int kCh = 8;

enum Op {
    OP_FMA32, OP_MAD64, OP_MULHI, OP_COS, OP_LOG, OP_SQRT, OP_EXP, OP_XOR3,
    OP_FMA64, OP_FMA32_COS, OP_MAD64_FMA32, OP_3FMA32_COS, OP_COUNT
};
const char* kNames[OP_COUNT] = {
    "v_fma_f32", "v_mad_u64_u32 + v_xor", "v_mul_hi_u32 + v_add", "v_cos_f32", "v_log_f32",
    "v_sqrt_f32", "v_exp_f32", "3-input xor (2 v_xor_b32)", "v_fma_f64", "v_fma_f32 + v_cos_f32",
    "v_mad_u64_u32 + v_xor + v_fma_f32", "3 v_fma_f32 + v_cos_f32"};

template <int OP>
__global__ __launch_bounds__(256) void k_rate(float* outf, uint32_t* outu, double* outd, float sf,
                                              uint32_t su) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    float f[kCh];
    uint32_t u[kCh];
    double d[kCh];
#pragma unroll
    for (int c = 0; c < kCh; ++c) {
        f[c] = sf + 0.001f * float(t + c);
        u[c] = su + t * 7u + c;
        d[c] = double(f[c]);
    }
    for (int i = 0; i < kIter; ++i) {
#pragma unroll
        for (int c = 0; c < kCh; ++c) {
            if constexpr (OP == OP_FMA32) f[c] = __builtin_fmaf(f[c], 0.999f, 0.0001f);
            if constexpr (OP == OP_MAD64 || OP == OP_MAD64_FMA32) {
                const uint64_t p = uint64_t(u[c]) * 0xD2511F53u;
                u[c] = uint32_t(p >> 32) ^ uint32_t(p);
            }
            if constexpr (OP == OP_MULHI) u[c] = __umulhi(u[c], 0xCD9E8D57u) + u[c];
            if constexpr (OP == OP_COS) f[c] = __builtin_amdgcn_cosf(f[c]);
            if constexpr (OP == OP_LOG) f[c] = __builtin_amdgcn_logf(f[c]);
            if constexpr (OP == OP_SQRT) f[c] = __builtin_amdgcn_sqrtf(f[c]);
            if constexpr (OP == OP_EXP) f[c] = __builtin_amdgcn_exp2f(f[c]);
            if constexpr (OP == OP_XOR3) u[c] = u[c] ^ u[(c + 1) % kCh] ^ u[(c + 3) % kCh];
            if constexpr (OP == OP_FMA64) d[c] = __builtin_fma(d[c], 0.999, 0.0001);
            if constexpr (OP == OP_FMA32_COS) {
                f[c] = __builtin_fmaf(f[c], 0.999f, 0.0001f);
                f[c] = __builtin_amdgcn_cosf(f[c]);
            }
            if constexpr (OP == OP_MAD64_FMA32) f[c] = __builtin_fmaf(f[c], 0.999f, 0.0001f);
            if constexpr (OP == OP_3FMA32_COS) {
                f[c] = __builtin_fmaf(f[c], 0.999f, 0.0001f);
                f[c] = __builtin_fmaf(f[c], 0.998f, 0.0002f);
                f[c] = __builtin_fmaf(f[c], 0.997f, 0.0003f);
                f[c] = __builtin_amdgcn_cosf(f[c]);
            }
        }
    }
    float fs = 0.0f;
    uint32_t us = 0u;
    double ds = 0.0;
#pragma unroll
    for (int c = 0; c < kCh; ++c) {
        fs += f[c];
        us ^= u[c];
        ds += d[c];
    }
    outf[t] = fs;
    outu[t] = us;
    outd[t] = ds;
}

template <int OP>
float time_op(int blocks, float* of, uint32_t* ou, double* od) {
    hipEvent_t a, b;
    CHECK(hipEventCreate(&a));
    CHECK(hipEventCreate(&b));
    hipLaunchKernelGGL(k_rate<OP>, dim3(blocks), dim3(256), 0, 0, of, ou, od, 0.5f, 1u);
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(a));
    for (int r = 0; r < 3; ++r)
        hipLaunchKernelGGL(k_rate<OP>, dim3(blocks), dim3(256), 0, 0, of, ou, od, 0.5f, 1u);
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(b));
    CHECK(hipEventSynchronize(b));
    float ms = 0.0f;
    CHECK(hipEventElapsedTime(&ms, a, b));
    CHECK(hipEventDestroy(a));
    CHECK(hipEventDestroy(b));
    return ms / 3.0f;
}

template <int... OPS>
void time_all(int blocks, float* of, uint32_t* ou, double* od, float* ms,
              std::integer_sequence<int, OPS...>) {
    ((ms[OPS] = time_op<OPS>(blocks, of, ou, od)), ...);
}

int main() {
    hipDeviceProp_t prop;
    CHECK(hipGetDeviceProperties(&prop, 0));
    const int cus = prop.multiProcessorCount;
    const int blocks = cus * 16;
    const size_t n = size_t(blocks) * 256;
    float* of;
    uint32_t* ou;
    double* od;
    CHECK(hipMalloc(&of, n * sizeof(float)));
    CHECK(hipMalloc(&ou, n * sizeof(uint32_t)));
    CHECK(hipMalloc(&od, n * sizeof(double)));
    float ms[OP_COUNT];
    time_all(blocks, of, ou, od, ms, std::make_integer_sequence<int, OP_COUNT>{});
    CHECK(hipDeviceSynchronize());
    const double simds = 4.0 * cus;
    const double groups = double(n / 64) * kIter * kCh / simds;  // per SIMD
    const double base = ms[OP_FMA32] * 1e6 / groups;
    std::printf("%s, %d CUs, %d workgroups x 256 threads, %d x %d operation groups per thread\n",
                prop.gcnArchName, cus, blocks, kIter, kCh);
    std::printf("| operation group (one wave64, one SIMD) | kernel ms | ns per group | FP32 FMA = 1 |\n");
    std::printf("|---|---|---|---|\n");
    for (int i = 0; i < OP_COUNT; ++i) {
        const double ns = ms[i] * 1e6 / groups;
        std::printf("| %s | %.3f | %.3f | %.2f |\n", kNames[i], ms[i], ns, ns / base);
    }
    CHECK(hipFree(of));
    CHECK(hipFree(ou));
    CHECK(hipFree(od));
    return 0;
}
