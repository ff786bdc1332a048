// qk_mlp.hip - gfx950 kernels of the Heston surrogate MLP (forward, backward,
// Adam).  GEMMs use v_mfma_f32_32x32x16_bf16 with its wave64 operand layout:
//   A[i][k] -> lane i + 32*(k/8), element k%8   (8 consecutive k per lane)
//   B[k][j] -> lane j + 32*(k/8), element k%8
//   D[i][j] -> lane j + 32*((i/4)%2), register (i%4) + 4*(i/8)
// so with K-contiguous rows each lane loads its 8 operand values with one
// 16-byte load.  Every reduction (bias and output-layer gradients, loss) runs
// in a fixed order, so training is bitwise reproducible.
#include "qk_mlp.hpp"

namespace qk {
namespace mlp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ float gelu(float x) {
    return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f));
}

__device__ __forceinline__ float gelu_grad(float x) {
    const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
    const float pdf = 0.39894228040143268f * __expf(-0.5f * x * x);
    return cdf + x * pdf;
}

// One wave per 32x32 output tile, 4 waves per 64x64 block tile.
__global__ __launch_bounds__(256) void k_gemm_nt(const __bf16* __restrict__ A, int lda,
                                                 const __bf16* __restrict__ B, int ldb,
                                                 float* __restrict__ C, int ldc, int M, int N,
                                                 int K) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int tm = blockIdx.x * 64 + (wave >> 1) * 32;
    const int tn = blockIdx.y * 64 + (wave & 1) * 32;
    if (tm >= M || tn >= N) return;
    const __bf16* a = A + size_t(tm + (lane & 31)) * lda + (lane >> 5) * 8;
    const __bf16* b = B + size_t(tn + (lane & 31)) * ldb + (lane >> 5) * 8;
    f32x16 acc = {};
    int k0 = 0;
    for (; k0 + 32 <= K; k0 += 32) {
        const bf16x8 a0 = *reinterpret_cast<const bf16x8*>(a + k0);
        const bf16x8 b0 = *reinterpret_cast<const bf16x8*>(b + k0);
        const bf16x8 a1 = *reinterpret_cast<const bf16x8*>(a + k0 + 16);
        const bf16x8 b1 = *reinterpret_cast<const bf16x8*>(b + k0 + 16);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b0, acc, 0, 0, 0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a1, b1, acc, 0, 0, 0);
    }
    if (k0 < K) {
        const bf16x8 a0 = *reinterpret_cast<const bf16x8*>(a + k0);
        const bf16x8 b0 = *reinterpret_cast<const bf16x8*>(b + k0);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a0, b0, acc, 0, 0, 0);
    }
    const int col = tn + (lane & 31);
    const int row0 = tm + 4 * (lane >> 5);
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        const int row = row0 + (r & 3) + 8 * (r >> 2);
        C[size_t(row) * ldc + col] = acc[r];
    }
}

// Fused inference: the four hidden layers and the output layer of 256 rows
// per workgroup in one launch.  Waves 2r and 2r+1 (r = 0..7) share rows
// 32r .. 32r+31 of the tile and compute output columns 0-127 and 128-255 of
// a layer for them (4 MFMA tiles of 32x32 each).  The layer's activations
// stay in LDS as BF16 (the A operands of the next layer); the weights of a
// layer pass through LDS in chunks of kInfKc inputs, staged once per
// workgroup, and the threads load the next chunk (the next layer's first
// after a layer's last) into registers while the current chunk's MFMAs run.
// At about 155 KB of LDS one workgroup of 16 waves fills a CU; a tile of 256
// rows stages each weight once per 256 rows (the network stage of 1M
// queries took 1.78, 1.64 and 1.12 ms at 64, 128 and 256 rows).  The hidden
// activations equal those of k_gemm_nt + k_act_forward bitwise (same MFMA
// sequence over K, same bias and GELU); the output dot product is summed in
// another order.
constexpr int kInfRows = 256;              // rows per workgroup: 32 per pair of waves
constexpr int kInfThreads = kInfRows * 4;  // waves (row block, column half)
constexpr int kInfKc = 32;  // every layer's K is a multiple of 32
constexpr int kActStride = kH + 8;  // bf16; the pad spreads rows over the LDS banks
constexpr int kWStride = kInfKc + 8;
// The wave layout assumes 16 waves: 8 row blocks of 32 rows x 2 column
// halves of 4 MFMA tiles of 32 columns.
static_assert(kInfRows == 8 * 32 && kInfThreads == 16 * 64, "k_mlp_infer: 16 waves of 32 rows");
static_assert(kH == 2 * 4 * 32, "k_mlp_infer: two column halves of four 32-wide MFMA tiles");
static_assert(kInPad % kInfKc == 0 && kH % kInfKc == 0, "k_mlp_infer: K must be whole chunks");
// Every thread stages the same whole number of 8-wide weight vectors per chunk.
static_assert((kH * (kInfKc / 8)) % kInfThreads == 0 && kH * (kInfKc / 8) / kInfThreads >= 1,
              "k_mlp_infer: weight chunk must split evenly over the threads");
// Static LDS of the kernel (activations, weight chunk, half dot products)
// within the 160 KB of a gfx950 CU.
static_assert(kInfRows * kActStride * 2 + kH * kWStride * 2 + kInfRows * 4 <= 160 * 1024,
              "k_mlp_infer: LDS over 160 KB");

__global__ __launch_bounds__(kInfThreads) void k_mlp_infer(const __bf16* __restrict__ X,
                                                   const __bf16* __restrict__ wbf,
                                                   const float* __restrict__ params,
                                                   float* __restrict__ y, int M) {
    __shared__ __attribute__((aligned(16))) __bf16 act[kInfRows * kActStride];
    __shared__ __attribute__((aligned(16))) __bf16 wch[kH * kWStride];
    __shared__ float s_half[kInfRows];  // the column-half 1 dot products
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int wr = wave >> 1, wc = wave & 1;  // row block, column half
    const int li = lane & 31, lk = (lane >> 5) * 8;
    const int row0 = blockIdx.x * kInfRows + wr * 32;  // the row block's first row
    const int arow = (wr * 32 + li) * kActStride;
    float ydot[16];
#pragma unroll
    for (int r = 0; r < 16; ++r) ydot[r] = 0.0f;
    // weight chunk c of layer l, Wt_l[n][32c .. 32c + 32) for the kH rows n,
    // as kH * 4 vectors of 8 bf16: thread t stages vectors
    // t * kPer .. t * kPer + kPer - 1 (vector v: row v / 4, columns 8 (v % 4))
    constexpr int kVec = kInfKc / 8;
    constexpr int kPer = kH * kVec / kInfThreads;
    bf16x8 pre[kPer];
    auto load_chunk = [&](int l, int c) {
#pragma unroll
        for (int i = 0; i < kPer; ++i) {
            const int v = int(threadIdx.x) * kPer + i;
            pre[i] = *reinterpret_cast<const bf16x8*>(wbf + kBfWt[l] + size_t(v / kVec) * kK[l] +
                                                       c * kInfKc + (v % kVec) * 8);
        }
    };
    load_chunk(0, 0);
    for (int l = 0; l < kHidden; ++l) {
        const int nch = kK[l] / kInfKc;
        f32x16 acc[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) acc[t] = f32x16{};
        for (int c = 0; c < nch; ++c) {
            // the previous chunk (and, at c = 0, the previous epilogue) done
            __syncthreads();
#pragma unroll
            for (int i = 0; i < kPer; ++i) {
                const int v = int(threadIdx.x) * kPer + i;
                *reinterpret_cast<bf16x8*>(wch + (v / kVec) * kWStride + (v % kVec) * 8) = pre[i];
            }
            __syncthreads();
            if (c + 1 < nch) load_chunk(l, c + 1);
            else if (l + 1 < kHidden) load_chunk(l + 1, 0);
            const int k0 = c * kInfKc;
#pragma unroll
            for (int ks = 0; ks < kInfKc; ks += 16) {
                bf16x8 a;
                if (l == 0) {
                    const int r = row0 + li;
                    a = r < M ? *reinterpret_cast<const bf16x8*>(X + size_t(r) * kInPad + k0 + ks + lk)
                              : bf16x8{};
                } else {
                    a = *reinterpret_cast<const bf16x8*>(act + arow + k0 + ks + lk);
                }
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    const int n = wc * 128 + t * 32 + li;
                    const bf16x8 b = *reinterpret_cast<const bf16x8*>(wch + n * kWStride + ks + lk);
                    acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[t], 0, 0, 0);
                }
            }
        }
        // both column halves have read this layer's A operands before either
        // overwrites them
        __syncthreads();
        // bias + GELU; D[i][j]: lane j + 32((i/4)%2), register (i%4) + 4(i/8)
        const float* bias = params + kBOff[l];
        const bool last = l == kHidden - 1;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int j = wc * 128 + t * 32 + li;
            const float bj = bias[j];
            const float w5 = last ? params[kW5 + j] : 0.0f;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const float h = gelu(acc[t][r] + bj);
                if (last) {
                    ydot[r] = __builtin_fmaf(h, w5, ydot[r]);
                } else {
                    const int i = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                    act[(wr * 32 + i) * kActStride + j] = __bf16(h);
                }
            }
        }
    }
    // row i's half dot product: the 32 lanes of one half-wave hold its terms
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        float v = ydot[r];
#pragma unroll
        for (int off = 16; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
        ydot[r] = v;
    }
    if (wc == 1 && li == 0) {
#pragma unroll
        for (int r = 0; r < 16; ++r) s_half[wr * 32 + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3)] = ydot[r];
    }
    __syncthreads();
    if (wc == 0 && li == 0) {
        const float b5 = params[kB5];
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int i = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            const int row = row0 + i;
            if (row < M) y[row] = ydot[r] + s_half[wr * 32 + i] + b5;
        }
    }
}

__global__ __launch_bounds__(256) void k_act_forward(const float* __restrict__ Z,
                                                     const float* __restrict__ bias, int M, int N,
                                                     float* __restrict__ Zb, __bf16* __restrict__ H,
                                                     __bf16* __restrict__ Ht, int ldt,
                                                     float* __restrict__ Hf) {
    const size_t total = size_t(M) * N;
    for (size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < total;
         idx += size_t(gridDim.x) * blockDim.x) {
        const int m = int(idx / N), n = int(idx % N);
        const float z = Z[idx] + bias[n];
        const float h = gelu(z);
        if (Zb) Zb[idx] = z;
        H[idx] = __bf16(h);
        if (Ht) Ht[size_t(n) * ldt + m] = __bf16(h);
        if (Hf) Hf[idx] = h;
    }
}

// one wave per row
__global__ __launch_bounds__(256) void k_out_forward(const float* __restrict__ Hf,
                                                     const float* __restrict__ params, int M,
                                                     int M_valid, const float* __restrict__ target,
                                                     float* __restrict__ y, float* __restrict__ dy,
                                                     float* __restrict__ sq) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (row >= M) return;
    const float* h = Hf + size_t(row) * kH;
    const float* w = params + kW5;
    float s = 0.0f;
#pragma unroll
    for (int j = 0; j < kH / 64; ++j) s = __builtin_fmaf(h[lane + 64 * j], w[lane + 64 * j], s);
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) {
        const float out = s + params[kB5];
        y[row] = out;
        if (target) {
            const bool valid = row < M_valid;
            const float diff = valid ? out - target[row] : 0.0f;
            dy[row] = valid ? 2.0f * diff / float(M_valid) : 0.0f;
            sq[row] = diff * diff;
        }
    }
}

__global__ __launch_bounds__(256) void k_act_backward(const float* __restrict__ dH,
                                                      const float* __restrict__ dy,
                                                      const float* __restrict__ w5,
                                                      const float* __restrict__ Zb, int M, int N,
                                                      float* __restrict__ dZf,
                                                      __bf16* __restrict__ dZ,
                                                      __bf16* __restrict__ dZt) {
    const size_t total = size_t(M) * N;
    for (size_t idx = size_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < total;
         idx += size_t(gridDim.x) * blockDim.x) {
        const int m = int(idx / N), n = int(idx % N);
        const float g = dH ? dH[idx] : dy[m] * w5[n];
        const float d = g * gelu_grad(Zb[idx]);
        dZf[idx] = d;
        dZ[idx] = __bf16(d);
        dZt[size_t(n) * M + m] = __bf16(d);
    }
}

constexpr int kColRows = 128;

__global__ __launch_bounds__(256) void k_colsum_part(const float* __restrict__ X, int M, int N,
                                                     const float* __restrict__ w,
                                                     float* __restrict__ part) {
    const int n = threadIdx.x;
    if (n >= N) return;
    const int m0 = blockIdx.x * kColRows, m1 = min(M, m0 + kColRows);
    float s = 0.0f;
    for (int m = m0; m < m1; ++m) {
        const float x = X[size_t(m) * N + n];
        s = w ? __builtin_fmaf(x, w[m], s) : s + x;
    }
    part[size_t(blockIdx.x) * N + n] = s;
}

__global__ __launch_bounds__(256) void k_colsum_final(const float* __restrict__ part, int G, int N,
                                                      float* __restrict__ out) {
    const int n = threadIdx.x;
    if (n >= N) return;
    float s = 0.0f;
    for (int g = 0; g < G; ++g) s += part[size_t(g) * N + n];
    out[n] = s;
}

__global__ __launch_bounds__(256) void k_adam(float* __restrict__ p, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v,
                                              int count, float lr, float b1, float b2, float eps,
                                              float bc1, float bc2) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const float gi = g[i];
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    }
}

__global__ __launch_bounds__(256) void k_refresh(const float* __restrict__ p,
                                                 __bf16* __restrict__ wbf) {
    // Wt copies (same layout as the master) and W = transposed copies
    const int total = kH * kInPad + 3 * kH * kH;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gridDim.x * blockDim.x) {
        int layer, off;
        if (i < kH * kInPad) {
            layer = 0;
            off = i;
        } else {
            layer = 1 + (i - kH * kInPad) / (kH * kH);
            off = (i - kH * kInPad) % (kH * kH);
        }
        const float w = p[kWOff[layer] + off];
        wbf[kBfWt[layer] + off] = __bf16(w);
        if (layer > 0) {
            const int o = off / kH, in = off % kH;  // Wt[o][in] -> W[in][o]
            wbf[kBfW[layer] + in * kH + o] = __bf16(w);
        }
    }
}

__global__ __launch_bounds__(256) void k_gather(const float* __restrict__ feat,
                                                const float* __restrict__ tgt,
                                                const int32_t* __restrict__ perm, int start,
                                                int M_valid, int M, __bf16* __restrict__ X,
                                                __bf16* __restrict__ Xt, float* __restrict__ t) {
    const int total = M * kInPad;
    for (int idx = blockIdx.x * blockDim.x + threadIdx.x; idx < total;
         idx += gridDim.x * blockDim.x) {
        const int m = idx / kInPad, c = idx % kInPad;
        float x = 0.0f;
        if (m < M_valid) {
            const int src = perm ? perm[start + m] : start + m;
            if (c < kIn) x = feat[size_t(src) * kIn + c];
            if (c == 0 && t) t[m] = tgt[src];
        } else if (c == 0 && t) {
            t[m] = 0.0f;
        }
        X[idx] = __bf16(x);
        if (Xt) Xt[size_t(c) * M + m] = __bf16(x);
    }
}

__global__ void k_zero(float* p, int count) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x)
        p[i] = 0.0f;
}

static int grid_for(size_t total) {
    const size_t g = (total + 255) / 256;
    return int(g < 4096 ? (g ? g : 1) : 4096);
}

hipError_t gemm_nt(hipStream_t st, const void* A, int lda, const void* B, int ldb, float* C, int ldc,
                   int M, int N, int K) {
    if (M % 32 || N % 32 || K % 16 || lda % 8 || ldb % 8) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_gemm_nt, dim3((M + 63) / 64, (N + 63) / 64), dim3(256), 0, st,
                       static_cast<const __bf16*>(A), lda, static_cast<const __bf16*>(B), ldb, C,
                       ldc, M, N, K);
    return hipGetLastError();
}

hipError_t infer(hipStream_t st, const void* X, const void* wbf, const float* params, float* y,
                 int M) {
    if (M <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_mlp_infer, dim3((M + kInfRows - 1) / kInfRows), dim3(kInfThreads), 0, st,
                       static_cast<const __bf16*>(X), static_cast<const __bf16*>(wbf), params, y, M);
    return hipGetLastError();
}

hipError_t act_forward(hipStream_t st, const float* Z, const float* bias, int M, int N, float* Zb,
                       void* H, void* Ht, int ldt, float* Hf) {
    hipLaunchKernelGGL(k_act_forward, dim3(grid_for(size_t(M) * N)), dim3(256), 0, st, Z, bias, M,
                       N, Zb, static_cast<__bf16*>(H), static_cast<__bf16*>(Ht), ldt, Hf);
    return hipGetLastError();
}

hipError_t out_forward(hipStream_t st, const float* Hf, const float* params, int M, int M_valid,
                       const float* target, float* y, float* dy, float* sq) {
    hipLaunchKernelGGL(k_out_forward, dim3((M + 3) / 4), dim3(256), 0, st, Hf, params, M, M_valid,
                       target, y, dy, sq);
    return hipGetLastError();
}

hipError_t act_backward(hipStream_t st, const float* dH, const float* dy, const float* w5,
                        const float* Zb, int M, int N, float* dZf, void* dZ, void* dZt) {
    hipLaunchKernelGGL(k_act_backward, dim3(grid_for(size_t(M) * N)), dim3(256), 0, st, dH, dy, w5,
                       Zb, M, N, dZf, static_cast<__bf16*>(dZ), static_cast<__bf16*>(dZt));
    return hipGetLastError();
}

hipError_t colsum(hipStream_t st, const float* X, int M, int N, const float* w, float* part,
                  float* out) {
    if (N > 256) return hipErrorInvalidValue;
    const int G = (M + kColRows - 1) / kColRows;
    hipLaunchKernelGGL(k_colsum_part, dim3(G), dim3(256), 0, st, X, M, N, w, part);
    hipLaunchKernelGGL(k_colsum_final, dim3(1), dim3(256), 0, st, part, G, N, out);
    return hipGetLastError();
}

hipError_t adam(hipStream_t st, float* p, const float* g, float* m, float* v, int count, float lr,
                float b1, float b2, float eps, float bc1, float bc2) {
    hipLaunchKernelGGL(k_adam, dim3(grid_for(size_t(count))), dim3(256), 0, st, p, g, m, v, count,
                       lr, b1, b2, eps, bc1, bc2);
    return hipGetLastError();
}

hipError_t refresh_bf16(hipStream_t st, const float* p, void* wbf) {
    hipLaunchKernelGGL(k_refresh, dim3(grid_for(size_t(kH * kInPad + 3 * kH * kH))), dim3(256), 0,
                       st, p, static_cast<__bf16*>(wbf));
    return hipGetLastError();
}

hipError_t gather(hipStream_t st, const float* feat, const float* tgt, const int32_t* perm,
                  int start, int M_valid, int M, void* X, void* Xt, float* t) {
    hipLaunchKernelGGL(k_gather, dim3(grid_for(size_t(M) * kInPad)), dim3(256), 0, st, feat, tgt,
                       perm, start, M_valid, M, static_cast<__bf16*>(X), static_cast<__bf16*>(Xt),
                       t);
    return hipGetLastError();
}

hipError_t zero_f32(hipStream_t st, float* p, int count) {
    hipLaunchKernelGGL(k_zero, dim3(grid_for(size_t(count))), dim3(256), 0, st, p, count);
    return hipGetLastError();
}

}  // namespace mlp
}  // namespace qk
