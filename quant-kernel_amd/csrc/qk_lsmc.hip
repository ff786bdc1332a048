// qk_lsmc.hip - Longstaff-Schwartz kernel for gfx950.
//
// One 1024-thread workgroup per option (longstaff_schwartz.cpp:46-130):
//   forward : each lane generates whole paths and stores S_1..S_N date-major
//             (S[(j-1)*P + p]) so every later sweep reads coalesced rows;
//   backward: for j = N-1..1 the ITM regression sums are accumulated per lane
//             in FP64, reduced over the wave64 and then the 16 waves in a
//             fixed order, and one lane solves the normal equations in FP64
//             (Gauss-Jordan, partial pivoting, |pivot| < 1e-14 -> beta = 0,
//             longstaff_schwartz.cpp:15-42); the next sweep applies the
//             exercise rule at j and discounts and accumulates the sums of
//             j-1 in the same pass over the cashflows;
//   final   : mean of the cashflows times e^{-r dt}, and its standard error.
// Basis: QK_LSMC_BASIS_REFERENCE {1,S,S^2}; QK_LSMC_BASIS_CUBIC
// {1,x,x^2,x^3}, x = S/K.
#include "qk_common.hpp"
#include "qk_kernels.hpp"
#include "qk_lsmc.hpp"
#include "qk_rng.hpp"

namespace qk {

namespace {

constexpr int kLsmcBlock = 1024;

template <int NB>
struct Sums {
    // upper triangle of X^T X (row-major) then X^T y
    static constexpr int kA = NB * (NB + 1) / 2;
    static constexpr int kN = kA + NB;
    double v[kN];
};

template <int NB>
__device__ __forceinline__ void basis(double s, double strike, double x[4]) {
    if (NB == 3) {
        x[0] = 1.0;
        x[1] = s;
        x[2] = s * s;
        x[3] = 0.0;
    } else {
        const double u = s / strike;
        x[0] = 1.0;
        x[1] = u;
        x[2] = u * u;
        x[3] = u * u * u;
    }
}

template <int NB>
__device__ __forceinline__ void accumulate(Sums<NB>& acc, double s, double strike, double y) {
    double x[4];
    basis<NB>(s, strike, x);
    int k = 0;
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
        for (int j = i; j < NB; ++j) acc.v[k++] += x[i] * x[j];
#pragma unroll
    for (int i = 0; i < NB; ++i) acc.v[Sums<NB>::kA + i] += x[i] * y;
}

template <int NB>
__device__ __forceinline__ double continuation(const double* beta, double s, double strike) {
    if (NB == 3) return beta[0] + beta[1] * s + beta[2] * s * s;
    const double u = s / strike;
    return beta[0] + beta[1] * u + beta[2] * u * u + beta[3] * u * u * u;
}

__device__ __forceinline__ double wave_sum64(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Gauss-Jordan on the augmented NB x (NB+1) system (one lane).
template <int NB>
__device__ void solve(double m[4][5], double beta[4]) {
    for (int col = 0; col < NB; ++col) {
        int piv = col;
        for (int row = col + 1; row < NB; ++row)
            if (fabs(m[row][col]) > fabs(m[piv][col])) piv = row;
        if (fabs(m[piv][col]) < 1e-14) {
            for (int i = 0; i < NB; ++i) beta[i] = 0.0;
            return;
        }
        if (piv != col)
            for (int j = 0; j <= NB; ++j) {
                const double tmp = m[piv][j];
                m[piv][j] = m[col][j];
                m[col][j] = tmp;
            }
        const double d = m[col][col];
        for (int j = col; j <= NB; ++j) m[col][j] /= d;
        for (int row = 0; row < NB; ++row) {
            if (row == col) continue;
            const double f = m[row][col];
            for (int j = col; j <= NB; ++j) m[row][j] -= f * m[col][j];
        }
    }
    for (int i = 0; i < NB; ++i) beta[i] = m[i][NB];
}

template <typename R>
__device__ __forceinline__ double payoff_d(R s, double k, int32_t ot) {
    const double x = double(s);
    const double d = ot == QK_CALL ? x - k : k - x;
    return d > 0.0 ? d : 0.0;
}

}  // namespace

template <typename R, int MODE, int NB>
__global__ __launch_bounds__(kLsmcBlock) void k_lsmc(const LsmcRec* __restrict__ recs,
                                                     R* __restrict__ pathbuf,
                                                     R* __restrict__ cfbuf,
                                                     const R* __restrict__ stream,
                                                     double* __restrict__ out_price,
                                                     double* __restrict__ out_se) {
    constexpr int kN = Sums<NB>::kN;
    constexpr int kWaves = kLsmcBlock / 64;
    __shared__ double s_red[kWaves][kN];
    __shared__ double s_beta[4];
    __shared__ int s_do;

    const LsmcRec& rec = recs[blockIdx.x];
    const int32_t P = rec.paths, N = rec.steps, ot = rec.ot;
    const double strike = rec.strike, disc = rec.disc;
    R* __restrict__ S = pathbuf + rec.path_off;
    R* __restrict__ cf = cfbuf + rec.cf_off;
    const R* __restrict__ zs = stream + rec.stream_off;
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;

    // block reduction of kN FP64 sums; lane 0 of the block gets the totals
    auto reduce = [&](Sums<NB>& acc, double* tot) {
#pragma unroll
        for (int i = 0; i < kN; ++i) {
            const double v = wave_sum64(acc.v[i]);
            if (lane == 0) s_red[wave][i] = v;
        }
        __syncthreads();
        if (tid == 0) {
            for (int i = 0; i < kN; ++i) {
                double a = 0.0;
                for (int w = 0; w < kWaves; ++w) a += s_red[w][i];
                tot[i] = a;
            }
        }
        __syncthreads();
    };

    Sums<NB> acc;
#pragma unroll
    for (int i = 0; i < kN; ++i) acc.v[i] = 0.0;
    double f1 = 0.0, f2 = 0.0;  // final cashflow sums (N == 1 case)

    // ---- forward: generate, store, and start date N-1 ----
    {
        const R spot = R(rec.spot);
        const R A = R(rec.drift * kLog2e), B = R(rec.vs * kLog2e);
        const uint32_t k0 = rec.seed_lo, k1 = rec.seed_hi;
        for (int32_t p = tid; p < P; p += kLsmcBlock) {
            R s = spot, x = R(0);
            for (int32_t b = 0; b * 4 < N; ++b) {
                R z[4];
                if constexpr (MODE == SIM_PHILOX_F32) {
                    const Normal4 nn = philox_normals(uint32_t(p), uint32_t(b), 0u, k0, k1);
                    z[0] = nn.z0; z[1] = nn.z1; z[2] = nn.z2; z[3] = nn.z3;
                } else {
#pragma unroll
                    for (int c = 0; c < 4; ++c) {
                        const int64_t d = int64_t(b) * 4 + c;
                        z[c] = d < N ? zs[d * P + p] : R(0);
                    }
                }
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const int32_t j = b * 4 + c + 1;  // date
                    if (j > N) break;
                    if constexpr (sizeof(R) == 4) {
                        x = x + __builtin_fmaf(B, z[c], A);
                        s = spot * __builtin_amdgcn_exp2f(x);
                    } else {
                        s = s * exp(rec.drift + rec.vs * z[c]);  // longstaff_schwartz.cpp:71-73
                    }
                    S[int64_t(j - 1) * P + p] = s;
                }
            }
            double c = payoff_d(s, strike, ot);
            if (N >= 2) {
                c *= disc;
                const R sp = S[int64_t(N - 2) * P + p];
                if (payoff_d(sp, strike, ot) > 0.0) accumulate<NB>(acc, double(sp), strike, c);
                cf[p] = R(c);
            } else {
                f1 += c;
                f2 += c * c;
            }
        }
    }

    // ---- backward induction ----
    double tot[Sums<NB>::kN];
    for (int32_t j = N - 1; j >= 1; --j) {
        reduce(acc, tot);
        if (tid == 0) {
            const double count = tot[0];
            s_do = count >= 3.0 ? 1 : 0;
            if (s_do) {
                double m[4][5];
                int k = 0;
                for (int a = 0; a < NB; ++a)
                    for (int b2 = a; b2 < NB; ++b2) {
                        m[a][b2] = tot[k];
                        m[b2][a] = tot[k];
                        ++k;
                    }
                for (int a = 0; a < NB; ++a) m[a][NB] = tot[Sums<NB>::kA + a];
                double beta[4] = {0, 0, 0, 0};
                solve<NB>(m, beta);
                for (int a = 0; a < 4; ++a) s_beta[a] = beta[a];
            }
        }
        __syncthreads();
        const bool do_ex = s_do != 0;
        double beta[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) beta[a] = s_beta[a];
#pragma unroll
        for (int i = 0; i < kN; ++i) acc.v[i] = 0.0;
        const R* Sj = S + int64_t(j - 1) * P;
        const R* Sm = S + int64_t(j >= 2 ? j - 2 : 0) * P;
        for (int32_t p = tid; p < P; p += kLsmcBlock) {
            double c = double(cf[p]);
            if (do_ex) {
                const R s = Sj[p];
                const double ex = payoff_d(s, strike, ot);
                if (ex > 0.0 && ex > continuation<NB>(beta, double(s), strike)) c = ex;
            }
            if (j >= 2) {
                c *= disc;
                const R sp = Sm[p];
                if (payoff_d(sp, strike, ot) > 0.0) accumulate<NB>(acc, double(sp), strike, c);
                cf[p] = R(c);
            } else {
                f1 += c;
                f2 += c * c;
            }
        }
        __syncthreads();  // s_beta / s_do are rewritten by the next date
    }

    // ---- final mean and standard error ----
    f1 = wave_sum64(f1);
    f2 = wave_sum64(f2);
    if (lane == 0) {
        s_red[wave][0] = f1;
        s_red[wave][1] = f2;
    }
    __syncthreads();
    if (tid == 0) {
        double a = 0.0, b = 0.0;
        for (int w = 0; w < kWaves; ++w) {
            a += s_red[w][0];
            b += s_red[w][1];
        }
        const double np = double(P);
        const double mean = a / np;
        double var = np > 1.0 ? (b - a * mean) / (np - 1.0) : 0.0;
        var = var > 0.0 ? var : 0.0;
        out_price[rec.out_index] = mean * disc;  // longstaff_schwartz.cpp:127-129
        out_se[rec.out_index] = disc * sqrt(var / np);
    }
}

hipError_t launch_lsmc(hipStream_t st, int32_t mode, int32_t basis, const LsmcRec* recs,
                       int32_t count, void* pathbuf, void* cfbuf, const void* stream,
                       double* out_price, double* out_se) {
    const dim3 grid(count), block(kLsmcBlock);
#define QK_LSMC_LAUNCH(R, M, NB)                                                               \
    hipLaunchKernelGGL((k_lsmc<R, M, NB>), grid, block, 0, st, recs, static_cast<R*>(pathbuf), \
                       static_cast<R*>(cfbuf), static_cast<const R*>(stream), out_price, out_se)
    const bool cubic = basis == QK_LSMC_BASIS_CUBIC;
    if (mode == SIM_PHILOX_F32) {
        if (cubic) QK_LSMC_LAUNCH(float, SIM_PHILOX_F32, 4);
        else QK_LSMC_LAUNCH(float, SIM_PHILOX_F32, 3);
    } else if (mode == SIM_STREAM_F64) {
        if (cubic) QK_LSMC_LAUNCH(double, SIM_STREAM_F64, 4);
        else QK_LSMC_LAUNCH(double, SIM_STREAM_F64, 3);
    } else {
        if (cubic) QK_LSMC_LAUNCH(float, SIM_STREAM_F32, 4);
        else QK_LSMC_LAUNCH(float, SIM_STREAM_F32, 3);
    }
#undef QK_LSMC_LAUNCH
    return hipGetLastError();
}

}  // namespace qk
