// qk_lsmc.hip - Longstaff-Schwartz kernels for gfx950.
//
// Grid: (path tile, option).  A tile is kLsmcBlock lanes x kLsmcPerLane
// paths of one option.  Per chunk of options the host launches
//   k_lsmc_forward           generate paths, store S_1..S_{N-1} date-major
//                            (S[(j-1)*P + p], FP32 in production), set the
//                            discounted terminal cashflow and accumulate the
//                            normal-equation sums of date N-1;
//   k_lsmc_solve / k_lsmc_sweep, once per exercise date j = N-1 .. 1:
//     solve: per option (one wave), sum the tiles' FP64 partials in a fixed
//            order (lane-strided, then a butterfly across the wave) and solve
//            the normal equations in FP64 (Gauss-Jordan, partial pivoting,
//            |pivot| < 1e-14 -> beta = 0; skip the date with < 3 ITM paths),
//            as longstaff_schwartz.cpp:15-42, 84-125;
//     sweep: exercise at j where payoff > continuation, then discount one step
//            and accumulate the sums of date j-1 (or, at j = 1, the final
//            cashflow sums);
//   k_lsmc_final             mean cashflow times e^{-r dt}
//                            (longstaff_schwartz.cpp:127-129) and its SE.
// Options with fewer dates finish early and are skipped by later launches.
// Basis: QK_LSMC_BASIS_REFERENCE {1,S,S^2}; QK_LSMC_BASIS_CUBIC
// {1,x,x^2,x^3}, x = S/K.
#include "qk_common.hpp"
#include "qk_kernels.hpp"
#include "qk_lsmc.hpp"
#include "qk_rng.hpp"


namespace qk {

namespace {

template <int NB>
struct Layout {
    static constexpr int kA = NB * (NB + 1) / 2;  // upper triangle of X^T X
    static constexpr int kN = kA + NB;            // + X^T y
    static constexpr int kRec = kN + 2;           // + final sum, sum of squares
};

template <typename R, int NB>
__device__ __forceinline__ void basis(R s, double strike, double inv_k, double x[4]) {
    if (NB == 3) {
        const double v = double(s);
        x[0] = 1.0;
        x[1] = v;
        x[2] = v * v;
        x[3] = 0.0;
    } else {
        // FP64 validation keeps the division of the oracle; FP32 paths use 1/K
        const double u = sizeof(R) == 8 ? double(s) / strike : double(s) * inv_k;
        x[0] = 1.0;
        x[1] = u;
        x[2] = u * u;
        x[3] = u * u * u;
    }
}

template <typename R, int NB>
__device__ __forceinline__ void accumulate(double* acc, R s, double strike, double inv_k,
                                           double y) {
    double x[4];
    basis<R, NB>(s, strike, inv_k, x);
    int k = 0;
#pragma unroll
    for (int i = 0; i < NB; ++i)
#pragma unroll
        for (int j = i; j < NB; ++j) acc[k++] += x[i] * x[j];
#pragma unroll
    for (int i = 0; i < NB; ++i) acc[Layout<NB>::kA + i] += x[i] * y;
}

template <typename R, int NB>
__device__ __forceinline__ double continuation(const double* beta, R s, double strike,
                                               double inv_k) {
    if (NB == 3) {
        const double v = double(s);
        return beta[0] + beta[1] * v + beta[2] * v * v;
    }
    const double u = sizeof(R) == 8 ? double(s) / strike : double(s) * inv_k;
    return beta[0] + beta[1] * u + beta[2] * u * u + beta[3] * u * u * u;
}

template <typename R>
__device__ __forceinline__ double payoff_d(R s, double k, int32_t ot) {
    const double x = double(s);
    const double d = ot == QK_CALL ? x - k : k - x;
    return d > 0.0 ? d : 0.0;
}

__device__ __forceinline__ double wave_sum64(double v) {
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Block reduction of kRec FP64 values into out[0..kRec), fixed order.
template <int NB>
__device__ __forceinline__ void block_store(double* acc, double* out,
                                            double (*red)[Layout<NB>::kRec]) {
    constexpr int kRec = Layout<NB>::kRec;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int i = 0; i < kRec; ++i) {
        const double v = wave_sum64(acc[i]);
        if (lane == 0) red[wave][i] = v;
    }
    __syncthreads();
    if (threadIdx.x < kRec) {
        double a = 0.0;
        for (int w = 0; w < kLsmcBlock / 64; ++w) a += red[w][threadIdx.x];
        out[threadIdx.x] = a;
    }
}

__device__ __forceinline__ int lsmc_tiles(int32_t paths) {
    return (paths + kLsmcTile - 1) / kLsmcTile;
}

}  // namespace

// 16 bytes of R (4 FP32 or 2 FP64 values) loaded or stored by one
// instruction, marked non-temporal: a sweep touches each of S and the
// cashflows once, and no cache holds them until the next sweep (2.7% faster
// than cached accesses for 1024 options x 100k paths x 50 dates)
template <typename R>
struct alignas(16) LsmcVec {
    static constexpr int N = int(16 / sizeof(R));
    typedef R vt __attribute__((ext_vector_type(N)));
    R v[N] = {};
    __device__ __forceinline__ static LsmcVec load(const R* p) {
        const vt x =
            __builtin_nontemporal_load(reinterpret_cast<const vt*>(__builtin_assume_aligned(p, 16)));
        LsmcVec o;
#pragma unroll
        for (int i = 0; i < N; ++i) o.v[i] = x[i];
        return o;
    }
    __device__ __forceinline__ void store(R* p) const {
        vt x;
#pragma unroll
        for (int i = 0; i < N; ++i) x[i] = v[i];
        __builtin_nontemporal_store(x, reinterpret_cast<vt*>(__builtin_assume_aligned(p, 16)));
    }
};
// 16-byte groups whose loads are issued before the first is used (two was
// slower: 24.8 against 22.4 ms for 1024 options x 100k paths x 50 dates)
constexpr int kLsmcAheadV = 1;

template <typename R, int MODE, int NB>
__global__ __launch_bounds__(kLsmcBlock) void k_lsmc_forward(const LsmcRec* __restrict__ recs,
                                                             int tiles_max, R* __restrict__ pathbuf,
                                                             R* __restrict__ cfbuf,
                                                             const R* __restrict__ stream,
                                                             double* __restrict__ part) {
    constexpr int kRec = Layout<NB>::kRec;
    __shared__ double s_red[kLsmcBlock / 64][kRec];
    const LsmcRec& rec = recs[blockIdx.y];
    const int32_t P = rec.paths, N = rec.steps, ot = rec.ot;
    const int tile = blockIdx.x;
    if (tile >= lsmc_tiles(P)) return;
    const double strike = rec.strike, disc = rec.disc, inv_k = 1.0 / rec.strike;
    R* __restrict__ S = pathbuf + rec.path_off;
    R* __restrict__ cf = cfbuf + rec.cf_off;
    const R* __restrict__ zs = stream + rec.stream_off;
    const R spot = R(rec.spot);
    const R A = R(rec.drift * kLog2e), B = R(rec.vs * kLog2e);
    double acc[kRec];
#pragma unroll
    for (int i = 0; i < kRec; ++i) acc[i] = 0.0;
    // draws 4b .. 4b+3 of path p
    auto draws = [&](int32_t p, int32_t b, R z[4]) {
        if constexpr (MODE == SIM_PHILOX_F32) {
            const Normal4 nn = philox_normals(uint32_t(p), uint32_t(b), 0u, rec.seed_lo,
                                              rec.seed_hi);
            z[0] = nn.z0;
            z[1] = nn.z1;
            z[2] = nn.z2;
            z[3] = nn.z3;
        } else {
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int64_t d = int64_t(b) * 4 + c;
                z[c] = d < N ? zs[d * P + p] : R(0);
            }
        }
    };
    // S at the next date from S (FP64) or from the log2 increment x (FP32)
    auto advance = [&](R& s, R& x, R z) {
        if constexpr (sizeof(R) == 4) {
            x = x + __builtin_fmaf(B, z, A);
            s = spot * __builtin_amdgcn_exp2f(x);
        } else {
            s = s * exp(rec.drift + rec.vs * z);  // longstaff_schwartz.cpp:71-73
        }
    };
    // terminal cashflow and the sums of date N-1; returns the cashflow
    auto finish = [&](R s, R sprev) -> double {
        double c = payoff_d(s, strike, ot);
        if (N >= 2) {
            c *= disc;
            if (payoff_d(sprev, strike, ot) > 0.0)
                accumulate<R, NB>(acc, sprev, strike, inv_k, c);
        } else {
            acc[kRec - 2] += c;
            acc[kRec - 1] += c * c;
        }
        return c;
    };
    const int32_t t0 = tile * kLsmcTile;
    using Vec = LsmcVec<R>;
    constexpr int V = Vec::N;
    const bool aligned = ((rec.path_off | rec.cf_off | int64_t(P)) & (V - 1)) == 0;
    if (aligned && t0 + kLsmcTile <= P) {
        // full tile, rows on 16-byte boundaries: lane l simulates V
        // consecutive paths of each group of kLsmcBlock * V and stores their
        // S with one 16-byte store per date
        for (int g = 0; g < kLsmcPerLane / V; ++g) {
            const int32_t pb = t0 + (g * kLsmcBlock + int32_t(threadIdx.x)) * V;
            Vec s, x, sprev;
#pragma unroll
            for (int e = 0; e < V; ++e) {
                s.v[e] = spot;
                sprev.v[e] = spot;
            }
            for (int32_t b = 0; b * 4 < N; ++b) {
                R z[V][4];
#pragma unroll
                for (int e = 0; e < V; ++e) draws(pb + e, b, z[e]);
#pragma unroll
                for (int c = 0; c < 4; ++c) {
                    const int32_t j = b * 4 + c + 1;  // date
                    if (j > N) break;
                    sprev = s;
#pragma unroll
                    for (int e = 0; e < V; ++e) advance(s.v[e], x.v[e], z[e][c]);
                    if (j < N) s.store(S + int64_t(j - 1) * P + pb);  // no sweep reads S_N
                }
            }
            Vec out;
#pragma unroll
            for (int e = 0; e < V; ++e) out.v[e] = R(finish(s.v[e], sprev.v[e]));
            if (N >= 2) out.store(cf + pb);
        }
        block_store<NB>(acc, part + (size_t(blockIdx.y) * tiles_max + tile) * kRec, s_red);
        return;
    }
    const int32_t p0 = t0 + threadIdx.x;
    for (int k = 0; k < kLsmcPerLane; ++k) {
        const int32_t p = p0 + k * kLsmcBlock;
        if (p >= P) break;
        R s = spot, x = R(0), sprev = spot;
        for (int32_t b = 0; b * 4 < N; ++b) {
            R z[4];
            draws(p, b, z);
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int32_t j = b * 4 + c + 1;  // date
                if (j > N) break;
                sprev = s;
                advance(s, x, z[c]);
                if (j < N) S[int64_t(j - 1) * P + p] = s;  // no sweep reads S_N
            }
        }
        const double c = finish(s, sprev);
        if (N >= 2) cf[p] = R(c);
    }
    block_store<NB>(acc, part + (size_t(blockIdx.y) * tiles_max + tile) * kRec, s_red);
}

// One wave per option: the lanes sum the tiles' partials, lane 0 solves.
// beta[o*5 + 0..3] = coefficients, beta[o*5 + 4] = 1 when the date has >= 3
// ITM paths (exercise allowed).
template <int NB>
__global__ __launch_bounds__(64) void k_lsmc_solve(const LsmcRec* __restrict__ recs, int count,
                                                   int tiles_max, int it,
                                                   const double* __restrict__ part,
                                                   double* __restrict__ beta_out) {
    constexpr int kN = Layout<NB>::kN, kRec = Layout<NB>::kRec, kA = Layout<NB>::kA;
    const int o = blockIdx.x;  // one wave per option
    if (o >= count) return;
    const LsmcRec& rec = recs[o];
    if (rec.steps - 1 - it < 1) return;  // finished
    const int nt = lsmc_tiles(rec.paths);
    // lane l sums tiles l, l + 64, ... in order; the wave then adds the 64
    // lane sums in a fixed butterfly order, so the result is deterministic
    double tot[kN];
#pragma unroll
    for (int i = 0; i < kN; ++i) tot[i] = 0.0;
    const double* pr = part + size_t(o) * tiles_max * kRec;
    for (int t = threadIdx.x; t < nt; t += 64)
#pragma unroll
        for (int i = 0; i < kN; ++i) tot[i] += pr[size_t(t) * kRec + i];
#pragma unroll
    for (int i = 0; i < kN; ++i) tot[i] = wave_sum64(tot[i]);
    if (threadIdx.x != 0) return;
    double* b = beta_out + size_t(o) * 5;
    if (!(tot[0] >= 3.0)) {
        b[0] = b[1] = b[2] = b[3] = 0.0;
        b[4] = 0.0;
        return;
    }
    // augmented system, fully unrolled so it stays in registers
    double m[NB][NB + 1];
    {
        int k = 0;
#pragma unroll
        for (int r = 0; r < NB; ++r)
#pragma unroll
            for (int c = r; c < NB; ++c) {
                m[r][c] = tot[k];
                m[c][r] = tot[k];
                ++k;
            }
#pragma unroll
        for (int r = 0; r < NB; ++r) m[r][NB] = tot[kA + r];
    }
    bool singular = false;
#pragma unroll
    for (int col = 0; col < NB; ++col) {
        // partial pivot: first row of maximal |m[row][col]| among row >= col
        int piv = col;
        double best = fabs(m[col][col]);
#pragma unroll
        for (int row = col + 1; row < NB; ++row) {
            const double v = fabs(m[row][col]);
            if (v > best) {
                best = v;
                piv = row;
            }
        }
        if (best < 1e-14) singular = true;
#pragma unroll
        for (int row = col + 1; row < NB; ++row) {
            if (row == piv) {
#pragma unroll
                for (int j = 0; j <= NB; ++j) {
                    const double tmp = m[row][j];
                    m[row][j] = m[col][j];
                    m[col][j] = tmp;
                }
            }
        }
        const double d = m[col][col];
#pragma unroll
        for (int j = col; j <= NB; ++j) m[col][j] /= d;
#pragma unroll
        for (int row = 0; row < NB; ++row) {
            if (row == col) continue;
            const double f = m[row][col];
#pragma unroll
            for (int j = col; j <= NB; ++j) m[row][j] -= f * m[col][j];
        }
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) b[r] = (!singular && r < NB) ? m[r][NB] : 0.0;
    b[4] = 1.0;
}


template <typename R, int NB>
__global__ __launch_bounds__(kLsmcBlock) void k_lsmc_sweep(const LsmcRec* __restrict__ recs,
                                                           int tiles_max, int it,
                                                           const R* __restrict__ pathbuf,
                                                           R* __restrict__ cfbuf,
                                                           const double* __restrict__ beta_in,
                                                           double* __restrict__ part) {
    constexpr int kRec = Layout<NB>::kRec;
    __shared__ double s_red[kLsmcBlock / 64][kRec];
    const LsmcRec& rec = recs[blockIdx.y];
    const int32_t P = rec.paths, ot = rec.ot;
    const int32_t j = rec.steps - 1 - it;  // exercise date of this launch
    const int tile = blockIdx.x;
    if (j < 1 || tile >= lsmc_tiles(P)) return;
    const double strike = rec.strike, disc = rec.disc, inv_k = 1.0 / rec.strike;
    const double* bsrc = beta_in + size_t(blockIdx.y) * 5;
    const double beta[4] = {bsrc[0], bsrc[1], bsrc[2], bsrc[3]};
    const bool do_ex = bsrc[4] != 0.0;
    const R* __restrict__ Sj = pathbuf + rec.path_off + int64_t(j - 1) * P;
    const R* __restrict__ Sm = pathbuf + rec.path_off + int64_t(j >= 2 ? j - 2 : 0) * P;
    R* __restrict__ cf = cfbuf + rec.cf_off;
    double acc[kRec];
#pragma unroll
    for (int i = 0; i < kRec; ++i) acc[i] = 0.0;
    // one path: exercise at j, discount, and the sums of date j - 1; returns
    // the path's new cashflow (stored by the caller for j >= 2)
    auto path_c = [&](R cv, R s, R sp) -> double {
        double c = double(cv);
        if (do_ex) {
            const double ex = payoff_d(s, strike, ot);
            if (ex > 0.0 && ex > continuation<R, NB>(beta, s, strike, inv_k)) c = ex;
        }
        if (j >= 2) {
            c *= disc;
            if (payoff_d(sp, strike, ot) > 0.0) accumulate<R, NB>(acc, sp, strike, inv_k, c);
        } else {
            acc[kRec - 2] += c;
            acc[kRec - 1] += c * c;
        }
        return c;
    };
    auto path = [&](int32_t p, R cv, R s, R sp) {
        const double c = path_c(cv, s, sp);
        if (j >= 2) cf[p] = R(c);
    };
    // Full tiles of options whose rows start on 16-byte boundaries read and
    // write 16 bytes per lane: lane l takes V consecutive paths of each of
    // the tile's kLsmcPerLane / V groups of kLsmcBlock * V paths, so one load
    // instruction moves 4x (FP32) the bytes of a per-path load.
    using Vec = LsmcVec<R>;
    constexpr int V = Vec::N;
    const int32_t t0 = tile * kLsmcTile;
    const bool aligned = ((rec.path_off | rec.cf_off | int64_t(P)) & (V - 1)) == 0;
    if (aligned && t0 + kLsmcTile <= P) {
        constexpr int kGroups = kLsmcPerLane / V;
        for (int g0 = 0; g0 < kGroups; g0 += kLsmcAheadV) {
            Vec cv[kLsmcAheadV], sj[kLsmcAheadV], sm[kLsmcAheadV];
#pragma unroll
            for (int k = 0; k < kLsmcAheadV; ++k) {
                const int32_t pb = t0 + ((g0 + k) * kLsmcBlock + int32_t(threadIdx.x)) * V;
                cv[k] = Vec::load(cf + pb);
                sj[k] = do_ex ? Vec::load(Sj + pb) : Vec{};
                sm[k] = j >= 2 ? Vec::load(Sm + pb) : Vec{};
            }
#pragma unroll
            for (int k = 0; k < kLsmcAheadV; ++k) {
                Vec out;
#pragma unroll
                for (int e = 0; e < V; ++e) out.v[e] = R(path_c(cv[k].v[e], sj[k].v[e], sm[k].v[e]));
                if (j >= 2) out.store(cf + t0 + ((g0 + k) * kLsmcBlock + int32_t(threadIdx.x)) * V);
            }
        }
        block_store<NB>(acc, part + (size_t(blockIdx.y) * tiles_max + tile) * kRec, s_red);
        return;
    }
    const int32_t p0 = t0 + threadIdx.x;
    if (p0 + (kLsmcPerLane - 1) * kLsmcBlock < P) {
        // a full tile: the loads of kAhead paths are issued before the first
        // of them is used, so more bytes are in flight per wave
        constexpr int kAhead = 2;
        for (int k0 = 0; k0 < kLsmcPerLane; k0 += kAhead) {
            R cv[kAhead], sj[kAhead], sm[kAhead];
#pragma unroll
            for (int k = 0; k < kAhead; ++k) {
                const int32_t p = p0 + (k0 + k) * kLsmcBlock;
                cv[k] = cf[p];
                sj[k] = do_ex ? Sj[p] : R(0);
                sm[k] = j >= 2 ? Sm[p] : R(0);
            }
#pragma unroll
            for (int k = 0; k < kAhead; ++k)
                path(p0 + (k0 + k) * kLsmcBlock, cv[k], sj[k], sm[k]);
        }
    } else {
        for (int k = 0; k < kLsmcPerLane; ++k) {
            const int32_t p = p0 + k * kLsmcBlock;
            if (p >= P) break;
            path(p, cf[p], do_ex ? Sj[p] : R(0), j >= 2 ? Sm[p] : R(0));
        }
    }
    block_store<NB>(acc, part + (size_t(blockIdx.y) * tiles_max + tile) * kRec, s_red);
}

template <int NB>
__global__ __launch_bounds__(64) void k_lsmc_final(const LsmcRec* __restrict__ recs, int count,
                                                   int tiles_max, const double* __restrict__ part,
                                                   double* __restrict__ out_price,
                                                   double* __restrict__ out_se) {
    constexpr int kRec = Layout<NB>::kRec;
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= count) return;
    const LsmcRec& rec = recs[o];
    const int nt = lsmc_tiles(rec.paths);
    const double* pr = part + size_t(o) * tiles_max * kRec;
    double a = 0.0, b = 0.0;
    for (int t = 0; t < nt; ++t) {
        a += pr[size_t(t) * kRec + kRec - 2];
        b += pr[size_t(t) * kRec + kRec - 1];
    }
    const double np = double(rec.paths);
    const double mean = a / np;
    double var = np > 1.0 ? (b - a * mean) / (np - 1.0) : 0.0;
    var = var > 0.0 ? var : 0.0;
    out_price[rec.out_index] = mean * rec.disc;
    out_se[rec.out_index] = rec.disc * sqrt(var / np);
}

int lsmc_record_doubles(int32_t basis) {
    return basis == QK_LSMC_BASIS_CUBIC ? Layout<4>::kRec : Layout<3>::kRec;
}

template <typename R, int MODE, int NB>
static hipError_t lsmc_run(hipStream_t st, const LsmcRec* recs, int32_t count, int32_t tiles_max,
                           int32_t max_steps, void* pathbuf, void* cfbuf, const void* stream,
                           double* part, double* beta, double* out_price, double* out_se) {
    const dim3 grid(tiles_max, count), block(kLsmcBlock);
    const dim3 small((count + 63) / 64), b64(64);
    hipLaunchKernelGGL((k_lsmc_forward<R, MODE, NB>), grid, block, 0, st, recs, tiles_max,
                       static_cast<R*>(pathbuf), static_cast<R*>(cfbuf),
                       static_cast<const R*>(stream), part);
    for (int it = 0; it < max_steps - 1; ++it) {
        hipLaunchKernelGGL((k_lsmc_solve<NB>), dim3(count), b64, 0, st, recs, count, tiles_max, it,
                           part, beta);
        hipLaunchKernelGGL((k_lsmc_sweep<R, NB>), grid, block, 0, st, recs, tiles_max, it,
                           static_cast<const R*>(pathbuf), static_cast<R*>(cfbuf), beta, part);
    }
    hipLaunchKernelGGL((k_lsmc_final<NB>), small, b64, 0, st, recs, count, tiles_max, part,
                       out_price, out_se);
    return hipGetLastError();
}

hipError_t launch_lsmc(hipStream_t st, int32_t mode, int32_t basis, const LsmcRec* recs,
                       int32_t count, int32_t tiles_max, int32_t max_steps, void* pathbuf,
                       void* cfbuf, const void* stream, double* part, double* beta,
                       double* out_price, double* out_se) {
    const bool cubic = basis == QK_LSMC_BASIS_CUBIC;
#define QK_LSMC_RUN(R, M, NB)                                                                   \
    return lsmc_run<R, M, NB>(st, recs, count, tiles_max, max_steps, pathbuf, cfbuf, stream, part, \
                              beta, out_price, out_se)
    if (mode == SIM_PHILOX_F32) {
        if (cubic) QK_LSMC_RUN(float, SIM_PHILOX_F32, 4);
        QK_LSMC_RUN(float, SIM_PHILOX_F32, 3);
    } else if (mode == SIM_STREAM_F64) {
        if (cubic) QK_LSMC_RUN(double, SIM_STREAM_F64, 4);
        QK_LSMC_RUN(double, SIM_STREAM_F64, 3);
    }
    if (cubic) QK_LSMC_RUN(float, SIM_STREAM_F32, 4);
    QK_LSMC_RUN(float, SIM_STREAM_F32, 3);
#undef QK_LSMC_RUN
}

}  // namespace qk
