// qk_surrogate.cpp - Heston price surrogate: GPU training with the project's
// own MFMA forward/backward kernels and Adam, GPU inference, a weights file,
// and pricing with plain Monte Carlo as the fallback and as the spot check of its error.
//
// Targets are price/S of European calls from the Heston full-truncation Euler
// kernel (256 steps).  Features are (log K/S, T, r, kappa, theta, sigma, rho,
// v0), affinely mapped from the training box to [-1, 1].  The model is
// heston_control(features) + network(features): training fits the network to
// target - control, prediction adds the control back.
#include "qk_surrogate.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "qk_device.hpp"
#include "qk_engine.hpp"
#include "qk_mlp.hpp"

namespace qk {

using namespace mlp;

namespace {

constexpr int kChunk = 65536;  // rows per gather + inference pass of sur_predict

struct Sur {
    bool have = false;
    DevBuf params, grads, am, av, wbf;
    // activations / gradients of one batch or chunk
    DevBuf X, Xt, tgt, Z, Zb[kHidden], H[kHidden], Ht[kHidden], H4f, dH, dZf, dZ, dZt, dy, y, sq,
        part, lossv, feat, tgts, perm;
    int64_t t = 0;  // Adam step
    int32_t spot_every = 128, spot_paths = 8192;
    double spot_k = 3.0;
    double target = 1e-3;  // RMS |price - MC expectation| / spot
    // pricing: pass A = spot checks + out-of-box queries, pass B = queries
    // whose spot check or region failed
    DevBuf sX, sy, act_a, paths_a, act_spot, act_b, status, region, closed, sur, a_p, a_se, b_p,
        b_se, judge_part, judge_stats;
    double ms_net = -1.0, ms_mc = -1.0;
    bool have_judge = false;
};

Sur g_sur;

int round_up(int x, int m) { return (x + m - 1) / m * m; }

void normalise(const double* f, float* out) {
    for (int c = 0; c < kIn; ++c) {
        const double mid = 0.5 * (kBox.lo[c] + kBox.hi[c]);
        const double half = 0.5 * (kBox.hi[c] - kBox.lo[c]);
        out[c] = float((f[c] - mid) / half);
    }
}

int32_t need_model(std::string* err) {
    if (!g_sur.have) {
        if (err) *err = "surrogate has no weights: call qk_sur_init or qk_sur_load first";
        return QK_ERR_INVALID_INPUT;
    }
    return QK_OK;
}

int32_t alloc_params(Device* d, std::string* err) {
    QK_HIP(g_sur.params.ensure(kParams * 4));
    QK_HIP(g_sur.grads.ensure(kParams * 4));
    QK_HIP(g_sur.am.ensure(kParams * 4));
    QK_HIP(g_sur.av.ensure(kParams * 4));
    QK_HIP(g_sur.wbf.ensure(size_t(kBfCount) * 2));
    (void)d;
    return QK_OK;
}

// Training buffers for batches of M rows.
int32_t alloc_work(int M, std::string* err) {
    const size_t MH = size_t(M) * kH;
    QK_HIP(g_sur.X.ensure(size_t(M) * kInPad * 2));
    QK_HIP(g_sur.Z.ensure(MH * 4));
    QK_HIP(g_sur.H4f.ensure(MH * 4));
    QK_HIP(g_sur.y.ensure(size_t(M) * 4));
    QK_HIP(g_sur.Xt.ensure(size_t(M) * kInPad * 2));
    QK_HIP(g_sur.tgt.ensure(size_t(M) * 4));
    for (int l = 0; l < kHidden; ++l) {
        QK_HIP(g_sur.H[l].ensure(MH * 2));
        QK_HIP(g_sur.Zb[l].ensure(MH * 4));
        QK_HIP(g_sur.Ht[l].ensure(MH * 2));
    }
    QK_HIP(g_sur.dH.ensure(MH * 4));
    QK_HIP(g_sur.dZf.ensure(MH * 4));
    QK_HIP(g_sur.dZ.ensure(MH * 2));
    QK_HIP(g_sur.dZt.ensure(MH * 2));
    QK_HIP(g_sur.dy.ensure(size_t(M) * 4));
    QK_HIP(g_sur.sq.ensure(size_t(M) * 4));
    QK_HIP(g_sur.part.ensure(size_t((M + 127) / 128 + 1) * kH * 4));
    return QK_OK;
}

template <typename T>
T* P(DevBuf& b) {
    return static_cast<T*>(b.p);
}

// Training forward pass of M rows already in X (M_valid of them real): keeps
// every layer's pre-activations and activations for the backward pass; y
// gets the outputs, dy and sq the loss terms.  Inference uses the fused
// kernel (infer) instead.
int32_t forward(hipStream_t st, int M, int M_valid, std::string* err) {
    (void)err;
    const float* prm = P<float>(g_sur.params);
    const __bf16* wbf = P<__bf16>(g_sur.wbf);
    const void* in = g_sur.X.p;
    for (int l = 0; l < kHidden; ++l) {
        const int K = kK[l];
        QK_HIP(gemm_nt(st, in, K, wbf + kBfWt[l], K, P<float>(g_sur.Z), kH, M, kH, K));
        QK_HIP(act_forward(st, P<float>(g_sur.Z), prm + kBOff[l], M, kH, P<float>(g_sur.Zb[l]),
                           g_sur.H[l].p, g_sur.Ht[l].p, M,
                           l == kHidden - 1 ? P<float>(g_sur.H4f) : nullptr));
        in = g_sur.H[l].p;
    }
    QK_HIP(out_forward(st, P<float>(g_sur.H4f), prm, M, M_valid, P<float>(g_sur.tgt),
                       P<float>(g_sur.y), P<float>(g_sur.dy), P<float>(g_sur.sq)));
    return QK_OK;
}

// Backward pass after a training forward: fills grads.
int32_t backward(hipStream_t st, int M, std::string* err) {
    const float* prm = P<float>(g_sur.params);
    float* G = P<float>(g_sur.grads);
    const __bf16* wbf = P<__bf16>(g_sur.wbf);
    float* part = P<float>(g_sur.part);
    QK_HIP(act_backward(st, nullptr, P<float>(g_sur.dy), prm + kW5, P<float>(g_sur.Zb[3]), M, kH,
                        P<float>(g_sur.dZf), g_sur.dZ.p, g_sur.dZt.p));
    QK_HIP(colsum(st, P<float>(g_sur.H4f), M, kH, P<float>(g_sur.dy), part, G + kW5));
    QK_HIP(colsum(st, P<float>(g_sur.dy), M, 1, nullptr, part, G + kB5));
    for (int l = kHidden - 1; l >= 0; --l) {
        const int K = kK[l];
        QK_HIP(colsum(st, P<float>(g_sur.dZf), M, kH, nullptr, part, G + kBOff[l]));
        const void* hin = l == 0 ? g_sur.Xt.p : g_sur.Ht[l - 1].p;
        QK_HIP(gemm_nt(st, g_sur.dZt.p, M, hin, M, G + kWOff[l], K, kH, K, M));
        if (l > 0) {
            QK_HIP(gemm_nt(st, g_sur.dZ.p, kH, wbf + kBfW[l], kH, P<float>(g_sur.dH), kH, M, kH,
                           kH));
            QK_HIP(act_backward(st, P<float>(g_sur.dH), nullptr, nullptr,
                                P<float>(g_sur.Zb[l - 1]), M, kH, P<float>(g_sur.dZf), g_sur.dZ.p,
                                g_sur.dZt.p));
        }
    }
    return QK_OK;
}

uint64_t fnv1a(const void* p, size_t n) {
    const unsigned char* c = static_cast<const unsigned char*>(p);
    uint64_t h = 1469598103934665603ULL;
    for (size_t i = 0; i < n; ++i) h = (h ^ c[i]) * 1099511628211ULL;
    return h;
}

const char kMagic[8] = {'Q', 'K', 'S', 'U', 'R', 'R', '0', '1'};

}  // namespace

int32_t sur_init(uint64_t seed, std::string* err) {
    Device* d = get_device(err);
    if (!d) return QK_ERR_RUNTIME_INIT;
    std::lock_guard<std::mutex> lock(d->mu);
    QK_HIP(hipSetDevice(d->dev));
    int32_t rc = alloc_params(d, err);
    if (rc != QK_OK) return rc;
    // He-normal hidden weights (std sqrt(2/fan_in)), zero biases, small output row.
    std::vector<float> p(kParams, 0.0f);
    std::mt19937_64 eng(seed);
    for (int l = 0; l < kHidden; ++l) {
        const int fan_in = l == 0 ? kIn : kH;
        std::normal_distribution<double> nd(0.0, std::sqrt(2.0 / fan_in));
        for (int o = 0; o < kH; ++o)
            for (int i = 0; i < fan_in; ++i) p[kWOff[l] + o * kK[l] + i] = float(nd(eng));
    }
    std::normal_distribution<double> nd5(0.0, std::sqrt(1.0 / kH));
    for (int i = 0; i < kH; ++i) p[kW5 + i] = float(nd5(eng));
    QK_HIP(hipMemcpyAsync(g_sur.params.p, p.data(), kParams * 4, hipMemcpyHostToDevice, d->stream));
    QK_HIP(zero_f32(d->stream, P<float>(g_sur.grads), kParams));
    QK_HIP(zero_f32(d->stream, P<float>(g_sur.am), kParams));
    QK_HIP(zero_f32(d->stream, P<float>(g_sur.av), kParams));
    QK_HIP(refresh_bf16(d->stream, P<float>(g_sur.params), g_sur.wbf.p));
    QK_HIP(hipStreamSynchronize(d->stream));
    g_sur.t = 0;
    g_sur.have = true;
    return QK_OK;
}

int32_t sur_set_params(const float* in, std::string* err) {
    Device* d = get_device(err);
    if (!d) return QK_ERR_RUNTIME_INIT;
    std::lock_guard<std::mutex> lock(d->mu);
    QK_HIP(hipSetDevice(d->dev));
    int32_t rc = alloc_params(d, err);
    if (rc != QK_OK) return rc;
    QK_HIP(hipMemcpyAsync(g_sur.params.p, in, kParams * 4, hipMemcpyHostToDevice, d->stream));
    QK_HIP(zero_f32(d->stream, P<float>(g_sur.grads), kParams));
    QK_HIP(zero_f32(d->stream, P<float>(g_sur.am), kParams));
    QK_HIP(zero_f32(d->stream, P<float>(g_sur.av), kParams));
    QK_HIP(refresh_bf16(d->stream, P<float>(g_sur.params), g_sur.wbf.p));
    QK_HIP(hipStreamSynchronize(d->stream));
    g_sur.t = 0;
    g_sur.have = true;
    return QK_OK;
}

int32_t sur_get_params(float* out, std::string* err) {
    Device* d = get_device(err);
    if (!d) return QK_ERR_RUNTIME_INIT;
    std::lock_guard<std::mutex> lock(d->mu);
    int32_t rc = need_model(err);
    if (rc != QK_OK) return rc;
    QK_HIP(hipMemcpyAsync(out, g_sur.params.p, kParams * 4, hipMemcpyDeviceToHost, d->stream));
    QK_HIP(hipStreamSynchronize(d->stream));
    return QK_OK;
}

int32_t sur_train(const double* feat, const double* target, int32_t n, int32_t epochs,
                  int32_t batch, double lr, double lr_final, uint64_t seed, double* epoch_loss,
                  std::string* err) {
    Device* d = get_device(err);
    if (!d) return QK_ERR_RUNTIME_INIT;
    std::lock_guard<std::mutex> lock(d->mu);
    QK_HIP(hipSetDevice(d->dev));
    int32_t rc = need_model(err);
    if (rc != QK_OK) return rc;
    if (batch < 64 || batch > 65536 || epochs <= 0) {
        if (err) *err = "training needs 64 <= batch <= 65536 and epochs > 0";
        return QK_ERR_INVALID_INPUT;
    }
    const int M = round_up(batch, 64);
    rc = alloc_work(M, err);
    if (rc != QK_OK) return rc;
    std::vector<float> hf(size_t(n) * kIn), ht(static_cast<size_t>(n));
    for (int32_t i = 0; i < n; ++i) {
        normalise(feat + size_t(i) * kIn, &hf[size_t(i) * kIn]);
        ht[i] = float(target[i] - heston_control(feat + size_t(i) * kIn));
    }
    QK_HIP(g_sur.feat.ensure(hf.size() * 4));
    QK_HIP(g_sur.tgts.ensure(ht.size() * 4));
    QK_HIP(g_sur.perm.ensure(size_t(n) * 4));
    const int steps_per_epoch = (n + batch - 1) / batch;
    QK_HIP(g_sur.lossv.ensure(size_t(steps_per_epoch) * 4 + 16));
    QK_HIP(hipMemcpyAsync(g_sur.feat.p, hf.data(), hf.size() * 4, hipMemcpyHostToDevice, d->stream));
    QK_HIP(hipMemcpyAsync(g_sur.tgts.p, ht.data(), ht.size() * 4, hipMemcpyHostToDevice, d->stream));
    std::vector<int32_t> perm(static_cast<size_t>(n));
    std::vector<float> lossh(static_cast<size_t>(steps_per_epoch));
    const int64_t total_steps = int64_t(epochs) * steps_per_epoch;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    std::mt19937_64 eng(seed);
    int64_t step = 0;
    for (int e = 0; e < epochs; ++e) {
        for (int32_t i = 0; i < n; ++i) perm[i] = i;
        std::shuffle(perm.begin(), perm.end(), eng);
        QK_HIP(hipMemcpyAsync(g_sur.perm.p, perm.data(), size_t(n) * 4, hipMemcpyHostToDevice,
                              d->stream));
        for (int s = 0; s < steps_per_epoch; ++s, ++step) {
            const int start = s * batch;
            const int Mv = std::min(batch, n - start);
            const int Mb = round_up(Mv, 64);
            QK_HIP(gather(d->stream, P<float>(g_sur.feat), P<float>(g_sur.tgts),
                          P<int32_t>(g_sur.perm), start, Mv, Mb, g_sur.X.p, g_sur.Xt.p,
                          P<float>(g_sur.tgt)));
            rc = forward(d->stream, Mb, Mv, err);
            if (rc != QK_OK) return rc;
            rc = backward(d->stream, Mb, err);
            if (rc != QK_OK) return rc;
            QK_HIP(colsum(d->stream, P<float>(g_sur.sq), Mb, 1, nullptr, P<float>(g_sur.part),
                          P<float>(g_sur.lossv) + s));
            const double frac = total_steps > 1 ? double(step) / double(total_steps - 1) : 1.0;
            const double lr_t = lr_final + 0.5 * (lr - lr_final) * (1.0 + std::cos(M_PI * frac));
            ++g_sur.t;
            const float bc1 = float(1.0 - std::pow(double(b1), double(g_sur.t)));
            const float bc2 = float(1.0 - std::pow(double(b2), double(g_sur.t)));
            QK_HIP(adam(d->stream, P<float>(g_sur.params), P<float>(g_sur.grads), P<float>(g_sur.am),
                        P<float>(g_sur.av), kParams, float(lr_t), b1, b2, eps, bc1, bc2));
            QK_HIP(refresh_bf16(d->stream, P<float>(g_sur.params), g_sur.wbf.p));
        }
        QK_HIP(hipMemcpyAsync(lossh.data(), g_sur.lossv.p, size_t(steps_per_epoch) * 4,
                              hipMemcpyDeviceToHost, d->stream));
        QK_HIP(hipStreamSynchronize(d->stream));
        double tot = 0.0;
        for (float v : lossh) tot += v;
        if (epoch_loss) epoch_loss[e] = tot / double(n);
    }
    return QK_OK;
}

int32_t sur_predict(const double* feat, int32_t n, double* out, std::string* err) {
    Device* d = get_device(err);
    if (!d) return QK_ERR_RUNTIME_INIT;
    std::lock_guard<std::mutex> lock(d->mu);
    QK_HIP(hipSetDevice(d->dev));
    int32_t rc = need_model(err);
    if (rc != QK_OK) return rc;
    QK_HIP(g_sur.X.ensure(size_t(kChunk) * kInPad * 2));
    QK_HIP(g_sur.y.ensure(size_t(kChunk) * 4));
    std::vector<float> hf(size_t(n) * kIn), hy(static_cast<size_t>(n));
    for (int32_t i = 0; i < n; ++i) normalise(feat + size_t(i) * kIn, &hf[size_t(i) * kIn]);
    QK_HIP(g_sur.feat.ensure(hf.size() * 4));
    QK_HIP(hipMemcpyAsync(g_sur.feat.p, hf.data(), hf.size() * 4, hipMemcpyHostToDevice, d->stream));
    for (int32_t start = 0; start < n; start += kChunk) {
        const int Mv = std::min(kChunk, n - start);
        const int Mb = round_up(Mv, 64);
        QK_HIP(gather(d->stream, P<float>(g_sur.feat), nullptr, nullptr, start, Mv, Mb, g_sur.X.p,
                      nullptr, nullptr));
        QK_HIP(infer(d->stream, g_sur.X.p, g_sur.wbf.p, P<float>(g_sur.params), P<float>(g_sur.y),
                     Mv));
        QK_HIP(hipMemcpyAsync(hy.data() + start, g_sur.y.p, size_t(Mv) * 4, hipMemcpyDeviceToHost,
                              d->stream));
    }
    QK_HIP(hipStreamSynchronize(d->stream));
    for (int32_t i = 0; i < n; ++i) out[i] = heston_control(feat + size_t(i) * kIn) + double(hy[i]);
    return QK_OK;
}

int32_t sur_gradients(const double* feat, const double* target, int32_t n, float* grads,
                      double* loss, std::string* err) {
    Device* d = get_device(err);
    if (!d) return QK_ERR_RUNTIME_INIT;
    std::lock_guard<std::mutex> lock(d->mu);
    QK_HIP(hipSetDevice(d->dev));
    int32_t rc = need_model(err);
    if (rc != QK_OK) return rc;
    if (n > 65536) {
        if (err) *err = "gradient check batch must be <= 65536 rows";
        return QK_ERR_INVALID_INPUT;
    }
    const int M = round_up(n, 64);
    rc = alloc_work(M, err);
    if (rc != QK_OK) return rc;
    std::vector<float> hf(size_t(n) * kIn), ht(static_cast<size_t>(n));
    for (int32_t i = 0; i < n; ++i) {
        normalise(feat + size_t(i) * kIn, &hf[size_t(i) * kIn]);
        ht[i] = float(target[i]);
    }
    QK_HIP(g_sur.feat.ensure(hf.size() * 4));
    QK_HIP(g_sur.tgts.ensure(ht.size() * 4));
    QK_HIP(g_sur.lossv.ensure(16));
    QK_HIP(hipMemcpyAsync(g_sur.feat.p, hf.data(), hf.size() * 4, hipMemcpyHostToDevice, d->stream));
    QK_HIP(hipMemcpyAsync(g_sur.tgts.p, ht.data(), ht.size() * 4, hipMemcpyHostToDevice, d->stream));
    QK_HIP(gather(d->stream, P<float>(g_sur.feat), P<float>(g_sur.tgts), nullptr, 0, n, M,
                  g_sur.X.p, g_sur.Xt.p, P<float>(g_sur.tgt)));
    rc = forward(d->stream, M, n, err);
    if (rc != QK_OK) return rc;
    rc = backward(d->stream, M, err);
    if (rc != QK_OK) return rc;
    QK_HIP(colsum(d->stream, P<float>(g_sur.sq), M, 1, nullptr, P<float>(g_sur.part),
                  P<float>(g_sur.lossv)));
    float lsum = 0.0f;
    QK_HIP(hipMemcpyAsync(grads, g_sur.grads.p, kParams * 4, hipMemcpyDeviceToHost, d->stream));
    QK_HIP(hipMemcpyAsync(&lsum, g_sur.lossv.p, 4, hipMemcpyDeviceToHost, d->stream));
    QK_HIP(hipStreamSynchronize(d->stream));
    if (loss) *loss = double(lsum) / double(n);
    return QK_OK;
}

int32_t sur_save(const char* path, std::string* err) {
    std::vector<float> p(kParams);
    int32_t rc = sur_get_params(p.data(), err);
    if (rc != QK_OK) return rc;
    FILE* f = std::fopen(path, "wb");
    if (!f) {
        if (err) *err = std::string("cannot open for writing: ") + path;
        return QK_ERR_INVALID_INPUT;
    }
    const uint32_t hdr[4] = {1u, uint32_t(kParams), uint32_t(kIn), uint32_t(kH)};
    const uint64_t sum = fnv1a(p.data(), p.size() * 4);
    bool ok = std::fwrite(kMagic, 1, 8, f) == 8 && std::fwrite(hdr, 4, 4, f) == 4 &&
              std::fwrite(kBox.lo, 8, kIn, f) == size_t(kIn) &&
              std::fwrite(kBox.hi, 8, kIn, f) == size_t(kIn) &&
              std::fwrite(p.data(), 4, p.size(), f) == p.size() &&
              std::fwrite(&sum, 8, 1, f) == 1;
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) {
        if (err) *err = std::string("write failed: ") + path;
        return QK_ERR_INVALID_INPUT;
    }
    return QK_OK;
}

int32_t sur_load(const char* path, std::string* err) {
    FILE* f = std::fopen(path, "rb");
    if (!f) {
        if (err) *err = std::string("cannot open weights file: ") + path;
        return QK_ERR_INVALID_INPUT;
    }
    char magic[8];
    uint32_t hdr[4];
    double lo[kIn], hi[kIn];
    std::vector<float> p(kParams);
    uint64_t sum = 0;
    bool ok = std::fread(magic, 1, 8, f) == 8 && std::memcmp(magic, kMagic, 8) == 0 &&
              std::fread(hdr, 4, 4, f) == 4 && hdr[0] == 1u && hdr[1] == uint32_t(kParams) &&
              hdr[2] == uint32_t(kIn) && hdr[3] == uint32_t(kH) &&
              std::fread(lo, 8, kIn, f) == size_t(kIn) && std::fread(hi, 8, kIn, f) == size_t(kIn) &&
              std::fread(p.data(), 4, p.size(), f) == p.size() && std::fread(&sum, 8, 1, f) == 1;
    std::fclose(f);
    if (ok) {
        ok = sum == fnv1a(p.data(), p.size() * 4);
        for (int c = 0; c < kIn; ++c) ok = ok && lo[c] == kBox.lo[c] && hi[c] == kBox.hi[c];
    }
    if (!ok) {
        if (err) *err = std::string("not a valid surrogate weights file: ") + path;
        return QK_ERR_INVALID_INPUT;
    }
    return sur_set_params(p.data(), err);
}

void sur_configure(int32_t spot_every, int32_t spot_paths, double spot_k_se) {
    g_sur.spot_every = spot_every;
    g_sur.spot_paths = spot_paths;
    g_sur.spot_k = spot_k_se;
}

int32_t sur_set_target(double target, std::string* err) {
    if (!(target > 0.0 && target < 1.0)) {
        if (err) *err = "surrogate accuracy target must lie in (0, 1) (per unit spot)";
        return QK_ERR_INVALID_INPUT;
    }
    g_sur.target = target;
    return QK_OK;
}

int32_t sur_last_judge(double* rows, int32_t max_rows, std::string* err) {
    Device* d = get_device(err);
    if (!d) return QK_ERR_RUNTIME_INIT;
    std::lock_guard<std::mutex> lock(d->mu);
    if (!g_sur.have_judge) {
        if (err) *err = "no surrogate pricing call with spot checks has run yet";
        return QK_ERR_INVALID_INPUT;
    }
    const int32_t m = std::min(max_rows, kSurRegions + 1);
    if (m <= 0) {
        if (err) *err = "max_rows must be positive";
        return QK_ERR_BAD_SIZE;
    }
    QK_HIP(hipMemcpy(rows, g_sur.judge_stats.p, size_t(m) * 4 * sizeof(double),
                     hipMemcpyDeviceToHost));
    return m;
}

bool sur_last_timing(double* ms_network, double* ms_mc) {
    if (g_sur.ms_net < 0.0) return false;
    if (ms_network) *ms_network = g_sur.ms_net;
    if (ms_mc) *ms_mc = g_sur.ms_mc;
    return true;
}

int32_t sur_heston_price(const BatchIn& in_arg, int32_t n, bool device, double* out_price,
                         int32_t* out_source, std::string* err) {
    Device* d = get_device(err);
    if (!d) return QK_ERR_RUNTIME_INIT;
    std::lock_guard<std::mutex> lock(d->mu);
    QK_HIP(hipSetDevice(d->dev));
    int32_t rc = need_model(err);
    if (rc != QK_OK) return rc;
    hipStream_t st = d->stream;

    BatchIn in = in_arg;
    const int32_t* n_dev = nullptr;
    if (!device) {
        HostBatch hb;
        hb.uniform_method = M_HESTON;
        hb.spot = in_arg.spot;
        hb.strike = in_arg.strike;
        hb.t = in_arg.t;
        hb.r = in_arg.r;
        hb.q = in_arg.q;
        hb.v0 = in_arg.v0;
        hb.kappa = in_arg.kappa;
        hb.theta = in_arg.theta;
        hb.sigma = in_arg.sigma;
        hb.rho = in_arg.rho;
        hb.ot = in_arg.ot;
        hb.paths = in_arg.paths;
        hb.steps = in_arg.steps;
        hb.seed = in_arg.seed;
        rc = stage_host_batch(d, hb, n, &in, &n_dev, err);
        if (rc != QK_OK) return rc;
    } else {
        QK_HIP(d->pin_in.ensure(64));
        QK_HIP(d->in.ensure(64));
        std::memcpy(d->pin_in.p, &n, 4);
        QK_HIP(hipMemcpyAsync(d->in.p, d->pin_in.p, 4, hipMemcpyHostToDevice, st));
        n_dev = static_cast<const int32_t*>(d->in.p);
    }
    in.uniform_method = M_HESTON;
    in.method = nullptr;
    in.method_is_model = 0;
    in.active = nullptr;
    in.paths_override = 0;

    const int Mpad = round_up(std::max(n, 1), 64);
    rc = ensure_mc_capacity(d, n, err);
    if (rc != QK_OK) return rc;
    QK_HIP(g_sur.sX.ensure(size_t(Mpad) * kInPad * 2));
    QK_HIP(g_sur.sy.ensure(size_t(Mpad) * 4));
    for (DevBuf* b : {&g_sur.act_a, &g_sur.act_spot, &g_sur.act_b, &g_sur.status, &g_sur.region})
        QK_HIP(b->ensure(size_t(Mpad)));
    QK_HIP(g_sur.paths_a.ensure(size_t(Mpad) * 4));
    for (DevBuf* b : {&g_sur.closed, &g_sur.sur, &g_sur.a_p, &g_sur.a_se, &g_sur.b_p, &g_sur.b_se})
        QK_HIP(b->ensure(size_t(Mpad) * 8));
    QK_HIP(g_sur.judge_part.ensure(size_t(sur_judge_blocks(std::max(n, 1))) * kSurRegions * 3 * 8));
    QK_HIP(g_sur.judge_stats.ensure(size_t(kSurRegions + 1) * 4 * 8));
    double* dprice = out_price;
    int32_t* dsrc = out_source;
    if (!device) {
        QK_HIP(d->out.ensure(size_t(Mpad) * 16 + 64));
        QK_HIP(d->pin_out.ensure(size_t(Mpad) * 16 + 64));
        dprice = P<double>(d->out);
        dsrc = reinterpret_cast<int32_t*>(dprice + Mpad);
    }
    uint8_t* act_a = P<uint8_t>(g_sur.act_a);
    uint8_t* act_spot = P<uint8_t>(g_sur.act_spot);
    uint8_t* act_b = P<uint8_t>(g_sur.act_b);
    int8_t* status = P<int8_t>(g_sur.status);
    int8_t* region = P<int8_t>(g_sur.region);
    hipEvent_t e0 = d->ev_sur[0], e1 = d->ev_sur[1], e3 = d->ev_sur[2];
    QK_HIP(hipEventRecord(e0, st));
    // 1. features, box test, spot-check sample, pass A's path counts
    QK_HIP(sur_prepare(st, in, n, Mpad, g_sur.spot_every, g_sur.spot_paths, g_sur.sX.p, act_a,
                       P<int32_t>(g_sur.paths_a), act_spot, act_b, status, region,
                       P<double>(g_sur.closed)));
    // 2. network: every row in one fused launch
    QK_HIP(infer(st, g_sur.sX.p, g_sur.wbf.p, P<float>(g_sur.params), P<float>(g_sur.sy), Mpad));
    QK_HIP(sur_price(st, in, n, P<float>(g_sur.sy), status, P<double>(g_sur.sur)));
    QK_HIP(hipEventRecord(e1, st));
    // 3. pass A: the spot checks (spot_paths paths) and the out-of-box
    //    queries (their own paths) in one MC pipeline
    BatchIn pa = in;
    pa.paths = P<int32_t>(g_sur.paths_a);
    pa.active = act_a;
    rc = mc_device_locked(d, pa, n, n_dev, P<double>(g_sur.a_p), P<double>(g_sur.a_se), err);
    if (rc != QK_OK) return rc;
    // 4. per-query checks, then the batch judge over all checks
    if (g_sur.spot_every > 0) {
        QK_HIP(sur_compare(st, in, n, g_sur.spot_k, 0.5 * g_sur.target, act_spot,
                           P<double>(g_sur.sur), P<double>(g_sur.a_p), P<double>(g_sur.a_se), act_b,
                           status));
        QK_HIP(sur_judge(st, in, n, act_spot, region, P<double>(g_sur.sur), P<double>(g_sur.a_p),
                         P<double>(g_sur.a_se), g_sur.target, kSurJudgeSigmas,
                         P<double>(g_sur.judge_part), P<double>(g_sur.judge_stats), act_b, status));
    }
    g_sur.have_judge = g_sur.spot_every > 0;
    // 5. pass B: plain MC for failed checks and rejected regions (their own
    //    paths); with none, its simulate launch finds no tile
    BatchIn pb = in;
    pb.active = act_b;
    rc = mc_device_locked(d, pb, n, n_dev, P<double>(g_sur.b_p), P<double>(g_sur.b_se), err);
    if (rc != QK_OK) return rc;
    QK_HIP(sur_combine(st, n, status, P<double>(g_sur.sur), P<double>(g_sur.a_p),
                       P<double>(g_sur.b_p), P<double>(g_sur.closed), dprice, dsrc));
    QK_HIP(hipEventRecord(e3, st));
    if (!device) {
        QK_HIP(hipMemcpyAsync(d->pin_out.p, d->out.p, size_t(Mpad) * 16, hipMemcpyDeviceToHost, st));
    }
    QK_HIP(hipStreamSynchronize(st));
    float a = 0.f, b = 0.f;
    if (hipEventElapsedTime(&a, e0, e1) == hipSuccess && hipEventElapsedTime(&b, e1, e3) == hipSuccess) {
        g_sur.ms_net = a;
        g_sur.ms_mc = b;
    }
    if (!device) {
        const double* hp = static_cast<const double*>(d->pin_out.p);
        std::memcpy(out_price, hp, size_t(n) * 8);
        std::memcpy(out_source, hp + Mpad, size_t(n) * 4);
    }
    return QK_OK;
}

}  // namespace qk
