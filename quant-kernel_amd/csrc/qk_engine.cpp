// qk_engine.cpp - host engine of the MI355X Monte Carlo batch-pricing path.
//
// One lazily created device context per process (std::call_once): the HIP
// device, one default-priority stream, timing events and grow-only device
// and pinned host buffers.  Device work is serialised by a mutex; the C ABI
// may be entered from several host threads.
#include "qk_engine.hpp"
#include "qk_device.hpp"
#include "qk_lsmc.hpp"
#include "qk_cf.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <random>
#include <fstream>
#include <sstream>
#include <thread>
#include <vector>

namespace qk {

namespace {

std::atomic<int32_t> g_rng_mode{QK_RNG_PHILOX};

// Fixed device scratch of the allocation-free MC entry point
// (run_mc_batch_device_noalloc): option records, per-option tile counts and
// offsets, the pipeline's counters and the tile partials, each 256-byte
// aligned.  Every device the library runs on holds its own copy until the
// process ends.  In host code the array's name is not a device address: the
// address comes from hipGetSymbolAddress in init_device, once per device.
constexpr size_t kNaRecsOff = 0;
constexpr size_t kNaNtilesOff = kNaRecsOff + size_t(65536) * sizeof(OptionRec);
constexpr size_t kNaStartOff = kNaNtilesOff + size_t(65536) * 4;
constexpr size_t kNaCountersOff = kNaStartOff + size_t(65536 + 64) * 4;
constexpr size_t kNaPartialsOff = kNaCountersOff + 256;
constexpr size_t kNaBytes = kNaPartialsOff + size_t(kNoAllocMaxTiles) * sizeof(Partial);
static_assert(kNaNtilesOff % 256 == 0 && kNaStartOff % 256 == 0 && kNaCountersOff % 256 == 0 &&
                  kNaPartialsOff % 256 == 0,
              "scratch sections are 256-byte aligned");
static_assert(kNoAllocMaxOptions < 65536, "options beyond the record section");

static __device__ __align__(256) char scratch[kNaBytes];

Device* g_dev = nullptr;
}  // namespace
void apply_cached_shapes(Device* d);
namespace {
std::string g_init_err;
std::once_flag g_once;

void init_device() {
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0) {
        g_init_err = std::string("no HIP device available (") +
                     (e != hipSuccess ? hipGetErrorString(e) : "0 devices") + ")";
        return;
    }
    const int want = env_int("QK_DEVICE", 0);
    if (want < 0 || want >= count) {
        g_init_err = "QK_DEVICE out of range";
        return;
    }
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, want);
    if (e != hipSuccess) {
        g_init_err = std::string("hipGetDeviceProperties: ") + hipGetErrorString(e);
        return;
    }
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
        g_init_err = std::string("device ") + prop.gcnArchName +
                     " is not gfx950 (this library is built for MI355X only)";
        return;
    }
    e = hipSetDevice(want);
    if (e != hipSuccess) {
        g_init_err = std::string("hipSetDevice: ") + hipGetErrorString(e);
        return;
    }
    Device* d = new Device();
    d->dev = want;
    d->cus = prop.multiProcessorCount;
    d->name = prop.name[0] ? prop.name : prop.gcnArchName;  // some runtimes leave the marketing name empty
    e = hipStreamCreateWithFlags(&d->stream, hipStreamNonBlocking);
    for (int i = 0; e == hipSuccess && i < 4; ++i) e = hipEventCreate(&d->ev[i]);
    for (int i = 0; e == hipSuccess && i < 3; ++i) e = hipEventCreate(&d->ev_sur[i]);
    if (e != hipSuccess) {
        g_init_err = std::string("stream/event creation: ") + hipGetErrorString(e);
        delete d;
        return;
    }
    e = upload_trig_table();
    if (e != hipSuccess) {
        g_init_err = std::string("angle table upload: ") + hipGetErrorString(e);
        delete d;
        return;
    }
    // this device's copy of the allocation-free entry point's scratch
    void* sp = nullptr;
    e = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
    if (e == hipSuccess && sp) {
        d->noalloc_scratch = static_cast<char*>(sp);
    } else {
        (void)hipGetLastError();
        d->noalloc_err = std::string("hipGetSymbolAddress(scratch): ") +
                         (e != hipSuccess ? hipGetErrorString(e) : "null address");
    }
    const int per_cu = std::max(1, env_int("QK_GRID_PER_CU", 8));
    d->shape_step.lane_units = std::max(1, env_int("QK_LANE_PATHS", 8));
    d->shape_step.grid = d->cus * per_cu;
    // 25 groups (200 antithetic pairs) per lane: the autotuner's pick for
    // 4096 x 100k-path options (one 51,200-pair tile covers an option's
    // 50,000 pairs), where the per-tile fetch and lookup otherwise weigh on
    // tiles of a few microseconds
    d->shape_term.lane_units = std::max(1, env_int("QK_LANE_GROUPS", 25));
    d->shape_term.grid = d->cus * per_cu;
    if (env_int("QK_AUTOTUNE", 0)) apply_cached_shapes(d);
    g_dev = d;
}

}  // namespace

Device* get_device(std::string* err) {
    std::call_once(g_once, init_device);
    if (!g_dev && err) *err = g_init_err;
    return g_dev;
}

namespace {

int32_t method_of(const HostBatch& hb, int32_t i) {
    if (!hb.model_ids) return hb.uniform_method;
    switch (hb.model_ids[i]) {
        case QK_MODEL_GBM_EULER: return M_EULER;
        case QK_MODEL_GBM_MILSTEIN: return M_MILSTEIN;
        case QK_MODEL_HESTON_EULER: return M_HESTON;
        case QK_MODEL_LOCAL_VOL: return M_LOCAL_VOL;
        default: return -1;
    }
}

bool is_terminal_method(int32_t m) { return method_is_terminal(m); }

size_t align64(size_t x) { return (x + 63) & ~size_t(63); }

// Device address of byte `offset` of a pinned host allocation (the pointer
// hipHostMalloc returned), or null.
int32_t* host_flags_ptr(void* pinned, size_t offset) {
    void* dp = nullptr;
    if (hipHostGetDevicePointer(&dp, pinned, 0) != hipSuccess || !dp) return nullptr;
    return reinterpret_cast<int32_t*>(static_cast<char*>(dp) + offset);
}

// Launch prepare -> scan -> simulate -> finalize on d's stream.  The caller
// holds d->mu.  `family` selects the simulation kernel.  With n_dev null the
// batch size is n, which prepare stores in counters[8] for the later kernels.
// host_flags (device address of pinned host memory, or null) receives the
// scan's counters[0..3] from the finalizer.
// Work buffers of one pipeline run: the context's grow-only buffers, or the
// fixed device scratch of the allocation-free entry point.
struct PipeBufs {
    OptionRec* recs;
    int32_t* ntiles;
    int32_t* tile_start;
    int32_t* counters;  // at least 9 ints
    void* scan;         // block sums of the multi-pass scan (n > 65535 only)
    Partial* partials;
};

PipeBufs context_bufs(Device* d) {
    return PipeBufs{static_cast<OptionRec*>(d->recs.p), static_cast<int32_t*>(d->ntiles.p),
                    static_cast<int32_t*>(d->tile_start.p), static_cast<int32_t*>(d->counters.p),
                    d->scan.p, static_cast<Partial*>(d->partials.p)};
}

// A terminal batch of one method that the pair-only kernel covers.
bool pair_only_batch(const BatchIn& in, int32_t family) {
    return family == FAM_TERMINAL && !in.method &&
           (in.uniform_method == M_STANDARD || in.uniform_method == M_ANTITHETIC);
}

int32_t launch_pipeline(Device* d, const PipeBufs& bufs, const BatchIn& in, int32_t n,
                        const int32_t* n_dev, const int64_t* stream_off, int32_t family,
                        int32_t mode, const void* stream, double* out_price, double* out_se,
                        std::string* err, bool timed = true, int32_t capacity = 0x7fffffff,
                        int32_t* host_flags = nullptr) {
    OptionRec* recs = bufs.recs;
    int32_t* ntiles = bufs.ntiles;
    int32_t* tstart = bufs.tile_start;
    int32_t* ctr = bufs.counters;
    Partial* parts = bufs.partials;
    const SimShape& shp = family == FAM_TERMINAL ? d->shape_term : d->shape_step;
    int32_t* n_slot = n_dev ? nullptr : ctr + 8;
    if (!n_dev) n_dev = n_slot;
    if (timed) QK_HIP(hipEventRecord(d->ev[0], d->stream));
    QK_HIP(launch_prepare(d->stream, in, n_slot ? nullptr : n_dev, n, n_slot, stream_off,
                          tile_groups_terminal(d->shape_term), tile_units_stepwise(d->shape_step),
                          recs, ntiles));
    QK_HIP(launch_scan(d->stream, n, n_dev, ntiles, recs, tstart, ctr, bufs.scan, capacity));
    if (timed) QK_HIP(hipEventRecord(d->ev[1], d->stream));
    QK_HIP(launch_simulate(d->stream, family, mode, shp, recs, tstart, n_dev, ctr, parts, stream,
                           pair_only_batch(in, family)));
    if (timed) QK_HIP(hipEventRecord(d->ev[2], d->stream));
    QK_HIP(launch_finalize(d->stream, n_dev, recs, parts, ctr, out_price, out_se, host_flags));
    if (timed) QK_HIP(hipEventRecord(d->ev[3], d->stream));
    return QK_OK;
}

void record_timing(Device* d) {
    float a = 0.f, b = 0.f;
    if (hipEventElapsedTime(&a, d->ev[0], d->ev[3]) == hipSuccess &&
        hipEventElapsedTime(&b, d->ev[1], d->ev[2]) == hipSuccess) {
        d->ms_total = a;
        d->ms_sim = b;
    }
}

// A replayed launch graph is timed as a whole (events recorded around it on
// the stream); its simulate kernel is not timed separately.
void record_timing_total(Device* d) {
    float a = 0.f;
    if (hipEventElapsedTime(&a, d->ev[0], d->ev[3]) == hipSuccess) {
        d->ms_total = a;
        d->ms_sim = std::nan("");
    }
}

// Tile-count bound from the raw path counts (valid elements can only need
// fewer tiles than this).
int64_t tile_bound(const HostBatch& hb, int32_t n, const Device* d) {
    const int64_t tu_s = tile_units_stepwise(d->shape_step);
    const int64_t tu_t = int64_t(tile_groups_terminal(d->shape_term)) * 4;  // smallest group
    int64_t total = 0;
    for (int32_t i = 0; i < n; ++i) {
        const int64_t p = std::max<int32_t>(hb.paths[i], 0);
        const int32_t m = method_of(hb, i);
        if (m < 0) continue;
        if (is_terminal_method(m)) total += (p + tu_t - 1) / tu_t + 1;
        else total += (p + tu_s - 1) / tu_s + 1;
    }
    return total;
}

}  // namespace

void set_rng_mode(int32_t mode) { g_rng_mode.store(mode); }
int32_t get_rng_mode() { return g_rng_mode.load(); }

void host_normal_stream(uint64_t seed, int64_t count, double* out) {
    std::mt19937_64 eng(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    for (int64_t i = 0; i < count; ++i) out[i] = nd(eng);
}

int32_t philox_probe(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c3, int32_t count,
                     uint32_t* words, float* normals, std::string* err) {
    Device* d = get_device(err);
    if (!d) return QK_ERR_RUNTIME_INIT;
    std::lock_guard<std::mutex> lock(d->mu);
    QK_HIP(hipSetDevice(d->dev));
    const size_t bytes = size_t(count) * 16;
    QK_HIP(d->out.ensure(2 * bytes));
    QK_HIP(d->pin_out.ensure(2 * bytes));
    uint32_t* dw = static_cast<uint32_t*>(d->out.p);
    float* dn = reinterpret_cast<float*>(static_cast<char*>(d->out.p) + bytes);
    QK_HIP(launch_philox_probe(d->stream, seed, c0, c1, c3, count, dw, dn));
    QK_HIP(hipMemcpyAsync(d->pin_out.p, d->out.p, 2 * bytes, hipMemcpyDeviceToHost, d->stream));
    QK_HIP(hipStreamSynchronize(d->stream));
    std::memcpy(words, d->pin_out.p, bytes);
    std::memcpy(normals, static_cast<char*>(d->pin_out.p) + bytes, bytes);
    return QK_OK;
}

int32_t philox_probe8(uint64_t seed, uint32_t c0, uint32_t c1, uint32_t c3, int32_t count,
                      float* normals, std::string* err) {
    Device* d = get_device(err);
    if (!d) return QK_ERR_RUNTIME_INIT;
    std::lock_guard<std::mutex> lock(d->mu);
    QK_HIP(hipSetDevice(d->dev));
    const size_t bytes = size_t(count) * 32;
    QK_HIP(d->out.ensure(bytes));
    QK_HIP(d->pin_out.ensure(bytes));
    QK_HIP(launch_philox_probe8(d->stream, seed, c0, c1, c3, count, static_cast<float*>(d->out.p)));
    QK_HIP(hipMemcpyAsync(d->pin_out.p, d->out.p, bytes, hipMemcpyDeviceToHost, d->stream));
    QK_HIP(hipStreamSynchronize(d->stream));
    std::memcpy(normals, d->pin_out.p, bytes);
    return QK_OK;
}

bool last_timing(double* ms_total, double* ms_simulate) {
    if (!g_dev || g_dev->ms_total < 0.0) return false;
    if (ms_total) *ms_total = g_dev->ms_total;
    if (ms_simulate) *ms_simulate = g_dev->ms_sim;
    return true;
}

bool device_info(std::string* name, int32_t* cus, std::string* err) {
    Device* d = get_device(err);
    if (!d) return false;
    if (name) *name = d->name;
    if (cus) *cus = d->cus;
    return true;
}

// Builds the validation-mode stream: per option the host mt19937_64 sequence
// in the reference's draw order (paths outer, draws inner; mc_engine.h:84-136),
// stored draw-major for the kernels.  Normals come from
// std::normal_distribution<double>; stratified sampling draws uniforms from
// std::uniform_real_distribution<double> (stratified_sampling.cpp:23-27).
// Elements flagged in hb.stream_continues extend the previous element's
// generator (multilevel_monte_carlo.cpp:20-22: one stream for all levels).
template <typename T>
static void fill_streams(const HostBatch& hb, int32_t n, const std::vector<OptionRec>& recs,
                         const std::vector<int64_t>& off, T* buf) {
    const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<int32_t> starts;
    for (int32_t i = 0; i < n; ++i)
        if (!hb.stream_continues || !hb.stream_continues[i] || i == 0) starts.push_back(i);
    starts.push_back(n);
    const int32_t groups = int32_t(starts.size()) - 1;
    std::atomic<int32_t> next{0};
    auto work = [&]() {
        for (;;) {
            const int32_t g = next.fetch_add(1);
            if (g >= groups) break;
            std::mt19937_64 eng(hb.seed[starts[g]]);
            std::normal_distribution<double> nd(0.0, 1.0);
            std::uniform_real_distribution<double> ud(0.0, 1.0);
            for (int32_t i = starts[g]; i < starts[g + 1]; ++i) {
                const OptionRec& rc = recs[i];
                if (rc.family == FAM_NONE) continue;
                T* base = buf + off[i];
                if (rc.family == FAM_TERMINAL) {
                    const int64_t cnt = stream_normals(rc);
                    if (rc.method == M_STRATIFIED)
                        for (int64_t k = 0; k < cnt; ++k) base[k] = T(ud(eng));
                    else
                        for (int64_t k = 0; k < cnt; ++k) base[k] = T(nd(eng));
                } else {
                    const int64_t P = stream_units(rc);
                    const int64_t D = stream_draws_per_unit(rc);
                    for (int64_t p = 0; p < P; ++p)
                        for (int64_t dd = 0; dd < D; ++dd) base[dd * P + p] = T(nd(eng));
                }
            }
        }
    };
    std::vector<std::thread> pool;
    for (unsigned k = 1; k < hw; ++k) pool.emplace_back(work);
    work();
    for (auto& th : pool) th.join();
}

// Packs a host SoA batch (stride align64(n)) into pinned memory and queues
// one H2D copy into the device staging buffer; fills *bi and *n_dev.
int32_t stage_host_batch(Device* d, const HostBatch& hb, int32_t n, BatchIn* bi_out,
                         const int32_t** n_dev_out, std::string* err, size_t* bytes_out) {
    const size_t ns = align64(size_t(n));
    const double* dsrc[12] = {hb.spot, hb.strike, hb.t, hb.vol, hb.r, hb.q,
                              hb.v0, hb.kappa, hb.theta, hb.sigma, hb.rho, hb.aux};
    const int32_t* isrc[4] = {hb.model_ids, hb.ot, hb.paths, hb.steps};
    const size_t in_bytes = ns * (12 * 8 + 8 + 4 * 4) + 64;
    QK_HIP(d->pin_in.ensure(in_bytes));
    QK_HIP(d->in.ensure(in_bytes));
    char* hp = static_cast<char*>(d->pin_in.p);
    char* dp = static_cast<char*>(d->in.p);
    BatchIn bi{};
    const double** dfield[12] = {&bi.spot, &bi.strike, &bi.t, &bi.vol, &bi.r, &bi.q,
                                 &bi.v0, &bi.kappa, &bi.theta, &bi.sigma, &bi.rho, &bi.aux};
    size_t off = 0;
    for (int k = 0; k < 12; ++k) {
        if (dsrc[k]) {
            std::memcpy(hp + off, dsrc[k], size_t(n) * 8);
            *dfield[k] = reinterpret_cast<const double*>(dp + off);
        } else {
            *dfield[k] = nullptr;
        }
        off += ns * 8;
    }
    std::memcpy(hp + off, hb.seed, size_t(n) * 8);
    bi.seed = reinterpret_cast<const uint64_t*>(dp + off);
    off += ns * 8;
    const int32_t** ifield[4] = {&bi.method, &bi.ot, &bi.paths, &bi.steps};
    for (int k = 0; k < 4; ++k) {
        if (isrc[k]) {
            std::memcpy(hp + off, isrc[k], size_t(n) * 4);
            *ifield[k] = reinterpret_cast<const int32_t*>(dp + off);
        } else {
            *ifield[k] = nullptr;
        }
        off += ns * 4;
    }
    std::memcpy(hp + off, &n, 4);
    const int32_t* n_dev = reinterpret_cast<const int32_t*>(dp + off);
    off += 64;
    bi.uniform_method = hb.uniform_method;
    bi.method_is_model = hb.model_ids ? 1 : 0;
    if (!bytes_out)
        QK_HIP(hipMemcpyAsync(d->in.p, d->pin_in.p, off, hipMemcpyHostToDevice, d->stream));
    else
        *bytes_out = off;  // the caller enqueues the copy (inside its graph)
    *bi_out = bi;
    *n_dev_out = n_dev;
    return QK_OK;
}

namespace {

// Runs `enqueue` on d->stream through a cached instantiated graph when
// possible.  The first call for a key captures and instantiates; later calls
// with the same key only launch.  Any capture or instantiation failure turns
// graphs off for the process and runs the work directly.
bool graphs_active(const Device* d) {
    static const bool disabled = env_int("QK_NO_GRAPH", 0) != 0;
    return !disabled && d->graphs_ok && d->graphs_enabled;
}

template <class F>
int32_t run_graphed(Device* d, const std::vector<uint64_t>& key, F&& enqueue, std::string* err) {
    if (!graphs_active(d)) return enqueue();
    for (auto& g : d->graphs)
        if (g.key == key) {
            QK_HIP(hipGraphLaunch(g.exec, d->stream));
            ++d->graph_launches;
            return QK_OK;
        }
    hipError_t e = hipStreamBeginCapture(d->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
        d->graphs_ok = false;
        return enqueue();
    }
    const int32_t rc = enqueue();
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(d->stream, &graph);
    hipGraphExec_t exec = nullptr;
    if (rc == QK_OK && e == hipSuccess && graph)
        e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (graph) (void)hipGraphDestroy(graph);
    if (rc != QK_OK || e != hipSuccess || !exec) {
        (void)hipGetLastError();
        d->graphs_ok = false;
        return enqueue();
    }
    if (d->graphs.size() >= 8) {
        (void)hipGraphExecDestroy(d->graphs.front().exec);
        d->graphs.erase(d->graphs.begin());
    }
    d->graphs.push_back(Device::Graph{key, exec});
    QK_HIP(hipGraphLaunch(exec, d->stream));
    ++d->graph_launches;
    return QK_OK;
}

}  // namespace


int32_t run_mc_batch(const HostBatch& hb, int32_t n, double* out_price, double* out_se,
                     std::string* err) {
    Device* d = get_device(err);
    if (!d) return QK_ERR_RUNTIME_INIT;
    std::lock_guard<std::mutex> lock(d->mu);
    QK_HIP(hipSetDevice(d->dev));

    const int32_t mode = get_rng_mode();
    // family of the launch: a batch is all-terminal or all-stepwise
    int32_t family = FAM_STEPWISE;
    if (!hb.model_ids && is_terminal_method(hb.uniform_method)) family = FAM_TERMINAL;

    const int64_t tiles = tile_bound(hb, n, d);
    if (tiles > 0x7fffffffLL) {
        if (err) *err = "batch too large: more than 2^31-1 path tiles";
        return QK_ERR_INVALID_INPUT;
    }

    const size_t ns = align64(size_t(n));
    BatchIn bi{};
    const int32_t* n_dev = nullptr;
    // Philox batches run as one cached graph: H2D, kernels, D2H
    const bool graphed = mode == QK_RNG_PHILOX;
    size_t in_copy = 0;
    int32_t prc = stage_host_batch(d, hb, n, &bi, &n_dev, err, graphed ? &in_copy : nullptr);
    if (prc != QK_OK) return prc;

    // ---- device work buffers ----
    QK_HIP(d->recs.ensure(ns * sizeof(OptionRec)));
    QK_HIP(d->ntiles.ensure(ns * 4));
    QK_HIP(d->tile_start.ensure((ns + 1) * 4));
    QK_HIP(d->counters.ensure(64));
    QK_HIP(d->scan.ensure((ns / 1024 + 2) * 8));
    QK_HIP(d->partials.ensure(size_t(std::max<int64_t>(tiles, 1)) * sizeof(Partial)));
    QK_HIP(d->out.ensure(ns * 16 + 64));
    QK_HIP(d->pin_out.ensure(ns * 16 + 64));
    double* dprice = static_cast<double*>(d->out.p);
    double* dse = dprice + ns;

    // ---- validation modes: host normal stream ----
    const int64_t* d_off = nullptr;
    const void* d_stream = nullptr;
    int32_t sim_mode = SIM_PHILOX_F32;
    if (mode == QK_RNG_MT19937_FP64 || mode == QK_RNG_MT19937_FP32) {
        sim_mode = mode == QK_RNG_MT19937_FP64 ? SIM_STREAM_F64 : SIM_STREAM_F32;
        std::vector<OptionRec> hrec(static_cast<size_t>(n));
        std::vector<int64_t> hoff(static_cast<size_t>(n) + 1, 0);
        for (int32_t i = 0; i < n; ++i) {
            RawOption o;
            o.method = method_of(hb, i);
            o.spot = hb.spot[i];
            o.strike = hb.strike[i];
            o.t = hb.t[i];
            o.vol = hb.vol ? hb.vol[i] : 0.0;
            o.r = hb.r[i];
            o.q = hb.q[i];
            o.v0 = hb.v0 ? hb.v0[i] : 0.0;
            o.kappa = hb.kappa ? hb.kappa[i] : 0.0;
            o.theta = hb.theta ? hb.theta[i] : 0.0;
            o.sigma = hb.sigma ? hb.sigma[i] : 0.0;
            o.rho = hb.rho ? hb.rho[i] : 0.0;
            o.ot = hb.ot[i];
            o.paths = hb.paths[i];
            o.steps = hb.steps ? hb.steps[i] : 1;
            o.seed = hb.seed[i];
            o.aux = hb.aux ? hb.aux[i] : 0.0;
            if (o.method < 0) {
                hrec[i].family = FAM_NONE;
            } else {
                prepare_option(o, 1, hrec[i]);
            }
            hoff[i + 1] = hoff[i] + stream_normals(hrec[i]);
        }
        const int64_t total = hoff[n];
        const size_t esz = sim_mode == SIM_STREAM_F64 ? 8 : 4;
        const int64_t limit = int64_t(env_int("QK_VALIDATION_STREAM_GB", 32)) << 30;
        if (total * int64_t(esz) > limit) {
            if (err)
                *err = "validation stream too large (" + std::to_string(total) +
                       " normals); reduce paths or raise QK_VALIDATION_STREAM_GB";
            return QK_ERR_INVALID_INPUT;
        }
        QK_HIP(d->stream_buf.ensure(size_t(std::max<int64_t>(total, 1)) * esz));
        QK_HIP(d->stream_off.ensure(ns * 8));
        std::vector<char> host(size_t(std::max<int64_t>(total, 1)) * esz);
        if (sim_mode == SIM_STREAM_F64)
            fill_streams<double>(hb, n, hrec, hoff, reinterpret_cast<double*>(host.data()));
        else
            fill_streams<float>(hb, n, hrec, hoff, reinterpret_cast<float*>(host.data()));
        QK_HIP(hipMemcpyAsync(d->stream_buf.p, host.data(), size_t(total) * esz,
                              hipMemcpyHostToDevice, d->stream));
        QK_HIP(hipMemcpyAsync(d->stream_off.p, hoff.data(), size_t(n) * 8, hipMemcpyHostToDevice,
                              d->stream));
        QK_HIP(hipStreamSynchronize(d->stream));  // host vectors die at scope end
        d_off = static_cast<const int64_t*>(d->stream_off.p);
        d_stream = d->stream_buf.p;
    }

    const bool use_graph = graphed && graphs_active(d);
    int32_t* hflags = host_flags_ptr(d->pin_out.p, ns * 16);
    if (!hflags) {
        if (err) *err = "pinned host memory has no device address";
        return QK_ERR_RUNTIME_INIT;
    }
    auto enqueue = [&]() -> int32_t {
        if (graphed)
            QK_HIP(hipMemcpyAsync(d->in.p, d->pin_in.p, in_copy, hipMemcpyHostToDevice, d->stream));
        // inside a graph the pipeline records no events (they are taken around it);
        // the finalizer stores the scan's counters after the outputs in pin_out
        int32_t rc = launch_pipeline(d, context_bufs(d), bi, n, n_dev, d_off, family, sim_mode,
                                     d_stream, dprice, dse, err, !use_graph, 0x7fffffff, hflags);
        if (rc != QK_OK) return rc;
        QK_HIP(hipMemcpyAsync(d->pin_out.p, d->out.p, ns * 16, hipMemcpyDeviceToHost, d->stream));
        return QK_OK;
    };
    int32_t rc;
    if (use_graph) {
        const SimShape& st = d->shape_step;
        const SimShape& tt = d->shape_term;
        uint64_t present = 0;
        const void* fields[] = {bi.method, bi.spot, bi.strike, bi.t, bi.vol, bi.r, bi.q, bi.v0,
                                bi.kappa, bi.theta, bi.sigma, bi.rho, bi.ot, bi.paths, bi.steps,
                                bi.seed, bi.aux};
        for (int k = 0; k < 17; ++k) present |= uint64_t(fields[k] != nullptr) << k;
        const std::vector<uint64_t> key = {
            1, uint64_t(n), uint64_t(family), uint64_t(bi.uniform_method), uint64_t(bi.method_is_model),
            present, uint64_t(st.lane_units), uint64_t(st.grid), uint64_t(tt.lane_units),
            uint64_t(tt.grid), uint64_t(in_copy), uint64_t(uintptr_t(d->in.p)),
            uint64_t(uintptr_t(d->pin_in.p)), uint64_t(uintptr_t(d->recs.p)),
            uint64_t(uintptr_t(d->ntiles.p)), uint64_t(uintptr_t(d->tile_start.p)),
            uint64_t(uintptr_t(d->counters.p)), uint64_t(uintptr_t(d->scan.p)),
            uint64_t(uintptr_t(d->partials.p)), uint64_t(uintptr_t(d->out.p)),
            uint64_t(uintptr_t(d->pin_out.p))};
        QK_HIP(hipEventRecord(d->ev[0], d->stream));
        rc = run_graphed(d, key, enqueue, err);
        if (rc == QK_OK) QK_HIP(hipEventRecord(d->ev[3], d->stream));
    } else {
        rc = enqueue();
    }
    if (rc != QK_OK) return rc;
    QK_HIP(hipStreamSynchronize(d->stream));
    if (use_graph) record_timing_total(d);
    else record_timing(d);
    const int32_t* hc = reinterpret_cast<const int32_t*>(static_cast<char*>(d->pin_out.p) + ns * 16);
    if (hc[2]) {
        if (err) *err = "batch too large: more than 2^31-1 path tiles";
        return QK_ERR_INVALID_INPUT;
    }
    const double* hpr = static_cast<const double*>(d->pin_out.p);
    std::memcpy(out_price, hpr, size_t(n) * 8);
    if (out_se) std::memcpy(out_se, hpr + ns, size_t(n) * 8);
    return QK_OK;
}

int32_t mc_device_locked(Device* d, const BatchIn& in, int32_t n, const int32_t* n_dev,
                         double* out_price_dev, double* out_se_dev, std::string* err) {
    // tile count is only known on the device: prepare + scan, read it back,
    // size the partials, then simulate.
    OptionRec* recs = static_cast<OptionRec*>(d->recs.p);
    int32_t* ntiles = static_cast<int32_t*>(d->ntiles.p);
    int32_t* tstart = static_cast<int32_t*>(d->tile_start.p);
    int32_t* ctr = static_cast<int32_t*>(d->counters.p);
    QK_HIP(hipEventRecord(d->ev[0], d->stream));
    QK_HIP(launch_prepare(d->stream, in, n_dev, 0, nullptr, nullptr,
                          tile_groups_terminal(d->shape_term), tile_units_stepwise(d->shape_step),
                          recs, ntiles));
    QK_HIP(launch_scan(d->stream, n, n_dev, ntiles, recs, tstart, ctr, d->scan.p, 0x7fffffff));
    QK_HIP(hipMemcpyAsync(d->pin_out.p, ctr, 16, hipMemcpyDeviceToHost, d->stream));
    QK_HIP(hipStreamSynchronize(d->stream));
    const int32_t* hc = static_cast<const int32_t*>(d->pin_out.p);
    if (hc[2]) {
        if (err) *err = "batch too large: more than 2^31-1 path tiles";
        return QK_ERR_INVALID_INPUT;
    }
    QK_HIP(d->partials.ensure(size_t(std::max(hc[1], 1)) * sizeof(Partial)));
    const int32_t family =
        (!in.method && is_terminal_method(in.uniform_method)) ? FAM_TERMINAL : FAM_STEPWISE;
    const SimShape& shp = family == FAM_TERMINAL ? d->shape_term : d->shape_step;
    Partial* parts = static_cast<Partial*>(d->partials.p);
    QK_HIP(hipEventRecord(d->ev[1], d->stream));
    QK_HIP(launch_simulate(d->stream, family, SIM_PHILOX_F32, shp, recs, tstart, n_dev, ctr, parts,
                           nullptr, pair_only_batch(in, family)));
    QK_HIP(hipEventRecord(d->ev[2], d->stream));
    QK_HIP(launch_finalize(d->stream, n_dev, recs, parts, ctr, out_price_dev, out_se_dev, nullptr));
    QK_HIP(hipEventRecord(d->ev[3], d->stream));
    return QK_OK;
}

int32_t ensure_mc_capacity(Device* d, int32_t n, std::string* err) {
    const size_t ns = align64(size_t(n));
    QK_HIP(d->recs.ensure(ns * sizeof(OptionRec)));
    QK_HIP(d->ntiles.ensure(ns * 4));
    QK_HIP(d->tile_start.ensure((ns + 1) * 4));
    QK_HIP(d->counters.ensure(64));
    QK_HIP(d->pin_out.ensure(64));
    QK_HIP(d->scan.ensure((ns / 1024 + 2) * 8));
    return QK_OK;
}

int32_t run_mc_batch_device(const BatchIn& in, int32_t n, double* out_price_dev,
                            double* out_se_dev, std::string* err) {
    Device* d = get_device(err);
    if (!d) return QK_ERR_RUNTIME_INIT;
    if (get_rng_mode() != QK_RNG_PHILOX) {
        if (err) *err = "device-resident batches support QK_RNG_PHILOX only";
        return QK_ERR_INVALID_INPUT;
    }
    std::lock_guard<std::mutex> lock(d->mu);
    QK_HIP(hipSetDevice(d->dev));
    int32_t rc = ensure_mc_capacity(d, n, err);
    if (rc != QK_OK) return rc;
    QK_HIP(d->pin_in.ensure(64));
    QK_HIP(d->in.ensure(64));
    // The tile count is only known on the device, so the partials buffer has a
    // capacity (grow-only, at least QK_DEVICE_TILES tiles); a batch that needs
    // more runs no tile, and is run again once after the buffer has grown.
    // Without a mid-batch synchronisation the whole pipeline is one graph.
    const int64_t min_tiles = std::max(1, env_int("QK_DEVICE_TILES", 1 << 20));
    QK_HIP(d->partials.ensure(size_t(min_tiles) * sizeof(Partial)));
    const int32_t family =
        (!in.method && is_terminal_method(in.uniform_method)) ? FAM_TERMINAL : FAM_STEPWISE;
    int32_t* hflags = host_flags_ptr(d->pin_out.p, 0);
    if (!hflags) {
        if (err) *err = "pinned host memory has no device address";
        return QK_ERR_RUNTIME_INIT;
    }
    for (int attempt = 0; attempt < 2; ++attempt) {
        const int32_t cap =
            int32_t(std::min<size_t>(d->partials.cap / sizeof(Partial), size_t(0x7fffffff)));
        const bool use_graph = attempt == 0 && graphs_active(d);
        // n travels as a kernel argument; the finalizer stores the scan's
        // counters in pin_out: the graph holds only the four kernels
        auto enqueue = [&]() -> int32_t {
            return launch_pipeline(d, context_bufs(d), in, n, nullptr, nullptr, family,
                                   SIM_PHILOX_F32, nullptr, out_price_dev, out_se_dev, err,
                                   !use_graph, cap, hflags);
        };
        if (use_graph) {
            const SimShape& st = d->shape_step;
            const SimShape& tt = d->shape_term;
            const void* fields[] = {in.method, in.spot, in.strike, in.t, in.vol, in.r, in.q,
                                    in.v0, in.kappa, in.theta, in.sigma, in.rho, in.ot,
                                    in.paths, in.steps, in.seed, in.aux, in.active,
                                    out_price_dev, out_se_dev};
            std::vector<uint64_t> key = {
                2, uint64_t(n), uint64_t(family), uint64_t(in.uniform_method),
                uint64_t(in.method_is_model), uint64_t(uint32_t(in.paths_override)),
                uint64_t(st.lane_units), uint64_t(st.grid), uint64_t(tt.lane_units),
                uint64_t(tt.grid), uint64_t(cap), uint64_t(uintptr_t(d->in.p)),
                uint64_t(uintptr_t(d->pin_in.p)), uint64_t(uintptr_t(d->recs.p)),
                uint64_t(uintptr_t(d->ntiles.p)), uint64_t(uintptr_t(d->tile_start.p)),
                uint64_t(uintptr_t(d->counters.p)), uint64_t(uintptr_t(d->scan.p)),
                uint64_t(uintptr_t(d->partials.p)), uint64_t(uintptr_t(d->pin_out.p))};
            for (const void* f : fields) key.push_back(uint64_t(uintptr_t(f)));
            QK_HIP(hipEventRecord(d->ev[0], d->stream));
            rc = run_graphed(d, key, enqueue, err);
            if (rc == QK_OK) QK_HIP(hipEventRecord(d->ev[3], d->stream));
        } else {
            rc = enqueue();
        }
        if (rc != QK_OK) return rc;
        QK_HIP(hipStreamSynchronize(d->stream));
        if (use_graph) record_timing_total(d);
        else record_timing(d);
        const int32_t* hc = static_cast<const int32_t*>(d->pin_out.p);
        if (hc[2] == 0) return QK_OK;
        if (hc[2] == 2) {
            if (err) *err = "batch too large: more than 2^31-1 path tiles";
            return QK_ERR_INVALID_INPUT;
        }
        QK_HIP(d->partials.ensure(size_t(hc[3]) * sizeof(Partial)));
    }
    if (err) *err = "tile partials could not be sized for the batch";
    return QK_ERR_RUNTIME_INIT;
}

int32_t run_mc_batch_device_noalloc(const BatchIn& in, int32_t n, double* out_price_dev,
                                    double* out_se_dev, std::string* err) {
    Device* d = get_device(err);
    if (!d) return QK_ERR_RUNTIME_INIT;
    if (n > kNoAllocMaxOptions) {
        if (err)
            *err = "batch of " + std::to_string(n) +
                   " options exceeds the allocation-free entry point's maximum of " +
                   std::to_string(kNoAllocMaxOptions);
        return QK_ERR_BAD_SIZE;
    }
    if (get_rng_mode() != QK_RNG_PHILOX) {
        if (err) *err = "device-resident batches support QK_RNG_PHILOX only";
        return QK_ERR_INVALID_INPUT;
    }
    if (!d->noalloc_scratch) {
        if (err) *err = d->noalloc_err;
        return QK_ERR_RUNTIME_INIT;
    }
    std::lock_guard<std::mutex> lock(d->mu);
    QK_HIP(hipSetDevice(d->dev));
    char* base = d->noalloc_scratch;
    // n <= 65535 always takes the one-pass scan, which needs no block sums
    const PipeBufs bufs{reinterpret_cast<OptionRec*>(base + kNaRecsOff),
                        reinterpret_cast<int32_t*>(base + kNaNtilesOff),
                        reinterpret_cast<int32_t*>(base + kNaStartOff),
                        reinterpret_cast<int32_t*>(base + kNaCountersOff), nullptr,
                        reinterpret_cast<Partial*>(base + kNaPartialsOff)};
    const int32_t family =
        (!in.method && is_terminal_method(in.uniform_method)) ? FAM_TERMINAL : FAM_STEPWISE;
    // direct launches (no graph: instantiating one allocates); a batch that
    // needs more tiles than the scratch holds runs no tile and gets NaN
    int32_t rc = launch_pipeline(d, bufs, in, n, nullptr, nullptr, family, SIM_PHILOX_F32, nullptr,
                                 out_price_dev, out_se_dev, err, true, kNoAllocMaxTiles, nullptr);
    if (rc != QK_OK) return rc;
    int32_t flags[4] = {0, 0, 0, 0};
    QK_HIP(hipMemcpyAsync(flags, bufs.counters, sizeof(flags), hipMemcpyDeviceToHost, d->stream));
    QK_HIP(hipStreamSynchronize(d->stream));
    record_timing(d);
    if (flags[2] == 2) {
        if (err) *err = "batch too large: more than 2^31-1 path tiles";
        return QK_ERR_INVALID_INPUT;
    }
    if (flags[2] != 0) {
        if (err)
            *err = "batch needs " + std::to_string(flags[3]) +
                   " path tiles; the allocation-free entry point holds at most " +
                   std::to_string(kNoAllocMaxTiles) + " (outputs are NaN)";
        return QK_ERR_INVALID_INPUT;
    }
    return QK_OK;
}

// Longstaff-Schwartz batch: host validation (longstaff_schwartz.cpp:50-55),
// then chunks of options sized to a memory budget, one workgroup per option.
int32_t run_lsmc_batch(const double* spot, const double* strike, const double* t,
                       const double* vol, const double* r, const double* q, const int32_t* ot,
                       const int32_t* paths, const int32_t* steps, const uint64_t* seed,
                       int32_t basis, int32_t n, double* out_price, double* out_se,
                       std::string* err) {
    Device* d = get_device(err);
    if (!d) return QK_ERR_RUNTIME_INIT;
    std::lock_guard<std::mutex> lock(d->mu);
    QK_HIP(hipSetDevice(d->dev));
    const int32_t mode = get_rng_mode();
    const int32_t sim_mode = mode == QK_RNG_MT19937_FP64   ? SIM_STREAM_F64
                             : mode == QK_RNG_MT19937_FP32 ? SIM_STREAM_F32
                                                           : SIM_PHILOX_F32;
    const size_t esz = sim_mode == SIM_STREAM_F64 ? 8 : 4;

    std::vector<LsmcRec> all;
    all.reserve(size_t(n));
    for (int32_t i = 0; i < n; ++i) {
        out_price[i] = qnan_d();
        if (out_se) out_se[i] = qnan_d();
        if (!common_ok(spot[i], strike[i], t[i], vol[i], r[i], q[i], ot[i]) || paths[i] <= 1 ||
            steps[i] <= 0)
            continue;
        if (t[i] <= kEps) {
            out_price[i] = intrinsic(spot[i], strike[i], ot[i]);
            if (out_se) out_se[i] = 0.0;
            continue;
        }
        LsmcRec rc{};
        rc.out_index = i;
        rc.ot = ot[i];
        rc.paths = paths[i];
        rc.steps = steps[i];
        rc.seed_lo = uint32_t(seed[i]);
        rc.seed_hi = uint32_t(seed[i] >> 32);
        const double dt = t[i] / double(steps[i]);
        rc.spot = spot[i];
        rc.strike = strike[i];
        rc.disc = std::exp(-r[i] * dt);
        rc.drift = (r[i] - q[i] - 0.5 * vol[i] * vol[i]) * dt;
        rc.vs = vol[i] * std::sqrt(dt);
        all.push_back(rc);
    }
    if (all.empty()) return QK_OK;

    const size_t ns = align64(size_t(n));
    QK_HIP(d->out.ensure(ns * 16 + 64));
    QK_HIP(d->pin_out.ensure(ns * 16 + 64));
    double* dprice = static_cast<double*>(d->out.p);
    double* dse = dprice + ns;
    const int64_t budget = int64_t(std::max(1, env_int("QK_LSMC_MEM_GB", 32))) << 30;
    QK_HIP(hipEventRecord(d->ev[0], d->stream));
    QK_HIP(hipEventRecord(d->ev[1], d->stream));
    size_t first = 0;
    while (first < all.size()) {
        // chunk [first, last): bounded by the memory budget and the grid's y extent
        size_t last = first;
        int64_t path_elems = 0, cf_elems = 0, draws = 0;
        while (last < all.size() && last - first < 65535) {
            const int64_t pe = int64_t(all[last].paths) * all[last].steps;
            const int64_t bytes = (path_elems + pe + cf_elems + all[last].paths) * int64_t(esz);
            if (last > first && bytes > budget) break;
            all[last].path_off = path_elems;
            all[last].cf_off = cf_elems;
            all[last].stream_off = draws;
            path_elems += pe;
            cf_elems += all[last].paths;
            draws += pe;
            ++last;
        }
        const int32_t count = int32_t(last - first);
        int32_t tiles_max = 1, max_steps = 1;
        for (size_t k = first; k < last; ++k) {
            tiles_max = std::max(tiles_max, (all[k].paths + kLsmcTile - 1) / kLsmcTile);
            max_steps = std::max(max_steps, all[k].steps);
        }
        QK_HIP(d->lsmc_part.ensure(size_t(count) * tiles_max * lsmc_record_doubles(basis) * 8));
        QK_HIP(d->lsmc_beta.ensure(size_t(count) * 5 * 8));
        QK_HIP(d->lsmc_path.ensure(size_t(path_elems) * esz));
        QK_HIP(d->lsmc_cf.ensure(size_t(cf_elems) * esz));
        QK_HIP(d->lsmc_recs.ensure(size_t(count) * sizeof(LsmcRec)));
        QK_HIP(d->pin_in.ensure(size_t(count) * sizeof(LsmcRec)));
        std::memcpy(d->pin_in.p, all.data() + first, size_t(count) * sizeof(LsmcRec));
        QK_HIP(hipMemcpyAsync(d->lsmc_recs.p, d->pin_in.p, size_t(count) * sizeof(LsmcRec),
                              hipMemcpyHostToDevice, d->stream));
        const void* dstream = nullptr;
        std::vector<char> host;
        if (sim_mode != SIM_PHILOX_F32) {
            // longstaff_schwartz.cpp:61-75: per path, dates 1..N in order
            host.resize(size_t(draws) * esz);
            const unsigned hw = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
            std::atomic<size_t> next{first};
            auto work = [&]() {
                for (;;) {
                    const size_t k = next.fetch_add(1);
                    if (k >= last) break;
                    const LsmcRec& rc = all[k];
                    std::mt19937_64 eng(seed[rc.out_index]);
                    std::normal_distribution<double> nd(0.0, 1.0);
                    const int64_t P = rc.paths, N = rc.steps;
                    for (int64_t p = 0; p < P; ++p)
                        for (int64_t j = 0; j < N; ++j) {
                            const double z = nd(eng);
                            const size_t at = size_t(rc.stream_off + j * P + p);
                            if (esz == 8) reinterpret_cast<double*>(host.data())[at] = z;
                            else reinterpret_cast<float*>(host.data())[at] = float(z);
                        }
                }
            };
            std::vector<std::thread> pool;
            for (unsigned k = 1; k < hw; ++k) pool.emplace_back(work);
            work();
            for (auto& th : pool) th.join();
            QK_HIP(d->stream_buf.ensure(host.size()));
            QK_HIP(hipMemcpyAsync(d->stream_buf.p, host.data(), host.size(),
                                  hipMemcpyHostToDevice, d->stream));
            dstream = d->stream_buf.p;
        }
        QK_HIP(launch_lsmc(d->stream, sim_mode, basis,
                           static_cast<const LsmcRec*>(d->lsmc_recs.p), count, tiles_max,
                           max_steps, d->lsmc_path.p, d->lsmc_cf.p, dstream,
                           static_cast<double*>(d->lsmc_part.p),
                           static_cast<double*>(d->lsmc_beta.p), dprice, dse));
        QK_HIP(hipStreamSynchronize(d->stream));  // pinned recs / host stream reused next chunk
        first = last;
    }
    QK_HIP(hipEventRecord(d->ev[2], d->stream));
    QK_HIP(hipMemcpyAsync(d->pin_out.p, d->out.p, ns * 16, hipMemcpyDeviceToHost, d->stream));
    QK_HIP(hipEventRecord(d->ev[3], d->stream));
    QK_HIP(hipStreamSynchronize(d->stream));
    record_timing(d);
    const double* hp = static_cast<const double*>(d->pin_out.p);
    for (const LsmcRec& rc : all) {
        out_price[rc.out_index] = hp[rc.out_index];
        if (out_se) out_se[rc.out_index] = hp[ns + rc.out_index];
    }
    return QK_OK;
}

// ---------------------------------------------------------------------------
// Shapes and the autotuner
// ---------------------------------------------------------------------------
namespace {

std::string autotune_path() {
    const char* p = std::getenv("QK_AUTOTUNE_CACHE");
    if (p && *p) return p;
    const char* home = std::getenv("HOME");
    return std::string(home && *home ? home : "/tmp") + "/.cache/quantkernel_mi355x_autotune.txt";
}

SimShape* shape_of(Device* d, int32_t family) {
    return family == FAM_TERMINAL ? &d->shape_term : family == FAM_STEPWISE ? &d->shape_step : nullptr;
}

// Synthetic batch of one family in device memory (freed by the caller).
struct TuneBatch {
    std::vector<void*> bufs;
    BatchIn in{};
    int32_t n = 0;
    double *price = nullptr, *se = nullptr;
    int32_t* n_dev = nullptr;
};

template <typename T>
T* upload(TuneBatch& tb, const std::vector<T>& v, std::string* err, int32_t* rc) {
    void* p = nullptr;
    if (hipMalloc(&p, v.size() * sizeof(T)) != hipSuccess ||
        hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) {
        if (err) *err = "autotune: device allocation failed";
        *rc = QK_ERR_RUNTIME_INIT;
        return nullptr;
    }
    tb.bufs.push_back(p);
    return static_cast<T*>(p);
}

int32_t make_tune_batch(int32_t family, TuneBatch& tb, std::string* err) {
    const int32_t n = 4096;
    std::vector<double> spot(n), strike(n), t(n), vol(n), r(n), q(n), v0(n), ka(n), th(n), sg(n), rho(n);
    std::vector<int32_t> model(n), ot(n), paths(n), steps(n);
    std::vector<uint64_t> seed(n);
    for (int32_t i = 0; i < n; ++i) {
        spot[i] = 100.0;
        strike[i] = 80.0 + 40.0 * (i % 97) / 96.0;
        t[i] = 0.25 + 1.5 * (i % 13) / 12.0;
        vol[i] = 0.15 + 0.3 * (i % 7) / 6.0;
        r[i] = 0.03;
        q[i] = 0.0;
        v0[i] = th[i] = 0.04;
        ka[i] = 2.0;
        sg[i] = 0.5;
        rho[i] = -0.6;
        model[i] = i % 4;
        ot[i] = i & 1;
        // C2-like imbalance for the stepwise family, C0 for the terminal one
        paths[i] = family == FAM_TERMINAL ? 100000 : (4096 << (i % 5));
        steps[i] = family == FAM_TERMINAL ? 1 : (16 << (i % 4));
        seed[i] = 1234567u + uint64_t(i);
    }
    int32_t rc = QK_OK;
    tb.n = n;
    tb.in.spot = upload(tb, spot, err, &rc);
    tb.in.strike = upload(tb, strike, err, &rc);
    tb.in.t = upload(tb, t, err, &rc);
    tb.in.vol = upload(tb, vol, err, &rc);
    tb.in.r = upload(tb, r, err, &rc);
    tb.in.q = upload(tb, q, err, &rc);
    tb.in.v0 = upload(tb, v0, err, &rc);
    tb.in.kappa = upload(tb, ka, err, &rc);
    tb.in.theta = upload(tb, th, err, &rc);
    tb.in.sigma = upload(tb, sg, err, &rc);
    tb.in.rho = upload(tb, rho, err, &rc);
    tb.in.ot = upload(tb, ot, err, &rc);
    tb.in.paths = upload(tb, paths, err, &rc);
    tb.in.steps = upload(tb, steps, err, &rc);
    tb.in.seed = upload(tb, seed, err, &rc);
    if (family == FAM_TERMINAL) {
        tb.in.uniform_method = M_STANDARD;
    } else {
        tb.in.method = upload(tb, model, err, &rc);
        tb.in.method_is_model = 1;
    }
    tb.n_dev = upload(tb, std::vector<int32_t>{n}, err, &rc);
    tb.price = upload(tb, std::vector<double>(static_cast<size_t>(n)), err, &rc);
    tb.se = upload(tb, std::vector<double>(static_cast<size_t>(n)), err, &rc);
    return rc;
}

void free_tune_batch(TuneBatch& tb) {
    for (void* p : tb.bufs) (void)hipFree(p);
    tb.bufs.clear();
}

}  // namespace

int32_t set_graphs(int32_t enable, int32_t* cached, int64_t* launches, std::string* err) {
    Device* d = get_device(err);
    if (!d) return QK_ERR_RUNTIME_INIT;
    std::lock_guard<std::mutex> lock(d->mu);
    if (enable >= 0) d->graphs_enabled = enable != 0;
    if (cached) *cached = d->graphs_ok ? int32_t(d->graphs.size()) : -1;
    if (launches) *launches = d->graph_launches;
    return QK_OK;
}

int32_t set_shape(int32_t family, int32_t lane_units, int32_t grid_per_cu, std::string* err) {
    Device* d = get_device(err);
    if (!d) return QK_ERR_RUNTIME_INIT;
    std::lock_guard<std::mutex> lock(d->mu);
    SimShape* sh = shape_of(d, family);
    if (!sh || lane_units < 1 || lane_units > 64 || grid_per_cu < 1 || grid_per_cu > 64) {
        if (err) *err = "shape: family 1 (terminal) or 2 (stepwise), 1 <= lane_units, grid_per_cu <= 64";
        return QK_ERR_INVALID_INPUT;
    }
    sh->lane_units = lane_units;
    sh->grid = d->cus * grid_per_cu;
    return QK_OK;
}

int32_t get_shape(int32_t family, int32_t* lane_units, int32_t* grid_per_cu, std::string* err) {
    Device* d = get_device(err);
    if (!d) return QK_ERR_RUNTIME_INIT;
    std::lock_guard<std::mutex> lock(d->mu);
    SimShape* sh = shape_of(d, family);
    if (!sh) {
        if (err) *err = "shape: family 1 (terminal) or 2 (stepwise)";
        return QK_ERR_INVALID_INPUT;
    }
    if (lane_units) *lane_units = sh->lane_units;
    if (grid_per_cu) *grid_per_cu = sh->grid / std::max(1, d->cus);
    return QK_OK;
}

// Times every candidate shape of a family on a representative synthetic batch
// (the simulate kernel alone, best of two runs), applies the fastest and
// appends it to the cache file.
int32_t autotune(int32_t family, int32_t* best_lane, int32_t* best_grid, double* best_ms,
                 std::string* err) {
    Device* d = get_device(err);
    if (!d) return QK_ERR_RUNTIME_INIT;
    std::lock_guard<std::mutex> lock(d->mu);
    QK_HIP(hipSetDevice(d->dev));
    SimShape* sh = shape_of(d, family);
    if (!sh) {
        if (err) *err = "autotune: family 1 (terminal) or 2 (stepwise)";
        return QK_ERR_INVALID_INPUT;
    }
    TuneBatch tb;
    int32_t rc = make_tune_batch(family, tb, err);
    if (rc == QK_OK) rc = ensure_mc_capacity(d, tb.n, err);
    const SimShape saved = *sh;
    SimShape best = saved;
    double best_t = 1e300;
    // lane counts between the powers of two let a tile divide an option's
    // units nearly evenly (50,000 pairs: 7 tiles of 7168 waste 0.3%, where
    // 7 tiles of 8192 waste 14%)
    const int lanes[] = {2, 3, 4, 5, 6, 7, 8, 12, 16, 20, 24, 25, 32};
    const int grids[] = {4, 8, 16};
    constexpr int kLanes = sizeof(lanes) / sizeof(lanes[0]);
    for (int li = 0; rc == QK_OK && li < kLanes; ++li)
        for (int gi = 0; rc == QK_OK && gi < 3; ++gi) {
            sh->lane_units = lanes[li];
            sh->grid = d->cus * grids[gi];
            double t = 1e300;
            for (int rep = 0; rep < 2 && rc == QK_OK; ++rep) {
                rc = mc_device_locked(d, tb.in, tb.n, tb.n_dev, tb.price, tb.se, err);
                if (rc != QK_OK) break;
                if (hipStreamSynchronize(d->stream) != hipSuccess) {
                    rc = QK_ERR_RUNTIME_INIT;
                    break;
                }
                float ms = 0.f;
                if (hipEventElapsedTime(&ms, d->ev[1], d->ev[2]) == hipSuccess)
                    t = std::min(t, double(ms));
            }
            if (t < best_t) {
                best_t = t;
                best = *sh;
            }
        }
    free_tune_batch(tb);
    if (rc != QK_OK) {
        *sh = saved;
        return rc;
    }
    *sh = best;
    if (best_lane) *best_lane = best.lane_units;
    if (best_grid) *best_grid = best.grid / std::max(1, d->cus);
    if (best_ms) *best_ms = best_t;
    std::ofstream f(autotune_path(), std::ios::app);
    if (f)
        f << d->name << "|" << d->cus << "|" << family << "|" << best.lane_units << "|"
          << best.grid / std::max(1, d->cus) << "|" << best_t << "\n";
    return QK_OK;
}

// Applies the last cached result for this device, if any (QK_AUTOTUNE=1).
void apply_cached_shapes(Device* d) {
    std::ifstream f(autotune_path());
    std::string line;
    while (std::getline(f, line)) {
        std::stringstream ss(line);
        std::string name, cus, fam, lane, grid;
        if (!std::getline(ss, name, '|') || !std::getline(ss, cus, '|') ||
            !std::getline(ss, fam, '|') || !std::getline(ss, lane, '|') ||
            !std::getline(ss, grid, '|'))
            continue;
        if (name != d->name || std::atoi(cus.c_str()) != d->cus) continue;
        SimShape* sh = shape_of(d, std::atoi(fam.c_str()));
        const int lu = std::atoi(lane.c_str()), g = std::atoi(grid.c_str());
        if (sh && lu >= 1 && lu <= 64 && g >= 1 && g <= 64) {
            sh->lane_units = lu;
            sh->grid = d->cus * g;
        }
    }
}

// ----------------------------------------------------------------------------
// Heston characteristic-function batch
// ----------------------------------------------------------------------------
namespace {

// The 64-point Gauss-Legendre rule on [-1, 1]: Newton iteration on the
// Legendre recurrence from Tricomi's initial guesses; the 32 positive nodes in
// ascending order, then their weights.
std::vector<double> gauss_legendre64() {
    const int n = 64;
    std::vector<double> gl(64);
    for (int i = 0; i < n / 2; ++i) {
        double z = std::cos(3.14159265358979323846 * (i + 0.75) / (n + 0.5)), dp = 1.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = z;
            for (int k = 2; k <= n; ++k) {
                const double pk = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
                p0 = p1;
                p1 = pk;
            }
            dp = n * (z * p1 - p0) / (z * z - 1.0);
            const double dz = p1 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-16) break;
        }
        double p0 = 1.0, p1 = z;
        for (int k = 2; k <= n; ++k) {
            const double pk = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
            p0 = p1;
            p1 = pk;
        }
        dp = n * (z * p1 - p0) / (z * z - 1.0);
        gl[31 - i] = z;
        gl[32 + 31 - i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
    return gl;
}

}  // namespace

int32_t ensure_gl(Device* d, std::string* err) {
    if (!d->gl_ready) {
        const std::vector<double> gl = gauss_legendre64();
        QK_HIP(d->gl.ensure(64 * sizeof(double)));
        QK_HIP(hipMemcpy(d->gl.p, gl.data(), 64 * sizeof(double), hipMemcpyHostToDevice));
        d->gl_ready = true;
    }
    return QK_OK;
}

int32_t run_heston_cf_batch_device(const HestonCfIn& in, int32_t n, double* out_dev,
                                   std::string* err) {
    Device* d = get_device(err);
    if (!d) return QK_ERR_RUNTIME_INIT;
    std::lock_guard<std::mutex> lock(d->mu);
    QK_HIP(hipSetDevice(d->dev));
    int32_t rc = ensure_gl(d, err);
    if (rc != QK_OK) return rc;
    const int32_t grid = int32_t(std::min<int64_t>((int64_t(n) + 3) / 4, int64_t(d->cus) * 16));
    QK_HIP(hipEventRecord(d->ev[0], d->stream));
    QK_HIP(hipEventRecord(d->ev[1], d->stream));
    QK_HIP(launch_heston_cf(d->stream, in, n, static_cast<const double*>(d->gl.p), out_dev, grid));
    QK_HIP(hipEventRecord(d->ev[2], d->stream));
    QK_HIP(hipEventRecord(d->ev[3], d->stream));
    QK_HIP(hipStreamSynchronize(d->stream));
    record_timing(d);
    return QK_OK;
}

int32_t run_heston_cf_batch(const double* spot, const double* strike, const double* t,
                            const double* r, const double* q, const double* v0,
                            const double* kappa, const double* theta, const double* sigma,
                            const double* rho, const int32_t* ot, const int32_t* steps,
                            const double* limit, int32_t n, double* out, std::string* err) {
    Device* d = get_device(err);
    if (!d) return QK_ERR_RUNTIME_INIT;
    std::lock_guard<std::mutex> lock(d->mu);
    QK_HIP(hipSetDevice(d->dev));
    int32_t rc = ensure_gl(d, err);
    if (rc != QK_OK) return rc;
    // stage the 11 double and 2 int arrays through the batch packer
    std::vector<uint64_t> no_seed(static_cast<size_t>(n), 0);
    HostBatch hb;
    hb.spot = spot;
    hb.strike = strike;
    hb.t = t;
    hb.r = r;
    hb.q = q;
    hb.v0 = v0;
    hb.kappa = kappa;
    hb.theta = theta;
    hb.sigma = sigma;
    hb.rho = rho;
    hb.ot = ot;
    hb.steps = steps;
    hb.aux = limit;
    hb.seed = no_seed.data();
    BatchIn bi{};
    const int32_t* n_dev = nullptr;
    rc = stage_host_batch(d, hb, n, &bi, &n_dev, err);
    if (rc != QK_OK) return rc;
    const size_t ns = align64(size_t(n));
    QK_HIP(d->out.ensure(ns * 8 + 64));
    QK_HIP(d->pin_out.ensure(ns * 8 + 64));
    HestonCfIn in{bi.spot, bi.strike, bi.t, bi.r, bi.q, bi.v0, bi.kappa, bi.theta, bi.sigma,
                  bi.rho, bi.aux, bi.ot, bi.steps};
    const int32_t grid = int32_t(std::min<int64_t>((int64_t(n) + 3) / 4, int64_t(d->cus) * 16));
    QK_HIP(hipEventRecord(d->ev[0], d->stream));
    QK_HIP(hipEventRecord(d->ev[1], d->stream));
    QK_HIP(launch_heston_cf(d->stream, in, n, static_cast<const double*>(d->gl.p),
                            static_cast<double*>(d->out.p), grid));
    QK_HIP(hipEventRecord(d->ev[2], d->stream));
    QK_HIP(hipMemcpyAsync(d->pin_out.p, d->out.p, size_t(n) * 8, hipMemcpyDeviceToHost,
                          d->stream));
    QK_HIP(hipEventRecord(d->ev[3], d->stream));
    QK_HIP(hipStreamSynchronize(d->stream));
    record_timing(d);
    std::memcpy(out, d->pin_out.p, size_t(n) * 8);
    return QK_OK;
}

}  // namespace qk
