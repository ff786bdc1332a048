"""Benchmark of the Heston pricing path on one MI355X.

Default workload (the headline metric): 1,048,576 seeded Heston European
call queries, 95% drawn from the surrogate's training box and 5% from the box
widened by 10% on every side.  One step prices the whole batch through
qk_sur_heston_price_batch_device: network for in-box queries, plain MC
(65536 paths x 256 steps, Philox, FP32 paths, FP64 sums) for queries outside
the box and for spot-check disagreements (every 128th answered query is
re-priced with 8192 paths).  Inputs are resident in HBM before timing.

Alongside, in the same run:
  * plain GPU MC at the same accuracy: the same MC kernel at 65536 paths
    (median SE ~6.7e-4 of spot, below the 1e-3 target) on a 4096-query slice;
  * accuracy of the served answers (network and MC fallback alike) on every
    512th query of the timed batch, and of plain MC at 65536 paths on the same
    queries, against MC at 4,194,304 paths; the batch judge's statistics;
  * the reference's CPU batch entry point (a build of its sources, -O3
    -fopenmp) on the first queries, same paths and steps, on all host threads
    and on one.

Other workloads (reported with the same fields; not the headline line):
  c0  4096 GBM calls/puts, standard MC (antithetic terminal), 100000 paths;
  c1  plain MC: 16384 Heston calls x 65536 paths x 256 steps;
  c2  50000 uneven mixed options (GBM Euler / Milstein / Heston / local vol),
      paths log-uniform in [4096, 1048576], steps log-uniform in [16, 512];
  c3  1024 American puts, Longstaff-Schwartz, 100000 paths x 50 dates,
      basis {1, x, x^2, x^3}, x = S/K;
  cf  1048576 Heston characteristic-function prices (1024 nodes, FP64).
c3 goes through the host-memory entry point (its inputs are 100 bytes per
option); the others use device-resident entry points, replayed as launch
graphs where the path has them.

Device memory comes from the HIP runtime the library links (no torch).
roofline.traffic and roofline.valu_issue are profiled constants: rocprofv3
--pmc runs of the same command, read from profiles/ (labelled as such).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload c0|c1|c2|c3|c4|cf]
       python bench.py --calibration-copy MIB   (the known-size copy for PMC scaling)
"""
import argparse
import ctypes as ct
import json
import math
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "quant-kernel_amd"))

from quantkernel_amd import devmem  # noqa: E402  (the library's own HIP runtime)

PEAK_FP32_TFLOPS = 157.3      # MI355X published FP32 vector peak
PEAK_BF16_TFLOPS = 2500.0     # MI355X published dense BF16 MFMA peak
FLOPS_PER_PATH_STEP = 37      # Heston step incl. two Box-Muller normals (algorithmic count)
MC_PATHS = 65536
SPOT_EVERY, SPOT_PATHS = 128, 8192
STEPS = 256
METRIC = "Heston prices/sec at \u22641e-3 rel err: surrogate+fallback vs plain MC, 1 GPU"
# HBM bytes per launch of each workload's dominant kernel, from rocprofv3
# --pmc FETCH_SIZE / WRITE_SIZE runs scaled by a copy kernel of known size
# (written by the profiling run; absent -> traffic null)
PMC_TRAFFIC = ROOT / "profiles" / "pmc_traffic.json"
# VALU issue of the dominant kernel (rocprofv3 --pmc SQ_ACTIVE_INST_VALU / SQ_BUSY_CYCLES style
# counters of the same command), as a fraction of the SIMDs' issue cycles
PMC_VALU = ROOT / "profiles" / "pmc_valu_issue.json"
ACC_STRIDE = 512  # C4 accuracy: every 512th query of the timed batch
KEYS = ("spot", "strike", "t", "r", "q", "v0", "kappa", "theta", "sigma", "rho")


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


_GOMP = None


def set_omp_threads(k):
    """Thread count of the OpenMP parallel regions the CPU baseline starts
    from this thread (the reference build links libgomp)."""
    global _GOMP
    if _GOMP is None:
        _GOMP = ct.CDLL("libgomp.so.1")
        _GOMP.omp_set_num_threads.argtypes = [ct.c_int]
    _GOMP.omp_set_num_threads(int(k))


def ref_fn(name):
    """A batch entry point of the reference build (oracle/_ref), or (None, None)."""
    ref = ROOT / "oracle" / "_ref" / "libqkref.so"
    if ref.exists():
        return "reference", getattr(ct.CDLL(str(ref)), name)
    return None, None


def time_cpu(fn, arrays, idx):
    """Wall seconds of one batch call of fn on the elements idx of arrays
    (the batch arguments in order, before n and the output)."""
    PD, PI, PU = ct.POINTER(ct.c_double), ct.POINTER(ct.c_int32), ct.POINTER(ct.c_uint64)
    conv = [np.ascontiguousarray(a[idx]) for a in arrays]
    ptrs = [a.ctypes.data_as(PD if a.dtype == np.float64 else PI if a.dtype == np.int32 else PU) for a in conv]
    k = len(idx)
    out = np.empty(k)
    t0 = time.perf_counter()
    fn(*ptrs, ct.c_int32(k), out.ctypes.data_as(PD))
    return time.perf_counter() - t0


def sample_rate(fn, arrays, n, threads, budget_s, full=False, stride=False):
    """Prices/s of fn on `threads` OpenMP threads: all n elements (full), or a
    sample grown from a probe to about budget_s seconds of wall time: the
    first k elements, or (stride) every (n/k)-th element, which keeps the
    batch's mix of sizes and models.  Returns (k, seconds, description)."""
    set_omp_threads(threads)
    try:
        def pick(k):
            if stride:
                return np.unique(np.linspace(0, n - 1, k).round().astype(np.int64))
            return np.arange(k)
        if full:
            idx = np.arange(n)
            return n, time_cpu(fn, arrays, idx), f"all {n}"
        k = min(n, max(1, threads))
        dt = time_cpu(fn, arrays, pick(k))
        if dt < budget_s / 3 and k < n:
            k = min(n, max(k + 1, int(k * budget_s / max(dt, 1e-3))))
            if not stride:
                k = max(threads, k // threads * threads)
            idx = pick(k)
            k = len(idx)
            dt = time_cpu(fn, arrays, idx)
        return k, dt, (f"every {n / k:.0f}th of {n}" if stride else f"first {k} of {n}")
    finally:
        set_omp_threads(host_threads())


def cpu_baseline(fn, kind, arrays, n, noun, what, budget_s=10.0, budget_1t_s=5.0, full=False,
                 stride=False):
    """The reference's CPU batch entry point on all host threads and on one."""
    threads = host_threads()
    k, dt, how = sample_rate(fn, arrays, n, threads, budget_s, full=full, stride=stride)
    log(f"cpu baseline: {k} {noun} in {dt:.2f}s on {threads} threads")
    k1, dt1, how1 = sample_rate(fn, arrays, n, 1, budget_1t_s, full=full, stride=stride)
    log(f"cpu baseline: {k1} {noun} in {dt1:.2f}s on 1 thread")
    return {"value": k / dt, "unit": "prices/s", "cores": threads, "kind": kind,
            "sample": f"{how} {noun} of the batch, {what}, OpenMP over options ({threads} threads, "
                      f"{cpu_model()}), wall time {dt:.2f} s; reference sources built -O3 -fopenmp",
            "one_thread": {"value": k1 / dt1, "unit": "prices/s", "cores": 1,
                           "sample": f"{how1} {noun}, wall time {dt1:.2f} s"}}


def heston_cpu_baseline(d, noun="options", note=""):
    """Reference Heston MC batch entry point (or, without the reference build,
    the oracle's port of it) on the first options of d."""
    kind, fn = ref_fn("qk_mcm_heston_monte_carlo_price_batch")
    if not fn:
        kind, fn = "port", ct.CDLL(str(ROOT / "oracle" / "liboracle.so")).orc_heston_batch
    arrays = [d[k] for k in KEYS] + [d["option_type"], d["paths"], d["steps"], d["seed"]]
    return cpu_baseline(fn, kind, arrays, len(d["spot"]), noun,
                        f"Heston MC batch entry point, {int(d['paths'][0])} paths x {int(d['steps'][0])} "
                        f"steps each{note}")


def bsm_price(s, k, t, vol, r, q, ot):
    """Black-Scholes-Merton closed form (Black & Scholes 1973, Merton 1973)."""
    from scipy.special import ndtr
    st = vol * np.sqrt(t)
    d1 = (np.log(s / k) + (r - q + 0.5 * vol * vol) * t) / st
    d2 = d1 - st
    call = s * np.exp(-q * t) * ndtr(d1) - k * np.exp(-r * t) * ndtr(d2)
    return np.where(ot == 0, call, call - s * np.exp(-q * t) + k * np.exp(-r * t))


def z_stats(price, se, ref, spot, against):
    """Errors of MC prices against reference prices, in standard errors and
    per unit spot.  z moments use the options whose SE is meaningful: with
    (price / SE)^2 >= 30, about 30 or more paths finish in the money (deep
    out-of-the-money options with a handful of hits have SE ~ price)."""
    ok = np.isfinite(price) & np.isfinite(ref) & (se > 0)
    z = (price[ok] - ref[ok]) / se[ok]
    e = np.abs(price[ok] - ref[ok]) / spot[ok]
    rel = (price[ok] / se[ok]) ** 2 >= 30
    zr = z[rel]
    return {"against": against, "options": int(ok.sum()), "frac_within_3se": float((np.abs(z) < 3).mean()),
            "z_options": int(rel.sum()), "mean_z": float(zr.mean()), "rms_z": float(np.sqrt((zr * zr).mean())),
            "median_abs_err_over_spot": float(np.median(e)), "max_abs_err_over_spot": float(e.max())}


def heston_cf_prices(qk, d, idx):
    """The library's FP64 Heston characteristic-function prices (1024
    Gauss-Legendre nodes on [1e-8, 120]) of the options idx of d."""
    m = len(idx)
    return qk.heston_price_cf_batch(*[d[k][idx] for k in KEYS], d["option_type"][idx],
                                    np.full(m, 1024, np.int32), np.full(m, 120.0))


def oracle_lib():
    """The FP64 oracle (a checker: used only inside the CPU-baseline legs)."""
    path = ROOT / "oracle" / "liboracle.so"
    return ct.CDLL(str(path)) if path.exists() else None


def threaded(fn, items, threads):
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(fn, items))


def to_device(d):
    """Device copies of the batch arrays (synchronous copies: complete before
    the library's stream starts on them)."""
    return devmem.to_device(d)


def ptrs_of(t):
    return {k: v.ptr for k, v in t.items()}


def direct_kernel_ms(qk, call, reps=1):
    """Simulate-kernel time from the HIP events the library records around the
    kernel.  A replayed launch graph is timed only as a whole, so these
    batches run with direct launches (same kernels, same shapes)."""
    qk.graphs(0)
    try:
        sims = []
        for _ in range(reps):
            call()
            sims.append(qk.last_timing()[1])
    finally:
        qk.graphs(1)
    return float(np.mean(sims))


def plain_mc_rate(qk, qka, d, n, reps=2):
    """Plain GPU MC (device-resident) on the first n queries: prices/s and
    the simulate-kernel time."""
    sl = {k: v[:n] for k, v in d.items()}
    sl["model"] = np.full(n, qka.QK_MODEL_HESTON_EULER, np.int32)
    sl["vol"] = np.zeros(n)
    t = to_device(sl)
    ptrs = ptrs_of(t)
    out, se = devmem.DevArray(n=n), devmem.DevArray(n=n)

    def call():
        qk.mixed_price_batch_device(n, ptrs, out.ptr, se.ptr)

    call()
    devmem.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        call()
    devmem.synchronize()
    wall = (time.perf_counter() - t0) / reps
    return n / wall, direct_kernel_ms(qk, call), out.numpy(), se.numpy()


def bench_c1(args, qk, qka, dev, name, cus):
    from quantkernel_amd import workloads as wl
    n = args.options or 16384
    d = wl.config1(n)
    d["model"] = np.full(n, qka.QK_MODEL_HESTON_EULER, np.int32)
    d["vol"] = np.zeros(n)
    t = to_device(d)
    ptrs = ptrs_of(t)
    out = devmem.DevArray(n=n)
    se = devmem.DevArray(n=n)
    def call():
        qk.mixed_price_batch_device(n, ptrs, out.ptr, se.ptr)

    for _ in range(args.warmup):
        call()
    devmem.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        call()  # a replayed launch graph
        log(f"step done: batch {qk.last_timing()[0]:.2f} ms")
    devmem.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    sim_ms = direct_kernel_ms(qk, call)
    log(f"simulate kernel {sim_ms:.2f} ms")
    flops = float(n) * MC_PATHS * STEPS * FLOPS_PER_PATH_STEP
    ach = flops / (sim_ms * 1e-3) / 1e12
    res = dict(value=n / (ms * 1e-3), ms_per_step=ms, dtype="fp32",
               config={"workload": f"plain GPU MC: {n} Heston full-truncation Euler calls x {MC_PATHS} paths x "
                                   f"{STEPS} steps (Philox, FP32 paths, FP64 sums)", "options": n,
                       "paths": MC_PATHS, "time_steps": STEPS, "device": name, "cus": cus,
                       "median_stderr_over_spot": float(np.median(se.numpy() / d["spot"]))},
               roofline={"bound": "valu", "achieved": ach, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                         "frac": ach / PEAK_FP32_TFLOPS, "traffic": None,
                         "kernel": "qk::k_sim_stepwise<float, PHILOX>", "kernel_ms": sim_ms,
                         "flops_per_launch": flops})
    price, sev = out.numpy(), se.numpy()
    idx = np.arange(min(n, 4096))
    res["config"]["accuracy"] = z_stats(price[idx], sev[idx], heston_cf_prices(qk, d, idx), d["spot"][idx],
                                        "Heston characteristic-function prices (FP64, this library); the "
                                        "difference includes the 256-step Euler bias")
    if not args.no_cpu:
        res["cpu_baseline"] = heston_cpu_baseline(d)
        orc = oracle_lib()
        if orc:
            f = orc.orc_px_heston
            f.restype = ct.c_double
            f.argtypes = [ct.c_double] * 10 + [ct.c_int32] * 3 + [ct.c_uint64, ct.POINTER(ct.c_double)]
            threads = host_threads()
            k = min(n, threads)

            def one(i):
                e = ct.c_double()
                return f(*[float(d[key][i]) for key in KEYS], int(d["option_type"][i]), MC_PATHS, STEPS,
                         int(d["seed"][i]), ct.byref(e))

            want = np.array(threaded(one, range(k), threads))
            res["cpu_baseline"]["fp64_oracle_check"] = dict(
                z_stats(price[:k], sev[:k], want, d["spot"][:k],
                        "FP64 oracle on the same Philox stream (first options of the batch)"),
                max_abs_err_over_stderr=float(np.max(np.abs(price[:k] - want) / sev[:k])))
    return res


def bench_c0(args, qk, qka, dev, name, cus):
    from quantkernel_amd import workloads as wl
    n = args.options or 4096
    d = wl.config0(n)
    a = [d[k] for k in ("spot", "strike", "t", "vol", "r", "q", "option_type", "paths")]
    t = to_device(d)
    ptrs = ptrs_of(t)
    out = devmem.DevArray(n=n)
    seo = devmem.DevArray(n=n)
    # tile shape from the autotuner (timed on a C0-shaped synthetic batch)
    lane, grid, tune_ms = qk.autotune(qk.FAMILY_TERMINAL)
    log(f"autotuned terminal shape: {lane} Philox groups/lane, {grid} workgroups/CU")

    def call():
        qk.standard_monte_carlo_price_batch_device(n, ptrs, out.ptr, seo.ptr)

    for _ in range(args.warmup):
        call()
    devmem.synchronize()
    tot = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        call()  # a replayed launch graph
        tot.append(qk.last_timing()[0])
        log(f"step done: device {tot[-1]:.3f} ms")
    devmem.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    sim_ms = direct_kernel_ms(qk, call, reps=5)
    se = seo.numpy()
    # the host-memory entry point on the same batch (adds the PCIe copies)
    reps = max(args.steps, 5)
    qk.standard_monte_carlo_price_batch(*a, d["seed"])
    t1 = time.perf_counter()
    for _ in range(reps):
        qk.standard_monte_carlo_price_batch(*a, d["seed"])
    host_rate = n * reps / (time.perf_counter() - t1)
    flops = float(n) * 100000 * 11
    ach = flops / (sim_ms * 1e-3) / 1e12
    res = dict(value=n / (ms * 1e-3), ms_per_step=ms, dtype="fp32",
               config={"workload": f"C0: {n} GBM European calls/puts, standard MC (antithetic terminal), 100000 paths "
                                   "(Philox, FP32 paths, FP64 sums); device-resident entry point, one launch graph "
                                   "per batch", "options": n,
                       "device_ms": float(np.mean(tot)), "device": name, "cus": cus,
                       "autotuned_shape": {"philox_groups_per_lane": lane, "workgroups_per_cu": grid},
                       "host_entry_prices_per_s_incl_pcie": host_rate,
                       "median_stderr_over_spot": float(np.median(se / d["spot"])),
                       "accuracy": z_stats(out.numpy(), se, bsm_price(*[d[k] for k in (
                           "spot", "strike", "t", "vol", "r", "q", "option_type")]), d["spot"],
                           "Black-Scholes-Merton closed form")},
               roofline={"bound": "valu", "achieved": ach, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                         "frac": ach / PEAK_FP32_TFLOPS, "traffic": None,
                         "kernel": "qk::k_sim_terminal<float, PHILOX>", "kernel_ms": sim_ms, "flops_per_launch": flops})
    if not args.no_cpu:
        kind, fn = ref_fn("qk_mcm_standard_monte_carlo_price_batch")
        if fn:
            # the whole batch, as the reference's designated CPU configuration
            res["cpu_baseline"] = cpu_baseline(fn, kind, a + [d["seed"]], n, "options",
                                               "standard MC batch entry point, 100000 paths each", full=True)
    return res


def bench_c2(args, qk, qka, dev, name, cus):
    from quantkernel_amd import workloads as wl
    n = args.options or 50000
    d = wl.config2(n)
    t = to_device(d)
    ptrs = ptrs_of(t)
    out = devmem.DevArray(n=n)
    se = devmem.DevArray(n=n)
    # tile shape from the autotuner (timed on C2-like synthetic batches)
    lane, grid, tune_ms = qk.autotune(qk.FAMILY_STEPWISE)
    log(f"autotuned stepwise shape: {lane} paths/lane, {grid} workgroups/CU")

    def call():
        qk.mixed_price_batch_device(n, ptrs, out.ptr, se.ptr)

    for _ in range(args.warmup):
        call()
    devmem.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        call()  # a replayed launch graph
        log(f"step done: batch {qk.last_timing()[0]:.2f} ms")
    devmem.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    sim_ms = direct_kernel_ms(qk, call)
    log(f"simulate kernel {sim_ms:.2f} ms")
    per_step = {0: 14, 1: 18, 2: FLOPS_PER_PATH_STEP, 3: 15}
    ps = d["paths"].astype(np.float64) * d["steps"]
    flops = float(sum(ps[d["model"] == m].sum() * f for m, f in per_step.items()))
    ach = flops / (sim_ms * 1e-3) / 1e12
    price, sea = out.numpy(), se.numpy()
    sev = sea / d["spot"]
    acc = {}
    for m, label in ((0, "gbm_euler"), (1, "gbm_milstein"), (3, "local_vol")):
        idx = np.nonzero(d["model"] == m)[0]
        acc[label] = z_stats(price[idx], sea[idx], bsm_price(*[d[k][idx] for k in (
            "spot", "strike", "t", "vol", "r", "q", "option_type")]), d["spot"][idx],
            "Black-Scholes-Merton closed form; includes the time-discretisation bias")
    idx = np.nonzero(d["model"] == 2)[0][:4096]
    acc["heston"] = z_stats(price[idx], sea[idx], heston_cf_prices(qk, d, idx), d["spot"][idx],
                            "Heston characteristic-function prices (FP64, this library); includes the "
                            "time-discretisation bias")
    res = dict(value=n / (ms * 1e-3), ms_per_step=ms, dtype="fp32",
               config={"workload": f"C2: {n} uneven options, 25% each GBM-Euler / GBM-Milstein / Heston-Euler / "
                                   "constant local vol; paths log-uniform [4096, 1048576], steps log-uniform "
                                   "[16, 512]; one persistent launch, replayed as a launch graph (device-resident "
                                   "entry point)",
                       "options": n, "path_steps": float(ps.sum()), "device": name, "cus": cus,
                       "autotuned_shape": {"paths_per_lane": lane, "workgroups_per_cu": grid},
                       "median_stderr_over_spot": float(np.median(sev)), "max_stderr_over_spot": float(np.max(sev)),
                       "accuracy": acc},
               roofline={"bound": "valu", "achieved": ach, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                         "frac": ach / PEAK_FP32_TFLOPS, "traffic": None,
                         "kernel": "qk::k_sim_stepwise<float, PHILOX>", "kernel_ms": sim_ms, "flops_per_launch": flops})
    if not args.no_cpu:
        res["cpu_baseline"] = c2_cpu_baseline(d)
    return res


C2_NAMES = {0: "qk_mcm_euler_maruyama_price_batch", 1: "qk_mcm_milstein_price_batch",
            2: "qk_mcm_heston_monte_carlo_price_batch", 3: "qk_mcm_local_vol_monte_carlo_price_batch"}


def c2_time(d, idx):
    """Wall seconds of the reference's batch entry points over the options
    idx of the mixed batch: one call per model (the reference has no mixed
    entry point), each OpenMP-parallel over its options."""
    total = 0.0
    for m, name in C2_NAMES.items():
        sel = idx[d["model"][idx] == m]
        if len(sel) == 0:
            continue
        _, fn = ref_fn(name)
        if m == 2:
            arrays = [d[k] for k in KEYS] + [d["option_type"], d["paths"], d["steps"], d["seed"]]
        else:
            arrays = [d[k] for k in ("spot", "strike", "t", "vol", "r", "q", "option_type", "paths", "steps",
                                     "seed")]
        total += time_cpu(fn, arrays, sel)
    return total


def c2_cpu_baseline(d, budget_s=10.0, budget_1t_s=5.0):
    """A measured run of the reference on an evenly strided sample of the
    mixed batch (every m-th option keeps its mix of models, paths and steps),
    on all host threads and on one."""
    if ref_fn(C2_NAMES[0])[1] is None:
        return None
    n = len(d["spot"])
    out = {}
    for threads, budget in ((host_threads(), budget_s), (1, budget_1t_s)):
        set_omp_threads(threads)
        try:
            k = 4 * threads
            idx = np.unique(np.linspace(0, n - 1, k).round().astype(np.int64))
            dt = c2_time(d, idx)
            if dt < budget / 3:
                k = min(n, int(len(idx) * budget / max(dt, 1e-3)))
                idx = np.unique(np.linspace(0, n - 1, k).round().astype(np.int64))
                dt = c2_time(d, idx)
        finally:
            set_omp_threads(host_threads())
        log(f"cpu baseline: {len(idx)} mixed options in {dt:.2f}s on {threads} threads")
        out[threads] = (len(idx), dt)
    threads = host_threads()
    k, dt = out[threads]
    k1, dt1 = out[1]
    return {"value": k / dt, "unit": "prices/s", "cores": threads, "kind": "reference",
            "sample": f"every {n / k:.0f}th of the {n} options ({k}: the batch's mix of models, paths and steps), "
                      f"one reference batch call per model, OpenMP over options ({threads} threads, {cpu_model()}), "
                      f"measured wall time {dt:.2f} s",
            "one_thread": {"value": k1 / dt1, "unit": "prices/s", "cores": 1,
                           "sample": f"every {n / k1:.0f}th option ({k1}), wall time {dt1:.2f} s"}}


def copy_bandwidth(mib=2048, reps=5):
    """What a plain streaming kernel moves on this GPU: the runtime's
    device-to-device copy of mib MiB, read + written bytes per second (best
    of reps synchronous copies after one warm-up)."""
    n = mib << 18  # fp32 elements
    x = devmem.DevArray(n=n, dtype=np.float32)
    y = devmem.DevArray(n=n, dtype=np.float32)
    x.fill_bytes(0)
    y.copy_from(x)
    devmem.synchronize()
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        y.copy_from(x)
        devmem.synchronize()
        best = min(best, time.perf_counter() - t0)
    del x, y
    return 2.0 * n * 4 / best / 1e9


def bench_c3(args, qk, qka, dev, name, cus):
    from quantkernel_amd import workloads as wl
    n = args.options or 1024
    d = wl.config3(n)
    a = [d[k] for k in ("spot", "strike", "t", "vol", "r", "q", "option_type", "paths", "steps")]
    for _ in range(args.warmup):
        qk.longstaff_schwartz_price_batch(*a, d["seed"], basis=1)
    times = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        p, se = qk.longstaff_schwartz_price_batch(*a, d["seed"], basis=1, return_stderr=True)
        times.append(qk.last_timing()[0])
        log(f"step done: device {times[-1]:.2f} ms")
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    dev_ms = float(np.mean(times))
    # algorithmic HBM bytes per path: the forward pass writes S at dates
    # 1 .. N-1 and the terminal cashflow (4 B x dates); the sweeps of dates
    # N-1 .. 2 read S twice (S_j for the exercise rule, S_{j-1} for the next
    # regression) and read and write the FP32 cashflow (16 B per path-date);
    # the sweep of date 1 reads S_1 and the cashflow (8 B): 976 B per path
    paths, dates = 100000, 50
    byts = float(n) * paths * (4 * dates + 16 * (dates - 2) + 8)
    ach = byts / (dev_ms * 1e-3) / 1e9
    copy_gbps = copy_bandwidth()
    res = dict(value=n / (ms * 1e-3), ms_per_step=ms, dtype="fp32",
               config={"workload": f"C3: {n} American puts, Longstaff-Schwartz, {paths} paths x {dates} dates, basis "
                                   "{1,x,x^2,x^3}, x=S/K (Philox, FP32 paths, FP64 normal equations); host entry",
                       "options": n, "device": name, "cus": cus,
                       "median_stderr_over_spot": float(np.median(se / d["spot"]))},
               roofline={"bound": "hbm", "achieved": ach, "peak": 8000.0, "unit": "GB/s", "frac": ach / 8000.0,
                         "traffic": None,
                         "kernel": "LSMC pipeline: k_lsmc_forward, 49 x (k_lsmc_solve, k_lsmc_sweep), k_lsmc_final",
                         "kernel_ms": dev_ms,
                         "bytes_per_launch": byts,
                         "plain_copy_gbps": copy_gbps,
                         "plain_copy": "hipMemcpy device-to-device of 2 GiB on the same box, "
                                       "read + written bytes per second (best of 5)"})
    if not args.no_cpu:
        kind, fn = ref_fn("qk_mcm_longstaff_schwartz_price_batch")
        if fn:
            threads = host_threads()
            res["cpu_baseline"] = cpu_baseline(fn, kind, a + [d["seed"]], n, "options",
                                               "Longstaff-Schwartz batch entry point (its {1, S, S^2} basis), "
                                               "100000 paths x 50 dates each")
            orc = oracle_lib()
            if orc:
                f = orc.orc_crr
                f.restype = ct.c_double
                f.argtypes = [ct.c_double] * 6 + [ct.c_int32] * 3
                m = min(n, 256)
                keys = ("spot", "strike", "t", "vol", "r", "q")
                tree = np.array(threaded(lambda i: f(*[float(d[key][i]) for key in keys],
                                                     int(d["option_type"][i]), 2000, 1), range(m), threads))
                res["cpu_baseline"]["fp64_oracle_check"] = z_stats(
                    p[:m], se[:m], tree, d["spot"][:m],
                    "American CRR binomial tree, 2000 steps (FP64 oracle); LSMC is biased low by its "
                    "sub-optimal exercise rule and the 50 exercise dates")
    return res


# FP64 flops of one Heston CF node: two characteristic-function evaluations
# (about 104 each: complex products, quotients, square root, two exponentials
# and a logarithm, with each transcendental counted once) and about 20 for
# the node's weights, cosine / sine and P1 / P2 terms
FLOPS_PER_CF_NODE = 230
PEAK_FP64_TFLOPS = 78.6


def bench_cf(args, qk, qka, dev, name, cus):
    """Heston characteristic-function prices of 1,048,576 queries from the C1
    parameter box (1024 Gauss-Legendre nodes on [1e-8, 120], the reference
    wrapper's defaults), device-resident."""
    from quantkernel_amd import workloads as wl
    n = args.options or 1 << 20
    d = wl.config1(n, paths=1, steps=1)
    d["option_type"] = (np.arange(n) % 2).astype(np.int32)
    d["integration_steps"] = np.full(n, 1024, np.int32)
    d["integration_limit"] = np.full(n, 120.0)
    keys = KEYS + ("option_type", "integration_steps", "integration_limit")
    t = to_device({k: d[k] for k in keys})
    ptrs = ptrs_of(t)
    out = devmem.DevArray(n=n)

    def call():
        qk.heston_price_cf_batch_device(n, ptrs, out.ptr)

    for _ in range(args.warmup):
        call()
    devmem.synchronize()
    ks = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        call()
        ks.append(qk.last_timing()[1])
        log(f"step done: kernel {ks[-1]:.2f} ms")
    devmem.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    k_ms = float(np.mean(ks))
    flops = float(n) * 1024 * FLOPS_PER_CF_NODE
    ach = flops / (k_ms * 1e-3) / 1e12
    res = dict(value=n / (ms * 1e-3), ms_per_step=ms, dtype="fp64",
               config={"workload": f"Heston characteristic-function prices: {n} calls/puts from the C1 parameter "
                                   "box, 1024 Gauss-Legendre nodes on [1e-8, 120] (FP64; device-resident)",
                       "options": n, "device": name, "cus": cus},
               roofline={"bound": "valu", "achieved": ach, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                         "frac": ach / PEAK_FP64_TFLOPS, "traffic": None, "kernel": "qk::k_heston_cf",
                         "kernel_ms": k_ms, "flops_per_launch": flops})
    if not args.no_cpu:
        kind, fn = ref_fn("qk_cf_heston_price_cf_batch")
        if fn:
            res["cpu_baseline"] = cpu_baseline(fn, kind, [d[k] for k in keys], n, "queries",
                                               "Heston characteristic-function batch entry point, 1024 nodes")
    return res


def bench_c4(args, qk, qka, dev, name, cus):
    from quantkernel_amd import surrogate as sur
    weights = Path(args.weights)
    qk.sur_load(weights)
    qk.sur_configure(spot_every=SPOT_EVERY, spot_paths=SPOT_PATHS, spot_k_se=3.0)
    n = args.options or 1 << 20
    n_out = n // 20
    d = sur.sample_queries(n - n_out, sur.MASTER_QUERY)
    e = sur.sample_queries(n_out, sur.MASTER_QUERY + 1, expand=0.1)
    d = {k: np.concatenate([d[k], e[k]]) for k in d}
    d["paths"] = np.full(n, MC_PATHS, np.int32)
    t = to_device(d)
    ptrs = ptrs_of(t)
    out = devmem.DevArray(n=n)
    src = devmem.DevArray(n=n, dtype=np.int32)

    def step():
        qk.sur_heston_price_batch_device(n, ptrs, out.ptr, src.ptr)

    for _ in range(args.warmup):
        step()
    devmem.synchronize()
    net_ms, mc_ms = [], []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        a, b = qk.sur_last_timing()
        net_ms.append(a)
        mc_ms.append(b)
        log(f"step done: network {a:.2f} ms, MC {b:.2f} ms")
    devmem.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    source = src.numpy()
    served = out.numpy()
    judge = qk.sur_last_judge()
    n_net = int(np.sum(source == 0))
    n_box = int(np.sum(source == 1))
    n_chk = int(np.sum(source == 2))
    n_judge = int(np.sum(source == 4))
    n_spot = int(np.sum((np.arange(n) % SPOT_EVERY == 0) & ((source == 0) | (source == 2) | (source == 4))))
    mc_flops = (float(n_box + n_chk + n_judge) * MC_PATHS + float(n_spot) * SPOT_PATHS) * STEPS * \
        FLOPS_PER_PATH_STEP
    mc_ms_avg = float(np.mean(mc_ms))
    ach = mc_flops / (mc_ms_avg * 1e-3) / 1e12
    net_flops = 2.0 * n * (32 * 256 + 3 * 256 * 256 + 256)
    net_ms_avg = float(np.mean(net_ms))
    g = judge[16]
    judge_summary = {"spot_checks": int(g[0]),
                     "implied_rms_err_over_spot": float(np.sqrt(max(g[1], 0.0))),
                     "rms_err_over_spot_3se_upper": float(np.sqrt(max(g[1] + 3 * g[2], 0.0))),
                     "regions_rejected": int(judge[:16, 3].sum()), "batch_rejected": bool(g[3])}

    # plain GPU MC at the same accuracy, and accuracy of both, same run
    plain_rate, plain_sim_ms, _, _ = plain_mc_rate(qk, qka, d, 4096)
    log(f"plain MC: {plain_rate:.0f} prices/s")
    na = args.accuracy_queries
    acc = {}
    if na > 0:
        # every (n/na)-th query of the timed batch: its 5% widened-box share
        # and the answers as served (network and MC fallback alike)
        idx = np.unique(np.linspace(0, n - 1, na).round().astype(np.int64))
        a = {k: v[idx] for k, v in d.items()}
        ref = dict(a, seed=a["seed"] ^ np.uint64(0xA5A5A5A55A5A5A5A))  # independent paths
        ref_p, ref_se = sur.mc_price(qk, ref, 1 << 22, chunk=1024)
        mc, mc_se = qk.heston_monte_carlo_price_batch(*[a[k] for k in KEYS], a["option_type"], a["paths"],
                                                      a["steps"], a["seed"], return_stderr=True)
        sp, ss = served[idx], source[idx]
        s_ = a["spot"]
        es, em = (sp - ref_p) / s_, (mc - ref_p) / s_
        ref_ms = float(np.mean((ref_se / s_) ** 2))
        net = ss == 0
        big = ref_p >= 0.01 * s_
        rel_s = np.abs(sp[big] - ref_p[big]) / ref_p[big]
        rel_m = np.abs(mc[big] - ref_p[big]) / ref_p[big]
        rms = lambda x: float(np.sqrt(np.mean(x ** 2))) if len(x) else None  # noqa: E731
        acc = {"queries": int(len(idx)), "sample": f"every {n / len(idx):.0f}th query of the timed batch",
               "outside_box": int(np.sum(ss == 1)), "reference_paths": 1 << 22,
               "reference_median_se_over_spot": float(np.median(ref_se / s_)),
               "served_rmse_over_spot": rms(es),
               "served_rmse_over_spot_net_of_reference_noise": float(np.sqrt(max(rms(es) ** 2 - ref_ms, 0.0))),
               "served_p99_abs_err_over_spot": float(np.quantile(np.abs(es), 0.99)),
               "network_answers_rmse_over_spot": rms(es[net]), "mc_answers_rmse_over_spot": rms(es[~net]),
               "plain_mc_rmse_over_spot": rms(em),
               "plain_mc_p99_abs_err_over_spot": float(np.quantile(np.abs(em), 0.99)),
               "price_relative_floor": "prices >= 1% of spot", "price_relative_queries": int(big.sum()),
               "served_rms_rel_err": rms(rel_s), "served_p99_rel_err": float(np.quantile(rel_s, 0.99)),
               "plain_mc_rms_rel_err": rms(rel_m), "plain_mc_p99_rel_err": float(np.quantile(rel_m, 0.99))}
        log(json.dumps(acc))
    value = n / (ms * 1e-3)
    res = dict(value=value, ms_per_step=ms, dtype="fp32",
               config={"workload": f"Heston surrogate + MC fallback: {n} call queries (95% in the training box, "
                                   f"5% from the box widened 10% per side); fallback and spot checks by plain MC "
                                   f"({MC_PATHS} / {SPOT_PATHS} paths x {STEPS} steps, Philox, FP32 paths)",
                       "queries": n, "answered_by_network": n_net, "mc_fallback_out_of_box": n_box,
                       "mc_fallback_spot_check": n_chk, "mc_fallback_batch_judge": n_judge,
                       "spot_checked": n_spot, "batch_judge": judge_summary, "device": name, "cus": cus,
                       "weights": weights.name, "network_ms": net_ms_avg, "mc_ms": mc_ms_avg,
                       "network_mfma_tflops": net_flops / (net_ms_avg * 1e-3) / 1e12,
                       "network_mfma_frac_bf16_peak": net_flops / (net_ms_avg * 1e-3) / 1e12 / PEAK_BF16_TFLOPS,
                       "plain_gpu_mc_prices_per_s": plain_rate, "gain_vs_plain_gpu_mc": value / plain_rate,
                       "accuracy": acc},
               roofline={"bound": "valu", "achieved": ach, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                         "frac": ach / PEAK_FP32_TFLOPS, "traffic": None,
                         "kernel": "qk::k_sim_stepwise<float, PHILOX> (MC fallback + spot checks)",
                         "kernel_ms": mc_ms_avg, "flops_per_launch": mc_flops})
    if not args.no_cpu:
        res["cpu_baseline"] = heston_cpu_baseline(
            d, noun="queries", note="; the reference has no surrogate, so this is its MC at the fallback's accuracy")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=["c0", "c1", "c2", "c3", "c4", "cf"], default=None)
    ap.add_argument("--options", type=int, default=0)
    ap.add_argument("--weights", default=str(ROOT / "quant-kernel_amd" / "quantkernel_amd" / "heston_surrogate.qkw"))
    ap.add_argument("--accuracy-queries", type=int, default=2048)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--calibration-copy", type=int, default=0, metavar="MIB",
                    help="only run one device-to-device copy of MIB MiB on the GPU (the runtime's copy "
                         "kernel): the known-size traffic that scales rocprofv3 FETCH_SIZE / WRITE_SIZE")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(options are independent and would shard across GPUs by option; "
              "the value reported is one GPU's throughput).", file=sys.stderr)
        sys.exit(2)

    if args.calibration_copy > 0:
        n = args.calibration_copy << 18  # fp32 elements
        x = devmem.DevArray(n=n, dtype=np.float32)
        y = devmem.DevArray(n=n, dtype=np.float32)
        x.fill_bytes(0)
        y.copy_from(x)
        devmem.synchronize()
        print(json.dumps({"calibration_copy_bytes_read": n * 4, "calibration_copy_bytes_written": n * 4}))
        return

    import quantkernel_amd as qka

    qk = qka.QuantKernel()
    name, cus = qk.device_info()
    dev = None  # device memory comes from devmem
    workload = args.workload or ("c4" if Path(args.weights).exists() else "c1")
    res = {"c0": bench_c0, "c1": bench_c1, "c2": bench_c2, "c3": bench_c3, "c4": bench_c4,
           "cf": bench_cf}[workload](
        args, qk, qka, dev, name, cus)
    # profiled constants, not measured in this run: rocprofv3 --pmc runs of
    # this command (one counter set per run), kept under profiles/
    if PMC_TRAFFIC.exists():
        entry = json.loads(PMC_TRAFFIC.read_text()).get(workload)
        if entry and entry.get("kernel") == res["roofline"]["kernel"]:
            res["roofline"]["traffic"] = entry["traffic_bytes"]
            res["roofline"]["traffic_source"] = (f"profiled constant: rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE runs of "
                                                 f"this command, profiles/{PMC_TRAFFIC.name}; not measured in this run")
    if PMC_VALU.exists():
        entry = json.loads(PMC_VALU.read_text()).get(workload)
        if entry and entry.get("kernel") == res["roofline"]["kernel"]:
            res["roofline"]["valu_issue_frac"] = entry["valu_issue_frac"]
            res["roofline"]["valu_issue_source"] = (f"profiled constant: {entry['definition']}; profiles/"
                                                    f"{PMC_VALU.name}; not measured in this run")
    line = {"metric": METRIC, "value": res["value"], "unit": "prices/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": res["ms_per_step"], "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": res["dtype"], "data": "synthetic", "config": res["config"],
            "roofline": res["roofline"]}
    if "cpu_baseline" in res:
        line["cpu_baseline"] = res["cpu_baseline"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
