"""Benchmark of the Heston Monte Carlo batch path on one MI355X.

Workload: 16384 Heston full-truncation Euler European calls, 65536 paths and
256 steps each (seeded synthetic parameters, see quantkernel_amd.workloads).
A step is one pass of the whole batch through the library's device-resident
entry point (inputs already in HBM): prepare -> tile scan -> persistent
Philox/FP32 path kernel -> finalize.  The CPU baseline times the reference's
own batch entry point (a build of its sources, -O3 -fopenmp) on the first
options of the same batch.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import ctypes as ct
import json
import os
import subprocess
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "quant-kernel_amd"))

PEAK_FP32_TFLOPS = 157.3       # MI355X published FP32 vector peak
FLOPS_PER_PATH_STEP = 37       # Heston step incl. two Box-Muller normals (algorithmic count)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_info():
    model = ""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return model, os.cpu_count() or 1


def cpu_baseline(d, budget_s=12.0):
    """Times the reference build (or, without it, the oracle port) on the
    first k options of the batch, full paths and steps each."""
    ref = ROOT / "oracle" / "_ref" / "libqkref.so"
    kind = "reference"
    if ref.exists():
        lib = ct.CDLL(str(ref))
        fn = lib.qk_mcm_heston_monte_carlo_price_batch
    else:
        kind = "port"
        lib = ct.CDLL(str(ROOT / "oracle" / "liboracle.so"))
        fn = lib.orc_heston_batch
    PD, PI, PU = ct.POINTER(ct.c_double), ct.POINTER(ct.c_int32), ct.POINTER(ct.c_uint64)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    keys = ("spot", "strike", "t", "r", "q", "v0", "kappa", "theta", "sigma", "rho")

    def run(k):
        arrs = [np.ascontiguousarray(d[key][:k]) for key in keys]
        ot = np.ascontiguousarray(d["option_type"][:k])
        p = np.ascontiguousarray(d["paths"][:k])
        st = np.ascontiguousarray(d["steps"][:k])
        sd = np.ascontiguousarray(d["seed"][:k])
        out = np.empty(k)
        args = [a.ctypes.data_as(PD) for a in arrs] + [ot.ctypes.data_as(PI), p.ctypes.data_as(PI),
                                                       st.ctypes.data_as(PI), sd.ctypes.data_as(PU),
                                                       ct.c_int32(k), out.ctypes.data_as(PD)]
        t0 = time.perf_counter()
        fn(*args)
        return time.perf_counter() - t0, out

    k = max(1, threads)
    dt, _ = run(k)
    log(f"cpu baseline probe: {k} options in {dt:.2f}s")
    if dt < budget_s / 3:
        k2 = int(k * budget_s / max(dt, 1e-3))
        k2 = max(k, (k2 // threads) * threads)
        dt, _ = run(k2)
        k = k2
        log(f"cpu baseline: {k} options in {dt:.2f}s")
    return {"value": k / dt, "unit": "prices/s", "cores": threads, "kind": kind,
            "sample": f"first {k} options of the batch, 65536 paths x 256 steps each, "
                      f"OpenMP over options ({threads} threads, {cpu_info()[0]})"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--options", type=int, default=16384)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path "
              "(options shard independently; see the design notes).", file=sys.stderr)
        sys.exit(2)

    import torch

    import quantkernel_amd as qka
    from quantkernel_amd import workloads as wl

    qk = qka.QuantKernel()
    name, cus = qk.device_info()
    n = args.options
    d = wl.config1(n)
    d["model"] = np.full(n, qka.QK_MODEL_HESTON_EULER, dtype=np.int32)
    d["vol"] = np.zeros(n)
    dev = torch.device("cuda:0")
    tens = {k: torch.from_numpy(np.ascontiguousarray(v)).to(dev) for k, v in d.items()}
    ptrs = {k: t.data_ptr() for k, t in tens.items()}
    out = torch.empty(n, dtype=torch.float64, device=dev)
    se = torch.empty(n, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()

    def step():
        qk.mixed_price_batch_device(n, ptrs, out.data_ptr(), se.data_ptr())

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    sims = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        sims.append(qk.last_timing()[1])
        log(f"step done: sim kernel {sims[-1]:.2f} ms")
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_step = wall * 1e3 / args.steps
    price = out.cpu().numpy()
    stderr = se.cpu().numpy()
    assert np.all(np.isfinite(price)), "non-finite prices"

    sim_ms = float(np.mean(sims))
    flops = float(n) * 65536 * 256 * FLOPS_PER_PATH_STEP
    achieved = flops / (sim_ms * 1e-3) / 1e12
    result = {
        "metric": "Heston prices/sec at <=1e-3 rel err: surrogate+fallback vs plain MC, 1 GPU",
        "value": n / (ms_step * 1e-3),
        "unit": "prices/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": f"plain GPU MC, Heston full-truncation Euler calls, {n} options x 65536 "
                               f"paths x 256 steps (Philox, FP32 paths, FP64 sums)",
                   "options": n, "paths": 65536, "time_steps": 256, "device": name, "cus": cus,
                   "median_stderr_over_spot": float(np.median(stderr / d["spot"]))},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32_TFLOPS, "traffic": None,
                     "kernel": "k_sim_stepwise<float,PHILOX>", "kernel_ms": sim_ms,
                     "flops_per_launch": flops},
    }
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(d)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
