"""Training and evaluation driver of the Heston price surrogate.

    python -m quantkernel_amd.surrogate train [--samples 2000000] [--paths 262144]
        [--epochs 40] [--out quantkernel_amd/heston_surrogate.qkw] [--report r.json]

1. Labels: price/S of European calls from the library's own Heston
   full-truncation Euler kernel (256 steps, Philox, FP32 paths, FP64 sums) at
   ``--paths`` paths, on seeded queries drawn uniformly from the training box
   (the C1 parameter ranges: K/S in [.7, 1.3], T in [.05, 2], r in [0, .08],
   q = 0, kappa in [.5, 5], theta in [.01, .16], sigma in [.1, 1],
   rho in [-.9, -.1], v0 in [.01, .16]).
2. Training: the library's MFMA forward/backward kernels and Adam on the GPU.
3. Evaluation: fresh queries, error of the network against MC at
   ``--ref-paths`` paths, per unit of spot.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import math
import sys
import time
from pathlib import Path

import numpy as np

from . import workloads as wl
from .engine import QuantKernel

DEFAULT_WEIGHTS = Path(__file__).resolve().parent / "heston_surrogate.qkw"
STEPS = 256
BOX_LO = np.array([math.log(0.7), 0.05, 0.0, 0.5, 0.01, 0.1, -0.9, 0.01])
BOX_HI = np.array([math.log(1.3), 2.0, 0.08, 5.0, 0.16, 1.0, -0.1, 0.16])
MASTER_TRAIN, MASTER_TEST, MASTER_QUERY = 0x5EED5000, 0x5EED6000, 0x5EED7000


# features that the Heston model needs non-negative: T, theta, v0
_NONNEG = np.array([False, True, False, False, True, False, False, True])


def sample_queries(n: int, master: int, expand: float = 0.0, start: int = 0) -> dict:
    """n Heston calls with features uniform in the box (each side widened by
    ``expand`` of its width, the lower sides of T, theta and v0 no further
    than 0, so every query is a valid Heston option); spot uniform in
    [50, 150]."""
    lo = BOX_LO - expand * (BOX_HI - BOX_LO)
    hi = BOX_HI + expand * (BOX_HI - BOX_LO)
    lo = np.where(_NONNEG, np.maximum(lo, 0.0), lo)
    u = np.stack([wl._uniform(master, n, c, 0.0, 1.0, start) for c in range(8)], axis=1)
    f = lo + (hi - lo) * u
    spot = wl._uniform(master, n, 8, 50.0, 150.0, start)
    return dict(spot=spot, strike=spot * np.exp(f[:, 0]), t=f[:, 1], r=f[:, 2], q=np.zeros(n), kappa=f[:, 3],
                theta=f[:, 4], sigma=f[:, 5], rho=f[:, 6], v0=f[:, 7], option_type=np.zeros(n, np.int32),
                steps=np.full(n, STEPS, np.int32), seed=wl.seeds(master, n, start))


def features(d: dict) -> np.ndarray:
    return np.stack([np.log(d["strike"] / d["spot"]), d["t"], d["r"], d["kappa"], d["theta"], d["sigma"],
                     d["rho"], d["v0"]], axis=1)


KEYS = ("spot", "strike", "t", "r", "q", "v0", "kappa", "theta", "sigma", "rho")


def mc_price(qk: QuantKernel, d: dict, paths: int, chunk: int = 65536, log=None):
    """Plain GPU MC price and standard error for every query of d."""
    n = d["spot"].shape[0]
    price, se = np.empty(n), np.empty(n)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        sl = {k: v[s:e] for k, v in d.items()}
        p = np.full(e - s, paths, np.int32)
        price[s:e], se[s:e] = qk.heston_monte_carlo_price_batch(*[sl[k] for k in KEYS], sl["option_type"], p,
                                                                sl["steps"], sl["seed"], return_stderr=True)
        if log:
            log(f"labels {e}/{n}")
    return price, se


def train(qk: QuantKernel, samples: int, paths: int, epochs: int, batch: int, lr: float, lr_final: float,
          seed: int, log=print) -> dict:
    t0 = time.perf_counter()
    d = sample_queries(samples, MASTER_TRAIN)
    price, se = mc_price(qk, d, paths, log=log)
    y = price / d["spot"]
    t_label = time.perf_counter() - t0
    log(f"labels done in {t_label:.1f}s; median SE/S {np.median(se / d['spot']):.2e}")
    qk.sur_init(seed)
    t1 = time.perf_counter()
    losses = []
    # train in slices of epochs so progress is printed as it goes
    done = 0
    while done < epochs:
        k = min(5, epochs - done)
        # one cosine schedule over all epochs: split it into slices
        a = done / epochs
        b = (done + k) / epochs
        lr_a = lr_final + 0.5 * (lr - lr_final) * (1 + math.cos(math.pi * a))
        lr_b = lr_final + 0.5 * (lr - lr_final) * (1 + math.cos(math.pi * b))
        losses += list(qk.sur_train(features(d), y, epochs=k, batch=batch, lr=lr_a, lr_final=lr_b,
                                    seed=seed + done))
        done += k
        log(f"epoch {done}/{epochs}: mse {losses[-1]:.3e} (rmse {math.sqrt(losses[-1]):.2e})")
    t_train = time.perf_counter() - t1
    return dict(samples=samples, label_paths=paths, label_seconds=t_label, train_seconds=t_train,
                epochs=epochs, batch=batch, lr=lr, lr_final=lr_final, epoch_mse=losses,
                label_median_se_over_spot=float(np.median(se / d["spot"])))


def evaluate(qk: QuantKernel, n: int, ref_paths: int, log=print) -> dict:
    """Network error on fresh in-box calls and puts against MC at ref_paths."""
    d = sample_queries(n, MASTER_TEST)
    d["option_type"] = (np.arange(n) % 2).astype(np.int32)
    ref, se = mc_price(qk, d, ref_paths, chunk=4096, log=log)
    qk.sur_configure(spot_every=0)
    sur, src = qk.sur_heston_price_batch(*[d[k] for k in KEYS], d["option_type"], np.full(n, 65536, np.int32),
                                         d["steps"], d["seed"])
    qk.sur_configure()
    err = (sur - ref) / d["spot"]
    ref_se = se / d["spot"]
    out = dict(n=n, ref_paths=ref_paths, ref_median_se_over_spot=float(np.median(ref_se)),
               rmse_over_spot=float(np.sqrt(np.mean(err ** 2))),
               rmse_over_spot_net_of_ref_noise=float(np.sqrt(max(np.mean(err ** 2) - np.mean(ref_se ** 2), 0.0))),
               p99_abs_err_over_spot=float(np.quantile(np.abs(err), 0.99)),
               max_abs_err_over_spot=float(np.max(np.abs(err))), all_network=bool(np.all(src == 0)))
    log(json.dumps(out))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--samples", type=int, default=2_000_000)
    t.add_argument("--paths", type=int, default=262_144)
    t.add_argument("--epochs", type=int, default=60)
    t.add_argument("--batch", type=int, default=1024)
    t.add_argument("--lr", type=float, default=1e-3)
    t.add_argument("--lr-final", type=float, default=1e-5)
    t.add_argument("--seed", type=int, default=1)
    t.add_argument("--eval", type=int, default=4096)
    t.add_argument("--ref-paths", type=int, default=4_194_304)
    t.add_argument("--out", default=str(DEFAULT_WEIGHTS))
    t.add_argument("--report", default=None)
    e = sub.add_parser("eval")
    e.add_argument("--weights", default=str(DEFAULT_WEIGHTS))
    e.add_argument("--eval", type=int, default=4096)
    e.add_argument("--ref-paths", type=int, default=4_194_304)
    e.add_argument("--report", default=None)
    a = ap.parse_args(argv)

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    qk = QuantKernel()
    command = "python -m quantkernel_amd.surrogate " + " ".join(sys.argv[1:] if argv is None else argv)
    if a.cmd == "train":
        rep = dict(command=command)
        rep.update(train(qk, a.samples, a.paths, a.epochs, a.batch, a.lr, a.lr_final, a.seed, log))
        qk.sur_save(a.out)
        rep["weights_sha256"] = hashlib.sha256(Path(a.out).read_bytes()).hexdigest()
        log(f"weights written to {a.out}")
    else:
        qk.sur_load(a.weights)
        rep = dict(command=command)
    rep["evaluation"] = evaluate(qk, a.eval, a.ref_paths, log) if a.eval > 0 else None
    if a.report:
        Path(a.report).write_text(json.dumps(rep, indent=1))
    print(json.dumps(rep.get("evaluation")))


if __name__ == "__main__":
    main()
