"""Seeded synthetic option batches for the benchmark workloads.

Every field of option i comes from SplitMix64 (Steele, Lea, Flood 2014) of
``master + i * 16 + field``, so any prefix of a batch is reproducible on its
own and the CPU baseline can time exactly the first k options of the GPU's
batch.  MC seeds are ``splitmix64(master ^ (i + 1) * golden)``.
"""
from __future__ import annotations

import numpy as np

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)
_GOLD = np.uint64(0x9E3779B97F4A7C15)


def splitmix64(x: np.ndarray) -> np.ndarray:
    x = np.asarray(x, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = x + _GOLD
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))


def _uniform(master: int, n: int, field: int, lo: float, hi: float, start: int = 0) -> np.ndarray:
    idx = np.arange(start, start + n, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = splitmix64(np.uint64(master) + idx * np.uint64(16) + np.uint64(field))
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return lo + (hi - lo) * u


def _loguniform_int(master, n, field, lo, hi, start=0):
    u = _uniform(master, n, field, np.log(lo), np.log(hi), start)
    return np.rint(np.exp(u)).astype(np.int32)


def seeds(master: int, n: int, start: int = 0) -> np.ndarray:
    idx = np.arange(start + 1, start + n + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return splitmix64(np.uint64(master) ^ (idx * _GOLD))


MASTER = {0: 0x5EED0000, 1: 0x5EED0001, 2: 0x5EED0002, 3: 0x5EED0003, 4: 0x5EED0004}


def gbm_common(master, n, start=0, k_lo=0.7, k_hi=1.3):
    s = _uniform(master, n, 0, 50.0, 150.0, start)
    return dict(
        spot=s,
        strike=s * _uniform(master, n, 1, k_lo, k_hi, start),
        t=_uniform(master, n, 2, 0.05, 2.0, start),
        vol=_uniform(master, n, 3, 0.1, 0.6, start),
        r=_uniform(master, n, 4, 0.0, 0.08, start),
        q=_uniform(master, n, 5, 0.0, 0.04, start),
        option_type=(_uniform(master, n, 6, 0.0, 1.0, start) >= 0.5).astype(np.int32),
    )


def heston_params(master, n, start=0):
    return dict(
        kappa=_uniform(master, n, 7, 0.5, 5.0, start),
        theta=_uniform(master, n, 8, 0.01, 0.16, start),
        sigma=_uniform(master, n, 9, 0.1, 1.0, start),
        rho=_uniform(master, n, 10, -0.9, -0.1, start),
        v0=_uniform(master, n, 11, 0.01, 0.16, start),
    )


def config0(n=4096, paths=100_000, start=0):
    """GBM European calls/puts, terminal antithetic estimator."""
    m = MASTER[0]
    d = gbm_common(m, n, start)
    d["paths"] = np.full(n, paths, dtype=np.int32)
    d["seed"] = seeds(m, n, start)
    return d


def config1(n=16384, paths=65536, steps=256, start=0):
    """Heston full-truncation Euler European calls (q = 0)."""
    m = MASTER[1]
    g = gbm_common(m, n, start)
    d = dict(spot=g["spot"], strike=g["strike"], t=g["t"], r=g["r"], q=np.zeros(n))
    d.update(heston_params(m, n, start))
    d["option_type"] = np.zeros(n, dtype=np.int32)
    d["paths"] = np.full(n, paths, dtype=np.int32)
    d["steps"] = np.full(n, steps, dtype=np.int32)
    d["seed"] = seeds(m, n, start)
    return d


def config2(n=50_000, start=0, p_lo=4096, p_hi=1_048_576, s_lo=16, s_hi=512):
    """Uneven mixed-model batch: 25% each GBM-Euler, GBM-Milstein, Heston,
    constant local vol; log-uniform path and step counts."""
    m = MASTER[2]
    d = gbm_common(m, n, start)
    d.update(heston_params(m, n, start))
    d["model"] = np.floor(_uniform(m, n, 12, 0.0, 4.0, start)).astype(np.int32).clip(0, 3)
    d["paths"] = _loguniform_int(m, n, 13, p_lo, p_hi, start)
    d["steps"] = _loguniform_int(m, n, 14, s_lo, s_hi, start)
    d["seed"] = seeds(m, n, start)
    return d


def config3(n=1024, paths=100_000, steps=50, start=0):
    """Longstaff-Schwartz American puts, K/S ~ U[0.8, 1.2]."""
    m = MASTER[3]
    d = gbm_common(m, n, start, 0.8, 1.2)
    d["option_type"] = np.ones(n, dtype=np.int32)
    d["paths"] = np.full(n, paths, dtype=np.int32)
    d["steps"] = np.full(n, steps, dtype=np.int32)
    d["seed"] = seeds(m, n, start)
    return d
