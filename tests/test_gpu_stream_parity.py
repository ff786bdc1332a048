"""The production Philox streams of the path-by-path and terminal models
against the reference's own mt19937 stream, over 1024 options each.

Validation mode (the mt19937_64 + normal_distribution stream, FP64 paths)
equals the reference to 1e-6 relative, so it stands in for the reference.
The Philox estimate and the validation estimate of one option are
independent, so their difference over the combined SE is N(0, 1): the
fraction within 3 SE is held to its one-sided 3-sigma binomial bound and
the mean z to 4 / sqrt(n).  Step counts cover every residue mod 8 (eight
GBM steps per Philox block); path counts both parities (the odd extra path
of the antithetic estimator of standard MC); control variates runs the
general terminal kernel, standard MC the pair-only one.
"""
import math

import numpy as np
import pytest

from quantkernel_amd import workloads as wl

pytestmark = pytest.mark.gpu

RNG_PHILOX, RNG_F64 = 0, 1
P3 = math.erf(3 / math.sqrt(2))


def frac_bound(n):
    return P3 - 3.0 * math.sqrt(P3 * (1 - P3) / n)


def _z_stats(p1, s1, p2, s2, spot):
    # combined SE, plus the FP32-path allowance the Philox kernels are held
    # to against the FP64 oracle on their own stream (2e-6 S): control
    # variates on deep in-the-money options have a residual SE near zero,
    # where only the FP32 / FP64 path rounding separates the two estimates
    sd = np.sqrt(s1 * s1 + s2 * s2 + (2e-6 * spot) ** 2)
    ok = (np.minimum(p1, p2) / np.where(sd > 0, sd, 1.0)) ** 2 >= 30  # about 30+ in-the-money paths
    z = (p1[ok] - p2[ok]) / sd[ok]
    return z, float(np.mean(np.abs(z) < 3.0))


def _both_modes(qk, fn):
    p1, s1 = fn()
    qk.set_rng_mode(RNG_F64)
    try:
        p2, s2 = fn()
    finally:
        qk.set_rng_mode(RNG_PHILOX)
    return p1, s1, p2, s2


@pytest.mark.parametrize("model", ["euler_maruyama", "milstein", "local_vol_monte_carlo"])
def test_stepwise_philox_vs_reference_fraction_within_3se(qk, model):
    n = 1024
    d = wl.config0(n, paths=4096, start=20000)
    d["steps"] = (9 + np.arange(n) % 8).astype(np.int32)  # 9 .. 16: every residue mod 8
    args = [d[k] for k in ("spot", "strike", "t", "vol", "r", "q", "option_type", "paths", "steps")]
    fn = getattr(qk, f"{model}_price_batch")
    p1, s1, p2, s2 = _both_modes(qk, lambda: fn(*args, d["seed"], return_stderr=True))
    z, within = _z_stats(p1, s1, p2, s2, d["spot"])
    print(f"{model}: {z.size} options, mean z {z.mean():.3f}, rms {np.sqrt(np.mean(z * z)):.3f}, "
          f"within 3 SE {within:.4f} (bound {frac_bound(z.size):.4f})")
    assert z.size > 800
    assert abs(z.mean()) < 4.0 / math.sqrt(z.size)
    assert within >= frac_bound(z.size)


@pytest.mark.parametrize("method", ["standard_monte_carlo", "control_variates"])
def test_terminal_philox_vs_reference_fraction_within_3se(qk, method):
    n = 1024
    d = wl.config0(n, paths=8192, start=30000)
    d["paths"] = (8192 + np.arange(n) % 2).astype(np.int32)  # even and odd path counts
    args = [d[k] for k in ("spot", "strike", "t", "vol", "r", "q", "option_type", "paths")]
    fn = getattr(qk, f"{method}_price_batch")
    p1, s1, p2, s2 = _both_modes(qk, lambda: fn(*args, d["seed"], return_stderr=True))
    z, within = _z_stats(p1, s1, p2, s2, d["spot"])
    print(f"{method}: {z.size} options, mean z {z.mean():.3f}, rms {np.sqrt(np.mean(z * z)):.3f}, "
          f"within 3 SE {within:.4f} (bound {frac_bound(z.size):.4f})")
    assert z.size > 800
    assert abs(z.mean()) < 4.0 / math.sqrt(z.size)
    assert within >= frac_bound(z.size)
