"""GPU Heston characteristic-function prices (qk_cf_heston_price_cf[_batch])
against the reference build and the CPU oracle on the same inputs, the
reference's validation branches, put-call parity, and agreement with the
GPU Heston MC at many paths."""
import math

import numpy as np
import pytest

from quantkernel_amd import workloads as wl

pytestmark = pytest.mark.gpu

HST = ("spot", "strike", "t", "r", "q", "v0", "kappa", "theta", "sigma", "rho")


def _batch(n, seed=0x5EED0011):
    d = wl.config1(n, paths=1, steps=1)
    d["option_type"] = (np.arange(n) % 2).astype(np.int32)
    d["q"] = np.linspace(0.0, 0.04, n)
    d["steps"] = np.array([(16, 64, 512, 1024, 2048)[i % 5] for i in range(n)], np.int32)
    d["limit"] = np.array([(50.0, 120.0, 200.0)[i % 3] for i in range(n)])
    return d


def test_matches_reference_build_and_oracle(qk, orc, ref):
    n = 200
    d = _batch(n)
    got = qk.heston_price_cf_batch(*[d[k] for k in HST], d["option_type"], d["steps"], d["limit"])
    worst_ref = worst_orc = 0.0
    for i in range(n):
        row = [float(d[k][i]) for k in HST]
        args = (*row, int(d["option_type"][i]), int(d["steps"][i]), float(d["limit"][i]))
        w_ref = ref.lib.qk_cf_heston_price_cf(*args)
        w_orc = orc.lib.orc_heston_cf(*args)
        worst_ref = max(worst_ref, abs(got[i] - w_ref) / row[0])
        worst_orc = max(worst_orc, abs(got[i] - w_orc) / row[0])
    print("worst |gpu - ref| / S", worst_ref, "worst |gpu - oracle| / S", worst_orc)
    assert worst_ref < 1e-10 and worst_orc < 1e-10


def test_validation_branches_and_parity(qk, orc):
    base = [100.0, 95.0, 1.0, 0.03, 0.01, 0.04, 1.5, 0.05, 0.5, -0.6]
    cases = []
    for j, bad in [(0, -1.0), (1, 0.0), (2, -0.5), (5, -0.01), (6, 0.0), (7, -0.1), (8, 0.0), (9, -1.0),
                   (9, 1.0), (0, math.nan), (3, math.inf)]:
        c = list(base)
        c[j] = bad
        cases.append((c, 0, 1024, 120.0))
    cases += [(base, 2, 1024, 120.0), (base, 0, 15, 120.0), (base, 0, 16, 120.0), (base, 0, 1024, 1e-6),
              (base[:2] + [0.0] + base[3:], 1, 1024, 120.0), (base[:2] + [1e-13] + base[3:], 0, 1024, 120.0)]
    cols = [np.array([c[0][j] for c in cases]) for j in range(10)]
    got = qk.heston_price_cf_batch(*cols, np.array([c[1] for c in cases], np.int32),
                                   np.array([c[2] for c in cases], np.int32), np.array([c[3] for c in cases]))
    for i, (c, ot, st, lim) in enumerate(cases):
        w = orc.lib.orc_heston_cf(*c, ot, st, lim)
        print(c, ot, st, lim, got[i], w)
        assert (math.isnan(got[i]) and math.isnan(w)) or abs(got[i] - w) < 1e-10 * c[0], (c, ot, st, lim)
    # put-call parity holds exactly by construction
    call = qk.heston_price_cf(*base, 0)
    put = qk.heston_price_cf(*base, 1)
    assert call - put == pytest.approx(100 * math.exp(-0.01) - 95 * math.exp(-0.03), abs=1e-10)
    with pytest.raises(ValueError):
        qk.heston_price_cf(*base[:6], 0.0, *base[7:], 0)


def test_agrees_with_gpu_heston_mc(qk):
    n = 8
    d = wl.config1(n, paths=1 << 20, steps=512)
    d["option_type"] = (np.arange(n) % 2).astype(np.int32)
    cf = qk.heston_price_cf_batch(*[d[k] for k in HST], d["option_type"], np.full(n, 2048, np.int32),
                                  np.full(n, 200.0))
    mc, se = qk.heston_monte_carlo_price_batch(*[d[k] for k in HST], d["option_type"], d["paths"], d["steps"],
                                               d["seed"], return_stderr=True)
    z = (mc - cf) / se
    print("z", np.round(z, 2), "cf", cf)
    # full-truncation Euler at 512 steps: bias well below 4 SE + 2e-3 S here
    assert np.all(np.abs(mc - cf) < 4 * se + 2e-3 * d["spot"])


def test_throughput_large_batch(qk):
    n = 65536
    d = _batch(n)
    steps = np.full(n, 1024, np.int32)
    qk.heston_price_cf_batch(*[d[k] for k in HST], d["option_type"], steps, np.full(n, 120.0))
    got = qk.heston_price_cf_batch(*[d[k] for k in HST], d["option_type"], steps, np.full(n, 120.0))
    total, kernel = qk.last_timing()
    print(f"{n} Heston CF prices (1024 nodes): kernel {kernel:.2f} ms -> {n / kernel * 1e3:.3e} prices/s")
    assert np.all(np.isfinite(got)) and kernel > 0


def test_device_entry_equals_host_entry(qk):
    from hip_mem import DevArray, to_device
    n = 300
    d = _batch(n)
    d["integration_steps"], d["integration_limit"] = d["steps"], d["limit"]
    host = qk.heston_price_cf_batch(*[d[k] for k in HST], d["option_type"], d["steps"], d["limit"])
    t = to_device(d, HST + ("option_type", "integration_steps", "integration_limit"))
    out = DevArray(n=n)
    qk.heston_price_cf_batch_device(n, {k: v.ptr for k, v in t.items()}, out.ptr)
    assert np.array_equal(out.numpy(), host, equal_nan=True)
