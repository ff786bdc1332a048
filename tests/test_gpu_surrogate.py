"""GPU tests of the Heston surrogate: forward and gradients against the CPU
MLP oracle, training, weights file, and pricing with MC fallback."""
import math
from pathlib import Path

import numpy as np
import pytest

import mlp_oracle as mo

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
WEIGHTS = ROOT / "quant-kernel_amd" / "quantkernel_amd" / "heston_surrogate.qkw"


def _box(n, seed):
    rng = np.random.default_rng(seed)
    return mo.BOX_LO + (mo.BOX_HI - mo.BOX_LO) * rng.random((n, 8))


def test_forward_and_gradients_match_cpu_oracle(qk):
    p = mo.init_params(5)
    qk.sur_set_params(p)
    f = _box(300, 1)
    t = np.random.default_rng(2).random(300) * 0.3
    y_cpu, loss_cpu, g_cpu = mo.forward_backward(p, f, t, exact=True)
    y_gpu = qk.sur_predict(f) - mo.control(f)  # predictions are control + network
    assert np.max(np.abs(y_gpu - y_cpu)) < 2e-4 * max(1.0, np.max(np.abs(y_cpu)))
    g_gpu, loss_gpu = qk.sur_gradients(f, t)
    assert loss_gpu == pytest.approx(loss_cpu, rel=2e-4)
    blocks = [(mo.KW[l], mo.K_H * mo.KK[l]) for l in range(4)] + [(mo.KB[l], mo.K_H) for l in range(4)] + [
        (mo.KW5, mo.K_H), (mo.KB5, 1)]
    for off, size in blocks:
        a, b = g_gpu[off:off + size].astype(np.float64), g_cpu[off:off + size]
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        print(off, size, err)
        assert err < 5e-3, (off, err)
    # padded input columns of layer 1 and the padded output slots stay zero
    w1 = g_gpu[mo.KW[0]:mo.KW[0] + mo.K_H * mo.K_INPAD].reshape(mo.K_H, mo.K_INPAD)
    assert np.all(w1[:, mo.K_IN:] == 0.0)


def test_fused_inference_matches_training_forward(qk):
    """Predictions come from the fused one-launch inference kernel, the loss
    of sur_gradients from the per-layer training forward: the same network
    in two implementations (hidden layers bitwise equal, the output dot
    product summed in another order)."""
    p = mo.init_params(9)
    qk.sur_set_params(p)
    # partial, full and one-past-full 256-row workgroup tiles (k_mlp_infer), and a
    # batch spanning several tiles with a ragged last one
    for n in (1, 127, 255, 256, 257, 513, 1000):
        f = _box(n, 10 + n)
        t = np.random.default_rng(n).random(n) * 0.3
        y = qk.sur_predict(f) - mo.control(f)
        _, loss = qk.sur_gradients(f, t)
        assert float(np.mean((y - t) ** 2)) == pytest.approx(loss, rel=1e-5, abs=1e-12), n


def test_training_fits_and_is_deterministic(qk):
    f = _box(16384, 7)
    # a smooth price-like target: BSM-style call value per unit spot
    k = np.exp(f[:, 0])
    vol = np.sqrt(f[:, 7])
    sd = vol * np.sqrt(f[:, 1])
    from scipy.special import ndtr
    d1 = (-np.log(k) + (f[:, 2] + 0.5 * vol ** 2) * f[:, 1]) / sd
    t = ndtr(d1) - k * np.exp(-f[:, 2] * f[:, 1]) * ndtr(d1 - sd)
    qk.sur_init(11)
    l1 = qk.sur_train(f, t, epochs=40, batch=256, lr=1e-3, lr_final=1e-5, seed=3)
    print("loss", l1)
    assert l1[-1] < l1[0] / 100
    p1 = qk.sur_get_params()
    qk.sur_init(11)
    l2 = qk.sur_train(f, t, epochs=40, batch=256, lr=1e-3, lr_final=1e-5, seed=3)
    assert np.array_equal(l1, l2) and np.array_equal(p1, qk.sur_get_params())
    pred = qk.sur_predict(f[:2000])
    rmse = np.sqrt(np.mean((pred - t[:2000]) ** 2))
    print("rmse", rmse)
    assert rmse < 5e-3


def test_weights_file_roundtrip(qk, tmp_path):
    p = mo.init_params(9)
    qk.sur_set_params(p)
    path = tmp_path / "w.qkw"
    qk.sur_save(path)
    qk.sur_set_params(mo.init_params(10))
    qk.sur_load(path)
    assert np.array_equal(qk.sur_get_params(), p)
    raw = bytearray(path.read_bytes())
    raw[200] ^= 0xFF
    bad = tmp_path / "bad.qkw"
    bad.write_bytes(bytes(raw))
    with pytest.raises(Exception):
        qk.sur_load(bad)
    assert np.array_equal(qk.sur_get_params(), p)  # a rejected file changes nothing


def _queries():
    from quantkernel_amd import workloads as wl
    d = wl.config1(64, paths=4096, steps=256)
    d["option_type"] = (np.arange(64) % 2).astype(np.int32)
    d["strike"][5] = d["spot"][5] * 1.6        # outside the box
    d["q"][6] = 0.01                            # q != 0
    d["steps"][7] = 128                         # other step count
    d["spot"][8] = -1.0                         # invalid
    d["t"][9] = 0.0                             # intrinsic
    d["kappa"][10] = 7.0                        # outside the box
    return d


KEYS = ("spot", "strike", "t", "r", "q", "v0", "kappa", "theta", "sigma", "rho")


def test_pricing_sources_and_fallback(qk):
    qk.sur_set_params(mo.init_params(4))
    qk.sur_configure(spot_every=0, spot_paths=4096, spot_k_se=3.0)
    d = _queries()
    price, src = qk.sur_heston_price_batch(*[d[k] for k in KEYS], d["option_type"], d["paths"], d["steps"],
                                           d["seed"])
    mc = qk.heston_monte_carlo_price_batch(*[d[k] for k in KEYS], d["option_type"], d["paths"], d["steps"],
                                           d["seed"])
    assert list(src[5:11]) == [1, 1, 1, 3, 3, 1]
    assert math.isnan(price[8]) and price[9] == max(0.0, d["spot"][9] - d["strike"][9]) * (d["option_type"][9] == 0) \
        + max(0.0, d["strike"][9] - d["spot"][9]) * (d["option_type"][9] == 1)
    for i in (5, 6, 7, 10):
        assert price[i] == mc[i]  # the same MC pipeline, bitwise
    with np.errstate(invalid="ignore"):  # query 8 has a negative spot (invalid, src 3)
        logm = np.log(d["strike"] / d["spot"])
    feats = np.stack([logm, d["t"], d["r"], d["kappa"], d["theta"], d["sigma"], d["rho"], d["v0"]], axis=1)
    net = src == 0
    assert net.sum() >= 50
    y = qk.sur_predict(feats[net])
    s, dk = d["spot"][net], d["strike"][net] * np.exp(-d["r"][net] * d["t"][net])
    # the served call is clamped to its no-arbitrage bounds; puts by parity
    call = np.minimum(s, np.maximum(np.maximum(s - dk, 0.0), s * y))
    want = np.where(d["option_type"][net] == 0, call, call - s + dk)
    assert np.allclose(price[net], want, rtol=1e-6, atol=1e-9 * 150)


def test_spot_check_sends_disagreements_to_mc(qk):
    p = mo.init_params(4)
    p[mo.KB5] = 5.0  # a deliberately wrong network: every price is far off
    qk.sur_set_params(p)
    qk.sur_configure(spot_every=1, spot_paths=4096, spot_k_se=3.0)
    d = _queries()
    price, src = qk.sur_heston_price_batch(*[d[k] for k in KEYS], d["option_type"], d["paths"], d["steps"],
                                           d["seed"])
    mc = qk.heston_monte_carlo_price_batch(*[d[k] for k in KEYS], d["option_type"], d["paths"], d["steps"],
                                           d["seed"])
    qk.sur_configure()
    assert not np.any(src == 0)
    checked = src == 2
    assert checked.sum() >= 50
    assert np.array_equal(price[checked], mc[checked])
