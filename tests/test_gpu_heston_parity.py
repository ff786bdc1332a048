"""Heston (and constant local-vol) parity beyond the random C1 box.

- Edge parameters the reference admits (heston_monte_carlo.cpp:18-24:
  v0, kappa, theta, sigma >= 0, rho in [-1, 1]; no finiteness check), in
  both RNG modes: validation mode (mt19937 stream, FP64 paths) against the
  reference build within 1e-6 relative, Philox mode against the FP64 oracle
  on the same Philox stream.
- Closed forms the scheme satisfies exactly: with sigma = 0 the variance is
  deterministic, v_{k+1} = max(0, v_k + kappa (theta - v_k) dt), and log S_T
  of the log-Euler scheme is Gaussian with variance sum_k v_k dt, so the
  price is Black-Scholes-Merton at that total variance (no time-step bias);
  v0 = theta = 0 makes S_T = S e^{(r-q)T} on every path.
- Independent estimates: Philox mode against the reference's own mt19937
  stream (validation mode) over 1024 options, the fraction within 3 SE held
  to its binomial bound.
- The reference's invariant suites (python/tests/test_heston_mc_invariants.py:
  33-170, test_local_vol_monte_carlo.py:62-125) on the production Philox
  path, with common seeds.
"""
import math

import numpy as np
import pytest

from quantkernel_amd import workloads as wl

pytestmark = pytest.mark.gpu

RNG_PHILOX, RNG_F64 = 0, 1
KEYS = ("spot", "strike", "t", "r", "q", "v0", "kappa", "theta", "sigma", "rho")
QK_CALL, QK_PUT = 0, 1
# within-3-SE fraction of a standard normal and its one-sided 3-sigma
# binomial lower bound for n options
P3 = math.erf(3 / math.sqrt(2))


def frac_bound(n):
    return P3 - 3.0 * math.sqrt(P3 * (1 - P3) / n)


# spot, strike, t, r, q, v0, kappa, theta, sigma, rho, option_type
EDGE = [
    (100.0, 100.0, 1.0, 0.05, 0.02, 0.04, 2.0, 0.04, 0.3, -1.0, QK_CALL),   # rho = -1
    (100.0, 100.0, 1.0, 0.05, 0.02, 0.04, 2.0, 0.04, 0.3, -0.99, QK_PUT),
    (100.0, 100.0, 1.0, 0.05, 0.02, 0.04, 2.0, 0.04, 0.3, 0.0, QK_CALL),
    (100.0, 100.0, 1.0, 0.05, 0.02, 0.04, 2.0, 0.04, 0.3, 0.99, QK_PUT),
    (100.0, 100.0, 1.0, 0.05, 0.02, 0.04, 2.0, 0.04, 0.3, 1.0, QK_CALL),    # rho = +1
    (100.0, 90.0, 1.0, 0.05, 0.02, 0.04, 2.0, 0.09, 0.0, -0.7, QK_CALL),    # sigma = 0
    (100.0, 110.0, 0.5, 0.03, 0.0, 0.06, 0.0, 0.04, 0.5, -0.5, QK_PUT),     # kappa = 0
    (100.0, 95.0, 1.0, 0.05, 0.01, 0.0, 2.0, 0.0, 0.4, -0.7, QK_CALL),      # v0 = theta = 0
    (300.0, 100.0, 1.0, 0.05, 0.02, 0.04, 2.0, 0.04, 0.3, -0.7, QK_CALL),   # spot 300, deep ITM
    (300.0, 100.0, 1.0, 0.05, 0.02, 0.04, 2.0, 0.04, 0.3, -0.7, QK_PUT),    # spot 300, deep OTM
    (110.0, 100.0, 1e-14, 0.05, 0.02, 0.04, 2.0, 0.04, 0.3, -0.7, QK_CALL),  # T = 1e-14: intrinsic
    (100.0, 100.0, 2.0, 0.05, 0.0, 0.16, 0.5, 0.16, 1.0, -0.9, QK_PUT),      # Feller violated hard
]


def _edge_batch(paths, steps, seed0=900):
    n = len(EDGE)
    cols = np.array([e[:10] for e in EDGE]).T
    d = {k: np.ascontiguousarray(cols[j]) for j, k in enumerate(KEYS)}
    d["option_type"] = np.array([e[10] for e in EDGE], np.int32)
    d["paths"] = np.full(n, paths, np.int32)
    d["steps"] = np.full(n, steps, np.int32)
    d["seed"] = (np.arange(n) + seed0).astype(np.uint64)
    return d


def _price(qk, d, se=True):
    return qk.heston_monte_carlo_price_batch(*[d[k] for k in KEYS], d["option_type"], d["paths"], d["steps"],
                                             d["seed"], return_stderr=se)


def test_edge_parameters_validation_mode_match_reference(qk, ref):
    d = _edge_batch(3000, 24)
    qk.set_rng_mode(RNG_F64)
    got = _price(qk, d, se=False)
    for i, e in enumerate(EDGE):
        want = ref.mcm_heston_monte_carlo_price(*e[:10], e[10], 3000, 24, int(d["seed"][i]))
        print(i, got[i], want)
        assert math.isfinite(got[i])
        assert abs(got[i] - want) <= 1e-6 * abs(want) + 1e-12, (i, got[i], want)
    assert got[10] == 10.0  # T <= 1e-12: undiscounted intrinsic


def test_edge_parameters_philox_match_fp64_oracle(qk, orc):
    d = _edge_batch(4000, 33)
    got, se = _price(qk, d)
    for i, e in enumerate(EDGE):
        want, wse = orc.px("heston", *e[:10], e[10], 4000, 33, int(d["seed"][i]))
        print(i, got[i], want, se[i], wse)
        assert math.isfinite(got[i]) and got[i] >= 0.0
        # FP32 paths against FP64 on the same draws, per unit spot
        assert abs(got[i] - want) / e[0] < 2e-5, (i, got[i], want)
        # the SE too; v0 = theta = 0 has none (all paths equal), where the
        # GPU's centred sums give 0 and the oracle's raw sums a rounding residue
        assert abs(se[i] - wse) <= 1e-3 * wse + 1e-7 * e[0]
    assert got[10] == 10.0


def _sigma0_total_variance(v0, kappa, theta, t, steps):
    dt = t / steps
    v, tot = v0, 0.0
    for _ in range(steps):
        tot += v * dt
        v = max(0.0, v + kappa * (theta - v) * dt)
    return tot


@pytest.mark.parametrize("mode", [RNG_PHILOX, RNG_F64])
def test_sigma0_heston_is_black_scholes_at_the_scheme_variance(qk, orc, mode):
    """sigma = 0: the scheme's price is BSM at total variance sum_k v_k dt,
    exactly (no discretisation bias), so z = (MC - BSM) / SE is N(0, 1)."""
    n = 1024 if mode == RNG_PHILOX else 128
    paths = 65536 if mode == RNG_PHILOX else 8192
    d = wl.config1(n, paths=paths, steps=48, start=5000)
    d["sigma"][:] = 0.0
    d["q"] = np.where(np.arange(n) % 3 == 0, 0.02, 0.0)
    d["option_type"] = (np.arange(n) % 2).astype(np.int32)
    qk.set_rng_mode(mode)
    got, se = _price(qk, d)
    bsm = np.empty(n)
    for i in range(n):
        var = _sigma0_total_variance(d["v0"][i], d["kappa"][i], d["theta"][i], d["t"][i], 48)
        bsm[i] = orc.lib.orc_bsm(float(d["spot"][i]), float(d["strike"][i]), float(d["t"][i]),
                                 math.sqrt(var / d["t"][i]), float(d["r"][i]), float(d["q"][i]),
                                 int(d["option_type"][i]))
    ok = (got / np.where(se > 0, se, 1.0)) ** 2 >= 30  # SE meaningful: >= ~30 paths in the money
    z = (got[ok] - bsm[ok]) / se[ok]
    within = float(np.mean(np.abs(z) < 3.0))
    print(f"mode {mode}: {ok.sum()} options, mean z {z.mean():.3f}, rms z {np.sqrt(np.mean(z * z)):.3f}, "
          f"within 3 SE {within:.4f} (bound {frac_bound(ok.sum()):.4f})")
    m = ok.sum()
    assert abs(z.mean()) < 4.0 / math.sqrt(m)
    assert abs(math.sqrt(np.mean(z * z)) - 1.0) < 4.0 / math.sqrt(2 * m)  # rms of m N(0,1): sd 1/sqrt(2m)
    assert within >= frac_bound(m)


def test_zero_variance_heston_is_deterministic_forward(qk):
    """v0 = theta = 0: every path ends at S e^{(r-q)T} (to FP32 rounding)."""
    n = 64
    d = wl.config1(n, paths=4096, steps=64, start=7000)
    d["v0"][:] = 0.0
    d["theta"][:] = 0.0
    d["option_type"] = (np.arange(n) % 2).astype(np.int32)
    got, se = _price(qk, d)
    fwd = d["spot"] * np.exp((d["r"] - d["q"]) * d["t"])
    disc = np.exp(-d["r"] * d["t"])
    want = disc * np.where(d["option_type"] == 0, np.maximum(fwd - d["strike"], 0), np.maximum(d["strike"] - fwd, 0))
    assert np.all(np.abs(got - want) < 2e-5 * d["spot"])
    assert np.all(se < 1e-5 * d["spot"])


def test_philox_vs_reference_stream_fraction_within_3se(qk):
    """1024 Heston options from the C1 box: the Philox estimate and the
    estimate on the reference's own mt19937 stream (validation mode, which
    equals the reference to 1e-6) are independent; their difference over
    the combined SE is N(0, 1)."""
    n = 1024
    d = wl.config1(n, paths=4096, steps=32, start=9000)
    d["option_type"] = (np.arange(n) % 2).astype(np.int32)
    p1, s1 = _price(qk, d)
    qk.set_rng_mode(RNG_F64)
    p2, s2 = _price(qk, d)
    sd = np.sqrt(s1 * s1 + s2 * s2)
    ok = (np.minimum(p1, p2) / np.where(sd > 0, sd, 1.0)) ** 2 >= 30
    z = (p1[ok] - p2[ok]) / sd[ok]
    within = float(np.mean(np.abs(z) < 3.0))
    print(f"{ok.sum()} options: mean z {z.mean():.3f}, rms {np.sqrt(np.mean(z * z)):.3f}, within 3 SE {within:.4f} "
          f"(bound {frac_bound(ok.sum()):.4f})")
    assert ok.sum() > 900
    assert abs(z.mean()) < 4.0 / math.sqrt(ok.sum())
    assert within >= frac_bound(ok.sum())


# ---- the reference's invariant suites, Philox mode --------------------------
_BASE = dict(strike=100.0, t=1.0, r=0.05, q=0.02, v0=0.04, kappa=2.0, theta=0.04, sigma=0.3, rho=-0.7,
             paths=20000, steps=80, seed=314)


def _h(qk, **kw):
    a = dict(_BASE, **kw)
    return qk.heston_monte_carlo_price(a["spot"], a["strike"], a["t"], a["r"], a["q"], a["v0"], a["kappa"],
                                       a["theta"], a["sigma"], a["rho"], a["option_type"], a["paths"],
                                       a["steps"], a["seed"])


def test_heston_invariants(qk):
    put_ub = _BASE["strike"] * math.exp(-_BASE["r"] * _BASE["t"])
    for spot in (80.0, 100.0, 120.0):  # bounds (test_heston_mc_invariants.py:33-58)
        c, p = _h(qk, spot=spot, option_type=QK_CALL), _h(qk, spot=spot, option_type=QK_PUT)
        assert 0.0 <= c <= spot and 0.0 <= p <= put_ub
    # monotone in spot with one seed (:63-94): S_T is spot times a
    # spot-free path factor in the log-Euler scheme, so the order holds
    # exactly, without the reference's 0.5 margin
    spots = [80.0, 90.0, 100.0, 110.0, 120.0]
    calls = [_h(qk, spot=s, option_type=QK_CALL) for s in spots]
    puts = [_h(qk, spot=s, option_type=QK_PUT) for s in spots]
    assert all(b > a for a, b in zip(calls, calls[1:]))
    assert all(b < a for a, b in zip(puts, puts[1:]))
    # non-decreasing in vol of vol at the money (:101-126), the reference's margin
    for ot in (QK_CALL, QK_PUT):
        assert _h(qk, spot=100.0, option_type=ot, sigma=0.5) >= _h(qk, spot=100.0, option_type=ot, sigma=0.2) - 0.5
    # finite output (:133-170)
    for spot in (50.0, 100.0, 200.0):
        for ot in (QK_CALL, QK_PUT):
            assert math.isfinite(_h(qk, spot=spot, option_type=ot))
    deep_itm = _h(qk, spot=300.0, option_type=QK_CALL)
    assert math.isfinite(deep_itm) and deep_itm > 0.0
    deep_otm = _h(qk, spot=300.0, option_type=QK_PUT)
    assert math.isfinite(deep_otm) and deep_otm >= 0.0
    assert _h(qk, spot=110.0, option_type=QK_CALL, t=1e-14) == 10.0
    for rho in (-0.99, 0.0, 0.99):
        v = _h(qk, spot=100.0, option_type=QK_CALL, rho=rho)
        assert math.isfinite(v) and v > 0.0


_LV = dict(spot=100.0, strike=100.0, t=1.0, vol=0.2, r=0.03, q=0.01, option_type=QK_CALL, paths=20000, steps=64,
           seed=42)


def _lv(qk, fn="local_vol_monte_carlo_price", **kw):
    a = dict(_LV, **kw)
    return getattr(qk, fn)(a["spot"], a["strike"], a["t"], a["vol"], a["r"], a["q"], a["option_type"], a["paths"],
                           a["steps"], a["seed"])


def test_local_vol_invariants(qk, orc):
    # constant vol equals Euler (test_local_vol_monte_carlo.py:30-48): the
    # same arithmetic on the same draws, apart from one final rounding
    for ot in (QK_CALL, QK_PUT):
        lv, em = _lv(qk, option_type=ot), _lv(qk, fn="euler_maruyama_price", option_type=ot)
        assert abs(lv - em) < 1e-9 * abs(em) + 1e-12
    # close to the closed form (:50-60)
    bsm = orc.lib.orc_bsm(100.0, 100.0, 1.0, 0.2, 0.03, 0.01, QK_CALL)
    assert abs(_lv(qk, paths=50000) - bsm) < 1.0
    # bounds (:66-79)
    for spot in (80.0, 100.0, 120.0):
        assert 0.0 <= _lv(qk, spot=spot) <= spot
        assert 0.0 <= _lv(qk, spot=spot, option_type=QK_PUT) <= 100.0 * math.exp(-0.03)
    # monotone in spot (:82-103), exactly with one seed: s (c0 + c1 z) is
    # linear in the spot path by path
    spots = [80.0, 90.0, 100.0, 110.0, 120.0]
    c = [_lv(qk, spot=s) for s in spots]
    p = [_lv(qk, spot=s, option_type=QK_PUT) for s in spots]
    assert all(b > a for a, b in zip(c, c[1:])) and all(b < a for a, b in zip(p, p[1:]))
    # determinism (:106-113) and batch == scalar (:116-129), bitwise
    assert _lv(qk, seed=777) == _lv(qk, seed=777)
    rows = [_LV, dict(_LV, spot=110.0, option_type=QK_PUT, seed=99)]
    batch = qk.local_vol_monte_carlo_price_batch(*[np.array([r[k] for r in rows]) for k in (
        "spot", "strike", "t", "vol", "r", "q")], np.array([r["option_type"] for r in rows], np.int32),
        np.array([r["paths"] for r in rows], np.int32), np.array([r["steps"] for r in rows], np.int32),
        np.array([r["seed"] for r in rows], np.uint64))
    for i, r in enumerate(rows):
        assert batch[i] == _lv(qk, **{k: r[k] for k in r})
