"""GPU tests of importance sampling and the Monte Carlo deltas (pathwise,
likelihood ratio, Heston likelihood ratio) against the CPU oracle:

  * validation mode (the reference's mt19937_64 stream, FP64 paths): the
    reference goldens (cpp/tests/test_fdm_iqm_ram_agm_mlm.cpp:204-214) and
    the oracle within 1e-6 relative, odd path counts and edge cases included;
  * Philox mode: the FP64 Philox oracle on the same draws (FP32 path error),
    standard errors, closed forms within their standard errors, batch ==
    scalar and determinism.
"""
import math

import numpy as np
import pytest

from quantkernel_amd import workloads as wl

pytestmark = pytest.mark.gpu

RNG_F64 = 1
GBM = ("spot", "strike", "t", "vol", "r", "q", "option_type", "paths")
HST = ("spot", "strike", "t", "r", "q", "v0", "kappa", "theta", "sigma", "rho")


def rel(a, b):
    if math.isnan(a) or math.isnan(b):
        return 0.0 if (math.isnan(a) and math.isnan(b)) else math.inf
    return abs(a - b) / max(abs(b), 1e-300) if b != 0 else abs(a)


def _ncdf(x):
    return 0.5 * math.erfc(-x / math.sqrt(2.0))


def bsm_delta(s, k, t, vol, r, q, ot):
    d1 = (math.log(s / k) + (r - q + 0.5 * vol * vol) * t) / (vol * math.sqrt(t))
    qf = math.exp(-q * t)
    return qf * _ncdf(d1) if ot == 0 else qf * (_ncdf(d1) - 1.0)


def _cases(n, paths_lo=256, span=2000):
    d = wl.config0(n, paths=1)
    d["paths"] = (paths_lo + (np.arange(n) * 977) % span).astype(np.int32)  # odd and even
    d["seed"] = (np.arange(n) * 104729 + 11).astype(np.uint64)
    d.update(wl.heston_params(0x5EED0010, n))
    d["shift"] = ((np.arange(n) % 5) - 2) * 0.35
    d["clip"] = np.array([(2.0, 6.0, 0.3)[i % 3] for i in range(n)])
    d["steps"] = (1 + np.arange(n) % 9).astype(np.int32)
    return d


def _row(d, i, keys):
    return [float(d[k][i]) if d[k].dtype.kind == "f" else int(d[k][i]) for k in keys]


def test_goldens_validation_mode(qk):
    qk.set_rng_mode(RNG_F64)
    pw = qk.pathwise_derivative_delta(100.0, 100.0, 1.0, 0.2, 0.05, 0.0, 0, 50000, 12345)
    lr = qk.likelihood_ratio_delta(100.0, 100.0, 1.0, 0.2, 0.05, 0.0, 0, 50000, 12345, 10.0)
    print("goldens", repr(pw), repr(lr))
    assert abs(pw - 0.63559188066751005) < 1e-12
    assert abs(lr - 0.64537180232018476) < 1e-12


@pytest.mark.parametrize("method", ["importance", "pathwise", "lr", "heston_lr"])
def test_validation_fp64_matches_oracle(qk, orc, method):
    qk.set_rng_mode(RNG_F64)
    n = 24
    d = _cases(n)
    g = [d[k] for k in GBM]
    if method == "importance":
        got = qk.importance_sampling_price_batch(*g, d["shift"], d["seed"])
        want = [orc.lib.orc_importance_sampling(*_row(d, i, GBM), float(d["shift"][i]), int(d["seed"][i]))
                for i in range(n)]
    elif method == "pathwise":
        got = qk.pathwise_derivative_delta_batch(*g, d["seed"])
        want = [orc.lib.orc_pathwise_delta(*_row(d, i, GBM), int(d["seed"][i])) for i in range(n)]
    elif method == "lr":
        got = qk.likelihood_ratio_delta_batch(*g, d["seed"], d["clip"])
        want = [orc.lib.orc_lr_delta(*_row(d, i, GBM), int(d["seed"][i]), float(d["clip"][i])) for i in range(n)]
    else:
        d["paths"] = (d["paths"] // 8).astype(np.int32)
        got = qk.heston_lr_delta_batch(*[d[k] for k in HST], d["option_type"], d["paths"], d["steps"],
                                       d["seed"], d["clip"])
        want = [orc.lib.orc_heston_lr_delta(*_row(d, i, HST), int(d["option_type"][i]), int(d["paths"][i]),
                                            int(d["steps"][i]), int(d["seed"][i]), float(d["clip"][i]))
                for i in range(n)]
    worst = max(rel(a, b) for a, b in zip(got, want))
    print(method, "worst rel", worst)
    assert worst < 1e-6


EDGE = [
    # (spot, strike, t, vol, r, q, ot, paths, aux)
    (100, 100, 0.0, 0.2, 0.05, 0.01, 0, 1000, 0.5),
    (100, 100, 1e-13, 0.2, 0.05, 0.01, 1, 1000, 0.5),
    (100, 90, 1.0, 0.0, 0.05, 0.01, 0, 1000, 0.5),
    (100, 100 * math.exp(0.04), 1.0, 0.0, 0.05, 0.01, 1, 1000, 0.5),
    (100, 100, 1.0, 0.2, 0.05, 0.01, 0, 255, 0.5),
    (100, 100, 1.0, 0.2, 0.05, 0.01, 0, 1, 0.5),
    (100, 100, 1.0, 0.2, 0.05, 0.01, 0, 1000, 0.0),
    (100, 100, 1.0, 0.2, 0.05, 0.01, 0, 1000, -1.0),
    (100, 100, 1.0, 0.2, 0.05, 0.01, 0, 1000, math.inf),
    (100, 100, 1.0, 0.2, 0.05, 0.01, 0, 1000, math.nan),
    (math.nan, 100, 1.0, 0.2, 0.05, 0.01, 0, 1000, 0.5),
    (100, 100, 1.0, 0.2, math.inf, 0.01, 0, 1000, 0.5),
    (-5, 100, 1.0, 0.2, 0.05, 0.01, 0, 1000, 0.5),
    (100, 100, 1.0, 0.2, 0.05, 0.01, 2, 1000, 0.5),
    (100, 100, 1.0, 0.2, 0.05, 0.01, 1, 1001, 2.0),
]


@pytest.mark.parametrize("mode", [0, RNG_F64])
def test_edge_cases_match_oracle(qk, orc, mode):
    """Invalid elements give NaN, degenerate branches the closed values, in
    both RNG modes (closed branches do not depend on the stream)."""
    qk.set_rng_mode(mode)
    a = np.array(EDGE, dtype=np.float64)
    s, k, t, v, r, q = (a[:, j].copy() for j in range(6))
    ot, p, aux = a[:, 6].astype(np.int32), a[:, 7].astype(np.int32), a[:, 8].copy()
    seed = np.full(len(EDGE), 77, np.uint64)
    got = {"is": qk.importance_sampling_price_batch(s, k, t, v, r, q, ot, p, aux, seed),
           "pw": qk.pathwise_derivative_delta_batch(s, k, t, v, r, q, ot, p, seed),
           "lr": qk.likelihood_ratio_delta_batch(s, k, t, v, r, q, ot, p, seed, aux)}
    for i, c in enumerate(EDGE):
        sc, kc, tc, vc, rc, qc, oc, pc, xc = c
        want = {"is": orc.lib.orc_importance_sampling(sc, kc, tc, vc, rc, qc, oc, pc, xc, 77),
                "pw": orc.lib.orc_pathwise_delta(sc, kc, tc, vc, rc, qc, oc, pc, 77),
                "lr": orc.lib.orc_lr_delta(sc, kc, tc, vc, rc, qc, oc, pc, 77, xc)}
        for name in got:
            w, gv = want[name], got[name][i]
            closed = tc <= 1e-12 or (vc <= 1e-12 and name != "is")
            if math.isnan(w) or closed:
                # NaN, or a closed value (device exp may differ from glibc in the last ulp)
                assert rel(gv, w) < 1e-13, (name, c, gv, w)
            elif mode == RNG_F64:
                assert rel(gv, w) < 1e-6, (name, c, gv, w)
            else:
                assert math.isfinite(gv), (name, c, gv)
    # Heston LR: validation and the intrinsic-delta branch
    hp = [(0.04, 2.0, 0.04, 0.5, -0.7), (0.0, 1.0, 0.04, 0.3, 0.2), (0.04, 0.0, 0.0, 0.0, -1.0),
          (-0.01, 2.0, 0.04, 0.5, -0.7), (0.04, 2.0, 0.04, 0.5, 1.5)]
    # non-finite spot/strike/t/r/q are NaN by the product's policy (the
    # reference's result there depends on std::max's NaN ordering)
    rows = [(c, h) for c in EDGE for h in hp if all(math.isfinite(x) for x in (c[0], c[1], c[2], c[4], c[5]))]
    cols = [np.array([c[j] for c, _ in rows]) for j in (0, 1, 2, 4, 5)]
    hcols = [np.array([h[j] for _, h in rows]) for j in range(5)]
    ot = np.array([c[6] for c, _ in rows], np.int32)
    pp = np.array([max(c[7] // 10, 1) if c[7] > 1 else c[7] for c, _ in rows], np.int32)
    wc = np.array([c[8] for c, _ in rows])
    st = np.full(len(rows), 7, np.int32)
    sd = np.full(len(rows), 77, np.uint64)
    got = qk.heston_lr_delta_batch(*cols, *hcols, ot, pp, st, sd, wc)
    for i, (c, h) in enumerate(rows):
        w = orc.lib.orc_heston_lr_delta(c[0], c[1], c[2], c[4], c[5], *h, c[6], int(pp[i]), 7, 77, c[8])
        if math.isnan(w) or c[2] <= 1e-12:
            assert rel(got[i], w) == 0.0, (c, h, got[i], w)
        elif mode == RNG_F64:
            assert rel(got[i], w) < 1e-6, (c, h, got[i], w)
        else:
            assert math.isfinite(got[i]), (c, h, got[i])


def test_philox_matches_fp64_oracle(qk, orc):
    n = 12
    d = _cases(n, paths_lo=3001, span=4000)
    g = [d[k] for k in GBM]
    res = {"importance_sampling": qk.importance_sampling_price_batch(*g, d["shift"], d["seed"], return_stderr=True),
           "pathwise_delta": qk.pathwise_derivative_delta_batch(*g, d["seed"], return_stderr=True),
           "lr_delta": qk.likelihood_ratio_delta_batch(*g, d["seed"], d["clip"], return_stderr=True)}
    for i in range(n):
        row, seed = _row(d, i, GBM), int(d["seed"][i])
        want = {"importance_sampling": orc.px("importance_sampling", *row, float(d["shift"][i]), seed),
                "pathwise_delta": orc.px("pathwise_delta", *row, seed),
                "lr_delta": orc.px("lr_delta", *row, seed, float(d["clip"][i]))}
        for name, (got, se) in res.items():
            w, wse = want[name]
            scale = d["spot"][i] if name == "importance_sampling" else 1.0
            print(name, i, got[i], w, se[i], wse)
            assert abs(got[i] - w) / scale < 2e-5, (name, i, got[i], w)
            assert abs(se[i] - wse) <= 2e-3 * wse + 1e-9, (name, i, se[i], wse)
    # Heston LR delta: FP32 pairs driven by one Box-Muller evaluation
    h = wl.config1(8, paths=2001, steps=33)
    h["option_type"] = (np.arange(8) % 2).astype(np.int32)
    h["steps"] = np.array([33, 34, 35, 1, 2, 3, 4, 64], np.int32)  # every remainder mod 4
    clip = np.full(8, 6.0)
    got, se = qk.heston_lr_delta_batch(*[h[k] for k in HST], h["option_type"], h["paths"], h["steps"],
                                       h["seed"], clip, return_stderr=True)
    for i in range(8):
        w, wse = orc.px("heston_lr_delta", *_row(h, i, HST), int(h["option_type"][i]), 2001,
                        int(h["steps"][i]), int(h["seed"][i]), 6.0)
        print("heston_lr", i, got[i], w, se[i], wse)
        assert abs(got[i] - w) < 2e-4, (i, got[i], w)
        assert abs(se[i] - wse) <= 2e-3 * wse + 1e-9


def test_philox_vs_closed_forms(qk, orc):
    S, vol, r, q, T = 100.0, 0.25, 0.03, 0.01, 1.0
    K = np.array([80.0, 100.0, 120.0, 160.0] * 2)
    ot = np.array([0, 0, 0, 0, 1, 1, 1, 1], np.int32)
    n = len(K)
    one = np.ones(n)
    paths = np.full(n, 200001, np.int32)
    seed = np.arange(n, dtype=np.uint64) + 5
    args = (S * one, K, T * one, vol * one, r * one, q * one, ot, paths)
    shift = np.where(ot == 0, np.log(K / S) / (vol * math.sqrt(T)), -0.5)
    price, pse = qk.importance_sampling_price_batch(*args, shift, seed, return_stderr=True)
    pw, pwse = qk.pathwise_derivative_delta_batch(*args, seed, return_stderr=True)
    lr, lrse = qk.likelihood_ratio_delta_batch(*args, seed, np.full(n, 50.0), return_stderr=True)
    for i in range(n):
        bsm = orc.lib.orc_bsm(S, K[i], T, vol, r, q, int(ot[i]))
        dlt = bsm_delta(S, K[i], T, vol, r, q, int(ot[i]))
        print(i, price[i], bsm, pse[i], pw[i], lr[i], dlt, pwse[i], lrse[i])
        assert abs(price[i] - bsm) < 4 * pse[i] + 1e-6 * S
        assert abs(pw[i] - dlt) < 4 * pwse[i] + 1e-6
        assert abs(lr[i] - dlt) < 4 * lrse[i] + 1e-6
    # the shift toward a deep OTM strike cuts the error against plain MC
    _, se0 = qk.importance_sampling_price_batch(*args, np.zeros(n), seed, return_stderr=True)
    assert pse[3] < 0.3 * se0[3]
    # Heston LR delta vs a central difference of the characteristic-function
    # price.  The reference's score z1 / (sqrt(v0) sqrt(dt) S0) is the exact
    # first-step score only for rho = 0 (with rho != 0 the first variance step
    # also depends on z1, which adds -rho z2 / sqrt(1 - rho^2) to the
    # numerator), so the tight check uses rho = 0 and rho = -0.7 is held to
    # the reference's own tolerance (python/tests/test_heston_lr_delta.py:54).
    m = 4
    hot = np.array([0, 1, 0, 1], np.int32)
    for rho, tol_se, tol_abs in ((0.0, 4.0, 0.005), (-0.7, 0.0, 0.15)):
        hp = (0.04, 2.0, 0.04, 0.3, rho)
        hs = [np.full(m, x) for x in (100.0, 100.0, 1.0, 0.05, 0.02, *hp)]
        d, dse = qk.heston_lr_delta_batch(*hs, hot, np.full(m, 400001, np.int32), np.full(m, 128, np.int32),
                                          np.arange(m, dtype=np.uint64) + 9, np.full(m, 1e9), return_stderr=True)
        for i in range(m):
            fd = (orc.lib.orc_heston_cf(100.5, 100.0, 1.0, 0.05, 0.02, *hp, int(hot[i]), 2048, 200.0)
                  - orc.lib.orc_heston_cf(99.5, 100.0, 1.0, 0.05, 0.02, *hp, int(hot[i]), 2048, 200.0))
            print("heston lr", rho, i, d[i], fd, dse[i])
            assert abs(d[i] - fd) < tol_se * dse[i] + tol_abs


def test_batch_equals_scalar_and_deterministic(qk):
    d = _cases(6, paths_lo=5001)
    g = [d[k] for k in GBM]
    b1 = qk.likelihood_ratio_delta_batch(*g, d["seed"], d["clip"])
    b2 = qk.likelihood_ratio_delta_batch(*g, d["seed"], d["clip"])
    assert np.array_equal(b1, b2)
    for i in range(6):
        assert qk.likelihood_ratio_delta(*_row(d, i, GBM), int(d["seed"][i]), float(d["clip"][i])) == b1[i]
        assert qk.pathwise_derivative_delta(*_row(d, i, GBM), int(d["seed"][i])) == \
            qk.pathwise_derivative_delta_batch(*[x[i:i + 1] for x in g], d["seed"][i:i + 1])[0]
        assert qk.importance_sampling_price(*_row(d, i, GBM), float(d["shift"][i]), int(d["seed"][i])) == \
            qk.importance_sampling_price_batch(*[x[i:i + 1] for x in g], d["shift"][i:i + 1],
                                               d["seed"][i:i + 1])[0]
    h = wl.config1(3, paths=3001, steps=20)
    hb = qk.heston_lr_delta_batch(*[h[k] for k in HST], h["option_type"], h["paths"], h["steps"], h["seed"],
                                  np.full(3, 6.0))
    for i in range(3):
        assert qk.heston_lr_delta(*_row(h, i, HST), int(h["option_type"][i]), 3001, 20, int(h["seed"][i]),
                                  6.0) == hb[i]
    # the reference wrapper raises on NaN scalars
    with pytest.raises(ValueError):
        qk.pathwise_derivative_delta(100.0, 100.0, 1.0, 0.2, 0.05, 0.0, 0, 100, 1)
