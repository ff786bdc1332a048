"""GPU tests of control variates, stratified sampling, Sobol / Halton
quasi-MC and multilevel MC against the CPU oracle: validation mode (the
reference's mt19937_64 stream, FP64 paths) within 1e-6 relative, Philox mode
against the FP64 Philox oracle (FP32 path error) with standard errors, edge
cases, closed forms and batch == scalar."""
import math

import numpy as np
import pytest

from quantkernel_amd import workloads as wl

pytestmark = pytest.mark.gpu

RNG_F64 = 1
GBM = ("spot", "strike", "t", "vol", "r", "q", "option_type", "paths")


def rel(a, b):
    if math.isnan(a) or math.isnan(b):
        return 0.0 if (math.isnan(a) and math.isnan(b)) else math.inf
    return abs(a - b) / max(abs(b), 1e-300) if b != 0 else abs(a)


def _cases(n, lo=3, span=3000):
    d = wl.config0(n, paths=1)
    d["paths"] = (lo + (np.arange(n) * 977) % span).astype(np.int32)
    d["seed"] = (np.arange(n) * 104729 + 11).astype(np.uint64)
    d["levels"] = (1 + np.arange(n) % 4).astype(np.int32)
    d["base_steps"] = (1 + np.arange(n) % 5).astype(np.int32)
    return d


def _row(d, i, keys):
    return [float(d[k][i]) if d[k].dtype.kind == "f" else int(d[k][i]) for k in keys]


def _gpu(qk, method, d, se=False):
    g = [d[k] for k in GBM]
    if method == "cv":
        return qk.control_variates_price_batch(*g, d["seed"], return_stderr=se)
    if method == "strat":
        return qk.stratified_sampling_price_batch(*g, d["seed"], return_stderr=se)
    if method == "sobol":
        return qk.quasi_monte_carlo_sobol_price_batch(*g, return_stderr=se)
    if method == "halton":
        return qk.quasi_monte_carlo_halton_price_batch(*g, return_stderr=se)
    return qk.multilevel_monte_carlo_price_batch(*g[:-1], d["paths"] // 8 + 2, d["levels"], d["base_steps"],
                                                 d["seed"], return_stderr=se)


def _orc(orc, method, d, i, px=False):
    row, seed = _row(d, i, GBM), int(d["seed"][i])
    if method in ("sobol", "halton"):
        return orc.lib.orc_qmc(0 if method == "sobol" else 1, *row)
    if method == "mlmc":
        args = row[:-1] + [int(d["paths"][i]) // 8 + 2, int(d["levels"][i]), int(d["base_steps"][i]), seed]
        return orc.px("mlmc", *args) if px else orc.lib.orc_mlmc(*args)
    name = {"cv": "control_variates", "strat": "stratified"}[method]
    return orc.px(name, *row, seed) if px else getattr(orc.lib, "orc_" + name)(*row, seed)


@pytest.mark.parametrize("method", ["cv", "strat", "sobol", "halton", "mlmc"])
def test_validation_fp64_matches_oracle(qk, orc, method):
    qk.set_rng_mode(RNG_F64)
    n = 24
    d = _cases(n)
    got = _gpu(qk, method, d)
    want = [_orc(orc, method, d, i) for i in range(n)]
    worst = max(rel(a, b) for a, b in zip(got, want))
    print(method, "worst rel", worst)
    assert worst < 1e-6


@pytest.mark.parametrize("method", ["cv", "strat", "sobol", "halton", "mlmc"])
def test_philox_matches_fp64_oracle(qk, orc, method):
    n = 12
    d = _cases(n, lo=2001, span=5000)
    got, se = _gpu(qk, method, d, se=True)
    for i in range(n):
        if method in ("sobol", "halton"):  # deterministic: the FP64 result itself
            w, wse = _orc(orc, method, d, i), math.nan
        else:
            w, wse = _orc(orc, method, d, i, px=True)
        print(method, i, got[i], w, se[i], wse)
        assert abs(got[i] - w) / d["spot"][i] < 2e-5, (i, got[i], w)
        if math.isnan(wse):
            assert math.isnan(se[i])
        else:
            assert abs(se[i] - wse) <= 2e-3 * wse + 1e-9, (i, se[i], wse)


EDGE = [
    (100, 100, 0.0, 0.2, 0.05, 0.01, 0, 1000), (100, 100, 1e-13, 0.2, 0.05, 0.01, 1, 1000),
    (100, 90, 1.0, 0.0, 0.05, 0.01, 0, 1000), (100, 100, 1.0, 0.2, 0.05, 0.01, 0, 2),
    (100, 100, 1.0, 0.2, 0.05, 0.01, 0, 3), (100, 100, 1.0, 0.2, 0.05, 0.01, 1, 1),
    (math.nan, 100, 1.0, 0.2, 0.05, 0.01, 0, 1000), (100, 100, 1.0, 0.2, math.inf, 0.01, 0, 1000),
    (-5, 100, 1.0, 0.2, 0.05, 0.01, 0, 1000), (100, 100, 1.0, 0.2, 0.05, 0.01, 2, 1000),
]
MLMC_EDGE = [(1000, 3, 4), (1, 2, 4), (1000, 0, 4), (1000, 2, 0), (1000, 1, 1), (5, 4, 2), (1000, 40, 4)]


@pytest.mark.parametrize("mode", [0, RNG_F64])
def test_edge_cases_match_oracle(qk, orc, mode):
    qk.set_rng_mode(mode)
    a = np.array(EDGE, dtype=np.float64)
    cols = [a[:, j].copy() for j in range(6)]
    ot, p = a[:, 6].astype(np.int32), a[:, 7].astype(np.int32)
    seed = np.full(len(EDGE), 5, np.uint64)
    got = {"cv": qk.control_variates_price_batch(*cols, ot, p, seed),
           "strat": qk.stratified_sampling_price_batch(*cols, ot, p, seed),
           "sobol": qk.quasi_monte_carlo_sobol_price_batch(*cols, ot, p),
           "halton": qk.quasi_monte_carlo_halton_price_batch(*cols, ot, p)}
    for i, c in enumerate(EDGE):
        want = {"cv": orc.lib.orc_control_variates(*c, 5), "strat": orc.lib.orc_stratified(*c, 5),
                "sobol": orc.lib.orc_qmc(0, *c), "halton": orc.lib.orc_qmc(1, *c)}
        for name, w in want.items():
            gv = got[name][i]
            if math.isnan(w) or c[2] <= 1e-12:
                assert rel(gv, w) < 1e-13, (name, c, gv, w)
            elif mode == RNG_F64 or name in ("sobol", "halton"):
                assert rel(gv, w) < (1e-6 if mode == RNG_F64 else 1e-5), (name, c, gv, w)
            else:
                assert math.isfinite(gv), (name, c, gv)
    # multilevel: level / path / step validation (40 levels overflow int32 steps: NaN)
    rows = [(c, m) for c in EDGE for m in MLMC_EDGE]
    cc = [np.array([c[j] for c, _ in rows]) for j in range(6)]
    mo = [np.array([c[6] for c, _ in rows], np.int32)]
    mm = [np.array([m[j] for _, m in rows], np.int32) for j in range(3)]
    got = qk.multilevel_monte_carlo_price_batch(*cc, *mo, *mm, np.full(len(rows), 5, np.uint64))
    for i, (c, m) in enumerate(rows):
        if m[1] > 31:
            assert math.isnan(got[i])
            continue
        w = orc.lib.orc_mlmc(*c[:7], *m, 5)
        if math.isnan(w) or c[2] <= 1e-12:
            assert rel(got[i], w) < 1e-13, (c, m, got[i], w)
        elif mode == RNG_F64:
            assert rel(got[i], w) < 1e-6, (c, m, got[i], w)
        else:
            assert math.isfinite(got[i]), (c, m, got[i])


def test_philox_vs_closed_forms(qk, orc):
    S, vol, r, q, T = 100.0, 0.25, 0.03, 0.01, 1.0
    K = np.array([80.0, 100.0, 120.0, 160.0] * 2)
    ot = np.array([0, 0, 0, 0, 1, 1, 1, 1], np.int32)
    n = len(K)
    one = np.ones(n)
    args = (S * one, K, T * one, vol * one, r * one, q * one, ot)
    paths = np.full(n, 200001, np.int32)
    seed = np.arange(n, dtype=np.uint64) + 3
    cv, cv_se = qk.control_variates_price_batch(*args, paths, seed, return_stderr=True)
    st, st_se = qk.stratified_sampling_price_batch(*args, paths, seed, return_stderr=True)
    so = qk.quasi_monte_carlo_sobol_price_batch(*args, np.full(n, 1 << 20, np.int32))
    ha = qk.quasi_monte_carlo_halton_price_batch(*args, np.full(n, 1 << 20, np.int32))
    ml, ml_se = qk.multilevel_monte_carlo_price_batch(*args, np.full(n, 100000, np.int32), np.full(n, 4, np.int32),
                                                      np.full(n, 8, np.int32), seed, return_stderr=True)
    # plain MC with one draw per path: importance sampling with a zero shift
    _, plain_se = qk.importance_sampling_price_batch(*args, paths, np.zeros(n), seed, return_stderr=True)
    for i in range(n):
        bsm = orc.lib.orc_bsm(S, K[i], T, vol, r, q, int(ot[i]))
        print(i, bsm, cv[i], cv_se[i], st[i], st_se[i], so[i], ha[i], ml[i], ml_se[i])
        assert abs(cv[i] - bsm) < 4 * cv_se[i] + 1e-6 * S
        assert abs(st[i] - bsm) < 4 * st_se[i] + 1e-6 * S
        assert abs(so[i] - bsm) < 1e-3 and abs(ha[i] - bsm) < 1e-3
        assert abs(ml[i] - bsm) < 4 * ml_se[i] + 0.02
    # an in-the-money call is nearly linear in S_T: the control removes most
    # of its variance, and a good part of an at-the-money call's
    assert cv_se[0] < 0.25 * plain_se[0] and cv_se[1] < 0.6 * plain_se[1]


def test_batch_equals_scalar_and_deterministic(qk):
    d = _cases(5, lo=5001)
    g = [d[k] for k in GBM]
    b = qk.control_variates_price_batch(*g, d["seed"])
    assert np.array_equal(b, qk.control_variates_price_batch(*g, d["seed"]))
    mb = qk.multilevel_monte_carlo_price_batch(*g[:-1], d["paths"] // 8 + 2, d["levels"], d["base_steps"], d["seed"])
    for i in range(5):
        row = _row(d, i, GBM)
        assert qk.control_variates_price(*row, int(d["seed"][i])) == b[i]
        assert qk.stratified_sampling_price(*row, int(d["seed"][i])) == \
            qk.stratified_sampling_price_batch(*[x[i:i + 1] for x in g], d["seed"][i:i + 1])[0]
        assert qk.quasi_monte_carlo_sobol_price(*row) == qk.quasi_monte_carlo_sobol_price_batch(
            *[x[i:i + 1] for x in g])[0]
        assert qk.multilevel_monte_carlo_price(*row[:-1], int(d["paths"][i]) // 8 + 2, int(d["levels"][i]),
                                               int(d["base_steps"][i]), int(d["seed"][i])) == mb[i]
