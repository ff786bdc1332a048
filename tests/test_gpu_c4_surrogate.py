"""The trained Heston surrogate (the committed weights) as it is served:
network inside the training box, plain MC outside it and where a spot check
disagrees.

- Accuracy of the whole query mix (in-box queries and the 5% drawn from the
  box widened by 10% per side, so network and MC-fallback answers alike)
  against MC at 2^22 paths, per unit spot, with the reference's own standard
  error reported; the same queries priced by plain MC at 65,536 paths, the
  path count the fallback uses, for the comparison the surrogate is there
  to win: a smaller error for far fewer simulated path-steps.
- The full 1,048,576-query mix: every answer finite, the answer sources
  where the box test puts them, and every MC answer bitwise equal to the
  Heston MC batch entry point's for the same query.

The reference has no surrogate; the estimator it replaces and falls back to
is its Heston MC (heston_monte_carlo.cpp:11-36), here the library's own
Philox kernel of the same scheme.
"""
import math
from pathlib import Path

import numpy as np
import pytest

from quantkernel_amd import surrogate as sur

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
WEIGHTS = ROOT / "quant-kernel_amd" / "quantkernel_amd" / "heston_surrogate.qkw"
KEYS = ("spot", "strike", "t", "r", "q", "v0", "kappa", "theta", "sigma", "rho")
MASTER_C4_TEST = 0x5EED9000   # fresh queries: neither training, evaluation nor bench seeds
MC_PATHS = 65536
SPOT_EVERY, SPOT_PATHS = 128, 8192
REF_PATHS = 1 << 22
TARGET = 1e-3                 # RMS |price - reference| / spot
PRICE_FLOOR = 0.01            # price-relative errors are reported for prices >= 1% of spot


def _mix(n, master, frac_out=0.05):
    """n call/put queries: (1 - frac_out) from the training box, the rest
    from the box widened by 10% per side, interleaved."""
    n_out = int(round(n * frac_out))
    d = sur.sample_queries(n - n_out, master)
    e = sur.sample_queries(n_out, master + 1, expand=0.1)
    order = np.argsort(np.concatenate([np.arange(n - n_out) * n / (n - n_out),
                                       np.arange(n_out) * n / max(n_out, 1) + 0.5]), kind="stable")
    d = {k: np.concatenate([d[k], e[k]])[order] for k in d}
    d["option_type"] = (np.arange(n) % 2).astype(np.int32)
    d["paths"] = np.full(n, MC_PATHS, np.int32)
    return d


def _in_box(d):
    f = sur.features(d)
    inside = np.all((f >= sur.BOX_LO) & (f <= sur.BOX_HI), axis=1)
    return inside & (d["q"] == 0.0) & (d["steps"] == sur.STEPS)


@pytest.fixture(scope="module")
def served(qk):
    assert WEIGHTS.exists(), WEIGHTS
    qk.sur_load(WEIGHTS)
    qk.sur_configure(spot_every=SPOT_EVERY, spot_paths=SPOT_PATHS, spot_k_se=3.0)
    yield qk
    qk.sur_configure()


def test_whole_mix_accuracy_and_gain_over_plain_mc(served):
    qk = served
    n = 2048
    d = _mix(n, MASTER_C4_TEST)
    print(f"C4 accuracy: {n} queries, {int((~_in_box(d)).sum())} outside the box", flush=True)
    price, src = qk.sur_heston_price_batch(*[d[k] for k in KEYS], d["option_type"], d["paths"], d["steps"],
                                           d["seed"])
    # reference: the same scheme at 2^22 paths, independent seeds
    ref = dict(d, seed=d["seed"] ^ np.uint64(0xA5A5A5A55A5A5A5A))
    ref_p, ref_se = sur.mc_price(qk, ref, REF_PATHS, chunk=1024)
    # plain MC at the fallback's path count, on the same queries and seeds
    mc, mc_se = qk.heston_monte_carlo_price_batch(*[d[k] for k in KEYS], d["option_type"], d["paths"],
                                                  d["steps"], d["seed"], return_stderr=True)
    s = d["spot"]
    e_sur = (price - ref_p) / s
    e_mc = (mc - ref_p) / s
    ref_ms = float(np.mean((ref_se / s) ** 2))
    rms_sur = math.sqrt(np.mean(e_sur ** 2))
    rms_mc = math.sqrt(np.mean(e_mc ** 2))
    net = src == 0
    big = ref_p >= PRICE_FLOOR * s
    rel = np.abs(price[big] - ref_p[big]) / ref_p[big]
    rel_mc = np.abs(mc[big] - ref_p[big]) / ref_p[big]
    # simulated path-steps: surrogate+fallback against plain MC of every query
    fb = (src == 1) | (src == 2) | (src == 4)
    spot_checked = np.zeros(n, bool)
    spot_checked[::SPOT_EVERY] = True
    spot_checked &= (src == 0) | (src == 2)
    work_sur = (fb.sum() * MC_PATHS + spot_checked.sum() * SPOT_PATHS) * sur.STEPS
    work_mc = n * MC_PATHS * sur.STEPS
    print(f"sources: network {net.sum()}, out-of-box MC {(src == 1).sum()}, spot-check MC {(src == 2).sum()}, "
          f"closed {(src == 3).sum()}, judge MC {(src == 4).sum()}")
    print(f"reference: {REF_PATHS} paths, median SE/S {np.median(ref_se / s):.2e}, RMS SE/S {math.sqrt(ref_ms):.2e}")
    print(f"whole mix RMS |err|/S: surrogate+fallback {rms_sur:.3e} "
          f"(net of reference noise {math.sqrt(max(rms_sur ** 2 - ref_ms, 0)):.3e}); "
          f"network answers {math.sqrt(np.mean(e_sur[net] ** 2)):.3e}, MC answers "
          f"{math.sqrt(np.mean(e_sur[~net] ** 2)) if (~net).any() else 0:.3e}; plain MC {rms_mc:.3e}")
    print(f"price-relative error, prices >= {PRICE_FLOOR:.0%} of spot ({big.sum()} queries): surrogate RMS "
          f"{math.sqrt(np.mean(rel ** 2)):.3e} p99 {np.quantile(rel, 0.99):.3e}; plain MC RMS "
          f"{math.sqrt(np.mean(rel_mc ** 2)):.3e}")
    print(f"path-steps simulated: surrogate+fallback {work_sur:.3e}, plain MC {work_mc:.3e} "
          f"({work_mc / work_sur:.1f}x fewer)")
    assert np.all(np.isfinite(price))
    assert (~_in_box(d)).sum() == (src == 1).sum()
    assert rms_sur <= TARGET
    # what the network gains: a smaller error than plain MC at 65,536 paths,
    # with over 10x fewer simulated path-steps
    assert rms_sur < 0.5 * rms_mc
    assert work_mc > 10 * work_sur


def test_full_million_query_mix_properties(served):
    qk = served
    n = 1 << 20
    d = _mix(n, sur.MASTER_QUERY + 16)
    print(f"C4 full mix: {n} queries", flush=True)
    price, src = qk.sur_heston_price_batch(*[d[k] for k in KEYS], d["option_type"], d["paths"], d["steps"],
                                           d["seed"])
    inside = _in_box(d)
    counts = {k: int((src == k).sum()) for k in range(5)}
    judge = qk.sur_last_judge()
    print("sources", counts, "outside the box", int((~inside).sum()))
    print(f"judge over {int(judge[16, 0])} checks: implied RMS error/S {math.sqrt(max(judge[16, 1], 0)):.2e} "
          f"(SE of the MSE {judge[16, 2]:.1e}); regions rejected {int(judge[:16, 3].sum())}")
    assert np.all(np.isfinite(price))
    # no invalid query in the mix, and the trained network passes the judge
    assert set(np.unique(src)) <= {0, 1, 2}
    # the box test: out-of-box queries go to MC, in-box queries to the
    # network unless their spot check failed (only checked queries can)
    assert np.array_equal(src == 1, ~inside)
    checked = np.zeros(n, bool)
    checked[::SPOT_EVERY] = True
    assert not np.any((src == 2) & ~checked)
    assert counts[2] <= 0.05 * (checked & inside).sum()
    # every MC answer is the Heston MC batch entry point's, bitwise
    fb = np.nonzero(src != 0)[0]
    sub = {k: v[fb] for k, v in d.items()}
    mc = qk.heston_monte_carlo_price_batch(*[sub[k] for k in KEYS], sub["option_type"], sub["paths"],
                                           sub["steps"], sub["seed"])
    assert np.array_equal(price[fb].view(np.uint64), mc.view(np.uint64))
    # network answers respect the no-arbitrage bounds of calls and puts
    net = src == 0
    s, k, t, r = d["spot"][net], d["strike"][net], d["t"][net], d["r"][net]
    dk = k * np.exp(-r * t)
    call = d["option_type"][net] == 0
    lo = np.where(call, np.maximum(s - dk, 0.0), np.maximum(dk - s, 0.0))
    hi = np.where(call, s, dk)
    tol = 1e-12 * s
    assert np.all(price[net] >= lo - tol) and np.all(price[net] <= hi + tol)


def test_batch_judge_catches_a_planted_bias_the_per_query_check_misses(served):
    """A bias of 2e-3 of spot planted in the network's output bias b5 (every
    call, and by parity every put, moves by 2e-3 S: twice the target).  The
    per-query check sees it only on checked queries whose MC standard error
    is small; the batch judge, pooling the checks of each region, rejects the
    batch and sends every network answer to MC."""
    import mlp_oracle as mo
    qk = served
    n, every = 65536, 8
    d = sur.sample_queries(n, MASTER_C4_TEST + 2)
    d["option_type"] = (np.arange(n) % 2).astype(np.int32)
    d["paths"] = np.full(n, MC_PATHS, np.int32)
    args = [d[k] for k in KEYS] + [d["option_type"], d["paths"], d["steps"], d["seed"]]
    clean_params = qk.sur_get_params()
    try:
        qk.sur_configure(spot_every=0)
        clean, _ = qk.sur_heston_price_batch(*args)
        # the trained network, judged: no region rejected, nothing sent to MC by the judge
        qk.sur_configure(spot_every=every, spot_paths=SPOT_PATHS, spot_k_se=3.0)
        price0, src0 = qk.sur_heston_price_batch(*args)
        stats0 = qk.sur_last_judge()
        planted = clean_params.copy()
        planted[mo.KB5] += 2e-3
        qk.sur_set_params(planted)
        qk.sur_configure(spot_every=0)
        biased, _ = qk.sur_heston_price_batch(*args)
        qk.sur_configure(spot_every=every, spot_paths=SPOT_PATHS, spot_k_se=3.0)
        print(f"judge plant: {n} queries, every {every}th checked", flush=True)
        price, src = qk.sur_heston_price_batch(*args)
        stats = qk.sur_last_judge()
    finally:
        qk.sur_set_params(clean_params)
        qk.sur_configure(spot_every=SPOT_EVERY, spot_paths=SPOT_PATHS, spot_k_se=3.0)
    s = d["spot"]
    checked = np.arange(n) % every == 0
    print("clean: judge rows (checks, implied RMS/S, SE of MSE, rejected):")
    for r in stats0:
        print(f"  {int(r[0])} {math.sqrt(max(r[1], 0)):.2e} {r[2]:.1e} {int(r[3])}")
    print("planted: judge rows:")
    for r in stats:
        print(f"  {int(r[0])} {math.sqrt(max(r[1], 0)):.2e} {r[2]:.1e} {int(r[3])}")
    # clean network: nothing rejected, few per-query failures
    assert not np.any(stats0[:, 3]) and not np.any(src0 == 4)
    assert (src0 == 2).sum() <= 0.01 * checked.sum()
    assert math.sqrt(max(stats0[16, 1], 0.0)) < 0.3e-3
    # the plant moved the answers by 2e-3 S (less where the no-arbitrage clamp
    # holds a deep in- or out-of-the-money answer at its bound)
    shift = (biased - clean) / s
    assert np.all(shift <= 2e-3 + 2e-4) and np.mean(np.abs(shift - 2e-3) < 2e-4) > 0.8
    # per-query check alone: it catches only some checked queries, and the
    # answers it leaves are off by twice the target
    fail_frac = (src == 2).sum() / checked.sum()
    per_query_only = np.where(src == 2, price, biased)
    rms_pq = math.sqrt(np.mean(((per_query_only - clean) / s) ** 2))
    rms_judged = math.sqrt(np.mean(((price - clean) / s) ** 2))
    print(f"per-query failures {fail_frac:.1%} of checks; served RMS vs the clean network: "
          f"per-query check only {rms_pq:.2e}, with the batch judge {rms_judged:.2e}")
    assert fail_frac < 0.5 and rms_pq > 1.5 * TARGET
    # the batch judge: the batch and every region rejected, no network answer left
    assert stats[16, 3] == 1 and stats[:16, 3].sum() >= 12
    assert not np.any(src == 0) and set(np.unique(src)) <= {2, 4}
    assert rms_judged <= TARGET
