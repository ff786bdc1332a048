"""Pins the CPU MLP oracle (CPU only): its FP64 backward pass against central
finite differences, its BF16 rounding helper against exact cases, and its
parameter layout against the library's."""
import numpy as np

import mlp_oracle as mo


def _data(n=64, seed=3):
    rng = np.random.default_rng(seed)
    f = mo.BOX_LO + (mo.BOX_HI - mo.BOX_LO) * rng.random((n, 8))
    t = rng.random(n) * 0.3
    return f, t


def test_fp64_backward_matches_finite_differences():
    p = mo.init_params(1).astype(np.float64)
    f, t = _data()
    _, _, g = mo.forward_backward(p, f, t, exact=False)
    rng = np.random.default_rng(0)
    idx = np.concatenate([rng.choice(np.arange(mo.KW[l], mo.KW[l] + mo.K_H * mo.KK[l]), 4) for l in range(4)] +
                         [np.arange(mo.KB[0], mo.KB[0] + 2), np.arange(mo.KW5, mo.KW5 + 3), [mo.KB5]])
    for i in idx:
        if mo.KW[0] <= i < mo.KB[0] and (i - mo.KW[0]) % mo.K_INPAD >= mo.K_IN:
            continue  # padded input column: no effect on the loss
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mo.forward_backward(pp, f, t, exact=False)[1] - mo.forward_backward(pm, f, t, exact=False)[1]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])) + 1e-9, (i, fd, g[i])


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -9, -2.5, 1e-30], dtype=np.float32)
    r = mo.bf16(x)
    # bf16 keeps 7 fraction bits: the ulp at 1 is 2^-7
    assert r[0] == 1.0
    assert r[1] == 1.0            # exactly half an ulp: ties to even
    assert r[2] == 1.0 + 2 ** -7  # three quarters of an ulp: rounds up
    assert r[3] == 1.0            # a quarter ulp: rounds down
    assert r[4] == -2.5


def test_layout_matches_library(qk):
    assert qk.sur_param_count() == mo.KPARAMS
    # padded slots: 24 zero input columns of layer 1 and 3 after the output bias
    assert mo.KPARAMS - 3 - (mo.K_INPAD - mo.K_IN) * mo.K_H == 199937
