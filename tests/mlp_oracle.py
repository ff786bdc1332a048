"""CPU oracle of the surrogate MLP (test infrastructure only).

FP64 numpy restatement of the library's forward and backward passes with the
same BF16 roundings the GPU applies (round-to-nearest-even on the GEMM
operands: inputs, weights, hidden activations, and the activation gradients
fed to the weight and data GEMMs).  The output layer is FP32/FP64 on both
sides.  ``exact=False`` drops every rounding: the plain FP64 network, whose
backward pass is pinned by central finite differences in the tests.
"""
import math

import numpy as np
from scipy.special import erf

K_IN, K_INPAD, K_H = 8, 32, 256
KW = [0]
KB = []
for _k in (K_INPAD, K_H, K_H, K_H):
    KB.append(KW[-1] + K_H * _k)
    KW.append(KB[-1] + K_H)
KW5 = KW.pop()  # the slot after b4 is the output row
KB5 = KW5 + K_H
KPARAMS = KB5 + 4
KK = [K_INPAD, K_H, K_H, K_H]

BOX_LO = np.array([math.log(0.7), 0.05, 0.0, 0.5, 0.01, 0.1, -0.9, 0.01])
BOX_HI = np.array([math.log(1.3), 2.0, 0.08, 5.0, 0.16, 1.0, -0.1, 0.16])


def bf16(x):
    x = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = x.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return r.astype(np.uint32).view(np.float32).astype(np.float64)


def normalise(feat):
    f = np.asarray(feat, dtype=np.float64)
    mid, half = 0.5 * (BOX_LO + BOX_HI), 0.5 * (BOX_HI - BOX_LO)
    return ((f - mid) / half).astype(np.float32).astype(np.float64)


def control(feat):
    """Black-Scholes call/S at the expected average Heston variance (q = 0)."""
    from scipy.special import ndtr
    f = np.asarray(feat, dtype=np.float64)
    lk, t, r, kappa, theta, v0 = f[:, 0], f[:, 1], f[:, 2], f[:, 3], f[:, 4], f[:, 7]
    kt = kappa * t
    w = np.where(kt > 1e-8, (1.0 - np.exp(-kt)) / np.maximum(kt, 1e-300), 1.0 - 0.5 * kt)
    var = theta + (v0 - theta) * w
    sd = np.sqrt(var * t)
    d1 = (-lk + (r + 0.5 * var) * t) / sd
    return ndtr(d1) - np.exp(lk - r * t) * ndtr(d1 - sd)


def gelu(x):
    return 0.5 * x * (1.0 + erf(x / math.sqrt(2.0)))


def gelu_grad(x):
    return 0.5 * (1.0 + erf(x / math.sqrt(2.0))) + x * np.exp(-0.5 * x * x) / math.sqrt(2 * math.pi)


def unpack(params):
    p = np.asarray(params, dtype=np.float64)
    Ws = [p[KW[l]:KW[l] + K_H * KK[l]].reshape(K_H, KK[l]) for l in range(4)]
    bs = [p[KB[l]:KB[l] + K_H] for l in range(4)]
    return Ws, bs, p[KW5:KW5 + K_H], p[KB5]


def forward_backward(params, feat, target, exact=True):
    """Returns (y, loss, grads[KPARAMS]) for raw features (n, 8)."""
    rnd = bf16 if exact else (lambda a: np.asarray(a, dtype=np.float64))
    Ws, bs, w5, b5 = unpack(params)
    n = feat.shape[0]
    X = np.zeros((n, K_INPAD))
    X[:, :K_IN] = normalise(feat)
    a = rnd(X)
    ins, Zb, Hf = [], [], None
    for l in range(4):
        ins.append(a)
        z = a @ rnd(Ws[l]).T + bs[l]
        if exact:
            z = z.astype(np.float32).astype(np.float64)
        h = gelu(z)
        Zb.append(z)
        Hf = h
        a = rnd(h)
    y = Hf @ w5 + b5
    t = np.asarray(target, dtype=np.float64)
    diff = y - t
    loss = float(np.mean(diff * diff))
    dy = 2.0 * diff / n
    g = np.zeros(KPARAMS)
    g[KW5:KW5 + K_H] = Hf.T @ dy
    g[KB5] = dy.sum()
    dh = np.outer(dy, w5)
    for l in range(3, -1, -1):
        dz = dh * gelu_grad(Zb[l])
        g[KB[l]:KB[l] + K_H] = dz.sum(axis=0)
        g[KW[l]:KW[l] + K_H * KK[l]] = (rnd(dz).T @ ins[l]).ravel()
        if l > 0:
            dh = rnd(dz) @ rnd(Ws[l])
    return y, loss, g


def init_params(seed=0):
    rng = np.random.default_rng(seed)
    p = np.zeros(KPARAMS, dtype=np.float32)
    for l in range(4):
        fan_in = K_IN if l == 0 else K_H
        w = rng.normal(0.0, math.sqrt(2.0 / fan_in), size=(K_H, fan_in))
        blk = np.zeros((K_H, KK[l]))
        blk[:, :fan_in] = w
        p[KW[l]:KW[l] + K_H * KK[l]] = blk.ravel()
        p[KB[l]:KB[l] + K_H] = rng.normal(0.0, 0.05, size=K_H)
    p[KW5:KW5 + K_H] = rng.normal(0.0, math.sqrt(1.0 / K_H), size=K_H)
    p[KB5] = 0.1
    return p
