"""The allocation-free device entry point, qk_mcm_mixed_price_batch_device_noalloc,
against qk_mcm_mixed_price_batch_device through the C ABI: same kernels and
tile shapes, so prices and standard errors must be bitwise equal.  Two calls
in a row, the second smaller (the fixed scratch keeps the first call's
contents), then the size limits."""
import ctypes as ct

import numpy as np
import pytest

from quantkernel_amd import workloads as wl

pytestmark = pytest.mark.gpu

ORDER = ("model", "spot", "strike", "t", "vol", "r", "q", "v0", "kappa", "theta", "sigma", "rho",
         "option_type", "paths", "steps", "seed")
MAX_OPTIONS = 65535
MAX_TILES = 1 << 19


def _batch(n, start, p_hi=20000, s_hi=64):
    d = wl.config2(n, start=start, p_lo=1000, p_hi=p_hi, s_lo=8, s_hi=s_hi)
    d["option_type"] = (np.arange(n) % 2).astype(np.int32)
    return d


def _call(qk, name, dev, n, out, se):
    fn = getattr(qk._lib, name)
    return fn(*[ct.c_void_p(dev[k].ptr) for k in ORDER], ct.c_int32(n), ct.c_void_p(out.ptr),
              ct.c_void_p(se.ptr))


def test_noalloc_equals_device_entry_on_two_calls(qk):
    from hip_mem import DevArray, to_device
    for n, start in ((3000, 0), (700, 50000)):  # the second call is smaller
        d = _batch(n, start)
        print(f"no-alloc parity: n={n} start={start}", flush=True)
        dev = to_device(d, ORDER)
        o1, s1, o2, s2 = DevArray(n=n), DevArray(n=n), DevArray(n=n), DevArray(n=n)
        for a in (o1, s1, o2, s2):
            a.fill_bytes(0xFF)  # NaN: every output must be written
        assert _call(qk, "qk_mcm_mixed_price_batch_device", dev, n, o1, s1) == 0, qk.get_last_error()
        assert _call(qk, "qk_mcm_mixed_price_batch_device_noalloc", dev, n, o2, s2) == 0, \
            qk.get_last_error()
        p1, e1, p2, e2 = o1.numpy(), s1.numpy(), o2.numpy(), s2.numpy()
        assert np.all(np.isfinite(p1)) and np.all(np.isfinite(e1))
        assert np.array_equal(p1.view(np.uint64), p2.view(np.uint64))
        assert np.array_equal(e1.view(np.uint64), e2.view(np.uint64))


def test_noalloc_rejects_batches_beyond_its_scratch(qk):
    from hip_mem import DevArray, to_device
    # more options than the record section holds: rejected before any device work
    n = MAX_OPTIONS + 1
    small = _batch(1, 0)
    dev = to_device(small, ORDER)
    o, s = DevArray(n=1), DevArray(n=1)
    qk.clear_last_error()
    assert _call(qk, "qk_mcm_mixed_price_batch_device_noalloc", dev, n, o, s) == -2
    assert "65535" in qk.get_last_error()
    # more path tiles than the partials section holds: no tile runs, NaN out
    lane, _ = qk.get_shape(qk.FAMILY_STEPWISE)
    per_option = -(-1_048_576 // (256 * lane))
    n = min(MAX_OPTIONS, MAX_TILES // per_option + 64)
    d = _batch(n, 0)
    d["paths"][:] = 1_048_576
    d["steps"][:] = 16
    print(f"no-alloc tile limit: n={n} tiles/option={per_option}", flush=True)
    dev = to_device(d, ORDER)
    o, s = DevArray(n=n), DevArray(n=n)
    qk.clear_last_error()
    assert _call(qk, "qk_mcm_mixed_price_batch_device_noalloc", dev, n, o, s) == -5
    assert "path tiles" in qk.get_last_error()
    assert np.all(np.isnan(o.numpy()))
    # and the next call that fits runs normally
    d = _batch(500, 7000)
    dev = to_device(d, ORDER)
    o1, s1, o2, s2 = DevArray(n=500), DevArray(n=500), DevArray(n=500), DevArray(n=500)
    assert _call(qk, "qk_mcm_mixed_price_batch_device", dev, 500, o1, s1) == 0
    assert _call(qk, "qk_mcm_mixed_price_batch_device_noalloc", dev, 500, o2, s2) == 0
    assert np.array_equal(o1.numpy().view(np.uint64), o2.numpy().view(np.uint64))
