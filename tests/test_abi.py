"""C ABI conventions of libquantkernel.so, CPU only (no compute calls).

Mirrors the reference's cpp/tests/test_error_handling.cpp:5-108 and
qk_api.cpp:19-117: null pointers first, in declaration order, with
"null pointer: <param>"; then n <= 0 with "bad batch size: <n>"; a sticky,
thread-local message cleared only by qk_clear_last_error; ABI negotiation.
"""
import ctypes as ct
import re
import subprocess
import threading
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
LIB = ROOT / "quant-kernel_amd" / "libquantkernel.so"


def declared_symbols():
    text = "".join(p.read_text() for p in (ROOT / "include").rglob("*.h"))
    return sorted(set(re.findall(r"QK_EXPORT[^;(]*?\b(qk_\w+)\s*\(", text, re.S)))


def test_every_declared_symbol_is_exported():
    syms = declared_symbols()
    assert len(syms) >= 30
    out = subprocess.run(["nm", "-D", "--defined-only", str(LIB)], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (qk_\w+)", out))
    missing = [s for s in syms if s not in exported]
    assert not missing, missing


def test_reference_mc_subset_present(qk):
    # the reference's MC entry points (qk_api.h:259-403) that callers bind
    for name in ["standard_monte_carlo", "antithetic_variates", "euler_maruyama", "milstein",
                 "longstaff_schwartz", "heston_monte_carlo", "local_vol_monte_carlo"]:
        getattr(qk._lib, f"qk_mcm_{name}_price")
        getattr(qk._lib, f"qk_mcm_{name}_price_batch")


def test_abi_version_and_plugin_negotiation(qk):
    assert qk.abi_version() == (2, 13)
    major, minor, name = qk.plugin_get_api(2, 12)
    assert (major, minor) == (2, 13) and "gfx950" in name
    lib = qk._lib
    from quantkernel_amd.engine import QKPluginAPI
    out = ct.POINTER(QKPluginAPI)()
    assert lib.qk_plugin_get_api(2, 14, ct.byref(out)) == -3 and not out
    assert lib.qk_plugin_get_api(3, 0, ct.byref(out)) == -3
    assert lib.qk_plugin_get_api(2, 13, None) == -1


def _gbm_args(n=2):
    return dict(spot=np.full(n, 100.0), strike=np.full(n, 100.0), t=np.ones(n), vol=np.full(n, 0.2),
                r=np.full(n, 0.05), q=np.zeros(n), option_type=np.zeros(n, np.int32),
                paths=np.full(n, 1000, np.int32), steps=np.full(n, 10, np.int32),
                seed=np.arange(n, dtype=np.uint64), out_prices=np.empty(n))


def _ptr(a):
    return a.ctypes.data_as(ct.c_void_p)


STEPWISE = ["qk_mcm_euler_maruyama_price_batch", "qk_mcm_milstein_price_batch",
            "qk_mcm_longstaff_schwartz_price_batch", "qk_mcm_local_vol_monte_carlo_price_batch"]
ORDER = ["spot", "strike", "t", "vol", "r", "q", "option_type", "paths", "steps", "seed"]


def _call(name, qk, argv, n):
    fn = getattr(ct.CDLL(str(LIB)), name)
    fn.restype = ct.c_int32
    return fn(*argv[:-1], ct.c_int32(n), argv[-1])


@pytest.mark.parametrize("name", STEPWISE)
def test_null_pointer_names_in_declaration_order(qk, name):
    a = _gbm_args()
    names = ORDER + ["out_prices"]
    for k in range(len(names)):
        argv = [None if j == k else _ptr(a[nm]) for j, nm in enumerate(names)]
        rc = _call(name, qk, argv, 2)
        assert rc == -1
        assert qk.get_last_error() == f"null pointer: {names[k]}"
    # two nulls: the first in declaration order is named (test_error_handling.cpp: "null" not "spot")
    argv = [_ptr(a[nm]) for nm in names]
    argv[1] = None
    argv[3] = None
    assert _call(name, qk, argv, 2) == -1 and qk.get_last_error() == "null pointer: strike"


def test_heston_null_order_and_bad_size(qk):
    d = {k: np.ones(3) for k in ("spot", "strike", "t", "r", "q", "v0", "kappa", "theta", "sigma", "rho")}
    ints = {k: np.ones(3, np.int32) for k in ("option_type", "paths", "steps")}
    names = list(d) + list(ints) + ["seed", "out_prices"]
    arrays = {**d, **ints, "seed": np.ones(3, np.uint64), "out_prices": np.empty(3)}
    for k in range(len(names)):
        argv = [None if j == k else _ptr(arrays[nm]) for j, nm in enumerate(names)]
        assert _call("qk_mcm_heston_monte_carlo_price_batch", qk, argv, 3) == -1
        assert qk.get_last_error() == f"null pointer: {names[k]}"
    argv = [_ptr(arrays[nm]) for nm in names]
    for n in (0, -5):
        assert _call("qk_mcm_heston_monte_carlo_price_batch", qk, argv, n) == -2
        assert qk.get_last_error() == f"bad batch size: {n}"


def test_error_is_sticky_thread_local_and_clearable(qk):
    a = _gbm_args()
    argv = [_ptr(a[nm]) for nm in ORDER + ["out_prices"]]
    assert _call("qk_mcm_euler_maruyama_price_batch", qk, argv, 0) == -2
    msg = qk.get_last_error()
    assert msg == "bad batch size: 0"
    qk.set_rng_mode(0)  # a successful call does not clear it
    assert qk.get_last_error() == msg
    seen = []
    th = threading.Thread(target=lambda: seen.append(qk.get_last_error()))
    th.start()
    th.join()
    assert seen == [""]
    qk.clear_last_error()
    assert qk.get_last_error() == ""


def test_rng_mode_validation(qk):
    assert qk._lib.qk_mc_set_rng_mode(9) == -5
    assert "rng mode" in qk.get_last_error()
    for m in (1, 2, 0):
        qk.set_rng_mode(m)
        assert qk.get_rng_mode() == m


def test_lsmc_basis_validation(qk):
    a = _gbm_args()
    argv = [_ptr(a[nm]) for nm in ORDER]
    fn = qk._lib.qk_mcm_longstaff_schwartz_price_batch_ex
    se = np.empty(2)
    rc = fn(*[ct.cast(x, ct.POINTER(ct.c_double)) for x in argv[:6]],
            *[ct.cast(x, ct.POINTER(ct.c_int32)) for x in argv[6:9]],
            ct.cast(argv[9], ct.POINTER(ct.c_uint64)), 7, 2, a["out_prices"].ctypes.data_as(ct.POINTER(ct.c_double)),
            se.ctypes.data_as(ct.POINTER(ct.c_double)))
    assert rc == -5 and "basis" in qk.get_last_error()


def _has_gpu():
    buf = ct.create_string_buffer(64)
    cus = ct.c_int32()
    return ct.CDLL(str(LIB)).qk_device_info(buf, 64, ct.byref(cus)) == 0


def test_no_device_means_runtime_init_not_fallback(qk):
    """Without a gfx950 device every compute entry point fails loudly."""
    if _has_gpu():
        pytest.skip("a device is present")
    a = _gbm_args()
    argv = [_ptr(a[nm]) for nm in ORDER + ["out_prices"]]
    assert _call("qk_mcm_euler_maruyama_price_batch", qk, argv, 2) == -4
    assert "device" in qk.get_last_error()
    with pytest.raises(ValueError, match="NaN"):  # scalar NaN raises, as in the reference wrapper
        qk.euler_maruyama_price(100, 100, 1, .2, .05, 0, 0, 1000, 10, 1)
    assert np.isnan(qk._lib.qk_mcm_euler_maruyama_price(100, 100, 1, .2, .05, 0, 0, 1000, 10, 1))


def test_c_consumer_compiles_and_links(tmp_path):
    """The header is plain C and a C program links against the library."""
    src = tmp_path / "consumer.c"
    src.write_text(r'''
#include <stdio.h>
#include "quantkernel/qk_api.h"
int main(void) {
    int32_t ma = 0, mi = 0; const QKPluginAPI* api = 0;
    qk_abi_version(&ma, &mi);
    if (qk_plugin_get_api(2, 12, &api) != QK_OK || !api) return 1;
    double out[1];
    int32_t rc = qk_mcm_heston_monte_carlo_price_batch(0,0,0,0,0,0,0,0,0,0,0,0,0,0,1,out);
    printf("%d.%d %s rc=%d %s\n", ma, mi, api->plugin_name, rc, qk_get_last_error());
    return rc == QK_ERR_NULL_PTR ? 0 : 2;
}''')
    exe = tmp_path / "consumer"
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", f"-I{ROOT / 'include'}", str(src), "-o", str(exe),
                    f"-L{LIB.parent}", "-lquantkernel", f"-Wl,-rpath,{LIB.parent}"], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "null pointer: spot" in out.stdout
