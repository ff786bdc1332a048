"""The reference's Python loader accepts libquantkernel.so unchanged, CPU only.

The loader (python/quantkernel/_loader.py:311-345) resolves every name of its
signature table with getattr(lib, name) and then checks the ABI: major equal,
minor at least its own, and qk_plugin_get_api(major, minor) == QK_OK with a
matching abi_major.  tests/golden/loader_symbols.json holds that table's names
and the required ABI, written by tests/golden/gen_loader_symbols.py from the
loader's own text.  The functions outside the Monte Carlo path are stubs that
compute nothing: scalars return NaN, batches QK_ERR_RUNTIME_INIT, each with a
"not provided" last error, so that the reference wrapper raises on them.
"""
import ctypes as ct
import json
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
GOLDEN = json.loads((ROOT / "tests" / "golden" / "loader_symbols.json").read_text())
STUB_HEADER = ROOT / "include" / "quantkernel" / "qk_api_stubs.h"


def test_every_loader_symbol_resolves(qk):
    lib = qk._lib
    assert len(GOLDEN["symbols"]) == 128
    missing = [n for n in GOLDEN["symbols"] if not hasattr(lib, n)]
    assert not missing, missing


def test_loader_abi_check_passes(qk):
    lib = qk._lib
    major, minor = ct.c_int32(-1), ct.c_int32(-1)
    lib.qk_abi_version(ct.byref(major), ct.byref(minor))
    assert major.value == GOLDEN["abi_major"] and minor.value >= GOLDEN["abi_minor"]
    from quantkernel_amd.engine import QKPluginAPI
    api = ct.POINTER(QKPluginAPI)()
    assert lib.qk_plugin_get_api(GOLDEN["abi_major"], GOLDEN["abi_minor"], ct.byref(api)) == 0
    assert api and api.contents.abi_major == GOLDEN["abi_major"]


def test_stubs_cover_exactly_the_non_mc_names(qk):
    import re
    stubs = set(re.findall(r"\b(qk_\w+)\s*\(", STUB_HEADER.read_text()))
    own = set(re.findall(r"QK_EXPORT[^;(]*?\b(qk_\w+)\s*\(",
                         (ROOT / "include" / "quantkernel" / "qk_api.h").read_text(), re.S))
    names = set(GOLDEN["symbols"])
    assert not (stubs & own)
    assert names <= stubs | own
    assert stubs <= names
    # the Monte Carlo path is implemented, never stubbed
    assert not [n for n in stubs if n.startswith("qk_mcm_")]
    assert "qk_cf_heston_price_cf" in own and "qk_cf_heston_price_cf_batch" in own


def test_stub_scalar_returns_nan_with_message(qk):
    lib = qk._lib
    fn = lib.qk_cf_black_scholes_merton_price
    fn.restype = ct.c_double
    fn.argtypes = [ct.c_double] * 6 + [ct.c_int32]
    lib.qk_clear_last_error()
    assert np.isnan(fn(100.0, 100.0, 1.0, 0.2, 0.05, 0.0, 0))
    assert qk.get_last_error() == ("qk_cf_black_scholes_merton_price: not provided by the MI355X "
                               "Monte Carlo library")


def test_stub_batch_returns_runtime_init_and_leaves_output(qk):
    lib = qk._lib
    fn = lib.qk_cf_black_scholes_merton_price_batch
    fn.restype = ct.c_int32
    pd = ct.POINTER(ct.c_double)
    fn.argtypes = [pd] * 6 + [ct.POINTER(ct.c_int32), ct.c_int32, pd]
    n = 3
    x = np.full(n, 1.0)
    ot = np.zeros(n, np.int32)
    out = np.full(n, 7.0)
    p = x.ctypes.data_as(pd)
    rc = fn(p, p, p, p, p, p, ot.ctypes.data_as(ct.POINTER(ct.c_int32)), n, out.ctypes.data_as(pd))
    assert rc == -4  # QK_ERR_RUNTIME_INIT
    assert np.all(out == 7.0)
    assert qk.get_last_error().startswith("qk_cf_black_scholes_merton_price_batch: not provided")
