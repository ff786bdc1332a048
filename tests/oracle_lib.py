"""ctypes bindings of the CPU oracle (oracle/liboracle.so) and of the
reference build (oracle/_ref/libqkref.so).  Test infrastructure only."""
import ctypes as ct
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
D, I, U = ct.c_double, ct.c_int32, ct.c_uint64
PD, PI, PU = ct.POINTER(ct.c_double), ct.POINTER(ct.c_int32), ct.POINTER(ct.c_uint64)
U32P = ct.POINTER(ct.c_uint32)

_ORC = {
    "orc_mt_normals": (None, [U, ct.c_int64, PD]),
    "orc_standard_mc": (D, [D] * 6 + [I, I, U]),
    "orc_antithetic": (D, [D] * 6 + [I, I, U]),
    "orc_euler": (D, [D] * 6 + [I, I, I, U]),
    "orc_milstein": (D, [D] * 6 + [I, I, I, U]),
    "orc_local_vol": (D, [D] * 6 + [I, I, I, U]),
    "orc_heston": (D, [D] * 10 + [I, I, I, U]),
    "orc_lsmc": (D, [D] * 6 + [I, I, I, U, I]),
    "orc_philox4x32_10": (None, [U32P, U32P, U32P]),
    "orc_box_muller": (None, [ct.c_uint32, ct.c_uint32, PD, PD]),
    "orc_philox_path_normal": (D, [U, ct.c_uint32, ct.c_uint32]),
    "orc_philox_gbm_normal": (D, [U, ct.c_uint32, ct.c_uint32]),
    "orc_philox_terminal_normal": (D, [U, ct.c_uint32]),
    "orc_philox_heston_normals": (None, [U, ct.c_uint32, ct.c_uint32, PD, PD]),
    "orc_px_terminal": (D, [D] * 6 + [I, I, U, I, PD]),
    "orc_px_stepwise": (D, [I] + [D] * 6 + [I, I, I, U, PD]),
    "orc_px_heston": (D, [D] * 10 + [I, I, I, U, PD]),
    "orc_px_lsmc": (D, [D] * 6 + [I, I, I, U, I, PD]),
    "orc_bsm": (D, [D] * 6 + [I]),
    "orc_bsm_as": (D, [D] * 6 + [I]),
    "orc_heston_cf": (D, [D] * 10 + [I, I, D]),
    "orc_crr": (D, [D] * 6 + [I, I, I]),
    "orc_importance_sampling": (D, [D] * 6 + [I, I, D, U]),
    "orc_pathwise_delta": (D, [D] * 6 + [I, I, U]),
    "orc_lr_delta": (D, [D] * 6 + [I, I, U, D]),
    "orc_heston_lr_delta": (D, [D] * 10 + [I, I, I, U, D]),
    "orc_px_importance_sampling": (D, [D] * 6 + [I, I, D, U, PD]),
    "orc_px_pathwise_delta": (D, [D] * 6 + [I, I, U, PD]),
    "orc_px_lr_delta": (D, [D] * 6 + [I, I, U, D, PD]),
    "orc_px_heston_lr_delta": (D, [D] * 10 + [I, I, I, U, D, PD]),
    "orc_control_variates": (D, [D] * 6 + [I, I, U]),
    "orc_stratified": (D, [D] * 6 + [I, I, U]),
    "orc_mt_uniforms": (None, [U, ct.c_int64, PD]),
    "orc_acklam": (D, [D]),
    "orc_sobol_point": (D, [ct.c_uint32]),
    "orc_halton2_point": (D, [ct.c_uint32]),
    "orc_qmc": (D, [I] + [D] * 6 + [I, I]),
    "orc_mlmc": (D, [D] * 6 + [I, I, I, I, U]),
    "orc_px_control_variates": (D, [D] * 6 + [I, I, U, PD]),
    "orc_px_stratified": (D, [D] * 6 + [I, I, U, PD]),
    "orc_px_mlmc": (D, [D] * 6 + [I, I, I, I, U, PD]),
    "orc_heston_batch": (None, [PD] * 10 + [PI, PI, PI, PU, I, PD]),
    "orc_standard_mc_batch": (None, [PD] * 6 + [PI, PI, PU, I, PD]),
}


class Oracle:
    def __init__(self, path=None):
        p = Path(path) if path else ROOT / "oracle" / "liboracle.so"
        if not p.exists():
            raise RuntimeError(f"oracle not built: {p} (run __graft_entry__.build())")
        self.lib = ct.CDLL(str(p))
        for name, (res, args) in _ORC.items():
            fn = getattr(self.lib, name)
            fn.restype, fn.argtypes = res, args

    def __getattr__(self, name):
        return getattr(self.lib, "orc_" + name)

    def mt_normals(self, seed, count):
        out = np.empty(count, dtype=np.float64)
        self.lib.orc_mt_normals(seed, count, out.ctypes.data_as(PD))
        return out

    def philox(self, ctr, key):
        a = (ct.c_uint32 * 4)(*ctr)
        b = (ct.c_uint32 * 2)(*key)
        o = (ct.c_uint32 * 4)()
        self.lib.orc_philox4x32_10(a, b, o)
        return list(o)

    def heston_normals(self, seed, path, step):
        a, b = ct.c_double(), ct.c_double()
        self.lib.orc_philox_heston_normals(seed, path, step, ct.byref(a), ct.byref(b))
        return a.value, b.value

    def px(self, fn, *args):
        """Call an orc_px_* function; returns (price, stderr)."""
        se = ct.c_double()
        v = getattr(self.lib, "orc_px_" + fn)(*args, ct.byref(se))
        return v, se.value


_REF = {
    "qk_mcm_standard_monte_carlo_price": (D, [D] * 6 + [I, I, U]),
    "qk_mcm_antithetic_variates_price": (D, [D] * 6 + [I, I, U]),
    "qk_mcm_euler_maruyama_price": (D, [D] * 6 + [I, I, I, U]),
    "qk_mcm_milstein_price": (D, [D] * 6 + [I, I, I, U]),
    "qk_mcm_local_vol_monte_carlo_price": (D, [D] * 6 + [I, I, I, U]),
    "qk_mcm_longstaff_schwartz_price": (D, [D] * 6 + [I, I, I, U]),
    "qk_mcm_heston_monte_carlo_price": (D, [D] * 10 + [I, I, I, U]),
    "qk_mcm_heston_monte_carlo_price_batch": (I, [PD] * 10 + [PI, PI, PI, PU, I, PD]),
    "qk_mcm_standard_monte_carlo_price_batch": (I, [PD] * 6 + [PI, PI, PU, I, PD]),
    "qk_cf_black_scholes_merton_price": (D, [D] * 6 + [I]),
    "qk_mcm_importance_sampling_price": (D, [D] * 6 + [I, I, D, U]),
    "qk_agm_pathwise_derivative_delta": (D, [D] * 6 + [I, I, U]),
    "qk_agm_likelihood_ratio_delta": (D, [D] * 6 + [I, I, U, D]),
    "qk_mcm_heston_lr_delta": (D, [D] * 10 + [I, I, I, U, D]),
    "qk_cf_heston_price_cf": (D, [D] * 10 + [I, I, D]),
    "qk_mcm_control_variates_price": (D, [D] * 6 + [I, I, U]),
    "qk_mcm_stratified_sampling_price": (D, [D] * 6 + [I, I, U]),
    "qk_mcm_quasi_monte_carlo_sobol_price": (D, [D] * 6 + [I, I]),
    "qk_mcm_quasi_monte_carlo_halton_price": (D, [D] * 6 + [I, I]),
    "qk_mcm_multilevel_monte_carlo_price": (D, [D] * 6 + [I, I, I, I, U]),
}


class Reference:
    """The reference library itself, built from its sources by oracle/Makefile."""

    def __init__(self, path):
        self.lib = ct.CDLL(str(path))
        for name, (res, args) in _REF.items():
            fn = getattr(self.lib, name)
            fn.restype, fn.argtypes = res, args

    def __getattr__(self, name):
        return getattr(self.lib, "qk_" + name)

    @staticmethod
    def load():
        p = ROOT / "oracle" / "_ref" / "libqkref.so"
        return Reference(p) if p.exists() else None
