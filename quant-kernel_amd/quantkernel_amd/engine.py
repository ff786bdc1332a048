"""Python host interface of the MI355X Monte Carlo batch-pricing library.

Mirrors the Monte Carlo part of the reference's ``QuantKernel`` class
(python/quantkernel/engine.py:21-41 error classes, :99-124 array coercion,
:163-194 ``_call_batch_ctypes``, :1352-1823 MC wrappers): the same method
names, argument order and meaning, NaN for invalid elements, and the same
rc -> exception mapping.  Every call goes through ``libquantkernel.so`` built
from ``quant-kernel_amd/csrc`` (HIP kernels for gfx950); there is no CPU
fallback, so a missing library or device raises.
"""
from __future__ import annotations

import ctypes as ct
import math
import os
from pathlib import Path

import numpy as np

QK_OK = 0
QK_ERR_NULL_PTR = -1
QK_ERR_BAD_SIZE = -2
QK_ERR_ABI_MISMATCH = -3
QK_ERR_RUNTIME_INIT = -4
QK_ERR_INVALID_INPUT = -5
QK_CALL = 0
QK_PUT = 1
ABI_MAJOR = 2
ABI_MINOR = 13

QK_RNG_PHILOX = 0
QK_RNG_MT19937_FP64 = 1
QK_RNG_MT19937_FP32 = 2

QK_MODEL_GBM_EULER = 0
QK_MODEL_GBM_MILSTEIN = 1
QK_MODEL_HESTON_EULER = 2
QK_MODEL_LOCAL_VOL = 3

QK_LSMC_BASIS_REFERENCE = 0
QK_LSMC_BASIS_CUBIC = 1


class QKError(RuntimeError):
    """Base exception for QuantKernel errors."""


class QKNullPointerError(QKError):
    """Raised when a null pointer is passed to a batch API."""


class QKBadSizeError(QKError):
    """Raised when an invalid batch size is passed."""


class QKInvalidInputError(QKError):
    """Raised when invalid input parameters are detected."""


class QKRuntimeInitError(QKError):
    """Raised when no usable gfx950 device is available."""


_ERROR_CODE_MAP = {
    QK_ERR_NULL_PTR: QKNullPointerError,
    QK_ERR_BAD_SIZE: QKBadSizeError,
    QK_ERR_INVALID_INPUT: QKInvalidInputError,
    QK_ERR_RUNTIME_INIT: QKRuntimeInitError,
}


class QKPluginAPI(ct.Structure):
    _fields_ = [("abi_major", ct.c_int32), ("abi_minor", ct.c_int32), ("plugin_name", ct.c_char_p)]


_D = ct.c_double
_I = ct.c_int32
_U = ct.c_uint64
_PD = ct.POINTER(ct.c_double)
_PI = ct.POINTER(ct.c_int32)
_PU = ct.POINTER(ct.c_uint64)

# Every exported symbol and its signature (restype, argtypes).
_SIGNATURES = {
    "qk_abi_version": (None, [_PI, _PI]),
    "qk_plugin_get_api": (_I, [_I, _I, ct.POINTER(ct.POINTER(QKPluginAPI))]),
    "qk_get_last_error": (ct.c_char_p, []),
    "qk_clear_last_error": (None, []),
    "qk_mcm_standard_monte_carlo_price": (_D, [_D] * 6 + [_I, _I, _U]),
    "qk_mcm_standard_monte_carlo_price_batch": (_I, [_PD] * 6 + [_PI, _PI, _PU, _I, _PD]),
    "qk_mcm_standard_monte_carlo_price_batch_se": (_I, [_PD] * 6 + [_PI, _PI, _PU, _I, _PD, _PD]),
    "qk_mcm_antithetic_variates_price": (_D, [_D] * 6 + [_I, _I, _U]),
    "qk_mcm_antithetic_variates_price_batch": (_I, [_PD] * 6 + [_PI, _PI, _PU, _I, _PD]),
    "qk_mcm_euler_maruyama_price": (_D, [_D] * 6 + [_I, _I, _I, _U]),
    "qk_mcm_euler_maruyama_price_batch": (_I, [_PD] * 6 + [_PI, _PI, _PI, _PU, _I, _PD]),
    "qk_mcm_euler_maruyama_price_batch_se": (_I, [_PD] * 6 + [_PI, _PI, _PI, _PU, _I, _PD, _PD]),
    "qk_mcm_milstein_price": (_D, [_D] * 6 + [_I, _I, _I, _U]),
    "qk_mcm_milstein_price_batch": (_I, [_PD] * 6 + [_PI, _PI, _PI, _PU, _I, _PD]),
    "qk_mcm_milstein_price_batch_se": (_I, [_PD] * 6 + [_PI, _PI, _PI, _PU, _I, _PD, _PD]),
    "qk_mcm_local_vol_monte_carlo_price": (_D, [_D] * 6 + [_I, _I, _I, _U]),
    "qk_mcm_local_vol_monte_carlo_price_batch": (_I, [_PD] * 6 + [_PI, _PI, _PI, _PU, _I, _PD]),
    "qk_mcm_local_vol_monte_carlo_price_batch_se": (_I, [_PD] * 6 + [_PI, _PI, _PI, _PU, _I, _PD, _PD]),
    "qk_mcm_longstaff_schwartz_price": (_D, [_D] * 6 + [_I, _I, _I, _U]),
    "qk_mcm_longstaff_schwartz_price_batch": (_I, [_PD] * 6 + [_PI, _PI, _PI, _PU, _I, _PD]),
    "qk_mcm_longstaff_schwartz_price_batch_ex": (_I, [_PD] * 6 + [_PI, _PI, _PI, _PU, _I, _I, _PD, _PD]),
    "qk_mcm_heston_monte_carlo_price": (_D, [_D] * 10 + [_I, _I, _I, _U]),
    "qk_mcm_heston_monte_carlo_price_batch": (_I, [_PD] * 10 + [_PI, _PI, _PI, _PU, _I, _PD]),
    "qk_mcm_heston_monte_carlo_price_batch_se": (_I, [_PD] * 10 + [_PI, _PI, _PI, _PU, _I, _PD, _PD]),
    "qk_mcm_importance_sampling_price": (_D, [_D] * 6 + [_I, _I, _D, _U]),
    "qk_mcm_importance_sampling_price_batch": (_I, [_PD] * 6 + [_PI, _PI, _PD, _PU, _I, _PD]),
    "qk_mcm_importance_sampling_price_batch_se": (_I, [_PD] * 6 + [_PI, _PI, _PD, _PU, _I, _PD, _PD]),
    "qk_mcm_heston_lr_delta": (_D, [_D] * 10 + [_I, _I, _I, _U, _D]),
    "qk_mcm_heston_lr_delta_batch": (_I, [_PD] * 10 + [_PI, _PI, _PI, _PU, _PD, _I, _PD]),
    "qk_mcm_heston_lr_delta_batch_se": (_I, [_PD] * 10 + [_PI, _PI, _PI, _PU, _PD, _I, _PD, _PD]),
    "qk_agm_pathwise_derivative_delta": (_D, [_D] * 6 + [_I, _I, _U]),
    "qk_agm_pathwise_derivative_delta_batch": (_I, [_PD] * 6 + [_PI, _PI, _PU, _I, _PD]),
    "qk_agm_pathwise_derivative_delta_batch_se": (_I, [_PD] * 6 + [_PI, _PI, _PU, _I, _PD, _PD]),
    "qk_agm_likelihood_ratio_delta": (_D, [_D] * 6 + [_I, _I, _U, _D]),
    "qk_agm_likelihood_ratio_delta_batch": (_I, [_PD] * 6 + [_PI, _PI, _PU, _PD, _I, _PD]),
    "qk_agm_likelihood_ratio_delta_batch_se": (_I, [_PD] * 6 + [_PI, _PI, _PU, _PD, _I, _PD, _PD]),
    "qk_mcm_control_variates_price": (_D, [_D] * 6 + [_I, _I, _U]),
    "qk_mcm_control_variates_price_batch": (_I, [_PD] * 6 + [_PI, _PI, _PU, _I, _PD]),
    "qk_mcm_control_variates_price_batch_se": (_I, [_PD] * 6 + [_PI, _PI, _PU, _I, _PD, _PD]),
    "qk_mcm_stratified_sampling_price": (_D, [_D] * 6 + [_I, _I, _U]),
    "qk_mcm_stratified_sampling_price_batch": (_I, [_PD] * 6 + [_PI, _PI, _PU, _I, _PD]),
    "qk_mcm_stratified_sampling_price_batch_se": (_I, [_PD] * 6 + [_PI, _PI, _PU, _I, _PD, _PD]),
    "qk_mcm_quasi_monte_carlo_sobol_price": (_D, [_D] * 6 + [_I, _I]),
    "qk_mcm_quasi_monte_carlo_sobol_price_batch": (_I, [_PD] * 6 + [_PI, _PI, _I, _PD]),
    "qk_mcm_quasi_monte_carlo_sobol_price_batch_se": (_I, [_PD] * 6 + [_PI, _PI, _I, _PD, _PD]),
    "qk_mcm_quasi_monte_carlo_halton_price": (_D, [_D] * 6 + [_I, _I]),
    "qk_mcm_quasi_monte_carlo_halton_price_batch": (_I, [_PD] * 6 + [_PI, _PI, _I, _PD]),
    "qk_mcm_quasi_monte_carlo_halton_price_batch_se": (_I, [_PD] * 6 + [_PI, _PI, _I, _PD, _PD]),
    "qk_mcm_multilevel_monte_carlo_price": (_D, [_D] * 6 + [_I, _I, _I, _I, _U]),
    "qk_mcm_multilevel_monte_carlo_price_batch": (_I, [_PD] * 6 + [_PI, _PI, _PI, _PI, _PU, _I, _PD]),
    "qk_mcm_multilevel_monte_carlo_price_batch_se": (_I, [_PD] * 6 + [_PI, _PI, _PI, _PI, _PU, _I, _PD, _PD]),
    "qk_cf_heston_price_cf": (_D, [_D] * 10 + [_I, _I, _D]),
    "qk_cf_heston_price_cf_batch": (_I, [_PD] * 10 + [_PI, _PI, _PD, _I, _PD]),
    "qk_cf_heston_price_cf_batch_device": (_I, [ct.c_void_p] * 13 + [_I, ct.c_void_p]),
    "qk_mcm_mixed_price_batch": (_I, [_PI] + [_PD] * 11 + [_PI, _PI, _PI, _PU, _I, _PD, _PD]),
    "qk_mcm_mixed_price_batch_device": (_I, [ct.c_void_p] * 16 + [_I, ct.c_void_p, ct.c_void_p]),
    "qk_mcm_mixed_price_batch_device_noalloc": (_I, [ct.c_void_p] * 16 + [_I, ct.c_void_p, ct.c_void_p]),
    "qk_mcm_standard_monte_carlo_price_batch_device": (_I, [ct.c_void_p] * 9 + [_I, ct.c_void_p, ct.c_void_p]),
    "qk_mc_set_rng_mode": (_I, [_I]),
    "qk_mc_get_rng_mode": (_I, []),
    "qk_mc_last_timing": (_I, [_PD, _PD]),
    "qk_mc_validation_stream": (_I, [_U, ct.c_int64, _PD]),
    "qk_device_info": (_I, [ct.c_char_p, _I, _PI]),
    "qk_mc_philox_probe8": (_I, [_U, ct.c_uint32, ct.c_uint32, ct.c_uint32, _I, ct.POINTER(ct.c_float)]),
    "qk_mc_philox_probe": (_I, [_U, ct.c_uint32, ct.c_uint32, ct.c_uint32, _I,
                                ct.POINTER(ct.c_uint32), ct.POINTER(ct.c_float)]),
    "qk_mc_graphs": (_I, [_I, _PI, ct.POINTER(ct.c_int64)]),
    "qk_mc_set_shape": (_I, [_I, _I, _I]),
    "qk_mc_get_shape": (_I, [_I, _PI, _PI]),
    "qk_mc_autotune": (_I, [_I, _PI, _PI, _PD]),
    "qk_sur_param_count": (_I, []),
    "qk_sur_init": (_I, [_U]),
    "qk_sur_train": (_I, [_PD, _PD, _I, _I, _I, _D, _D, _U, _PD]),
    "qk_sur_predict": (_I, [_PD, _I, _PD]),
    "qk_sur_gradients": (_I, [_PD, _PD, _I, ct.POINTER(ct.c_float), _PD]),
    "qk_sur_get_params": (_I, [ct.POINTER(ct.c_float)]),
    "qk_sur_set_params": (_I, [ct.POINTER(ct.c_float)]),
    "qk_sur_save": (_I, [ct.c_char_p]),
    "qk_sur_load": (_I, [ct.c_char_p]),
    "qk_sur_configure": (_I, [_I, _I, _D]),
    "qk_sur_heston_price_batch": (_I, [_PD] * 10 + [_PI, _PI, _PI, _PU, _I, _PD, _PI]),
    "qk_sur_heston_price_batch_device": (_I, [ct.c_void_p] * 14 + [_I, ct.c_void_p, ct.c_void_p]),
    "qk_sur_last_timing": (_I, [_PD, _PD]),
    "qk_sur_set_target": (_I, [_D]),
    "qk_sur_last_judge": (_I, [_PD, _I]),
}


def _default_lib_path() -> Path:
    env = os.environ.get("QK_LIB_PATH")
    if env:
        p = Path(env)
        return p / "libquantkernel.so" if p.is_dir() else p
    return Path(__file__).resolve().parent.parent / "libquantkernel.so"


def load_library(path: str | os.PathLike | None = None) -> ct.CDLL:
    """Load libquantkernel.so and declare every signature; raises if absent."""
    p = Path(path) if path is not None else _default_lib_path()
    if not p.exists():
        raise QKRuntimeInitError(
            f"libquantkernel.so not found at {p}; build it with __graft_entry__.build()"
        )
    lib = ct.CDLL(str(p))
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(lib, name)  # AttributeError if a symbol is missing
        fn.restype = res
        fn.argtypes = args
    return lib


class QuantKernel:
    """Monte Carlo batch pricing on MI355X behind the QuantKernel C ABI."""

    CALL = QK_CALL
    PUT = QK_PUT

    def __init__(self, lib_path: str | os.PathLike | None = None):
        self._lib = load_library(lib_path)
        major, minor = self.abi_version()
        if major != ABI_MAJOR or minor < 12:
            raise QKError(f"ABI mismatch: library {major}.{minor}")

    # ---- ABI / errors (engine.py:130-161) --------------------------------
    def abi_version(self) -> tuple[int, int]:
        major, minor = ct.c_int32(-1), ct.c_int32(-1)
        self._lib.qk_abi_version(ct.byref(major), ct.byref(minor))
        return int(major.value), int(minor.value)

    def plugin_get_api(self, host_abi_major: int = ABI_MAJOR, host_abi_minor: int = ABI_MINOR):
        out = ct.POINTER(QKPluginAPI)()
        rc = self._lib.qk_plugin_get_api(int(host_abi_major), int(host_abi_minor), ct.byref(out))
        if rc != QK_OK:
            raise QKError(f"qk_plugin_get_api failed (rc={rc}, host_abi={host_abi_major}.{host_abi_minor})")
        info = out.contents
        return int(info.abi_major), int(info.abi_minor), info.plugin_name.decode()

    def get_last_error(self) -> str:
        raw = self._lib.qk_get_last_error()
        return raw.decode("utf-8", errors="replace") if raw else ""

    def clear_last_error(self) -> None:
        self._lib.qk_clear_last_error()

    # ---- mode / device ---------------------------------------------------
    def set_rng_mode(self, mode: int) -> None:
        self._check(self._lib.qk_mc_set_rng_mode(int(mode)), "qk_mc_set_rng_mode")

    def get_rng_mode(self) -> int:
        return int(self._lib.qk_mc_get_rng_mode())

    def last_timing(self) -> tuple[float, float]:
        a, b = ct.c_double(), ct.c_double()
        self._check(self._lib.qk_mc_last_timing(ct.byref(a), ct.byref(b)), "qk_mc_last_timing")
        return a.value, b.value

    def device_info(self) -> tuple[str, int]:
        buf = ct.create_string_buffer(256)
        cus = ct.c_int32()
        self._check(self._lib.qk_device_info(buf, 256, ct.byref(cus)), "qk_device_info")
        return buf.value.decode(), int(cus.value)

    def graphs(self, enable: int = -1) -> tuple[int, int]:
        c, n = ct.c_int32(), ct.c_int64()
        self._check(self._lib.qk_mc_graphs(int(enable), ct.byref(c), ct.byref(n)), "qk_mc_graphs")
        return c.value, n.value

    FAMILY_TERMINAL = 1
    FAMILY_STEPWISE = 2

    def set_shape(self, family: int, lane_units: int, grid_per_cu: int) -> None:
        self._check(self._lib.qk_mc_set_shape(int(family), int(lane_units), int(grid_per_cu)), "qk_mc_set_shape")

    def get_shape(self, family: int) -> tuple[int, int]:
        a, b = ct.c_int32(), ct.c_int32()
        self._check(self._lib.qk_mc_get_shape(int(family), ct.byref(a), ct.byref(b)), "qk_mc_get_shape")
        return a.value, b.value

    def autotune(self, family: int) -> tuple[int, int, float]:
        a, b, t = ct.c_int32(), ct.c_int32(), ct.c_double()
        self._check(self._lib.qk_mc_autotune(int(family), ct.byref(a), ct.byref(b), ct.byref(t)), "qk_mc_autotune")
        return a.value, b.value, t.value

    def validation_stream(self, seed: int, count: int) -> np.ndarray:
        out = np.empty(int(count), dtype=np.float64)
        self._check(self._lib.qk_mc_validation_stream(int(seed), int(count), out.ctypes.data_as(_PD)),
                    "qk_mc_validation_stream")
        return out

    def philox_probe(self, seed: int, c0: int, c1: int, c3: int, count: int):
        words = np.empty(4 * count, dtype=np.uint32)
        normals = np.empty(4 * count, dtype=np.float32)
        self._check(self._lib.qk_mc_philox_probe(int(seed), int(c0), int(c1), int(c3), int(count),
                                                 words.ctypes.data_as(ct.POINTER(ct.c_uint32)),
                                                 normals.ctypes.data_as(ct.POINTER(ct.c_float))),
                    "qk_mc_philox_probe")
        return words.reshape(count, 4), normals.reshape(count, 4)

    def philox_probe8(self, seed: int, c0: int, c1: int, c3: int, count: int):
        """Eight normals per block of the four-pair layout, shape (count, 8)."""
        normals = np.empty(8 * count, dtype=np.float32)
        self._check(self._lib.qk_mc_philox_probe8(int(seed), int(c0), int(c1), int(c3), int(count),
                                                  normals.ctypes.data_as(ct.POINTER(ct.c_float))),
                    "qk_mc_philox_probe8")
        return normals.reshape(count, 8)

    # ---- Heston surrogate (ABI 2.13) --------------------------------------
    def sur_param_count(self) -> int:
        return int(self._lib.qk_sur_param_count())

    def sur_init(self, seed: int = 0) -> None:
        self._check(self._lib.qk_sur_init(int(seed)), "qk_sur_init")

    @staticmethod
    def _features(f) -> np.ndarray:
        f = np.ascontiguousarray(np.asarray(f, dtype=np.float64))
        if f.ndim != 2 or f.shape[1] != 8:
            raise ValueError("features must be an (n, 8) array")
        return f

    def sur_train(self, features, targets, epochs=10, batch=8192, lr=1e-3, lr_final=1e-5, seed=0):
        f = self._features(features)
        t = self._as_f64_array(targets, "targets")
        self._check_same_length(f.shape[0], targets=t)
        loss = np.empty(int(epochs))
        self._check(self._lib.qk_sur_train(f.ctypes.data_as(_PD), t.ctypes.data_as(_PD), f.shape[0],
                                           int(epochs), int(batch), float(lr), float(lr_final), int(seed),
                                           loss.ctypes.data_as(_PD)), "qk_sur_train")
        return loss

    def sur_predict(self, features) -> np.ndarray:
        f = self._features(features)
        out = np.empty(f.shape[0])
        self._check(self._lib.qk_sur_predict(f.ctypes.data_as(_PD), f.shape[0], out.ctypes.data_as(_PD)),
                    "qk_sur_predict")
        return out

    def sur_gradients(self, features, targets):
        f = self._features(features)
        t = self._as_f64_array(targets, "targets")
        g = np.empty(self.sur_param_count(), dtype=np.float32)
        loss = ct.c_double()
        self._check(self._lib.qk_sur_gradients(f.ctypes.data_as(_PD), t.ctypes.data_as(_PD), f.shape[0],
                                               g.ctypes.data_as(ct.POINTER(ct.c_float)), ct.byref(loss)),
                    "qk_sur_gradients")
        return g, loss.value

    def sur_get_params(self) -> np.ndarray:
        p = np.empty(self.sur_param_count(), dtype=np.float32)
        self._check(self._lib.qk_sur_get_params(p.ctypes.data_as(ct.POINTER(ct.c_float))), "qk_sur_get_params")
        return p

    def sur_set_params(self, params) -> None:
        p = np.ascontiguousarray(np.asarray(params, dtype=np.float32))
        if p.shape != (self.sur_param_count(),):
            raise ValueError("wrong parameter vector length")
        self._check(self._lib.qk_sur_set_params(p.ctypes.data_as(ct.POINTER(ct.c_float))), "qk_sur_set_params")

    def sur_save(self, path) -> None:
        self._check(self._lib.qk_sur_save(str(path).encode()), "qk_sur_save")

    def sur_load(self, path) -> None:
        self._check(self._lib.qk_sur_load(str(path).encode()), "qk_sur_load")

    def sur_configure(self, spot_every=128, spot_paths=8192, spot_k_se=3.0) -> None:
        self._check(self._lib.qk_sur_configure(int(spot_every), int(spot_paths), float(spot_k_se)),
                    "qk_sur_configure")

    def sur_heston_price_batch(self, spot, strike, t, r, q, v0, kappa, theta, sigma, rho, option_type,
                               paths, steps, seed):
        f = [self._as_f64_array(x, nm) for x, nm in
             zip((spot, strike, t, r, q, v0, kappa, theta, sigma, rho),
                 ("spot", "strike", "t", "r", "q", "v0", "kappa", "theta", "sigma", "rho"))]
        n = f[0].shape[0]
        ot, p, st, s = (self._as_i32_array(option_type, "option_type"), self._as_i32_array(paths, "paths"),
                        self._as_i32_array(steps, "steps"), self._as_u64_array(seed, "seed"))
        self._check_same_length(n, option_type=ot, paths=p, steps=st, seed=s)
        out = np.empty(n)
        src = np.empty(n, dtype=np.int32)
        self._check(self._lib.qk_sur_heston_price_batch(*[self._ptr(a) for a in f + [ot, p, st, s]], n,
                                                        out.ctypes.data_as(_PD), src.ctypes.data_as(_PI)),
                    "qk_sur_heston_price_batch")
        return out, src

    def sur_heston_price_batch_device(self, n: int, ptrs: dict, out_price_ptr: int, out_source_ptr: int):
        order = ("spot", "strike", "t", "r", "q", "v0", "kappa", "theta", "sigma", "rho", "option_type",
                 "paths", "steps", "seed")
        args = [ct.c_void_p(int(ptrs[k])) for k in order]
        self._check(self._lib.qk_sur_heston_price_batch_device(*args, int(n), ct.c_void_p(int(out_price_ptr)),
                                                               ct.c_void_p(int(out_source_ptr))),
                    "qk_sur_heston_price_batch_device")

    def sur_set_target(self, target_rms_over_spot: float = 1e-3) -> None:
        self._check(self._lib.qk_sur_set_target(float(target_rms_over_spot)), "qk_sur_set_target")

    def sur_last_judge(self) -> np.ndarray:
        """Batch judge of the last pricing call: rows {checks, implied MSE/S^2, SE, rejected};
        rows 0-15 the box regions, row 16 all checks."""
        out = np.empty((17, 4))
        rc = int(self._lib.qk_sur_last_judge(out.ctypes.data_as(_PD), 17))
        if rc < 0:
            self._check(rc, "qk_sur_last_judge")
        return out[:rc]

    def sur_last_timing(self) -> tuple[float, float]:
        a, b = ct.c_double(), ct.c_double()
        self._check(self._lib.qk_sur_last_timing(ct.byref(a), ct.byref(b)), "qk_sur_last_timing")
        return a.value, b.value

    # ---- helpers (engine.py:99-124, 163-194) -----------------------------
    @staticmethod
    def _as_f64_array(x, name: str) -> np.ndarray:
        arr = np.asarray(x, dtype=np.float64)
        if arr.ndim != 1:
            raise ValueError(f"{name} must be a 1D array-like")
        return np.ascontiguousarray(arr)

    @staticmethod
    def _as_i32_array(x, name: str) -> np.ndarray:
        arr = np.asarray(x, dtype=np.int32)
        if arr.ndim != 1:
            raise ValueError(f"{name} must be a 1D array-like")
        return np.ascontiguousarray(arr)

    @staticmethod
    def _as_u64_array(x, name: str) -> np.ndarray:
        arr = np.asarray(x, dtype=np.uint64)
        if arr.ndim != 1:
            raise ValueError(f"{name} must be a 1D array-like")
        return np.ascontiguousarray(arr)

    @staticmethod
    def _check_same_length(n: int, **arrays) -> None:
        for name, arr in arrays.items():
            if arr.shape[0] != n:
                raise ValueError(f"{name} length mismatch: expected {n}, got {arr.shape[0]}")

    def _check(self, rc: int, fn_name: str) -> None:
        if rc != QK_OK:
            detail = self.get_last_error()
            exc = _ERROR_CODE_MAP.get(rc, QKError)
            msg = f"{fn_name} failed (rc={rc})"
            if detail:
                msg += f": {detail}"
            raise exc(msg)

    @staticmethod
    def _ptr(arr: np.ndarray):
        if arr.dtype == np.float64:
            return arr.ctypes.data_as(_PD)
        if arr.dtype == np.int32:
            return arr.ctypes.data_as(_PI)
        if arr.dtype == np.uint64:
            return arr.ctypes.data_as(_PU)
        raise TypeError(f"Unsupported dtype for batch call: {arr.dtype}")

    def _call_batch(self, fn_name: str, *arrays, extra_int=None, with_se=False):
        n = arrays[0].shape[0]
        out = np.empty(n, dtype=np.float64)
        se = np.empty(n, dtype=np.float64) if with_se else None
        args = [self._ptr(a) for a in arrays]
        if extra_int is not None:
            args.append(int(extra_int))
        args.append(int(n))
        args.append(out.ctypes.data_as(_PD))
        if with_se:
            args.append(se.ctypes.data_as(_PD))
        rc = getattr(self._lib, fn_name)(*args)
        self._check(rc, fn_name)
        return (out, se) if with_se else out

    def _gbm_arrays(self, spot, strike, t, vol, r, q, option_type):
        spot = self._as_f64_array(spot, "spot")
        n = spot.shape[0]
        arrs = dict(
            strike=self._as_f64_array(strike, "strike"), t=self._as_f64_array(t, "t"),
            vol=self._as_f64_array(vol, "vol"), r=self._as_f64_array(r, "r"),
            q=self._as_f64_array(q, "q"), option_type=self._as_i32_array(option_type, "option_type"),
        )
        self._check_same_length(n, **arrs)
        return [spot] + list(arrs.values())

    # ---- standard / antithetic (engine.py:1352-1393, 1466-1483, 1616-1647) -
    def standard_monte_carlo_price(self, spot, strike, t, vol, r, q, option_type, paths=100000, seed=42):
        return self._call_checked("qk_mcm_standard_monte_carlo_price", spot, strike, t, vol, r, q,
                                  int(option_type), int(paths), int(seed))

    def standard_monte_carlo_price_batch(self, spot, strike, t, vol, r, q, option_type, paths, seed,
                                         return_stderr=False):
        a = self._gbm_arrays(spot, strike, t, vol, r, q, option_type)
        p, s = self._as_i32_array(paths, "paths"), self._as_u64_array(seed, "seed")
        self._check_same_length(a[0].shape[0], paths=p, seed=s)
        if return_stderr:
            return self._call_batch("qk_mcm_standard_monte_carlo_price_batch_se", *a, p, s, with_se=True)
        return self._call_batch("qk_mcm_standard_monte_carlo_price_batch", *a, p, s)

    def antithetic_variates_price(self, spot, strike, t, vol, r, q, option_type, paths=100000, seed=42):
        return self._call_checked("qk_mcm_antithetic_variates_price", spot, strike, t, vol, r, q,
                                  int(option_type), int(paths), int(seed))

    def antithetic_variates_price_batch(self, spot, strike, t, vol, r, q, option_type, paths, seed):
        a = self._gbm_arrays(spot, strike, t, vol, r, q, option_type)
        p, s = self._as_i32_array(paths, "paths"), self._as_u64_array(seed, "seed")
        self._check_same_length(a[0].shape[0], paths=p, seed=s)
        return self._call_batch("qk_mcm_antithetic_variates_price_batch", *a, p, s)

    # ---- stepwise GBM (engine.py:1394-1411, 1484-1517, 1777-1823) ---------
    def _call_checked(self, fn_name: str, *args) -> float:
        """Scalar call; NaN raises ValueError as in the reference (engine.py:93-97)."""
        before = self.get_last_error()
        out = float(getattr(self._lib, fn_name)(*args))
        if math.isnan(out):
            msg = f"{fn_name} returned NaN; check inputs."
            after = self.get_last_error()
            if after and after != before:
                msg += f" ({after})"
            raise ValueError(msg)
        return out

    def _stepwise_scalar(self, fn, spot, strike, t, vol, r, q, option_type, paths, steps, seed):
        return self._call_checked(fn, spot, strike, t, vol, r, q, int(option_type), int(paths), int(steps),
                                  int(seed))

    def _stepwise_batch(self, fn, spot, strike, t, vol, r, q, option_type, paths, steps, seed, se):
        a = self._gbm_arrays(spot, strike, t, vol, r, q, option_type)
        p, st, s = (self._as_i32_array(paths, "paths"), self._as_i32_array(steps, "steps"),
                    self._as_u64_array(seed, "seed"))
        self._check_same_length(a[0].shape[0], paths=p, steps=st, seed=s)
        if se:
            return self._call_batch(fn + "_se", *a, p, st, s, with_se=True)
        return self._call_batch(fn, *a, p, st, s)

    def euler_maruyama_price(self, spot, strike, t, vol, r, q, option_type, paths=10000, steps=100, seed=42):
        return self._stepwise_scalar("qk_mcm_euler_maruyama_price", spot, strike, t, vol, r, q, option_type,
                                     paths, steps, seed)

    def euler_maruyama_price_batch(self, spot, strike, t, vol, r, q, option_type, paths, steps, seed,
                                   return_stderr=False):
        return self._stepwise_batch("qk_mcm_euler_maruyama_price_batch", spot, strike, t, vol, r, q,
                                    option_type, paths, steps, seed, return_stderr)

    def milstein_price(self, spot, strike, t, vol, r, q, option_type, paths=10000, steps=100, seed=42):
        return self._stepwise_scalar("qk_mcm_milstein_price", spot, strike, t, vol, r, q, option_type,
                                     paths, steps, seed)

    def milstein_price_batch(self, spot, strike, t, vol, r, q, option_type, paths, steps, seed,
                             return_stderr=False):
        return self._stepwise_batch("qk_mcm_milstein_price_batch", spot, strike, t, vol, r, q,
                                    option_type, paths, steps, seed, return_stderr)

    def local_vol_monte_carlo_price(self, spot, strike, t, vol, r, q, option_type, paths=50000, steps=200,
                                    seed=42):
        return self._stepwise_scalar("qk_mcm_local_vol_monte_carlo_price", spot, strike, t, vol, r, q,
                                     option_type, paths, steps, seed)

    def local_vol_monte_carlo_price_batch(self, spot, strike, t, vol, r, q, option_type, paths, steps, seed,
                                          return_stderr=False):
        return self._stepwise_batch("qk_mcm_local_vol_monte_carlo_price_batch", spot, strike, t, vol, r, q,
                                    option_type, paths, steps, seed, return_stderr)

    # ---- LSMC (engine.py:1412-1420, 1518-1534) ---------------------------
    def longstaff_schwartz_price(self, spot, strike, t, vol, r, q, option_type, paths=10000, steps=50,
                                 seed=42):
        return self._stepwise_scalar("qk_mcm_longstaff_schwartz_price", spot, strike, t, vol, r, q,
                                     option_type, paths, steps, seed)

    def longstaff_schwartz_price_batch(self, spot, strike, t, vol, r, q, option_type, paths, steps, seed,
                                       basis=None, return_stderr=False):
        a = self._gbm_arrays(spot, strike, t, vol, r, q, option_type)
        p, st, s = (self._as_i32_array(paths, "paths"), self._as_i32_array(steps, "steps"),
                    self._as_u64_array(seed, "seed"))
        self._check_same_length(a[0].shape[0], paths=p, steps=st, seed=s)
        if basis is None and not return_stderr:
            return self._call_batch("qk_mcm_longstaff_schwartz_price_batch", *a, p, st, s)
        out, se = self._call_batch("qk_mcm_longstaff_schwartz_price_batch_ex", *a, p, st, s,
                                   extra_int=QK_LSMC_BASIS_REFERENCE if basis is None else basis, with_se=True)
        return (out, se) if return_stderr else out

    # ---- Heston (engine.py:1648-1718) ------------------------------------
    def heston_monte_carlo_price(self, spot, strike, t, r, q, v0, kappa, theta, sigma, rho, option_type,
                                 paths=50000, steps=100, seed=42):
        return self._call_checked("qk_mcm_heston_monte_carlo_price", spot, strike, t, r, q, v0, kappa, theta,
                                  sigma, rho, int(option_type), int(paths), int(steps), int(seed))

    def heston_monte_carlo_price_batch(self, spot, strike, t, r, q, v0, kappa, theta, sigma, rho,
                                       option_type, paths, steps, seed, return_stderr=False):
        f = [self._as_f64_array(x, nm) for x, nm in
             zip((spot, strike, t, r, q, v0, kappa, theta, sigma, rho),
                 ("spot", "strike", "t", "r", "q", "v0", "kappa", "theta", "sigma", "rho"))]
        n = f[0].shape[0]
        ot, p, st, s = (self._as_i32_array(option_type, "option_type"), self._as_i32_array(paths, "paths"),
                        self._as_i32_array(steps, "steps"), self._as_u64_array(seed, "seed"))
        self._check_same_length(n, **{nm: x for nm, x in zip(
            ("strike", "t", "r", "q", "v0", "kappa", "theta", "sigma", "rho"), f[1:])},
            option_type=ot, paths=p, steps=st, seed=s)
        if return_stderr:
            return self._call_batch("qk_mcm_heston_monte_carlo_price_batch_se", *f, ot, p, st, s, with_se=True)
        return self._call_batch("qk_mcm_heston_monte_carlo_price_batch", *f, ot, p, st, s)

    # ---- mixed-model batch (ABI 2.13) ------------------------------------
    # ---- Heston characteristic-function price (engine.py:362-372, 422-445) --
    def heston_price_cf(self, spot, strike, t, r, q, v0, kappa, theta, sigma, rho, option_type,
                        integration_steps=1024, integration_limit=120.0):
        return self._call_checked("qk_cf_heston_price_cf", spot, strike, t, r, q, v0, kappa, theta, sigma, rho,
                                  int(option_type), int(integration_steps), float(integration_limit))

    def heston_price_cf_batch(self, spot, strike, t, r, q, v0, kappa, theta, sigma, rho, option_type,
                              integration_steps, integration_limit):
        f = [self._as_f64_array(x, nm) for x, nm in
             zip((spot, strike, t, r, q, v0, kappa, theta, sigma, rho),
                 ("spot", "strike", "t", "r", "q", "v0", "kappa", "theta", "sigma", "rho"))]
        n = f[0].shape[0]
        ot, st, lim = (self._as_i32_array(option_type, "option_type"),
                       self._as_i32_array(integration_steps, "integration_steps"),
                       self._as_f64_array(integration_limit, "integration_limit"))
        self._check_same_length(n, **{nm: x for nm, x in zip(
            ("strike", "t", "r", "q", "v0", "kappa", "theta", "sigma", "rho"), f[1:])},
            option_type=ot, integration_steps=st, integration_limit=lim)
        return self._call_batch("qk_cf_heston_price_cf_batch", *f, ot, st, lim)

    def heston_price_cf_batch_device(self, n: int, ptrs: dict, out_price_ptr: int) -> None:
        """Device-resident Heston CF batch: ``ptrs`` maps argument names to device addresses."""
        order = ("spot", "strike", "t", "r", "q", "v0", "kappa", "theta", "sigma", "rho", "option_type",
                 "integration_steps", "integration_limit")
        self._check(self._lib.qk_cf_heston_price_cf_batch_device(
            *[ct.c_void_p(int(ptrs[k])) for k in order], int(n), ct.c_void_p(int(out_price_ptr))),
            "qk_cf_heston_price_cf_batch_device")

    # ---- control variates, stratified sampling, quasi-MC, multilevel MC
    #      (engine.py:1421-1482, 1535-1647) -----------------------------------
    def _seeded_batch(self, fn, spot, strike, t, vol, r, q, option_type, paths, seed, se):
        a = self._gbm_arrays(spot, strike, t, vol, r, q, option_type)
        p, s = self._as_i32_array(paths, "paths"), self._as_u64_array(seed, "seed")
        self._check_same_length(a[0].shape[0], paths=p, seed=s)
        if se:
            return self._call_batch(fn + "_se", *a, p, s, with_se=True)
        return self._call_batch(fn, *a, p, s)

    def _qmc_batch(self, fn, spot, strike, t, vol, r, q, option_type, paths, se):
        a = self._gbm_arrays(spot, strike, t, vol, r, q, option_type)
        p = self._as_i32_array(paths, "paths")
        self._check_same_length(a[0].shape[0], paths=p)
        if se:
            return self._call_batch(fn + "_se", *a, p, with_se=True)
        return self._call_batch(fn, *a, p)

    def control_variates_price(self, spot, strike, t, vol, r, q, option_type, paths, seed=42):
        return self._call_checked("qk_mcm_control_variates_price", spot, strike, t, vol, r, q,
                                  int(option_type), int(paths), int(seed))

    def control_variates_price_batch(self, spot, strike, t, vol, r, q, option_type, paths, seed,
                                     return_stderr=False):
        return self._seeded_batch("qk_mcm_control_variates_price_batch", spot, strike, t, vol, r, q,
                                  option_type, paths, seed, return_stderr)

    def stratified_sampling_price(self, spot, strike, t, vol, r, q, option_type, paths, seed=42):
        return self._call_checked("qk_mcm_stratified_sampling_price", spot, strike, t, vol, r, q,
                                  int(option_type), int(paths), int(seed))

    def stratified_sampling_price_batch(self, spot, strike, t, vol, r, q, option_type, paths, seed,
                                        return_stderr=False):
        return self._seeded_batch("qk_mcm_stratified_sampling_price_batch", spot, strike, t, vol, r, q,
                                  option_type, paths, seed, return_stderr)

    def quasi_monte_carlo_sobol_price(self, spot, strike, t, vol, r, q, option_type, paths):
        return self._call_checked("qk_mcm_quasi_monte_carlo_sobol_price", spot, strike, t, vol, r, q,
                                  int(option_type), int(paths))

    def quasi_monte_carlo_sobol_price_batch(self, spot, strike, t, vol, r, q, option_type, paths,
                                            return_stderr=False):
        return self._qmc_batch("qk_mcm_quasi_monte_carlo_sobol_price_batch", spot, strike, t, vol, r, q,
                               option_type, paths, return_stderr)

    def quasi_monte_carlo_halton_price(self, spot, strike, t, vol, r, q, option_type, paths):
        return self._call_checked("qk_mcm_quasi_monte_carlo_halton_price", spot, strike, t, vol, r, q,
                                  int(option_type), int(paths))

    def quasi_monte_carlo_halton_price_batch(self, spot, strike, t, vol, r, q, option_type, paths,
                                             return_stderr=False):
        return self._qmc_batch("qk_mcm_quasi_monte_carlo_halton_price_batch", spot, strike, t, vol, r, q,
                               option_type, paths, return_stderr)

    def multilevel_monte_carlo_price(self, spot, strike, t, vol, r, q, option_type, base_paths, levels,
                                     base_steps, seed=42):
        return self._call_checked("qk_mcm_multilevel_monte_carlo_price", spot, strike, t, vol, r, q,
                                  int(option_type), int(base_paths), int(levels), int(base_steps), int(seed))

    def multilevel_monte_carlo_price_batch(self, spot, strike, t, vol, r, q, option_type, base_paths, levels,
                                           base_steps, seed, return_stderr=False):
        a = self._gbm_arrays(spot, strike, t, vol, r, q, option_type)
        bp, lv, bs, s = (self._as_i32_array(base_paths, "base_paths"), self._as_i32_array(levels, "levels"),
                         self._as_i32_array(base_steps, "base_steps"), self._as_u64_array(seed, "seed"))
        self._check_same_length(a[0].shape[0], base_paths=bp, levels=lv, base_steps=bs, seed=s)
        fn = "qk_mcm_multilevel_monte_carlo_price_batch"
        if return_stderr:
            return self._call_batch(fn + "_se", *a, bp, lv, bs, s, with_se=True)
        return self._call_batch(fn, *a, bp, lv, bs, s)

    # ---- importance sampling and Greeks on the MC engine (engine.py:965-982,
    #      1448-1455, 1583-1598, 1720-1775, 2090-2121) ------------------------
    def importance_sampling_price(self, spot, strike, t, vol, r, q, option_type, paths, shift, seed=42):
        return self._call_checked("qk_mcm_importance_sampling_price", spot, strike, t, vol, r, q,
                                  int(option_type), int(paths), float(shift), int(seed))

    def importance_sampling_price_batch(self, spot, strike, t, vol, r, q, option_type, paths, shift, seed,
                                        return_stderr=False):
        a = self._gbm_arrays(spot, strike, t, vol, r, q, option_type)
        p, h, s = (self._as_i32_array(paths, "paths"), self._as_f64_array(shift, "shift"),
                   self._as_u64_array(seed, "seed"))
        self._check_same_length(a[0].shape[0], paths=p, shift=h, seed=s)
        fn = "qk_mcm_importance_sampling_price_batch"
        if return_stderr:
            return self._call_batch(fn + "_se", *a, p, h, s, with_se=True)
        return self._call_batch(fn, *a, p, h, s)

    def pathwise_derivative_delta(self, spot, strike, t, vol, r, q, option_type, paths=20000, seed=42):
        return self._call_checked("qk_agm_pathwise_derivative_delta", spot, strike, t, vol, r, q,
                                  int(option_type), int(paths), int(seed))

    def pathwise_derivative_delta_batch(self, spot, strike, t, vol, r, q, option_type, paths, seed,
                                        return_stderr=False):
        a = self._gbm_arrays(spot, strike, t, vol, r, q, option_type)
        p, s = self._as_i32_array(paths, "paths"), self._as_u64_array(seed, "seed")
        self._check_same_length(a[0].shape[0], paths=p, seed=s)
        fn = "qk_agm_pathwise_derivative_delta_batch"
        if return_stderr:
            return self._call_batch(fn + "_se", *a, p, s, with_se=True)
        return self._call_batch(fn, *a, p, s)

    def likelihood_ratio_delta(self, spot, strike, t, vol, r, q, option_type, paths=20000, seed=42,
                               weight_clip=6.0):
        return self._call_checked("qk_agm_likelihood_ratio_delta", spot, strike, t, vol, r, q,
                                  int(option_type), int(paths), int(seed), float(weight_clip))

    def likelihood_ratio_delta_batch(self, spot, strike, t, vol, r, q, option_type, paths, seed, weight_clip,
                                     return_stderr=False):
        a = self._gbm_arrays(spot, strike, t, vol, r, q, option_type)
        p, s, w = (self._as_i32_array(paths, "paths"), self._as_u64_array(seed, "seed"),
                   self._as_f64_array(weight_clip, "weight_clip"))
        self._check_same_length(a[0].shape[0], paths=p, seed=s, weight_clip=w)
        fn = "qk_agm_likelihood_ratio_delta_batch"
        if return_stderr:
            return self._call_batch(fn + "_se", *a, p, s, w, with_se=True)
        return self._call_batch(fn, *a, p, s, w)

    def heston_lr_delta(self, spot, strike, t, r, q, v0, kappa, theta, sigma, rho, option_type, paths, steps,
                        seed=42, weight_clip=6.0):
        return self._call_checked("qk_mcm_heston_lr_delta", spot, strike, t, r, q, v0, kappa, theta, sigma, rho,
                                  int(option_type), int(paths), int(steps), int(seed), float(weight_clip))

    def heston_lr_delta_batch(self, spot, strike, t, r, q, v0, kappa, theta, sigma, rho, option_type, paths,
                              steps, seed, weight_clip, return_stderr=False):
        f = [self._as_f64_array(x, nm) for x, nm in
             zip((spot, strike, t, r, q, v0, kappa, theta, sigma, rho),
                 ("spot", "strike", "t", "r", "q", "v0", "kappa", "theta", "sigma", "rho"))]
        n = f[0].shape[0]
        ot, p, st, s, w = (self._as_i32_array(option_type, "option_type"), self._as_i32_array(paths, "paths"),
                           self._as_i32_array(steps, "steps"), self._as_u64_array(seed, "seed"),
                           self._as_f64_array(weight_clip, "weight_clip"))
        self._check_same_length(n, **{nm: x for nm, x in zip(
            ("strike", "t", "r", "q", "v0", "kappa", "theta", "sigma", "rho"), f[1:])},
            option_type=ot, paths=p, steps=st, seed=s, weight_clip=w)
        fn = "qk_mcm_heston_lr_delta_batch"
        if return_stderr:
            return self._call_batch(fn + "_se", *f, ot, p, st, s, w, with_se=True)
        return self._call_batch(fn, *f, ot, p, st, s, w)

    def mixed_price_batch(self, model, spot, strike, t, vol, r, q, v0, kappa, theta, sigma, rho, option_type,
                          paths, steps, seed):
        m = self._as_i32_array(model, "model")
        f = [self._as_f64_array(x, nm) for x, nm in
             zip((spot, strike, t, vol, r, q, v0, kappa, theta, sigma, rho),
                 ("spot", "strike", "t", "vol", "r", "q", "v0", "kappa", "theta", "sigma", "rho"))]
        ot, p, st, s = (self._as_i32_array(option_type, "option_type"), self._as_i32_array(paths, "paths"),
                        self._as_i32_array(steps, "steps"), self._as_u64_array(seed, "seed"))
        n = m.shape[0]
        self._check_same_length(n, spot=f[0], option_type=ot, paths=p, steps=st, seed=s)
        return self._call_batch("qk_mcm_mixed_price_batch", m, *f, ot, p, st, s, with_se=True)

    def standard_monte_carlo_price_batch_device(self, n: int, ptrs: dict, out_price_ptr: int,
                                                out_se_ptr: int) -> None:
        """Device-resident standard MC: ``ptrs`` maps argument names to device addresses."""
        order = ("spot", "strike", "t", "vol", "r", "q", "option_type", "paths", "seed")
        args = [ct.c_void_p(int(ptrs[k])) for k in order]
        self._check(self._lib.qk_mcm_standard_monte_carlo_price_batch_device(
            *args, int(n), ct.c_void_p(int(out_price_ptr)), ct.c_void_p(int(out_se_ptr))),
            "qk_mcm_standard_monte_carlo_price_batch_device")

    def mixed_price_batch_device(self, n: int, ptrs: dict, out_price_ptr: int, out_se_ptr: int) -> None:
        """Device-resident variant: ``ptrs`` maps argument names to device addresses."""
        order = ("model", "spot", "strike", "t", "vol", "r", "q", "v0", "kappa", "theta", "sigma", "rho",
                 "option_type", "paths", "steps", "seed")
        args = [ct.c_void_p(int(ptrs[k])) for k in order]
        rc = self._lib.qk_mcm_mixed_price_batch_device(*args, int(n), ct.c_void_p(int(out_price_ptr)),
                                                       ct.c_void_p(int(out_se_ptr)))
        self._check(rc, "qk_mcm_mixed_price_batch_device")

    def mixed_price_batch_device_noalloc(self, n: int, ptrs: dict, out_price_ptr: int, out_se_ptr: int) -> None:
        """Allocation-free variant of mixed_price_batch_device (fixed device scratch; at most
        65535 options and 2^19 path tiles)."""
        order = ("model", "spot", "strike", "t", "vol", "r", "q", "v0", "kappa", "theta", "sigma", "rho",
                 "option_type", "paths", "steps", "seed")
        args = [ct.c_void_p(int(ptrs[k])) for k in order]
        rc = self._lib.qk_mcm_mixed_price_batch_device_noalloc(*args, int(n), ct.c_void_p(int(out_price_ptr)),
                                                               ct.c_void_p(int(out_se_ptr)))
        self._check(rc, "qk_mcm_mixed_price_batch_device_noalloc")


