"""Device memory through the HIP runtime the library itself links
(libamdhip64.so.7), for callers of the device-resident entry points that do
not bring their own: no second runtime (such as the one torch bundles under
the same soname) is loaded into the process."""
from __future__ import annotations

import ctypes as ct

import numpy as np

_H2D, _D2H, _D2D = 1, 2, 3
_hip = None


def hip() -> ct.CDLL:
    global _hip
    if _hip is None:
        _hip = ct.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [ct.POINTER(ct.c_void_p), ct.c_size_t]
        _hip.hipFree.argtypes = [ct.c_void_p]
        _hip.hipMemcpy.argtypes = [ct.c_void_p, ct.c_void_p, ct.c_size_t, ct.c_int]
        _hip.hipMemset.argtypes = [ct.c_void_p, ct.c_int, ct.c_size_t]
        for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipMemset", "hipDeviceSynchronize"):
            getattr(_hip, f).restype = ct.c_int
    return _hip


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed: hipError {rc}")


def synchronize() -> None:
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DevArray:
    """A device copy of a 1-D numpy array (or an uninitialised one)."""

    def __init__(self, arr=None, *, n=None, dtype=np.float64):
        h = hip()
        if arr is not None:
            arr = np.ascontiguousarray(arr)
            self.dtype, self.n = arr.dtype, arr.shape[0]
        else:
            self.dtype, self.n = np.dtype(dtype), n
        self.nbytes = max(1, self.n * self.dtype.itemsize)
        self._p = ct.c_void_p()
        _ok(h.hipMalloc(ct.byref(self._p), self.nbytes), "hipMalloc")
        if arr is not None:
            _ok(h.hipMemcpy(self._p, arr.ctypes.data, self.nbytes, _H2D), "hipMemcpy H2D")

    @property
    def ptr(self) -> int:
        return self._p.value

    def numpy(self) -> np.ndarray:
        out = np.empty(self.n, dtype=self.dtype)
        _ok(hip().hipMemcpy(out.ctypes.data, self._p, self.nbytes, _D2H), "hipMemcpy D2H")
        return out

    def fill_bytes(self, byte: int) -> None:
        h = hip()
        _ok(h.hipMemset(self._p, int(byte), self.nbytes), "hipMemset")
        _ok(h.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def copy_from(self, other: "DevArray") -> None:
        """Device-to-device copy of other's bytes (same size)."""
        if other.nbytes != self.nbytes:
            raise ValueError("size mismatch")
        _ok(hip().hipMemcpy(self._p, other._p, self.nbytes, _D2D), "hipMemcpy D2D")

    def __del__(self):
        if getattr(self, "_p", None) is not None and self._p.value and _hip is not None:
            _hip.hipFree(self._p)
            self._p = ct.c_void_p()


def to_device(d: dict, keys=None) -> dict:
    """Device copies of the arrays of d (all of them, or those named)."""
    return {k: DevArray(d[k]) for k in (keys if keys is not None else d)}
