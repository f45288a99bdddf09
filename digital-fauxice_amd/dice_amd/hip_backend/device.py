"""Device (HBM) buffers for callers that keep frames resident on the GPU.

Allocated through the same HIP runtime that libdice_hip.so links, so the
pointers can be handed to ``dice_process_job`` / ``dice_process_batch``
with ``DICE_MAIN_ON_DEVICE`` / ``DICE_OUT_ON_DEVICE`` without a second
runtime in the process.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from .engine import HipBackendUnavailable, library

_HIP: list[C.CDLL] = []
_H2D, _D2H = 1, 2


def _hip() -> C.CDLL:
    if not _HIP:
        library()  # loads libamdhip64 as a dependency of libdice_hip.so
        try:
            hip = C.CDLL("libamdhip64.so.7")
        except OSError as error:
            raise HipBackendUnavailable(f"HIP runtime not loadable: {error}") from error
        hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        hip.hipFree.argtypes = [C.c_void_p]
        hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        hip.hipDeviceSynchronize.argtypes = []
        hip.hipGetErrorString.restype = C.c_char_p
        hip.hipGetErrorString.argtypes = [C.c_int]
        _HIP.append(hip)
    return _HIP[0]


def _check(code: int, what: str) -> None:
    if code != 0:
        message = _hip().hipGetErrorString(code).decode(errors="replace")
        raise HipBackendUnavailable(f"{what} failed: {message}")


class DeviceArray:
    """A device allocation with host upload/download (C-contiguous arrays)."""

    def __init__(self, nbytes: int) -> None:
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        _check(_hip().hipMalloc(C.byref(self.ptr), max(self.nbytes, 1)), "hipMalloc")

    @classmethod
    def from_host(cls, array: np.ndarray) -> "DeviceArray":
        array = np.ascontiguousarray(array)
        out = cls(array.nbytes)
        out.upload(array)
        return out

    @property
    def address(self) -> int:
        return int(self.ptr.value or 0)

    def upload(self, array: np.ndarray) -> None:
        array = np.ascontiguousarray(array)
        if array.nbytes > self.nbytes:
            raise ValueError("upload larger than the allocation")
        _check(_hip().hipMemcpy(self.ptr, array.ctypes.data, array.nbytes, _H2D), "hipMemcpy")

    def download(self, array: np.ndarray) -> np.ndarray:
        if not array.flags.c_contiguous or array.nbytes > self.nbytes:
            raise ValueError("download target must be C-contiguous and fit the allocation")
        _check(_hip().hipMemcpy(array.ctypes.data, self.ptr, array.nbytes, _D2H), "hipMemcpy")
        return array

    def free(self) -> None:
        if self.ptr.value:
            _check(_hip().hipFree(self.ptr), "hipFree")
            self.ptr = C.c_void_p()

    def __del__(self) -> None:
        try:
            self.free()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass


def synchronize() -> None:
    _check(_hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
