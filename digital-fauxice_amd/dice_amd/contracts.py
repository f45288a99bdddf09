"""Typed acquisition contracts (mirror of the reference's ``contracts.py``).

Same names, argument meaning and error behaviour as
``src/portable_digital_ice/contracts.py:25-74, 162-176`` of
rohanpandula/digital-fauxice: frames are copied, frozen and validated on
construction.
"""

from __future__ import annotations

import hashlib
from dataclasses import dataclass
from enum import Enum

import numpy as np
import numpy.typing as npt

UInt16Array = npt.NDArray[np.uint16]


class AcquisitionEpoch(str, Enum):
    PREPASS = "prepass"
    MAIN = "main"


def _frozen_uint16_image(pixels: npt.ArrayLike, *, lanes: int, label: str) -> UInt16Array:
    """contracts.py:25-40"""

    array = np.asarray(pixels)
    if array.dtype != np.dtype(np.uint16):
        raise TypeError(f"{label} must have dtype uint16")
    if array.ndim != 3 or array.shape[2] != lanes:
        raise ValueError(f"{label} must have shape HxWx{lanes}")
    if array.shape[0] < 1 or array.shape[1] < 1:
        raise ValueError(f"{label} cannot have an empty dimension")
    owned = np.array(array, dtype=np.uint16, order="C", copy=True)
    owned.flags.writeable = False
    return owned


@dataclass(frozen=True)
class RGBI16Frame:
    """contracts.py:44-74"""

    pixels: UInt16Array
    epoch: AcquisitionEpoch
    resolution_dpi: int
    evidence_id: str

    def __post_init__(self) -> None:
        if self.resolution_dpi <= 0:
            raise ValueError("resolution_dpi must be positive")
        if not self.evidence_id.strip():
            raise ValueError("evidence_id must be non-empty")
        object.__setattr__(
            self, "pixels", _frozen_uint16_image(self.pixels, lanes=4, label="RGBI frame")
        )

    @property
    def height(self) -> int:
        return int(self.pixels.shape[0])

    @property
    def width(self) -> int:
        return int(self.pixels.shape[1])

    @property
    def sha256(self) -> str:
        return hashlib.sha256(self.pixels.astype("<u2", copy=False).tobytes()).hexdigest()


@dataclass(frozen=True)
class DualRGBIAcquisition:
    """contracts.py:162-176"""

    prepass: RGBI16Frame
    main: RGBI16Frame
    same_frame_id: str

    def __post_init__(self) -> None:
        if self.prepass.epoch is not AcquisitionEpoch.PREPASS:
            raise ValueError("prepass frame has the wrong acquisition epoch")
        if self.main.epoch is not AcquisitionEpoch.MAIN:
            raise ValueError("main frame has the wrong acquisition epoch")
        if not self.same_frame_id.strip():
            raise ValueError("same_frame_id must be non-empty")
