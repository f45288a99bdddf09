"""HIP producer schedule (mirror of ``cuda_backend/producer.py:38-246``).

``derive_producer_record_schedule_hip`` returns the same fields as the
reference's ``ProducerRecordSchedule`` (producer_parameters.py:52-129):
per-row binary64 cumulative sums, float32 means, accepted counts, the
eight-row scale accumulators and the 0x40-byte payload of every row.
"""

from __future__ import annotations

import ctypes as C
import hashlib
from dataclasses import dataclass

import numpy as np
import numpy.typing as npt

from .._abi import DiceSchedule, ptr
from ..parameters import generate_nikon_response_lut
from .engine import context, library, raise_for


@dataclass(frozen=True)
class ProducerMeanSchedule:
    weighted_visible_sum: npt.NDArray[np.float64]
    weighted_infrared_sum: npt.NDArray[np.float64]
    weight_sum: npt.NDArray[np.float64]
    calibration_mean_visible: npt.NDArray[np.float32]
    calibration_mean_infrared: npt.NDArray[np.float32]
    accepted_samples: npt.NDArray[np.uint32]
    active_width: int
    infrared_threshold: float


@dataclass(frozen=True)
class ProducerRecordSchedule:
    mean_schedule: ProducerMeanSchedule
    scaled_input: npt.NDArray[np.float32]
    input_scale: npt.NDArray[np.float32]
    scale_denominator: npt.NDArray[np.float32]
    scale_numerator: npt.NDArray[np.float32]
    scale_add_denominator: npt.NDArray[np.float32]
    scale_add_numerator: npt.NDArray[np.float32]
    record_payloads: tuple[bytes, ...]
    scale_multiplier: float
    source_rgbi_sha256: str
    response_lut_sha256: str

    @property
    def row_count(self) -> int:
        return len(self.record_payloads)

    def record_payload(self, row: int) -> bytes:
        if not 0 <= row < self.row_count:
            raise IndexError(row)
        return self.record_payloads[row]


def schedule_arrays(rows: int) -> dict[str, np.ndarray]:
    return {
        "weighted_visible_sum": np.zeros(rows, np.float64),
        "weighted_infrared_sum": np.zeros(rows, np.float64),
        "weight_sum": np.zeros(rows, np.float64),
        "mean_visible": np.zeros(rows, np.float32),
        "mean_infrared": np.zeros(rows, np.float32),
        "accepted": np.zeros(rows, np.uint32),
        "scaled_input": np.zeros(rows, np.float32),
        "input_scale": np.zeros(rows, np.float32),
        "scale_denominator": np.zeros(rows, np.float32),
        "scale_numerator": np.zeros(rows, np.float32),
        "scale_add_denominator": np.zeros(rows, np.float32),
        "scale_add_numerator": np.zeros(rows, np.float32),
        "payloads": np.zeros((rows, 64), np.uint8),
    }


def schedule_struct(arrays: dict[str, np.ndarray]) -> DiceSchedule:
    s = DiceSchedule()
    s.rows = arrays["weight_sum"].shape[0]
    for name, _ in DiceSchedule._fields_[2:]:
        array = arrays[name]
        setattr(s, name, ptr(array, np.ctypeslib.as_ctypes_type(array.dtype)))
    return s


def derive_producer_record_schedule_hip(
    rgbi16: npt.ArrayLike,
    response_lut: npt.ArrayLike | None = None,
) -> ProducerRecordSchedule:
    pixels = np.asarray(rgbi16)
    if pixels.dtype != np.dtype(np.uint16):
        raise TypeError("producer RGBI input must have dtype uint16")
    if pixels.ndim != 3 or pixels.shape[2] != 4:
        raise ValueError("producer RGBI input must have shape HxWx4")
    table = generate_nikon_response_lut() if response_lut is None else np.asarray(response_lut)
    if table.dtype != np.dtype(np.float32) or table.shape != (65536,) or not np.all(np.isfinite(table)):
        raise ValueError("producer response LUT must be finite float32[65536]")
    height, width = pixels.shape[:2]
    if height <= 0:
        raise ValueError("producer record schedule cannot be empty")
    if (width // 8) * 8 <= 0:
        raise ValueError("producer active width is outside the RGBI row")
    source = np.ascontiguousarray(pixels, dtype=np.uint16)
    lut = np.ascontiguousarray(table, dtype=np.float32)
    arrays = schedule_arrays(height)
    sched = schedule_struct(arrays)
    code = library().dice_producer_schedule(context(), ptr(source, C.c_uint16), height, width,
                                            ptr(lut, C.c_float), C.byref(sched))
    raise_for(code)
    for array in arrays.values():
        array.flags.writeable = False
    means = ProducerMeanSchedule(
        weighted_visible_sum=arrays["weighted_visible_sum"],
        weighted_infrared_sum=arrays["weighted_infrared_sum"],
        weight_sum=arrays["weight_sum"],
        calibration_mean_visible=arrays["mean_visible"],
        calibration_mean_infrared=arrays["mean_infrared"],
        accepted_samples=arrays["accepted"],
        active_width=int(sched.active_width),
        infrared_threshold=float(np.float32(8847.2255859375)),
    )
    return ProducerRecordSchedule(
        mean_schedule=means,
        scaled_input=arrays["scaled_input"],
        input_scale=arrays["input_scale"],
        scale_denominator=arrays["scale_denominator"],
        scale_numerator=arrays["scale_numerator"],
        scale_add_denominator=arrays["scale_add_denominator"],
        scale_add_numerator=arrays["scale_add_numerator"],
        record_payloads=tuple(bytes(row) for row in arrays["payloads"]),
        scale_multiplier=float(np.float32(1.5)),
        source_rgbi_sha256=hashlib.sha256(source.astype("<u2").tobytes()).hexdigest(),
        response_lut_sha256=hashlib.sha256(lut.astype("<f4").tobytes()).hexdigest(),
    )
