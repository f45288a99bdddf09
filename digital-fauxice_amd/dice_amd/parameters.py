"""Replay parameter types and the two hash-pinned tables.

Mirrors the reference dataclasses that parameterize the streaming path
(rohanpandula/digital-fauxice, ``src/portable_digital_ice/``):
``AuxiliaryParameters`` / ``ScoreParameters`` / ``DecisionParameters``
(x3a.py:166-349), ``ReconstructionParameters`` (reconstruction.py:613-670),
``DitherBounds`` (dither.py:18-56) and the response LUT / inverse factor
generators (x3a.py:111-132, output.py:39-58), with the same validation
messages.  ``to_dice_params`` quantizes every field to the binary32 value
the reference uses at the point of consumption and packs the C ABI struct.
"""

from __future__ import annotations

import hashlib
import math
import struct
from dataclasses import dataclass
from enum import Enum

import numpy as np
import numpy.typing as npt

from ._abi import DiceParams

NIKON_RESPONSE_LUT_SHA256 = "5fd225e25719a544df9475315e3017f9b8027d593ac43e392245de59f49f3fcb"
H_FACTOR_SHA256 = "fdc378493f13916e36d3dbd9b69b714e888d7993680509df4550486db6fcf328"
L_FACTOR_SHA256 = "b7f3b1b5c0ea816ddb20196e1cf7614d364ea4e164295e0dcb78119c88ef404d"
LCG24_MASK = 0x00FF_FFFF
NIKON_PE_INITIAL_STATE = 0x3045
NIKON_NORMALIZATION = float.fromhex("0x1.fffffep-25")


class UnboundBehaviorError(RuntimeError):
    """Raised instead of silently substituting an unproven algorithm value."""


def _finite(value: float, label: str) -> float:
    converted = float(value)
    if not math.isfinite(converted):
        raise ValueError(f"{label} must be finite")
    return converted


def f32_from_bits(bits: int) -> float:
    return struct.unpack("<f", struct.pack("<I", bits))[0]


_LUT_CACHE: list[np.ndarray] = []


def generate_nikon_response_lut() -> npt.NDArray[np.float32]:
    """x3a.py:111-132: f32(65535*log2(i+1)/16), hash-checked."""

    if _LUT_CACHE:
        return _LUT_CACHE[0]
    domain = np.arange(65536, dtype=np.float64)
    table = np.asarray(65535.0 * np.log2(domain + np.float64(1.0)) / 16.0, dtype=np.float32)
    digest = hashlib.sha256(table.astype("<f4", copy=False).tobytes()).hexdigest()
    if digest != NIKON_RESPONSE_LUT_SHA256:
        raise UnboundBehaviorError(
            "platform logarithm does not reproduce Nikon's response LUT bit-exact"
        )
    table.flags.writeable = False
    _LUT_CACHE.append(table)
    return table


def generate_inverse_factors() -> tuple[np.ndarray, np.ndarray]:
    """output.py:39-58"""

    domain = np.arange(256, dtype=np.float64)
    scale = 16.0 * np.log10(np.float64(2.0)) / 65535.0
    high = np.floor(16.0 * np.power(10.0, 256.0 * domain * scale) + 0.5).astype(np.uint32)
    low = np.floor(65536.0 * np.power(10.0, domain * scale) + 0.5).astype(np.uint32)
    hashes = (
        hashlib.sha256(high.astype("<u4").tobytes()).hexdigest(),
        hashlib.sha256(low.astype("<u4").tobytes()).hexdigest(),
    )
    if hashes != (H_FACTOR_SHA256, L_FACTOR_SHA256):
        raise UnboundBehaviorError(
            "platform power function did not reproduce the recovered "
            "inverse-response factor tables"
        )
    return high, low


@dataclass(frozen=True)
class AuxiliaryParameters:
    """x3a.py:166-179"""

    selected_visible_channel: int
    alpha: float
    calibration_offset: float
    alpha_one_replacement: float | None

    def __post_init__(self) -> None:
        if self.selected_visible_channel not in (0, 1, 2):
            raise ValueError("selected_visible_channel must be 0, 1, or 2")
        _finite(self.alpha, "alpha")
        _finite(self.calibration_offset, "calibration_offset")
        if self.alpha_one_replacement is not None:
            _finite(self.alpha_one_replacement, "alpha_one_replacement")


@dataclass(frozen=True)
class ScoreParameters:
    """x3a.py:224-242"""

    base_primary: float
    base_addend: float
    scale: float
    offset: float
    floor: float
    resolution_metric: int
    horizontal_minimum_resolution_cutoff: int

    def __post_init__(self) -> None:
        for label in ("base_primary", "base_addend", "scale", "offset", "floor"):
            _finite(getattr(self, label), label)
        if self.floor > 1.0:
            raise ValueError("score floor cannot exceed the recovered ceiling 1.0")
        if self.resolution_metric < 0:
            raise ValueError("resolution_metric cannot be negative")
        if self.horizontal_minimum_resolution_cutoff < 0:
            raise ValueError("horizontal minimum resolution cutoff cannot be negative")


@dataclass(frozen=True)
class DecisionParameters:
    """x3a.py:337-348"""

    sample_threshold: float
    count_limit: int
    perpendicular_radius: int

    def __post_init__(self) -> None:
        _finite(self.sample_threshold, "sample_threshold")
        if not 0 <= self.count_limit <= 9:
            raise ValueError("count_limit must be in [0, 9]")
        if not 1 <= self.perpendicular_radius <= 4:
            raise ValueError("perpendicular_radius must be in [1, 4]")


class FeatureBandExtremaMode(str, Enum):
    CENTER_ONLY = "center_only"
    CROSS_NEIGHBOR = "cross_neighbor"


def feature_band_extrema_mode(*, resolution_metric: int, cross_neighbor_cutoff: int) -> FeatureBandExtremaMode:
    """reconstruction.py:561-572"""

    if resolution_metric < 0 or cross_neighbor_cutoff < 0:
        raise ValueError("resolution metric and cutoff cannot be negative")
    return (
        FeatureBandExtremaMode.CROSS_NEIGHBOR
        if resolution_metric > cross_neighbor_cutoff
        else FeatureBandExtremaMode.CENTER_ONLY
    )


@dataclass(frozen=True)
class ReconstructionParameters:
    """reconstruction.py:613-670"""

    resolution_metric: int
    cross_neighbor_cutoff: int
    coarse_enabled: bool
    coarse_reference: float
    coarse_slopes: tuple[float, float, float]
    band_enabled: tuple[bool, bool, bool]
    band_scales: tuple[float, float, float]
    factors_a: tuple[tuple[float, float, float], ...]
    factors_b: tuple[tuple[float, float, float], ...]
    configured_strengths: tuple[float, float, float]
    driver_gate_primary: bool
    driver_gate_secondary: bool
    row_reconstruction_gate: int
    dither_scales: tuple[float, float, float]

    def __post_init__(self) -> None:
        if self.resolution_metric < 0 or self.cross_neighbor_cutoff < 0:
            raise ValueError("reconstruction resolution fields cannot be negative")
        if self.row_reconstruction_gate < 0:
            raise ValueError("row reconstruction gate cannot be negative")
        if len(self.coarse_slopes) != 3:
            raise ValueError("coarse correction requires three channel slopes")
        if len(self.band_enabled) != 3 or len(self.band_scales) != 3:
            raise ValueError("reconstruction requires exactly three bands")
        if len(self.factors_a) != 3 or len(self.factors_b) != 3:
            raise ValueError("reconstruction requires three factor rows")
        if any(len(row) != 3 for row in (*self.factors_a, *self.factors_b)):
            raise ValueError("each reconstruction factor row requires R/G/B values")
        if len(self.configured_strengths) != 3 or len(self.dither_scales) != 3:
            raise ValueError("strength and dither tuples require three values")
        numeric = (
            self.coarse_reference,
            *self.coarse_slopes,
            *self.band_scales,
            *(value for row in self.factors_a for value in row),
            *(value for row in self.factors_b for value in row),
            *self.configured_strengths,
            *self.dither_scales,
        )
        if not all(math.isfinite(value) for value in numeric):
            raise ValueError("reconstruction parameters must be finite")


LOW_FRACTION = np.float32(0.01)
HIGH_FRACTION = np.float32(0.99)


def percentile_bound_indices(maximum_index: int) -> tuple[int, int]:
    """dither.py:18-31"""

    if not 0 <= maximum_index <= 0xFFFF:
        raise ValueError("maximum_index must fit the recovered unsigned-16 field")
    n = np.float32(maximum_index)
    low = min(math.trunc(float(np.multiply(n, LOW_FRACTION))), maximum_index)
    high = min(math.trunc(float(np.multiply(n, HIGH_FRACTION))), maximum_index)
    return low, high


@dataclass(frozen=True)
class DitherBounds:
    """dither.py:34-56"""

    low: np.float32
    high: np.float32
    low_index: int
    high_index: int

    @classmethod
    def from_lookup(cls, lookup: npt.ArrayLike, *, maximum_index: int) -> "DitherBounds":
        table = np.asarray(lookup)
        if table.dtype != np.dtype(np.float32) or table.ndim != 1:
            raise TypeError("dither lookup must be a float32 vector")
        if len(table) <= maximum_index or not np.all(np.isfinite(table)):
            raise ValueError("dither lookup does not contain every index through N")
        low_index, high_index = percentile_bound_indices(maximum_index)
        return cls(np.float32(table[low_index]), np.float32(table[high_index]),
                   low_index, high_index)


def to_dice_params(
    *,
    auxiliary_parameters: AuxiliaryParameters,
    score_parameters: ScoreParameters,
    decision_parameters: DecisionParameters,
    reconstruction_parameters: ReconstructionParameters,
    dither_bounds: DitherBounds,
) -> DiceParams:
    """Pack the stable parameters, quantized to binary32 as the reference does."""

    f = np.float32
    rp = reconstruction_parameters
    sp = score_parameters
    params = DiceParams()
    params.visible_channel = int(auxiliary_parameters.selected_visible_channel)
    params.score_addend = f(sp.base_addend)
    params.score_scale = f(sp.scale)
    params.score_offset = f(sp.offset)
    params.score_floor = f(sp.floor)
    params.horizontal_minimum = int(
        sp.horizontal_minimum_resolution_cutoff < sp.resolution_metric
    )
    params.decision_threshold = f(decision_parameters.sample_threshold)
    params.decision_count_limit = int(decision_parameters.count_limit)
    params.decision_radius = int(decision_parameters.perpendicular_radius)
    params.cross_neighbor = int(
        feature_band_extrema_mode(
            resolution_metric=rp.resolution_metric,
            cross_neighbor_cutoff=rp.cross_neighbor_cutoff,
        )
        is FeatureBandExtremaMode.CROSS_NEIGHBOR
    )
    params.coarse_enabled = int(bool(rp.coarse_enabled))
    for i in range(3):
        params.coarse_slopes[i] = f(rp.coarse_slopes[i])
        params.band_enabled[i] = int(bool(rp.band_enabled[i]))
        params.band_scales[i] = f(rp.band_scales[i])
        params.configured_strengths[i] = f(rp.configured_strengths[i])
        params.dither_scales[i] = f(rp.dither_scales[i])
        for c in range(3):
            params.factors_a[i * 3 + c] = f(rp.factors_a[i][c])
            params.factors_b[i * 3 + c] = f(rp.factors_b[i][c])
    params.gate_primary = int(bool(rp.driver_gate_primary))
    params.dither_low = f(dither_bounds.low)
    params.dither_high = f(dither_bounds.high)
    return params
