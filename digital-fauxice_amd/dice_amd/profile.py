"""Fail-closed LS-5000 selector-8 Normal profile.

The reference's profile object (``src/portable_digital_ice/profile.py:19-176``
of rohanpandula/digital-fauxice) is an interface its callers bind to: the
class and method names, the job fields, the order of the job checks and their
exact messages (its contract tests match on them, ``tests/test_contract.py``).
Those are kept.  Inside, the profile here is data:

* every single-precision constant it hands the replay is one entry of a
  bit-pattern table (the replay's decisions depend on the exact bits, so
  none is written as a decimal literal), and each parameter object is
  assembled from its table;
* the job check is an ordered list of rules, each a test and the exception
  it raises; the first rule a job breaks is raised.

The record-derived factories take the prepass 0x40 record; on the supported
path the per-row provider replaces every record-derived field, so the
record's observable effect is the fail-closed check on its alpha.
"""

from __future__ import annotations

import struct
from dataclasses import dataclass
from enum import Enum
from typing import Callable, Iterator

import numpy as np
import numpy.typing as npt

from ._abi import DicePrepassParams
from .contracts import AcquisitionEpoch, DualRGBIAcquisition, RGBI16Frame
from .parameters import (
    AuxiliaryParameters,
    DecisionParameters,
    DitherBounds,
    ReconstructionParameters,
    ScoreParameters,
)


class ScannerModel(str, Enum):
    NIKON_SUPER_COOLSCAN_5000_ED = "nikon-super-coolscan-5000-ed"


class ProcessingMode(str, Enum):
    NORMAL = "normal"


@dataclass(frozen=True)
class ProcessingJob:
    """profile.py:31-41"""

    acquisition: DualRGBIAcquisition
    scanner_model: ScannerModel
    mode: ProcessingMode
    selector: int
    resolution_metric: int
    bit_depth: int
    focus_exposure_locked: bool


@dataclass(frozen=True)
class PrepassRecord:
    """The 0x40-byte prepass frame record fields the profile reads
    (contracts.py:179-221)."""

    payload: bytes

    def _f32(self, offset: int) -> np.float32:
        return np.float32(struct.unpack_from("<f", self.payload, offset)[0])

    @property
    def observed_main_alpha(self) -> np.float32:
        return self._f32(0x04)

    @property
    def observed_calibration_mean_r(self) -> np.float32:
        return self._f32(0x28)

    @property
    def observed_calibration_mean_ir(self) -> np.float32:
        return self._f32(0x2C)


# ---------------------------------------------------------------------------
# Constants, as IEEE-754 single bit patterns (profile.py:116-176)
# ---------------------------------------------------------------------------

_SCORE_BITS = {"base_addend": 0xC2EED000, "scale": 0xBA887952,
               "offset": 0x3F800000, "floor": 0x3CA3D70A}
_SAMPLE_THRESHOLD_BITS = 0x4740E73C
# one value for all three channels
_COARSE_SLOPE_BITS = 0x3F8CCCCD
_BAND_SCALE_BITS = 0x3FA00000
# [table a|b][channel][band]
_FACTOR_BITS = np.array(
    [[[0x3F9AE148, 0x3F9D70A4, 0x3F90A3D7],
      [0x3F95C28F, 0x3F91EB85, 0x3F8A3D71],
      [0x3F851EB8, 0x3F6E147B, 0x3F7851EC]],
     [[0x3F8B851F, 0x3F90A3D7, 0x3F851EB8],
      [0x3F8A3D71, 0x3F866666, 0x3F828F5C],
      [0x3F75C28F, 0x3F570A3D, 0x3F63D70A]]], dtype="<u4")
_DITHER_SCALE_BITS = np.array([0x3C75C28F, 0x3C75C28F, 0x3CCCCCCD], dtype="<u4")

_HORIZONTAL_MINIMUM_RESOLUTION_CUTOFF = 550
_CROSS_NEIGHBOR_CUTOFF = 1_600
_COUNT_LIMIT = 8
_PERPENDICULAR_RADIUS = 4

# prepass.py:42-53 (PrepassParameters.captured_ls5000_selector8): these are
# exact singles, so the decimal literal is the value.
_PREPASS = {"valid_ir_threshold": 8847.2255859375, "alpha_scale": 1.5,
            "ratio_minimum": -0.2, "ratio_maximum": 0.2,
            "zero_denominator_alpha": 0.17}
_PREPASS_COMPLETION_MARKER = 16


def _single(bits) -> float | tuple:
    """The single(s) with these bit patterns, as Python floats (nested
    tuples for an array)."""

    values = np.asarray(bits, dtype="<u4").view("<f4")
    if values.ndim == 0:
        return float(values)
    return _as_tuples(values.tolist())


def _as_tuples(value):
    return tuple(_as_tuples(v) for v in value) if isinstance(value, list) else value


def _base_primary(record: PrepassRecord) -> np.float32:
    """profile.py:48-55: the prepass IR calibration mean with alpha times the
    red mean taken out, over 1 - alpha.  The record's fields are singles, the
    arithmetic is double, the result is rounded to single; alpha == 1 has no
    base and fails closed."""

    alpha, mean_r, mean_ir = (float(v) for v in (record.observed_main_alpha,
                                                  record.observed_calibration_mean_r,
                                                  record.observed_calibration_mean_ir))
    remaining = 1.0 - alpha
    if remaining == 0.0:
        raise ValueError("prepass alpha cannot equal one")
    return np.float32((mean_ir - alpha * mean_r) / remaining)


# ---------------------------------------------------------------------------
# The job check: ordered rules (profile.py:75-109)
# ---------------------------------------------------------------------------

# Job fields that must equal the profile's own value, with their names in
# the message, in the order they are checked.
_FIXED_FIELDS = (("scanner_model", "scanner model"), ("mode", "mode"),
                 ("selector", "selector"), ("resolution_metric", "resolution metric"),
                 ("bit_depth", "bit depth"))


def _broken_rules(profile: "LS5000Selector8NormalProfile", job) -> Iterator[Exception]:
    """The exception of each rule ``job`` breaks, lazily and in check order;
    a rule runs only when every earlier one held, so later rules may rely on
    the types the earlier ones established."""

    if not isinstance(job, ProcessingJob):
        yield TypeError("processing requires a ProcessingJob")
    for field, label in _FIXED_FIELDS:
        actual, supported = getattr(job, field), getattr(profile, field)
        if actual != supported:
            yield ValueError(f"unsupported {label} {actual!r}; this profile requires {supported!r}")
    if job.focus_exposure_locked is not True:
        yield ValueError("main capture must retain prepass focus and exposure")
    acquisition = job.acquisition
    if not isinstance(acquisition, DualRGBIAcquisition):
        yield TypeError("processing requires a dual RGBI acquisition")
    captures: tuple[tuple[RGBI16Frame, int, str], ...] = (
        (acquisition.prepass, profile.prepass_dpi, "prepass"),
        (acquisition.main, profile.main_dpi, "main frame"))
    for frame, dpi, what in captures:
        if frame.resolution_dpi != dpi:
            yield ValueError(f"{what} must be captured at {dpi} dpi")
    # A main frame of at least 8 columns always holds one whole 8-column
    # producer cell, so this rule also covers the reference's cell check.
    rows, columns = acquisition.main.pixels.shape[:2]
    if rows < 5 or columns < 8:
        yield ValueError("main RGBI frame must be at least 5 rows by 8 columns")


@dataclass(frozen=True, init=False)
class LS5000Selector8NormalProfile:
    """profile.py:58-176: the one scanner, mode and resolution supported."""

    scanner_model: ScannerModel = ScannerModel.NIKON_SUPER_COOLSCAN_5000_ED
    mode: ProcessingMode = ProcessingMode.NORMAL
    selector: int = 8
    resolution_metric: int = 4_000
    prepass_dpi: int = 285
    main_dpi: int = 4_000
    bit_depth: int = 16
    auxiliary_factor_b: float = 1.0

    @property
    def profile_id(self) -> str:
        return "nikon-ls5000-selector8-normal-metric4000"

    def validate_job(self, job: ProcessingJob) -> None:
        """Raises the first rule ``job`` breaks; returns if it breaks none."""

        first = next(_broken_rules(self, job), None)
        if first is not None:
            raise first

    def prepass_parameters(self) -> DicePrepassParams:
        params = DicePrepassParams()
        for name, value in _PREPASS.items():
            setattr(params, name, np.float32(value))
        params.completion_marker = _PREPASS_COMPLETION_MARKER
        return params

    def auxiliary_parameters(self, record: PrepassRecord) -> AuxiliaryParameters:
        """Channel 0 carries the IR crosstalk; alpha is the record's, scaled
        in single precision by the profile's factor b."""

        scaled_alpha = record.observed_main_alpha * np.float32(self.auxiliary_factor_b)
        return AuxiliaryParameters(
            selected_visible_channel=0,
            alpha=float(np.float32(scaled_alpha)),
            calibration_offset=1.0,
            alpha_one_replacement=float(_base_primary(record)),
        )

    def score_parameters(self, record: PrepassRecord) -> ScoreParameters:
        singles = {name: _single(bits) for name, bits in _SCORE_BITS.items()}
        return ScoreParameters(
            base_primary=float(_base_primary(record)),
            resolution_metric=self.resolution_metric,
            horizontal_minimum_resolution_cutoff=_HORIZONTAL_MINIMUM_RESOLUTION_CUTOFF,
            **singles,
        )

    def decision_parameters(self) -> DecisionParameters:
        return DecisionParameters(sample_threshold=_single(_SAMPLE_THRESHOLD_BITS),
                                  count_limit=_COUNT_LIMIT,
                                  perpendicular_radius=_PERPENDICULAR_RADIUS)

    def reconstruction_parameters(self, record: PrepassRecord) -> ReconstructionParameters:
        factors_a, factors_b = _single(_FACTOR_BITS)
        per_channel: Callable[[int], tuple] = lambda bits: (_single(bits),) * 3
        return ReconstructionParameters(
            resolution_metric=self.resolution_metric,
            cross_neighbor_cutoff=_CROSS_NEIGHBOR_CUTOFF,
            coarse_enabled=True,
            coarse_reference=float(_base_primary(record)),
            coarse_slopes=per_channel(_COARSE_SLOPE_BITS),
            band_enabled=(True,) * 3,
            band_scales=per_channel(_BAND_SCALE_BITS),
            factors_a=factors_a,
            factors_b=factors_b,
            configured_strengths=(0.0,) * 3,
            driver_gate_primary=True,
            driver_gate_secondary=True,
            row_reconstruction_gate=0,
            dither_scales=_single(_DITHER_SCALE_BITS),
        )

    def dither_bounds(self, lookup: npt.ArrayLike) -> DitherBounds:
        return DitherBounds.from_lookup(lookup, maximum_index=65_535)


DEFAULT_PROFILE = LS5000Selector8NormalProfile()


def neutral_prepass_record() -> PrepassRecord:
    """A record whose derived parameters are finite; used when the stable
    replay parameters are needed before a prepass has been reduced (the
    provider replaces every record-derived field per row)."""

    payload = bytearray(0x40)
    struct.pack_into("<f", payload, 0x04, 0.17)
    return PrepassRecord(bytes(payload))


def selector8_job(main: npt.NDArray[np.uint16], prepass: npt.NDArray[np.uint16],
                  frame_id: str = "job") -> ProcessingJob:
    """The supported acquisition of one frame: a 4000 dpi main pass and its
    285 dpi prepass (contracts.py:25-74), LS-5000 selector-8 Normal."""

    return ProcessingJob(
        acquisition=DualRGBIAcquisition(
            prepass=RGBI16Frame(prepass, AcquisitionEpoch.PREPASS, 285, f"{frame_id}-prepass"),
            main=RGBI16Frame(main, AcquisitionEpoch.MAIN, 4000, f"{frame_id}-main"),
            same_frame_id=frame_id,
        ),
        scanner_model=ScannerModel.NIKON_SUPER_COOLSCAN_5000_ED,
        mode=ProcessingMode.NORMAL,
        selector=8,
        resolution_metric=4000,
        bit_depth=16,
        focus_exposure_locked=True,
    )
