"""Fail-closed LS-5000 selector-8 Normal profile (mirror of ``profile.py``).

Same job contract, constants and messages as
``src/portable_digital_ice/profile.py:19-176`` of rohanpandula/digital-fauxice.
The record-derived parameter factories take the prepass 0x40 record; on the
supported path the per-row provider replaces every record-derived field, so
the record's observable effect is the fail-closed validation done here.
"""

from __future__ import annotations

import struct
from dataclasses import dataclass
from enum import Enum

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
    f32_from_bits,
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


def _base_primary(record: PrepassRecord) -> np.float32:
    """profile.py:48-55"""

    alpha = np.float32(record.observed_main_alpha)
    mean_r = np.float32(record.observed_calibration_mean_r)
    mean_ir = np.float32(record.observed_calibration_mean_ir)
    denominator = 1.0 - float(alpha)
    if denominator == 0.0:
        raise ValueError("prepass alpha cannot equal one")
    return np.float32((float(mean_ir) - float(alpha) * float(mean_r)) / denominator)


@dataclass(frozen=True, init=False)
class LS5000Selector8NormalProfile:
    """profile.py:58-176"""

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
        if not isinstance(job, ProcessingJob):
            raise TypeError("processing requires a ProcessingJob")
        expected = (
            ("scanner model", job.scanner_model, self.scanner_model),
            ("mode", job.mode, self.mode),
            ("selector", job.selector, self.selector),
            ("resolution metric", job.resolution_metric, self.resolution_metric),
            ("bit depth", job.bit_depth, self.bit_depth),
        )
        for label, actual, supported in expected:
            if actual != supported:
                raise ValueError(
                    f"unsupported {label} {actual!r}; this profile requires {supported!r}"
                )
        if job.focus_exposure_locked is not True:
            raise ValueError("main capture must retain prepass focus and exposure")
        acquisition = job.acquisition
        if not isinstance(acquisition, DualRGBIAcquisition):
            raise TypeError("processing requires a dual RGBI acquisition")
        if acquisition.prepass.resolution_dpi != self.prepass_dpi:
            raise ValueError(f"prepass must be captured at {self.prepass_dpi} dpi")
        if acquisition.main.resolution_dpi != self.main_dpi:
            raise ValueError(f"main frame must be captured at {self.main_dpi} dpi")
        main = acquisition.main.pixels
        if main.shape[0] < 5 or main.shape[1] < 8:
            raise ValueError("main RGBI frame must be at least 5 rows by 8 columns")
        if (main.shape[1] // 8) * 8 == 0:
            raise ValueError("main RGBI frame has no complete producer cell")

    def prepass_parameters(self) -> DicePrepassParams:
        """prepass.py:42-53 (captured_ls5000_selector8)."""

        params = DicePrepassParams()
        params.valid_ir_threshold = np.float32(8847.2255859375)
        params.alpha_scale = np.float32(1.5)
        params.ratio_minimum = np.float32(-0.2)
        params.ratio_maximum = np.float32(0.2)
        params.zero_denominator_alpha = np.float32(0.17)
        params.completion_marker = 16
        return params

    def auxiliary_parameters(self, record: PrepassRecord) -> AuxiliaryParameters:
        base = _base_primary(record)
        alpha = np.float32(np.float32(record.observed_main_alpha) * np.float32(self.auxiliary_factor_b))
        return AuxiliaryParameters(
            selected_visible_channel=0,
            alpha=float(alpha),
            calibration_offset=float(np.float32(1.0)),
            alpha_one_replacement=float(base),
        )

    def score_parameters(self, record: PrepassRecord) -> ScoreParameters:
        return ScoreParameters(
            base_primary=float(_base_primary(record)),
            base_addend=f32_from_bits(0xC2EED000),
            scale=f32_from_bits(0xBA887952),
            offset=f32_from_bits(0x3F800000),
            floor=f32_from_bits(0x3CA3D70A),
            resolution_metric=self.resolution_metric,
            horizontal_minimum_resolution_cutoff=550,
        )

    def decision_parameters(self) -> DecisionParameters:
        return DecisionParameters(
            sample_threshold=f32_from_bits(0x4740E73C),
            count_limit=8,
            perpendicular_radius=4,
        )

    def reconstruction_parameters(self, record: PrepassRecord) -> ReconstructionParameters:
        f = f32_from_bits
        return ReconstructionParameters(
            resolution_metric=self.resolution_metric,
            cross_neighbor_cutoff=1_600,
            coarse_enabled=True,
            coarse_reference=float(_base_primary(record)),
            coarse_slopes=(f(0x3F8CCCCD), f(0x3F8CCCCD), f(0x3F8CCCCD)),
            band_enabled=(True, True, True),
            band_scales=(f(0x3FA00000), f(0x3FA00000), f(0x3FA00000)),
            factors_a=(
                (f(0x3F9AE148), f(0x3F9D70A4), f(0x3F90A3D7)),
                (f(0x3F95C28F), f(0x3F91EB85), f(0x3F8A3D71)),
                (f(0x3F851EB8), f(0x3F6E147B), f(0x3F7851EC)),
            ),
            factors_b=(
                (f(0x3F8B851F), f(0x3F90A3D7), f(0x3F851EB8)),
                (f(0x3F8A3D71), f(0x3F866666), f(0x3F828F5C)),
                (f(0x3F75C28F), f(0x3F570A3D), f(0x3F63D70A)),
            ),
            configured_strengths=(0.0, 0.0, 0.0),
            driver_gate_primary=True,
            driver_gate_secondary=True,
            row_reconstruction_gate=0,
            dither_scales=(f(0x3C75C28F), f(0x3C75C28F), f(0x3CCCCCCD)),
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
