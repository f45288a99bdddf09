"""The job and frame contracts the HIP drop-in validates before any device work.

Mirrors the reference's tests/test_contract.py:17-83 (every unsupported job
field fails closed with its message, wrong capture resolutions, uint16
HxWx4 frames, a profile that cannot be reconfigured) for dice_amd's
profile and contracts, plus the frame-size and producer-cell branches of
validate_job (profile.py:75-106) and the frozen copies of contracts.py:25-40.
"""

from __future__ import annotations

from dataclasses import replace

import numpy as np
import pytest

import cases
from dice_amd.contracts import AcquisitionEpoch, DualRGBIAcquisition, RGBI16Frame
from dice_amd.profile import DEFAULT_PROFILE, ProcessingJob
from test_hip_parity import make_job


@pytest.fixture
def supported_job() -> ProcessingJob:
    main, prepass = cases.supported_job_frames()
    return make_job(main, prepass, "contract")


@pytest.mark.parametrize(
    ("field", "value", "message"),
    (
        ("scanner_model", "other-scanner", "scanner model"),
        ("mode", "fine", "mode"),
        ("selector", 7, "selector"),
        ("resolution_metric", 500, "resolution metric"),
        ("bit_depth", 8, "bit depth"),
        ("focus_exposure_locked", False, "focus and exposure"),
        ("focus_exposure_locked", 1, "focus and exposure"),
    ),
)
def test_profile_fails_closed_for_unsupported_job_fields(supported_job, field, value, message):
    with pytest.raises(ValueError, match=message):
        DEFAULT_PROFILE.validate_job(replace(supported_job, **{field: value}))


def test_unsupported_value_is_named_in_the_message(supported_job):
    with pytest.raises(ValueError, match=r"unsupported selector 7; this profile requires 8"):
        DEFAULT_PROFILE.validate_job(replace(supported_job, selector=7))


@pytest.mark.parametrize(("epoch_field", "dpi", "message"),
                         (("prepass", 500, "285 dpi"), ("main", 2000, "4000 dpi")))
def test_profile_rejects_wrong_capture_resolutions(supported_job, epoch_field, dpi, message):
    acquisition = supported_job.acquisition
    frame = getattr(acquisition, epoch_field)
    wrong = RGBI16Frame(frame.pixels, frame.epoch, dpi, "wrong-dpi")
    fields = {"prepass": acquisition.prepass, "main": acquisition.main, epoch_field: wrong}
    job = replace(supported_job,
                  acquisition=DualRGBIAcquisition(fields["prepass"], fields["main"],
                                                  acquisition.same_frame_id))
    with pytest.raises(ValueError, match=message):
        DEFAULT_PROFILE.validate_job(job)


@pytest.mark.parametrize(("shape", "message"), (((4, 8), "at least 5 rows by 8 columns"),
                                                ((5, 7), "at least 5 rows by 8 columns")))
def test_profile_rejects_frames_below_the_minimum(supported_job, shape, message):
    main = np.full((*shape, 4), 40000, np.uint16)
    acquisition = supported_job.acquisition
    job = replace(supported_job, acquisition=DualRGBIAcquisition(
        acquisition.prepass, RGBI16Frame(main, AcquisitionEpoch.MAIN, 4000, "small"),
        acquisition.same_frame_id))
    with pytest.raises(ValueError, match=message):
        DEFAULT_PROFILE.validate_job(job)


def test_first_broken_rule_in_check_order_is_raised(supported_job):
    """profile.py:75-109 checks scanner model, mode, selector, resolution
    metric, bit depth, focus lock, prepass dpi, main dpi, then frame size. A
    job that breaks all of them reports them one at a time, in that order,
    as each is mended."""

    acquisition = supported_job.acquisition
    small = np.full((4, 7, 4), 40000, np.uint16)

    def captures(prepass_dpi, main_dpi, main_pixels):
        return DualRGBIAcquisition(
            RGBI16Frame(acquisition.prepass.pixels, AcquisitionEpoch.PREPASS, prepass_dpi, "p"),
            RGBI16Frame(main_pixels, AcquisitionEpoch.MAIN, main_dpi, "m"),
            acquisition.same_frame_id)

    job = replace(supported_job, scanner_model="other-scanner", mode="fine", selector=7,
                  resolution_metric=500, bit_depth=8, focus_exposure_locked=False,
                  acquisition=captures(500, 2000, small))
    mends = (
        ("unsupported scanner model", dict(scanner_model=supported_job.scanner_model)),
        ("unsupported mode", dict(mode=supported_job.mode)),
        ("unsupported selector", dict(selector=8)),
        ("unsupported resolution metric", dict(resolution_metric=4000)),
        ("unsupported bit depth", dict(bit_depth=16)),
        ("focus and exposure", dict(focus_exposure_locked=True)),
        ("prepass must be captured at 285 dpi", dict(acquisition=captures(285, 2000, small))),
        ("main frame must be captured at 4000 dpi", dict(acquisition=captures(285, 4000, small))),
        ("at least 5 rows by 8 columns", dict(acquisition=acquisition)),
    )
    for message, mend in mends:
        with pytest.raises(ValueError, match=message):
            DEFAULT_PROFILE.validate_job(job)
        job = replace(job, **mend)
    DEFAULT_PROFILE.validate_job(job)


def test_profile_requires_a_job_and_a_dual_acquisition(supported_job):
    with pytest.raises(TypeError, match="ProcessingJob"):
        DEFAULT_PROFILE.validate_job("not a job")
    with pytest.raises(TypeError, match="dual RGBI acquisition"):
        DEFAULT_PROFILE.validate_job(replace(supported_job, acquisition=object()))


def test_frame_contract_requires_uint16_rgbi():
    with pytest.raises(TypeError, match="dtype uint16"):
        RGBI16Frame(np.zeros((8, 8, 4), dtype=np.float32), AcquisitionEpoch.MAIN, 4_000,
                    "wrong-dtype")
    with pytest.raises(ValueError, match="HxWx4"):
        RGBI16Frame(np.zeros((8, 8, 3), dtype=np.uint16), AcquisitionEpoch.MAIN, 4_000,
                    "wrong-lanes")
    with pytest.raises(ValueError, match="resolution_dpi must be positive"):
        RGBI16Frame(np.zeros((8, 8, 4), dtype=np.uint16), AcquisitionEpoch.MAIN, 0, "dpi")
    with pytest.raises(ValueError, match="evidence_id must be non-empty"):
        RGBI16Frame(np.zeros((8, 8, 4), dtype=np.uint16), AcquisitionEpoch.MAIN, 4_000, " ")


def test_frames_are_frozen_copies():
    """contracts.py:25-40: a frame keeps its own read-only copy."""

    pixels = np.full((8, 8, 4), 40000, np.uint16)
    frame = RGBI16Frame(pixels, AcquisitionEpoch.MAIN, 4_000, "frozen")
    pixels[0, 0, 0] = 1
    assert frame.pixels[0, 0, 0] == 40000
    assert not frame.pixels.flags.writeable
    with pytest.raises(ValueError):
        frame.pixels[0, 0, 0] = 2


def test_supported_profile_cannot_be_reconfigured():
    profile_type = type(DEFAULT_PROFILE)
    with pytest.raises(TypeError):
        profile_type(selector=7)


def test_supported_job_passes(supported_job):
    DEFAULT_PROFILE.validate_job(supported_job)
    assert DEFAULT_PROFILE.profile_id == "nikon-ls5000-selector8-normal-metric4000"
