"""Level 2 + producer parity: the HIP path against the reference vectors.

Every job case of tests/cases.py (the reference's self-test job, the ten
adversarial tiles with fixed seeds, minimal/ragged geometries, dense dust,
threshold extremes) runs through ``process_hip`` with diagnostics and must
reproduce the reference CPU path's output bytes, replay counters, startup
receipt and diagnostics planes (tests/test_cuda_level2_tiles.py:66-166 of
the reference); the streaming-level cases (CENTER_ONLY, writer gates,
alpha==1, visible channel 2, configured strengths) run through
``run_streaming_replay_hip``; the producer geometries through
``derive_producer_record_schedule_hip`` (tests/test_cuda_producer_parity.py).
Repeat runs must be identical (test_cuda_level2_tiles.py:80-82).
"""

from __future__ import annotations

import hashlib

import numpy as np
import pytest

import cases
from conftest import hip_device_or_skip
from dice_amd import AcquisitionEpoch, DualRGBIAcquisition, RGBI16Frame
from dice_amd.engine import LCG24
from dice_amd.hip_backend import (
    derive_producer_record_schedule_hip,
    process_hip,
    run_streaming_replay_hip,
)
from dice_amd.parameters import (
    AuxiliaryParameters,
    DecisionParameters,
    DitherBounds,
    ReconstructionParameters,
    ScoreParameters,
    generate_nikon_response_lut,
)
from dice_amd.profile import ProcessingJob, ProcessingMode, ScannerModel

pytestmark = pytest.mark.gpu


def sha(array: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(array).tobytes()).hexdigest()


def make_job(main: np.ndarray, prepass: np.ndarray, frame_id: str = "parity") -> ProcessingJob:
    return ProcessingJob(
        acquisition=DualRGBIAcquisition(
            prepass=RGBI16Frame(prepass, AcquisitionEpoch.PREPASS, 285, f"{frame_id}-pp"),
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


def replay_fingerprint(replay) -> dict:
    return {
        "output_sha256": replay.output_sha256,
        "attempted_pixels": replay.attempted_pixels,
        "written_pixels": replay.written_pixels,
        "public_rng_advances": replay.public_rng_advances,
        "final_rng_state": replay.final_rng_state,
        "changed_pixels": replay.changed_pixels,
        "startup_attempted": list(replay.startup.attempted_per_stage),
        "startup_advances": list(replay.startup.rng_advances_per_stage),
        "startup_final_rng_state": replay.startup.final_rng_state,
    }


@pytest.fixture(scope="module", autouse=True)
def _device():
    hip_device_or_skip()


@pytest.mark.parametrize("name", list(cases.job_cases()))
def test_job_matches_reference(name, reference_vectors):
    expected = reference_vectors["jobs"][name]
    main, prepass = cases.job_cases()[name]
    assert sha(main.astype("<u2")) == expected["main_sha256"]
    result = process_hip(make_job(main, prepass, name), export_diagnostics=True)
    got = replay_fingerprint(result.replay)
    assert got == {k: expected[k] for k in got}, name
    assert sha(result.output_rgb16.astype("<u2")) == expected["output_samples_sha256"]
    d = result.diagnostics
    assert sha(d.score_plane.astype("<f4")) == expected["score_plane_sha256"]
    assert int(np.asarray(d.score_floor).view(np.uint32)) == expected["score_floor_bits"]
    assert sha(d.at_floor_mask) == expected["at_floor_sha256"]
    assert sha(d.changed_mask) == expected["changed_sha256"]
    repeat = process_hip(make_job(main, prepass, name))
    assert repeat.replay.output_sha256 == result.replay.output_sha256
    assert repeat.replay.final_rng_state == result.replay.final_rng_state


def test_synthetic_lama_golden():
    main, prepass = cases.lama_frames()
    result = process_hip(make_job(main, prepass), export_diagnostics=True)
    assert result.replay.output_sha256 == (
        "780257c669f8df3f3f47768dbd99ffe61354de5f3b66e1b2c8bdedd0cc0e4ad8")
    d = result.diagnostics
    assert sha(d.score_plane.astype("<f4")) == (
        "566c3b0648d5c91566f2a4db3834d39fa9be934ab596ef4c9d188a9c8a3d9baa")
    assert sha(d.at_floor_mask) == (
        "ac7c5f191ce8f1150114da25edaa0a14ca1ba1c67a679bf980900fa56148af94")
    assert sha(d.changed_mask) == (
        "b652da2639bc8ee709d3cd026598f9acc085bc6ad464671507f8eecc150b40c5")


@pytest.mark.parametrize("name", list(cases.stream_cases()))
def test_stream_matches_reference(name, reference_vectors):
    expected = reference_vectors["streams"][name]
    pixels, spec, state = cases.stream_cases()[name]
    output = np.empty((*pixels.shape[:2], 3), np.uint16)
    generator = LCG24(state)
    replay = run_streaming_replay_hip(
        pixels,
        output,
        auxiliary_parameters=AuxiliaryParameters(**spec["auxiliary"]),
        score_parameters=ScoreParameters(**spec["score"]),
        decision_parameters=DecisionParameters(**spec["decision"]),
        reconstruction_parameters=ReconstructionParameters(**spec["reconstruction"]),
        dither_bounds=DitherBounds.from_lookup(generate_nikon_response_lut(), maximum_index=65535),
        generator=generator,
    )
    got = replay_fingerprint(replay)
    assert got == {k: expected[k] for k in got}, name
    assert sha(output.astype("<u2")) == expected["output_samples_sha256"]
    assert generator.state == expected["final_rng_state"]


@pytest.mark.parametrize("shape", cases.PRODUCER_CASES)
def test_producer_matches_reference(shape, reference_vectors):
    height, width = shape
    expected = reference_vectors["producer"][f"{height}x{width}"]
    pixels = cases.producer_frames(height, width)
    s = derive_producer_record_schedule_hip(pixels)
    m = s.mean_schedule
    assert m.active_width == expected["active_width"]
    got = {
        "weighted_visible_sum": sha(m.weighted_visible_sum.astype("<f8")),
        "weighted_infrared_sum": sha(m.weighted_infrared_sum.astype("<f8")),
        "weight_sum": sha(m.weight_sum.astype("<f8")),
        "mean_visible": sha(m.calibration_mean_visible.astype("<f4")),
        "mean_infrared": sha(m.calibration_mean_infrared.astype("<f4")),
        "accepted": sha(m.accepted_samples.astype("<u4")),
        "scaled_input": sha(s.scaled_input.astype("<f4")),
        "input_scale": sha(s.input_scale.astype("<f4")),
        "scale_denominator": sha(s.scale_denominator.astype("<f4")),
        "scale_numerator": sha(s.scale_numerator.astype("<f4")),
        "scale_add_denominator": sha(s.scale_add_denominator.astype("<f4")),
        "scale_add_numerator": sha(s.scale_add_numerator.astype("<f4")),
        "payloads": hashlib.sha256(b"".join(s.record_payloads)).hexdigest(),
    }
    assert got == {k: expected[k] for k in got}
