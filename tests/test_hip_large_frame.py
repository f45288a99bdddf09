"""Large-format frames: 8818 x 13228 RGBI16 (56 x 84 mm at 4000 dpi).

The frame carries 1500 dust disks up to 30 px in radius (wide blobs whose
score reaches the floor) and 120 scratches.  Its output must match the oracle
bit for bit, alone and when two such frames go through the serving engine
together.  There, each frame's serial writer walks on the host while the
other's producers and analysis run on the device.
"""

from __future__ import annotations

import numpy as np
import pytest

import oracle_lib
from conftest import hip_device_or_skip
from dice_amd import synth
from dice_amd.hip_backend import process_hip
from dice_amd.hip_backend.process import stable_profile_params
from dice_amd.hip_backend.serving import HipServingEngine
from dice_amd.profile import DEFAULT_PROFILE
from test_hip_parity import make_job

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _device():
    hip_device_or_skip()


@pytest.fixture(scope="module")
def large():
    params = stable_profile_params()
    prepass_params = DEFAULT_PROFILE.prepass_parameters()
    frames = []
    for seed in (4, 5):
        main = synth.frame(synth.config4_large(seed=seed))
        prepass = synth.prepass_for(main)
        cpu = oracle_lib.process_job(main, prepass, params, prepass_params, threads=16)
        frames.append((main, prepass, cpu))
    return frames


def fingerprint(replay) -> tuple:
    return (replay.output_sha256, replay.attempted_pixels, replay.written_pixels,
            replay.public_rng_advances, replay.final_rng_state, replay.changed_pixels)


def oracle_fingerprint(cpu) -> tuple:
    return (cpu.output_sha256, cpu.attempted_pixels, cpu.written_pixels,
            cpu.public_rng_advances, cpu.final_rng_state, cpu.changed_pixels)


def test_large_frame_matches_oracle(large):
    main, prepass, cpu = large[0]
    assert main.shape == (8818, 13228, 4)
    result = process_hip(make_job(main, prepass, "large"))
    assert fingerprint(result.replay) == oracle_fingerprint(cpu)
    assert np.array_equal(result.output_rgb16, cpu.output)
    assert result.replay.startup.final_rng_state == cpu.startup_final_rng_state


def test_two_large_frames_overlapped_in_the_serving_engine(large):
    with HipServingEngine(drivers=2, max_batch=1) as engine:
        tickets = [engine.submit(make_job(main, prepass, f"large-{k}"))
                   for k, (main, prepass, _) in enumerate(large)]
        results = [engine.wait(t) for t in tickets]
    for result, (_, _, cpu) in zip(results, large):
        assert fingerprint(result.replay) == oracle_fingerprint(cpu)
        assert np.array_equal(result.output_rgb16, cpu.output)
