"""The full-frame and serving workloads of the benchmark, against the oracle.

* A 5782x3946 RGBI16 frame with 400 dust disks and 40 scratches up to
  400 px, through ``process_hip`` and through graph-replayed batches whose
  later runs hold a *different* frame of the same geometry: every output
  value, the five replay counters and the hidden-startup receipt equal the
  oracle's (the reference's repeat determinism check is
  tests/test_cuda_level2_tiles.py:80-82).
* A seeded Poisson arrival schedule of 24 scan jobs, 70 % 1024-2048 px crops
  and 30 % full frames, through the serving engine while earlier jobs are
  on the device: each result equals the oracle and the same job alone.
"""

from __future__ import annotations

import hashlib
import threading
import time

import numpy as np
import pytest

import oracle_lib
from conftest import hip_device_or_skip
from dice_amd import synth
from dice_amd.hip_backend import process_hip
from dice_amd.hip_backend.batch import process_batch_hip
from dice_amd.hip_backend.process import stable_profile_params
from dice_amd.hip_backend.serving import HipServingEngine
from dice_amd.profile import DEFAULT_PROFILE
from test_hip_parity import make_job, replay_fingerprint

pytestmark = pytest.mark.gpu

ORACLE_THREADS = 16


@pytest.fixture(scope="module", autouse=True)
def _device():
    hip_device_or_skip()


def oracle_of(main, prepass):
    return oracle_lib.process_job(main, prepass, stable_profile_params(),
                                  DEFAULT_PROFILE.prepass_parameters(), threads=ORACLE_THREADS)


def oracle_fingerprint(r) -> dict:
    return {
        "output_sha256": r.output_sha256,
        "attempted_pixels": r.attempted_pixels,
        "written_pixels": r.written_pixels,
        "public_rng_advances": r.public_rng_advances,
        "final_rng_state": r.final_rng_state,
        "changed_pixels": r.changed_pixels,
        "startup_attempted": list(r.startup_attempted),
        "startup_advances": list(r.startup_advances),
        "startup_final_rng_state": r.startup_final_rng_state,
    }


@pytest.fixture(scope="module")
def full_frames():
    out = []
    for seed in (2, 12):
        main = synth.frame(synth.config2_full_frame(seed=seed))
        prepass = synth.prepass_for(main)
        out.append((main, prepass, oracle_of(main, prepass)))
    return out


def test_full_frame_matches_oracle(full_frames):
    main, prepass, cpu = full_frames[0]
    assert main.shape == (5782, 3946, 4)
    print(f"full frame: {cpu.attempted_pixels} attempted sites", flush=True)
    result = process_hip(make_job(main, prepass, "config2"))
    assert result.output_rgb16.size == 68_447_316
    mismatched = int(np.count_nonzero(result.output_rgb16 != cpu.output))
    assert mismatched == 0
    assert replay_fingerprint(result.replay) == oracle_fingerprint(cpu)


def test_full_frame_graph_replays_with_another_frame_match_oracle(full_frames):
    """Direct run, capturing run, then replays: the second geometry-equal
    frame goes through the captured graphs and must still match."""

    replayed = []
    for attempt, index in enumerate((0, 1, 0, 1)):
        main, prepass, cpu = full_frames[index]
        outcome = process_batch_hip([make_job(main, prepass, f"config2-graph-{attempt}")],
                                    graph=True)
        outcome.raise_first()
        result = outcome.results[0]
        assert np.array_equal(result.output_rgb16, cpu.output), (attempt, index)
        assert replay_fingerprint(result.replay) == oracle_fingerprint(cpu), (attempt, index)
        replayed.append(outcome.report["graph_replayed"])
    assert replayed[0] == 0 and replayed[-1] == 1


def test_serving_workload_matches_oracle():
    specs = synth.config5_serving(count=24, seed=55)
    full = sum(1 for s in specs if (s.height, s.width) == synth.FULL_FRAME)
    assert 4 <= full <= 12, full
    frames = []
    for spec in specs:
        main = synth.frame(spec)
        frames.append((main, synth.prepass_for(main)))
    expected = [oracle_fingerprint(oracle_of(m, p)) for m, p in frames]
    alone = [replay_fingerprint(process_hip(make_job(m, p, f"alone-{i}")).replay)
             for i, (m, p) in enumerate(frames)]
    assert alone == expected
    rng = np.random.default_rng(555)
    gaps = rng.exponential(0.02, size=len(frames))  # Poisson arrivals, 50 jobs/s
    tickets = [None] * len(frames)
    with HipServingEngine(drivers=2) as engine:
        def client():
            for i, (main, prepass) in enumerate(frames):
                time.sleep(gaps[i])
                tickets[i] = engine.submit(make_job(main, prepass, f"served-{i}"))

        feeder = threading.Thread(target=client)
        feeder.start()
        feeder.join()
        for i, ticket in enumerate(tickets):
            result = engine.wait(ticket, timeout=300)
            assert replay_fingerprint(result.replay) == expected[i], f"job {i} {specs[i]}"
            digest = hashlib.sha256(result.output_rgb16.astype("<u2").tobytes()).hexdigest()
            assert digest == expected[i]["output_sha256"], f"job {i} output array"
        stats = engine.stats()
    assert stats["completed"] == len(frames) and stats["failed"] == 0
    assert stats["latency_p99_ms"] >= stats["latency_p50_ms"] > 0
