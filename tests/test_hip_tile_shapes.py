"""Every analysis tile shape and site-kernel workgroup variant, against the oracle.

The tuner times these configurations and keeps one only when its results
digest equal to the default's.  Here each is forced through a tuning-cache
entry for the batch's bucket, on frames whose sizes leave partial tiles in
both directions and whose tile rows run into the strips the scan and site
kernels deal (four vertically adjacent tiles, the last quarter of the rows
one tile at a time), and every job must equal the oracle: output hash,
counters and the hidden-startup receipt.
"""

from __future__ import annotations

import math

import pytest

import oracle_lib
from conftest import hip_device_or_skip
from dice_amd import synth
from dice_amd.hip_backend import engine
from dice_amd.hip_backend.batch import process_batch_hip
from dice_amd.hip_backend.process import stable_profile_params
from dice_amd.profile import DEFAULT_PROFILE
from test_hip_parity import make_job, replay_fingerprint

pytestmark = pytest.mark.gpu

TILE_SHAPES = [(16, 64), (8, 128), (32, 32), (16, 32), (32, 64), (8, 64)]
# 37 and 11 rows of 16-row tiles: whole strips, singles, partial last tiles
FRAMES = [(590, 517), (181, 1000)]


@pytest.fixture(scope="module", autouse=True)
def _device():
    hip_device_or_skip()


@pytest.fixture(scope="module")
def batch():
    frames, expected = [], []
    params = stable_profile_params()
    for i, (h, w) in enumerate(FRAMES):
        main = synth.frame(synth.scaled_spec(h, w, seed=950 + i))
        prepass = synth.prepass_for(main)
        cpu = oracle_lib.process_job(main, prepass, params, DEFAULT_PROFILE.prepass_parameters())
        frames.append((main, prepass))
        expected.append({
            "output_sha256": cpu.output_sha256,
            "attempted_pixels": cpu.attempted_pixels,
            "written_pixels": cpu.written_pixels,
            "public_rng_advances": cpu.public_rng_advances,
            "final_rng_state": cpu.final_rng_state,
            "changed_pixels": cpu.changed_pixels,
            "startup_attempted": list(cpu.startup_attempted),
            "startup_advances": list(cpu.startup_advances),
            "startup_final_rng_state": cpu.startup_final_rng_state,
        })
    return frames, expected


def bucket(frames) -> str:
    """The tuner's bucket: job count, total and largest pixels, each rounded
    down to a power of two."""

    pixels = [m.shape[0] * m.shape[1] for m, _ in frames]
    live = len(pixels)
    return (f"n{1 << int(math.log2(2 * live - 1))}-t{int(math.log2(sum(pixels)))}"
            f"-m{int(math.log2(max(pixels)))}")


@pytest.mark.parametrize("variant", (0, 1))
@pytest.mark.parametrize(("rows", "cols"), TILE_SHAPES)
def test_forced_configuration_matches_oracle(batch, tmp_path, rows, cols, variant):
    frames, expected = batch
    lib = engine.library()
    cache = tmp_path / "tuning.tsv"
    cache.write_text(f"gfx950\t{bucket(frames)}\t{rows}\t{cols}\t{variant}\t1.0\n")
    assert lib.dice_tuning_cache(str(cache).encode()) == 0
    try:
        jobs = [make_job(m, p, f"shape-{rows}x{cols}-{variant}-{i}")
                for i, (m, p) in enumerate(frames)]
        outcome = process_batch_hip(jobs, autotune=True)
        outcome.raise_first()
        report = outcome.report
        assert report["tune_source"] == 3, report
        assert (report["tile_rows"], report["tile_cols"], report["sites_variant"]) == (
            rows, cols, variant)
        for i, result in enumerate(outcome.results):
            assert replay_fingerprint(result.replay) == expected[i], (rows, cols, variant, i)
    finally:
        lib.dice_tuning_cache(None)
