"""Batched execution: many uneven jobs in one set of persistent launches.

Every job of a batch must reproduce, byte for byte, both the oracle (the C
restatement of the reference CPU path) and the same job run alone through
``process_hip``.  Graph replays are checked replay by replay against the
eager batch, with unrelated eager work and synchronization between the
replays; autotuned tile shapes must not change any result; a job that fails
validation inside a batch fails alone.
"""

from __future__ import annotations

import ctypes as C
import hashlib

import numpy as np
import pytest

import cases
import oracle_lib
from conftest import hip_device_or_skip
from dice_amd import _abi, synth
from dice_amd._abi import DiceBatchReport, DiceJob, DiceReplayResult
from dice_amd.hip_backend import engine, process_hip
from dice_amd.hip_backend.batch import process_batch_hip
from dice_amd.hip_backend.process import stable_profile_params
from dice_amd.profile import DEFAULT_PROFILE
from test_hip_parity import make_job, replay_fingerprint

pytestmark = pytest.mark.gpu

# Uneven geometries: ragged tiles in both directions, a tiny frame, tall and
# wide shapes, at the full-frame defect density.
SHAPES = [(97, 211), (256, 256), (300, 517), (64, 72), (513, 129), (181, 640)]


@pytest.fixture(scope="module", autouse=True)
def _device():
    hip_device_or_skip()


@pytest.fixture(scope="module")
def frames():
    out = []
    for i, (h, w) in enumerate(SHAPES):
        spec = synth.scaled_spec(h, w, seed=700 + i)
        main = synth.frame(spec)
        out.append((main, synth.prepass_for(main)))
    out.append(cases.supported_job_frames())
    return out


@pytest.fixture(scope="module")
def oracle_fingerprints(frames):
    params = stable_profile_params()
    prepass_params = DEFAULT_PROFILE.prepass_parameters()
    prints = []
    for main, prepass in frames:
        cpu = oracle_lib.process_job(main, prepass, params, prepass_params)
        prints.append({
            "output_sha256": cpu.output_sha256,
            "attempted_pixels": cpu.attempted_pixels,
            "written_pixels": cpu.written_pixels,
            "public_rng_advances": cpu.public_rng_advances,
            "final_rng_state": cpu.final_rng_state,
            "changed_pixels": cpu.changed_pixels,
        })
    return prints


def _core(fingerprint: dict) -> dict:
    keys = ("output_sha256", "attempted_pixels", "written_pixels", "public_rng_advances",
            "final_rng_state", "changed_pixels")
    return {k: fingerprint[k] for k in keys}


def _jobs(frames):
    return [make_job(main, prepass, f"batch-{i}") for i, (main, prepass) in enumerate(frames)]


def test_batch_matches_oracle_and_single_runs(frames, oracle_fingerprints):
    outcome = process_batch_hip(_jobs(frames))
    outcome.raise_first()
    for i, (main, prepass) in enumerate(frames):
        got = replay_fingerprint(outcome.results[i].replay)
        assert _core(got) == oracle_fingerprints[i], f"job {i} {main.shape}"
        alone = process_hip(make_job(main, prepass, f"alone-{i}"))
        assert replay_fingerprint(alone.replay) == got, f"job {i} batch vs alone"
        assert np.array_equal(alone.output_rgb16, outcome.results[i].output_rgb16)


@pytest.mark.parametrize("mode", [1, 0], ids=["pairs", "one-per-job"])
def test_same_geometry_jobs_hashed_in_pairs_match_oracle(mode):
    """With pairs forced on, jobs of one geometry in a batch are hashed two
    by two, the two messages interleaved on one thread, and a third job of
    that geometry hashes alone; with them off, every job has its own hash
    thread.  Every job has its own content, so a digest handed to the wrong
    member of a pair fails; the host outputs are compared as well."""

    lib = engine.library()
    params = stable_profile_params()
    prepass_params = DEFAULT_PROFILE.prepass_parameters()
    shapes = [(97, 211)] * 3 + [(256, 256)] * 2
    frames = []
    for i, (h, w) in enumerate(shapes):
        main = synth.frame(synth.scaled_spec(h, w, seed=910 + i))
        frames.append((main, synth.prepass_for(main)))
    jobs = [make_job(main, prepass, f"pair-{i}") for i, (main, prepass) in enumerate(frames)]
    previous = C.c_int32(-1)
    assert lib.dice_t_hash_pairs(mode, C.byref(previous)) == 0
    try:
        outcome = process_batch_hip(jobs)
    finally:
        assert lib.dice_t_hash_pairs(previous.value, None) == 0
    outcome.raise_first()
    shas = set()
    for i, (main, prepass) in enumerate(frames):
        cpu = oracle_lib.process_job(main, prepass, params, prepass_params)
        result = outcome.results[i]
        assert result.replay.output_sha256 == cpu.output_sha256, f"job {i} {main.shape}"
        assert np.array_equal(result.output_rgb16, cpu.output), f"job {i}"
        assert result.replay.final_rng_state == cpu.final_rng_state
        shas.add(cpu.output_sha256)
    assert len(shas) == len(frames)


def test_hash_pair_mode_rejects_unknown_modes():
    lib = engine.library()
    assert lib.dice_t_hash_pairs(3, None) != 0
    assert lib.dice_t_hash_pairs(-1, None) != 0


def test_batch_graph_replays_match_eager(frames, oracle_fingerprints):
    jobs = _jobs(frames)
    eager = process_batch_hip(jobs)
    eager.raise_first()
    expected = [replay_fingerprint(r.replay) for r in eager.results]
    replayed = []
    for attempt in range(4):
        outcome = process_batch_hip(jobs, graph=True)
        outcome.raise_first()
        got = [replay_fingerprint(r.replay) for r in outcome.results]
        assert got == expected, f"graph run {attempt}"
        for i, r in enumerate(outcome.results):
            assert np.array_equal(r.output_rgb16, eager.results[i].output_rgb16)
        replayed.append(outcome.report["graph_replayed"])
        # unrelated eager work and a synchronization between replays
        main, prepass = frames[1]
        alone = process_hip(make_job(main, prepass, "between"))
        assert _core(replay_fingerprint(alone.replay)) == oracle_fingerprints[1]
    assert replayed[0] == 0 and all(replayed[2:]), replayed


def test_device_chain_matches_host_chain(frames, oracle_fingerprints):
    """The single-wavefront device chain reproduces every counter, the
    startup receipt and the output of the host chain."""

    jobs = _jobs(frames)
    host = process_batch_hip(jobs)
    host.raise_first()
    for attempt in range(2):
        device = process_batch_hip(jobs, device_chain=True, graph=attempt == 1)
        device.raise_first()
        for i in range(len(frames)):
            got = replay_fingerprint(device.results[i].replay)
            assert got == replay_fingerprint(host.results[i].replay), f"job {i} run {attempt}"
            assert _core(got) == oracle_fingerprints[i]
            assert np.array_equal(device.results[i].output_rgb16, host.results[i].output_rgb16)


def test_batch_autotune_keeps_results(frames, oracle_fingerprints):
    outcome = process_batch_hip(_jobs(frames), autotune=True)
    outcome.raise_first()
    shape = (outcome.report["tile_rows"], outcome.report["tile_cols"])
    assert shape in {(16, 64), (8, 128), (32, 32), (16, 32), (32, 64), (8, 64)}
    for i in range(len(frames)):
        assert _core(replay_fingerprint(outcome.results[i].replay)) == oracle_fingerprints[i]


def test_autotune_winner_is_reused_by_bucket_and_persisted(tmp_path):
    """The tuner keys its winner by (architecture, bucket), not by exact
    geometry: a second batch of different crops in the same bucket reuses
    it, and after the in-memory winners are forgotten it is read back from
    the cache file.  Every configuration's results were checked against the
    default's (none rejected), and every run equals the oracle."""

    from dice_amd import synth

    lib = engine.library()
    cache = tmp_path / "tuning.tsv"
    assert lib.dice_tuning_cache(str(cache).encode()) == 0
    try:
        def crops(shapes, seed):
            out = []
            for i, (h, w) in enumerate(shapes):
                main = synth.frame(synth.scaled_spec(h, w, seed=seed + i))
                out.append((main, synth.prepass_for(main)))
            return out

        def check(batch, outcome):
            outcome.raise_first()
            params = stable_profile_params()
            for (main, prepass), result in zip(batch, outcome.results):
                cpu = oracle_lib.process_job(main, prepass, params,
                                             DEFAULT_PROFILE.prepass_parameters())
                assert result.replay.output_sha256 == cpu.output_sha256

        first = crops([(700, 900), (800, 1000)], 7100)
        tuned = process_batch_hip([make_job(m, p, "tune-a") for m, p in first], autotune=True)
        check(first, tuned)
        report = tuned.report
        assert report["tune_source"] == 1
        assert report["tune_candidates"] == 12 and report["tune_rejected"] == 0
        winner = (report["tile_rows"], report["tile_cols"], report["sites_variant"])
        # different geometries, same bucket: n2, total and largest in the same octaves
        second = crops([(650, 950), (820, 980)], 7200)
        hit = process_batch_hip([make_job(m, p, "tune-b") for m, p in second], autotune=True)
        check(second, hit)
        assert hit.report["tune_source"] == 2
        assert (hit.report["tile_rows"], hit.report["tile_cols"],
                hit.report["sites_variant"]) == winner
        lines = cache.read_text().splitlines()
        assert any(line.split("\t")[1] == "n2-t20-m19" for line in lines), lines
        assert lib.dice_tuning_cache(str(cache).encode()) == 0  # forget, read back
        again = process_batch_hip([make_job(m, p, "tune-c") for m, p in second],
                                  autotune=True)
        check(second, again)
        assert again.report["tune_source"] == 3
        assert (again.report["tile_rows"], again.report["tile_cols"],
                again.report["sites_variant"]) == winner
    finally:
        lib.dice_tuning_cache(None)


def test_failing_job_fails_alone(frames, oracle_fingerprints):
    """A job rejected by validation inside the batch does not disturb the others."""

    lib = engine.library()
    params = stable_profile_params()
    prepass_params = DEFAULT_PROFILE.prepass_parameters()
    picks = [frames[0], frames[1]]
    bad = np.full((4, 16, 4), 40000, np.uint16)  # fewer than five rows
    keep, outs, results = [], [], []
    array = (DiceJob * 3)()
    for i, (main, prepass) in enumerate([picks[0], (bad, bad), picks[1]]):
        main = np.ascontiguousarray(main)
        prepass = np.ascontiguousarray(prepass)
        out = np.zeros(main.shape[:2] + (3,), np.uint16)
        res = DiceReplayResult()
        keep += [main, prepass]
        outs.append(out)
        results.append(res)
        e = array[i]
        e.main_rgbi = main.ctypes.data
        e.height, e.width = main.shape[:2]
        e.prepass_rgbi = prepass.ctypes.data
        e.prepass_height, e.prepass_width = prepass.shape[:2]
        e.params = C.pointer(params)
        e.prepass_params = C.pointer(prepass_params)
        e.out_rgb16 = out.ctypes.data
        e.result = C.pointer(res)
    report = DiceBatchReport()
    code = lib.dice_process_batch(engine.context(), array, 3, 0, C.byref(report))
    assert code == _abi.DICE_INVALID, lib.dice_last_error()
    assert array[1].status == _abi.DICE_INVALID
    assert b"at least 5 rows" in array[1].error
    for slot, index in ((0, 0), (2, 1)):
        assert array[slot].status == _abi.DICE_OK
        assert results[slot].output_sha256.decode() == oracle_fingerprints[index]["output_sha256"]
        assert results[slot].attempted_pixels == oracle_fingerprints[index]["attempted_pixels"]


def test_batch_device_resident_buffers(frames, oracle_fingerprints):
    """Inputs and outputs already in HBM (the serving layout)."""

    import hashlib

    from dice_amd.hip_backend.device import DeviceArray

    lib = engine.library()
    params = stable_profile_params()
    prepass_params = DEFAULT_PROFILE.prepass_parameters()
    picks = frames[:3]
    keep, results, outs = [], [], []
    array = (DiceJob * len(picks))()
    for i, (main, prepass) in enumerate(picks):
        d_main = DeviceArray.from_host(main)
        d_out = DeviceArray(main.shape[0] * main.shape[1] * 6)
        prepass = np.ascontiguousarray(prepass)
        res = DiceReplayResult()
        keep += [d_main, prepass]
        outs.append(d_out)
        results.append(res)
        e = array[i]
        e.main_rgbi = d_main.address
        e.height, e.width = main.shape[:2]
        e.prepass_rgbi = prepass.ctypes.data
        e.prepass_height, e.prepass_width = prepass.shape[:2]
        e.params = C.pointer(params)
        e.prepass_params = C.pointer(prepass_params)
        e.flags = _abi.DICE_MAIN_ON_DEVICE | _abi.DICE_OUT_ON_DEVICE
        e.out_rgb16 = d_out.address
        e.result = C.pointer(res)
    code = lib.dice_process_batch(engine.context(), array, len(picks), 0, None)
    assert code == 0, lib.dice_last_error()
    for i, (main, _) in enumerate(picks):
        assert results[i].output_sha256.decode() == oracle_fingerprints[i]["output_sha256"]
        host = outs[i].download(np.empty(main.shape[:2] + (3,), np.uint16))
        assert hashlib.sha256(host.tobytes()).hexdigest() == oracle_fingerprints[i]["output_sha256"]


def test_uneven_crops_workload_matches_oracle():
    """The 16-crop workload (sides 512-3000 px, aspect 0.5-2, defects scaled
    to area) as one batch with graph replay and autotuned tiles: every crop
    equals the oracle, on the run that captures the graphs and on a replay."""
    specs = synth.config3_crops()
    frames = [(synth.frame(s), None) for s in specs]
    frames = [(main, synth.prepass_for(main)) for main, _ in frames]
    params = stable_profile_params()
    prepass_params = DEFAULT_PROFILE.prepass_parameters()
    expected = []
    for main, prepass in frames:
        cpu = oracle_lib.process_job(main, prepass, params, prepass_params, threads=16)
        expected.append({
            "output_sha256": cpu.output_sha256,
            "attempted_pixels": cpu.attempted_pixels,
            "written_pixels": cpu.written_pixels,
            "public_rng_advances": cpu.public_rng_advances,
            "final_rng_state": cpu.final_rng_state,
            "changed_pixels": cpu.changed_pixels,
        })
    jobs = [make_job(main, prepass, f"crop-{i}") for i, (main, prepass) in enumerate(frames)]
    for run in range(3):  # direct, capture, replay
        outcome = process_batch_hip(jobs, graph=True, autotune=True)
        outcome.raise_first()
        for i, result in enumerate(outcome.results):
            assert _core(replay_fingerprint(result.replay)) == expected[i], (run, i)
            assert hashlib.sha256(result.output_rgb16.tobytes()).hexdigest() == \
                expected[i]["output_sha256"]
    assert outcome.report["graph_replayed"] == 1
    # the same batch with its 16 writer chains walked together on the device
    chained = process_batch_hip(jobs, device_chain=True)
    chained.raise_first()
    for i, result in enumerate(chained.results):
        assert _core(replay_fingerprint(result.replay)) == expected[i], ("device chain", i)


def test_seed0_crop_matches_oracle():
    """The 256 x 256 seed-0 crop (12 disks r 2-12, 3 scratches 50-200 px)."""
    main = synth.frame(synth.config1_crop())
    prepass = synth.prepass_for(main)
    cpu = oracle_lib.process_job(main, prepass, stable_profile_params(),
                                 DEFAULT_PROFILE.prepass_parameters(), diagnostics=True)
    result = process_hip(make_job(main, prepass, "seed0"), export_diagnostics=True)
    assert result.replay.output_sha256 == cpu.output_sha256
    assert np.array_equal(result.output_rgb16, cpu.output)
    assert result.replay.attempted_pixels == cpu.attempted_pixels
    assert np.array_equal(result.diagnostics.score_plane.view(np.uint32),
                          cpu.score_plane.view(np.uint32))
    assert np.array_equal(result.diagnostics.changed_mask, cpu.changed_mask)
