"""Frame streams: the main pass arrives as 8-row scanner blocks.

A frame pushed block by block through ``HipFrameStream`` must give exactly
what the assembled frame gives: the reference's own job vectors (output
bytes, counters, startup receipt, diagnostics planes) for every job case,
the oracle's result for uneven synthetic frames at the full-frame defect
density and for every block size, and ``process_hip``'s result on a full
5782 x 3946 frame.  While the blocks arrive, the output rows that are final
trail the rows received by a bounded number of rows (the analysis halo,
the 16-row row-sum units and one tile row), and a stream is reusable.
Block-order errors are refused the way assemble_rgbi16_blocks refuses them
(contracts.py:237-261).
"""

from __future__ import annotations

import ctypes as C
import hashlib

import numpy as np
import pytest

import cases
import oracle_lib
from conftest import hip_device_or_skip
from dice_amd import AcquisitionEpoch, RGBI16Frame, _abi, synth
from dice_amd.contracts import BlockSchedule, RGBI16Block, split_rgbi16_frame
from dice_amd.hip_backend import process_hip
from dice_amd.hip_backend.engine import library
from dice_amd.hip_backend.process import stable_profile_params
from dice_amd.hip_backend.stream import HipFrameStream
from dice_amd.profile import DEFAULT_PROFILE
from test_hip_parity import make_job, replay_fingerprint, sha

pytestmark = pytest.mark.gpu

# rows the final output may trail the received rows by: 7 rows of a partial
# block, 15 of a partial row-sum unit, the 6-row halo of the last tile row
# and one 16-row tile row (the default analysis tile is 16 x 64)
MAX_LAG = 7 + 15 + 6 + 16


@pytest.fixture(scope="module", autouse=True)
def _device():
    hip_device_or_skip()


@pytest.fixture(scope="module")
def stream():
    with HipFrameStream() as s:
        yield s


def prepass_frame(prepass: np.ndarray) -> RGBI16Frame:
    return RGBI16Frame(prepass, AcquisitionEpoch.PREPASS, 285, "stream-pp")


def main_frame(main: np.ndarray) -> RGBI16Frame:
    return RGBI16Frame(main, AcquisitionEpoch.MAIN, 4000, "stream-main")


def run_stream(stream, main, prepass, rows_per_block=8, diagnostics=False, check_lag=True):
    height, width = main.shape[:2]
    schedule = BlockSchedule.for_frame(height, width, rows_per_block)
    stream.begin(prepass_frame(prepass), schedule, export_diagnostics=diagnostics)
    blocks = split_rgbi16_frame(main_frame(main), rows_per_block)
    assert len(blocks) == schedule.block_count
    final = 0
    for block in blocks:
        progress = stream.push(block)
        received = block.start_row + block.valid_rows
        assert progress["rows_received"] == received
        assert final <= progress["rows_final"] <= received
        final = progress["rows_final"]
        if check_lag:
            assert progress["rows_final"] >= min(height, received - MAX_LAG), progress
    assert final == height
    return stream.finish()


@pytest.mark.parametrize("name", list(cases.job_cases()))
def test_stream_matches_reference_vectors(name, stream, reference_vectors):
    expected = reference_vectors["jobs"][name]
    main, prepass = cases.job_cases()[name]
    result = run_stream(stream, main, prepass, diagnostics=True)
    got = replay_fingerprint(result.replay)
    assert got == {k: expected[k] for k in got}, name
    assert sha(result.output_rgb16.astype("<u2")) == expected["output_samples_sha256"]
    d = result.diagnostics
    assert sha(d.score_plane.astype("<f4")) == expected["score_plane_sha256"]
    assert sha(d.at_floor_mask) == expected["at_floor_sha256"]
    assert sha(d.changed_mask) == expected["changed_sha256"]


SHAPES = [(262, 300), (97, 211), (181, 640), (513, 129), (40, 1000)]


@pytest.fixture(scope="module")
def synthetic():
    out = []
    params = stable_profile_params()
    prepass_params = DEFAULT_PROFILE.prepass_parameters()
    for i, (h, w) in enumerate(SHAPES):
        main = synth.frame(synth.scaled_spec(h, w, seed=900 + i))
        prepass = synth.prepass_for(main)
        cpu = oracle_lib.process_job(main, prepass, params, prepass_params, diagnostics=True)
        out.append((main, prepass, cpu))
    return out


@pytest.mark.parametrize("rows_per_block", [8, 1, 3, 16, 64, 1000])
def test_stream_matches_oracle_for_every_block_size(rows_per_block, stream, synthetic):
    for main, prepass, cpu in synthetic:
        result = run_stream(stream, main, prepass, rows_per_block, diagnostics=True,
                            check_lag=rows_per_block <= 8)
        replay = result.replay
        assert replay.output_sha256 == cpu.output_sha256
        assert np.array_equal(result.output_rgb16, cpu.output)
        assert (replay.attempted_pixels, replay.written_pixels, replay.public_rng_advances,
                replay.final_rng_state, replay.changed_pixels) == (
            cpu.attempted_pixels, cpu.written_pixels, cpu.public_rng_advances,
            cpu.final_rng_state, cpu.changed_pixels)
        assert replay.startup.attempted_per_stage == cpu.startup_attempted
        assert replay.startup.rng_advances_per_stage == cpu.startup_advances
        assert replay.startup.final_rng_state == cpu.startup_final_rng_state
        d = result.diagnostics
        assert np.array_equal(d.score_plane.view(np.uint32), cpu.score_plane.view(np.uint32))
        assert np.array_equal(d.at_floor_mask, cpu.at_floor_mask)
        assert np.array_equal(d.changed_mask, cpu.changed_mask)


def test_full_frame_stream_matches_whole_frame(stream):
    main = synth.frame(synth.config2_full_frame())
    prepass = synth.prepass_for(main)
    whole = process_hip(make_job(main, prepass, "whole"))
    streamed = run_stream(stream, main, prepass)
    assert replay_fingerprint(streamed.replay) == replay_fingerprint(whole.replay)
    assert hashlib.sha256(streamed.output_rgb16.tobytes()).hexdigest() == \
        whole.replay.output_sha256


def test_stream_is_reusable_and_a_failed_frame_does_not_leak(stream, synthetic):
    main, prepass, cpu = synthetic[0]
    small_main, small_prepass, small_cpu = synthetic[1]
    assert run_stream(stream, main, prepass).replay.output_sha256 == cpu.output_sha256
    # an abandoned frame: begin, two blocks, then an out-of-order block
    schedule = BlockSchedule.for_frame(*main.shape[:2])
    stream.begin(prepass_frame(prepass), schedule)
    blocks = split_rgbi16_frame(main_frame(main), 8)
    stream.push(blocks[0])
    stream.push(blocks[1])
    with pytest.raises(ValueError, match="ordered, gap-free"):
        stream.push(blocks[3])
    with pytest.raises(ValueError, match="incomplete"):
        stream.finish()
    result = run_stream(stream, small_main, small_prepass)
    assert result.replay.output_sha256 == small_cpu.output_sha256


def test_block_rules_at_the_c_boundary(stream, synthetic):
    main, prepass, _ = synthetic[1]
    height, width = main.shape[:2]
    lib = library()
    schedule = BlockSchedule.for_frame(height, width)
    stream.begin(prepass_frame(prepass), schedule)
    handle = stream._handle
    rows = np.ascontiguousarray(main[8:16])
    p = rows.ctypes.data_as(C.POINTER(C.c_uint16))
    assert lib.dice_stream_push(handle, p, 8, 8) == _abi.DICE_INVALID  # gap
    # a refused block leaves the frame as it was
    first = np.ascontiguousarray(main[:8])
    assert lib.dice_stream_push(handle, first.ctypes.data_as(C.POINTER(C.c_uint16)), 0, 8) == 0
    tail = np.ascontiguousarray(main[8:])
    assert lib.dice_stream_push(handle, tail.ctypes.data_as(C.POINTER(C.c_uint16)), 8,
                                height) == _abi.DICE_INVALID  # past the frame
    assert lib.dice_stream_push(handle, tail.ctypes.data_as(C.POINTER(C.c_uint16)), 8,
                                height - 8) == 0
    assert lib.dice_stream_push(handle, p, height, 1) == _abi.DICE_INVALID  # after the end
    stream._frame["next_block"] = schedule.block_count
    result = stream.finish()
    whole = process_hip(make_job(main, prepass, "boundary"))
    assert result.replay.output_sha256 == whole.replay.output_sha256


def test_block_rules_in_python(stream, synthetic):
    main, prepass, _ = synthetic[1]
    schedule = BlockSchedule.for_frame(*main.shape[:2])
    stream.begin(prepass_frame(prepass), schedule)
    blocks = split_rgbi16_frame(main_frame(main), 8)
    with pytest.raises(ValueError, match="ordered, gap-free"):
        stream.push(blocks[1])
    with pytest.raises(ValueError, match="must hold"):
        stream.push(RGBI16Block(main[:5], 0))
    with pytest.raises(ValueError, match="same width"):
        stream.push(RGBI16Block(main[:8, :-1], 0))
    with pytest.raises(TypeError):
        stream.push(main[:8])
    for block in blocks:
        stream.push(block)
    with pytest.raises(ValueError, match="every block"):
        stream.push(blocks[0])
    stream.finish()
    with pytest.raises(ValueError, match="no frame"):
        stream.finish()


def c_stream_job(main, prepass, out_address, flags, result):
    """A dice_job for the C stream entry points (main rows come by push)."""

    from dice_amd._abi import DiceJob

    job = DiceJob()
    job.main_rgbi = None
    job.height, job.width = main.shape[:2]
    job.prepass_rgbi = prepass.ctypes.data
    job.prepass_height, job.prepass_width = prepass.shape[:2]
    job.params = C.pointer(stable_profile_params())
    job.prepass_params = C.pointer(DEFAULT_PROFILE.prepass_parameters())
    job.flags = flags
    job.out_rgb16 = out_address
    job.result = C.pointer(result)
    return job


def push_all(lib, handle, main, rows_per_block=8):
    for row in range(0, main.shape[0], rows_per_block):
        block = np.ascontiguousarray(main[row:row + rows_per_block])
        code = lib.dice_stream_push(handle, block.ctypes.data_as(C.POINTER(C.c_uint16)), row,
                                    block.shape[0])
        assert code == 0, lib.dice_last_error()


def test_stream_flags_device_output_and_no_hash(synthetic):
    """DICE_OUT_ON_DEVICE writes the output into the caller's device buffer;
    DICE_SKIP_SHA256 leaves the hash empty and everything else unchanged."""
    from dice_amd._abi import DiceReplayResult
    from dice_amd.hip_backend.device import DeviceArray
    from dice_amd.hip_backend.engine import context

    main, prepass, cpu = synthetic[2]
    prepass = np.ascontiguousarray(prepass)
    lib = library()
    handle = C.c_void_p()
    assert lib.dice_stream_create(context(), C.byref(handle)) == 0
    try:
        for flags in (_abi.DICE_OUT_ON_DEVICE, _abi.DICE_OUT_ON_DEVICE | _abi.DICE_SKIP_SHA256,
                      _abi.DICE_SKIP_SHA256):
            device_out = DeviceArray(cpu.output.nbytes)
            host_out = np.zeros_like(cpu.output)
            on_device = bool(flags & _abi.DICE_OUT_ON_DEVICE)
            result = DiceReplayResult()
            job = c_stream_job(main, prepass,
                               device_out.address if on_device else host_out.ctypes.data,
                               flags, result)
            assert lib.dice_stream_begin(handle, C.byref(job), None) == 0, lib.dice_last_error()
            push_all(lib, handle, main)
            assert lib.dice_stream_finish(handle) == 0, lib.dice_last_error()
            assert job.status == 0
            got = device_out.download(np.empty_like(cpu.output)) if on_device else host_out
            assert np.array_equal(got, cpu.output), flags
            assert (result.attempted_pixels, result.written_pixels, result.final_rng_state,
                    result.changed_pixels) == (cpu.attempted_pixels, cpu.written_pixels,
                                               cpu.final_rng_state, cpu.changed_pixels)
            want = b"" if flags & _abi.DICE_SKIP_SHA256 else cpu.output_sha256.encode()
            assert result.output_sha256 == want
    finally:
        lib.dice_stream_destroy(handle)


def test_streams_on_two_threads(synthetic):
    """Two streams, each on its own thread (each has its own runner, HIP
    streams and workspace), give the results each gives alone."""
    import threading

    outcomes = {}

    def run(k):
        main, prepass, cpu = synthetic[k]
        with HipFrameStream() as s:
            results = [run_stream(s, main, prepass, check_lag=False) for _ in range(2)]
        outcomes[k] = (cpu, results)

    threads = [threading.Thread(target=run, args=(k,)) for k in (0, 3)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert sorted(outcomes) == [0, 3]
    for cpu, results in outcomes.values():
        for result in results:
            assert result.replay.output_sha256 == cpu.output_sha256
            assert np.array_equal(result.output_rgb16, cpu.output)


def test_refused_begin_abandons_the_frame_in_progress(stream, synthetic):
    """A begin that fails validation still ends the frame before it: later
    pushes are refused instead of feeding the old frame."""
    from dice_amd._abi import DiceReplayResult

    main, prepass, _ = synthetic[1]
    stream.begin(prepass_frame(prepass), BlockSchedule.for_frame(*main.shape[:2]))
    stream.push(split_rgbi16_frame(main_frame(main), 8)[0])
    lib = library()
    result = DiceReplayResult()
    bad = c_stream_job(main, np.ascontiguousarray(prepass), 0, 0, result)
    bad.height = 3  # below the 5-row minimum: refused before the device
    assert lib.dice_stream_begin(stream._handle, C.byref(bad), None) == _abi.DICE_INVALID
    assert bad.status == _abi.DICE_INVALID
    block = np.ascontiguousarray(main[8:16])
    assert lib.dice_stream_push(stream._handle, block.ctypes.data_as(C.POINTER(C.c_uint16)),
                                8, 8) == _abi.DICE_INVALID
    stream._frame = None
