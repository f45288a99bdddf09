"""Backend-selection contract of the HIP drop-in, on the CPU.

Mirrors the reference's tests/test_backend_contract.py:110-128 (an invalid
job fails before any backend probe), :196-216 (the self-test outcome is
cached), :219-254 (the parity comparator catches tampered results) and
:257-290 (a failure is cached and fails closed), for
``dice_amd.backend``.  ``process_hip`` is replaced by a stand-in: either
the oracle's result for the job (so the real comparator and the shipped
self-test fingerprint are exercised end to end), a tampered copy of it, or
a stand-in that fails.  The same tests on the GPU, with the real
``process_hip``, are in tests/test_hip_contract.py.
"""

from __future__ import annotations

import json
from contextlib import contextmanager
from dataclasses import replace

import numpy as np
import pytest

import oracle_lib as O
from dice_amd import backend as backend_module
from dice_amd.backend import ComputeBackend, process
from dice_amd.contracts import AcquisitionEpoch, DualRGBIAcquisition, RGBI16Frame
from dice_amd.engine import ProcessingDiagnostics, ProcessingResult, StartupReplayResult
from dice_amd.engine import StreamingReplayResult
from dice_amd.hip_backend import HipBackendUnavailable
from dice_amd.hip_backend import engine as hip_engine
from dice_amd.hip_backend.process import stable_profile_params
from dice_amd.profile import DEFAULT_PROFILE, ProcessingJob, ProcessingMode, ScannerModel


@contextmanager
def self_test_cache_snapshot():
    """Isolate self-test cache mutations so tests stay order-independent."""

    saved = dict(backend_module._SELF_TEST_CACHE)
    try:
        backend_module._SELF_TEST_CACHE.clear()
        yield backend_module._SELF_TEST_CACHE
    finally:
        backend_module._SELF_TEST_CACHE.clear()
        backend_module._SELF_TEST_CACHE.update(saved)


def tiny_job(frame_id: str = "backend-contract", **overrides) -> ProcessingJob:
    """tests/test_backend_contract.py:72-91 of the reference."""

    rng = np.random.default_rng(11)
    main = rng.integers(18000, 62000, size=(12, 24, 4), dtype=np.uint16)
    main[5:8, 9:14, 3] = 700
    prepass = rng.integers(18000, 62000, size=(8, 16, 4), dtype=np.uint16)
    fields = dict(
        acquisition=DualRGBIAcquisition(
            prepass=RGBI16Frame(prepass, AcquisitionEpoch.PREPASS, 285, "pp"),
            main=RGBI16Frame(main, AcquisitionEpoch.MAIN, 4000, "main"),
            same_frame_id=frame_id,
        ),
        scanner_model=ScannerModel.NIKON_SUPER_COOLSCAN_5000_ED,
        mode=ProcessingMode.NORMAL,
        selector=8,
        resolution_metric=4000,
        bit_depth=16,
        focus_exposure_locked=True,
    )
    fields.update(overrides)
    return ProcessingJob(**fields)


def oracle_process(job, *, output_rgb16=None, progress=None, cancelled=None,
                   export_diagnostics=False, stage_timings=None) -> ProcessingResult:
    """A stand-in for process_hip that returns the oracle's result."""

    r = O.process_job(job.acquisition.main.pixels, job.acquisition.prepass.pixels,
                      stable_profile_params(), DEFAULT_PROFILE.prepass_parameters(),
                      diagnostics=export_diagnostics)
    height, width = r.output.shape[:2]
    replay = StreamingReplayResult(
        shape=(height, width, 3),
        startup=StartupReplayResult(r.startup_attempted, r.startup_advances,
                                    r.startup_final_rng_state),
        attempted_pixels=r.attempted_pixels, written_pixels=r.written_pixels,
        public_rng_advances=r.public_rng_advances, final_rng_state=r.final_rng_state,
        output_sha256=r.output_sha256, changed_pixels=r.changed_pixels)
    diagnostics = None
    if export_diagnostics:
        diagnostics = ProcessingDiagnostics(
            score_plane=r.score_plane, score_floor=np.float32(0.02),
            at_floor_mask=r.at_floor_mask, changed_mask=r.changed_mask)
    output = r.output
    if output_rgb16 is not None:
        np.copyto(output_rgb16, r.output)
        output = output_rgb16
    return ProcessingResult(output_rgb16=output, replay=replay,
                            profile_id=DEFAULT_PROFILE.profile_id, diagnostics=diagnostics)


def test_backend_enum_values():
    assert ComputeBackend.AUTO.value == "auto"
    assert ComputeBackend.CPU.value == "cpu"
    assert ComputeBackend.CPU_FAST.value == "cpu-fast"
    assert ComputeBackend.CUDA.value == "cuda"
    assert ComputeBackend.METAL.value == "metal"
    assert ComputeBackend.HIP.value == "hip"


@pytest.mark.parametrize("field,value", [("selector", 9), ("resolution_metric", 500),
                                          ("bit_depth", 8), ("focus_exposure_locked", False),
                                          ("mode", "fast")])
def test_unsupported_job_fails_before_backend_probe(monkeypatch, field, value):
    """tests/test_backend_contract.py:110-128: validation precedes the probe."""

    def probe(*args, **kwargs):
        raise AssertionError("the backend was probed before the job was validated")

    monkeypatch.setattr(backend_module, "process_hip", probe)
    monkeypatch.setattr(backend_module, "hip_self_test", probe)
    bad = tiny_job(**{field: value})
    for backend in ("hip", "auto", "cpu", "cuda", "metal"):
        with pytest.raises((ValueError, TypeError)):
            process(bad, backend=backend)


def test_unknown_backend_name_is_rejected():
    with pytest.raises(ValueError):
        process(tiny_job(), backend="opencl")


def test_explicit_hip_without_library_raises_and_never_falls_back(monkeypatch, tmp_path):
    monkeypatch.setattr(hip_engine, "LIBRARY_PATH", tmp_path / "libdice_hip.so")
    monkeypatch.setattr(hip_engine, "_LIB", [])
    monkeypatch.setattr(hip_engine, "_CONTEXTS", {})
    with self_test_cache_snapshot():
        with pytest.raises(HipBackendUnavailable, match="HIP library not built"):
            process(tiny_job(), backend=ComputeBackend.HIP)
        # AUTO has no CPU backend in this package: it raises with the reason
        with pytest.raises(HipBackendUnavailable, match="HIP unavailable .*not built"):
            process(tiny_job(), backend=ComputeBackend.AUTO)


def test_non_hip_backends_are_not_served_here():
    for backend in ("cpu", "cpu-fast", "cuda", "metal"):
        with pytest.raises(HipBackendUnavailable, match="not part of this package"):
            process(tiny_job(), backend=backend)


def test_self_test_passes_on_exact_results_and_names_the_reason(monkeypatch):
    monkeypatch.setattr(backend_module, "process_hip", oracle_process)
    with self_test_cache_snapshot() as cache:
        routed = process(tiny_job(), backend="hip")
        assert routed.selection.used is ComputeBackend.HIP
        assert routed.selection.requested is ComputeBackend.HIP
        assert routed.selection.reason == "explicit HIP request; self-test passed"
        assert cache["hip-outcome"] is None
        auto = process(tiny_job(), backend=ComputeBackend.AUTO)
        assert auto.selection.used is ComputeBackend.HIP
        assert auto.selection.reason == "startup self-test passed byte parity"
        assert auto.result.replay.output_sha256 == routed.result.replay.output_sha256


def test_self_test_outcome_is_cached(monkeypatch):
    """tests/test_backend_contract.py:196-216."""

    calls = {"n": 0}

    def counting(job, **kwargs):
        calls["n"] += 1
        return oracle_process(job, **kwargs)

    monkeypatch.setattr(backend_module, "process_hip", counting)
    with self_test_cache_snapshot():
        backend_module.hip_self_test()
        assert calls["n"] == 1
        backend_module.hip_self_test()  # cached success: no rerun
        assert calls["n"] == 1
        process(tiny_job(), backend="hip")  # self-test cached, one processing call
        assert calls["n"] == 2


def test_parity_comparison_catches_tampered_results(monkeypatch):
    """tests/test_backend_contract.py:219-254."""

    def tampered(job, **kwargs):
        result = oracle_process(job, **kwargs)
        wrong_output = result.output_rgb16.copy()
        wrong_output[0, 0, 0] ^= 1
        wrong_mask = result.diagnostics.at_floor_mask.copy()
        wrong_mask[0, 0] = not wrong_mask[0, 0]
        return replace(result, output_rgb16=wrong_output,
                       diagnostics=replace(result.diagnostics, at_floor_mask=wrong_mask))

    monkeypatch.setattr(backend_module, "process_hip", tampered)
    with self_test_cache_snapshot():
        with pytest.raises(HipBackendUnavailable) as caught:
            backend_module.hip_self_test()
        assert "failed parity" in str(caught.value)
        assert "output sample mismatch" in str(caught.value)
        assert "at-floor mask mismatch" in str(caught.value)


@pytest.mark.parametrize("field", ["attempted_pixels", "written_pixels", "public_rng_advances",
                                   "final_rng_state", "changed_pixels", "output_sha256"])
def test_parity_comparison_catches_each_counter(field):
    expected = json.loads(backend_module.SELFTEST_EXPECTED.read_text())
    candidate = dict(expected)
    candidate[field] = "tampered" if field == "output_sha256" else expected[field] + 1
    failures = backend_module.parity_failures(expected, candidate, "hip")
    assert len(failures) == 1 and field.split("_")[0] in failures[0]
    assert backend_module.parity_failures(expected, dict(expected), "hip") == []
    no_planes = {k: v for k, v in expected.items() if "sha256" not in k or k.startswith("output")}
    no_planes.pop("score_floor_bits", None)
    assert backend_module.parity_failures(expected, no_planes, "hip") == [
        "diagnostics planes missing from self-test run"]


def test_tampered_self_test_fixture_is_rejected(monkeypatch, tmp_path):
    """A fixture whose recorded input does not match the synthetic job fails
    closed instead of comparing against the wrong expectations."""

    expected = json.loads(backend_module.SELFTEST_EXPECTED.read_text())
    expected["main_sha256"] = "0" * 64
    fake = tmp_path / "selftest_expected.json"
    fake.write_text(json.dumps(expected))
    monkeypatch.setattr(backend_module, "SELFTEST_EXPECTED", fake)
    monkeypatch.setattr(backend_module, "process_hip", oracle_process)
    with self_test_cache_snapshot():
        with pytest.raises(HipBackendUnavailable, match="input differs"):
            backend_module.hip_self_test()
        with pytest.raises(HipBackendUnavailable, match="input differs"):
            process(tiny_job(), backend="hip")


def test_tampered_expected_output_fails_parity(monkeypatch, tmp_path):
    expected = json.loads(backend_module.SELFTEST_EXPECTED.read_text())
    expected["changed_sha256"] = "f" * 64
    fake = tmp_path / "selftest_expected.json"
    fake.write_text(json.dumps(expected))
    monkeypatch.setattr(backend_module, "SELFTEST_EXPECTED", fake)
    monkeypatch.setattr(backend_module, "process_hip", oracle_process)
    with self_test_cache_snapshot():
        with pytest.raises(HipBackendUnavailable, match="changed mask mismatch"):
            backend_module.hip_self_test()


def test_failure_is_cached_and_fails_closed(monkeypatch):
    """tests/test_backend_contract.py:257-290."""

    calls = {"n": 0}

    def broken(*args, **kwargs):
        calls["n"] += 1
        raise HipBackendUnavailable("HIP is not usable: injected for test")

    monkeypatch.setattr(backend_module, "process_hip", broken)
    with self_test_cache_snapshot():
        with pytest.raises(HipBackendUnavailable):
            process(tiny_job(), backend=ComputeBackend.HIP)
        assert calls["n"] == 1
        with pytest.raises(HipBackendUnavailable) as second:
            process(tiny_job(), backend=ComputeBackend.HIP)
        assert calls["n"] == 1
        assert "HIP is not usable" in str(second.value)
        with pytest.raises(HipBackendUnavailable) as auto:
            process(tiny_job(), backend=ComputeBackend.AUTO)
        assert calls["n"] == 1
        assert "HIP unavailable" in str(auto.value)


def test_caller_buffer_and_callbacks_reach_the_backend(monkeypatch):
    seen = {}

    def recording(job, **kwargs):
        seen.update(kwargs)
        return oracle_process(job, **kwargs)

    monkeypatch.setattr(backend_module, "process_hip", recording)
    buffer = np.zeros((12, 24, 3), np.uint16)

    def progress(event):
        return None

    def cancelled():
        return False

    with self_test_cache_snapshot():
        routed = process(tiny_job(), backend="hip", output_rgb16=buffer, progress=progress,
                         cancelled=cancelled, export_diagnostics=True)
    assert routed.result.output_rgb16 is buffer
    assert seen["output_rgb16"] is buffer
    assert seen["progress"] is progress and seen["cancelled"] is cancelled
    assert seen["export_diagnostics"] is True
