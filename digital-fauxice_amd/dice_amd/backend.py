"""Fail-closed backend selection with the HIP backend.

Mirrors ``src/portable_digital_ice/backend.py`` of
rohanpandula/digital-fauxice: ``ComputeBackend`` gains ``HIP = "hip"``;
``process(job, backend=...)`` validates the job before probing any backend,
never falls back when a backend is requested explicitly, and runs a cached
per-process parity self-test before first use (backend.py:57-174, 275-443).

The self-test runs the reference's synthetic 48x64 acquisition
(backend.py:60-87) with diagnostics and compares every output sample, the
five replay counters and the three diagnostics planes against the expected
values of the CPU path, which ship with this package as
``selftest_expected.json`` (comparator: backend.py:90-141).  The CPU
backends themselves (cpu, cpu-fast) stay in the reference package; this
package ships only the HIP branch, so ``auto`` resolves to HIP or raises.
"""

from __future__ import annotations

import hashlib
import json
from dataclasses import dataclass
from enum import Enum
from pathlib import Path

import numpy as np

from .contracts import AcquisitionEpoch, DualRGBIAcquisition, RGBI16Frame
from .engine import ProcessingResult
from .hip_backend import HipBackendUnavailable, process_hip
from .profile import DEFAULT_PROFILE, ProcessingJob, ProcessingMode, ScannerModel

SELFTEST_EXPECTED = Path(__file__).resolve().parent / "selftest_expected.json"


class ComputeBackend(str, Enum):
    AUTO = "auto"
    CPU = "cpu"
    CPU_FAST = "cpu-fast"
    CUDA = "cuda"
    METAL = "metal"
    HIP = "hip"


@dataclass(frozen=True)
class BackendSelection:
    requested: ComputeBackend
    used: ComputeBackend
    reason: str


@dataclass(frozen=True)
class BackendProcessingResult:
    result: ProcessingResult
    selection: BackendSelection


_SELF_TEST_CACHE: dict[str, str | None] = {}


def synthetic_self_test_job() -> ProcessingJob:
    """backend.py:60-87 -- deterministic small acquisition covering repair
    and dither paths (numpy default_rng(20260716))."""

    rng = np.random.default_rng(20260716)
    height, width = 48, 64
    main = rng.integers(20000, 60000, size=(height, width, 4), dtype=np.uint16)
    main[18:26, 30:39, 3] = rng.integers(200, 2000, size=(8, 9), dtype=np.uint16)
    main[0:2, 0:6, 3] = 500
    main[-2:, -6:, 3] = 700
    main[3, width - 3, 3] = 900
    prepass = rng.integers(20000, 60000, size=(24, 32, 4), dtype=np.uint16)
    prepass[10:14, 12:17, 3] = 300
    acquisition = DualRGBIAcquisition(
        prepass=RGBI16Frame(prepass, AcquisitionEpoch.PREPASS, 285, "selftest-prepass"),
        main=RGBI16Frame(main, AcquisitionEpoch.MAIN, 4000, "selftest-main"),
        same_frame_id="hip-backend-self-test",
    )
    return ProcessingJob(
        acquisition=acquisition,
        scanner_model=ScannerModel.NIKON_SUPER_COOLSCAN_5000_ED,
        mode=ProcessingMode.NORMAL,
        selector=8,
        resolution_metric=4000,
        bit_depth=16,
        focus_exposure_locked=True,
    )


def _sha(array: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(array).tobytes()).hexdigest()


def result_fingerprint(result: ProcessingResult) -> dict:
    """Everything the parity comparator looks at, in hashable form."""

    diagnostics = result.diagnostics
    fingerprint = {
        "output_sha256": result.replay.output_sha256,
        "output_samples_sha256": _sha(result.output_rgb16.astype("<u2")),
        "attempted_pixels": result.replay.attempted_pixels,
        "written_pixels": result.replay.written_pixels,
        "public_rng_advances": result.replay.public_rng_advances,
        "final_rng_state": result.replay.final_rng_state,
        "changed_pixels": result.replay.changed_pixels,
    }
    if diagnostics is not None:
        fingerprint.update(
            score_plane_sha256=_sha(diagnostics.score_plane.astype("<f4")),
            score_floor_bits=int(np.asarray(diagnostics.score_floor).view(np.uint32)),
            at_floor_sha256=_sha(diagnostics.at_floor_mask.astype(np.bool_)),
            changed_sha256=_sha(diagnostics.changed_mask.astype(np.bool_)),
        )
    return fingerprint


def parity_failures(expected: dict, candidate: dict, label: str) -> list[str]:
    """backend.py:90-141, over fingerprints."""

    failures: list[str] = []
    if expected["output_sha256"] != candidate["output_sha256"]:
        failures.append(
            f"output hash mismatch cpu={expected['output_sha256']} "
            f"{label}={candidate['output_sha256']}"
        )
    if expected["output_samples_sha256"] != candidate["output_samples_sha256"]:
        failures.append("output sample mismatch")
    for field in ("attempted_pixels", "written_pixels", "public_rng_advances",
                  "final_rng_state", "changed_pixels"):
        if expected[field] != candidate[field]:
            failures.append(f"{field} mismatch cpu={expected[field]} {label}={candidate[field]}")
    if "score_plane_sha256" not in candidate:
        failures.append("diagnostics planes missing from self-test run")
    else:
        for field, message in (("score_plane_sha256", "score plane mismatch"),
                               ("score_floor_bits", "score floor mismatch"),
                               ("at_floor_sha256", "at-floor mask mismatch"),
                               ("changed_sha256", "changed mask mismatch")):
            if expected[field] != candidate[field]:
                failures.append(message)
    return failures


def hip_self_test() -> None:
    """Prove HIP/CPU byte parity on the synthetic job or raise.

    The (successful or failed) outcome is cached per process; a failure
    reason is re-raised on later calls.
    """

    cached = _SELF_TEST_CACHE.get("hip-outcome", "unset")
    if cached is None:
        return
    if cached != "unset":
        raise HipBackendUnavailable(str(cached))
    try:
        expected = json.loads(SELFTEST_EXPECTED.read_text())
        job = synthetic_self_test_job()
        if job.acquisition.main.sha256 != expected["main_sha256"]:
            reason = "HIP self-test input differs from the recorded synthetic job"
            _SELF_TEST_CACHE["hip-outcome"] = reason
            raise HipBackendUnavailable(reason)
        candidate = result_fingerprint(process_hip(job, export_diagnostics=True))
        failures = parity_failures(expected, candidate, "hip")
        if failures:
            reason = "HIP self-test failed parity: " + "; ".join(failures)
            _SELF_TEST_CACHE["hip-outcome"] = reason
            raise HipBackendUnavailable(reason)
    except Exception as error:
        if _SELF_TEST_CACHE.get("hip-outcome", "unset") == "unset":
            _SELF_TEST_CACHE["hip-outcome"] = f"HIP self-test raised: {error!r}"
        raise
    _SELF_TEST_CACHE["hip-outcome"] = None


def process(
    job: ProcessingJob,
    *,
    backend: ComputeBackend | str = ComputeBackend.AUTO,
    output_rgb16=None,
    progress=None,
    cancelled=None,
    export_diagnostics: bool = False,
) -> BackendProcessingResult:
    """Process one job on the requested backend with fail-closed semantics."""

    DEFAULT_PROFILE.validate_job(job)
    requested = ComputeBackend(backend)
    if requested in (ComputeBackend.HIP, ComputeBackend.AUTO):
        try:
            hip_self_test()
        except Exception as error:
            if requested is ComputeBackend.HIP:
                raise
            raise HipBackendUnavailable(
                f"HIP unavailable ({error}); the CPU backends are provided by "
                "portable_digital_ice itself"
            ) from error
        result = process_hip(
            job,
            output_rgb16=output_rgb16,
            progress=progress,
            cancelled=cancelled,
            export_diagnostics=export_diagnostics,
        )
        reason = (
            "explicit HIP request; self-test passed"
            if requested is ComputeBackend.HIP
            else "startup self-test passed byte parity"
        )
        return BackendProcessingResult(result, BackendSelection(requested, ComputeBackend.HIP, reason))
    raise HipBackendUnavailable(
        f"backend {requested.value!r} is not part of this package; "
        "it is provided by portable_digital_ice itself"
    )
