"""``process_hip`` -- one supported job on the HIP backend.

Mirrors ``process_cuda`` (rohanpandula/digital-fauxice,
``src/portable_digital_ice/cuda_backend/process.py:25-159``) and the shared
orchestration ``engine.py:110-267``: validate, prepass, producer, provider,
replay, diagnostics, commit the caller-owned output only on success.  The
whole pipeline after validation runs in one call into libdice_hip.so.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import numpy.typing as npt

from .._abi import DiceDiagnostics, DiceReplayResult, DiceStageTimings, ptr
from ..engine import (
    PHASES,
    CancellationCallback,
    ProcessingCancelled,
    ProcessingDiagnostics,
    ProcessingProgress,
    ProcessingResult,
    ProgressCallback,
)
from ..parameters import generate_nikon_response_lut, to_dice_params
from ..profile import DEFAULT_PROFILE, ProcessingJob, neutral_prepass_record
from .engine import CallbackBridge, _result_from, context, library


def stable_profile_params():
    """The profile's stable replay parameters (the provider replaces every
    record-derived field per row, so a neutral record suffices here)."""

    record = neutral_prepass_record()
    lut = generate_nikon_response_lut()
    return to_dice_params(
        auxiliary_parameters=DEFAULT_PROFILE.auxiliary_parameters(record),
        score_parameters=DEFAULT_PROFILE.score_parameters(record),
        decision_parameters=DEFAULT_PROFILE.decision_parameters(),
        reconstruction_parameters=DEFAULT_PROFILE.reconstruction_parameters(record),
        dither_bounds=DEFAULT_PROFILE.dither_bounds(lut),
    )


def process_hip(
    job: ProcessingJob,
    *,
    output_rgb16: npt.NDArray[np.uint16] | None = None,
    progress: ProgressCallback | None = None,
    cancelled: CancellationCallback | None = None,
    stage_timings: dict | None = None,
    export_diagnostics: bool = False,
) -> ProcessingResult:
    """Process one validated LS-5000 selector-8 Normal acquisition on HIP.

    The result is only correct if it is byte-identical to the CPU path; the
    parity ladders enforce that.  Any HIP problem raises
    ``HipBackendUnavailable``; nothing falls back to another numeric path.
    """

    def notify(phase: int, completed: int, total: int, attempted: int = 0, written: int = 0):
        if progress is not None:
            progress(ProcessingProgress(PHASES[phase], completed, total, attempted, written))

    DEFAULT_PROFILE.validate_job(job)
    main = job.acquisition.main.pixels
    prepass = job.acquisition.prepass.pixels
    height, width = main.shape[:2]
    expected_shape = (height, width, 3)
    destination = None
    if output_rgb16 is not None:
        destination = np.asarray(output_rgb16)
        if destination.dtype != np.dtype(np.uint16) or destination.shape != expected_shape:
            raise ValueError(f"output must be writable uint16 with shape {expected_shape}")
        if not destination.flags.writeable:
            raise ValueError("output must be writable")
    if cancelled is not None and cancelled():
        raise ProcessingCancelled("portable correction was cancelled")

    params = stable_profile_params()
    prepass_params = DEFAULT_PROFILE.prepass_parameters()
    ctx = context()
    # The library writes a host output buffer only after the whole job
    # succeeded, so a C-contiguous caller buffer is handed over directly.
    if destination is not None and destination.flags.c_contiguous:
        working_output = destination
    else:
        working_output = np.empty(expected_shape, dtype=np.uint16)
    diag = None
    if export_diagnostics:
        score = np.empty((height, width), np.float32)
        at_floor = np.empty((height, width), np.uint8)
        changed = np.empty((height, width), np.uint8)
        diag = DiceDiagnostics(ptr(score, C.c_float), ptr(at_floor, C.c_uint8),
                               ptr(changed, C.c_uint8))
    bridge = CallbackBridge(notify, cancelled)
    timings = DiceStageTimings()
    res = DiceReplayResult()
    main_c = np.ascontiguousarray(main)
    prepass_c = np.ascontiguousarray(prepass)
    code = library().dice_process_job(
        ctx, main_c.ctypes.data_as(C.c_void_p), height, width, ptr(prepass_c, C.c_uint16),
        prepass_c.shape[0], prepass_c.shape[1], C.byref(params), C.byref(prepass_params), 0,
        working_output.ctypes.data_as(C.c_void_p), C.byref(res),
        C.byref(diag) if diag is not None else None, C.byref(bridge.struct),
        C.byref(timings))
    bridge.finish(code)
    if stage_timings is not None:
        stage_timings.update(timings.as_dict())
    replay = _result_from(res, height, width)
    diagnostics = None
    if export_diagnostics:
        score.flags.writeable = False
        floor_mask = at_floor.astype(bool)
        changed_mask = changed.astype(bool)
        floor_mask.flags.writeable = False
        changed_mask.flags.writeable = False
        diagnostics = ProcessingDiagnostics(
            score_plane=score,
            score_floor=np.float32(params.score_floor),
            at_floor_mask=floor_mask,
            changed_mask=changed_mask,
        )
    if destination is None:
        output = working_output
    else:
        if working_output is not destination:
            np.copyto(destination, working_output)
        output = destination
    return ProcessingResult(
        output_rgb16=output,
        replay=replay,
        profile_id=DEFAULT_PROFILE.profile_id,
        diagnostics=diagnostics,
    )
