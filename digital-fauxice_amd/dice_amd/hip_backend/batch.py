"""``process_batch_hip`` -- many independent jobs in one set of device launches.

The reference processes one job per call (``engine.py:110-267``,
``cuda_backend/process.py:25-159``).  Here a list of jobs (crops or frames
of uneven size) runs through ``dice_process_batch``: each device phase is
one persistent launch over every job, optionally replayed from a captured
hipGraph, with the analysis tile shape optionally picked by timing every
compiled shape on the batch.  Each job's result is the one ``process_hip``
returns for it alone, byte for byte; a job that fails raises its own
exception without touching the others.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from .. import _abi
from .._abi import DiceBatchReport, DiceJob, DiceReplayResult, DiceStageTimings, ptr
from ..engine import ProcessingCancelled, ProcessingResult
from ..profile import DEFAULT_PROFILE, ProcessingJob
from .engine import HipBackendUnavailable, _result_from, context, library
from .process import stable_profile_params


@dataclass
class BatchOutcome:
    """Per-job results (or exceptions) of one batch and its report."""

    results: list = field(default_factory=list)
    report: dict = field(default_factory=dict)
    timings: list = field(default_factory=list)

    def raise_first(self) -> None:
        for item in self.results:
            if isinstance(item, BaseException):
                raise item


def _exception_for(code: int, message: str) -> BaseException:
    if code in (_abi.DICE_INVALID, _abi.DICE_NONFINITE_AUX, _abi.DICE_NONFINITE_WORK):
        return ValueError(message)
    if code == _abi.DICE_CANCELLED:
        return ProcessingCancelled(message)
    return HipBackendUnavailable(message)


def process_batch_hip(jobs: list[ProcessingJob], *, graph: bool = False,
                      autotune: bool = False, device_chain: bool = False) -> BatchOutcome:
    """Run ``jobs`` (1..64) as one batch on the HIP device.

    ``device_chain`` runs each job's serial writer chain on one device
    wavefront instead of a host core (same results; frees the host and the
    host link for throughput).
    """

    if not 1 <= len(jobs) <= 64:
        raise ValueError("a batch holds 1..64 jobs")
    for job in jobs:
        DEFAULT_PROFILE.validate_job(job)
    params = stable_profile_params()
    prepass_params = DEFAULT_PROFILE.prepass_parameters()
    keep = []
    array = (DiceJob * len(jobs))()
    outputs, results, timings = [], [], []
    for i, job in enumerate(jobs):
        main = np.ascontiguousarray(job.acquisition.main.pixels)
        prepass = np.ascontiguousarray(job.acquisition.prepass.pixels)
        height, width = main.shape[:2]
        out = np.empty((height, width, 3), np.uint16)
        res = DiceReplayResult()
        tim = DiceStageTimings()
        keep += [main, prepass]
        outputs.append(out)
        results.append(res)
        timings.append(tim)
        entry = array[i]
        entry.main_rgbi = main.ctypes.data
        entry.height = height
        entry.width = width
        entry.prepass_rgbi = prepass.ctypes.data
        entry.prepass_height = prepass.shape[0]
        entry.prepass_width = prepass.shape[1]
        entry.params = C.pointer(params)
        entry.prepass_params = C.pointer(prepass_params)
        entry.flags = 0
        entry.out_rgb16 = out.ctypes.data
        entry.result = C.pointer(res)
        entry.timings = C.pointer(tim)
    flags = ((_abi.DICE_BATCH_GRAPH if graph else 0) | (_abi.DICE_BATCH_AUTOTUNE if autotune else 0)
             | (_abi.DICE_BATCH_DEVICE_CHAIN if device_chain else 0))
    report = DiceBatchReport()
    code = library().dice_process_batch(context(), array, len(jobs), flags, C.byref(report))
    outcome = BatchOutcome(report={name: getattr(report, name) for name, _ in report._fields_})
    if code != _abi.DICE_OK and all(array[i].status == _abi.DICE_OK for i in range(len(jobs))):
        raise _exception_for(code, library().dice_last_error().decode(errors="replace"))
    for i, job in enumerate(jobs):
        entry = array[i]
        if entry.status != _abi.DICE_OK:
            outcome.results.append(_exception_for(entry.status, entry.error.decode(errors="replace")))
            outcome.timings.append(None)
            continue
        height, width = outputs[i].shape[:2]
        outcome.results.append(ProcessingResult(
            output_rgb16=outputs[i],
            replay=_result_from(results[i], height, width),
            profile_id=DEFAULT_PROFILE.profile_id,
            diagnostics=None,
        ))
        outcome.timings.append(timings[i].as_dict())
    return outcome
