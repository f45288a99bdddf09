"""``HipServingEngine`` -- continuous batching of scan jobs on one MI355X.

A thin owner of ``dice_engine`` (libdice_hip.so): jobs are submitted as
they arrive and processed in batches formed by free drivers, admitted by an
HBM budget; ``wait`` returns each job's ``ProcessingResult``, byte-identical
to ``process_hip`` on the same job.  The reference has no serving layer; the
per-job contract (validation, exceptions, result fields) is the one of
``process_cuda`` (``cuda_backend/process.py:25-159``).
"""

from __future__ import annotations

import ctypes as C
import threading

import numpy as np

from .. import _abi
from .._abi import DiceEngineConfig, DiceEngineStats, DiceJob, DiceReplayResult, DiceStageTimings
from ..engine import ProcessingResult
from ..profile import DEFAULT_PROFILE, ProcessingJob
from .batch import _exception_for
from .engine import HipBackendUnavailable, _result_from, library
from .process import stable_profile_params


class _Pending:
    """Everything a submitted job points to, kept alive until it is waited for."""

    def __init__(self, job_struct, keep, output, result, timings, shape) -> None:
        self.job = job_struct
        self.keep = keep
        self.output = output
        self.result = result
        self.timings = timings
        self.shape = shape


class HipServingEngine:
    """Submit / wait front end of the serving engine."""

    def __init__(self, *, device: int = 0, drivers: int = 3, max_batch: int = 8,
                 hbm_budget: int = 0, graph: bool = True, autotune: bool = False,
                 linger_ms: float = 2.0, device_chain: bool = False) -> None:
        self._lib = library()
        config = DiceEngineConfig(device, drivers, max_batch,
                                  (_abi.DICE_BATCH_GRAPH if graph else 0)
                                  | (_abi.DICE_BATCH_AUTOTUNE if autotune else 0)
                                  | (_abi.DICE_BATCH_DEVICE_CHAIN if device_chain else 0),
                                  hbm_budget, linger_ms)
        handle = C.c_void_p()
        code = self._lib.dice_engine_create(C.byref(config), C.byref(handle))
        if code != _abi.DICE_OK:
            raise HipBackendUnavailable(self._lib.dice_last_error().decode(errors="replace"))
        self._handle = handle
        self._params = stable_profile_params()
        self._prepass_params = DEFAULT_PROFILE.prepass_parameters()
        self._pending: dict[int, _Pending] = {}
        self._lock = threading.Lock()

    def submit(self, job: ProcessingJob) -> int:
        """Queue one validated job; returns its ticket."""

        DEFAULT_PROFILE.validate_job(job)
        main = np.ascontiguousarray(job.acquisition.main.pixels)
        prepass = np.ascontiguousarray(job.acquisition.prepass.pixels)
        height, width = main.shape[:2]
        output = np.empty((height, width, 3), np.uint16)
        return self._submit(main.ctypes.data, height, width, prepass, output.ctypes.data, 0,
                            keep=[main, prepass], output=output)

    def submit_device(self, main_address: int, height: int, width: int,
                      prepass: np.ndarray, out_address: int, *, sha256: bool = True) -> int:
        """Queue a job whose main frame and output are device (HBM) buffers."""

        flags = _abi.DICE_MAIN_ON_DEVICE | _abi.DICE_OUT_ON_DEVICE
        if not sha256:
            flags |= _abi.DICE_SKIP_SHA256
        prepass = np.ascontiguousarray(prepass)
        return self._submit(main_address, height, width, prepass, out_address, flags,
                            keep=[prepass], output=None)

    def _submit(self, main_address, height, width, prepass, out_address, flags, *, keep,
                output) -> int:
        result = DiceReplayResult()
        timings = DiceStageTimings()
        entry = DiceJob()
        entry.main_rgbi = main_address
        entry.height = height
        entry.width = width
        entry.prepass_rgbi = prepass.ctypes.data
        entry.prepass_height, entry.prepass_width = prepass.shape[:2]
        entry.params = C.pointer(self._params)
        entry.prepass_params = C.pointer(self._prepass_params)
        entry.flags = flags
        entry.out_rgb16 = out_address
        entry.result = C.pointer(result)
        entry.timings = C.pointer(timings)
        ticket = C.c_uint64()
        with self._lock:
            code = self._lib.dice_engine_submit(self._handle, C.byref(entry), C.byref(ticket))
            if code != _abi.DICE_OK:
                raise _exception_for(code, self._lib.dice_last_error().decode(errors="replace"))
            self._pending[ticket.value] = _Pending(entry, keep, output, result, timings,
                                                   (height, width))
        return ticket.value

    def reservation(self, job: ProcessingJob) -> int:
        """The HBM bytes ``job`` is admitted against (its workspace bound)."""

        entry = DiceJob()
        entry.height, entry.width = job.acquisition.main.pixels.shape[:2]
        entry.flags = 0
        value = C.c_uint64()
        code = self._lib.dice_engine_reservation(C.byref(entry), C.byref(value))
        if code != _abi.DICE_OK:
            raise _exception_for(code, self._lib.dice_last_error().decode(errors="replace"))
        return int(value.value)

    def wait(self, ticket: int, timeout: float | None = None):
        """The job's ``ProcessingResult`` (its output is None for device jobs)."""

        code = self._lib.dice_engine_wait(self._handle, ticket,
                                          -1.0 if timeout is None else float(timeout))
        if code == _abi.DICE_PENDING:
            raise TimeoutError(f"job {ticket} has not finished")
        with self._lock:
            pending = self._pending.pop(ticket)
        if code != _abi.DICE_OK:
            raise _exception_for(code, pending.job.error.decode(errors="replace"))
        height, width = pending.shape
        return ProcessingResult(
            output_rgb16=pending.output,
            replay=_result_from(pending.result, height, width),
            profile_id=DEFAULT_PROFILE.profile_id,
            diagnostics=None,
        )

    def stats(self) -> dict:
        report = DiceEngineStats()
        code = self._lib.dice_engine_stats(self._handle, C.byref(report))
        if code != _abi.DICE_OK:
            raise HipBackendUnavailable(self._lib.dice_last_error().decode(errors="replace"))
        return report.as_dict()

    def reset_stats(self) -> None:
        """Start a new measurement window for ``stats``."""

        self._lib.dice_engine_reset_stats(self._handle)

    def close(self) -> None:
        if self._handle is not None:
            self._lib.dice_engine_destroy(self._handle)
            self._handle = None

    def __enter__(self) -> "HipServingEngine":
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self) -> None:
        try:
            self.close()
        except Exception:  # noqa: BLE001 - interpreter shutdown
            pass
