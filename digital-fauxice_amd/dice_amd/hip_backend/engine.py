"""ctypes binding of libdice_hip.so and the HIP streaming replay.

Mirrors the CUDA backend's engine surface (rohanpandula/digital-fauxice,
``src/portable_digital_ice/cuda_backend/engine.py:58-590``):
``HipBackendUnavailable``, ``hip_device_summary()`` and
``run_streaming_replay_hip(...)`` with the same arguments and results as
``run_streaming_replay``.  Every numeric step runs in the HIP library; this
module only validates arguments, marshals buffers and maps return codes to
the reference's exceptions.  There is no fallback: a missing library or
device raises ``HipBackendUnavailable``.
"""

from __future__ import annotations

import ctypes as C
import os
import threading
from pathlib import Path

import numpy as np
import numpy.typing as npt

from .. import _abi
from .._abi import (
    DiceCallbacks,
    DiceDeviceInfo,
    DiceDiagnostics,
    DiceParams,
    DiceReplayResult,
    DiceStageTimings,
    ptr,
)
from ..engine import (
    LCG24,
    ProcessingCancelled,
    StartupReplayResult,
    StreamingReplayResult,
)
from ..parameters import (
    AuxiliaryParameters,
    DecisionParameters,
    DitherBounds,
    ReconstructionParameters,
    ScoreParameters,
    to_dice_params,
)
from ..stages import StageTable, constant_stage_table

LIBRARY_PATH = Path(__file__).resolve().parents[1] / "libdice_hip.so"


class HipBackendUnavailable(RuntimeError):
    """HIP execution was requested but cannot run exactly here."""


_LOCK = threading.RLock()  # context() -> hip_device_summary() -> library() re-enters
_LIB: list[C.CDLL] = []
_CONTEXTS: dict[int, C.c_void_p] = {}


def library() -> C.CDLL:
    """Load libdice_hip.so (fails closed when it is missing)."""

    with _LOCK:
        if _LIB:
            return _LIB[0]
        if not LIBRARY_PATH.exists():
            raise HipBackendUnavailable(
                f"HIP library not built: {LIBRARY_PATH} is missing "
                "(run __graft_entry__.build())"
            )
        # HIP maps all of a process's streams onto GPU_MAX_HW_QUEUES hardware
        # queues (4 by default); the batch runtime's runners drive three
        # streams each, which four queues serialize.  Read when HIP starts, so
        # this default only takes effect if nothing in the process has used
        # HIP yet; an explicit setting wins.
        os.environ.setdefault("GPU_MAX_HW_QUEUES", "16")
        try:
            lib = C.CDLL(str(LIBRARY_PATH))
        except OSError as error:
            raise HipBackendUnavailable(f"cannot load {LIBRARY_PATH}: {error}") from error
        _abi.declare(lib, "dice_", _abi.HIP_SIGNATURES)
        if lib.dice_abi_version() != _abi.ABI_VERSION:
            raise HipBackendUnavailable("libdice_hip.so ABI version mismatch")
        _LIB.append(lib)
        return lib


def last_error() -> str:
    return library().dice_last_error().decode(errors="replace")


def raise_for(code: int) -> None:
    """Map a C ABI return code onto the reference's exception types."""

    if code == _abi.DICE_OK:
        return
    message = last_error()
    if code in (_abi.DICE_INVALID, _abi.DICE_NONFINITE_AUX, _abi.DICE_NONFINITE_WORK):
        raise ValueError(message)
    if code == _abi.DICE_CANCELLED:
        raise ProcessingCancelled(message)
    raise HipBackendUnavailable(message)


def hip_device_summary() -> dict:
    """Describe the HIP device, failing closed when absent (engine.py:77-100)."""

    info = DiceDeviceInfo()
    code = library().dice_probe(C.byref(info))
    if code != _abi.DICE_OK:
        raise HipBackendUnavailable(last_error())
    return {
        "name": info.name.decode(),
        "arch": info.arch.decode(),
        "compute_units": int(info.compute_units),
        "device_count": int(info.device_count),
        "total_hbm_bytes": int(info.total_hbm_bytes),
        "free_hbm_bytes": int(info.free_hbm_bytes),
        "build_flags": info.build_flags.decode(),
    }


def context(device: int = 0) -> C.c_void_p:
    """The per-process, per-device library context."""

    lib = library()
    with _LOCK:
        ctx = _CONTEXTS.get(device)
        if ctx is None:
            hip_device_summary()
            handle = C.c_void_p()
            code = lib.dice_ctx_create(device, 0, C.byref(handle))
            if code != _abi.DICE_OK:
                raise HipBackendUnavailable(last_error())
            ctx = handle
            _CONTEXTS[device] = ctx
        return ctx


class CallbackBridge:
    """Exposes Python progress/cancel callables as C function pointers.

    Exceptions raised by the Python callables are stored and re-raised after
    the C call returns (the C side sees a cancellation request instead).
    """

    def __init__(self, progress=None, cancelled=None) -> None:
        self.error: BaseException | None = None
        self._progress = progress
        self._cancelled = cancelled

        def on_progress(_user, phase, completed, total, attempted, written):
            if self.error is not None or self._progress is None:
                return
            try:
                self._progress(int(phase), int(completed), int(total), int(attempted),
                               int(written))
            except BaseException as error:  # noqa: BLE001 - re-raised after the call
                self.error = error

        def on_cancelled(_user):
            if self.error is not None:
                return 1
            if self._cancelled is None:
                return 0
            try:
                return 1 if self._cancelled() else 0
            except BaseException as error:  # noqa: BLE001
                self.error = error
                return 1

        self._p = _abi.ProgressFn(on_progress)
        self._c = _abi.CancelledFn(on_cancelled)
        self.struct = DiceCallbacks(self._p, self._c, None)

    def finish(self, code: int) -> None:
        if self.error is not None:
            raise self.error
        raise_for(code)


def _stage_table_from_provider(provider, height: int) -> StageTable:
    """Evaluate a stage parameter provider for stages 1..H+6."""

    count = height + 6
    table = StageTable.empty(count)
    for stage in range(1, count + 1):
        calibration = provider(stage)
        if getattr(calibration, "stage_hit", stage) != stage:
            raise ValueError(
                "stage parameter provider returned calibration for "
                f"stage {calibration.stage_hit}, expected {stage}"
            )
        i = stage - 1
        table.aux_alpha[i] = np.float32(calibration.auxiliary_alpha)
        table.aux_offset[i] = np.float32(calibration.calibration_offset)
        table.aux_replacement[i] = np.float32(calibration.base_primary)
        table.aux_has_replacement[i] = 1
        table.score_base[i] = np.float32(calibration.base_primary)
        table.coarse_reference[i] = np.float32(calibration.base_primary)
        table.gate_secondary[i] = int(bool(calibration.writer_gate_secondary))
        table.row_gate[i] = int(calibration.row_reconstruction_gate)
    return table


def _result_from(res: DiceReplayResult, height: int, width: int) -> StreamingReplayResult:
    return StreamingReplayResult(
        shape=(height, width, 3),
        startup=StartupReplayResult(
            attempted_per_stage=tuple(int(v) for v in res.startup_attempted),
            rng_advances_per_stage=tuple(int(v) for v in res.startup_advances),
            final_rng_state=int(res.startup_final_rng_state),
        ),
        attempted_pixels=int(res.attempted_pixels),
        written_pixels=int(res.written_pixels),
        public_rng_advances=int(res.public_rng_advances),
        final_rng_state=int(res.final_rng_state),
        output_sha256=res.output_sha256.decode(),
        changed_pixels=int(res.changed_pixels),
    )


def run_streaming_replay_hip(
    main_rgbi: npt.ArrayLike,
    output_rgb16: npt.NDArray[np.uint16],
    *,
    response=None,
    auxiliary_parameters: AuxiliaryParameters,
    score_parameters: ScoreParameters,
    decision_parameters: DecisionParameters,
    reconstruction_parameters: ReconstructionParameters,
    dither_bounds: DitherBounds,
    generator: LCG24 | None = None,
    stage_parameter_provider=None,
    progress=None,
    cancelled=None,
    stage_timings: dict | None = None,
    _diagnostic_score_plane: npt.NDArray[np.float32] | None = None,
    _diagnostic_at_floor_mask: npt.NDArray[np.bool_] | None = None,
    _diagnostic_changed_mask: npt.NDArray[np.bool_] | None = None,
) -> StreamingReplayResult:
    """Byte-exact HIP mirror of ``streaming.run_streaming_replay``."""

    pixels = np.asarray(main_rgbi)
    output = np.asarray(output_rgb16)
    if pixels.dtype != np.dtype(np.uint16) or pixels.ndim != 3 or pixels.shape[2] != 4:
        raise ValueError("streaming main input must be uint16 HxWx4")
    height, width, _ = pixels.shape
    if height < 4:
        raise ValueError("streaming startup requires at least four main rows")
    if output.dtype != np.dtype(np.uint16) or output.shape != (height, width, 3):
        raise ValueError("streaming output must be uint16 HxWx3")
    if not output.flags.writeable:
        raise ValueError("streaming output must be writable")
    if isinstance(stage_parameter_provider, StageTable):
        table = stage_parameter_provider
    elif stage_parameter_provider is not None:
        table = _stage_table_from_provider(stage_parameter_provider, height)
    else:
        table = constant_stage_table(
            height,
            auxiliary_parameters=auxiliary_parameters,
            score_parameters=score_parameters,
            reconstruction_parameters=reconstruction_parameters,
        )
    params: DiceParams = to_dice_params(
        auxiliary_parameters=auxiliary_parameters,
        score_parameters=score_parameters,
        decision_parameters=decision_parameters,
        reconstruction_parameters=reconstruction_parameters,
        dither_bounds=dither_bounds,
    )
    active_generator = generator or LCG24.from_nikon_pe_initial_state()
    ctx = context()
    source = np.ascontiguousarray(pixels, dtype=np.uint16)
    staging = np.empty((height, width, 3), dtype=np.uint16)
    diag = None
    if _diagnostic_score_plane is not None:
        if _diagnostic_at_floor_mask is None or _diagnostic_changed_mask is None:
            raise ValueError("all HIP diagnostic buffers must be provided")
        floor_u8 = np.empty((height, width), np.uint8)
        changed_u8 = np.empty((height, width), np.uint8)
        score = np.empty((height, width), np.float32)
        diag = DiceDiagnostics(ptr(score, C.c_float), ptr(floor_u8, C.c_uint8),
                               ptr(changed_u8, C.c_uint8))

    def row_progress(phase, completed, total, attempted, written):
        if progress is not None and phase == 3:
            progress(completed, total, attempted, written)

    bridge = CallbackBridge(row_progress, cancelled)
    timings = DiceStageTimings()
    res = DiceReplayResult()
    st = table.as_struct()
    code = library().dice_replay(
        ctx, source.ctypes.data_as(C.c_void_p), height, width, None, C.byref(params),
        C.byref(st), int(active_generator.state), 0,
        staging.ctypes.data_as(C.c_void_p), C.byref(res),
        C.byref(diag) if diag is not None else None, C.byref(bridge.struct),
        C.byref(timings))
    bridge.finish(code)
    if stage_timings is not None:
        for key, value in timings.as_dict().items():
            stage_timings[key] = stage_timings.get(key, 0.0) + value
    if diag is not None:
        _diagnostic_score_plane[...] = score
        _diagnostic_at_floor_mask[...] = floor_u8.astype(bool)
        _diagnostic_changed_mask[...] = changed_u8.astype(bool)
    output[...] = staging
    active_generator.state = int(res.final_rng_state)
    return _result_from(res, height, width)
