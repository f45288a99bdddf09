"""ctypes bindings of the CPU oracle (oracle/libdice_oracle.so).

Test infrastructure: the oracle is the checker the HIP path is compared
against.  The same structs (dice_amd._abi) describe both libraries, so one
set of inputs drives both.
"""

from __future__ import annotations

import ctypes as C
import hashlib
import subprocess
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from dice_amd import _abi
from dice_amd._abi import (
    DiceDiagnostics,
    DiceParams,
    DicePrepassParams,
    DicePrepassResult,
    DiceReplayResult,
    DiceSchedule,
    DiceStageTable,
    f32p,
    f64p,
    i32p,
    i64,
    ptr,
    u8p,
    u16p,
    u32p,
    u64p,
)

ROOT = Path(__file__).resolve().parents[1]
ORACLE_DIR = ROOT / "oracle"
ORACLE_SO = ORACLE_DIR / "libdice_oracle.so"

ORACLE_SIGNATURES = {
    "last_error": (C.c_char_p, []),
    "response_lut": (None, [f32p]),
    "inverse_factors": (None, [u32p, u32p]),
    "dither_bounds": (None, [f32p, f32p, f32p, i32p, i32p]),
    "prepass": (C.c_int32, [u16p, i64, i64, f32p, C.POINTER(DicePrepassParams),
                            C.POINTER(DicePrepassResult)]),
    "producer_schedule": (C.c_int32, [u16p, i64, i64, f32p, C.POINTER(DiceSchedule)]),
    "stage_table_from_schedule": (
        C.c_int32,
        [C.POINTER(DiceSchedule), C.c_float, i64, f32p, f32p, f32p, u8p, f32p,
         f32p, u8p, i32p],
    ),
    "replay": (
        C.c_int32,
        [u16p, i64, i64, f32p, C.POINTER(DiceParams), C.POINTER(DiceStageTable),
         C.c_uint32, u16p, C.POINTER(DiceReplayResult), C.POINTER(DiceDiagnostics),
         C.c_int32],
    ),
    "process_job": (
        C.c_int32,
        [u16p, i64, i64, u16p, i64, i64, C.POINTER(DiceParams),
         C.POINTER(DicePrepassParams), u16p, C.POINTER(DiceReplayResult),
         C.POINTER(DiceDiagnostics), C.c_int32],
    ),
    "unscaled_averages": (None, [f32p, i64, f64p]),
    "rgb_unscaled_averages": (None, [f32p, i64, f64p]),
    "dither_chain": (None, [f64p, f32p, i64, C.c_uint32, C.c_float, C.c_float,
                            f64p, u32p, u64p]),
    "write_pixels": (None, [f64p, f32p, u8p, i64, C.c_float, C.c_float, f32p, C.c_uint32,
                            u32p, f32p, u32p, u64p]),
    "write_pixel_states": (None, [f64p, f32p, C.c_uint8, C.c_float, C.c_float, f32p, u32p, i64,
                                  u32p, u8p]),
    "emit": (C.c_int32, [f32p, i64, u16p]),
    "planes": (C.c_int32, [u16p, i64, i64, f32p, C.POINTER(DiceParams), f32p,
                           u8p, f32p, f32p, f32p, f32p, f32p, f32p, f32p, f32p]),
    "decision": (None, [f32p, i64, i64, C.POINTER(DiceParams), u8p]),
}

_LIB: list[C.CDLL] = []


class OracleError(RuntimeError):
    def __init__(self, code: int, message: str) -> None:
        super().__init__(f"oracle error {code}: {message}")
        self.code = code
        self.message = message


def build() -> None:
    subprocess.run(["make", "-s", "-C", str(ORACLE_DIR)], check=True)


def lib() -> C.CDLL:
    if not _LIB:
        if not ORACLE_SO.exists():
            build()
        handle = C.CDLL(str(ORACLE_SO))
        _abi.declare(handle, "dor_", ORACLE_SIGNATURES)
        _LIB.append(handle)
    return _LIB[0]


def _check(code: int) -> None:
    if code != 0:
        raise OracleError(code, lib().dor_last_error().decode())


@dataclass
class Replay:
    output: np.ndarray
    attempted_pixels: int
    written_pixels: int
    public_rng_advances: int
    final_rng_state: int
    changed_pixels: int
    startup_attempted: tuple[int, ...]
    startup_advances: tuple[int, ...]
    startup_final_rng_state: int
    score_plane: np.ndarray | None = None
    at_floor_mask: np.ndarray | None = None
    changed_mask: np.ndarray | None = None

    @property
    def output_sha256(self) -> str:
        return hashlib.sha256(self.output.astype("<u2").tobytes()).hexdigest()

    def counters(self) -> dict:
        return dict(
            attempted_pixels=self.attempted_pixels,
            written_pixels=self.written_pixels,
            public_rng_advances=self.public_rng_advances,
            final_rng_state=self.final_rng_state,
            changed_pixels=self.changed_pixels,
            startup_attempted=self.startup_attempted,
            startup_advances=self.startup_advances,
            startup_final_rng_state=self.startup_final_rng_state,
        )


def _replay_from(out, res, diag) -> Replay:
    return Replay(
        output=out,
        attempted_pixels=int(res.attempted_pixels),
        written_pixels=int(res.written_pixels),
        public_rng_advances=int(res.public_rng_advances),
        final_rng_state=int(res.final_rng_state),
        changed_pixels=int(res.changed_pixels),
        startup_attempted=tuple(int(v) for v in res.startup_attempted),
        startup_advances=tuple(int(v) for v in res.startup_advances),
        startup_final_rng_state=int(res.startup_final_rng_state),
        score_plane=None if diag is None else diag[0],
        at_floor_mask=None if diag is None else diag[1].astype(bool),
        changed_mask=None if diag is None else diag[2].astype(bool),
    )


def _diag_buffers(height: int, width: int):
    score = np.zeros((height, width), np.float32)
    at_floor = np.zeros((height, width), np.uint8)
    changed = np.zeros((height, width), np.uint8)
    d = DiceDiagnostics(ptr(score, C.c_float), ptr(at_floor, C.c_uint8),
                        ptr(changed, C.c_uint8))
    return (score, at_floor, changed), d


def lut() -> np.ndarray:
    table = np.empty(65536, np.float32)
    lib().dor_response_lut(ptr(table, C.c_float))
    return table


def process_job(main: np.ndarray, prepass: np.ndarray, params: DiceParams,
                prepass_params: DicePrepassParams, *, diagnostics: bool = False,
                threads: int = 0) -> Replay:
    main = np.ascontiguousarray(main, dtype=np.uint16)
    prepass = np.ascontiguousarray(prepass, dtype=np.uint16)
    height, width = main.shape[:2]
    out = np.zeros((height, width, 3), np.uint16)
    res = DiceReplayResult()
    diag = None
    dstruct = None
    if diagnostics:
        diag, dstruct = _diag_buffers(height, width)
    code = lib().dor_process_job(
        ptr(main, C.c_uint16), height, width, ptr(prepass, C.c_uint16),
        prepass.shape[0], prepass.shape[1], C.byref(params),
        C.byref(prepass_params), ptr(out, C.c_uint16), C.byref(res),
        C.byref(dstruct) if dstruct is not None else None, threads)
    _check(code)
    return _replay_from(out, res, diag)


def replay(main: np.ndarray, params: DiceParams, stages, *, rng_state: int = 0x3045,
           diagnostics: bool = False, threads: int = 0,
           table: np.ndarray | None = None) -> Replay:
    main = np.ascontiguousarray(main, dtype=np.uint16)
    height, width = main.shape[:2]
    table = lut() if table is None else table
    out = np.zeros((height, width, 3), np.uint16)
    res = DiceReplayResult()
    diag = None
    dstruct = None
    if diagnostics:
        diag, dstruct = _diag_buffers(height, width)
    st = stages.as_struct()
    code = lib().dor_replay(
        ptr(main, C.c_uint16), height, width, ptr(table, C.c_float),
        C.byref(params), C.byref(st), rng_state, ptr(out, C.c_uint16),
        C.byref(res), C.byref(dstruct) if dstruct is not None else None, threads)
    _check(code)
    return _replay_from(out, res, diag)


def producer_schedule(main: np.ndarray, table: np.ndarray | None = None) -> dict:
    main = np.ascontiguousarray(main, dtype=np.uint16)
    height, width = main.shape[:2]
    table = lut() if table is None else table
    arrays = new_schedule_arrays(height)
    sched = schedule_struct(arrays)
    _check(lib().dor_producer_schedule(ptr(main, C.c_uint16), height, width,
                                       ptr(table, C.c_float), C.byref(sched)))
    arrays["active_width"] = int(sched.active_width)
    return arrays


def new_schedule_arrays(rows: int) -> dict:
    return {
        "weighted_visible_sum": np.zeros(rows, np.float64),
        "weighted_infrared_sum": np.zeros(rows, np.float64),
        "weight_sum": np.zeros(rows, np.float64),
        "mean_visible": np.zeros(rows, np.float32),
        "mean_infrared": np.zeros(rows, np.float32),
        "accepted": np.zeros(rows, np.uint32),
        "scaled_input": np.zeros(rows, np.float32),
        "input_scale": np.zeros(rows, np.float32),
        "scale_denominator": np.zeros(rows, np.float32),
        "scale_numerator": np.zeros(rows, np.float32),
        "scale_add_denominator": np.zeros(rows, np.float32),
        "scale_add_numerator": np.zeros(rows, np.float32),
        "payloads": np.zeros((rows, 64), np.uint8),
    }


def schedule_struct(arrays: dict) -> DiceSchedule:
    s = DiceSchedule()
    s.rows = arrays["weight_sum"].shape[0]
    s.active_width = int(arrays.get("active_width", 0))
    for name, _ in DiceSchedule._fields_[2:]:
        array = arrays[name]
        ctype = np.ctypeslib.as_ctypes_type(array.dtype)
        setattr(s, name, ptr(array, ctype))
    return s


def prepass_params(values) -> DicePrepassParams:
    """DicePrepassParams from (valid_ir_threshold, alpha_scale, ratio_minimum,
    ratio_maximum, zero_denominator_alpha, completion_marker)."""

    params = DicePrepassParams()
    (params.valid_ir_threshold, params.alpha_scale, params.ratio_minimum,
     params.ratio_maximum, params.zero_denominator_alpha) = (float(v) for v in values[:5])
    params.completion_marker = int(values[5])
    return params


def prepass(pixels: np.ndarray, params: DicePrepassParams,
            table: np.ndarray | None = None) -> DicePrepassResult:
    pixels = np.ascontiguousarray(pixels, dtype=np.uint16)
    table = lut() if table is None else table
    res = DicePrepassResult()
    _check(lib().dor_prepass(ptr(pixels, C.c_uint16), pixels.shape[0],
                             pixels.shape[1], ptr(table, C.c_float),
                             C.byref(params), C.byref(res)))
    return res


def stage_table_from_schedule(arrays: dict, count: int, factor_b: float = 1.0):
    from dice_amd.stages import StageTable

    table = StageTable.empty(count)
    sched = schedule_struct(arrays)
    _check(lib().dor_stage_table_from_schedule(
        C.byref(sched), factor_b, count,
        ptr(table.aux_alpha, C.c_float), ptr(table.aux_offset, C.c_float),
        ptr(table.aux_replacement, C.c_float),
        ptr(table.aux_has_replacement, C.c_uint8),
        ptr(table.score_base, C.c_float), ptr(table.coarse_reference, C.c_float),
        ptr(table.gate_secondary, C.c_uint8), ptr(table.row_gate, C.c_int32)))
    return table
