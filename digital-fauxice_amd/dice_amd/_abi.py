"""ctypes mirror of ``include/dice_hip.h``.

The structures here are byte-for-byte the C declarations; ``declare()``
attaches argument and result types to every exported entry point of a
loaded library, given the symbol prefix (``dice_`` for the HIP library).
"""

from __future__ import annotations

import ctypes as C

ABI_VERSION = 1

DICE_OK = 0
DICE_UNAVAILABLE = 1
DICE_INVALID = 2
DICE_NONFINITE_AUX = 3
DICE_NONFINITE_WORK = 4
DICE_CANCELLED = 5
DICE_OVER_BUDGET = 6
DICE_INTERNAL = 7
DICE_PENDING = 8
DICE_NOALLOC_SCRATCH_BYTES = 512 << 20

DICE_MAIN_ON_DEVICE = 0x1
DICE_OUT_ON_DEVICE = 0x2
DICE_SKIP_SHA256 = 0x4
DICE_BATCH_GRAPH = 0x100
DICE_BATCH_AUTOTUNE = 0x200
DICE_BATCH_DEVICE_CHAIN = 0x400

f32p = C.POINTER(C.c_float)
f64p = C.POINTER(C.c_double)
u8p = C.POINTER(C.c_uint8)
u16p = C.POINTER(C.c_uint16)
u32p = C.POINTER(C.c_uint32)
u64p = C.POINTER(C.c_uint64)
i32p = C.POINTER(C.c_int32)
i64 = C.c_int64


class DiceParams(C.Structure):
    _fields_ = [
        ("visible_channel", C.c_int32),
        ("score_addend", C.c_float),
        ("score_scale", C.c_float),
        ("score_offset", C.c_float),
        ("score_floor", C.c_float),
        ("horizontal_minimum", C.c_int32),
        ("decision_threshold", C.c_float),
        ("decision_count_limit", C.c_int32),
        ("decision_radius", C.c_int32),
        ("cross_neighbor", C.c_int32),
        ("coarse_enabled", C.c_int32),
        ("coarse_slopes", C.c_float * 3),
        ("band_enabled", C.c_int32 * 3),
        ("band_scales", C.c_float * 3),
        ("factors_a", C.c_float * 9),
        ("factors_b", C.c_float * 9),
        ("configured_strengths", C.c_float * 3),
        ("gate_primary", C.c_int32),
        ("dither_scales", C.c_float * 3),
        ("dither_low", C.c_float),
        ("dither_high", C.c_float),
    ]


class DiceStageTable(C.Structure):
    _fields_ = [
        ("count", C.c_int64),
        ("aux_alpha", f32p),
        ("aux_offset", f32p),
        ("aux_replacement", f32p),
        ("aux_has_replacement", u8p),
        ("score_base", f32p),
        ("coarse_reference", f32p),
        ("gate_secondary", u8p),
        ("row_gate", i32p),
    ]


class DiceReplayResult(C.Structure):
    _fields_ = [
        ("attempted_pixels", C.c_uint64),
        ("written_pixels", C.c_uint64),
        ("public_rng_advances", C.c_uint64),
        ("changed_pixels", C.c_uint64),
        ("final_rng_state", C.c_uint32),
        ("startup_final_rng_state", C.c_uint32),
        ("startup_attempted", C.c_int64 * 6),
        ("startup_advances", C.c_int64 * 6),
        ("output_sha256", C.c_char * 72),
    ]


class DiceDiagnostics(C.Structure):
    _fields_ = [
        ("score_plane", f32p),
        ("at_floor_mask", u8p),
        ("changed_mask", u8p),
    ]


class DiceSchedule(C.Structure):
    _fields_ = [
        ("rows", C.c_int64),
        ("active_width", C.c_int64),
        ("weighted_visible_sum", f64p),
        ("weighted_infrared_sum", f64p),
        ("weight_sum", f64p),
        ("mean_visible", f32p),
        ("mean_infrared", f32p),
        ("accepted", u32p),
        ("scaled_input", f32p),
        ("input_scale", f32p),
        ("scale_denominator", f32p),
        ("scale_numerator", f32p),
        ("scale_add_denominator", f32p),
        ("scale_add_numerator", f32p),
        ("payloads", u8p),
    ]


class DiceFpCanary(C.Structure):
    _fields_ = [
        ("count", C.c_int64),
        ("a", C.POINTER(C.c_double)),
        ("b", C.POINTER(C.c_double)),
        ("c", C.POINTER(C.c_double)),
        ("fa", C.POINTER(C.c_float)),
        ("fb", C.POINTER(C.c_float)),
        ("fc", C.POINTER(C.c_float)),
        ("mul_add", C.POINTER(C.c_double)),
        ("quotient", C.POINTER(C.c_double)),
        ("narrowed", C.POINTER(C.c_float)),
        ("f32_mul_add", C.POINTER(C.c_float)),
        ("f32_quotient", C.POINTER(C.c_float)),
        ("f32_product", C.POINTER(C.c_float)),
    ]


class DicePrepassParams(C.Structure):
    _fields_ = [
        ("valid_ir_threshold", C.c_float),
        ("alpha_scale", C.c_float),
        ("ratio_minimum", C.c_float),
        ("ratio_maximum", C.c_float),
        ("zero_denominator_alpha", C.c_float),
        ("completion_marker", C.c_uint32),
    ]


class DicePrepassResult(C.Structure):
    _fields_ = [
        ("payload", C.c_uint8 * 64),
        ("valid_pixels", C.c_int64),
        ("fully_valid_tiles", C.c_int64),
    ]


ProgressFn = C.CFUNCTYPE(None, C.c_void_p, C.c_int32, C.c_int64, C.c_int64,
                         C.c_int64, C.c_int64)
CancelledFn = C.CFUNCTYPE(C.c_int32, C.c_void_p)


class DiceCallbacks(C.Structure):
    _fields_ = [
        ("progress", ProgressFn),
        ("cancelled", CancelledFn),
        ("user", C.c_void_p),
    ]


class DiceStageTimings(C.Structure):
    _fields_ = [
        ("prepass", C.c_double),
        ("producer_device", C.c_double),
        ("upload", C.c_double),
        ("analysis_device", C.c_double),
        ("download_sites", C.c_double),
        ("writer_host", C.c_double),
        ("finish_device", C.c_double),
        ("download_output", C.c_double),
        ("sha256_host", C.c_double),
        ("total", C.c_double),
        ("producer_kernels_ms", C.c_double),
        ("scan_kernel_ms", C.c_double),
        ("sites_kernel_ms", C.c_double),
        ("startup_kernels_ms", C.c_double),
        ("segments_kernel_ms", C.c_double),
        ("writer_prep_kernel_ms", C.c_double),
        ("writer_values_kernel_ms", C.c_double),
    ]

    def as_dict(self) -> dict[str, float]:
        return {name: float(getattr(self, name)) for name, _ in self._fields_}


class DiceJob(C.Structure):
    _fields_ = [
        ("main_rgbi", C.c_void_p),
        ("height", C.c_int64),
        ("width", C.c_int64),
        ("prepass_rgbi", C.c_void_p),
        ("prepass_height", C.c_int64),
        ("prepass_width", C.c_int64),
        ("params", C.POINTER(DiceParams)),
        ("prepass_params", C.POINTER(DicePrepassParams)),
        ("flags", C.c_uint32),
        ("out_rgb16", C.c_void_p),
        ("result", C.POINTER(DiceReplayResult)),
        ("timings", C.POINTER(DiceStageTimings)),
        ("status", C.c_int32),
        ("error", C.c_char * 256),
    ]


class DiceBatchReport(C.Structure):
    _fields_ = [
        ("tile_rows", C.c_int32),
        ("tile_cols", C.c_int32),
        ("graph_replayed", C.c_int32),
        ("phase_a_ms", C.c_double),
        ("phase_b_ms", C.c_double),
        ("workspace_bytes", C.c_uint64),
        ("wall_seconds", C.c_double),
        ("sites_variant", C.c_int32),
        ("tune_source", C.c_int32),
        ("tune_candidates", C.c_int32),
        ("tune_rejected", C.c_int32),
    ]


class DiceEngineConfig(C.Structure):
    _fields_ = [
        ("device", C.c_int32),
        ("drivers", C.c_int32),
        ("max_batch", C.c_int32),
        ("batch_flags", C.c_uint32),
        ("hbm_budget", C.c_uint64),
        ("linger_ms", C.c_double),
    ]


class DiceEngineStats(C.Structure):
    _fields_ = [
        ("submitted", C.c_uint64),
        ("completed", C.c_uint64),
        ("failed", C.c_uint64),
        ("queued", C.c_uint64),
        ("running", C.c_uint64),
        ("batches", C.c_uint64),
        ("mean_batch", C.c_double),
        ("mpix_completed", C.c_double),
        ("latency_p50_ms", C.c_double),
        ("latency_p99_ms", C.c_double),
        ("latency_max_ms", C.c_double),
        ("queue_p50_ms", C.c_double),
        ("admitted_bytes", C.c_uint64),
        ("admitted_bytes_peak", C.c_uint64),
        ("hbm_budget", C.c_uint64),
        ("workspace_bytes", C.c_uint64),
        ("elapsed_seconds", C.c_double),
        ("workspace_bytes_peak", C.c_uint64),
        ("workspace_releases", C.c_uint64),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}


class DiceDeviceInfo(C.Structure):
    _fields_ = [
        ("name", C.c_char * 128),
        ("arch", C.c_char * 64),
        ("compute_units", C.c_int32),
        ("device_count", C.c_int32),
        ("total_hbm_bytes", C.c_uint64),
        ("free_hbm_bytes", C.c_uint64),
        ("build_flags", C.c_char * 256),
        ("reason", C.c_char * 256),
    ]


class DiceRoutingSummary(C.Structure):
    _fields_ = [
        ("height", C.c_int64),
        ("width", C.c_int64),
        ("provisional_regions", C.c_int64),
        ("final_regions", C.c_int64),
        ("pairs", C.c_int64),
        ("at_floor_pixels", C.c_int64),
        ("synthesis_pixels", C.c_int64),
        ("non_floor_synthesis_pixels", C.c_int64),
        ("perimeter_excluded_id", C.c_int64),
        ("perimeter_suppressed_halo_pixels", C.c_int64),
    ]


class DiceProvisionalRegion(C.Structure):
    _fields_ = [
        ("y0", C.c_int32), ("x0", C.c_int32), ("y1", C.c_int32), ("x1", C.c_int32),
        ("area", C.c_int64),
        ("max_chessboard_radius", C.c_int64),
        ("synthesized_pixels", C.c_int64),
        ("routed", C.c_int32),
        ("reserved", C.c_int32),
    ]


class DiceFinalRegion(C.Structure):
    _fields_ = [
        ("y0", C.c_int32), ("x0", C.c_int32), ("y1", C.c_int32), ("x1", C.c_int32),
        ("area", C.c_int64),
        ("at_floor_pixels", C.c_int64),
    ]


DICE_ROUTING_MASK_ON_DEVICE = 0x1


class DiceStreamProgress(C.Structure):
    _fields_ = [
        ("rows_received", C.c_int64),
        ("rows_analyzed", C.c_int64),
        ("rows_final", C.c_int64),
        ("sites_written", C.c_int64),
        ("windows", C.c_int64),
        ("push_seconds", C.c_double),
    ]

    def as_dict(self) -> dict:
        return {name: getattr(self, name) for name, _ in self._fields_}

VOIDP = C.c_void_p

# Signatures of the HIP library entry points (symbol prefix "dice_").
HIP_SIGNATURES: dict[str, tuple[object, list[object]]] = {
    "abi_version": (C.c_int32, []),
    "last_error": (C.c_char_p, []),
    "probe": (C.c_int32, [C.POINTER(DiceDeviceInfo)]),
    "ctx_create": (C.c_int32, [C.c_int32, C.c_uint64, C.POINTER(VOIDP)]),
    "ctx_destroy": (None, [VOIDP]),
    "response_lut": (C.c_int32, [f32p]),
    "inverse_factors": (C.c_int32, [u32p, u32p]),
    "prepass": (C.c_int32, [u16p, i64, i64, f32p, C.POINTER(DicePrepassParams),
                            C.POINTER(DicePrepassResult)]),
    "producer_schedule": (C.c_int32, [VOIDP, u16p, i64, i64, f32p,
                                      C.POINTER(DiceSchedule)]),
    "stage_table_from_schedule": (
        C.c_int32,
        [C.POINTER(DiceSchedule), C.c_float, i64, f32p, f32p, f32p, u8p, f32p,
         f32p, u8p, i32p],
    ),
    "replay": (
        C.c_int32,
        [VOIDP, VOIDP, i64, i64, f32p, C.POINTER(DiceParams),
         C.POINTER(DiceStageTable), C.c_uint32, C.c_uint32, VOIDP,
         C.POINTER(DiceReplayResult), C.POINTER(DiceDiagnostics),
         C.POINTER(DiceCallbacks), C.POINTER(DiceStageTimings)],
    ),
    "process_job": (
        C.c_int32,
        [VOIDP, VOIDP, i64, i64, u16p, i64, i64, C.POINTER(DiceParams),
         C.POINTER(DicePrepassParams), C.c_uint32, VOIDP,
         C.POINTER(DiceReplayResult), C.POINTER(DiceDiagnostics),
         C.POINTER(DiceCallbacks), C.POINTER(DiceStageTimings)],
    ),
    "process_job_noalloc": (
        C.c_int32,
        [VOIDP, VOIDP, i64, i64, u16p, i64, i64, C.POINTER(DiceParams),
         C.POINTER(DicePrepassParams), VOIDP, C.POINTER(DiceReplayResult)],
    ),
    "process_batch": (
        C.c_int32,
        [VOIDP, C.POINTER(DiceJob), C.c_int32, C.c_uint32, C.POINTER(DiceBatchReport)],
    ),
    "engine_create": (C.c_int32, [C.POINTER(DiceEngineConfig), C.POINTER(VOIDP)]),
    "engine_destroy": (None, [VOIDP]),
    "engine_submit": (C.c_int32, [VOIDP, C.POINTER(DiceJob), C.POINTER(C.c_uint64)]),
    "engine_wait": (C.c_int32, [VOIDP, C.c_uint64, C.c_double]),
    "tuning_cache": (C.c_int32, [C.c_char_p]),
    "engine_reservation": (C.c_int32, [C.POINTER(DiceJob), C.POINTER(C.c_uint64)]),
    "engine_stats": (C.c_int32, [VOIDP, C.POINTER(DiceEngineStats)]),
    "engine_reset_stats": (C.c_int32, [VOIDP]),
    "t_unscaled_averages": (C.c_int32, [VOIDP, f32p, i64, f64p]),
    "t_rgb_unscaled_averages": (C.c_int32, [VOIDP, f32p, i64, f64p]),
    "t_dither_chain": (
        C.c_int32,
        [VOIDP, f64p, f32p, i64, C.c_uint32, C.c_float, C.c_float, f64p, u32p,
         u64p],
    ),
    "t_emit": (C.c_int32, [VOIDP, f32p, i64, u16p]),
    "t_planes": (
        C.c_int32,
        [VOIDP, u16p, i64, i64, f32p, C.POINTER(DiceParams), f32p, u8p, f32p,
         f32p, f32p, f32p, f32p, f32p, f32p, f32p, u32p],
    ),
    "t_decision": (C.c_int32, [VOIDP, f32p, i64, i64, C.POINTER(DiceParams),
                               u8p]),
    "t_copy": (C.c_int32, [VOIDP, C.c_uint64, C.c_int32, f64p]),
    "t_fp_canaries": (C.c_int32, [VOIDP, C.POINTER(DiceFpCanary)]),
    "t_writer_chain_host": (C.c_int32, [f64p, f32p, u8p, i64, C.c_float, C.c_float, f32p,
                                        C.c_uint32, C.c_int32, u32p, u32p, u64p]),
    "route_at_floor": (C.c_int32, [VOIDP, VOIDP, i64, i64, C.c_uint32, i64, i64, i64,
                                   C.POINTER(VOIDP)]),
    "routing_summary_get": (C.c_int32, [VOIDP, C.POINTER(DiceRoutingSummary)]),
    "routing_tables": (C.c_int32, [VOIDP, C.POINTER(DiceProvisionalRegion),
                                   C.POINTER(DiceFinalRegion), i32p]),
    "routing_planes": (C.c_int32, [VOIDP, i32p, u8p, i32p]),
    "routing_free": (None, [VOIDP]),
    "stream_create": (C.c_int32, [VOIDP, C.POINTER(VOIDP)]),
    "stream_begin": (C.c_int32, [VOIDP, C.POINTER(DiceJob), C.POINTER(DiceDiagnostics)]),
    "stream_push": (C.c_int32, [VOIDP, u16p, i64, i64]),
    "stream_finish": (C.c_int32, [VOIDP]),
    "stream_progress_get": (C.c_int32, [VOIDP, C.POINTER(DiceStreamProgress)]),
    "stream_destroy": (None, [VOIDP]),
    "t_writer_site_host": (C.c_int32, [f64p, f32p, C.c_uint8, C.c_float, C.c_float, f32p, u32p,
                                       i64, C.c_int32, u32p, u8p]),
    "t_sha256_pieces": (C.c_int32, [C.POINTER(C.c_void_p), C.POINTER(C.c_int64), C.c_int32,
                                    i64, C.c_int32, C.c_char_p, C.POINTER(C.c_int32)]),
    "t_hash_pairs": (C.c_int32, [C.c_int32, C.POINTER(C.c_int32)]),
}


def declare(lib: C.CDLL, prefix: str, signatures: dict) -> None:
    """Attach restype/argtypes to ``prefix + name`` for every signature."""

    for name, (restype, argtypes) in signatures.items():
        function = getattr(lib, prefix + name)
        function.restype = restype
        function.argtypes = argtypes


def ptr(array, ctype):
    """Pointer to a C-contiguous numpy array (or None)."""

    if array is None:
        return None
    if not array.flags.c_contiguous:
        raise ValueError("array must be C-contiguous")
    return array.ctypes.data_as(C.POINTER(ctype))
