"""HIP backend for MI355X (gfx950): ctypes bindings over libdice_hip.so."""

from .engine import (
    HipBackendUnavailable,
    hip_device_summary,
    run_streaming_replay_hip,
)
from .process import process_hip
from .producer import derive_producer_record_schedule_hip

__all__ = [
    "HipBackendUnavailable",
    "derive_producer_record_schedule_hip",
    "hip_device_summary",
    "process_hip",
    "run_streaming_replay_hip",
]
