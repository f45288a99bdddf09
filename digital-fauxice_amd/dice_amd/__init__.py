"""MI355X HIP backend for the portable Digital ICE processing core.

A drop-in beside the reference's cpu / cpu-fast / cuda / metal backends:
the public surface mirrors ``portable_digital_ice`` (contracts, profile,
processing results, ``process(job, backend=...)``) with ``ComputeBackend.HIP``.
"""

from .contracts import AcquisitionEpoch, DualRGBIAcquisition, RGBI16Frame
from .engine import (
    LCG24,
    ProcessingCancelled,
    ProcessingDiagnostics,
    ProcessingPhase,
    ProcessingProgress,
    ProcessingResult,
    StartupReplayResult,
    StreamingReplayResult,
)
from .profile import DEFAULT_PROFILE, ProcessingJob, ProcessingMode, ScannerModel

__all__ = [
    "AcquisitionEpoch",
    "DEFAULT_PROFILE",
    "DualRGBIAcquisition",
    "LCG24",
    "ProcessingCancelled",
    "ProcessingDiagnostics",
    "ProcessingJob",
    "ProcessingMode",
    "ProcessingPhase",
    "ProcessingProgress",
    "ProcessingResult",
    "RGBI16Frame",
    "ScannerModel",
    "StartupReplayResult",
    "StreamingReplayResult",
]
