"""Result and progress types of a processing job.

Mirrors ``src/portable_digital_ice/engine.py:22-78`` (phases, progress,
diagnostics, result, cancellation) and the replay receipts
``streaming.py:63-74`` / ``startup.py:37-43`` of rohanpandula/digital-fauxice.
"""

from __future__ import annotations

from collections.abc import Callable
from dataclasses import dataclass
from enum import Enum

import numpy as np
import numpy.typing as npt


class ProcessingPhase(str, Enum):
    VALIDATING = "validating"
    PREPASS = "prepass"
    PRODUCER = "producer"
    RECONSTRUCTION = "reconstruction"
    COMPLETE = "complete"


PHASES = (
    ProcessingPhase.VALIDATING,
    ProcessingPhase.PREPASS,
    ProcessingPhase.PRODUCER,
    ProcessingPhase.RECONSTRUCTION,
    ProcessingPhase.COMPLETE,
)


@dataclass(frozen=True)
class ProcessingProgress:
    phase: ProcessingPhase
    completed: int
    total: int
    attempted_pixels: int = 0
    written_pixels: int = 0


@dataclass(frozen=True)
class ProcessingDiagnostics:
    score_plane: npt.NDArray[np.float32]
    score_floor: np.float32
    at_floor_mask: npt.NDArray[np.bool_]
    changed_mask: npt.NDArray[np.bool_]

    def __post_init__(self) -> None:
        if self.score_plane.dtype != np.dtype(np.float32):
            raise ValueError("diagnostic score plane must be float32")
        if self.score_plane.ndim != 2:
            raise ValueError("diagnostic score plane must be HxW")
        floor = np.asarray(self.score_floor)
        if floor.dtype != np.dtype(np.float32) or floor.ndim != 0:
            raise ValueError("diagnostic score floor must be a float32 scalar")
        if not np.isfinite(floor) or floor <= np.float32(0.0):
            raise ValueError("diagnostic score floor must be finite and positive")
        for name, mask in (("at-floor", self.at_floor_mask), ("changed", self.changed_mask)):
            if mask.dtype != np.dtype(np.bool_):
                raise ValueError(f"diagnostic {name} mask must be bool")
            if mask.shape != self.score_plane.shape:
                raise ValueError(f"diagnostic {name} mask must have shape {self.score_plane.shape}")


@dataclass(frozen=True)
class StartupReplayResult:
    attempted_per_stage: tuple[int, int, int, int, int, int]
    rng_advances_per_stage: tuple[int, int, int, int, int, int]
    final_rng_state: int


@dataclass(frozen=True)
class StreamingReplayResult:
    shape: tuple[int, int, int]
    startup: StartupReplayResult | None
    attempted_pixels: int
    written_pixels: int
    public_rng_advances: int
    final_rng_state: int
    output_sha256: str
    changed_pixels: int


@dataclass(frozen=True)
class ProcessingResult:
    output_rgb16: npt.NDArray[np.uint16]
    replay: StreamingReplayResult
    profile_id: str
    diagnostics: ProcessingDiagnostics | None = None


class ProcessingCancelled(RuntimeError):
    """Raised when a caller requests cooperative cancellation."""


ProgressCallback = Callable[[ProcessingProgress], None]
CancellationCallback = Callable[[], bool]


@dataclass
class LCG24:
    """rng.py:17-51 -- the per-job 24-bit generator state owner."""

    state: int

    def __post_init__(self) -> None:
        if not 0 <= self.state <= 0x00FF_FFFF:
            raise ValueError("LCG24 state must fit in 24 bits")

    @classmethod
    def from_nikon_pe_initial_state(cls) -> "LCG24":
        return cls(0x3045)

    def advance(self) -> int:
        self.state = (125 * self.state + 1) & 0x00FF_FFFF
        return self.state
