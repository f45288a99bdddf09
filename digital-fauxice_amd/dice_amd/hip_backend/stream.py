"""``HipFrameStream`` -- one frame processed while its main pass arrives.

The reference replays a frame once all of it is in memory
(``streaming.py:365-552``); its scanner transport delivers the main pass as
valid-row-only blocks with an absolute start row (``contracts.py:108-158``,
``RGBI16Block`` / ``BlockSchedule``; 8 rows per block on the LS-5000
profile, a shorter final block).  A stream takes those blocks in order and
runs every phase whose inputs are complete as each one arrives (producer
blocks, the cumulative chains, stage entries, analysis tile rows, sites,
the serial writer over the new sites, the copy and hash of the rows that are
final), so only the last rows are left when the last block lands.  The
result is the one ``process_hip`` gives for the assembled frame, byte for
byte.

The block rules are ``assemble_rgbi16_blocks``'s (``contracts.py:237-261``):
ordered, gap-free, non-overlapping, one width; here each block must also
have the row count the schedule allocates it.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _abi
from .._abi import (DiceDiagnostics, DiceJob, DiceReplayResult, DiceStageTimings,
                    DiceStreamProgress, ptr)
from ..contracts import AcquisitionEpoch, BlockSchedule, RGBI16Block, RGBI16Frame
from ..engine import ProcessingDiagnostics, ProcessingResult
from ..profile import DEFAULT_PROFILE
from .batch import _exception_for
from .engine import _result_from, context, library
from .process import stable_profile_params


class HipFrameStream:
    """A reusable stream: ``begin`` a frame, ``push`` its blocks, ``finish``."""

    def __init__(self, device: int = 0):
        self._lib = library()
        handle = C.c_void_p()
        self._check(self._lib.dice_stream_create(context(device), C.byref(handle)))
        self._handle = handle
        self._frame = None

    def _check(self, code: int) -> None:
        if code != _abi.DICE_OK:
            raise _exception_for(code, self._lib.dice_last_error().decode(errors="replace"))

    def close(self) -> None:
        if self._handle:
            self._lib.dice_stream_destroy(self._handle)
            self._handle = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc) -> None:
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def begin(self, prepass: RGBI16Frame, schedule: BlockSchedule, *,
              export_diagnostics: bool = False) -> None:
        """Start a frame of ``schedule``'s geometry whose prepass is complete
        (profile.py:75-106's checks on what is known before the main pass)."""

        if not isinstance(prepass, RGBI16Frame):
            raise TypeError("the prepass must be an RGBI16Frame")
        if prepass.epoch is not AcquisitionEpoch.PREPASS:
            raise ValueError("prepass frame has the wrong acquisition epoch")
        if prepass.resolution_dpi != DEFAULT_PROFILE.prepass_dpi:
            raise ValueError(f"prepass must be captured at {DEFAULT_PROFILE.prepass_dpi} dpi")
        if not isinstance(schedule, BlockSchedule):
            raise TypeError("a stream needs the frame's BlockSchedule")
        if schedule.height < 5 or schedule.width < 8:
            raise ValueError("main RGBI frame must be at least 5 rows by 8 columns")
        height, width = schedule.height, schedule.width
        frame = {
            "schedule": schedule,
            "next_block": 0,
            "prepass": np.ascontiguousarray(prepass.pixels),
            "params": stable_profile_params(),
            "prepass_params": DEFAULT_PROFILE.prepass_parameters(),
            "out": np.empty((height, width, 3), np.uint16),
            "result": DiceReplayResult(),
            "timings": DiceStageTimings(),
            "diag": None,
        }
        if export_diagnostics:
            planes = (np.empty((height, width), np.float32), np.empty((height, width), np.uint8),
                      np.empty((height, width), np.uint8))
            frame["planes"] = planes
            frame["diag"] = DiceDiagnostics(ptr(planes[0], C.c_float), ptr(planes[1], C.c_uint8),
                                            ptr(planes[2], C.c_uint8))
        job = DiceJob()
        job.main_rgbi = None
        job.height = height
        job.width = width
        job.prepass_rgbi = frame["prepass"].ctypes.data
        job.prepass_height = frame["prepass"].shape[0]
        job.prepass_width = frame["prepass"].shape[1]
        job.params = C.pointer(frame["params"])
        job.prepass_params = C.pointer(frame["prepass_params"])
        job.flags = 0
        job.out_rgb16 = frame["out"].ctypes.data
        job.result = C.pointer(frame["result"])
        job.timings = C.pointer(frame["timings"])
        frame["job"] = job
        self._frame = frame
        diag = C.byref(frame["diag"]) if frame["diag"] is not None else None
        self._check(self._lib.dice_stream_begin(self._handle, C.byref(job), diag))

    def push(self, block: RGBI16Block) -> dict:
        """Hand over the next block; returns the stream's progress."""

        frame = self._frame
        if frame is None:
            raise ValueError("the stream has no frame in progress")
        if not isinstance(block, RGBI16Block):
            raise TypeError("a stream takes RGBI16Block rows")
        schedule = frame["schedule"]
        index = frame["next_block"]
        if index >= schedule.block_count:
            raise ValueError("every block of the schedule has arrived")
        if block.width != schedule.width:
            raise ValueError("all RGBI blocks must have the same width")
        if block.start_row != index * schedule.rows_per_block:
            raise ValueError("RGBI blocks must be ordered, gap-free, and non-overlapping")
        if block.valid_rows != schedule.valid_rows_for(index):
            raise ValueError(f"block {index} must hold {schedule.valid_rows_for(index)} rows")
        pixels = np.ascontiguousarray(block.pixels)
        self._check(self._lib.dice_stream_push(self._handle, ptr(pixels, C.c_uint16),
                                               block.start_row, block.valid_rows))
        frame["next_block"] = index + 1
        return self.progress()

    def progress(self) -> dict:
        report = DiceStreamProgress()
        self._check(self._lib.dice_stream_progress_get(self._handle, C.byref(report)))
        return report.as_dict()

    def finish(self) -> ProcessingResult:
        """The frame's result once every block has arrived."""

        frame = self._frame
        if frame is None:
            raise ValueError("the stream has no frame in progress")
        self._frame = None
        self._check(self._lib.dice_stream_finish(self._handle))
        schedule = frame["schedule"]
        diagnostics = None
        if frame["diag"] is not None:
            score, at_floor, changed = frame["planes"]
            score.flags.writeable = False
            floor_mask = at_floor.astype(bool)
            changed_mask = changed.astype(bool)
            floor_mask.flags.writeable = False
            changed_mask.flags.writeable = False
            diagnostics = ProcessingDiagnostics(
                score_plane=score,
                score_floor=np.float32(frame["params"].score_floor),
                at_floor_mask=floor_mask,
                changed_mask=changed_mask,
            )
        return ProcessingResult(
            output_rgb16=frame["out"],
            replay=_result_from(frame["result"], schedule.height, schedule.width),
            profile_id=DEFAULT_PROFILE.profile_id,
            diagnostics=diagnostics,
        )
