"""Typed acquisition contracts (mirror of the reference's ``contracts.py``).

Same names, argument meaning and error behaviour as
``src/portable_digital_ice/contracts.py:25-74, 108-176, 224-261`` of
rohanpandula/digital-fauxice: frames and blocks are copied, frozen and
validated on construction; a frame travels as valid-row-only blocks with an
absolute start row, and joins back only when its blocks are ordered,
gap-free, non-overlapping and of one width.
"""

from __future__ import annotations

import hashlib
from dataclasses import dataclass
from enum import Enum
from typing import Iterable

import numpy as np
import numpy.typing as npt

UInt16Array = npt.NDArray[np.uint16]


class AcquisitionEpoch(str, Enum):
    PREPASS = "prepass"
    MAIN = "main"


def _frozen_uint16_image(pixels: npt.ArrayLike, *, lanes: int, label: str) -> UInt16Array:
    """contracts.py:25-40"""

    array = np.asarray(pixels)
    if array.dtype != np.dtype(np.uint16):
        raise TypeError(f"{label} must have dtype uint16")
    if array.ndim != 3 or array.shape[2] != lanes:
        raise ValueError(f"{label} must have shape HxWx{lanes}")
    if array.shape[0] < 1 or array.shape[1] < 1:
        raise ValueError(f"{label} cannot have an empty dimension")
    owned = np.array(array, dtype=np.uint16, order="C", copy=True)
    owned.flags.writeable = False
    return owned


@dataclass(frozen=True)
class RGBI16Frame:
    """contracts.py:44-74"""

    pixels: UInt16Array
    epoch: AcquisitionEpoch
    resolution_dpi: int
    evidence_id: str

    def __post_init__(self) -> None:
        if self.resolution_dpi <= 0:
            raise ValueError("resolution_dpi must be positive")
        if not self.evidence_id.strip():
            raise ValueError("evidence_id must be non-empty")
        object.__setattr__(
            self, "pixels", _frozen_uint16_image(self.pixels, lanes=4, label="RGBI frame")
        )

    @property
    def height(self) -> int:
        return int(self.pixels.shape[0])

    @property
    def width(self) -> int:
        return int(self.pixels.shape[1])

    @property
    def sha256(self) -> str:
        return hashlib.sha256(self.pixels.astype("<u2", copy=False).tobytes()).hexdigest()


@dataclass(frozen=True)
class DualRGBIAcquisition:
    """contracts.py:162-176"""

    prepass: RGBI16Frame
    main: RGBI16Frame
    same_frame_id: str

    def __post_init__(self) -> None:
        if self.prepass.epoch is not AcquisitionEpoch.PREPASS:
            raise ValueError("prepass frame has the wrong acquisition epoch")
        if self.main.epoch is not AcquisitionEpoch.MAIN:
            raise ValueError("main frame has the wrong acquisition epoch")
        if not self.same_frame_id.strip():
            raise ValueError("same_frame_id must be non-empty")


@dataclass(frozen=True)
class RGBI16Block:
    """contracts.py:108-130: rows of a frame in transport, with the absolute
    index of their first row."""

    pixels: UInt16Array
    start_row: int

    def __post_init__(self) -> None:
        if self.start_row < 0:
            raise ValueError("start_row cannot be negative")
        object.__setattr__(
            self, "pixels", _frozen_uint16_image(self.pixels, lanes=4, label="RGBI block")
        )

    @property
    def valid_rows(self) -> int:
        return int(self.pixels.shape[0])

    @property
    def width(self) -> int:
        return int(self.pixels.shape[1])


@dataclass(frozen=True)
class BlockSchedule:
    """contracts.py:133-158: the allocated block geometry of a frame; every
    block holds ``rows_per_block`` rows except the last, which holds
    ``final_valid_rows``."""

    width: int
    height: int
    rows_per_block: int
    block_count: int
    final_valid_rows: int

    def __post_init__(self) -> None:
        if min(self.width, self.height, self.rows_per_block, self.block_count) <= 0:
            raise ValueError("schedule dimensions must be positive")
        if not 1 <= self.final_valid_rows <= self.rows_per_block:
            raise ValueError("final_valid_rows is outside the allocated block")
        if (self.block_count - 1) * self.rows_per_block + self.final_valid_rows != self.height:
            raise ValueError("block schedule does not cover exactly height rows")

    @classmethod
    def for_frame(cls, height: int, width: int, rows_per_block: int = 8) -> "BlockSchedule":
        """The schedule that cuts ``height`` rows into blocks of
        ``rows_per_block`` (the scanner's 8 on the supported profile)."""

        if rows_per_block <= 0 or height <= 0:
            raise ValueError("schedule dimensions must be positive")
        count = -(-height // rows_per_block)
        return cls(width, height, rows_per_block, count, height - (count - 1) * rows_per_block)

    def valid_rows_for(self, block_index: int) -> int:
        if not 0 <= block_index < self.block_count:
            raise IndexError(block_index)
        if block_index == self.block_count - 1:
            return self.final_valid_rows
        return self.rows_per_block


def split_rgbi16_frame(frame: RGBI16Frame, rows_per_block: int) -> tuple[RGBI16Block, ...]:
    """contracts.py:224-234: consecutive valid-row blocks, no padding."""

    if rows_per_block <= 0:
        raise ValueError("rows_per_block must be positive")
    pixels = frame.pixels
    return tuple(
        RGBI16Block(pixels[row:row + rows_per_block], row)
        for row in range(0, frame.height, rows_per_block)
    )


def assemble_rgbi16_blocks(
    blocks: Iterable[RGBI16Block],
    *,
    epoch: AcquisitionEpoch,
    resolution_dpi: int,
    evidence_id: str,
) -> RGBI16Frame:
    """contracts.py:237-261: the frame of an ordered, gap-free row stream."""

    blocks = tuple(blocks)
    if not blocks:
        raise ValueError("at least one RGBI block is required")
    width = blocks[0].width
    next_row = 0
    for block in blocks:
        if block.start_row != next_row:
            raise ValueError("RGBI blocks must be ordered, gap-free, and non-overlapping")
        if block.width != width:
            raise ValueError("all RGBI blocks must have the same width")
        next_row += block.valid_rows
    return RGBI16Frame(
        np.concatenate([block.pixels for block in blocks], axis=0), epoch, resolution_dpi,
        evidence_id,
    )
