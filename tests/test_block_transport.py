"""Block transport contracts (contracts.py:108-158, 224-261 of the reference).

Known answer: a 5782-row frame arrives as 723 allocated 8-row blocks whose
last one holds 6 valid rows (docs/reverse-engineering.md:81-88).  The split
must round-trip through the assembly, and the assembly must refuse what the
reference refuses: reordered, gapped or overlapping blocks and mixed widths.
"""

from __future__ import annotations

import numpy as np
import pytest

from dice_amd import AcquisitionEpoch, RGBI16Frame
from dice_amd.contracts import (
    BlockSchedule,
    RGBI16Block,
    assemble_rgbi16_blocks,
    split_rgbi16_frame,
)


def test_full_frame_schedule_known_answer():
    schedule = BlockSchedule.for_frame(5782, 3946)
    assert schedule.block_count == 723
    assert schedule.final_valid_rows == 6
    assert schedule.valid_rows_for(0) == 8
    assert schedule.valid_rows_for(722) == 6
    with pytest.raises(IndexError):
        schedule.valid_rows_for(723)
    exact = BlockSchedule.for_frame(64, 9)
    assert (exact.block_count, exact.final_valid_rows) == (8, 8)


def test_schedule_validation():
    with pytest.raises(ValueError, match="positive"):
        BlockSchedule(0, 8, 8, 1, 8)
    with pytest.raises(ValueError, match="outside the allocated block"):
        BlockSchedule(4, 16, 8, 2, 9)
    with pytest.raises(ValueError, match="exactly height rows"):
        BlockSchedule(4, 17, 8, 2, 8)


def frame(height: int, width: int, seed: int = 0) -> RGBI16Frame:
    pixels = np.random.default_rng(seed).integers(0, 65536, (height, width, 4), dtype=np.uint16)
    return RGBI16Frame(pixels, AcquisitionEpoch.MAIN, 4000, "blocks")


@pytest.mark.parametrize("height,rows", [(5782, 8), (13, 8), (8, 8), (5, 8), (21, 1), (30, 7)])
def test_split_assemble_round_trip(height, rows):
    source = frame(height, 11)
    blocks = split_rgbi16_frame(source, rows)
    schedule = BlockSchedule.for_frame(height, 11, rows)
    assert len(blocks) == schedule.block_count
    assert [b.valid_rows for b in blocks] == [
        schedule.valid_rows_for(i) for i in range(schedule.block_count)]
    assert [b.start_row for b in blocks] == list(range(0, height, rows))
    joined = assemble_rgbi16_blocks(blocks, epoch=AcquisitionEpoch.MAIN, resolution_dpi=4000,
                                    evidence_id="joined")
    assert np.array_equal(joined.pixels, source.pixels)
    assert joined.sha256 == source.sha256


def test_assembly_refusals():
    blocks = split_rgbi16_frame(frame(24, 9), 8)
    kwargs = dict(epoch=AcquisitionEpoch.MAIN, resolution_dpi=4000, evidence_id="x")
    for broken in ((blocks[1], blocks[0], blocks[2]), (blocks[0], blocks[2]),
                   (blocks[0], blocks[0], blocks[1], blocks[2])):
        with pytest.raises(ValueError, match="ordered, gap-free"):
            assemble_rgbi16_blocks(broken, **kwargs)
    narrow = RGBI16Block(frame(8, 8).pixels, 8)
    with pytest.raises(ValueError, match="same width"):
        assemble_rgbi16_blocks((blocks[0], narrow), **kwargs)
    with pytest.raises(ValueError, match="at least one"):
        assemble_rgbi16_blocks((), **kwargs)


def test_block_validation_and_freezing():
    with pytest.raises(ValueError, match="negative"):
        RGBI16Block(np.zeros((8, 4, 4), np.uint16), -1)
    with pytest.raises(TypeError):
        RGBI16Block(np.zeros((8, 4, 4), np.int32), 0)
    with pytest.raises(ValueError):
        RGBI16Block(np.zeros((8, 4, 3), np.uint16), 0)
    source = np.zeros((8, 4, 4), np.uint16)
    block = RGBI16Block(source, 0)
    source[0, 0, 0] = 7
    assert block.pixels[0, 0, 0] == 0
    assert not block.pixels.flags.writeable
