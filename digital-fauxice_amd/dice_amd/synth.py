"""Synthetic RGBI uint16 scans (libdice_synth.so) for parity tests and benchmarks.

The five workloads of the benchmark contract:
  1. 256x256 crop, seed 0, 12 dust disks (r 2-12), 3 scratches (50-200 px);
  2. 5782x3946 full 35 mm frame (4000 dpi), 400 disks, 40 scratches <= 400 px;
  3. 16 crops, sides uniform in 512-3000 px, aspect 0.5-2, defects scaled to
     area at the full-frame density;
  4. 8818x13228 large format (56x84 mm), 1500 disks incl. blobs r <= 30,
     120 scratches;
  5. a serving mix of 1024-2048 px crops (70 %) and full frames (30 %).
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from pathlib import Path

import numpy as np

LIBRARY_PATH = Path(__file__).resolve().parent / "libdice_synth.so"
FULL_FRAME = (5782, 3946)
LARGE_FRAME = (8818, 13228)
FULL_FRAME_DISKS = 400
FULL_FRAME_SCRATCHES = 40

_LIB: list[C.CDLL] = []


def _lib() -> C.CDLL:
    if not _LIB:
        if not LIBRARY_PATH.exists():
            raise RuntimeError(f"{LIBRARY_PATH} is missing (run __graft_entry__.build())")
        lib = C.CDLL(str(LIBRARY_PATH))
        lib.dice_synth_frame.restype = C.c_int32
        lib.dice_synth_frame.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_uint64,
                                         C.c_int32, C.c_float, C.c_float, C.c_int32,
                                         C.c_float, C.c_float]
        lib.dice_synth_prepass.restype = C.c_int32
        lib.dice_synth_prepass.argtypes = [C.c_void_p, C.c_int64, C.c_int64, C.c_int32,
                                           C.c_void_p, C.POINTER(C.c_int64),
                                           C.POINTER(C.c_int64)]
        _LIB.append(lib)
    return _LIB[0]


@dataclass(frozen=True)
class FrameSpec:
    height: int
    width: int
    seed: int
    disks: int
    disk_radius: tuple[float, float] = (2.0, 12.0)
    scratches: int = 0
    scratch_length: tuple[float, float] = (50.0, 200.0)

    @property
    def pixels(self) -> int:
        return self.height * self.width


def frame(spec: FrameSpec) -> np.ndarray:
    out = np.empty((spec.height, spec.width, 4), dtype=np.uint16)
    code = _lib().dice_synth_frame(out.ctypes.data, spec.height, spec.width, spec.seed,
                                   spec.disks, spec.disk_radius[0], spec.disk_radius[1],
                                   spec.scratches, spec.scratch_length[0],
                                   spec.scratch_length[1])
    if code != 0:
        raise RuntimeError(f"synthetic frame generation failed ({code})")
    return out


def prepass_for(main: np.ndarray, factor: int = 14) -> np.ndarray:
    height, width = main.shape[:2]
    out = np.empty((max(1, height // factor), max(1, width // factor), 4), dtype=np.uint16)
    ph, pw = C.c_int64(), C.c_int64()
    src = np.ascontiguousarray(main)
    code = _lib().dice_synth_prepass(src.ctypes.data, height, width, factor, out.ctypes.data,
                                     C.byref(ph), C.byref(pw))
    if code != 0:
        raise RuntimeError("synthetic prepass generation failed")
    return out[: ph.value, : pw.value]


def config1_crop() -> FrameSpec:
    return FrameSpec(256, 256, seed=0, disks=12, scratches=3, scratch_length=(50.0, 200.0))


def config2_full_frame(seed: int = 2) -> FrameSpec:
    return FrameSpec(*FULL_FRAME, seed=seed, disks=FULL_FRAME_DISKS,
                     scratches=FULL_FRAME_SCRATCHES, scratch_length=(50.0, 400.0))


def scaled_spec(height: int, width: int, seed: int) -> FrameSpec:
    """Defect counts at the full-frame density."""

    area = height * width / float(FULL_FRAME[0] * FULL_FRAME[1])
    return FrameSpec(height, width, seed=seed,
                     disks=max(1, int(round(FULL_FRAME_DISKS * area))),
                     scratches=max(1, int(round(FULL_FRAME_SCRATCHES * area))),
                     scratch_length=(50.0, 400.0))


def config3_crops(count: int = 16, seed: int = 3) -> list[FrameSpec]:
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        side = int(rng.integers(512, 3001))
        aspect = float(rng.uniform(0.5, 2.0))
        height = side
        width = int(min(3000, max(512, round(side * aspect))))
        specs.append(scaled_spec(height, width, seed=1000 + i))
    return specs


def config4_large(seed: int = 4) -> FrameSpec:
    return FrameSpec(*LARGE_FRAME, seed=seed, disks=1500, disk_radius=(2.0, 30.0),
                     scratches=120, scratch_length=(50.0, 400.0))


def config5_serving(count: int = 200, seed: int = 5) -> list[FrameSpec]:
    rng = np.random.default_rng(seed)
    specs = []
    for i in range(count):
        if rng.random() < 0.7:
            height = int(rng.integers(1024, 2049))
            width = int(rng.integers(1024, 2049))
            specs.append(scaled_spec(height, width, seed=5000 + i))
        else:
            specs.append(config2_full_frame(seed=5000 + i))
    return specs
