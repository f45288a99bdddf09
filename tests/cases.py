"""Deterministic parity cases shared by the tests and tests/golden scripts.

Inputs come from fixed-seed numpy generators (the reference's own L2 tile
shapes, tests/test_cuda_level2_tiles.py:91-319, with crc32(case) seeds in
place of the per-process ``hash(case)``), so every case is reproducible and
its input SHA-256 is recorded next to the expected outputs.
"""

from __future__ import annotations

import zlib

import numpy as np


def _seed(name: str) -> int:
    return zlib.crc32(name.encode())


def ramp_rgbi(height: int, width: int, *, offset: int) -> np.ndarray:
    """tests/conftest.py:16-25 of the reference."""

    y, x = np.indices((height, width), dtype=np.uint32)
    p = np.empty((height, width, 4), dtype=np.uint16)
    p[:, :, 0] = (12_000 + offset + 53 * y + 97 * x).astype(np.uint16)
    p[:, :, 1] = (16_000 + offset + 67 * y + 71 * x).astype(np.uint16)
    p[:, :, 2] = (20_000 + offset + 89 * y + 43 * x).astype(np.uint16)
    p[:, :, 3] = (44_000 + offset + 31 * y + 29 * x).astype(np.uint16)
    return p


def supported_job_frames() -> tuple[np.ndarray, np.ndarray]:
    """The reference's 8x8 ``supported_job`` (tests/conftest.py:28-53)."""

    main = ramp_rgbi(8, 8, offset=150)
    main[3:5, 3:5, 3] = np.uint16(9_000)
    return main, ramp_rgbi(8, 8, offset=0)


def lama_frames() -> tuple[np.ndarray, np.ndarray]:
    """hybrid/examples/make_synthetic_inputs.py:21-46 (256x256)."""

    def frame(offset: int) -> np.ndarray:
        y, x = np.indices((256, 256), dtype=np.uint32)
        f = np.empty((256, 256, 4), dtype=np.uint16)
        f[:, :, 0] = (12_000 + offset + 31 * y + 47 * x).astype(np.uint16)
        f[:, :, 1] = (16_000 + offset + 43 * y + 29 * x).astype(np.uint16)
        f[:, :, 2] = (20_000 + offset + 37 * y + 41 * x).astype(np.uint16)
        f[:, :, 3] = (45_000 + offset + 17 * y + 13 * x).astype(np.uint16)
        return f

    main = frame(120)
    main[110:130, 118:138, 3] = np.uint16(0)
    return main, frame(0)


def selftest_frames() -> tuple[np.ndarray, np.ndarray]:
    """backend.py:60-87 of the reference (seed 20260716, 48x64)."""

    rng = np.random.default_rng(20260716)
    height, width = 48, 64
    main = rng.integers(20000, 60000, size=(height, width, 4), dtype=np.uint16)
    main[18:26, 30:39, 3] = rng.integers(200, 2000, size=(8, 9), dtype=np.uint16)
    main[0:2, 0:6, 3] = 500
    main[-2:, -6:, 3] = 700
    main[3, width - 3, 3] = 900
    prepass = rng.integers(20000, 60000, size=(24, 32, 4), dtype=np.uint16)
    prepass[10:14, 12:17, 3] = 300
    return main, prepass


TILE_CASES = (
    "uniform",
    "impulse",
    "edge",
    "checkerboard",
    "high_dynamic_range",
    "corner_dust",
    "edge_band_dust",
    "epoch_boundary_dust",
    "threshold_straddle",
    "final_block_six_rows",
)


def tile_frames(case: str) -> tuple[np.ndarray, np.ndarray]:
    """The ten adversarial L2 tiles (41x64 / 30x64) with a fixed seed."""

    rng = np.random.default_rng(_seed(case))
    height, width = 41, 64
    if case == "final_block_six_rows":
        height = 30
    main = rng.integers(18000, 62000, size=(height, width, 4), dtype=np.uint16)
    if case == "uniform":
        main[:] = 43210
        main[:, :, 3] = 51234
    elif case == "impulse":
        main[:] = 40000
        main[height // 2, width // 2] = (65535, 0, 65535, 60000)
        main[height // 2, width // 2 + 1, 3] = 100
    elif case == "edge":
        main[:, : width // 2] = (15000, 16000, 17000, 52000)
        main[:, width // 2 :] = (60000, 59000, 58000, 52000)
        main[:, width // 2 - 1 : width // 2 + 1, 3] = 800
    elif case == "checkerboard":
        yy, xx = np.mgrid[0:height, 0:width]
        board = ((yy + xx) % 2).astype(np.uint16)
        for c in range(3):
            main[:, :, c] = 20000 + board * 30000
        main[:, :, 3] = 50000 + board * 10000
    elif case == "high_dynamic_range":
        main[:, :, :3] = rng.choice(
            np.array([0, 1, 2, 32768, 65533, 65534, 65535], dtype=np.uint16),
            size=(height, width, 3),
        )
        main[:, :, 3] = rng.choice(
            np.array([0, 8847, 8848, 8849, 65535], dtype=np.uint16), size=(height, width)
        )
    elif case == "corner_dust":
        for (y0, x0) in ((0, 0), (0, width - 3), (height - 3, 0), (height - 3, width - 3)):
            main[y0 : y0 + 3, x0 : x0 + 3, 3] = 500
    elif case == "edge_band_dust":
        main[0:4, :, 3] = rng.integers(100, 3000, size=(4, width), dtype=np.uint16)
        main[:, 0:4, 3] = rng.integers(100, 3000, size=(height, 4), dtype=np.uint16)
        main[:, width - 4 :, 3] = rng.integers(100, 3000, size=(height, 4), dtype=np.uint16)
        main[height - 4 :, :, 3] = rng.integers(100, 3000, size=(4, width), dtype=np.uint16)
    elif case == "epoch_boundary_dust":
        main[7:9, :, 3] = 600
        main[15:17, 24:40, 3] = 700
        main[23:25, :, 3] = rng.integers(100, 9000, size=(2, width), dtype=np.uint16)
    elif case == "threshold_straddle":
        main[:, :, 3] = rng.choice(
            np.array([8846, 8847, 8848, 8849, 8850], dtype=np.uint16), size=(height, width)
        )
    elif case == "final_block_six_rows":
        main[height - 6 :, :, 3] = rng.integers(100, 4000, size=(6, width), dtype=np.uint16)
    prepass = rng.integers(15000, 62000, size=(24, 32, 4), dtype=np.uint16)
    prepass[6:10, 8:13, 3] = 400
    return main, prepass


GEOMETRY_CASES = ((5, 8), (6, 9), (7, 16), (9, 17), (13, 24), (16, 33), (21, 40))


def geometry_frames(height: int, width: int) -> tuple[np.ndarray, np.ndarray]:
    """Minimal and ragged geometries (partial 8-row block, trailing columns)
    with a dust spot clipped by the frame edge."""

    rng = np.random.default_rng(_seed(f"geometry-{height}x{width}"))
    main = rng.integers(20000, 62000, size=(height, width, 4), dtype=np.uint16)
    y0, y1 = max(0, height // 2 - 2), min(height, height // 2 + 2)
    x0 = max(0, width - 5)
    main[y0:y1, x0:, 3] = rng.integers(300, 3000, size=(y1 - y0, width - x0), dtype=np.uint16)
    prepass = rng.integers(20000, 60000, size=(7, 9, 4), dtype=np.uint16)
    return main, prepass


def special_frames(name: str) -> tuple[np.ndarray, np.ndarray]:
    rng = np.random.default_rng(_seed(name))
    if name == "all_below_threshold":
        main = rng.integers(20000, 60000, size=(24, 40, 4), dtype=np.uint16)
        main[:, :, 3] = rng.integers(0, 8848, size=(24, 40), dtype=np.uint16)
    elif name == "uniform_high_ir":
        main = rng.integers(20000, 60000, size=(24, 40, 4), dtype=np.uint16)
        main[:, :, 3] = 65535
    elif name == "dense_dust":
        main = rng.integers(30000, 60000, size=(40, 72, 4), dtype=np.uint16)
        yy, xx = np.mgrid[0:40, 0:72]
        for cy, cx, r in ((10, 15, 7), (25, 40, 9), (33, 60, 5), (5, 66, 4)):
            disk = (yy - cy) ** 2 + (xx - cx) ** 2 <= r * r
            main[disk, 3] = (main[disk, 3] * 0.3).astype(np.uint16)
            for c in range(3):
                main[disk, c] = (main[disk, c] * 0.9).astype(np.uint16)
        main[18:20, :, 3] = 2000
    else:
        raise KeyError(name)
    prepass = rng.integers(20000, 60000, size=(12, 16, 4), dtype=np.uint16)
    return main, prepass


SPECIAL_CASES = ("all_below_threshold", "uniform_high_ir", "dense_dust")


def job_cases() -> dict[str, tuple[np.ndarray, np.ndarray]]:
    """Every full-job parity case, by name."""

    cases: dict[str, tuple[np.ndarray, np.ndarray]] = {
        "supported_job_8x8": supported_job_frames(),
        "selftest_48x64": selftest_frames(),
    }
    for case in TILE_CASES:
        cases[f"tile_{case}"] = tile_frames(case)
    for height, width in GEOMETRY_CASES:
        cases[f"geometry_{height}x{width}"] = geometry_frames(height, width)
    for name in SPECIAL_CASES:
        cases[name] = special_frames(name)
    return cases


# ---------------------------------------------------------------------------
# Streaming-level cases with explicit parameters (provider=None).
# ---------------------------------------------------------------------------

def f32_bits(bits: int) -> float:
    return float(np.frombuffer(np.uint32(bits).tobytes(), dtype=np.float32)[0])


def base_stream_parameters(*, resolution_metric: int = 4000, radius: int = 4,
                           alpha: float = 0.4, replacement=None, visible_channel: int = 0,
                           count_limit: int = 8, **recon_overrides) -> dict:
    """Plain dict of the reference dataclass fields (tests/test_cuda_level2_
    tiles.py:169-319 parameter shapes)."""

    recon = dict(
        resolution_metric=resolution_metric,
        cross_neighbor_cutoff=1600,
        coarse_enabled=True,
        coarse_reference=float(np.float32(52215.4)),
        coarse_slopes=(1.1, 1.1, 1.1),
        band_enabled=(True, True, True),
        band_scales=(1.25, 1.25, 1.25),
        factors_a=((1.21, 1.23, 1.13), (1.17, 1.14, 1.08), (1.04, 0.93, 0.97)),
        factors_b=((1.09, 1.13, 1.04), (1.08, 1.05, 1.02), (0.96, 0.84, 0.89)),
        configured_strengths=(0.0, 0.0, 0.0),
        driver_gate_primary=True,
        driver_gate_secondary=True,
        row_reconstruction_gate=0,
        dither_scales=(0.015, 0.015, 0.025),
    )
    recon.update(recon_overrides)
    return dict(
        auxiliary=dict(selected_visible_channel=visible_channel,
                       alpha=float(np.float32(alpha)), calibration_offset=1.0,
                       alpha_one_replacement=replacement),
        score=dict(base_primary=float(np.float32(52215.4)),
                   base_addend=float(np.float32(-119.40625)),
                   scale=float(np.float32(-0.0010412337)), offset=1.0,
                   floor=float(np.float32(0.02)), resolution_metric=resolution_metric,
                   horizontal_minimum_resolution_cutoff=550),
        decision=dict(sample_threshold=float(np.float32(49383.234)), count_limit=count_limit,
                      perpendicular_radius=radius),
        reconstruction=recon,
    )


def stream_cases() -> dict[str, tuple[np.ndarray, dict, int]]:
    """name -> (pixels, parameter dict, initial LCG state)."""

    cases: dict[str, tuple[np.ndarray, dict, int]] = {}
    rng = np.random.default_rng(500500)
    pixels = rng.integers(15000, 64000, size=(26, 48, 4), dtype=np.uint16)
    pixels[9:14, 20:29, 3] = 700
    cases["metric500_center_only"] = (
        pixels, base_stream_parameters(resolution_metric=500, radius=1, alpha=0.32), 0x3045)
    rng = np.random.default_rng(99)
    pixels = rng.integers(15000, 64000, size=(17, 40, 4), dtype=np.uint16)
    pixels[5:9, 10:19, 3] = 900
    cases["floor_gates_off"] = (
        pixels, base_stream_parameters(driver_gate_primary=False, driver_gate_secondary=False),
        0x3045)
    cases["row_gate_on"] = (pixels, base_stream_parameters(row_reconstruction_gate=3), 0x3045)
    cases["secondary_gate_only"] = (
        pixels, base_stream_parameters(driver_gate_primary=False), 0x3045)
    rng = np.random.default_rng(_seed("alpha_one"))
    pixels = rng.integers(15000, 64000, size=(19, 33, 4), dtype=np.uint16)
    pixels[6:11, 12:20, 3] = 1100
    cases["alpha_one_replacement"] = (
        pixels, base_stream_parameters(alpha=1.0, replacement=float(np.float32(51000.25))),
        0x3045)
    cases["visible_channel_2_radius_2"] = (
        pixels, base_stream_parameters(visible_channel=2, radius=2, count_limit=6), 0x123456)
    cases["configured_strengths"] = (
        pixels,
        base_stream_parameters(configured_strengths=(0.5, 0.25, 0.75),
                               band_enabled=(True, False, True), coarse_enabled=False),
        0x0ABCDE)
    return cases


PRODUCER_CASES = ((16, 32), (41, 64), (30, 48), (9, 24), (26, 41), (7, 16), (9, 17),
                  (8, 8), (33, 70))


def producer_frames(height: int, width: int) -> np.ndarray:
    """Producer geometries with dense threshold straddles
    (tests/test_cuda_producer_parity.py:31-80 shapes)."""

    rng = np.random.default_rng(_seed(f"producer-{height}x{width}"))
    pixels = rng.integers(0, 65536, size=(height, width, 4), dtype=np.uint16)
    straddle = rng.random((height, width)) < 0.15
    pixels[:, :, 3] = np.where(
        straddle,
        rng.choice(np.array([8846, 8847, 8848, 8849], dtype=np.uint16), size=(height, width)),
        rng.integers(8848, 65536, size=(height, width), dtype=np.uint16),
    )
    return pixels


# ---------------------------------------------------------------------------
# Prepass reductions (prepass.py:96-236) on their own.
# ---------------------------------------------------------------------------

# (valid_ir_threshold, alpha_scale, ratio_minimum, ratio_maximum,
#  zero_denominator_alpha, completion_marker); every float is binary32-exact.
PREPASS_DEFAULT = (8847.2255859375, 1.5, float(np.float32(-0.2)), float(np.float32(0.2)),
                   float(np.float32(0.17)), 16)
PREPASS_CUSTOM = (30000.0, 0.75, -1.0, 1.0, 0.5, 7)
# zero_denominator_alpha = 1 on a frame with no fully valid tile: the record's
# alpha is exactly 1, which the profile must refuse (profile.py:48-55).
PREPASS_ALPHA_ONE = (8847.2255859375, 1.5, float(np.float32(-0.2)), float(np.float32(0.2)),
                     1.0, 16)


def _smooth_field(rng, height: int, width: int, low: int, high: int) -> np.ndarray:
    y, x = np.indices((height, width), dtype=np.float64)
    field = np.zeros((height, width))
    for _ in range(4):
        fy, fx = rng.uniform(0.2, 3.0, 2) * 2 * np.pi / np.array([height, width])
        field += np.cos(fy * y + fx * x + rng.uniform(0, 2 * np.pi))
    field = (field - field.min()) / max(np.ptp(field), 1e-9)
    return low + (high - low) * field


def prepass_cases() -> dict[str, tuple[np.ndarray, tuple]]:
    """Prepass frames and parameters: every job case's prepass, ragged and
    degenerate geometries (zero padding to a multiple of 8), threshold
    straddles, flat tiles (zero deviations), a large smooth frame with many
    in-window ratios, non-default parameters, and a record whose alpha is 1."""

    out: dict[str, tuple[np.ndarray, tuple]] = {}
    for name, (_, prepass) in job_cases().items():
        out[f"job_{name}"] = (prepass, PREPASS_DEFAULT)

    def random_frame(name, height, width, invalid=0.0, low=20000, high=60000):
        rng = np.random.default_rng(_seed(f"prepass-{name}"))
        px = rng.integers(low, high, size=(height, width, 4), dtype=np.uint16)
        if invalid:
            mask = rng.random((height, width)) < invalid
            px[:, :, 3] = np.where(mask, rng.integers(0, 8848, size=(height, width)),
                                   px[:, :, 3]).astype(np.uint16)
        return px

    one = np.zeros((1, 1, 4), np.uint16)
    one[0, 0] = (31000, 27000, 22000, 50000)
    out["one_by_one_valid"] = (one, PREPASS_DEFAULT)
    dark = one.copy()
    dark[0, 0, 3] = 100
    out["one_by_one_invalid"] = (dark, PREPASS_DEFAULT)
    out["ragged_7x13"] = (random_frame("7x13", 7, 13), PREPASS_DEFAULT)
    out["ragged_3x100"] = (random_frame("3x100", 3, 100), PREPASS_DEFAULT)
    out["ragged_13x21_mixed"] = (random_frame("13x21", 13, 21, invalid=0.02), PREPASS_DEFAULT)
    out["all_invalid_16x24"] = (random_frame("invalid", 16, 24, low=0, high=8848),
                                PREPASS_DEFAULT)
    flat = np.empty((16, 16, 4), np.uint16)
    flat[...] = (30000, 31000, 32000, 50000)
    out["flat_tiles_16x16"] = (flat, PREPASS_DEFAULT)
    edge = random_frame("edge", 16, 16)
    edge[0:8, 0:8, 3] = 8848           # lowest valid raw IR: the tile is fully valid
    edge[8:16, 8:16, 3] = 8848
    edge[12, 12, 3] = 8847             # highest invalid raw IR: the tile is not
    out["threshold_edge_16x16"] = (edge, PREPASS_DEFAULT)
    rng = np.random.default_rng(_seed("prepass-smooth"))
    smooth = np.empty((413, 282, 4), np.uint16)
    for lane, (low, high) in enumerate(((9000, 50000), (9000, 50000), (9000, 50000),
                                        (40000, 60000))):
        smooth[:, :, lane] = np.clip(np.rint(_smooth_field(rng, 413, 282, low, high)
                                             + rng.normal(0, 40, (413, 282))), 0, 65535)
    out["smooth_413x282"] = (smooth, PREPASS_DEFAULT)
    out["custom_params_40x56"] = (random_frame("custom", 40, 56, invalid=0.01, low=30001),
                                  PREPASS_CUSTOM)
    out["alpha_one_zero_denominator"] = (random_frame("alpha-one", 9, 17, low=0, high=8000),
                                         PREPASS_ALPHA_ONE)
    return out
