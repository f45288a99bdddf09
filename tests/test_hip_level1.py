"""Level 1: HIP device primitives against the oracle, bit for bit.

Mirrors tests/test_cuda_level1_primitives.py:112-460 of the reference:
response/aux planes on a random 37x53 frame, the score equation with
weighted planes on 23x61, the 9x9 averages on 4096 patches (with all-zero,
all-65535 and unit patches), the RGB-candidate averages on 2048 patches, a
50k-draw dither chain with the interval edges, the output conversion on
100k values plus edges, the nonfinite-aux flag, and the decision bars at
every edge for radii 1-4.  Comparisons are on raw bit patterns.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from conftest import hip_device_or_skip
from dice_amd._abi import ptr
from dice_amd.hip_backend import engine as H
from dice_amd.parameters import f32_from_bits
from test_oracle import profile_params

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    hip_device_or_skip()
    return H.context()


def _params(**overrides):
    p = profile_params()
    for key, value in overrides.items():
        setattr(p, key, value)
    return p


def test_device_summary_reports_gfx950(ctx):
    summary = H.hip_device_summary()
    assert summary["arch"].startswith("gfx950")
    assert summary["total_hbm_bytes"] > 0
    assert "-ffp-contract=off" in summary["build_flags"]


@pytest.mark.parametrize("alpha,vis,horizontal", [(0.4375, 0, 1), (0.5, 2, 1), (1.0, 0, 1),
                                                  (0.32, 1, 0)])
def test_planes_bits(ctx, alpha, vis, horizontal):
    rng = np.random.default_rng(1)
    height, width = 37, 53
    pixels = rng.integers(0, 65536, size=(height, width, 4), dtype=np.uint16)
    lut = O.lut()
    params = _params(visible_channel=vis, horizontal_minimum=horizontal)
    row_alpha = np.full(height, np.float32(alpha), np.float32)
    row_alpha[::5] = np.float32(0.25)
    is_one = (row_alpha == np.float32(1.0)).astype(np.uint8)
    repl = np.full(height, np.float32(52011.4), np.float32)
    offset = np.full(height, np.float32(1.0), np.float32)
    base = np.linspace(50000, 54000, height).astype(np.float32)
    outs = {}
    for label in ("hip", "oracle"):
        o = dict(working=np.zeros((height, width, 4), np.float32),
                 aux=np.zeros((height, width), np.float32),
                 score=np.zeros((height, width), np.float32),
                 waux=np.zeros((height, width), np.float32),
                 wrgb=np.zeros((height, width, 3), np.float32))
        args = [ptr(pixels, C.c_uint16), height, width, ptr(lut, C.c_float), C.byref(params),
                ptr(row_alpha, C.c_float), ptr(is_one, C.c_uint8), ptr(repl, C.c_float),
                ptr(offset, C.c_float), ptr(base, C.c_float)] + [
                    ptr(o[k], C.c_float) for k in ("working", "aux", "score", "waux", "wrgb")]
        if label == "hip":
            flags = C.c_uint32(0)
            assert H.library().dice_t_planes(ctx, *args, C.byref(flags)) == 0
            assert flags.value == 0
        else:
            assert O.lib().dor_planes(*args) == 0
        outs[label] = o
    for key in outs["hip"]:
        assert np.array_equal(outs["hip"][key].view(np.uint32),
                              outs["oracle"][key].view(np.uint32)), key


def test_nonfinite_auxiliary_fails_closed_like_cpu(ctx):
    """tests/test_cuda_level1_primitives.py:342-388: flag, never launder."""

    height, width = 5, 16
    pixels = np.full((height, width, 4), 30000, np.uint16)
    lut = O.lut()
    params = _params()
    nan = np.float32(np.nan)
    row_alpha = np.full(height, np.float32(0.4), np.float32)
    is_one = np.zeros(height, np.uint8)
    is_one[2] = 1
    repl = np.full(height, nan, np.float32)
    offset = np.ones(height, np.float32)
    base = np.full(height, np.float32(50000), np.float32)
    bufs = [np.zeros((height, width, 4), np.float32)] + [
        np.zeros((height, width), np.float32) for _ in range(3)] + [
        np.zeros((height, width, 3), np.float32)]
    flags = C.c_uint32(0)
    code = H.library().dice_t_planes(
        ctx, ptr(pixels, C.c_uint16), height, width, ptr(lut, C.c_float), C.byref(params),
        ptr(row_alpha, C.c_float), ptr(is_one, C.c_uint8), ptr(repl, C.c_float),
        ptr(offset, C.c_float), ptr(base, C.c_float), *[ptr(b, C.c_float) for b in bufs],
        C.byref(flags))
    assert code == 0
    assert flags.value & 1


def test_unscaled_averages_bits(ctx):
    rng = np.random.default_rng(3)
    count = 4096
    patches = rng.uniform(0.0, 66000.0, size=(count, 9, 9)).astype(np.float32)
    patches[0] = 0.0
    patches[1] = 65535.0
    patches[2] = rng.uniform(0.0, 1.0, size=(9, 9)).astype(np.float32)
    got = np.zeros((count, 3))
    want = np.zeros((count, 3))
    assert H.library().dice_t_unscaled_averages(ctx, ptr(patches, C.c_float), count,
                                                ptr(got, C.c_double)) == 0
    O.lib().dor_unscaled_averages(ptr(patches, C.c_float), count, ptr(want, C.c_double))
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))


def test_rgb_candidate_averages_bits(ctx):
    rng = np.random.default_rng(4)
    count = 2048
    patches = rng.uniform(0.0, 66000.0, size=(count, 9, 9, 3)).astype(np.float32)
    got = np.zeros((count, 3, 3))
    want = np.zeros((count, 3, 3))
    assert H.library().dice_t_rgb_unscaled_averages(ctx, ptr(patches, C.c_float), count,
                                                    ptr(got, C.c_double)) == 0
    O.lib().dor_rgb_unscaled_averages(ptr(patches, C.c_float), count, ptr(want, C.c_double))
    assert np.array_equal(got.view(np.uint64), want.view(np.uint64))


def test_conditional_dither_chain_bits(ctx):
    lut = O.lut()
    low, high = float(lut[655]), float(lut[64879])
    rng = np.random.default_rng(5)
    count = 50000
    values = np.concatenate([
        rng.uniform(low - 100.0, high + 100.0, size=count - 6),
        np.array([low, high, np.nextafter(low, high), np.nextafter(high, low), 0.0, 65535.0]),
    ]).astype(np.float64)
    scales = rng.choice(np.array([0.015, 0.025, 0.0], np.float32), size=count).astype(np.float32)
    results = {}
    for label in ("hip", "oracle"):
        deltas = np.zeros(count)
        state = C.c_uint32()
        adv = C.c_uint64()
        args = (ptr(values, C.c_double), ptr(scales, C.c_float), count, 0x3045,
                np.float32(low), np.float32(high), ptr(deltas, C.c_double), C.byref(state),
                C.byref(adv))
        if label == "hip":
            assert H.library().dice_t_dither_chain(ctx, *args) == 0
        else:
            O.lib().dor_dither_chain(*args)
        results[label] = (deltas, state.value, adv.value)
    assert np.array_equal(results["hip"][0].view(np.uint64), results["oracle"][0].view(np.uint64))
    assert results["hip"][1:] == results["oracle"][1:]


def test_output_conversion_bits(ctx):
    rng = np.random.default_rng(6)
    values = np.concatenate([
        rng.uniform(-10.0, 65545.0, size=100000),
        np.array([-0.5, -0.4999999, 0.0, 0.49999997, 0.5, 65534.5, 65535.0, 65540.0]),
    ]).astype(np.float32)
    got = np.zeros(values.size, np.uint16)
    want = np.zeros(values.size, np.uint16)
    assert H.library().dice_t_emit(ctx, ptr(values, C.c_float), values.size,
                                   ptr(got, C.c_uint16)) == 0
    assert O.lib().dor_emit(ptr(values, C.c_float), values.size, ptr(want, C.c_uint16)) == 0
    assert np.array_equal(got, want)
    bad = np.array([1.0, np.inf], np.float32)
    assert H.library().dice_t_emit(ctx, ptr(bad, C.c_float), 2, ptr(got, C.c_uint16)) == 4


@pytest.mark.parametrize("radius", [1, 2, 3, 4])
@pytest.mark.parametrize("limit", [8, 3, 0])
def test_decision_all_edges(ctx, radius, limit):
    rng = np.random.default_rng(7)
    height, width = 21, 33
    aux = rng.uniform(40000.0, 60000.0, size=(height, width)).astype(np.float32)
    aux[0, :5] = 100.0
    aux[-1, -7:] = 200.0
    aux[:, 0] = rng.uniform(0, 60000, size=height).astype(np.float32)
    params = _params(decision_radius=radius, decision_count_limit=limit,
                     decision_threshold=np.float32(f32_from_bits(0x4740E73C)))
    got = np.zeros((height, width), np.uint8)
    want = np.zeros((height, width), np.uint8)
    assert H.library().dice_t_decision(ctx, ptr(aux, C.c_float), height, width,
                                       C.byref(params), ptr(got, C.c_uint8)) == 0
    O.lib().dor_decision(ptr(aux, C.c_float), height, width, C.byref(params),
                         ptr(want, C.c_uint8))
    assert np.array_equal(got, want)
