"""The integer writer chain against the oracle's serial writer (CPU only).

The replay splits the reference writer (reconstruction.py:916-982,
dither.py:59-104, in the strict row-major order of host_writer.py:73-93)
into an integer draw schedule per site and a serial walk over those
schedules.  `dice_t_writer_chain_host` runs exactly that code on the host
(the schedule builder the kernel uses, compiled for the host, and the walk
the runner uses, with its three-channel fast path or the per-channel path
forced); the oracle's `dor_write_pixels` runs the writer itself.  Every
site's start state, the final state and the advance count must agree.

The sites are built to sit on the decision edges: candidates within a
dither amplitude of the original (the floor test splits the state range),
at and around the dither bounds (acceptance splits it), invalid candidates
(NaN, infinities, zero, negative: no draws), the floor gate on and off,
negative and zero dither scales (the mirrored and the constant predicate).
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
from dice_amd._abi import ptr
from dice_amd.hip_backend import engine as hip_engine


def _sites(seed: int, count: int, low: float, high: float):
    rng = np.random.default_rng(seed)
    orig = rng.uniform(low - 500.0, high + 500.0, size=(count, 3)).astype(np.float32)
    kind = rng.integers(0, 6, size=count)
    cand = orig.astype(np.float64).copy()
    # 0: within a dither amplitude of the original
    near = kind == 0
    cand[near] += rng.uniform(-400.0, 400.0, size=(near.sum(), 3))
    # 1: anywhere in and around the bounds
    wide = kind == 1
    cand[wide] = rng.uniform(low - 200.0, high + 200.0, size=(wide.sum(), 3))
    # 2: next to the bounds
    edge = kind == 2
    picks = rng.choice(np.array([low, high]), size=(edge.sum(), 3))
    cand[edge] = picks + rng.uniform(-2.0, 2.0, size=(edge.sum(), 3))
    # 3: exactly on the original or on a bound
    exact = kind == 3
    cand[exact] = np.where(rng.random((exact.sum(), 3)) < 0.5, orig[exact].astype(np.float64),
                           rng.choice(np.array([low, high]), size=(exact.sum(), 3)))
    # 4: invalid candidates
    bad = np.flatnonzero(kind == 4)
    values = np.array([np.nan, np.inf, -np.inf, 0.0, -1.0, -0.0])
    cand[bad, rng.integers(0, 3, size=bad.size)] = rng.choice(values, size=bad.size)
    # 5: far above / below the original
    far = kind == 5
    cand[far] += rng.choice(np.array([-5000.0, 5000.0]), size=(far.sum(), 3))
    floor = (rng.random(count) < 0.8).astype(np.uint8)
    return np.ascontiguousarray(cand), np.ascontiguousarray(orig), floor


def _oracle(cand, orig, floor, low, high, scales, state):
    n = floor.size
    states = np.zeros(n, np.uint32)
    values = np.zeros((n, 3), np.float32)
    out, adv = C.c_uint32(), C.c_uint64()
    O.lib().dor_write_pixels(ptr(cand, C.c_double), ptr(orig, C.c_float), ptr(floor, C.c_uint8),
                             n, np.float32(low), np.float32(high), ptr(scales, C.c_float), state,
                             ptr(states, C.c_uint32), ptr(values, C.c_float), C.byref(out),
                             C.byref(adv))
    return states, out.value, adv.value


def _chain(cand, orig, floor, low, high, scales, state, generic):
    n = floor.size
    states = np.zeros(n, np.uint32)
    out, adv = C.c_uint32(), C.c_uint64()
    code = hip_engine.library().dice_t_writer_chain_host(
        ptr(cand, C.c_double), ptr(orig, C.c_float), ptr(floor, C.c_uint8), n, np.float32(low),
        np.float32(high), ptr(scales, C.c_float), state, int(generic), ptr(states, C.c_uint32),
        C.byref(out), C.byref(adv))
    assert code == 0
    return states, out.value, adv.value


SCALES = {
    "profile": (0.015, 0.015, 0.025),
    "mirrored": (-0.015, 0.02, -0.025),
    "zero": (0.0, 0.015, 0.0),
}


@pytest.mark.parametrize("scales", sorted(SCALES))
@pytest.mark.parametrize("generic", [0, 1, 2, 3], ids=["fast", "per-channel", "lanes", "grouped"])
def test_writer_chain_matches_serial_writer(scales, generic):
    lut = O.lut()
    low, high = float(lut[655]), float(lut[64879])
    scale = np.array(SCALES[scales], np.float32)
    cand, orig, floor = _sites(11 + len(scales), 60000, low, high)
    for state in (0x3045, 0xFFFFFF, 0):
        want = _oracle(cand, orig, floor, low, high, scale, state)
        got = _chain(cand, orig, floor, low, high, scale, state, generic)
        assert np.array_equal(got[0], want[0]), np.flatnonzero(got[0] != want[0])[:5]
        assert got[1:] == want[1:]


def test_writer_chain_with_empty_dither_band():
    """low >= high: nothing draws, the state never moves."""
    lut = O.lut()
    cand, orig, floor = _sites(3, 5000, float(lut[655]), float(lut[64879]))
    scale = np.array(SCALES["profile"], np.float32)
    for low, high in ((float(lut[64879]), float(lut[655])), (40000.0, 40000.0)):
        want = _oracle(cand, orig, floor, low, high, scale, 0x3045)
        got = _chain(cand, orig, floor, low, high, scale, 0x3045, False)
        assert np.array_equal(got[0], want[0])
        assert got[1:] == want[1:] == (0x3045, 0)


def _every_state(cand, orig, floor, scales, generic, low, high):
    states = np.arange(1 << 24, dtype=np.uint32)
    out = np.zeros_like(states)
    adv = np.zeros(states.size, np.uint8)
    code = hip_engine.library().dice_t_writer_site_host(
        ptr(cand, C.c_double), ptr(orig, C.c_float), int(floor), np.float32(low), np.float32(high),
        ptr(scales, C.c_float), ptr(states, C.c_uint32), states.size, int(generic),
        ptr(out, C.c_uint32), ptr(adv, C.c_uint8))
    assert code == 0
    return out, adv


def _every_state_oracle(cand, orig, floor, scales, low, high):
    states = np.arange(1 << 24, dtype=np.uint32)
    out = np.zeros_like(states)
    adv = np.zeros(states.size, np.uint8)
    O.lib().dor_write_pixel_states(ptr(cand, C.c_double), ptr(orig, C.c_float), int(floor),
                                   np.float32(low), np.float32(high), ptr(scales, C.c_float),
                                   ptr(states, C.c_uint32), states.size, ptr(out, C.c_uint32),
                                   ptr(adv, C.c_uint8))
    return out, adv


# (candidate offsets from the original, original, floor, scales): each site
# puts at least one channel on a state-dependent decision; the profile's
# scales are 0.015 / 0.015 / 0.025, the others reach the mirrored (negative
# span) and the acceptance-limited branches
EDGE_SITES = {
    "above-original": ((100.0, 50.0, 200.0), 52000.0, 1, (0.015, 0.015, 0.025)),
    "below-original": ((-100.0, -30.0, -250.0), 52000.0, 1, (0.015, 0.015, 0.025)),
    "above-mirrored": ((100.0, 50.0, 200.0), 52000.0, 1, (-0.015, -0.015, -0.025)),
    "below-mirrored": ((-100.0, -30.0, -250.0), 52000.0, 1, (-0.015, -0.015, -0.025)),
    # acceptance low < c + delta < high only splits the states when the
    # dither span exceeds half the band, hence the large scales
    "near-high-bound": ((-10.0, 30.0, 0.0), None, 1, (0.5, -0.6, 0.9)),
    "near-low-bound": ((5.0, -5.0, 60.0), "low", 1, (0.7, -0.8, 1.0)),
    "floor-off": ((100.0, -100.0, 0.0), 52000.0, 0, (0.015, 0.015, 0.025)),
    "mixed-draws": ((20000.0, 80.0, -60.0), 52000.0, 1, (0.015, 0.0, 0.025)),
}


@pytest.mark.parametrize("name", sorted(EDGE_SITES))
def test_writer_site_from_every_start_state(name):
    """Exhaustive over the 2^24 LCG states: every threshold the integer
    schedule found (including off-by-one at the interval ends) is exact."""
    lut = O.lut()
    low, high = float(lut[655]), float(lut[64879])
    offsets, base, floor, scales = EDGE_SITES[name]
    if base is None:
        orig = np.full(3, high - 40.0, np.float32)
    elif base == "low":
        orig = np.full(3, low + 20.0, np.float32)
    else:
        orig = np.full(3, base, np.float32)
    cand = orig.astype(np.float64) + np.array(offsets)
    scale = np.array(scales, np.float32)
    want = _every_state_oracle(cand, orig, floor, scale, low, high)
    if floor:
        assert np.unique(want[1]).size > 1, "site must depend on the state"
    for generic in (0, 1, 2):  # host fast path, per channel, device lane form
        got = _every_state(cand, orig, floor, scale, generic, low, high)
        bad = np.flatnonzero((got[0] != want[0]) | (got[1] != want[1]))
        assert bad.size == 0, (generic, bad[:8])
