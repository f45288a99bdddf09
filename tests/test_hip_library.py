"""CPU-side checks of libdice_hip.so (no GPU needed).

The library loads, exports every entry point include/dice_hip.h declares,
reports its ABI version, fails closed when no device exists, and its
host-only functions (hash-pinned tables, prepass, setter/stage table)
agree bit for bit with the oracle.
"""

from __future__ import annotations

import ctypes as C
import hashlib
import re
from pathlib import Path

import numpy as np
import pytest

import cases
import oracle_lib as O
from dice_amd import _abi
from dice_amd._abi import DicePrepassResult, ptr
from dice_amd.hip_backend import engine as hip_engine
from dice_amd.hip_backend.producer import schedule_arrays, schedule_struct
from dice_amd.profile import DEFAULT_PROFILE
from dice_amd.stages import StageTable

HEADER = Path(__file__).resolve().parents[1] / "include" / "dice_hip.h"


def declared_symbols() -> list[str]:
    text = HEADER.read_text()
    return sorted(set(re.findall(r"\b(dice_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = hip_engine.library()
    names = declared_symbols()
    assert "dice_process_job" in names and "dice_replay" in names
    for name in names:
        assert hasattr(lib, name), name
    assert lib.dice_abi_version() == _abi.ABI_VERSION
    for name in _abi.HIP_SIGNATURES:
        assert f"dice_{name}" in names, name


def test_tables_match_pinned_hashes():
    lib = hip_engine.library()
    lut = np.zeros(65536, np.float32)
    assert lib.dice_response_lut(ptr(lut, C.c_float)) == 0
    assert hashlib.sha256(lut.astype("<f4").tobytes()).hexdigest() == (
        "5fd225e25719a544df9475315e3017f9b8027d593ac43e392245de59f49f3fcb")
    high = np.zeros(256, np.uint32)
    low = np.zeros(256, np.uint32)
    assert lib.dice_inverse_factors(ptr(high, C.c_uint32), ptr(low, C.c_uint32)) == 0
    oh = np.zeros(256, np.uint32)
    ol = np.zeros(256, np.uint32)
    O.lib().dor_inverse_factors(ptr(oh, C.c_uint32), ptr(ol, C.c_uint32))
    assert np.array_equal(high, oh) and np.array_equal(low, ol)


@pytest.mark.parametrize("name", ["supported_job_8x8", "selftest_48x64", "tile_edge",
                                  "geometry_9x17", "all_below_threshold"])
def test_prepass_matches_oracle(name):
    _, prepass = cases.job_cases()[name]
    lut = O.lut()
    params = DEFAULT_PROFILE.prepass_parameters()
    got = DicePrepassResult()
    px = np.ascontiguousarray(prepass)
    assert hip_engine.library().dice_prepass(ptr(px, C.c_uint16), px.shape[0], px.shape[1],
                                            ptr(lut, C.c_float), C.byref(params),
                                            C.byref(got)) == 0
    want = O.prepass(prepass, params)
    assert bytes(got.payload) == bytes(want.payload)
    assert got.valid_pixels == want.valid_pixels
    assert got.fully_valid_tiles == want.fully_valid_tiles


@pytest.mark.parametrize("name", sorted(cases.prepass_cases()))
def test_prepass_matches_reference_records(name, reference_vectors):
    """dice_prepass against the reference's 0x40 records (prepass.py:96-236)."""

    pixels, params = cases.prepass_cases()[name]
    want = reference_vectors["prepass"][name]
    got = DicePrepassResult()
    px = np.ascontiguousarray(pixels)
    assert hip_engine.library().dice_prepass(ptr(px, C.c_uint16), px.shape[0], px.shape[1],
                                            ptr(O.lut(), C.c_float),
                                            C.byref(O.prepass_params(params)),
                                            C.byref(got)) == 0
    assert bytes(got.payload).hex() == want["payload_hex"]
    assert got.valid_pixels == want["valid_pixels"]
    assert got.fully_valid_tiles == want["fully_valid_tiles"]


@pytest.mark.parametrize("shape", cases.PRODUCER_CASES)
def test_stage_table_from_schedule_matches_oracle(shape):
    pixels = cases.producer_frames(*shape)
    sched = O.producer_schedule(pixels)
    count = shape[0] + 6
    want = O.stage_table_from_schedule(sched, count)
    got = StageTable.empty(count)
    arrays = schedule_arrays(shape[0])
    for key in arrays:
        arrays[key][...] = sched[key]
    s = schedule_struct(arrays)
    code = hip_engine.library().dice_stage_table_from_schedule(
        C.byref(s), 1.0, count, ptr(got.aux_alpha, C.c_float), ptr(got.aux_offset, C.c_float),
        ptr(got.aux_replacement, C.c_float), ptr(got.aux_has_replacement, C.c_uint8),
        ptr(got.score_base, C.c_float), ptr(got.coarse_reference, C.c_float),
        ptr(got.gate_secondary, C.c_uint8), ptr(got.row_gate, C.c_int32))
    assert code == 0
    for field in ("aux_alpha", "aux_offset", "aux_replacement", "aux_has_replacement",
                  "score_base", "coarse_reference", "gate_secondary", "row_gate"):
        assert np.array_equal(getattr(got, field).view(np.uint8),
                              getattr(want, field).view(np.uint8)), field


def test_no_device_fails_closed():
    try:
        summary = hip_engine.hip_device_summary()
    except hip_engine.HipBackendUnavailable as error:
        assert str(error)
        return
    assert summary["arch"].startswith("gfx950")


def test_context_lock_is_reentrant():
    """context() calls hip_device_summary() -> library() under the same lock."""

    assert hip_engine._LOCK.acquire(timeout=1)
    try:
        assert hip_engine._LOCK.acquire(timeout=1), "module lock must be reentrant"
        hip_engine._LOCK.release()
    finally:
        hip_engine._LOCK.release()
    try:
        hip_engine.context()
    except hip_engine.HipBackendUnavailable:
        pass
