"""The allocation-free job entry point against the allocating one.

``dice_process_job_noalloc`` runs the job of ``dice_process_job`` (device
input and output, no SHA-256) with every intermediate in the library's
file-scope device scratch.  Two calls in a row, the second on a smaller
frame, must each equal ``dice_process_job``'s output array, five counters
and startup receipt, and the oracle's; a frame whose intermediates cannot
fit the scratch is refused before any device work.
"""

from __future__ import annotations

import ctypes as C
import hashlib

import numpy as np
import pytest

import oracle_lib
from conftest import hip_device_or_skip
from dice_amd import _abi, synth
from dice_amd._abi import DiceReplayResult, ptr
from dice_amd.hip_backend import engine as hip_engine
from dice_amd.hip_backend.device import DeviceArray
from dice_amd.hip_backend.process import stable_profile_params
from dice_amd.profile import DEFAULT_PROFILE

pytestmark = pytest.mark.gpu

DEVICE_FLAGS = _abi.DICE_MAIN_ON_DEVICE | _abi.DICE_OUT_ON_DEVICE | _abi.DICE_SKIP_SHA256
COUNTERS = ("attempted_pixels", "written_pixels", "public_rng_advances", "final_rng_state",
            "changed_pixels", "startup_final_rng_state")


@pytest.fixture(scope="module", autouse=True)
def _device():
    hip_device_or_skip()


def receipt(res: DiceReplayResult) -> dict:
    out = {name: int(getattr(res, name)) for name in COUNTERS}
    out["startup_attempted"] = list(res.startup_attempted)
    out["startup_advances"] = list(res.startup_advances)
    return out


def run(entry: str, main: np.ndarray, prepass: np.ndarray):
    lib = hip_engine.library()
    ctx = hip_engine.context()
    params = stable_profile_params()
    prepass_params = DEFAULT_PROFILE.prepass_parameters()
    height, width = main.shape[:2]
    d_main = DeviceArray.from_host(main)
    d_out = DeviceArray(height * width * 6)
    prepass = np.ascontiguousarray(prepass)
    res = DiceReplayResult()
    common = (ctx, C.c_void_p(d_main.address), height, width, ptr(prepass, C.c_uint16),
              prepass.shape[0], prepass.shape[1], C.byref(params), C.byref(prepass_params))
    if entry == "noalloc":
        code = lib.dice_process_job_noalloc(*common, C.c_void_p(d_out.address), C.byref(res))
    else:
        code = lib.dice_process_job(*common, DEVICE_FLAGS, C.c_void_p(d_out.address),
                                    C.byref(res), None, None, None)
    assert code == 0, (entry, hip_engine.last_error())
    output = d_out.download(np.empty((height, width, 3), np.uint16))
    return output, receipt(res), res.output_sha256.decode()


@pytest.mark.parametrize("order", ["full-then-crop"])
def test_noalloc_matches_allocating_entry_point_twice(order):
    specs = [synth.config2_full_frame(seed=7), synth.config1_crop()]
    for call, spec in enumerate(specs):
        main = synth.frame(spec)
        prepass = synth.prepass_for(main)
        print(f"call {call}: {spec.height}x{spec.width}", flush=True)
        want_out, want, want_sha = run("alloc", main, prepass)
        got_out, got, got_sha = run("noalloc", main, prepass)
        assert want_sha == "" and got_sha == ""
        assert np.array_equal(got_out, want_out), call
        assert got == want, call
        cpu = oracle_lib.process_job(main, prepass, stable_profile_params(),
                                     DEFAULT_PROFILE.prepass_parameters(), threads=16)
        assert hashlib.sha256(got_out.astype("<u2").tobytes()).hexdigest() == cpu.output_sha256
        assert got["attempted_pixels"] == cpu.attempted_pixels
        assert got["final_rng_state"] == cpu.final_rng_state


def test_noalloc_refuses_a_frame_beyond_its_rows_before_device_work():
    lib = hip_engine.library()
    ctx = hip_engine.context()
    height, width = 16385, 8
    prepass = np.full((4, 4, 4), 50000, np.uint16)
    res = DiceReplayResult()
    # full-size device buffers, although the frame is refused on its shape
    d_main = DeviceArray(height * width * 8)
    d_out = DeviceArray(height * width * 6)
    code = lib.dice_process_job_noalloc(
        ctx, C.c_void_p(d_main.address), height, width, ptr(prepass, C.c_uint16), 4, 4,
        C.byref(stable_profile_params()), C.byref(DEFAULT_PROFILE.prepass_parameters()),
        C.c_void_p(d_out.address), C.byref(res))
    assert code == _abi.DICE_OVER_BUDGET
    assert "at most 16384 rows" in hip_engine.last_error()
