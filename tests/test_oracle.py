"""Pins the CPU oracle (oracle/libdice_oracle.so) before it checks anything.

Every check here is independent of the oracle's own formulas:
  * hashes and known answers published by the reference (LUT, factor
    tables, LCG states, normalization bytes, the 8x8 job, the 256x256
    synthetic-lama output / score plane / masks);
  * reference_vectors.json, written by tests/golden/make_reference_vectors.py
    from the reference's own CPU path, for the functions the published
    goldens leave open (startup receipts, off-profile branches, producer
    edge geometries, the self-test job, adversarial tiles);
  * closed forms: impulse responses of the 9x9 averages against the
    reference's mask tables, constant patches, dither bound indices.
"""

from __future__ import annotations

import ctypes as C
import hashlib
import struct

import numpy as np
import pytest

import cases
import oracle_lib as O
from dice_amd._abi import ptr
from dice_amd.parameters import (
    AuxiliaryParameters,
    DecisionParameters,
    DitherBounds,
    NIKON_NORMALIZATION,
    ReconstructionParameters,
    ScoreParameters,
    generate_nikon_response_lut,
    to_dice_params,
)
from dice_amd.profile import DEFAULT_PROFILE, PrepassRecord, _base_primary, neutral_prepass_record
from dice_amd.stages import constant_stage_table


def sha(array: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(array).tobytes()).hexdigest()


def profile_params():
    record = neutral_prepass_record()
    return to_dice_params(
        auxiliary_parameters=DEFAULT_PROFILE.auxiliary_parameters(record),
        score_parameters=DEFAULT_PROFILE.score_parameters(record),
        decision_parameters=DEFAULT_PROFILE.decision_parameters(),
        reconstruction_parameters=DEFAULT_PROFILE.reconstruction_parameters(record),
        dither_bounds=DEFAULT_PROFILE.dither_bounds(generate_nikon_response_lut()),
    )


def stream_inputs(spec: dict, height: int):
    aux = AuxiliaryParameters(**spec["auxiliary"])
    score = ScoreParameters(**spec["score"])
    decision = DecisionParameters(**spec["decision"])
    recon = ReconstructionParameters(**spec["reconstruction"])
    bounds = DitherBounds.from_lookup(generate_nikon_response_lut(), maximum_index=65535)
    params = to_dice_params(auxiliary_parameters=aux, score_parameters=score,
                            decision_parameters=decision, reconstruction_parameters=recon,
                            dither_bounds=bounds)
    stages = constant_stage_table(height, auxiliary_parameters=aux, score_parameters=score,
                                  reconstruction_parameters=recon)
    return params, stages


def job_fingerprint(r) -> dict:
    return {
        "output_sha256": r.output_sha256,
        "attempted_pixels": r.attempted_pixels,
        "written_pixels": r.written_pixels,
        "public_rng_advances": r.public_rng_advances,
        "final_rng_state": r.final_rng_state,
        "changed_pixels": r.changed_pixels,
        "startup_attempted": list(r.startup_attempted),
        "startup_advances": list(r.startup_advances),
        "startup_final_rng_state": r.startup_final_rng_state,
    }


# ---------------------------------------------------------------------------
# Tables and known answers (tests/test_core.py:17-30 of the reference)
# ---------------------------------------------------------------------------

def test_response_lut_and_factor_hashes():
    lut = O.lut()
    assert sha(lut.astype("<f4")) == (
        "5fd225e25719a544df9475315e3017f9b8027d593ac43e392245de59f49f3fcb")
    high = np.zeros(256, np.uint32)
    low = np.zeros(256, np.uint32)
    O.lib().dor_inverse_factors(ptr(high, C.c_uint32), ptr(low, C.c_uint32))
    assert sha(high.astype("<u4")) == (
        "fdc378493f13916e36d3dbd9b69b714e888d7993680509df4550486db6fcf328")
    assert sha(low.astype("<u4")) == (
        "b7f3b1b5c0ea816ddb20196e1cf7614d364ea4e164295e0dcb78119c88ef404d")


def test_lcg_known_answers_and_normalization_bytes():
    assert struct.pack("<f", NIKON_NORMALIZATION) == bytes.fromhex("ffff7f33")
    lut = O.lut()
    low, high = float(lut[655]), float(lut[64879])
    expected = [0x1791B2, 0x8223EB, 0x8B89C0, 0x2242C1]
    for n, state in enumerate(expected, start=1):
        values = np.full(n, 50000.0)
        scales = np.full(n, 0.015, np.float32)
        deltas = np.zeros(n)
        out = C.c_uint32()
        adv = C.c_uint64()
        O.lib().dor_dither_chain(ptr(values, C.c_double), ptr(scales, C.c_float), n, 0x3045,
                                 low, high, ptr(deltas, C.c_double), C.byref(out),
                                 C.byref(adv))
        assert out.value == state
        assert adv.value == n


def test_dither_bounds_indices():
    lut = O.lut()
    lo, hi = C.c_float(), C.c_float()
    li, hi_i = C.c_int32(), C.c_int32()
    O.lib().dor_dither_bounds(ptr(lut, C.c_float), C.byref(lo), C.byref(hi), C.byref(li),
                              C.byref(hi_i))
    assert (li.value, hi_i.value) == (655, 64879)
    assert lo.value == lut[655] and hi.value == lut[64879]


# The reference's mask tables (reconstruction.py:31-36): row widths of the
# rounded 9x9 (69 samples) and 5x5 (21 samples) masks, the [1,2,1]^2 binomial.
MASK_69 = (5, 7, 9, 9, 9, 9, 9, 7, 5)
MASK_21 = (3, 5, 5, 5, 3)


def _mask(widths, size):
    m = np.zeros((size, size), bool)
    c = size // 2
    for row, w in enumerate(widths):
        m[row, c - w // 2 : c + w // 2 + 1] = True
    return m


def test_unscaled_average_impulse_responses_match_reference_masks():
    m69 = _mask(MASK_69, 9)
    m21 = np.zeros((9, 9), bool)
    m21[2:7, 2:7] = _mask(MASK_21, 5)
    binomial = np.zeros((9, 9))
    binomial[3:6, 3:6] = np.outer([1, 2, 1], [1, 2, 1])
    patches = np.zeros((81, 9, 9), np.float32)
    for i in range(81):
        patches[i].flat[i] = 1.0
    out = np.zeros((81, 3))
    O.lib().dor_unscaled_averages(ptr(patches, C.c_float), 81, ptr(out, C.c_double))
    c69 = float(np.float32(1.0 / 69.0))
    c21 = float(np.float32(1.0 / 21.0))
    for i in range(81):
        assert out[i, 0] == (c69 if m69.flat[i] else 0.0)
        assert out[i, 1] == (c21 if m21.flat[i] else 0.0)
        assert out[i, 2] == binomial.flat[i] / 16.0
    rgb = np.zeros((81, 9, 9, 3), np.float32)
    for i in range(81):
        rgb[i].reshape(81, 3)[i, :] = 1.0
    out9 = np.zeros((81, 3, 3))
    O.lib().dor_rgb_unscaled_averages(ptr(rgb, C.c_float), 81, ptr(out9, C.c_double))
    for i in range(81):
        for c in range(3):
            assert out9[i, 0, c] == (c69 if m69.flat[i] else 0.0)
            assert out9[i, 1, c] == (c21 if m21.flat[i] else 0.0)
            assert out9[i, 2, c] == binomial.flat[i] / 16.0


def test_unscaled_averages_of_constant_patches():
    patches = np.stack([np.full((9, 9), v, np.float32) for v in (0.0, 1.0, 65535.0)])
    out = np.zeros((3, 3))
    O.lib().dor_unscaled_averages(ptr(patches, C.c_float), 3, ptr(out, C.c_double))
    c69 = float(np.float32(1.0 / 69.0))
    c21 = float(np.float32(1.0 / 21.0))
    assert list(out[0]) == [0.0, 0.0, 0.0]
    assert list(out[1]) == [69.0 * c69, 21.0 * c21, 1.0]
    assert out[2, 2] == 65535.0


def test_emit_known_answers():
    values = np.array([-5.0, 0.0, 0.49999997, 65535.0, 70000.0], np.float32)
    out = np.zeros(values.size, np.uint16)
    assert O.lib().dor_emit(ptr(values, C.c_float), values.size, ptr(out, C.c_uint16)) == 0
    assert out[0] == 0 and out[1] == 0 and out[2] == 0
    assert out[3] == out[4]
    bad = np.array([np.nan], np.float32)
    assert O.lib().dor_emit(ptr(bad, C.c_float), 1, ptr(out, C.c_uint16)) == 4


# ---------------------------------------------------------------------------
# Published end-to-end goldens
# ---------------------------------------------------------------------------

def test_supported_job_golden():
    """tests/test_engine.py:28-33 of the reference."""

    main, prepass = cases.supported_job_frames()
    r = O.process_job(main, prepass, profile_params(), DEFAULT_PROFILE.prepass_parameters())
    assert r.output_sha256 == (
        "8d0992222a7523de293d69de38d560131ccf9bef2b90be9f7942406743a07941")
    assert r.attempted_pixels == 8
    assert r.written_pixels == r.changed_pixels == 2
    assert r.public_rng_advances == 26


def test_synthetic_lama_golden():
    """hybrid/examples/synthetic-lama/run/diagnostics-cache (256x256)."""

    main, prepass = cases.lama_frames()
    r = O.process_job(main, prepass, profile_params(), DEFAULT_PROFILE.prepass_parameters(),
                      diagnostics=True)
    assert r.output_sha256 == (
        "780257c669f8df3f3f47768dbd99ffe61354de5f3b66e1b2c8bdedd0cc0e4ad8")
    assert sha(r.score_plane.astype("<f4")) == (
        "566c3b0648d5c91566f2a4db3834d39fa9be934ab596ef4c9d188a9c8a3d9baa")
    assert sha(r.at_floor_mask) == (
        "ac7c5f191ce8f1150114da25edaa0a14ca1ba1c67a679bf980900fa56148af94")
    assert sha(r.changed_mask) == (
        "b652da2639bc8ee709d3cd026598f9acc085bc6ad464671507f8eecc150b40c5")
    assert int(r.at_floor_mask.sum()) == 440


def test_prepass_and_producer_invariants():
    """tests/test_core.py:33-56 of the reference."""

    main, prepass = cases.supported_job_frames()
    res = O.prepass(prepass, DEFAULT_PROFILE.prepass_parameters())
    assert res.valid_pixels == 64 and res.fully_valid_tiles == 1
    s = O.producer_schedule(main)
    assert s["active_width"] == 8
    assert np.all(s["accepted"] == 8)


PREPASS_CASES = cases.prepass_cases()


@pytest.mark.parametrize("name", sorted(PREPASS_CASES))
def test_prepass_record_matches_reference(name, reference_vectors):
    """The whole 0x40 record, valid_pixels and fully_valid_tiles against the
    reference's reduce_prepass_frame (prepass.py:96-236) on ragged, 1x1,
    all-invalid, flat, threshold-edge, large and non-default-parameter
    frames; and the profile's fail-closed base (profile.py:48-55)."""

    pixels, params = PREPASS_CASES[name]
    want = reference_vectors["prepass"][name]
    assert want["prepass_sha256"] == sha(pixels.astype("<u2"))
    assert want["parameters"] == list(params)
    got = O.prepass(pixels, O.prepass_params(params))
    assert bytes(got.payload).hex() == want["payload_hex"]
    assert got.valid_pixels == want["valid_pixels"]
    assert got.fully_valid_tiles == want["fully_valid_tiles"]
    record = PrepassRecord(bytes(got.payload))
    if "base_primary_error" in want:
        with pytest.raises(ValueError, match=want["base_primary_error"]):
            _base_primary(record)
    else:
        assert int(np.asarray(_base_primary(record), np.float32).view(np.uint32)) == (
            want["base_primary_bits"])


def test_prepass_vectors_cover_the_fail_closed_and_degenerate_cases(reference_vectors):
    vectors = reference_vectors["prepass"]
    assert vectors["alpha_one_zero_denominator"]["payload_hex"][8:16] == "0000803f"
    assert "base_primary_error" in vectors["alpha_one_zero_denominator"]
    assert vectors["all_invalid_16x24"]["valid_pixels"] == 0
    assert vectors["flat_tiles_16x16"]["fully_valid_tiles"] == 4
    assert vectors["threshold_edge_16x16"]["fully_valid_tiles"] == 3
    assert vectors["smooth_413x282"]["fully_valid_tiles"] > 1000


def test_all_below_threshold_keeps_initial_means():
    """tests/test_fast_cpu_producer_parity.py:93-102 of the reference."""

    main, _ = cases.special_frames("all_below_threshold")
    s = O.producer_schedule(main)
    assert np.all(s["accepted"] == 0)
    assert np.all(s["mean_visible"] == np.float32(46618.78515625))
    assert np.all(s["mean_infrared"] == np.float32(56239.65625))


# ---------------------------------------------------------------------------
# Reference vectors (reference CPU path outputs)
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("name", list(cases.job_cases()))
def test_oracle_job_matches_reference_vectors(name, reference_vectors):
    expected = reference_vectors["jobs"][name]
    main, prepass = cases.job_cases()[name]
    assert sha(main.astype("<u2")) == expected["main_sha256"], "case input drifted"
    r = O.process_job(main, prepass, profile_params(), DEFAULT_PROFILE.prepass_parameters(),
                      diagnostics=True)
    got = job_fingerprint(r)
    assert got == {k: expected[k] for k in got}
    assert sha(r.score_plane.astype("<f4")) == expected["score_plane_sha256"]
    assert sha(r.at_floor_mask) == expected["at_floor_sha256"]
    assert sha(r.changed_mask) == expected["changed_sha256"]


@pytest.mark.parametrize("name", list(cases.stream_cases()))
def test_oracle_stream_matches_reference_vectors(name, reference_vectors):
    expected = reference_vectors["streams"][name]
    pixels, spec, state = cases.stream_cases()[name]
    assert sha(pixels.astype("<u2")) == expected["main_sha256"]
    params, stages = stream_inputs(spec, pixels.shape[0])
    r = O.replay(pixels, params, stages, rng_state=state)
    got = job_fingerprint(r)
    assert got == {k: expected[k] for k in got}


@pytest.mark.parametrize("shape", cases.PRODUCER_CASES)
def test_oracle_producer_matches_reference_vectors(shape, reference_vectors):
    height, width = shape
    expected = reference_vectors["producer"][f"{height}x{width}"]
    pixels = cases.producer_frames(height, width)
    assert sha(pixels.astype("<u2")) == expected["main_sha256"]
    s = O.producer_schedule(pixels)
    assert s["active_width"] == expected["active_width"]
    dtypes = {"weighted_visible_sum": "<f8", "weighted_infrared_sum": "<f8",
              "weight_sum": "<f8", "accepted": "<u4"}
    for key in ("weighted_visible_sum", "weighted_infrared_sum", "weight_sum", "mean_visible",
                "mean_infrared", "accepted", "scaled_input", "input_scale",
                "scale_denominator", "scale_numerator", "scale_add_denominator",
                "scale_add_numerator"):
        assert sha(s[key].astype(dtypes.get(key, "<f4"))) == expected[key], key
    assert sha(s["payloads"]) == expected["payloads"]


def test_oracle_thread_count_determinism():
    """tests/test_fast_cpu_differential.py:430-460 of the reference."""

    main, prepass = cases.special_frames("dense_dust")
    one = O.process_job(main, prepass, profile_params(), DEFAULT_PROFILE.prepass_parameters(),
                        threads=1)
    many = O.process_job(main, prepass, profile_params(),
                         DEFAULT_PROFILE.prepass_parameters(), threads=4)
    assert one.output_sha256 == many.output_sha256
    assert one.counters() == many.counters()
