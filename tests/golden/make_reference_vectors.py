"""Write tests/golden/reference_vectors.json from the reference CPU path.

Runs only in the build container, where the reference checkout is mounted
at /root/reference.  It loads the reference's pure-NumPy CPU modules
(src/portable_digital_ice/, without the package __init__, which needs
Python >= 3.11) and records, for every case of tests/cases.py, the input
SHA-256 and the reference's outputs: output SHA-256, replay counters,
startup receipt, diagnostics plane hashes, and producer schedule hashes.
These pin the oracle for the functions the published goldens leave open
(startup receipts, off-profile branches, producer edge geometries,
self-test job, adversarial tiles).  The vectors are data; no reference
source is copied.
"""

from __future__ import annotations

import hashlib
import importlib
import json
import sys
import types
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
import cases  # noqa: E402

REFERENCE_SRC = Path("/root/reference/src/portable_digital_ice")


def load_reference():
    pkg = types.ModuleType("portable_digital_ice")
    pkg.__path__ = [str(REFERENCE_SRC)]
    sys.modules["portable_digital_ice"] = pkg
    names = ["engine", "streaming", "x3a", "reconstruction", "dither", "rng",
             "contracts", "profile", "producer_parameters", "prepass"]
    return {name: importlib.import_module(f"portable_digital_ice.{name}") for name in names}


def sha(array: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(array).tobytes()).hexdigest()


def replay_fields(replay) -> dict:
    return {
        "output_sha256": replay.output_sha256,
        "attempted_pixels": int(replay.attempted_pixels),
        "written_pixels": int(replay.written_pixels),
        "public_rng_advances": int(replay.public_rng_advances),
        "final_rng_state": int(replay.final_rng_state),
        "changed_pixels": int(replay.changed_pixels),
        "startup_attempted": [int(v) for v in replay.startup.attempted_per_stage],
        "startup_advances": [int(v) for v in replay.startup.rng_advances_per_stage],
        "startup_final_rng_state": int(replay.startup.final_rng_state),
    }


def job_vector(R, main, prepass) -> dict:
    contracts, profile, engine = R["contracts"], R["profile"], R["engine"]
    job = profile.ProcessingJob(
        acquisition=contracts.DualRGBIAcquisition(
            prepass=contracts.RGBI16Frame(prepass, contracts.AcquisitionEpoch.PREPASS, 285, "pp"),
            main=contracts.RGBI16Frame(main, contracts.AcquisitionEpoch.MAIN, 4000, "main"),
            same_frame_id="golden",
        ),
        scanner_model=profile.ScannerModel.NIKON_SUPER_COOLSCAN_5000_ED,
        mode=profile.ProcessingMode.NORMAL,
        selector=8,
        resolution_metric=4000,
        bit_depth=16,
        focus_exposure_locked=True,
    )
    result = engine.process_cpu(job, export_diagnostics=True)
    d = result.diagnostics
    vector = {
        "main_sha256": sha(main.astype("<u2")),
        "prepass_sha256": sha(prepass.astype("<u2")),
        **replay_fields(result.replay),
        "output_samples_sha256": sha(result.output_rgb16.astype("<u2")),
        "score_plane_sha256": sha(d.score_plane.astype("<f4")),
        "score_floor_bits": int(np.asarray(d.score_floor).view(np.uint32)),
        "at_floor_sha256": sha(d.at_floor_mask.astype(np.bool_)),
        "changed_sha256": sha(d.changed_mask.astype(np.bool_)),
    }
    return vector


def stream_vector(R, pixels, spec, state) -> dict:
    x3a, rec, dither, rng, streaming = R["x3a"], R["reconstruction"], R["dither"], R["rng"], R["streaming"]
    response = x3a.SharedLookupInputResponse.nikon_logarithmic()
    recon = dict(spec["reconstruction"])
    output = np.empty((*pixels.shape[:2], 3), dtype=np.uint16)
    replay = streaming.run_streaming_replay(
        pixels,
        output,
        response=response,
        auxiliary_parameters=x3a.AuxiliaryParameters(**spec["auxiliary"]),
        score_parameters=x3a.ScoreParameters(**spec["score"]),
        decision_parameters=x3a.DecisionParameters(**spec["decision"]),
        reconstruction_parameters=rec.ReconstructionParameters(**recon),
        dither_bounds=dither.DitherBounds.from_lookup(response.table, maximum_index=65535),
        generator=rng.LCG24(state),
    )
    return {"main_sha256": sha(pixels.astype("<u2")), **replay_fields(replay),
            "output_samples_sha256": sha(output.astype("<u2"))}


def producer_vector(R, pixels) -> dict:
    x3a, pp = R["x3a"], R["producer_parameters"]
    table = x3a.SharedLookupInputResponse.nikon_logarithmic().table
    s = pp.derive_producer_record_schedule(pixels, table)
    m = s.mean_schedule
    return {
        "main_sha256": sha(pixels.astype("<u2")),
        "active_width": int(m.active_width),
        "weighted_visible_sum": sha(m.weighted_visible_sum.astype("<f8")),
        "weighted_infrared_sum": sha(m.weighted_infrared_sum.astype("<f8")),
        "weight_sum": sha(m.weight_sum.astype("<f8")),
        "mean_visible": sha(m.calibration_mean_visible.astype("<f4")),
        "mean_infrared": sha(m.calibration_mean_infrared.astype("<f4")),
        "accepted": sha(m.accepted_samples.astype("<u4")),
        "scaled_input": sha(s.scaled_input.astype("<f4")),
        "input_scale": sha(s.input_scale.astype("<f4")),
        "scale_denominator": sha(s.scale_denominator.astype("<f4")),
        "scale_numerator": sha(s.scale_numerator.astype("<f4")),
        "scale_add_denominator": sha(s.scale_add_denominator.astype("<f4")),
        "scale_add_numerator": sha(s.scale_add_numerator.astype("<f4")),
        "payloads": hashlib.sha256(b"".join(s.record_payloads)).hexdigest(),
    }


def prepass_vector(R, pixels, params) -> dict:
    """The reference's 0x40 record (prepass.py:96-236) and the profile's
    fail-closed base (profile.py:48-55) for one prepass frame."""

    contracts, prepass, profile = R["contracts"], R["prepass"], R["profile"]
    frame = contracts.RGBI16Frame(pixels, contracts.AcquisitionEpoch.PREPASS, 285, "pp")
    keys = ("valid_ir_threshold", "alpha_scale", "ratio_minimum", "ratio_maximum",
            "zero_denominator_alpha", "completion_marker")
    reduction = prepass.reduce_prepass_frame(
        frame, parameters=prepass.PrepassParameters(**dict(zip(keys, params))),
        frame_index=0, evidence_id="pp")
    vector = {
        "prepass_sha256": sha(pixels.astype("<u2")),
        "parameters": list(params),
        "payload_hex": reduction.record.payload.hex(),
        "valid_pixels": int(reduction.valid_pixels),
        "fully_valid_tiles": int(reduction.fully_valid_tiles),
    }
    try:
        base = profile._base_primary(reduction.record)
        vector["base_primary_bits"] = int(np.asarray(base, np.float32).view(np.uint32))
    except ValueError as error:
        vector["base_primary_error"] = str(error)
    return vector


def prepass_vectors(R) -> dict:
    out = {}
    for name, (pixels, params) in cases.prepass_cases().items():
        out[name] = prepass_vector(R, pixels, params)
        print("prepass", name, out[name]["payload_hex"][:16], flush=True)
    return out


def main(argv: list[str]) -> int:
    R = load_reference()
    target = HERE / "reference_vectors.json"
    if argv == ["--prepass-only"]:
        # Refresh only the prepass records; the other vectors are unchanged.
        out = json.loads(target.read_text())
        out["prepass"] = prepass_vectors(R)
        target.write_text(json.dumps(out, indent=1, sort_keys=True) + "\n")
        return 0
    out = {"jobs": {}, "streams": {}, "producer": {}}
    out["prepass"] = prepass_vectors(R)
    for name, (main_px, prepass) in cases.job_cases().items():
        out["jobs"][name] = job_vector(R, main_px, prepass)
        print("job", name, out["jobs"][name]["output_sha256"][:16], flush=True)
    for name, (pixels, spec, state) in cases.stream_cases().items():
        out["streams"][name] = stream_vector(R, pixels, spec, state)
        print("stream", name, out["streams"][name]["output_sha256"][:16], flush=True)
    for height, width in cases.PRODUCER_CASES:
        key = f"{height}x{width}"
        out["producer"][key] = producer_vector(R, cases.producer_frames(height, width))
        print("producer", key, flush=True)
    target.write_text(json.dumps(out, indent=1, sort_keys=True) + "\n")
    return 0


if __name__ == "__main__":
    raise SystemExit(main(sys.argv[1:]))
