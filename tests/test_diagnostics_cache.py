"""The diagnostics-cache export against the reference's own cache.

`tests/golden/synthetic_lama_diagnostics_cache.json` is the manifest the
reference wrote for `hybrid/examples/synthetic-lama` (256x256, cpu-fast,
core 0.3.0).  For the same arrays and the same binding, the exported
manifest must be that file byte for byte.  This pins all of the following:

* the artifact documents, including the `.npy` file hashes;
* the score-floor bits;
* the canonicalization table;
* the canonical JSON encoding.

On the GPU, the HIP path's own planes export to the same four artifacts;
only the backend fields of the binding differ.
"""

from __future__ import annotations

import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

import cases
import oracle_lib as O
from dice_amd import diagnostics_cache as dc
from dice_amd.engine import ProcessingDiagnostics, ProcessingResult, StreamingReplayResult
from dice_amd.hip_backend.process import stable_profile_params
from dice_amd.profile import DEFAULT_PROFILE

GOLDEN = Path(__file__).resolve().parent / "golden" / "synthetic_lama_diagnostics_cache.json"


def golden_document() -> dict:
    return json.loads(GOLDEN.read_bytes())


def golden_binding(**override) -> dc.CacheBinding:
    b = golden_document()["binding"]
    main, prepass = cases.lama_frames()
    fields = dict(
        profile_id=b["profile_id"], requested_backend=b["requested_backend"],
        used_backend=b["used_backend"], core_version=b["core_version"],
        core_source_manifest_sha256=b["core_source_manifest_sha256"],
        provenance_assertion_id=b["provenance_assertion_id"],
        provenance_assertion_sha256=b["provenance_assertion_sha256"])
    fields.update(override)
    return dc.binding_for(main, prepass, **fields)


def oracle_result() -> ProcessingResult:
    main, prepass = cases.lama_frames()
    r = O.process_job(main, prepass, stable_profile_params(),
                      DEFAULT_PROFILE.prepass_parameters(), diagnostics=True)
    replay = StreamingReplayResult(
        shape=r.output.shape, startup=None, attempted_pixels=r.attempted_pixels,
        written_pixels=r.written_pixels, public_rng_advances=r.public_rng_advances,
        final_rng_state=r.final_rng_state, output_sha256=r.output_sha256,
        changed_pixels=r.changed_pixels)
    diagnostics = ProcessingDiagnostics(
        score_plane=r.score_plane.astype(np.float32), score_floor=np.float32(0.02),
        at_floor_mask=r.at_floor_mask.astype(bool), changed_mask=r.changed_mask.astype(bool))
    return ProcessingResult(output_rgb16=r.output, replay=replay,
                            profile_id=DEFAULT_PROFILE.profile_id, diagnostics=diagnostics)


def test_binding_hashes_the_example_inputs():
    """The regenerated example inputs are the reference's (acquisition.json)."""

    b = golden_binding().document()
    assert b == golden_document()["binding"]


def test_manifest_is_the_reference_manifest_byte_for_byte():
    document, _ = dc.manifest(oracle_result(), golden_binding())
    assert dc.canonical_json_bytes(document) == GOLDEN.read_bytes()


def test_saved_cache_files_match_their_manifest(tmp_path):
    target = tmp_path / "cache"
    document = dc.save_diagnostics_cache(target, processing_result=oracle_result(),
                                         binding=golden_binding())
    names = sorted(p.name for p in target.iterdir())
    assert names == sorted([dc.MANIFEST_FILENAME] + [f for f, _, _ in dc.ARTIFACTS.values()])
    for name, artifact in document["artifacts"].items():
        data = (target / artifact["filename"]).read_bytes()
        assert hashlib.sha256(data).hexdigest() == artifact["file_sha256"]
        array = np.load(target / artifact["filename"], allow_pickle=False)
        assert dc.raw_sha256(array) == artifact["raw_sha256"]
    assert (target / dc.MANIFEST_FILENAME).read_bytes() == GOLDEN.read_bytes()
    assert (target.stat().st_mode & 0o777) == 0o700


def test_cache_refuses_to_overwrite(tmp_path):
    (tmp_path / "cache").mkdir()
    (tmp_path / "cache" / "other").write_bytes(b"x")
    with pytest.raises(dc.DiagnosticsCacheError, match="overwrite"):
        dc.save_diagnostics_cache(tmp_path / "cache", processing_result=oracle_result(),
                                  binding=golden_binding())


def test_cache_rejects_inconsistent_inputs():
    result = oracle_result()
    bad_floor = ProcessingDiagnostics(
        score_plane=result.diagnostics.score_plane, score_floor=np.float32(0.02),
        at_floor_mask=~result.diagnostics.at_floor_mask,
        changed_mask=result.diagnostics.changed_mask)
    with pytest.raises(dc.DiagnosticsCacheError, match="at-floor"):
        dc.manifest(ProcessingResult(result.output_rgb16, result.replay, result.profile_id,
                                     bad_floor), golden_binding())
    with pytest.raises(dc.DiagnosticsCacheError, match="impossible backend"):
        golden_binding(requested_backend="cpu", used_backend="hip").document()
    with pytest.raises(dc.DiagnosticsCacheError, match="SHA-256"):
        golden_binding(core_source_manifest_sha256="not-a-digest").document()


@pytest.mark.gpu
def test_hip_planes_export_to_the_reference_artifacts(tmp_path):
    from dice_amd.hip_backend.process import process_hip
    from test_hip_parity import make_job

    main, prepass = cases.lama_frames()
    result = process_hip(make_job(main, prepass), export_diagnostics=True)
    binding = golden_binding(requested_backend="hip", used_backend="hip")
    document = dc.save_diagnostics_cache(tmp_path / "hip", processing_result=result,
                                         binding=binding)
    golden = golden_document()
    assert document["artifacts"] == golden["artifacts"]
    for key in ("diagnostics", "hash_canonicalization", "pure_output", "schema"):
        assert document[key] == golden[key]
    differs = {k for k in golden["binding"] if document["binding"][k] != golden["binding"][k]}
    assert differs == {"backend_reason", "requested_backend", "used_backend"}
    assert document["binding"]["backend_reason"] == "explicit HIP request; self-test passed"
