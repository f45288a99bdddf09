"""The diagnostics-cache export (``fauxce-hybrid-diagnostics-cache-v1``).

The hybrid consumer of the processing core reads a job's pure output and its
three diagnostics planes from a cache directory
(``hybrid/src/fauxce_hybrid/cache.py:651-789``, reader ``:1123-1361``):

* four non-object ``.npy`` files: the output, the score plane, the at-floor
  and the changed masks;
* ``diagnostics-cache.json``, a canonical JSON manifest.

The manifest binds the files to the inputs, the caller's provenance
assertion, the core build, the profile and the backend that produced them.
The HIP path already produces the planes on the device, in the same replay
that writes the output (``dice_diagnostics``).  This module writes them in
that format, byte for byte what the reference writer produces for the same
arrays and binding.

Format rules kept identical, because the reader checks them:

* artifact names, file names, dtypes and ``source`` strings;
* raw hashes over little-endian C-order bytes, and file hashes over the
  ``.npy`` bytes NumPy writes;
* the score floor as the little-endian bits of its float32;
* the hash-canonicalization table;
* JSON with sorted keys, compact separators, ASCII only and a trailing
  newline.

The cache never overwrites: the destination must be absent or empty.  Files
are created exclusively with mode 0600 and the directory is restricted to
0700.
"""

from __future__ import annotations

import hashlib
import io
import json
import os
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from .engine import ProcessingResult

CACHE_SCHEMA = "fauxce-hybrid-diagnostics-cache-v1"
MANIFEST_FILENAME = "diagnostics-cache.json"

# name -> (file name, dtype string, producing field)
ARTIFACTS = {
    "pure_output_rgb16": ("output.rgb16.npy", "<u2",
                          "portable_digital_ice.ProcessingResult.output_rgb16"),
    "score_plane": ("score-plane.npy", "<f4",
                    "portable_digital_ice.ProcessingDiagnostics.score_plane"),
    "at_floor_mask": ("at-floor-mask.npy", "|b1",
                      "portable_digital_ice.ProcessingDiagnostics.at_floor_mask"),
    "changed_mask": ("changed-mask.npy", "|b1",
                     "portable_digital_ice.ProcessingDiagnostics.changed_mask"),
}

HASH_CANONICALIZATION = {
    "algorithm": "sha256",
    "array_order": "C",
    "bool": "one byte per element, NumPy |b1",
    "float32": "IEEE-754 little-endian float32 bytes",
    "json": "UTF-8, sorted keys, compact separators, ensure_ascii=true, newline",
    "uint16": "little-endian unsigned 16-bit bytes",
}

# The reference's path-free selection reasons (cache.py:80-107), plus the HIP
# backend's (the reasons dice_amd.backend.process reports).
BACKEND_REASONS = {
    ("cpu", "cpu"): "explicit CPU request",
    ("cpu-fast", "cpu-fast"): "explicit cpu-fast request; self-test passed byte parity",
    ("cuda", "cuda"): "explicit CUDA request; self-test passed",
    ("auto", "cpu-fast"): "CUDA unavailable; cpu-fast startup self-test passed byte parity",
    ("auto", "cuda"): "startup self-test passed byte parity",
    ("auto", "cpu"): "CUDA unavailable; complete job ran on exact CPU reference",
    ("hip", "hip"): "explicit HIP request; self-test passed",
    ("auto", "hip"): "startup self-test passed byte parity",
}


class DiagnosticsCacheError(ValueError):
    """The arrays or the binding cannot form a trustworthy cache."""


def backend_reason(requested_backend: str, used_backend: str) -> str:
    try:
        return BACKEND_REASONS[(requested_backend, used_backend)]
    except KeyError:
        raise DiagnosticsCacheError(
            f"impossible backend selection: requested={requested_backend!r}, "
            f"used={used_backend!r}") from None


def raw_sha256(array: np.ndarray) -> str:
    """SHA-256 of an array's C-order bytes (the cache's raw hash)."""

    return hashlib.sha256(np.ascontiguousarray(array).tobytes(order="C")).hexdigest()


def canonical_json_bytes(document: object) -> bytes:
    text = json.dumps(document, allow_nan=False, ensure_ascii=True, separators=(",", ":"),
                      sort_keys=True)
    return (text + "\n").encode("ascii")


def _sha256_hex(value: str, field: str) -> str:
    if not (isinstance(value, str) and len(value) == 64
            and all(c in "0123456789abcdef" for c in value)):
        raise DiagnosticsCacheError(f"{field} must be a lowercase 64-character SHA-256 hex digest")
    return value


@dataclass(frozen=True)
class CacheBinding:
    """What the cache is bound to (the manifest's ``binding`` object).

    The core build (``core_version``, ``core_source_manifest_sha256``) and
    the provenance assertion come from the caller, as in the reference; the
    input hashes and shapes come from the job (:func:`binding_for`)."""

    requested_backend: str
    used_backend: str
    core_source_manifest_sha256: str
    core_version: str
    main_raw_sha256: str
    main_shape: tuple[int, ...]
    prepass_raw_sha256: str
    prepass_shape: tuple[int, ...]
    profile_id: str
    provenance_assertion_id: str
    provenance_assertion_sha256: str

    def document(self) -> dict:
        for field in ("core_source_manifest_sha256", "main_raw_sha256", "prepass_raw_sha256",
                      "provenance_assertion_sha256"):
            _sha256_hex(getattr(self, field), field)
        for field in ("core_version", "profile_id", "provenance_assertion_id"):
            value = getattr(self, field)
            if not isinstance(value, str) or not value or value.strip() != value:
                raise DiagnosticsCacheError(f"{field} must be a non-empty trimmed string")
        return {
            "backend_reason": backend_reason(self.requested_backend, self.used_backend),
            "core_source_manifest_sha256": self.core_source_manifest_sha256,
            "core_version": self.core_version,
            "main_raw_sha256": self.main_raw_sha256,
            "main_shape": [int(d) for d in self.main_shape],
            "prepass_raw_sha256": self.prepass_raw_sha256,
            "prepass_shape": [int(d) for d in self.prepass_shape],
            "profile_id": self.profile_id,
            "provenance_assertion_id": self.provenance_assertion_id,
            "provenance_assertion_sha256": self.provenance_assertion_sha256,
            "requested_backend": self.requested_backend,
            "used_backend": self.used_backend,
        }


def binding_for(main_rgbi: np.ndarray, prepass_rgbi: np.ndarray, *, profile_id: str,
                requested_backend: str, used_backend: str, core_version: str,
                core_source_manifest_sha256: str, provenance_assertion_id: str,
                provenance_assertion_sha256: str) -> CacheBinding:
    """The binding of one job: its two RGBI frames hashed as little-endian
    uint16 in C order."""

    main = np.asarray(main_rgbi)
    prepass = np.asarray(prepass_rgbi)
    for role, frame in (("main", main), ("prepass", prepass)):
        if frame.dtype.kind != "u" or frame.dtype.itemsize != 2 or frame.ndim != 3 \
                or frame.shape[2] != 4:
            raise DiagnosticsCacheError(f"{role} input must be uint16 HxWx4")
    return CacheBinding(
        requested_backend=requested_backend, used_backend=used_backend,
        core_source_manifest_sha256=core_source_manifest_sha256, core_version=core_version,
        main_raw_sha256=raw_sha256(main.astype("<u2", copy=False)), main_shape=main.shape,
        prepass_raw_sha256=raw_sha256(prepass.astype("<u2", copy=False)),
        prepass_shape=prepass.shape, profile_id=profile_id,
        provenance_assertion_id=provenance_assertion_id,
        provenance_assertion_sha256=provenance_assertion_sha256)


def _npy_bytes(array: np.ndarray) -> bytes:
    buffer = io.BytesIO()
    np.save(buffer, array, allow_pickle=False)
    return buffer.getvalue()


def _artifacts(result: ProcessingResult, binding: CacheBinding) -> tuple[dict, int]:
    diagnostics = result.diagnostics
    if diagnostics is None:
        raise DiagnosticsCacheError("processing result has no exported diagnostics")
    output = np.asarray(result.output_rgb16)
    if output.dtype.kind != "u" or output.dtype.itemsize != 2 or output.ndim != 3 \
            or output.shape[2] != 3:
        raise DiagnosticsCacheError("pure output must be uint16 HxWx3")
    arrays = {
        "pure_output_rgb16": np.array(output, dtype="<u2", order="C"),
        "score_plane": np.array(diagnostics.score_plane, dtype="<f4", order="C"),
        "at_floor_mask": np.array(diagnostics.at_floor_mask, dtype="|b1", order="C"),
        "changed_mask": np.array(diagnostics.changed_mask, dtype="|b1", order="C"),
    }
    image = arrays["pure_output_rgb16"].shape[:2]
    if tuple(binding.main_shape[:2]) != image:
        raise DiagnosticsCacheError("pure output geometry does not match the bound main input")
    for name in ("score_plane", "at_floor_mask", "changed_mask"):
        if arrays[name].shape != image:
            raise DiagnosticsCacheError(f"{name} shape does not match the pure output")
    floor = np.asarray(diagnostics.score_floor)
    if floor.dtype != np.float32 or floor.ndim != 0:
        raise DiagnosticsCacheError("score floor must be a float32 scalar")
    floor_bits = int(np.asarray(floor, dtype="<f4").view("<u4")[()])
    if not np.array_equal(arrays["at_floor_mask"], arrays["score_plane"] == floor):
        raise DiagnosticsCacheError("at-floor mask does not equal score_plane == score_floor")
    if result.replay.output_sha256 != raw_sha256(arrays["pure_output_rgb16"]):
        raise DiagnosticsCacheError("replay output hash does not match the pure output bytes")
    if result.profile_id != binding.profile_id:
        raise DiagnosticsCacheError("processing result profile does not match the binding")
    return arrays, floor_bits


def manifest(result: ProcessingResult, binding: CacheBinding) -> tuple[dict, dict]:
    """The manifest document and the ``.npy`` payloads, without writing."""

    arrays, floor_bits = _artifacts(result, binding)
    payloads = {name: _npy_bytes(array) for name, array in arrays.items()}
    documents = {}
    for name in sorted(arrays):
        filename, dtype, source = ARTIFACTS[name]
        documents[name] = {
            "dtype": dtype,
            "file_sha256": hashlib.sha256(payloads[name]).hexdigest(),
            "filename": filename,
            "raw_sha256": raw_sha256(arrays[name]),
            "shape": list(arrays[name].shape),
            "source": source,
        }
    document = {
        "artifacts": documents,
        "binding": binding.document(),
        "diagnostics": {"score_floor_u32_le_bits": floor_bits},
        "hash_canonicalization": dict(HASH_CANONICALIZATION),
        "pure_output": {
            "artifact": "pure_output_rgb16",
            "generative_modification": False,
            "output_sha256": documents["pure_output_rgb16"]["raw_sha256"],
            "source": ARTIFACTS["pure_output_rgb16"][2],
        },
        "schema": CACHE_SCHEMA,
    }
    return document, payloads


def _write_new(path: Path, payload: bytes) -> None:
    flags = os.O_WRONLY | os.O_CREAT | os.O_EXCL | getattr(os, "O_NOFOLLOW", 0)
    descriptor = os.open(path, flags, 0o600)
    with os.fdopen(descriptor, "wb") as handle:
        handle.write(payload)
        handle.flush()
        os.fsync(handle.fileno())


def save_diagnostics_cache(directory: str | Path, *, processing_result: ProcessingResult,
                           binding: CacheBinding) -> dict:
    """Write one diagnostics cache; returns its manifest document.

    The destination must not exist or must be an empty directory."""

    document, payloads = manifest(processing_result, binding)
    destination = Path(directory)
    if os.path.lexists(destination):
        if destination.is_symlink() or not destination.is_dir():
            raise DiagnosticsCacheError("cache destination is not a directory")
        if any(destination.iterdir()):
            raise DiagnosticsCacheError(f"refusing to overwrite non-empty cache path: {destination}")
    else:
        destination.mkdir(mode=0o700, parents=True)
    if os.name != "nt":
        destination.chmod(0o700)
    for name, payload in payloads.items():
        _write_new(destination / ARTIFACTS[name][0], payload)
    _write_new(destination / MANIFEST_FILENAME, canonical_json_bytes(document))
    return document
