"""pytest configuration: markers, import paths and shared fixtures."""

from __future__ import annotations

import json
import os
import sys
import tempfile
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "digital-fauxice_amd"))
sys.path.insert(0, str(ROOT / "tests"))

GOLDEN = ROOT / "tests" / "golden"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) device")
    # the tuner's winners persist in a file: keep the test session's apart
    if "DICE_TUNING_CACHE" not in os.environ:
        os.environ["DICE_TUNING_CACHE"] = str(
            Path(tempfile.mkdtemp(prefix="dice-tuning-")) / "tuning.tsv")


@pytest.fixture(scope="session")
def reference_vectors() -> dict:
    return json.loads((GOLDEN / "reference_vectors.json").read_text())


def hip_device_or_skip():
    """Skip (with the reason) when no HIP device is usable."""

    from dice_amd.hip_backend import HipBackendUnavailable, hip_device_summary

    try:
        return hip_device_summary()
    except HipBackendUnavailable as error:
        pytest.skip(f"HIP device unavailable: {error}")
