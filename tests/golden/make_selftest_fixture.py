"""Write dice_amd/selftest_expected.json from reference_vectors.json.

The HIP self-test compares the synthetic 48x64 job against the reference
CPU path's recorded result (tests/golden/reference_vectors.json, entry
selftest_48x64, produced by make_reference_vectors.py).
"""

from __future__ import annotations

import json
from pathlib import Path

HERE = Path(__file__).resolve().parent
TARGET = HERE.parents[1] / "digital-fauxice_amd" / "dice_amd" / "selftest_expected.json"
FIELDS = ("main_sha256", "output_sha256", "output_samples_sha256", "attempted_pixels",
          "written_pixels", "public_rng_advances", "final_rng_state", "changed_pixels",
          "score_plane_sha256", "score_floor_bits", "at_floor_sha256", "changed_sha256")


def main() -> int:
    vectors = json.loads((HERE / "reference_vectors.json").read_text())
    entry = vectors["jobs"]["selftest_48x64"]
    TARGET.write_text(json.dumps({k: entry[k] for k in FIELDS}, indent=1, sort_keys=True) + "\n")
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
