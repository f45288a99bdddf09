"""Pins the routing oracle (oracle/routing_oracle.py) before it checks the
device routing.

* The reference's routing census of `hybrid/examples/synthetic-lama`
  (tests/golden/synthetic_lama_routing.json), recomputed from the at-floor
  mask of that example.  That census predates the `final_component_partition`
  policy entry; everything else must match.
* The known answers of the reference's routing tests
  (hybrid/tests/test_routing.py): reason attribution and radii of reviewed
  rectangles, chessboard radius boundaries, diagonal connectivity,
  single-pixel dilation, frame clipping, the full-perimeter exclusion and its
  halo suppression, partial absorption, a bridge bound to two final regions,
  bbox-ordered final ids, the oversized-region partition, and the rational
  budget boundaries.
"""

from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np
import pytest

import cases
import oracle_lib as O
from dice_amd.hip_backend.process import stable_profile_params
from dice_amd.profile import DEFAULT_PROFILE

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import routing_oracle as R  # noqa: E402

GOLDEN = Path(__file__).resolve().parent / "golden" / "synthetic_lama_routing.json"


def lama_at_floor() -> np.ndarray:
    main, prepass = cases.lama_frames()
    r = O.process_job(main, prepass, stable_profile_params(),
                      DEFAULT_PROFILE.prepass_parameters(), diagnostics=True)
    return r.at_floor_mask.astype(bool)


def test_lama_routing_census():
    document, _, synth, final = R.route(lama_at_floor())
    golden = json.loads(GOLDEN.read_text())
    partition = document["policy"].pop("final_component_partition")
    assert partition == {"algorithm": "bbox_grid_8_connected_fragments", "crop_margin": 128,
                         "max_crop_pixels": 4_194_304, "max_core_tile_edge": 1_792}
    assert document == golden
    assert int(synth.sum()) == 840 and int(final.max()) == 1


def rectangle(height, width):
    mask = np.zeros((height + 4, width + 4), bool)
    mask[2:2 + height, 2:2 + width] = True
    return mask


@pytest.mark.parametrize("height,width,area,radius,reason", [
    (3, 3, 9, 2, "not_routed"),
    (1, 400, 400, 1, "area"),
    (9, 9, 81, 5, "chessboard_radius"),
    (30, 30, 900, 15, "area+chessboard_radius"),
    (500, 12, 6000, 6, "area+chessboard_radius"),
    (12, 8, 96, 4, "not_routed"),
    (12, 10, 120, 5, "chessboard_radius"),
])
def test_reviewed_rectangles(height, width, area, radius, reason):
    doc, *_ = R.route(rectangle(height, width), margin=0, max_synth_fraction=1.0)
    (region,) = doc["provisional_regions"]
    assert (region["area"], region["max_chessboard_radius"], region["routed_reason"]) == (
        area, radius, reason)
    assert doc["counts"]["synthesis_pixels"] == (0 if reason == "not_routed" else area)


def test_diagonal_and_edge_geometry():
    diagonal = np.zeros((34, 34), bool)
    diagonal[np.arange(2, 32), np.arange(2, 32)] = True
    doc, *_ = R.route(diagonal, margin=0, max_synth_fraction=1.0)
    assert doc["provisional_regions"][0]["bbox_yxyx_half_open"] == [2, 2, 32, 32]
    assert doc["provisional_regions"][0]["max_chessboard_radius"] == 1

    corner = np.zeros((30, 30), bool)
    corner[:9, :9] = True
    doc, _, synth, _ = R.route(corner, max_synth_fraction=1.0)
    expected = np.zeros_like(corner)
    expected[:13, :13] = True
    assert np.array_equal(synth, expected)
    assert doc["provisional_regions"][0]["max_chessboard_radius"] == 9

    dot = np.zeros((21, 21), bool)
    dot[10, 10] = True
    _, _, synth, _ = R.route(dot, min_area=1, min_radius=100, max_synth_fraction=1.0)
    expected = np.zeros_like(dot)
    expected[6:15, 6:15] = True
    assert np.array_equal(synth, expected)


def perimeter_ring():
    mask = np.zeros((30, 40), bool)
    mask[0, :] = mask[-1, :] = True
    mask[:, 0] = mask[:, -1] = True
    return mask


def test_perimeter_rules():
    doc, *_ = R.route(perimeter_ring(), min_area=1, min_radius=100, max_synth_fraction=1.0)
    assert doc["provisional_regions"][0]["disposition"] == "perimeter_excluded"
    assert doc["perimeter_excluded_region"]["perimeter_pixel_count"] == 136
    assert doc["counts"]["synthesis_pixels"] == 0

    gap = perimeter_ring()
    gap[0, 20] = False
    doc, *_ = R.route(gap, min_area=1, min_radius=100, margin=0, max_synth_fraction=1.0)
    assert doc["provisional_regions"][0]["disposition"] == "routed"

    halo = perimeter_ring()
    halo[2:11, 15:24] = True
    doc, labels, synth, _ = R.route(halo, max_synth_fraction=1.0)
    assert doc["counts"]["perimeter_suppressed_halo_pixels"] > 0
    assert not np.any(synth & (labels == doc["perimeter_excluded_region"]["id"]))


def test_absorption_and_final_ids():
    mask = np.zeros((50, 70), bool)
    mask[10:19, 10:19] = mask[10:19, 25:34] = True
    mask[22:25, 14:17] = True
    doc, *_ = R.route(mask, max_synth_fraction=1.0)
    absorbed = [r for r in doc["provisional_regions"] if r["disposition"] == "absorbed"]
    assert len(doc["final_regions"]) == 1 and len(absorbed) == 1
    assert absorbed[0]["absorbed_pixel_count"] == 3

    bridge = np.zeros((40, 70), bool)
    bridge[10:19, 10:19] = bridge[10:19, 40:49] = True
    bridge[22, 22:37] = True
    doc, *_ = R.route(bridge, max_synth_fraction=1.0)
    b = [r for r in doc["provisional_regions"] if r["routed_reason"] == "not_routed"][0]
    assert (b["disposition"], b["absorbed_pixel_count"], b["final_component_ids"]) == (
        "absorbed", 2, [1, 2])

    two = np.zeros((50, 50), bool)
    two[30:39, 35:44] = two[4:13, 5:14] = True
    doc, _, _, final = R.route(two, margin=0, max_synth_fraction=1.0)
    assert [f["bbox_yxyx_half_open"] for f in doc["final_regions"]] == [
        [4, 5, 13, 14], [30, 35, 39, 44]]
    assert final[4, 5] == 1 and final[30, 35] == 2


def test_oversized_partition():
    mask = np.zeros((2200, 2200), bool)
    mask[100:2100, 100:2100] = True
    doc, _, synth, final = R.route(mask, min_area=1, min_radius=100, margin=0,
                                   max_synth_fraction=1.0)
    assert np.array_equal(synth, mask)
    assert len(doc["final_regions"]) == 4
    assert doc["provisional_regions"][0]["final_component_ids"] == [1, 2, 3, 4]
    assert sum(f["area"] for f in doc["final_regions"]) == int(mask.sum())


def test_budget_boundaries():
    two = np.zeros((10, 10), bool)
    two[1, 1] = two[8, 8] = True
    doc, *_ = R.route(two, min_area=1, min_radius=100, margin=0)
    assert doc["synthesis_fraction"] == 0.02 and doc["within_synthesis_budget"]
    three = two.copy()
    three[5, 5] = True
    assert not R.route(three, min_area=1, min_radius=100, margin=0)[0]["within_synthesis_budget"]
    one = np.zeros((7, 7), bool)
    one[3, 3] = True
    doc, *_ = R.route(one, min_area=1, min_radius=100, margin=0)
    assert doc["synthesis_fraction"] == 1 / 49 and not doc["within_synthesis_budget"]
    empty, *_ = R.route(np.zeros((10, 10), bool), max_synth_fraction=0.0)
    assert empty["within_synthesis_budget"] and empty["provisional_regions"] == []
