"""At-floor routing on the device against the routing oracle.

The oracle (oracle/routing_oracle.py, pinned by test_routing_oracle.py) and
`dice_amd.hip_backend.routing` must produce the same census document,
identical region tables, id lists and counts, and identical planes
(provisional labels, synthesis mask, final labels).  The masks cover the
synthetic-lama example (whose census is also the reference's golden), the
reference's reviewed geometries, perimeter cases, the oversized-region
partition and random masks with many components.
"""

from __future__ import annotations

import json
import sys
from pathlib import Path

import numpy as np
import pytest

import cases
from dice_amd.hip_backend import routing as hip_routing
from dice_amd.hip_backend.process import process_hip

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import routing_oracle as R  # noqa: E402

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden" / "synthetic_lama_routing.json"


def run_both(mask, **policy):
    want, labels, synth, final = R.route(mask, **policy)
    result = hip_routing.route_at_floor_mask(mask, hip_routing.RoutingPolicy(**policy))
    got = hip_routing.routing_json_document(result)
    assert got == want
    p, s, f = result.planes()
    assert np.array_equal(p, labels)
    assert np.array_equal(s, synth)
    assert np.array_equal(f, final)
    return result


def test_lama_at_floor_mask_from_the_device():
    from test_hip_parity import make_job

    main, prepass = cases.lama_frames()
    mask = process_hip(make_job(main, prepass), export_diagnostics=True).diagnostics.at_floor_mask
    result = run_both(mask)
    document = hip_routing.routing_json_document(result)
    document["policy"].pop("final_component_partition")
    assert document == json.loads(GOLDEN.read_text())


def blobs(shape, seed, threshold, salt=0.0):
    rng = np.random.default_rng(seed)
    field = rng.standard_normal(shape)
    for axis in (0, 1):  # cheap low-pass: running means along both axes
        field = np.cumsum(field, axis=axis)
        k = 6
        field = (np.roll(field, -k, axis=axis) - np.roll(field, k, axis=axis)) / (2 * k)
    mask = field > threshold * field.std()
    if salt:
        mask |= rng.random(shape) < salt
    return mask


@pytest.mark.parametrize("seed,shape,threshold,salt", [
    (1, (300, 400), 1.2, 0.0),
    (2, (257, 333), 0.8, 0.002),
    (3, (512, 128), 1.5, 0.01),
    (4, (64, 1000), 1.0, 0.0),
    (5, (1, 300), 0.5, 0.05),
    (6, (300, 1), 0.5, 0.05),
])
def test_random_masks(seed, shape, threshold, salt):
    mask = blobs(shape, seed, threshold, salt)
    if mask.all():
        mask[0, 0] = False
    run_both(mask)
    run_both(mask, min_area=20, min_radius=2, margin=2, max_synth_fraction=0.5)
    run_both(mask, min_area=1, min_radius=100, margin=0, max_synth_fraction=1.0)


def test_reviewed_geometries():
    rect = np.zeros((504, 16), bool)
    rect[2:502, 2:14] = True
    run_both(rect, margin=0, max_synth_fraction=1.0)
    diagonal = np.zeros((34, 34), bool)
    diagonal[np.arange(2, 32), np.arange(2, 32)] = True
    run_both(diagonal, margin=0, max_synth_fraction=1.0)
    corner = np.zeros((30, 30), bool)
    corner[:9, :9] = True
    run_both(corner, max_synth_fraction=1.0)
    absorb = np.zeros((50, 70), bool)
    absorb[10:19, 10:19] = absorb[10:19, 25:34] = True
    absorb[22:25, 14:17] = True
    run_both(absorb, max_synth_fraction=1.0)
    bridge = np.zeros((40, 70), bool)
    bridge[10:19, 10:19] = bridge[10:19, 40:49] = True
    bridge[22, 22:37] = True
    run_both(bridge, max_synth_fraction=1.0)
    run_both(np.zeros((10, 10), bool))


def test_perimeter_cases():
    ring = np.zeros((30, 40), bool)
    ring[0, :] = ring[-1, :] = True
    ring[:, 0] = ring[:, -1] = True
    run_both(ring, min_area=1, min_radius=100, max_synth_fraction=1.0)
    halo = ring.copy()
    halo[2:11, 15:24] = True
    r = run_both(halo, max_synth_fraction=1.0)
    assert r.perimeter_suppressed_halo_pixel_count > 0
    gap = ring.copy()
    gap[0, 20] = False
    run_both(gap, min_area=1, min_radius=100, margin=0, max_synth_fraction=1.0)
    near_all = np.ones((11, 13), bool)
    near_all[5, 6] = False
    run_both(near_all)


def test_oversized_partition():
    square = np.zeros((2200, 2200), bool)
    square[100:2100, 100:2100] = True
    square[50, 50] = True
    r = run_both(square, min_area=1, min_radius=100, margin=0, max_synth_fraction=1.0)
    assert len(r.final_regions) == 5
    # a comb joined only below the first tile row: its teeth become separate
    # fragments inside one tile
    comb = np.zeros((2200, 2200), bool)
    comb[2050:2100, 100:2100] = True
    for x0 in (100, 600, 1100, 1600, 2050):
        comb[100:2100, x0:x0 + 50] = True
    r = run_both(comb, min_area=1, min_radius=100, margin=2, max_synth_fraction=1.0)
    assert len(r.final_regions) > 4


def test_contract_errors():
    with pytest.raises(hip_routing.NoHealthyContextError):
        hip_routing.route_at_floor_mask(np.ones((6, 9), bool))
    with pytest.raises(TypeError):
        hip_routing.route_at_floor_mask(np.zeros((4, 4), np.uint8))
    with pytest.raises(ValueError):
        hip_routing.RoutingPolicy(min_area=0)
