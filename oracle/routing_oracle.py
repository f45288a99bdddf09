"""Oracle (test infrastructure only): at-floor routing on the CPU.

A restatement of `route_at_floor_mask` and `routing_json_document`
(hybrid/src/fauxce_hybrid/routing.py:447-687, 690-783) over SciPy's
`ndimage` routines, the same library calls the reference makes:

* provisional regions are the 8-connected components of the at-floor mask,
  numbered in raster order of their first pixel (`ndimage.label`);
* each region has an area, a half-open bounding box, and the largest
  chessboard distance to in-frame background of any of its pixels
  (`distance_transform_cdt`, unpadded);
* a region is routed by `area >= min_area` or `radius >= min_radius`.  A
  component that holds every perimeter pixel is measured but never routed;
* the routed union is dilated by a square of radius `margin` (separable
  maximum filters clipped to the frame), minus the perimeter component;
* final regions are the 8-connected components of that mask.  Those whose
  box plus the 128 px crop margin exceeds 2048 x 2048 are split into
  1792 px grid tiles and re-labelled per tile.  Fragments are numbered by
  (bbox, creation order).

Only tests import this module, and only as the checker.
"""

from __future__ import annotations

import math
from fractions import Fraction

import numpy as np
from scipy import ndimage

CROP_MARGIN = 128
MAX_CROP_PIXELS = 4_194_304
TILE_EDGE = math.isqrt(MAX_CROP_PIXELS) - 2 * CROP_MARGIN
EIGHT = np.ones((3, 3), dtype=bool)


def _boxes(labels, count):
    out = []
    for s in ndimage.find_objects(labels, max_label=count):
        out.append((s[0].start, s[1].start, s[0].stop, s[1].stop))
    return out


def _final_labels(synth):
    raw = np.zeros(synth.shape, np.int32)
    count = int(ndimage.label(synth, structure=EIGHT, output=raw))
    height, width = synth.shape
    fragments = np.zeros(synth.shape, np.int32)
    boxes = []
    for raw_id, (y0, x0, y1, x1) in enumerate(_boxes(raw, count), start=1):
        crop = ((min(height, y1 + CROP_MARGIN) - max(0, y0 - CROP_MARGIN))
                * (min(width, x1 + CROP_MARGIN) - max(0, x0 - CROP_MARGIN)))
        if crop <= MAX_CROP_PIXELS:
            view = fragments[y0:y1, x0:x1]
            view[raw[y0:y1, x0:x1] == raw_id] = len(boxes) + 1
            boxes.append((y0, x0, y1, x1))
            continue
        for ty in range(y0, y1, TILE_EDGE):
            for tx in range(x0, x1, TILE_EDGE):
                ey, ex = min(ty + TILE_EDGE, y1), min(tx + TILE_EDGE, x1)
                part = raw[ty:ey, tx:ex] == raw_id
                local = np.zeros(part.shape, np.int32)
                n = int(ndimage.label(part, structure=EIGHT, output=local))
                for lid, (a, b, c, d) in enumerate(_boxes(local, n), start=1):
                    fragments[ty:ey, tx:ex][local == lid] = len(boxes) + 1
                    boxes.append((ty + a, tx + b, ty + c, tx + d))
    order = sorted(range(1, len(boxes) + 1), key=lambda f: (*boxes[f - 1], f))
    remap = np.zeros(len(boxes) + 1, np.int32)
    for new, old in enumerate(order, start=1):
        remap[old] = new
    return remap[fragments], [boxes[f - 1] for f in order]


def route(mask, *, min_area=400, min_radius=5, margin=4, max_synth_fraction=0.02):
    """Returns (routing document, provisional labels, synthesis mask, final labels)."""

    mask = np.ascontiguousarray(mask, dtype=bool)
    height, width = mask.shape
    if mask.all():
        raise ValueError("at_floor_mask covers the entire frame")
    labels = np.zeros(mask.shape, np.int32)
    count = int(ndimage.label(mask, structure=EIGHT, output=labels))
    areas = np.bincount(labels.ravel(), minlength=count + 1)[1:]
    if count:
        distance = ndimage.distance_transform_cdt(mask, metric="chessboard")
        radii = np.asarray(ndimage.maximum(distance, labels=labels,
                                           index=np.arange(1, count + 1)), np.int64)
        radii = radii.reshape(count)
    else:
        radii = np.zeros(0, np.int64)
    boxes = _boxes(labels, count)

    def reason(area, radius):
        a, r = area >= min_area, radius >= min_radius
        return ("area+chessboard_radius" if a and r else "area" if a
                else "chessboard_radius" if r else "not_routed")

    reasons = [reason(int(a), int(r)) for a, r in zip(areas, radii)]
    corner = int(labels[0, 0])
    perimeter = None
    if corner and all(np.all(edge == corner) for edge in
                      (labels[0], labels[-1], labels[:, 0], labels[:, -1])):
        perimeter = corner
    routed = [reasons[i] != "not_routed" and i + 1 != perimeter for i in range(count)]
    lookup = np.zeros(count + 1, bool)
    lookup[1:] = routed
    union = lookup[labels]
    rv, rh = min(margin, height - 1), min(margin, width - 1)
    synth = ndimage.maximum_filter1d(union, size=2 * rv + 1, axis=0, mode="constant", cval=0)
    synth = ndimage.maximum_filter1d(synth, size=2 * rh + 1, axis=1, mode="constant", cval=0)
    synth = np.ascontiguousarray(synth, dtype=bool)
    halo = 0
    if perimeter is not None:
        excluded = labels == perimeter
        halo = int(np.count_nonzero(synth & excluded))
        synth[excluded] = False
    final, final_boxes = _final_labels(synth)
    nfinal = len(final_boxes)

    inside = final > 0
    prov_in, fin_in = labels[inside], final[inside]
    floor_in = prov_in > 0
    overlap = np.bincount(prov_in[floor_in], minlength=count + 1)
    final_floor = np.bincount(fin_in[floor_in], minlength=nfinal + 1)
    pairs = sorted({(int(p), int(f)) for p, f in zip(prov_in[floor_in], fin_in[floor_in])})
    finals_of = [[] for _ in range(count)]
    direct = [[] for _ in range(nfinal)]
    absorbed = [[] for _ in range(nfinal)]
    for p, f in pairs:
        finals_of[p - 1].append(f)
        (direct if routed[p - 1] else absorbed)[f - 1].append(p)

    regions = []
    for i in range(count):
        rid = i + 1
        synthesized = int(overlap[rid])
        if rid == perimeter:
            disposition, absorbed_count = "perimeter_excluded", 0
        elif routed[i]:
            disposition, absorbed_count = "routed", 0
        elif synthesized:
            disposition, absorbed_count = "absorbed", synthesized
        else:
            disposition, absorbed_count = "pristine", 0
        regions.append({
            "id": rid, "bbox_yxyx_half_open": list(boxes[i]), "area": int(areas[i]),
            "max_chessboard_radius": int(radii[i]), "routed_reason": reasons[i],
            "disposition": disposition, "final_component_ids": finals_of[i],
            "absorbed_pixel_count": absorbed_count})
    final_area = np.bincount(final.ravel(), minlength=nfinal + 1)
    finals = [{
        "id": f, "bbox_yxyx_half_open": list(final_boxes[f - 1]), "area": int(final_area[f]),
        "direct_region_ids": direct[f - 1], "absorbed_region_ids": absorbed[f - 1],
        "non_floor_synthesis_pixels": int(final_area[f] - final_floor[f])}
        for f in range(1, nfinal + 1)]
    synth_count = int(np.count_nonzero(synth))
    limit = Fraction(str(float(max_synth_fraction)))
    perimeter_pixels = (width if height == 1 else height if width == 1
                        else 2 * (height + width) - 4)
    excluded_doc = None
    if perimeter is not None:
        r = regions[perimeter - 1]
        excluded_doc = {
            "id": perimeter, "bbox_yxyx_half_open": r["bbox_yxyx_half_open"],
            "area": r["area"], "max_chessboard_radius": r["max_chessboard_radius"],
            "routed_reason_by_thresholds": r["routed_reason"],
            "perimeter_pixel_count": perimeter_pixels,
            "predicate": "contains_every_frame_perimeter_pixel",
            "content_interpretation": "unverified"}
    document = {
        "schema": "fauxce-hybrid-routing-v1",
        "image_shape": [height, width],
        "policy": {
            "connectivity": 8, "distance_metric": "chessboard",
            "distance_border_semantics": "unpadded_in_frame_background",
            "min_area": min_area, "min_radius": min_radius, "margin": margin,
            "dilation": "single_square_chebyshev_footprint",
            "max_synth_fraction": float(max_synth_fraction), "thresholds_provisional": True,
            "perimeter_rule": "exclude_component_covering_full_perimeter",
            "final_id_sort_key": ["y0", "x0", "y1", "x1", "raw_id"],
            "final_component_partition": {
                "algorithm": "bbox_grid_8_connected_fragments", "crop_margin": CROP_MARGIN,
                "max_crop_pixels": MAX_CROP_PIXELS, "max_core_tile_edge": TILE_EDGE}},
        "counts": {
            "frame_pixels": height * width,
            "at_floor_pixels": int(np.count_nonzero(mask)),
            "directly_routed_pixels": int(sum(r["area"] for r in regions
                                              if r["disposition"] == "routed")),
            "perimeter_excluded_regions": int(perimeter is not None),
            "perimeter_excluded_pixels": int(areas[perimeter - 1]) if perimeter else 0,
            "perimeter_suppressed_halo_pixels": halo,
            "provisional_regions": count, "final_regions": nfinal,
            "synthesis_pixels": synth_count,
            "non_floor_synthesis_pixels": int(np.count_nonzero(synth & ~mask))},
        "synthesis_fraction": synth_count / (height * width),
        "within_synthesis_budget": synth_count * limit.denominator
        <= height * width * limit.numerator,
        "perimeter_excluded_region": excluded_doc,
        "provisional_regions": regions,
        "final_regions": finals,
    }
    return document, labels, synth, final.astype(np.int32)
