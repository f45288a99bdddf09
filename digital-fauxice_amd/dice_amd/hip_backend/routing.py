"""At-floor region routing on the HIP device.

The hybrid consumer of the core routes the at-floor mask to synthesis
regions with `route_at_floor_mask` (hybrid/src/fauxce_hybrid/routing.py:
447-687) and records the decision as a `fauxce-hybrid-routing-v1` census
(`routing_json_document`, :690-783).  Here the per-pixel work runs in
libdice_hip.so (`dice_route_at_floor`):

* the 8-connected components of the mask and of the synthesis mask;
* their areas and boxes, and the chessboard radii;
* the perimeter test, the square dilation and the final partition;
* the region pairs.

This module turns the resulting tables into the reference's region records
and census: reasons, dispositions, id lists, counts, the rational budget.
The three planes stay in HBM until they are asked for.
"""

from __future__ import annotations

import ctypes as C
import json
import math
from dataclasses import dataclass, field
from fractions import Fraction

import numpy as np

from .. import _abi
from .._abi import DiceFinalRegion, DiceProvisionalRegion, DiceRoutingSummary
from .engine import HipBackendUnavailable, context, library

INPAINT_CROP_MARGIN = 128
MAX_INPAINT_CROP_PIXELS = 4_194_304
MAX_SYNTHESIS_TILE_EDGE = math.isqrt(MAX_INPAINT_CROP_PIXELS) - 2 * INPAINT_CROP_MARGIN


class SynthesisBudgetExceeded(RuntimeError):
    """The routed mask exceeds the policy's synthesis budget."""


class NoHealthyContextError(RuntimeError):
    """The whole frame is at floor: no in-frame context to synthesize from."""

    def __init__(self, image_shape: tuple[int, int]) -> None:
        self.image_shape = image_shape
        self.finite_radius_sentinel = max(image_shape)
        super().__init__("at_floor_mask covers the entire frame; no healthy in-frame "
                         "context exists for bounded synthesis")


def _positive_int(value, name, minimum):
    if isinstance(value, bool) or not isinstance(value, int) or value < minimum:
        raise ValueError(f"{name} must be an integer >= {minimum}")


@dataclass(frozen=True)
class RoutingPolicy:
    """Provisional thresholds and the fail-closed synthesis budget
    (routing.py:84-107 defaults)."""

    min_area: int = 400
    min_radius: int = 5
    margin: int = 4
    max_synth_fraction: float = 0.02

    def __post_init__(self) -> None:
        _positive_int(self.min_area, "min_area", 1)
        _positive_int(self.min_radius, "min_radius", 1)
        _positive_int(self.margin, "margin", 0)
        f = self.max_synth_fraction
        if isinstance(f, bool) or not isinstance(f, (int, float)) or not math.isfinite(float(f)) \
                or not 0.0 <= float(f) <= 1.0:
            raise ValueError("max_synth_fraction must be finite and in [0, 1]")
        object.__setattr__(self, "max_synth_fraction", float(f))


@dataclass(frozen=True)
class ProvisionalRegion:
    id: int
    bbox: tuple[int, int, int, int]  # half-open y0, x0, y1, x1
    area: int
    max_chessboard_radius: int
    routed_reason: str
    disposition: str
    final_component_ids: tuple[int, ...]
    absorbed_pixel_count: int


@dataclass(frozen=True)
class FinalSynthesisRegion:
    id: int
    bbox: tuple[int, int, int, int]
    area: int
    direct_region_ids: tuple[int, ...]
    absorbed_region_ids: tuple[int, ...]
    non_floor_synth_pixel_count: int


@dataclass
class RoutingResult:
    image_shape: tuple[int, int]
    policy: RoutingPolicy
    provisional_regions: tuple[ProvisionalRegion, ...]
    final_regions: tuple[FinalSynthesisRegion, ...]
    at_floor_pixel_count: int
    directly_routed_pixel_count: int
    perimeter_excluded_region_count: int
    perimeter_excluded_pixel_count: int
    perimeter_suppressed_halo_pixel_count: int
    synthesis_pixel_count: int
    non_floor_synth_pixel_count: int
    synthesis_fraction: float
    within_synthesis_budget: bool
    _handle: object = field(default=None, repr=False)

    def planes(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(provisional labels int32, synthesis mask bool, final labels int32)."""

        height, width = self.image_shape
        provisional = np.empty((height, width), np.int32)
        synthesis = np.empty((height, width), np.uint8)
        final = np.empty((height, width), np.int32)
        code = library().dice_routing_planes(
            self._handle.value, provisional.ctypes.data_as(_abi.i32p),
            synthesis.ctypes.data_as(_abi.u8p), final.ctypes.data_as(_abi.i32p))
        if code != _abi.DICE_OK:
            raise HipBackendUnavailable(library().dice_last_error().decode(errors="replace"))
        return provisional, synthesis.astype(bool), final

    def require_safe_for_inpainting(self) -> None:
        if not self.within_synthesis_budget:
            height, width = self.image_shape
            raise SynthesisBudgetExceeded(
                f"synthesis mask exceeds configured budget: {self.synthesis_pixel_count}/"
                f"{height * width} pixels")

    def __del__(self) -> None:
        if self._handle is not None and self._handle.value:
            try:
                library().dice_routing_free(self._handle.value)
            except Exception:  # noqa: BLE001 - interpreter shutdown
                pass
            self._handle = None


def _reason(area: int, radius: int, policy: RoutingPolicy) -> str:
    by_area, by_radius = area >= policy.min_area, radius >= policy.min_radius
    if by_area and by_radius:
        return "area+chessboard_radius"
    return "area" if by_area else "chessboard_radius" if by_radius else "not_routed"


def route_at_floor_mask(at_floor_mask, policy: RoutingPolicy = RoutingPolicy()) -> RoutingResult:
    """Measure and route an at-floor mask on the device (no inpainting)."""

    if not isinstance(policy, RoutingPolicy):
        raise TypeError("policy must be a RoutingPolicy")
    mask = np.asarray(at_floor_mask)
    if mask.dtype != np.bool_:
        raise TypeError("at_floor_mask must have dtype bool")
    if mask.ndim != 2:
        raise ValueError("at_floor_mask must have shape HxW")
    if 0 in mask.shape:
        raise ValueError("at_floor_mask cannot have an empty dimension")
    height, width = (int(mask.shape[0]), int(mask.shape[1]))
    if bool(np.all(mask)):
        raise NoHealthyContextError((height, width))
    raw = np.ascontiguousarray(mask, dtype=np.uint8)
    lib = library()
    handle = C.c_void_p()
    # a margin beyond the frame dilates no further (the reference clips it)
    margin = min(policy.margin, max(height, width))
    code = lib.dice_route_at_floor(context(), raw.ctypes.data, height, width, 0,
                                   policy.min_area, policy.min_radius, margin,
                                   C.byref(handle))
    if code != _abi.DICE_OK:
        message = lib.dice_last_error().decode(errors="replace")
        if code == _abi.DICE_INVALID:
            raise ValueError(message)
        raise HipBackendUnavailable(message)
    try:
        summary = DiceRoutingSummary()
        lib.dice_routing_summary_get(handle.value, C.byref(summary))
        provisional = (DiceProvisionalRegion * max(1, summary.provisional_regions))()
        finals = (DiceFinalRegion * max(1, summary.final_regions))()
        pairs = np.zeros(max(2, 2 * summary.pairs), np.int32)
        lib.dice_routing_tables(handle.value, provisional, finals, pairs.ctypes.data_as(_abi.i32p))
    except BaseException:
        lib.dice_routing_free(handle.value)
        raise

    count, nfinal = summary.provisional_regions, summary.final_regions
    excluded = summary.perimeter_excluded_id
    finals_of: list[list[int]] = [[] for _ in range(count)]
    direct: list[list[int]] = [[] for _ in range(nfinal)]
    absorbed: list[list[int]] = [[] for _ in range(nfinal)]
    for k in range(summary.pairs):
        p, f = int(pairs[2 * k]), int(pairs[2 * k + 1])
        if p == excluded:
            raise RuntimeError("perimeter-excluded region entered synthesis mask")
        finals_of[p - 1].append(f)
        (direct if provisional[p - 1].routed else absorbed)[f - 1].append(p)

    regions = []
    for i in range(count):
        r = provisional[i]
        rid = i + 1
        if rid == excluded:
            disposition, absorbed_count = "perimeter_excluded", 0
        elif r.routed:
            disposition, absorbed_count = "routed", 0
        elif r.synthesized_pixels:
            disposition, absorbed_count = "absorbed", int(r.synthesized_pixels)
        else:
            disposition, absorbed_count = "pristine", 0
        regions.append(ProvisionalRegion(
            id=rid, bbox=(r.y0, r.x0, r.y1, r.x1), area=int(r.area),
            max_chessboard_radius=int(r.max_chessboard_radius),
            routed_reason=_reason(int(r.area), int(r.max_chessboard_radius), policy),
            disposition=disposition, final_component_ids=tuple(finals_of[i]),
            absorbed_pixel_count=absorbed_count))
    final_rows = tuple(FinalSynthesisRegion(
        id=k + 1, bbox=(finals[k].y0, finals[k].x0, finals[k].y1, finals[k].x1),
        area=int(finals[k].area), direct_region_ids=tuple(direct[k]),
        absorbed_region_ids=tuple(absorbed[k]),
        non_floor_synth_pixel_count=int(finals[k].area - finals[k].at_floor_pixels))
        for k in range(nfinal))
    synthesized = int(summary.synthesis_pixels)
    limit = Fraction(str(policy.max_synth_fraction))
    result = RoutingResult(
        image_shape=(height, width), policy=policy, provisional_regions=tuple(regions),
        final_regions=final_rows, at_floor_pixel_count=int(summary.at_floor_pixels),
        directly_routed_pixel_count=sum(r.area for r in regions if r.disposition == "routed"),
        perimeter_excluded_region_count=int(excluded > 0),
        perimeter_excluded_pixel_count=regions[excluded - 1].area if excluded else 0,
        perimeter_suppressed_halo_pixel_count=int(summary.perimeter_suppressed_halo_pixels),
        synthesis_pixel_count=synthesized,
        non_floor_synth_pixel_count=int(summary.non_floor_synthesis_pixels),
        synthesis_fraction=synthesized / (height * width),
        within_synthesis_budget=synthesized * limit.denominator
        <= height * width * limit.numerator,
        _handle=handle)
    absorbed_total = sum(r.absorbed_pixel_count for r in regions if r.disposition == "absorbed")
    if synthesized != (result.directly_routed_pixel_count + absorbed_total
                       + result.non_floor_synth_pixel_count):
        raise RuntimeError("synthesis pixel accounting is not conservative")
    if sum(r.area for r in final_rows) != synthesized:
        raise RuntimeError("final region areas do not equal synthesis pixel count")
    return result


def routing_json_document(result: RoutingResult) -> dict:
    """The versioned routing census (fauxce-hybrid-routing-v1)."""

    height, width = result.image_shape
    excluded = next((r for r in result.provisional_regions
                     if r.disposition == "perimeter_excluded"), None)
    perimeter_pixels = width if height == 1 else height if width == 1 else 2 * (height + width) - 4
    return {
        "schema": "fauxce-hybrid-routing-v1",
        "image_shape": [height, width],
        "policy": {
            "connectivity": 8,
            "distance_metric": "chessboard",
            "distance_border_semantics": "unpadded_in_frame_background",
            "min_area": result.policy.min_area,
            "min_radius": result.policy.min_radius,
            "margin": result.policy.margin,
            "dilation": "single_square_chebyshev_footprint",
            "max_synth_fraction": result.policy.max_synth_fraction,
            "thresholds_provisional": True,
            "perimeter_rule": "exclude_component_covering_full_perimeter",
            "final_id_sort_key": ["y0", "x0", "y1", "x1", "raw_id"],
            "final_component_partition": {
                "algorithm": "bbox_grid_8_connected_fragments",
                "crop_margin": INPAINT_CROP_MARGIN,
                "max_crop_pixels": MAX_INPAINT_CROP_PIXELS,
                "max_core_tile_edge": MAX_SYNTHESIS_TILE_EDGE,
            },
        },
        "counts": {
            "frame_pixels": height * width,
            "at_floor_pixels": result.at_floor_pixel_count,
            "directly_routed_pixels": result.directly_routed_pixel_count,
            "perimeter_excluded_regions": result.perimeter_excluded_region_count,
            "perimeter_excluded_pixels": result.perimeter_excluded_pixel_count,
            "perimeter_suppressed_halo_pixels": result.perimeter_suppressed_halo_pixel_count,
            "provisional_regions": len(result.provisional_regions),
            "final_regions": len(result.final_regions),
            "synthesis_pixels": result.synthesis_pixel_count,
            "non_floor_synthesis_pixels": result.non_floor_synth_pixel_count,
        },
        "synthesis_fraction": result.synthesis_fraction,
        "within_synthesis_budget": result.within_synthesis_budget,
        "perimeter_excluded_region": None if excluded is None else {
            "id": excluded.id,
            "bbox_yxyx_half_open": list(excluded.bbox),
            "area": excluded.area,
            "max_chessboard_radius": excluded.max_chessboard_radius,
            "routed_reason_by_thresholds": excluded.routed_reason,
            "perimeter_pixel_count": perimeter_pixels,
            "predicate": "contains_every_frame_perimeter_pixel",
            "content_interpretation": "unverified",
        },
        "provisional_regions": [{
            "id": r.id, "bbox_yxyx_half_open": list(r.bbox), "area": r.area,
            "max_chessboard_radius": r.max_chessboard_radius, "routed_reason": r.routed_reason,
            "disposition": r.disposition, "final_component_ids": list(r.final_component_ids),
            "absorbed_pixel_count": r.absorbed_pixel_count,
        } for r in result.provisional_regions],
        "final_regions": [{
            "id": f.id, "bbox_yxyx_half_open": list(f.bbox), "area": f.area,
            "direct_region_ids": list(f.direct_region_ids),
            "absorbed_region_ids": list(f.absorbed_region_ids),
            "non_floor_synthesis_pixels": f.non_floor_synth_pixel_count,
        } for f in result.final_regions],
    }


def serialize_routing_json(result: RoutingResult) -> bytes:
    """Canonical census bytes: UTF-8, sorted keys, compact, trailing newline."""

    text = json.dumps(routing_json_document(result), ensure_ascii=False, allow_nan=False,
                      sort_keys=True, separators=(",", ":"))
    return text.encode("utf-8") + b"\n"
