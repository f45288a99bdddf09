// dice_routing.hpp -- at-floor region routing on the device
// (hybrid/src/fauxce_hybrid/routing.py:447-687).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/dice_hip.h"
#include "dice_common.hpp"

namespace dice {

// Everything one routing decision produces: the per-region tables on the
// host, the three planes on the device.
struct RoutingOutput {
  int64_t height = 0, width = 0;
  dice_routing_summary summary{};
  std::vector<dice_provisional_region> provisional;
  std::vector<dice_final_region> finals;
  std::vector<int32_t> pairs;  // (provisional id, final id), sorted, distinct
  DeviceBuffer provisional_labels;  // int32 H*W
  DeviceBuffer synthesis_mask;      // uint8 H*W
  DeviceBuffer final_labels;        // int32 H*W
};

// d_mask: uint8 H*W on the device (nonzero = at floor).  The frame must not
// be entirely at floor (the caller checks).
int32_t route_at_floor(hipStream_t stream, const uint8_t* d_mask, int64_t height, int64_t width,
                       int64_t min_area, int64_t min_radius, int64_t margin, RoutingOutput& out);

}  // namespace dice
