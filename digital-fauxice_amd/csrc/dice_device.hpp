// dice_device.hpp -- device-side job layout shared by the kernels and the
// host orchestration.  One JobDev describes one frame's buffers in HBM; the
// kernels index everything from it, so a batch is an array of JobDev.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dice_numerics.hpp"

namespace dice {

// One attempted reconstruction site handed to the serial writer.
// pixel: y*W + x in the low 31 bits; bit 31 = writer floor gate enabled.
struct alignas(8) Site {
  double cand[3];
  float orig[3];
  uint32_t pixel;
};
static_assert(sizeof(Site) == 40, "Site record must stay 40 bytes");

constexpr uint32_t kSitePixelMask = 0x7FFFFFFFu;
constexpr uint32_t kSiteFloorBit = 0x80000000u;

// Device error flags (atomicOr): the host maps them to the reference's
// fail-closed exceptions.
constexpr uint32_t kErrNonfiniteAux = 1u;    // continuous_score ValueError
constexpr uint32_t kErrSetterScaleOne = 4u;  // derive_x3a_setter_fields
constexpr uint32_t kErrStageNonfinite = 8u;  // StageCalibration finiteness

// Stable replay parameters (dice_params, binary32-quantized).
struct DevParams {
  int visible_channel;
  float score_addend, score_scale, score_offset, score_floor;
  int horizontal_minimum;
  float decision_threshold;
  int count_limit, radius;
  int cross_neighbor;
  int gate_primary;
  CombineParams combine;
};

// Per-stage calibration arrays in HBM, index s-1 for stage s = 1..H+6.
struct StageDev {
  float* alpha;
  float* offset;
  float* replacement;
  uint8_t* has_replacement;
  float* score_base;
  float* coarse;
  uint8_t* gate_secondary;
  int32_t* row_gate;
};

struct JobDev {
  int32_t height, width;
  int32_t tiles_x, tiles_y;     // analysis tiles (TY x TX)
  int32_t tile_rows, tile_cols; // TY, TX of the launch
  const uint16_t* rgbi;         // H*W*4
  uint16_t* out;                // H*W*3 (noop emit; patched after the writer)
  const float* lut;             // 65536
  const uint32_t* factor_high;  // 256
  const uint32_t* factor_low;   // 256
  StageDev stages;
  uint16_t* site_lists;         // tiles * TY*TX local indices, row-major
  uint32_t* tile_count;         // sites per tile
  uint32_t* seg_count;          // (tiles_y*TY) * tiles_x
  uint32_t* seg_offset;         // (tiles_y*TY) * tiles_x
  Site* sites;                  // compacted, row-major
  uint32_t* counters;           // [0] total sites
  uint32_t* error_flags;
  float* diag_score;            // optional H*W
  uint8_t* diag_floor;          // optional H*W
  // hidden startup (rows 0..4 planes and 6*W dense site slots)
  float* su_aux;                // 5*W
  float* su_score;              // 5*W
  float* su_work;               // 5*W*3
  Site* su_sites;               // 6*W
  uint8_t* su_attempted;        // 6*W
  // producer
  uint8_t* failpos;             // ceil(H/8) * (W/8)
  double* row_sums;             // H * 3 (visible, infrared, weight)
  uint32_t* accepted;           // H
  double* epoch_terms;          // epochs * cells * 8
  double* epoch_adds;           // epochs * 2 (den, num)
  double* cumulative;           // H * 3
  double* epoch_wide;           // epochs * 2 (den_w, num_w)
  float* epoch_store;           // epochs * 2 (den32, num32)
};

}  // namespace dice
