// dice_numerics.hpp -- exact scalar numerics of the Digital ICE processing
// core, shared by the HIP kernels (device) and the host writer (host).
//
// Each function follows the reference CPU path of
// rohanpandula/digital-fauxice (file:line in src/portable_digital_ice/):
// every binary64 operation rounds once in source order and float32
// narrowing happens only at the stores the reference makes.  The translation
// units that include this header are compiled with -ffp-contract=off and
// `#pragma clang fp contract(off)`, so no multiply-add is fused.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

#define DICE_HD __host__ __device__ __forceinline__

namespace dice {

constexpr uint32_t kLcgMask = 0xFFFFFFu;               // rng.py:8
constexpr double kNikonNormalization = 0x1.fffffep-25;  // rng.py:14
// reconstruction.py:229, 247-248: float32-quantized, widened.
constexpr double kCoeff69 = (double)(float)(1.0 / 69.0);
constexpr double kCoeff21 = (double)(float)(1.0 / 21.0);
constexpr double kCoeff16 = (double)(float)(1.0 / 16.0);

DICE_HD float fmin_plain(float a, float b) { return a < b ? a : b; }

// x3a.py:193-221: aux = f32((ir - a*vis)/(1-a) - offset) in binary64.
DICE_HD float auxiliary(float visible, float infrared, float alpha32,
                        float offset32) {
  double alpha = (double)alpha32;
  double offset = (double)offset32;
  return (float)(((double)infrared - alpha * (double)visible) / (1.0 - alpha) -
                 offset);
}

// x3a.py:277-286: one final narrowing of the clamped affine score.
DICE_HD float score_value(float sample, float base, float addend, float scale,
                          float offset, float lower32) {
  double score64 = (((double)base + (double)addend) - (double)sample) *
                       (double)scale +
                   (double)offset;
  score64 = score64 < 1.0 ? score64 : 1.0;
  double lower = (double)lower32;
  score64 = score64 > lower ? score64 : lower;
  return (float)score64;
}

// output.py:61-94: half-up index, clamp, factorized inverse response.
DICE_HD uint16_t emit_value(float value, const uint32_t* high,
                            const uint32_t* low) {
  double widened = (double)value + 0.5;
  double clamped = widened < 0.0 ? 0.0 : widened;
  clamped = clamped > 65535.0 ? 65535.0 : clamped;
  uint32_t index = (uint32_t)clamped;
  uint64_t product = (uint64_t)high[index >> 8] * (uint64_t)low[index & 255u];
  return (uint16_t)((product >> 20) - 1u);
}

// Vertical scratch of one 9-sample column (reconstruction.py:187-237):
// S3/S5/S7/S9 stored float32 in the recovered add order, and the
// vertically filtered binomial half narrowed after the 1/16 scale.
struct Column {
  float s3, s5, s7, s9, b16;
};

DICE_HD Column column_scratch(float p0, float p1, float p2, float p3, float p4,
                              float p5, float p6, float p7, float p8) {
  Column c;
  double accumulator = (double)p3 + (double)p4;
  accumulator += (double)p5;
  c.s3 = (float)accumulator;
  accumulator += (double)p2;
  accumulator += (double)p6;
  c.s5 = (float)accumulator;
  accumulator += (double)p1;
  accumulator += (double)p7;
  c.s7 = (float)accumulator;
  accumulator += (double)p0;
  accumulator += (double)p8;
  c.s9 = (float)accumulator;
  double center = (double)p4;
  double binomial = (double)p3 + center;
  binomial += (double)p5;
  binomial += center;
  c.b16 = (float)(binomial * kCoeff16);
  return c;
}

// reconstruction.py:240-268: Q69 / Q21 / Q16 from nine column scratches.
DICE_HD void combine_columns(const Column col[9], double q[3]) {
  double total_21 = (double)col[6].s3 + (double)col[2].s3;
  total_21 += (double)col[3].s5;
  total_21 += (double)col[4].s5;
  total_21 += (double)col[5].s5;
  double total_69 = (double)col[8].s5 + (double)col[0].s5;
  total_69 += (double)col[1].s7;
  total_69 += (double)col[7].s7;
  total_69 += (double)col[2].s9;
  total_69 += (double)col[3].s9;
  total_69 += (double)col[4].s9;
  total_69 += (double)col[5].s9;
  total_69 += (double)col[6].s9;
  double total_16 = (double)col[4].b16 + (double)col[3].b16;
  total_16 += (double)col[5].b16;
  total_16 += (double)col[4].b16;
  q[0] = total_69 * kCoeff69;
  q[1] = total_21 * kCoeff21;
  q[2] = total_16;
}

// Q69/Q21/Q16 of a 9x9 patch read through `at(dy, dx)`, dy, dx in 0..8.
template <class At>
DICE_HD void unscaled_averages(const At& at, double q[3]) {
  Column col[9];
#pragma unroll
  for (int x = 0; x < 9; ++x)
    col[x] = column_scratch(at(0, x), at(1, x), at(2, x), at(3, x), at(4, x),
                            at(5, x), at(6, x), at(7, x), at(8, x));
  combine_columns(col, q);
}

// reconstruction.py:271-299: the RGB path's Q16 keeps three widened
// horizontal [1,2,1] rows and one vertical combine before the 1/16 scale.
template <class At>
DICE_HD double rgb_q16(const At& at) {
  double rows[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    int y = 3 + i;
    double center = (double)at(y, 4);
    double horizontal = (double)at(y, 3) + (double)at(y, 5);
    horizontal += center;
    horizontal += center;
    rows[i] = horizontal;
  }
  double total = rows[0] + rows[1];
  total += rows[1];
  total += rows[2];
  return total * kCoeff16;
}

// streaming.py:236-256: W record + A record of one 9x9 neighbourhood.
// The A lanes divide by the UNROUNDED binary64 W (reconstruction.py:365-371).
template <class ScoreAt, class WauxAt>
DICE_HD void feature_record(const ScoreAt& score_at, const WauxAt& waux_at,
                            float point_aux, float point_score, float fallback,
                            float weights[4], float feature[4]) {
  double unrounded[3], numerator[3];
  unscaled_averages(score_at, unrounded);
  weights[0] = (float)unrounded[0];
  weights[1] = (float)unrounded[1];
  weights[2] = (float)unrounded[2];
  weights[3] = point_score;
  unscaled_averages(waux_at, numerator);
#pragma unroll
  for (int l = 0; l < 3; ++l)
    feature[l] = unrounded[l] > 0.0 ? (float)(numerator[l] / unrounded[l])
                                    : fallback;
  feature[3] = weights[3] > 0.0f ? point_aux : fallback;
}

// Stable reconstruction constants packed for the combiner.
struct CombineParams {
  int coarse_enabled;
  float coarse_slopes[3];
  int band_enabled[3];
  float band_scales[3];
  float factors_a[9];
  float factors_b[9];
  float configured_strengths[3];
};

// reconstruction.py:672-887 (feature_band_ranges, _automatic_strengths,
// combine_recovered_candidate with asymmetric_dead_zone_residual), in three
// parts: combine_shared, combine_channel and combine_candidate over them.
//
// The parts shared by the three channels: the feature band ranges and the
// band strengths.
struct CombineShared {
  double range_min[3], range_max[3], strengths[3];
  double coarse_delta;
};

DICE_HD void combine_shared(const float weights[4], const float features[5][4], int record_count,
                            float coarse_reference, const CombineParams& P, CombineShared& out) {
#pragma unroll
  for (int t = 0; t < 3; ++t) {
    double d0 = (double)features[0][t + 1] - (double)features[0][t];
    double minimum = d0, maximum = d0;
    // (unrolled over the at most five records, so the records stay in
    // registers)
#pragma unroll
    for (int r = 1; r < 5; ++r) {
      if (r >= record_count) break;
      double d = (double)features[r][t + 1] - (double)features[r][t];
      if (d < minimum) minimum = d;
      if (d > maximum) maximum = d;
    }
    out.range_min[t] = (double)(float)minimum;
    out.range_max[t] = (double)(float)maximum;
  }
  if (P.configured_strengths[0] == 0.0f) {
    // min(1, max(0, x)) with the literal first: NaN collapses to 0.
    double doubled = (double)(float)(2.0 * (double)weights[1]);
    double lower_clamped = 0.0;
    if (doubled > 0.0) lower_clamped = doubled;
    double clamped = 1.0;
    if (lower_clamped < 1.0) clamped = lower_clamped;
    out.strengths[0] = (double)(float)clamped;
  } else {
    out.strengths[0] = (double)P.configured_strengths[0];
  }
  if (P.configured_strengths[1] == 0.0f) {
    float m2 = weights[2];
    if (m2 < 0.0f) m2 = 0.0f;
    out.strengths[1] = (double)m2;
  } else {
    out.strengths[1] = (double)P.configured_strengths[1];
  }
  if (P.configured_strengths[2] == 0.0f)
    out.strengths[2] = (double)weights[3] * (double)weights[3];
  else
    out.strengths[2] = (double)P.configured_strengths[2];
  out.coarse_delta = (double)coarse_reference - (double)features[0][0];
}

// Channel c of the combined candidate: the C69 candidate, the coarse slope
// term, then the dead-zone residual of each enabled band (coarse -> fine:
// C69 -> C21, C21 -> C16, C16 -> the original), in that order.
DICE_HD double combine_channel(int c, float c69, float c21, float c16, float source,
                               const CombineShared& S, const CombineParams& P) {
  double candidate = (double)c69;
  if (P.coarse_enabled) candidate = candidate + (double)P.coarse_slopes[c] * S.coarse_delta;
#pragma unroll
  for (int band = 0; band < 3; ++band) {
    if (!P.band_enabled[band]) continue;
    double scale = (double)P.band_scales[band];
    double edge_min = S.range_min[band], edge_max = S.range_max[band];
    bool negative_band = edge_min < 0.0 && edge_max < 0.0;
    const float coarse = band == 0 ? c69 : (band == 1 ? c21 : c16);
    const float fine = band == 0 ? c21 : (band == 1 ? c16 : source);
    double difference = scale * ((double)fine - (double)coarse);
    double upper, lower;
    if (negative_band) {
      upper = (double)P.factors_b[band * 3 + c] * edge_max;
      lower = (double)P.factors_a[band * 3 + c] * edge_min;
    } else {
      upper = (double)P.factors_a[band * 3 + c] * edge_max;
      lower = (double)P.factors_b[band * 3 + c] * edge_min;
    }
    double residual;
    if (difference > upper)
      residual = difference - upper;
    else if (difference < lower)
      residual = difference - lower;
    else
      residual = 0.0;
    candidate = candidate + S.strengths[band] * residual;
  }
  return candidate;
}

DICE_HD void combine_candidate(const float c69[3], const float c21[3],
                               const float c16[3], const float source[3],
                               const float weights[4],
                               const float features[5][4], int record_count,
                               float coarse_reference, const CombineParams& P,
                               double candidate[3]) {
  CombineShared shared;
  combine_shared(weights, features, record_count, coarse_reference, P, shared);
#pragma unroll
  for (int c = 0; c < 3; ++c)
    candidate[c] = combine_channel(c, c69[c], c21[c], c16[c], source[c], shared, P);
}

// dither.py:59-104 with the state threaded explicitly (rng.py:159-172).
struct WriterConsts {
  double low, high;
  bool low_lt_high;
  float scales[3];
};

DICE_HD double dither_delta(double value, const WriterConsts& K,
                            float scale32, uint32_t& state) {
  float candidate32 = (float)value;
  double candidate = (double)candidate32;
  if (!K.low_lt_high) return 0.0;
  if (!(K.low < candidate && candidate < K.high)) return 0.0;
  double width = K.high - K.low;
  double coefficient = 4.0 / (width * width);
  double envelope = ((K.high - candidate) * (candidate - K.low)) * coefficient;
  float random_span = (float)((double)scale32 * candidate);
  state = (125u * state + 1u) & kLcgMask;
  double centered = (double)(state + 1u) * kNikonNormalization - 0.5;
  double random_value = 0.0 + centered * (double)random_span;
  double delta = envelope * random_value;
  double changed = candidate + delta;
  if (K.low < changed && changed < K.high) return delta;
  return 0.0;
}

DICE_HD bool finite64(double v) { return __builtin_isfinite(v); }

// reconstruction.py:916-982: validity, original floor, redraw, f32 store.
// Returns the RNG advances counted as `state != before` per draw.
DICE_HD int write_pixel(const double candidate[3], const float original[3],
                        bool floor_enabled, const WriterConsts& K,
                        uint32_t& state, float values[3]) {
  bool valid = finite64(candidate[0]) && finite64(candidate[1]) &&
               finite64(candidate[2]) && candidate[0] > 0.0 &&
               candidate[1] > 0.0 && candidate[2] > 0.0;
  if (!valid) {
    values[0] = original[0];
    values[1] = original[1];
    values[2] = original[2];
    return 0;
  }
  int advances = 0;
  for (int c = 0; c < 3; ++c) {
    uint32_t before = state;
    double first_delta = dither_delta(candidate[c], K, K.scales[c], state);
    if (state != before) advances += 1;
    double source = (double)original[c];
    if (floor_enabled && candidate[c] + first_delta <= source) {
      values[c] = original[c];
    } else {
      double delta = first_delta;
      if (floor_enabled) {
        uint32_t before_redraw = state;
        delta = dither_delta(candidate[c], K, K.scales[c], state);
        if (state != before_redraw) advances += 1;
      }
      values[c] = (float)(candidate[c] + delta);
    }
  }
  return advances;
}

}  // namespace dice
