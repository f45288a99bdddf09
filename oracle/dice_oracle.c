/*
 * dice_oracle.c -- CPU oracle for the Digital ICE processing core.
 *
 * TEST INFRASTRUCTURE ONLY (see dice_oracle.h).  This is a plain-C
 * restatement of the reference CPU path of rohanpandula/digital-fauxice,
 * written from its Python sources; file:line citations below are relative
 * to that repository.  The compiled-CPU backend (src/portable_digital_ice/
 * fast_cpu/) is the structural template: whole-image planes, row-parallel
 * analysis in 128-row bands (OpenMP, disjoint writes, no reductions), and
 * the strictly serial writer.
 *
 * Numeric discipline (the reference's own rules): every binary64 operation
 * is rounded once, in source order; float32 narrowing happens only at the
 * cited stores.  Build with -ffp-contract=off and without fast-math.
 */
#define _GNU_SOURCE
#include "dice_oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static __thread char g_error[512];

static int32_t fail(int32_t code, const char* fmt, ...) {
  va_list args;
  va_start(args, fmt);
  vsnprintf(g_error, sizeof g_error, fmt, args);
  va_end(args);
  return code;
}

const char* dor_last_error(void) { return g_error; }

/* rng.py:9-14 */
#define LCG24_MASK 0xFFFFFFu
#define NIKON_NORMALIZATION 0x1.fffffep-25

/* reconstruction.py:229, 247-248: float32-quantized coefficients, widened. */
static const double COEFF_69 = (double)(float)(1.0 / 69.0);
static const double COEFF_21 = (double)(float)(1.0 / 21.0);
static const double COEFF_16 = (double)(float)(1.0 / 16.0);

/* producer_parameters.py:35-49 */
static const float PRODUCER_IR_THRESHOLD = 8847.2255859375f;
static const float PRODUCER_INITIAL_P1 = 0.15f;
static const float PRODUCER_INITIAL_P2 = 0.15f;
static const float PRODUCER_INITIAL_VISIBLE_MEAN = 46618.78515625f;
static const float PRODUCER_INITIAL_INFRARED_MEAN = 56239.65625f;
static const float PRODUCER_MULTIPLIER = 1.5f;

/* ------------------------------------------------------------------------
 * Tables
 * ---------------------------------------------------------------------- */

/* x3a.py:111-132: f32(65535 * log2(i + 1) / 16) in binary64. */
void dor_response_lut(float* lut) {
  for (int i = 0; i < 65536; ++i) {
    double value = 65535.0 * log2((double)i + 1.0) / 16.0;
    lut[i] = (float)value;
  }
}

/* output.py:39-58 */
void dor_inverse_factors(uint32_t* high, uint32_t* low) {
  double scale = 16.0 * log10(2.0) / 65535.0;
  for (int d = 0; d < 256; ++d) {
    double dd = (double)d;
    high[d] = (uint32_t)floor(16.0 * pow(10.0, 256.0 * dd * scale) + 0.5);
    low[d] = (uint32_t)floor(65536.0 * pow(10.0, dd * scale) + 0.5);
  }
}

/* dither.py:18-56 with maximum_index 65535 (profile.py:175-176). */
void dor_dither_bounds(const float* lut, float* low, float* high,
                       int32_t* low_index, int32_t* high_index) {
  const int maximum_index = 65535;
  float n = (float)maximum_index;
  float low_product = n * 0.01f;
  float high_product = n * 0.99f;
  int li = (int)truncf(low_product);
  int hi = (int)truncf(high_product);
  if (li > maximum_index) li = maximum_index;
  if (hi > maximum_index) hi = maximum_index;
  *low = lut[li];
  *high = lut[hi];
  if (low_index) *low_index = li;
  if (high_index) *high_index = hi;
}

/* ------------------------------------------------------------------------
 * Prepass (prepass.py:67-236)
 * ---------------------------------------------------------------------- */

static void put_f32(uint8_t* payload, int offset, float value) {
  memcpy(payload + offset, &value, 4);
}
static void put_f64(uint8_t* payload, int offset, double value) {
  memcpy(payload + offset, &value, 8);
}
static void put_u32(uint8_t* payload, int offset, uint32_t value) {
  memcpy(payload + offset, &value, 4);
}
static float get_f32(const uint8_t* payload, int offset) {
  float value;
  memcpy(&value, payload + offset, 4);
  return value;
}

/* prepass.py:67-83: four sequential row-major 4x4 sums. */
static void prepass_quadrant_means(const float values[8][8], float out[4]) {
  int index = 0;
  for (int fy = 0; fy < 8; fy += 4) {
    for (int fx = 0; fx < 8; fx += 4) {
      double total = 0.0;
      for (int y = fy; y < fy + 4; ++y)
        for (int x = fx; x < fx + 4; ++x) total += (double)values[y][x];
      out[index++] = (float)(total * 0.0625);
    }
  }
}

/* prepass.py:86-93 */
static float prepass_global_mean(const float means[4]) {
  double total = (double)means[0];
  total += (double)means[1];
  total += (double)means[2];
  total += (double)means[3];
  return (float)(total * 0.25);
}

int32_t dor_prepass(const uint16_t* pixels, int64_t height, int64_t width,
                    const float* lut, const dice_prepass_params* params,
                    dice_prepass_result* out) {
  if (height < 1 || width < 1)
    return fail(DICE_INVALID, "RGBI frame cannot have an empty dimension");
  int64_t padded_height = (height + 7) & ~(int64_t)7;
  int64_t padded_width = (width + 7) & ~(int64_t)7;
  float threshold = params->valid_ir_threshold;
  double ratio_minimum = (double)params->ratio_minimum;
  double ratio_maximum = (double)params->ratio_maximum;
  double alpha_scale = (double)params->alpha_scale;
  double calibration_r = 0.0, calibration_ir = 0.0, calibration_square = 0.0;
  float regression_denominator = 0.0f, regression_numerator = 0.0f;
  int64_t valid_pixels = 0, fully_valid_tiles = 0;

  for (int64_t fy = 0; fy < padded_height; fy += 8) {
    for (int64_t fx = 0; fx < padded_width; fx += 8) {
      float response_r[8][8], response_ir[8][8];
      memset(response_r, 0, sizeof response_r);
      memset(response_ir, 0, sizeof response_ir);
      float sum_raw_ir = 0.0f;
      int tile_valid = 1;
      for (int y = 0; y < 8; ++y) {
        for (int x = 0; x < 8; ++x) {
          int64_t py = fy + y, px = fx + x;
          uint16_t raw[4] = {0, 0, 0, 0};
          if (py < height && px < width)
            memcpy(raw, pixels + (py * width + px) * 4, sizeof raw);
          int raw_ir = raw[3];
          if ((float)raw_ir <= threshold) {
            tile_valid = 0;
            continue;
          }
          valid_pixels += 1;
          float converted_r = lut[raw[0]];
          float converted_ir = lut[raw_ir];
          response_r[y][x] = converted_r;
          response_ir[y][x] = converted_ir;
          double square = (double)((int64_t)raw_ir * raw_ir);
          calibration_r += square * (double)converted_r;
          calibration_ir += square * (double)converted_ir;
          calibration_square += square;
          sum_raw_ir = (float)((double)sum_raw_ir + (double)raw_ir);
        }
      }
      if (!tile_valid) continue;
      fully_valid_tiles += 1;
      float r_means[4], ir_means[4];
      prepass_quadrant_means(response_r, r_means);
      prepass_quadrant_means(response_ir, ir_means);
      float r_global = prepass_global_mean(r_means);
      float ir_global = prepass_global_mean(ir_means);
      double denominator = (double)regression_denominator;
      double numerator = (double)regression_numerator;
      double scalar = (double)sum_raw_ir;
      for (int q = 0; q < 4; ++q) {
        float r_dev = (float)((double)r_means[q] - (double)r_global);
        float ir_dev = (float)((double)ir_means[q] - (double)ir_global);
        if (r_dev == 0.0f) continue;
        double ratio = (double)ir_dev / (double)r_dev;
        if (ratio < ratio_minimum || ratio > ratio_maximum) continue;
        double weight = (double)r_dev * (double)r_dev * scalar * scalar;
        denominator += weight;
        numerator += ratio * weight;
      }
      regression_denominator = (float)denominator;
      regression_numerator = (float)numerator;
    }
  }

  double unrounded_alpha;
  if (regression_denominator == 0.0f)
    unrounded_alpha = (double)params->zero_denominator_alpha;
  else
    unrounded_alpha =
        (double)regression_numerator / (double)regression_denominator;
  float alpha = (float)unrounded_alpha;
  float scaled_alpha = (float)(unrounded_alpha * alpha_scale);

  memset(out->payload, 0, 64);
  put_u32(out->payload, 0x00, 1);
  put_f32(out->payload, 0x04, alpha);
  put_f32(out->payload, 0x08, scaled_alpha);
  put_f64(out->payload, 0x10, calibration_r);
  put_f64(out->payload, 0x18, calibration_ir);
  put_f64(out->payload, 0x20, calibration_square);
  if (calibration_square != 0.0) {
    put_f32(out->payload, 0x28, (float)(calibration_r / calibration_square));
    put_f32(out->payload, 0x2C, (float)(calibration_ir / calibration_square));
  }
  put_f32(out->payload, 0x30, regression_denominator);
  put_f32(out->payload, 0x34, regression_numerator);
  put_u32(out->payload, 0x38, params->completion_marker);
  out->valid_pixels = valid_pixels;
  out->fully_valid_tiles = fully_valid_tiles;
  return DICE_OK;
}

/* ------------------------------------------------------------------------
 * Producer schedule (producer_parameters.py:136-495)
 * ---------------------------------------------------------------------- */

/* producer_parameters.py:136-167: the asymmetric 4x4 sum tree. */
static void producer_quadrant_means(const double block[8][8], float out[4]) {
  static const int starts[4][2] = {{0, 0}, {0, 4}, {4, 0}, {4, 4}};
  for (int q = 0; q < 4; ++q) {
    const int r0 = starts[q][0], c0 = starts[q][1];
#define V(r, c) block[r0 + (r)][c0 + (c)]
    double first = V(0, 0);
    first += V(0, 1);
    first += V(0, 2);
    first += V(0, 3);
    first += V(1, 0);
    first += V(1, 1);
    double second_tail = V(1, 2) + V(1, 3);
    first += second_tail;
    double third = V(2, 0) + V(2, 1);
    double third_tail = V(2, 2) + V(2, 3);
    third += third_tail;
    double fourth = V(3, 0) + V(3, 1);
    double fourth_tail = V(3, 2) + V(3, 3);
    fourth += fourth_tail;
    third += fourth;
    double total = first + third;
#undef V
    out[q] = (float)(total * 0.0625);
  }
}

/* producer_parameters.py:170-180 */
static void producer_center_quadrants(const float means[4], float out[4]) {
  double total = (double)means[0];
  total += (double)means[1];
  total += (double)means[2];
  total += (double)means[3];
  float mean = (float)(total * 0.25);
  for (int q = 0; q < 4; ++q)
    out[q] = (float)((double)means[q] - (double)mean);
}

/* producer_parameters.py:183-263 */
static void scale_epoch_additions(const uint16_t* rgbi, int64_t width,
                                  int64_t row0, int64_t active_width,
                                  const float* lut, float* add_den,
                                  float* add_num) {
  const double low = (double)-0.125f;
  const double full_low = (double)0.0f;
  const double full_high = (double)0.3f;
  const double high = (double)0.425f;
  const double outer_weight = (double)0.70710677f;
  const double threshold = (double)PRODUCER_IR_THRESHOLD;
  double denominator = 0.0, numerator = 0.0;
  for (int64_t column = 0; column < active_width; column += 8) {
    int ok = 1;
    for (int r = 0; r < 8 && ok; ++r)
      for (int c = 0; c < 8; ++c)
        if ((double)rgbi[((row0 + r) * width + column + c) * 4 + 3] <=
            threshold) {
          ok = 0;
          break;
        }
    if (!ok) continue;
    double visible_block[8][8], infrared_block[8][8];
    for (int r = 0; r < 8; ++r)
      for (int c = 0; c < 8; ++c) {
        const uint16_t* px = rgbi + ((row0 + r) * width + column + c) * 4;
        visible_block[r][c] = (double)lut[px[0]];
        infrared_block[r][c] = (double)lut[px[3]];
      }
    float visible_means[4], infrared_means[4];
    float visible_deviations[4], infrared_deviations[4];
    producer_quadrant_means(visible_block, visible_means);
    producer_quadrant_means(infrared_block, infrared_means);
    producer_center_quadrants(visible_means, visible_deviations);
    producer_center_quadrants(infrared_means, infrared_deviations);
    float raw_sum = 0.0f;
    for (int r = 0; r < 8; ++r)
      for (int c = 0; c < 8; ++c)
        raw_sum = (float)((double)raw_sum +
                          (double)rgbi[((row0 + r) * width + column + c) * 4 +
                                       3]);
    double raw_sum_wide = (double)raw_sum;
    for (int q = 0; q < 4; ++q) {
      double visible_wide = (double)visible_deviations[q];
      double ratio, weight;
      if (visible_wide == 0.0) {
        ratio = 0.0;
        weight = 0.0;
      } else {
        ratio = (double)infrared_deviations[q] / visible_wide;
        if (ratio < low || ratio > high)
          weight = 0.0;
        else if (ratio < full_low || ratio > full_high)
          weight = visible_wide * outer_weight;
        else
          weight = visible_wide;
      }
      double term = weight * weight;
      term *= raw_sum_wide;
      term *= raw_sum_wide;
      denominator += term;
      numerator += ratio * term;
    }
  }
  *add_den = (float)denominator;
  *add_num = (float)numerator;
}

int32_t dor_producer_schedule(const uint16_t* rgbi, int64_t height,
                              int64_t width, const float* lut,
                              dice_schedule* out) {
  if (height <= 0) return fail(DICE_INVALID, "producer record schedule cannot be empty");
  int64_t active_width = (width / 8) * 8;
  if (!(0 < active_width && active_width <= width))
    return fail(DICE_INVALID, "producer active width is outside the RGBI row");
  out->rows = height;
  out->active_width = active_width;

  /* producer_parameters.py:266-377: per-row column-ordered sums with the
   * eight-row cell-validity latch. */
  const double threshold = (double)PRODUCER_IR_THRESHOLD;
  int64_t cell_count = active_width / 8;
  uint8_t* valid_cells = (uint8_t*)malloc((size_t)cell_count);
  double total_visible = 0.0, total_infrared = 0.0, total_weight = 0.0;
  float visible_mean = PRODUCER_INITIAL_VISIBLE_MEAN;
  float infrared_mean = PRODUCER_INITIAL_INFRARED_MEAN;
  for (int64_t row = 0; row < height; ++row) {
    if (row % 8 == 0) memset(valid_cells, 1, (size_t)cell_count);
    double row_visible = 0.0, row_infrared = 0.0, row_weight = 0.0;
    uint32_t accepted = 0;
    const uint16_t* line = rgbi + row * width * 4;
    for (int64_t column = 0; column < active_width; ++column) {
      int64_t cell = column / 8;
      if (!valid_cells[cell]) continue;
      int64_t raw_infrared = line[column * 4 + 3];
      if (!(threshold < (double)raw_infrared)) {
        valid_cells[cell] = 0;
        continue;
      }
      int64_t weight = raw_infrared * raw_infrared;
      int64_t raw_visible = line[column * 4 + 0];
      row_visible += (double)weight * (double)lut[raw_visible];
      row_infrared += (double)weight * (double)lut[raw_infrared];
      row_weight += (double)weight;
      accepted += 1;
    }
    total_visible += row_visible;
    total_infrared += row_infrared;
    total_weight += row_weight;
    if (total_weight != 0.0) {
      visible_mean = (float)(total_visible / total_weight);
      infrared_mean = (float)(total_infrared / total_weight);
    }
    out->weighted_visible_sum[row] = total_visible;
    out->weighted_infrared_sum[row] = total_infrared;
    out->weight_sum[row] = total_weight;
    out->mean_visible[row] = visible_mean;
    out->mean_infrared[row] = infrared_mean;
    out->accepted[row] = accepted;
  }
  free(valid_cells);

  /* producer_parameters.py:418-479 */
  float multiplier = PRODUCER_MULTIPLIER;
  float current_p1 = PRODUCER_INITIAL_P1;
  float current_p2 = PRODUCER_INITIAL_P2;
  float denominator = 0.0f, numerator = 0.0f;
  for (int64_t row = 0; row < height; ++row) {
    float add_den = 0.0f, add_num = 0.0f;
    if (row % 8 == 7) {
      int64_t start = row + 1 - 8;
      scale_epoch_additions(rgbi, width, start, active_width, lut, &add_den,
                            &add_num);
      double denominator_wide = (double)denominator + (double)add_den;
      double numerator_wide = (double)numerator + (double)add_num;
      denominator = (float)denominator_wide;
      numerator = (float)numerator_wide;
      if (denominator_wide != 0.0) {
        double ratio = numerator_wide / denominator_wide;
        current_p2 = (float)ratio;
        current_p1 = (float)((double)multiplier * ratio);
      }
    }
    out->scale_add_denominator[row] = add_den;
    out->scale_add_numerator[row] = add_num;
    out->scaled_input[row] = current_p1;
    out->input_scale[row] = current_p2;
    out->scale_denominator[row] = denominator;
    out->scale_numerator[row] = numerator;
  }
  for (int64_t row = 0; row < height; ++row) {
    uint8_t* payload = out->payloads + row * 64;
    memset(payload, 0, 64);
    put_f32(payload, 0x04, out->input_scale[row]);
    put_f32(payload, 0x08, out->scaled_input[row]);
    put_f64(payload, 0x10, out->weighted_visible_sum[row]);
    put_f64(payload, 0x18, out->weighted_infrared_sum[row]);
    put_f64(payload, 0x20, out->weight_sum[row]);
    put_f32(payload, 0x28, out->mean_visible[row]);
    put_f32(payload, 0x2C, out->mean_infrared[row]);
    put_f32(payload, 0x30, out->scale_denominator[row]);
    put_f32(payload, 0x34, out->scale_numerator[row]);
  }
  return DICE_OK;
}

/* producer_parameters.py:498-541 + stage_parameters.py:93-197 */
int32_t dor_stage_table_from_schedule(
    const dice_schedule* schedule, float auxiliary_factor_b,
    int64_t count, float* aux_alpha, float* aux_offset, float* aux_replacement,
    uint8_t* aux_has_replacement, float* score_base, float* coarse_reference,
    uint8_t* gate_secondary, int32_t* row_gate) {
  if (count < schedule->rows)
    return fail(DICE_INVALID, "producer hold cannot end before its source rows");
  if (!isfinite(auxiliary_factor_b))
    return fail(DICE_INVALID, "auxiliary factor B must be finite float32");
  for (int64_t stage = 1; stage <= count; ++stage) {
    int64_t row = stage - 1 < schedule->rows - 1 ? stage - 1 : schedule->rows - 1;
    const uint8_t* payload = schedule->payloads + row * 64;
    float p1 = get_f32(payload, 0x08);
    float p2 = get_f32(payload, 0x04);
    float p3 = get_f32(payload, 0x28);
    float p4 = get_f32(payload, 0x2C);
    if (!isfinite(p1)) return fail(DICE_INVALID, "scaled_input must be finite float32");
    if (!isfinite(p2)) return fail(DICE_INVALID, "input_scale must be finite float32");
    if (!isfinite(p3))
      return fail(DICE_INVALID, "calibration_mean_visible must be finite float32");
    if (!isfinite(p4))
      return fail(DICE_INVALID, "calibration_mean_infrared must be finite float32");
    if (p2 < 0.0f) {
      p1 = -p1;
      p2 = -p2;
    }
    double denominator_p1 = 1.0 - (double)p1;
    double denominator_p2 = 1.0 - (double)p2;
    if (denominator_p1 == 0.0 || denominator_p2 == 0.0)
      return fail(DICE_INVALID, "X3A setter scale cannot equal one");
    float base = (float)(((double)p4 - (double)p3 * (double)p2) / denominator_p2);
    int gate = !(p3 != 0.0f && p2 <= 0.075f);
    float alpha = p2 * auxiliary_factor_b;
    if (!isfinite(alpha))
      return fail(DICE_INVALID, "auxiliary_alpha must be finite float32");
    if (!isfinite(base))
      return fail(DICE_INVALID, "base_primary must be finite float32");
    int64_t i = stage - 1;
    aux_alpha[i] = alpha;
    aux_offset[i] = 1.0f;
    aux_replacement[i] = base;
    aux_has_replacement[i] = 1;
    score_base[i] = base;
    coarse_reference[i] = base;
    gate_secondary[i] = (uint8_t)gate;
    row_gate[i] = 0;
  }
  return DICE_OK;
}

/* ------------------------------------------------------------------------
 * Planes (x3a.py:193-286 via streaming.py:103-149)
 * ---------------------------------------------------------------------- */

static int compute_plane_row(const uint16_t* rgbi, int64_t width, int64_t y,
                             const float* lut, const dice_params* P,
                             float alpha32, int alpha_is_one,
                             float replacement, float offset32,
                             float score_base, float* working, float* aux,
                             float* score, float* waux, float* wrgb) {
  const uint16_t* line = rgbi + y * width * 4;
  float* work_row = working + y * width * 4;
  float* aux_row = aux + y * width;
  for (int64_t x = 0; x < width * 4; ++x) work_row[x] = lut[line[x]];
  if (alpha_is_one) {
    for (int64_t x = 0; x < width; ++x) aux_row[x] = replacement;
  } else {
    double alpha = (double)alpha32;
    double offset = (double)offset32;
    for (int64_t x = 0; x < width; ++x) {
      double visible = (double)work_row[x * 4 + P->visible_channel];
      double infrared = (double)work_row[x * 4 + 3];
      aux_row[x] = (float)((infrared - alpha * visible) / (1.0 - alpha) - offset);
    }
  }
  for (int64_t x = 0; x < width; ++x)
    if (!isfinite(aux_row[x])) return DICE_NONFINITE_AUX;
  double primary = (double)score_base;
  double addend = (double)P->score_addend;
  double scale = (double)P->score_scale;
  double offset = (double)P->score_offset;
  double lower = (double)P->score_floor;
  float* score_row = score + y * width;
  for (int64_t x = 0; x < width; ++x) {
    float sample = aux_row[x];
    if (P->horizontal_minimum && width > 2 && x >= 1 && x <= width - 2) {
      float a = aux_row[x - 1], b = aux_row[x], c = aux_row[x + 1];
      float m = a < b ? a : b;
      sample = m < c ? m : c;
    }
    double score64 = ((primary + addend) - (double)sample) * scale + offset;
    score64 = score64 < 1.0 ? score64 : 1.0;
    score64 = score64 > lower ? score64 : lower;
    score_row[x] = (float)score64;
  }
  for (int64_t x = 0; x < width; ++x) {
    float s = score_row[x];
    waux[y * width + x] = s * aux_row[x];
    for (int c = 0; c < 3; ++c)
      wrgb[(y * width + x) * 3 + c] = s * work_row[x * 4 + c];
  }
  return DICE_OK;
}

int32_t dor_planes(const uint16_t* rgbi, int64_t height, int64_t width,
                   const float* lut, const dice_params* P,
                   const float* row_alpha, const uint8_t* row_alpha_is_one,
                   const float* row_replacement, const float* row_offset,
                   const float* row_score_base, float* working, float* aux,
                   float* score, float* waux, float* wrgb) {
  int32_t status = DICE_OK;
#pragma omp parallel for schedule(static)
  for (int64_t y = 0; y < height; ++y) {
    int rc = compute_plane_row(rgbi, width, y, lut, P, row_alpha[y],
                               row_alpha_is_one[y], row_replacement[y],
                               row_offset[y], row_score_base[y], working, aux,
                               score, waux, wrgb);
    if (rc != DICE_OK) {
#pragma omp atomic write
      status = rc;
    }
  }
  if (status == DICE_NONFINITE_AUX)
    return fail(status, "auxiliary score input must be a finite HxW plane");
  return status;
}

/* ------------------------------------------------------------------------
 * 9x9 averages (reconstruction.py:187-299)
 * ---------------------------------------------------------------------- */

static void unscaled_averages(const float p[9][9], double q[3]) {
  float s3[9], s5[9], s7[9], s9[9], b16[9];
  for (int x = 0; x < 9; ++x) {
    double accumulator = (double)p[3][x] + (double)p[4][x];
    accumulator += (double)p[5][x];
    s3[x] = (float)accumulator;
    accumulator += (double)p[2][x];
    accumulator += (double)p[6][x];
    s5[x] = (float)accumulator;
    accumulator += (double)p[1][x];
    accumulator += (double)p[7][x];
    s7[x] = (float)accumulator;
    accumulator += (double)p[0][x];
    accumulator += (double)p[8][x];
    s9[x] = (float)accumulator;
    double center = (double)p[4][x];
    double binomial = (double)p[3][x] + center;
    binomial += (double)p[5][x];
    binomial += center;
    b16[x] = (float)(binomial * COEFF_16);
  }
  double total_21 = (double)s3[6] + (double)s3[2];
  total_21 += (double)s5[3];
  total_21 += (double)s5[4];
  total_21 += (double)s5[5];
  double total_69 = (double)s5[8] + (double)s5[0];
  total_69 += (double)s7[1];
  total_69 += (double)s7[7];
  for (int x = 2; x < 7; ++x) total_69 += (double)s9[x];
  double total_16 = (double)b16[4] + (double)b16[3];
  total_16 += (double)b16[5];
  total_16 += (double)b16[4];
  q[0] = total_69 * COEFF_69;
  q[1] = total_21 * COEFF_21;
  q[2] = total_16;
}

static void rgb_unscaled_averages(const float p[9][9][3], double q[3][3]) {
  for (int c = 0; c < 3; ++c) {
    float lane[9][9];
    for (int y = 0; y < 9; ++y)
      for (int x = 0; x < 9; ++x) lane[y][x] = p[y][x][c];
    double averages[3];
    unscaled_averages(lane, averages);
    q[0][c] = averages[0];
    q[1][c] = averages[1];
    double rows[3];
    for (int i = 0; i < 3; ++i) {
      int y = 3 + i;
      double center = (double)lane[y][4];
      double horizontal = (double)lane[y][3] + (double)lane[y][5];
      horizontal += center;
      horizontal += center;
      rows[i] = horizontal;
    }
    double total = rows[0] + rows[1];
    total += rows[1];
    total += rows[2];
    q[2][c] = total * COEFF_16;
  }
}

void dor_unscaled_averages(const float* patches, int64_t count, double* out) {
  for (int64_t i = 0; i < count; ++i)
    unscaled_averages((const float(*)[9])(patches + i * 81), out + i * 3);
}

void dor_rgb_unscaled_averages(const float* patches, int64_t count,
                               double* out) {
  for (int64_t i = 0; i < count; ++i)
    rgb_unscaled_averages((const float(*)[9][3])(patches + i * 243),
                          (double(*)[3])(out + i * 9));
}

/* ------------------------------------------------------------------------
 * Dither + writer (dither.py:59-104, rng.py:17-51, reconstruction.py:916-982)
 * ---------------------------------------------------------------------- */

typedef struct writer_consts {
  double low, high;
  int low_lt_high;
  float scales[3];
} writer_consts;

static double dither_delta(double value, const writer_consts* K, float scale32,
                           uint32_t* state) {
  float candidate32 = (float)value;
  double candidate = (double)candidate32;
  if (!K->low_lt_high) return 0.0;
  if (!(K->low < candidate && candidate < K->high)) return 0.0;
  double width = K->high - K->low;
  double coefficient = 4.0 / (width * width);
  double envelope = ((K->high - candidate) * (candidate - K->low)) * coefficient;
  float random_span = (float)((double)scale32 * candidate);
  *state = (125u * *state + 1u) & LCG24_MASK;
  double centered = (double)(*state + 1u) * NIKON_NORMALIZATION - 0.5;
  double random_value = 0.0 + centered * (double)random_span;
  double delta = envelope * random_value;
  double changed = candidate + delta;
  if (K->low < changed && changed < K->high) return delta;
  return 0.0;
}

void dor_dither_chain(const double* values, const float* scales, int64_t count,
                      uint32_t state_in, float low, float high, double* deltas,
                      uint32_t* state_out, uint64_t* advances_out) {
  writer_consts K;
  K.low = (double)low;
  K.high = (double)high;
  K.low_lt_high = K.low < K.high;
  uint32_t state = state_in;
  uint64_t advances = 0;
  for (int64_t i = 0; i < count; ++i) {
    uint32_t before = state;
    deltas[i] = dither_delta(values[i], &K, scales[i], &state);
    advances += state != before;
  }
  *state_out = state;
  *advances_out = advances;
}

/* fast_cpu/kernels.py:388-433 */
static int64_t write_pixel(const double candidate[3], const float original[3],
                           int floor_enabled, const writer_consts* K,
                           uint32_t* state, float values[3]) {
  int valid = isfinite(candidate[0]) && isfinite(candidate[1]) &&
              isfinite(candidate[2]) && candidate[0] > 0.0 &&
              candidate[1] > 0.0 && candidate[2] > 0.0;
  if (!valid) {
    values[0] = original[0];
    values[1] = original[1];
    values[2] = original[2];
    return 0;
  }
  int64_t advances = 0;
  for (int c = 0; c < 3; ++c) {
    uint32_t before = *state;
    double first_delta = dither_delta(candidate[c], K, K->scales[c], state);
    if (*state != before) advances += 1;
    double source = (double)original[c];
    if (floor_enabled && candidate[c] + first_delta <= source) {
      values[c] = original[c];
    } else {
      double delta = first_delta;
      if (floor_enabled) {
        uint32_t before_redraw = *state;
        delta = dither_delta(candidate[c], K, K->scales[c], state);
        if (*state != before_redraw) advances += 1;
      }
      values[c] = (float)(candidate[c] + delta);
    }
  }
  return advances;
}

/* The serial writer over n sites (host_writer.py:73-93 order): each site's
 * start state, its written values, the final state and the advances. */
void dor_write_pixels(const double* cand, const float* orig, const uint8_t* floor,
                      int64_t count, float low, float high, const float* scales,
                      uint32_t state_in, uint32_t* states, float* values,
                      uint32_t* state_out, uint64_t* advances_out) {
  writer_consts K;
  K.low = (double)low;
  K.high = (double)high;
  K.low_lt_high = K.low < K.high;
  for (int c = 0; c < 3; ++c) K.scales[c] = scales[c];
  uint32_t state = state_in & LCG24_MASK;
  uint64_t advances = 0;
  for (int64_t i = 0; i < count; ++i) {
    states[i] = state;
    advances += (uint64_t)write_pixel(cand + i * 3, orig + i * 3, floor[i], &K, &state,
                                      values + i * 3);
  }
  *state_out = state;
  *advances_out = advances;
}

/* One site's writer step from each of n start states. */
void dor_write_pixel_states(const double* cand, const float* orig, uint8_t floor, float low,
                            float high, const float* scales, const uint32_t* states_in,
                            int64_t count, uint32_t* states_out, uint8_t* advances) {
  writer_consts K;
  K.low = (double)low;
  K.high = (double)high;
  K.low_lt_high = K.low < K.high;
  for (int c = 0; c < 3; ++c) K.scales[c] = scales[c];
  for (int64_t i = 0; i < count; ++i) {
    uint32_t state = states_in[i] & LCG24_MASK;
    float values[3];
    advances[i] = (uint8_t)write_pixel(cand, orig, floor, &K, &state, values);
    states_out[i] = state;
  }
}

/* ------------------------------------------------------------------------
 * Emit (output.py:61-94)
 * ---------------------------------------------------------------------- */

static uint16_t emit_one(float v, const uint32_t* high, const uint32_t* low) {
  double widened = (double)v + 0.5;
  double clamped = widened < 0.0 ? 0.0 : widened;
  clamped = clamped > 65535.0 ? 65535.0 : clamped;
  uint32_t index = (uint32_t)clamped;
  uint64_t product = (uint64_t)high[index >> 8] * (uint64_t)low[index & 255u];
  return (uint16_t)((product >> 20) - 1u);
}

int32_t dor_emit(const float* values, int64_t count, uint16_t* out) {
  uint32_t high[256], low[256];
  dor_inverse_factors(high, low);
  for (int64_t i = 0; i < count; ++i)
    if (!isfinite(values[i])) return fail(DICE_NONFINITE_WORK, "work RGB must be finite");
  for (int64_t i = 0; i < count; ++i) out[i] = emit_one(values[i], high, low);
  return DICE_OK;
}

/* ------------------------------------------------------------------------
 * Features, candidates, combiner (streaming.py:236-308,
 * reconstruction.py:425-462, 672-904; fast_cpu/kernels.py:143-356)
 * ---------------------------------------------------------------------- */

static void feature_record(const float sp[9][9], const float wp[9][9],
                           float point_aux, float point_score, float fallback,
                           float weights[4], float feature[4]) {
  double unrounded[3], numerator[3];
  unscaled_averages(sp, unrounded);
  weights[0] = (float)unrounded[0];
  weights[1] = (float)unrounded[1];
  weights[2] = (float)unrounded[2];
  weights[3] = point_score;
  unscaled_averages(wp, numerator);
  for (int l = 0; l < 3; ++l)
    feature[l] = unrounded[l] > 0.0 ? (float)(numerator[l] / unrounded[l]) : fallback;
  feature[3] = weights[3] > 0.0f ? point_aux : fallback;
}

static void rgb_candidates(const float wrgb[9][9][3], const float w[3],
                           float out[3][3]) {
  double averages[3][3];
  rgb_unscaled_averages(wrgb, averages);
  for (int s = 0; s < 3; ++s) {
    if (w[s] > 0.0f) {
      double reciprocal = 1.0 / (double)w[s];
      for (int c = 0; c < 3; ++c) out[s][c] = (float)(averages[s][c] * reciprocal);
    } else {
      for (int c = 0; c < 3; ++c) out[s][c] = (float)averages[s][c];
    }
  }
}

static void combine(const float c69[3], const float c21[3], const float c16[3],
                    const float source[3], const float weights[4],
                    const float features[5][4], int record_count,
                    float coarse_reference, const dice_params* P,
                    double candidate[3]) {
  /* feature_band_ranges (reconstruction.py:672-705) */
  double range_min[3], range_max[3];
  for (int t = 0; t < 3; ++t) {
    double d0 = (double)features[0][t + 1] - (double)features[0][t];
    double minimum = d0, maximum = d0;
    for (int r = 1; r < record_count; ++r) {
      double d = (double)features[r][t + 1] - (double)features[r][t];
      if (d < minimum) minimum = d;
      if (d > maximum) maximum = d;
    }
    range_min[t] = (double)(float)minimum;
    range_max[t] = (double)(float)maximum;
  }
  /* _automatic_strengths (reconstruction.py:745-771) */
  double strengths[3];
  if (P->configured_strengths[0] == 0.0f) {
    double doubled = (double)(float)(2.0 * (double)weights[1]);
    double lower_clamped = 0.0;
    if (doubled > 0.0) lower_clamped = doubled;
    double clamped = 1.0;
    if (lower_clamped < 1.0) clamped = lower_clamped;
    strengths[0] = (double)(float)clamped;
  } else {
    strengths[0] = (double)P->configured_strengths[0];
  }
  if (P->configured_strengths[1] == 0.0f) {
    float m2 = weights[2];
    if (m2 < 0.0f) m2 = 0.0f;
    strengths[1] = (double)m2;
  } else {
    strengths[1] = (double)P->configured_strengths[1];
  }
  if (P->configured_strengths[2] == 0.0f)
    strengths[2] = (double)weights[3] * (double)weights[3];
  else
    strengths[2] = (double)P->configured_strengths[2];

  /* combine_recovered_candidate (reconstruction.py:807-887) */
  for (int c = 0; c < 3; ++c) candidate[c] = (double)c69[c];
  if (P->coarse_enabled) {
    double coarse_delta = (double)coarse_reference - (double)features[0][0];
    for (int c = 0; c < 3; ++c)
      candidate[c] = candidate[c] + (double)P->coarse_slopes[c] * coarse_delta;
  }
  for (int band = 0; band < 3; ++band) {
    if (!P->band_enabled[band]) continue;
    double scale = (double)P->band_scales[band];
    double edge_min = range_min[band], edge_max = range_max[band];
    int negative_band = edge_min < 0.0 && edge_max < 0.0;
    for (int c = 0; c < 3; ++c) {
      double coarse_value, fine_value;
      if (band == 0) {
        coarse_value = (double)c69[c];
        fine_value = (double)c21[c];
      } else if (band == 1) {
        coarse_value = (double)c21[c];
        fine_value = (double)c16[c];
      } else {
        coarse_value = (double)c16[c];
        fine_value = (double)source[c];
      }
      double difference = scale * (fine_value - coarse_value);
      double upper, lower;
      if (negative_band) {
        upper = (double)P->factors_b[band * 3 + c] * edge_max;
        lower = (double)P->factors_a[band * 3 + c] * edge_min;
      } else {
        upper = (double)P->factors_a[band * 3 + c] * edge_max;
        lower = (double)P->factors_b[band * 3 + c] * edge_min;
      }
      double residual;
      if (difference > upper)
        residual = difference - upper;
      else if (difference < lower)
        residual = difference - lower;
      else
        residual = 0.0;
      candidate[c] = candidate[c] + strengths[band] * residual;
    }
  }
}

/* ------------------------------------------------------------------------
 * History loaders (streaming.py:201-233; fast_cpu/kernels.py:699-773)
 * ---------------------------------------------------------------------- */

typedef struct planes {
  int64_t height, width;
  const float* working; /* H*W*4 */
  const float* aux;
  const float* score;
  const float* waux;
  const float* wrgb; /* H*W*3 */
  float floor;
} planes;

static float load_score(const planes* P, int64_t y, int64_t x) {
  if (x < 0 || x >= P->width || y < -1) return 0.0f;
  if (y == -1) return P->floor;
  if (y >= P->height) y = P->height - 1;
  return P->score[y * P->width + x];
}

static float load_waux(const planes* P, int64_t y, int64_t x) {
  if (x < 0 || x >= P->width || y < -1) return 0.0f;
  if (y == -1) return P->aux[x] * P->floor;
  if (y >= P->height) y = P->height - 1;
  return P->waux[y * P->width + x];
}

static float load_wrgb(const planes* P, int64_t y, int64_t x, int c) {
  if (x < 0 || x >= P->width || y < -1) return 0.0f;
  if (y == -1) return P->working[x * 4 + c] * P->floor;
  if (y >= P->height) y = P->height - 1;
  return P->wrgb[(y * P->width + x) * 3 + c];
}

static void gather_patches(const planes* P, int64_t cy, int64_t cx,
                           float sp[9][9], float wp[9][9]) {
  for (int dy = -4; dy <= 4; ++dy)
    for (int dx = -4; dx <= 4; ++dx) {
      sp[dy + 4][dx + 4] = load_score(P, cy + dy, cx + dx);
      wp[dy + 4][dx + 4] = load_waux(P, cy + dy, cx + dx);
    }
}

/* ------------------------------------------------------------------------
 * Decision (streaming.py:157-198; fast_cpu/kernels.py:776-855)
 * ---------------------------------------------------------------------- */

static int decision_fallback(const float* aux, int64_t height, int64_t width,
                             int64_t y, int64_t x, float threshold, int radius,
                             int count_limit) {
  int64_t lx = x - radius;
  if (lx < 0) lx = 0;
  else if (lx >= width) lx = width - 1;
  int64_t rx = x + radius;
  if (rx < 0) rx = 0;
  else if (rx >= width) rx = width - 1;
  int vertical_left = 0, vertical_right = 0;
  for (int dy = -4; dy <= 4; ++dy) {
    int64_t yy = y + dy;
    if (yy < 0) yy = 0;
    else if (yy >= height) yy = height - 1;
    if (aux[yy * width + lx] < threshold) vertical_left += 1;
    if (aux[yy * width + rx] < threshold) vertical_right += 1;
  }
  int64_t ay = y - radius;
  if (ay < 0) ay = 0;
  int64_t by = y + radius;
  if (by >= height) by = height - 1;
  int horizontal_above = 0, horizontal_below = 0;
  for (int dx = -4; dx <= 4; ++dx) {
    int64_t xx = x + dx;
    if (xx < 0) xx = 0;
    else if (xx >= width) xx = width - 1;
    if (aux[ay * width + xx] < threshold) horizontal_above += 1;
    if (aux[by * width + xx] < threshold) horizontal_below += 1;
  }
  return vertical_left > count_limit || horizontal_above > count_limit ||
         vertical_right > count_limit || horizontal_below > count_limit;
}

void dor_decision(const float* aux, int64_t height, int64_t width,
                  const dice_params* P, uint8_t* fallback) {
  for (int64_t y = 0; y < height; ++y)
    for (int64_t x = 0; x < width; ++x)
      fallback[y * width + x] = (uint8_t)decision_fallback(
          aux, height, width, y, x, P->decision_threshold, P->decision_radius,
          P->decision_count_limit);
}

/* ------------------------------------------------------------------------
 * Row analysis (fast_cpu/kernels.py:858-1067)
 * ---------------------------------------------------------------------- */

static int64_t analyze_row(const planes* Q, int64_t y, const dice_params* P,
                           int floor_enabled, int32_t row_gate,
                           float fallback, uint8_t* attempted_row,
                           double* candidates_row) {
  const int64_t height = Q->height, width = Q->width;
  const int record_count = P->cross_neighbor ? 5 : 1;
  int64_t attempted = 0;
  float sp[9][9], wp[9][9], wrgb[9][9][3];
  for (int64_t x = 0; x < width; ++x) {
    attempted_row[x] = 0;
    if (row_gate != 0) continue;
    if (decision_fallback(Q->aux, height, width, y, x, P->decision_threshold,
                          P->decision_radius, P->decision_count_limit))
      continue;
    if (floor_enabled && Q->score[y * width + x] >= 1.0f) continue;

    float weights[4], features[5][4];
    memset(features, 0, sizeof features);
    gather_patches(Q, y, x, sp, wp);
    feature_record(sp, wp, Q->aux[y * width + x], Q->score[y * width + x],
                   fallback, weights, features[0]);
    /* driver_forces_fallback (reconstruction.py:890-904) */
    if (row_gate != 0 || (floor_enabled && weights[3] >= 1.0f)) continue;
    attempted += 1;
    if (record_count == 5) {
      float w_unused[4];
      if (x > 0) {
        gather_patches(Q, y, x - 1, sp, wp);
        feature_record(sp, wp, Q->aux[y * width + x - 1],
                       Q->score[y * width + x - 1], fallback, w_unused,
                       features[1]);
      }
      if (x < width - 1) {
        gather_patches(Q, y, x + 1, sp, wp);
        feature_record(sp, wp, Q->aux[y * width + x + 1],
                       Q->score[y * width + x + 1], fallback, w_unused,
                       features[2]);
      }
      int64_t up = y - 1;
      gather_patches(Q, up, x, sp, wp);
      float up_aux = up < 0 ? Q->aux[x] : Q->aux[up * width + x];
      float up_score = up < 0 ? Q->floor : Q->score[up * width + x];
      feature_record(sp, wp, up_aux, up_score, fallback, w_unused, features[3]);
      int64_t down = y + 1;
      gather_patches(Q, down, x, sp, wp);
      int64_t clipped = down < height ? down : height - 1;
      feature_record(sp, wp, Q->aux[clipped * width + x],
                     Q->score[clipped * width + x], fallback, w_unused,
                     features[4]);
    }
    for (int dy = -4; dy <= 4; ++dy)
      for (int dx = -4; dx <= 4; ++dx)
        for (int c = 0; c < 3; ++c)
          wrgb[dy + 4][dx + 4][c] = load_wrgb(Q, y + dy, x + dx, c);
    float w3[3] = {weights[0], weights[1], weights[2]};
    float candidates[3][3];
    rgb_candidates(wrgb, w3, candidates);
    const float* px = Q->working + (y * width + x) * 4;
    float original[3] = {px[0], px[1], px[2]};
    combine(candidates[0], candidates[1], candidates[2], original, weights,
            (const float(*)[4])features, record_count, fallback, P,
            candidates_row + x * 3);
    attempted_row[x] = 1;
  }
  return attempted;
}

/* ------------------------------------------------------------------------
 * Hidden startup (startup.py:74-353; fast_cpu/kernels.py:448-680)
 * ---------------------------------------------------------------------- */

#define STARTUP_MIN_Y (-10)
#define STARTUP_GUARD 4
#define STARTUP_HIDDEN_FLOOR (-6)

typedef struct startup_hist {
  int64_t rows, padded_width, width;
  float* score;
  float* waux;
  float* wrgb;
  float* raw;
} startup_hist;

static float* hist_at(float* base, const startup_hist* S, int64_t y,
                      int64_t padded_x) {
  return base + (y - STARTUP_MIN_Y) * S->padded_width + padded_x;
}

static float* hist_rgb(const startup_hist* S, int64_t y, int64_t padded_x,
                        int c) {
  return S->wrgb + ((y - STARTUP_MIN_Y) * S->padded_width + padded_x) * 3 + c;
}

static void startup_record(const startup_hist* S, int64_t logical_y, int64_t x,
                           float fallback, float weights[4], float feature[4]) {
  memset(weights, 0, 4 * sizeof(float));
  memset(feature, 0, 4 * sizeof(float));
  if (logical_y <= STARTUP_HIDDEN_FLOOR || x < 0 || x >= S->width) return;
  int64_t padded_x = x + STARTUP_GUARD;
  float sp[9][9], wp[9][9];
  for (int dy = -4; dy <= 4; ++dy)
    for (int dx = -4; dx <= 4; ++dx) {
      sp[dy + 4][dx + 4] = *hist_at(S->score, S, logical_y + dy, padded_x + dx);
      wp[dy + 4][dx + 4] = *hist_at(S->waux, S, logical_y + dy, padded_x + dx);
    }
  double unrounded[3], numerator[3];
  unscaled_averages(sp, unrounded);
  weights[0] = (float)unrounded[0];
  weights[1] = (float)unrounded[1];
  weights[2] = (float)unrounded[2];
  weights[3] = *hist_at(S->score, S, logical_y, padded_x);
  unscaled_averages(wp, numerator);
  for (int l = 0; l < 3; ++l)
    feature[l] = unrounded[l] > 0.0 ? (float)(numerator[l] / unrounded[l]) : fallback;
  feature[3] = weights[3] > 0.0f ? *hist_at(S->raw, S, logical_y, padded_x) : fallback;
}

static int stage_floor_enabled(const dice_params* P, const dice_stage_table* st,
                               int64_t stage) {
  return P->gate_primary || st->gate_secondary[stage - 1];
}

static int64_t max_i64(int64_t a, int64_t b) { return a > b ? a : b; }

static void startup_replay(const planes* Q, const dice_params* P,
                           const dice_stage_table* st, const writer_consts* K,
                           uint32_t* state, dice_replay_result* result) {
  const int cross = P->cross_neighbor != 0;
  const int64_t width = Q->width;
  const int64_t first_rows = cross ? 5 : 4;
  const int64_t max_y = first_rows - 1;
  startup_hist S;
  S.rows = max_y - STARTUP_MIN_Y + 1;
  S.padded_width = width + 2 * STARTUP_GUARD;
  S.width = width;
  size_t cells = (size_t)(S.rows * S.padded_width);
  S.score = (float*)calloc(cells, sizeof(float));
  S.waux = (float*)calloc(cells, sizeof(float));
  S.wrgb = (float*)calloc(cells * 3, sizeof(float));
  S.raw = (float*)calloc(cells, sizeof(float));
  const float floor = Q->floor;
  for (int64_t y = STARTUP_MIN_Y; y <= max_y; ++y) {
    for (int64_t x = 0; x < width; ++x) {
      int64_t px = x + STARTUP_GUARD;
      if (y == -1) {
        *hist_at(S.score, &S, y, px) = floor;
        *hist_at(S.waux, &S, y, px) = Q->aux[x] * floor;
        for (int c = 0; c < 3; ++c)
          *hist_rgb(&S, y, px, c) = Q->working[x * 4 + c] * floor;
      } else if (y >= 0) {
        float s = Q->score[y * width + x];
        *hist_at(S.score, &S, y, px) = s;
        *hist_at(S.waux, &S, y, px) = s * Q->aux[y * width + x];
        for (int c = 0; c < 3; ++c)
          *hist_rgb(&S, y, px, c) = s * Q->working[(y * width + x) * 4 + c];
      }
    }
    const float* source = y < 0 ? Q->aux : Q->aux + y * width;
    float* raw_row = hist_at(S.raw, &S, y, 0);
    for (int g = 0; g < STARTUP_GUARD; ++g) raw_row[g] = source[0];
    for (int64_t x = 0; x < width; ++x) raw_row[STARTUP_GUARD + x] = source[x];
    if (y >= 0)
      for (int g = 0; g < STARTUP_GUARD; ++g)
        raw_row[STARTUP_GUARD + width + g] = source[width - 1];
  }

  const int radius = P->decision_radius;
  const int record_count = cross ? 5 : 1;
  const float* original_row = Q->working; /* row 0, stride 4 */
  for (int stage = 1; stage <= 6; ++stage) {
    int64_t center_y = stage - 7;
    int64_t writer_width = (cross && stage <= 2) ? width - 4 : width;
    int floor_enabled = stage_floor_enabled(P, st, stage);
    int32_t row_gate = st->row_gate[stage - 1];
    float combiner_reference = st->coarse_reference[stage - 1];
    float fallback_center = st->coarse_reference[max_i64(1, center_y + 6) - 1];
    float fallback_up = st->coarse_reference[max_i64(1, center_y + 5) - 1];
    float fallback_down = st->coarse_reference[max_i64(1, center_y + 7) - 1];
    int64_t attempted = 0, advances = 0;
    for (int64_t x = 0; x < writer_width; ++x) {
      int64_t padded_x = x + STARTUP_GUARD;
      int fallback = 0;
      for (int group = 0; group < 4 && !fallback; ++group) {
        int count = 0;
        for (int sample = 0; sample < 9; ++sample) {
          int axis = sample - 4, dy, dx;
          if (group == 0) {
            dy = axis;
            dx = -radius;
          } else if (group == 1) {
            dy = -radius;
            dx = axis;
          } else if (group == 2) {
            dy = axis;
            dx = radius;
          } else {
            dy = radius;
            dx = axis;
          }
          if (*hist_at(S.raw, &S, center_y + dy, padded_x + dx) < P->decision_threshold)
            count += 1;
        }
        if (count > P->decision_count_limit) fallback = 1;
      }
      float weights[4], features[5][4];
      memset(features, 0, sizeof features);
      startup_record(&S, center_y, x, fallback_center, weights, features[0]);
      if (fallback || row_gate != 0 || (floor_enabled && weights[3] >= 1.0f))
        continue;
      attempted += 1;
      if (record_count == 5) {
        float w_unused[4];
        startup_record(&S, center_y, x - 1, fallback_center, w_unused, features[1]);
        startup_record(&S, center_y, x + 1, fallback_center, w_unused, features[2]);
        startup_record(&S, center_y - 1, x, fallback_up, w_unused, features[3]);
        startup_record(&S, center_y + 1, x, fallback_down, w_unused, features[4]);
      }
      float wrgb[9][9][3];
      for (int dy = -4; dy <= 4; ++dy)
        for (int dx = -4; dx <= 4; ++dx)
          for (int c = 0; c < 3; ++c)
            wrgb[dy + 4][dx + 4][c] = *hist_rgb(&S, center_y + dy, padded_x + dx, c);
      float w3[3] = {weights[0], weights[1], weights[2]};
      float candidates[3][3];
      rgb_candidates(wrgb, w3, candidates);
      float original[3] = {original_row[x * 4 + 0], original_row[x * 4 + 1],
                           original_row[x * 4 + 2]};
      double candidate[3];
      combine(candidates[0], candidates[1], candidates[2], original, weights,
              (const float(*)[4])features, record_count, combiner_reference, P,
              candidate);
      float values[3];
      advances += write_pixel(candidate, original, floor_enabled, K, state, values);
    }
    result->startup_attempted[stage - 1] = attempted;
    result->startup_advances[stage - 1] = advances;
  }
  result->startup_final_rng_state = *state;
  free(S.score);
  free(S.waux);
  free(S.wrgb);
  free(S.raw);
}

/* ------------------------------------------------------------------------
 * Replay (streaming.py:365-552; fast_cpu/engine.py:248-534)
 * ---------------------------------------------------------------------- */

#define BAND_ROWS 128

static int all_finite(const float* v, int n) {
  for (int i = 0; i < n; ++i)
    if (!isfinite(v[i])) return 0;
  return 1;
}

int32_t dor_replay(const uint16_t* rgbi, int64_t height, int64_t width,
                   const float* lut, const dice_params* P,
                   const dice_stage_table* st, uint32_t rng_state,
                   uint16_t* out, dice_replay_result* result,
                   dice_diagnostics* diagnostics, int32_t threads) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#else
  (void)threads;
#endif
  memset(result, 0, sizeof *result);
  if (height < 4) return fail(DICE_INVALID, "streaming startup requires at least four main rows");
  if (width < 1) return fail(DICE_INVALID, "RGBI frame cannot have an empty dimension");
  if (P->visible_channel < 0 || P->visible_channel > 2)
    return fail(DICE_INVALID, "selected_visible_channel must be 0, 1, or 2");
  if (st->count < height + 6)
    return fail(DICE_INVALID, "stage table must cover stages 1..H+6");
  if (rng_state > LCG24_MASK) return fail(DICE_INVALID, "LCG24 state must fit in 24 bits");
  if (P->cross_neighbor && height < 5)
    return fail(DICE_INVALID, "cross-neighbor startup requires the first five real rows");
  if (P->cross_neighbor && width < 4)
    return fail(DICE_INVALID, "cross-neighbor startup requires at least four columns");
  if (!(isfinite(P->score_floor) && P->score_floor > 0.0f))
    return fail(DICE_INVALID, "startup score floor must be finite and positive");
  if (!(isfinite(P->dither_low) && isfinite(P->dither_high) &&
        all_finite(P->dither_scales, 3)))
    return fail(DICE_INVALID, "dither inputs must be finite");
  if (!(all_finite(P->coarse_slopes, 3) && all_finite(P->band_scales, 3) &&
        all_finite(P->factors_a, 9) && all_finite(P->factors_b, 9) &&
        all_finite(P->configured_strengths, 3)))
    return fail(DICE_INVALID, "reconstruction parameters must be finite");
  for (int64_t i = 0; i < height + 6; ++i) {
    if (!(isfinite(st->aux_alpha[i]) && isfinite(st->aux_offset[i]) &&
          isfinite(st->score_base[i]) && isfinite(st->coarse_reference[i])))
      return fail(DICE_INVALID, "stage calibration contains a non-finite value");
    if (i < height && st->aux_alpha[i] == 1.0f && !st->aux_has_replacement[i])
      return fail(DICE_INVALID,
                  "alpha==1 requires the runtime-configured auxiliary replacement");
  }

  const int64_t pixels = height * width;
  float* working = (float*)malloc((size_t)pixels * 4 * sizeof(float));
  float* aux = (float*)malloc((size_t)pixels * sizeof(float));
  float* score = (float*)malloc((size_t)pixels * sizeof(float));
  float* waux = (float*)malloc((size_t)pixels * sizeof(float));
  float* wrgb = (float*)malloc((size_t)pixels * 3 * sizeof(float));
  float* row_alpha = (float*)malloc((size_t)height * sizeof(float));
  uint8_t* row_is_one = (uint8_t*)malloc((size_t)height);
  float* row_replacement = (float*)malloc((size_t)height * sizeof(float));
  float* row_offset = (float*)malloc((size_t)height * sizeof(float));
  float* row_score_base = (float*)malloc((size_t)height * sizeof(float));
  int32_t status = DICE_OK;
  if (!working || !aux || !score || !waux || !wrgb) {
    status = fail(DICE_INVALID, "oracle allocation failed");
    goto done;
  }
  /* cuda_backend/rowparams.py:60-139 */
  for (int64_t y = 0; y < height; ++y) {
    row_alpha[y] = st->aux_alpha[y];
    row_is_one[y] = st->aux_alpha[y] == 1.0f;
    row_replacement[y] = st->aux_replacement[y];
    row_offset[y] = st->aux_offset[y];
    row_score_base[y] = st->score_base[y + 1];
  }
  status = dor_planes(rgbi, height, width, lut, P, row_alpha, row_is_one,
                      row_replacement, row_offset, row_score_base, working, aux,
                      score, waux, wrgb);
  if (status != DICE_OK) goto done;

  planes Q = {height, width, working, aux, score, waux, wrgb, P->score_floor};
  writer_consts K;
  K.low = (double)P->dither_low;
  K.high = (double)P->dither_high;
  K.low_lt_high = K.low < K.high;
  memcpy(K.scales, P->dither_scales, sizeof K.scales);
  uint32_t state = rng_state;
  startup_replay(&Q, P, st, &K, &state, result);

  uint32_t high[256], low[256];
  dor_inverse_factors(high, low);
  int64_t band_capacity = height < BAND_ROWS ? height : BAND_ROWS;
  uint8_t* band_attempted = (uint8_t*)malloc((size_t)(band_capacity * width));
  double* band_candidates =
      (double*)malloc((size_t)(band_capacity * width * 3) * sizeof(double));
  int64_t* band_counts = (int64_t*)malloc((size_t)band_capacity * sizeof(int64_t));
  float* row_values = (float*)malloc((size_t)width * 3 * sizeof(float));
  uint64_t attempted = 0, written = 0, advances = 0, changed = 0;
  for (int64_t y0 = 0; y0 < height && status == DICE_OK; y0 += band_capacity) {
    int64_t band_rows = height - y0 < band_capacity ? height - y0 : band_capacity;
#pragma omp parallel for schedule(dynamic, 1)
    for (int64_t i = 0; i < band_rows; ++i) {
      int64_t y = y0 + i;
      int64_t stage_index = y + 6; /* writer stage y+7 */
      band_counts[i] = analyze_row(
          &Q, y, P, stage_floor_enabled(P, st, y + 7), st->row_gate[stage_index],
          st->coarse_reference[stage_index], band_attempted + i * width,
          band_candidates + i * width * 3);
    }
    for (int64_t i = 0; i < band_rows && status == DICE_OK; ++i) {
      int64_t y = y0 + i;
      int floor_enabled = stage_floor_enabled(P, st, y + 7);
      for (int64_t x = 0; x < width; ++x) {
        const float* px = working + (y * width + x) * 4;
        row_values[x * 3 + 0] = px[0];
        row_values[x * 3 + 1] = px[1];
        row_values[x * 3 + 2] = px[2];
      }
      for (int64_t x = 0; x < width; ++x) {
        if (!band_attempted[i * width + x]) continue;
        const float* px = working + (y * width + x) * 4;
        float original[3] = {px[0], px[1], px[2]};
        float values[3];
        advances += (uint64_t)write_pixel(band_candidates + (i * width + x) * 3,
                                          original, floor_enabled, &K, &state,
                                          values);
        memcpy(row_values + x * 3, values, sizeof values);
        if (values[0] != original[0] || values[1] != original[1] ||
            values[2] != original[2])
          written += 1;
      }
      attempted += (uint64_t)band_counts[i];
      for (int64_t x = 0; x < width * 3; ++x)
        if (!isfinite(row_values[x])) {
          status = fail(DICE_NONFINITE_WORK, "work RGB must be finite");
          break;
        }
      if (status != DICE_OK) break;
      for (int64_t x = 0; x < width; ++x) {
        int row_changed = 0;
        const float* px = working + (y * width + x) * 4;
        for (int c = 0; c < 3; ++c) {
          uint16_t rendered = emit_one(row_values[x * 3 + c], high, low);
          uint16_t noop = emit_one(px[c], high, low);
          out[(y * width + x) * 3 + c] = rendered;
          row_changed |= rendered != noop;
        }
        changed += (uint64_t)row_changed;
        if (diagnostics && diagnostics->changed_mask)
          diagnostics->changed_mask[y * width + x] = (uint8_t)row_changed;
      }
    }
  }
  free(band_attempted);
  free(band_candidates);
  free(band_counts);
  free(row_values);
  if (status == DICE_OK) {
    result->attempted_pixels = attempted;
    result->written_pixels = written;
    result->public_rng_advances = advances;
    result->changed_pixels = changed;
    result->final_rng_state = state;
    if (diagnostics) {
      for (int64_t i = 0; i < pixels; ++i) {
        if (diagnostics->score_plane) diagnostics->score_plane[i] = score[i];
        if (diagnostics->at_floor_mask)
          diagnostics->at_floor_mask[i] = (uint8_t)(score[i] == P->score_floor);
      }
    }
  }
done:
  free(working);
  free(aux);
  free(score);
  free(waux);
  free(wrgb);
  free(row_alpha);
  free(row_is_one);
  free(row_replacement);
  free(row_offset);
  free(row_score_base);
  return status;
}

/* ------------------------------------------------------------------------
 * Whole job (engine.py:110-267)
 * ---------------------------------------------------------------------- */

int32_t dor_process_job(const uint16_t* main_rgbi, int64_t height,
                        int64_t width, const uint16_t* prepass_rgbi,
                        int64_t prepass_height, int64_t prepass_width,
                        const dice_params* params,
                        const dice_prepass_params* prepass_params,
                        uint16_t* out_rgb16, dice_replay_result* result,
                        dice_diagnostics* diagnostics, int32_t threads) {
  if (height < 5 || width < 8)
    return fail(DICE_INVALID, "main RGBI frame must be at least 5 rows by 8 columns");
  float* lut = (float*)malloc(65536 * sizeof(float));
  dor_response_lut(lut);
  dice_prepass_result prepass;
  int32_t status = dor_prepass(prepass_rgbi, prepass_height, prepass_width, lut,
                               prepass_params, &prepass);
  if (status != DICE_OK) {
    free(lut);
    return status;
  }
  /* profile.py:48-55, 108-173: record-derived parameters are validated even
   * though the provider replaces them per row. */
  float alpha = get_f32(prepass.payload, 0x04);
  float mean_r = get_f32(prepass.payload, 0x28);
  float mean_ir = get_f32(prepass.payload, 0x2C);
  double denominator = 1.0 - (double)alpha;
  if (denominator == 0.0) {
    free(lut);
    return fail(DICE_INVALID, "prepass alpha cannot equal one");
  }
  float base = (float)(((double)mean_ir - (double)alpha * (double)mean_r) / denominator);
  if (!isfinite(alpha) || !isfinite(base)) {
    free(lut);
    return fail(DICE_INVALID, "alpha must be finite");
  }

  int64_t rows = height;
  dice_schedule S;
  memset(&S, 0, sizeof S);
  S.weighted_visible_sum = (double*)malloc((size_t)rows * sizeof(double));
  S.weighted_infrared_sum = (double*)malloc((size_t)rows * sizeof(double));
  S.weight_sum = (double*)malloc((size_t)rows * sizeof(double));
  S.mean_visible = (float*)malloc((size_t)rows * sizeof(float));
  S.mean_infrared = (float*)malloc((size_t)rows * sizeof(float));
  S.accepted = (uint32_t*)malloc((size_t)rows * sizeof(uint32_t));
  S.scaled_input = (float*)malloc((size_t)rows * sizeof(float));
  S.input_scale = (float*)malloc((size_t)rows * sizeof(float));
  S.scale_denominator = (float*)malloc((size_t)rows * sizeof(float));
  S.scale_numerator = (float*)malloc((size_t)rows * sizeof(float));
  S.scale_add_denominator = (float*)malloc((size_t)rows * sizeof(float));
  S.scale_add_numerator = (float*)malloc((size_t)rows * sizeof(float));
  S.payloads = (uint8_t*)malloc((size_t)rows * 64);
  status = dor_producer_schedule(main_rgbi, height, width, lut, &S);

  int64_t count = height + 6;
  float* aux_alpha = (float*)malloc((size_t)count * sizeof(float));
  float* aux_offset = (float*)malloc((size_t)count * sizeof(float));
  float* aux_replacement = (float*)malloc((size_t)count * sizeof(float));
  uint8_t* aux_has = (uint8_t*)malloc((size_t)count);
  float* score_base = (float*)malloc((size_t)count * sizeof(float));
  float* coarse = (float*)malloc((size_t)count * sizeof(float));
  uint8_t* gate_secondary = (uint8_t*)malloc((size_t)count);
  int32_t* row_gate = (int32_t*)malloc((size_t)count * sizeof(int32_t));
  if (status == DICE_OK)
    status = dor_stage_table_from_schedule(&S, 1.0f, count, aux_alpha, aux_offset,
                                           aux_replacement, aux_has, score_base,
                                           coarse, gate_secondary, row_gate);
  if (status == DICE_OK) {
    dice_stage_table st = {count,      aux_alpha, aux_offset,     aux_replacement,
                           aux_has,    score_base, coarse,        gate_secondary,
                           row_gate};
    status = dor_replay(main_rgbi, height, width, lut, params, &st, 0x3045u,
                        out_rgb16, result, diagnostics, threads);
  }
  free(S.weighted_visible_sum);
  free(S.weighted_infrared_sum);
  free(S.weight_sum);
  free(S.mean_visible);
  free(S.mean_infrared);
  free(S.accepted);
  free(S.scaled_input);
  free(S.input_scale);
  free(S.scale_denominator);
  free(S.scale_numerator);
  free(S.scale_add_denominator);
  free(S.scale_add_numerator);
  free(S.payloads);
  free(aux_alpha);
  free(aux_offset);
  free(aux_replacement);
  free(aux_has);
  free(score_base);
  free(coarse);
  free(gate_secondary);
  free(row_gate);
  free(lut);
  return status;
}
