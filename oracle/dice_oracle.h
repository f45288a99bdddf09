/*
 * dice_oracle.h -- CPU oracle for the Digital ICE processing core.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker (or the timed CPU baseline).  The HIP product never links,
 * imports or calls it.
 *
 * Every function restates the reference algorithm of
 * rohanpandula/digital-fauxice (cited file:line in dice_oracle.c), using the
 * struct layouts of include/dice_hip.h so tests can hand the same arguments
 * to the oracle and to the HIP library.
 */
#ifndef DICE_ORACLE_H
#define DICE_ORACLE_H

#include "../include/dice_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

const char* dor_last_error(void);

void dor_response_lut(float* lut65536);
void dor_inverse_factors(uint32_t* high256, uint32_t* low256);
void dor_dither_bounds(const float* lut, float* low, float* high,
                       int32_t* low_index, int32_t* high_index);

int32_t dor_prepass(const uint16_t* pixels, int64_t height, int64_t width,
                    const float* lut, const dice_prepass_params* params,
                    dice_prepass_result* out);

int32_t dor_producer_schedule(const uint16_t* rgbi, int64_t height,
                              int64_t width, const float* lut,
                              dice_schedule* out);

int32_t dor_stage_table_from_schedule(
    const dice_schedule* schedule, float auxiliary_factor_b,
    int64_t count, float* aux_alpha, float* aux_offset, float* aux_replacement,
    uint8_t* aux_has_replacement, float* score_base, float* coarse_reference,
    uint8_t* gate_secondary, int32_t* row_gate);

/* threads <= 0: use the OpenMP default. */
int32_t dor_replay(const uint16_t* rgbi, int64_t height, int64_t width,
                   const float* lut, const dice_params* params,
                   const dice_stage_table* stages, uint32_t rng_state,
                   uint16_t* out_rgb16, dice_replay_result* result,
                   dice_diagnostics* diagnostics, int32_t threads);

int32_t dor_process_job(const uint16_t* main_rgbi, int64_t height,
                        int64_t width, const uint16_t* prepass_rgbi,
                        int64_t prepass_height, int64_t prepass_width,
                        const dice_params* params,
                        const dice_prepass_params* prepass_params,
                        uint16_t* out_rgb16, dice_replay_result* result,
                        dice_diagnostics* diagnostics, int32_t threads);

/* Level-1 primitives. */
void dor_unscaled_averages(const float* patches, int64_t count, double* out3);
void dor_rgb_unscaled_averages(const float* patches, int64_t count,
                               double* out9);
void dor_dither_chain(const double* values, const float* scales, int64_t count,
                      uint32_t state_in, float low, float high, double* deltas,
                      uint32_t* state_out, uint64_t* advances_out);
void dor_write_pixels(const double* cand, const float* orig, const uint8_t* floor,
                      int64_t count, float low, float high, const float* scales,
                      uint32_t state_in, uint32_t* states, float* values,
                      uint32_t* state_out, uint64_t* advances_out);
void dor_write_pixel_states(const double* cand, const float* orig, uint8_t floor, float low,
                            float high, const float* scales, const uint32_t* states_in,
                            int64_t count, uint32_t* states_out, uint8_t* advances);
int32_t dor_emit(const float* values, int64_t count, uint16_t* out);
int32_t dor_planes(const uint16_t* rgbi, int64_t height, int64_t width,
                   const float* lut, const dice_params* params,
                   const float* row_alpha, const uint8_t* row_alpha_is_one,
                   const float* row_replacement, const float* row_offset,
                   const float* row_score_base, float* working, float* aux,
                   float* score, float* waux, float* wrgb);
void dor_decision(const float* aux, int64_t height, int64_t width,
                  const dice_params* params, uint8_t* fallback);

#ifdef __cplusplus
}
#endif

#endif
