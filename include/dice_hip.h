/*
 * dice_hip.h -- C ABI of the MI355X HIP backend for the portable Digital ICE
 * processing core (LS-5000, selector 8, ICE Normal, metric 4000).
 *
 * The library (libdice_hip.so) is loaded by Python through ctypes.  Every
 * entry point takes plain pointers and sizes; no torch or HIP types cross
 * this boundary.  Host buffers are C-contiguous and owned by the caller; the
 * library keeps no pointer after a call returns.
 *
 * Reference interfaces replaced (paths relative to the reference repo
 * rohanpandula/digital-fauxice):
 *   dice_process_job          process_cuda(job, ...)
 *                             src/portable_digital_ice/cuda_backend/process.py:25-159
 *                             (same orchestration as engine.py:110-267)
 *   dice_replay               run_streaming_replay_cuda(main_rgbi, output, ...)
 *                             src/portable_digital_ice/cuda_backend/engine.py:183-590
 *                             (semantics: streaming.py:365-552)
 *   dice_producer_schedule    derive_producer_record_schedule_cuda(module, rgbi, lut)
 *                             src/portable_digital_ice/cuda_backend/producer.py:38-246
 *                             (semantics: producer_parameters.py:380-495)
 *   dice_prepass              reduce_prepass_frame(frame, ...)
 *                             src/portable_digital_ice/prepass.py:96-236
 *   dice_stage_table_from_schedule
 *                             ContentDerivedStageParameterProvider
 *                             src/portable_digital_ice/producer_parameters.py:498-541
 *                             + StageCalibration.from_setter_inputs
 *                             src/portable_digital_ice/stage_parameters.py:93-197
 *   dice_probe                cuda_device_summary()
 *                             src/portable_digital_ice/cuda_backend/engine.py:77-100
 *   dice_response_lut         generate_nikon_response_lut()  x3a.py:111-132
 *   dice_inverse_factors      InverseResponseFactors.recovered_16bit()  output.py:39-58
 *   dice_t_*                  the level-1 test wrappers appended to the CUDA
 *                             kernel source in tests/test_cuda_level1_primitives.py:45-99
 *                             and the level-1 kernels k_convert_and_auxiliary,
 *                             k_score_and_weighted, k_decision_eligibility
 *                             (cuda_backend/kernels.py:219-332)
 *
 * Return codes (every int-returning entry point):
 *   0 DICE_OK
 *   1 DICE_UNAVAILABLE       -> HipBackendUnavailable (no device, no HBM, ...)
 *   2 DICE_INVALID           -> ValueError (message from dice_last_error)
 *   3 DICE_NONFINITE_AUX     -> ValueError("auxiliary score input must be a finite HxW plane")
 *   4 DICE_NONFINITE_WORK    -> ValueError("work RGB must be finite")
 *   5 DICE_CANCELLED         -> ProcessingCancelled
 *   6 DICE_OVER_BUDGET       -> HipBackendUnavailable (HBM budget)
 *   7 DICE_INTERNAL          -> HipBackendUnavailable (HIP runtime error)
 *   8 DICE_PENDING           -> TimeoutError (dice_engine_wait: not finished yet)
 */
#ifndef DICE_HIP_H
#define DICE_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DICE_ABI_VERSION 1

enum dice_status {
  DICE_OK = 0,
  DICE_UNAVAILABLE = 1,
  DICE_INVALID = 2,
  DICE_NONFINITE_AUX = 3,
  DICE_NONFINITE_WORK = 4,
  DICE_CANCELLED = 5,
  DICE_OVER_BUDGET = 6,
  DICE_INTERNAL = 7,
  DICE_PENDING = 8
};

/* Stable (non-stage) replay parameters.  Every float is the binary32 value
 * the reference quantizes the dataclass field to before use
 * (x3a.py:193-286, reconstruction.py:614-670, dither.py:34-56,
 * profile.py:108-173). */
typedef struct dice_params {
  int32_t visible_channel;      /* AuxiliaryParameters.selected_visible_channel */
  float score_addend;           /* ScoreParameters.base_addend */
  float score_scale;            /* ScoreParameters.scale */
  float score_offset;           /* ScoreParameters.offset */
  float score_floor;            /* ScoreParameters.floor */
  int32_t horizontal_minimum;   /* cutoff < resolution_metric (x3a.py:266-268) */
  float decision_threshold;     /* DecisionParameters.sample_threshold */
  int32_t decision_count_limit; /* DecisionParameters.count_limit */
  int32_t decision_radius;      /* DecisionParameters.perpendicular_radius */
  int32_t cross_neighbor;       /* metric > cross_neighbor_cutoff (reconstruction.py:561-572) */
  int32_t coarse_enabled;
  float coarse_slopes[3];
  int32_t band_enabled[3];
  float band_scales[3];
  float factors_a[9];           /* [band][channel] */
  float factors_b[9];
  float configured_strengths[3];
  int32_t gate_primary;         /* driver_gate_primary */
  float dither_scales[3];
  float dither_low;             /* DitherBounds.low */
  float dither_high;            /* DitherBounds.high */
} dice_params;

/* Per-stage calibration, index s-1 for scheduler stage s = 1..count.
 * Row y consumes aux stage y+1, score stage y+2 and writer stage y+7
 * (stage_parameters.py:207-232); the hidden startup consumes stages 1..6.
 * A replay of H rows therefore needs count >= H + 6. */
typedef struct dice_stage_table {
  int64_t count;
  const float* aux_alpha;
  const float* aux_offset;
  const float* aux_replacement;
  const uint8_t* aux_has_replacement;
  const float* score_base;
  const float* coarse_reference;
  const uint8_t* gate_secondary;
  const int32_t* row_gate;
} dice_stage_table;

/* StreamingReplayResult (streaming.py:63-74) + StartupReplayResult
 * (startup.py:37-43). */
typedef struct dice_replay_result {
  uint64_t attempted_pixels;
  uint64_t written_pixels;
  uint64_t public_rng_advances;
  uint64_t changed_pixels;
  uint32_t final_rng_state;
  uint32_t startup_final_rng_state;
  int64_t startup_attempted[6];
  int64_t startup_advances[6];
  char output_sha256[72]; /* lowercase hex, NUL-terminated */
} dice_replay_result;

/* ProcessingDiagnostics planes (engine.py:40-66); any pointer may be NULL. */
typedef struct dice_diagnostics {
  float* score_plane;    /* H*W float32 */
  uint8_t* at_floor_mask; /* H*W bool */
  uint8_t* changed_mask;  /* H*W bool */
} dice_diagnostics;

/* ProducerRecordSchedule (producer_parameters.py:84-129).  Every array has
 * `rows` elements except payloads (rows * 64 bytes).  Caller-allocated. */
typedef struct dice_schedule {
  int64_t rows;
  int64_t active_width;
  double* weighted_visible_sum;
  double* weighted_infrared_sum;
  double* weight_sum;
  float* mean_visible;
  float* mean_infrared;
  uint32_t* accepted;
  float* scaled_input;
  float* input_scale;
  float* scale_denominator;
  float* scale_numerator;
  float* scale_add_denominator;
  float* scale_add_numerator;
  uint8_t* payloads;
} dice_schedule;

/* Prepass constants (prepass.py:42-53). */
typedef struct dice_prepass_params {
  float valid_ir_threshold;
  float alpha_scale;
  float ratio_minimum;
  float ratio_maximum;
  float zero_denominator_alpha;
  uint32_t completion_marker;
} dice_prepass_params;

typedef struct dice_prepass_result {
  uint8_t payload[64];
  int64_t valid_pixels;
  int64_t fully_valid_tiles;
} dice_prepass_result;

/* Progress / cancellation hooks, called between stages on the calling
 * thread.  phase: 0 validating, 1 prepass, 2 producer, 3 reconstruction,
 * 4 complete (engine.py:22-37).  cancelled() returning nonzero aborts with
 * DICE_CANCELLED before any caller-owned output is written. */
typedef void (*dice_progress_fn)(void* user, int32_t phase, int64_t completed,
                                 int64_t total, int64_t attempted,
                                 int64_t written);
typedef int32_t (*dice_cancelled_fn)(void* user);
typedef struct dice_callbacks {
  dice_progress_fn progress;
  dice_cancelled_fn cancelled;
  void* user;
} dice_callbacks;

/* Stage timings in seconds (cuda_backend/engine.py:119-165 keys). */
typedef struct dice_stage_timings {
  double prepass;
  double producer_device;
  double upload;
  double analysis_device;
  double download_sites;
  double writer_host;
  double finish_device;
  double download_output;
  double sha256_host;
  double total;
  /* Device durations (milliseconds) measured with HIP events on the
   * library's stream around each kernel group. */
  double producer_kernels_ms;
  double scan_kernel_ms;
  double sites_kernel_ms;
  double startup_kernels_ms;
  double segments_kernel_ms;
  double writer_prep_kernel_ms;
  double writer_values_kernel_ms;
} dice_stage_timings;

typedef struct dice_device_info {
  char name[128];
  char arch[64];
  int32_t compute_units;
  int32_t device_count;
  uint64_t total_hbm_bytes;
  uint64_t free_hbm_bytes;
  char build_flags[256];
  char reason[256];
} dice_device_info;

typedef struct dice_ctx dice_ctx;

/* Job flags for dice_process_job / dice_replay. */
#define DICE_MAIN_ON_DEVICE 0x1u /* main RGBI pointer is device memory */
#define DICE_OUT_ON_DEVICE 0x2u  /* output RGB16 pointer is device memory */
#define DICE_SKIP_SHA256 0x4u    /* leave output_sha256 empty */

/* ---- Batches ------------------------------------------------------------
 * No counterpart in the reference, which processes one job per call
 * (engine.py:110-267): many independent jobs (crops or frames of uneven
 * size) run through the same device phases together, one persistent kernel
 * launch per phase for the whole batch.  Each job is exactly the job
 * dice_process_job would run (the profile's prepass, producer and replay);
 * its result is bit-identical to running it alone. */
typedef struct dice_job {
  const uint16_t* main_rgbi;          /* H*W*4, host or device (flags) */
  int64_t height, width;
  const uint16_t* prepass_rgbi;       /* host */
  int64_t prepass_height, prepass_width;
  const dice_params* params;
  const dice_prepass_params* prepass_params;
  uint32_t flags;                     /* DICE_MAIN_ON_DEVICE | DICE_OUT_ON_DEVICE | DICE_SKIP_SHA256 */
  uint16_t* out_rgb16;                /* H*W*3, host or device (flags) */
  dice_replay_result* result;         /* required */
  dice_stage_timings* timings;        /* optional */
  int32_t status;                     /* out: this job's return code */
  char error[256];                    /* out: this job's message when status != 0 */
} dice_job;

#define DICE_BATCH_GRAPH 0x100u    /* replay the device phases from captured hipGraphs */
#define DICE_BATCH_AUTOTUNE 0x200u /* time every tile shape on this batch geometry first */
#define DICE_BATCH_DEVICE_CHAIN 0x400u /* run each job's serial writer chain on one device
                                          wavefront instead of a host core */

typedef struct dice_batch_report {
  int32_t tile_rows, tile_cols;   /* analysis tile shape used */
  int32_t graph_replayed;         /* 1 when both device phases replayed a captured graph */
  double phase_a_ms, phase_b_ms;  /* device time of the two batched phases */
  uint64_t workspace_bytes;       /* HBM held by the context's job workspaces */
  double wall_seconds;
  int32_t sites_variant;          /* site-kernel workgroup variant used (0 wide, 1 256 threads) */
  int32_t tune_source;            /* DICE_BATCH_AUTOTUNE: 1 timed on this batch, 2 a winner
                                     held for its bucket, 3 read from the tuning cache */
  int32_t tune_candidates;        /* configurations timed (tune_source 1) */
  int32_t tune_rejected;          /* of those, configurations whose results differed */
} dice_batch_report;

/* Where the tuner's winners persist (no reference counterpart: the CUDA
 * backend uses fixed launch shapes, cuda_backend/engine.py:298-423).  A
 * winner is kept per (gfx architecture, bucket: job count and the total and
 * largest frame sizes, each to a power of two).  NULL selects the default,
 * $DICE_TUNING_CACHE or ~/.cache/dice_hip/tuning.tsv.  Forgets the winners
 * held in memory, so the next tuned batch reads the file again. */
int32_t dice_tuning_cache(const char* path);

int32_t dice_abi_version(void);
const char* dice_last_error(void);

int32_t dice_probe(dice_device_info* info);
int32_t dice_ctx_create(int32_t device, uint64_t hbm_budget, dice_ctx** out);
void dice_ctx_destroy(dice_ctx* ctx);

int32_t dice_response_lut(float* lut65536);
int32_t dice_inverse_factors(uint32_t* high256, uint32_t* low256);

int32_t dice_prepass(const uint16_t* pixels, int64_t height, int64_t width,
                     const float* lut, const dice_prepass_params* params,
                     dice_prepass_result* out);

int32_t dice_producer_schedule(dice_ctx* ctx, const uint16_t* rgbi,
                               int64_t height, int64_t width,
                               const float* lut, dice_schedule* out);

/* Arrays of `count` stages derived from a producer schedule of `rows` rows;
 * stage s reads record row min(s-1, rows-1). */
int32_t dice_stage_table_from_schedule(
    const dice_schedule* schedule, float auxiliary_factor_b,
    int64_t count, float* aux_alpha, float* aux_offset, float* aux_replacement,
    uint8_t* aux_has_replacement, float* score_base, float* coarse_reference,
    uint8_t* gate_secondary, int32_t* row_gate);

/* ---- Serving engine -----------------------------------------------------
 * No counterpart in the reference (one job per call).  Jobs are submitted as
 * they arrive; drivers (threads, each with its own HIP stream and job
 * workspaces) take batches from the head of the FIFO queue as soon as they
 * are free, admitting a job only while the HBM reservations of the jobs on
 * the device fit hbm_budget (continuous batching).  A submitted dice_job and
 * every buffer it points to stay owned by the caller and must stay valid
 * until dice_engine_wait has returned for its ticket; the engine writes only
 * its result, timings, status and error. */
typedef struct dice_engine dice_engine;

typedef struct dice_engine_config {
  int32_t device;
  int32_t drivers;          /* concurrent batches, 1..16 (0: 3) */
  int32_t max_batch;        /* jobs per batch, 1..64 (0: 8) */
  uint32_t batch_flags;     /* DICE_BATCH_GRAPH | DICE_BATCH_AUTOTUNE */
  uint64_t hbm_budget;      /* HBM the drivers' job workspaces may hold (0: 80% of free HBM);
                               jobs are admitted by their dice_engine_reservation */
  double linger_ms;         /* a partial batch waits this long for more jobs while another
                               batch is running (0: never waits) */
} dice_engine_config;

typedef struct dice_engine_report {
  uint64_t submitted, completed, failed, queued, running;
  uint64_t batches;
  double mean_batch;              /* jobs per batch */
  double mpix_completed;
  double latency_p50_ms, latency_p99_ms, latency_max_ms;  /* submit -> finished */
  double queue_p50_ms;            /* submit -> started */
  uint64_t admitted_bytes, admitted_bytes_peak, hbm_budget;  /* the most the drivers'
                                     workspaces can hold (now, peak), and its limit */
  uint64_t workspace_bytes;       /* HBM the idle drivers' workspaces hold */
  double elapsed_seconds;         /* first submit -> last finish */
  uint64_t workspace_bytes_peak;  /* the largest workspace_bytes seen after a batch */
  uint64_t workspace_releases;    /* times an idle driver freed its workspaces for the queue */
} dice_engine_report;

int32_t dice_engine_create(const dice_engine_config* config, dice_engine** out);
void dice_engine_destroy(dice_engine* engine);
/* DICE_OVER_BUDGET when the job alone cannot fit the budget. */
int32_t dice_engine_submit(dice_engine* engine, dice_job* job, uint64_t* ticket);
/* The bytes of HBM a job is admitted against: an upper bound of what it can
 * add to a driver's workspaces (every pixel an attempted site, buffer
 * headroom and rounding included). */
int32_t dice_engine_reservation(const dice_job* job, uint64_t* bytes);
/* Returns the job's status (DICE_PENDING when timeout_seconds >= 0 runs out;
 * a negative timeout waits until the job is done).  A finished ticket is
 * forgotten once waited for. */
int32_t dice_engine_wait(dice_engine* engine, uint64_t ticket, double timeout_seconds);
int32_t dice_engine_stats(dice_engine* engine, dice_engine_report* stats);
/* Starts a new measurement window: counters, latencies and the admission
 * peak restart from now (jobs in flight stay in flight). */
int32_t dice_engine_reset_stats(dice_engine* engine);

/* Runs jobs[0..count) (1..64) as one batch.  Returns 0 when every job
 * succeeded, else the first failing job's code (each job's own outcome is in
 * jobs[i].status / .error; other jobs are unaffected). */
int32_t dice_process_batch(dice_ctx* ctx, dice_job* jobs, int32_t count, uint32_t batch_flags,
                           dice_batch_report* report);

int32_t dice_replay(dice_ctx* ctx, const uint16_t* rgbi, int64_t height,
                    int64_t width, const float* lut, const dice_params* params,
                    const dice_stage_table* stages, uint32_t rng_state,
                    uint32_t flags, uint16_t* out_rgb16,
                    dice_replay_result* result, dice_diagnostics* diagnostics,
                    const dice_callbacks* callbacks,
                    dice_stage_timings* timings);

int32_t dice_process_job(dice_ctx* ctx, const uint16_t* main_rgbi,
                         int64_t height, int64_t width,
                         const uint16_t* prepass_rgbi, int64_t prepass_height,
                         int64_t prepass_width, const dice_params* params,
                         const dice_prepass_params* prepass_params,
                         uint32_t flags, uint16_t* out_rgb16,
                         dice_replay_result* result,
                         dice_diagnostics* diagnostics,
                         const dice_callbacks* callbacks,
                         dice_stage_timings* timings);

/* The allocation-free variant of dice_process_job.  It has no counterpart in
 * the reference; it is for callers that cannot allocate device memory: code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc.  The job is dice_process_job's with flags DICE_MAIN_ON_DEVICE |
 * DICE_OUT_ON_DEVICE | DICE_SKIP_SHA256 (identical output and counters,
 * output_sha256 left empty), except that:
 *   * every device intermediate lives in one file-scope device array of
 *     DICE_NOALLOC_SCRATCH_BYTES that each GPU holds from the library's load
 *     until the process ends, and the writer chain runs on one device
 *     wavefront, so the call allocates nothing (no hipMalloc, no pinned or
 *     heap memory);
 *   * a job of more than 16384 rows, or whose intermediates (about 3 B per
 *     pixel plus 68 B per attempted site) do not fit the array, is refused
 *     with DICE_OVER_BUDGET;
 *   * on failure the device output holds partial values.
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (from different contexts of that device) race and must not be made.
 * main_rgbi and out_rgb16 are device pointers; the prepass is host memory. */
#define DICE_NOALLOC_SCRATCH_BYTES (512ull << 20)
int32_t dice_process_job_noalloc(dice_ctx* ctx, const uint16_t* main_rgbi,
                                 int64_t height, int64_t width,
                                 const uint16_t* prepass_rgbi,
                                 int64_t prepass_height, int64_t prepass_width,
                                 const dice_params* params,
                                 const dice_prepass_params* prepass_params,
                                 uint16_t* out_rgb16, dice_replay_result* result);

/* Level-1 test entry points: the same device functions the replay kernels
 * use, launched over host-provided vectors. */
int32_t dice_t_unscaled_averages(dice_ctx* ctx, const float* patches,
                                 int64_t count, double* out3);
int32_t dice_t_rgb_unscaled_averages(dice_ctx* ctx, const float* patches,
                                     int64_t count, double* out9);
int32_t dice_t_dither_chain(dice_ctx* ctx, const double* values,
                            const float* scales, int64_t count,
                            uint32_t state_in, float low, float high,
                            double* deltas, uint32_t* state_out,
                            uint64_t* advances_out);
int32_t dice_t_emit(dice_ctx* ctx, const float* values, int64_t count,
                    uint16_t* out);
/* Planes for every row with explicit per-row parameter arrays:
 * working H*W*4, aux/score/waux H*W, wrgb H*W*3. */
int32_t dice_t_planes(dice_ctx* ctx, const uint16_t* rgbi, int64_t height,
                      int64_t width, const float* lut,
                      const dice_params* params, const float* row_alpha,
                      const uint8_t* row_alpha_is_one,
                      const float* row_replacement, const float* row_offset,
                      const float* row_score_base, float* working, float* aux,
                      float* score, float* waux, float* wrgb,
                      uint32_t* error_flags);
/* Decision fallback (1) / eligible (0) for every pixel of an aux plane,
 * with the floor gate disabled and row gate 0. */
int32_t dice_t_decision(dice_ctx* ctx, const float* aux, int64_t height,
                        int64_t width, const dice_params* params,
                        uint8_t* fallback);

/* Level-0 compiler canaries (no reference entry point: the reference's CUDA
 * backend runs the same probes at module load, docs/cuda-decision-record.md:
 * 23-40, and tests/test_fast_cpu_bitexact.py:347-375 checks its CPU build).
 * Element-wise over `count` operands read from device memory:
 *   mul_add = a*b + c (binary64, two roundings), quotient = a/b,
 *   narrowed = (float)a, f32_mul_add = fa*fb + fc (binary32, two roundings),
 *   f32_quotient = fa/fb, f32_product = fa*fb (denormal results kept). */
typedef struct dice_fp_canary {
  int64_t count;
  const double* a;
  const double* b;
  const double* c;
  const float* fa;
  const float* fb;
  const float* fc;
  double* mul_add;
  double* quotient;
  float* narrowed;
  float* f32_mul_add;
  float* f32_quotient;
  float* f32_product;
} dice_fp_canary;

int32_t dice_t_fp_canaries(dice_ctx* ctx, const dice_fp_canary* io);

/* The serial writer on the host, no device involved: the integer draw
 * schedule of each site (the kernel's writer_record, compiled for the host)
 * and the chain over them, exactly as the replay runs it; `generic` 1
 * forces the per-channel path for every site, 2 takes the device chain's
 * lane-parallel form (evaluated lane by lane) wherever all three channels
 * draw, 3 keeps every 16th start state only and re-walks each group from it
 * (the host walk and the device expansion of the replay).  Sites: cand n*3
 * (binary64),
 * orig n*3, floor n (writer floor gate on/off).  Writes each site's start
 * state.  Stands in for the writer loop of write_band
 * (fast_cpu/kernels.py:1139-1204) in the L1 tests. */
int32_t dice_t_writer_chain_host(const double* cand, const float* orig,
                                 const uint8_t* floor, int64_t count, float low,
                                 float high, const float* scales,
                                 uint32_t state_in, int32_t generic,
                                 uint32_t* states, uint32_t* state_out,
                                 uint64_t* advances_out);

/* One site's serial writer step from each of `count` start states (the
 * same schedule builder and walk as dice_t_writer_chain_host): the end
 * state and the advances taken from each. */
int32_t dice_t_writer_site_host(const double* cand, const float* orig,
                                uint8_t floor, float low, float high,
                                const float* scales, const uint32_t* states_in,
                                int64_t count, int32_t generic,
                                uint32_t* states_out, uint8_t* advances);

/* SHA-256 of `count` (1..64) messages fed in pieces, as the batch runtime
 * hashes the replay outputs band by band (the reference: one
 * hashlib.sha256() per replay updated row by row, streaming.py:425, :550).
 * Message i is cut into pieces of piece*(1 + i%3) + i bytes.  mode 0: groups
 * of four messages, each round of pieces compressed interleaved; 1: each
 * piece on its own.  Writes each digest as 64 lowercase hex digits plus a
 * NUL at digests + 72*i, and whether the x86 SHA extensions were used. */
int32_t dice_t_sha256_pieces(const uint8_t* const* messages,
                             const int64_t* lengths, int32_t count,
                             int64_t piece, int32_t mode, char* digests,
                             int32_t* extensions);

/* When the batch runtime hashes the jobs of one geometry in pairs, one
 * thread per pair compressing the two outputs interleaved: 0 never, 1
 * always, 2 when the messages being hashed in the process would outnumber
 * the host's cores (the default).  Applies to batches that start afterwards;
 * writes the previous mode.  Lets the tests reach both hash paths. */
int32_t dice_t_hash_pairs(int32_t mode, int32_t* previous);

/* ---- At-floor routing ------------------------------------------------------
 * route_at_floor_mask (hybrid/src/fauxce_hybrid/routing.py:447-687) on the
 * device, for the hybrid step after the core:
 *   * provisional regions: the 8-connected components of the at-floor mask,
 *     numbered in raster order of their first pixel, with area, half-open
 *     bounding box and largest chessboard distance to in-frame background;
 *   * routed = (area >= min_area || radius >= min_radius), except for a
 *     component holding every perimeter pixel;
 *   * the synthesis mask: the routed union dilated by a square of radius
 *     `margin` (clipped to the frame), minus that perimeter component;
 *   * final regions: its 8-connected components.  Those whose box grown by
 *     128 px passes 2048 x 2048 pixels are cut on a 1792 px grid from the
 *     box origin.  They are numbered by (y0, x0, y1, x1, creation order).
 * The selection reasons, dispositions, JSON census and synthesis budget are
 * assembled from these tables by the caller (dice_amd/routing.py).  The
 * mask must have at least one pixel off the floor.  A routing holds its
 * planes in HBM until dice_routing_free. */
#define DICE_ROUTING_MASK_ON_DEVICE 0x1u

typedef struct dice_routing dice_routing;

typedef struct dice_routing_summary {
  int64_t height, width;
  int64_t provisional_regions, final_regions, pairs;
  int64_t at_floor_pixels, synthesis_pixels, non_floor_synthesis_pixels;
  int64_t perimeter_excluded_id;              /* 0: none */
  int64_t perimeter_suppressed_halo_pixels;   /* dilated halo removed from it */
} dice_routing_summary;

typedef struct dice_provisional_region {      /* id = index + 1 */
  int32_t y0, x0, y1, x1;
  int64_t area;
  int64_t max_chessboard_radius;
  int64_t synthesized_pixels;                 /* its pixels inside the synthesis mask */
  int32_t routed;                             /* directly routed */
  int32_t reserved;
} dice_provisional_region;

typedef struct dice_final_region {            /* id = index + 1 */
  int32_t y0, x0, y1, x1;
  int64_t area;
  int64_t at_floor_pixels;
} dice_final_region;

int32_t dice_route_at_floor(dice_ctx* ctx, const uint8_t* at_floor, int64_t height,
                            int64_t width, uint32_t flags, int64_t min_area,
                            int64_t min_radius, int64_t margin, dice_routing** out);
int32_t dice_routing_summary_get(const dice_routing* routing, dice_routing_summary* summary);
/* provisional[provisional_regions], finals[final_regions], pairs[2 * pairs]
 * ((provisional id, final id) of every pair that shares an at-floor pixel
 * of the synthesis mask, sorted); any pointer may be null. */
int32_t dice_routing_tables(const dice_routing* routing, dice_provisional_region* provisional,
                            dice_final_region* finals, int32_t* pairs);
/* Host buffers of H*W; any may be null. */
int32_t dice_routing_planes(const dice_routing* routing, int32_t* provisional_labels,
                            uint8_t* synthesis_mask, int32_t* final_labels);
void dice_routing_free(dice_routing* routing);

/* ---- Frame streams (8-row scanner blocks) ---------------------------------
 * The reference replays a frame only once all of it is in memory
 * (streaming.py:365-552 takes main_rgbi whole); its transport hands the main
 * pass over as valid-row-only blocks with an absolute start row
 * (contracts.py:108-158 RGBI16Block / BlockSchedule, 8 rows per block, the
 * last one shorter) that assemble_rgbi16_blocks (contracts.py:224-261)
 * joins when they are ordered, gap-free, non-overlapping and of one width.
 * A stream takes those blocks as they arrive and runs every device phase
 * whose inputs are complete: producer blocks, row sums and the cumulative
 * chains, stage entries, analysis tile rows, sites, the writer draw
 * schedules, the serial writer chain over the new sites, their values, and
 * the copy and SHA-256 of the output rows that are final.  The result is
 * the one dice_process_batch gives for the whole frame (bit-exact output,
 * hash, counters and receipts).
 *
 * dice_stream_begin takes a dice_job whose main_rgbi is null (the frame
 * arrives through dice_stream_push) and whose prepass frame is complete;
 * DICE_MAIN_ON_DEVICE is rejected.  height/width are the block schedule's.
 * Pushes must continue at the first row not yet received; a push that
 * completes the frame runs the rest of it.  dice_stream_finish waits for
 * nothing more: it fills job->result, status and error (and the optional
 * diagnostics) and returns the job's status.  After an error every later
 * call returns it (a refused block is not an error of the frame).  A stream
 * is reusable: begin may follow finish, or abandon a frame in progress.  One
 * caller thread at a time per stream. */
typedef struct dice_stream dice_stream;

typedef struct dice_stream_progress {
  int64_t rows_received;   /* main rows pushed so far */
  int64_t rows_analyzed;   /* rows whose analysis tiles and sites are done */
  int64_t rows_final;      /* output rows written, copied and hashed */
  int64_t sites_written;   /* main sites through the writer */
  int64_t windows;         /* device windows run */
  double push_seconds;     /* wall time inside the last push */
} dice_stream_progress;

int32_t dice_stream_create(dice_ctx* ctx, dice_stream** out);
int32_t dice_stream_begin(dice_stream* stream, dice_job* job, dice_diagnostics* diagnostics);
int32_t dice_stream_push(dice_stream* stream, const uint16_t* rgbi, int64_t start_row,
                         int64_t rows);
int32_t dice_stream_finish(dice_stream* stream);
int32_t dice_stream_progress_get(const dice_stream* stream, dice_stream_progress* progress);
void dice_stream_destroy(dice_stream* stream);

/* Calibration: `repeats` launches of a plain 16-byte-vector copy kernel over
 * `bytes` (device buffers owned by the call); average milliseconds per copy. */
int32_t dice_t_copy(dice_ctx* ctx, uint64_t bytes, int32_t repeats,
                    double* ms_per_copy);

#ifdef __cplusplus
}
#endif

#endif /* DICE_HIP_H */
