// dice_kernels.hpp -- launchers of the gfx950 kernels (dice_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dice_device.hpp"

namespace dice {

// Device copies of the ProducerRecordSchedule arrays (producer API only).
struct ScheduleDev {
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
};

// Batched launches: d_jobs is the device array of n descriptors, h_jobs its
// host copy (launch geometry only).  Every job of one launch uses the same
// analysis tile shape.
hipError_t launch_zero_jobs(const JobDev* d_jobs, const JobDev* h_jobs, int n,
                            hipStream_t stream);
hipError_t launch_producer(const JobDev* d_jobs, const JobDev* h_jobs, int n,
                           hipStream_t stream);
hipError_t launch_schedule_rows(const JobDev& J, const ScheduleDev& S,
                                hipStream_t stream);
bool tile_shape_supported(int ty, int tx);
// The compiled analysis tile shapes {rows, cols}; returns how many exist.
int tile_shapes(int (*out)[2], int max);
hipError_t launch_scan(const JobDev* d_jobs, const JobDev* h_jobs, int n, hipStream_t stream);
hipError_t launch_scan_segments(const JobDev* d_jobs, const JobDev* h_jobs, int n,
                                hipStream_t stream);
// `queue`: two zeroed uint32 the launch uses as its ticket counter (and
// leaves zeroed); one pair per stream.  `variant` selects the workgroup
// of k_sites (0: 512 or 1024 threads at 4 waves per SIMD; 1: 256 threads
// at the LDS occupancy of the tile shape); every variant computes the same
// sites.
constexpr int kSitesVariants = 2;
hipError_t launch_sites(const JobDev* d_jobs, const JobDev* h_jobs, int n, uint32_t* queue,
                        hipStream_t stream, int variant = 0);
// Adds a 64-bit digest of `bytes` (even) of device memory to *out (device).
hipError_t launch_digest(const void* data, size_t bytes, unsigned long long* out,
                         hipStream_t stream);
hipError_t launch_startup(const JobDev* d_jobs, const JobDev* h_jobs, int n,
                          hipStream_t stream);
hipError_t launch_writer_prep(const JobDev* d_jobs, int n, hipStream_t stream);
// Start states of sites [begin, end) from J.group_state (the state at every
// kStateGroup-th site), written to J.site_state.
hipError_t launch_writer_expand(const JobDev& J, uint32_t begin, uint32_t end,
                                hipStream_t stream);
// Per job: the writer over sites [begin, end) and the startup check; the
// state after the range is end_state, or *end_state_at when that is given
// (device memory written by the device chain).
hipError_t launch_writer_values(const JobDev& J, uint32_t begin, uint32_t end,
                                uint32_t end_state, const uint32_t* end_state_at,
                                hipStream_t stream);
hipError_t launch_writer_startup_check(const JobDev& J, uint32_t final_state,
                                       const uint32_t* final_at, hipStream_t stream);
// The serial writer chain on one wavefront per job of the device array
// d_jobs[0, n) (jobs marked skip return at once), carrying each job's state
// in its J.chain: the 6*W hidden startup slots, or the main sites of band
// `band` of min(bands, height) (rows [H*b/bands, H*(b+1)/bands)).
hipError_t launch_writer_startup_chain(const JobDev* d_jobs, int n, hipStream_t stream);
hipError_t launch_writer_chain(const JobDev* d_jobs, int n, int band, int bands,
                               hipStream_t stream);
hipError_t launch_fill_u8(uint8_t* p, size_t n, uint8_t value, hipStream_t stream);

hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s);
// dice_canary.hip: the level-0 exactness canaries, element-wise over n.
hipError_t launch_fp_canaries(int64_t n, const double* a, const double* b, const double* c,
                              const float* fa, const float* fb, const float* fc,
                              double* mul_add, double* quotient, float* narrowed,
                              float* f32_mul_add, float* f32_quotient, float* f32_product,
                              hipStream_t s);
hipError_t launch_t_unscaled_averages(const float* patches, long count, double* out,
                                      hipStream_t s);
hipError_t launch_t_rgb_unscaled_averages(const float* patches, long count,
                                          double* out, hipStream_t s);
hipError_t launch_t_dither_chain(const double* values, const float* scales, long count,
                                 uint32_t state_in, float low, float high,
                                 double* deltas, uint32_t* state_out,
                                 unsigned long long* advances, hipStream_t s);
hipError_t launch_t_emit(const float* values, long count, uint16_t* out,
                         const uint32_t* high, const uint32_t* low, uint32_t* flags,
                         hipStream_t s);
hipError_t launch_t_planes(const uint16_t* rgbi, int height, int width,
                           const float* lut, const DevParams& P, const float* alpha,
                           const uint8_t* is_one, const float* replacement,
                           const float* offset, const float* score_base,
                           float* working, float* aux, float* score, float* waux,
                           float* wrgb, uint32_t* flags, hipStream_t s);
hipError_t launch_t_decision(const float* aux, int height, int width,
                             const DevParams& P, uint8_t* fallback, hipStream_t s);

}  // namespace dice
