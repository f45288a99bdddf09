// dice_kernels.hip -- CDNA4 (gfx950) kernels of the Digital ICE hot path.
//
// Layout in HBM (per job): interleaved RGBI uint16 input (8 B/px), RGB
// uint16 output (6 B/px); every intermediate plane of the reference
// (working, aux, score, weighted aux/RGB) lives only in LDS tiles.  The
// analysis kernel reads the input once per tile (plus its 9x9-window halo),
// writes the no-op emit of every pixel and one 40-byte Site record per
// attempted pixel; the serial writer runs on the host over those records.
//
// Reference semantics (rohanpandula/digital-fauxice,
// src/portable_digital_ice/): streaming.py (row loop), x3a.py (planes,
// decision), reconstruction.py (averages, candidates, combiner),
// startup.py (hidden rows), producer_parameters.py + stage_parameters.py
// (producer schedule and setter).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <type_traits>
#include <utility>

#include "dice_device.hpp"
#include "dice_kernels.hpp"

#pragma clang fp contract(off)

namespace dice {

namespace {

constexpr int kThreads = 256;
constexpr float kProducerThreshold = 8847.2255859375f;  // producer_parameters.py:36

__device__ __forceinline__ int clampi(int v, int lo, int hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

// Loads the four uint16 lanes of one pixel with one 64-bit load.
// The value of the other lane of this lane's pair (lanes 2m, 2m+1): a DPP
// quad permutation [1, 0, 3, 2] on each half, no LDS round trip.  Both
// lanes of the pair must be active.
__device__ __forceinline__ double swap_pair(double v) {
  const unsigned long long bits = (unsigned long long)__double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp((int)(uint32_t)bits, (int)(uint32_t)bits, 0xB1, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp((int)(uint32_t)(bits >> 32), (int)(uint32_t)(bits >> 32), 0xB1, 0xF, 0xF, false);
  return __longlong_as_double((long long)(((unsigned long long)(uint32_t)hi << 32) | (uint32_t)lo));
}

__device__ __forceinline__ void load_pixel(const uint16_t* rgbi, size_t index,
                                           uint32_t lanes[4]) {
  uint2 q = *reinterpret_cast<const uint2*>(rgbi + index * 4);
  lanes[0] = q.x & 0xFFFFu;
  lanes[1] = q.x >> 16;
  lanes[2] = q.y & 0xFFFFu;
  lanes[3] = q.y >> 16;
}

// Blocks are dispatched round-robin over the 8 XCDs; give each XCD a
// contiguous run of work items so neighbouring tiles share an L2.
__device__ __forceinline__ int xcd_contiguous_index(int block, int blocks) {
  int per = blocks / 8, rem = blocks % 8;
  int xcd = block % 8, slot = block / 8;
  int start = xcd * per + (xcd < rem ? xcd : rem);
  return start + slot;
}

// Exclusive scan of one value per thread over a block of `Threads`.
template <int Threads = kThreads>
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t value,
                                                         uint32_t* wave_sums,
                                                         uint32_t* total) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  uint32_t inclusive = value;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    uint32_t t = __shfl_up(inclusive, d, 64);
    if (lane >= d) inclusive += t;
  }
  if (lane == 63) wave_sums[wave] = inclusive;
  __syncthreads();
  uint32_t offset = 0, sum = 0;
#pragma unroll
  for (int w = 0; w < Threads / 64; ++w) {
    uint32_t s = wave_sums[w];
    offset += w < wave ? s : 0u;
    sum += s;
  }
  *total = sum;
  __syncthreads();
  return offset + inclusive - value;
}

// The job table is written before a launch and only read by kernels.
// Reading it through the constant address space lets the compiler load its
// fields with scalar loads and treat the buffer pointers read from it as
// global memory: global rather than flat loads and stores, whose shared
// completion counter would make every LDS wait also wait for memory.
__device__ __forceinline__ const JobDev& job_at(const JobSet& set, int j) {
  using ConstantJob = const __attribute__((address_space(4))) JobDev;
  return *(const JobDev*)((ConstantJob*)set.jobs + j);
}

// Job of work unit u: the last j with base[j] <= u (empty ranges skipped).
__device__ __forceinline__ int job_of(const JobSet& set, uint32_t u) {
  int lo = 0, hi = set.count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (set.base[mid] <= u)
      lo = mid;
    else
      hi = mid - 1;
  }
  return lo;
}

// Persistent loop over the work units of a job set.  The grid is sized to
// the machine, not to the work; each round hands every XCD a contiguous run
// of units (neighbouring tiles share halo rows in that XCD's L2).  Units are
// uniform per workgroup, so bodies may synchronize; the barrier after each
// unit lets the next one reuse LDS.
template <class Body>
__device__ __forceinline__ void for_each_unit(const JobSet& set, Body&& body) {
  const uint32_t total = set.base[set.count];
  const uint32_t me = (uint32_t)xcd_contiguous_index(blockIdx.x, gridDim.x);
  for (uint32_t first = 0; first < total; first += gridDim.x) {
    const uint32_t u = first + me;
    if (u < total) {
      const int j = job_of(set, u);
      const JobDev& J = job_at(set, j);
      if (!J.skip) body(J, (int)(u - set.base[j]));
    }
    __syncthreads();
  }
}

// Persistent loop that deals the units out one at a time from a ticket
// counter, for kernels whose units differ widely in cost (a tile of the
// reconstruction kernel holds anywhere from 0 to TY*TX sites): a workgroup
// takes the next unit as soon as it is free, so the launch ends about one
// unit after the average workgroup instead of with the unluckiest one.
// Every workgroup leaves once the tickets pass the unit count; the last one
// to leave resets the counter pair for the next launch on the stream.
template <class Body>
__device__ __forceinline__ void for_each_unit_dynamic(const JobSet& set, Body&& body) {
  __shared__ uint32_t ticket;
  const uint32_t total = set.base[set.count];
  uint32_t* queue = set.queue;
  for (;;) {
    if (threadIdx.x == 0) ticket = atomicAdd(&queue[0], 1u);
    __syncthreads();
    const uint32_t u = ticket;
    if (u >= total) break;
    const int j = job_of(set, u);
    const JobDev& J = job_at(set, j);
    if (!J.skip) body(J, (int)(u - set.base[j]));
    __syncthreads();  // every thread has read the ticket and left the unit's LDS
  }
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&queue[1], 1u) == gridDim.x - 1) {
      atomicExch(&queue[0], 0u);
      atomicExch(&queue[1], 0u);
    }
  }
}

// streaming.py:157-198: four nine-sample bars, edge-clamped by `aux_at`.
template <class AuxAt>
__device__ __forceinline__ bool decision_fallback(const AuxAt& aux_at,
                                                  float threshold, int radius,
                                                  int limit) {
  int vertical_left = 0, vertical_right = 0, above = 0, below = 0;
#pragma unroll
  for (int d = -4; d <= 4; ++d) {
    vertical_left += aux_at(d, -radius) < threshold;
    vertical_right += aux_at(d, radius) < threshold;
    above += aux_at(-radius, d) < threshold;
    below += aux_at(radius, d) < threshold;
  }
  return vertical_left > limit || above > limit || vertical_right > limit ||
         below > limit;
}

// reconstruction.py:425-462: C69/C21/C16 with the explicit float32 W.
template <class WrgbAt>
__device__ __forceinline__ void rgb_candidates(const WrgbAt& at,
                                               const float weights[4],
                                               float out[3][3]) {
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    auto lane = [&](int dy, int dx) { return at(dy, dx, k); };
    double q[3];
    unscaled_averages(lane, q);
    q[2] = rgb_q16(lane);
#pragma unroll
    for (int s = 0; s < 3; ++s) {
      if (weights[s] > 0.0f) {
        double reciprocal = 1.0 / (double)weights[s];
        out[s][k] = (float)(q[s] * reciprocal);
      } else {
        out[s][k] = (float)q[s];
      }
    }
  }
}

}  // namespace

// ==========================================================================
// Producer (producer_parameters.py:266-495)
// ==========================================================================

namespace {
// producer_parameters.py:136-167, evaluated one row of a 4x4 quadrant at a
// time: rows 0-1 feed `first` (six single adds, then the paired tail),
// rows 2 and 3 form `third` and `fourth` from pairs; the quadrant total is
// first + (third + fourth) and its mean f32(total * 0.0625).
struct QuadrantTree {
  double first, third;
  __device__ __forceinline__ void row(int local_row, double a0, double a1, double a2,
                                      double a3, float* mean) {
    if (local_row == 0) {
      first = a0;
      first += a1;
      first += a2;
      first += a3;
    } else if (local_row == 1) {
      first += a0;
      first += a1;
      const double tail = a2 + a3;
      first += tail;
    } else if (local_row == 2) {
      third = a0 + a1;
      const double tail = a2 + a3;
      third += tail;
    } else {
      double fourth = a0 + a1;
      const double tail = a2 + a3;
      fourth += tail;
      third += fourth;
      const double total = first + third;
      *mean = (float)(total * 0.0625);
    }
  }
};

// producer_parameters.py:170-180
__device__ __forceinline__ void producer_center(const float m[4], float out[4]) {
  double total = (double)m[0];
  total += (double)m[1];
  total += (double)m[2];
  total += (double)m[3];
  float mean = (float)(total * 0.25);
#pragma unroll
  for (int q = 0; q < 4; ++q) out[q] = (float)((double)m[q] - (double)mean);
}
}  // namespace

// One workgroup per 8-row block b.  Every thread owns one 8x8 cell (the
// 64 B of each of its rows are contiguous, so a wave reads 4 KB runs):
//   * failpos[b][cell]: first pixel with IR <= threshold in row-major order
//     within the cell, 64 when none -- the 8-row validity latch the row
//     sums apply (producer_parameters.py:303-340);
//   * for a complete epoch, the four (term, ratio*term) pairs of a block
//     whose 64 IR samples all pass (producer_parameters.py:209-262), zero
//     otherwise -- the sums start at +0.0 and can never become -0.0, so a
//     +0.0 term is the exact identity of the skipped block;
//   * thread 0 then adds the pairs in cell-then-quadrant order and narrows
//     the epoch's two additions once.
constexpr int kProducerThreads = 256;

__global__ __launch_bounds__(kProducerThreads) void k_producer_blocks(JobSet set) {
  __shared__ double pairs[kProducerThreads][8];
  for_each_unit(set, [&](const JobDev& J, const int unit) {
    const int H = J.height, W = J.width;
    const int cells = W / 8;
    const int b = J.win.block_lo + unit;
    const bool epoch = b < H / 8;
    const int tid = threadIdx.x;
    const double threshold = (double)kProducerThreshold;
    double denominator = 0.0, numerator = 0.0;
    for (int c0 = 0; c0 < cells; c0 += kProducerThreads) {
      const int c = c0 + tid;
      if (c < cells) {
        int fail = 64;
        float raw_sum = 0.0f;
        QuadrantTree vis[2], inf[2];
        float vm[4] = {0.0f, 0.0f, 0.0f, 0.0f}, im[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int r = 0; r < 8; ++r) {
          const int y = b * 8 + r;
          if (y >= H) break;
          const uint2* row = reinterpret_cast<const uint2*>(J.rgbi + ((size_t)y * W + c * 8) * 4);
          uint2 q[8];
#pragma unroll
          for (int k = 0; k < 8; ++k) q[k] = row[k];
          double v[8], ir[8];
#pragma unroll
          for (int k = 0; k < 8; ++k) {
            const uint32_t red = q[k].x & 0xFFFFu, infrared = q[k].y >> 16;
            if (fail == 64 && !(threshold < (double)infrared)) fail = r * 8 + k;
            raw_sum = (float)((double)raw_sum + (double)infrared);
            v[k] = (double)J.lut[red];
            ir[k] = (double)J.lut[infrared];
          }
          const int local = r & 3, half = (r >> 2) * 2;
          vis[0].row(local, v[0], v[1], v[2], v[3], &vm[half]);
          vis[1].row(local, v[4], v[5], v[6], v[7], &vm[half + 1]);
          inf[0].row(local, ir[0], ir[1], ir[2], ir[3], &im[half]);
          inf[1].row(local, ir[4], ir[5], ir[6], ir[7], &im[half + 1]);
        }
        J.failpos[(size_t)b * cells + c] = (uint8_t)fail;
        if (epoch) {
          double* out = pairs[tid];
          if (fail != 64) {
#pragma unroll
            for (int j = 0; j < 8; ++j) out[j] = 0.0;
          } else {
            float vd[4], id[4];
            producer_center(vm, vd);
            producer_center(im, id);
            const double low = (double)-0.125f, full_low = (double)0.0f;
            const double full_high = (double)0.3f, high = (double)0.425f;
            const double outer_weight = (double)0.70710677f;
            const double raw_wide = (double)raw_sum;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const double visible_wide = (double)vd[q];
              double ratio, weight;
              if (visible_wide == 0.0) {
                ratio = 0.0;
                weight = 0.0;
              } else {
                ratio = (double)id[q] / visible_wide;
                if (ratio < low || ratio > high)
                  weight = 0.0;
                else if (ratio < full_low || ratio > full_high)
                  weight = visible_wide * outer_weight;
                else
                  weight = visible_wide;
              }
              double term = weight * weight;
              term *= raw_wide;
              term *= raw_wide;
              out[q * 2 + 0] = term;
              out[q * 2 + 1] = ratio * term;
            }
          }
        }
      }
      if (!epoch) continue;
      __syncthreads();
      if (tid == 0) {
        const int n = min(kProducerThreads, cells - c0);
        for (int j = 0; j < n; ++j) {
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            denominator += pairs[j][q * 2 + 0];
            numerator += pairs[j][q * 2 + 1];
          }
        }
      }
      __syncthreads();
    }
    if (epoch && tid == 0) {
      J.epoch_adds[b * 2 + 0] = (double)(float)denominator;
      J.epoch_adds[b * 2 + 1] = (double)(float)numerator;
    }
  });
}

// Per-row binary64 sums in column order (producer_parameters.py:303-360).
// A workgroup owns kRowSumRows rows and pipelines 128-column chunks through
// three kinds of waves, one barrier per step:
//   loader waves     take turns: wave w holds chunk k = w (mod 4) in
//                    registers, moves it into a two-slot LDS ring at step k
//                    and requests chunk k + 4, so four chunks are in flight
//                    and each wave only ever waits for its own loads;
//   converter groups alternate chunks: at step c+1 a group reads chunk c
//                    from the ring and requests the latch byte and the two
//                    LUT values of every pixel; at step c+2 it turns them
//                    into the three binary64 terms of every pixel --
//                    weight * lut(visible), weight * lut(infrared) and
//                    weight, each one rounded product whichever thread
//                    forms it -- stored column-major;
//   the summing wave adds chunk c at step c+3: lane r*3+q carries chain q
//                    (visible, infrared, weight) of row r, with the LDS
//                    reads of the next 16 columns in flight while the
//                    current 16 are added.
// A rejected pixel has weight 0 and three +0.0 terms, the exact identity of
// the non-negative sums; so do the padding columns and rows past the frame,
// whose loads are clamped into the frame.
constexpr int kRowSumLanes = kRowSumRows * 3;
constexpr int kRowSumPitch = kRowSumLanes + 1;  // odd pitch: conflict-free column stores
constexpr int kRowSumCols = 128;
constexpr int kRowSumChunk = kRowSumRows * kRowSumCols;
constexpr int kRowSumLoaderWaves = 4;
constexpr int kRowSumGroup = 256;  // threads of one converter group
constexpr int kRowSumThreads = 64 * (1 + kRowSumLoaderWaves) + 2 * kRowSumGroup;
constexpr int kRowSumLoad = kRowSumChunk / 64;
constexpr int kRowSumConvert = kRowSumChunk / kRowSumGroup;
constexpr int kRowSumBatch = 16;
static_assert(kRowSumChunk % kRowSumGroup == 0, "whole pixels per converter");
static_assert(kRowSumCols % kRowSumBatch == 0, "whole summing batches");

__global__ __launch_bounds__(kRowSumThreads) void k_row_sums(JobSet set) {
  __shared__ uint2 s_raw[2][kRowSumChunk];
  __shared__ double s_terms[2][kRowSumCols][kRowSumPitch];
  for_each_unit(set, [&](const JobDev& J, const int unit) {
    const int H = J.height, W = J.width;
    const int cells = W / 8, active = cells * 8;
    const int row0 = (J.win.rowsum_lo + unit) * kRowSumRows;
    const int tid = threadIdx.x;
    const int chunks = (active + kRowSumCols - 1) / kRowSumCols;
    const int steps = chunks == 0 ? 0 : chunks + 3;
    if (tid < 64) {
      double sum = 0.0;
      uint32_t accepted = 0;
      for (int k = 0; k < steps; ++k) {
        if (k >= 3 && tid < kRowSumLanes) {
          const double(*terms)[kRowSumPitch] = s_terms[(k - 3) & 1];
          double now[kRowSumBatch];
#pragma unroll
          for (int i = 0; i < kRowSumBatch; ++i) now[i] = terms[i][tid];
#pragma unroll
          for (int c0 = kRowSumBatch; c0 <= kRowSumCols; c0 += kRowSumBatch) {
            double next[kRowSumBatch];
            if (c0 < kRowSumCols) {
#pragma unroll
              for (int i = 0; i < kRowSumBatch; ++i) next[i] = terms[c0 + i][tid];
            }
#pragma unroll
            for (int i = 0; i < kRowSumBatch; ++i) {
              sum += now[i];
              accepted += now[i] != 0.0;
            }
            if (c0 < kRowSumCols) {
#pragma unroll
              for (int i = 0; i < kRowSumBatch; ++i) now[i] = next[i];
            }
          }
        }
        __syncthreads();
      }
      const int y = row0 + tid / 3;
      if (tid < kRowSumLanes && y < H) {
        J.row_sums[(size_t)row0 * 3 + tid] = sum;
        if (tid % 3 == 2) J.accepted[y] = accepted;
      }
    } else if (tid < 64 * (1 + kRowSumLoaderWaves)) {
      const int wave = (tid >> 6) - 1, lane = tid & 63;
      uint2 raw[kRowSumLoad];
      auto fetch = [&](int chunk) {
#pragma unroll
        for (int i = 0; i < kRowSumLoad; ++i) {
          const int e = lane + i * 64;
          const int y = min(row0 + e / kRowSumCols, H - 1);
          const int x = min(chunk * kRowSumCols + e % kRowSumCols, active - 1);
          raw[i] = *reinterpret_cast<const uint2*>(J.rgbi + ((size_t)y * W + x) * 4);
        }
      };
      if (wave < chunks) fetch(wave);
      for (int k = 0; k < steps; ++k) {
        if (k < chunks && k % kRowSumLoaderWaves == wave) {
          uint2* ring = s_raw[k & 1];
#pragma unroll
          for (int i = 0; i < kRowSumLoad; ++i) ring[lane + i * 64] = raw[i];
          if (k + kRowSumLoaderWaves < chunks) fetch(k + kRowSumLoaderWaves);
        }
        __syncthreads();
      }
    } else {
      const int converter = tid - 64 * (1 + kRowSumLoaderWaves);
      const int group = converter / kRowSumGroup, t = converter % kRowSumGroup;
      uint2 pixel[kRowSumConvert];
      uint32_t fail[kRowSumConvert];
      float lut_visible[kRowSumConvert], lut_infrared[kRowSumConvert];
      for (int k = 0; k < steps; ++k) {
        if (k >= 1 && k - 1 < chunks && ((k - 1) & 1) == group) {
          const int chunk = k - 1;
          const uint2* ring = s_raw[chunk & 1];
#pragma unroll
          for (int i = 0; i < kRowSumConvert; ++i) {
            const int e = t + i * kRowSumGroup;
            const int y = min(row0 + e / kRowSumCols, H - 1);
            const int x = min(chunk * kRowSumCols + e % kRowSumCols, active - 1);
            pixel[i] = ring[e];
            fail[i] = J.failpos[(size_t)(y >> 3) * cells + (x >> 3)];
            lut_visible[i] = J.lut[pixel[i].x & 0xFFFFu];
            lut_infrared[i] = J.lut[pixel[i].y >> 16];
          }
        } else if (k >= 2 && k - 2 < chunks && ((k - 2) & 1) == group) {
          const int chunk = k - 2;
          double(*terms)[kRowSumPitch] = s_terms[chunk & 1];
#pragma unroll
          for (int i = 0; i < kRowSumConvert; ++i) {
            const int e = t + i * kRowSumGroup;
            const int r = e / kRowSumCols, c = e % kRowSumCols;
            const int y = row0 + r, x = chunk * kRowSumCols + c;
            const int position = (y & 7) * 8 + (x & 7);
            const bool keep = y < H && x < active && position < (int)fail[i];
            const uint32_t infrared = pixel[i].y >> 16;
            const double weight = keep ? (double)(infrared * infrared) : 0.0;
            terms[c][r * 3 + 0] = weight * (double)lut_visible[i];
            terms[c][r * 3 + 1] = weight * (double)lut_infrared[i];
            terms[c][r * 3 + 2] = weight;
          }
        }
        __syncthreads();
      }
    }
  });
}

// The strictly sequential cross-row chains (producer_parameters.py:
// 358-360, 437-456).  Workgroup 0: the cumulative row sums (lanes 0..2);
// workgroup 1: the epoch accumulators (lane 0).  Each chain reads its
// inputs through LDS in chunks the whole wave loads (the next chunk in
// flight while the current one is summed).  The chains cover the window's
// rows [chain_lo, chain_hi) and its whole epochs, resuming from their own
// outputs of the row / epoch before (a streamed frame extends them as its
// rows arrive).
constexpr int kChainChunk = 512;

__global__ __launch_bounds__(64) void k_producer_chain(JobSet set) {
  __shared__ double in_buffer[kChainChunk * 3];
  __shared__ double out_buffer[kChainChunk * 3];
  __shared__ float out_store[kChainChunk * 2];
  for_each_unit(set, [&](const JobDev& J, const int unit) {
    const int lane = threadIdx.x;
    constexpr int kPer = kChainChunk * 3 / 64;
    if (unit == 0) {
      const int H = J.win.chain_hi;
      const double* __restrict__ source = J.row_sums;
      const int first = J.win.chain_lo;
      double total = first > 0 && lane < 3 ? J.cumulative[(size_t)(first - 1) * 3 + lane] : 0.0;
      double next[kPer];
      auto load = [&](int base) {
        const int n = min(kChainChunk, H - base) * 3;
#pragma unroll
        for (int j = 0; j < kPer; ++j) {
          const int i = lane + j * 64;
          next[j] = i < n ? source[(size_t)base * 3 + i] : 0.0;
        }
      };
      if (first < H) load(first);
      for (int base = first; base < H; base += kChainChunk) {
        const int n = min(kChainChunk, H - base);
#pragma unroll
        for (int j = 0; j < kPer; ++j) in_buffer[lane + j * 64] = next[j];
        __syncthreads();
        if (base + kChainChunk < H) load(base + kChainChunk);
        if (lane < 3) {
#pragma unroll 16
          for (int r = 0; r < n; ++r) {
            total += in_buffer[r * 3 + lane];
            out_buffer[r * 3 + lane] = total;
          }
        }
        __syncthreads();
        for (int i = lane; i < n * 3; i += 64) J.cumulative[(size_t)base * 3 + i] = out_buffer[i];
        __syncthreads();
      }
    } else {
      const int epochs = J.win.chain_hi / 8;
      const int first = J.win.chain_lo / 8;
      const double* __restrict__ source = J.epoch_adds;
      float denominator = first > 0 ? J.epoch_store[(size_t)(first - 1) * 2 + 0] : 0.0f;
      float numerator = first > 0 ? J.epoch_store[(size_t)(first - 1) * 2 + 1] : 0.0f;
      for (int base = first; base < epochs; base += kChainChunk) {
        const int n = min(kChainChunk, epochs - base);
        for (int i = lane; i < n * 2; i += 64) in_buffer[i] = source[(size_t)base * 2 + i];
        __syncthreads();
        if (lane == 0) {
#pragma unroll 8
          for (int e = 0; e < n; ++e) {
            const double denominator_wide = (double)denominator + in_buffer[e * 2 + 0];
            const double numerator_wide = (double)numerator + in_buffer[e * 2 + 1];
            denominator = (float)denominator_wide;
            numerator = (float)numerator_wide;
            out_buffer[e * 2 + 0] = denominator_wide;
            out_buffer[e * 2 + 1] = numerator_wide;
            out_store[e * 2 + 0] = denominator;
            out_store[e * 2 + 1] = numerator;
          }
        }
        __syncthreads();
        for (int i = lane; i < n * 2; i += 64) {
          J.epoch_wide[(size_t)base * 2 + i] = out_buffer[i];
          J.epoch_store[(size_t)base * 2 + i] = out_store[i];
        }
        __syncthreads();
      }
    }
  });
}

namespace {
struct RowRecord {
  float mean_visible, mean_infrared, p1, p2, denominator, numerator;
};

// The record fields of row y after the chains (producer_parameters.py:
// 358-360, 437-460): means update whenever the cumulative weight is
// nonzero; p1/p2 update at every epoch end with a nonzero denominator.
__device__ __forceinline__ RowRecord row_record(const JobDev& J, int y) {
  RowRecord r;
  const double total_weight = J.cumulative[(size_t)y * 3 + 2];
  if (total_weight != 0.0) {
    r.mean_visible = (float)(J.cumulative[(size_t)y * 3 + 0] / total_weight);
    r.mean_infrared = (float)(J.cumulative[(size_t)y * 3 + 1] / total_weight);
  } else {
    r.mean_visible = 46618.78515625f;
    r.mean_infrared = 56239.65625f;
  }
  r.p1 = 0.15f;
  r.p2 = 0.15f;
  r.denominator = 0.0f;
  r.numerator = 0.0f;
  const int last_epoch = (y + 1) / 8 - 1;
  if (last_epoch >= 0) {
    r.denominator = J.epoch_store[last_epoch * 2 + 0];
    r.numerator = J.epoch_store[last_epoch * 2 + 1];
    const double denominator_wide = J.epoch_wide[last_epoch * 2 + 0];
    if (denominator_wide != 0.0) {
      const double ratio = J.epoch_wide[last_epoch * 2 + 1] / denominator_wide;
      r.p2 = (float)ratio;
      r.p1 = (float)((double)1.5f * ratio);
    }
  }
  return r;
}
}  // namespace

__global__ __launch_bounds__(kThreads) void k_schedule_rows(JobDev J,
                                                            ScheduleDev S) {
  const int y = blockIdx.x * blockDim.x + threadIdx.x;
  if (y >= J.height) return;
  RowRecord r = row_record(J, y);
  S.weighted_visible_sum[y] = J.cumulative[(size_t)y * 3 + 0];
  S.weighted_infrared_sum[y] = J.cumulative[(size_t)y * 3 + 1];
  S.weight_sum[y] = J.cumulative[(size_t)y * 3 + 2];
  S.mean_visible[y] = r.mean_visible;
  S.mean_infrared[y] = r.mean_infrared;
  S.accepted[y] = J.accepted[y];
  S.scaled_input[y] = r.p1;
  S.input_scale[y] = r.p2;
  S.scale_denominator[y] = r.denominator;
  S.scale_numerator[y] = r.numerator;
  const bool epoch_end = (y % 8) == 7;
  S.scale_add_denominator[y] = epoch_end ? (float)J.epoch_adds[(y / 8) * 2 + 0] : 0.0f;
  S.scale_add_numerator[y] = epoch_end ? (float)J.epoch_adds[(y / 8) * 2 + 1] : 0.0f;
}

// ContentDerivedStageParameterProvider (producer_parameters.py:498-541) +
// derive_x3a_setter_fields / StageCalibration (stage_parameters.py:93-197):
// stage s reads record row min(s-1, H-1).
__global__ __launch_bounds__(kThreads) void k_stage_table(JobSet set) {
  for_each_unit(set, [&](const JobDev& J, const int unit) {
    const float factor_b = J.factor_b;
    const int i = J.win.stage_lo + unit * blockDim.x + threadIdx.x;
    if (i >= J.win.stage_hi) return;
    const int row = i < J.height - 1 ? i : J.height - 1;
    RowRecord r = row_record(J, row);
    float p1 = r.p1, p2 = r.p2, p3 = r.mean_visible, p4 = r.mean_infrared;
    uint32_t error = 0;
    if (!(__builtin_isfinite(p1) && __builtin_isfinite(p2) &&
          __builtin_isfinite(p3) && __builtin_isfinite(p4)))
      error |= kErrStageNonfinite;
    if (p2 < 0.0f) {
      p1 = -p1;
      p2 = -p2;
    }
    const double denominator_p1 = 1.0 - (double)p1;
    const double denominator_p2 = 1.0 - (double)p2;
    if (denominator_p1 == 0.0 || denominator_p2 == 0.0) error |= kErrSetterScaleOne;
    const float base = (float)(((double)p4 - (double)p3 * (double)p2) / denominator_p2);
    const bool gate = !(p3 != 0.0f && p2 <= 0.075f);
    const float alpha = p2 * factor_b;
    if (!(__builtin_isfinite(alpha) && __builtin_isfinite(base))) error |= kErrStageNonfinite;
    if (error) atomicOr(J.error_flags, error);
    J.stages.alpha[i] = alpha;
    J.stages.offset[i] = 1.0f;
    J.stages.replacement[i] = base;
    J.stages.has_replacement[i] = 1;
    J.stages.score_base[i] = base;
    J.stages.coarse[i] = base;
    J.stages.gate_secondary[i] = gate ? 1 : 0;
    J.stages.row_gate[i] = 0;
  });
}

// ==========================================================================
// Analysis, in two tile kernels over the same TY x TX tiling:
//   k_scan   (all tiles, light): aux plane on the edge-clamped halo, the
//            centre scores, decision + eligibility, no-op emit of every
//            pixel, and a row-major list of the tile's sites.
//   k_sites  (tiles with sites, f64-heavy): the full history halo in LDS
//            (score, weighted aux, weighted RGB) and, per site, the W/A
//            records, candidates and combiner; each Site record goes
//            straight to its row-major position in the compacted array.
// ==========================================================================

// Row parameters of the halo rows: aux stage y+1 (index y), score stage y+2
// (index y+1), writer stage y+7 (index y+6) -- stage_parameters.py:207-232.
struct RowParams {
  float alpha, offset, replacement, score_base;
  uint8_t is_one;
};

__device__ __forceinline__ RowParams load_row_params(const JobDev& J, int sy) {
  RowParams p;
  p.alpha = J.stages.alpha[sy];
  p.is_one = p.alpha == 1.0f;
  p.offset = J.stages.offset[sy];
  p.replacement = J.stages.replacement[sy];
  p.score_base = J.stages.score_base[sy + 1];
  return p;
}

// k_scan runs 512-thread workgroups: each stages the 32 KB no-op table once
// for 8 waves.  The register target is the occupancy the LDS allows (four
// workgroups, 8 waves per SIMD at 64 registers, at the 16x64 tile).
constexpr int kScanThreads = 512;
template <int TY, int TX>
struct ScanShared {
  static constexpr int AR = TY + 8, AC = TX + 8;  // aux, edge-clamped, +-4
  uint32_t noop_delta[8192];  // the no-op emit table as 4-bit deltas (32 KB)
  float aux[AR][AC];
  // aux < decision threshold: row-major over the halo (bit r * AC + c), and
  // one word per halo column (bit r; 32 bits while the halo has 32 rows)
  static_assert(AR <= 64, "one 64-bit word per halo column");
  using ColumnWord = typename std::conditional<(AR <= 32), uint32_t, unsigned long long>::type;
  unsigned long long under_flat[(AR * AC + 63) / 64];
  ColumnWord under_column[AC];
  RowParams rows[AR];
  uint8_t writer_floor[TY];
  int32_t writer_gate[TY];
  uint32_t row_count[TY];
  // eligible centres per wave in each slab of kScanThreads centres
  uint32_t slab_wave[TY * TX / kScanThreads][kScanThreads / 64];
};

// Tiles per work unit of k_scan: a vertical strip of kScanStrip tiles, each
// after the first keeping the 8 halo rows the tile above staged.  Strips
// cover the first three quarters of a job's tile rows; the rest go one tile
// per unit, so that the units dealt last are short.
constexpr int kScanStrip = 4;
__host__ __device__ inline int scan_strips(int trow_n) { return (trow_n * 3 / 4) / kScanStrip; }
__host__ __device__ inline int scan_units(int trow_n, int tiles_x) {
  const int strips = scan_strips(trow_n);
  return (strips + trow_n - strips * kScanStrip) * tiles_x;
}

template <int TY, int TX>
constexpr int scan_waves() {
  const int blocks = (int)((160 * 1024) / sizeof(ScanShared<TY, TX>));
  const int waves = blocks * (kScanThreads / 64) / 4;
  return waves < 8 ? waves : 8;
}

template <int TY, int TX>
__global__ __launch_bounds__(kScanThreads, (scan_waves<TY, TX>())) void k_scan(JobSet set) {
  using Sh = ScanShared<TY, TX>;
  constexpr int AR = Sh::AR, AC = Sh::AC;
  static_assert((TY * TX) % kScanThreads == 0, "tile must be a multiple of the block");
  __shared__ Sh S;
  bool table_staged = false;  // uniform: every job of a device shares the tables
  // One tile; `staged`: the top AR - TY halo rows are those the tile above
  // left in LDS.
  auto scan_tile = [&](const JobDev& J, const int tile_y, const int tile_x, const bool staged) {
    const DevParams& P = J.P;
    const int t = tile_y * J.tiles_x + tile_x;
    const int H = J.height, W = J.width;
    const int y0 = tile_y * TY, x0 = tile_x * TX;
    const int tid = threadIdx.x;
    const float floor = P.score_floor;

    if (!table_staged) {  // once per workgroup; the barrier below orders it
      const uint4* table = reinterpret_cast<const uint4*>(J.noop_delta);
      for (int i = tid; i < 8192 / 4; i += kScanThreads)
        reinterpret_cast<uint4*>(S.noop_delta)[i] = table[i];
      table_staged = true;
    }
    for (int r = tid; r < AR; r += kScanThreads) S.rows[r] = load_row_params(J, clampi(y0 - 4 + r, 0, H - 1));
    for (int i = tid; i < TY; i += kScanThreads) {
      const int y = min(y0 + i, H - 1);
      S.writer_floor[i] = (P.gate_primary || J.stages.gate_secondary[y + 6]) ? 1 : 0;
      S.writer_gate[i] = J.stages.row_gate[y + 6];
      S.row_count[i] = 0;
    }
    // Below a staged tile, halo rows [TY, AR) become rows [0, AR - TY) (the
    // ranges are disjoint: AR - TY = 8 <= TY), in the aux plane and in both
    // bit layouts (TY * AC is a multiple of 64).
    static_assert(AR - TY <= TY && (TY * AC) % 64 == 0, "halo rows shift by whole words");
    for (int i = tid; i < AC; i += kScanThreads)
      S.under_column[i] = staged ? S.under_column[i] >> TY : 0;
    if (staged) {
      for (int i = tid; i < (AR - TY) * AC; i += kScanThreads)
        (&S.aux[0][0])[i] = (&S.aux[0][0])[i + TY * AC];
      for (int i = tid; i < (AR - TY) * AC / 64; i += kScanThreads)
        S.under_flat[i] = S.under_flat[i + TY * AC / 64];
    }
    __syncthreads();

    // x3a.py:193-221 on the clamped halo; no-op emit (output.py:76-94) of the
    // centres.
    uint32_t aux_error = 0;
    auto noop = [&](uint32_t raw) {
      const uint32_t delta = (S.noop_delta[raw >> 3] >> ((raw & 7u) * 4u)) & 15u;
      return (uint16_t)(raw + delta - (uint32_t)kNoopDeltaBias);
    };
    // Below a staged tile only halo rows [AR - TY, AR) are loaded; the tile
    // rows above them (halo rows 4 .. AR - TY - 1) still need their output.
    if (staged) {
      for (int i = tid; i < (AR - TY - 4) * TX; i += kScanThreads) {
        const int ly = y0 + i / TX, lx = x0 + i % TX;
        if (ly < H && lx < W) {
          uint32_t lanes[4];
          load_pixel(J.rgbi, (size_t)ly * W + lx, lanes);
          uint16_t* o = J.out + ((size_t)ly * W + lx) * 3;
          o[0] = noop(lanes[0]);
          o[1] = noop(lanes[1]);
          o[2] = noop(lanes[2]);
        }
      }
    }
    for (int i = tid + (staged ? (AR - TY) * AC : 0); i < AR * AC; i += kScanThreads) {
      const int r = i / AC, c = i % AC;
      const int ly = y0 - 4 + r, lx = x0 - 4 + c;
      uint32_t lanes[4];
      load_pixel(J.rgbi, (size_t)clampi(ly, 0, H - 1) * W + clampi(lx, 0, W - 1), lanes);
      const uint32_t raw_visible =
          P.visible_channel == 0 ? lanes[0] : (P.visible_channel == 1 ? lanes[1] : lanes[2]);
      const float visible = J.lut[raw_visible], infrared = J.lut[lanes[3]];
      const RowParams& rp = S.rows[r];
      const float a = rp.is_one ? rp.replacement : auxiliary(visible, infrared, rp.alpha, rp.offset);
      if (!__builtin_isfinite(a)) aux_error = kErrNonfiniteAux;
      S.aux[r][c] = a;
      // the decision bars' samples as bits: a wave holds 64 consecutive
      // cells (i of its lane 0 is a multiple of 64)
      const bool under = a < P.decision_threshold;
      const unsigned long long under_wave = __ballot(under);
      if ((tid & 63) == 0) S.under_flat[i >> 6] = under_wave;
      if (under) atomicOr(&S.under_column[c], (typename Sh::ColumnWord)1 << r);
      const int ty = r - 4, tx = c - 4;
      if (ty >= 0 && ty < TY && tx >= 0 && tx < TX && ly < H && lx < W) {
        // emit(LUT[raw]) is a function of the raw sample alone
        uint16_t* o = J.out + ((size_t)ly * W + lx) * 3;
        o[0] = noop(lanes[0]);
        o[1] = noop(lanes[1]);
        o[2] = noop(lanes[2]);
      }
    }
    if (aux_error) atomicOr(J.error_flags, aux_error);
    __syncthreads();

    // Centre score (x3a.py:245-286), decision bars (streaming.py:157-198) and
    // eligibility (streaming.py:461-466).  Centre k * kScanThreads + tid: the 64
    // lanes of a wave read 64 consecutive LDS words (no bank conflicts) and
    // store consecutive diagnostics.
    constexpr int kPer = TY * TX / kScanThreads;
    uint32_t eligible_bits = 0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const int li = k * kScanThreads + tid;
      const int ty = li / TX, tx = li % TX;
      const int y = y0 + ty, x = x0 + tx;
      if (y >= H || x >= W) continue;
      const int ar = ty + 4, ac = tx + 4;
      float sample = S.aux[ar][ac];
      if (P.horizontal_minimum && W > 2 && x >= 1 && x <= W - 2)
        sample = fmin_plain(fmin_plain(S.aux[ar][ac - 1], S.aux[ar][ac]), S.aux[ar][ac + 1]);
      const float sc = score_value(sample, S.rows[ar].score_base, P.score_addend,
                                   P.score_scale, P.score_offset, floor);
      if (J.diag_score) {
        J.diag_score[(size_t)y * W + x] = sc;
        J.diag_floor[(size_t)y * W + x] = sc == floor ? 1 : 0;
      }
      if (S.writer_gate[ty] != 0) continue;
      if (S.writer_floor[ty] && sc >= 1.0f) continue;
      // decision_fallback's four nine-sample bars, counted from the bits
      const int R = P.radius;
      auto row_bar = [&](int r) {  // columns ac-4 .. ac+4 of halo row r
        const int f = r * AC + ac - 4, w = f >> 6, o = f & 63;
        unsigned long long bits = S.under_flat[w] >> o;
        if (o > 55) bits |= S.under_flat[w + 1] << (64 - o);
        return __popcll(bits & 0x1FFull);
      };
      auto column_bar = [&](int c) {  // rows ar-4 .. ar+4 of halo column c
        return __popcll(((unsigned long long)S.under_column[c] >> (ar - 4)) & 0x1FFull);
      };
      if (column_bar(ac - R) > P.count_limit || row_bar(ar - R) > P.count_limit ||
          column_bar(ac + R) > P.count_limit || row_bar(ar + R) > P.count_limit)
        continue;
      eligible_bits |= 1u << k;
    }
    // The queue in row-major order: slab by slab, and within a slab wave by
    // wave and lane by lane (counts by ballot, one barrier for all slabs).
    const int lane = tid & 63, wave = tid >> 6;
    const unsigned long long below = (1ull << lane) - 1ull;
    uint32_t rank[kPer];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      const unsigned long long mask = __ballot((eligible_bits >> k) & 1u);
      rank[k] = (uint32_t)__popcll(mask & below);
      if (lane == 0) S.slab_wave[k][wave] = (uint32_t)__popcll(mask);
    }
    __syncthreads();
    uint16_t* list = J.site_lists + (size_t)t * (TY * TX);
    uint32_t total = 0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
      uint32_t before = 0, slab = 0;
#pragma unroll
      for (int w = 0; w < kScanThreads / 64; ++w) {
        const uint32_t c = S.slab_wave[k][w];
        before += w < wave ? c : 0u;
        slab += c;
      }
      if ((eligible_bits >> k) & 1u) {
        const int li = k * kScanThreads + tid;
        list[total + before + rank[k]] = (uint16_t)li;
        atomicAdd(&S.row_count[li / TX], 1u);
      }
      total += slab;
    }
    __syncthreads();
    for (int i = tid; i < TY; i += kScanThreads)
      J.seg_count[(size_t)(y0 + i) * J.tiles_x + tile_x] = S.row_count[i];
    if (tid == 0) J.tile_count[t] = total;
  };
  for_each_unit(set, [&](const JobDev& J, const int unit) {
    const int strips = scan_strips(J.win.trow_n);
    const int strip_units = strips * J.tiles_x;
    int tile_x, trow_first, trow_end;
    if (unit < strip_units) {
      tile_x = unit % J.tiles_x;
      trow_first = J.win.trow_lo + (unit / J.tiles_x) * kScanStrip;
      trow_end = trow_first + kScanStrip;
    } else {
      const int u = unit - strip_units;
      tile_x = u % J.tiles_x;
      trow_first = J.win.trow_lo + strips * kScanStrip + u / J.tiles_x;
      trow_end = trow_first + 1;
    }
    for (int tile_y = trow_first; tile_y < trow_end; ++tile_y) {
      scan_tile(J, tile_y, tile_x, tile_y > trow_first);
      if (tile_y + 1 < trow_end) __syncthreads();  // the next tile reuses the LDS
    }
  });
}

template <int TY, int TX, int Threads>
struct SitesShared {
  static constexpr int AR = TY + 10, AC = TX + 12;  // aux, edge-clamped, +-5/+-6
  static constexpr int HR = TY + 10, HC = TX + 10;  // score / waux history, +-5
  static constexpr int RR = TY + 8, RC = TX + 8;    // weighted RGB history, +-4
  // feature-record positions: the tile and a one-pixel ring (the cross
  // neighbours of its sites)
  static constexpr int PR = TY + 2, PC = TX + 2, NP = PR * PC;
  // record runs: kRun adjacent table columns of one table row (5: every run
  // pair of a 16x64 tile fits one round of 512 lanes)
  static constexpr int kRun = 5, kRowRuns = (PC + kRun - 1) / kRun, kRuns = PR * kRowRuns;
  float aux[AR][AC];
  float score[HR][HC];
  // the two planes sit 16 banks apart (of 32), so that the score and
  // weighted-aux lanes of the feature records read disjoint banks
  float pad_planes[((16 - (HR * HC) % 32) + 32) % 32];
  float waux[HR][HC];
  float wrgb[3][RR][RC];
  float record_w[3][NP];      // f32 of the unrounded score averages
  float record_ratio[3][NP];  // f32(waux average / unrounded score average)
  uint8_t record_valid[NP];   // bit l: score average l > 0 (else the fallback)
  uint32_t needed_bits[(NP + 31) / 32];
  // needed positions (at most Threads / 2 of them) or record runs
  uint16_t work_list[kRuns > Threads / 2 ? kRuns : Threads / 2];
  uint32_t run_totals;  // runs (low 16 bits) and needed positions (high 16 bits)
  RowParams rows[AR];
  float writer_coarse[TY];
  uint8_t writer_floor[TY];
  uint32_t row_start[TY];
  uint32_t row_offset[TY];
  uint32_t wave_sums[Threads / 64];
};

// Reconstruction sites of one tile.  A feature record (streaming.py:236-308)
// is a function of its position alone, up to the fallback that replaces an
// empty average (the centre row's writer coarse reference); so every
// position a site of the tile needs -- the site itself and, in cross-
// neighbour mode, its four neighbours -- gets its averages once, in a table
// of raw ratios and validity bits, and each site applies its own fallback
// when it reads them.  Adjacent sites share most of their records.
// Workgroups per CU the LDS of a tile shape allows (160 KB per CU).
template <int TY, int TX, int Threads>
constexpr int sites_occupancy() {
  return sizeof(SitesShared<TY, TX, Threads>) * 3 <= 160 * 1024   ? 3
         : sizeof(SitesShared<TY, TX, Threads>) * 2 <= 160 * 1024 ? 2
                                                                  : 1;
}

// The wide workgroup of a tile shape: as many threads as give the CU 16
// waves (4 per SIMD) at the workgroups its LDS allows -- 1024 threads where
// one tile fills the LDS, else 512 (two workgroups; where the LDS would
// allow three, the 128-register budget of 4 waves per SIMD admits two).
template <int TY, int TX>
constexpr int sites_wide_threads() {
  return sites_occupancy<TY, TX, 512>() == 1 ? 1024 : 512;
}

// The reconstruction sites of a tile with at most Threads / 3 of them, three
// lanes per site: lane k takes channel k.  The candidates
// (reconstruction.py:425-462) and the combiner (:672-887) are per channel
// once the site's weights and feature records are known, so each lane runs
// a third of the site's chain of binary64 operations and stores its
// channel's fields of the site.  Lanes [0, total) take channel 0 of the
// listed sites in order, the next `total` channel 1, the last channel 2, so
// a wave reads one plane of the weighted RGB at a time.  Eligibility
// already implies the driver gate passes (reconstruction.py:890-904 re-tests
// the same row gate and the same point score), so every listed centre is
// attempted.  The same records and averages as the one-lane loop of
// k_sites, whose comments say where each comes from.
template <int TY, int TX, int Threads>
__device__ __forceinline__ void sites_channel(SitesShared<TY, TX, Threads>& S, const JobDev& J,
                                              const uint16_t* list, int y0, int x0,
                                              uint32_t total, uint32_t u) {
  using Sh = SitesShared<TY, TX, Threads>;
  constexpr int PC = Sh::PC;
  const DevParams& P = J.P;
  const int W = J.width;
  const uint64_t* rgbi = reinterpret_cast<const uint64_t*>(J.rgbi);
  const float* lut = J.lut;
  const uint32_t k = u / total, q = u - k * total;
  const int li = list[q];
  const int ty = li / TX, tx = li % TX;
  const int y = y0 + ty, x = x0 + tx;
  const float fallback = S.writer_coarse[ty];
  const uint64_t pixel = rgbi[(size_t)y * W + x];  // R, G, B, IR (16 bits each)
  const float original = lut[(uint32_t)(pixel >> (16 * k)) & 0xFFFFu];
  const int cr = ty + 5, cc = tx + 5, ar = ty + 5, ac = tx + 6;
  const int centre = (ty + 1) * PC + tx + 1;
  auto feature = [&](int i, float point_aux, float point_score, float f[4]) {
    const uint8_t valid = S.record_valid[i];
#pragma unroll
    for (int l = 0; l < 3; ++l) f[l] = (valid >> l) & 1u ? S.record_ratio[l][i] : fallback;
    f[3] = point_score > 0.0f ? point_aux : fallback;
  };
  float weights[4], features[5][4];
  weights[0] = S.record_w[0][centre];
  weights[1] = S.record_w[1][centre];
  weights[2] = S.record_w[2][centre];
  weights[3] = S.score[cr][cc];
  feature(centre, S.aux[ar][ac], S.score[cr][cc], features[0]);
  int record_count = 1;
  if (P.cross_neighbor) {
    record_count = 5;
    if (x > 0) {
      feature(centre - 1, S.aux[ar][ac - 1], S.score[cr][cc - 1], features[1]);
    } else {
      features[1][0] = features[1][1] = features[1][2] = features[1][3] = 0.0f;
    }
    if (x < W - 1) {
      feature(centre + 1, S.aux[ar][ac + 1], S.score[cr][cc + 1], features[2]);
    } else {
      features[2][0] = features[2][1] = features[2][2] = features[2][3] = 0.0f;
    }
    feature(centre - PC, S.aux[ar - 1][ac], S.score[cr - 1][cc], features[3]);
    feature(centre + PC, S.aux[ar + 1][ac], S.score[cr + 1][cc], features[4]);
  }
  auto lane = [&](int dy, int dx) { return S.wrgb[k][ty + dy][tx + dx]; };
  double averages[3];
  unscaled_averages(lane, averages);
  averages[2] = rgb_q16(lane);
  float candidates[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    if (weights[c] > 0.0f) {
      double reciprocal = 1.0 / (double)weights[c];
      candidates[c] = (float)(averages[c] * reciprocal);
    } else {
      candidates[c] = (float)averages[c];
    }
  }
  CombineShared shared;
  combine_shared(weights, features, record_count, fallback, P.combine, shared);
  Site* site = J.sites + S.row_offset[ty] + (q - S.row_start[ty]);
  site->cand[k] = combine_channel((int)k, candidates[0], candidates[1], candidates[2], original,
                                  shared, P.combine);
  site->orig[k] = original;
  if (k == 0)
    site->pixel = (uint32_t)((size_t)y * W + x) | (S.writer_floor[ty] ? kSiteFloorBit : 0u);
}

// Tiles per work unit of k_sites: a workgroup takes a vertical strip of
// this many tiles, and each tile after the first keeps the bottom rows of
// the halo the tile above staged (10 of the 26 aux and history rows, 8 of
// the 24 weighted-RGB rows at 16-row tiles) instead of loading and
// recomputing them.  Tiles shorter than the 10-row halo overlap do not
// strip.  Strips cover the first three quarters of a job's tile rows; the
// rest go one tile per unit, so that the units handed out last are short.
__host__ __device__ constexpr int sites_strip(int tile_rows) { return tile_rows >= 10 ? 4 : 1; }
__host__ __device__ inline int sites_strips(int tile_rows, int trow_n) {
  const int k = sites_strip(tile_rows);
  return k > 1 ? (trow_n * 3 / 4) / k : 0;
}
__host__ __device__ inline int sites_units(int tile_rows, int trow_n, int tiles_x) {
  const int strips = sites_strips(tile_rows, trow_n);
  return (strips + trow_n - strips * sites_strip(tile_rows)) * tiles_x;
}

// `Threads` per workgroup and `Waves` per SIMD, the occupancy the register
// allocation targets.  Variant 0 is the wide workgroup at 4 waves per SIMD:
// every phase of a tile takes fewer rounds of the workgroup, and the other
// tile on the CU runs while one waits.  Variant 1 is 256 threads at the
// LDS occupancy of the shape (up to 3 waves per SIMD, more registers per
// wave), for the tuner to weigh.
template <int TY, int TX, int Threads, int Waves>
__global__ __launch_bounds__(Threads, Waves) void k_sites(JobSet set) {
  using Sh = SitesShared<TY, TX, Threads>;
  constexpr int AR = Sh::AR, AC = Sh::AC, HR = Sh::HR, HC = Sh::HC;
  constexpr int RR = Sh::RR, RC = Sh::RC, PC = Sh::PC, NP = Sh::NP;
  constexpr int kWords = (NP + 31) / 32;
  // tiles of at most kChannelRounds * Threads / 3 sites take three lanes
  // per site
  constexpr uint32_t kChannelRounds = 1;
  __shared__ Sh S;
  // One tile with sites; `staged`: rows of its halo are those the tile
  // above left in LDS.
  auto sites_tile = [&](const JobDev& J, const int t, const int tile_y, const int tile_x,
                        const uint32_t total, const bool staged) {
    const DevParams& P = J.P;
    const int H = J.height, W = J.width;
    const int y0 = tile_y * TY, x0 = tile_x * TX;
    // threadIdx.x made opaque once per tile, so that the LDS addresses
    // derived from it are recomputed in each tile instead of being held
    // across the persistent loop (at 128 registers they were spilled)
    int tid = threadIdx.x;
    asm volatile("" : "+v"(tid));
    const float floor = P.score_floor;
    const uint16_t* list = J.site_lists + (size_t)t * (TY * TX);

    // The tile's tables in registers first: its halo rows' parameters (one
    // row per thread) and, per tile row (one lane each, in the first wave),
    // the row's site count, its first place in the site array and its
    // writer stage.
    static_assert(AR <= Threads, "one halo row per thread");
    static_assert(TY <= 64, "one lane per tile row");
    RowParams row_params{};
    if (tid < AR) row_params = load_row_params(J, clampi(y0 - 5 + tid, 0, H - 1));
    uint32_t head_count = 0, head_offset = 0;
    float head_coarse = 0.0f;
    bool head_floor = false;
    if (tid < TY) {
      const size_t seg = (size_t)(y0 + tid) * J.tiles_x + tile_x;
      const int y = min(y0 + tid, H - 1);
      head_count = J.seg_count[seg];
      head_offset = J.seg_offset[seg];
      head_coarse = J.stages.coarse[y + 6];
      head_floor = P.gate_primary || J.stages.gate_secondary[y + 6];
    }
    // The tile's site list entries this thread marks.
    constexpr int kListRounds = (TY * TX + Threads - 1) / Threads;
    uint32_t listed[kListRounds];
#pragma unroll
    for (int u = 0; u < kListRounds; ++u) {
      const uint32_t q = (uint32_t)tid + (uint32_t)u * Threads;
      listed[u] = q < total ? (uint32_t)list[q] : 0u;
    }
    // Then the raw halo pixels this thread stages: requested after the
    // tables and the list, so that waiting for those does not wait for
    // them, and in flight through the tables' stores, the strip shift and
    // the site marks.  Below a staged tile, from aux row AR - TY - 1 on: the first
    // row whose working RGB the weighted-RGB rows still need (its aux,
    // rewritten, is the value already there).
    constexpr int kStageRounds = (AR * AC + Threads - 1) / Threads;
    const int stage_from = staged ? (AR - TY - 1) * AC : 0;
    uint2 halo_raw[kStageRounds];
#pragma unroll
    for (int u = 0; u < kStageRounds; ++u) {
      const int i = stage_from + tid + u * Threads;
      halo_raw[u] = make_uint2(0u, 0u);
      if (i < AR * AC) {
        const int r = i / AC, c = i % AC;
        const size_t pixel = (size_t)clampi(y0 - 5 + r, 0, H - 1) * W + clampi(x0 - 6 + c, 0, W - 1);
        halo_raw[u] = reinterpret_cast<const uint2*>(J.rgbi)[pixel];
      }
    }

    if (tid < AR) S.rows[tid] = row_params;
    // the rows' first entries in the site list: an exclusive scan of the
    // row counts
    if (tid < 64) {
      const int i = tid;
      uint32_t inclusive = head_count;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(inclusive, d, 64);
        if (i >= d) inclusive += t;
      }
      if (i < TY) {
        S.row_start[i] = inclusive - head_count;
        S.row_offset[i] = head_offset;
        S.writer_coarse[i] = head_coarse;
        S.writer_floor[i] = head_floor ? 1 : 0;
      }
    }
    // The record positions the sites need (record index = (ty+1)*PC + tx+1),
    // into bits the previous tile's site phase cleared.  Neighbours outside
    // the frame's columns are all-zero records, not averages
    // (streaming.py:259-308).
#pragma unroll
    for (int u = 0; u < kListRounds; ++u) {
      const uint32_t q = (uint32_t)tid + (uint32_t)u * Threads;
      if (q >= total) continue;
      const int li = (int)listed[u];
      const int ty = li / TX, tx = li % TX;
      const int x = x0 + tx;
      const int centre = (ty + 1) * PC + tx + 1;
      auto mark = [&](int i) { atomicOr(&S.needed_bits[i >> 5], 1u << (i & 31)); };
      mark(centre);
      if (P.cross_neighbor) {
        if (x > 0) mark(centre - 1);
        if (x < W - 1) mark(centre + 1);
        mark(centre - PC);
        mark(centre + PC);
      }
    }
    // A tile below a staged one takes its top halo rows from the bottom rows
    // the tile above left (rows [TY, R) -> [0, R - TY); the two ranges are
    // disjoint, so no barrier separates the reads from the writes).
    static_assert(sites_strip(TY) == 1 || (AR - TY <= TY && HR - TY <= TY && RR - TY <= TY),
                  "striped tiles must be at least as tall as their halo overlap");
    if (sites_strip(TY) > 1 && staged) {
      auto shift_up = [&](float* plane, int rows, int cols) {
        for (int i = tid; i < (rows - TY) * cols; i += Threads) plane[i] = plane[i + TY * cols];
      };
      shift_up(&S.aux[0][0], AR, AC);
      shift_up(&S.score[0][0], HR, HC);
      shift_up(&S.waux[0][0], HR, HC);
      shift_up(&S.wrgb[0][0][0], RR, RC);
      shift_up(&S.wrgb[1][0][0], RR, RC);
      shift_up(&S.wrgb[2][0][0], RR, RC);
    }
    __syncthreads();

    // Working lanes and aux on the clamped halo; working RGB staged for the
    // weighted-RGB history.  Every round's table gathers are requested
    // before any is used.
    float halo_work[kStageRounds][4];
#pragma unroll
    for (int u = 0; u < kStageRounds; ++u) {
      const uint2 q = halo_raw[u];
      halo_work[u][0] = J.lut[q.x & 0xFFFFu];
      halo_work[u][1] = J.lut[q.x >> 16];
      halo_work[u][2] = J.lut[q.y & 0xFFFFu];
      halo_work[u][3] = J.lut[q.y >> 16];
    }
#pragma unroll
    for (int u = 0; u < kStageRounds; ++u) {
      const int i = stage_from + tid + u * Threads;
      if (i >= AR * AC) continue;
      const int r = i / AC, c = i % AC;
      const float work0 = halo_work[u][0], work1 = halo_work[u][1];
      const float work2 = halo_work[u][2], work3 = halo_work[u][3];
      const float visible =
          P.visible_channel == 0 ? work0 : (P.visible_channel == 1 ? work1 : work2);
      const RowParams& rp = S.rows[r];
      S.aux[r][c] = rp.is_one ? rp.replacement : auxiliary(visible, work3, rp.alpha, rp.offset);
      const int rr = r - 1, rc = c - 2;
      if (rr >= 0 && rr < RR && rc >= 0 && rc < RC) {
        S.wrgb[0][rr][rc] = work0;
        S.wrgb[1][rr][rc] = work1;
        S.wrgb[2][rr][rc] = work2;
      }
    }
    __syncthreads();

    // Score / weighted-aux history with the asymmetric guards
    // (streaming.py:201-233): zero outside the row and above row -1, the
    // floor pseudo-row at -1, the last row repeated below; each cell also
    // weights its pixel's working RGB.  Below a staged tile, from history
    // row HR - TY - 1 on: that row's scores weight the first restaged
    // weighted-RGB row (recomputed, they equal the values already there).
    for (int i = tid + (staged ? (HR - TY - 1) * HC : 0); i < HR * HC; i += Threads) {
      const int r = i / HC, c = i % HC;
      const int ly = y0 - 5 + r, lx = x0 - 5 + c;
      float s, wa;
      if (lx < 0 || lx >= W || ly < -1) {
        s = 0.0f;
        wa = 0.0f;
      } else if (ly == -1) {
        s = floor;
        wa = S.aux[r][c + 1] * floor;
      } else {
        const float a = S.aux[r][c + 1];
        float sample = a;
        if (P.horizontal_minimum && W > 2 && lx >= 1 && lx <= W - 2)
          sample = fmin_plain(fmin_plain(S.aux[r][c], a), S.aux[r][c + 2]);
        s = score_value(sample, S.rows[r].score_base, P.score_addend, P.score_scale,
                        P.score_offset, floor);
        wa = s * a;
      }
      S.score[r][c] = s;
      S.waux[r][c] = wa;
      // the weighted-RGB cell of the same pixel (history row r is weighted-
      // RGB row r - 1, column c is column c - 1), zero outside the frame
      const int rr = r - 1, rc = c - 1;
      if (rr >= 0 && rr < RR && rc >= 0 && rc < RC) {
        if (lx < 0 || lx >= W || ly < -1) {
          S.wrgb[0][rr][rc] = 0.0f;
          S.wrgb[1][rr][rc] = 0.0f;
          S.wrgb[2][rr][rc] = 0.0f;
        } else {
          S.wrgb[0][rr][rc] = s * S.wrgb[0][rr][rc];
          S.wrgb[1][rr][rc] = s * S.wrgb[1][rr][rc];
          S.wrgb[2][rr][rc] = s * S.wrgb[2][rr][rc];
        }
      }
    }
    // The needed positions by runs: a run is an aligned group of kRun
    // adjacent table columns of one table row.
    constexpr int kRun = Sh::kRun, kRowRuns = Sh::kRowRuns, kRuns = Sh::kRuns;
    // the needed bits of run (row r, run b), bit j = table column b * kRun + j
    auto run_bits = [&](int r, int b) -> uint32_t {
      const int f = r * PC + b * kRun;
      const int w = f >> 5, o = f & 31;
      uint32_t bits = S.needed_bits[w] >> o;
      if (o > 32 - kRun && w + 1 < kWords) bits |= S.needed_bits[w + 1] << (32 - o);
      const int n = min(kRun, PC - b * kRun);
      return bits & ((1u << n) - 1u);
    };
    // The first wave lists them, after its share of the pass above, with
    // one wave-level scan that counts both: runs in the low 16 bits,
    // positions in the high 16 bits.  Where the positions fit one lane pair
    // each in a single round, the list holds positions (a run's lanes would
    // walk longer); otherwise it holds runs.
    if (tid < 64) {
      constexpr int kPerLane = (kRuns + 63) / 64;
      const int first = tid * kPerLane;
      uint32_t count = 0;
#pragma unroll
      for (int k = 0; k < kPerLane; ++k) {
        const int e = first + k;
        if (e < kRuns) {
          const uint32_t bits = run_bits(e / kRowRuns, e % kRowRuns);
          if (bits != 0u) count += 1u + ((uint32_t)__builtin_popcount(bits) << 16);
        }
      }
      uint32_t inclusive = count;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(inclusive, d, 64);
        if (tid >= d) inclusive += t;
      }
      const uint32_t totals = __shfl(inclusive, 63, 64);
      const uint32_t at = inclusive - count;
      const bool by_position = 2 * (totals >> 16) <= (uint32_t)Threads;
      uint32_t slot = by_position ? at >> 16 : at & 0xFFFFu;
#pragma unroll
      for (int k = 0; k < kPerLane; ++k) {
        const int e = first + k;
        if (e >= kRuns) continue;
        const int r = e / kRowRuns, b = e % kRowRuns;
        uint32_t bits = run_bits(r, b);
        if (bits == 0u) continue;
        if (!by_position) {
          S.work_list[slot++] = (uint16_t)e;
          continue;
        }
        while (bits != 0u) {
          const int j = __builtin_ctz(bits);
          bits &= bits - 1u;
          S.work_list[slot++] = (uint16_t)(r * PC + b * kRun + j);
        }
      }
      if (tid == 0) S.run_totals = totals;
    }
    __syncthreads();
    const uint32_t runs = S.run_totals & 0xFFFFu, needed = S.run_totals >> 16;

    // The averages of every needed position (streaming.py:236-256; the A
    // lanes divide by the unrounded binary64 W, reconstruction.py:365-371).
    // Position (pr, pc) has its patch at history rows pr..pr+8, columns
    // pc..pc+8.  Two lanes per position or run: the even lane averages the
    // score plane, the odd lane the weighted-aux plane, and the pair swaps
    // its sums; the even lane stores the f32 weights and the first ratio,
    // the odd lane the other two.
    auto column_at = [&](const float* c) {
      return column_scratch(c[0], c[HC], c[2 * HC], c[3 * HC], c[4 * HC], c[5 * HC], c[6 * HC],
                            c[7 * HC], c[8 * HC]);
    };
    auto store_record = [&](int i, bool odd, const double mine[3], const double other[3]) {
      const double* unrounded = odd ? other : mine;
      const double* numerator = odd ? mine : other;
      auto ratio = [&](int l) {
        S.record_ratio[l][i] = unrounded[l] > 0.0 ? (float)(numerator[l] / unrounded[l]) : 0.0f;
      };
      if (!odd) {
        uint8_t valid = 0;
#pragma unroll
        for (int l = 0; l < 3; ++l) {
          S.record_w[l][i] = (float)unrounded[l];
          valid |= unrounded[l] > 0.0 ? (uint8_t)(1u << l) : (uint8_t)0;
        }
        S.record_valid[i] = valid;
        ratio(0);
      } else {
        ratio(1);
        ratio(2);
      }
    };
    if (2 * needed <= (uint32_t)Threads) {
      // one lane pair per position, in one round
      const uint32_t k = tid >> 1;
      const bool odd = tid & 1;
      if (k < needed) {
        const int i = S.work_list[k];
        const float* base = (odd ? &S.waux[0][0] : &S.score[0][0]) + (i / PC) * HC + i % PC;
        Column col[9];
#pragma unroll
        for (int x = 0; x < 9; ++x) col[x] = column_at(base + x);
        double mine[3], other[3];
        combine_columns(col, mine);
#pragma unroll
        for (int l = 0; l < 3; ++l) other[l] = swap_pair(mine[l]);
        store_record(i, odd, mine, other);
      }
    } else {
      // One lane pair per run: a column scratch (reconstruction.py:187-237)
      // is a function of its nine samples alone, so the positions of a run
      // share theirs.  The pair walks the run from its first needed position
      // to its last with nine column scratches in registers, each step
      // adding one column, and combines the nine for every needed position
      // (reconstruction.py:240-268).
      for (uint32_t k = tid; k < 2 * runs; k += Threads) {
        const bool odd = k & 1u;
        const int e = S.work_list[k >> 1];
        const int r = e / kRowRuns, b = e % kRowRuns;
        uint32_t bits = run_bits(r, b);
        const int lead = __builtin_ctz(bits);
        bits >>= lead;
        const int span = 32 - __builtin_clz(bits);  // 1 .. kRun
        const int i0 = r * PC + b * kRun + lead;      // record index of the first
        const float* base = (odd ? &S.waux[0][0] : &S.score[0][0]) + r * HC + b * kRun + lead;
        Column col[9];
#pragma unroll
        for (int x = 0; x < 9; ++x) col[x] = column_at(base + x);
#pragma unroll
        for (int j = 0; j < kRun; ++j) {
          if (j >= span) break;  // the pair walks the same run
          if (j > 0) {
#pragma unroll
            for (int x = 0; x < 8; ++x) col[x] = col[x + 1];
            col[8] = column_at(base + j + 8);
          }
          double mine[3], other[3];
          combine_columns(col, mine);
#pragma unroll
          for (int l = 0; l < 3; ++l) other[l] = swap_pair(mine[l]);
          if ((bits >> j) & 1u) store_record(i0 + j, odd, mine, other);
        }
      }
    }
    __syncthreads();

    // The needed bits are read for the last time above: clear them for the
    // next tile's marks (a barrier follows every tile).
    for (int i = tid; i < kWords; i += Threads) S.needed_bits[i] = 0u;

    if (3 * total <= (uint32_t)(kChannelRounds * Threads)) {
      // Few sites: three lanes per site, lane k taking channel k
      // (sites_channel), in at most kChannelRounds rounds of the workgroup.
      // The lane index made opaque here, so that nothing of this path is
      // computed ahead of the branch and held through the other one.
      int lane_id = tid;
      asm volatile("" : "+v"(lane_id));
      for (uint32_t u = (uint32_t)lane_id; u < 3 * total; u += Threads)
        sites_channel<TY, TX, Threads>(S, J, list, y0, x0, total, u);
    } else {
      // One site per thread.  Eligibility already implies the driver gate
      // passes (reconstruction.py:890-904 re-tests the same row gate and the
      // same point score), so every listed centre is attempted.
      for (uint32_t q = tid; q < total; q += Threads) {
        const int li = list[q];
        const int ty = li / TX, tx = li % TX;
        const int y = y0 + ty, x = x0 + tx;
        const int cr = ty + 5, cc = tx + 5;  // history index of the centre
        const int ar = ty + 5, ac = tx + 6;  // aux index of the centre
        const int centre = (ty + 1) * PC + tx + 1;
        const float fallback = S.writer_coarse[ty];
        // the site's own working RGB, loaded first so that its latency
        // overlaps the averaging below
        uint32_t lanes[4];
        load_pixel(J.rgbi, (size_t)y * W + x, lanes);
        const float original[3] = {J.lut[lanes[0]], J.lut[lanes[1]], J.lut[lanes[2]]};
        // record at table index i with point values (aux, score) of its pixel
        auto feature = [&](int i, float point_aux, float point_score, float f[4]) {
          const uint8_t valid = S.record_valid[i];
#pragma unroll
          for (int l = 0; l < 3; ++l) f[l] = (valid >> l) & 1u ? S.record_ratio[l][i] : fallback;
          f[3] = point_score > 0.0f ? point_aux : fallback;
        };
        float weights[4], features[5][4];
        weights[0] = S.record_w[0][centre];
        weights[1] = S.record_w[1][centre];
        weights[2] = S.record_w[2][centre];
        weights[3] = S.score[cr][cc];
        feature(centre, S.aux[ar][ac], S.score[cr][cc], features[0]);
        int record_count = 1;
        if (P.cross_neighbor) {
          // streaming.py:259-308: neighbours share the centre row's fallback;
          // out-of-row neighbours are all-zero records; rows -1 and H take the
          // pseudo-row / repeated-row point values already in the halo.
          record_count = 5;
          if (x > 0) {
            feature(centre - 1, S.aux[ar][ac - 1], S.score[cr][cc - 1], features[1]);
          } else {
            features[1][0] = features[1][1] = features[1][2] = features[1][3] = 0.0f;
          }
          if (x < W - 1) {
            feature(centre + 1, S.aux[ar][ac + 1], S.score[cr][cc + 1], features[2]);
          } else {
            features[2][0] = features[2][1] = features[2][2] = features[2][3] = 0.0f;
          }
          feature(centre - PC, S.aux[ar - 1][ac], S.score[cr - 1][cc], features[3]);
          feature(centre + PC, S.aux[ar + 1][ac], S.score[cr + 1][cc], features[4]);
        }
        auto wrgb_at = [&](int dy, int dx, int k) { return S.wrgb[k][ty + dy][tx + dx]; };
        float candidates[3][3];
        rgb_candidates(wrgb_at, weights, candidates);
        Site site;
        combine_candidate(candidates[0], candidates[1], candidates[2], original, weights,
                          features, record_count, fallback, P.combine, site.cand);
        site.orig[0] = original[0];
        site.orig[1] = original[1];
        site.orig[2] = original[2];
        site.pixel = (uint32_t)((size_t)y * W + x) | (S.writer_floor[ty] ? kSiteFloorBit : 0u);
        J.sites[S.row_offset[ty] + (q - S.row_start[ty])] = site;
      }
    }
  };
  constexpr int kStrip = sites_strip(TY);
  // the needed bits start clear; each tile's site phase clears them again
  for (int i = threadIdx.x; i < kWords; i += Threads) S.needed_bits[i] = 0u;
  for_each_unit_dynamic(set, [&](const JobDev& J, const int unit) {
    const int strips = sites_strips(TY, J.win.trow_n);
    const int strip_units = strips * J.tiles_x;
    int tile_x, trow_first, trow_end;
    if (unit < strip_units) {
      tile_x = unit % J.tiles_x;
      trow_first = J.win.trow_lo + (unit / J.tiles_x) * kStrip;
      trow_end = trow_first + kStrip;
    } else {
      const int u = unit - strip_units;
      tile_x = u % J.tiles_x;
      trow_first = J.win.trow_lo + strips * kStrip + u / J.tiles_x;
      trow_end = trow_first + 1;
    }
    bool staged = false;  // the LDS holds the halo of the tile above
    for (int tile_y = trow_first; tile_y < trow_end; ++tile_y) {
      const int t = tile_y * J.tiles_x + tile_x;
      const uint32_t total = J.tile_count[t];
      if (total == 0) {
        staged = false;
        continue;
      }
      sites_tile(J, t, tile_y, tile_x, total, staged);
      staged = kStrip > 1;
      __syncthreads();  // the next tile reuses the LDS
    }
  });
}

// Exclusive scan of the (row, tile-column) segment counts in row-major
// order: the site order the serial writer requires (host_writer.py:73-93).
// Three passes: per-1024-segment sums, one workgroup scanning those sums
// (and writing the total), and the per-block rescan with its offset.
constexpr int kSegPerBlock = 1024;

// The window's segments: rows [trow_lo, trow_lo + trow_n) x TY, every tile
// column; the scan continues from the sites of the rows before
// (counters[0], zero for a whole frame).
__device__ __forceinline__ long window_segment_lo(const JobDev& J) {
  return (long)J.win.trow_lo * J.tile_rows * J.tiles_x;
}
__device__ __forceinline__ long window_segments(const JobDev& J) {
  return (long)J.win.trow_n * J.tile_rows * J.tiles_x;
}

__global__ __launch_bounds__(kThreads) void k_seg_partials(JobSet set) {
  __shared__ uint32_t wave_sums[kThreads / 64];
  for_each_unit(set, [&](const JobDev& J, const int unit) {
    const long n = window_segment_lo(J) + window_segments(J);
    const long base = window_segment_lo(J) + (long)unit * kSegPerBlock;
    uint32_t sum = 0;
#pragma unroll
    for (int j = 0; j < kSegPerBlock / kThreads; ++j) {
      const long i = base + threadIdx.x + j * kThreads;
      if (i < n) sum += J.seg_count[i];
    }
    uint32_t total;
    block_exclusive_scan(sum, wave_sums, &total);
    if (threadIdx.x == 0) J.seg_block_sums[unit] = total;
  });
}

__global__ __launch_bounds__(kThreads) void k_seg_top(JobSet set) {
  __shared__ uint32_t wave_sums[kThreads / 64];
  for_each_unit(set, [&](const JobDev& J, const int unit) {
    const long segments = window_segments(J);
    const int blocks = (int)((segments + kSegPerBlock - 1) / kSegPerBlock);
    uint32_t carry = J.counters[0];
    __syncthreads();  // every thread has read the carry before thread 0 replaces it
    for (int base = 0; base < blocks; base += kThreads) {
      const int i = base + threadIdx.x;
      const uint32_t value = i < blocks ? J.seg_block_sums[i] : 0u;
      uint32_t total;
      const uint32_t offset = block_exclusive_scan(value, wave_sums, &total);
      if (i < blocks) J.seg_block_sums[i] = carry + offset;
      carry += total;
    }
    if (threadIdx.x == 0) {
      J.counters[0] = carry;
      J.row_start[J.height] = carry;
    }
  });
}

__global__ __launch_bounds__(kThreads) void k_seg_offsets(JobSet set) {
  __shared__ uint32_t wave_sums[kThreads / 64];
  for_each_unit(set, [&](const JobDev& J, const int unit) {
    const long n = window_segment_lo(J) + window_segments(J);
    constexpr int kPer = kSegPerBlock / kThreads;
    // each thread owns kPer consecutive segments so the rescan stays ordered
    const long first =
        window_segment_lo(J) + (long)unit * kSegPerBlock + (long)threadIdx.x * kPer;
    uint32_t values[kPer];
    uint32_t sum = 0;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      values[j] = first + j < n ? J.seg_count[first + j] : 0u;
      sum += values[j];
    }
    uint32_t total;
    uint32_t offset = J.seg_block_sums[unit] + block_exclusive_scan(sum, wave_sums, &total);
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
      const long i = first + j;
      if (i < n) {
        J.seg_offset[i] = offset;
        // first site of row y: its tile-column-0 segment
        const long y = i / J.tiles_x;
        if (i % J.tiles_x == 0 && y < J.height) J.row_start[y] = offset;
      }
      offset += values[j];
    }
  });
}

// ==========================================================================
// Hidden startup (startup.py:74-353)
// ==========================================================================

// Planes of the first rows (5 in cross-neighbour mode) for the six hidden
// writer stages.
__global__ __launch_bounds__(kThreads) void k_startup_rows(JobSet set) {
  for_each_unit(set, [&](const JobDev& J, const int unit) {
    const DevParams& P = J.P;
    const int rows = J.height < 5 ? J.height : 5;
    const int W = J.width;
    const int i = unit * blockDim.x + threadIdx.x;
    if (i >= rows * W) return;
    const int y = i / W, x = i % W;
    const float alpha = J.stages.alpha[y];
    const bool is_one = alpha == 1.0f;
    float aux3[3];
#pragma unroll
    for (int d = -1; d <= 1; ++d) {
      const int xx = clampi(x + d, 0, W - 1);
      uint32_t lanes[4];
      load_pixel(J.rgbi, (size_t)y * W + xx, lanes);
      const float visible = J.lut[lanes[P.visible_channel]];
      aux3[d + 1] = is_one ? J.stages.replacement[y]
                           : auxiliary(visible, J.lut[lanes[3]], alpha, J.stages.offset[y]);
    }
    float sample = aux3[1];
    if (P.horizontal_minimum && W > 2 && x >= 1 && x <= W - 2)
      sample = fmin_plain(fmin_plain(aux3[0], aux3[1]), aux3[2]);
    const float s = score_value(sample, J.stages.score_base[y + 1], P.score_addend,
                                P.score_scale, P.score_offset, P.score_floor);
    uint32_t lanes[4];
    load_pixel(J.rgbi, (size_t)y * W + x, lanes);
    J.su_aux[i] = aux3[1];
    J.su_score[i] = s;
    J.su_work[(size_t)i * 3 + 0] = J.lut[lanes[0]];
    J.su_work[(size_t)i * 3 + 1] = J.lut[lanes[1]];
    J.su_work[(size_t)i * 3 + 2] = J.lut[lanes[2]];
  });
}

namespace {
// k_startup_sites: 16 sites per workgroup, 16 lanes per site.  Lanes 0-4
// average the score windows of the site's five feature records, lanes 5-9
// their weighted-aux windows, lanes 10-12 the three weighted-RGB windows,
// lane 13 counts the decision bars; lane 0 then forms the site from them.
constexpr int kStartupSites = kThreads / 16;
// The startup histories of the workgroup's columns, staged in LDS: real
// rows 0..4 (the pseudo-row -1 derives from row 0), columns x0 - 6 ..
// x0 + kStartupSites + 5 clamped into the frame.
constexpr int kStartupCols = kStartupSites + 12;
struct StartupShared {
  float score[5][kStartupCols];
  float aux[5][kStartupCols];
  float work[3][5][kStartupCols];
  double sums[kStartupSites][13][3];  // the lanes' averages
  uint8_t fallback[kStartupSites];    // the decision bars' verdict
};

// _startup_histories (startup.py:74-128) as point accessors over the staged
// columns (x0 - 6 + c for column c).
struct StartupView {
  const StartupShared& S;
  int W, x0;
  float floor;
  __device__ int col(int x) const { return x - x0 + 6; }
  __device__ float score(int y, int x) const {
    if (x < 0 || x >= W || y < -1) return 0.0f;
    if (y == -1) return floor;
    return S.score[y][col(x)];
  }
  __device__ float waux(int y, int x) const {
    if (x < 0 || x >= W || y < -1) return 0.0f;
    if (y == -1) return S.aux[0][col(x)] * floor;
    return S.score[y][col(x)] * S.aux[y][col(x)];
  }
  __device__ float wrgb(int y, int x, int k) const {
    if (x < 0 || x >= W || y < -1) return 0.0f;
    if (y == -1) return S.work[k][0][col(x)] * floor;
    return S.score[y][col(x)] * S.work[k][y][col(x)];
  }
  // Raw decision history: first real row for y < 0; left guard replicates;
  // the right guard is replicated for real rows and zero for y < 0 (the
  // staged columns are clamped into the frame).
  __device__ float raw(int y, int x) const {
    if (x >= W && y < 0) return 0.0f;
    return S.aux[y < 0 ? 0 : y][col(x)];
  }
  // startup.py:131-183: rows at or below the first hidden centre and
  // out-of-row columns hold allocator zero.
  __device__ bool has_record(int y, int x) const { return y > -6 && x >= 0 && x < W; }
};
}  // namespace

// The hidden-startup sites: one workgroup per (stage, run of kStartupSites
// columns), its columns' startup histories staged in LDS first.
__global__ __launch_bounds__(kThreads) void k_startup_sites(JobSet set) {
  __shared__ StartupShared S;
  for_each_unit(set, [&](const JobDev& J, const int unit) {
    const DevParams& P = J.P;
    const int W = J.width;
    const int chunks = (W + kStartupSites - 1) / kStartupSites;
    const int stage = unit / chunks + 1;
    const int x0 = (unit % chunks) * kStartupSites;
    const int site = threadIdx.x >> 4, job = threadIdx.x & 15;
    const int x = x0 + site;
    const int center_y = stage - 7;
    for (int i = threadIdx.x; i < 5 * kStartupCols; i += kThreads) {
      const int y = i / kStartupCols, c = i % kStartupCols;
      const size_t at = (size_t)y * W + clampi(x0 - 6 + c, 0, W - 1);
      S.score[y][c] = J.su_score[at];
      S.aux[y][c] = J.su_aux[at];
      S.work[0][y][c] = J.su_work[at * 3 + 0];
      S.work[1][y][c] = J.su_work[at * 3 + 1];
      S.work[2][y][c] = J.su_work[at * 3 + 2];
    }
    __syncthreads();
    StartupView V{S, W, x0, P.score_floor};
    // the five records: centre, left, right, up, down
    const int ry[5] = {center_y, center_y, center_y, center_y - 1, center_y + 1};
    const int rx[5] = {x, x - 1, x + 1, x, x};
    if (x < W && job < 13) {
      double q[3] = {0.0, 0.0, 0.0};
      if (job < 10) {
        const int r = job < 5 ? job : job - 5;
        const int y = ry[r], xr = rx[r];
        if (V.has_record(y, xr)) {
          if (job < 5)
            unscaled_averages([&](int dy, int dx) { return V.score(y - 4 + dy, xr - 4 + dx); }, q);
          else
            unscaled_averages([&](int dy, int dx) { return V.waux(y - 4 + dy, xr - 4 + dx); }, q);
        }
      } else {
        const int k = job - 10;
        auto lane = [&](int dy, int dx) { return V.wrgb(center_y - 4 + dy, x - 4 + dx, k); };
        unscaled_averages(lane, q);
        q[2] = rgb_q16(lane);
      }
      S.sums[site][job][0] = q[0];
      S.sums[site][job][1] = q[1];
      S.sums[site][job][2] = q[2];
    } else if (x < W && job == 13) {
      auto aux_at = [&](int dy, int dx) { return V.raw(center_y + dy, x + dx); };
      S.fallback[site] =
          decision_fallback(aux_at, P.decision_threshold, P.radius, P.count_limit) ? 1 : 0;
    }
    __syncthreads();
    if (x >= W || job != 0) return;
    const int i = (stage - 1) * W + x;
    const bool cross = P.cross_neighbor != 0;
    const int writer_width = (cross && stage <= 2) ? W - 4 : W;
    J.su_attempted[i] = 0;
    if (x >= writer_width) return;
    const StageDev& st = J.stages;
    const bool floor_enabled = P.gate_primary || st.gate_secondary[stage - 1];
    const int32_t row_gate = st.row_gate[stage - 1];
    const float combiner_reference = st.coarse[stage - 1];
    const float fallback_center = st.coarse[max(1, center_y + 6) - 1];
    const float fallback_up = st.coarse[max(1, center_y + 5) - 1];
    const float fallback_down = st.coarse[max(1, center_y + 7) - 1];
    const float record_fallback[5] = {fallback_center, fallback_center, fallback_center,
                                      fallback_up, fallback_down};
    // feature_record (streaming.py:236-256) from the lanes' sums
    float weights[4], features[5][4];
    auto record = [&](int r, float w[4], float f[4]) {
      const int y = ry[r], xr = rx[r];
      if (!V.has_record(y, xr)) {
        w[0] = w[1] = w[2] = w[3] = 0.0f;
        f[0] = f[1] = f[2] = f[3] = 0.0f;
        return;
      }
      const double* unrounded = S.sums[site][r];
      const double* numerator = S.sums[site][r + 5];
#pragma unroll
      for (int l = 0; l < 3; ++l) {
        w[l] = (float)unrounded[l];
        f[l] = unrounded[l] > 0.0 ? (float)(numerator[l] / unrounded[l]) : record_fallback[r];
      }
      w[3] = V.score(y, xr);
      f[3] = w[3] > 0.0f ? V.raw(y, xr) : record_fallback[r];
    };
    record(0, weights, features[0]);
    if (S.fallback[site] || row_gate != 0 || (floor_enabled && weights[3] >= 1.0f)) return;
    int record_count = 1;
    if (cross) {
      record_count = 5;
      float unused[4];
      for (int r = 1; r < 5; ++r) record(r, unused, features[r]);
    }
    // rgb_candidates (reconstruction.py:425-462) from the lanes' sums
    float candidates[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      const double* q = S.sums[site][10 + k];
#pragma unroll
      for (int s = 0; s < 3; ++s) {
        if (weights[s] > 0.0f) {
          double reciprocal = 1.0 / (double)weights[s];
          candidates[s][k] = (float)(q[s] * reciprocal);
        } else {
          candidates[s][k] = (float)q[s];
        }
      }
    }
    const float original[3] = {S.work[0][0][V.col(x)], S.work[1][0][V.col(x)],
                               S.work[2][0][V.col(x)]};
    Site out;
    combine_candidate(candidates[0], candidates[1], candidates[2], original, weights,
                      features, record_count, combiner_reference, P.combine, out.cand);
    out.orig[0] = original[0];
    out.orig[1] = original[1];
    out.orig[2] = original[2];
    out.pixel = (uint32_t)x | (floor_enabled ? kSiteFloorBit : 0u);
    J.su_sites[i] = out;
    J.su_attempted[i] = 1;
  });
}

// ==========================================================================
// Writer support (reconstruction.py:916-982, dither.py:59-104)
// ==========================================================================

// Per-job flags and counters to zero (one unit per job).  A kernel, not a
// memset, so that it can sit inside a captured graph.
__global__ __launch_bounds__(kThreads) void k_zero_jobs(JobSet set) {
  for_each_unit(set, [&](const JobDev& J, const int unit) {
    for (int i = threadIdx.x; i < kCountSlots * kCountStride; i += blockDim.x)
      J.writer_counts[i] = 0ull;
    if (threadIdx.x < 16) J.counters[threadIdx.x] = 0u;
    if (threadIdx.x == 0) *J.error_flags = 0u;
  });
}

// Integer draw schedule (writer_record) of the main sites [win.site_lo,
// counters[0]) and, when the window holds the start of the frame, of every
// hidden-startup slot of every job; the main site count is read on the
// device, so the launch geometry does not depend on it.
__global__ __launch_bounds__(kThreads) void k_writer_prep(JobSet set) {
  const uint32_t stride = gridDim.x * blockDim.x;
  const uint32_t first = blockIdx.x * blockDim.x + threadIdx.x;
  for (int j = 0; j < set.count; ++j) {
    const JobDev& J = job_at(set, j);
    if (J.skip) continue;
    const uint32_t lo = J.win.site_lo;
    const uint32_t count = J.counters[0] - lo;
    const uint32_t total = count + (J.win.startup ? 6u * (uint32_t)J.width : 0u);
    for (uint32_t i = first; i < total; i += stride) {
      if (i < count) {
        J.writer_records[lo + i] = writer_record(J.sites[lo + i], true, J.K);
      } else {
        const uint32_t k = i - count;
        J.su_writer_records[k] = writer_record(J.su_sites[k], J.su_attempted[k] != 0, J.K);
      }
    }
  }
}

// Start states of sites [begin, end) of one job from those the host walk
// kept at every kStateGroup-th site: one thread per group re-walks the
// integer chain (the walk's own functions) from the group's start.
__global__ __launch_bounds__(kThreads) void k_writer_expand(JobDev J, uint32_t begin,
                                                            uint32_t end) {
  const uint32_t g = begin / kStateGroup + blockIdx.x * blockDim.x + threadIdx.x;
  const uint32_t first = g * kStateGroup;
  if (first >= end) return;
  const uint32_t last = min(first + kStateGroup, end);
  // the group's records first, so that their loads overlap
  WriterRecord r[kStateGroup];
#pragma unroll
  for (uint32_t k = 0; k < kStateGroup; ++k)
    if (first + k < last) r[k] = J.writer_records[first + k];
  uint32_t s = J.group_state[g];
  uint64_t draws = 0;
#pragma unroll
  for (uint32_t k = 0; k < kStateGroup; ++k) {
    const uint32_t i = first + k;
    if (i >= last) break;
    if (i >= begin) J.site_state[i] = s;
    s = draws_all_channels(r[k]) ? chain_site_drawing(r[k], s, draws) : chain_site(r[k], s, draws);
  }
}

// The writer itself over sites [begin, end) of one job, every site in
// parallel from the LCG state the serial integer chain recorded at its
// start: values, written / changed counts, the patch of the output, and a
// check that each site ends in the state the chain recorded for the next
// one (`end_state` after the last site of the range).
__global__ __launch_bounds__(kThreads) void k_writer_values(JobDev J, uint32_t begin,
                                                            uint32_t end,
                                                            uint32_t end_state_value,
                                                            const uint32_t* end_state_at) {
  const uint32_t end_state = end_state_at ? *end_state_at : end_state_value;
  const uint32_t i = begin + blockIdx.x * blockDim.x + threadIdx.x;
  bool written = false, changed = false;
  if (i < end) {
    const Site site = J.sites[i];
    uint32_t state = J.site_state[i];
    float values[3];
    write_pixel(site.cand, site.orig, (site.pixel & kSiteFloorBit) != 0, J.K, state, values);
    const uint32_t expected = i + 1 < end ? J.site_state[i + 1] : end_state;
    uint32_t error = state != expected ? kErrWriterChain : 0u;
    written = values[0] != site.orig[0] || values[1] != site.orig[1] ||
              values[2] != site.orig[2];
    if (!(__builtin_isfinite(values[0]) && __builtin_isfinite(values[1]) &&
          __builtin_isfinite(values[2]))) {
      error |= kErrNonfiniteWork;
    } else {
      uint16_t rendered[3];
      for (int c = 0; c < 3; ++c) {
        rendered[c] = emit_value(values[c], J.factor_high, J.factor_low);
        changed |= rendered[c] != emit_value(site.orig[c], J.factor_high, J.factor_low);
      }
      const uint32_t pixel = site.pixel & kSitePixelMask;
      if (changed) {
        uint16_t* o = J.out + (size_t)pixel * 3;
        o[0] = rendered[0];
        o[1] = rendered[1];
        o[2] = rendered[2];
      }
      if (J.diag_changed) J.diag_changed[pixel] = changed ? 1 : 0;
    }
    if (error) atomicOr(J.error_flags, error);
  }
  __shared__ uint32_t wave_counts[kThreads / 64][2];
  const unsigned long long written_mask = __ballot(written);
  const unsigned long long changed_mask = __ballot(changed);
  const int wave = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    wave_counts[wave][0] = (uint32_t)__popcll(written_mask);
    wave_counts[wave][1] = (uint32_t)__popcll(changed_mask);
  }
  __syncthreads();
  if (threadIdx.x < 2) {
    uint32_t total = 0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) total += wave_counts[w][threadIdx.x];
    if (total)
      atomicAdd(&J.writer_counts[(blockIdx.x % kCountSlots) * kCountStride + threadIdx.x],
                (unsigned long long)total);
  }
}

// The same check over the six hidden startup rows, which write no output:
// each attempted site, run from its recorded start state, must end in the
// state recorded for the next index (a skipped index passes the state on).
__global__ __launch_bounds__(kThreads) void k_writer_startup_check(JobDev J,
                                                                   uint32_t final_value,
                                                                   const uint32_t* final_at) {
  const uint32_t final_state = final_at ? *final_at : final_value;
  const uint32_t count = 6u * (uint32_t)J.width;
  const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  uint32_t state = J.su_site_state[i];
  if (J.su_attempted[i]) {
    const Site site = J.su_sites[i];
    float values[3];
    write_pixel(site.cand, site.orig, (site.pixel & kSiteFloorBit) != 0, J.K, state, values);
  }
  const uint32_t expected = i + 1 < count ? J.su_site_state[i + 1] : final_state;
  if (state != expected) atomicOr(J.error_flags, kErrWriterChain);
}

// The writer's serial LCG chain on one wavefront per job (host_writer.py:73-93,
// reconstruction.py:916-982): records [begin, end) of the job, in order,
// from the state carried in J.chain.  The 64 lanes stage the records
// through LDS (the next chunk loads while the current one is chained) and
// collect the start state of every site, one per lane, to store 64 at a
// time.  The chain value itself is wave-uniform and lives in a scalar
// register.
constexpr int kWriterChunk = 512;
constexpr int kChunkWords = kWriterChunk * 6;
constexpr int kChunkPerLane = kChunkWords / 64;

__device__ __forceinline__ uint32_t uniform32(uint32_t v) {
  return (uint32_t)__builtin_amdgcn_readfirstlane((int)v);
}

// v_writelane_b32: `value` (wave-uniform) into lane `lane` of `old` (the
// compiler intrinsic; this compiler has no builtin spelling for it).
extern "C" __device__ int dice_writelane(int value, int lane, int old) __asm("llvm.amdgcn.writelane");

// Records [first, first + min(kWriterChunk, end - first)) as words, kChunkPerLane per lane.
__device__ __forceinline__ void fetch_chunk(const WriterRecord* records, uint32_t first,
                                            uint32_t end, int lane, uint32_t next[kChunkPerLane]) {
  const uint32_t* source = reinterpret_cast<const uint32_t*>(records + first);
  const uint32_t n = min((uint32_t)kWriterChunk, end - first) * 6;
#pragma unroll
  for (int j = 0; j < kChunkPerLane; ++j) {
    const uint32_t e = (uint32_t)(lane + j * 64);
    next[j] = e < n ? source[e] : 0u;
  }
}

// The six hidden startup rows of every job (6*W slots, launched with one
// workgroup per job), per channel, with the per-stage receipt.
__global__ __launch_bounds__(64) void k_writer_startup_chain(const JobDev* jobs) {
  const JobDev& J = jobs[blockIdx.x];
  if (J.skip) return;
  __shared__ uint32_t words[2][kChunkWords];
  const int lane = threadIdx.x;
  const WriterRecord* records = J.su_writer_records;
  uint32_t* states = J.su_site_state;
  const uint32_t W = (uint32_t)J.width;
  const uint32_t end = 6u * W;
  uint32_t state = uniform32(J.chain->state);
  long long stage_attempted = 0, stage_draws = 0;
  uint32_t stage = 0, stage_end = W;
  uint32_t next[kChunkPerLane];
  fetch_chunk(records, 0u, end, lane, next);
  int buffer = 0;
  for (uint32_t first = 0; first < end; first += kWriterChunk) {
    const uint32_t n = min((uint32_t)kWriterChunk, end - first);
#pragma unroll
    for (int j = 0; j < kChunkPerLane; ++j) words[buffer][lane + j * 64] = next[j];
    __syncthreads();
    if (first + kWriterChunk < end) fetch_chunk(records, first + kWriterChunk, end, lane, next);
    const uint32_t* w = words[buffer];
    uint32_t collected = 0;
    for (uint32_t k = 0; k < n; ++k) {
      if (first + k == stage_end) {  // next hidden stage
        if (lane == 0) {
          J.chain->startup_attempted[stage] = stage_attempted;
          J.chain->startup_advances[stage] = stage_draws;
        }
        stage_attempted = 0;
        stage_draws = 0;
        ++stage;
        stage_end += W;
      }
      const uint32_t* r = w + k * 6;
      collected = lane == (int)(k & 63u) ? state : collected;
      const uint32_t l0 = uniform32(r[0]);
      uint32_t d0, d1, d2;
      state = chain_channel(l0, uniform32(r[1]), state, d0);
      state = chain_channel(uniform32(r[2]), uniform32(r[3]), state, d1);
      state = chain_channel(uniform32(r[4]), uniform32(r[5]), state, d2);
      stage_draws += (long long)(d0 + d1 + d2);
      stage_attempted += (l0 >> 31) & 1u;
      if ((k & 63u) == 63u || k + 1 == n) {
        const uint32_t base = first + (k & ~63u);
        if (lane <= (int)(k & 63u)) states[base + lane] = collected;
      }
    }
    __syncthreads();
    buffer ^= 1;
  }
  if (lane == 0) {
    J.chain->startup_attempted[stage] = stage_attempted;
    J.chain->startup_advances[stage] = stage_draws;
    J.chain->startup_final = state;
    J.chain->state = state;
  }
}

// Band `band` of `bands` of every job's main sites (one workgroup per job;
// a job with fewer rows than bands has one band per row).  Where all three
// channels of a site draw, the lanes evaluate chain_site_drawing's six
// tests at once (chain_site_lanes): per lane, t - low from one 24-bit
// multiply-add (add - low prepared with the site's words), a 24-bit mask and
// a compare with the interval length, then the ballot of the six keep bits reads
// the site's draws and end lane from a 64-entry table held one entry per
// lane, and that lane's state becomes the chain value.  Other sites take
// the per-channel path on the scalar unit.  The chain value is kept
// unmasked (only its low 24 bits are ever read) and masked when stored.
__global__ __launch_bounds__(64) void k_writer_chain(const JobDev* jobs, int band, int bands) {
  const JobDev& J = jobs[blockIdx.x];
  if (J.skip) return;
  const int64_t H = J.height;
  const int job_bands = (int)max((int64_t)1, min((int64_t)bands, H));
  if (band >= job_bands) return;
  // (read through readfirstlane, so that the loops below are uniform and
  // the chain value stays in a scalar register)
  const uint32_t begin = uniform32(J.row_start[H * band / job_bands]);
  const uint32_t end = uniform32(J.row_start[H * (band + 1) / job_bands]);
  __shared__ __align__(16) uint32_t words[2][kChunkWords];
  const int lane = threadIdx.x;
  // this lane's speculation: steps, channel word, and its entry of the
  // keep-bits table (entry m = end lane | draws << 8 for keep bits m)
  const int spec = lane < kSpecLanes ? lane : kSpecLanes - 1;
  uint32_t mul = 0, add = 0;
  switch (spec_steps(spec)) {
    case 1: mul = lcg_pow(1); add = lcg_add(1); break;
    case 2: mul = lcg_pow(2); add = lcg_add(2); break;
    case 3: mul = lcg_pow(3); add = lcg_add(3); break;
    case 4: mul = lcg_pow(4); add = lcg_add(4); break;
    case 5: mul = lcg_pow(5); add = lcg_add(5); break;
    default: mul = lcg_pow(6); add = lcg_add(6); break;
  }
  const uint32_t word = 2u * (uint32_t)spec_channel(spec);
  const uint32_t draws_of = spec_draws((uint32_t)lane);
  const uint32_t table = (uint32_t)spec_end_lane(draws_of) | (draws_of << 8);
  const WriterRecord* records = J.writer_records;
  uint32_t* states = J.site_state;
  uint32_t s = uniform32(J.chain->state);
  unsigned long long advances = 0;
  uint32_t next[kChunkPerLane];
  if (begin < end) fetch_chunk(records, begin, end, lane, next);
  int buffer = 0;
  for (uint32_t first = begin; first < end; first += kWriterChunk) {
    const uint32_t n = min((uint32_t)kWriterChunk, end - first);
#pragma unroll
    for (int j = 0; j < kChunkPerLane; ++j) words[buffer][lane + j * 64] = next[j];
    __syncthreads();
    if (first + kWriterChunk < end) fetch_chunk(records, first + kWriterChunk, end, lane, next);
    const uint32_t* w = words[buffer];
    for (uint32_t g = 0; g < n; g += 64) {
      const uint32_t gn = min(64u, n - g);
      // the sites of this group whose three channels all draw
      const uint32_t mine = min(g + (uint32_t)lane, n - 1u) * 6u;
      const bool drawing =
          g + (uint32_t)lane < n && ((w[mine + 1] & w[mine + 3] & w[mine + 5]) & kChanDraws) != 0;
      const unsigned long long fast = __ballot(drawing);
      // this lane's channel words for every site of the group, in registers
      // before the chain starts (no LDS read is then in flight on it)
      // (.x holds add - low, so that one multiply-add gives t - low, and .y
      // the interval length alone: the fast path only sees drawing channels)
      uint2 cw[64];
#pragma unroll
      for (uint32_t j = 0; j < 64; ++j) {
        cw[j] = *reinterpret_cast<const uint2*>(w + min(g + j, n - 1u) * 6u + word);
        cw[j].x = add - cw[j].x;
        cw[j].y &= kChanLen;
      }
      uint32_t collected = 0;
      if (gn == 64u && fast == ~0ull) {  // a full group whose sites all draw
        uint32_t group_draws = 0;
#pragma unroll
        for (uint32_t j = 0; j < 64; ++j) {
          collected = (uint32_t)dice_writelane((int)(s & kLcgMask), (int)j, (int)collected);
          const uint32_t t = __umul24(mul, s) + add;
          const uint32_t d = __umul24(mul, s) + cw[j].x;
          const uint32_t m = (uint32_t)__ballot((d & kLcgMask) < cw[j].y) & 63u;
          const uint32_t e = (uint32_t)__builtin_amdgcn_readlane((int)table, (int)m);
          s = (uint32_t)__builtin_amdgcn_readlane((int)t, (int)(e & 0xFFu));
          group_draws += e >> 8;
        }
        advances += group_draws;
        states[first + g + lane] = collected;
        continue;
      }
#pragma unroll
      for (uint32_t j = 0; j < 64; ++j) {
        if (j >= gn) continue;  // uniform: the group's last sites
        collected = (uint32_t)dice_writelane((int)(s & kLcgMask), (int)j, (int)collected);
        if ((fast >> j) & 1ull) {
          const uint32_t t = __umul24(mul, s) + add;
          // keeps(t, low, len), with t - low from its own multiply-add
          const uint32_t d = __umul24(mul, s) + cw[j].x;
          const uint32_t m = (uint32_t)__ballot((d & kLcgMask) < cw[j].y) & 63u;
          const uint32_t e = (uint32_t)__builtin_amdgcn_readlane((int)table, (int)m);
          s = (uint32_t)__builtin_amdgcn_readlane((int)t, (int)(e & 0xFFu));
          advances += e >> 8;
        } else {
          const uint32_t* r = w + (g + j) * 6u;
          uint32_t d0, d1, d2;
          s = chain_channel(uniform32(r[0]), uniform32(r[1]), s, d0);
          s = chain_channel(uniform32(r[2]), uniform32(r[3]), s, d1);
          s = chain_channel(uniform32(r[4]), uniform32(r[5]), s, d2);
          advances += d0 + d1 + d2;
        }
      }
      if ((uint32_t)lane < gn) states[first + g + lane] = collected;
    }
    __syncthreads();
    buffer ^= 1;
  }
  if (lane == 0) {
    J.chain->advances += advances;
    J.band_end[band] = s & kLcgMask;
    J.chain->state = s & kLcgMask;
  }
}

__global__ __launch_bounds__(kThreads) void k_fill_u8(uint8_t* p, size_t n, uint8_t v) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}



// ==========================================================================
// Level-1 test kernels: the device functions above, over host vectors.
// ==========================================================================

__global__ void t_unscaled_averages(const float* patches, long count, double* out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const float* p = patches + i * 81;
  auto at = [&](int dy, int dx) { return p[dy * 9 + dx]; };
  double q[3];
  unscaled_averages(at, q);
  out[i * 3 + 0] = q[0];
  out[i * 3 + 1] = q[1];
  out[i * 3 + 2] = q[2];
}

__global__ void t_rgb_unscaled_averages(const float* patches, long count, double* out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const float* p = patches + i * 243;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    auto lane = [&](int dy, int dx) { return p[(dy * 9 + dx) * 3 + k]; };
    double q[3];
    unscaled_averages(lane, q);
    q[2] = rgb_q16(lane);
    for (int s = 0; s < 3; ++s) out[(i * 3 + s) * 3 + k] = q[s];
  }
}

__global__ void t_dither_chain(const double* values, const float* scales, long count,
                               uint32_t state_in, float low, float high,
                               double* deltas, uint32_t* state_out,
                               unsigned long long* advances_out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  WriterConsts K;
  K.low = (double)low;
  K.high = (double)high;
  K.low_lt_high = K.low < K.high;
  uint32_t state = state_in;
  unsigned long long advances = 0;
  for (long i = 0; i < count; ++i) {
    uint32_t before = state;
    deltas[i] = dither_delta(values[i], K, scales[i], state);
    advances += state != before;
  }
  *state_out = state;
  *advances_out = advances;
}

__global__ void t_emit(const float* values, long count, uint16_t* out,
                       const uint32_t* high, const uint32_t* low, uint32_t* flags) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  if (!__builtin_isfinite(values[i])) atomicOr(flags, 2u);
  out[i] = emit_value(values[i], high, low);
}

__global__ void t_planes_aux(const uint16_t* rgbi, long pixels, int width,
                             const float* lut, DevParams P, const float* alpha,
                             const uint8_t* is_one, const float* replacement,
                             const float* offset, float* working, float* aux,
                             uint32_t* flags) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= pixels) return;
  const int y = (int)(i / width);
  uint32_t lanes[4];
  load_pixel(rgbi, (size_t)i, lanes);
  float w[4];
  for (int k = 0; k < 4; ++k) {
    w[k] = lut[lanes[k]];
    working[i * 4 + k] = w[k];
  }
  const float a = is_one[y] ? replacement[y]
                            : auxiliary(w[P.visible_channel], w[3], alpha[y], offset[y]);
  if (!__builtin_isfinite(a)) atomicOr(flags, kErrNonfiniteAux);
  aux[i] = a;
}

__global__ void t_planes_score(long pixels, int width, DevParams P,
                               const float* score_base, const float* working,
                               const float* aux, float* score, float* waux,
                               float* wrgb) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= pixels) return;
  const int y = (int)(i / width), x = (int)(i % width);
  float sample = aux[i];
  if (P.horizontal_minimum && width > 2 && x >= 1 && x <= width - 2)
    sample = fmin_plain(fmin_plain(aux[i - 1], aux[i]), aux[i + 1]);
  const float s = score_value(sample, score_base[y], P.score_addend, P.score_scale,
                              P.score_offset, P.score_floor);
  score[i] = s;
  waux[i] = s * aux[i];
  for (int k = 0; k < 3; ++k) wrgb[i * 3 + k] = s * working[i * 4 + k];
}

__global__ void t_decision(const float* aux, int height, int width, DevParams P,
                           uint8_t* fallback) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long)height * width) return;
  const int y = (int)(i / width), x = (int)(i % width);
  auto aux_at = [&](int dy, int dx) {
    return aux[(size_t)clampi(y + dy, 0, height - 1) * width + clampi(x + dx, 0, width - 1)];
  };
  fallback[i] = decision_fallback(aux_at, P.decision_threshold, P.radius, P.count_limit);
}

// Plain 16-byte-vector copy: the calibration kernel for HBM counters and
// attainable bandwidth.
__global__ __launch_bounds__(kThreads) void k_copy(const uint4* src, uint4* dst, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
    dst[i] = src[i];
}

// ==========================================================================
// Launchers
// ==========================================================================

hipError_t launch_copy(const void* src, void* dst, size_t bytes, hipStream_t s) {
  const size_t n = bytes / 16;
  hipLaunchKernelGGL(k_copy, dim3(256 * 8 * 4), dim3(kThreads), 0, s,
                     static_cast<const uint4*>(src), static_cast<uint4*>(dst), n);
  return hipGetLastError();
}

static inline unsigned blocks_for(long n, int threads) {
  return (unsigned)((n + threads - 1) / threads);
}

namespace {
// Workgroups that stay resident on the whole device for one kernel: the
// grid of a persistent launch (never more than the work needs).
unsigned resident_blocks(const void* kernel, int threads) {
  static std::mutex mutex;
  static std::map<std::pair<int, const void*>, unsigned> cache;
  int device = 0;
  (void)hipGetDevice(&device);
  std::lock_guard<std::mutex> lock(mutex);
  auto key = std::make_pair(device, kernel);
  auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int cus = 0, per_cu = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess ||
      per_cu < 1 || cus < 1) {
    per_cu = 1;
    cus = 256;
  }
  const unsigned blocks = (unsigned)(per_cu * cus);
  cache[key] = blocks;
  return blocks;
}

template <class Units>
JobSet make_set(const JobDev* d_jobs, const JobDev* h_jobs, int n, Units&& units) {
  JobSet set;
  set.jobs = d_jobs;
  set.count = n;
  set.base[0] = 0;
  for (int j = 0; j < n; ++j) {
    const uint32_t u = h_jobs[j].skip ? 0u : (uint32_t)units(h_jobs[j]);
    set.base[j + 1] = set.base[j] + u;
  }
  for (int j = n + 1; j <= kMaxBatch; ++j) set.base[j] = set.base[n];
  set.queue = nullptr;
  return set;
}

template <class Kernel, class... Args>
void launch_persistent(Kernel kernel, const JobSet& set, int threads, hipStream_t stream,
                       Args... args) {
  const uint32_t total = set.base[set.count];
  if (total == 0) return;
  const unsigned grid =
      std::min<unsigned>(total, resident_blocks(reinterpret_cast<const void*>(kernel), threads));
  hipLaunchKernelGGL(kernel, dim3(grid), dim3(threads), 0, stream, set, args...);
}

uint32_t segment_blocks(const JobDev& J) {
  const long segments = (long)J.win.trow_n * J.tile_rows * J.tiles_x;
  return (uint32_t)((segments + kSegPerBlock - 1) / kSegPerBlock);
}
}  // namespace

hipError_t launch_zero_jobs(const JobDev* d_jobs, const JobDev* h_jobs, int n,
                            hipStream_t stream) {
  const JobSet set = make_set(d_jobs, h_jobs, n, [](const JobDev&) { return 1; });
  launch_persistent(k_zero_jobs, set, kThreads, stream);
  return hipGetLastError();
}

hipError_t launch_producer(const JobDev* d_jobs, const JobDev* h_jobs, int n,
                           hipStream_t stream) {
  auto when = [](const JobDev& J, long units) { return J.run_producer ? units : 0L; };
  launch_persistent(k_producer_blocks, make_set(d_jobs, h_jobs, n, [&](const JobDev& J) {
                      return when(J, J.win.block_n);
                    }), kProducerThreads, stream);
  launch_persistent(k_row_sums, make_set(d_jobs, h_jobs, n, [&](const JobDev& J) {
                      return when(J, J.win.rowsum_n);
                    }), kRowSumThreads, stream);
  launch_persistent(k_producer_chain, make_set(d_jobs, h_jobs, n, [&](const JobDev& J) {
                      return when(J, J.win.chain_hi > J.win.chain_lo ? 2 : 0);
                    }), 64, stream);
  launch_persistent(k_stage_table, make_set(d_jobs, h_jobs, n, [&](const JobDev& J) {
                      return when(J, blocks_for(J.win.stage_hi - J.win.stage_lo, kThreads));
                    }), kThreads, stream);
  return hipGetLastError();
}

hipError_t launch_schedule_rows(const JobDev& J, const ScheduleDev& S,
                                hipStream_t stream) {
  hipLaunchKernelGGL(k_schedule_rows, dim3(blocks_for(J.height, kThreads)),
                     dim3(kThreads), 0, stream, J, S);
  return hipGetLastError();
}

#define DICE_TILE_SHAPES(X) X(16, 64) X(8, 128) X(32, 32) X(16, 32) X(32, 64) X(8, 64)

int tile_shapes(int (*out)[2], int max) {
  int n = 0;
#define DICE_LIST(a, b)   \
  if (n < max) {          \
    out[n][0] = a;        \
    out[n][1] = b;        \
  }                       \
  ++n;
  DICE_TILE_SHAPES(DICE_LIST)
#undef DICE_LIST
  return n;
}

bool tile_shape_supported(int ty, int tx) {
#define DICE_MATCH(a, b) if (ty == a && tx == b) return true;
  DICE_TILE_SHAPES(DICE_MATCH)
#undef DICE_MATCH
  return false;
}

namespace {
template <int TY, int TX>
void launch_scan_shape(const JobSet& set, hipStream_t stream) {
  launch_persistent(k_scan<TY, TX>, set, kScanThreads, stream);
}
template <int TY, int TX>
void launch_sites_shape(const JobSet& set, int variant, hipStream_t stream) {
  constexpr int kWide = sites_wide_threads<TY, TX>();
  static_assert(sites_occupancy<TY, TX, kWide>() * kWide >= 1024,
                "the wide workgroup must reach 16 waves per CU");
  constexpr int kBlocks = sites_occupancy<TY, TX, kThreads>();
  if (variant == 1)
    launch_persistent(k_sites<TY, TX, kThreads, kBlocks>, set, kThreads, stream);
  else
    launch_persistent(k_sites<TY, TX, kWide, 4>, set, kWide, stream);
}
}  // namespace

namespace {
// The tile shape every live job of the batch shares (false: none is live,
// or two live jobs disagree).
bool batch_tile(const JobDev* h_jobs, int n, int* ty, int* tx, bool* any) {
  *any = false;
  for (int j = 0; j < n; ++j) {
    if (h_jobs[j].skip) continue;
    if (!*any) {
      *ty = h_jobs[j].tile_rows;
      *tx = h_jobs[j].tile_cols;
      *any = true;
    } else if (h_jobs[j].tile_rows != *ty || h_jobs[j].tile_cols != *tx) {
      return false;
    }
  }
  return true;
}
}  // namespace

hipError_t launch_scan(const JobDev* d_jobs, const JobDev* h_jobs, int n, hipStream_t stream) {
  int ty = 0, tx = 0;
  bool any = false;
  if (!batch_tile(h_jobs, n, &ty, &tx, &any)) return hipErrorInvalidValue;
  if (!any) return hipSuccess;
  // one unit per strip of kScanStrip tiles, then one per tile
  const JobSet set = make_set(d_jobs, h_jobs, n, [](const JobDev& J) {
    return scan_units(J.win.trow_n, J.tiles_x);
  });
#define DICE_MATCH(a, b)                 \
  if (ty == a && tx == b) {              \
    launch_scan_shape<a, b>(set, stream); \
    return hipGetLastError();            \
  }
  DICE_TILE_SHAPES(DICE_MATCH)
#undef DICE_MATCH
  return hipErrorInvalidValue;
}

hipError_t launch_scan_segments(const JobDev* d_jobs, const JobDev* h_jobs, int n,
                                hipStream_t stream) {
  const JobSet blocks = make_set(d_jobs, h_jobs, n, segment_blocks);
  launch_persistent(k_seg_partials, blocks, kThreads, stream);
  launch_persistent(k_seg_top, make_set(d_jobs, h_jobs, n, [](const JobDev&) { return 1; }),
                    kThreads, stream);
  launch_persistent(k_seg_offsets, blocks, kThreads, stream);
  return hipGetLastError();
}

hipError_t launch_sites(const JobDev* d_jobs, const JobDev* h_jobs, int n, uint32_t* queue,
                        hipStream_t stream, int variant) {
  if (variant < 0 || variant >= kSitesVariants) return hipErrorInvalidValue;
  int ty = 0, tx = 0;
  bool any = false;
  if (!batch_tile(h_jobs, n, &ty, &tx, &any)) return hipErrorInvalidValue;
  if (!any) return hipSuccess;
  // one unit per strip of sites_strip(ty) tiles, then one per tile
  JobSet set = make_set(d_jobs, h_jobs, n, [ty](const JobDev& J) {
    return sites_units(ty, J.win.trow_n, J.tiles_x);
  });
  set.queue = queue;
#define DICE_MATCH(a, b)                           \
  if (ty == a && tx == b) {                        \
    launch_sites_shape<a, b>(set, variant, stream); \
    return hipGetLastError();                      \
  }
  DICE_TILE_SHAPES(DICE_MATCH)
#undef DICE_MATCH
  return hipErrorInvalidValue;
}

namespace {
// A position-sensitive, order-independent 64-bit digest of n 16-bit words:
// the wrapping sum of a 64-bit mix of (index, word).
__global__ __launch_bounds__(kThreads) void k_digest(const uint16_t* __restrict__ words, size_t n,
                                                     unsigned long long* out) {
  unsigned long long sum = 0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x) {
    unsigned long long z = ((unsigned long long)i << 16 | words[i]) + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    sum += z ^ (z >> 31);
  }
  for (int d = 32; d > 0; d >>= 1) sum += __shfl_down(sum, d, 64);
  if ((threadIdx.x & 63) == 0) atomicAdd(out, sum);
}
}  // namespace

hipError_t launch_digest(const void* data, size_t bytes, unsigned long long* out,
                         hipStream_t stream) {
  const size_t n = bytes / 2;
  if (n == 0) return hipSuccess;
  const unsigned blocks = (unsigned)std::min<size_t>((n + kThreads - 1) / kThreads, 4096);
  hipLaunchKernelGGL(k_digest, dim3(blocks), dim3(kThreads), 0, stream,
                     static_cast<const uint16_t*>(data), n, out);
  return hipGetLastError();
}

hipError_t launch_startup(const JobDev* d_jobs, const JobDev* h_jobs, int n,
                          hipStream_t stream) {
  launch_persistent(k_startup_rows, make_set(d_jobs, h_jobs, n, [](const JobDev& J) {
                      return J.win.startup ? blocks_for((long)(J.height < 5 ? J.height : 5) *
                                                            J.width, kThreads)
                                           : 0u;
                    }), kThreads, stream);
  launch_persistent(k_startup_sites, make_set(d_jobs, h_jobs, n, [](const JobDev& J) {
                      return J.win.startup ? 6u * blocks_for(J.width, kStartupSites) : 0u;
                    }), kThreads, stream);
  return hipGetLastError();
}

hipError_t launch_writer_prep(const JobDev* d_jobs, int n, hipStream_t stream) {
  JobSet set;
  set.jobs = d_jobs;
  set.count = n;
  for (int j = 0; j <= kMaxBatch; ++j) set.base[j] = 0;
  set.queue = nullptr;
  const unsigned grid = resident_blocks(reinterpret_cast<const void*>(k_writer_prep), kThreads);
  hipLaunchKernelGGL(k_writer_prep, dim3(grid), dim3(kThreads), 0, stream, set);
  return hipGetLastError();
}

hipError_t launch_writer_expand(const JobDev& J, uint32_t begin, uint32_t end,
                                hipStream_t stream) {
  if (end <= begin) return hipSuccess;
  const long groups = (long)((end - 1) / kStateGroup - begin / kStateGroup + 1);
  hipLaunchKernelGGL(k_writer_expand, dim3(blocks_for(groups, kThreads)), dim3(kThreads), 0,
                     stream, J, begin, end);
  return hipGetLastError();
}

hipError_t launch_writer_values(const JobDev& J, uint32_t begin, uint32_t end,
                                uint32_t end_state, const uint32_t* end_state_at,
                                hipStream_t stream) {
  if (end <= begin) return hipSuccess;
  hipLaunchKernelGGL(k_writer_values, dim3(blocks_for(end - begin, kThreads)), dim3(kThreads),
                     0, stream, J, begin, end, end_state, end_state_at);
  return hipGetLastError();
}

hipError_t launch_writer_startup_check(const JobDev& J, uint32_t final_state,
                                       const uint32_t* final_at, hipStream_t stream) {
  hipLaunchKernelGGL(k_writer_startup_check, dim3(blocks_for(6L * J.width, kThreads)),
                     dim3(kThreads), 0, stream, J, final_state, final_at);
  return hipGetLastError();
}

hipError_t launch_writer_startup_chain(const JobDev* d_jobs, int n, hipStream_t stream) {
  hipLaunchKernelGGL(k_writer_startup_chain, dim3(n), dim3(64), 0, stream, d_jobs);
  return hipGetLastError();
}

hipError_t launch_writer_chain(const JobDev* d_jobs, int n, int band, int bands,
                               hipStream_t stream) {
  hipLaunchKernelGGL(k_writer_chain, dim3(n), dim3(64), 0, stream, d_jobs, band, bands);
  return hipGetLastError();
}

hipError_t launch_fill_u8(uint8_t* p, size_t n, uint8_t value, hipStream_t stream) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u8, dim3(std::min<size_t>(blocks_for((long)n, kThreads), 4096)),
                     dim3(kThreads), 0, stream, p, n, value);
  return hipGetLastError();
}

hipError_t launch_t_unscaled_averages(const float* patches, long count, double* out,
                                      hipStream_t s) {
  hipLaunchKernelGGL(t_unscaled_averages, dim3(blocks_for(count, 128)), dim3(128), 0, s,
                     patches, count, out);
  return hipGetLastError();
}

hipError_t launch_t_rgb_unscaled_averages(const float* patches, long count, double* out,
                                          hipStream_t s) {
  hipLaunchKernelGGL(t_rgb_unscaled_averages, dim3(blocks_for(count, 128)), dim3(128), 0,
                     s, patches, count, out);
  return hipGetLastError();
}

hipError_t launch_t_dither_chain(const double* values, const float* scales, long count,
                                 uint32_t state_in, float low, float high, double* deltas,
                                 uint32_t* state_out, unsigned long long* advances,
                                 hipStream_t s) {
  hipLaunchKernelGGL(t_dither_chain, dim3(1), dim3(1), 0, s, values, scales, count,
                     state_in, low, high, deltas, state_out, advances);
  return hipGetLastError();
}

hipError_t launch_t_emit(const float* values, long count, uint16_t* out,
                         const uint32_t* high, const uint32_t* low, uint32_t* flags,
                         hipStream_t s) {
  hipLaunchKernelGGL(t_emit, dim3(blocks_for(count, kThreads)), dim3(kThreads), 0, s,
                     values, count, out, high, low, flags);
  return hipGetLastError();
}

hipError_t launch_t_planes(const uint16_t* rgbi, int height, int width, const float* lut,
                           const DevParams& P, const float* alpha, const uint8_t* is_one,
                           const float* replacement, const float* offset,
                           const float* score_base, float* working, float* aux,
                           float* score, float* waux, float* wrgb, uint32_t* flags,
                           hipStream_t s) {
  const long pixels = (long)height * width;
  hipLaunchKernelGGL(t_planes_aux, dim3(blocks_for(pixels, kThreads)), dim3(kThreads), 0,
                     s, rgbi, pixels, width, lut, P, alpha, is_one, replacement, offset,
                     working, aux, flags);
  hipLaunchKernelGGL(t_planes_score, dim3(blocks_for(pixels, kThreads)), dim3(kThreads),
                     0, s, pixels, width, P, score_base, working, aux, score, waux, wrgb);
  return hipGetLastError();
}

hipError_t launch_t_decision(const float* aux, int height, int width, const DevParams& P,
                             uint8_t* fallback, hipStream_t s) {
  hipLaunchKernelGGL(t_decision, dim3(blocks_for((long)height * width, kThreads)),
                     dim3(kThreads), 0, s, aux, height, width, P, fallback);
  return hipGetLastError();
}

}  // namespace dice
