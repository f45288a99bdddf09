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

// The writer's draw schedule of one site in integer form, one 64-bit word
// (two 32-bit halves) per channel.  A channel that draws takes the LCG step
// s -> t and keeps t (the value is decided by this draw) when
// ((t - a) mod 2^24) < len, and otherwise redraws once more.  The keep set
// is an interval of the 2^24 states, possibly wrapping: the writer's own
// predicate is an interval of the dither variable u (t, or its mirror
// 2^24-1-t when the dither span is negative, so that the delta is
// non-decreasing in u), or its complement, and channel_word maps it onto t.
// A channel without the floor gate always keeps t (len = 2^24); one that
// can never satisfy the floor test always redraws (len = 0).
// Low half: bits 0-23 a, bit 31 (channel 0 only) the site is attempted
// (always for main-frame sites).  High half: bits 0-24 len, 27 the channel
// draws.
struct WriterRecord {
  uint64_t ch[3];
};
static_assert(sizeof(WriterRecord) == 24, "WriterRecord must stay 24 bytes");

constexpr uint32_t kChanLen = 0x1FFFFFFu;
constexpr uint32_t kChanDraws = 1u << 27;
constexpr uint64_t kRecordAttemptedBit = 1ull << 31;
constexpr uint32_t kStates = 1u << 24;

// keep <=> (u in [a, a + len)) != inv, with u = t or, mirrored, 2^24-1-t.
DICE_HD uint64_t channel_word(uint32_t a, uint32_t len, bool inv, bool mirror) {
  if (inv) {  // the complement [a + len, a + 2^24) mod 2^24
    a = (a + len) & kLcgMask;
    len = kStates - len;
  }
  if (mirror) a = (kStates - a - len) & kLcgMask;  // u in [a, a+len) <=> t in [2^24-a-len, ...)
  return (uint64_t)(a & kLcgMask) | ((uint64_t)(len | kChanDraws) << 32);
}

DICE_HD bool keeps(uint32_t t, uint32_t a, uint32_t len) { return ((t - a) & kLcgMask) < len; }

// One channel of the chain on the two halves of its word: the next state,
// and the draws it took (0, 1 or 2).
DICE_HD uint32_t chain_channel(uint32_t low, uint32_t high, uint32_t s, uint32_t& draws) {
  const uint32_t a = low & kLcgMask;
  const uint32_t len = high & kChanLen;
  const bool drawn = (high & kChanDraws) != 0;
  const uint32_t t = (125u * s + 1u) & kLcgMask;
  const uint32_t t2 = (125u * t + 1u) & kLcgMask;
  const bool keep = keeps(t, a, len);
  draws = drawn ? (keep ? 1u : 2u) : 0u;
  return drawn ? (keep ? t : t2) : s;
}

// The no-op emit of raw value r is r + d - kNoopDeltaBias with the 4-bit d
// stored as nibble r & 7 of word r >> 3 of the delta table.
constexpr int kNoopDeltaBias = 6;

constexpr uint32_t kSitePixelMask = 0x7FFFFFFFu;
constexpr uint32_t kSiteFloorBit = 0x80000000u;

// Written/changed counters are spread over slots 128 B apart so that the
// per-workgroup atomics do not all hit one address.
constexpr int kCountSlots = 64;
constexpr int kCountStride = 16;  // u64 per slot

// Number of u in [0, 2^24) on the true prefix of a monotone predicate
// (the two ends first: most predicates hold everywhere or nowhere).
template <class F>
DICE_HD uint32_t prefix_count(const F& holds) {
  if (!holds(0u)) return 0;
  if (holds(kLcgMask)) return 1u << 24;
  uint32_t lo = 1, hi = kLcgMask;
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (holds(mid))
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// prefix_count started from an estimate of the count: it gallops from the
// estimate to a bracket and bisects the bracket.  For a monotone predicate
// the result is prefix_count's; an estimate within k states of the count
// takes about 2 log2(k) + 4 evaluations instead of 26.  A non-finite
// estimate bisects the whole range.
template <class F>
DICE_HD uint32_t prefix_count_near(const F& holds, double estimate) {
  if (!(estimate == estimate) || estimate - estimate != 0.0) return prefix_count(holds);
  if (!holds(0u)) return 0;
  if (holds(kLcgMask)) return 1u << 24;
  // the count c lies in [1, 2^24 - 1]: holds(u) <=> u < c
  const double top = (double)(kLcgMask - 1);
  const uint32_t g = estimate <= 1.0 ? 1u : (estimate >= top ? kLcgMask - 1 : (uint32_t)estimate);
  uint32_t lo, hi;
  if (holds(g)) {  // c > g
    lo = g + 1;
    hi = kLcgMask;
    for (uint32_t step = 1;; step <<= 1) {
      const uint32_t p = g + step;
      if (p >= kLcgMask) break;
      if (!holds(p)) {
        hi = p;
        break;
      }
      lo = p + 1;
    }
  } else {  // c <= g
    lo = 1;
    hi = g;
    for (uint32_t step = 1; step < g; step <<= 1) {
      const uint32_t p = g - step;
      if (holds(p)) {
        lo = p + 1;
        break;
      }
      hi = p;
    }
  }
  while (lo < hi) {
    const uint32_t mid = lo + (hi - lo) / 2;
    if (holds(mid))
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

// The integer draw schedule of one site (reconstruction.py:916-982,
// dither.py:59-104).  Every operation between the LCG state and the two
// comparisons the writer makes -- acceptance low < cand32 + delta < high
// and the original floor cand + delta <= orig -- is a monotone IEEE
// rounding of a monotone function of the (mirrored) state, so each
// comparison holds on a prefix of the state range and binary searches
// recover the exact thresholds.  A non-monotone case (a NaN delta) cannot
// pass silently: the per-site writer re-run from the recorded states
// checks the whole chain.
DICE_HD WriterRecord writer_record(const Site& site, bool attempted, const WriterConsts& K) {
  WriterRecord r;
  r.ch[0] = r.ch[1] = r.ch[2] = 0ull;
  if (!attempted) return r;
  const bool valid = finite64(site.cand[0]) && finite64(site.cand[1]) &&
                     finite64(site.cand[2]) && site.cand[0] > 0.0 && site.cand[1] > 0.0 &&
                     site.cand[2] > 0.0;
  if (valid) {
    const bool floor = (site.pixel & kSiteFloorBit) != 0;
    for (int c = 0; c < 3; ++c) {
      const double cd = (double)(float)site.cand[c];
      if (!(K.low_lt_high && K.low < cd && cd < K.high)) continue;
      if (!floor) {  // one draw, whatever its value
        r.ch[c] = channel_word(0u, 1u << 24, false, false);
        continue;
      }
      const double width = K.high - K.low;
      const double coefficient = 4.0 / (width * width);
      const double envelope = ((K.high - cd) * (cd - K.low)) * coefficient;
      const double span = (double)(float)((double)K.scales[c] * cd);
      const bool mirrored = span < 0.0;
      auto delta = [&](uint32_t u) {
        const uint32_t state = mirrored ? kLcgMask - u : u;
        const double centered = (double)(state + 1u) * kNikonNormalization - 0.5;
        const double random_value = 0.0 + centered * span;
        return envelope * random_value;
      };
      const double cand = site.cand[c];
      const double source = (double)site.orig[c];
      // where x + delta(u) crosses t in real arithmetic: the estimate each
      // threshold search starts from (only the search's path depends on
      // it, so one reciprocal serves the three)
      const double per_delta = 1.0 / (envelope * span);
      auto crossing = [&](double x, double t) {
        const double state = ((t - x) * per_delta + 0.5) * (1.0 / kNikonNormalization) - 1.0;
        return mirrored ? (double)kLcgMask - state : state;
      };
      // acceptance low < cd + delta < high holds for u in [lo, hi); when
      // accepted, the floor test cand + delta <= orig holds for u < fe; a
      // rejected draw (delta = 0) keeps the value when cand <= orig
      const uint32_t lo = prefix_count_near([&](uint32_t u) { return cd + delta(u) <= K.low; },
                                            crossing(cd, K.low));
      const uint32_t hi = prefix_count_near([&](uint32_t u) { return cd + delta(u) < K.high; },
                                            crossing(cd, K.high));
      const uint32_t fe = prefix_count_near([&](uint32_t u) { return cand + delta(u) <= source; },
                                            crossing(cand, source));
      const bool keep_when_rejected = cand + 0.0 <= source;
      if (!keep_when_rejected) {
        // keep <=> lo <= u < min(hi, fe)
        const uint32_t end = hi < fe ? hi : fe;
        r.ch[c] = channel_word(lo, end > lo ? end - lo : 0u, false, mirrored);
      } else {
        // redraw <=> max(lo, fe) <= u < hi
        const uint32_t begin = lo > fe ? lo : fe;
        r.ch[c] = channel_word(begin, hi > begin ? hi - begin : 0u, true, mirrored);
      }
    }
  }
  r.ch[0] |= kRecordAttemptedBit;
  return r;
}

DICE_HD uint32_t lcg_step(uint32_t s) { return (125u * s + 1u) & kLcgMask; }

// The serial part of the writer on one record, branch-free: per drawing
// channel one LCG step, and a second one unless the draw is kept.  The LCG
// has no fixed point (124 s = -1 mod 2^24 has no solution), so every draw
// is one counted advance.
DICE_HD uint32_t chain_site(const WriterRecord& r, uint32_t s, uint64_t& draws) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    uint32_t d;
    s = chain_channel((uint32_t)r.ch[c], (uint32_t)(r.ch[c] >> 32), s, d);
    draws += d;
  }
  return s;
}

// k LCG steps as one affine map: lcg^k(s) = (lcg_pow(k) * s + lcg_add(k)) mod 2^24
// (evaluated at compile time).
constexpr uint32_t lcg_pow(int k) { return k == 0 ? 1u : (lcg_pow(k - 1) * 125u) & kLcgMask; }
constexpr uint32_t lcg_add(int k) { return k == 0 ? 0u : (lcg_add(k - 1) * 125u + 1u) & kLcgMask; }

DICE_HD bool draws_all_channels(const WriterRecord& r) {
  return ((r.ch[0] & r.ch[1] & r.ch[2]) >> 32) & kChanDraws;
}

// chain_site for a record whose three channels all draw (nearly every
// site), with a short dependency chain: the six states the site can reach,
// lcg^1..6(s), come straight from s; channel 1 tests t1, channel 2 tests t2
// or t3, channel 3 one of t3..t5, and the site ends on t3..t6 -- every test
// is evaluated at once and the outcome selected.
DICE_HD uint32_t chain_site_drawing(const WriterRecord& r, uint32_t s, uint64_t& draws) {
  constexpr uint32_t p1 = lcg_pow(1), p2 = lcg_pow(2), p3 = lcg_pow(3);
  constexpr uint32_t p4 = lcg_pow(4), p5 = lcg_pow(5), p6 = lcg_pow(6);
  constexpr uint32_t c1_ = lcg_add(1), c2_ = lcg_add(2), c3_ = lcg_add(3);
  constexpr uint32_t c4_ = lcg_add(4), c5_ = lcg_add(5), c6_ = lcg_add(6);
  const uint32_t t[7] = {s,
                         (p1 * s + c1_) & kLcgMask,
                         (p2 * s + c2_) & kLcgMask,
                         (p3 * s + c3_) & kLcgMask,
                         (p4 * s + c4_) & kLcgMask,
                         (p5 * s + c5_) & kLcgMask,
                         (p6 * s + c6_) & kLcgMask};
  const uint32_t a0 = (uint32_t)r.ch[0] & kLcgMask, l0 = (uint32_t)(r.ch[0] >> 32) & kChanLen;
  const uint32_t a1 = (uint32_t)r.ch[1] & kLcgMask, l1 = (uint32_t)(r.ch[1] >> 32) & kChanLen;
  const uint32_t a2 = (uint32_t)r.ch[2] & kLcgMask, l2 = (uint32_t)(r.ch[2] >> 32) & kChanLen;
  const bool c1 = keeps(t[1], a0, l0);
  const bool m22 = keeps(t[2], a1, l1), m23 = keeps(t[3], a1, l1);
  const bool m33 = keeps(t[3], a2, l2), m34 = keeps(t[4], a2, l2), m35 = keeps(t[5], a2, l2);
  const bool c2 = c1 ? m22 : m23;  // channel 2 draws t2 after a kept t1, else t3
  const bool c3 = c1 ? (c2 ? m33 : m34) : (c2 ? m34 : m35);
  const uint32_t kept = c1 ? (c2 ? t[3] : t[4]) : (c2 ? t[4] : t[5]);
  const uint32_t redrawn = c1 ? (c2 ? t[4] : t[5]) : (c2 ? t[5] : t[6]);
  draws += 6u - c1 - c2 - c3;
  return c3 ? kept : redrawn;
}

// chain_site_drawing spread over the lanes of one wavefront, as the device
// chain runs it.  Lane i < 6 tests channel spec_channel(i) on the state
// spec_steps(i) LCG steps on (one 24-bit multiply-add from s): channel 1 on
// t1, channel 2 on t2 / t3, channel 3 on t3 / t4 / t5; lane 6 carries t6.
// The six keep bits fix the site's draws (spec_draws) and the lane holding
// its end state (spec_end_lane).  Two exact shortcuts keep a lane at four
// operations: the state is never masked (the 24-bit multiply reads only
// its low 24 bits and the test masks the difference), and the masked
// difference, with the draws flag set, is compared with the high word as
// stored, len | kChanDraws (both channels draw, len <= 2^24 < kChanDraws).
constexpr int kSpecLanes = 7;
DICE_HD constexpr int spec_steps(int lane) { return lane < 3 ? lane + 1 : lane; }
DICE_HD constexpr int spec_channel(int lane) { return lane == 0 ? 0 : (lane < 3 ? 1 : 2); }
DICE_HD constexpr uint32_t spec_draws(uint32_t keep_bits) {
  const uint32_t d0 = (keep_bits & 1u) ? 1u : 2u;
  const uint32_t d1 = ((keep_bits >> d0) & 1u) ? 1u : 2u;             // lane 1 or 2
  const uint32_t d2 = ((keep_bits >> (d0 + d1 + 1u)) & 1u) ? 1u : 2u;  // lane 3, 4 or 5
  return d0 + d1 + d2;
}
DICE_HD constexpr int spec_end_lane(uint32_t draws) { return draws == 3u ? 2 : (int)draws; }
DICE_HD bool spec_keep(uint32_t t, uint32_t low, uint32_t high) {
  return (((t - low) & kLcgMask) | kChanDraws) < high;
}

// The lane form evaluated on the host, lane by lane (the low 32 bits of the
// product of the low 24 bits, as the device multiply gives them).
inline uint32_t chain_site_lanes(const WriterRecord& r, uint32_t s, uint64_t& draws) {
  uint32_t t[kSpecLanes], keep_bits = 0;
  for (int i = 0; i < kSpecLanes; ++i) {
    const int n = spec_steps(i);
    t[i] = (lcg_pow(n) & kLcgMask) * (s & kLcgMask) + lcg_add(n);
    const uint64_t w = r.ch[spec_channel(i)];
    if (i < 6 && spec_keep(t[i], (uint32_t)w, (uint32_t)(w >> 32))) keep_bits |= 1u << i;
  }
  const uint32_t d = spec_draws(keep_bits);
  draws += d;
  return t[spec_end_lane(d)] & kLcgMask;
}

// The chain over n consecutive records: each record's start state into
// states[i]; returns the state after the last.
inline uint32_t chain_records(const WriterRecord* r, size_t n, uint32_t s, uint32_t* states,
                              uint64_t& draws) {
  for (size_t i = 0; i < n; ++i) {
    states[i] = s;
    s = draws_all_channels(r[i]) ? chain_site_drawing(r[i], s, draws) : chain_site(r[i], s, draws);
  }
  return s;
}

// The walk keeps the start state of every kStateGroup-th site only
// (group_states[g] = state at site g * kStateGroup); the device re-walks each
// group from it (k_writer_expand).  Sites [first, first + n) of the job's
// main chain; returns the state after the last.
constexpr uint32_t kStateGroup = 16;

inline uint32_t chain_records_grouped(const WriterRecord* r, size_t first, size_t n, uint32_t s,
                                      uint32_t* group_states, uint64_t& draws) {
  for (size_t i = first; i < first + n; ++i) {
    if (i % kStateGroup == 0) group_states[i / kStateGroup] = s;
    s = draws_all_channels(r[i]) ? chain_site_drawing(r[i], s, draws) : chain_site(r[i], s, draws);
  }
  return s;
}

DICE_HD bool record_attempted(const WriterRecord& r) {
  return (r.ch[0] & kRecordAttemptedBit) != 0;
}

// Device error flags (atomicOr): the host maps them to the reference's
// fail-closed exceptions.
constexpr uint32_t kErrNonfiniteAux = 1u;    // continuous_score ValueError
constexpr uint32_t kErrSetterScaleOne = 4u;  // derive_x3a_setter_fields
constexpr uint32_t kErrStageNonfinite = 8u;  // StageCalibration finiteness
constexpr uint32_t kErrNonfiniteWork = 16u;  // work_value_indices ValueError
constexpr uint32_t kErrWriterChain = 32u;    // integer chain disagrees with write_pixel

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

// The writer chain's carried state and receipt when it runs on the device
// (one wavefront): the LCG state between launches, the public advances of
// the main frame and the hidden-startup receipt (startup.py:37-43).
struct ChainState {
  uint32_t state;
  uint32_t startup_final;
  unsigned long long advances;
  long long startup_attempted[6];
  long long startup_advances[6];
};

struct JobDev {
  DevParams P;                  // stable replay parameters of this job
  WriterConsts K;               // dither bounds and scales of this job
  float factor_b;               // auxiliary factor of the setter (stage_parameters.py)
  int32_t run_producer;         // 1: producer + setter on device; 0: stage table given
  int32_t skip;                 // 1: the job already failed; later phases leave it
  int32_t height, width;
  int32_t tiles_x, tiles_y;     // analysis tiles (TY x TX)
  int32_t tile_rows, tile_cols; // TY, TX of the launch
  const uint16_t* rgbi;         // H*W*4
  uint16_t* out;                // H*W*3 (noop emit; patched after the writer)
  const float* lut;             // 65536
  const uint32_t* noop_delta;   // 8192: emit of LUT[i], the output of an unchanged value,
                                // as 4-bit deltas from i (kNoopDeltaBias)
  const uint32_t* factor_high;  // 256
  const uint32_t* factor_low;   // 256
  StageDev stages;
  uint16_t* site_lists;         // tiles * TY*TX local indices, row-major
  uint32_t* tile_count;         // sites per tile
  uint32_t* seg_count;          // (tiles_y*TY) * tiles_x
  uint32_t* seg_offset;         // (tiles_y*TY) * tiles_x
  uint32_t* seg_block_sums;     // one per 1024 segments
  uint32_t* row_start;          // H+1: first site of each row; [H] = total
  Site* sites;                  // compacted, row-major
  WriterRecord* writer_records; // one per site, row-major
  WriterRecord* su_writer_records; // 6*W
  uint32_t* site_state;         // LCG state at the start of each site
  uint32_t* group_state;        // the same at every kStateGroup-th site (host walk)
  uint32_t* su_site_state;      // 6*W, the same for the hidden startup rows
  ChainState* chain;            // the device writer chain's state and receipt
  uint32_t* band_end;           // LCG state after each writer band
  unsigned long long* writer_counts; // kCountSlots slots of {written, changed}
  uint8_t* diag_changed;        // optional H*W
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
  double* epoch_adds;           // epochs * 2 (den, num)
  double* cumulative;           // H * 3
  double* epoch_wide;           // epochs * 2 (den_w, num_w)
  float* epoch_store;           // epochs * 2 (den32, num32)
  // The part of the job the device phases cover.  A whole-frame run covers
  // everything; a frame streamed in as row blocks advances the window as its
  // rows arrive (every phase only reads rows, records and stages that are
  // complete when it runs).
  struct Window {
    int32_t block_lo, block_n;    // 8-row producer blocks
    int32_t rowsum_lo, rowsum_n;  // row-sum units of kRowSumRows rows
    int32_t chain_lo, chain_hi;   // rows of the cumulative chain (epochs follow)
    int32_t stage_lo, stage_hi;   // stage-table entries
    int32_t trow_lo, trow_n;      // analysis tile rows
    uint32_t site_lo;             // first main site of the draw schedules
    int32_t startup;              // 1: the hidden startup runs in this window
  } win;
};

// Rows per k_row_sums work unit.
constexpr int kRowSumRows = 16;

// The window of a whole-frame run.
inline JobDev::Window full_window(int32_t height, int32_t tiles_y) {
  JobDev::Window w{};
  w.block_lo = 0;
  w.block_n = (height + 7) / 8;
  w.rowsum_lo = 0;
  w.rowsum_n = (height + kRowSumRows - 1) / kRowSumRows;
  w.chain_lo = 0;
  w.chain_hi = height;
  w.stage_lo = 0;
  w.stage_hi = height + 6;
  w.trow_lo = 0;
  w.trow_n = tiles_y;
  w.site_lo = 0;
  w.startup = 1;
  return w;
}

// A batch of jobs for one launch: work unit u of the launch belongs to the
// job j with base[j] <= u < base[j+1] (empty ranges allowed), and is unit
// u - base[j] of that job.
constexpr int kMaxBatch = 64;

struct JobSet {
  const JobDev* jobs;  // device array of `count` descriptors
  int32_t count;
  uint32_t base[kMaxBatch + 1];
  // {next ticket, finished workgroups} of a dynamically scheduled launch:
  // zero before the launch, and zero again after it (the last workgroup
  // resets them); one pair per stream that launches such kernels
  uint32_t* queue;
};

}  // namespace dice
