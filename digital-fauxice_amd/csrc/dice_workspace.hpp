// dice_workspace.hpp -- the device and pinned buffers of one job slot of a
// Runner, shared by the batch runtime and the frame stream.
#pragma once

#include <hip/hip_runtime.h>

#include <condition_variable>
#include <mutex>

#include "dice_common.hpp"
#include "dice_device.hpp"

namespace dice {

constexpr int kMaxBands = 64;

// The status code and message of the first error flag a device phase raised.
inline int32_t flag_error(uint32_t flags) {
  if (flags & kErrNonfiniteAux)
    return set_error(DICE_NONFINITE_AUX, "auxiliary score input must be a finite HxW plane");
  if (flags & kErrSetterScaleOne) return set_error(DICE_INVALID, "X3A setter scale cannot equal one");
  if (flags & kErrStageNonfinite)
    return set_error(DICE_INVALID, "stage calibration contains a non-finite value");
  if (flags & kErrNonfiniteWork) return set_error(DICE_NONFINITE_WORK, "work RGB must be finite");
  if (flags & kErrWriterChain)
    return set_error(DICE_INTERNAL, "writer state chain disagrees with the per-site writer");
  return DICE_OK;
}

// The serial chain over the hidden startup's 6 x W draw schedules
// (startup.py:74-353, host_writer.py:73-93): the LCG state at the start of
// every slot into `states`, the per-stage receipt into `res`; returns the
// state after the last slot.
inline uint32_t walk_startup(const WriterRecord* records, int64_t W, uint32_t state,
                             uint32_t* states, dice_replay_result* res) {
  for (int stage = 0; stage < 6; ++stage) {
    uint64_t draws = 0;
    int64_t attempted = 0;
    for (int64_t x = 0; x < W; ++x) {
      const size_t i = (size_t)stage * W + x;
      states[i] = state;
      attempted += record_attempted(records[i]) ? 1 : 0;
      state = chain_site(records[i], state, draws);
    }
    res->startup_attempted[stage] = attempted;
    res->startup_advances[stage] = (int64_t)draws;
  }
  res->startup_final_rng_state = state;
  return state;
}

// Band b of H rows in `bands` equal parts.
inline int64_t band_row(int64_t H, int bands, int b) { return H * b / bands; }

// --------------------------------------------------------------------------
// Workspace: the device and pinned buffers of one job slot, grown on demand
// and kept for the next batch.
// --------------------------------------------------------------------------

struct Workspace {
  DeviceBuffer rgbi, out, site_lists, tile_count, seg_count, seg_offset, seg_block_sums,
      row_start, counters, flags;
  DeviceBuffer diag_score, diag_floor, diag_changed;
  DeviceBuffer su_aux, su_score, su_work, su_sites, su_attempted;
  DeviceBuffer failpos, row_sums, accepted, epoch_adds, cumulative, epoch_wide, epoch_store;
  DeviceBuffer st_alpha, st_offset, st_replacement, st_has, st_score_base, st_coarse, st_gate,
      st_row_gate;
  DeviceBuffer sites, records, su_records, states, writer_counts;
  DeviceBuffer chain;        // ChainState, then the LCG state after each band
  PinnedBuffer h_chain;      // the same on the host
  PinnedBuffer h_small;      // [0,64) counters, [64,128) flags, [128, ...) writer slots
  PinnedBuffer h_records;    // 6W startup records, then the main records
  PinnedBuffer h_row_start;  // H+1
  PinnedBuffer h_states;     // 6W startup states, then the main states
  PinnedBuffer h_out;        // staging of the output rows on their way to the hash and caller
  hipEvent_t band[kMaxBands] = {};     // writer values of band b done (runner stream)
  hipEvent_t fetched[kMaxBands] = {};  // draw schedules of band b on the host (copy stream)
  hipEvent_t copied[kMaxBands] = {};   // output rows of band b on the host (output stream)
  hipEvent_t startup_fetched = nullptr;
  hipEvent_t tail = nullptr;  // the job's last device work and result copies (runner stream)
  int bands = 1;
  bool events = false;
  // chain -> copy scheduler -> hash hand-off: `published` counts the bands
  // whose output copy is queued, `aborted` that the chain failed
  std::mutex mutex;
  std::condition_variable cv;
  int published = 0;
  bool aborted = false;

  ~Workspace() {
    for (hipEvent_t* set : {band, fetched, copied})
      for (int b = 0; b < kMaxBands; ++b)
        if (set[b]) (void)hipEventDestroy(set[b]);
    if (startup_fetched) (void)hipEventDestroy(startup_fetched);
    if (tail) (void)hipEventDestroy(tail);
  }
  int32_t init_events() {
    if (events) return DICE_OK;
    // worker threads wait on these through wait_event (polling with sleeps:
    // hipEventSynchronize spins even on blocking-sync events), so waiting
    // threads leave the cores to the chains and hashes
    const unsigned flags = hipEventDisableTiming | hipEventBlockingSync;
    for (hipEvent_t* set : {band, fetched, copied})
      for (int b = 0; b < kMaxBands; ++b) HIP_TRY(hipEventCreateWithFlags(&set[b], flags));
    HIP_TRY(hipEventCreateWithFlags(&startup_fetched, flags));
    HIP_TRY(hipEventCreateWithFlags(&tail, flags));
    events = true;
    return DICE_OK;
  }
  uint64_t bytes() const {
    const DeviceBuffer* all[] = {&rgbi, &out, &site_lists, &tile_count, &seg_count, &seg_offset,
                                 &seg_block_sums, &row_start, &counters, &flags, &diag_score,
                                 &diag_floor, &diag_changed, &su_aux, &su_score, &su_work,
                                 &su_sites, &su_attempted, &failpos, &row_sums, &accepted,
                                 &epoch_adds, &cumulative, &epoch_wide, &epoch_store,
                                 &st_alpha, &st_offset, &st_replacement, &st_has,
                                 &st_score_base, &st_coarse, &st_gate, &st_row_gate, &sites,
                                 &records, &su_records, &states, &writer_counts, &chain};
    uint64_t total = 0;
    for (const DeviceBuffer* b : all) total += b->bytes;
    return total;
  }
};

}  // namespace dice
