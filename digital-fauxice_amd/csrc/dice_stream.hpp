// dice_stream.hpp -- one frame replayed while its main pass arrives as row
// blocks (contracts.py:108-158, 224-261).
//
// Every device phase of the whole-frame runtime works on a window of the
// job (JobDev::Window); a push advances the window to everything whose
// inputs are complete and runs it:
//
//   producer block b        rows [8b, 8b+8) received (the last block may be short)
//   row-sum unit u          producer blocks 2u, 2u+1 done
//   cumulative chain        rows with row sums; epochs follow the rows
//   stage entry i           chain through row min(i, H-1)
//   analysis tile row t     rows through min(H, t*TY + TY + 5) received and
//                           stage entries through min(t*TY + TY + 5, H + 5)
//   sites, draw schedules   the new tile rows (the segment scan continues
//                           from the sites of the rows before)
//   hidden startup          with the first tile row
//   writer chain, values    the new sites, in raster order, on the host chain
//   output copy, SHA-256    rows [0, min(H, TY * tile rows done))
//
// A window opens when a push completes a tile row (or the frame); producer
// work alone waits for it, since nothing downstream could use it sooner.
// Each push ends with the stream idle, so the caller's block buffer and the
// descriptors are free again when push returns.
#pragma once

#include <memory>

#include "dice_runtime.hpp"

namespace dice {

class Sha256;

class FrameStream {
 public:
  FrameStream(int device, const DeviceTables* tables);
  ~FrameStream();
  FrameStream(const FrameStream&) = delete;
  FrameStream& operator=(const FrameStream&) = delete;
  int32_t init();
  // Starts a frame (R.main is not read; the rows come through push).
  int32_t begin(const JobRequest& R);
  int32_t push(const uint16_t* rgbi, int64_t start_row, int64_t rows);
  // Completes the result of a fully received frame; returns its status.
  int32_t finish();
  // Drops a frame in progress (the stream is idle between calls).
  void abandon() { active_ = false; }
  const JobRequest& request() const { return R_; }
  const dice_stream_progress& progress() const { return progress_; }

 private:
  int32_t advance();
  int32_t fail(int32_t rc);
  int32_t upload_job();

  int device_;
  Runner runner_;
  JobRequest R_;
  bool active_ = false;
  std::unique_ptr<Sha256> sha_;
  int64_t rows_in_ = 0, rows_out_ = 0;
  int32_t blocks_done_ = 0, rowsums_done_ = 0, chain_done_ = 0, stages_done_ = 0;
  int32_t trows_done_ = 0;
  uint32_t sites_done_ = 0;
  bool startup_done_ = false;
  uint32_t state_ = 0;
  uint64_t advances_ = 0;
  dice_stream_progress progress_{};
};

}  // namespace dice
