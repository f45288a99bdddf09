// dice_stream.cpp -- FrameStream: the whole-frame device phases run window
// by window as the main pass arrives in row blocks.
#include "dice_stream.hpp"

#include <algorithm>
#include <cstring>

#include "dice_kernels.hpp"
#include "dice_workspace.hpp"

#pragma clang fp contract(off)

namespace dice {

FrameStream::FrameStream(int device, const DeviceTables* tables)
    : device_(device), runner_(device, tables) {}

FrameStream::~FrameStream() = default;

int32_t FrameStream::init() { return runner_.init(); }

// Records the first failure of the frame; every later call reports it.
int32_t FrameStream::fail(int32_t rc) {
  if (rc != DICE_OK && R_.code == DICE_OK) {
    R_.code = rc;
    R_.error = error_buffer();
  }
  return rc;
}

int32_t FrameStream::upload_job() {
  HIP_TRY(hipMemcpyAsync(runner_.d_jobs_.ptr, runner_.h_jobs_.ptr, sizeof(JobDev),
                         hipMemcpyHostToDevice, runner_.stream_));
  return DICE_OK;
}

int32_t FrameStream::begin(const JobRequest& R) {
  HIP_TRY(hipSetDevice(device_));
  active_ = false;  // a frame in progress is abandoned
  if (runner_.stream_) DICE_TRY(runner_.wait_stream(runner_.stream_));
  if (R.flags & DICE_MAIN_ON_DEVICE)
    return set_error(DICE_INVALID, "a streamed frame arrives from the host, block by block");
  R_ = R;
  R_.main = nullptr;
  R_.stages = nullptr;
  R_.streamed = true;
  R_.code = DICE_OK;
  R_.error.clear();
  memset(R_.result, 0, sizeof *R_.result);
  rows_in_ = rows_out_ = 0;
  blocks_done_ = rowsums_done_ = chain_done_ = stages_done_ = trows_done_ = 0;
  sites_done_ = 0;
  startup_done_ = false;
  state_ = R_.rng_state;
  advances_ = 0;
  progress_ = dice_stream_progress{};
  sha_ = std::make_unique<Sha256>();
  DICE_TRY(runner_.prepare(&R_, 1, nullptr));
  if (R_.code != DICE_OK) return set_error(R_.code, "%s", R_.error.c_str());
  JobDev& J = runner_.h_jobs_.as<JobDev>()[0];
  J.win = JobDev::Window{};  // nothing has arrived
  const hipStream_t s = runner_.stream_;
  DICE_TRY(upload_job());
  HIP_TRY(launch_zero_jobs(runner_.d_jobs_.as<JobDev>(), &J, 1, s));
  if (J.diag_changed)
    HIP_TRY(launch_fill_u8(J.diag_changed, (size_t)R_.height * R_.width, 0, s));
  DICE_TRY(runner_.wait_stream(s));
  active_ = true;
  return DICE_OK;
}

int32_t FrameStream::push(const uint16_t* rgbi, int64_t start_row, int64_t rows) {
  if (!active_) return set_error(DICE_INVALID, "the stream has no frame in progress");
  if (R_.code != DICE_OK) return set_error(R_.code, "%s", R_.error.c_str());
  HIP_TRY(hipSetDevice(device_));
  const double t0 = now_seconds();
  if (!rgbi || rows <= 0) return set_error(DICE_INVALID, "an RGBI block needs at least one row");
  if (start_row != rows_in_)
    return set_error(DICE_INVALID,
                     "RGBI blocks must be ordered, gap-free, and non-overlapping (expected "
                     "start row %lld, got %lld)",
                     (long long)rows_in_, (long long)start_row);
  if (start_row + rows > R_.height)
    return set_error(DICE_INVALID, "RGBI block rows [%lld, %lld) run past the frame height %lld",
                     (long long)start_row, (long long)(start_row + rows), (long long)R_.height);
  Workspace& w = *runner_.ws_[0];
  const size_t row_bytes = (size_t)R_.width * 8;
  auto body = [&]() -> int32_t {
    HIP_TRY(hipMemcpyAsync(w.rgbi.as<uint8_t>() + (size_t)start_row * row_bytes, rgbi,
                           (size_t)rows * row_bytes, hipMemcpyHostToDevice, runner_.stream_));
    rows_in_ += rows;
    return advance();
  };
  int32_t rc = fail(body());
  // the caller's block and the descriptors are free again once the stream is idle
  const int32_t idle = runner_.wait_stream(runner_.stream_);
  if (rc == DICE_OK) rc = fail(idle);
  progress_.rows_received = rows_in_;
  progress_.push_seconds = now_seconds() - t0;
  return rc;
}

int32_t FrameStream::advance() {
  JobDev& J = runner_.h_jobs_.as<JobDev>()[0];
  Workspace& w = *runner_.ws_[0];
  const hipStream_t s = runner_.stream_;
  const JobDev* d = runner_.d_jobs_.as<JobDev>();
  const int32_t H = J.height, W = J.width, TY = J.tile_rows;
  const int64_t received = rows_in_;

  // ---- the window: everything whose inputs are complete
  const bool all_rows = received == H;
  const int32_t blocks = all_rows ? (H + 7) / 8 : (int32_t)(received / 8);
  const int32_t blocked_rows = std::min(H, 8 * blocks);
  const int32_t rowsums =
      blocked_rows == H ? (H + kRowSumRows - 1) / kRowSumRows : blocked_rows / kRowSumRows;
  const int32_t chained = blocked_rows == H ? H : rowsums * kRowSumRows;
  const int32_t staged = chained == H ? H + 6 : chained;
  int32_t trows = trows_done_;
  while (trows < J.tiles_y) {
    const int64_t y_end = (int64_t)trows * TY + TY;  // one past the tile row's last row
    const int64_t need_rows = std::min<int64_t>(H, y_end + 5);
    const int64_t need_stages = std::min<int64_t>(y_end + 6, (int64_t)H + 6);
    if (received < need_rows || staged < need_stages) break;
    ++trows;
  }
  // Producer work alone waits for the next window that completes a tile row
  // (or the frame): nothing downstream could use it sooner.
  if (trows == trows_done_ && !all_rows) return DICE_OK;
  if (blocks == blocks_done_ && trows == trows_done_ && rowsums == rowsums_done_) return DICE_OK;

  JobDev::Window& win = J.win;
  win.block_lo = blocks_done_;
  win.block_n = blocks - blocks_done_;
  win.rowsum_lo = rowsums_done_;
  win.rowsum_n = rowsums - rowsums_done_;
  win.chain_lo = chain_done_;
  win.chain_hi = chained;
  win.stage_lo = stages_done_;
  win.stage_hi = staged;
  win.trow_lo = trows_done_;
  win.trow_n = trows - trows_done_;
  win.site_lo = sites_done_;
  win.startup = !startup_done_ && win.trow_n > 0 ? 1 : 0;
  DICE_TRY(upload_job());
  HIP_TRY(launch_producer(d, &J, 1, s));
  progress_.windows += 1;
  blocks_done_ = blocks;
  rowsums_done_ = rowsums;
  chain_done_ = chained;
  stages_done_ = staged;
  if (win.trow_n == 0) return DICE_OK;

  // ---- analysis of the new tile rows and their site count
  HIP_TRY(launch_scan(d, &J, 1, s));
  HIP_TRY(launch_scan_segments(d, &J, 1, s));
  uint32_t* small = w.h_small.as<uint32_t>();
  HIP_TRY(hipMemcpyAsync(small, w.counters.ptr, 64, hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(small + 16, w.flags.ptr, 4, hipMemcpyDeviceToHost, s));
  DICE_TRY(runner_.wait_stream(s));
  DICE_TRY(flag_error(small[16]));
  const uint32_t count = small[0];
  const uint32_t fresh = count - sites_done_;
  const size_t su = 6 * (size_t)W;

  // ---- sites, hidden startup, draw schedules.  Sites of earlier windows
  // are finished, so the site arrays may be replaced when they grow.
  HIP_TRY(w.sites.ensure(((size_t)count + 1) * sizeof(Site), s));
  HIP_TRY(w.records.ensure(((size_t)count + 1) * sizeof(WriterRecord), s));
  HIP_TRY(w.states.ensure((su + count + 1) * 4, s));
  HIP_TRY(w.h_records.ensure((su + fresh + 1) * sizeof(WriterRecord)));
  HIP_TRY(w.h_states.ensure((su + fresh + 1) * 4));
  J.sites = w.sites.as<Site>();
  J.writer_records = w.records.as<WriterRecord>();
  J.su_site_state = w.states.as<uint32_t>();
  J.site_state = J.su_site_state + su;
  DICE_TRY(upload_job());
  HIP_TRY(launch_sites(d, &J, 1, runner_.queue_.as<uint32_t>(), s, runner_.sites_variant_));
  if (win.startup) HIP_TRY(launch_startup(d, &J, 1, s));
  HIP_TRY(launch_writer_prep(d, 1, s));
  WriterRecord* records = w.h_records.as<WriterRecord>();
  uint32_t* states = w.h_states.as<uint32_t>();
  if (win.startup)
    HIP_TRY(hipMemcpyAsync(records, J.su_writer_records, su * sizeof(WriterRecord),
                           hipMemcpyDeviceToHost, s));
  if (fresh)
    HIP_TRY(hipMemcpyAsync(records + su, J.writer_records + sites_done_,
                           (size_t)fresh * sizeof(WriterRecord), hipMemcpyDeviceToHost, s));
  HIP_TRY(hipMemcpyAsync(small + 16, w.flags.ptr, 4, hipMemcpyDeviceToHost, s));
  DICE_TRY(runner_.wait_stream(s));
  DICE_TRY(flag_error(small[16]));

  // ---- the serial chain over the new sites, then their values
  if (win.startup) {
    state_ = walk_startup(records, W, state_, states, R_.result);
    HIP_TRY(hipMemcpyAsync(J.su_site_state, states, su * 4, hipMemcpyHostToDevice, s));
    HIP_TRY(launch_writer_startup_check(J, state_, nullptr, s));
    startup_done_ = true;
  }
  if (fresh) {
    state_ = chain_records(records + su, fresh, state_, states + su, advances_);
    HIP_TRY(hipMemcpyAsync(J.site_state + sites_done_, states + su, (size_t)fresh * 4,
                           hipMemcpyHostToDevice, s));
    HIP_TRY(launch_writer_values(J, sites_done_, count, state_, nullptr, s));
  }
  sites_done_ = count;
  trows_done_ = trows;
  progress_.rows_analyzed = std::min<int64_t>(H, (int64_t)trows * TY);
  progress_.sites_written = count;

  // ---- output rows that are final: to the caller and the hash
  const int64_t final_rows = std::min<int64_t>(H, (int64_t)trows * TY);
  const bool want_sha = (R_.flags & DICE_SKIP_SHA256) == 0;
  const bool host_out = (R_.flags & DICE_OUT_ON_DEVICE) == 0;
  if (final_rows > rows_out_ && (want_sha || host_out)) {
    const size_t row_bytes = (size_t)W * 6;
    const size_t begin = (size_t)rows_out_ * row_bytes;
    const size_t bytes = (size_t)(final_rows - rows_out_) * row_bytes;
    HIP_TRY(w.h_out.ensure(bytes));
    HIP_TRY(hipMemcpyAsync(w.h_out.ptr, reinterpret_cast<const uint8_t*>(J.out) + begin, bytes,
                           hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(small + 16, w.flags.ptr, 4, hipMemcpyDeviceToHost, s));
    DICE_TRY(runner_.wait_stream(s));
    DICE_TRY(flag_error(small[16]));
    if (want_sha) sha_->update(w.h_out.ptr, bytes);
    if (host_out) memcpy(reinterpret_cast<uint8_t*>(R_.out) + begin, w.h_out.ptr, bytes);
  }
  rows_out_ = std::max<int64_t>(rows_out_, final_rows);
  progress_.rows_final = rows_out_;
  return DICE_OK;
}

int32_t FrameStream::finish() {
  if (!active_) return set_error(DICE_INVALID, "the stream has no frame in progress");
  active_ = false;
  if (R_.code != DICE_OK) return set_error(R_.code, "%s", R_.error.c_str());
  HIP_TRY(hipSetDevice(device_));
  if (rows_in_ != R_.height)
    return fail(set_error(DICE_INVALID, "the frame is incomplete: %lld of %lld rows received",
                          (long long)rows_in_, (long long)R_.height));
  auto body = [&]() -> int32_t {
    const JobDev& J = runner_.h_jobs_.as<JobDev>()[0];
    Workspace& w = *runner_.ws_[0];
    const hipStream_t s = runner_.stream_;
    uint8_t* small = w.h_small.as<uint8_t>();
    HIP_TRY(hipMemcpyAsync(small + 128, J.writer_counts, kCountSlots * kCountStride * 8,
                           hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(small + 64, J.error_flags, 4, hipMemcpyDeviceToHost, s));
    const size_t pixels = (size_t)R_.height * R_.width;
    if (R_.diag) {
      if (R_.diag->score_plane)
        HIP_TRY(hipMemcpyAsync(R_.diag->score_plane, J.diag_score, pixels * 4,
                               hipMemcpyDeviceToHost, s));
      if (R_.diag->at_floor_mask)
        HIP_TRY(hipMemcpyAsync(R_.diag->at_floor_mask, J.diag_floor, pixels,
                               hipMemcpyDeviceToHost, s));
      if (R_.diag->changed_mask)
        HIP_TRY(hipMemcpyAsync(R_.diag->changed_mask, J.diag_changed, pixels,
                               hipMemcpyDeviceToHost, s));
    }
    DICE_TRY(runner_.wait_stream(s));
    DICE_TRY(flag_error(*reinterpret_cast<const uint32_t*>(small + 64)));
    const unsigned long long* slots = reinterpret_cast<const unsigned long long*>(small + 128);
    uint64_t written = 0, changed = 0;
    for (int slot = 0; slot < kCountSlots; ++slot) {
      written += slots[slot * kCountStride + 0];
      changed += slots[slot * kCountStride + 1];
    }
    dice_replay_result& res = *R_.result;
    res.written_pixels = written;
    res.changed_pixels = changed;
    res.attempted_pixels = sites_done_;
    res.public_rng_advances = advances_;
    res.final_rng_state = state_;
    if ((R_.flags & DICE_SKIP_SHA256) == 0 && !sha_->hex(res.output_sha256))
      return set_error(DICE_INTERNAL, "SHA-256 failed");
    return DICE_OK;
  };
  return fail(body());
}

}  // namespace dice
