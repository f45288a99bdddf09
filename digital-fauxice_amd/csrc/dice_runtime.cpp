// dice_runtime.cpp -- batch execution of Digital ICE jobs on one MI355X.
//
// The reference runs one job at a time (engine.py:110-267,
// cuda_backend/engine.py:183-590); here a Runner takes 1..kMaxBatch
// independent jobs and runs every device phase as one persistent launch
// over all of them, so small crops fill the 256 CUs together.  The strictly
// serial writer (host_writer.py:73-93) stays one chain per job, on one host
// core; chains of different jobs run on different cores, and each job's
// SHA-256 runs beside its chain, band by band.
#include "dice_runtime.hpp"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>

#include <sys/stat.h>
#include <unistd.h>

#include "dice_kernels.hpp"
#include "dice_sha.hpp"
#include "dice_workspace.hpp"

#pragma clang fp contract(off)

namespace dice {

namespace {

// run_in_arena's pinned host area: the row starts of up to kArenaMaxRows
// rows, then the chain state, the counters and flags, the writer slots.
constexpr size_t kArenaRowStartBytes = (size_t)(Runner::kArenaMaxRows + 1) * 4;
constexpr size_t kArenaChainAt = (kArenaRowStartBytes + 255) & ~(size_t)255;
constexpr size_t kArenaSmallAt = kArenaChainAt + 256;
constexpr size_t kArenaSlotsAt = kArenaSmallAt + 128;
constexpr size_t kArenaIoBytes = kArenaSlotsAt + (size_t)kCountSlots * kCountStride * 8;

// DICE_TRACE=1 prints one timeline line per batch to stderr.
bool trace_enabled() {
  static const bool on = [] {
    const char* v = getenv("DICE_TRACE");
    return v && v[0] == '1';
  }();
  return on;
}

double trace_origin() {
  static const double origin = now_seconds();
  return origin;
}

// The jobs of a batch that hash may be taken two by two, each pair's
// outputs hashed by one thread, the two messages interleaved: about 40 ms of
// a core per full frame instead of 56-62 for one message on its own, at a
// frame's hash latency of about 80 ms instead of 60.  That pays only when
// the hashes outnumber the host's cores, so a batch pairs its jobs when the
// messages being hashed in the process (every runner's), with its own,
// would: a batch alone on an idle host keeps the shorter latency.
// DICE_HASH_PAIRS=0 never pairs, =1 always does (set_hash_pair_mode
// changes it at run time).
std::atomic<int> g_messages_hashing{0};
std::atomic<int> g_pair_mode{-1};  // -1: not read from the environment yet

int hash_pair_mode() {
  int mode = g_pair_mode.load();
  if (mode < 0) {
    const char* v = getenv("DICE_HASH_PAIRS");
    int expected = -1;
    g_pair_mode.compare_exchange_strong(
        expected, v && v[0] == '0' ? kHashPairsNever : (v && v[0] == '1' ? kHashPairsAlways
                                                                          : kHashPairsWhenBusy));
    mode = g_pair_mode.load();
  }
  return mode;
}

// The messages being hashed in the process, with `messages` more, as a
// share of the host's cores (1: one per core).
double hash_load(int messages) {
  static const int cores = std::max(1u, std::thread::hardware_concurrency());
  return (double)(g_messages_hashing.load() + messages) / cores;
}

bool hash_in_pairs(int messages) {
  const int mode = hash_pair_mode();
  if (mode != kHashPairsWhenBusy) return mode == kHashPairsAlways;
  return hash_load(messages) > 1.0;
}
}  // namespace

int set_hash_pair_mode(int mode) {
  const int previous = hash_pair_mode();
  g_pair_mode.store(mode);
  return previous;
}

DevParams device_params(const dice_params& p) {
  DevParams P;
  P.visible_channel = p.visible_channel;
  P.score_addend = p.score_addend;
  P.score_scale = p.score_scale;
  P.score_offset = p.score_offset;
  P.score_floor = p.score_floor;
  P.horizontal_minimum = p.horizontal_minimum;
  P.decision_threshold = p.decision_threshold;
  P.count_limit = p.decision_count_limit;
  P.radius = p.decision_radius;
  P.cross_neighbor = p.cross_neighbor;
  P.gate_primary = p.gate_primary;
  P.combine.coarse_enabled = p.coarse_enabled;
  for (int i = 0; i < 3; ++i) {
    P.combine.coarse_slopes[i] = p.coarse_slopes[i];
    P.combine.band_enabled[i] = p.band_enabled[i];
    P.combine.band_scales[i] = p.band_scales[i];
    P.combine.configured_strengths[i] = p.configured_strengths[i];
  }
  for (int i = 0; i < 9; ++i) {
    P.combine.factors_a[i] = p.factors_a[i];
    P.combine.factors_b[i] = p.factors_b[i];
  }
  return P;
}

WriterConsts writer_consts(const dice_params& p) {
  WriterConsts K;
  K.low = (double)p.dither_low;
  K.high = (double)p.dither_high;
  K.low_lt_high = K.low < K.high;
  K.scales[0] = p.dither_scales[0];
  K.scales[1] = p.dither_scales[1];
  K.scales[2] = p.dither_scales[2];
  return K;
}

int32_t DeviceTables::upload() {
  DICE_TRY(require_tables());
  const Tables& t = tables();
  if (noop_delta.bytes >= 8192 * 4) return DICE_OK;
  HIP_TRY(lut.ensure(65536 * 4));
  HIP_TRY(high.ensure(256 * 4));
  HIP_TRY(low.ensure(256 * 4));
  HIP_TRY(noop_delta.ensure(8192 * 4));
  HIP_TRY(hipMemcpy(lut.ptr, t.lut.data(), 65536 * 4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(high.ptr, t.high.data(), 256 * 4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(low.ptr, t.low.data(), 256 * 4, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(noop_delta.ptr, t.noop_delta.data(), 8192 * 4, hipMemcpyHostToDevice));
  return DICE_OK;
}

// --------------------------------------------------------------------------
// Runner
// --------------------------------------------------------------------------

Runner::Runner(int device, const DeviceTables* tables) : device_(device), tables_(tables) {}

Runner::~Runner() {
  (void)hipSetDevice(device_);
  if (stream_) (void)hipStreamSynchronize(stream_);
  for (Graph* g : {&graph_a_, &graph_b_})
    if (g->exec) (void)hipGraphExecDestroy(g->exec);
  for (hipEvent_t e : ev_)
    if (e) (void)hipEventDestroy(e);
  for (hipEvent_t e : fine_)
    if (e) (void)hipEventDestroy(e);
  if (sync_event_) (void)hipEventDestroy(sync_event_);
  if (fetch_stream_) (void)hipStreamSynchronize(fetch_stream_);
  if (out_stream_) (void)hipStreamSynchronize(out_stream_);
  ws_.clear();
  if (fetch_stream_) (void)hipStreamDestroy(fetch_stream_);
  if (out_stream_) (void)hipStreamDestroy(out_stream_);
  if (stream_) (void)hipStreamDestroy(stream_);
}

int32_t Runner::init() {
  HIP_TRY(hipSetDevice(device_));
  HIP_TRY(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
  HIP_TRY(hipStreamCreateWithFlags(&fetch_stream_, hipStreamNonBlocking));
  HIP_TRY(hipStreamCreateWithFlags(&out_stream_, hipStreamNonBlocking));
  for (hipEvent_t& e : ev_) HIP_TRY(hipEventCreate(&e));
  for (hipEvent_t& e : fine_) HIP_TRY(hipEventCreate(&e));
  HIP_TRY(hipEventCreateWithFlags(&sync_event_, hipEventDisableTiming | hipEventBlockingSync));
  HIP_TRY(d_jobs_.ensure(sizeof(JobDev) * kMaxBatch));
  HIP_TRY(queue_.ensure(256));
  HIP_TRY(hipMemset(queue_.ptr, 0, 256));
  HIP_TRY(h_jobs_.ensure(sizeof(JobDev) * kMaxBatch));
  HIP_TRY(h_io_.ensure(kArenaIoBytes));
  HIP_TRY(digest_.ensure(8 * kMaxBatch));
  hipDeviceProp_t prop;
  HIP_TRY(hipGetDeviceProperties(&prop, device_));
  arch_ = prop.gcnArchName;
  arch_ = arch_.substr(0, arch_.find(':'));
  return DICE_OK;
}

uint64_t Runner::workspace_bytes() const {
  uint64_t total = 0;
  for (const auto& w : ws_) total += w->bytes();
  return total;
}

uint64_t job_reservation(int64_t height, int64_t width, uint32_t flags, bool diagnostics) {
  const uint64_t H = (uint64_t)std::max<int64_t>(height, 1);
  const uint64_t W = (uint64_t)std::max<int64_t>(width, 1);
  const uint64_t pixels = H * W;
  // DeviceBuffer::ensure: the need plus a quarter, rounded up to 1 MiB
  auto grown = [](uint64_t need) {
    need = std::max<uint64_t>(need, 1);
    return (uint64_t)((need + need / 4 + 0xFFFFFull) & ~0xFFFFFull);
  };
  uint64_t tiled = 0;  // the buffers whose size follows the tile shape: the largest shape
  int shapes[16][2];
  const int count = std::min(16, tile_shapes(shapes, 16));
  for (int k = 0; k < count; ++k) {
    const uint64_t TY = (uint64_t)shapes[k][0], TX = (uint64_t)shapes[k][1];
    const uint64_t tiles_x = (W + TX - 1) / TX, tiles_y = (H + TY - 1) / TY;
    const uint64_t tiles = tiles_x * tiles_y, segments = tiles_y * TY * tiles_x;
    tiled = std::max(tiled, grown(tiles * TY * TX * 2) + grown(tiles * 4) +
                                2 * grown(segments * 4) + grown((segments / 1024 + 2) * 4));
  }
  const uint64_t stages = H + 6, su = 6 * W, sites = pixels + 1;
  uint64_t bytes = tiled;
  for (uint64_t need : {(H + 1) * 4, (uint64_t)64, (uint64_t)64,
                        (uint64_t)kCountSlots * kCountStride * 8, 5 * W * 4, 5 * W * 4,
                        5 * W * 12, su * sizeof(Site), su, ((H + 7) / 8) * (W / 8) + 8, H * 24,
                        H * 4, (H / 8) * 16 + 16, H * 24, (H / 8) * 16 + 16, (H / 8) * 8 + 8,
                        su * sizeof(WriterRecord), sizeof(ChainState) + kMaxBands * 4,
                        sites * sizeof(Site), sites * sizeof(WriterRecord),
                        (su + sites + pixels / kStateGroup + 2) * 4})
    bytes += grown(need);
  bytes += 6 * grown(stages * 4) + 2 * grown(stages);
  if (diagnostics) bytes += grown(pixels * 4) + 2 * grown(pixels);
  if (!(flags & DICE_MAIN_ON_DEVICE)) bytes += grown(pixels * 8);
  if (!(flags & DICE_OUT_ON_DEVICE)) bytes += grown(pixels * 6);
  return bytes;
}

int32_t Runner::release_workspaces() {
  HIP_TRY(hipSetDevice(device_));
  DICE_TRY(wait_stream(stream_));
  DICE_TRY(wait_stream(fetch_stream_));
  DICE_TRY(wait_stream(out_stream_));
  ws_.clear();
  return DICE_OK;
}

// Sizes every job's workspace, places its input, stage table and output,
// and uploads the descriptors.
int32_t Runner::prepare(JobRequest* jobs, int n, bool* any_host_work) {
  std::vector<Upload> uploads;
  (void)any_host_work;
  while ((int)ws_.size() < n) ws_.push_back(std::make_unique<Workspace>());
  JobDev* H_jobs = h_jobs_.as<JobDev>();
  const int TY = tile_rows_, TX = tile_cols_;
  for (int j = 0; j < n; ++j) {
    JobRequest& R = jobs[j];
    Workspace& w = *ws_[j];
    DICE_TRY(w.init_events());
    JobDev& J = H_jobs[j];
    memset(&J, 0, sizeof J);
    if (R.code != DICE_OK) {  // failed before the device: no work, no buffers
      J.skip = 1;
      continue;
    }
    const int64_t H = R.height, W = R.width;
    J.P = device_params(R.params);
    J.K = writer_consts(R.params);
    J.factor_b = R.factor_b;
    J.run_producer = R.stages ? 0 : 1;
    J.height = (int32_t)H;
    J.width = (int32_t)W;
    J.tile_rows = TY;
    J.tile_cols = TX;
    J.tiles_x = (int32_t)((W + TX - 1) / TX);
    J.tiles_y = (int32_t)((H + TY - 1) / TY);
    J.win = full_window(J.height, J.tiles_y);
    const size_t pixels = (size_t)H * W;
    const size_t tiles = (size_t)J.tiles_x * J.tiles_y;
    const size_t segments = (size_t)J.tiles_y * TY * J.tiles_x;
    const size_t stages = (size_t)H + 6;
    const size_t cells = (size_t)W / 8;
    HIP_TRY(w.site_lists.ensure(tiles * TY * TX * sizeof(uint16_t), stream_));
    HIP_TRY(w.tile_count.ensure(tiles * 4, stream_));
    HIP_TRY(w.seg_count.ensure(segments * 4, stream_));
    HIP_TRY(w.seg_offset.ensure(segments * 4, stream_));
    HIP_TRY(w.seg_block_sums.ensure((segments / 1024 + 2) * 4, stream_));
    HIP_TRY(w.row_start.ensure((H + 1) * 4, stream_));
    HIP_TRY(w.counters.ensure(64, stream_));
    HIP_TRY(w.flags.ensure(64, stream_));
    HIP_TRY(w.writer_counts.ensure(kCountSlots * kCountStride * 8, stream_));
    HIP_TRY(w.su_aux.ensure(5 * W * 4, stream_));
    HIP_TRY(w.su_score.ensure(5 * W * 4, stream_));
    HIP_TRY(w.su_work.ensure(5 * W * 12, stream_));
    HIP_TRY(w.su_sites.ensure(6 * W * sizeof(Site), stream_));
    HIP_TRY(w.su_attempted.ensure(6 * W, stream_));
    HIP_TRY(w.failpos.ensure(((H + 7) / 8) * cells + 8, stream_));
    HIP_TRY(w.row_sums.ensure(H * 3 * 8, stream_));
    HIP_TRY(w.accepted.ensure(H * 4, stream_));
    HIP_TRY(w.epoch_adds.ensure((H / 8) * 16 + 16, stream_));
    HIP_TRY(w.cumulative.ensure(H * 24, stream_));
    HIP_TRY(w.epoch_wide.ensure((H / 8) * 16 + 16, stream_));
    HIP_TRY(w.epoch_store.ensure((H / 8) * 8 + 8, stream_));
    HIP_TRY(w.st_alpha.ensure(stages * 4, stream_));
    HIP_TRY(w.st_offset.ensure(stages * 4, stream_));
    HIP_TRY(w.st_replacement.ensure(stages * 4, stream_));
    HIP_TRY(w.st_has.ensure(stages, stream_));
    HIP_TRY(w.st_score_base.ensure(stages * 4, stream_));
    HIP_TRY(w.st_coarse.ensure(stages * 4, stream_));
    HIP_TRY(w.st_gate.ensure(stages, stream_));
    HIP_TRY(w.st_row_gate.ensure(stages * 4, stream_));
    HIP_TRY(w.su_records.ensure(6 * W * sizeof(WriterRecord), stream_));
    HIP_TRY(w.chain.ensure(sizeof(ChainState) + kMaxBands * 4, stream_));
    HIP_TRY(w.h_chain.ensure(sizeof(ChainState) + kMaxBands * 4));
    HIP_TRY(w.h_small.ensure(128 + kCountSlots * kCountStride * 8));
    HIP_TRY(w.h_row_start.ensure((H + 1) * 4));
    const bool want_scores = R.diag && (R.diag->score_plane || R.diag->at_floor_mask);
    if (want_scores) {
      HIP_TRY(w.diag_score.ensure(pixels * 4, stream_));
      HIP_TRY(w.diag_floor.ensure(pixels, stream_));
    }
    const bool want_changed = R.diag && R.diag->changed_mask;
    if (want_changed) HIP_TRY(w.diag_changed.ensure(pixels, stream_));
    J.lut = tables_->lut.as<float>();
    J.noop_delta = tables_->noop_delta.as<uint32_t>();
    J.factor_high = tables_->high.as<uint32_t>();
    J.factor_low = tables_->low.as<uint32_t>();
    J.stages.alpha = w.st_alpha.as<float>();
    J.stages.offset = w.st_offset.as<float>();
    J.stages.replacement = w.st_replacement.as<float>();
    J.stages.has_replacement = w.st_has.as<uint8_t>();
    J.stages.score_base = w.st_score_base.as<float>();
    J.stages.coarse = w.st_coarse.as<float>();
    J.stages.gate_secondary = w.st_gate.as<uint8_t>();
    J.stages.row_gate = w.st_row_gate.as<int32_t>();
    J.site_lists = w.site_lists.as<uint16_t>();
    J.tile_count = w.tile_count.as<uint32_t>();
    J.seg_count = w.seg_count.as<uint32_t>();
    J.seg_offset = w.seg_offset.as<uint32_t>();
    J.seg_block_sums = w.seg_block_sums.as<uint32_t>();
    J.row_start = w.row_start.as<uint32_t>();
    J.counters = w.counters.as<uint32_t>();
    J.error_flags = w.flags.as<uint32_t>();
    J.writer_counts = w.writer_counts.as<unsigned long long>();
    J.diag_score = want_scores ? w.diag_score.as<float>() : nullptr;
    J.diag_floor = want_scores ? w.diag_floor.as<uint8_t>() : nullptr;
    J.diag_changed = want_changed ? w.diag_changed.as<uint8_t>() : nullptr;
    J.su_aux = w.su_aux.as<float>();
    J.su_score = w.su_score.as<float>();
    J.su_work = w.su_work.as<float>();
    J.su_sites = w.su_sites.as<Site>();
    J.su_attempted = w.su_attempted.as<uint8_t>();
    J.su_writer_records = w.su_records.as<WriterRecord>();
    J.chain = w.chain.as<ChainState>();
    J.band_end = reinterpret_cast<uint32_t*>(w.chain.as<uint8_t>() + sizeof(ChainState));
    J.failpos = w.failpos.as<uint8_t>();
    J.row_sums = w.row_sums.as<double>();
    J.accepted = w.accepted.as<uint32_t>();
    J.epoch_adds = w.epoch_adds.as<double>();
    J.cumulative = w.cumulative.as<double>();
    J.epoch_wide = w.epoch_wide.as<double>();
    J.epoch_store = w.epoch_store.as<float>();
    // input and output
    if (R.flags & DICE_MAIN_ON_DEVICE) {
      J.rgbi = R.main;
    } else {
      HIP_TRY(w.rgbi.ensure(pixels * 8, stream_));
      if (!R.streamed)
        uploads.push_back({reinterpret_cast<const uint8_t*>(R.main), w.rgbi.as<uint8_t>(),
                           pixels * 8, &w});
      J.rgbi = w.rgbi.as<uint16_t>();
    }
    if (R.flags & DICE_OUT_ON_DEVICE) {
      J.out = R.out;
    } else {
      HIP_TRY(w.out.ensure(pixels * 6, stream_));
      J.out = w.out.as<uint16_t>();
    }
    if (R.stages) {
      const dice_stage_table* st = R.stages;
      const size_t c = (size_t)H + 6;
      HIP_TRY(hipMemcpyAsync(J.stages.alpha, st->aux_alpha, c * 4, hipMemcpyHostToDevice, stream_));
      HIP_TRY(hipMemcpyAsync(J.stages.offset, st->aux_offset, c * 4, hipMemcpyHostToDevice, stream_));
      HIP_TRY(hipMemcpyAsync(J.stages.replacement, st->aux_replacement, c * 4,
                             hipMemcpyHostToDevice, stream_));
      HIP_TRY(hipMemcpyAsync(J.stages.has_replacement, st->aux_has_replacement, c,
                             hipMemcpyHostToDevice, stream_));
      HIP_TRY(hipMemcpyAsync(J.stages.score_base, st->score_base, c * 4, hipMemcpyHostToDevice,
                             stream_));
      HIP_TRY(hipMemcpyAsync(J.stages.coarse, st->coarse_reference, c * 4,
                             hipMemcpyHostToDevice, stream_));
      HIP_TRY(hipMemcpyAsync(J.stages.gate_secondary, st->gate_secondary, c,
                             hipMemcpyHostToDevice, stream_));
      HIP_TRY(hipMemcpyAsync(J.stages.row_gate, st->row_gate, c * 4, hipMemcpyHostToDevice,
                             stream_));
    }
  }
  DICE_TRY(upload_frames(uploads));
  HIP_TRY(hipMemcpyAsync(d_jobs_.ptr, h_jobs_.ptr, sizeof(JobDev) * n, hipMemcpyHostToDevice,
                         stream_));
  return DICE_OK;
}

// Host frames go up through pinned staging: helper threads copy them in
// 8 MB chunks into each job's output staging, which is free until the job's
// outputs come down after phase B (stream order: these copies, then the
// phases, then the band copies), while this thread queues each chunk's DMA
// as soon as the chunk is staged.  A DMA straight from pageable memory is
// staged by the runtime one piece at a time, well below the link's rate.
int32_t Runner::upload_frames(const std::vector<Upload>& uploads) {
  if (uploads.empty()) return DICE_OK;
  struct Chunk {
    const uint8_t* src;
    uint8_t* staged;
    uint8_t* dst;
    size_t bytes;
  };
  constexpr size_t kChunk = 8u << 20;
  std::vector<Chunk> chunks;
  for (const Upload& u : uploads) {
    HIP_TRY(u.w->h_out.ensure(u.bytes));
    uint8_t* staging = u.w->h_out.as<uint8_t>();
    for (size_t at = 0; at < u.bytes; at += kChunk)
      chunks.push_back({u.src + at, staging + at, u.dst + at, std::min(kChunk, u.bytes - at)});
  }
  std::unique_ptr<std::atomic<int>[]> staged(new std::atomic<int>[chunks.size()]);
  for (size_t i = 0; i < chunks.size(); ++i) staged[i].store(0);
  std::atomic<size_t> next{0};
  auto copier = [&]() {
    for (size_t i; (i = next.fetch_add(1)) < chunks.size();) {
      memcpy(chunks[i].staged, chunks[i].src, chunks[i].bytes);
      staged[i].store(1, std::memory_order_release);
    }
  };
  std::vector<std::thread> helpers;
  const size_t count = std::min<size_t>(chunks.size(), 6);
  for (size_t t = 0; t < count; ++t) helpers.emplace_back(copier);
  int32_t rc = DICE_OK;
  for (size_t i = 0; i < chunks.size(); ++i) {
    while (!staged[i].load(std::memory_order_acquire)) std::this_thread::yield();
    if (rc != DICE_OK) continue;
    const hipError_t e = hipMemcpyAsync(chunks[i].dst, chunks[i].staged, chunks[i].bytes,
                                        hipMemcpyHostToDevice, stream_);
    if (e != hipSuccess)
      rc = set_error(DICE_INTERNAL, "hipMemcpyAsync (frame upload) failed: %s",
                     hipGetErrorString(e));
  }
  for (std::thread& t : helpers) t.join();
  return rc;
}

// Waits for everything queued on `stream` without spinning a core.
int32_t Runner::wait_stream(hipStream_t stream) {
  HIP_TRY(hipEventRecord(sync_event_, stream));
  HIP_TRY(wait_event(sync_event_));
  return DICE_OK;
}

int32_t Runner::mark(int slot) {
  if (fine_on_) HIP_TRY(hipEventRecord(fine_[slot], stream_));
  return DICE_OK;
}

int32_t Runner::launch_a(int n) {
  const JobDev* d = d_jobs_.as<JobDev>();
  const JobDev* h = h_jobs_.as<JobDev>();
  HIP_TRY(launch_zero_jobs(d, h, n, stream_));
  DICE_TRY(mark(0));
  HIP_TRY(launch_producer(d, h, n, stream_));
  DICE_TRY(mark(1));
  HIP_TRY(launch_scan(d, h, n, stream_));
  DICE_TRY(mark(2));
  HIP_TRY(launch_scan_segments(d, h, n, stream_));
  DICE_TRY(mark(3));
  return DICE_OK;
}

int32_t Runner::launch_b(int n) {
  const JobDev* d = d_jobs_.as<JobDev>();
  const JobDev* h = h_jobs_.as<JobDev>();
  DICE_TRY(mark(4));
  HIP_TRY(launch_sites(d, h, n, queue_.as<uint32_t>(), stream_, sites_variant_));
  DICE_TRY(mark(5));
  HIP_TRY(launch_startup(d, h, n, stream_));
  DICE_TRY(mark(6));
  HIP_TRY(launch_writer_prep(d, n, stream_));
  DICE_TRY(mark(7));
  return DICE_OK;
}

double Runner::fine_ms(int a, int b) const {
  float ms = 0.0f;
  if (!fine_on_ || hipEventElapsedTime(&ms, fine_[a], fine_[b]) != hipSuccess) return 0.0;
  return (double)ms;
}

std::string Runner::geometry_key(int n, char phase) const {
  const JobDev* h = h_jobs_.as<JobDev>();
  std::string key(1, phase);
  char part[96];
  snprintf(part, sizeof part, "|%d|%dx%d/%d", n, tile_rows_, tile_cols_, sites_variant_);
  key += part;
  for (int j = 0; j < n; ++j) {
    snprintf(part, sizeof part, "|%dx%d:%d%d%d%d", h[j].height, h[j].width, h[j].run_producer,
             h[j].skip, h[j].diag_score ? 1 : 0, h[j].diag_changed ? 1 : 0);
    key += part;
  }
  return key;
}

// Launch directly, or replay the graph captured for this geometry (the
// first run of a geometry is direct; the second captures it).
int32_t Runner::replay_or_capture(Graph& g, const std::string& key, bool graph, bool* replayed,
                                  int32_t (Runner::*launch)(int), int n) {
  if (!graph) return (this->*launch)(n);
  if (g.key != key) {
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
    g.exec = nullptr;
    g.key = key;
    g.seen = 0;
  }
  g.seen += 1;
  if (!g.exec) {
    if (g.seen < 2) return (this->*launch)(n);
    HIP_TRY(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    const int32_t rc = (this->*launch)(n);
    hipGraph_t captured = nullptr;
    const hipError_t e = hipStreamEndCapture(stream_, &captured);
    if (rc != DICE_OK) {
      if (captured) (void)hipGraphDestroy(captured);
      return rc;
    }
    if (e != hipSuccess)
      return set_error(DICE_INTERNAL, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
    const hipError_t ie = hipGraphInstantiate(&g.exec, captured, nullptr, nullptr, 0);
    (void)hipGraphDestroy(captured);
    if (ie != hipSuccess) {
      g.exec = nullptr;
      return set_error(DICE_INTERNAL, "hipGraphInstantiate failed: %s", hipGetErrorString(ie));
    }
  }
  HIP_TRY(hipGraphLaunch(g.exec, stream_));
  *replayed = true;
  return DICE_OK;
}

int32_t Runner::phase_a(int n, bool graph, bool* replayed) {
  return replay_or_capture(graph_a_, geometry_key(n, 'A'), graph, replayed, &Runner::launch_a, n);
}

int32_t Runner::phase_b(int n, bool graph, bool* replayed) {
  return replay_or_capture(graph_b_, geometry_key(n, 'B'), graph, replayed, &Runner::launch_b, n);
}

namespace {

// The first failure of a job's worker threads (chain, copy scheduler, hash)
// with the failing thread's message.
struct JobThreads {
  std::mutex mutex;
  int32_t code = DICE_OK;
  std::string error;
  void fail(int32_t rc) {
    std::lock_guard<std::mutex> lock(mutex);
    if (code == DICE_OK) {
      code = rc;
      error = error_buffer();
    }
  }
};

}  // namespace

int32_t Runner::run(JobRequest* jobs, int n, const RunOptions& options, RunReport* report) {
  if (n < 1 || n > kMaxBatch) return set_error(DICE_INVALID, "a batch holds 1..%d jobs", kMaxBatch);
  HIP_TRY(hipSetDevice(device_));
  for (int j = 0; j < n; ++j) {
    if (jobs[j].result) memset(jobs[j].result, 0, sizeof *jobs[j].result);
  }
  DICE_TRY(prepare(jobs, n, nullptr));
  JobDev* H_jobs = h_jobs_.as<JobDev>();
  const double t_start = now_seconds();
  fine_on_ = options.fine_timing && !options.graph;
  bool replayed_a = false, replayed_b = false;

  // ---- phase A: producer, analysis scan, segment scan
  for (int j = 0; j < n; ++j)
    if (jobs[j].code == DICE_OK) jobs[j].cb.progress(2, 0, 1, 0, 0);
  HIP_TRY(hipEventRecord(ev_[0], stream_));
  DICE_TRY(phase_a(n, options.graph, &replayed_a));
  HIP_TRY(hipEventRecord(ev_[1], stream_));
  for (int j = 0; j < n; ++j) {
    if (H_jobs[j].skip) continue;  // failed before the device: no workspace
    Workspace& w = *ws_[j];
    HIP_TRY(hipMemcpyAsync(w.h_small.ptr, w.counters.ptr, 64, hipMemcpyDeviceToHost, stream_));
    HIP_TRY(hipMemcpyAsync(w.h_small.as<uint8_t>() + 64, w.flags.ptr, 4, hipMemcpyDeviceToHost,
                           stream_));
  }
  DICE_TRY(wait_stream(stream_));
  const double t_a = now_seconds();
  float ms_a = 0.0f;
  (void)hipEventElapsedTime(&ms_a, ev_[0], ev_[1]);

  // ---- size the site arrays
  for (int j = 0; j < n; ++j) {
    JobRequest& R = jobs[j];
    JobDev& J = H_jobs[j];
    Workspace& w = *ws_[j];
    if (R.code != DICE_OK) {
      J.skip = 1;
      continue;
    }
    const uint32_t flags = *reinterpret_cast<uint32_t*>(w.h_small.as<uint8_t>() + 64);
    if (int32_t rc = flag_error(flags)) {
      R.code = rc;
      R.error = error_buffer();
      J.skip = 1;
      continue;
    }
    if (R.cb.cancelled()) {
      R.code = cancelled_error();
      R.error = error_buffer();
      J.skip = 1;
      continue;
    }
    R.cb.progress(2, 1, 1, 0, 0);
    R.cb.progress(3, 0, R.height, 0, 0);
    const size_t count = w.h_small.as<uint32_t>()[0];
    const size_t su = 6 * (size_t)R.width;
    HIP_TRY(w.sites.ensure((count + 1) * sizeof(Site), stream_));
    HIP_TRY(w.records.ensure((count + 1) * sizeof(WriterRecord), stream_));
    // device: startup states, site states, then the group start states;
    // host: startup states, then the group start states
    const size_t groups = count / kStateGroup + 2;
    HIP_TRY(w.states.ensure((su + count + 1 + groups) * 4, stream_));
    HIP_TRY(w.h_records.ensure((count + su + 1) * sizeof(WriterRecord)));
    HIP_TRY(w.h_states.ensure((su + std::max(count + 1, groups)) * 4));
    J.sites = w.sites.as<Site>();
    J.writer_records = w.records.as<WriterRecord>();
    J.su_site_state = w.states.as<uint32_t>();
    J.site_state = J.su_site_state + su;
    J.group_state = J.site_state + count + 1;
  }
  HIP_TRY(hipMemcpyAsync(d_jobs_.ptr, h_jobs_.ptr, sizeof(JobDev) * n, hipMemcpyHostToDevice,
                         stream_));
  for (int j = 0; j < n; ++j)
    if (!H_jobs[j].skip && H_jobs[j].diag_changed)
      HIP_TRY(launch_fill_u8(H_jobs[j].diag_changed, (size_t)jobs[j].height * jobs[j].width, 0,
                             stream_));

  // ---- phase B: sites, hidden startup, writer draw schedules
  HIP_TRY(hipEventRecord(ev_[2], stream_));
  DICE_TRY(phase_b(n, options.graph, &replayed_b));
  HIP_TRY(hipEventRecord(ev_[3], stream_));
  // row tables and flags first (small), then the draw schedules band by
  // band, interleaved across jobs on the fetch stream, so every chain can
  // start as soon as its first band is on the host (output rows go back on
  // each job's own copy stream, which does not queue behind these)
  for (int j = 0; j < n; ++j) {
    const JobDev& J = H_jobs[j];
    if (J.skip) continue;
    Workspace& w = *ws_[j];
    HIP_TRY(hipMemcpyAsync(w.h_row_start.ptr, J.row_start, ((size_t)J.height + 1) * 4,
                           hipMemcpyDeviceToHost, stream_));
    HIP_TRY(hipMemcpyAsync(w.h_small.as<uint8_t>() + 64, w.flags.ptr, 4, hipMemcpyDeviceToHost,
                           stream_));
  }
  DICE_TRY(wait_stream(stream_));
  const double t_b = now_seconds();
  float ms_b = 0.0f;
  (void)hipEventElapsedTime(&ms_b, ev_[2], ev_[3]);
  int max_bands = 1;
  // Bands of the host chain and hash: with at most one hash for every two
  // cores, twice the default, so that each job's hash starts sooner after
  // its chain (plain p50 -1.5 ms, one batch of 8 +7 %); a busier host keeps
  // the default, which costs less per band (the serving engine under load
  // lost 12 % of its capacity with the finer bands).
  int band_count = options.bands;
  if (!options.device_chain) {
    int messages = 0;
    for (int j = 0; j < n; ++j)
      if (!H_jobs[j].skip && (jobs[j].flags & DICE_SKIP_SHA256) == 0) ++messages;
    if (hash_load(messages) <= 0.5) band_count = 2 * options.bands;
  }
  for (int j = 0; j < n; ++j) {
    JobDev& J = H_jobs[j];
    if (J.skip) continue;
    Workspace& w = *ws_[j];
    const uint32_t flags = *reinterpret_cast<uint32_t*>(w.h_small.as<uint8_t>() + 64);
    if (int32_t rc = flag_error(flags)) {
      jobs[j].code = rc;
      jobs[j].error = error_buffer();
      J.skip = 1;
      continue;
    }
    w.bands = std::max(1, std::min<int>(std::min(band_count, kMaxBands), J.height));
    max_bands = std::max(max_bands, w.bands);
    if (options.device_chain) continue;  // the schedules stay on the device
    const size_t su = 6 * (size_t)J.width;
    HIP_TRY(hipMemcpyAsync(w.h_records.ptr, J.su_writer_records, su * sizeof(WriterRecord),
                           hipMemcpyDeviceToHost, fetch_stream_));
    HIP_TRY(hipEventRecord(w.startup_fetched, fetch_stream_));
  }
  // The first kFetchAhead bands of every job now; each chain then asks for
  // band b + kFetchAhead when it starts band b, so the schedules come down
  // just ahead of need and never queue the output copies behind them.
  auto fetch_band = [this](const JobDev& J, Workspace& w, int b) -> int32_t {
    const uint32_t* row_start = w.h_row_start.as<uint32_t>();
    const uint32_t s0 = row_start[band_row(J.height, w.bands, b)];
    const uint32_t s1 = row_start[band_row(J.height, w.bands, b + 1)];
    const size_t su = 6 * (size_t)J.width;
    if (s1 > s0)
      HIP_TRY(hipMemcpyAsync(w.h_records.as<WriterRecord>() + su + s0, J.writer_records + s0,
                             (size_t)(s1 - s0) * sizeof(WriterRecord), hipMemcpyDeviceToHost,
                             fetch_stream_));
    HIP_TRY(hipEventRecord(w.fetched[b], fetch_stream_));
    return DICE_OK;
  };
  constexpr int kFetchAhead = 2;
  for (int b = 0; b < std::min(max_bands, kFetchAhead) && !options.device_chain; ++b) {
    for (int j = 0; j < n; ++j) {
      const JobDev& J = H_jobs[j];
      Workspace& w = *ws_[j];
      if (J.skip || b >= w.bands) continue;
      DICE_TRY(fetch_band(J, w, b));
    }
  }

  // ---- host: per job, the serial chain and the hash, overlapped by band.
  // Output rows always come down into pinned staging (a DMA straight into
  // a pageable caller buffer is staged by the runtime and far slower).  The
  // chains hand each band whose writer values are queued to one copy
  // scheduler, which queues the output copies on one stream, lowest band
  // first across the jobs: every job's first band arrives early, so every
  // hash starts early, and a job whose chain runs ahead does not take the
  // host link from the others.
  struct CopyQueue {
    std::mutex mutex;
    std::condition_variable cv;
    std::vector<int> ready, queued, bands;
    std::vector<bool> closed;  // chain failed or nothing to copy
  } cq;
  cq.ready.assign(n, 0);
  cq.queued.assign(n, 0);
  cq.bands.assign(n, 0);
  cq.closed.assign(n, true);
  for (int j = 0; j < n; ++j) {
    if (H_jobs[j].skip) continue;
    const JobRequest& R = jobs[j];
    const bool want_sha = (R.flags & DICE_SKIP_SHA256) == 0;
    const bool host_out = (R.flags & DICE_OUT_ON_DEVICE) == 0;
    if (!want_sha && !host_out) continue;
    HIP_TRY(ws_[j]->h_out.ensure((size_t)R.height * R.width * 6));
    cq.bands[j] = ws_[j]->bands;
    cq.closed[j] = false;
  }
  auto values_queued = [&cq](int j, int bands_ready) {
    {
      std::lock_guard<std::mutex> lock(cq.mutex);
      cq.ready[j] = bands_ready;
    }
    cq.cv.notify_one();
  };
  auto chain_failed = [&cq](int j) {
    {
      std::lock_guard<std::mutex> lock(cq.mutex);
      cq.closed[j] = true;
    }
    cq.cv.notify_one();
  };
  std::vector<JobThreads> outcome(n);
  // Set by the calling thread when the job's cancel callback says so; the
  // job's chain and hash threads stop at their next band.
  std::unique_ptr<std::atomic<int>[]> cancel(new std::atomic<int>[n]);
  for (int j = 0; j < n; ++j) cancel[j].store(0);
  std::vector<std::thread> threads;
  std::atomic<int> threads_done{0};
  auto spawn = [&threads, &threads_done](auto body) {
    threads.emplace_back([body, &threads_done]() mutable {
      body();
      threads_done.fetch_add(1);
    });
  };
  std::vector<double> chain_seconds(n, 0.0), hash_seconds(n, 0.0), copy_seconds(n, 0.0);
  std::vector<double> chain_waits(n, 0.0), band_waits(n, 0.0);

  // Per job, once its threads are done: the device error flags, the final
  // cancellation check, counts, diagnostics, timings and the output commit.
  // A job without callbacks is finalized by the last of its own threads
  // (so with a completion hook it completes without waiting for the rest
  // of the batch); every other job on the calling thread after the batch.
  std::unique_ptr<std::atomic<int>[]> live(new std::atomic<int>[n]);
  std::vector<char> early(n, 0), finalized(n, 0);
  std::vector<int> commit;  // calling-thread jobs whose host output is copied in parallel
  auto finalize = [&](int j, bool on_caller) {
    JobRequest& R = jobs[j];
    const JobDev& J = H_jobs[j];
    Workspace& w = *ws_[j];
    const double t_fin = now_seconds();
    auto body = [&]() -> int32_t {
      if (outcome[j].code != DICE_OK) return set_error(outcome[j].code, "%s",
                                                      outcome[j].error.c_str());
      HIP_TRY(hipSetDevice(device_));
      if (!on_caller) HIP_TRY(wait_event(w.tail));
      const uint32_t flags = *reinterpret_cast<uint32_t*>(w.h_small.as<uint8_t>() + 64);
      DICE_TRY(flag_error(flags));
      if (cancel[j].load() || R.cb.cancelled()) return cancelled_error();
      const unsigned long long* slots =
          reinterpret_cast<const unsigned long long*>(w.h_small.as<uint8_t>() + 128);
      uint64_t written = 0, changed = 0;
      for (int slot = 0; slot < kCountSlots; ++slot) {
        written += slots[slot * kCountStride + 0];
        changed += slots[slot * kCountStride + 1];
      }
      R.result->written_pixels = written;
      R.result->changed_pixels = changed;
      if (options.device_chain) {
        const ChainState& chain = *w.h_chain.as<ChainState>();
        for (int stage = 0; stage < 6; ++stage) {
          R.result->startup_attempted[stage] = chain.startup_attempted[stage];
          R.result->startup_advances[stage] = chain.startup_advances[stage];
        }
        R.result->startup_final_rng_state = chain.startup_final;
        R.result->attempted_pixels = w.h_row_start.as<uint32_t>()[R.height];
        R.result->public_rng_advances = chain.advances;
        R.result->final_rng_state = chain.state;
      }
      const size_t pixels = (size_t)R.height * R.width;
      if (R.diag) {
        if (R.diag->score_plane)
          HIP_TRY(hipMemcpy(R.diag->score_plane, J.diag_score, pixels * 4, hipMemcpyDeviceToHost));
        if (R.diag->at_floor_mask)
          HIP_TRY(hipMemcpy(R.diag->at_floor_mask, J.diag_floor, pixels, hipMemcpyDeviceToHost));
        if (R.diag->changed_mask)
          HIP_TRY(hipMemcpy(R.diag->changed_mask, J.diag_changed, pixels,
                            hipMemcpyDeviceToHost));
      }
      R.cb.progress(3, R.height, R.height, (int64_t)R.result->attempted_pixels,
                    (int64_t)written);
      if (R.timings) {
        dice_stage_timings& T = *R.timings;
        T.producer_device += t_a - t_start;
        T.analysis_device += t_b - t_a;
        T.writer_host += chain_seconds[j];
        T.download_sites += chain_waits[j];  // chain waiting for its draw schedules
        T.upload += band_waits[j];           // hash waiting for the chain
        T.download_output += copy_seconds[j];
        T.sha256_host += hash_seconds[j];
        T.finish_device += now_seconds() - t_fin;
        if (fine_on_) {
          T.producer_kernels_ms += fine_ms(0, 1);
          T.scan_kernel_ms += fine_ms(1, 2);
          T.segments_kernel_ms += fine_ms(2, 3);
          T.sites_kernel_ms += fine_ms(4, 5);
          T.startup_kernels_ms += fine_ms(5, 6);
          T.writer_prep_kernel_ms += fine_ms(6, 7);
        } else {
          T.producer_kernels_ms += ms_a;
          T.sites_kernel_ms += ms_b;
        }
      }
      // The caller's host buffer is written only now, after every check
      // of its job passed (engine.py:256-260 writes the caller's output
      // once, on success).
      if ((R.flags & DICE_OUT_ON_DEVICE) == 0) {
        if (on_caller)
          commit.push_back(j);
        else
          memcpy(R.out, w.h_out.ptr, pixels * 6);
      }
      return DICE_OK;
    };
    if (int32_t rc = body()) {
      R.code = rc;
      R.error = error_buffer();
    }
    finalized[j] = 1;
  };
  auto job_exit = [&](int j) {
    if (live[j].fetch_sub(1) == 1 && early[j]) {
      finalize(j, false);
      options.on_job_done(j);
    }
  };
  // The device writer chain (opt-in): after phase B, one wavefront per job
  // walks its chain, one launch per band for the whole batch, on the fetch
  // stream (it fetches nothing in this mode), each band followed by its
  // writer values.  The job descriptors go up again first: jobs that failed
  // after phase B are marked skip on the host only.
  int32_t device_chain_rc = DICE_OK;
  double device_chain_seconds = 0.0;
  if (options.device_chain) {
    const double t0 = now_seconds();
    hipStream_t cs = fetch_stream_;
    const int bands = std::max(1, std::min(options.bands, kMaxBands));
    auto enqueue = [&]() -> int32_t {
      HIP_TRY(hipStreamWaitEvent(cs, ev_[3], 0));
      for (int j = 0; j < n; ++j) {
        const JobDev& J = H_jobs[j];
        if (J.skip) continue;
        ChainState* init = ws_[j]->h_chain.as<ChainState>();
        memset(init, 0, sizeof(ChainState));
        init->state = jobs[j].rng_state;
        HIP_TRY(hipMemcpyAsync(J.chain, init, sizeof(ChainState), hipMemcpyHostToDevice, cs));
      }
      HIP_TRY(hipMemcpyAsync(d_jobs_.ptr, h_jobs_.ptr, sizeof(JobDev) * n, hipMemcpyHostToDevice,
                             cs));
      const JobDev* d_jobs = d_jobs_.as<JobDev>();
      HIP_TRY(launch_writer_startup_chain(d_jobs, n, cs));
      for (int j = 0; j < n; ++j) {
        const JobDev& J = H_jobs[j];
        if (!J.skip) HIP_TRY(launch_writer_startup_check(J, 0, &J.chain->startup_final, cs));
      }
      for (int b = 0; b < bands; ++b) {
        HIP_TRY(launch_writer_chain(d_jobs, n, b, bands, cs));
        for (int j = 0; j < n; ++j) {
          const JobDev& J = H_jobs[j];
          if (J.skip) continue;
          Workspace& w = *ws_[j];
          if (b >= w.bands) continue;
          const uint32_t* row_start = w.h_row_start.as<uint32_t>();
          const uint32_t s0 = row_start[band_row(J.height, w.bands, b)];
          const uint32_t s1 = row_start[band_row(J.height, w.bands, b + 1)];
          HIP_TRY(launch_writer_values(J, s0, s1, 0, J.band_end + b, cs));
          HIP_TRY(hipEventRecord(w.band[b], cs));
        }
      }
      for (int j = 0; j < n; ++j) {
        const JobDev& J = H_jobs[j];
        if (J.skip) continue;
        Workspace& w = *ws_[j];
        HIP_TRY(hipMemcpyAsync(w.h_chain.ptr, J.chain, sizeof(ChainState), hipMemcpyDeviceToHost,
                               cs));
        HIP_TRY(hipMemcpyAsync(w.h_small.as<uint8_t>() + 128, J.writer_counts,
                               kCountSlots * kCountStride * 8, hipMemcpyDeviceToHost, cs));
        HIP_TRY(hipMemcpyAsync(w.h_small.as<uint8_t>() + 64, J.error_flags, 4,
                               hipMemcpyDeviceToHost, cs));
        HIP_TRY(hipEventRecord(w.tail, cs));
      }
      return DICE_OK;
    };
    device_chain_rc = enqueue();  // a failure fails every job; no hash waits for it
    device_chain_seconds = now_seconds() - t0;
  }
  for (int j = 0; j < n; ++j) {
    live[j].store(0);
    if (H_jobs[j].skip) continue;
    const bool hashes = (jobs[j].flags & DICE_SKIP_SHA256) == 0 ||
                        (jobs[j].flags & DICE_OUT_ON_DEVICE) == 0;
    live[j].store((options.device_chain ? 0 : 1) + (hashes ? 1 : 0));
    early[j] = options.on_job_done && !jobs[j].cb.cb && live[j].load() > 0;
  }
  // Jobs whose outputs are hashed in pairs (hash_in_pairs): a pair
  // thread, spawned after this loop, stands in for their hash threads.
  // Only jobs of the same geometry pair up: their bands arrive together and
  // they finish together, so a small job never waits for a large one.
  std::vector<int> paired;
  int messages = 0;
  for (int j = 0; j < n; ++j)
    if (!H_jobs[j].skip && (jobs[j].flags & DICE_SKIP_SHA256) == 0) ++messages;
  if (hash_in_pairs(messages)) {
    std::vector<char> taken(n, 0);
    for (int a = 0; a < n; ++a) {
      if (taken[a] || H_jobs[a].skip || (jobs[a].flags & DICE_SKIP_SHA256) != 0) continue;
      for (int b = a + 1; b < n; ++b) {
        if (taken[b] || H_jobs[b].skip || (jobs[b].flags & DICE_SKIP_SHA256) != 0) continue;
        if (jobs[b].height != jobs[a].height || jobs[b].width != jobs[a].width) continue;
        taken[a] = taken[b] = 1;
        paired.push_back(a);
        paired.push_back(b);
        break;
      }
    }
  }
  std::vector<char> in_pair(n, 0);
  for (int j : paired) in_pair[j] = 1;
  g_messages_hashing.fetch_add(messages);  // each hash thread takes its own off when done
  for (int j = 0; j < n; ++j) {
    if (H_jobs[j].skip) continue;
    Workspace& w = *ws_[j];
    {
      std::lock_guard<std::mutex> lock(w.mutex);
      w.published = 0;
      w.aborted = false;
    }
    const int bands = w.bands;
    if (options.device_chain) {
      // every band is published at once: its event orders the hash
      // thread's copy behind the chain and values of that band
      if (device_chain_rc != DICE_OK) {
        outcome[j].fail(device_chain_rc);
        {
          std::lock_guard<std::mutex> lock(w.mutex);
          w.aborted = true;
        }
        chain_failed(j);
      } else {
        values_queued(j, bands);
      }
      chain_seconds[j] = device_chain_seconds;
    } else {
    // chain thread
    spawn([this, j, bands, jobs, H_jobs, &outcome, &chain_seconds,
                          &chain_waits, &values_queued, &chain_failed, &fetch_band,
                          &cancel, &job_exit]() {
      Workspace& w = *ws_[j];
      JobRequest& R = jobs[j];
      const JobDev& J = H_jobs[j];
      JobThreads& out = outcome[j];
      const double t0 = now_seconds();
      double fetch_wait = 0.0, walking = 0.0;
      auto body = [&]() -> int32_t {
        HIP_TRY(hipSetDevice(device_));
        const int64_t H = R.height, W = R.width;
        const size_t su = 6 * (size_t)W;
        const WriterRecord* records = w.h_records.as<WriterRecord>();
        const uint32_t* row_start = w.h_row_start.as<uint32_t>();
        uint32_t* states = w.h_states.as<uint32_t>();
        dice_replay_result* res = R.result;
        uint32_t state = R.rng_state;
        const double f0 = now_seconds();
        HIP_TRY(wait_event(w.startup_fetched));
        fetch_wait += now_seconds() - f0;
        state = walk_startup(records, W, state, states, res);
        HIP_TRY(hipMemcpyAsync(J.su_site_state, states, su * 4, hipMemcpyHostToDevice, stream_));
        HIP_TRY(launch_writer_startup_check(J, state, nullptr, stream_));
        uint64_t advances = 0;
        const WriterRecord* main = records + su;
        uint32_t* group_states = states + su;
        for (int b = 0; b < bands; ++b) {
          if (cancel[j].load()) return cancelled_error();
          const uint32_t s0 = row_start[band_row(H, bands, b)];
          const uint32_t s1 = row_start[band_row(H, bands, b + 1)];
          if (b + kFetchAhead < bands) DICE_TRY(fetch_band(J, w, b + kFetchAhead));
          const double f0 = now_seconds();
          HIP_TRY(wait_event(w.fetched[b]));
          fetch_wait += now_seconds() - f0;
          const double k0 = now_seconds();
          state = chain_records_grouped(main, s0, s1 - s0, state, group_states, advances);
          walking += now_seconds() - k0;
          if (s1 > s0) {
            // the group starts that fall in this band; the device expands
            // them to the start state of every site of the band
            const uint32_t g0 = (s0 + kStateGroup - 1) / kStateGroup;
            const uint32_t g1 = (s1 + kStateGroup - 1) / kStateGroup;
            if (g1 > g0)
              HIP_TRY(hipMemcpyAsync(J.group_state + g0, group_states + g0, (size_t)(g1 - g0) * 4,
                                     hipMemcpyHostToDevice, stream_));
            HIP_TRY(launch_writer_expand(J, s0, s1, stream_));
            HIP_TRY(launch_writer_values(J, s0, s1, state, nullptr, stream_));
          }
          HIP_TRY(hipEventRecord(w.band[b], stream_));
          values_queued(j, b + 1);
        }
        res->attempted_pixels = row_start[H];
        res->public_rng_advances = advances;
        res->final_rng_state = state;
        HIP_TRY(hipMemcpyAsync(w.h_small.as<uint8_t>() + 128, J.writer_counts,
                               kCountSlots * kCountStride * 8, hipMemcpyDeviceToHost, stream_));
        HIP_TRY(hipMemcpyAsync(w.h_small.as<uint8_t>() + 64, J.error_flags, 4,
                               hipMemcpyDeviceToHost, stream_));
        HIP_TRY(hipEventRecord(w.tail, stream_));
        return DICE_OK;
      };
      if (int32_t rc = body()) {
        out.fail(rc);
        {
          std::lock_guard<std::mutex> lock(w.mutex);
          w.aborted = true;
        }
        chain_failed(j);
      }
      w.cv.notify_all();
      chain_seconds[j] = now_seconds() - t0;
      chain_waits[j] = fetch_wait;
      if (trace_enabled())
        fprintf(stderr, "dice-trace chain job=%d total=%.2fms fetch_wait=%.2fms walk=%.2fms\n", j,
                chain_seconds[j] * 1e3, fetch_wait * 1e3, walking * 1e3);
      job_exit(j);
    });
    }
    // hash / download thread
    if ((jobs[j].flags & DICE_SKIP_SHA256) != 0 && (jobs[j].flags & DICE_OUT_ON_DEVICE) != 0)
      continue;
    if (in_pair[j]) continue;
    spawn([this, j, bands, jobs, &outcome, &hash_seconds, &copy_seconds,
                          &band_waits, &cancel, &job_exit]() {
      Workspace& w = *ws_[j];
      JobRequest& R = jobs[j];
      const bool want_sha = (R.flags & DICE_SKIP_SHA256) == 0;
      double hashing = 0.0, copying = 0.0;
      auto body = [&]() -> int32_t {
        HIP_TRY(hipSetDevice(device_));
        const int64_t H = R.height, W = R.width;
        uint8_t* staging = w.h_out.as<uint8_t>();
        Sha256 sha;
        for (int b = 0; b < bands; ++b) {
          if (cancel[j].load()) return cancelled_error();
          const double p0 = now_seconds();
          {
            std::unique_lock<std::mutex> lock(w.mutex);
            w.cv.wait(lock, [&] { return w.published > b || w.aborted; });
            if (w.published <= b) return DICE_OK;  // the chain failed; it reports
          }
          const double c0 = now_seconds();
          band_waits[j] += c0 - p0;
          HIP_TRY(wait_event(w.copied[b]));
          const double c1 = now_seconds();
          if (trace_enabled())
            fprintf(stderr, "dice-trace band job=%d b=%d published=%.2f copied=%.2f\n", j, b,
                    (c0 - trace_origin()) * 1e3, (c1 - trace_origin()) * 1e3);
          const size_t begin = (size_t)band_row(H, bands, b) * W * 6;
          const size_t end = (size_t)band_row(H, bands, b + 1) * W * 6;
          constexpr size_t kPiece = 256 << 10;
          if (want_sha)
            for (size_t at = begin; at < end; at += kPiece)
              sha.update(staging + at, std::min(kPiece, end - at));
          copying += c1 - c0;
          hashing += now_seconds() - c1;
        }
        if (want_sha && !sha.hex(R.result->output_sha256))
          return set_error(DICE_INTERNAL, "SHA-256 failed");
        return DICE_OK;
      };
      if (int32_t rc = body()) outcome[j].fail(rc);
      if (want_sha) g_messages_hashing.fetch_sub(1);
      hash_seconds[j] = hashing;
      copy_seconds[j] = copying;
      if (trace_enabled())
        fprintf(stderr, "dice-trace hash job=%d band_wait=%.2fms copy_wait=%.2fms hash=%.2fms\n",
                j, band_waits[j] * 1e3, copying * 1e3, hashing * 1e3);
      job_exit(j);
    });
  }
  // Pair threads: band b of both jobs once both are on the host, the two
  // messages compressed interleaved (Sha256Stream::update_interleaved); a
  // job whose chain fails or that is cancelled drops out and the other
  // goes on alone.
  for (size_t p = 0; p + 1 < paired.size(); p += 2) {
    const int pair[2] = {paired[p], paired[p + 1]};
    spawn([this, pair, jobs, &outcome, &hash_seconds, &copy_seconds, &band_waits, &cancel,
           &job_exit]() {
      struct Member {
        int j;
        bool active = true, done = false;
        double hashing = 0.0, copying = 0.0;
        Sha256Stream sha;
      };
      Member m[2];
      m[0].j = pair[0];
      m[1].j = pair[1];
      auto drop = [&](Member& x, int32_t rc) {
        x.active = false;
        if (rc != DICE_OK) outcome[x.j].fail(rc);
      };
      auto finish = [&](Member& x) {
        if (x.done) return;
        x.done = true;
        g_messages_hashing.fetch_sub(1);
        if (x.active && !x.sha.hex(jobs[x.j].result->output_sha256))
          drop(x, set_error(DICE_INTERNAL, "SHA-256 failed"));
        x.active = false;
        hash_seconds[x.j] = x.hashing;
        copy_seconds[x.j] = x.copying;
        if (trace_enabled())
          fprintf(stderr,
                  "dice-trace hash job=%d band_wait=%.2fms copy_wait=%.2fms hash=%.2fms pair\n",
                  x.j, band_waits[x.j] * 1e3, x.copying * 1e3, x.hashing * 1e3);
        job_exit(x.j);
      };
      if (hipSetDevice(device_) != hipSuccess) {
        for (Member& x : m)
          drop(x, set_error(DICE_INTERNAL, "hipSetDevice failed in the hash thread"));
      }
      const int bands = std::max(ws_[m[0].j]->bands, ws_[m[1].j]->bands);
      for (int b = 0; b < bands; ++b) {
        Sha256Stream* streams[2];
        const uint8_t* data[2];
        size_t bytes[2];
        Member* fed[2];
        int k = 0;
        for (Member& x : m) {
          if (!x.active) continue;
          Workspace& w = *ws_[x.j];
          const JobRequest& R = jobs[x.j];
          if (b >= w.bands) continue;
          if (cancel[x.j].load()) {
            drop(x, cancelled_error());
            continue;
          }
          const double p0 = now_seconds();
          {
            std::unique_lock<std::mutex> lock(w.mutex);
            w.cv.wait(lock, [&] { return w.published > b || w.aborted; });
            if (w.published <= b) {  // the chain failed; it reports
              x.active = false;
              continue;
            }
          }
          const double c0 = now_seconds();
          band_waits[x.j] += c0 - p0;
          const hipError_t e = wait_event(w.copied[b]);
          if (e != hipSuccess) {
            drop(x, set_error(DICE_INTERNAL, "band copy failed: %s", hipGetErrorString(e)));
            continue;
          }
          x.copying += now_seconds() - c0;
          const size_t begin = (size_t)band_row(R.height, w.bands, b) * R.width * 6;
          const size_t end = (size_t)band_row(R.height, w.bands, b + 1) * R.width * 6;
          streams[k] = &x.sha;
          data[k] = w.h_out.as<uint8_t>() + begin;
          bytes[k] = end - begin;
          fed[k++] = &x;
        }
        const double h0 = now_seconds();
        if (k > 0) Sha256Stream::update_interleaved(streams, data, bytes, k);
        const double share = (now_seconds() - h0) / std::max(k, 1);
        for (int i = 0; i < k; ++i) fed[i]->hashing += share;
        // a member whose last band this was completes now
        for (int i = 0; i < k; ++i)
          if (b + 1 == ws_[fed[i]->j]->bands) finish(*fed[i]);
      }
      for (Member& x : m) finish(x);  // the members that dropped out
    });
  }
  // the copy scheduler: the lowest band ready across all jobs goes next
  spawn([this, n, jobs, H_jobs, &cq, &outcome]() {
    (void)hipSetDevice(device_);
    for (;;) {
      int j = -1, b = 0;
      {
        std::unique_lock<std::mutex> lock(cq.mutex);
        cq.cv.wait(lock, [&] {
          j = -1;
          bool finished = true;
          for (int k = 0; k < n; ++k) {
            if (cq.closed[k] || cq.queued[k] == cq.bands[k]) continue;
            finished = false;
            if (cq.queued[k] < cq.ready[k] && (j < 0 || cq.queued[k] < cq.queued[j])) j = k;
          }
          return j >= 0 || finished;
        });
        if (j < 0) break;
        b = cq.queued[j];
      }
      Workspace& w = *ws_[j];
      const JobRequest& R = jobs[j];
      auto submit = [&]() -> int32_t {
        const size_t begin = (size_t)band_row(R.height, w.bands, b) * R.width * 6;
        const size_t end = (size_t)band_row(R.height, w.bands, b + 1) * R.width * 6;
        HIP_TRY(hipStreamWaitEvent(out_stream_, w.band[b], 0));
        if (end > begin)
          HIP_TRY(hipMemcpyAsync(w.h_out.as<uint8_t>() + begin,
                                 reinterpret_cast<const uint8_t*>(H_jobs[j].out) + begin,
                                 end - begin, hipMemcpyDeviceToHost, out_stream_));
        HIP_TRY(hipEventRecord(w.copied[b], out_stream_));
        return DICE_OK;
      };
      const int32_t rc = submit();
      if (rc != DICE_OK) outcome[j].fail(rc);
      {
        std::lock_guard<std::mutex> lock(cq.mutex);
        cq.queued[j] = b + 1;
        if (rc != DICE_OK) cq.closed[j] = true;
      }
      {
        std::lock_guard<std::mutex> lock(w.mutex);
        if (rc != DICE_OK)
          w.aborted = true;
        else
          w.published = b + 1;
      }
      w.cv.notify_all();
    }
  });
  // The calling thread reports each job's reconstruction progress band by
  // band and polls its cancel callback once per band (the callbacks belong
  // to the caller's thread).  Bands a job finished before this thread saw
  // them are reported in order, so every job sees exactly one poll per
  // band, then one more before its result is committed.
  std::vector<int> reported(n, 0);
  auto report_bands = [&](int j) {
    JobRequest& R = jobs[j];
    Workspace& w = *ws_[j];
    int published;
    {
      std::lock_guard<std::mutex> lock(w.mutex);
      published = w.published;
    }
    while (reported[j] < published && !cancel[j].load()) {
      reported[j] += 1;
      R.cb.progress(3, band_row(R.height, w.bands, reported[j]), R.height, 0, 0);
      if (R.cb.cancelled()) {
        cancel[j].store(1);
        std::lock_guard<std::mutex> lock(w.mutex);
        w.cv.notify_all();
      }
    }
  };
  bool any_callbacks = false;
  for (int j = 0; j < n; ++j)
    if (!H_jobs[j].skip && jobs[j].cb.cb) any_callbacks = true;
  if (any_callbacks) {
    while (threads_done.load() < (int)threads.size()) {
      for (int j = 0; j < n; ++j)
        if (!H_jobs[j].skip) report_bands(j);
      std::this_thread::sleep_for(std::chrono::microseconds(500));
    }
  }
  for (std::thread& t : threads) t.join();
  for (int j = 0; j < n; ++j)
    if (!H_jobs[j].skip && outcome[j].code == DICE_OK && !finalized[j]) report_bands(j);
  DICE_TRY(wait_stream(stream_));
  DICE_TRY(wait_stream(out_stream_));
  DICE_TRY(wait_stream(fetch_stream_));
  const double t_host = now_seconds();
  for (int j = 0; j < n; ++j)
    if (!H_jobs[j].skip && !finalized[j]) finalize(j, true);
  // Host outputs of the jobs finalized here: pinned staging -> caller, in
  // parallel pieces.
  if (!commit.empty()) {
    struct Piece {
      uint8_t* dst;
      const uint8_t* src;
      size_t len;
    };
    std::vector<Piece> pieces;
    constexpr size_t kCommitPiece = 8u << 20;
    for (int j : commit) {
      const size_t bytes = (size_t)jobs[j].height * jobs[j].width * 6;
      uint8_t* dst = reinterpret_cast<uint8_t*>(jobs[j].out);
      const uint8_t* src = ws_[j]->h_out.as<uint8_t>();
      for (size_t at = 0; at < bytes; at += kCommitPiece)
        pieces.push_back({dst + at, src + at, std::min(kCommitPiece, bytes - at)});
    }
    std::atomic<size_t> next{0};
    auto copier = [&]() {
      for (size_t i; (i = next.fetch_add(1)) < pieces.size();)
        memcpy(pieces[i].dst, pieces[i].src, pieces[i].len);
    };
    const size_t workers = std::min<size_t>(pieces.size(), 8);
    std::vector<std::thread> copiers;
    for (size_t t = 1; t < workers; ++t) copiers.emplace_back(copier);
    copier();
    for (std::thread& t : copiers) t.join();
  }
  if (options.on_job_done)
    for (int j = 0; j < n; ++j)
      if (!early[j] || !finalized[j]) options.on_job_done(j);
  if (trace_enabled()) {
    const double t_end = now_seconds();
    fprintf(stderr,
            "dice-trace runner=%p jobs=%d start=%.4f phaseA=%.2fms phaseB+fetch-setup=%.2fms "
            "host=%.2fms finish=%.2fms\n",
            (void*)this, n, t_start - trace_origin(), (t_a - t_start) * 1e3, (t_b - t_a) * 1e3,
            (t_host - t_b) * 1e3, (t_end - t_host) * 1e3);
  }
  if (report) {
    report->tile_rows = tile_rows_;
    report->tile_cols = tile_cols_;
    report->sites_variant = sites_variant_;
    report->graph_replayed = replayed_a && replayed_b;
    report->phase_a_ms = ms_a;
    report->phase_b_ms = ms_b;
    report->workspace_bytes = workspace_bytes();
  }
  return DICE_OK;
}

int32_t Runner::producer_schedule(const uint16_t* rgbi, int64_t height, int64_t width,
                                  const dice_params& params, dice_schedule* out) {
  HIP_TRY(hipSetDevice(device_));
  JobRequest R;
  R.main = rgbi;
  R.height = height;
  R.width = width;
  R.params = params;
  fine_on_ = false;
  DICE_TRY(prepare(&R, 1, nullptr));
  const JobDev* d = d_jobs_.as<JobDev>();
  const JobDev* h = h_jobs_.as<JobDev>();
  HIP_TRY(launch_zero_jobs(d, h, 1, stream_));
  HIP_TRY(launch_producer(d, h, 1, stream_));
  const size_t H = (size_t)height;
  const size_t sizes[12] = {8, 8, 8, 4, 4, 4, 4, 4, 4, 4, 4, 4};
  for (int i = 0; i < 12; ++i) HIP_TRY(sched_[i].ensure(H * sizes[i]));
  ScheduleDev S;
  S.weighted_visible_sum = sched_[0].as<double>();
  S.weighted_infrared_sum = sched_[1].as<double>();
  S.weight_sum = sched_[2].as<double>();
  S.mean_visible = sched_[3].as<float>();
  S.mean_infrared = sched_[4].as<float>();
  S.accepted = sched_[5].as<uint32_t>();
  S.scaled_input = sched_[6].as<float>();
  S.input_scale = sched_[7].as<float>();
  S.scale_denominator = sched_[8].as<float>();
  S.scale_numerator = sched_[9].as<float>();
  S.scale_add_denominator = sched_[10].as<float>();
  S.scale_add_numerator = sched_[11].as<float>();
  HIP_TRY(launch_schedule_rows(h[0], S, stream_));
  void* dst[12] = {out->weighted_visible_sum, out->weighted_infrared_sum, out->weight_sum,
                   out->mean_visible,         out->mean_infrared,         out->accepted,
                   out->scaled_input,         out->input_scale,           out->scale_denominator,
                   out->scale_numerator,      out->scale_add_denominator, out->scale_add_numerator};
  for (int i = 0; i < 12; ++i)
    HIP_TRY(hipMemcpyAsync(dst[i], sched_[i].ptr, H * sizes[i], hipMemcpyDeviceToHost, stream_));
  DICE_TRY(wait_stream(stream_));
  return DICE_OK;
}

namespace {

// Consecutive 256-byte-aligned pieces of one device block; `used` keeps
// counting past the end, so a failed fit reports what the job needs.
struct Arena {
  uint8_t* base;
  size_t size;
  size_t used = 0;
  template <class T>
  T* take(size_t count) {
    const size_t at = (used + 255) & ~(size_t)255;
    used = at + std::max<size_t>(count, 1) * sizeof(T);
    return used <= size ? reinterpret_cast<T*>(base + at) : nullptr;
  }
  bool fits() const { return used <= size; }
};

}  // namespace

int32_t Runner::run_in_arena(JobRequest& R, uint8_t* arena, size_t arena_bytes) {
  HIP_TRY(hipSetDevice(device_));
  memset(R.result, 0, sizeof *R.result);
  const int64_t H = R.height, W = R.width;
  if (H > kArenaMaxRows)
    return set_error(DICE_OVER_BUDGET, "the allocation-free entry point takes at most %lld rows",
                     (long long)kArenaMaxRows);
  const int TY = tile_rows_, TX = tile_cols_;
  JobDev& J = *h_jobs_.as<JobDev>();
  memset(&J, 0, sizeof J);
  J.P = device_params(R.params);
  J.K = writer_consts(R.params);
  J.factor_b = R.factor_b;
  J.run_producer = 1;
  J.height = (int32_t)H;
  J.width = (int32_t)W;
  J.tile_rows = TY;
  J.tile_cols = TX;
  J.tiles_x = (int32_t)((W + TX - 1) / TX);
  J.tiles_y = (int32_t)((H + TY - 1) / TY);
  J.win = full_window(J.height, J.tiles_y);
  const size_t tiles = (size_t)J.tiles_x * J.tiles_y;
  const size_t segments = (size_t)J.tiles_y * TY * J.tiles_x;
  const size_t stages = (size_t)H + 6;
  const size_t su = 6 * (size_t)W;
  Arena A{arena, arena_bytes};
  J.rgbi = R.main;
  J.out = R.out;
  J.lut = tables_->lut.as<float>();
  J.noop_delta = tables_->noop_delta.as<uint32_t>();
  J.factor_high = tables_->high.as<uint32_t>();
  J.factor_low = tables_->low.as<uint32_t>();
  J.stages.alpha = A.take<float>(stages);
  J.stages.offset = A.take<float>(stages);
  J.stages.replacement = A.take<float>(stages);
  J.stages.has_replacement = A.take<uint8_t>(stages);
  J.stages.score_base = A.take<float>(stages);
  J.stages.coarse = A.take<float>(stages);
  J.stages.gate_secondary = A.take<uint8_t>(stages);
  J.stages.row_gate = A.take<int32_t>(stages);
  J.site_lists = A.take<uint16_t>(tiles * TY * TX);
  J.tile_count = A.take<uint32_t>(tiles);
  J.seg_count = A.take<uint32_t>(segments);
  J.seg_offset = A.take<uint32_t>(segments);
  J.seg_block_sums = A.take<uint32_t>(segments / 1024 + 2);
  J.row_start = A.take<uint32_t>(H + 1);
  J.counters = A.take<uint32_t>(16);
  J.error_flags = A.take<uint32_t>(16);
  J.writer_counts = A.take<unsigned long long>((size_t)kCountSlots * kCountStride);
  J.su_aux = A.take<float>(5 * W);
  J.su_score = A.take<float>(5 * W);
  J.su_work = A.take<float>(15 * W);
  J.su_sites = A.take<Site>(su);
  J.su_attempted = A.take<uint8_t>(su);
  J.su_writer_records = A.take<WriterRecord>(su);
  J.failpos = A.take<uint8_t>(((H + 7) / 8) * (W / 8) + 8);
  J.row_sums = A.take<double>(H * 3);
  J.accepted = A.take<uint32_t>(H);
  J.epoch_adds = A.take<double>((H / 8) * 2 + 2);
  J.cumulative = A.take<double>(H * 3);
  J.epoch_wide = A.take<double>((H / 8) * 2 + 2);
  J.epoch_store = A.take<float>((H / 8) * 2 + 2);
  uint8_t* chain = A.take<uint8_t>(sizeof(ChainState) + kMaxBands * 4);
  J.chain = reinterpret_cast<ChainState*>(chain);
  J.band_end = reinterpret_cast<uint32_t*>(chain + sizeof(ChainState));
  if (!A.fits())
    return set_error(DICE_OVER_BUDGET,
                     "a %lldx%lld job needs %zu bytes of scratch before its sites; the "
                     "allocation-free entry point holds %zu",
                     (long long)H, (long long)W, A.used, arena_bytes);
  const JobDev* d = d_jobs_.as<JobDev>();
  uint8_t* io = h_io_.as<uint8_t>();
  uint32_t* h_row_start = reinterpret_cast<uint32_t*>(io);
  ChainState* h_chain = reinterpret_cast<ChainState*>(io + kArenaChainAt);
  uint32_t* h_counters = reinterpret_cast<uint32_t*>(io + kArenaSmallAt);
  uint32_t* h_flags = reinterpret_cast<uint32_t*>(io + kArenaSmallAt + 64);
  const unsigned long long* h_slots =
      reinterpret_cast<const unsigned long long*>(io + kArenaSlotsAt);
  fine_on_ = false;

  // phase A: zero, producer + stage table, analysis scan, segment scan
  HIP_TRY(hipMemcpyAsync(d_jobs_.ptr, h_jobs_.ptr, sizeof(JobDev), hipMemcpyHostToDevice,
                         stream_));
  DICE_TRY(launch_a(1));
  HIP_TRY(hipMemcpyAsync(h_counters, J.counters, 64, hipMemcpyDeviceToHost, stream_));
  HIP_TRY(hipMemcpyAsync(h_flags, J.error_flags, 4, hipMemcpyDeviceToHost, stream_));
  DICE_TRY(wait_stream(stream_));
  DICE_TRY(flag_error(*h_flags));
  if (R.cb.cancelled()) return cancelled_error();

  // the site arrays, now that the site count is known
  const size_t count = h_counters[0];
  const size_t groups = count / kStateGroup + 2;
  J.sites = A.take<Site>(count + 1);
  J.writer_records = A.take<WriterRecord>(count + 1);
  J.su_site_state = A.take<uint32_t>(su + count + 1 + groups);
  J.site_state = J.su_site_state + su;
  J.group_state = J.site_state + count + 1;
  if (!A.fits())
    return set_error(DICE_OVER_BUDGET,
                     "a %lldx%lld job with %zu attempted sites needs %zu bytes of scratch; the "
                     "allocation-free entry point holds %zu",
                     (long long)H, (long long)W, count, A.used, arena_bytes);
  HIP_TRY(hipMemcpyAsync(d_jobs_.ptr, h_jobs_.ptr, sizeof(JobDev), hipMemcpyHostToDevice,
                         stream_));

  // phase B: sites, hidden startup, draw schedules; then the serial chain on
  // one device wavefront, band by band, each band followed by its writer
  DICE_TRY(launch_b(1));
  HIP_TRY(hipMemcpyAsync(h_row_start, J.row_start, ((size_t)H + 1) * 4, hipMemcpyDeviceToHost,
                         stream_));
  memset(h_chain, 0, sizeof(ChainState));
  h_chain->state = R.rng_state;
  HIP_TRY(hipMemcpyAsync(J.chain, h_chain, sizeof(ChainState), hipMemcpyHostToDevice, stream_));
  DICE_TRY(wait_stream(stream_));
  HIP_TRY(launch_writer_startup_chain(d, 1, stream_));
  HIP_TRY(launch_writer_startup_check(J, 0, &J.chain->startup_final, stream_));
  const int bands = std::max<int>(1, std::min<int64_t>(std::min(8, kMaxBands), H));
  for (int b = 0; b < bands; ++b) {
    HIP_TRY(launch_writer_chain(d, 1, b, bands, stream_));
    const uint32_t s0 = h_row_start[band_row(H, bands, b)];
    const uint32_t s1 = h_row_start[band_row(H, bands, b + 1)];
    HIP_TRY(launch_writer_values(J, s0, s1, 0, J.band_end + b, stream_));
  }
  HIP_TRY(hipMemcpyAsync(h_chain, J.chain, sizeof(ChainState), hipMemcpyDeviceToHost, stream_));
  HIP_TRY(hipMemcpyAsync(io + kArenaSlotsAt, J.writer_counts,
                         (size_t)kCountSlots * kCountStride * 8, hipMemcpyDeviceToHost, stream_));
  HIP_TRY(hipMemcpyAsync(h_flags, J.error_flags, 4, hipMemcpyDeviceToHost, stream_));
  DICE_TRY(wait_stream(stream_));
  DICE_TRY(flag_error(*h_flags));
  if (R.cb.cancelled()) return cancelled_error();
  uint64_t written = 0, changed = 0;
  for (int slot = 0; slot < kCountSlots; ++slot) {
    written += h_slots[slot * kCountStride + 0];
    changed += h_slots[slot * kCountStride + 1];
  }
  dice_replay_result& res = *R.result;
  res.attempted_pixels = h_row_start[H];
  res.written_pixels = written;
  res.changed_pixels = changed;
  res.public_rng_advances = h_chain->advances;
  res.final_rng_state = h_chain->state;
  for (int stage = 0; stage < 6; ++stage) {
    res.startup_attempted[stage] = h_chain->startup_attempted[stage];
    res.startup_advances[stage] = h_chain->startup_advances[stage];
  }
  res.startup_final_rng_state = h_chain->startup_final;
  return DICE_OK;
}

namespace {

struct TuneEntry {
  int rows = 16, cols = 64, variant = 0;
  double ms = 0.0;
  bool from_file = false;
};

struct TuningTable {
  std::mutex mutex;
  std::string path;
  bool path_set = false, loaded = false;
  std::map<std::string, TuneEntry> entries;  // "arch|bucket"
};

TuningTable& tuning() {
  static TuningTable table;
  return table;
}

std::string tuning_path_locked(TuningTable& t) {
  if (!t.path_set) {
    t.path_set = true;
    if (const char* env = getenv("DICE_TUNING_CACHE")) {
      t.path = env;
    } else if (const char* home = getenv("HOME")) {
      if (*home) t.path = std::string(home) + "/.cache/dice_hip/tuning.tsv";
    }
  }
  return t.path;
}

// One winner per line: arch, bucket, rows, cols, sites variant, ms.
void load_tuning_locked(TuningTable& t) {
  if (t.loaded) return;
  t.loaded = true;
  const std::string path = tuning_path_locked(t);
  if (path.empty()) return;
  FILE* f = fopen(path.c_str(), "r");
  if (!f) return;
  char arch[64], bucket[96];
  TuneEntry e;
  while (fscanf(f, "%63s %95s %d %d %d %lf", arch, bucket, &e.rows, &e.cols, &e.variant,
                &e.ms) == 6) {
    if (!tile_shape_supported(e.rows, e.cols) || e.variant < 0 || e.variant >= kSitesVariants)
      continue;
    e.from_file = true;
    t.entries[std::string(arch) + "|" + bucket] = e;
  }
  fclose(f);
}

// Rewrites the cache file with every winner held (a temporary file renamed
// over it); a cache that cannot be written leaves the winners in memory.
void save_tuning_locked(TuningTable& t) {
  const std::string path = tuning_path_locked(t);
  if (path.empty()) return;
  for (size_t at = path.find('/', 1); at != std::string::npos; at = path.find('/', at + 1))
    (void)mkdir(path.substr(0, at).c_str(), 0755);
  const std::string tmp = path + ".tmp" + std::to_string((long)getpid());
  FILE* f = fopen(tmp.c_str(), "w");
  if (!f) return;
  for (const auto& kv : t.entries) {
    const size_t bar = kv.first.find('|');
    fprintf(f, "%s\t%s\t%d\t%d\t%d\t%.4f\n", kv.first.substr(0, bar).c_str(),
            kv.first.substr(bar + 1).c_str(), kv.second.rows, kv.second.cols,
            kv.second.variant, kv.second.ms);
  }
  const bool ok = fclose(f) == 0;
  if (!ok || rename(tmp.c_str(), path.c_str()) != 0) (void)remove(tmp.c_str());
}

int log2_floor(uint64_t v) {
  int b = 0;
  while (v >> (b + 1)) ++b;
  return b;
}

}  // namespace

void set_tuning_cache(const char* path) {
  TuningTable& t = tuning();
  std::lock_guard<std::mutex> lock(t.mutex);
  t.path_set = path != nullptr;
  t.path = path ? path : "";
  t.entries.clear();
  t.loaded = false;
}

int32_t Runner::autotune(JobRequest* jobs, int n, RunReport* report) {
  if (n < 1 || n > kMaxBatch) return set_error(DICE_INVALID, "a batch holds 1..%d jobs", kMaxBatch);
  HIP_TRY(hipSetDevice(device_));
  uint64_t total = 0, largest = 0;
  int live = 0;
  for (int j = 0; j < n; ++j) {
    if (jobs[j].code != DICE_OK) continue;
    const uint64_t pixels = (uint64_t)jobs[j].height * (uint64_t)jobs[j].width;
    total += pixels;
    largest = std::max(largest, pixels);
    live += 1;
  }
  if (live == 0) return DICE_OK;
  char bucket[96];
  snprintf(bucket, sizeof bucket, "n%d-t%d-m%d", 1 << log2_floor((uint64_t)(2 * live - 1)),
           log2_floor(total), log2_floor(largest));
  const std::string key = arch_ + "|" + bucket;
  TuningTable& table = tuning();
  {
    std::lock_guard<std::mutex> lock(table.mutex);
    load_tuning_locked(table);
    auto it = table.entries.find(key);
    if (it != table.entries.end()) {
      set_tile(it->second.rows, it->second.cols, it->second.variant);
      if (report) {
        report->tile_rows = tile_rows_;
        report->tile_cols = tile_cols_;
        report->sites_variant = sites_variant_;
        report->tune_source = it->second.from_file ? 3 : 2;
      }
      return DICE_OK;
    }
  }
  // Every configuration, the default (16 x 64, variant 0) first: its
  // results are the reference the others must reproduce.
  struct Candidate {
    int rows, cols, variant;
  };
  std::vector<Candidate> candidates;
  candidates.push_back({16, 64, 0});
  int shapes[16][2];
  const int count = std::min(16, tile_shapes(shapes, 16));
  for (int s = 0; s < count; ++s)
    for (int v = 0; v < kSitesVariants; ++v)
      if (!(shapes[s][0] == 16 && shapes[s][1] == 64 && v == 0))
        candidates.push_back({shapes[s][0], shapes[s][1], v});
  // trial copies: results, diagnostics and outputs of the timing runs are
  // not the caller's
  std::vector<JobRequest> trial(jobs, jobs + n);
  std::vector<dice_replay_result> scratch(n);
  for (int j = 0; j < n; ++j) {
    trial[j].result = &scratch[j];
    trial[j].diag = nullptr;
    trial[j].timings = nullptr;
    trial[j].cb = Callbacks{};
  }
  std::vector<unsigned long long> reference(n, 0), digest(n, 0);
  std::vector<uint32_t> reference_sites(n, 0);
  double default_ms = 0.0, best_ms = 1e30;
  Candidate best = candidates[0];
  int rejected = 0;
  for (size_t c = 0; c < candidates.size(); ++c) {
    set_tile(candidates[c].rows, candidates[c].cols, candidates[c].variant);
    double shape_ms = 1e30;
    bool agrees = true;
    for (int rep = 0; rep < 3 && agrees; ++rep) {
      fine_on_ = false;
      DICE_TRY(prepare(trial.data(), n, nullptr));
      HIP_TRY(hipEventRecord(ev_[4], stream_));
      DICE_TRY(launch_a(n));
      JobDev* H_jobs = h_jobs_.as<JobDev>();
      for (int j = 0; j < n; ++j)
        if (!H_jobs[j].skip)
          HIP_TRY(hipMemcpyAsync(ws_[j]->h_small.ptr, ws_[j]->counters.ptr, 64,
                                 hipMemcpyDeviceToHost, stream_));
      HIP_TRY(hipEventRecord(ev_[5], stream_));
      DICE_TRY(wait_stream(stream_));
      for (int j = 0; j < n; ++j) {
        if (H_jobs[j].skip) continue;
        Workspace& w = *ws_[j];
        const size_t sites = w.h_small.as<uint32_t>()[0];
        HIP_TRY(w.sites.ensure((sites + 1) * sizeof(Site), stream_));
        HIP_TRY(w.records.ensure((sites + 1) * sizeof(WriterRecord), stream_));
        H_jobs[j].sites = w.sites.as<Site>();
        H_jobs[j].writer_records = w.records.as<WriterRecord>();
      }
      HIP_TRY(hipMemcpyAsync(d_jobs_.ptr, h_jobs_.ptr, sizeof(JobDev) * n,
                             hipMemcpyHostToDevice, stream_));
      HIP_TRY(hipEventRecord(ev_[6], stream_));
      DICE_TRY(launch_b(n));
      HIP_TRY(hipEventRecord(ev_[7], stream_));
      if (rep == 0) {  // this configuration's results, digested per job
        unsigned long long* d_digest = digest_.as<unsigned long long>();
        HIP_TRY(launch_fill_u8(reinterpret_cast<uint8_t*>(d_digest), 8 * (size_t)n, 0, stream_));
        for (int j = 0; j < n; ++j) {
          const JobDev& J = H_jobs[j];
          if (J.skip) continue;
          const size_t sites = ws_[j]->h_small.as<uint32_t>()[0];
          HIP_TRY(launch_digest(J.sites, sites * sizeof(Site), d_digest + j, stream_));
          HIP_TRY(launch_digest(J.writer_records, sites * sizeof(WriterRecord), d_digest + j,
                                stream_));
          HIP_TRY(launch_digest(J.su_writer_records, 6 * (size_t)J.width * sizeof(WriterRecord),
                                d_digest + j, stream_));
          HIP_TRY(launch_digest(J.row_start, ((size_t)J.height + 1) * 4, d_digest + j, stream_));
          HIP_TRY(launch_digest(J.out, (size_t)J.height * J.width * 6, d_digest + j, stream_));
        }
        HIP_TRY(hipMemcpyAsync(digest.data(), d_digest, 8 * (size_t)n, hipMemcpyDeviceToHost,
                               stream_));
      }
      DICE_TRY(wait_stream(stream_));
      if (rep == 0) {
        for (int j = 0; j < n; ++j) {
          if (H_jobs[j].skip) continue;
          const uint32_t sites = ws_[j]->h_small.as<uint32_t>()[0];
          if (c == 0) {
            reference[j] = digest[j];
            reference_sites[j] = sites;
          } else if (digest[j] != reference[j] || sites != reference_sites[j]) {
            agrees = false;
          }
        }
        if (!agrees) break;
      }
      float a = 0.0f, b = 0.0f;
      (void)hipEventElapsedTime(&a, ev_[4], ev_[5]);
      (void)hipEventElapsedTime(&b, ev_[6], ev_[7]);
      shape_ms = std::min(shape_ms, (double)a + (double)b);
    }
    if (!agrees) {
      rejected += 1;
      continue;
    }
    if (c == 0) default_ms = shape_ms;
    if (shape_ms < best_ms) {
      best_ms = shape_ms;
      best = candidates[c];
    }
  }
  // keep the default unless a configuration is clearly faster
  if (best_ms > 0.98 * default_ms) {
    best = candidates[0];
    best_ms = default_ms;
  }
  set_tile(best.rows, best.cols, best.variant);
  {
    std::lock_guard<std::mutex> lock(table.mutex);
    TuneEntry e;
    e.rows = best.rows;
    e.cols = best.cols;
    e.variant = best.variant;
    e.ms = best_ms;
    table.entries[key] = e;
    save_tuning_locked(table);
  }
  if (report) {
    report->tile_rows = tile_rows_;
    report->tile_cols = tile_cols_;
    report->sites_variant = sites_variant_;
    report->tune_source = 1;
    report->tune_candidates = (int)candidates.size();
    report->tune_rejected = rejected;
  }
  return DICE_OK;
}

}  // namespace dice
