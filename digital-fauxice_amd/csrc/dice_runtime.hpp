// dice_runtime.hpp -- the job runtime of libdice_hip.so: one Runner executes
// a batch of 1..kMaxBatch independent jobs on one HIP stream.
//
//   phase A (device, one persistent launch per kernel for the whole batch):
//     zero flags -> producer + setter -> analysis scan -> segment scan
//   host: site counts, workspace sizing
//   phase B (device): sites -> hidden startup -> writer draw schedules
//   host, per job:
//     chain thread   the serial LCG chain over the draw schedules, in row
//                    bands; after each band its start states go to the
//                    device and the band's writer values are launched
//     hash thread    waits for each band, downloads its output rows and
//                    feeds SHA-256 (so the chain and the hash overlap); on a
//                    busy host, one thread per pair of jobs of one geometry,
//                    the two messages compressed interleaved
//
// Phases A and B have static launch geometry (site counts are read on the
// device), so a repeated batch replays them from captured hipGraphs.
#pragma once

#include <hip/hip_runtime.h>

#include <functional>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "dice_common.hpp"
#include "dice_device.hpp"

namespace dice {

DevParams device_params(const dice_params& p);
WriterConsts writer_consts(const dice_params& p);

// When the jobs of one geometry in a batch are hashed in pairs: never,
// always, or when the messages being hashed in the process would outnumber
// the host's cores (the default; DICE_HASH_PAIRS=0 / =1 select the others).
constexpr int kHashPairsNever = 0, kHashPairsAlways = 1, kHashPairsWhenBusy = 2;
// Sets the mode for the batches that start from now on; returns the old one.
int set_hash_pair_mode(int mode);

// One job as the caller described it, plus its outcome.
struct JobRequest {
  const uint16_t* main = nullptr;
  int64_t height = 0, width = 0;
  uint32_t flags = 0;
  uint16_t* out = nullptr;
  dice_params params{};
  const dice_stage_table* stages = nullptr;  // nullptr: producer on device
  float factor_b = 1.0f;
  uint32_t rng_state = 0x3045u;
  dice_diagnostics* diag = nullptr;
  Callbacks cb;
  dice_replay_result* result = nullptr;
  dice_stage_timings* timings = nullptr;
  bool streamed = false;  // main rows arrive later, block by block (FrameStream)
  // outcome
  int32_t code = DICE_OK;
  std::string error;
};

struct Workspace;

struct RunOptions {
  bool graph = false;        // replay phases A/B from captured graphs
  bool fine_timing = false;  // HIP events between kernel groups (direct launches only)
  bool device_chain = false; // the serial writer chain on one device wavefront per job
  int bands = 8;             // writer/hash pipeline bands per job
  // Called once per job as soon as that job's outcome (jobs[j].code, its
  // result and output) is final; from a worker thread for a job without
  // callbacks, else from the calling thread before run() returns.
  std::function<void(int)> on_job_done;
};

struct RunReport {
  int tile_rows = 0, tile_cols = 0;
  int sites_variant = 0;
  // how the launch configuration was chosen: 0 not tuned, 1 timed now,
  // 2 a winner this process already holds for the batch's bucket, 3 a
  // winner loaded from the tuning cache file
  int tune_source = 0;
  int tune_candidates = 0, tune_rejected = 0;
  bool graph_replayed = false;
  double phase_a_ms = 0.0, phase_b_ms = 0.0;  // device time (HIP events)
  uint64_t workspace_bytes = 0;
};

// Device-resident tables shared by every job of a device.
struct DeviceTables {
  DeviceBuffer lut, high, low, noop_delta;
  int32_t upload();
};

class Runner {
 public:
  Runner(int device, const DeviceTables* tables);
  ~Runner();
  Runner(const Runner&) = delete;
  Runner& operator=(const Runner&) = delete;
  int32_t init();
  // Runs jobs[0..n) as one batch.  Returns DICE_OK when the batch ran
  // (each job's own status is in jobs[i].code / .error), or the code of a
  // failure that prevented the batch from running at all.
  int32_t run(JobRequest* jobs, int n, const RunOptions& options, RunReport* report);
  // The launch configuration (tile shape, k_sites variant) for this batch's
  // bucket (the device architecture, the job count and the total and
  // largest frame sizes, each to a power of two): a winner already held or
  // persisted in the tuning cache, or else every configuration timed on
  // this batch (phases A+B) with its sites, draw schedules, row starts and
  // output digested and compared with the default configuration's -- a
  // configuration whose results differ is rejected -- and the fastest kept
  // when it beats the default by more than 2 %.  The timing runs' results
  // are discarded.
  int32_t autotune(JobRequest* jobs, int n, RunReport* report);
  // The producer record schedule of one frame (arrays only; the caller
  // packs the payloads).
  int32_t producer_schedule(const uint16_t* rgbi, int64_t height, int64_t width,
                            const dice_params& params, dice_schedule* out);
  // One job whose device intermediates are all carved from `arena`
  // (`arena_bytes` long, device memory the caller keeps) instead of this
  // runner's workspaces.  The main frame and the output are device memory
  // (R.main, R.out); the writer chain runs on one device wavefront; no
  // SHA-256 and no diagnostics.  Allocates nothing: the runner's own
  // buffers come from init().  Frames of more than kArenaMaxRows rows, or
  // whose intermediates do not fit the arena, are refused with
  // DICE_OVER_BUDGET before the device phase that would need the space.
  int32_t run_in_arena(JobRequest& R, uint8_t* arena, size_t arena_bytes);
  static constexpr int64_t kArenaMaxRows = 16384;
  void set_tile(int rows, int cols, int sites_variant = 0) {
    tile_rows_ = rows;
    tile_cols_ = cols;
    sites_variant_ = sites_variant;
  }
  int tile_rows() const { return tile_rows_; }
  int tile_cols() const { return tile_cols_; }
  hipStream_t stream() const { return stream_; }
  uint64_t workspace_bytes() const;
  // Frees every job workspace (the next batch grows them again).
  int32_t release_workspaces();

 private:
  friend class FrameStream;
  struct Graph {
    std::string key;
    hipGraphExec_t exec = nullptr;
    int seen = 0;
  };
  int32_t prepare(JobRequest* jobs, int n, bool* any_host_work);
  // Host frames to device buffers through pinned staging (prepare).
  struct Upload {
    const uint8_t* src;
    uint8_t* dst;
    size_t bytes;
    struct Workspace* w;
  };
  int32_t upload_frames(const std::vector<Upload>& uploads);
  int32_t phase_a(int n, bool graph, bool* replayed);
  int32_t phase_b(int n, bool graph, bool* replayed);
  int32_t replay_or_capture(Graph& g, const std::string& key, bool graph, bool* replayed,
                            int32_t (Runner::*launch)(int), int n);
  int32_t launch_a(int n);
  int32_t launch_b(int n);
  int32_t mark(int slot);
  int32_t wait_stream(hipStream_t stream);
  double fine_ms(int a, int b) const;
  std::string geometry_key(int n, char phase) const;

  int device_;
  const DeviceTables* tables_;
  hipStream_t stream_ = nullptr;
  hipStream_t fetch_stream_ = nullptr;  // draw schedules to the host (chain)
  hipStream_t out_stream_ = nullptr;    // output rows to the host, lowest band first
  hipEvent_t ev_[8] = {};
  hipEvent_t fine_[8] = {};
  bool fine_on_ = false;
  hipEvent_t sync_event_ = nullptr;
  int tile_rows_ = 16, tile_cols_ = 64;
  int sites_variant_ = 0;
  std::string arch_;  // gfx architecture of the device, e.g. "gfx950"
  std::vector<std::unique_ptr<Workspace>> ws_;
  DeviceBuffer d_jobs_;
  DeviceBuffer queue_;  // ticket counter of the dynamically scheduled launches (zero at rest)
  PinnedBuffer h_jobs_;
  PinnedBuffer h_io_;  // run_in_arena's host side: row starts, chain state, counts, flags
  Graph graph_a_, graph_b_;
  DeviceBuffer digest_;  // the tuner's result digests, one u64 per job
  DeviceBuffer sched_[12];
};

// Validation, prepass and one batch run of caller-described jobs (the body
// of dice_process_batch, also used by the serving engine).
// An upper bound of the device bytes a job can add to the workspace of the
// Runner slot it runs in (every buffer grown with its headroom and 1 MiB
// rounding, every pixel an attempted site, any compiled tile shape): the
// unit of the serving engine's HBM admission.
uint64_t job_reservation(int64_t height, int64_t width, uint32_t flags, bool diagnostics);

// The tuner's winners, process-wide, persisted in `path` (empty: the
// default, $DICE_TUNING_CACHE or ~/.cache/dice_hip/tuning.tsv); setting it
// forgets the winners held in memory, so the next lookup reads the file.
void set_tuning_cache(const char* path);

// `on_done(j)`, when given, is called once per job as soon as jobs[j]'s
// status, error and result are final (possibly from a worker thread, before
// the batch ends); the job's dice_job is not touched after that.
int32_t process_jobs(Runner& runner, dice_job* jobs, int32_t count, uint32_t batch_flags,
                     dice_batch_report* report,
                     const std::function<void(int)>& on_done = nullptr);

}  // namespace dice
