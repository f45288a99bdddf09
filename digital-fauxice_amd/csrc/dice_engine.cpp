// dice_engine.cpp -- the serving loop of libdice_hip.so (no counterpart in
// the reference, which processes one job per call).
//
// Jobs are submitted as they arrive and wait in a FIFO queue.  Several
// drivers -- each a thread with its own Runner, i.e. its own HIP streams and
// job workspaces -- take batches from the head of the queue whenever they
// are free, so jobs that arrive while others run join the next batch
// (continuous batching), and the device phases of one batch overlap the
// host chain and hash of the others.  Each job completes on its own, as
// soon as its own hash and checks are done, without waiting for the rest of
// its batch.  Every job's result is the one it gets alone: jobs share
// nothing but the device (rng.py:17-24: each job owns its generator).
//
// HBM admission.  A job's reservation bounds what it can add to the
// workspace of the slot it runs in (job_reservation).  A driver's Runner
// holds its workspaces between batches (`held`); while it runs a batch it
// can hold at most held + the batch's reservations (`bound`).  A batch is
// admitted only while the sum over drivers of bound (running) or held
// (idle) stays within the budget, so the workspaces the engine holds never
// exceed it.  When the queue's head does not fit, idle drivers free their
// workspaces, so a job that fits the budget alone is always admitted
// eventually.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/dice_hip.h"
#include "dice_common.hpp"
#include "dice_runtime.hpp"

#pragma clang fp contract(off)

using namespace dice;

namespace {

struct Ticket {
  dice_job* job = nullptr;
  uint64_t id = 0;
  uint64_t reserve = 0;
  double submitted = 0.0, started = 0.0, finished = 0.0;
  bool done = false;
};

uint64_t reservation(const dice_job& job) {
  return job_reservation(job.height, job.width, job.flags, false);
}

}  // namespace

struct dice_engine {
  dice_engine_config config{};
  DeviceTables tables;
  std::vector<std::unique_ptr<Runner>> runners;
  std::vector<std::thread> drivers;
  std::mutex mutex;
  std::condition_variable work, finished;
  std::deque<Ticket*> queue;
  std::map<uint64_t, std::unique_ptr<Ticket>> tickets;
  uint64_t next_id = 1;
  bool stopping = false;
  // HBM accounting, per driver
  std::vector<uint64_t> held;   // workspace bytes its Runner holds while idle
  std::vector<uint64_t> bound;  // held + reservations of the batch it runs
  std::vector<char> busy;
  uint64_t committed_peak = 0, held_peak = 0, releases = 0;
  // accounting
  uint64_t submitted = 0, completed = 0, failed = 0, running = 0, batches = 0, batch_jobs = 0;
  double mpix = 0.0, first_submit = 0.0, last_finish = 0.0;
  std::vector<double> latency_ms, queue_ms;

  // The most HBM the drivers' workspaces can hold right now (mutex held).
  uint64_t committed() const {
    uint64_t total = 0;
    for (size_t i = 0; i < held.size(); ++i) total += busy[i] ? bound[i] : held[i];
    return total;
  }
  uint64_t held_total() const {
    uint64_t total = 0;
    for (uint64_t h : held) total += h;
    return total;
  }
  // One job's outcome is final (mutex held).
  void complete(Ticket* t, const dice_job& outcome, double now) {
    dice_job& caller = *t->job;
    caller.status = outcome.status;
    memcpy(caller.error, outcome.error, sizeof caller.error);
    t->finished = now;
    t->done = true;
    running -= 1;
    if (caller.status == DICE_OK) {
      completed += 1;
      mpix += (double)caller.height * (double)caller.width / 1e6;
    } else {
      failed += 1;
    }
    latency_ms.push_back((t->finished - t->submitted) * 1e3);
    queue_ms.push_back((t->started - t->submitted) * 1e3);
    last_finish = now;
  }
  void drive(int index);
};

void dice_engine::drive(int index) {
  Runner& runner = *runners[index];
  (void)hipSetDevice(config.device);
  std::vector<Ticket*> batch;
  std::vector<dice_job> jobs;
  std::vector<char> done;
  for (;;) {
    batch.clear();
    bool release = false;
    {
      std::unique_lock<std::mutex> lock(mutex);
      for (;;) {
        if (stopping && queue.empty()) return;
        if (!queue.empty()) {
          if (committed() + queue.front()->reserve <= config.hbm_budget) break;
          if (held[index] > 0) {  // the head does not fit: free what this idle driver holds
            release = true;
            break;
          }
        }
        work.wait(lock);
      }
      if (!release) {
        // While other batches keep the device busy, give a partial batch a
        // short linger to fill up (more jobs per persistent launch); an
        // idle device starts at once.
        if ((int)queue.size() < config.max_batch && running > 0 && config.linger_ms > 0.0) {
          work.wait_for(lock, std::chrono::duration<double, std::milli>(config.linger_ms), [&] {
            return stopping || (int)queue.size() >= config.max_batch;
          });
        }
        uint64_t reserved = 0;
        const uint64_t others = committed();
        while (!queue.empty() && (int)batch.size() < config.max_batch) {
          Ticket* t = queue.front();
          if (others + reserved + t->reserve > config.hbm_budget) break;
          queue.pop_front();
          reserved += t->reserve;
          batch.push_back(t);
        }
        if (batch.empty()) continue;  // another driver took the head meanwhile
        busy[index] = 1;
        bound[index] = held[index] + reserved;
        committed_peak = std::max(committed_peak, committed());
        running += batch.size();
        const double now = now_seconds();
        for (Ticket* t : batch) t->started = now;
      }
    }
    if (release) {
      const int32_t rc = runner.release_workspaces();
      {
        std::lock_guard<std::mutex> lock(mutex);
        held[index] = rc == DICE_OK ? runner.workspace_bytes() : held[index];
        releases += 1;
      }
      work.notify_all();
      continue;
    }
    jobs.resize(batch.size());
    done.assign(batch.size(), 0);
    for (size_t i = 0; i < batch.size(); ++i) jobs[i] = *batch[i]->job;
    auto on_done = [&](int j) {
      {
        std::lock_guard<std::mutex> lock(mutex);
        complete(batch[j], jobs[j], now_seconds());
        done[j] = 1;
      }
      finished.notify_all();
    };
    const int32_t rc = process_jobs(runner, jobs.data(), (int32_t)jobs.size(),
                                    config.batch_flags, nullptr, on_done);
    const std::string batch_error = rc != DICE_OK ? error_buffer() : std::string();
    const uint64_t now_held = runner.workspace_bytes();
    {
      std::lock_guard<std::mutex> lock(mutex);
      // a batch that could not run at all fails every job not yet completed
      const double now = now_seconds();
      for (size_t i = 0; i < batch.size(); ++i) {
        if (done[i]) continue;
        if (jobs[i].status == DICE_OK || jobs[i].status == DICE_PENDING) {
          jobs[i].status = rc != DICE_OK ? rc : DICE_INTERNAL;
          snprintf(jobs[i].error, sizeof jobs[i].error, "%s",
                   rc != DICE_OK ? batch_error.c_str() : "job did not complete");
        }
        complete(batch[i], jobs[i], now);
      }
      held[index] = now_held;
      busy[index] = 0;
      bound[index] = 0;
      held_peak = std::max(held_peak, held_total());
      batches += 1;
      batch_jobs += batch.size();
    }
    finished.notify_all();
    work.notify_all();
  }
}

namespace {
double percentile(std::vector<double> v, double q) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const double pos = q * (double)(v.size() - 1);
  const size_t lo = (size_t)pos;
  const size_t hi = std::min(lo + 1, v.size() - 1);
  return v[lo] + (v[hi] - v[lo]) * (pos - (double)lo);
}
}  // namespace

extern "C" {

int32_t dice_engine_create(const dice_engine_config* config, dice_engine** out) {
  *out = nullptr;
  DICE_TRY(require_tables());
  dice_engine_config c = *config;
  if (c.drivers < 1) c.drivers = 3;  // full frames in batches of 8: the measured best
  if (c.drivers > 16) return set_error(DICE_INVALID, "at most 16 drivers");
  if (c.max_batch < 1) c.max_batch = 8;
  if (c.max_batch > kMaxBatch) return set_error(DICE_INVALID, "max_batch must be 1..%d", kMaxBatch);
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || c.device < 0 || c.device >= count)
    return set_error(DICE_UNAVAILABLE, "HIP device %d is not available", c.device);
  HIP_TRY(hipSetDevice(c.device));
  if (c.hbm_budget == 0) {
    size_t free_bytes = 0, total_bytes = 0;
    HIP_TRY(hipMemGetInfo(&free_bytes, &total_bytes));
    c.hbm_budget = (uint64_t)(0.8 * (double)free_bytes);
  }
  std::unique_ptr<dice_engine> engine(new dice_engine());
  engine->config = c;
  engine->held.assign(c.drivers, 0);
  engine->bound.assign(c.drivers, 0);
  engine->busy.assign(c.drivers, 0);
  if (int32_t rc = engine->tables.upload()) return rc;
  for (int i = 0; i < c.drivers; ++i) {
    engine->runners.emplace_back(new Runner(c.device, &engine->tables));
    DICE_TRY(engine->runners.back()->init());
  }
  for (int i = 0; i < c.drivers; ++i)
    engine->drivers.emplace_back([e = engine.get(), i] { e->drive(i); });
  *out = engine.release();
  return DICE_OK;
}

void dice_engine_destroy(dice_engine* engine) {
  if (!engine) return;
  {
    std::lock_guard<std::mutex> lock(engine->mutex);
    engine->stopping = true;
  }
  engine->work.notify_all();
  for (std::thread& t : engine->drivers) t.join();
  (void)hipSetDevice(engine->config.device);
  engine->runners.clear();
  delete engine;
}

int32_t dice_engine_submit(dice_engine* engine, dice_job* job, uint64_t* ticket) {
  if (!engine || !job || !ticket) return set_error(DICE_INVALID, "null engine, job or ticket");
  const uint64_t reserve = reservation(*job);
  if (reserve > engine->config.hbm_budget)
    return set_error(DICE_OVER_BUDGET,
                     "job needs up to %llu bytes of HBM; the engine budget is %llu",
                     (unsigned long long)reserve,
                     (unsigned long long)engine->config.hbm_budget);
  {
    std::lock_guard<std::mutex> lock(engine->mutex);
    if (engine->stopping) return set_error(DICE_UNAVAILABLE, "engine is shutting down");
    std::unique_ptr<Ticket> t(new Ticket());
    t->job = job;
    t->id = engine->next_id++;
    t->reserve = reserve;
    t->submitted = now_seconds();
    if (engine->first_submit == 0.0) engine->first_submit = t->submitted;
    engine->submitted += 1;
    job->status = DICE_PENDING;
    job->error[0] = 0;
    *ticket = t->id;
    engine->queue.push_back(t.get());
    engine->tickets.emplace(t->id, std::move(t));
  }
  engine->work.notify_all();
  return DICE_OK;
}

int32_t dice_engine_reservation(const dice_job* job, uint64_t* bytes) {
  if (!job || !bytes) return set_error(DICE_INVALID, "null job or bytes");
  *bytes = reservation(*job);
  return DICE_OK;
}

int32_t dice_engine_wait(dice_engine* engine, uint64_t ticket, double timeout_seconds) {
  if (!engine) return set_error(DICE_INVALID, "null engine");
  std::unique_lock<std::mutex> lock(engine->mutex);
  auto it = engine->tickets.find(ticket);
  if (it == engine->tickets.end()) return set_error(DICE_INVALID, "unknown ticket %llu",
                                                    (unsigned long long)ticket);
  Ticket* t = it->second.get();
  auto ready = [&] { return t->done; };
  if (timeout_seconds < 0) {
    engine->finished.wait(lock, ready);
  } else if (!engine->finished.wait_for(lock, std::chrono::duration<double>(timeout_seconds),
                                        ready)) {
    return set_error(DICE_PENDING, "job %llu has not finished", (unsigned long long)ticket);
  }
  const int32_t status = t->job->status;
  if (status != DICE_OK) set_error(status, "%s", t->job->error);
  engine->tickets.erase(it);
  return status;
}

int32_t dice_engine_reset_stats(dice_engine* engine) {
  if (!engine) return set_error(DICE_INVALID, "null engine");
  std::lock_guard<std::mutex> lock(engine->mutex);
  engine->completed = engine->failed = engine->batches = engine->batch_jobs = 0;
  engine->submitted = engine->queue.size() + engine->running;
  engine->mpix = 0.0;
  engine->latency_ms.clear();
  engine->queue_ms.clear();
  engine->committed_peak = engine->committed();
  engine->held_peak = engine->held_total();
  engine->releases = 0;
  engine->first_submit = now_seconds();
  engine->last_finish = engine->first_submit;
  return DICE_OK;
}

int32_t dice_engine_stats(dice_engine* engine, dice_engine_report* stats) {
  if (!engine || !stats) return set_error(DICE_INVALID, "null engine or stats");
  memset(stats, 0, sizeof *stats);
  std::lock_guard<std::mutex> lock(engine->mutex);
  stats->submitted = engine->submitted;
  stats->completed = engine->completed;
  stats->failed = engine->failed;
  stats->queued = engine->queue.size();
  stats->running = engine->running;
  stats->batches = engine->batches;
  stats->mean_batch = engine->batches ? (double)engine->batch_jobs / (double)engine->batches : 0.0;
  stats->mpix_completed = engine->mpix;
  stats->latency_p50_ms = percentile(engine->latency_ms, 0.50);
  stats->latency_p99_ms = percentile(engine->latency_ms, 0.99);
  stats->latency_max_ms = percentile(engine->latency_ms, 1.0);
  stats->queue_p50_ms = percentile(engine->queue_ms, 0.50);
  stats->admitted_bytes = engine->committed();
  stats->admitted_bytes_peak = engine->committed_peak;
  stats->hbm_budget = engine->config.hbm_budget;
  stats->workspace_bytes = engine->held_total();
  stats->workspace_bytes_peak = engine->held_peak;
  stats->workspace_releases = engine->releases;
  stats->elapsed_seconds =
      engine->submitted ? std::max(0.0, engine->last_finish - engine->first_submit) : 0.0;
  return DICE_OK;
}

}  // extern "C"
