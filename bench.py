"""Benchmark: bit-exact Mpix/s of the Digital ICE processing core on MI355X.

Workload: synthetic 5782x3946 RGBI uint16 frames (400 dust disks, 40
scratches <= 400 px; a different seed per frame), resident in HBM before
the timed region.  One job covers:
* the producer schedule;
* planes / decision / features / candidates / combiner;
* the hidden startup and the serial conditional-dither writer chain;
* the output patch;
* the SHA-256 of the public RGB16 output, which needs the output bytes on
  the host.

Legs:
  plain    one job per call (dice_process_job), W warm-up + K timed jobs:
           p50/p99 latency and Mpix/s of the plain single-frame path;
  batched  B frames per dice_process_batch call: persistent launches over
           the batch, device phases replayed from captured hipGraphs, tile
           shape autotuned on the batch; W warm-up + K timed batches;
  pipelined  the headline `value`: K x B x rounds frames through the serving
           engine (batches of up to --pipeline-batch, `drivers` batches in
           flight, so the device phases of one batch overlap the host work
           of the others);
  pipelined_without_sha256  the same jobs with the output SHA-256 skipped
           (outputs stay in HBM): what the device path and the host writer
           walk sustain without the hashing; informational only;
  crops    16 crops of uneven size (sides 512-3000 px, aspect 0.5-2) as one
           batch (same machinery): Mpix/s and batch latency.
  routing  the hybrid's at-floor routing of frame 0's at-floor mask on the
           device (components, radii, dilation, partition), checked against
           the SciPy oracle.
  stream   frame 0 handed over as 723 blocks of 8 rows (the scanner's
           transport) through a frame stream: the device phases and the
           serial writer run as the blocks arrive; the latency from the
           last block to the result, against the plain whole-frame latency.
  device_chain  frame 0 with its serial writer chain on one device
           wavefront instead of a host core (the opt-in mode): ms per frame
           and ns per site beside the host walk, checked against the oracle.
  large    one 8818x13228 large-format frame (1500 disks up to r 30 px, 120
           scratches), one job per call: p50 latency and Mpix/s, checked
           against the oracle in the CPU leg.
  serving  Poisson arrivals (70 % 1024-2048 px crops, 30 % full frames) at
           80 % of the batched capacity into the serving engine (continuous
           batching, HBM-budget admission, eight drivers with batches of up
           to two): p50/p99 latency and
           Mpix/s, against the same jobs processed one at a time.

Every timed job's output SHA-256 and counters are compared with the compiled
CPU oracle (the C restatement of the reference CPU path) run on the same
frames.  That oracle run is also the CPU baseline, timed on this host's
cores.  The roofline is that of the kernel with the largest share of device
time, the reconstruction-site kernel k_sites: its algorithmic bytes (8 B
RGBI read per pixel of the frame, one 40 B site record written per attempted
site) over its HIP-event time on the library's stream, with its binary64
work (about 1750 operations per site: 13 nine-column averages, the record
divisions and the combiner) against the FP64 vector peak beside it.  The
streaming kernel k_scan (14 B per pixel: 8 B in, 6 B out) is reported too.
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import statistics
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "digital-fauxice_amd"))

METRIC = "bit-exact Mpix/s at 5782x3946 RGBI16 (batched vs plain); p50/p99 job latency"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
BYTES_PER_PIXEL = 14  # 8 B RGBI uint16 in + 6 B RGB uint16 out
SITE_READ_BYTES_PER_PIXEL = 8  # k_sites: the RGBI of every pixel of its tiles
SITE_RECORD_BYTES = 40  # k_sites: one Site record per attempted site
# binary64 operations per attempted site of the reference algorithm: 13
# unscaled 9x9 averages (9 column scratches of 12 operations + 17 to combine
# them), the RGB Q16 rows (13 each), 15 record divisions, 9 candidate
# scalings and the combiner (about 50).  k_sites shares feature records
# between neighbouring sites and executes about half of these; the rate is
# quoted on the algorithmic count.
SITE_FP64_OPS = 13 * (9 * 12 + 17) + 3 * 13 + 15 + 9 + 50
# HBM bytes one k_sites<16, 64, 512, 4> launch moves on this workload:
# rocprofv3 FETCH_SIZE and WRITE_SIZE (two --pmc runs over the plain leg, one
# frame per dispatch), each scaled by what it reads for the plain copy kernel
# k_copy over 1 GiB (x2.000 and x1.000)
SITES_TRAFFIC_BYTES = 364_240_000
TRAFFIC_SOURCE = ("profiles/kernel_stats_full_frame_5782x3946.md, k_sites<16, 64, 512, 4> row: "
                  "FETCH_SIZE x 2.000 + WRITE_SIZE x 1.000 per dispatch (not measured in this run)")


def parse_args() -> argparse.Namespace:
    parser = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    parser.add_argument("--gpus", type=int, default=1)
    parser.add_argument("--hw-queues", type=int, default=16,
                        help="GPU_MAX_HW_QUEUES for this process (0: keep the environment's)")
    parser.add_argument("--steps", type=int, default=6, help="timed batches (and plain jobs)")
    parser.add_argument("--warmup", type=int, default=2)
    parser.add_argument("--batch", type=int, default=8, help="full frames per batch")
    parser.add_argument("--device-chain", action="store_true",
                        help="batched legs run each job's writer chain on one device wavefront")
    parser.add_argument("--drivers", type=int, default=2,
                        help="concurrent batches in the pipelined leg")
    parser.add_argument("--pipeline-batch", type=int, default=16,
                        help="largest batch the pipelined leg's engine forms (0: --batch)")
    parser.add_argument("--pipeline-rounds", type=int, default=8,
                        help="the pipelined leg runs steps x batch x this many frames")
    parser.add_argument("--no-hashless", action="store_true",
                        help="skip the pipelined leg without the output SHA-256")
    parser.add_argument("--no-crops", action="store_true", help="skip the uneven-crops leg")
    parser.add_argument("--no-end-to-end", action="store_true",
                        help="skip the host-buffer (PCIe-inclusive) leg")
    parser.add_argument("--no-serving", action="store_true", help="skip the serving leg")
    parser.add_argument("--serving-jobs", type=int, default=200)
    parser.add_argument("--serving-drivers", type=int, default=8,
                        help="drivers of the serving leg's engine")
    parser.add_argument("--serving-batch", type=int, default=2,
                        help="largest batch of the serving leg's engine")
    parser.add_argument("--no-routing", action="store_true",
                        help="skip the at-floor routing leg")
    parser.add_argument("--no-stream", action="store_true",
                        help="skip the 8-row block streaming leg")
    parser.add_argument("--no-large", action="store_true",
                        help="skip the 8818x13228 large-format leg")
    parser.add_argument("--no-device-chain-leg", action="store_true",
                        help="skip the single-wavefront device writer chain leg")
    parser.add_argument("--no-cpu-baseline", action="store_true",
                        help="skip the CPU oracle leg (and its bit-exact check)")
    parser.add_argument("--cpu-seconds", type=float, default=20.0,
                        help="approximate CPU time budget of the baseline sample")
    parser.add_argument("--no-one-thread", action="store_true",
                        help="skip the one-thread oracle frame of the CPU leg")
    parser.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                        help="only run the plain copy kernel over BYTES (counter calibration)")
    return parser.parse_args()


def log(message: str) -> None:
    print(f"bench: {message}", file=sys.stderr, flush=True)


def cpu_model() -> str | None:
    """The host CPU's model name, from /proc/cpuinfo (None where it is not there)."""
    try:
        with open("/proc/cpuinfo", encoding="utf-8", errors="replace") as info:
            for line in info:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def copy_calibration(nbytes: int) -> int:
    """Plain device copy of a known size: the reference point for HBM counters."""

    from dice_amd.hip_backend import engine

    ms = C.c_double()
    engine.raise_for(engine.library().dice_t_copy(engine.context(), nbytes, 20, C.byref(ms)))
    gbs = 2 * nbytes / (ms.value / 1e3) / 1e9
    print(json.dumps({"copy_bytes": nbytes, "ms_per_copy": round(ms.value, 4),
                      "read_plus_write_GBs": round(gbs, 1)}), flush=True)
    return 0


class Frames:
    """Synthetic frames, their prepasses and resident device buffers."""

    def __init__(self, specs) -> None:
        from dice_amd import synth
        from dice_amd.hip_backend.device import DeviceArray

        self.specs = specs
        self.mains = [synth.frame(s) for s in specs]
        self.prepasses = [np.ascontiguousarray(synth.prepass_for(m)) for m in self.mains]
        self.d_main = [DeviceArray.from_host(m) for m in self.mains]
        self.d_out = [DeviceArray(s.height * s.width * 6) for s in specs]

    @property
    def mpix(self) -> list[float]:
        return [s.height * s.width / 1e6 for s in self.specs]


def fingerprint(res) -> tuple:
    return (res.output_sha256.decode(), res.attempted_pixels, res.written_pixels,
            res.public_rng_advances, res.final_rng_state, res.changed_pixels)


def main() -> int:
    args = parse_args()
    if args.hw_queues > 0:
        # HIP maps every stream of the process onto GPU_MAX_HW_QUEUES hardware
        # queues (4 by default).  Three runners drive nine streams; on four
        # queues one runner's kernels wait behind another's copies.  HIP reads
        # it when it starts, which is below.
        os.environ["GPU_MAX_HW_QUEUES"] = str(args.hw_queues)
    if args.gpus != 1:
        print(f"bench.py: this library runs on one GPU; --gpus {args.gpus} is not supported "
              "(jobs would shard across GPUs by job, one process per GPU, without "
              "collectives)", file=sys.stderr)
        return 2
    if args.copy_calibration:
        return copy_calibration(args.copy_calibration)

    from dice_amd import _abi, synth
    from dice_amd._abi import DiceBatchReport, DiceJob, DiceReplayResult, DiceStageTimings, ptr
    from dice_amd.hip_backend import engine
    from dice_amd.hip_backend.device import synchronize
    from dice_amd.hip_backend.process import stable_profile_params
    from dice_amd.profile import DEFAULT_PROFILE

    lib = engine.library()
    ctx = engine.context()
    params = stable_profile_params()
    prepass_params = DEFAULT_PROFILE.prepass_parameters()
    device_flags = _abi.DICE_MAIN_ON_DEVICE | _abi.DICE_OUT_ON_DEVICE

    t0 = time.perf_counter()
    full = Frames([synth.config2_full_frame(seed=2 + i) for i in range(args.batch)])
    log(f"{args.batch} full frames generated and resident in {time.perf_counter() - t0:.1f} s")
    frame_mpix = full.mpix[0]
    height, width = full.specs[0].height, full.specs[0].width

    # ---- plain leg: one job per call on frame 0
    def plain_step(timings: DiceStageTimings) -> DiceReplayResult:
        res = DiceReplayResult()
        code = lib.dice_process_job(ctx, C.c_void_p(full.d_main[0].address), height, width,
                                    ptr(full.prepasses[0], C.c_uint16),
                                    full.prepasses[0].shape[0], full.prepasses[0].shape[1],
                                    C.byref(params), C.byref(prepass_params), device_flags,
                                    C.c_void_p(full.d_out[0].address), C.byref(res), None,
                                    None, C.byref(timings))
        engine.raise_for(code)
        return res

    for _ in range(args.warmup):
        plain_step(DiceStageTimings())
    synchronize()
    plain_lat, stage_rows, plain_results = [], [], []
    start = time.perf_counter()
    for _ in range(args.steps):
        timings = DiceStageTimings()
        a = time.perf_counter()
        plain_results.append(fingerprint(plain_step(timings)))
        plain_lat.append((time.perf_counter() - a) * 1e3)
        stage_rows.append(timings.as_dict())
    synchronize()
    plain_elapsed = time.perf_counter() - start
    log(f"plain: {args.steps} jobs in {plain_elapsed:.3f} s")
    stage_ms = {k: statistics.mean(r[k] for r in stage_rows) * (1.0 if k.endswith("_ms") else 1e3)
                for k in stage_rows[0]}
    scan_s = stage_ms["scan_kernel_ms"] / 1e3
    achieved = BYTES_PER_PIXEL * height * width / scan_s / 1e9 if scan_s > 0 else 0.0
    sites_s = stage_ms["sites_kernel_ms"] / 1e3
    sites0 = int(plain_results[0][1])
    sites_bytes = SITE_READ_BYTES_PER_PIXEL * height * width + SITE_RECORD_BYTES * sites0
    sites_achieved = sites_bytes / sites_s / 1e9 if sites_s > 0 else 0.0
    sites_tflops = SITE_FP64_OPS * sites0 / sites_s / 1e12 if sites_s > 0 else 0.0

    # ---- batched leg: B full frames per call
    batch_timings: list[dict] = []

    def batch_jobs(frames: Frames):
        array = (DiceJob * len(frames.specs))()
        results = [DiceReplayResult() for _ in frames.specs]
        timings = [DiceStageTimings() for _ in frames.specs]
        batch_jobs.timings = timings
        for i, spec in enumerate(frames.specs):
            e = array[i]
            e.main_rgbi = frames.d_main[i].address
            e.height, e.width = spec.height, spec.width
            e.prepass_rgbi = frames.prepasses[i].ctypes.data
            e.prepass_height, e.prepass_width = frames.prepasses[i].shape[:2]
            e.params = C.pointer(params)
            e.prepass_params = C.pointer(prepass_params)
            e.flags = device_flags
            e.out_rgb16 = frames.d_out[i].address
            e.result = C.pointer(results[i])
            e.timings = C.pointer(timings[i])
        return array, results

    def run_batch(frames: Frames, flags: int) -> tuple[list, DiceBatchReport]:
        array, results = batch_jobs(frames)
        report = DiceBatchReport()
        code = lib.dice_process_batch(ctx, array, len(frames.specs), flags, C.byref(report))
        engine.raise_for(code)
        return [fingerprint(r) for r in results], report

    batch_flags = _abi.DICE_BATCH_GRAPH | _abi.DICE_BATCH_AUTOTUNE
    if args.device_chain:
        batch_flags |= _abi.DICE_BATCH_DEVICE_CHAIN
    for _ in range(max(args.warmup, 2)):  # the second run captures the graphs
        run_batch(full, batch_flags)
    synchronize()
    batch_lat, batch_results = [], []
    start = time.perf_counter()
    for _ in range(args.steps):
        a = time.perf_counter()
        prints, batch_report = run_batch(full, batch_flags)
        batch_lat.append((time.perf_counter() - a) * 1e3)
        batch_results.append(prints)
        batch_timings.extend(t.as_dict() for t in batch_jobs.timings)
    synchronize()
    batch_elapsed = time.perf_counter() - start
    log(f"batched (one call per batch): {args.steps} batches of {args.batch} in "
        f"{batch_elapsed:.3f} s (tile {batch_report.tile_rows}x{batch_report.tile_cols}, "
        f"graph {batch_report.graph_replayed})")
    single_call_value = args.steps * sum(full.mpix) / batch_elapsed

    # ---- end to end: host numpy frames in, host numpy outputs back (PCIe
    # and the pageable-memory staging included), as process_hip callers see
    end_to_end = None
    if not args.no_end_to_end:
        host_outs = [np.empty((height, width, 3), np.uint16) for _ in full.specs]

        def host_job(i: int) -> tuple:
            res = DiceReplayResult()
            code = lib.dice_process_job(ctx, C.c_void_p(full.mains[i].ctypes.data), height,
                                        width, ptr(full.prepasses[i], C.c_uint16),
                                        full.prepasses[i].shape[0], full.prepasses[i].shape[1],
                                        C.byref(params), C.byref(prepass_params), 0,
                                        C.c_void_p(host_outs[i].ctypes.data), C.byref(res),
                                        None, None, None)
            engine.raise_for(code)
            return fingerprint(res)

        def host_batch() -> list:
            array, results = batch_jobs(full)
            for i in range(len(full.specs)):
                array[i].main_rgbi = full.mains[i].ctypes.data
                array[i].out_rgb16 = host_outs[i].ctypes.data
                array[i].flags = 0
            code = lib.dice_process_batch(ctx, array, len(full.specs), batch_flags, None)
            engine.raise_for(code)
            return [fingerprint(r) for r in results]

        host_job(0)
        host_batch()
        e2e_plain, e2e_prints = [], []
        for _ in range(args.steps):
            a = time.perf_counter()
            e2e_prints.append(host_job(0))
            e2e_plain.append((time.perf_counter() - a) * 1e3)
        e2e_batch = []
        for _ in range(max(2, args.steps // 2)):
            a = time.perf_counter()
            batch_prints = host_batch()
            e2e_batch.append((time.perf_counter() - a) * 1e3)
        plain_p50 = float(np.percentile(e2e_plain, 50))
        batch_p50 = float(np.percentile(e2e_batch, 50))
        end_to_end = {
            "note": "host numpy RGBI in and RGB16 out per call (PCIe and pageable staging "
                    "included); never the headline value",
            "plain": {"value": round(frame_mpix / plain_p50 * 1e3, 2), "unit": "Mpix/s",
                      "p50_ms": round(plain_p50, 3),
                      "p99_ms": round(float(np.percentile(e2e_plain, 99)), 3)},
            "batch": {"jobs": len(full.specs),
                      "value": round(sum(full.mpix) / batch_p50 * 1e3, 2), "unit": "Mpix/s",
                      "p50_ms": round(batch_p50, 3)},
            "identical_to_resident": (all(p == plain_results[0] for p in e2e_prints)
                                      and batch_prints == batch_results[-1]),
        }
        log(f"end to end: {end_to_end}")
        del host_outs

    # ---- pipelined batches (the headline): the same frames through the
    # serving engine, batches of B formed by `drivers` concurrent drivers, so
    # the device phases of one batch overlap the host chain and hash of others
    from dice_amd.hip_backend.device import DeviceArray
    from dice_amd.hip_backend.serving import HipServingEngine

    jobs_total = args.steps * args.batch * max(1, args.pipeline_rounds)
    pipe_batch = args.pipeline_batch or args.batch
    # output buffers are recycled: a batch takes jobs in submission order and at
    # most `drivers` batches run at once, so a buffer is free again long before
    # its next job starts
    pipe_outs = [DeviceArray(height * width * 6)
                 for _ in range(min(jobs_total, max(64, 4 * args.drivers * pipe_batch)))]
    with HipServingEngine(drivers=args.drivers, max_batch=pipe_batch, graph=True,
                          device_chain=args.device_chain,
                          autotune=True) as served:
        warm = [served.submit_device(full.d_main[k % args.batch].address, height, width,
                                     full.prepasses[k % args.batch],
                                     pipe_outs[k % len(pipe_outs)].address)
                for k in range(min(jobs_total, 2 * args.drivers * pipe_batch))]
        for t in warm:
            served.wait(t)
        synchronize()
        served.reset_stats()
        cpu_start = os.times()
        start = time.perf_counter()
        tickets = [served.submit_device(full.d_main[k % args.batch].address, height, width,
                                        full.prepasses[k % args.batch],
                                        pipe_outs[k % len(pipe_outs)].address)
                   for k in range(jobs_total)]
        pipe_prints = []
        for t in tickets:
            r = served.wait(t).replay
            pipe_prints.append((r.output_sha256, r.attempted_pixels, r.written_pixels,
                                r.public_rng_advances, r.final_rng_state, r.changed_pixels))
        pipe_elapsed = time.perf_counter() - start
        cpu_end = os.times()
        pipe_stats = served.stats()
    # host cores the pipelined leg kept busy (user + system time of every
    # thread of this process over the leg's wall time)
    pipe_cores = ((cpu_end.user - cpu_start.user) + (cpu_end.system - cpu_start.system)) / pipe_elapsed
    value = jobs_total * frame_mpix / pipe_elapsed
    log(f"pipelined batches: {jobs_total} frames in {pipe_elapsed:.3f} s "
        f"({value:.1f} Mpix/s, mean batch {pipe_stats['mean_batch']:.2f})")

    # ---- the same pipeline without the output SHA-256 (the outputs stay in
    # HBM): what the device path and the host writer walk sustain without
    # the host hashing 137 MB per frame.  Informational, never the headline;
    # every job's counters must equal the hashed leg's.
    hashless = None
    if not args.no_hashless:
        with HipServingEngine(drivers=args.drivers, max_batch=pipe_batch, graph=True,
                              autotune=True) as served:
            for t in [served.submit_device(full.d_main[k % args.batch].address, height, width,
                                           full.prepasses[k % args.batch],
                                           pipe_outs[k % len(pipe_outs)].address, sha256=False)
                      for k in range(min(jobs_total, 2 * args.drivers * pipe_batch))]:
                served.wait(t)
            synchronize()
            served.reset_stats()
            cpu_start = os.times()
            start = time.perf_counter()
            tickets = [served.submit_device(full.d_main[k % args.batch].address, height, width,
                                            full.prepasses[k % args.batch],
                                            pipe_outs[k % len(pipe_outs)].address, sha256=False)
                       for k in range(jobs_total)]
            counters = []
            for t in tickets:
                r = served.wait(t).replay
                counters.append((r.attempted_pixels, r.written_pixels, r.public_rng_advances,
                                 r.final_rng_state, r.changed_pixels))
            elapsed = time.perf_counter() - start
            cpu_end = os.times()
        hashless = {
            "note": "the pipelined leg's jobs without the output SHA-256 (outputs stay in HBM); "
                    "never the headline value",
            "jobs": jobs_total,
            "value": round(jobs_total * frame_mpix / elapsed, 2), "unit": "Mpix/s",
            "host_cores_busy": round(((cpu_end.user - cpu_start.user)
                                      + (cpu_end.system - cpu_start.system)) / elapsed, 2),
            "counters_equal_hashed": counters == [p[1:] for p in pipe_prints],
        }
        log(f"pipelined without the SHA-256: {hashless}")
    del pipe_outs

    # ---- crops leg: 16 uneven crops as one batch
    crops = None
    if not args.no_crops:
        crop_frames = Frames(synth.config3_crops())
        for _ in range(2):
            run_batch(crop_frames, batch_flags)
        synchronize()
        lat, crop_runs = [], []
        start = time.perf_counter()
        for _ in range(args.steps):
            a = time.perf_counter()
            crop_prints, crop_report = run_batch(crop_frames, batch_flags)
            lat.append((time.perf_counter() - a) * 1e3)
            crop_runs.append(crop_prints)
        crop_elapsed = time.perf_counter() - start
        crops = {
            "jobs": len(crop_frames.specs),
            "mpix_per_batch": round(sum(crop_frames.mpix), 3),
            "value": round(args.steps * sum(crop_frames.mpix) / crop_elapsed, 2),
            "unit": "Mpix/s",
            "batch_p50_ms": round(float(np.percentile(lat, 50)), 3),
            "tile": [crop_report.tile_rows, crop_report.tile_cols],
            "sites_variant": crop_report.sites_variant,
            "tune_source": crop_report.tune_source,
            "graph_replayed": bool(crop_report.graph_replayed),
        }
        log(f"crops: {crops}")

    # ---- serving leg: Poisson arrivals into the continuous-batching engine
    serving = None
    if not args.no_serving:
        from dice_amd.hip_backend.device import DeviceArray
        from dice_amd.hip_backend.serving import HipServingEngine

        rng = np.random.default_rng(5)
        crop_pool = Frames([synth.scaled_spec(int(rng.integers(1024, 2049)),
                                              int(rng.integers(1024, 2049)), seed=5000 + i)
                            for i in range(8)])
        pool = [(full, i) for i in range(min(4, args.batch))] + [(crop_pool, i) for i in range(8)]
        picks = []
        for _ in range(args.serving_jobs):
            if rng.random() < 0.3:
                picks.append(pool[int(rng.integers(0, min(4, args.batch)))])
            else:
                picks.append(pool[min(4, args.batch) + int(rng.integers(0, 8))])
        job_mpix = [frames.mpix[i] for frames, i in picks]
        outs = [DeviceArray(frames.specs[i].height * frames.specs[i].width * 6)
                for frames, i in picks]
        gaps = rng.exponential(1.0, size=len(picks))
        # the same jobs one at a time (plain path)
        plain_prints, plain_job_lat = [], []
        a0 = time.perf_counter()
        for (frames, i), out in zip(picks, outs):
            spec = frames.specs[i]
            res = DiceReplayResult()
            a = time.perf_counter()
            engine.raise_for(lib.dice_process_job(
                ctx, C.c_void_p(frames.d_main[i].address), spec.height, spec.width,
                ptr(frames.prepasses[i], C.c_uint16), frames.prepasses[i].shape[0],
                frames.prepasses[i].shape[1], C.byref(params), C.byref(prepass_params),
                device_flags, C.c_void_p(out.address), C.byref(res), None, None, None))
            plain_job_lat.append((time.perf_counter() - a) * 1e3)
            plain_prints.append(fingerprint(res))
        plain_mix_elapsed = time.perf_counter() - a0
        with HipServingEngine(drivers=args.serving_drivers, max_batch=args.serving_batch,
                              graph=False, autotune=True,
                              device_chain=args.device_chain) as served:
            def closed_loop():
                a = time.perf_counter()
                ts = []
                for (frames, i), out in zip(picks, outs):
                    spec = frames.specs[i]
                    ts.append(served.submit_device(frames.d_main[i].address, spec.height,
                                                   spec.width, frames.prepasses[i],
                                                   out.address))
                for t in ts:
                    served.wait(t)
                return time.perf_counter() - a

            closed_loop()  # warm every driver's workspaces on the mix
            capacity = sum(job_mpix) / closed_loop()
            rate = 0.8 * capacity / statistics.mean(job_mpix)  # jobs per second
            arrivals = np.cumsum(gaps / rate)
            served.reset_stats()
            tickets = []
            t_start = time.perf_counter()
            for (frames, i), out, at in zip(picks, outs, arrivals):
                delay = at - (time.perf_counter() - t_start)
                if delay > 0:
                    time.sleep(delay)
                spec = frames.specs[i]
                tickets.append(served.submit_device(frames.d_main[i].address, spec.height,
                                                    spec.width, frames.prepasses[i],
                                                    out.address))
            served_prints = []
            for t in tickets:
                r = served.wait(t).replay
                served_prints.append((r.output_sha256, r.attempted_pixels, r.written_pixels,
                                      r.public_rng_advances, r.final_rng_state,
                                      r.changed_pixels))
            serve_elapsed = time.perf_counter() - t_start
            stats = served.stats()
        lat = []  # the timed jobs only (the warm-up jobs are the first latencies)
        serving = {
            "jobs": len(picks),
            "arrival_rate_jobs_s": round(rate, 2),
            "value": round(sum(job_mpix) / serve_elapsed, 2),
            "unit": "Mpix/s",
            "latency_p50_ms": round(stats["latency_p50_ms"], 3),
            "latency_p99_ms": round(stats["latency_p99_ms"], 3),
            "queue_p50_ms": round(stats["queue_p50_ms"], 3),
            "mean_batch": round(stats["mean_batch"], 2),
            "drivers": args.serving_drivers,
            "max_batch": args.serving_batch,
            "hbm_budget": int(stats["hbm_budget"]),
            "admitted_bytes_peak": int(stats["admitted_bytes_peak"]),
            "workspace_bytes": int(stats["workspace_bytes"]),
            "workspace_bytes_peak": int(stats["workspace_bytes_peak"]),
            "capacity_vs_plain": round(capacity / (sum(job_mpix) / plain_mix_elapsed), 3),
            "plain_one_at_a_time": {
                "value": round(sum(job_mpix) / plain_mix_elapsed, 2),
                "unit": "Mpix/s",
                "p50_ms": round(float(np.percentile(plain_job_lat, 50)), 3),
                "p99_ms": round(float(np.percentile(plain_job_lat, 99)), 3),
            },
            "identical_to_plain": served_prints == [tuple(p) for p in plain_prints],
            "closed_loop_capacity": round(capacity, 2),
        }
        del lat
        log(f"serving: {serving}")

    # ---- at-floor routing of frame 0's at-floor mask (the hybrid's next
    # step), on the device; the oracle (SciPy) checks the census
    routing = None
    if not args.no_routing:
        from dice_amd.hip_backend import routing as hip_routing
        from dice_amd.hip_backend.process import process_hip
        from dice_amd.profile import selector8_job

        at_floor = process_hip(selector8_job(full.mains[0], full.prepasses[0]),
                               export_diagnostics=True).diagnostics.at_floor_mask
        for _ in range(2):
            hip_routing.route_at_floor_mask(at_floor)
        route_lat = []
        for _ in range(max(3, args.steps)):
            a = time.perf_counter()
            routed = hip_routing.route_at_floor_mask(at_floor)
            route_lat.append((time.perf_counter() - a) * 1e3)
        route_ms = statistics.median(route_lat)
        routing = {
            "at_floor_pixels": routed.at_floor_pixel_count,
            "provisional_regions": len(routed.provisional_regions),
            "final_regions": len(routed.final_regions),
            "ms": round(route_ms, 3),
            "mpix_per_s": round(frame_mpix / route_ms * 1e3, 2),
        }
        log(f"routing: {routing}")

    # ---- frame 0 arriving as 8-row blocks through a frame stream
    streamed = None
    if not args.no_stream:
        handle = C.c_void_p()
        engine.raise_for(lib.dice_stream_create(ctx, C.byref(handle)))
        main0 = np.ascontiguousarray(full.mains[0])
        out0 = np.empty((height, width, 3), np.uint16)
        prepass0 = np.ascontiguousarray(full.prepasses[0])
        blocks = [(row, np.ascontiguousarray(main0[row:row + 8])) for row in range(0, height, 8)]

        def stream_frame():
            res = DiceReplayResult()
            job = DiceJob()
            job.main_rgbi = None
            job.height, job.width = height, width
            job.prepass_rgbi = prepass0.ctypes.data
            job.prepass_height, job.prepass_width = prepass0.shape[:2]
            job.params = C.pointer(params)
            job.prepass_params = C.pointer(prepass_params)
            job.flags = 0
            job.out_rgb16 = out0.ctypes.data
            job.result = C.pointer(res)
            engine.raise_for(lib.dice_stream_begin(handle, C.byref(job), None))
            pushes = []
            frame_start = time.perf_counter()
            for row, block in blocks:
                a = time.perf_counter()
                engine.raise_for(lib.dice_stream_push(handle, ptr(block, C.c_uint16), row,
                                                      block.shape[0]))
                pushes.append(time.perf_counter() - a)
            a = time.perf_counter()
            engine.raise_for(lib.dice_stream_finish(handle))
            done = time.perf_counter()
            return (fingerprint(res), pushes, (done - a + pushes[-1]) * 1e3,
                    (done - frame_start) * 1e3)

        stream_frame()
        runs = [stream_frame() for _ in range(max(3, min(args.steps, 5)))]
        progress = _abi.DiceStreamProgress()
        lib.dice_stream_progress_get(handle, C.byref(progress))
        lib.dice_stream_destroy(handle)
        stream_print = runs[0][0]
        tails = [r[2] for r in runs]
        frames_ms = [r[3] for r in runs]
        pushes = [p * 1e3 for r in runs for p in r[1]]
        streamed = {
            "blocks": len(blocks),
            "rows_per_block": 8,
            "windows": int(progress.windows),
            "last_block_to_result_ms_p50": round(statistics.median(tails), 3),
            "last_block_to_result_ms_max": round(max(tails), 3),
            "push_ms_mean": round(statistics.mean(pushes), 4),
            "push_ms_p99": round(float(np.percentile(pushes, 99)), 4),
            "frame_ms_back_to_back": round(statistics.median(frames_ms), 2),
            "ingest_mpix_per_s": round(frame_mpix / statistics.median(frames_ms) * 1e3, 2),
            "plain_p50_ms": round(float(np.percentile(plain_lat, 50)), 3),
            "identical_to_plain": all(r[0] == plain_results[0] for r in runs),
        }
        log(f"stream: {streamed}")

    # ---- frame 0 with the writer chain on one device wavefront
    device_chain = None
    if not args.no_device_chain_leg:
        chain_prints, chain_ms = [], []
        for _ in range(3):
            array, results = batch_jobs(full)  # the first job of the array: frame 0
            chain_report = DiceBatchReport()
            a = time.perf_counter()
            code = lib.dice_process_batch(ctx, array, 1, _abi.DICE_BATCH_DEVICE_CHAIN,
                                          C.byref(chain_report))
            chain_ms.append((time.perf_counter() - a) * 1e3)
            engine.raise_for(code)
            chain_prints.append(fingerprint(results[0]))
        sites0 = int(plain_results[0][1])
        device_chain = {
            "ms_per_frame": round(min(chain_ms), 2),
            "ns_per_site": round(min(chain_ms) * 1e6 / sites0, 1),
            "host_walk_ms_plain": round(stage_ms["writer_host"], 2),
            "identical_to_plain": all(p == plain_results[0] for p in chain_prints),
        }
        log(f"device chain: {device_chain}")

    # ---- one large-format frame per call (device-resident, like plain)
    large = None
    if not args.no_large:
        big = Frames([synth.config4_large()])
        bh, bw = big.specs[0].height, big.specs[0].width

        def large_step() -> DiceReplayResult:
            res = DiceReplayResult()
            code = lib.dice_process_job(ctx, C.c_void_p(big.d_main[0].address), bh, bw,
                                        ptr(big.prepasses[0], C.c_uint16),
                                        big.prepasses[0].shape[0], big.prepasses[0].shape[1],
                                        C.byref(params), C.byref(prepass_params), device_flags,
                                        C.c_void_p(big.d_out[0].address), C.byref(res), None,
                                        None, None)
            engine.raise_for(code)
            return res

        large_step()
        large_lat, large_prints = [], []
        for _ in range(max(3, min(args.steps, 5))):
            a = time.perf_counter()
            large_prints.append(fingerprint(large_step()))
            large_lat.append((time.perf_counter() - a) * 1e3)
        large = {
            "height": bh,
            "width": bw,
            "attempted_sites": int(large_prints[0][1]),
            "p50_ms": round(float(np.percentile(large_lat, 50)), 3),
            "mpix_per_s": round(big.mpix[0] / float(np.percentile(large_lat, 50)) * 1e3, 2),
            "repeatable": all(p == large_prints[0] for p in large_prints),
        }
        log(f"large: {large}")

    # ---- CPU baseline (the oracle), which also checks every output
    cpu_baseline = None
    bit_exact = None
    if not args.no_cpu_baseline:
        sys.path.insert(0, str(ROOT / "tests"))
        import oracle_lib

        threads = min(os.cpu_count() or 1, 16)
        expected, runs = [], []
        budget_start = time.perf_counter()
        for i in range(args.batch):
            a = time.perf_counter()
            cpu = oracle_lib.process_job(full.mains[i], full.prepasses[i], params, prepass_params,
                                         threads=threads)
            runs.append(time.perf_counter() - a)
            expected.append((cpu.output_sha256, cpu.attempted_pixels, cpu.written_pixels,
                             cpu.public_rng_advances, cpu.final_rng_state, cpu.changed_pixels))
            log(f"cpu oracle frame {i}: {runs[-1]:.2f} s")
            if i >= 1 and time.perf_counter() - budget_start > args.cpu_seconds:
                break
        checked = len(expected)
        bit_exact = (all(p == expected[0] for p in plain_results)
                     and all(prints[:checked] == expected for prints in batch_results)
                     and all(p == expected[k % args.batch] for k, p in enumerate(pipe_prints)
                             if k % args.batch < checked)
                     and (hashless is None or hashless["counters_equal_hashed"]))
        if serving is not None:
            # every served job against the oracle on its pool frame
            oracle_prints = {}
            for frames, i in set(picks):
                if frames is full and i < checked:
                    oracle_prints[(id(frames), i)] = expected[i]
                else:
                    cpu = oracle_lib.process_job(frames.mains[i], frames.prepasses[i], params,
                                                 prepass_params, threads=threads)
                    oracle_prints[(id(frames), i)] = (
                        cpu.output_sha256, cpu.attempted_pixels, cpu.written_pixels,
                        cpu.public_rng_advances, cpu.final_rng_state, cpu.changed_pixels)
            serving_exact = all(p == oracle_prints[(id(f), i)]
                                for p, (f, i) in zip(served_prints, picks))
            serving["bit_exact"] = serving_exact
            bit_exact = bit_exact and serving_exact
        if streamed is not None:
            streamed["bit_exact"] = stream_print == expected[0]
            bit_exact = bit_exact and streamed["bit_exact"]
        if crops is not None:
            # every crop of every timed batch against the oracle on that crop
            crop_expected = []
            for main, prepass in zip(crop_frames.mains, crop_frames.prepasses):
                cpu = oracle_lib.process_job(main, prepass, params, prepass_params,
                                             threads=threads)
                crop_expected.append((cpu.output_sha256, cpu.attempted_pixels,
                                      cpu.written_pixels, cpu.public_rng_advances,
                                      cpu.final_rng_state, cpu.changed_pixels))
            crops["bit_exact"] = all(run == crop_expected for run in crop_runs)
            bit_exact = bit_exact and crops["bit_exact"]
        if device_chain is not None:
            device_chain["bit_exact"] = chain_prints[0] == expected[0]
            bit_exact = bit_exact and device_chain["bit_exact"]
        if large is not None:
            a = time.perf_counter()
            cpu = oracle_lib.process_job(big.mains[0], big.prepasses[0], params, prepass_params,
                                         threads=threads)
            large["cpu_s"] = round(time.perf_counter() - a, 3)
            large["bit_exact"] = large_prints[0] == (
                cpu.output_sha256, cpu.attempted_pixels, cpu.written_pixels,
                cpu.public_rng_advances, cpu.final_rng_state, cpu.changed_pixels)
            bit_exact = bit_exact and large["bit_exact"]
        if routing is not None:
            # the routing census against the SciPy oracle of the same mask
            sys.path.insert(0, str(ROOT / "oracle"))
            import routing_oracle

            a = time.perf_counter()
            want = routing_oracle.route(at_floor)[0]
            routing["cpu_ms"] = round((time.perf_counter() - a) * 1e3, 1)
            routing["bit_exact"] = want == hip_routing.routing_json_document(routed)
            bit_exact = bit_exact and routing["bit_exact"]
        timed = runs[1:] if len(runs) > 1 else runs
        median = statistics.median(timed)
        one_thread_s = None
        if not args.no_one_thread:
            a = time.perf_counter()
            cpu = oracle_lib.process_job(full.mains[0], full.prepasses[0], params,
                                         prepass_params, threads=1)
            one_thread_s = time.perf_counter() - a
            bit_exact = bit_exact and cpu.output_sha256 == expected[0][0]
            log(f"cpu oracle frame 0 on one thread: {one_thread_s:.2f} s")
        cpu_baseline = {
            "value": round(frame_mpix / median, 4),
            "unit": "Mpix/s",
            "cores": threads,
            "kind": "port",
            "sample": (f"{len(timed)} full 5782x3946 frames (after 1 warm-up) of the compiled C "
                       f"oracle, OpenMP row-parallel analysis + serial writer; median "
                       f"{median:.3f} s; it also checked {checked} frames of every timed batch "
                       f"and every plain job bit for bit"),
            "cpu_model": cpu_model(),
            "nproc": os.cpu_count(),
            "affinity": len(os.sched_getaffinity(0)),
            "one_thread_s": None if one_thread_s is None else round(one_thread_s, 3),
            "one_thread_mpix_per_s": (None if one_thread_s is None
                                      else round(frame_mpix / one_thread_s, 4)),
        }

    plain_value = args.steps * frame_mpix / plain_elapsed
    batched_single_call = {
        "value": round(single_call_value, 2),
        "job_stage_ms": {k: round(statistics.mean(t[k] for t in batch_timings) * 1e3, 3)
                         for k in ("prepass", "producer_device", "analysis_device",
                                   "writer_host", "download_sites", "upload",
                                   "download_output", "sha256_host", "finish_device",
                                   "total")},
        "unit": "Mpix/s",
        "batch_p50_ms": round(float(np.percentile(batch_lat, 50)), 3),
        "batch_p99_ms": round(float(np.percentile(batch_lat, 99)), 3),
        "tile": [batch_report.tile_rows, batch_report.tile_cols],
        "sites_variant": batch_report.sites_variant,
        "tune_source": batch_report.tune_source,
        "graph_replayed": bool(batch_report.graph_replayed),
    }
    line = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "Mpix/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(pipe_elapsed / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"{jobs_total} full 5782x3946 RGBI16 frames (steps {args.steps} x batch "
                         f"{args.batch} x {max(1, args.pipeline_rounds)} rounds; {args.batch} "
                         "distinct frames with 400 disks + 40 scratches each), resident in HBM, "
                         f"through the batching engine: batches of up to {pipe_batch}, {args.drivers} "
                         "concurrent drivers, persistent launches over each batch, graph replay, "
                         "autotuned tiles; SHA-256 of every output"),
            "frames_timed": jobs_total,
            "drivers": args.drivers,
            "pipeline_batch": pipe_batch,
            "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES"),
            "writer_chain": "device wavefront" if args.device_chain else "host core",
            "global_batch": args.batch,
            "height": height,
            "width": width,
            "attempted_sites_frame0": int(plain_results[0][1]),
            "parallelism": "single GPU",
        },
        "pipelined": {
            "jobs": jobs_total,
            "mean_batch": round(pipe_stats["mean_batch"], 2),
            "note": ("throughput leg: every job is submitted at once, so per-job times "
                     "are queue positions, not latencies (see plain and serving)"),
            "host_cores_busy": round(pipe_cores, 2),
            "host_cores": len(os.sched_getaffinity(0)),
        },
        "pipelined_without_sha256": hashless,
        "batched_single_call": batched_single_call,
        "plain": {
            "value": round(plain_value, 2),
            "unit": "Mpix/s",
            "p50_ms": round(float(np.percentile(plain_lat, 50)), 3),
            "p99_ms": round(float(np.percentile(plain_lat, 99)), 3),
        },
        "batched_vs_plain": round(value / plain_value, 3),
        "end_to_end": end_to_end,
        "crops": crops,
        "serving": serving,
        "routing": routing,
        "stream": streamed,
        "large": large,
        "device_chain": device_chain,
        "bit_exact": bit_exact,
        "stage_ms_plain": {k: round(v, 3) for k, v in stage_ms.items()},
        "writer_share_plain": round(stage_ms["writer_host"] / stage_ms["total"], 4)
        if stage_ms["total"] else None,
        "roofline": {
            "bound": "hbm",
            "kernel": "k_sites",
            "achieved": round(sites_achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(sites_achieved / HBM_PEAK_GBS, 4),
            "traffic": SITES_TRAFFIC_BYTES,
            "traffic_source": TRAFFIC_SOURCE,
            "kernel_ms": round(stage_ms["sites_kernel_ms"], 4),
            "algorithmic_bytes": sites_bytes,
            "fp64_tflops": round(sites_tflops, 3),
            "fp64_peak_tflops": FP64_PEAK_TFLOPS,
            "fp64_frac": round(sites_tflops / FP64_PEAK_TFLOPS, 4),
        },
        "roofline_scan": {
            "bound": "hbm",
            "kernel": "k_scan",
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "kernel_ms": round(stage_ms["scan_kernel_ms"], 4),
        },
        "cpu_baseline": cpu_baseline,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
