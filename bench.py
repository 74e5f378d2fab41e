"""Benchmark: 4K NV12 frames/s at distance 1.0 on one MI355X.

Workload: 3840x2160 seeded synthetic NV12 frames (cujpegxl_amd/synth.py, seeds
0-7) resident in HBM, encoded to `.jxl` files on the host by the plain HIP
VarDCT path (RD selection over DCT8/16/32, CfL, adaptive quantization,
32-cluster prefix coding).  One step is one synchronous cujpegxl_encoder_encode
call: NV12 in HBM -> file bytes in a caller-owned host buffer.  The timed steps
cycle through 24 device frames (the 8 seeds, 3 copies each: 299 MB, more than
the 256 MB last-level cache), as the reference's benchmark alternates inputs
(tools/benchmark/encode_bench.cc:512-518), back to back in native code (the
stage library's cjx_stage_encode_cycle, as the reference's C++ benchmark
drives the C API).  The same steps from a Python loop are reported as
`python_loop_fps`, and one resident frame encoded repeatedly as
`single_frame_fps`.  The timed buffer is page-locked (allocated through the
library), into which the assembly kernel writes the file over the host link;
the rate with an ordinary pageable buffer (the library's pinned staging copy
plus a Python bytes copy) is `pageable_output_fps`.

Prints one JSON line.  `roofline` is the dominant kernel's achieved
algorithmic HBM bandwidth (HIP events bound to the kernel on the encoder's
stream, over REGION_FRAMES separate frames), with the instruction-issue
floor from the committed SQ counter summary beside it; `cpu_baseline` is the
C oracle (a restatement of the reference path) encoding the same frames on
all host cores the process may use (up to 16), with its single-thread rate
beside it.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)

WIDTH, HEIGHT, DISTANCE, SEED = 3840, 2160, 1.0, 0
SEEDS = 8    # distinct seeded frames in the timed cycle
COPIES = 3   # device copies of each (8 x 3 x 12.4 MB = 299 MB)
HBM_PEAK_GBS = 8000.0
REGION_FRAMES = 48  # frames whose region kernel is timed by bound events
TRAFFIC_FILE = os.path.join(HERE, "profiles", "hbm_traffic_4k_d1.txt")
COUNTER_FILE = os.path.join(HERE, "profiles", "sq_counters_4k_d1.txt")
CUS, SCLK_GHZ = 256, 2.4  # MI355X compute units and peak engine clock
# Forms of the path timed beside the plain one (encoder mode bits).
FORMS = (("mfma16", 1), ("learned", 2), ("learned+mfma16", 3))
BATCH_FORMS = (("batch8", 0), ("batch8 learned+mfma16", 3))


def region_valu_instructions() -> int | None:
    """VALU instructions per region_kernel launch (SQ_INSTS_VALU summed over
    the device), from the committed SQ counter summary."""
    try:
        with open(COUNTER_FILE) as f:
            for line in f:
                if line.startswith("region_kernel valu_insts="):
                    return int(line.split("=", 1)[1])
    except OSError:
        pass
    return None


def region_traffic_bytes() -> int | None:
    """HBM bytes per region_kernel launch on this workload, from the committed
    FETCH_SIZE/WRITE_SIZE summary (counters scaled by a copy of known size)."""
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("region_kernel traffic_bytes="):
                    return int(line.split("=", 1)[1])
    except OSError:
        pass
    return None


QUALITY_FILE = os.path.join(HERE, "profiles", "quality_4k.json")


def quality_summary() -> dict | None:
    """Rate at equal quality of each form against the plain path, from the
    committed summary of tests/test_gpu_quality.py at 3840x2160 (decoded
    files scored with the Butteraugli-structured distance of
    tests/butteraugli.py, not libjxl's)."""
    try:
        with open(QUALITY_FILE) as f:
            q = json.load(f)
    except (OSError, ValueError):
        return None
    return {"source": "profiles/quality_4k.json (tests/test_gpu_quality.py, CJX_QUALITY_FULL=1)",
            "metric": q.get("metric"), "size": q.get("size"),
            "bpp_increase_at_plain_score": q.get("mean_bpp_increase"),
            "network_decided": [f.get("network_decided") for f in q.get("frames", [])],
            "region_ms_3840x2160": q.get("region_ms_4k")}


MIXED_SIZES = [(640, 360), (1280, 720), (1920, 1080), (2560, 1440), (3840, 2160)]


def mixed_batch(cj, synthetic_nv12, frames: int = 64, seed: int = 4) -> dict:
    """64 frames with sizes drawn uniformly from MIXED_SIZES and distances
    uniform in [0.8, 3.0] (seeded); one seeded synthetic frame per size,
    resident in HBM.  Sequential: one encoder per size, reconfigured for each
    frame's distance, encoded one after another into a page-locked buffer
    (rate and per-frame synchronous latency).  Batched: the same frames as
    one Batch (rate over replays of the whole batch)."""
    rng = np.random.default_rng(seed)
    picks = rng.integers(0, len(MIXED_SIZES), frames)
    dists = rng.uniform(0.8, 3.0, frames)
    encs, bufs = [], []
    for i, (w, h) in enumerate(MIXED_SIZES):
        luma, chroma = synthetic_nv12(w, h, 100 + i)
        bufs.append(cj.DeviceBuffer.from_array(np.concatenate([luma.ravel(), chroma.ravel()])))
        encs.append(cj.Encoder(w, h, 1.0))
    pinned = cj.PinnedBuffer(max(e.capacity for e in encs))

    def run(k: int) -> int:
        w, h = MIXED_SIZES[picks[k]]
        e, b = encs[picks[k]], bufs[picks[k]]
        e.configure(float(dists[k]))
        return e.encode_into(b.ptr, b.ptr + w * h, w, w, pinned.ptr, pinned.nbytes)

    for k in range(min(frames, 8)):  # warm every size
        run(k)
    cj.stages_lib().cjx_device_synchronize()
    lat, nbytes = [], 0
    t0 = time.perf_counter()
    for k in range(frames):
        t = time.perf_counter()
        nbytes += run(k)
        lat.append(time.perf_counter() - t)
    total = time.perf_counter() - t0
    pinned.close()
    for x in encs:
        x.close()
    lat_ms = np.array(lat) * 1e3
    pixels = sum(MIXED_SIZES[p][0] * MIXED_SIZES[p][1] for p in picks)
    sequential = {"fps": round(frames / total, 2), "megapixels_per_s": round(pixels / total / 1e6, 1),
                  "p50_ms": round(float(np.percentile(lat_ms, 50)), 4),
                  "p99_ms": round(float(np.percentile(lat_ms, 99)), 4), "bytes": nbytes,
                  "launches": "one synchronous encode per frame (14 launches each)"}
    # The same 64 frames as one batch: every stage one launch over all
    # frames, the 14 launches replayed from a captured hipGraph, launch
    # shapes chosen by the autotuner first.
    batch = cj.Batch([MIXED_SIZES[p] for p in picks], [float(d) for d in dists])
    batch.set_inputs([(bufs[p].ptr, bufs[p].ptr + MIXED_SIZES[p][0] * MIXED_SIZES[p][1], MIXED_SIZES[p][0],
                       MIXED_SIZES[p][0]) for p in picks])
    batch.run(2)
    tune = batch.tune(3)
    batch.run(2)
    reps = 20
    sec = batch.run(reps)
    # Per-batch latency: every frame of a batch completes with the batch.
    lat_b = np.array([batch.run(1) for _ in range(reps)]) * 1e3
    batch_bytes = sum(len(f) for f in batch.files())
    batch.close()
    for x in bufs:
        x.close()
    batched = {"fps": round(frames * reps / sec, 2),
               "megapixels_per_s": round(pixels * reps / sec / 1e6, 1),
               "ms_per_batch": round(1e3 * sec / reps, 4),
               "frame_latency_p50_ms": round(float(np.percentile(lat_b, 50)), 4),
               "frame_latency_p99_ms": round(float(np.percentile(lat_b, 99)), 4), "bytes": batch_bytes,
               "bytes_equal_sequential": batch_bytes == nbytes,
               "tuned": tune.strip().splitlines()[-1],
               "launches": "one launch per stage over the 64 frames (14 per batch), replayed from a hipGraph"}
    return {"frames": frames, "sizes": "640x360 .. 3840x2160 uniform, d ~ U[0.8, 3.0], seed 4",
            "sequential": sequential, "batched": batched,
            "batched_vs_sequential": round(batched["fps"] / sequential["fps"], 3)}


def host_info() -> dict:
    info = {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    info["model"] = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return info


def cpu_baseline(frames: list, budget_s: float = 10.0) -> dict:
    """The C oracle encoding the bench's 4K frames (cycling through them):
    single-threaded, then with its row/tile/region/block loops on all the
    cores this process may use (at most 16).  Each leg runs whole frames
    until `budget_s` has passed (at least one, at most 8)."""
    import oracle

    oracle.lib()

    def leg(threads: int) -> tuple[float, int, float]:
        oracle.set_threads(threads)
        n, t0 = 0, time.perf_counter()
        while True:
            luma, chroma = frames[n % len(frames)]
            oracle.encode_nv12(luma, chroma, WIDTH, HEIGHT, DISTANCE)
            n += 1
            el = time.perf_counter() - t0
            if el >= budget_s or n >= 8:
                break
        oracle.set_threads(1)
        return n / el, n, el

    cores = min(16, len(os.sched_getaffinity(0)))
    one = leg(1)
    many = leg(cores) if cores > 1 else one
    return {"value": round(many[0], 4), "unit": "frames/s", "cores": cores, "kind": "port",
            "sample": f"{many[1]} full 3840x2160 frame(s) of the bench's seeds, d=1.0, {cores} threads "
                      f"(OpenMP over rows, AQ tiles, selection regions and DCT blocks; entropy coding "
                      f"serial), {many[2]:.1f} s",
            "single_thread": {"value": round(one[0], 4), "cores": 1,
                              "sample": f"{one[1]} frame(s), {one[2]:.1f} s"},
            "host": host_info()}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-forms", action="store_true",
                    help="skip the matrix-core, learned and batched forms (plain path only)")
    ap.add_argument("--no-mixed-batch", action="store_true",
                    help="skip the mixed-resolution batch (4K frames only, e.g. under a profiler)")
    ap.add_argument("--copy-bytes", type=int, default=0,
                    help="only run device-to-device copies of this size (HBM counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the encoder runs on a single MI355X "
              "(sharding a frame's AC-group rows over several GPUs is not implemented)", file=sys.stderr)
        return 2

    import cujpegxl_amd as cj
    from cujpegxl_amd.synth import synthetic_nv12

    if args.copy_bytes:
        src, dst = cj.DeviceBuffer(args.copy_bytes), cj.DeviceBuffer(args.copy_bytes)
        for _ in range(max(1, args.steps)):
            cj.stages_lib().cjx_device_copy(dst.ptr, src.ptr, args.copy_bytes)
        cj.stages_lib().cjx_device_synchronize()
        print(json.dumps({"copy_bytes": args.copy_bytes, "copies": max(1, args.steps)}))
        return 0

    host_frames = [synthetic_nv12(WIDTH, HEIGHT, SEED + k) for k in range(SEEDS)]
    devs = []
    for _ in range(COPIES):
        for luma, chroma in host_frames:
            devs.append(cj.DeviceBuffer.from_array(np.concatenate([luma.ravel(), chroma.ravel()])))
    inputs = [(d.ptr, d.ptr + WIDTH * HEIGHT, WIDTH, WIDTH) for d in devs]
    enc = cj.Encoder(WIDTH, HEIGHT, DISTANCE, stages=True)
    lib = enc.lib

    pinned = cj.PinnedBuffer(enc.capacity)
    for k in range(max(args.warmup, len(inputs))):
        enc.encode_into(*inputs[k % len(inputs)], pinned.ptr, enc.capacity)
    # Timed steps: `steps` synchronous encodes back to back in native code
    # (the reference's benchmark drives the C API from a C++ loop), cycling
    # through the frames, no events.  The same steps from a Python loop (one
    # ctypes call per frame) are reported beside them.
    lib.cjx_device_synchronize()
    t0 = time.perf_counter()
    _, size = enc.encode_cycle(inputs, pinned.ptr, enc.capacity, args.steps)
    lib.cjx_device_synchronize()
    elapsed = time.perf_counter() - t0
    last = (args.steps - 1) % len(inputs)
    data = pinned.array[:size].tobytes()
    sizes = []
    for k in range(SEEDS):
        sizes.append(enc.encode_into(*inputs[k], pinned.ptr, enc.capacity))
    t1 = time.perf_counter()
    for k in range(args.steps):
        enc.encode_into(*inputs[k % len(inputs)], pinned.ptr, enc.capacity)
    lib.cjx_device_synchronize()
    python_loop_fps = args.steps / (time.perf_counter() - t1)
    t1 = time.perf_counter()
    enc.encode_repeat(*inputs[0], pinned.ptr, enc.capacity, args.steps)
    lib.cjx_device_synchronize()
    single_frame_fps = args.steps / (time.perf_counter() - t1)

    # The other forms of the path on the same frames and loop: FP16-operand
    # matrix-core DCTs, the learned transform selection, and both.
    forms = {}
    for name, mode in () if args.no_forms else FORMS:
        enc.set_mode(mode)
        for k in range(len(inputs)):
            enc.encode_into(*inputs[k], pinned.ptr, enc.capacity)
        lib.cjx_device_synchronize()
        t1 = time.perf_counter()
        enc.encode_cycle(inputs, pinned.ptr, enc.capacity, args.steps)
        lib.cjx_device_synchronize()
        f_fps = args.steps / (time.perf_counter() - t1)
        f_sizes = [enc.encode_into(*inputs[k], pinned.ptr, enc.capacity) for k in range(SEEDS)]
        enc.set_profiling(True, region_only=True)
        rsum = 0.0
        nreg = min(args.steps, REGION_FRAMES // 2)
        for k in range(nreg):
            enc.encode_into(*inputs[k % len(inputs)], pinned.ptr, enc.capacity)
            rsum += enc.region_ms()
        enc.set_profiling(False)
        forms[name] = {"fps": round(f_fps, 3), "bpp": round(8.0 * float(np.mean(f_sizes)) / (WIDTH * HEIGHT), 4),
                       "region_ms": round(rsum / max(nreg, 1), 4)}
    enc.set_mode(0)

    # Batched forms: the eight distinct frames as one batch (every stage one
    # launch over the eight frames, replayed from a captured hipGraph, launch
    # shapes from the autotuner), outputs into mapped page-locked memory.
    # One batch at a time, then two batches (over the frames' first and
    # second device copies) in flight on their own streams.
    for name, mode in () if args.no_forms else BATCH_FORMS:
        pair = [cj.Batch([(WIDTH, HEIGHT)] * SEEDS, [DISTANCE] * SEEDS, mode=mode) for _ in range(2)]
        tuned, report = [], []
        for j, batch in enumerate(pair):
            batch.set_inputs(inputs[j * SEEDS:(j + 1) * SEEDS])
            batch.run(2)
            lines = batch.tune(2).strip().splitlines()
            tuned.append(lines[-1])
            if j == 0:
                report = lines[:-1]
            batch.run(1)
        reps = max(1, args.steps // SEEDS)
        sec = pair[0].run(reps)
        b_sizes = [len(f) for f in pair[0].files()]
        reps2 = max(1, reps // 2)
        lib.cjx_device_synchronize()
        t1 = time.perf_counter()
        for _ in range(reps2):
            for batch in pair:
                batch.launch(1)
        for batch in pair:
            batch.wait()
        sec2 = time.perf_counter() - t1
        b2_sizes = [len(f) for f in pair[1].files()]
        for batch in pair:
            batch.close()
        forms[name] = {"fps": round(SEEDS * reps / sec, 3),
                       "bpp": round(8.0 * float(np.mean(b_sizes)) / (WIDTH * HEIGHT), 4),
                       "frames_per_batch": SEEDS, "batches_timed": reps, "tuned": tuned[0],
                       "tune_report": report,
                       "two_in_flight_fps": round(2 * SEEDS * reps2 / sec2, 3),
                       "two_in_flight_sizes_equal": b2_sizes == b_sizes}
        if mode == 0:
            forms[name]["sizes_equal_single_frame"] = b_sizes == sizes

    # The region kernel's duration (the roofline figure), from HIP events
    # bound to the kernel's own start and end on the encoder's stream, over
    # separate frames (a start event adds a few microseconds of idle device
    # time after the kernel).  The other stages are timed in separate
    # frames below.
    region_frames = min(args.steps, REGION_FRAMES)
    region_sum = 0.0
    enc.set_profiling(True, region_only=True)
    for k in range(region_frames):
        enc.encode_into(*inputs[k % len(inputs)], pinned.ptr, enc.capacity)
        region_sum += enc.region_ms()
    enc.set_profiling(False)

    enc.set_profiling(True)
    stage_frames = min(args.steps, 16)
    stage_sum = {k: 0.0 for k in cj.STAGE_NAMES}
    for k in range(stage_frames):
        enc.encode_into(*inputs[k % len(inputs)], pinned.ptr, enc.capacity)
        for name, v in enc.timings().items():
            stage_sum[name] += v
    enc.set_profiling(False)

    # The same steps into a pageable numpy buffer, returned as Python bytes.
    lib.cjx_device_synchronize()
    t1 = time.perf_counter()
    for k in range(args.steps):
        paged = enc.encode_device(*inputs[k % len(inputs)])
    lib.cjx_device_synchronize()
    pageable_fps = args.steps / (time.perf_counter() - t1)
    if enc.encode_device(*inputs[last]) != data:
        raise RuntimeError("pinned and pageable outputs differ")

    ms_per_step = 1e3 * elapsed / args.steps
    fps = args.steps / elapsed
    stage_ms = {k: v / stage_frames for k, v in stage_sum.items()}
    region_ms = region_sum / region_frames
    nv12_bytes = 1.5 * WIDTH * HEIGHT
    achieved = nv12_bytes / (region_ms * 1e-3) / 1e9
    mean_bytes = float(np.mean(sizes))
    valu = region_valu_instructions()
    limiter = None
    if valu is not None:
        floor_us = valu / (CUS * SCLK_GHZ * 1e3)  # one wave64 VALU instruction per CU per clock
        limiter = {"kind": "valu-issue", "valu_insts_per_launch": valu,
                   "issue_floor_us": round(floor_us, 1),
                   "frac": round(floor_us / (region_ms * 1e3), 4),
                   "source": "profiles/sq_counters_4k_d1.txt (SQ_INSTS_VALU over the device; "
                             "256 CUs x 2.4 GHz, one wave64 VALU instruction per CU per clock)"}
    out = {
        "metric": "4K NV12 frames/s at d=1.0 on 1 MI355X",
        "value": round(fps, 3),
        "unit": "frames/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "3840x2160 NV12 8-bit seeded synthetic frames (seeds 0-7, 3 device copies each, "
                        "cycled), distance 1.0, plain HIP VarDCT (DCT8/16/32 RD selection, CfL, adaptive "
                        "quant, 32-cluster prefix coding), NV12 resident in HBM, .jxl bytes to host",
            "path": "plain",
            "output": "caller-owned page-locked host buffer",
            "single_frame_fps": round(single_frame_fps, 3),
            "pageable_output_fps": round(pageable_fps, 3),
            "python_loop_fps": round(python_loop_fps, 3),
            "bytes_per_frame": round(mean_bytes, 1),
            "bpp": round(8.0 * mean_bytes / (WIDTH * HEIGHT), 4),
            "stage_ms": {k: round(v, 4) for k, v in stage_ms.items()},
            "stage_ms_frames": stage_frames,
            "frame_alg_bytes_frac": round((nv12_bytes + mean_bytes) * fps / (HBM_PEAK_GBS * 1e9), 6),
        },
        "roofline": {
            "bound": "hbm",
            "kernel": "region_kernel",
            "kernel_ms": round(region_ms, 4),
            "kernel_frames_timed": region_frames,
            "achieved": round(achieved, 2),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 6),
            "traffic": region_traffic_bytes(),
            "limiter": limiter,
        },
    }
    for name, f in forms.items():
        f["vs_plain_fps"] = round(f["fps"] / fps, 4)
        f["bpp_vs_plain"] = round(f["bpp"] / out["config"]["bpp"] - 1.0, 5)
    out["forms"] = forms
    quality = quality_summary()
    if quality is not None:
        out["quality"] = quality
    if "learned+mfma16" in forms and "batch8 learned+mfma16" in forms:
        # The comparison the metric names: the learned + matrix-core form
        # against the plain path, in speed and in rate at equal quality.
        lm, blm, bp = forms["learned+mfma16"], forms["batch8 learned+mfma16"], forms["batch8"]
        out["learned_mfma_vs_plain"] = {
            "single_frame_speedup": lm["vs_plain_fps"],
            "batched_speedup_vs_plain_single_frame": blm["vs_plain_fps"],
            "batched_speedup_vs_plain_batched": round(blm["fps"] / bp["fps"], 4),
            "bpp_vs_plain_same_distance": lm["bpp_vs_plain"],
            "bpp_increase_at_equal_score": (quality or {}).get("bpp_increase_at_plain_score", {}).get("learned+mfma16"),
            "target": ">= 1.3x the plain GPU form at <= 2 % bpp increase at equal distance-1.0 score"}
    if not args.no_mixed_batch:
        out["mixed_batch"] = mixed_batch(cj, synthetic_nv12)
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(host_frames)
    pinned.close()
    for d in devs:
        d.close()
    enc.close()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
