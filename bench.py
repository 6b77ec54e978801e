#!/usr/bin/env python3
"""Benchmark of the raster-scan generalised geodesic distance transform on one
MI355X.

Workload (the reference's headline 3D size): ``generalised_geodesic3d`` on a
1x1x512x512x512 float32 blob field (12 Gaussian blobs, sigma 8-40 voxels,
+U[0,0.05] noise, default_rng(512)), softmask of ones with 16 seeded zero
voxels, spacing [1,1,1], v=1e10, lamb=1.0, iter=4.  Inputs are resident in
HBM before the timed region.  One step = one full call (6 passes x 4
iterations = 3.22e9 voxel updates).

Prints ONE JSON line on stdout; progress goes to stderr.
  value      voxel-updates/s of the whole call (N * 6 * iter / wall time)
  roofline   the dominant kernel (persistent tiled sweep): algorithmic bytes
             (12 B per voxel update at C=1) per launch / its mean duration,
             measured with HIP events around each launch on its own stream
             in calls after the timed region (the timed calls carry no
             events), against MI355X's 8 TB/s HBM peak;
             traffic = HBM bytes per launch from rocprofv3 FETCH_SIZE +
             WRITE_SIZE (profiles/traffic_sweep_512.md), or null
  hbm_copy   measured device-to-device copy rate (1 GiB torch copy_, read +
             write bytes / time), the practical ceiling beside the nominal peak
  cpu_baseline  the C restatement of the reference's OpenMP CPU path
             (oracle/fg_oracle.c, bit-exact with the reference) on the same
             volume and call: 1 warm-up, median of 5 on all host threads, plus
             a 1-thread rate on a 16-plane slab; host fields (nproc, affinity,
             CPU model, OMP_NUM_THREADS)
  small_2d_call  the 512^2 2D call (C=1, lamb 1, iter 2) on that CPU path
             (median of 10, all threads and 1 thread) and on the GPU (median
             of 50)
  click_latency  p50/p99 of the serving loop (64 resident sessions, sides
             96-320, signed iter 2) at 50 and 200 offered clicks/s
  batched_vs_per_call  one batched call against a loop of plain calls on 256
             uneven 2D + 32 uneven 3D instances (outputs compared bit for bit)
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def read_traffic(path: str):
    """HBM bytes per sweep launch from a committed rocprofv3 summary."""
    try:
        with open(path) as f:
            text = f.read()
    except OSError:
        return None
    m = re.search(r"traffic_bytes_per_launch:\s*([0-9.eE+]+)", text)
    return float(m.group(1)) if m else None


def _median_time(fn, runs: int) -> float:
    import numpy as np

    fn()  # warm-up
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        fn()
        times.append(time.perf_counter() - t0)
    return float(np.median(times))


def exchange_summary(hist, khz) -> dict:
    """Percentiles of the per-wave halo-exchange time from the log2 histogram
    of device-clock ticks (a bucket's upper edge is reported)."""
    total = sum(hist)
    if total == 0 or khz <= 0:
        return {"exchanges": 0}
    us_per_tick = 1e3 / khz

    def pct(q):
        acc = 0
        for b, c in enumerate(hist):
            acc += c
            if acc >= q * total:
                return (2 ** b) * us_per_tick
        return (2 ** (len(hist) - 1)) * us_per_tick

    mean = sum(c * (0 if b == 0 else 1.5 * 2 ** (b - 1)) for b, c in enumerate(hist)) / total * us_per_tick
    return {"wave_exchanges": total, "p50_us_le": round(pct(0.5), 3), "p90_us_le": round(pct(0.9), 3),
            "p99_us_le": round(pct(0.99), 3), "mean_us_approx": round(mean, 3),
            "histogram_log2_ticks": hist, "clock_khz": khz,
            "what": "per wave and exchange: time from publishing its halo band to receiving the last "
                    "neighbour value (device constant-rate clock, log2 buckets; p = bucket upper edge)"}


def host_fields() -> dict:
    try:
        affinity = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        affinity = os.cpu_count() or 1
    return {"nproc": os.cpu_count(), "affinity_cores": affinity, "cpu_model": cpu_model(),
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}


def cpu_baseline(img_gpu, mask_gpu, iters: int, runs: int = 5, slab_1t: int = 16):
    """The C restatement of the reference's OpenMP CPU path (oracle/fg_oracle.c,
    layout-faithful form: per-pass contiguous copies like the reference's
    transposes), bit-exact with the reference: 1 warm-up then the median of
    `runs` calls of the full volume on all threads, and a 1-thread figure on a
    `slab_1t`-deep slab of the same volume (median of 3)."""
    sys.path.insert(0, ROOT)
    import oracle

    img = img_gpu.cpu().numpy()
    mask = mask_gpu.cpu().numpy()
    S = img.shape[2:]
    n = int(img[0, 0].size)
    sp = [1.0, 1.0, 1.0]
    threads = oracle.num_threads()
    t_all = _median_time(lambda: oracle.geodesic3d_layout(img, mask, sp, 1e10, 1.0, 0.0, iters), runs)
    slab_i = img[:, :, :slab_1t].copy()
    slab_m = mask[:, :, :slab_1t].copy()
    oracle.set_num_threads(1)
    try:
        t_one = _median_time(lambda: oracle.geodesic3d_layout(slab_i, slab_m, sp, 1e10, 1.0, 0.0, iters), 3)
    finally:
        oracle.set_num_threads(threads)
    rate = n * 6 * iters / t_all
    rate_1t = int(slab_i[0, 0].size) * 6 * iters / t_one
    res = {"value": rate, "unit": "voxel-updates/s", "cores": threads, "kind": "port",
           "sample": (f"the same {S[0]}x{S[1]}x{S[2]} call (lamb=1.0, iter={iters}): 1 warm-up, median of "
                      f"{runs} = {t_all:.2f} s on {threads} threads; 1 thread: {slab_1t}x{S[1]}x{S[2]} slab, "
                      f"median of 3 = {t_one:.2f} s"),
           "value_1thread": rate_1t, "seconds_per_call": round(t_all, 3),
           "what": "oracle/fg_oracle.c layout-faithful restatement of fastgeodis_cpu.cpp, -O3 -fopenmp "
                   "-ffp-contract=off, bit-exact with the reference build on the golden fixtures"}
    res.update(host_fields())
    return res


def c1_cpu_and_gpu():
    """The small 2D call (512^2, C=1, lamb 1, iter 2): the CPU restatement
    (1 warm-up, median of 10, all threads and 1 thread) and the GPU (median of
    50 calls, each synchronised)."""
    sys.path.insert(0, ROOT)
    import oracle
    from fastgeodis_amd import synthetic, workloads

    img, mask = synthetic.config1("cpu")
    img, mask = img.numpy(), mask.numpy()
    threads = oracle.num_threads()
    fn = lambda: oracle.geodesic2d_layout(img, mask, 1e10, 1.0, 0.0, 2)  # noqa: E731
    t_cpu = _median_time(fn, 10)
    oracle.set_num_threads(1)
    try:
        t_one = _median_time(fn, 10)
    finally:
        oracle.set_num_threads(threads)
    gpu = workloads.run_c1(graphs=True)
    direct = workloads.run_c1(graphs=False)
    n_upd = 512 * 512 * 4 * 2
    return {"workload": "1x1x512x512 blob image, 4 seeds, lamb 1, iter 2",
            "cpu_ms": round(t_cpu * 1e3, 3), "cpu_vups": n_upd / t_cpu, "cpu_kind": "port",
            "cpu_threads": threads, "cpu_1thread_ms": round(t_one * 1e3, 3),
            "gpu_median_ms": round(gpu["median_ms"], 4), "gpu_vups": gpu["vups"],
            "gpu_median_ms_without_graphs": round(direct["median_ms"], 4)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--size", type=int, default=512)
    ap.add_argument("--iters", type=int, default=4)
    ap.add_argument("--halo", type=int, default=None, help="override the halo depth K")
    ap.add_argument("--lamb", type=float, default=1.0,
                    help="lamb of the timed call (the headline workload is 1.0; other values "
                         "run the general candidate form, for profiling)")
    ap.add_argument("--no-histogram", action="store_true",
                    help="skip the exchange-wait histogram pass (its instrumented launches are "
                         "slower and would mix into a kernel profile)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-serving", action="store_true",
                    help="skip the click-latency sample (64 resident sessions, Poisson arrivals)")
    ap.add_argument("--no-batch", action="store_true",
                    help="skip the batched-vs-per-call comparison on the uneven 2D/3D workload")
    ap.add_argument("--calibration-copy-mib", type=int, default=0,
                    help="after the timed region, run one plain device copy of this many MiB "
                         "(reference point for rocprofv3 byte counters)")
    ap.add_argument("--no-copy-rate", action="store_true",
                    help="skip the measured device-copy bandwidth")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the transform runs on a single MI355X "
              "(no multi-GPU path); rerun with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import torch

    sys.path.insert(0, ROOT)
    import fastgeodis_amd as FG
    from fastgeodis_amd import synthetic

    if not torch.cuda.is_available():
        print("bench.py needs a ROCm GPU", file=sys.stderr)
        sys.exit(2)
    if args.halo is not None:
        t = FG.get_tuning()
        FG.set_tuning(t["engine"], args.halo, t["max_groups"], t["shape"])
    tuning = FG.get_tuning()

    S = args.size
    log(f"[bench] generating {S}^3 input")
    img, mask = synthetic.headline_512("cuda", S)
    torch.cuda.synchronize()
    spacing = [1.0, 1.0, 1.0]
    N = S ** 3
    updates = N * 6 * args.iters

    def step():
        return FG.generalised_geodesic3d(img, mask, spacing, 1e10, args.lamb, args.iters)

    for w in range(args.warmup):
        step()
        torch.cuda.synchronize()
        log(f"[bench] warmup {w + 1}/{args.warmup} done")

    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k in range(args.steps):
        out = step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    # an expired halo wait in any timed call would have left the device's
    # status word set: raise instead of reporting invalid work
    FG.check_status()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    value = updates / (ms_per_step * 1e-3)
    log(f"[bench] {args.steps} steps: {ms_per_step:.3f} ms/step, {value / 1e9:.1f} Gvu/s")

    # per-launch durations of the sweep kernel: HIP events around each launch
    # on its own stream, in separate (untimed) calls
    prof_steps = max(1, min(args.steps, 3))
    FG.FastGeodisCpp.profile_enable(1)
    for k in range(prof_steps):
        step()
    FG.FastGeodisCpp.profile_enable(0)
    launches, kern_ms, kern_bytes = FG.FastGeodisCpp.profile_collect()
    # halo-exchange waits of the sweeps (device timestamps), in their own call
    exchange = None
    if not args.no_histogram:
        FG.FastGeodisCpp.profile_enable(2)
        step()
        FG.FastGeodisCpp.profile_enable(0)
        hist, khz = FG.FastGeodisCpp.profile_exchange_histogram()
        FG.check_status()
        exchange = exchange_summary(hist, khz)

    avg_ms = kern_ms / max(launches, 1)
    bytes_per_launch = kern_bytes / max(launches, 1)
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9 if launches else 0.0
    traffic = read_traffic(os.path.join(ROOT, "profiles", "traffic_sweep_512.md"))
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "kernel": "sweep_tiled_kernel<16 waves, 2 rows, 5-slice ring, C=1, "
                          + ("lamb=1" if args.lamb == 1.0 else "general lamb")
                          + "> (own-layout, H-backward B-contiguous and W-backward A-contiguous output forms)",
                "launches": launches,
                "mean_launch_ms": round(avg_ms, 4), "alg_bytes_per_launch": bytes_per_launch}
    whole_gbs = updates * 12 / (ms_per_step * 1e-3) / 1e9

    if args.calibration_copy_mib > 0:
        n = args.calibration_copy_mib * (1 << 20) // 4
        src = torch.ones(n, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] calibration copy of {args.calibration_copy_mib} MiB done")
        del src, dst

    # measured device-to-device copy rate beside the nominal 8 TB/s: 1 GiB
    # read + 1 GiB written per copy, 10 copies timed with events on the stream
    copy_res = None
    if not args.no_copy_rate:
        n = (1 << 30) // 4
        src = torch.ones(n, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(10):
            dst.copy_(src)
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / 10
        copy_res = {"bytes_per_copy": 2 * n * 4, "ms": round(ms, 4),
                    "GBps": round(2 * n * 4 / (ms * 1e-3) / 1e9, 1),
                    "what": "torch copy_ of 1 GiB float32 (read + write bytes / time)"}
        log(f"[bench] device copy {copy_res['GBps']} GB/s")
        del src, dst

    cpu = c1 = None
    if not args.no_cpu_baseline:
        log("[bench] timing the CPU baseline sample")
        cpu = cpu_baseline(img, mask, args.iters)
        log(f"[bench] cpu baseline {cpu['value'] / 1e9:.3f} Gvu/s ({cpu['kind']})")
        c1 = c1_cpu_and_gpu()
        log(f"[bench] 512^2 2D call: cpu {c1['cpu_ms']} ms, gpu {c1['gpu_median_ms']} ms")

    serving_res = None
    if not args.no_serving:
        from fastgeodis_amd import workloads

        log("[bench] click latency: 64 resident sessions, signed iter 2, Poisson arrivals")
        c5 = workloads.run_c5(rates=(50, 200), requests=1000)
        serving_res = {"workload": c5["workload"], "mean_session_voxels": c5["mean_session_voxels"],
                       "rows": [{k: (round(v, 2) if isinstance(v, float) else v) for k, v in r.items()}
                                for r in c5["rows"]]}
    batch_res = None
    if not args.no_batch:
        from fastgeodis_amd import workloads

        log("[bench] batched vs per-call: 256 uneven 2D + 32 uneven 3D instances")
        c4 = workloads.run_c4(reps=2)
        batch_res = {"workload": c4["workload"], "identical": c4["identical"],
                     "per_call_ms": round(c4["loop_ms"], 2), "batched_ms": round(c4["batch_ms"], 2),
                     "speedup": round(c4["speedup"], 2)}

    line = {
        "metric": "voxel-updates/s + % HBM peak, 512^3 geodesic3d; p50/p99 click latency",
        "value": value,
        "unit": "voxel-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        # reference's own 512^3 GPU figure: 0.92998 s for 512^3 x 6 x 2 updates
        # on a Titan X (figures/experiment_3d.json:18) = 1.73e9 voxel-updates/s
        "vs_baseline": value / (512 ** 3 * 12 / 0.92998),
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"generalised_geodesic3d {S}^3 C=1 lamb=1.0 iter={args.iters} "
                               "spacing=[1,1,1] v=1e10, 12-blob field, 16 seeds",
                   "shape": [1, 1, S, S, S], "iter": args.iters, "engine": tuning["engine"],
                   "halo": tuning["halo"], "tile_shape": tuning["shape"]},
        "whole_call_alg_GBps": round(whole_gbs, 1),
        "roofline": roofline,
        "exchange_wait": exchange,
        "hbm_copy": copy_res,
        "cpu_baseline": cpu,
        "small_2d_call": c1,
        "click_latency": serving_res,
        "batched_vs_per_call": batch_res,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
