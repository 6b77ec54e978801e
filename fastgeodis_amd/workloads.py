"""Throughput and latency workloads beside the 512^3 headline.

  c1  one 512^2 image (C=1), generalised_geodesic2d lamb 1, iter 2: the small
      2D call (median of repeated calls)
  c2  plain 3D, 1x1x128x256x256, spacing [2.5, 1, 1], unsigned and signed,
      lamb in {0, 0.5, 1}, iter 4
  c3  32 RGB 1024^2 images: generalised_geodesic2d (lamb 0.5, iter 2) and
      GSF2d (theta 0.1, lamb 1, iter 4), per-image calls against one batched call
  c4  256 uneven 2D + 32 uneven 3D instances (lamb 1, iter 2): a per-instance
      loop of plain calls against one batched call; outputs must be identical
  c5  interactive serving: 64 resident sessions, Poisson clicks, p50/p99

Each workload prints one JSON line.  Voxel-updates: N x passes x iter per
transform (passes = 4 in 2D, 6 in 3D; signed = 2 transforms, GSF = 6).

    python -m fastgeodis_amd.workloads c4 c5 --rates 50 200 1000
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np
import torch

from . import (GSF2d, Instance, generalised_geodesic2d, generalised_geodesic3d, geodesic_batch,
               signed_generalised_geodesic3d)
from . import serving, synthetic


def _updates(img: torch.Tensor, iters: int, transforms: int = 1) -> float:
    nd = img.dim() - 2
    return float(np.prod(img.shape[2:])) * (4 if nd == 2 else 6) * iters * transforms


def _timed(fn, reps: int):
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        out = fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps, out


def _same(a: torch.Tensor, b: torch.Tensor) -> bool:
    return a.shape == b.shape and torch.equal(a.contiguous().view(torch.int32), b.contiguous().view(torch.int32))


def run_c1(device="cuda", reps: int = 50, graphs: bool = True):
    """The 512^2 call, each timed alone (synchronised before and after); with
    graph replay on (the default) or off."""
    from . import set_graphs

    img, mask = synthetic.config1(device)
    set_graphs(graphs)
    try:
        for _ in range(3):
            generalised_geodesic2d(img, mask, 1e10, 1.0, 2)
        times = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            generalised_geodesic2d(img, mask, 1e10, 1.0, 2)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t0)
    finally:
        set_graphs(True)
    t = float(np.median(times))
    return {"workload": "c1 1x1x512x512 lamb 1 iter 2", "graphs": graphs, "median_ms": t * 1e3,
            "min_ms": float(np.min(times)) * 1e3, "vups": _updates(img, 2) / t}


def run_c2(device="cuda", reps: int = 5):
    img, mask, smask = synthetic.config2(device)
    sp = [2.5, 1.0, 1.0]
    rows = []
    for lamb in (0.0, 0.5, 1.0):
        t, _ = _timed(lambda: generalised_geodesic3d(img, mask, sp, 1e10, lamb, 4), reps)
        ts, _ = _timed(lambda: signed_generalised_geodesic3d(img, smask, sp, 1e10, lamb, 4), reps)
        rows.append({"lamb": lamb, "unsigned_ms": t * 1e3, "unsigned_vups": _updates(img, 4) / t,
                     "signed_ms": ts * 1e3, "signed_vups": _updates(img, 4, 2) / ts})
    return {"workload": "c2 128x256x256 spacing [2.5,1,1] iter 4", "rows": rows}


def run_c3(device="cuda", reps: int = 3, batch: int = 32):
    images, masks = synthetic.config3(device, batch)
    n_upd = sum(_updates(i, 2) for i in images)
    n_gsf = sum(_updates(i, 4, 6) for i in images)
    t_loop, outs_loop = _timed(lambda: [generalised_geodesic2d(i, m, 1e10, 0.5, 2) for i, m in zip(images, masks)], reps)
    insts = [Instance(i, m, 1e10, 0.5, iter=2) for i, m in zip(images, masks)]
    t_batch, outs_batch = _timed(lambda: geodesic_batch(insts), reps)
    g_loop, gl = _timed(lambda: [GSF2d(i, m, 0.1, 1e10, 1.0, 4) for i, m in zip(images, masks)], 1)
    ginsts = [Instance(i, m, 1e10, 1.0, iter=4, op="gsf", theta=0.1) for i, m in zip(images, masks)]
    g_batch, gb = _timed(lambda: geodesic_batch(ginsts), 1)
    same = all(_same(a, b) for a, b in zip(outs_loop, outs_batch)) and all(_same(a, b) for a, b in zip(gl, gb))
    return {"workload": f"c3 {batch}x(1x3x1024x1024)", "identical": same,
            "geodesic2d_loop_ms": t_loop * 1e3, "geodesic2d_batch_ms": t_batch * 1e3,
            "geodesic2d_loop_vups": n_upd / t_loop, "geodesic2d_batch_vups": n_upd / t_batch,
            "gsf2d_loop_ms": g_loop * 1e3, "gsf2d_batch_ms": g_batch * 1e3,
            "gsf2d_loop_vups": n_gsf / g_loop, "gsf2d_batch_vups": n_gsf / g_batch}


def run_c4(device="cuda", reps: int = 3):
    items = synthetic.config4(device)
    n_upd = sum(_updates(i, 2) for i, _, _ in items)

    def single(i, m, sp):
        if sp is None:
            return generalised_geodesic2d(i, m, 1e10, 1.0, 2)
        return generalised_geodesic3d(i, m, sp, 1e10, 1.0, 2)

    from . import autotune

    t_loop, outs_loop = _timed(lambda: [single(*it) for it in items], reps)
    insts = [Instance(i, m, 1e10, 1.0, iter=2, spacing=sp) for i, m, sp in items]
    t_batch, outs_batch = _timed(lambda: geodesic_batch(insts), reps)
    tuning = autotune.tune_batch(insts, reps=max(1, reps - 1))
    t_tuned, outs_tuned = _timed(lambda: geodesic_batch(insts, tuned=True), reps)
    same = all(_same(a, b) for a, b in zip(outs_loop, outs_batch)) and \
        all(_same(a, b) for a, b in zip(outs_batch, outs_tuned))
    return {"workload": "c4 256 uneven 2D (64-2048, C 1|3) + 32 uneven 3D (64-256), lamb 1, iter 2",
            "identical": same, "voxel_updates": n_upd,
            "loop_ms": t_loop * 1e3, "batch_ms": t_batch * 1e3, "tuned_batch_ms": t_tuned * 1e3,
            "tuning": tuning["best"], "tuning_table": tuning["table"],
            "loop_vups": n_upd / t_loop, "batch_vups": n_upd / t_batch, "tuned_vups": n_upd / t_tuned,
            "speedup": t_loop / t_batch, "tuned_speedup": t_loop / t_tuned}


def run_c5(rates=(50, 200, 500, 1000, 2000), requests: int = 1000, sessions: int = 64,
           budget_gb: float = 8.0, device="cuda"):
    srv = serving.Server(budget_bytes=int(budget_gb * (1 << 30)), device=device)
    serving.make_sessions(srv, sessions, 96, 320, seed=5, spacing=(2.0, 1.0, 1.0))
    torch.cuda.synchronize()
    rows = []
    srv.start()
    try:
        serving.poisson_load(srv, 20.0, 20)  # warm-up
        for q in rates:
            print(f"[workloads] c5: {requests} requests at {q:g} clicks/s", file=sys.stderr, flush=True)
            lat, wall, bs = serving.poisson_load(srv, float(q), requests, seed=int(q))
            lat = np.sort(np.asarray(lat)) * 1e3
            rows.append({"qps_offered": q, "requests": len(lat), "throughput_rps": len(lat) / wall,
                         "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                         "mean_batch": float(np.mean(bs))})
    finally:
        srv.stop()
    vox = [s.voxels for s in srv.sessions.values()]
    return {"workload": f"c5 {sessions} sessions (sides 96-320, spacing [2,1,1]), signed iter 2, "
                        f"budget {budget_gb} GiB", "mean_session_voxels": float(np.mean(vox)), "rows": rows}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("which", nargs="+", choices=["c1", "c2", "c3", "c4", "c5"])
    ap.add_argument("--rates", type=float, nargs="+", default=[50, 200, 500, 1000, 2000])
    ap.add_argument("--requests", type=int, default=1000)
    args = ap.parse_args(argv)
    if not torch.cuda.is_available():
        print("workloads need a ROCm GPU", file=sys.stderr)
        return 2
    for w in args.which:
        print(f"[workloads] {w}", file=sys.stderr, flush=True)
        if w == "c5":
            res = run_c5(args.rates, args.requests)
        else:
            res = {"c1": run_c1, "c2": run_c2, "c3": run_c3, "c4": run_c4}[w]()
        print(json.dumps(res), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
