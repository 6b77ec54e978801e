"""Empirical tuner for the persistent sweep engine.

The engine's own defaults are analytic: the 16 x 2 workgroup with a 5-slice
prefetch ring and the deepest halo whose tiles fit the device (per lone
plane, or over the planes that share a pass).  For other problem sizes a different workgroup shape or halo
depth can be faster; this module measures that directly: it builds a
synthetic blob-field input of the requested size on the device, times the
plain transform under each candidate ``(engine, halo, shape)`` (see
:func:`fastgeodis_amd.set_tuning`), checks that every candidate returns the
same bits, and keeps the fastest.  Results are cached per problem signature
in memory and, when ``FASTGEODIS_AMD_TUNE_CACHE`` names a file, in that JSON
file.

    from fastgeodis_amd import autotune
    best = autotune.tune((128, 256, 256), channels=1, spacing=(2.5, 1, 1))
    autotune.apply(best)            # tuning for later calls on this thread

Batched calls are tuned as a whole: :func:`tune_batch` times the complete
batched call under each candidate tuning (halo depth, 3D workgroup shape,
cap on co-resident workgroups per launch, which splits a large batch into
more launches) and caches the fastest per *size class* of the batch
(:func:`batch_signature`: the multiset of instance kinds with extents rounded
to powers of two, and the device).  ``geodesic_batch(..., tuned=True)``
applies the cached tuning of its size class, measuring it first when absent.
"""
from __future__ import annotations

import json
import os
import time
from typing import Dict, Iterable, List, Optional, Sequence, Tuple

import torch

from . import generalised_geodesic2d, generalised_geodesic3d, get_tuning, set_tuning

# (engine, halo, shape): halo 0 = automatic depth, shape 0 = default
DEFAULT_CANDIDATES_3D: Tuple[Tuple[int, int, int], ...] = (
    (1, 0, 5), (1, 0, 4), (1, 0, 2), (1, 0, 3), (1, 0, 1),
    (1, 3, 2), (1, 6, 2), (1, 10, 4), (1, 8, 1), (1, 0, 7),
)
DEFAULT_CANDIDATES_2D: Tuple[Tuple[int, int, int], ...] = (
    (1, 0, 0), (1, 6, 0), (1, 12, 0), (1, 24, 0), (1, 30, 0),
)

_cache: Dict[str, dict] = {}


def variant_class(lamb: float) -> str:
    """The kernel variant a lamb selects: "l1" (lamb == 1: the multiply-free
    specialisation), "fast" (0 <= lamb <= 1: every step weight non-negative,
    minimum over candidates) or "exact" (candidates folded one by one)."""
    lamb = float(lamb)
    if lamb == 1.0:
        return "l1"
    if 0.0 <= lamb <= 1.0:
        return "fast"
    return "exact"


def device_tag(device="cuda") -> str:
    """Device name and compute-unit count (tunings do not carry over between
    devices of different size); "nodev" without a GPU."""
    if not torch.cuda.is_available():
        return "nodev"
    p = torch.cuda.get_device_properties(torch.device(device))
    name = getattr(p, "gcnArchName", "") or p.name
    return f"{name.split(':')[0]}-{p.multi_processor_count}cu"


def signature(shape: Sequence[int], channels: int, lamb: float = 1.0, iters: int = 1,
              device="cuda") -> str:
    return (f"{len(shape)}d:{'x'.join(str(int(s)) for s in shape)}:c{int(channels)}:"
            f"{variant_class(lamb)}:it{int(iters)}:{device_tag(device)}")


def _cache_file() -> Optional[str]:
    return os.environ.get("FASTGEODIS_AMD_TUNE_CACHE") or None


def load_cache() -> Dict[str, dict]:
    path = _cache_file()
    if path and os.path.isfile(path):
        with open(path) as f:
            _cache.update(json.load(f))
    return _cache


def _save_cache() -> None:
    path = _cache_file()
    if path:
        tmp = path + ".tmp"
        with open(tmp, "w") as f:
            json.dump(_cache, f, indent=1, sort_keys=True)
        os.replace(tmp, path)


def _input(shape: Sequence[int], channels: int, device):
    from .synthetic import blob_field, seed_mask

    chans = [blob_field(shape, 8, (8.0, 40.0), 9000 + c, device) for c in range(channels)]
    img = torch.stack(chans)[None]
    return img, seed_mask(shape, 8, 9000, device)


def tune(shape: Sequence[int], channels: int = 1, spacing: Sequence[float] = (1.0, 1.0, 1.0),
         lamb: float = 1.0, iters: int = 1, reps: int = 3,
         candidates: Optional[Iterable[Tuple[int, int, int]]] = None, device="cuda",
         use_cache: bool = True) -> dict:
    """Measures candidate tunings for one problem size; returns
    ``{"signature", "best": {"engine", "halo", "shape"}, "ms", "table"}``."""
    shape = [int(s) for s in shape]
    if len(shape) not in (2, 3):
        raise ValueError("shape must be (H, W) or (D, H, W)")
    sig = signature(shape, channels, lamb, iters, device)
    if use_cache and sig in load_cache():
        return _cache[sig]
    if candidates is None:
        candidates = DEFAULT_CANDIDATES_3D if len(shape) == 3 else DEFAULT_CANDIDATES_2D
    img, mask = _input(shape, channels, device)

    def call():
        if len(shape) == 2:
            return generalised_geodesic2d(img, mask, 1e10, lamb, iters)
        return generalised_geodesic3d(img, mask, list(spacing), 1e10, lamb, iters)

    candidates = list(candidates)
    old = get_tuning()
    table: List[dict] = []
    ref = None
    try:
        # first pass: every candidate once (warm-up, and the results must agree)
        for (engine, halo, shp) in candidates:
            set_tuning(engine, halo, 0, shp)
            out = call()
            torch.cuda.synchronize()
            if ref is None:
                ref = out.clone()
            elif not torch.equal(out.view(torch.int32), ref.view(torch.int32)):
                raise RuntimeError(f"tuning {(engine, halo, shp)} changed the result")
        # second pass: timing
        for (engine, halo, shp) in candidates:
            set_tuning(engine, halo, 0, shp)
            call()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                call()
            torch.cuda.synchronize()
            table.append({"engine": engine, "halo": halo, "shape": shp,
                          "ms": (time.perf_counter() - t0) * 1e3 / reps})
    finally:
        set_tuning(**old)
    best = min(table, key=lambda r: r["ms"])
    res = {"signature": sig, "best": {k: best[k] for k in ("engine", "halo", "shape")},
           "ms": best["ms"], "table": table}
    _cache[sig] = res
    _save_cache()
    return res


def apply(result: dict) -> None:
    """Makes a :func:`tune` result the tuning of later calls on this thread."""
    b = result["best"]
    set_tuning(int(b["engine"]), int(b["halo"]), 0, int(b["shape"]))


# ---- whole batches ----------------------------------------------------------

# (engine, halo, shape, max_groups): halo 0 = automatic, shape 0 = default,
# max_groups 0 = every co-resident workgroup, < 0 = that fraction of the
# compute units (-2: half)
DEFAULT_BATCH_CANDIDATES: Tuple[Tuple[int, int, int, int], ...] = (
    (1, 0, 0, 0), (1, 0, 4, 0), (1, 0, 2, 0), (1, 5, 0, 0), (1, 10, 0, 0), (1, 0, 0, -2),
)


def _bucket(x: int) -> int:
    """Power of two nearest to x (in log scale)."""
    import math

    return 1 << max(0, int(round(math.log2(max(int(x), 1)))))


def batch_signature(instances, device="cuda") -> str:
    """Size class of a batch: sorted (dims, C, op, variant class, iterations,
    extents rounded to powers of two) of every instance, and the device."""
    items = []
    for i in instances:
        ext = tuple(_bucket(x) for x in i.image.shape[2:])
        items.append(f"{len(ext)}d{'x'.join(map(str, ext))}c{int(i.image.shape[1])}{i.op[0]}"
                     f"{variant_class(i.lamb)}i{int(i.iter)}")
    items.sort()
    import hashlib

    h = hashlib.sha1(",".join(items).encode()).hexdigest()[:16]
    return f"batch:{len(items)}:{h}:{device_tag(device)}"


def _resolve_groups(g: int, device) -> int:
    if g >= 0:
        return g
    cus = torch.cuda.get_device_properties(torch.device(device)).multi_processor_count
    return max(1, cus // (-g))


def tune_batch(instances, reps: int = 2, candidates=None, device="cuda", use_cache: bool = True) -> dict:
    """Times the batched call of `instances` under each candidate tuning and
    returns ``{"signature", "best": {"engine", "halo", "shape", "max_groups"},
    "ms", "analytic_ms", "table"}``; every candidate must return the same
    bits as the analytic default (the first candidate)."""
    from . import geodesic_batch

    sig = batch_signature(instances, device)
    if use_cache and sig in load_cache():
        return _cache[sig]
    cands = list(candidates if candidates is not None else DEFAULT_BATCH_CANDIDATES)
    old = get_tuning()
    table: List[dict] = []
    ref = None
    try:
        for (engine, halo, shp, groups) in cands:
            g = _resolve_groups(groups, device)
            set_tuning(engine, halo, g, shp)
            outs = geodesic_batch(instances)
            torch.cuda.synchronize()
            if ref is None:
                ref = [o.clone() for o in outs]
            elif not all(torch.equal(a.view(torch.int32), b.view(torch.int32)) for a, b in zip(ref, outs)):
                raise RuntimeError(f"batch tuning {(engine, halo, shp, g)} changed a result")
            del outs
        for (engine, halo, shp, groups) in cands:
            g = _resolve_groups(groups, device)
            set_tuning(engine, halo, g, shp)
            geodesic_batch(instances)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                geodesic_batch(instances)
            torch.cuda.synchronize()
            table.append({"engine": engine, "halo": halo, "shape": shp, "max_groups": g,
                          "ms": (time.perf_counter() - t0) * 1e3 / reps})
    finally:
        set_tuning(**old)
    best = min(table, key=lambda r: r["ms"])
    res = {"signature": sig, "best": {k: best[k] for k in ("engine", "halo", "shape", "max_groups")},
           "ms": best["ms"], "analytic_ms": table[0]["ms"], "table": table}
    _cache[sig] = res
    _save_cache()
    return res


def tuning_for_batch(instances, device="cuda", measure: bool = True) -> Optional[dict]:
    """The cached tuning of the batch's size class ({"engine", "halo",
    "shape", "max_groups"}); measured now when absent and `measure`."""
    sig = batch_signature(instances, device)
    if sig in load_cache():
        return _cache[sig]["best"]
    if not measure:
        return None
    return tune_batch(instances, device=device)["best"]
