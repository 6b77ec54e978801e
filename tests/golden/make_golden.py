"""Writes tests/golden/fixtures.npz + cases.json from the reference CPU path.

The expected outputs come from the reference's own ``FastGeodisCpp`` CPU
extension, compiled from ``/root/reference/FastGeodis/*.cpp`` into
``oracle/_ref`` by ``oracle/build.py``.  It is called exactly as the Python
wrappers call it (``FastGeodis/__init__.py:63-65, 96-98, 127-129, 160-162,
553, 585``): ``l_grad = lamb`` and ``l_eucl = 1 - lamb`` in double for the
unsigned/signed entry points, ``lamb`` alone for GSF.  The fixtures are data
(inputs and expected outputs); nothing from the reference is stored here.

Run in the build container (where /root/reference exists):
    python oracle/build.py ref && python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)


def blob_field(rng, shape, nblobs, smin, smax):
    """Sum of Gaussian blobs with uniform centres plus U[0, 0.05] noise."""
    grids = np.meshgrid(*[np.arange(s, dtype=np.float64) for s in shape], indexing="ij")
    f = np.zeros(shape, dtype=np.float64)
    for _ in range(nblobs):
        c = [rng.uniform(0, s) for s in shape]
        sig = rng.uniform(smin, smax)
        amp = rng.uniform(0, 1)
        r2 = sum((g - ci) ** 2 for g, ci in zip(grids, c))
        f += amp * np.exp(-r2 / (2 * sig * sig))
    f += rng.uniform(0, 0.05, size=shape)
    return f.astype(np.float32)


def seeds_mask(rng, shape, nseeds, channels=1):
    m = np.ones((1, channels, *shape), dtype=np.float32)
    for _ in range(nseeds):
        idx = tuple(int(rng.integers(0, s)) for s in shape)
        m[(0, 0) + idx] = 0.0
    return m


def cases():
    """(name, kind, nd, C, Cm, shape, spacing, v, lamb, iter, theta, maskkind)."""
    out = []
    k = 0
    for nd, C, shape, sp in [
        (2, 1, (48, 40), None),
        (2, 3, (33, 45), None),
        (3, 1, (12, 16, 20), (2.5, 1.0, 1.0)),
        (3, 2, (9, 13, 11), (0.7, 1.9, 1.3)),
    ]:
        for lamb in (0.0, 0.5, 1.0):
            for it in (1, 2, 4):
                out.append(dict(name=f"c{k}", kind="unsigned", nd=nd, C=C, Cm=1,
                                shape=shape, spacing=sp, v=1e10, lamb=lamb, iter=it,
                                theta=0.0, mask="seeds"))
                k += 1
        for lamb in (0.3, 1.0):
            out.append(dict(name=f"c{k}", kind="signed", nd=nd, C=C, Cm=1, shape=shape,
                            spacing=sp, v=1e10, lamb=lamb, iter=2, theta=0.0,
                            mask="binary"))
            k += 1
        out.append(dict(name=f"c{k}", kind="gsf", nd=nd, C=C, Cm=1, shape=shape,
                        spacing=sp, v=1e10, lamb=1.0, iter=2, theta=0.1, mask="binary"))
        k += 1
        out.append(dict(name=f"c{k}", kind="gsf", nd=nd, C=C, Cm=1, shape=shape,
                        spacing=sp, v=1000.0, lamb=0.7, iter=4, theta=0.5, mask="soft"))
        k += 1
    # edge shapes and odd arguments
    extra = [
        dict(kind="unsigned", nd=2, C=1, Cm=2, shape=(1, 37), spacing=None, v=1e10,
             lamb=0.5, iter=2, theta=0.0, mask="seeds"),
        dict(kind="unsigned", nd=2, C=2, Cm=1, shape=(29, 1), spacing=None, v=1e10,
             lamb=0.5, iter=2, theta=0.0, mask="seeds"),
        dict(kind="unsigned", nd=3, C=1, Cm=1, shape=(1, 17, 23), spacing=(1.0, 1.0, 1.0),
             v=1e10, lamb=0.25, iter=2, theta=0.0, mask="seeds"),
        dict(kind="unsigned", nd=3, C=1, Cm=2, shape=(7, 1, 9), spacing=(3.0, 0.5, 1.5),
             v=1e10, lamb=0.75, iter=3, theta=0.0, mask="seeds"),
        dict(kind="unsigned", nd=3, C=3, Cm=1, shape=(6, 5, 1), spacing=(1.0, 2.0, 3.0),
             v=1e10, lamb=0.6, iter=2, theta=0.0, mask="seeds"),
        dict(kind="unsigned", nd=2, C=1, Cm=1, shape=(20, 24), spacing=None, v=1e10,
             lamb=0.5, iter=0, theta=0.0, mask="seeds"),
        dict(kind="unsigned", nd=2, C=1, Cm=1, shape=(31, 27), spacing=None, v=7.5,
             lamb=0.1, iter=3, theta=0.0, mask="soft"),
        dict(kind="unsigned", nd=3, C=1, Cm=1, shape=(10, 11, 12), spacing=(1.0, 1.0, 1.0),
             v=50.0, lamb=0.9, iter=2, theta=0.0, mask="soft"),
        dict(kind="unsigned", nd=2, C=1, Cm=1, shape=(25, 30), spacing=None, v=1e10,
             lamb=0.5, iter=2, theta=0.0, mask="seeds_nan"),
        dict(kind="signed", nd=3, C=1, Cm=1, shape=(8, 9, 10), spacing=(1.5, 1.0, 0.8),
             v=1e10, lamb=0.0, iter=4, theta=0.0, mask="binary"),
    ]
    for e in extra:
        e["name"] = f"c{k}"
        k += 1
        out.append(e)
    return out


def make_inputs(rng, case):
    shape = tuple(case["shape"])
    nd, C, Cm = case["nd"], case["C"], case["Cm"]
    smax = max(2.0, max(shape) / 3.0)
    image = np.stack([blob_field(rng, shape, 4, 1.0, smax) for _ in range(C)])[None]
    mk = case["mask"]
    if mk in ("seeds", "seeds_nan"):
        mask = seeds_mask(rng, shape, 3, Cm)
        if Cm > 1:
            mask[:, 1:] = rng.random(mask[:, 1:].shape).astype(np.float32)
        if mk == "seeds_nan":
            image = image.copy()
            image[(0, 0) + tuple(s // 2 for s in shape)] = np.nan
    elif mk == "binary":
        f = blob_field(rng, shape, 3, 1.0, smax)
        f = (f - f.min()) / max(f.max() - f.min(), 1e-6)
        mask = (f > 0.5).astype(np.float32)[None, None]
        mask = np.repeat(mask, Cm, axis=1)
    else:  # soft
        mask = rng.random((1, Cm, *shape)).astype(np.float32)
        mask[(0, 0) + tuple(s // 3 for s in shape)] = 0.0
    return image.astype(np.float32), mask.astype(np.float32)


def run_reference(ref, case, image, mask):
    import torch

    ti, tm = torch.from_numpy(image), torch.from_numpy(mask)
    v, lamb, it = case["v"], case["lamb"], case["iter"]
    sp = list(case["spacing"]) if case["spacing"] else None
    kind, nd = case["kind"], case["nd"]
    if kind == "unsigned":
        r = (ref.generalised_geodesic2d(ti, tm, v, lamb, 1 - lamb, it) if nd == 2 else
             ref.generalised_geodesic3d(ti, tm, sp, v, lamb, 1 - lamb, it))
    elif kind == "signed":
        r = (ref.signed_generalised_geodesic2d(ti, tm, v, lamb, 1 - lamb, it) if nd == 2 else
             ref.signed_generalised_geodesic3d(ti, tm, sp, v, lamb, 1 - lamb, it))
    else:
        r = (ref.GSF2d(ti, tm, case["theta"], v, lamb, it) if nd == 2 else
             ref.GSF3d(ti, tm, case["theta"], sp, v, lamb, it))
    return r.contiguous().numpy().astype(np.float32)


def main():
    import oracle

    ref = oracle.reference_module()
    if ref is None:
        raise SystemExit("reference build missing: run python oracle/build.py ref")
    rng = np.random.default_rng(2024)
    arrays = {}
    meta = []
    for case in cases():
        image, mask = make_inputs(rng, case)
        out = run_reference(ref, case, image, mask)
        n = case["name"]
        arrays[n + "_image"] = image
        arrays[n + "_mask"] = mask
        arrays[n + "_out"] = out
        c = dict(case)
        c["shape"] = list(c["shape"])
        c["spacing"] = list(c["spacing"]) if c["spacing"] else None
        meta.append(c)
    np.savez_compressed(os.path.join(HERE, "fixtures.npz"), **arrays)
    with open(os.path.join(HERE, "cases.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(f"wrote {len(meta)} cases")


if __name__ == "__main__":
    main()
