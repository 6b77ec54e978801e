"""Writes tests/golden/crosscheck.npz + crosscheck.json: outputs of the
reference's two other CPU transforms, used as accuracy cross-checks of the
raster-scan path (the reference's own accuracy study compares against fast
marching, ``docs/source/methodology.rst:110-126``).

* Toivanen's raster scan (``FastGeodisCpp.generalised_geodesic{2,3}d_toivanen``,
  ``geodis_toivanen.cpp``), called as its Python wrappers call it
  (``FastGeodis/__init__.py:196-198, 235-237``: ``l_grad = lamb``,
  ``l_eucl = 1 - lamb``).  These pin the oracle's restatement.
* Fast marching in 2D (``FastGeodisCpp.geodesic2d_fastmarch``,
  ``geodis_fastmarch.cpp:163-248``, wrapper ``__init__.py:440-442``), which
  solves the eikonal equation with the third-party header library vendored
  as ``dependency/fmm`` (thinks/fast_marching_method, high-accuracy
  varying-speed solver).  Not restated: the outputs are the data.  The 3D
  form is not used: its gradient filter setup writes the z filter through the
  y filter's accessor, past that 2-element tensor
  (``geodis_fastmarch.cpp:265, 273-274``).

The expected outputs come from the reference extension compiled from its
sources into ``oracle/_ref`` by ``oracle/build.py``; the fixtures are data
(inputs and outputs).  Run in the build container:
    python oracle/build.py ref && python tests/golden/make_crosscheck.py
"""
from __future__ import annotations

import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

from make_golden import blob_field, seeds_mask  # noqa: E402


def cases():
    out = []
    k = 0
    for nd, C, shape, sp in [
        (2, 1, (40, 33), None),
        (2, 3, (27, 36), None),
        (3, 1, (10, 12, 11), (1.0, 1.0, 1.0)),
        (3, 2, (9, 13, 8), (0.7, 1.9, 1.3)),
        (3, 1, (8, 10, 12), (2.5, 1.0, 1.0)),
    ]:
        for lamb in (0.0, 0.5, 1.0):
            for it in (1, 3):
                out.append(dict(name=f"t{k}", method="toivanen", nd=nd, C=C, shape=shape,
                                spacing=sp, v=1e10, lamb=lamb, iter=it, nseeds=3))
                k += 1
    for C, shape in [(1, (64, 80)), (3, (72, 56))]:
        for lamb in (0.0, 0.5, 1.0):
            out.append(dict(name=f"f{k}", method="fastmarch", nd=2, C=C, shape=shape,
                            spacing=None, v=None, lamb=lamb, iter=None, nseeds=3))
            k += 1
    return out


def make_inputs(rng, case):
    shape = tuple(case["shape"])
    smax = max(2.0, max(shape) / 3.0)
    image = np.stack([blob_field(rng, shape, 5, 1.0, smax) for _ in range(case["C"])])[None]
    mask = seeds_mask(rng, shape, case["nseeds"])
    return image.astype(np.float32), mask


def run_reference(ref, case, image, mask):
    import torch

    ti, tm = torch.from_numpy(image), torch.from_numpy(mask)
    lamb = case["lamb"]
    if case["method"] == "toivanen":
        if case["nd"] == 2:
            r = ref.generalised_geodesic2d_toivanen(ti, tm, case["v"], lamb, 1 - lamb, case["iter"])
        else:
            r = ref.generalised_geodesic3d_toivanen(ti, tm, list(case["spacing"]), case["v"], lamb,
                                                    1 - lamb, case["iter"])
    else:
        r = ref.geodesic2d_fastmarch(ti, tm, lamb, 1 - lamb)
    return r.contiguous().numpy().astype(np.float32)


def main():
    import oracle

    ref = oracle.reference_module()
    if ref is None:
        raise SystemExit("reference build missing: run python oracle/build.py ref")
    rng = np.random.default_rng(7)
    arrays, meta = {}, []
    for case in cases():
        image, mask = make_inputs(rng, case)
        n = case["name"]
        arrays[n + "_image"] = image
        arrays[n + "_mask"] = mask
        arrays[n + "_out"] = run_reference(ref, case, image, mask)
        c = dict(case)
        c["shape"] = list(c["shape"])
        c["spacing"] = list(c["spacing"]) if c["spacing"] else None
        meta.append(c)
    np.savez_compressed(os.path.join(HERE, "crosscheck.npz"), **arrays)
    with open(os.path.join(HERE, "crosscheck.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(f"wrote {len(meta)} cases")


if __name__ == "__main__":
    main()
