"""Accuracy cross-checks of the HIP raster-scan path against the reference's
other two transforms (its accuracy study: docs/source/methodology.rst:110-126).

* Toivanen's raster scan (geodis_toivanen.cpp, restated in the oracle and
  pinned there against the reference's outputs): at convergence the HIP
  transform equals it bit for bit, in 2D and 3D, isotropic and anisotropic,
  at sizes that run the tiled engine with several tiles per plane.
* Euclidean mode (lamb = 0) against the exact Euclidean distance
  (scipy.ndimage.distance_transform_edt): the HIP result lies between it and
  the chamfer bound (1.0824 in 2D, 1.1281 in 3D) at full image sizes.
* Fast marching (geodis_fastmarch.cpp with its vendored eikonal solver;
  outputs written by the reference into tests/golden/crosscheck.npz): the
  agreement statistics the reference reports as joint histograms, here as a
  correlation and a median ratio away from the seeds.  For lamb = 1 the two
  methods model the image term differently (per-step |dI| against a
  forward-difference gradient magnitude), so only a loose correlation holds.
"""
import json
import os

import numpy as np
import pytest
import torch
from scipy import ndimage

import fastgeodis_amd as FG
import oracle

from test_crosscheck_oracle import bits_equal, chamfer_ratio_bound, converged, seeded

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
with open(os.path.join(HERE, "golden", "crosscheck.json")) as f:
    XCASES = json.load(f)
XDATA = np.load(os.path.join(HERE, "golden", "crosscheck.npz"))
FASTMARCH = [c for c in XCASES if c["method"] == "fastmarch"]


def say(*a):
    print(*a, flush=True)


def hip2d(image, mask, lamb, iters):
    return FG.generalised_geodesic2d(torch.from_numpy(image).cuda(), torch.from_numpy(mask).cuda(), 1e10,
                                     lamb, iters).cpu().numpy()


def hip3d(image, mask, spacing, lamb, iters):
    return FG.generalised_geodesic3d(torch.from_numpy(image).cuda(), torch.from_numpy(mask).cuda(),
                                     spacing, 1e10, lamb, iters).cpu().numpy()


@pytest.mark.parametrize("lamb", [0.0, 0.5, 1.0])
@pytest.mark.parametrize("C", [1, 3])
def test_converged_hip_equals_converged_toivanen_2d(device, lamb, C):
    rng = np.random.default_rng(200 + C + int(10 * lamb))
    image, mask = seeded(rng, (150, 211), C, 4)
    say(f"toivanen 2d C={C} lamb={lamb} shape=(150, 211)")
    lg, le = oracle.wrapper_weights(lamb)
    want, n = converged(lambda it: oracle.toivanen2d(image, mask, 1e10, lg, le, it))
    got = hip2d(image, mask, lamb, n + 2)
    assert bits_equal(got, want)


@pytest.mark.parametrize("spacing", [[1.0, 1.0, 1.0], [2.5, 1.0, 1.0], [0.7, 1.9, 1.3]])
@pytest.mark.parametrize("lamb", [0.0, 0.5, 1.0])
def test_converged_hip_equals_converged_toivanen_3d(device, lamb, spacing):
    rng = np.random.default_rng(300 + int(10 * lamb) + int(10 * spacing[0]))
    image, mask = seeded(rng, (19, 52, 67), 1, 3)
    say(f"toivanen 3d lamb={lamb} spacing={spacing} shape=(19, 52, 67)")
    lg, le = oracle.wrapper_weights(lamb)
    want, n = converged(lambda it: oracle.toivanen3d(image, mask, spacing, 1e10, lg, le, it))
    got = hip3d(image, mask, spacing, lamb, n + 2)
    assert bits_equal(got, want)


@pytest.mark.parametrize("nd", [2, 3])
def test_euclidean_mode_within_chamfer_bounds_full_size(device, nd):
    rng = np.random.default_rng(500 + nd)
    shape = (512, 512) if nd == 2 else (96, 128, 112)
    image, mask = seeded(rng, shape, 1, 12)
    say(f"euclidean {nd}d shape={shape}")
    if nd == 2:
        a, b = hip2d(image, mask, 0.0, 4), hip2d(image, mask, 0.0, 5)
    else:
        a, b = hip3d(image, mask, [1.0, 1.0, 1.0], 0.0, 4), hip3d(image, mask, [1.0, 1.0, 1.0], 0.0, 5)
    assert bits_equal(a, b)  # converged
    edt = ndimage.distance_transform_edt(mask[0, 0])
    sel = edt > 0
    r = a[0, 0][sel].astype(np.float64) / edt[sel]
    say(f"  ratio to Euclidean: min {r.min():.7f} max {r.max():.5f} bound {chamfer_ratio_bound(nd):.5f}")
    assert r.min() >= 1 - 1e-6
    assert r.max() <= chamfer_ratio_bound(nd) + 1e-6
    assert np.all(a[0, 0][~sel] == 0)


@pytest.mark.parametrize("case", FASTMARCH, ids=[c["name"] for c in FASTMARCH])
def test_agreement_with_fast_marching_2d(device, case):
    n = case["name"]
    image, mask, fmm = XDATA[n + "_image"], XDATA[n + "_mask"], XDATA[n + "_out"]
    got = hip2d(image, mask, case["lamb"], 8)[0, 0].astype(np.float64)
    f = fmm[0, 0].astype(np.float64)
    far = ndimage.distance_transform_edt(mask[0, 0]) > 3
    corr = float(np.corrcoef(got[far], f[far])[0, 1])
    med = float(np.median(got[far] / f[far]))
    say(f"fast marching {n} C={case['C']} lamb={case['lamb']}: correlation {corr:.4f}, median ratio {med:.4f}")
    if case["lamb"] < 1.0:
        assert corr >= 0.995
        assert 0.95 <= med <= 1.1
    else:
        assert corr >= 0.5
