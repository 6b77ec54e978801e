"""Accuracy cross-checks, CPU side: the oracle's restatement of Toivanen's
raster scan (geodis_toivanen.cpp) against the reference's own outputs, and
the facts the GPU cross-checks rely on, checked on the oracle:

* at convergence the raster-scan transform and Toivanen's scan give the same
  values bit for bit (both relax the same 8- / 26-neighbour graph with the
  same two-rounding candidate, and the limit from the same initialisation is
  that graph's shortest-path fold, whatever the visiting order);
* with lamb = 0 the converged transform is the chamfer distance of that
  graph, which lies between the Euclidean distance and a fixed multiple of it
  (1.0824 in 2D, 1.1281 in 3D for unit spacing), checked against
  scipy.ndimage.distance_transform_edt.
"""
import json
import os

import numpy as np
import pytest
from scipy import ndimage

import oracle

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")

with open(os.path.join(GOLDEN, "crosscheck.json")) as f:
    XCASES = json.load(f)
XDATA = np.load(os.path.join(GOLDEN, "crosscheck.npz"))
TOIVANEN = [c for c in XCASES if c["method"] == "toivanen"]


def bits_equal(a, b):
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def toivanen(case, image, mask, iters):
    lg, le = oracle.wrapper_weights(case["lamb"])
    if case["nd"] == 2:
        return oracle.toivanen2d(image, mask, case["v"], lg, le, iters)
    return oracle.toivanen3d(image, mask, case["spacing"], case["v"], lg, le, iters)


@pytest.mark.parametrize("case", TOIVANEN, ids=[c["name"] for c in TOIVANEN])
def test_toivanen_restatement_matches_reference_outputs(case):
    n = case["name"]
    out = toivanen(case, XDATA[n + "_image"], XDATA[n + "_mask"], case["iter"])
    assert bits_equal(out, XDATA[n + "_out"])


def test_crosscheck_fixtures_cover_both_methods():
    kinds = {(c["method"], c["nd"]) for c in XCASES}
    assert kinds == {("toivanen", 2), ("toivanen", 3), ("fastmarch", 2)}
    assert any(c["method"] == "toivanen" and c["nd"] == 3 and len(set(c["spacing"])) == 3 for c in XCASES)
    assert {c["lamb"] for c in XCASES} == {0.0, 0.5, 1.0}


def converged(fn, max_iter=64):
    """fn(iters) at the first iteration count after which it stops changing."""
    prev = fn(1)
    for it in range(2, max_iter):
        cur = fn(it)
        if bits_equal(cur, prev):
            return cur, it - 1
        prev = cur
    raise AssertionError("no convergence")


def seeded(rng, shape, C, nseeds):
    image = rng.random((1, C, *shape), dtype=np.float32)
    mask = np.ones((1, 1, *shape), np.float32)
    for _ in range(nseeds):
        mask[(0, 0) + tuple(int(rng.integers(0, s)) for s in shape)] = 0.0
    return image, mask


@pytest.mark.parametrize("lamb", [0.0, 0.3, 1.0])
@pytest.mark.parametrize("C", [1, 3])
def test_converged_raster_equals_converged_toivanen_2d(lamb, C):
    rng = np.random.default_rng(int(lamb * 10) + C)
    image, mask = seeded(rng, (37, 29), C, 3)
    lg, le = oracle.wrapper_weights(lamb)
    a, _ = converged(lambda it: oracle.geodesic2d(image, mask, 1e10, lg, le, it))
    b, _ = converged(lambda it: oracle.toivanen2d(image, mask, 1e10, lg, le, it))
    assert bits_equal(a, b)


@pytest.mark.parametrize("spacing", [[1.0, 1.0, 1.0], [2.5, 1.0, 1.0], [0.7, 1.9, 1.3]])
@pytest.mark.parametrize("lamb", [0.0, 0.5, 1.0])
def test_converged_raster_equals_converged_toivanen_3d(lamb, spacing):
    rng = np.random.default_rng(int(lamb * 10) + int(spacing[0] * 10))
    image, mask = seeded(rng, (9, 11, 10), 1, 2)
    lg, le = oracle.wrapper_weights(lamb)
    a, _ = converged(lambda it: oracle.geodesic3d(image, mask, spacing, 1e10, lg, le, it))
    b, _ = converged(lambda it: oracle.toivanen3d(image, mask, spacing, 1e10, lg, le, it))
    assert bits_equal(a, b)


def test_a_planted_wrong_step_length_breaks_the_equality():
    """The equality is sensitive to the step lengths: Toivanen's side run
    with l_eucl one ulp larger is caught."""
    rng = np.random.default_rng(3)
    image, mask = seeded(rng, (21, 19), 1, 2)
    a, _ = converged(lambda it: oracle.geodesic2d(image, mask, 1e10, np.float32(0), np.float32(1), it))
    le = np.nextafter(np.float32(1), np.float32(2))
    b, _ = converged(lambda it: oracle.toivanen2d(image, mask, 1e10, np.float32(0), le, it))
    assert not bits_equal(a, b)


def chamfer_ratio_bound(nd):
    """Largest ratio of the unit-spacing chamfer metric (steps 1, sqrt2[, sqrt3])
    to the Euclidean norm, over directions (dense sampling of the sorted
    sector; the maximum is interior, so the sampled maximum is within 1e-6)."""
    t = np.linspace(0.0, 1.0, 2001)
    if nd == 2:
        a = t  # y / x with x >= y >= 0
        return float(np.max((1 - a + np.sqrt(2) * a) / np.sqrt(1 + a * a)))
    a, b = np.meshgrid(t, t, indexing="ij")
    ok = b <= a
    r = (1 - a + np.sqrt(2) * (a - b) + np.sqrt(3) * b) / np.sqrt(1 + a * a + b * b)
    return float(np.max(r[ok]))


def test_chamfer_bounds_are_the_known_constants():
    assert abs(chamfer_ratio_bound(2) - np.cos(np.pi / 8) - (np.sqrt(2) - 1) * np.sin(np.pi / 8)) < 1e-6
    assert 1.128 < chamfer_ratio_bound(3) < 1.1282


@pytest.mark.parametrize("nd", [2, 3])
def test_euclidean_mode_lies_within_the_chamfer_bounds(nd):
    rng = np.random.default_rng(40 + nd)
    shape = (61, 47) if nd == 2 else (17, 21, 19)
    image, mask = seeded(rng, shape, 1, 4)
    lg, le = oracle.wrapper_weights(0.0)
    if nd == 2:
        d = oracle.geodesic2d(image, mask, 1e10, lg, le, 4)
    else:
        d = oracle.geodesic3d(image, mask, [1.0, 1.0, 1.0], 1e10, lg, le, 4)
    edt = ndimage.distance_transform_edt(mask[0, 0])
    sel = edt > 0
    r = d[0, 0][sel].astype(np.float64) / edt[sel]
    assert r.min() >= 1 - 1e-6
    assert r.max() <= chamfer_ratio_bound(nd) + 1e-6
