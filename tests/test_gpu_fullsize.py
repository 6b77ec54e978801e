"""GPU parity at the benchmarked sizes.

The headline input (1x1x512^3, iter 4) and the plain-3D, RGB-batch and 2D
configurations run through the reference-compatible API and are compared bit
for bit with the oracle on the host (its OpenMP build finishes the 512^3 call
in about ten seconds), and with the independent per-slice engine and a second
tiling of the tiled engine (different tile seams, exchange depths and
workgroup shapes must not change a bit).
"""
import sys

import numpy as np
import pytest
import torch

import fastgeodis_amd as FG
import oracle
from fastgeodis_amd import synthetic

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits_equal(a, b):
    a = np.ascontiguousarray(a, dtype=np.float32).view(np.int32)
    b = np.ascontiguousarray(b, dtype=np.float32).view(np.int32)
    return a.shape == b.shape and np.array_equal(a, b)


def with_tuning(engine, halo, shape, fn):
    old = FG.get_tuning()
    try:
        FG.set_tuning(engine, halo, 0, shape)
        return fn()
    finally:
        FG.set_tuning(**old)


@pytest.fixture(scope="module")
def headline(device):
    say("[fullsize] building the 512^3 headline input")
    img, mask = synthetic.headline_512("cuda")
    torch.cuda.synchronize()
    return img, mask


def test_headline_512_matches_oracle(headline):
    img, mask = headline
    say("[fullsize] 512^3 geodesic3d iter 4 on the GPU")
    out = FG.generalised_geodesic3d(img, mask, [1.0, 1.0, 1.0], 1e10, 1.0, 4).cpu().numpy()
    say(f"[fullsize] oracle on {oracle.num_threads()} host threads")
    ref = oracle.py_geodesic3d(img.cpu().numpy(), mask.cpu().numpy(), [1.0, 1.0, 1.0], 1e10, 1.0, 4)
    assert bits_equal(out, ref)
    # seeds stay at zero, everything else is positive and finite
    seeds = mask.cpu().numpy() == 0
    assert np.all(out[seeds] == 0) and np.all(out[~seeds] > 0) and np.all(np.isfinite(out))


def test_headline_512_engines_and_tilings_agree(headline):
    img, mask = headline
    call = lambda: FG.generalised_geodesic3d(img, mask, [1.0, 1.0, 1.0], 1e10, 1.0, 2)  # noqa: E731
    say("[fullsize] 512^3 iter 2: default tiling")
    base = call().cpu().numpy()
    for engine, halo, shape, name in [(1, 10, 0, "tiled K=10"), (1, 0, 2, "tiled 8x4/3"), (1, 0, 6, "tiled 8x4 exchange waves"),
                                      (0, 0, 0, "per-slice")]:
        say(f"[fullsize] 512^3 iter 2: {name}")
        other = with_tuning(engine, halo, shape, call).cpu().numpy()
        assert bits_equal(base, other), name


def test_plain_3d_signed_anisotropic_matches_oracle(device):
    img, _, smask = synthetic.config2("cuda")
    for lamb in (0.0, 0.5, 1.0):
        say(f"[fullsize] 128x256x256 signed lamb {lamb}")
        out = FG.signed_generalised_geodesic3d(img, smask, [2.5, 1.0, 1.0], 1e10, lamb, 4).cpu().numpy()
        ref = oracle.py_signed_geodesic3d(img.cpu().numpy(), smask.cpu().numpy(), [2.5, 1.0, 1.0], 1e10, lamb, 4)
        assert bits_equal(out, ref), lamb


def test_rgb_1024_geodesic_and_gsf_match_oracle(device):
    images, masks = synthetic.config3("cuda", batch=2)
    for i, m in zip(images, masks):
        say("[fullsize] 1x3x1024x1024 geodesic2d lamb 0.5 iter 2 and GSF2d theta 0.1 iter 4")
        out = FG.generalised_geodesic2d(i, m, 1e10, 0.5, 2).cpu().numpy()
        ref = oracle.py_geodesic2d(i.cpu().numpy(), m.cpu().numpy(), 1e10, 0.5, 2)
        assert bits_equal(out, ref)
        g = FG.GSF2d(i, m, 0.1, 1e10, 1.0, 4).cpu().numpy()
        gref = oracle.py_gsf2d(i.cpu().numpy(), m.cpu().numpy(), 0.1, 1e10, 1.0, 4)
        assert bits_equal(g, gref)


def test_2d_512_and_wide_match_oracle(device):
    img, mask = synthetic.config1("cuda")
    say("[fullsize] 512^2 geodesic2d iter 2")
    out = FG.generalised_geodesic2d(img, mask, 1e10, 1.0, 2).cpu().numpy()
    assert bits_equal(out, oracle.py_geodesic2d(img.cpu().numpy(), mask.cpu().numpy(), 1e10, 1.0, 2))
    # a wide image: row tiles beyond one per CU at the deepest halo
    say("[fullsize] 1x1x600x9000 geodesic2d iter 2")
    big = synthetic.blob_field((600, 9000), 8, (20.0, 80.0), 77, "cuda")[None, None]
    bmask = synthetic.seed_mask((600, 9000), 6, 77, "cuda")
    out = FG.generalised_geodesic2d(big, bmask, 1e10, 1.0, 2).cpu().numpy()
    assert bits_equal(out, oracle.py_geodesic2d(big.cpu().numpy(), bmask.cpu().numpy(), 1e10, 1.0, 2))
