"""Calls from several host threads on several streams at once.

A tiled sweep launch needs all of its workgroups resident together.  The
library orders the calls that launch tiled sweeps per device, so calls made
concurrently on different streams must each return exactly what the same
call returns alone.
"""
import threading

import numpy as np
import pytest
import torch

import fastgeodis_amd as FG

pytestmark = pytest.mark.gpu


def bits_equal(a, b):
    return torch.equal(a.contiguous().view(torch.int32), b.contiguous().view(torch.int32))


def make_inputs(rng, shape, C):
    img = torch.from_numpy(rng.random((1, C, *shape), dtype=np.float32)).cuda()
    m = np.ones((1, 1, *shape), np.float32)
    for _ in range(4):
        m[(0, 0) + tuple(int(rng.integers(0, s)) for s in shape)] = 0
    return img, torch.from_numpy(m).cuda()


def test_concurrent_streams_match_serial(device):
    rng = np.random.default_rng(11)
    calls = []
    # 3D planes of 160 x 200 and 256 x 256 use many tiles per launch; the 2D
    # image and the signed call add other launch mixes
    for shape in [(48, 160, 200), (40, 256, 256), (64, 96, 300)]:
        img, mask = make_inputs(rng, shape, 1)
        calls.append(lambda i=img, m=mask: FG.generalised_geodesic3d(i, m, [1.5, 1.0, 1.0], 1e10, 1.0, 2))
    img, mask = make_inputs(rng, (600, 700), 3)
    calls.append(lambda i=img, m=mask: FG.generalised_geodesic2d(i, m, 1e10, 0.5, 2))
    img, mask = make_inputs(rng, (32, 200, 180), 1)
    calls.append(lambda i=img, m=(mask > 0.5).float(): FG.signed_generalised_geodesic3d(
        i, m, [1.0, 1.0, 1.0], 1e10, 1.0, 2))
    torch.cuda.synchronize()
    expected = [f() for f in calls]
    torch.cuda.synchronize()

    results = [None] * len(calls)
    errors = []

    def worker(k):
        try:
            torch.cuda.set_device(0)
            s = torch.cuda.Stream()
            with torch.cuda.stream(s):
                outs = [calls[k]() for _ in range(3)]
            s.synchronize()
            results[k] = outs
        except BaseException as exc:  # reported by the main thread
            errors.append((k, exc))

    threads = [threading.Thread(target=worker, args=(k,)) for k in range(len(calls))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=90)
    assert not any(t.is_alive() for t in threads), "a worker did not finish"
    assert not errors, errors
    for k, outs in enumerate(results):
        for o in outs:
            assert bits_equal(o, expected[k]), f"call {k} differs when run concurrently"
