"""Graph replay: a repeated call runs as one launch of a cached graph whose
kernel nodes get each call's arguments.  Every replay must equal the call
issued launch by launch, with other work (direct calls of other shapes, torch
kernels, synchronisations) between the replays, and with the arguments
(buffers, values) changing from one replay to the next.
"""
import numpy as np
import pytest
import torch

import fastgeodis_amd as FG

pytestmark = pytest.mark.gpu


def bits_equal(a, b):
    return torch.equal(a.contiguous().view(torch.int32), b.contiguous().view(torch.int32))


def make(seed, shape, C=1, nseeds=4):
    rng = np.random.default_rng(seed)
    img = torch.from_numpy(rng.random((1, C, *shape), dtype=np.float32)).cuda()
    mask = torch.ones((1, 1, *shape), device="cuda")
    for _ in range(nseeds):
        mask[(0, 0) + tuple(int(rng.integers(0, s)) for s in shape)] = 0
    return img, mask


CALLS = [
    ("2d", lambda i, m, lamb: FG.generalised_geodesic2d(i, m, 1e10, lamb, 2), (200, 300)),
    ("3d", lambda i, m, lamb: FG.generalised_geodesic3d(i, m, [1.5, 1.0, 0.8], 1e10, lamb, 2), (30, 70, 90)),
    ("signed3d", lambda i, m, lamb: FG.signed_generalised_geodesic3d(i, m, [1.0, 1.0, 1.0], 1e10, lamb, 2),
     (20, 64, 80)),
    ("gsf2d", lambda i, m, lamb: FG.GSF2d(i, m, 0.1, 1e10, lamb, 2), (128, 160)),
]


@pytest.mark.parametrize("name,fn,shape", CALLS, ids=[c[0] for c in CALLS])
def test_graph_replays_equal_direct_launches(device, name, fn, shape):
    inputs = [make(100 + k, shape) for k in range(3)]
    lambs = [1.0, 0.5, 1.0]
    FG.set_graphs(False)
    want = [fn(i, m, lamb).clone() for (i, m), lamb in zip(inputs, lambs)]
    FG.set_graphs(True)
    try:
        other_i, other_m = make(7, (37, 53))
        for rep in range(6):
            k = rep % 3
            print(f"[graphs] {name} replay {rep} (input {k})", flush=True)
            got = fn(inputs[k][0], inputs[k][1], lambs[k])
            # other work between replays: an eager call of another shape, torch kernels, a sync
            FG.set_graphs(False)
            FG.generalised_geodesic2d(other_i, other_m, 1e10, 1.0, 1)
            FG.set_graphs(True)
            _ = (inputs[k][0] * 2).sum()
            torch.cuda.synchronize()
            assert bits_equal(got, want[k]), f"{name}: replay {rep} differs from the direct launches"
    finally:
        FG.set_graphs(True)
    FG.check_status()
