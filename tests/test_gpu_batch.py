"""Batched form: many uneven 2D/3D instances in the same launches.

Each instance's output must equal the single-instance call with the same
arguments bit for bit (and therefore the oracle), whatever else shares the
launches: other sizes, channel counts, operations, iteration counts.
"""
import numpy as np
import pytest
import torch

import fastgeodis_amd as FG
import oracle

pytestmark = pytest.mark.gpu


def bits_equal(a, b):
    return torch.equal(a.contiguous().view(torch.int32), b.contiguous().view(torch.int32))


def rand_instance(rng, nd, C, shape, op, lamb, it, spacing=None):
    img = torch.from_numpy(rng.random((1, C, *shape), dtype=np.float32)).cuda()
    if op == "unsigned":
        m = np.ones((1, 1, *shape), np.float32)
        for _ in range(int(rng.integers(1, 6))):
            m[(0, 0) + tuple(int(rng.integers(0, s)) for s in shape)] = 0
    else:
        m = (rng.random((1, 1, *shape)) > 0.6).astype(np.float32)
    mask = torch.from_numpy(m).cuda()
    return FG.Instance(img, mask, 1e10, lamb, iter=it, spacing=spacing, op=op, theta=0.1)


def single(inst):
    i, m = inst.image, inst.softmask
    if inst.image.dim() == 4:
        if inst.op == "unsigned":
            return FG.generalised_geodesic2d(i, m, inst.v, inst.lamb, inst.iter)
        if inst.op == "signed":
            return FG.signed_generalised_geodesic2d(i, m, inst.v, inst.lamb, inst.iter)
        return FG.GSF2d(i, m, inst.theta, inst.v, inst.lamb, inst.iter)
    if inst.op == "unsigned":
        return FG.generalised_geodesic3d(i, m, inst.spacing, inst.v, inst.lamb, inst.iter)
    if inst.op == "signed":
        return FG.signed_generalised_geodesic3d(i, m, inst.spacing, inst.v, inst.lamb, inst.iter)
    return FG.GSF3d(i, m, inst.theta, inst.spacing, inst.v, inst.lamb, inst.iter)


def test_mixed_batch_equals_single_calls(device):
    rng = np.random.default_rng(42)
    insts = []
    for k in range(10):
        shape = (int(rng.integers(20, 200)), int(rng.integers(20, 200)))
        insts.append(rand_instance(rng, 2, int(rng.choice([1, 3])), shape,
                                   ["unsigned", "signed", "gsf"][k % 3], float(rng.choice([0.0, 0.5, 1.0])),
                                   int(rng.integers(1, 4))))
    for k in range(6):
        shape = tuple(int(x) for x in rng.integers(8, 90, size=3))
        sp = [float(x) for x in rng.uniform(0.5, 3.0, size=3)]
        insts.append(rand_instance(rng, 3, int(rng.choice([1, 2])), shape,
                                   ["unsigned", "signed", "gsf"][k % 3], float(rng.choice([0.3, 1.0])),
                                   int(rng.integers(1, 4)), sp))
    print("batch of", len(insts), flush=True)
    outs = FG.geodesic_batch(insts)
    torch.cuda.synchronize()
    for k, (inst, out) in enumerate(zip(insts, outs)):
        alone = single(inst)
        assert bits_equal(out, alone), f"instance {k} ({inst.op}, {tuple(inst.image.shape)}) differs"


def test_batch_matches_oracle_small(device):
    rng = np.random.default_rng(7)
    insts, want = [], []
    for k in range(6):
        if k % 2 == 0:
            shape = (int(rng.integers(10, 80)), int(rng.integers(10, 80)))
            inst = rand_instance(rng, 2, 1 + k % 3, shape, "unsigned", 0.5, 2)
            want.append(oracle.py_geodesic2d(inst.image.cpu().numpy(), inst.softmask.cpu().numpy(), 1e10, 0.5, 2))
        else:
            shape = tuple(int(x) for x in rng.integers(5, 40, size=3))
            inst = rand_instance(rng, 3, 1, shape, "signed", 1.0, 2, [1.5, 1.0, 0.7])
            want.append(oracle.py_signed_geodesic3d(inst.image.cpu().numpy(), inst.softmask.cpu().numpy(),
                                                    [1.5, 1.0, 0.7], 1e10, 1.0, 2))
        insts.append(inst)
    outs = FG.geodesic_batch(insts)
    for out, w in zip(outs, want):
        o = out.cpu().numpy()
        assert np.array_equal(o.view(np.uint32), w.view(np.uint32))


def test_batch_tuning_invariance(device):
    """The same batch under different halo depths and workgroup shapes."""
    rng = np.random.default_rng(3)
    insts = [rand_instance(rng, 3, 1, (40, 130, 90), "unsigned", 1.0, 2, [1.0, 1.0, 1.0]),
             rand_instance(rng, 2, 3, (300, 257), "unsigned", 0.5, 2),
             rand_instance(rng, 3, 3, (33, 70, 71), "signed", 0.5, 2, [2.0, 1.0, 1.0])]
    old = FG.get_tuning()
    try:
        FG.set_tuning(1, 4, 0, 0)
        ref = FG.geodesic_batch(insts)
        for e, k, s in [(1, 2, 1), (1, 6, 3), (0, 4, 0), (1, 3, 2), (1, 5, 4), (1, 0, 4), (1, 0, 5), (1, 0, 0)]:
            FG.set_tuning(e, k, 0, s)
            outs = FG.geodesic_batch(insts)
            for a, b in zip(outs, ref):
                assert bits_equal(a, b)
    finally:
        FG.set_tuning(**old)


def test_empty_and_validation(device):
    assert FG.geodesic_batch([]) == []
    x = torch.rand(1, 1, 8, 8, device="cuda")
    bad = FG.Instance(x, torch.rand(1, 1, 8, 9, device="cuda"), 1e10, 1.0)
    with pytest.raises(ValueError):
        FG.geodesic_batch([bad])


def test_empty_instance_in_batch_and_zero_channel_image(device):
    rng = np.random.default_rng(11)
    full = rand_instance(rng, 2, 1, (40, 70), "unsigned", 1.0, 2)
    empty = FG.Instance(torch.rand(1, 1, 0, 9, device="cuda"), torch.ones(1, 1, 0, 9, device="cuda"), 1e10, 1.0)
    outs = FG.geodesic_batch([empty, full])
    assert outs[0].shape == (1, 1, 0, 9)
    assert bits_equal(outs[1], single(full))
    # an image with no channels cannot define a gradient term: refused, never garbage
    with pytest.raises(ValueError):
        FG.generalised_geodesic2d(torch.rand(1, 0, 8, 8, device="cuda"), torch.ones(1, 1, 8, 8, device="cuda"),
                                  1e10, 1.0, 2)


@pytest.mark.parametrize("seed", [21, 22])
def test_shared_halo_planning_mixed_rings(device, seed):
    """3D planes of 1-4 channels (two slab-ring lengths) and 2D rows share the
    passes of one batch; enough of them that their tiles exceed the device, so
    the shared halo depth differs from each plane's own: every result must
    still equal its single call."""
    rng = np.random.default_rng(seed)
    insts = []
    for k in range(14):
        shape = tuple(int(x) for x in rng.integers(24, 120, size=3))
        sp = [float(x) for x in rng.uniform(0.5, 3.0, size=3)]
        insts.append(rand_instance(rng, 3, 1 + k % 4, shape, ["unsigned", "signed", "gsf"][k % 3],
                                   float(rng.choice([0.0, 0.5, 1.0])), int(rng.integers(1, 3)), sp))
    for k in range(12):
        shape = (int(rng.integers(30, 700)), int(rng.integers(30, 900)))
        insts.append(rand_instance(rng, 2, 1 + k % 3, shape, ["unsigned", "signed"][k % 2], 1.0, 2))
    print("shared-halo batch of", len(insts), flush=True)
    outs = FG.geodesic_batch(insts)
    torch.cuda.synchronize()
    for k, (inst, out) in enumerate(zip(insts, outs)):
        assert bits_equal(out, single(inst)), f"instance {k} ({inst.op}, {tuple(inst.image.shape)}) differs"


def test_transposed_image_given_equals_computed(device):
    """Instances that bring their image already transposed (as the serving
    loop's sessions do) give the same bits as instances that do not."""
    rng = np.random.default_rng(21)
    insts, with_t = [], []
    for k in range(4):
        if k % 2:
            shape = (int(rng.integers(10, 60)), int(rng.integers(10, 90)))
            inst = rand_instance(rng, 2, 1 + k % 2, shape, ["unsigned", "signed"][k % 2], 0.5, 2)
        else:
            shape = tuple(int(x) for x in rng.integers(6, 40, size=3))
            inst = rand_instance(rng, 3, 1 + k % 2, shape, ["signed", "gsf"][k // 2 % 2], 1.0, 2,
                                 [1.5, 1.0, 0.7])
        insts.append(inst)
        with_t.append(FG.Instance(inst.image, inst.softmask, inst.v, inst.lamb, iter=inst.iter,
                                  spacing=inst.spacing, op=inst.op, theta=inst.theta,
                                  image_t=FG.transposed_image(inst.image)))
    a = FG.geodesic_batch(insts)
    b = FG.geodesic_batch(with_t)
    for k, (x, y) in enumerate(zip(a, b)):
        assert bits_equal(x, y), f"instance {k}"
    bad = FG.Instance(insts[0].image, insts[0].softmask, 1e10, 1.0, spacing=insts[0].spacing,
                      image_t=insts[0].image.clone())  # wrong layout
    if insts[0].image.shape[2] != insts[0].image.shape[4]:
        with pytest.raises(ValueError):
            FG.geodesic_batch([bad])


def test_set_voxels_writes_single_values(device):
    a = torch.zeros(3, 40, 50, device="cuda")
    b = torch.full((1, 1, 7, 9, 11), 0.5, device="cuda")
    idx = [0, 5, 119, 3 * 40 * 50 - 1]
    FG.FastGeodisCpp.set_voxels([a] * 4 + [b] * 2, idx + [0, 7 * 9 * 11 - 1], [1.0, 2.0, 3.0, 4.0, 0.0, 1.0])
    want = torch.zeros(3, 40, 50)
    want.view(-1)[idx] = torch.tensor([1.0, 2.0, 3.0, 4.0])
    assert torch.equal(a.cpu(), want)
    wb = torch.full((1, 1, 7, 9, 11), 0.5)
    wb.view(-1)[0], wb.view(-1)[-1] = 0.0, 1.0
    assert torch.equal(b.cpu(), wb)
    with pytest.raises(ValueError):
        FG.FastGeodisCpp.set_voxels([a], [3 * 40 * 50], [1.0])
    # more than one launch's worth of items
    c = torch.zeros(1000, device="cuda")
    FG.FastGeodisCpp.set_voxels([c] * 300, list(range(0, 600, 2)), [float(i) for i in range(300)])
    wc = torch.zeros(1000)
    wc[0:600:2] = torch.arange(300, dtype=torch.float32)
    assert torch.equal(c.cpu(), wc)
