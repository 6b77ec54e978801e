"""Stepper (fg_stepper_*): transforms that advance one pass pair per step
and join between steps.  Every result must equal the same instance computed
alone by the plain call, bit for bit, whichever transforms shared its steps
and whenever it joined; a full exchange buffer refuses an instance instead
of failing a step."""
import numpy as np
import pytest
import torch

import fastgeodis_amd as FG

pytestmark = pytest.mark.gpu

OP_UNSIGNED, OP_SIGNED = 0, 1


def bits_equal(a, b):
    return torch.equal(a.contiguous().view(torch.int32), b.contiguous().view(torch.int32))


def instance(rng, nd, signed):
    shape = tuple(int(x) for x in rng.integers(20, 70, size=nd))
    C = int(rng.integers(1, 3))
    img = torch.from_numpy(rng.random((1, C, *shape), dtype=np.float32)).cuda()
    if signed:
        m = torch.from_numpy((rng.random((1, 1, *shape)) > 0.6).astype(np.float32)).cuda()
    else:
        m = torch.ones((1, 1, *shape), device="cuda")
        for _ in range(3):
            m[(0, 0) + tuple(int(rng.integers(0, s)) for s in shape)] = 0
    sp = [float(x) for x in rng.uniform(0.5, 2.5, size=3)]
    lamb = float(rng.choice([0.0, 0.5, 1.0]))
    it = int(rng.integers(1, 4))
    return img, m, sp, lamb, it


def alone(img, m, sp, lamb, it, signed):
    if img.dim() == 4:
        fn = FG.signed_generalised_geodesic2d if signed else FG.generalised_geodesic2d
        return fn(img, m, 1e10, lamb, it)
    fn = FG.signed_generalised_geodesic3d if signed else FG.generalised_geodesic3d
    return fn(img, m, sp, 1e10, lamb, it)


@pytest.mark.parametrize("seed", range(3))
def test_instances_joining_mid_flight_equal_plain_calls(device, seed):
    rng = np.random.default_rng(70 + seed)
    st = FG.FastGeodisCpp.Stepper(64 << 20, 64, torch.cuda.current_device())
    specs = [instance(rng, int(rng.choice([2, 3])), bool(k % 2)) for k in range(10)]
    print(f"stepper seed {seed}: {[tuple(s[0].shape) for s in specs]}", flush=True)
    outs, done = {}, set()
    pending = list(range(len(specs)))
    steps = 0
    while pending or st.active():
        # join at different steps: two new instances before every other step
        if pending and steps % 2 == 0:
            for _ in range(min(2, len(pending))):
                k = pending.pop(0)
                img, m, sp, lamb, it = specs[k]
                ticket, out = st.add(img, m, sp, OP_SIGNED if k % 2 else OP_UNSIGNED, 1e10, lamb, 1.0 - lamb, it)
                assert ticket > 0
                outs[ticket] = (k, out)
        done.update(st.step())
        steps += 1
        assert steps < 200
    torch.cuda.synchronize()
    FG.check_status()
    assert done == set(outs)
    for ticket, (k, out) in outs.items():
        img, m, sp, lamb, it = specs[k]
        assert bits_equal(out, alone(img, m, sp, lamb, it, bool(k % 2))), f"instance {k}"


def test_full_exchange_buffer_refuses_instead_of_failing(device):
    rng = np.random.default_rng(5)
    # one 48 x 96 x 96 unsigned transform takes at most 153600 bytes of
    # exchange entries: this buffer holds two of them
    st = FG.FastGeodisCpp.Stepper(400000, 8, torch.cuda.current_device())
    img = torch.from_numpy(rng.random((1, 1, 48, 96, 96), dtype=np.float32)).cuda()
    m = torch.ones((1, 1, 48, 96, 96), device="cuda")
    m[0, 0, 3, 4, 5] = 0
    tickets = [st.add(img, m, [1.0, 1.0, 1.0], OP_UNSIGNED, 1e10, 1.0, 0.0, 1)[0] for _ in range(3)]
    assert tickets[0] > 0 and tickets[1] > 0 and tickets[2] == -1
    while st.active():
        st.step()
    t, out = st.add(img, m, [1.0, 1.0, 1.0], OP_UNSIGNED, 1e10, 1.0, 0.0, 1)  # room again
    assert t > 0
    while st.active():
        st.step()
    torch.cuda.synchronize()
    FG.check_status()
    assert bits_equal(out, FG.generalised_geodesic3d(img, m, [1.0, 1.0, 1.0], 1e10, 1.0, 1))


def test_rejects_gsf_and_zero_iterations(device):
    st = FG.FastGeodisCpp.Stepper(1 << 20, 4, torch.cuda.current_device())
    img = torch.rand(1, 1, 16, 16, 16, device="cuda")
    m = torch.ones(1, 1, 16, 16, 16, device="cuda")
    with pytest.raises(ValueError):
        st.add(img, m, [1.0, 1.0, 1.0], 2, 1e10, 1.0, 0.0, 2)  # GSF
    with pytest.raises(ValueError):
        st.add(img, m, [1.0, 1.0, 1.0], OP_UNSIGNED, 1e10, 1.0, 0.0, 0)
    assert st.active() == 0
