"""Parity at the full sizes of the benchmark configurations that run as
batches or served requests.

* plain 3D, unsigned: 1x1x128x256x256, spacing [2.5, 1, 1], 16 seeds,
  lamb in {0, 0.5, 1}, iter 4, bit-exact against the oracle;
* the uneven batch: 256 2D instances (64-2048 px per side, C in {1, 3}) and 32
  volumes (64-256 per side, anisotropic spacing) in one batched call; every
  instance equals its own plain call bit for bit, and a seeded sample that
  always includes the largest 2D C=3 member and the largest volume equals the
  oracle;
* the serving loop: 64 resident sessions (96-320 per side, spacing [2, 1, 1]),
  240 clicks submitted at once; every result equals the same request served
  alone, and a seeded sample equals the oracle's signed transform of the mask
  the request saw.
"""
import sys

import numpy as np
import pytest
import torch

import fastgeodis_amd as FG
import oracle
from fastgeodis_amd import serving, synthetic

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def bits_equal_np(a, b):
    a = np.ascontiguousarray(a, dtype=np.float32).view(np.int32)
    b = np.ascontiguousarray(b, dtype=np.float32).view(np.int32)
    return a.shape == b.shape and np.array_equal(a, b)


def bits_equal(a, b):
    return a.shape == b.shape and torch.equal(a.contiguous().view(torch.int32), b.contiguous().view(torch.int32))


@pytest.mark.parametrize("lamb", [0.0, 0.5, 1.0])
def test_plain_3d_unsigned_full_size(device, lamb):
    img, mask, _ = synthetic.config2("cuda")
    say(f"[configs] 128x256x256 unsigned lamb={lamb} iter 4")
    out = FG.generalised_geodesic3d(img, mask, [2.5, 1.0, 1.0], 1e10, lamb, 4).cpu().numpy()
    want = oracle.py_geodesic3d(img.cpu().numpy(), mask.cpu().numpy(), [2.5, 1.0, 1.0], 1e10, lamb, 4)
    assert bits_equal_np(out, want)


def test_uneven_batch_full_size(device):
    say("[configs] building 256 uneven 2D + 32 uneven 3D instances")
    items = synthetic.config4("cuda")
    insts = [FG.Instance(i, m, 1e10, 1.0, iter=2, spacing=sp) for i, m, sp in items]
    say("[configs] batched call")
    outs = FG.geodesic_batch(insts)
    torch.cuda.synchronize()
    say("[configs] per-instance plain calls")
    for k, ((i, m, sp), out) in enumerate(zip(items, outs)):
        alone = (FG.generalised_geodesic2d(i, m, 1e10, 1.0, 2) if sp is None
                 else FG.generalised_geodesic3d(i, m, sp, 1e10, 1.0, 2))
        assert bits_equal(out, alone), f"instance {k} {tuple(i.shape)} differs from its plain call"
    # oracle sample: the largest 2D C=3 member, the largest volume, and 8 more
    two = [k for k, (i, _, sp) in enumerate(items) if sp is None]
    three = [k for k, (_, _, sp) in enumerate(items) if sp is not None]
    big2 = max((k for k in two if items[k][0].shape[1] == 3), key=lambda k: items[k][0].numel())
    big3 = max(three, key=lambda k: items[k][0].numel())
    rng = np.random.default_rng(44)
    rest = [int(k) for k in rng.choice([k for k in range(len(items)) if k not in (big2, big3)], 8, replace=False)]
    for k in [big2, big3] + rest:
        i, m, sp = items[k]
        say(f"[configs] oracle check of instance {k} {tuple(i.shape)}")
        ii, mm = i.cpu().numpy(), m.cpu().numpy()
        want = (oracle.py_geodesic2d(ii, mm, 1e10, 1.0, 2) if sp is None
                else oracle.py_geodesic3d(ii, mm, sp, 1e10, 1.0, 2))
        assert bits_equal_np(outs[k].cpu().numpy(), want), f"instance {k} differs from the oracle"


def test_serving_full_size(device):
    srv = serving.Server(budget_bytes=8 << 30)
    say("[configs] 64 sessions, sides 96-320")
    serving.make_sessions(srv, 64, 96, 320, seed=5, spacing=(2.0, 1.0, 1.0))
    torch.cuda.synchronize()
    rng = np.random.default_rng(55)
    state = {sid: s.mask.clone() for sid, s in srv.sessions.items()}
    clicks, before = [], []
    for k in range(240):
        sid = int(rng.integers(0, 64))
        shape = srv.sessions[sid].mask.shape[2:]
        vox = [tuple(int(rng.integers(0, s)) for s in shape) for _ in range(int(rng.integers(1, 6)))]
        c = serving.Click(sid, vox, bool(rng.integers(0, 2)))
        before.append(state[sid].clone())
        for (z, y, x) in vox:
            state[sid][0, 0, z, y, x] = 1.0 if c.foreground else 0.0
        clicks.append(c)
    srv.start()
    try:
        say("[configs] 240 clicks submitted at once")
        futs = [srv.submit(c) for c in clicks]
        res = [f.result(timeout=300) for f in futs]
    finally:
        srv.stop()
    assert max(r.batch_size for r in res) > 1, "requests were never batched together"
    say("[configs] each result against the request served alone")
    for k, (c, b, r) in enumerate(zip(clicks, before, res)):
        alone = serving.run_alone(srv, c, b)
        assert bits_equal(r.distance, alone), f"click {k} on session {c.session} differs from alone"
    for sid, s in srv.sessions.items():
        assert bits_equal(s.mask, state[sid])
    for k in [int(x) for x in rng.choice(len(clicks), 4, replace=False)]:
        c = clicks[k]
        s = srv.sessions[c.session]
        m = before[k].clone()
        for (z, y, x) in c.voxels:
            m[0, 0, z, y, x] = 1.0 if c.foreground else 0.0
        say(f"[configs] oracle check of click {k}, session {c.session} {tuple(s.image.shape[2:])}")
        want = oracle.py_signed_geodesic3d(s.image.cpu().numpy(), m.cpu().numpy(), list(s.spacing),
                                           srv.v, srv.lamb, srv.iterations)
        assert bits_equal_np(res[k].distance.cpu().numpy(), want), f"click {k} differs from the oracle"
