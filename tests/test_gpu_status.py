"""The device status word on the GPU: an error left by one call is raised by
the next call (or by check_status), once.

The expiry itself is forced from the host with a wait limit of one check, so
every kernel still exits normally: a tile that finds a neighbour's halo not
yet published at its first check stops waiting, records the expiry and
finishes its sweep with stale halo values (all accesses stay in bounds).
"""
import numpy as np
import pytest
import torch

import fastgeodis_amd as FG

pytestmark = pytest.mark.gpu


def small_case():
    x = torch.rand(1, 1, 40, 48, device="cuda")
    m = torch.ones_like(x)
    m[0, 0, 3, 5] = 0
    return x, m


def test_injected_status_raises_on_the_next_call(device):
    x, m = small_case()
    want = FG.generalised_geodesic2d(x, m, 1e10, 1.0, 2)
    FG.check_status()
    dev = torch.cuda.current_device()
    FG.FastGeodisCpp.inject_status(dev, -3)
    with pytest.raises(RuntimeError, match="halo wait expired"):
        FG.generalised_geodesic2d(x, m, 1e10, 1.0, 2)
    # reported once: the next call runs and is correct
    again = FG.generalised_geodesic2d(x, m, 1e10, 1.0, 2)
    assert torch.equal(want.view(torch.int32), again.view(torch.int32))
    FG.FastGeodisCpp.inject_status(dev, -3)
    with pytest.raises(RuntimeError, match="halo wait expired"):
        FG.check_status()
    FG.check_status()


def test_expired_wait_is_reported(device, monkeypatch):
    # no per-call synchronising check here: the expiry must surface by itself
    monkeypatch.setenv("FASTGEODIS_SYNC_CHECK", "0")
    rng = np.random.default_rng(1)
    img = torch.from_numpy(rng.random((1, 1, 160, 200, 200), dtype=np.float32)).cuda()
    mask = torch.ones_like(img)
    mask[0, 0, 7, 8, 9] = 0
    FG.check_status()
    print("forced expiry: 160x200x200, wait limit 1", flush=True)
    FG.FastGeodisCpp.set_wait_limit(1)
    try:
        raised = False
        for _ in range(3):
            FG.generalised_geodesic3d(img, mask, [1.0, 1.0, 1.0], 1e10, 1.0, 2)
            try:
                FG.check_status()
            except RuntimeError as e:
                assert "halo wait expired" in str(e)
                raised = True
                break
    finally:
        FG.FastGeodisCpp.set_wait_limit(0)
    assert raised, "no tile found a halo value missing at its first check in three calls"
    FG.check_status()
    good = FG.generalised_geodesic3d(img, mask, [1.0, 1.0, 1.0], 1e10, 1.0, 2)
    again = FG.generalised_geodesic3d(img, mask, [1.0, 1.0, 1.0], 1e10, 1.0, 2)
    FG.check_status()
    assert torch.equal(good.view(torch.int32), again.view(torch.int32))
