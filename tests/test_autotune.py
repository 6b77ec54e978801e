"""Empirical tuner: cache behaviour on the CPU, measured choice on the GPU."""
import json

import pytest
import torch

import fastgeodis_amd as FG
from fastgeodis_amd import autotune


def test_signature_and_cache_file_roundtrip(tmp_path, monkeypatch):
    path = tmp_path / "tune.json"
    monkeypatch.setenv("FASTGEODIS_AMD_TUNE_CACHE", str(path))
    sig = autotune.signature((8, 16, 32), 3)
    assert sig.startswith("3d:8x16x32:c3:l1:it1:")
    # the kernel-variant class and the iteration count are part of the key
    assert autotune.signature((8, 16, 32), 3, lamb=0.5) != sig
    assert autotune.signature((8, 16, 32), 3, lamb=1.5) != autotune.signature((8, 16, 32), 3, lamb=0.5)
    assert autotune.signature((8, 16, 32), 3, iters=2) != sig
    entry = {"signature": sig, "best": {"engine": 1, "halo": 5, "shape": 4}, "ms": 1.0, "table": []}
    path.write_text(json.dumps({sig: entry}))
    autotune._cache.clear()
    assert autotune.load_cache()[sig]["best"]["halo"] == 5
    # a cached signature is returned without measuring (no GPU needed)
    assert autotune.tune((8, 16, 32), channels=3) == entry
    old = FG.get_tuning()
    try:
        autotune.apply(entry)
        assert FG.get_tuning() == {"engine": 1, "halo": 5, "max_groups": 0, "shape": 4}
    finally:
        FG.set_tuning(**old)
        autotune._cache.clear()


def test_rejects_bad_shape():
    with pytest.raises(ValueError):
        autotune.tune((4,), use_cache=False)


@pytest.mark.gpu
def test_tune_measures_and_keeps_results_identical(device, monkeypatch):
    monkeypatch.delenv("FASTGEODIS_AMD_TUNE_CACHE", raising=False)
    autotune._cache.clear()
    old = FG.get_tuning()
    res = autotune.tune((40, 96, 130), channels=1, spacing=(2.0, 1.0, 1.0), reps=1, use_cache=False)
    assert FG.get_tuning() == old  # tuning restored after measuring
    assert len(res["table"]) == len(autotune.DEFAULT_CANDIDATES_3D)
    assert res["ms"] == min(r["ms"] for r in res["table"])
    res2 = autotune.tune((300, 257), channels=3, reps=1, use_cache=False)
    assert res2["best"]["shape"] == 0


@pytest.mark.gpu
def test_tuned_choice_matches_the_oracle(device, monkeypatch):
    """A tuned configuration's output equals the oracle's (not only the other
    candidates')."""
    import numpy as np

    import oracle

    monkeypatch.delenv("FASTGEODIS_AMD_TUNE_CACHE", raising=False)
    autotune._cache.clear()
    shape = (24, 70, 90)
    res = autotune.tune(shape, channels=1, spacing=(1.5, 1.0, 0.8), lamb=0.5, reps=1, use_cache=False)
    rng = np.random.default_rng(11)
    img = rng.random((1, 1, *shape), dtype=np.float32)
    mask = np.ones((1, 1, *shape), np.float32)
    mask[0, 0, 3, 4, 5] = 0
    old = FG.get_tuning()
    try:
        autotune.apply(res)
        out = FG.generalised_geodesic3d(torch.from_numpy(img).cuda(), torch.from_numpy(mask).cuda(),
                                        [1.5, 1.0, 0.8], 1e10, 0.5, 2).cpu().numpy()
    finally:
        FG.set_tuning(**old)
    want = oracle.py_geodesic3d(img, mask, [1.5, 1.0, 0.8], 1e10, 0.5, 2)
    assert np.array_equal(out.view(np.uint32), want.view(np.uint32))


def test_batch_signature_is_a_size_class():
    """Extents rounded to powers of two, order-independent, op and variant aware."""
    a = FG.Instance(torch.zeros(1, 1, 100, 200), torch.ones(1, 1, 100, 200), 1e10, 1.0)
    b = FG.Instance(torch.zeros(1, 1, 96, 210), torch.ones(1, 1, 96, 210), 1e10, 1.0)
    c = FG.Instance(torch.zeros(1, 3, 30, 40, 50), torch.ones(1, 1, 30, 40, 50), 1e10, 0.5, op="signed")
    assert autotune.batch_signature([a, c]) == autotune.batch_signature([c, b])
    assert autotune.batch_signature([a, c]) != autotune.batch_signature([a, a])
    d = FG.Instance(torch.zeros(1, 3, 30, 40, 50), torch.ones(1, 1, 30, 40, 50), 1e10, 1.0, op="signed")
    assert autotune.batch_signature([a, c]) != autotune.batch_signature([a, d])


@pytest.mark.gpu
def test_tuned_batch_equals_analytic_and_uses_the_cache(device, monkeypatch, tmp_path):
    import numpy as np

    monkeypatch.setenv("FASTGEODIS_AMD_TUNE_CACHE", str(tmp_path / "t.json"))
    autotune._cache.clear()
    rng = np.random.default_rng(5)
    insts = []
    for k in range(5):
        shape = (int(rng.integers(20, 90)), int(rng.integers(20, 130))) if k % 2 else \
            tuple(int(x) for x in rng.integers(10, 50, size=3))
        img = torch.from_numpy(rng.random((1, 1, *shape), dtype=np.float32)).cuda()
        m = torch.ones((1, 1, *shape), device="cuda")
        m[(0, 0) + tuple(int(rng.integers(0, s)) for s in shape)] = 0
        insts.append(FG.Instance(img, m, 1e10, 1.0, iter=2, spacing=[1.5, 1.0, 1.0]))
    old = FG.get_tuning()
    res = autotune.tune_batch(insts, reps=1)
    assert FG.get_tuning() == old
    assert len(res["table"]) == len(autotune.DEFAULT_BATCH_CANDIDATES)
    assert (tmp_path / "t.json").exists()
    a = FG.geodesic_batch(insts)
    b = FG.geodesic_batch(insts, tuned=True)
    assert FG.get_tuning() == old
    for x, y in zip(a, b):
        assert torch.equal(x.view(torch.int32), y.view(torch.int32))
    autotune._cache.clear()
    assert autotune.tuning_for_batch(insts, measure=False) == res["best"]  # reloaded from the file
    autotune._cache.clear()
