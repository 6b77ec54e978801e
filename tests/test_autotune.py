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
