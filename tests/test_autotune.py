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
    assert sig == "3d:8x16x32:c3"
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
