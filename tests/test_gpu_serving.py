"""Serving loop: continuous batching of click requests on resident sessions.

Every request's result must equal the same request served alone (bit for
bit), whatever shared its steps; requests for one session keep their order;
the HBM budget caps what is in flight at once.
"""
import threading
import time

import numpy as np
import pytest
import torch

import fastgeodis_amd as FG
from fastgeodis_amd import serving

pytestmark = pytest.mark.gpu


def bits_equal(a, b):
    return torch.equal(a.contiguous().view(torch.int32), b.contiguous().view(torch.int32))


def test_each_result_equals_alone(device):
    srv = serving.Server(budget_bytes=2 << 30, max_batch=8)
    serving.make_sessions(srv, 6, lo=24, hi=72, seed=11)
    rng = np.random.default_rng(1)
    clicks, before = [], []
    # all masks as they were before each click (sequential per session)
    state = {sid: s.mask.clone() for sid, s in srv.sessions.items()}
    for k in range(30):
        sid = int(rng.integers(0, 6))
        shape = srv.sessions[sid].mask.shape[2:]
        vox = [tuple(int(rng.integers(0, s)) for s in shape) for _ in range(int(rng.integers(1, 6)))]
        c = serving.Click(sid, vox, bool(k % 2))
        before.append(state[sid].clone())
        for (z, y, x) in vox:
            state[sid][0, 0, z, y, x] = 1.0 if c.foreground else 0.0
        clicks.append(c)
    srv.start()
    try:
        futs = [srv.submit(c) for c in clicks]  # all at once: they get batched
        res = [f.result(timeout=60) for f in futs]
    finally:
        srv.stop()
    assert max(r.batch_size for r in res) > 1, "requests were never batched together"
    for c, b, r in zip(clicks, before, res):
        alone = serving.run_alone(srv, c, b)
        assert bits_equal(r.distance, alone)
    for sid, s in srv.sessions.items():
        assert bits_equal(s.mask, state[sid])


def test_budget_limits_batches(device):
    srv = serving.Server(budget_bytes=1 << 30, max_batch=64)
    serving.make_sessions(srv, 8, lo=40, hi=48, seed=3)
    per = max(srv.request_bytes(s) for s in srv.sessions)
    srv.budget = 3 * per  # room for three requests at a time
    srv.start()
    try:
        futs = [srv.submit(serving.Click(sid, [(1, 1, 1)], True)) for sid in range(8)]
        res = [f.result(timeout=60) for f in futs]
    finally:
        srv.stop()
    assert max(srv.batches) <= 3  # requests in flight at each step
    assert len(res) == 8 and max(r.batch_size for r in res) <= 3
    assert all(r.latency_s > 0 for r in res)


def test_request_larger_than_budget_is_rejected(device):
    srv = serving.Server(budget_bytes=1 << 20)
    serving.make_sessions(srv, 1, lo=64, hi=64, seed=1)
    with pytest.raises(ValueError):
        srv.submit(serving.Click(0, [(0, 0, 0)], True))


def test_poisson_load_reports_latency(device):
    srv = serving.Server(budget_bytes=2 << 30)
    serving.make_sessions(srv, 4, lo=32, hi=64, seed=9)
    srv.start()
    try:
        lat, wall, bs = serving.poisson_load(srv, 200.0, 40)
    finally:
        srv.stop()
    assert len(lat) == 40 and wall > 0
    assert min(lat) > 0
