"""Interactive-segmentation serving loop.

Volumes stay resident on the device as sessions.  A click adds 1-5 seed
voxels to a session's soft mask (foreground -> 1, background -> 0; the mask
starts at 0.5, which makes ``v * 0.5`` effectively infinite) and asks for
``signed_generalised_geodesic3d(image, mask, spacing, v, lamb, iter)``.

Continuous batching at pass granularity: a scheduler thread keeps a queue
of requests and a stepper (``FastGeodisCpp.Stepper``, the C ABI's
``fg_stepper_*``) that holds the transforms in flight.  Each step advances
every transform in flight by one pass pair, in launches shared by all of
them; before each step the scheduler admits queued requests, in arrival
order, while their estimated device memory fits the HBM budget, at most one
new request per session per step.  So a request that arrives while others
are on the GPU starts at the next step (a few hundred microseconds), and its
sweeps share the launches of the requests already running; a request leaves
with the step that runs its last pass.  Every request's result equals the
same request served alone, bit for bit (a transform computes exactly the
launches of its plain call, only in the company of others).

Everything runs on the server's stream, in order: the click edits admitted
before a step (one launch for all of their voxels) are written into the
resident masks before that step, and a transform reads its mask only in its
first step, so a session's next click may be admitted while its previous
request is still running, and the masks need no per-request copies.  Each
session keeps its image in the transposed layout the transform's W pass
reads, so no click re-transposes it.  At most ``inflight`` steps are
enqueued ahead of the device.

Latency is measured from a request's arrival to the completion of the HIP
event recorded after the step that finished it.
"""
from __future__ import annotations

import collections
import math
import threading
import time
from concurrent.futures import Future
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import FastGeodisCpp, Instance, geodesic_batch, transposed_image

OP_SIGNED = 1


@dataclass
class Session:
    image: torch.Tensor            # (1, C, D, H, W) float32, on device
    mask: torch.Tensor             # (1, 1, D, H, W) float32, on device
    spacing: Tuple[float, float, float]
    image_t: Optional[torch.Tensor] = None  # (1, C, W, H, D): the image in the W pass layout

    @property
    def voxels(self) -> int:
        return int(self.mask.numel())


@dataclass
class Click:
    session: int
    voxels: Sequence[Tuple[int, int, int]]
    foreground: bool
    arrival: float = field(default_factory=time.perf_counter)
    # the submitting thread's current stream (set by Server.submit): the
    # result is handed over ordered before later work on it
    stream: Optional[torch.cuda.Stream] = None


@dataclass
class Result:
    """distance is complete when the future resolves.  It is recorded as in
    use on the stream that was current in the submitting thread (so freeing it
    there is safe), and that stream waits for the step that produced it; on
    any other stream, call ``distance.record_stream(stream)`` before use.
    batch_size: the requests in flight (its own included) when it was
    admitted."""
    distance: torch.Tensor         # (1, 1, D, H, W) signed distance, on device
    latency_s: float
    batch_size: int


class Server:
    """Continuous-batching server for click requests on resident sessions.

    budget_bytes: device memory the in-flight requests may use (outputs and
    workspaces, estimated per request); max_batch: requests in flight at
    once; inflight: steps enqueued ahead of the device; halo_bytes: the
    stepper's halo-exchange buffer (a request that does not fit waits).
    """

    def __init__(self, budget_bytes: int = 8 << 30, max_batch: int = 64, v: float = 1e10,
                 lamb: float = 1.0, iterations: int = 2, device: str = "cuda", inflight: int = 2,
                 halo_bytes: int = 256 << 20):
        self.budget = int(budget_bytes)
        self.max_batch = int(max_batch)
        self.v, self.lamb, self.iterations = float(v), float(lamb), int(iterations)
        dev = torch.device(device)
        if dev.type != "cuda":
            raise ValueError("the serving loop runs on a HIP device")
        if dev.index is None:
            dev = torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        self.inflight_limit = max(1, int(inflight))
        self.sessions: Dict[int, Session] = {}
        self._queue: collections.deque = collections.deque()
        self._cv = threading.Condition()
        self._stop = False
        self._thread: Optional[threading.Thread] = None
        self._error: Optional[BaseException] = None
        self._stream = torch.cuda.Stream(device=self.device)
        self._stepper = FastGeodisCpp.Stepper(int(halo_bytes), 2 * self.max_batch, self.device.index)
        self.batches: List[int] = []   # requests in flight at each step
        self.rejected = 0

    # ---- sessions -------------------------------------------------------
    def add_session(self, sid: int, image: torch.Tensor, spacing=(1.0, 1.0, 1.0)) -> None:
        if image.dim() != 5 or image.shape[0] != 1:
            raise ValueError("session image must be (1, C, D, H, W)")
        image = image.to(self.device, torch.float32).contiguous()
        mask = torch.full((1, 1) + tuple(image.shape[2:]), 0.5, device=self.device)
        # both were written on the caller's stream; the server's stream reads
        # and edits them from now on
        ready = torch.cuda.Event()
        ready.record(torch.cuda.current_stream(self.device))
        self._stream.wait_event(ready)
        image.record_stream(self._stream)
        mask.record_stream(self._stream)
        with torch.cuda.stream(self._stream):
            image_t = transposed_image(image)
        self.sessions[sid] = Session(image, mask, tuple(float(s) for s in spacing), image_t)

    def request_bytes(self, sid: int) -> int:
        """Device bytes one request of this session holds while in flight:
        output + its share of the batched workspace (second distance, two
        transposed distances; the session keeps its transposed image)."""
        s = self.sessions[sid]
        n = s.voxels
        return 4 * n * (1 + 3) + (1 << 20)

    # ---- requests -------------------------------------------------------
    def submit(self, click: Click) -> Future:
        if click.session not in self.sessions:
            raise KeyError(f"unknown session {click.session}")
        if self.request_bytes(click.session) > self.budget:
            raise ValueError("request larger than the HBM budget")
        if click.stream is None:
            click.stream = torch.cuda.current_stream(self.device)
        fut: Future = Future()
        with self._cv:
            if self._error is not None:
                raise RuntimeError("the serving loop has failed") from self._error
            self._queue.append((click, fut))
            self._cv.notify()
        return fut

    def start(self) -> "Server":
        self._stop = False
        self._thread = threading.Thread(target=self._run, name="fastgeodis-serving", daemon=True)
        self._thread.start()
        return self

    def stop(self) -> None:
        with self._cv:
            self._stop = True
            self._cv.notify()
        if self._thread is not None:
            self._thread.join()
            self._thread = None

    # ---- scheduler ------------------------------------------------------
    def _admit(self, used: int, inflight: int) -> List[Tuple[Click, Future]]:
        """Takes queued requests in arrival order while they fit the budget and
        the in-flight limit, at most one per session (this step)."""
        batch, seen, keep = [], set(), collections.deque()
        while self._queue and inflight + len(batch) < self.max_batch:
            click, fut = self._queue.popleft()
            need = self.request_bytes(click.session)
            if click.session in seen or used + need > self.budget:
                keep.append((click, fut))
                if used + need > self.budget:
                    break
                continue
            seen.add(click.session)
            used += need
            batch.append((click, fut))
        # requests left behind keep their order in front of later arrivals
        while keep:
            self._queue.appendleft(keep.pop())
        return batch

    def _add(self, batch, inflight: int):
        """Adds the admitted requests' transforms to the stepper and writes
        their clicks' edits (on the server's stream, before the next step,
        which is the first to read the masks).  Returns the requests taken and
        those the stepper could not take yet (its exchange buffer is full)."""
        for click, _ in batch:
            shape = self.sessions[click.session].mask.shape[2:]
            for vox in click.voxels:
                if not all(0 <= c < n for c, n in zip(vox, shape)):
                    raise ValueError(f"click voxel {tuple(vox)} outside session {click.session}")
        taken, later = [], []
        masks, idx, vals = [], [], []
        with torch.cuda.stream(self._stream):
            for click, fut in batch:
                s = self.sessions[click.session]
                ticket, out = self._stepper.add(s.image, s.mask, list(s.spacing), OP_SIGNED, self.v,
                                                self.lamb, 1.0 - self.lamb, self.iterations, s.image_t)
                if ticket < 0:
                    later.append((click, fut))
                    continue
                taken.append((ticket, click, fut, out))
                _, _, D, H, W = s.mask.shape
                for (z, y, x) in click.voxels:
                    masks.append(s.mask)
                    idx.append((z * H + y) * W + x)
                    vals.append(1.0 if click.foreground else 0.0)
            if masks:
                FastGeodisCpp.set_voxels(masks, idx, vals)
        n = inflight + len(taken)
        for ticket, click, fut, out in taken:
            self._inflight[ticket] = (click, fut, out, n)
        return taken, later

    def _run(self):
        pending = collections.deque()  # (event, tickets finished by that step)
        self._inflight: Dict[int, tuple] = {}
        try:
            self._loop(pending)
        except BaseException as exc:  # fail every outstanding request instead of hanging it
            with self._cv:
                self._error = exc
                waiting = [f for _, f in self._queue]
                self._queue.clear()
            waiting.extend(f for _, f, _, _ in self._inflight.values())
            for f in waiting:
                if not f.done():
                    f.set_exception(exc)

    def _retire(self, pending, block: bool) -> None:
        ev, done = pending[0]
        if not ev.query():
            if not block:
                return
            ev.synchronize()
        pending.popleft()
        t = time.perf_counter()
        for ticket in done:
            click, fut, out, n = self._inflight.pop(ticket)
            consumer = click.stream
            if consumer is not None and consumer != self._stream:
                consumer.wait_event(ev)
                out.record_stream(consumer)
            fut.set_result(Result(out, t - click.arrival, n))

    def _loop(self, pending):
        torch.cuda.set_device(self.device)
        used = 0      # budget bytes of the requests in the stepper
        held = {}     # ticket -> its budget bytes
        phase = 0     # pass (0 D, 1 H, 2 W) of the next step of every request in flight
        while True:
            with self._cv:
                while not self._queue and not self._inflight and not pending and not self._stop:
                    self._cv.wait(timeout=0.05)
                if self._stop and not self._queue and not self._inflight and not pending:
                    return
                active = self._stepper.active()
                if active == 0:
                    phase = 0
                # requests join at iteration boundaries, so that every step
                # runs one pass for all of them in shared launches
                batch = self._admit(used, active) if len(pending) < self.inflight_limit and phase == 0 else []
            if batch:
                try:
                    taken, later = self._add(batch, active)
                except BaseException as exc:
                    for _, f in batch:
                        f.set_exception(exc)
                    raise
                for ticket, click, _, _ in taken:
                    held[ticket] = self.request_bytes(click.session)
                    used += held[ticket]
                if later:
                    with self._cv:
                        for item in reversed(later):
                            self._queue.appendleft(item)
            if self._stepper.active() and len(pending) < self.inflight_limit:
                with torch.cuda.stream(self._stream):
                    done = self._stepper.step()
                    ev = torch.cuda.Event()
                    ev.record(self._stream)
                pending.append((ev, done))
                self.batches.append(self._stepper.active() + len(done))
                phase = (phase + 1) % 3
                for ticket in done:
                    used -= held.pop(ticket)
                continue
            if pending:
                # the device is busy: retire finished steps, or wait a little
                # for new arrivals before blocking on the oldest step
                before = len(pending)
                self._retire(pending, block=False)
                if len(pending) == before:
                    time.sleep(20e-6)
                    if len(pending) >= self.inflight_limit or not self._stepper.active():
                        self._retire(pending, block=True)


def run_alone(server: Server, click: Click, mask_before: torch.Tensor) -> torch.Tensor:
    """The same request served alone (for checks): applies the click to a copy
    of the session's mask as it was before and runs the single-instance call."""
    from . import signed_generalised_geodesic3d

    s = server.sessions[click.session]
    m = mask_before.clone()
    val = 1.0 if click.foreground else 0.0
    for (z, y, x) in click.voxels:
        m[0, 0, z, y, x] = val
    return signed_generalised_geodesic3d(s.image, m, list(s.spacing), server.v, server.lamb, server.iterations)


def poisson_load(server: Server, qps: float, nreq: int, seed: int = 5,
                 sessions: Optional[Sequence[int]] = None):
    """Submits `nreq` clicks with exponential inter-arrival times at `qps` and
    returns (latencies_s, wall_s, batch sizes)."""
    rng = np.random.default_rng(seed)
    sids = list(sessions if sessions is not None else server.sessions.keys())
    futs = []
    t0 = time.perf_counter()
    next_t = t0
    for _ in range(nreq):
        next_t += rng.exponential(1.0 / qps)
        while True:
            now = time.perf_counter()
            if now >= next_t:
                break
            time.sleep(min(next_t - now, 0.0005))
        sid = int(rng.choice(sids))
        shape = server.sessions[sid].mask.shape[2:]
        k = int(rng.integers(1, 6))
        vox = [tuple(int(rng.integers(0, s)) for s in shape) for _ in range(k)]
        futs.append(server.submit(Click(sid, vox, bool(rng.integers(0, 2)))))
    res = [f.result() for f in futs]
    wall = time.perf_counter() - t0
    return [r.latency_s for r in res], wall, [r.batch_size for r in res]


def make_sessions(server: Server, n: int, lo: int = 96, hi: int = 320, seed: int = 5,
                  spacing=(2.0, 1.0, 1.0)) -> None:
    """n sessions with sides uniform in [lo, hi] and blob-field images."""
    from .synthetic import blob_field

    rng = np.random.default_rng(seed)
    for sid in range(n):
        shape = tuple(int(x) for x in rng.integers(lo, hi + 1, size=3))
        img = blob_field(shape, 8, (6.0, 30.0), seed * 1000 + sid, server.device)[None, None]
        server.add_session(sid, img, spacing)


def main(argv=None):
    import argparse
    import json

    ap = argparse.ArgumentParser(description="click-serving load test (Poisson arrivals)")
    ap.add_argument("--sessions", type=int, default=64)
    ap.add_argument("--lo", type=int, default=96)
    ap.add_argument("--hi", type=int, default=320)
    ap.add_argument("--qps", type=float, nargs="+", default=[50, 200, 1000])
    ap.add_argument("--requests", type=int, default=300)
    ap.add_argument("--budget-gb", type=float, default=8.0)
    args = ap.parse_args(argv)
    srv = Server(budget_bytes=int(args.budget_gb * (1 << 30)))
    make_sessions(srv, args.sessions, args.lo, args.hi)
    torch.cuda.synchronize()
    srv.start()
    try:
        poisson_load(srv, 20.0, 10)  # warm-up
        for q in args.qps:
            lat, wall, bs = poisson_load(srv, q, args.requests)
            lat = np.sort(np.asarray(lat))
            print(json.dumps({"qps_offered": q, "requests": len(lat), "throughput_rps": len(lat) / wall,
                              "p50_ms": float(np.percentile(lat, 50) * 1e3),
                              "p99_ms": float(np.percentile(lat, 99) * 1e3),
                              "mean_batch": float(np.mean(bs))}), flush=True)
    finally:
        srv.stop()


if __name__ == "__main__":
    main()
