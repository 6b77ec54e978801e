"""Seeded synthetic inputs for the benchmark and serving workloads.

Images are sums of Gaussian blobs plus U[0, 0.05] noise, built on the target
device from separable 1-D profiles (a 3-D Gaussian is the outer product of
three 1-D ones), so a 512^3 field costs a few outer products instead of
134 M exp() calls on the host.  Blob parameters come from
``numpy.random.default_rng(seed)`` and are therefore reproducible.
"""
from __future__ import annotations

from typing import Sequence, Tuple

import numpy as np
import torch


def blob_field(shape: Sequence[int], nblobs: int, sigma: Tuple[float, float], seed: int,
               device="cuda", noise: float = 0.05, normalise: bool = False) -> torch.Tensor:
    """float32 tensor of `shape` (spatial only) = sum of `nblobs` Gaussian blobs."""
    rng = np.random.default_rng(seed)
    shape = [int(s) for s in shape]
    out = torch.zeros(shape, dtype=torch.float32, device=device)
    axes = [torch.arange(s, dtype=torch.float32, device=device) for s in shape]
    for _ in range(nblobs):
        c = [rng.uniform(0, s) for s in shape]
        sig = rng.uniform(*sigma)
        amp = rng.uniform(0, 1)
        prof = [torch.exp(-((ax - ci) ** 2) / (2 * sig * sig)) for ax, ci in zip(axes, c)]
        blob = prof[0]
        for p in prof[1:]:
            blob = blob.unsqueeze(-1) * p
        out += amp * blob
    if noise > 0:
        g = torch.Generator(device=device)
        g.manual_seed(seed)
        out += noise * torch.rand(shape, generator=g, device=device)
    if normalise:
        lo, hi = out.min(), out.max()
        out = (out - lo) / torch.clamp(hi - lo, min=1e-12)
    return out


def seed_mask(shape: Sequence[int], nseeds: int, seed: int, device="cuda") -> torch.Tensor:
    """(1, 1, *shape) mask of ones with `nseeds` seeded-random voxels set to 0."""
    rng = np.random.default_rng(seed + 7919)
    m = torch.ones([1, 1] + [int(s) for s in shape], dtype=torch.float32, device=device)
    for _ in range(nseeds):
        idx = tuple(int(rng.integers(0, s)) for s in shape)
        m[(0, 0) + idx] = 0.0
    return m


def headline_512(device="cuda", size: int = 512):
    """512^3 geodesic3d input: 12 blobs (sigma 8-40 voxels) + noise, 16 seeds,
    default_rng(512)."""
    shape = (size, size, size)
    img = blob_field(shape, 12, (8.0, 40.0), 512, device)[None, None]
    return img, seed_mask(shape, 16, 512, device)


def config2(device="cuda"):
    """1x1x128x256x256 volume, spacing [2.5, 1, 1]: 12 blobs + noise, 16 seeds;
    signed mask from a second field thresholded at 0.5 (default_rng(2), (3))."""
    shape = (128, 256, 256)
    img = blob_field(shape, 12, (8.0, 40.0), 2, device)[None, None]
    mask = seed_mask(shape, 16, 2, device)
    f = blob_field(shape, 12, (8.0, 40.0), 3, device, normalise=True)
    smask = (f > 0.5).float()[None, None]
    return img, mask, smask


def config1(device="cpu"):
    """512x512 image: 8 blobs (sigma 20-80 px) + noise, 4 seeds (default_rng(1))."""
    shape = (512, 512)
    img = blob_field(shape, 8, (20.0, 80.0), 1, device)[None, None]
    return img, seed_mask(shape, 4, 1, device)


def config3(device="cuda", batch: int = 32, size: int = 1024):
    """B RGB images (per-channel blob fields normalised to [0, 1]) with disc
    masks (0 inside a disc of radius U[50, 300] px at a uniform centre, 1
    outside) and 2 % of pixels flipped (default_rng(3))."""
    rng = np.random.default_rng(3)
    images, masks = [], []
    yy, xx = torch.meshgrid(torch.arange(size, device=device, dtype=torch.float32),
                            torch.arange(size, device=device, dtype=torch.float32), indexing="ij")
    for b in range(batch):
        chans = [blob_field((size, size), 8, (20.0, 80.0), 3000 + 3 * b + c, device, normalise=True)
                 for c in range(3)]
        images.append(torch.stack(chans)[None])
        r = rng.uniform(50, 300)
        cy, cx = rng.uniform(0, size), rng.uniform(0, size)
        m = (((yy - cy) ** 2 + (xx - cx) ** 2) > r * r).float()
        g = torch.Generator(device=device)
        g.manual_seed(int(rng.integers(0, 2 ** 31)))
        flip = torch.rand((size, size), generator=g, device=device) < 0.02
        m = torch.where(flip, 1.0 - m, m)
        masks.append(m[None, None])
    return images, masks


def config4(device="cuda", n2d: int = 256, n3d: int = 32):
    """Uneven batch (default_rng(4)): n2d 2D instances with H, W log-uniform in
    [64, 2048] and C in {1, 3}, and n3d 3D volumes with sides uniform in
    [64, 256] and per-axis spacing U[0.5, 3.0]; blob-field images, 1-32 seeds.
    Returns a list of (image, mask, spacing or None)."""
    rng = np.random.default_rng(4)
    out = []
    for k in range(n2d):
        H, W = (int(round(np.exp(rng.uniform(np.log(64), np.log(2048))))) for _ in range(2))
        C = int(rng.choice([1, 3]))
        img = torch.stack([blob_field((H, W), 8, (8.0, 80.0), 40000 + 4 * k + c, device) for c in range(C)])
        out.append((img[None], seed_mask((H, W), int(rng.integers(1, 33)), 40000 + k, device), None))
    for k in range(n3d):
        shape = tuple(int(x) for x in rng.integers(64, 257, size=3))
        sp = [float(x) for x in rng.uniform(0.5, 3.0, size=3)]
        img = blob_field(shape, 12, (8.0, 40.0), 50000 + k, device)[None, None]
        out.append((img, seed_mask(shape, int(rng.integers(1, 33)), 50000 + k, device), sp))
    return out
