"""Float32 Dijkstra brute force for tiny grids (test infrastructure only).

The raster scan relaxes, in every pass, each element p from each neighbour q in
the previous slice with ``cur = fl(fl(d_q + fl(l_eucl*LD)) + fl(l_grad*l(p,q)))``
(FastGeodis/fastgeodis_cpu.cpp:106-110, 275-277).  With l_eucl, l_grad >= 0
that relax is monotone and never decreases d_q, so once the raster iterations
stop changing anything its output equals the shortest-path fixpoint over the
union of every pass's edge set.  This module computes that fixpoint directly,
which pins the neighbour sets, the per-pass step tables (including the
per-pass summation order of the 3D diagonal), the spacing-to-axis mapping and
the channel sum, independently of the sweep order.

Edge sets (offset (dz, dh, dw) from q to p):
* 2D: the 8-neighbourhood; LD = 1 for axis steps and sqrt(2f) for diagonals
  (fastgeodis_cpu.cpp:67).
* 3D: the 26-neighbourhood.  An offset with dz != 0 is relaxed by the D pass
  with table(S_D, S_H, S_W)[dh, dw]; dh != 0 by the H pass with
  table(S_H, S_D, S_W)[dz, dw]; dw != 0 by the W pass with
  table(S_W, S_H, S_D)[dh, dz] (fastgeodis_cpu.cpp:220-231, 365, 373-379,
  391-397).  An offset covered by several passes is relaxed with each table.
"""
from __future__ import annotations

import heapq
import itertools

import numpy as np

f32 = np.float32


def _table(s0, s1, s2):
    t = np.zeros((3, 3), dtype=np.float32)
    for i in range(3):
        for j in range(3):
            v = f32(s0)
            v = f32(v + f32(abs(i - 1)) * f32(s1))
            v = f32(v + f32(abs(j - 1)) * f32(s2))
            t[i, j] = np.sqrt(v, dtype=np.float32)
    return t


def _ldist(img, p, q):
    C = img.shape[0]
    if C == 1:
        return f32(abs(f32(img[(0,) + p] - img[(0,) + q])))
    acc = f32(0.0)
    for c in range(C):
        acc = f32(acc + abs(f32(img[(c,) + p] - img[(c,) + q])))
    return acc


def geodesic_fixpoint(image, mask, v, l_grad, l_eucl, spacing=None):
    """Shortest-path fixpoint of the raster relaxations.

    image: (1, C, H, W) or (1, C, D, H, W); mask: matching spatial shape
    (channel 0 is used).  Returns distance channel 0 as float32 array of the
    spatial shape.
    """
    img = np.asarray(image, dtype=np.float32)[0]
    m = np.asarray(mask, dtype=np.float32)[0, 0]
    shape = m.shape
    nd = len(shape)
    lg, le = f32(l_grad), f32(l_eucl)
    if nd == 2:
        r2 = np.sqrt(f32(2.0), dtype=np.float32)
        edges = []
        for dh, dw in itertools.product((-1, 0, 1), repeat=2):
            if dh == 0 and dw == 0:
                continue
            ld = f32(1.0) if (dh == 0 or dw == 0) else r2
            edges.append(((dh, dw), f32(le * ld)))
    else:
        sq = [f32(f32(s) * f32(s)) for s in spacing]
        tD = _table(sq[0], sq[1], sq[2])
        tH = _table(sq[1], sq[0], sq[2])
        tW = _table(sq[2], sq[1], sq[0])
        edges = []
        for dz, dh, dw in itertools.product((-1, 0, 1), repeat=3):
            if dz == 0 and dh == 0 and dw == 0:
                continue
            if dz != 0:
                edges.append(((dz, dh, dw), f32(le * tD[dh + 1, dw + 1])))
            if dh != 0:
                edges.append(((dz, dh, dw), f32(le * tH[dz + 1, dw + 1])))
            if dw != 0:
                edges.append(((dz, dh, dw), f32(le * tW[dh + 1, dz + 1])))

    d = (f32(v) * m).astype(np.float32)
    heap = [(float(d[idx]), idx) for idx in np.ndindex(*shape)]
    heapq.heapify(heap)
    done = np.zeros(shape, dtype=bool)
    while heap:
        dq, q = heapq.heappop(heap)
        if done[q] or dq != float(d[q]):
            continue
        done[q] = True
        for off, e1 in edges:
            p = tuple(q[k] + off[k] for k in range(nd))
            if any(p[k] < 0 or p[k] >= shape[k] for k in range(nd)) or done[p]:
                continue
            e2 = f32(lg * _ldist(img, p, q))
            cur = f32(f32(d[q] + e1) + e2)
            if cur < d[p]:
                d[p] = cur
                heapq.heappush(heap, (float(cur), p))
    return d
