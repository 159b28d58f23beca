"""Synthetic workloads of the benchmark, deterministic (SplitMix64 seeds).

batch64: 64 displaced geodesic icospheres, T = 20 n^2 with T log-uniform on [1e3, 5e5]
(n = round(sqrt(T / 20))), radial displacement r = 1 + 0.1 f / max|f| with f a seeded polynomial
of degree <= 8; each mesh's query grid is side^3 nodes over its bounding box padded by 10% per side,
side log-uniform on [32, 192] (about 84 M queries in all)."""
from __future__ import annotations

import math

import numpy as np

from . import meshgen as mg


def batch64(nmeshes: int = 64, seed: int = 3, tmin: float = 1e3, tmax: float = 5e5, smin: float = 32,
            smax: float = 192):
    """(meshes, dims [n,3], boxes [n,6] = min xyz, max xyz)."""
    rng = mg.SplitMix64(seed)
    u_t = rng.uniform(nmeshes)
    u_s = rng.uniform(nmeshes)
    meshes, dims, boxes = [], [], []
    for i in range(nmeshes):
        T = math.exp(math.log(tmin) + u_t[i] * (math.log(tmax) - math.log(tmin)))
        m = mg.displaced_icosphere(int(round(T)), seed=1000 + i)
        side = int(round(math.exp(math.log(smin) + u_s[i] * (math.log(smax) - math.log(smin)))))
        lo, hi = m.bbox()
        pad = 0.1 * (hi - lo)
        meshes.append(m)
        dims.append((side, side, side))
        boxes.append(np.concatenate([lo - pad, hi + pad]).astype(np.float64))
    return meshes, np.array(dims, np.int64), np.array(boxes, np.float64)
