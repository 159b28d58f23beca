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


def serve_pool(nmeshes: int = 32, seed: int = 5, tmin: float = 5e3, tmax: float = 2e6):
    """Serving mesh pool: half displaced icospheres, half noisy tori, T log-uniform on [tmin, tmax]."""
    rng = mg.SplitMix64(seed)
    u = rng.uniform(nmeshes)
    meshes = []
    for i in range(nmeshes):
        T = int(round(math.exp(math.log(tmin) + u[i] * (math.log(tmax) - math.log(tmin)))))
        if i < nmeshes // 2:
            meshes.append(mg.displaced_icosphere(T, seed=2000 + i))
        else:
            meshes.append(mg.displaced_torus(T, seed=3000 + i))
    return meshes


def query_cloud(m, n: int, seed: int, pad: float = 0.1, offset: float = 0.02):
    """n FP32 points: half uniform in the bounding box padded by `pad` per side, half at area-weighted
    random surface points moved along the face normal by N(0, (offset / 3)^2) truncated to
    +-offset."""
    rng = np.random.default_rng(seed)
    lo, hi = m.bbox()
    ext = hi - lo
    nu = n // 2
    P = np.empty((n, 3))
    P[:nu] = rng.uniform(lo - pad * ext, hi + pad * ext, (nu, 3))
    v = np.stack([m.sx, m.sy, m.sz], 1).astype(np.float64)
    a, b, c = v[m.s0], v[m.s1], v[m.s2]
    cr = np.cross(b - a, c - a)
    area = 0.5 * np.linalg.norm(cr, axis=1)
    cdf = np.cumsum(area)
    t = np.minimum(np.searchsorted(cdf, rng.uniform(0, cdf[-1], n - nu)), m.ntris - 1)
    r1, r2 = rng.uniform(size=(2, n - nu))
    s = np.sqrt(r1)
    w0, w1, w2 = 1 - s, s * (1 - r2), s * r2
    p = w0[:, None] * a[t] + w1[:, None] * b[t] + w2[:, None] * c[t]
    nrm = cr[t] / np.maximum(np.linalg.norm(cr[t], axis=1, keepdims=True), 1e-300)
    off = np.clip(rng.normal(0.0, offset / 3.0, n - nu), -offset, offset)
    P[nu:] = p + off[:, None] * nrm
    return [np.ascontiguousarray(P[:, d], dtype=np.float32) for d in range(3)]


def poisson_requests(rate: float, duration: float, nmeshes: int, seed: int = 11, qmin: float = 1e4,
                     qmax: float = 1e7):
    """Arrival times (Poisson, `rate` per second over `duration` s), mesh ids (uniform) and query
    counts (log-uniform on [qmin, qmax])."""
    rng = mg.SplitMix64(seed)
    times = []
    t = 0.0
    while True:
        t += -math.log(1.0 - float(rng.uniform(1)[0])) / rate
        if t >= duration:
            break
        times.append(t)
    k = len(times)
    mesh = (rng.uniform(k) * nmeshes).astype(np.int64) % nmeshes
    nq = np.exp(math.log(qmin) + rng.uniform(k) * (math.log(qmax) - math.log(qmin))).astype(np.int64)
    return np.array(times), mesh, nq
