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


def _seg_seg_dist(p0, p1, q0, q1) -> float:
    """Distance between segments [p0, p1] and [q0, q1] (closest points by clamped parameters)."""
    d1, d2, r = p1 - p0, q1 - q0, p0 - q0
    a, e, f = d1 @ d1, d2 @ d2, d2 @ r
    c, b = d1 @ r, d1 @ d2
    den = a * e - b * b
    s = np.clip((b * f - c * e) / den, 0.0, 1.0) if den > 1e-18 else 0.0
    t = (b * s + f) / e
    if t < 0.0:
        t, s = 0.0, np.clip(-c / a, 0.0, 1.0)
    elif t > 1.0:
        t, s = 1.0, np.clip((b - c) / a, 0.0, 1.0)
    return float(np.linalg.norm((p0 + d1 * s) - (q0 + d2 * t)))


def capped_cylinder(p0, p1, r: float, n_circ: int, n_axial: int):
    """Closed, outward-oriented cylinder of radius r around the segment p0 -> p1: n_axial rings of
    n_circ quads (two triangles each) and a triangle fan on each end: V = n_circ (n_axial + 1) + 2,
    T = 2 n_circ (n_axial + 1)."""
    axis = p1 - p0
    L = float(np.linalg.norm(axis))
    w = axis / L
    u = np.cross(w, [1.0, 0.0, 0.0] if abs(w[0]) < 0.9 else [0.0, 1.0, 0.0])
    u /= np.linalg.norm(u)
    v = np.cross(w, u)
    th = 2.0 * np.pi * np.arange(n_circ) / n_circ
    ring = np.cos(th)[:, None] * u + np.sin(th)[:, None] * v  # outward unit directions
    t = np.arange(n_axial + 1) / n_axial
    verts = (p0 + t[:, None, None] * axis + r * ring[None, :, :]).reshape(-1, 3)
    verts = np.vstack([verts, p0, p1])
    c0, c1 = verts.shape[0] - 2, verts.shape[0] - 1
    i = np.arange(n_axial)[:, None]
    j = np.arange(n_circ)[None, :]
    a = i * n_circ + j
    b = i * n_circ + (j + 1) % n_circ
    c = (i + 1) * n_circ + (j + 1) % n_circ
    d = (i + 1) * n_circ + j
    side = np.concatenate([np.stack([a, b, c], -1).reshape(-1, 3), np.stack([a, c, d], -1).reshape(-1, 3)])
    jj = np.arange(n_circ)
    cap0 = np.stack([np.full(n_circ, c0), (jj + 1) % n_circ, jj], 1)  # facing -w
    top = n_axial * n_circ
    cap1 = np.stack([np.full(n_circ, c1), top + jj, top + (jj + 1) % n_circ], 1)  # facing +w
    return verts, np.concatenate([side, cap0, cap1])


def beams(nbeams: int = 200, seed: int = 17, n_circ: int = 64, n_axial: int = 194,
          hole_fraction: float = 0.0):
    """Capped cylinders with radius U[0.01, 0.03], length U[0.2, 0.8] and uniformly random
    orientation, kept only if their axes stay further apart than the sum of the radii (plus 0.005)
    and inside the unit cube; n_circ x n_axial side quads plus two fans, 24,960 triangles each
    by default (4.99 M for 200 beams).  hole_fraction > 0 deletes that seeded
    fraction of the triangles (an open surface)."""
    rng = np.random.default_rng(seed)
    segs, verts, faces = [], [], []
    off = 0
    tries = 0
    while len(segs) < nbeams:
        tries += 1
        if tries > 200 * nbeams:
            raise RuntimeError("beams: could not place the cylinders")
        r = rng.uniform(0.01, 0.03)
        L = rng.uniform(0.2, 0.8)
        w = rng.normal(size=3)
        w /= np.linalg.norm(w)
        c = rng.uniform(0.0, 1.0, 3)
        p0, p1 = c - 0.5 * L * w, c + 0.5 * L * w
        if np.any(np.minimum(p0, p1) < r) or np.any(np.maximum(p0, p1) > 1.0 - r):
            continue
        if any(_seg_seg_dist(p0, p1, q0, q1) <= r + s + 0.005 for q0, q1, s in segs):
            continue
        segs.append((p0, p1, r))
        v, f = capped_cylinder(p0, p1, r, n_circ, n_axial)
        verts.append(v)
        faces.append(f + off)
        off += v.shape[0]
    V = np.concatenate(verts)
    F = np.concatenate(faces)
    if hole_fraction > 0:
        keep = np.random.default_rng(seed + 1).random(F.shape[0]) >= hole_fraction
        F = F[keep]
    return mg._mesh(V, F, f"beams_{nbeams}_s{seed}_h{hole_fraction}")
