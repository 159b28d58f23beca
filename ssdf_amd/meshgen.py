"""Deterministic synthetic meshes and query sets for the distance-field workloads.

Every random draw comes from an in-repo SplitMix64 stream (Steele, Lea and Flood, OOPSLA 2014)
turned into 53-bit uniform doubles and, where needed, Box-Muller Gaussians, so a (generator, seed)
pair names the same input on every machine.  Coordinates are returned as float32 SoA arrays and
triangles as three int32 index arrays, the layout ssdf's C ABI takes (serial/ssdf.cpp:36-48).
All meshes except deep_chain (a triangle soup for tree-depth tests) are closed and consistently
oriented: counter-clockwise seen from outside.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


@dataclass
class Mesh:
    sx: np.ndarray
    sy: np.ndarray
    sz: np.ndarray
    s0: np.ndarray
    s1: np.ndarray
    s2: np.ndarray
    name: str = "mesh"

    @property
    def nverts(self) -> int:
        return int(self.sx.shape[0])

    @property
    def ntris(self) -> int:
        return int(self.s0.shape[0])

    def bbox(self):
        lo = np.array([self.sx.min(), self.sy.min(), self.sz.min()], dtype=np.float64)
        hi = np.array([self.sx.max(), self.sy.max(), self.sz.max()], dtype=np.float64)
        return lo, hi

    def as_dtype(self, dtype) -> "Mesh":
        return Mesh(self.sx.astype(dtype), self.sy.astype(dtype), self.sz.astype(dtype), self.s0, self.s1,
                    self.s2, self.name)

    def checksum(self) -> str:
        import hashlib
        h = hashlib.sha256()
        for a in (self.sx, self.sy, self.sz, self.s0, self.s1, self.s2):
            h.update(np.ascontiguousarray(a).tobytes())
        return h.hexdigest()[:16]


class SplitMix64:
    """Counter-based SplitMix64: draw i of stream `seed` is mix(seed + (i + 1) * golden)."""

    def __init__(self, seed: int):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.count = 0

    def next_u64(self, n: int) -> np.ndarray:
        with np.errstate(over="ignore"):
            k = np.arange(self.count + 1, self.count + 1 + n, dtype=np.uint64)
            z = self.state + k * _GOLDEN
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            z = z ^ (z >> np.uint64(31))
        self.count += n
        return z

    def uniform(self, n: int) -> np.ndarray:
        """n doubles in [0, 1) with 53 random bits each."""
        return (self.next_u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)

    def normal(self, n: int) -> np.ndarray:
        """n standard normals by Box-Muller (cosine branch only)."""
        u1 = 1.0 - self.uniform(n)  # (0, 1]
        u2 = self.uniform(n)
        return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


def _mesh(v: np.ndarray, f: np.ndarray, name: str) -> Mesh:
    v32 = np.ascontiguousarray(v.astype(np.float32))
    f = np.ascontiguousarray(f.astype(np.int32))
    return Mesh(np.ascontiguousarray(v32[:, 0]), np.ascontiguousarray(v32[:, 1]), np.ascontiguousarray(v32[:, 2]),
                np.ascontiguousarray(f[:, 0]), np.ascontiguousarray(f[:, 1]), np.ascontiguousarray(f[:, 2]), name)


def uv_sphere(radius: float = 0.5, nlon: int = 48, nlat: int = 24) -> Mesh:
    """UV sphere: two poles and nlat-1 rings of nlon vertices; V = nlon (nlat-1) + 2 and
    T = 2 nlon (nlat-2) + 2 nlon (1,106 and 2,208 for 48 x 24)."""
    verts = [(0.0, 0.0, radius)]
    for k in range(1, nlat):
        th = math.pi * k / nlat
        for j in range(nlon):
            ph = 2.0 * math.pi * j / nlon
            verts.append((radius * math.sin(th) * math.cos(ph), radius * math.sin(th) * math.sin(ph),
                          radius * math.cos(th)))
    verts.append((0.0, 0.0, -radius))
    south = len(verts) - 1

    def ring(k, j):
        return 1 + (k - 1) * nlon + (j % nlon)

    faces = []
    for j in range(nlon):
        faces.append((0, ring(1, j), ring(1, j + 1)))
    for k in range(1, nlat - 1):
        for j in range(nlon):
            a, b, c, d = ring(k, j), ring(k + 1, j), ring(k + 1, j + 1), ring(k, j + 1)
            faces.append((a, b, c))
            faces.append((a, c, d))
    for j in range(nlon):
        faces.append((ring(nlat - 1, j), south, ring(nlat - 1, j + 1)))
    return _mesh(np.array(verts, dtype=np.float64), np.array(faces), f"uv_sphere_{nlon}x{nlat}")


def torus(R: float = 0.6, r: float = 0.2, nu: int = 500, nv: int = 200, sigma: float = 0.0,
          seed: int = 2) -> Mesh:
    """Torus around z with nu x nv quads split in two; each vertex displaced along the analytic
    outward normal by N(0, sigma^2) from SplitMix64(seed)."""
    u = 2.0 * np.pi * np.arange(nu, dtype=np.float64) / nu
    v = 2.0 * np.pi * np.arange(nv, dtype=np.float64) / nv
    U, V = np.meshgrid(u, v, indexing="ij")  # vertex (i, j) -> i * nv + j
    nx, ny, nz = np.cos(V) * np.cos(U), np.cos(V) * np.sin(U), np.sin(V)
    px, py, pz = (R + r * np.cos(V)) * np.cos(U), (R + r * np.cos(V)) * np.sin(U), r * np.sin(V)
    if sigma > 0:
        g = SplitMix64(seed).normal(nu * nv).reshape(nu, nv) * sigma
        px, py, pz = px + g * nx, py + g * ny, pz + g * nz
    verts = np.stack([px.ravel(), py.ravel(), pz.ravel()], axis=1)
    i = np.arange(nu)[:, None]
    j = np.arange(nv)[None, :]
    a = i * nv + j
    b = ((i + 1) % nu) * nv + j
    c = ((i + 1) % nu) * nv + (j + 1) % nv
    d = i * nv + (j + 1) % nv
    f1 = np.stack([a, b, c], axis=-1).reshape(-1, 3)
    f2 = np.stack([a, c, d], axis=-1).reshape(-1, 3)
    faces = np.empty((2 * nu * nv, 3), dtype=np.int64)
    faces[0::2] = f1
    faces[1::2] = f2
    return _mesh(verts, faces, f"torus_{nu}x{nv}_s{sigma}")


_ICO_V = None


def _icosahedron():
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
                  (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
                  (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5),
                  (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)])
    return v, f


def icosphere(freq: int) -> Mesh:
    """Frequency-`freq` geodesic icosphere (class I): each icosahedron face split into freq^2
    triangles, vertices projected to the unit sphere; V = 10 freq^2 + 2, T = 20 freq^2."""
    v0, f0 = _icosahedron()
    n = int(freq)
    # barycentric lattice of one face: (i, j) with i + j <= n; point = (l*A + i*B + j*C) / n
    ii, jj = [], []
    for i in range(n + 1):
        for j in range(n + 1 - i):
            ii.append(i)
            jj.append(j)
    ii = np.array(ii)
    jj = np.array(jj)
    ll = n - ii - jj
    lattice_index = -np.ones((n + 1, n + 1), dtype=np.int64)
    lattice_index[ii, jj] = np.arange(ii.size)
    tri = []
    for i in range(n):
        for j in range(n - i):
            tri.append((lattice_index[i, j], lattice_index[i + 1, j], lattice_index[i, j + 1]))
            if i + j + 1 < n:
                tri.append((lattice_index[i + 1, j], lattice_index[i + 1, j + 1], lattice_index[i, j + 1]))
    tri = np.array(tri)
    A, B, C = v0[f0[:, 0]], v0[f0[:, 1]], v0[f0[:, 2]]
    # a zero weight contributes an exact zero, so shared edge points come out bit-identical from
    # both faces (two-term sums commute)
    pts = (ll[None, :, None] * A[:, None, :] + ii[None, :, None] * B[:, None, :] + jj[None, :, None] * C[:, None, :])
    pts = pts / float(n)
    pts = pts / np.linalg.norm(pts, axis=2, keepdims=True)
    allp = pts.reshape(-1, 3)
    uniq, inv = np.unique(allp, axis=0, return_inverse=True)
    inv = inv.reshape(20, -1)
    faces = np.concatenate([inv[f][tri] for f in range(20)], axis=0)
    if uniq.shape[0] != 10 * n * n + 2:
        raise RuntimeError(f"icosphere dedup failed: {uniq.shape[0]} vertices for freq {n}")
    return _mesh(uniq, faces, f"icosphere_{n}")


def displaced_icosphere(ntris_target: int, seed: int, amplitude: float = 0.1, degree: int = 8) -> Mesh:
    """Icosphere with T = 20 n^2 close to ntris_target, radially displaced r = 1 + a f / max|f|, f a
    seeded polynomial of total degree <= `degree` in (x, y, z) — on the sphere, exactly a
    spherical-harmonic expansion of degree <= `degree`."""
    n = max(1, int(round(math.sqrt(ntris_target / 20.0))))
    m = icosphere(n)
    v = np.stack([m.sx, m.sy, m.sz], axis=1).astype(np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    rng = SplitMix64(seed)
    terms = [(a, b, c) for a in range(degree + 1) for b in range(degree + 1 - a) for c in range(degree + 1 - a - b)]
    coef = rng.normal(len(terms))
    # power tables: x^a by repeated products
    P = [[np.ones(v.shape[0])] for _ in range(3)]
    for d in range(3):
        for _ in range(degree):
            P[d].append(P[d][-1] * v[:, d])
    f = np.zeros(v.shape[0])
    for (a, b, c), w in zip(terms, coef):
        f += w * (P[0][a] * P[1][b] * P[2][c])
    f /= np.max(np.abs(f))
    v *= (1.0 + amplitude * f)[:, None]
    faces = np.stack([m.s0, m.s1, m.s2], axis=1)
    out = _mesh(v, faces, f"dispico_{n}_s{seed}")
    return out


def displaced_torus(ntris_target: int, seed: int, R: float = 0.6, r: float = 0.2, sigma_rel: float = 0.02) -> Mesh:
    """Torus with nu x nv ~ ntris_target / 2 (nu = 2.5 nv) and normal noise sigma = sigma_rel * r."""
    nv = max(8, int(round(math.sqrt(ntris_target / 5.0))))
    nu = max(8, int(round(2.5 * nv)))
    return torus(R, r, nu, nv, sigma=sigma_rel * r, seed=seed)


def grid_axis(n: int, lo: float, hi: float, dtype=np.float32) -> np.ndarray:
    """Grid coordinates x_i = lo + i * h, h = (hi - lo) / (n - 1), every operation rounded in
    `dtype` (no fused multiply-add), as serial/bvh/bvh.cpp:1665-1667 and 1744 evaluate them."""
    dt = np.dtype(dtype).type
    lo_, hi_ = dt(lo), dt(hi)
    h = (hi_ - lo_) / dt(n - 1) if n > 1 else dt(0)
    i = np.arange(n).astype(dtype)
    return (lo_ + i * h).astype(dtype)


def grid_points(nx: int, ny: int, nz: int, lo, hi, dtype=np.float32):
    """SoA coordinates of all grid nodes, x fastest: index i + j nx + k nx ny."""
    ax = grid_axis(nx, lo[0], hi[0], dtype)
    ay = grid_axis(ny, lo[1], hi[1], dtype)
    az = grid_axis(nz, lo[2], hi[2], dtype)
    z, y, x = np.meshgrid(az, ay, ax, indexing="ij")
    return np.ascontiguousarray(x.ravel()), np.ascontiguousarray(y.ravel()), np.ascontiguousarray(z.ravel())


def staircase(N: int = 4, H: float = 0.25) -> tuple[Mesh, list]:
    """The voxel staircase of tests/sample_sdf_bvh_test.cpp:18-90 (reference): lattice cells (i, j, k)
    with i <= j and 1 <= k <= 2, boundary faces split into two outward triangles.  Returns the
    mesh and the list of boundary faces (axis, sign, lo[3], hi[3]) for exact distances."""
    def filled(i, j, k):
        if i < 0 or j < 0 or k < 0 or i >= N or j >= N or k >= N:
            return False
        return i <= j and 1 <= k <= 2

    M = N + 1
    vid = lambda i, j, k: i + M * (j + M * k)  # noqa: E731
    verts = [(i * H, j * H, k * H) for k in range(M) for j in range(M) for i in range(M)]
    tris, faces = [], []
    for k in range(N):
        for j in range(N):
            for i in range(N):
                if not filled(i, j, k):
                    continue
                for axis in range(3):
                    for sign in (-1, 1):
                        nb = [i, j, k]
                        nb[axis] += sign
                        if filled(*nb):
                            continue
                        u, v = (axis + 1) % 3, (axis + 2) % 3
                        c = [[i, j, k] for _ in range(4)]
                        for q in range(4):
                            c[q][axis] += 1 if sign > 0 else 0
                        c[1][u] += 1
                        c[2][u] += 1
                        c[2][v] += 1
                        c[3][v] += 1
                        q = [0, 1, 2, 3]
                        if sign < 0:
                            q[1], q[3] = q[3], q[1]
                        a, b, d, e = (vid(*c[q[0]]), vid(*c[q[1]]), vid(*c[q[2]]), vid(*c[q[3]]))
                        tris.append((a, b, d))
                        tris.append((a, d, e))
                        lo = [min(c[0][dd], c[2][dd]) * H for dd in range(3)]
                        hi = [max(c[0][dd], c[2][dd]) * H for dd in range(3)]
                        faces.append((axis, sign, lo, hi))
    return _mesh(np.array(verts, dtype=np.float64), np.array(tris), f"staircase_{N}"), faces


def deep_chain(ndup: int = 8) -> Mesh:
    """Open triangle soup whose LBVH is more than 64 levels deep (for the two-VGPR traversal stack).

    Small triangles are centred at the origin (ndup copies), at (1, 1, 1), and at one point per
    Morton bit: bit i of the 63-bit key is bit i // 3 of axis 2 - i % 3, so a centre at 2^(i // 3 - 21)
    on that axis alone has the key 2^i once the centre bounds are [0, 1].  The sorted keys
    0, 1, 2, 4, ..., 2^62 split one bit per level, and the ndup equal keys at 0 add log2(ndup) levels
    of index splits below that.  Every coordinate is exact in float32, so are the box centres.
    """
    centres = [(0.0, 0.0, 0.0)] * ndup + [(1.0, 1.0, 1.0)]
    for i in range(63):
        c = [0.0, 0.0, 0.0]
        c[2 - i % 3] = 2.0 ** (i // 3 - 21)
        centres.append(tuple(c))
    verts, tris = [], []
    for c in centres:
        d = max(c) * 2.0 ** -10 if max(c) > 0 else 2.0 ** -31
        b = len(verts)
        verts += [(c[0] - d, c[1] - d, c[2] - d), (c[0] + d, c[1] + d, c[2] + d), (c[0] + d, c[1] - d, c[2])]
        tris.append((b, b + 1, b + 2))
    return _mesh(np.array(verts, dtype=np.float64), np.array(tris), f"deep_chain_{ndup}")


def padded_bbox(mesh: Mesh, pad: float = 0.1):
    lo, hi = mesh.bbox()
    ext = hi - lo
    return lo - pad * ext, hi + pad * ext
