"""A mesh past 2^26 triangles: the traversal whose node records are addressed by 32-bit byte offsets
serves trees of fewer than 2^26 internal nodes only, and larger meshes take the 64-bit one even when
the tree is shallow.  A UV sphere is queried alone and again inside a mesh that adds a flat grid of
2^26 + 2^20 - 2,208 triangles far away (x in [10, 11]); FP32 + FP64 fixup results are the canonical
FP64 argmin, so every grid node near the sphere must get the same signed distance and triangle bit
for bit.  About 35 GB of device memory."""
import numpy as np
import pytest

import ssdf_amd as S
from ssdf_amd import meshgen as mg

pytestmark = pytest.mark.gpu


def _flat_grid(ntri_min: int, v0: int):
    """A flat N x N quad grid on [10, 11] x [-0.5, 0.5] at z = 0.25 with at least ntri_min triangles;
    vertex indices start at v0."""
    n = int(np.ceil(np.sqrt(ntri_min / 2)))
    t = np.linspace(0.0, 1.0, n + 1, dtype=np.float64)
    gx, gy = np.meshgrid(10.0 + t, t - 0.5, indexing="ij")
    sx = gx.ravel().astype(np.float32)
    sy = gy.ravel().astype(np.float32)
    sz = np.full(sx.size, 0.25, np.float32)
    i, j = np.meshgrid(np.arange(n, dtype=np.int32), np.arange(n, dtype=np.int32), indexing="ij")
    a = (i * (n + 1) + j).ravel() + np.int32(v0)
    del i, j
    b, c, d = a + np.int32(n + 1), a + np.int32(1), a + np.int32(n + 2)
    s0 = np.empty(2 * a.size, np.int32)
    s1 = np.empty_like(s0)
    s2 = np.empty_like(s0)
    s0[0::2], s1[0::2], s2[0::2] = a, b, c
    s0[1::2], s1[1::2], s2[1::2] = c, b, d
    return sx, sy, sz, s0, s1, s2


def test_mesh_past_2_26_triangles_matches_the_small_mesh():
    sphere = mg.uv_sphere(0.5, 48, 24)
    n = (16, 16, 16)
    lo, hi = (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0)
    N = n[0] * n[1] * n[2]

    def query(m):
        rm = S.ResidentMesh.from_mesh(m, with_normals=True)
        out = S.DeviceArray(N, np.float32)
        tri = S.DeviceArray(N, np.int32)
        rm.query_grid(S.MODE_SDF, n, lo, hi, out, tri, flags=S.FLAG_FIXUP)
        _, err = rm.last_stats()
        info = rm.info()
        res = out.download(), tri.download()
        rm.close()
        return res, err, info

    (d_small, t_small), err, _ = query(sphere)
    assert err == 0
    gx, gy, gz, g0, g1, g2 = _flat_grid((1 << 26) + (1 << 20) - sphere.ntris, sphere.nverts)
    big = mg.Mesh(np.concatenate([sphere.sx, gx]), np.concatenate([sphere.sy, gy]), np.concatenate([sphere.sz, gz]),
                  np.concatenate([sphere.s0, g0]), np.concatenate([sphere.s1, g1]), np.concatenate([sphere.s2, g2]),
                  name="sphere + far grid")
    del gx, gy, gz, g0, g1, g2
    assert big.ntris >= (1 << 26) + 1
    print(f"big mesh: {big.ntris} triangles", flush=True)
    (d_big, t_big), err, info = query(big)
    assert err == 0
    assert info["maxdepth"] <= 64  # shallow: the size alone routes it to the 64-bit traversal
    assert np.array_equal(t_big, t_small)
    assert np.array_equal(d_big.view(np.int32), d_small.view(np.int32))
