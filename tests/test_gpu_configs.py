"""GPU parity at the benchmark workloads' full sizes, against the oracle's exact FP64 argmin.

Each workload runs in full on the GPU; a seeded subsample of its queries is checked against the
oracle (exact FP64 argmin through its CPU BVH, which equals its brute force, plus the reference's
pseudonormal sign rule), and properties that hold for every query are checked over all of them.

  * noisy torus, 200,000 triangles, 256^3 grid on [-1, 1]^3: FP32 + fixup on every one of the
    16.7 M nodes against the oracle, and every node's sign against the analytic torus outside the
    band that the vertex noise can reach; plain FP32 on 2^18 nodes.
  * 64 displaced icospheres (1 k - 500 k triangles, 4.1 M in all) with their own grids (84.8 M
    nodes) in one batch: 4096 nodes per mesh against the oracle.
  * the serving pool (32 displaced icospheres and tori, 5 k - 2 M triangles): one request per mesh,
    all in flight together, each with its own points (half uniform in the padded box, half within
    0.02 of the surface); 8192 points per request against the oracle, and every result equal to the
    same request served alone.
  * 200 capped beams (4.99 M triangles) on the 512^3 grid over the unit cube and at 10^8 uniform
    points, intact and with 1% of the triangles deleted (signed by winding numbers through
    SSDF_FLAG_SIGN_AUTO): 2^16 queries per set against the oracle; the holed mesh's signs equal the
    intact mesh's at every node farther than 0.01 from it.

Bars (tests/bars.py): FP32 + fixup indices identical to the FP64 argmin, distances within 1e-6
relative, no sign mismatch outside the 1e-6 * bbox band; plain FP32 within 1e-6 * bbox diagonal
with its relative error reported.
"""
import numpy as np
import pytest

import ssdf_amd as S
from ssdf_amd import meshgen as mg
from ssdf_amd import workloads as W
from tests import bars

pytestmark = pytest.mark.gpu


def _diag(m):
    lo, hi = m.bbox()
    return float(np.linalg.norm(hi - lo))


def _grid_sample(n, lo, hi, k, seed):
    """k distinct seeded node indices of an n[0] x n[1] x n[2] grid and their FP32 coordinates."""
    N = n[0] * n[1] * n[2]
    rng = np.random.default_rng(seed)
    idx = np.sort(rng.choice(N, size=min(k, N), replace=False))
    i, j, kk = idx % n[0], (idx // n[0]) % n[1], idx // (n[0] * n[1])
    ax = [mg.grid_axis(n[d], lo[d], hi[d]) for d in range(3)]
    return idx, ax[0][i], ax[1][j], ax[2][kk]


def _check_fixup(what, m, out, tri, x, y, z, oracle, normals=None, signed=True):
    s, rtri, _, _, _ = oracle.sdf(x, y, z, m, normals=normals)
    assert np.array_equal(tri, rtri), f"{what}: {int((tri != rtri).sum())} of {tri.size} indices differ"
    bars.assert_fixup_distances(out, np.abs(s), what)
    if signed:
        band = 1e-6 * _diag(m)
        assert bars.sign_mismatches_outside_band(out, s, band) == 0, what
    return s


# ----------------------------------------------------------------------------------------------
# noisy torus 200 k x 256^3
# ----------------------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def torus200k():
    m = mg.torus(0.6, 0.2, 500, 200, sigma=0.005, seed=2)
    return m, S.ResidentMesh.from_mesh(m)


def _torus_grid(rm, flags):
    n = (256, 256, 256)
    N = 256 ** 3
    out = S.DeviceArray(N, np.float32)
    tri = S.DeviceArray(N, np.int32)
    rm.query_grid(S.MODE_SDF, n, (-1, -1, -1), (1, 1, 1), out, tri, flags=flags)
    q, err = rm.last_stats()
    assert err == 0
    return out.download(), tri.download(), q


def test_torus200k_grid256_fixup(oracle, torus200k):
    """Every one of the 16.7 M nodes against the oracle (FP64 argmin through its CPU BVH, the
    reference's pseudonormal sign), in slabs of 32 z slices."""
    m, rm = torus200k
    out, tri, q = _torus_grid(rm, S.FLAG_FIXUP)
    ax = mg.grid_axis(256, -1.0, 1.0)
    yy, xx = np.meshgrid(ax, ax, indexing="ij")
    normals = oracle.pseudonormals(m)
    worst = 0.0
    for k0 in range(0, 256, 32):
        zz = np.repeat(ax[k0:k0 + 32], 65536)
        x, y = np.tile(xx.ravel(), 32), np.tile(yy.ravel(), 32)
        sl = slice(k0 * 65536, (k0 + 32) * 65536)
        s = _check_fixup(f"torus200k fixup, z slices {k0}-{k0 + 31}", m, out[sl], tri[sl], x, y, zz, oracle,
                         normals=normals)
        worst = max(worst, float(np.max(np.abs(np.abs(out[sl].astype(np.float64)) - np.abs(s)) /
                                       np.maximum(np.abs(s), 1e-30))))
    print(f"[torus200k fixup] all {256 ** 3} nodes: indices equal the FP64 argmin, worst relative error {worst:.3g}")
    # every node: the analytic torus decides inside/outside beyond the band the vertex noise reaches
    R, r = 0.6, 0.2
    vt = np.sqrt((np.hypot(m.sx.astype(np.float64), m.sy.astype(np.float64)) - R) ** 2 + m.sz.astype(np.float64) ** 2) - r
    band = float(np.abs(vt).max()) + 1e-3
    ax = mg.grid_axis(256, -1.0, 1.0).astype(np.float64)
    mism = 0
    for k in range(256):  # one z slice at a time
        zz = ax[k]
        yy, xx = np.meshgrid(ax, ax, indexing="ij")
        dt = np.sqrt((np.hypot(xx, yy) - R) ** 2 + zz * zz) - r
        sl = out[k * 65536:(k + 1) * 65536].reshape(256, 256)
        far = np.abs(dt) > band
        mism += int(np.sum(((sl < 0) != (dt < 0)) & far))
    assert mism == 0
    print(f"[torus200k fixup] queued {q}, analytic-sign band {band:.4f}")


def test_torus200k_grid256_plain(oracle, torus200k):
    m, rm = torus200k
    out, tri, q = _torus_grid(rm, 0)
    idx, x, y, z = _grid_sample((256, 256, 256), (-1, -1, -1), (1, 1, 1), 1 << 18, seed=13)
    d2, rtri = oracle.bvh_argmin(x, y, z, m)
    bars.plain_report(out[idx], np.sqrt(d2), _diag(m), "torus200k grid256")
    # index differs only where the FP32 distances of the two triangles tie to within FP32 rounding
    diff = tri[idx] != rtri
    print(f"[torus200k plain] index differs from the FP64 argmin at {int(diff.sum())} of {idx.size} nodes")


# ----------------------------------------------------------------------------------------------
# 64 displaced icospheres, one batch
# ----------------------------------------------------------------------------------------------
def test_batch64_full(oracle):
    meshes, dims, boxes = W.batch64()
    off, total = S.Batch.grid_layout(dims)
    b = S.Batch(meshes)
    out = S.DeviceArray(total, np.float32)
    tri = S.DeviceArray(total, np.int32)
    b.query_grids(S.MODE_SDF, dims, boxes, off, out, tri, flags=S.FLAG_FIXUP)
    q, err = b.last_stats()
    assert err == 0
    ho, ht = out.download(), tri.download()
    worst = 0.0
    for i, m in enumerate(meshes):
        n = tuple(int(v) for v in dims[i])
        idx, x, y, z = _grid_sample(n, boxes[i][:3], boxes[i][3:], 4096, seed=100 + i)
        s, rtri, _, _, _ = oracle.sdf(x, y, z, m)
        o, t = ho[off[i] + idx], ht[off[i] + idx]
        assert np.array_equal(t, rtri), f"mesh {i}: {int((t != rtri).sum())} indices differ"
        worst = max(worst, bars.assert_fixup_distances(o, np.abs(s), f"batch mesh {i}"))
        assert bars.sign_mismatches_outside_band(o, s, 1e-6 * _diag(m)) == 0, f"mesh {i}"
    print(f"[batch64] {total} nodes, {q} queued, worst relative error {worst:.3g}")
    b.close()


# ----------------------------------------------------------------------------------------------
# serving pool: one request per mesh, all in flight together
# ----------------------------------------------------------------------------------------------
def test_serve32_one_request_per_mesh(oracle):
    meshes = W.serve_pool()
    rng = np.random.default_rng(21)
    counts = np.exp(rng.uniform(np.log(1e4), np.log(1e6), len(meshes))).astype(np.int64)
    clouds = [W.query_cloud(m, int(c), seed=500 + i) for i, (m, c) in enumerate(zip(meshes, counts))]
    srv = S.Server(48 << 30, max_inflight=16)
    for i, m in enumerate(meshes):
        srv.add_mesh(i, m)

    def serve(ids):
        outs = {i: (np.empty(counts[i], np.float32), np.empty(counts[i], np.int32)) for i in ids}
        tk = [srv.submit(i, S.MODE_SDF, *clouds[i], outs[i][0], outs[i][1], flags=S.FLAG_FIXUP) for i in ids]
        for t in tk:
            srv.wait(t)
        return outs

    mixed = serve(range(len(meshes)))
    for i, m in enumerate(meshes):
        x, y, z = clouds[i]
        sel = np.sort(np.random.default_rng(700 + i).choice(x.size, size=min(8192, x.size), replace=False))
        _check_fixup(f"serve mesh {i}", m, mixed[i][0][sel], mixed[i][1][sel], x[sel], y[sel], z[sel], oracle)
    # each result is the same alone as in the mix
    for i in (0, 7, 15, 16, 31):
        alone = serve([i])[i]
        assert np.array_equal(alone[0], mixed[i][0]) and np.array_equal(alone[1], mixed[i][1]), i
    st = srv.stats()
    assert st["errors"] == 0
    srv.close()


# ----------------------------------------------------------------------------------------------
# 200 capped beams, 4.99 M triangles
# ----------------------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def beams200():
    full = W.beams()
    holed = W.beams(hole_fraction=0.01)
    return full, holed


def _beams_grid(m, flags, g=512):
    rm = S.ResidentMesh.from_mesh(m)
    N = g ** 3
    out = S.DeviceArray(N, np.float32)
    tri = S.DeviceArray(N, np.int32)
    rm.query_grid(S.MODE_SDF, (g, g, g), (0, 0, 0), (1, 1, 1), out, tri, flags=flags)
    q, err = rm.last_stats()
    assert err == 0
    res = out.download(), tri.download(), rm.info()
    out.free(), tri.free()
    return res, rm


def test_beams200_grid512_intact_and_holed(oracle, beams200):
    full, holed = beams200
    g = 512
    (of, tf, info_f), rm_f = _beams_grid(full, S.FLAG_FIXUP)
    assert info_f["open_edges"] == 0
    idx, x, y, z = _grid_sample((g, g, g), (0, 0, 0), (1, 1, 1), 1 << 16, seed=31)
    _check_fixup("beams grid", full, of[idx], tf[idx], x, y, z, oracle)
    rm_f.close()
    (oh, th, info_h), rm_h = _beams_grid(holed, S.FLAG_FIXUP | S.FLAG_SIGN_AUTO)
    assert info_h["open_edges"] > 0
    # distances and indices of the holed mesh against its own FP64 argmin
    d2, rtri = oracle.bvh_argmin(x, y, z, holed)
    assert np.array_equal(th[idx], rtri)
    bars.assert_fixup_distances(oh[idx], np.sqrt(d2), "holed beams grid")
    # winding-number signs of the holed mesh = pseudonormal signs of the intact one, beyond 0.01
    far = np.abs(of) > 0.01
    assert int(np.sum((np.sign(of) != np.sign(oh)) & far)) == 0
    rm_h.close()


def test_beams200_scattered_1e8(oracle, beams200):
    full, _ = beams200
    n = 100_000_000
    rng = np.random.default_rng(5)
    rm = S.ResidentMesh.from_mesh(full)
    dev = []
    for _ in range(3):
        dev.append(S.DeviceArray.from_host(rng.random(n, dtype=np.float32)))
    out = S.DeviceArray(n, np.float32)
    tri = S.DeviceArray(n, np.int32)
    rm.query_points(S.MODE_SDF, *dev, out, tri, flags=S.FLAG_FIXUP)
    q, err = rm.last_stats()
    assert err == 0
    sel = np.sort(np.random.default_rng(77).choice(n, size=1 << 16, replace=False))
    pts = [d.download()[sel] for d in dev]
    o, t = out.download()[sel], tri.download()[sel]
    _check_fixup("beams scattered", full, o, t, *pts, oracle)
