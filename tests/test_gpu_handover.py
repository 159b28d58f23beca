"""Hand-over of packet waves to the per-query kernel (one wave per query).

With the leaf budget forced to 1 every packet wave that needs more than one leaf hands its queries
over after its first, until the hand-over half of the queue (an eighth of the queries, at least
64 Ki entries, capped at the query count) is full: about half of the queries of these tests are
answered by the per-query kernel, the rest in their packet waves; budget 0 keeps every query in
its packet wave.  The two must agree:
  FP32 + FP64 fixup: bit-identical (both canonical) and equal to the oracle's indices;
  plain FP32: neither walk is rounding-safe and the packet walk keeps the first of exact FP32
  ties where the per-query walk keeps the lowest index, so distances within the plain contract
  (1e-6 of the bounding-box diagonal) and nearly all bit-equal, the same signs (pseudonormals
  give tied triangles one sign); the indices may differ on ties.
Queries beyond the queue's capacity stay in their (partly handed-over) waves.
"""
import contextlib

import numpy as np
import pytest

import ssdf_amd as S
from ssdf_amd import meshgen as mg
from ssdf_amd import workloads as W

pytestmark = pytest.mark.gpu


@contextlib.contextmanager
def heavy(leaves):
    old = S.set_heavy_leaves(leaves)
    try:
        yield
    finally:
        S.set_heavy_leaves(old)


def _same(o1, o0, t1, t0):
    assert np.array_equal(o1.view(np.uint32), o0.view(np.uint32))
    assert np.array_equal(t1, t0)


def _plain_close(m, o1, o0):
    lo, hi = m.bbox()
    eps = 1e-6 * float(np.linalg.norm(hi - lo))
    a1, a0 = np.abs(o1), np.abs(o0)
    fin = np.isfinite(a0)
    assert np.array_equal(np.isfinite(a1), fin)
    assert np.abs(a1[fin] - a0[fin]).max(initial=0.0) <= eps
    assert np.mean(a1 == a0) > 0.999
    assert np.mean(np.sign(o1) == np.sign(o0)) > 0.999


def _dev(*arrs):
    return [S.DeviceArray.from_host(a) for a in arrs]


def _grid(rm, n, lo, hi, flags, leaves, mode=S.MODE_SDF):
    N = n[0] * n[1] * n[2]
    out = S.DeviceArray(N, np.float32)
    tri = S.DeviceArray(N, np.int32)
    with heavy(leaves):
        rm.query_grid(mode, n, lo, hi, out, tri, flags=flags)
        q, err = rm.last_stats()
    assert err == 0
    return out.download(), tri.download(), q


def _points(rm, mode, x, y, z, flags, leaves, seed=None):
    n = x.size
    dx, dy, dz = _dev(x, y, z)
    out = S.DeviceArray.from_host(seed) if seed is not None else S.DeviceArray(n, np.float32)
    tri = S.DeviceArray.from_host(np.full(n, -7, np.int32))
    cx, cy, cz = (S.DeviceArray.from_host(np.zeros(n, np.float32)) for _ in range(3))
    with heavy(leaves):
        if mode == S.MODE_CLOSEST:
            rm.query_points(mode, dx, dy, dz, out, tri, cx, cy, cz, flags=flags)
        else:
            rm.query_points(mode, dx, dy, dz, out, tri, flags=flags)
        q, err = rm.last_stats()
    assert err == 0
    return out.download(), tri.download(), (cx.download(), cy.download(), cz.download()), q


@pytest.fixture(scope="module")
def torus():
    m = mg.torus(0.6, 0.2, 125, 50, sigma=0.005, seed=2)
    return m, S.ResidentMesh.from_mesh(m, with_normals=True)


def test_grid_fixup_handed_over(torus, oracle):
    m, rm = torus
    n, lo, hi = (40, 40, 40), (-1, -1, -0.6), (1, 1, 0.6)
    o0, t0, q0 = _grid(rm, n, lo, hi, S.FLAG_FIXUP, 0)
    o1, t1, q1 = _grid(rm, n, lo, hi, S.FLAG_FIXUP, 1)
    assert q1 >= 0.45 * 40 ** 3 and q0 < 100, (q0, q1)
    assert np.array_equal(o1.view(np.uint32), o0.view(np.uint32))
    assert np.array_equal(t1, t0)
    x, y, z = mg.grid_points(*n, lo, hi)
    _, rtri, _, _, _ = oracle.sdf(x, y, z, m)
    assert np.array_equal(t1, rtri)


def test_grid_plain_handed_over(torus):
    m, rm = torus
    n, lo, hi = (40, 40, 40), (-1, -1, -0.6), (1, 1, 0.6)
    o0, t0, _ = _grid(rm, n, lo, hi, 0, 0)
    o1, t1, q1 = _grid(rm, n, lo, hi, 0, 1)
    assert q1 >= 0.45 * 40 ** 3
    _plain_close(m, o1, o0)


@pytest.mark.parametrize("flags", [S.FLAG_FIXUP, 0])
@pytest.mark.parametrize("mode", [S.MODE_EDF, S.MODE_SDF, S.MODE_SQEDF, S.MODE_CLOSEST])
def test_points_modes(torus, mode, flags):
    """Scattered points (Morton-sorted: the queue carries original indices), seeded modes with
    half the seeds below the distance (those queries must stay untouched)."""
    m, rm = torus
    x, y, z = W.query_cloud(m, 30000, seed=4)
    seed = None
    if mode in (S.MODE_EDF, S.MODE_CLOSEST, S.MODE_SQEDF):
        rng = np.random.default_rng(6)
        seed = np.where(rng.random(x.size) < 0.5, np.float32(3.0), np.float32(0.01)).astype(np.float32)
    o0, t0, c0, q0 = _points(rm, mode, x, y, z, flags, 0, seed)
    o1, t1, c1, q1 = _points(rm, mode, x, y, z, flags, 1, seed)
    # seeds of 0.01 prune most of their queries at the root (no leaf, nothing to hand over)
    assert q1 >= x.size * (0.2 if seed is not None else 0.45), q1
    if flags:
        _same(o1, o0, t1, t0)
        for a, b in zip(c0, c1):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    else:
        _plain_close(m, o1, o0)
    if seed is not None:
        kept = t0 == -7  # seed already below the distance: untouched by both
        assert kept.any() and np.array_equal(t1 == -7, kept)


@pytest.mark.parametrize("flags", [S.FLAG_FIXUP, 0])
def test_queue_overflow_keeps_the_rest_in_the_wave(torus, flags):
    """100 k points against a queue of 65536 entries, half of them for hand-overs: the queries
    that find that half full finish in their packet waves, next to lanes of the same waves that
    were handed over."""
    m, rm = torus
    x, y, z = W.query_cloud(m, 100_000, seed=8)
    o0, t0, _, _ = _points(rm, S.MODE_SDF, x, y, z, flags, 0)
    o1, t1, _, q1 = _points(rm, S.MODE_SDF, x, y, z, flags, 1)
    assert 32768 <= q1 < 32768 + 100
    if flags:
        _same(o1, o0, t1, t0)
    else:
        _plain_close(m, o1, o0)


def test_bumpy_sphere_interior(oracle):
    """Coarse grid (a 4x4x4 brick spans 0.15) over a bumpy near-sphere with a budget of 256
    leaves: most waves exceed it and are handed over, some do not; the fixup result stays the
    canonical one."""
    m = mg.displaced_icosphere(80_000, seed=1)
    rm = S.ResidentMesh.from_mesh(m, with_normals=True)
    n, lo, hi = (48, 48, 48), (-1.2, -1.2, -1.2), (1.2, 1.2, 1.2)
    o0, t0, _ = _grid(rm, n, lo, hi, S.FLAG_FIXUP, 0)
    o1, t1, q1 = _grid(rm, n, lo, hi, S.FLAG_FIXUP, 256)
    assert 0 < q1 < 48 ** 3, q1
    assert np.array_equal(o1.view(np.uint32), o0.view(np.uint32)) and np.array_equal(t1, t0)
    x, y, z = mg.grid_points(*n, lo, hi)
    sel = np.random.default_rng(2).choice(x.size, 4000, replace=False)
    _, rtri, _, _, _ = oracle.sdf(x[sel], y[sel], z[sel], m)
    assert np.array_equal(t1[sel], rtri)


def test_one_triangle_mesh():
    """The root is the leaf: every wave is done after it, nothing is handed over."""
    m = mg._mesh(np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float64), np.array([[0, 1, 2]]), "tri1")
    rm = S.ResidentMesh.from_mesh(m, with_normals=True)
    n, lo, hi = (9, 9, 9), (-1, -1, -1), (1, 1, 1)
    for flags in (S.FLAG_FIXUP, 0):
        o0, t0, _ = _grid(rm, n, lo, hi, flags, 0)
        o1, t1, q1 = _grid(rm, n, lo, hi, flags, 1)
        assert q1 == 0
        assert np.array_equal(o1.view(np.uint32), o0.view(np.uint32)) and np.all(t1 == 0)


@pytest.mark.parametrize("flags", [S.FLAG_FIXUP, 0])
def test_batch_grids(flags):
    meshes = [mg.displaced_icosphere(3000, seed=1), mg.torus(0.6, 0.2, 60, 24, sigma=0.005, seed=3),
              mg.uv_sphere(0.5, 24, 12)]
    dims = np.array([(20, 20, 20), (24, 16, 12), (17, 17, 17)], np.int64)
    boxes = np.array([np.concatenate([lo - 0.1, hi + 0.1]) for lo, hi in (mm.bbox() for mm in meshes)])
    off, total = S.Batch.grid_layout(dims)
    b = S.Batch(meshes)
    res = []
    for leaves in (0, 1):
        out = S.DeviceArray(total, np.float32)
        tri = S.DeviceArray(total, np.int32)
        with heavy(leaves):
            b.query_grids(S.MODE_SDF, dims, boxes, off, out, tri, flags=flags)
            q, err = b.last_stats()
        assert err == 0
        res.append((out.download(), tri.download(), q))
    (o0, t0, _), (o1, t1, q1) = res
    assert q1 >= 0.45 * total
    if flags:
        _same(o1, o0, t1, t0)
    else:
        for i, mm in enumerate(meshes):
            sl = slice(off[i], off[i] + int(np.prod(dims[i])))
            _plain_close(mm, o1[sl], o0[sl])
