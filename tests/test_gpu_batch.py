"""GPU batches: many meshes in shared arenas, one persistent launch per call.  Every mesh's batched
result must be bit-identical to the single-mesh path (same kernels, per-mesh offsets), which the
parity tests pin to the oracle; one mesh is also checked against the oracle directly.  Captured
graphs are replayed several times, with eager work between replays, against the eager result."""
import numpy as np
import pytest

import ssdf_amd as S
from tests import bars
from ssdf_amd import meshgen as mg

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def meshes():
    ms = [mg.uv_sphere(0.5, 24, 12), mg.torus(0.6, 0.2, 80, 30, sigma=0.004, seed=3),
          mg.displaced_icosphere(3000, seed=1), mg.displaced_icosphere(12000, seed=2),
          mg.Mesh(np.zeros(0, np.float32), np.zeros(0, np.float32), np.zeros(0, np.float32),
                  np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.int32)),
          mg.displaced_icosphere(1000, seed=3)]
    return ms


def _grids(ms):
    dims, boxes = [], []
    for i, m in enumerate(ms):
        if m.ntris:
            lo, hi = m.bbox()
        else:
            lo, hi = np.array([-1.0, -1, -1]), np.array([1.0, 1, 1])
        pad = 0.1 * (hi - lo)
        dims.append((9 + 7 * i, 13 + 3 * i, 11 + 5 * i))
        boxes.append(np.concatenate([lo - pad, hi + pad]))
    return np.array(dims), np.array(boxes)


def _single_grid(m, n, box, flags):
    N = n[0] * n[1] * n[2]
    if m.ntris == 0:
        return None, None
    rm = S.ResidentMesh.from_mesh(m)
    o = S.DeviceArray(N, np.float32)
    t = S.DeviceArray(N, np.int32)
    rm.query_grid(S.MODE_SDF, tuple(int(v) for v in n), box[:3], box[3:], o, t, flags=flags)
    return o.download(), t.download()


@pytest.mark.parametrize("flags", [S.FLAG_FIXUP, 0])
def test_batch_grids_equal_single_mesh(meshes, flags):
    dims, boxes = _grids(meshes)
    off, total = S.Batch.grid_layout(dims)
    b = S.Batch(meshes)
    out = S.DeviceArray(total, np.float32)
    tri = S.DeviceArray(total, np.int32)
    out.upload(np.full(total, 123.0, np.float32))
    b.query_grids(S.MODE_SDF, dims, boxes, off, out, tri, flags=flags)
    q, err = b.last_stats()
    assert err == 0
    o, t = out.download(), tri.download()
    for i, m in enumerate(meshes):
        n = dims[i]
        N = int(n.prod())
        so, st = _single_grid(m, n, boxes[i], flags)
        if so is None:
            assert np.all(o[off[i]:off[i] + N] == 123.0)  # empty mesh: untouched
            continue
        assert np.array_equal(o[off[i]:off[i] + N], so), i
        assert np.array_equal(t[off[i]:off[i] + N], st), i


def test_batch_mesh_against_oracle(meshes, oracle):
    dims, boxes = _grids(meshes)
    off, total = S.Batch.grid_layout(dims)
    b = S.Batch(meshes)
    out = S.DeviceArray(total, np.float32)
    tri = S.DeviceArray(total, np.int32)
    b.query_grids(S.MODE_SDF, dims, boxes, off, out, tri, flags=S.FLAG_FIXUP)
    i = 3
    m, n = meshes[i], dims[i]
    x, y, z = mg.grid_points(*n, boxes[i][:3].astype(np.float32), boxes[i][3:].astype(np.float32))
    s, rtri, _, _, _ = oracle.sdf(x, y, z, m)
    N = int(n.prod())
    assert np.array_equal(tri.download()[off[i]:off[i] + N], rtri)
    lo, hi = m.bbox()
    o = out.download()[off[i]:off[i] + N]
    bars.assert_fixup_distances(o, np.abs(s), "batch mesh 3")
    assert bars.sign_mismatches_outside_band(o, s, 1e-6 * float(np.linalg.norm(hi - lo))) == 0


def test_batch_points_closest_and_edf(meshes):
    rng = np.random.default_rng(4)
    counts = [300, 1000, 65, 0, 500, 129]
    po = np.zeros(len(meshes) + 1, np.int64)
    po[1:] = np.cumsum(counts)
    P = rng.uniform(-1.2, 1.2, (int(po[-1]), 3)).astype(np.float32)
    x, y, z = (S.DeviceArray.from_host(np.ascontiguousarray(P[:, d])) for d in range(3))
    b = S.Batch(meshes)
    Np = int(po[-1])
    init = np.full(Np, 0.35, np.float32)
    od = S.DeviceArray.from_host(init)
    ot = S.DeviceArray.from_host(np.full(Np, -5, np.int32))
    cx, cy, cz = (S.DeviceArray.from_host(np.zeros(Np, np.float32)) for _ in range(3))
    b.query_points(S.MODE_CLOSEST, po, x, y, z, od, ot, cx, cy, cz, flags=S.FLAG_FIXUP)
    e = S.DeviceArray.from_host(init)
    b.query_points(S.MODE_EDF, po, x, y, z, e, None, flags=0)
    got_d, got_t, got_e = od.download(), ot.download(), e.download()
    for i, m in enumerate(meshes):
        sl = slice(int(po[i]), int(po[i + 1]))
        if m.ntris == 0 or counts[i] == 0:
            assert np.all(got_d[sl] == 0.35) and np.all(got_t[sl] == -5)
            continue
        rm = S.ResidentMesh.from_mesh(m)
        px, py, pz = (S.DeviceArray.from_host(np.ascontiguousarray(P[sl, d])) for d in range(3))
        d1 = S.DeviceArray.from_host(init[sl])
        t1 = S.DeviceArray.from_host(np.full(counts[i], -5, np.int32))
        c1 = [S.DeviceArray.from_host(np.zeros(counts[i], np.float32)) for _ in range(3)]
        rm.query_points(S.MODE_CLOSEST, px, py, pz, d1, t1, *c1, flags=S.FLAG_FIXUP)
        assert np.array_equal(got_d[sl], d1.download()) and np.array_equal(got_t[sl], t1.download())
        e1 = S.DeviceArray.from_host(init[sl])
        rm.query_points(S.MODE_EDF, px, py, pz, e1, None, flags=0)
        assert np.array_equal(got_e[sl], e1.download())


def test_batch_graph_replays_match_eager(meshes):
    dims, boxes = _grids(meshes)
    off, total = S.Batch.grid_layout(dims)
    b = S.Batch(meshes)
    st = S.Stream()
    eo = S.DeviceArray.from_host(np.zeros(total, np.float32))  # the empty mesh's slice stays untouched
    et = S.DeviceArray.from_host(np.zeros(total, np.int32))
    b.query_grids(S.MODE_SDF, dims, boxes, off, eo, et, flags=S.FLAG_FIXUP, stream=st)
    st.synchronize()
    want_o, want_t = eo.download(), et.download()
    go = S.DeviceArray(total, np.float32)
    gt = S.DeviceArray(total, np.int32)
    b.capture(True, S.MODE_SDF, dims, boxes, off, go, gt, flags=S.FLAG_FIXUP, stream=st)
    for rep in range(3):
        go.upload(np.zeros(total, np.float32))
        gt.upload(np.zeros(total, np.int32))
        b.replay(st)
        st.synchronize()
        assert np.array_equal(go.download(), want_o), rep
        assert np.array_equal(gt.download(), want_t), rep
        # eager work between replays
        b.query_grids(S.MODE_SDF, dims, boxes, off, eo, et, flags=0, stream=st)
        st.synchronize()


@pytest.mark.parametrize("flags", [S.FLAG_FIXUP, 0])
def test_batch_cost_ordered_units(meshes, flags):
    """Each query takes its work units in the order the previous query with the same plan measured
    (costliest first within each XCD range), eager and graph alike: every later query, interleaved
    with queries of another plan and with graph replays, writes every output (sentinel-filled
    before each call) bit for bit as the first one, in grid and point queries."""
    dims, boxes = _grids(meshes)
    off, total = S.Batch.grid_layout(dims)
    dims2 = dims + 3  # another plan: other units
    off2, total2 = S.Batch.grid_layout(dims2)
    b = S.Batch(meshes)
    st = S.Stream()
    out, tri = S.DeviceArray(total, np.float32), S.DeviceArray(total, np.int32)
    out2, tri2 = S.DeviceArray(total2, np.float32), S.DeviceArray(total2, np.int32)

    def query(d, bx, o, t, n, tot):
        o.upload(np.full(tot, np.nan, np.float32))
        t.upload(np.full(tot, -7, np.int32))
        b.query_grids(S.MODE_SDF, d, bx, n, o, t, flags=flags, stream=st)
        st.synchronize()
        assert b.last_stats()[1] == 0
        return o.download(), t.download()

    want = query(dims, boxes, out, tri, off, total)
    want2 = query(dims2, boxes, out2, tri2, off2, total2)
    for rep in range(3):
        got = query(dims, boxes, out, tri, off, total)
        assert np.array_equal(got[0], want[0], equal_nan=True) and np.array_equal(got[1], want[1]), rep
        if rep == 1:
            got2 = query(dims2, boxes, out2, tri2, off2, total2)
            assert np.array_equal(got2[0], want2[0], equal_nan=True) and np.array_equal(got2[1], want2[1])
    # the empty mesh's slice is the only one left at the sentinel
    e = 4
    assert np.all(np.isnan(want[0][off[e]:off[e] + int(dims[e].prod())]))
    assert np.count_nonzero(np.isnan(want[0])) == int(dims[e].prod())
    # graph replays keep an order of their own, eager queries of another plan between them
    go, gt = S.DeviceArray(total, np.float32), S.DeviceArray(total, np.int32)
    b.capture(False, S.MODE_SDF, dims, boxes, off, go, gt, flags=flags, stream=st)
    for rep in range(4):
        go.upload(np.full(total, np.nan, np.float32))
        gt.upload(np.full(total, -7, np.int32))
        b.replay(st)
        st.synchronize()
        assert np.array_equal(go.download(), want[0], equal_nan=True) and np.array_equal(gt.download(), want[1]), rep
        got2 = query(dims2, boxes, out2, tri2, off2, total2)
        assert np.array_equal(got2[0], want2[0], equal_nan=True)
    # points: the same across calls
    rng = np.random.default_rng(9)
    counts = [3000, 10000, 650, 5000, 0, 1290]  # none for the empty mesh (4)
    po = np.zeros(len(meshes) + 1, np.int64)
    po[1:] = np.cumsum(counts)
    Np = int(po[-1])
    P = rng.uniform(-1.2, 1.2, (Np, 3)).astype(np.float32)
    x, y, z = (S.DeviceArray.from_host(np.ascontiguousarray(P[:, d])) for d in range(3))
    od, ot = S.DeviceArray(Np, np.float32), S.DeviceArray(Np, np.int32)
    first = None
    for rep in range(3):
        od.upload(np.full(Np, np.nan, np.float32))
        ot.upload(np.full(Np, -7, np.int32))
        b.query_points(S.MODE_SDF, po, x, y, z, od, ot, flags=flags, stream=st)
        st.synchronize()
        got = od.download(), ot.download()
        if first is None:
            first = got
            assert not np.any(np.isnan(got[0])) and np.all(got[1] >= 0)
        assert np.array_equal(got[0], first[0]) and np.array_equal(got[1], first[1]), rep


def test_batch_autotune_and_schedule(meshes):
    dims, boxes = _grids(meshes)
    off, total = S.Batch.grid_layout(dims)
    b = S.Batch(meshes)
    out = S.DeviceArray.from_host(np.zeros(total, np.float32))  # the empty mesh's slice stays untouched
    tri = S.DeviceArray.from_host(np.zeros(total, np.int32))
    b.query_grids(S.MODE_SDF, dims, boxes, off, out, tri, flags=S.FLAG_FIXUP)
    want = out.download(), tri.download()
    bpc, ch, ms = b.autotune(S.MODE_SDF, dims, boxes, off, out, tri, flags=S.FLAG_FIXUP)
    assert bpc > 0 and ch > 0 and ms > 0
    for sched in [(1, 1), (8, 16), (0, 0)]:
        b.set_schedule(*sched)
        out.upload(np.zeros(total, np.float32))
        b.query_grids(S.MODE_SDF, dims, boxes, off, out, tri, flags=S.FLAG_FIXUP)
        assert np.array_equal(out.download(), want[0]) and np.array_equal(tri.download(), want[1])
    info = b.info()
    assert info["nselements"] == sum(m.ntris for m in meshes)
    with pytest.raises(S.SsdfError):
        b.set_schedule(99, 1)
    # the autotuner's per-mesh choices are among its candidates, or, where they did not beat the
    # default schedule, the default schedule is back in place
    scheds = [b.mesh_schedule(i) for i in range(len(meshes))]
    if all(h == -1 for _, h in scheds):
        assert all(bk == (4, 4, 4) for bk, _ in scheds) and (bpc, ch) == (8, 1)
    for brick, heavy in scheds:
        assert brick in [(4, 4, 4), (8, 4, 2), (4, 8, 2), (2, 4, 8), (8, 8, 1)] and heavy in (-1, 4096, 8192, 16384)


@pytest.mark.parametrize("seed", [1, 2])
def test_batch_per_mesh_schedules_bit_identical(meshes, seed):
    """Any per-mesh brick shape and leaf budget gives the same FP32 + fixup results as the default
    schedule, bit for bit, and the same as each mesh queried alone (budget 1 hands most waves over)."""
    dims, boxes = _grids(meshes)
    off, total = S.Batch.grid_layout(dims)
    b = S.Batch(meshes)
    out = S.DeviceArray.from_host(np.zeros(total, np.float32))
    tri = S.DeviceArray.from_host(np.zeros(total, np.int32))
    b.query_grids(S.MODE_SDF, dims, boxes, off, out, tri, flags=S.FLAG_FIXUP)
    want = out.download(), tri.download()
    rng = np.random.default_rng(seed)
    shapes = [(4, 4, 4), (8, 4, 2), (2, 4, 8), (8, 8, 1), (16, 4, 1), (1, 1, 64), (64, 1, 1), (2, 32, 1)]
    for i in range(len(meshes)):
        b.set_mesh_schedule(i, shapes[int(rng.integers(len(shapes)))], int(rng.choice([-1, 0, 1, 64, 8192])))
    out.upload(np.zeros(total, np.float32))
    b.query_grids(S.MODE_SDF, dims, boxes, off, out, tri, flags=S.FLAG_FIXUP)
    q, err = b.last_stats()
    assert err == 0
    assert np.array_equal(out.download(), want[0]) and np.array_equal(tri.download(), want[1])
    for i, m in enumerate(meshes):
        so, st = _single_grid(m, dims[i], boxes[i], S.FLAG_FIXUP)
        if so is not None:
            N = int(np.prod(dims[i]))
            assert np.array_equal(want[0][off[i]:off[i] + N], so) and np.array_equal(want[1][off[i]:off[i] + N], st)
    with pytest.raises(S.SsdfError):
        b.set_mesh_schedule(0, (4, 4, 2))
    with pytest.raises(S.SsdfError):
        b.set_mesh_schedule(0, (3, 3, 7))


def test_destroy_after_the_stream_is_gone(meshes):
    """Destroying a batch or a resident mesh does not touch the streams it was created or queried
    on: the caller may have destroyed them first (as Python's shutdown can)."""
    st = S.Stream()
    b = S.Batch(meshes[:3], stream=st)
    dims, boxes = _grids(meshes[:3])
    off, total = S.Batch.grid_layout(dims)
    out = S.DeviceArray(total, np.float32)
    b.query_grids(S.MODE_EDF, dims, boxes, off, out, None, flags=S.FLAG_FIXUP, stream=st)
    rm = S.ResidentMesh.from_mesh(meshes[1], stream=st)
    o = S.DeviceArray(8 ** 3, np.float32)
    rm.query_grid(S.MODE_SDF, (8, 8, 8), (-1, -1, -1), (1, 1, 1), o, None, flags=S.FLAG_FIXUP, stream=st)
    st.synchronize()
    S._check(S.lib().ssdf_stream_destroy(st.ptr), "ssdf_stream_destroy")
    st.ptr = None
    b.close()
    rm.close()
    # the device is still usable afterwards
    rm2 = S.ResidentMesh.from_mesh(meshes[0])
    o2 = S.DeviceArray(8 ** 3, np.float32)
    rm2.query_grid(S.MODE_SDF, (8, 8, 8), (-1, -1, -1), (1, 1, 1), o2, None, flags=S.FLAG_FIXUP)
    assert np.all(np.isfinite(o2.download()))
    rm2.close()


@pytest.mark.parametrize("flags", [S.FLAG_FIXUP, 0])
def test_batch_with_a_tree_deeper_than_64(flags):
    """A batch whose deepest tree passes 64 levels runs the two-VGPR-stack kernel for every mesh;
    each result stays bit-identical to the single-mesh path."""
    ms = [mg.deep_chain(8), mg.torus(0.6, 0.2, 60, 24, sigma=0.004, seed=5)]
    dims, boxes = _grids(ms)
    off, total = S.Batch.grid_layout(dims)
    b = S.Batch(ms)
    assert 64 < b.info()["maxdepth"] <= 128
    out = S.DeviceArray(total, np.float32)
    tri = S.DeviceArray(total, np.int32)
    b.query_grids(S.MODE_SDF, dims, boxes, off, out, tri, flags=flags)
    q, err = b.last_stats()
    assert err == 0
    o, t = out.download(), tri.download()
    for i, m in enumerate(ms):
        so, st = _single_grid(m, dims[i], boxes[i], flags)
        n = int(np.prod(dims[i]))
        assert np.array_equal(o[off[i]:off[i] + n].view(np.int32), so.view(np.int32))
        assert np.array_equal(t[off[i]:off[i] + n], st)
