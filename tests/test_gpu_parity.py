"""GPU parity: the HIP path through the C ABI against the oracle on the same seeded inputs.

Bars (from the engine's precision contract, tests/bars.py):
  FP64 path (ssdf_edf_d, FP64 meshes): distance bit-exact to the FP64 argmin, index identical;
  FP32 plain: |d - d_oracle| <= 1e-6 * bbox diagonal (relative error reported); index may differ
  only on near-ties;
  FP32 + FP64 fixup: closest-triangle index identical to the FP64 argmin (lowest index on exact
  ties), |d - d_oracle| <= 1e-6 * d_oracle, sign identical outside the 1e-6 * bbox band.
"""
import math

import numpy as np
import pytest

import ssdf_amd as S
from ssdf_amd import meshgen as mg
from tests import bars
from tests.test_oracle import brute_force_case, single_triangle_case

pytestmark = pytest.mark.gpu


def _diag(m):
    lo, hi = m.bbox()
    return float(np.linalg.norm(hi - lo))


def _dev(*arrs):
    return [S.DeviceArray.from_host(a) for a in arrs]


def _sdf_grid(m, n, lo, hi, flags, dtype=np.float32):
    rm = S.ResidentMesh.from_mesh(m, with_normals=True, dtype=dtype)
    N = n[0] * n[1] * n[2]
    out = S.DeviceArray(N, dtype)
    tri = S.DeviceArray(N, np.int32)
    rm.query_grid(S.MODE_SDF, n, lo, hi, out, tri, flags=flags)
    q, err = rm.last_stats()
    assert err == 0
    return out.download(), tri.download(), rm, q


# ----------------------------------------------------------------------------------------------
# the reference's own tests, through the C ABI
# ----------------------------------------------------------------------------------------------
def test_reference_single_triangle(gpu_lib):
    m, x, y, z = single_triangle_case()
    out = np.full(5, 1111111, np.float32)
    assert S.edf(x, y, z, m.s0, m.s1, m.s2, m.sx, m.sy, m.sz, out) == 0
    assert np.all(np.isfinite(out)) and np.all(out >= 0)
    assert np.abs(out - np.array([0, 0, 0, 1, 0])).max() < 1e-5  # tests/ssdf_test.cpp:61


def test_reference_brute_force_case(gpu_lib, oracle):
    m, x, y, z = brute_force_case()
    out = np.full(20, 1e9)
    S.edf(x, y, z, m.s0, m.s1, m.s2, m.sx, m.sy, m.sz, out)
    d2, tri, _, _ = oracle.brute_argmin(x, y, z, m)
    assert np.abs(out - np.sqrt(d2)).max() <= 1e-7  # tests/ssdf_brute_force_test.cpp:84,91
    assert np.array_equal(out, np.sqrt(d2))  # FP64 path: bit-exact


# ----------------------------------------------------------------------------------------------
# UV sphere 48x24, 64^3 grid on [-1,1]^3
# ----------------------------------------------------------------------------------------------
@pytest.fixture(scope="module")
def sphere48(oracle):
    m = mg.uv_sphere()
    x, y, z = mg.grid_points(64, 64, 64, (-1, -1, -1), (1, 1, 1))
    s, tri, sg, rg, dot = oracle.sdf(x, y, z, m)
    d2, _ = oracle.bvh_argmin(x, y, z, m)
    return m, (x, y, z), s, tri, d2


def test_sphere48_fp64_edf_bit_exact(sphere48):
    m, (x, y, z), s, tri, d2 = sphere48
    md = m.as_dtype(np.float64)
    out = np.full(x.size, 1e9)
    S.edf(x.astype(np.float64), y.astype(np.float64), z.astype(np.float64), md.s0, md.s1, md.s2, md.sx, md.sy, md.sz,
          out)
    assert np.array_equal(out, np.sqrt(d2))


def test_sphere48_fp64_closest_point_index(sphere48):
    m, (x, y, z), s, tri, d2 = sphere48
    md = m.as_dtype(np.float64)
    n = x.size
    outd = np.full(n, np.inf)
    cx, cy, cz = np.zeros(n), np.zeros(n), np.zeros(n)
    ot = np.full(n, -7, np.int32)
    S.closest_point_bvh(x.astype(np.float64), y.astype(np.float64), z.astype(np.float64), md.s0, md.s1, md.s2, md.sx,
                        md.sy, md.sz, outd, cx, cy, cz, ot)
    assert np.array_equal(ot, tri)
    assert np.array_equal(outd, np.sqrt(d2))
    # closest point lies at the returned distance
    dd = np.sqrt((cx - x) ** 2 + (cy - y) ** 2 + (cz - z) ** 2)
    assert np.abs(dd - outd).max() < 1e-12


def test_sphere48_fp32_plain_edf(sphere48):
    m, (x, y, z), s, tri, d2 = sphere48
    out = np.full(x.size, np.finfo(np.float32).max, np.float32)
    S.edf(x, y, z, m.s0, m.s1, m.s2, m.sx, m.sy, m.sz, out)
    bars.plain_report(out, np.sqrt(d2), _diag(m), "sphere48 edf")


@pytest.mark.parametrize("flags", [S.FLAG_FIXUP, S.FLAG_FIXUP | S.FLAG_SIGN_FP64])
def test_sphere48_sdf_fixup_grid(sphere48, flags):
    m, (x, y, z), s, tri, d2 = sphere48
    out, otri, rm, nq = _sdf_grid(m, (64, 64, 64), (-1, -1, -1), (1, 1, 1), flags)
    assert np.array_equal(otri, tri), f"{(otri != tri).sum()} index mismatches"
    bars.assert_fixup_distances(out, np.sqrt(d2), "sphere48 grid")
    assert bars.sign_mismatches_outside_band(out, s, 1e-6 * _diag(m)) == 0


def test_sphere48_sdf_plain_grid_sign_and_distance(sphere48):
    m, (x, y, z), s, tri, d2 = sphere48
    out, otri, rm, nq = _sdf_grid(m, (64, 64, 64), (-1, -1, -1), (1, 1, 1), 0)
    eps = 1e-6 * _diag(m)
    bars.plain_report(out, np.sqrt(d2), _diag(m), "sphere48 sdf grid")
    assert np.sum((np.sign(out) != np.sign(s)) & (np.sqrt(d2) > eps)) == 0
    assert nq == 0


def test_sphere48_points_equal_grid(sphere48):
    m, (x, y, z), s, tri, d2 = sphere48
    out_g, tri_g, rm, _ = _sdf_grid(m, (64, 64, 64), (-1, -1, -1), (1, 1, 1), S.FLAG_FIXUP)
    out_p = np.empty(x.size, np.float32)
    tri_p = np.empty(x.size, np.int32)
    S.sdf(x, y, z, m.s0, m.s1, m.s2, m.sx, m.sy, m.sz, out_p, tri_p, flags=S.FLAG_FIXUP)
    assert np.array_equal(out_p, out_g) and np.array_equal(tri_p, tri_g)


def test_gpu_brute_equals_oracle(oracle):
    m = mg.uv_sphere(0.5, 24, 12)
    x, y, z = mg.grid_points(30, 30, 30, (-1, -1, -1), (1, 1, 1))
    rm = S.ResidentMesh.from_mesh(m, with_normals=False)
    dx, dy, dz = _dev(x, y, z)
    d2 = S.DeviceArray(x.size, np.float64)
    t = S.DeviceArray(x.size, np.int32)
    rm.brute_points(dx, dy, dz, d2, t)
    rd2, rtri, _, _ = oracle.brute_argmin(x, y, z, m)
    assert np.array_equal(d2.download(), rd2) and np.array_equal(t.download(), rtri)


def test_lbvh_depth_matches_canonical_replay(oracle):
    m = mg.torus(0.6, 0.2, 100, 40, sigma=0.005, seed=9)
    x, y, z = mg.grid_points(8, 8, 8, (-1, -1, -1), (1, 1, 1))
    _, _, md, _ = oracle.lbvh_counts(x, y, z, m)
    rm = S.ResidentMesh.from_mesh(m, with_normals=True)
    info = rm.info()
    assert info["maxdepth"] == md
    assert info["open_edges"] == 0 and info["flipped_edges"] == 0


@pytest.mark.parametrize("nu,nv", [(4, 3), (100, 40), (250, 200), (1000, 400)])
def test_lbvh_leaf_order_is_the_stable_morton_sort(oracle, nu, nv):
    """The build's radix sort against the oracle's stable sort by Morton code: 24 triangles (one
    partial tile), 8 k, 100 k (tiles with the scan folded into hist and scatter) and 800 k (more
    tiles than that: the separate row-scan kernel).  The pseudonormal pass sorts 3 T edge keys with
    the same sort; a misordered edge sort splits shared edges, which the open-edge count shows."""
    m = mg.torus(0.6, 0.2, nu, nv, sigma=0.005, seed=17)
    rm = S.ResidentMesh.from_mesh(m, with_normals=True)
    got = S.DeviceArray(m.s0.size, np.int32)
    rm.leaf_order(got)
    assert np.array_equal(got.download(), oracle.lbvh_order(m))
    info = rm.info()
    assert info["open_edges"] == 0 and info["flipped_edges"] == 0


@pytest.mark.parametrize("ntri", [1, 2, 4095, 4096, 4097, 786432, 786433])
def test_lbvh_leaf_order_at_tile_boundaries(oracle, ntri):
    """Random triangle soups of exactly these sizes around the sort's 4096-key tiles and its
    192-tile limit for the folded scan (786,432 keys: folded; one more: the row-scan kernel)."""
    rng = np.random.default_rng(ntri)
    v = rng.uniform(-1.0, 1.0, (3 * ntri, 3)).astype(np.float32)
    f = np.arange(3 * ntri, dtype=np.int32).reshape(ntri, 3)
    m = mg.Mesh(*(np.ascontiguousarray(v[:, d]) for d in range(3)), *(np.ascontiguousarray(f[:, k]) for k in range(3)))
    rm = S.ResidentMesh.from_mesh(m, with_normals=False)
    got = S.DeviceArray(ntri, np.int32)
    rm.leaf_order(got)
    assert np.array_equal(got.download(), oracle.lbvh_order(m))


def test_deep_lbvh_two_vgpr_stack(oracle):
    """A tree deeper than 64 levels takes the traversal whose stack spans two VGPRs; every wave's
    first descent (bound +inf) pushes at every level, so entries past 64 are written and popped.
    Points mode (uniform, and clustered at the chain's deep end) and grid mode, FP32 + fixup and
    plain, against the FP64 brute force."""
    m = mg.deep_chain(8)
    rm = S.ResidentMesh.from_mesh(m, with_normals=True)
    assert 64 < rm.info()["maxdepth"] <= 128
    diag = _diag(m)
    rng = np.random.default_rng(7)
    pts = np.concatenate([rng.uniform(-0.2, 1.2, (6000, 3)), rng.uniform(-1e-5, 1e-5, (3000, 3))]).astype(np.float32)
    x, y, z = (np.ascontiguousarray(pts[:, i]) for i in range(3))
    d2, tri, _, _ = oracle.brute_argmin(x, y, z, m)
    dx, dy, dz = _dev(x, y, z)
    for flags in (S.FLAG_FIXUP, 0):
        o = S.DeviceArray(x.size, np.float32)
        t = S.DeviceArray(x.size, np.int32)
        o.upload(np.full(x.size, np.inf, np.float32))
        rm.query_points(S.MODE_EDF, dx, dy, dz, o, t, flags=flags)
        q, err = rm.last_stats()
        assert err == 0
        got, gt = o.download(), t.download()
        if flags:
            assert np.array_equal(gt, tri)
            bars.assert_fixup_distances(got, np.sqrt(d2), "deep chain points")
        else:
            bars.plain_report(got, np.sqrt(d2), diag, "deep chain points")
    n = (21, 19, 23)
    gx, gy, gz = mg.grid_points(*n, (-0.1, -0.1, -0.1), (1.1, 1.1, 1.1))
    gd2, gtri, _, _ = oracle.brute_argmin(gx, gy, gz, m)
    o = S.DeviceArray(gx.size, np.float32)
    t = S.DeviceArray(gx.size, np.int32)
    rm.query_grid(S.MODE_SDF, n, (-0.1, -0.1, -0.1), (1.1, 1.1, 1.1), o, t, flags=S.FLAG_FIXUP)
    q, err = rm.last_stats()
    assert err == 0
    assert np.array_equal(t.download(), gtri)
    bars.assert_fixup_distances(o.download(), np.sqrt(gd2), "deep chain grid")


# ----------------------------------------------------------------------------------------------
# noisy torus (the 200 k benchmark torus at test size), and the staircase sign fixture
# ----------------------------------------------------------------------------------------------
def test_noisy_torus_fixup_parity(oracle):
    m = mg.torus(0.6, 0.2, 125, 50, sigma=0.005, seed=2)
    n = (48, 48, 40)
    lo, hi = (-1, -1, -0.5), (1, 1, 0.5)
    x, y, z = mg.grid_points(*n, lo, hi)
    s, tri, _, _, _ = oracle.sdf(x, y, z, m)
    out, otri, rm, nq = _sdf_grid(m, n, lo, hi, S.FLAG_FIXUP)
    assert np.array_equal(otri, tri), f"{(otri != tri).sum()} index mismatches"
    bars.assert_fixup_distances(out, np.abs(s), "noisy torus")
    assert bars.sign_mismatches_outside_band(out, s, 1e-6 * _diag(m)) == 0


def test_staircase_sample_sdf(gpu_lib):
    """tests/sample_sdf_bvh_test.cpp through ssdf_sample_sdf_f: zero wrong signs, |err| < 1e-5."""
    m, faces = mg.staircase()
    N, H = 4, 0.25
    n = 5 * (N + 2) + 1
    lo, hi = np.float32(-H), np.float32((N + 1) * H)
    out = np.empty(n ** 3, np.float32)
    S.sample_sdf_bvh(m.s0, m.s1, m.s2, m.sx, m.sy, m.sz, n, n, n, lo, lo, lo, hi, hi, hi, out)
    hh = (float(hi) - float(lo)) / (n - 1)
    g = np.float64(lo) + np.arange(n) * hh
    Z, Y, X = np.meshgrid(g, g, g, indexing="ij")
    P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    d2 = np.full(P.shape[0], np.inf)
    for axis, sign, flo, fhi in faces:
        c = np.clip(P, flo, fhi)
        d2 = np.minimum(d2, ((P - c) ** 2).sum(1))
    d = np.sqrt(d2)
    off = d >= 1e-6
    ci, cj, ck = (np.floor(P[:, k] / H).astype(int) for k in range(3))
    inside = (ci >= 0) & (cj >= 0) & (ck >= 0) & (ci < N) & (cj < N) & (ck < N) & (ci <= cj) & (ck >= 1) & (ck <= 2)
    assert np.sum(((out < 0) != inside) & off) == 0
    assert np.abs(np.abs(out) - d)[off].max() < 1e-5


# ----------------------------------------------------------------------------------------------
# semantics and edge cases
# ----------------------------------------------------------------------------------------------
def test_empty_mesh_leaves_out_untouched(gpu_lib):
    x = np.zeros(10, np.float32)
    out = np.full(10, 3.5, np.float32)
    e = np.zeros(0, np.int32)
    S.edf(x, x, x, e, e, e, np.zeros(0, np.float32), np.zeros(0, np.float32), np.zeros(0, np.float32), out)
    assert np.all(out == 3.5)


def test_min_accumulate_and_strict_improvement(gpu_lib, oracle):
    m = mg.uv_sphere(0.5, 16, 8)
    x, y, z = mg.grid_points(9, 9, 9, (-1, -1, -1), (1, 1, 1))
    d2, tri, _, _ = oracle.brute_argmin(x, y, z, m)
    cap = np.float32(0.3)
    out = np.full(x.size, cap, np.float32)
    S.edf(x, y, z, m.s0, m.s1, m.s2, m.sx, m.sy, m.sz, out)
    exp = np.minimum(np.sqrt(d2), cap)
    assert np.abs(out - exp).max() <= 1e-6
    # closest point: outputs untouched where the caller's distance is already smaller
    outd = np.full(x.size, cap, np.float32)
    cx = np.full(x.size, 9, np.float32)
    cy, cz = cx.copy(), cx.copy()
    ot = np.full(x.size, -5, np.int32)
    S.closest_point_bvh(x, y, z, m.s0, m.s1, m.s2, m.sx, m.sy, m.sz, outd, cx, cy, cz, ot, flags=S.FLAG_FIXUP)
    far = np.sqrt(d2) > cap * 1.001
    assert np.all(ot[far] == -5) and np.all(cx[far] == 9) and np.all(outd[far] == cap)
    near = np.sqrt(d2) < cap * 0.999
    assert np.array_equal(ot[near], tri[near])


@pytest.mark.parametrize("npts", [1, 63, 64, 65, 1000])
def test_ragged_point_counts(gpu_lib, oracle, npts):
    m = mg.icosphere(6)
    rng = np.random.default_rng(npts)
    x, y, z = (rng.uniform(-1.5, 1.5, npts).astype(np.float32) for _ in range(3))
    out = np.empty(npts, np.float32)
    ot = np.empty(npts, np.int32)
    S.sdf(x, y, z, m.s0, m.s1, m.s2, m.sx, m.sy, m.sz, out, ot, flags=S.FLAG_FIXUP)
    s, tri, _, _, _ = oracle.sdf(x, y, z, m, use_bvh=False)
    assert np.array_equal(ot, tri)
    bars.assert_fixup_distances(out, np.abs(s), f"{npts} points")
    assert np.array_equal(np.sign(out), np.sign(s))


@pytest.mark.parametrize("n", [(1, 1, 1), (5, 3, 7), (17, 1, 9), (33, 34, 35)])
def test_ragged_grids(gpu_lib, oracle, n):
    m = mg.uv_sphere(0.5, 20, 10)
    lo, hi = (-0.9, -0.8, -0.7), (0.9, 0.8, 0.7)
    out = np.empty(n[0] * n[1] * n[2], np.float32)
    ot = np.empty_like(out, dtype=np.int32)
    S.sample_sdf_bvh(m.s0, m.s1, m.s2, m.sx, m.sy, m.sz, *n, *lo, *hi, out, ot, flags=S.FLAG_FIXUP)
    x, y, z = mg.grid_points(*n, lo, hi)
    s, tri, _, _, _ = oracle.sdf(x, y, z, m, use_bvh=False)
    assert np.array_equal(ot, tri)
    bars.assert_fixup_distances(out, np.abs(s), f"grid {n}")
    assert bars.sign_mismatches_outside_band(out, s, 1e-6 * _diag(m)) == 0


def test_single_and_degenerate_triangles(gpu_lib, oracle):
    # one triangle (the LBVH root is a leaf)
    m, x, y, z = single_triangle_case()
    out = np.full(5, 1e30, np.float32)
    S.edf(x, y, z, m.s0, m.s1, m.s2, m.sx, m.sy, m.sz, out)
    assert np.abs(out - np.array([0, 0, 0, 1, 0])).max() < 1e-6
    # a soup with zero-area triangles and many coincident centroids (equal Morton keys)
    rng = np.random.default_rng(4)
    v = rng.uniform(-1, 1, (300, 3)).astype(np.float32)
    v = np.concatenate([v, np.repeat(v[:10], 3, axis=0), np.zeros((60, 3), np.float32)])
    t = np.arange(v.shape[0] // 3, dtype=np.int32)
    soup = mg.Mesh(v[:, 0].copy(), v[:, 1].copy(), v[:, 2].copy(), 3 * t, 3 * t + 1, 3 * t + 2)
    px, py, pz = (rng.uniform(-1.2, 1.2, 4000).astype(np.float32) for _ in range(3))
    d2, tri, _, _ = oracle.brute_argmin(px, py, pz, soup)
    sd = soup.as_dtype(np.float64)
    out64 = np.full(4000, np.inf)
    S.edf(px.astype(np.float64), py.astype(np.float64), pz.astype(np.float64), sd.s0, sd.s1, sd.s2, sd.sx, sd.sy,
          sd.sz, out64)
    assert np.array_equal(out64, np.sqrt(d2))
    out32 = np.full(4000, np.inf, np.float32)
    S.edf(px, py, pz, soup.s0, soup.s1, soup.s2, soup.sx, soup.sy, soup.sz, out32)
    bars.plain_report(out32, np.sqrt(d2), 3.5, "degenerate soup")


def test_resident_mesh_repeatable(gpu_lib):
    m = mg.torus(0.6, 0.2, 60, 24, sigma=0.004, seed=1)
    rm = S.ResidentMesh.from_mesh(m)
    n = (20, 20, 12)
    outs = []
    for _ in range(3):
        o = S.DeviceArray(np.prod(n), np.float32)
        t = S.DeviceArray(np.prod(n), np.int32)
        rm.query_grid(S.MODE_SDF, n, (-1, -1, -0.4), (1, 1, 0.4), o, t, flags=S.FLAG_FIXUP)
        outs.append((o.download(), t.download()))
    for o, t in outs[1:]:
        assert np.array_equal(o, outs[0][0]) and np.array_equal(t, outs[0][1])


def test_errors_are_reported(gpu_lib):
    m = mg.uv_sphere(0.5, 8, 4)
    with pytest.raises(TypeError):
        S.edf(np.zeros(2, np.float32), np.zeros(2, np.float32), np.zeros(2, np.float64), m.s0, m.s1, m.s2, m.sx, m.sy,
              m.sz, np.zeros(2, np.float32))
    rm = S.ResidentMesh.from_mesh(m, with_normals=False)
    o = S.DeviceArray(8, np.float32)
    with pytest.raises(S.SsdfError):
        rm.query_grid(S.MODE_SDF, (2, 2, 2), (-1, -1, -1), (1, 1, 1), o)  # signed needs normals


def test_moving_mesh_update_vertices(oracle):
    """A resident handle rebuilt in place after its vertices move answers for the moved mesh."""
    m0 = mg.torus(0.6, 0.2, 80, 32, sigma=0.004, seed=3)
    m1 = mg.torus(0.6, 0.2, 80, 32, sigma=0.004, seed=4)  # same topology, other vertices
    rm = S.ResidentMesh.from_mesh(m0)
    n, lo, hi = (24, 24, 12), (-1, -1, -0.4), (1, 1, 0.4)
    x, y, z = mg.grid_points(*n, lo, hi)
    for m in (m1, m0, m1):
        rm.update_vertices(m.sx, m.sy, m.sz)
        o = S.DeviceArray(x.size, np.float32)
        t = S.DeviceArray(x.size, np.int32)
        rm.query_grid(S.MODE_SDF, n, lo, hi, o, t, flags=S.FLAG_FIXUP)
        s, tri, _, _, _ = oracle.sdf(x, y, z, m)
        assert np.array_equal(t.download(), tri)
        bars.assert_fixup_distances(o.download(), np.abs(s), "moved mesh")
    assert rm.info()["open_edges"] == 0


def _odd_topology(seed):
    """A noisy torus with 2% of its faces deleted (open edges), one face repeated (three faces on its
    edges: non-manifold) and one face reversed (flipped edges)."""
    m = mg.torus(0.6, 0.2, 60, 24, sigma=0.004, seed=seed)
    f = np.stack([m.s0, m.s1, m.s2], 1)
    rng = np.random.default_rng(seed)
    keep = rng.uniform(size=len(f)) >= 0.02
    f = f[keep]
    f = np.concatenate([f, f[5:6]])            # repeated face
    f[17] = f[17][::-1]                        # reversed face
    v = np.stack([m.sx, m.sy, m.sz], 1)
    return mg._mesh(v.astype(np.float64), f, "odd_torus")


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_rebuild_normals_equal_a_fresh_build(dtype):
    """A handle sorts its corners by vertex and by edge once (the connectivity never changes) and
    reuses them when it rebuilds after its vertices move: every rebuilt handle answers SDF queries
    bit for bit like a handle built fresh at the same positions, with the same open and flipped
    edge counts, on a mesh with holes, a non-manifold edge and a reversed face."""
    m0 = _odd_topology(1)
    rng = np.random.default_rng(2)
    n, lo, hi = (20, 20, 10), (-1, -1, -0.4), (1, 1, 0.4)
    N = n[0] * n[1] * n[2]
    m0 = m0.as_dtype(dtype)
    rm = S.ResidentMesh.from_mesh(m0, dtype=dtype)
    for step in range(3):
        jit = (0.004 * rng.normal(size=(m0.sx.size, 3))).astype(dtype)
        sx, sy, sz = (np.ascontiguousarray(a + jit[:, d]) for d, a in enumerate((m0.sx, m0.sy, m0.sz)))
        rm.update_vertices(sx, sy, sz)
        moved = mg.Mesh(sx, sy, sz, m0.s0, m0.s1, m0.s2, "moved")
        fresh = S.ResidentMesh.from_mesh(moved, dtype=dtype)
        res = []
        for h in (rm, fresh):
            o = S.DeviceArray(N, dtype)
            t = S.DeviceArray(N, np.int32)
            h.query_grid(S.MODE_SDF, n, lo, hi, o, t, flags=S.FLAG_FIXUP)
            res.append((o.download(), t.download(), h.info()["open_edges"], h.info()["flipped_edges"]))
        assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1]), step
        assert res[0][2:] == res[1][2:] and res[0][2] > 0 and res[0][3] > 0, res[0][2:]
        fresh.close()


def test_fixup_queue_many_ties(oracle):
    """Points on the axis of a UV sphere are equidistant from every triangle of a latitude ring (48
    exact ties, beyond the 24-entry candidate list): they take the fixup queue, whose FP64 search must
    return the oracle's canonical (lowest) index."""
    m = mg.uv_sphere(0.5, 48, 24)
    rng = np.random.default_rng(5)
    zs = np.linspace(-0.3, 0.3, 61)
    P = np.concatenate([np.stack([np.zeros_like(zs), np.zeros_like(zs), zs], 1),
                        rng.normal(0, 1e-3, (6000, 3))]).astype(np.float32)
    x, y, z = (np.ascontiguousarray(P[:, d]) for d in range(3))
    rm = S.ResidentMesh.from_mesh(m)
    dx, dy, dz = _dev(x, y, z)
    out = S.DeviceArray(x.size, np.float32)
    tri = S.DeviceArray(x.size, np.int32)
    rm.query_points(S.MODE_SDF, dx, dy, dz, out, tri, flags=S.FLAG_FIXUP)
    q, err = rm.last_stats()
    assert err == 0 and q > 61, q
    s, rtri, _, _, _ = oracle.sdf(x, y, z, m)
    assert np.array_equal(tri.download(), rtri)
    bars.assert_fixup_distances(out.download(), np.abs(s), "axis ties")
    # grid form through the same queue: a grid whose middle node is the centre
    n = (9, 9, 61)
    out_g = S.DeviceArray(9 * 9 * 61, np.float32)
    tri_g = S.DeviceArray(9 * 9 * 61, np.int32)
    rm.query_grid(S.MODE_SDF, n, (-0.01, -0.01, -0.3), (0.01, 0.01, 0.3), out_g, tri_g, flags=S.FLAG_FIXUP)
    q, err = rm.last_stats()
    assert err == 0 and q > 0
    gx, gy, gz = mg.grid_points(*n, (-0.01, -0.01, -0.3), (0.01, 0.01, 0.3))
    s, rtri, _, _, _ = oracle.sdf(gx, gy, gz, m)
    assert np.array_equal(tri_g.download(), rtri)
    bars.assert_fixup_distances(out_g.download(), np.abs(s), "axis ties grid")


def test_sample_sqedf(oracle):
    """sample_sqedf_bvh (serial/bvh/bvh.cpp:1575-1639): out holds squared distances in/out; FP64 is the
    exact FP64 argmin's d2, FP32 + fixup within 1e-6 relative of it; seeds below d^2 stay."""
    m = mg.torus(0.6, 0.2, 60, 24, sigma=0.004, seed=8)
    n = (21, 17, 13)
    lo, hi = (-1.0, -1.0, -0.5), (1.0, 1.0, 0.5)
    x, y, z = mg.grid_points(*n, lo, hi)
    d2, _, _, _ = oracle.brute_argmin(x, y, z, m)
    seed = np.where(np.arange(x.size) % 5 == 0, 0.001, np.finfo(np.float32).max)
    md = m.as_dtype(np.float64)
    gx, gy, gz = mg.grid_points(*n, lo, hi, dtype=np.float64)
    d2d, _, _, _ = oracle.brute_argmin(gx, gy, gz, md)
    out = seed.astype(np.float64).copy()
    S.sample_sqedf_bvh(md.s0, md.s1, md.s2, md.sx, md.sy, md.sz, *n, *lo, *hi, out)
    assert np.array_equal(out, np.minimum(seed, d2d))
    outf = seed.astype(np.float32).copy()
    S.sample_sqedf_bvh(m.s0, m.s1, m.s2, m.sx, m.sy, m.sz, *n, *(np.float32(v) for v in lo),
                       *(np.float32(v) for v in hi), outf, flags=S.FLAG_FIXUP)
    want = np.minimum(seed, d2)
    assert np.all(np.abs(outf.astype(np.float64) - want) <= bars.REL * want + bars.FLOOR)
    stay = (seed < 1) & (d2 > 0.0011)  # seeds clearly below d^2 are kept bit for bit
    assert stay.sum() > 100 and np.all(outf[stay] == np.float32(0.001))


def _seg_d(p, u, v):
    e = v - u
    t = np.clip((p - u) @ e / (e @ e), 0.0, 1.0)
    return float(np.linalg.norm(p - u - t * e))


def _sliver_traps(oracle, ntraps, seed):
    """A mesh of `ntraps` spread-out traps and the query points that spring them.  A trap is a
    needle-shaped sliver (0.1 to 0.5 long, height 1e-8 to 1e-10 of that, flagged degenerate) on whose face the
    reference's FP64 routine overestimates the distance of a point p by more than five times the
    fixup band, plus a small triangle at a distance from p between the sliver's true distance (that
    of its nearest edge, to within its height) and that FP64 value: the FP64 argmin (the oracle's
    brute force) is the small triangle, while the FP32 search finds the sliver first."""
    rng = np.random.default_rng(seed)
    verts, tris, pts = [], [], []
    while len(pts) < ntraps:
        q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
        length = 10 ** rng.uniform(-1, -0.3)
        k = len(pts)
        origin = np.array([k % 4, k // 4 % 4, k // 16], np.float64) - 1.5  # 1 apart: traps never meet
        shape = np.array([[0, 0, 0], [1, 0, 0], [0.02 * rng.uniform(), 10.0 ** -rng.uniform(8, 10), 0]]) * length
        a, b, c = (np.asarray(origin + q @ v, np.float32) for v in shape)
        af, bf, cf = (v.astype(np.float64) for v in (a, b, c))
        n = np.cross(bf - af, cf - af)
        if not np.linalg.norm(n) > 0:
            continue  # collinear once rounded to FP32
        n /= np.linalg.norm(n)
        p = ((af + bf + cf) / 3 + 1e-4 * length * n).astype(np.float32)
        pf = p.astype(np.float64)
        d64 = np.sqrt(oracle.pt_tri_d2(pf, af, bf, cf))
        true = min(_seg_d(pf, af, bf), _seg_d(pf, bf, cf), _seg_d(pf, cf, af))
        if not 2e-5 <= d64 - true <= 1e-3:
            continue
        r = 0.5 * (d64 + true)  # the small triangle's distance from p
        dirn = n if rng.random() < 0.5 else np.cross(n, bf - af) / np.linalg.norm(bf - af)
        centre = pf + (r + 1e-7) * dirn
        w1 = np.cross(dirn, [1.0, 0, 0] if abs(dirn[0]) < 0.9 else [0, 1.0, 0])
        w1 /= np.linalg.norm(w1)
        w2 = np.cross(dirn, w1)
        base = len(verts)
        verts += [a, b, c]
        verts += [np.asarray(centre + 1e-7 * w, np.float32) for w in (w1, -0.5 * w1 + 0.87 * w2, -0.5 * w1 - 0.87 * w2)]
        tris += [(base, base + 1, base + 2), (base + 3, base + 4, base + 5)]
        pts.append(p)
    V = np.array(verts, np.float32)
    F = np.array(tris, np.int32)
    m = mg.Mesh(*(np.ascontiguousarray(V[:, k]) for k in range(3)), *(np.ascontiguousarray(F[:, k]) for k in range(3)))
    P = np.array(pts, np.float32)
    return m, [np.ascontiguousarray(P[:, k]) for k in range(3)]


@pytest.mark.parametrize("leaves", [0, 1])
def test_fixup_slivers_follow_the_fp64_routine(oracle, leaves):
    """Where the reference's FP64 routine overestimates a sliver's distance beyond the band, the fixup
    still returns the FP64 argmin: the query is searched again in FP64 without a bound (in the packet
    epilogue, and for waves handed over past their leaf budget)."""
    m, (x, y, z) = _sliver_traps(oracle, 40, seed=3)
    d2, tri, _, _ = oracle.brute_argmin(x, y, z, m)
    assert np.all(tri % 2 == 1)  # the small triangles, not the slivers
    rm = S.ResidentMesh.from_mesh(m, with_normals=False)
    dx, dy, dz = _dev(x, y, z)
    out = S.DeviceArray.from_host(np.full(x.size, np.inf, np.float32))
    t = S.DeviceArray.from_host(np.full(x.size, -1, np.int32))
    prev = S.set_heavy_leaves(leaves)
    try:
        rm.query_points(S.MODE_EDF, dx, dy, dz, out, t, flags=S.FLAG_FIXUP)
        q, err = rm.last_stats()
    finally:
        S.set_heavy_leaves(prev)
    assert err == 0 and q >= x.size
    assert np.array_equal(t.download(), tri)
    bars.assert_fixup_distances(out.download(), np.sqrt(d2), "sliver traps")


@pytest.mark.parametrize("brick", [(8, 4, 2), (2, 4, 8), (8, 8, 1), (1, 64, 1), (64, 1, 1)])
def test_mesh_brick_schedules_bit_identical(brick):
    """A handle's grid waves of any 64-node brick give the same FP32 + fixup results, bit for bit, as
    the default 4 x 4 x 4 bricks, on a ragged grid (partial bricks on every face)."""
    m = mg.torus(0.6, 0.2, 80, 32, sigma=0.004, seed=3)
    rm = S.ResidentMesh.from_mesh(m)
    n, lo, hi = (37, 29, 21), (-1, -1, -0.45), (1, 1, 0.45)
    N = n[0] * n[1] * n[2]
    res = []
    for sched in ((4, 4, 4), brick):
        rm.set_schedule(sched)
        o = S.DeviceArray.from_host(np.full(N, 7.0, np.float32))
        t = S.DeviceArray.from_host(np.full(N, -9, np.int32))
        rm.query_grid(S.MODE_SDF, n, lo, hi, o, t, flags=S.FLAG_FIXUP)
        q, err = rm.last_stats()
        assert err == 0
        res.append((o.download(), t.download()))
    assert np.array_equal(res[0][0], res[1][0]) and np.array_equal(res[0][1], res[1][1])
    assert np.all(res[1][1] >= 0)  # every node written


@pytest.mark.parametrize("mode", [S.MODE_SDF, S.MODE_EDF])
def test_persistent_launch_equals_small_launches(mode):
    """A point query of more waves than the GPU holds at once (over 2048 blocks of 4 waves, i.e.
    more than 524,288 points) runs as one resident set that takes its waves from per-XCD queues.
    FP32 + fixup results are canonical, so they equal, bit for bit, the same points answered in
    chunks small enough for one block per 4 waves (each chunk Morton-sorted on its own, so its
    waves hold other points).  A bumpy sphere with points inside it: some waves hand their queries
    over to the per-query kernel in both runs (a leaf budget of 256 leaves)."""
    prev = S.set_heavy_leaves(256)
    try:
        _persistent_vs_chunks(mode)
    finally:
        S.set_heavy_leaves(prev)


def _persistent_vs_chunks(mode):
    m = mg.displaced_icosphere(20000, seed=5)
    rm = S.ResidentMesh.from_mesh(m)
    rng = np.random.default_rng(11)
    n = 600_000
    P = rng.uniform(-1.25, 1.25, (n, 3)).astype(np.float32)
    x, y, z = _dev(*(np.ascontiguousarray(P[:, d]) for d in range(3)))
    init = np.full(n, 0.8, np.float32)  # EDF min-accumulates into the seed
    o = S.DeviceArray.from_host(init)
    t = S.DeviceArray.from_host(np.full(n, -9, np.int32))
    rm.query_points(mode, x, y, z, o, t, flags=S.FLAG_FIXUP)
    q_all, err = rm.last_stats()
    assert err == 0
    want_o, want_t = o.download(), t.download()
    got_o, got_t = np.empty_like(want_o), np.empty_like(want_t)
    q_chunks = 0
    for c0 in range(0, n, 60_000):
        sl = slice(c0, min(n, c0 + 60_000))
        cx, cy, cz = _dev(*(np.ascontiguousarray(P[sl, d]) for d in range(3)))
        co = S.DeviceArray.from_host(init[sl])
        ct = S.DeviceArray.from_host(np.full(sl.stop - sl.start, -9, np.int32))
        rm.query_points(mode, cx, cy, cz, co, ct, flags=S.FLAG_FIXUP)
        q, err = rm.last_stats()
        assert err == 0
        q_chunks += q
        got_o[sl], got_t[sl] = co.download(), ct.download()
    assert np.array_equal(want_o, got_o) and np.array_equal(want_t, got_t)
    assert q_all > 0 and q_chunks > 0  # the per-query kernel ran in both
    if mode == S.MODE_SDF:
        assert np.all(want_t >= 0)


def test_cell_validate_environment(capfd, monkeypatch):
    """SSDF_CELL_VALIDATE=1 (serial/ssdf.hpp:991-1010): the C symbols check their result against the
    exact FP64 argmin of the same inputs, print the maximum and the norm of the relative differences
    to stderr when the maximum exceeds 1e-5, and return the engine's own result."""
    m = mg.uv_sphere(0.5, 24, 12)
    v = np.stack([m.sx, m.sy, m.sz], 1).astype(np.float64)
    cen = ((v[m.s0] + v[m.s1] + v[m.s2]) / 3).astype(np.float32)  # on the faces: FP32 error dominates
    x, y, z = (np.ascontiguousarray(cen[:, d]) for d in range(3))
    base = np.full(x.size, 1e9, np.float32)
    S.edf(x, y, z, m.s0, m.s1, m.s2, m.sx, m.sy, m.sz, base)
    capfd.readouterr()
    monkeypatch.setenv("SSDF_CELL_VALIDATE", "1")
    out = np.full(x.size, 1e9, np.float32)
    S.edf(x, y, z, m.s0, m.s1, m.s2, m.sx, m.sy, m.sz, out)
    err = capfd.readouterr().err
    assert np.array_equal(out, base)
    assert "EDF cell-list validation max rel diff:" in err and "EDF cell-list validation rel diff norm:" in err
    # far from the surface the FP32 result agrees to 1e-5: nothing printed
    far = [np.ascontiguousarray(a * np.float32(3.0)) for a in (x, y, z)]
    out = np.full(x.size, 1e9, np.float32)
    S.edf(*far, m.s0, m.s1, m.s2, m.sx, m.sy, m.sz, out)
    assert "validation" not in capfd.readouterr().err
    # FP64 inputs: the exact path checked against itself, silent and unchanged
    md = m.as_dtype(np.float64)
    o64 = np.full(x.size, 1e9)
    S.edf(x.astype(np.float64), y.astype(np.float64), z.astype(np.float64), md.s0, md.s1, md.s2, md.sx, md.sy, md.sz,
          o64)
    assert "validation" not in capfd.readouterr().err
