"""The sdf command on the GPU against the oracle: distance to the union of several surfaces,
FP64 mode bit-exact (per-surface rounding to float as drivers/sdf.exe.cpp:370-417), warm start
from SSDF_INPUT_SDF, the SSDF_INIT_WITH_VOLUME_BOX seed, and .raw file names."""
import os
import subprocess

import numpy as np
import pytest

from ssdf_amd import meshgen as mg
from ssdf_amd import rawio
from tests import bars

pytestmark = pytest.mark.gpu


def _union_diag():
    """Bounding-box diagonal of the scene's two surfaces (the sphere and the shifted torus)."""
    a = mg.uv_sphere(0.5, 32, 16)
    b = mg.torus(0.6, 0.2, 60, 24, sigma=0.003, seed=4)
    lo = np.minimum(a.bbox()[0], b.bbox()[0] + [1.5, 0, 0])
    hi = np.maximum(a.bbox()[1], b.bbox()[1] + [1.5, 0, 0])
    return float(np.linalg.norm(hi - lo))


DIAG = _union_diag()


@pytest.fixture(scope="module")
def scene(tmp_path_factory, oracle):
    d = tmp_path_factory.mktemp("sdf")
    a = mg.uv_sphere(0.5, 32, 16)
    b = mg.torus(0.6, 0.2, 60, 24, sigma=0.003, seed=4)
    b = mg.Mesh(b.sx + np.float32(1.5), b.sy, b.sz, b.s0, b.s1, b.s2)
    rawio.write_surface(d / "a", a)
    rawio.write_surface(d / "b", b)
    rawio.write_surface(d / "b_raw", b, ext="raw")
    rng = np.random.default_rng(3)
    P = rng.uniform([-1.0, -1.0, -1.0], [2.5, 1.0, 1.0], (20000, 3)).astype(np.float32)
    x, y, z = (np.ascontiguousarray(P[:, k]) for k in range(3))
    rawio.write_points(d / "p", x, y, z)
    da = np.sqrt(oracle.brute_argmin(x, y, z, a)[0])
    db = np.sqrt(oracle.brute_argmin(x, y, z, b)[0])
    return d, (x, y, z), da, db


def _run(d, *args, **env):
    e = dict(os.environ)
    e.update({k: str(v) for k, v in env.items()})
    r = subprocess.run([str(rawio.SDF_EXE), *map(str, args)], capture_output=True, text=True, env=e, timeout=300)
    assert r.returncode == 0, r.stderr
    return r


def test_union_of_surfaces_fp32(scene):
    d, _, da, db = scene
    _run(d, d / "a", d / "b", d / "p", d / "o.raw")
    out = rawio.read_field(d / "o.raw")
    bars.plain_report(out, np.minimum(da, db), DIAG, "sdf command")
    # .raw file names give the same field
    _run(d, d / "a", d / "b_raw", d / "p", d / "o2.raw")
    assert np.array_equal(rawio.read_field(d / "o2.raw"), out)


def test_double_precision_bit_exact(scene):
    d, _, da, db = scene
    _run(d, d / "a", d / "b", d / "p", d / "o64.raw", SSDF_DOUBLE_PRECISION=1)
    f1 = np.minimum(np.float64(np.finfo(np.float32).max), da).astype(np.float32)
    want = np.minimum(f1.astype(np.float64), db).astype(np.float32)
    assert np.array_equal(rawio.read_field(d / "o64.raw"), want)


def test_warm_start_and_volume_box(scene, oracle):
    d, (x, y, z), da, db = scene
    seed = np.full(x.size, 0.05, np.float32)
    seed.tofile(d / "seed.raw")
    _run(d, d / "a", d / "p", d / "ow.raw", SSDF_INPUT_SDF=d / "seed.raw")
    bars.plain_report(rawio.read_field(d / "ow.raw"), np.minimum(np.float64(np.float32(0.05)), da), DIAG, "warm start")
    _run(d, d / "b", d / "p", d / "ov.raw", SSDF_INIT_WITH_VOLUME_BOX=1)
    box = np.abs(oracle.all_points_aabb_sdist(x, y, z, oracle.compute_aabb(x, y, z)))
    bars.plain_report(rawio.read_field(d / "ov.raw"), np.minimum(box, db), DIAG, "volume box")


def test_wrong_seed_size_fails(scene):
    d, (x, _, _), _, _ = scene
    np.zeros(x.size - 1, np.float32).tofile(d / "bad.raw")
    e = dict(os.environ, SSDF_INPUT_SDF=str(d / "bad.raw"))
    r = subprocess.run([str(rawio.SDF_EXE), str(d / "a"), str(d / "p"), str(d / "o.raw")], capture_output=True,
                       text=True, env=e, timeout=120)
    assert r.returncode == 1 and "expected" in r.stderr


def _expand_box(lo, hi, scale, margins):
    """demo/mesh_to_sdf.exe.cpp:30-54 in float32: grow by the margins, scale about the centre."""
    f = np.float32
    out_lo, out_hi = [], []
    for d in range(3):
        c = (f(lo[d]) + f(hi[d])) * f(0.5)
        h = ((f(hi[d]) - f(lo[d])) * f(0.5) + f(margins[d])) * f(scale)
        out_lo.append(f(c - h))
        out_hi.append(f(c + h))
    return out_lo, out_hi


def test_mesh_to_sdf(tmp_path, oracle):
    import math
    import ssdf_amd as S
    m = mg.uv_sphere(0.5, 32, 16)
    rawio.write_surface(tmp_path / "s", m)
    lo = [m.sx.min(), m.sy.min(), m.sz.min()]
    hi = [m.sx.max(), m.sy.max(), m.sz.max()]
    env = {k: v for k, v in os.environ.items() if not k.startswith("SSDF_")}

    def run(*args, **extra):
        r = subprocess.run([str(rawio.MESH_TO_SDF_EXE), *map(str, args)], capture_output=True, text=True,
                           env=dict(env, **{k: str(v) for k, v in extra.items()}), timeout=300)
        assert r.returncode == 0, r.stderr
        return rawio.read_grid(args[-1])

    # margins (one per axis) and a box scale; distances negated
    field, dims, box = run(tmp_path / "s", 17, 19, 21, tmp_path / "o1", SSDF_MARGIN="0.1", SSDF_MARGIN_Z="0.2",
                           SSDF_SCALE_BOX="1.5", SSDF_SCALE="-1")
    blo, bhi = _expand_box(lo, hi, 1.5, (0.1, 0.1, 0.2))
    assert dims == (17, 19, 21)
    assert [box[i] for i in range(3)] == blo and [box[3 + i] for i in range(3)] == bhi
    x, y, z = mg.grid_points(17, 19, 21, blo, bhi)
    s, _, _, _, _ = oracle.sdf(x, y, z, m)
    diag = float(np.linalg.norm(np.array(bhi, np.float64) - np.array(blo, np.float64)))
    got = -field.reshape(-1).astype(np.float64)
    bars.plain_report(got, np.abs(s), diag, "mesh_to_sdf")
    far = np.abs(s) > 1e-6 * diag
    assert np.array_equal(np.sign(got[far]), np.sign(s[far]))

    # SSDF_HMAX: the node counts follow from the spacing; the field equals sample_sdf_bvh's
    field, dims, box = run(tmp_path / "s", 0, 0, 0, tmp_path / "o2", SSDF_HMAX="0.07")
    blo, bhi = _expand_box(lo, hi, 1.0, (0.0, 0.0, 0.0))
    want_dims = tuple(int(math.ceil(float(np.float32(bhi[d] - blo[d])) / 0.07)) + 1 for d in range(3))
    assert dims == want_dims
    direct = np.empty(int(np.prod(dims)), np.float32)
    S.sample_sdf_bvh(m.s0, m.s1, m.s2, m.sx, m.sy, m.sz, *dims, *box[:3], *box[3:], direct, None, flags=0)
    assert np.array_equal(field.reshape(-1), direct)

    # SSDF_BOX_MESH: the box of another folder's points
    pts = np.array([[-1.0, -0.5, -0.25], [1.0, 0.75, 0.5], [0.2, 0.1, 0.3]], np.float32)
    rawio.write_points(tmp_path / "b", pts[:, 0], pts[:, 1], pts[:, 2])
    field, dims, box = run(tmp_path / "s", 9, 8, 7, tmp_path / "o3", SSDF_BOX_MESH=tmp_path / "b")
    assert box == tuple(np.float32(v) for v in (-1.0, -0.5, -0.25, 1.0, 0.75, 0.5))
    direct = np.empty(9 * 8 * 7, np.float32)
    S.sample_sdf_bvh(m.s0, m.s1, m.s2, m.sx, m.sy, m.sz, 9, 8, 7, *box[:3], *box[3:], direct, None, flags=0)
    assert np.array_equal(field.reshape(-1), direct)


def test_points_inside_and_sdf_sign_chain(scene, oracle):
    """drivers/points_inside.exe.cpp (surfaces concatenated, indices offset) and sdf_sign.exe.cpp
    after the sdf command: the signed distance to the union, negative inside.  On closed surfaces
    the exact winding number is 0 or 1, so every flag equals the oracle's (w >= 1/2)."""
    d, (x, y, z), da, db = scene
    a = mg.uv_sphere(0.5, 32, 16)
    b = mg.torus(0.6, 0.2, 60, 24, sigma=0.003, seed=4)
    b = mg.Mesh(b.sx + np.float32(1.5), b.sy, b.sz, b.s0, b.s1, b.s2)
    nv = a.sx.size
    union = mg.Mesh(np.concatenate([a.sx, b.sx]), np.concatenate([a.sy, b.sy]), np.concatenate([a.sz, b.sz]),
                    np.concatenate([a.s0, b.s0 + nv]), np.concatenate([a.s1, b.s1 + nv]),
                    np.concatenate([a.s2, b.s2 + nv]))
    inside_want = (oracle.gwn(x, y, z, union) >= 0.5).astype(np.uint8)
    assert 0 < inside_want.sum() < inside_want.size

    def run(exe, *args):
        r = subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True, timeout=300)
        assert r.returncode == 0, r.stderr

    run(rawio.SDF_EXE, d / "a", d / "b", d / "p", d / "u.raw")
    run(rawio.POINTS_INSIDE_EXE, d / "a", d / "b_raw", d / "p", d / "in.raw")
    inside = np.fromfile(d / "in.raw", np.uint8)
    assert np.array_equal(inside, inside_want)
    run(rawio.SDF_SIGN_EXE, d / "u.raw", d / "in.raw", d / "s.raw")
    signed = np.fromfile(d / "s.raw", np.float32)
    udf = np.minimum(da, db)
    bars.plain_report(signed, udf, DIAG, "sdf_sign chain")
    assert np.array_equal(signed < 0, (inside_want == 1) & (udf > 0))


def test_sqedf_command(scene, oracle):
    """drivers/sqedf.exe.cpp: squared distances to the union, min-accumulated per surface; FP64 mode
    exact per surface and rounded to float (sqedf.exe.cpp:383-420); SSDF_INPUT_SDF holds squared
    seeds."""
    d, (x, y, z), da, db = scene

    def run(*args, **env):
        e = {k: v for k, v in os.environ.items() if not k.startswith("SSDF_")}
        e.update({k: str(v) for k, v in env.items()})
        r = subprocess.run([str(rawio.SQEDF_EXE), *map(str, args)], capture_output=True, text=True, env=e,
                           timeout=300)
        assert r.returncode == 0, r.stderr

    d2a, d2b = da ** 2, db ** 2
    run(d / "a", d / "b_raw", d / "p", d / "q.raw")
    got = rawio.read_field(d / "q.raw").astype(np.float64)
    want = np.minimum(d2a, d2b)
    assert np.abs(got - want).max() <= 2 * 1e-6 * 4.0 * np.sqrt(want.max())
    # FP64: per surface, min(out, exact d^2) in double, then rounded to float
    run(d / "a", d / "b", d / "p", d / "q64.raw", SSDF_DOUBLE_PRECISION=1)
    f1 = np.minimum(np.float64(np.finfo(np.float32).max), oracle.brute_argmin(x, y, z, mg.uv_sphere(0.5, 32, 16))[0])
    f1 = f1.astype(np.float32)
    b = mg.torus(0.6, 0.2, 60, 24, sigma=0.003, seed=4)
    b = mg.Mesh(b.sx + np.float32(1.5), b.sy, b.sz, b.s0, b.s1, b.s2)
    f2 = np.minimum(f1.astype(np.float64), oracle.brute_argmin(x, y, z, b)[0]).astype(np.float32)
    assert np.array_equal(rawio.read_field(d / "q64.raw"), f2)
    # warm start from squared seeds
    seed = np.full(x.size, 0.01, np.float32)
    seed.tofile(d / "qseed.raw")
    run(d / "a", d / "p", d / "qw.raw", SSDF_INPUT_SDF=d / "qseed.raw")
    got = rawio.read_field(d / "qw.raw").astype(np.float64)
    assert np.abs(got - np.minimum(np.float64(np.float32(0.01)), d2a)).max() <= 2 * 1e-6 * 4.0
