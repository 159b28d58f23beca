"""GPU parity of the driver's seeding helpers (serial/edf/edf.hpp:500-557, drivers/sdf.exe.cpp:317-327)
against the oracle: bounds and signed box distances bit-exact in both precisions."""
import numpy as np
import pytest

import ssdf_amd as S

pytestmark = pytest.mark.gpu


def _cloud(n, dt, seed):
    rng = np.random.default_rng(seed)
    p = rng.normal(size=(3, n)) * np.array([[0.5], [2.0], [7.0]]) + np.array([[1.0], [-3.0], [0.25]])
    return [np.ascontiguousarray(v, dtype=dt) for v in p]


@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("n", [1, 63, 64, 257, 100003, 1 << 20])
def test_compute_aabb_and_box_sdist(oracle, dt, n):
    x, y, z = _cloud(n, dt, n)
    box = S.compute_aabb(x, y, z)
    rbox = oracle.compute_aabb(x, y, z)
    assert np.array_equal(np.array(box, dt), rbox)
    out = np.empty(n, dt)
    S.all_points_aabb_signed_distance(x, y, z, *rbox, out)
    assert np.array_equal(out, oracle.all_points_aabb_sdist(x, y, z, rbox))
    # a box that leaves points outside: positive distances there
    small = rbox.copy()
    small[1::2] = (rbox[0::2] + rbox[1::2]) / 2
    S.all_points_aabb_signed_distance(x, y, z, *small, out)
    assert np.array_equal(out, oracle.all_points_aabb_sdist(x, y, z, small))


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_init_volume_box_device(oracle, dt):
    n = 300001
    x, y, z = _cloud(n, dt, 5)
    dx, dy, dz = (S.DeviceArray.from_host(v) for v in (x, y, z))
    out = S.DeviceArray(n, dt)
    st = S.Stream()
    S.init_volume_box(dx, dy, dz, out, st)
    st.synchronize()
    want = np.abs(oracle.all_points_aabb_sdist(x, y, z, oracle.compute_aabb(x, y, z)))
    assert np.array_equal(out.download(), want)


def test_seed_then_edf_min_accumulates(oracle):
    """The driver's flow: seed out with the volume-box distance, then min-accumulate the surface."""
    from ssdf_amd import meshgen as mg
    m = mg.uv_sphere(0.5, 24, 12)
    x, y, z = _cloud(5000, np.float64, 9)
    box = oracle.compute_aabb(x, y, z)
    seed = np.abs(oracle.all_points_aabb_sdist(x, y, z, box))
    md = m.as_dtype(np.float64)
    out = seed.copy()
    S.edf(x, y, z, md.s0, md.s1, md.s2, md.sx, md.sy, md.sz, out)
    d2, _, _, _ = oracle.brute_argmin(x, y, z, m)
    assert np.array_equal(out, np.minimum(seed, np.sqrt(d2)))


def test_compute_aabb_rejects_empty():
    e = np.empty(0, np.float32)
    with pytest.raises(S.SsdfError):
        S.compute_aabb(e, e, e)
    out = np.empty(0, np.float32)
    assert S.all_points_aabb_signed_distance(e, e, e, 0, 1, 0, 1, 0, 1, out) == 0
