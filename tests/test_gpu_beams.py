"""The capped-cylinder stress family at test size: closed beams against the oracle (canonical
indices, distances, signs outside the band), and the 1%-holed variant signed through the
winding-number fallback (SSDF_FLAG_SIGN_AUTO) against the intact mesh's signs away from the surface."""
import numpy as np
import pytest

import ssdf_amd as S
from tests import bars
from ssdf_amd import meshgen as mg
from ssdf_amd import workloads as W

pytestmark = pytest.mark.gpu


def _grid(m, n, flags):
    rm = S.ResidentMesh.from_mesh(m)
    N = n ** 3
    o = S.DeviceArray(N, np.float32)
    t = S.DeviceArray(N, np.int32)
    rm.query_grid(S.MODE_SDF, (n, n, n), (0, 0, 0), (1, 1, 1), o, t, flags=flags)
    q, e = rm.last_stats()
    assert e == 0
    return o.download(), t.download(), rm.info()


def test_beams_closed_against_oracle(oracle):
    m = W.beams(10, seed=21, n_circ=24, n_axial=40)
    n = 40
    o, t, info = _grid(m, n, S.FLAG_FIXUP)
    assert info["open_edges"] == 0
    x, y, z = mg.grid_points(n, n, n, (0, 0, 0), (1, 1, 1))
    s, rtri, _, _, _ = oracle.sdf(x, y, z, m)
    assert np.array_equal(t, rtri)
    bars.assert_fixup_distances(o, np.abs(s), "beams grid")
    lo, hi = m.bbox()
    assert bars.sign_mismatches_outside_band(o, s, 1e-6 * float(np.linalg.norm(hi - lo))) == 0
    assert np.any(o < 0)  # some nodes fall inside a beam


def test_beams_holed_gwn_signs(oracle):
    full = W.beams(10, seed=21, n_circ=24, n_axial=40)
    holed = W.beams(10, seed=21, n_circ=24, n_axial=40, hole_fraction=0.01)
    n = 40
    a, _, _ = _grid(full, n, S.FLAG_FIXUP)
    b, _, info = _grid(holed, n, S.FLAG_FIXUP | S.FLAG_SIGN_AUTO)
    assert info["open_edges"] > 0
    away = np.abs(a) > 0.01
    assert np.array_equal(np.sign(a[away]), np.sign(b[away]))
