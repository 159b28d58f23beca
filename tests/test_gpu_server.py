"""Serving loop: results of served requests are bit-identical to direct resident-mesh queries;
budget admission evicts idle meshes and rejects requests that can never fit; statistics add up."""
import numpy as np
import pytest

import ssdf_amd as S
from ssdf_amd import meshgen as mg

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pool():
    return [mg.uv_sphere(0.5, 32, 16), mg.displaced_icosphere(8000, seed=7),
            mg.torus(0.6, 0.2, 100, 40, sigma=0.004, seed=5)]


def _direct(m, mode, x, y, z, init, flags):
    rm = S.ResidentMesh.from_mesh(m)
    dx, dy, dz = (S.DeviceArray.from_host(v) for v in (x, y, z))
    d = S.DeviceArray.from_host(init.copy())
    t = S.DeviceArray.from_host(np.full(x.size, -1, np.int32))
    rm.query_points(mode, dx, dy, dz, d, t, flags=flags)
    return d.download(), t.download()


def _requests(pool, n_req, seed):
    rng = np.random.default_rng(seed)
    reqs = []
    for k in range(n_req):
        mid = int(rng.integers(len(pool)))
        n = int(rng.integers(1, 40000))
        P = rng.uniform(-1.2, 1.2, (n, 3)).astype(np.float32)
        mode = S.MODE_SDF if k % 3 else S.MODE_EDF
        flags = S.FLAG_FIXUP if k % 2 else 0
        init = np.full(n, 0.3 if mode == S.MODE_EDF else 0.0, np.float32)
        reqs.append((mid, mode, flags, [np.ascontiguousarray(P[:, d]) for d in range(3)], init))
    return reqs


def test_served_results_equal_direct(pool):
    srv = S.Server(8 << 30)
    for i, m in enumerate(pool):
        srv.add_mesh(i, m)
    reqs = _requests(pool, 24, 1)
    tickets, outs = [], []
    for k, (mid, mode, flags, (x, y, z), init) in enumerate(reqs):
        if k % 2:  # pinned buffers for half of the requests
            bufs = [S.PinnedArray(x.size, np.float32) for _ in range(4)] + [S.PinnedArray(x.size, np.int32)]
            for b, v in zip(bufs[:3], (x, y, z)):
                b.array[:] = v
            bufs[3].array[:] = init
            bufs[4].array[:] = -1
            arrs = [b.array for b in bufs]
        else:
            bufs = None
            arrs = [x, y, z, init.copy(), np.full(x.size, -1, np.int32)]
        tickets.append(srv.submit(mid, mode, *arrs[:3], arrs[3], arrs[4], flags=flags))
        outs.append((arrs, bufs))
    lat = [srv.wait(t) for t in tickets]
    assert all(v >= 0 for v in lat)
    for (mid, mode, flags, (x, y, z), init), (arrs, bufs) in zip(reqs, outs):
        d, t = _direct(pool[mid], mode, x, y, z, init, flags)
        assert np.array_equal(arrs[3], d)
        assert np.array_equal(arrs[4], t)
    st = srv.stats()
    assert st["served"] == len(reqs) and st["errors"] == 0
    assert st["queries"] == sum(r[3][0].size for r in reqs)
    assert st["mesh_loads"] == len(pool) and st["evictions"] == 0
    assert 0 <= st["p50_s"] <= st["p99_s"] <= st["max_s"]
    assert st["pending"] == 0 and st["inflight"] == 0 and st["bytes_in_use"] > 0
    srv.close()


def test_budget_evicts_idle_meshes(pool):
    reqs = _requests(pool, 12, 2)
    # room for the largest mesh plus the largest request, not for two meshes
    probe = S.Server(1 << 40)
    for i, m in enumerate(pool):
        probe.add_mesh(i, m)
    charges = sorted(probe.charge(i, 1)[0] for i in range(len(pool)))
    req_b = max(probe.charge(0, r[3][0].size)[1] for r in reqs)
    probe.close()
    budget = charges[-1] + req_b + 4096
    assert budget < charges[-1] + charges[-2]  # the two larger meshes never fit together
    srv = S.Server(budget, max_inflight=2)
    for i, m in enumerate(pool):
        srv.add_mesh(i, m)
    res = []
    for mid, mode, flags, (x, y, z), init in reqs:
        o, t = init.copy(), np.full(x.size, -1, np.int32)
        res.append((srv.submit(mid, mode, x, y, z, o, t, flags=flags), o, t))
    for tk, o, t in res:
        srv.wait(tk)
    for (mid, mode, flags, (x, y, z), init), (_, o, t) in zip(reqs, res):
        d, tt = _direct(pool[mid], mode, x, y, z, init, flags)
        assert np.array_equal(o, d) and np.array_equal(t, tt)
    st = srv.stats()
    assert st["evictions"] > 0 and st["mesh_loads"] > len(pool) - 1
    assert st["peak_bytes"] <= budget
    srv.close()


def test_rejects_what_can_never_fit(pool):
    srv = S.Server(4 << 20)
    srv.add_mesh(0, pool[0])
    x = np.zeros(1 << 20, np.float32)
    with pytest.raises(S.SsdfError):
        srv.submit(0, S.MODE_SDF, x, x, x, np.zeros_like(x))
    with pytest.raises(S.SsdfError):
        srv.submit(5, S.MODE_SDF, x[:10], x[:10], x[:10], np.zeros(10, np.float32))
    srv.close()


def test_submitters_on_several_threads(pool):
    """Requests submitted and waited for from four host threads at once (the C calls release the
    interpreter lock, so the server sees concurrent submits): every result equals the direct
    query's, and the statistics count every request once."""
    import threading
    import time

    srv = S.Server(8 << 30)
    for i, m in enumerate(pool):
        srv.add_mesh(i, m)
    per_thread = [_requests(pool, 8, 100 + k) for k in range(4)]
    results = [None] * 4
    errors = []

    def worker(k):
        try:
            mine = []
            for mid, mode, flags, (x, y, z), init in per_thread[k]:
                out = init.copy()
                tri = np.full(x.size, -1, np.int32)
                mine.append((srv.submit(mid, mode, x, y, z, out, tri, flags=flags), out, tri))
            for t, _, _ in mine:
                while not srv.poll(t):
                    time.sleep(0.0005)
                assert srv.wait(t) >= 0
            results[k] = [(out, tri) for _, out, tri in mine]
        except Exception as e:  # reported by the main thread
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(k,)) for k in range(4)]
    for th in threads:
        th.start()
    for th in threads:
        th.join()
    assert not errors, errors
    for k in range(4):
        for (mid, mode, flags, (x, y, z), init), (out, tri) in zip(per_thread[k], results[k]):
            d, t = _direct(pool[mid], mode, x, y, z, init, flags)
            assert np.array_equal(out, d) and np.array_equal(tri, t)
    st = srv.stats()
    assert st["served"] == 32 and st["errors"] == 0 and st["pending"] == 0 and st["inflight"] == 0
    srv.close()
