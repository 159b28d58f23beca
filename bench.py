#!/usr/bin/env python3
"""Benchmark of the closest-point / signed-distance hot path on one MI355X.

Workload (default): noisy torus, R=0.6 r=0.2, 500 x 200 segments = 200,000 triangles, every vertex
displaced along its normal by N(0, 0.005^2) (SplitMix64 seed 2); query set: the 256^3 grid on
[-1,1]^3 (16,777,216 FP32 nodes); output: signed distance + closest triangle per node.

A step is one full job on inputs already resident in HBM: LBVH build + FP64 pseudonormals from the
device mesh arrays, then the 256^3 signed query in FP32 + FP64 fixup, outputs written to HBM.
`value` = queries / step time (device-resident).  Also reported: the same step in plain FP32, the
query alone on a resident BVH, the reference-shaped host calls (upload, build, query, download) in
plain FP32 and with the fixup, the roofline of the query kernel, and the other workloads:
  batch64   64 displaced icospheres (1 k - 500 k triangles) with their own grids (84.8 M nodes),
            one persistent launch per call, against 64 reference-shaped per-mesh calls;
  served    Poisson arrivals (200 requests/s for 60 s) over a pool of 32 resident meshes
            (5 k - 2 M triangles), 1e4 - 1e7 distinct points per request, plus a sweep of lower
            arrival rates for the highest one served with bounded latency;
  beams200  200 capped cylinders (4.99 M triangles), 512^3 grid and 1e8 scattered points, intact
            and with 1% of the triangles deleted (winding-number signs).
The cpu_baseline leg times the oracle's restatement of edf_celllist<float> (the reference's CPU
engine) on the host's physical cores, and is where every leg's sampled results are checked against
the oracle's exact FP64 argmin: each leg reports its max and RMS distance error, index mismatches
and sign mismatches outside the band (`error_vs_fp64`).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mqueries/s @256^3×200k tris (FP32+fixup vs plain); served p50/p99 latency"
FLOPS_BOX, FLOPS_TRI, FLOPS_SIGN = 8, 81, 8
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0
HOST_THREADS_MAX = 16  # host threads this process may use per GPU (the host's CPU share)


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _rnd(v, k=4):
    return round(float(v), k)


class Samples:
    """Seeded subsamples of every leg's results, kept for the FP64 check of the cpu_baseline leg.
    Each entry: leg name, mesh, FP32 query coordinates, the GPU's (signed) distance and triangle,
    and how its sign is judged (the mesh's own pseudonormal sign outside the 1e-6 * bbox band, or
    another mesh's sign outside a given band)."""

    def __init__(self):
        self.items = []

    def add(self, leg, mesh, x, y, z, d, tri, signed=True, sign_mesh=None, sign_band=None):
        self.items.append(dict(leg=leg, mesh=mesh, x=np.ascontiguousarray(x, np.float32),
                               y=np.ascontiguousarray(y, np.float32), z=np.ascontiguousarray(z, np.float32),
                               d=np.asarray(d, np.float32).copy(), tri=np.asarray(tri, np.int32).copy(),
                               signed=signed, sign_mesh=sign_mesh, sign_band=sign_band))


def _pick(n, k, seed):
    return np.sort(np.random.default_rng(seed).choice(n, size=min(k, n), replace=False))


def _grid_coords(idx, n, lo, hi):
    from ssdf_amd import meshgen as mg
    ax = [mg.grid_axis(n[d], lo[d], hi[d]) for d in range(3)]
    i, j, k = idx % n[0], (idx // n[0]) % n[1], idx // (n[0] * n[1])
    return ax[0][i], ax[1][j], ax[2][k]


def make_workload(args):
    from ssdf_amd import meshgen as mg
    m = mg.torus(0.6, 0.2, 500, 200, sigma=0.005, seed=2)
    n = (args.grid,) * 3
    return m, n, (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0)


def run_gpu(args, m, n, lo, hi, samples):
    import ctypes as C

    import ssdf_amd as S
    L = S.lib()
    stream = S.Stream()
    sp = C.c_void_p(stream.ptr)
    dev = {k: S.DeviceArray.from_host(getattr(m, k)) for k in ("s0", "s1", "s2", "sx", "sy", "sz")}
    N = n[0] * n[1] * n[2]
    out = S.DeviceArray(N, np.float32)
    tri = S.DeviceArray(N, np.int32)
    ev = [S.Event() for _ in range(4)]

    def create():
        h = C.c_void_p()
        rc = L.ssdf_mesh_create(0, m.ntris, C.c_void_p(dev["s0"].ptr), C.c_void_p(dev["s1"].ptr),
                                C.c_void_p(dev["s2"].ptr), m.nverts, C.c_void_p(dev["sx"].ptr),
                                C.c_void_p(dev["sy"].ptr), C.c_void_p(dev["sz"].ptr), 1, 1, sp, C.byref(h))
        S._check(rc, "ssdf_mesh_create")
        return h

    def query(h, flags):
        rc = L.ssdf_mesh_query_grid(h, S.MODE_SDF, n[0], n[1], n[2], lo[0], lo[1], lo[2], hi[0], hi[1], hi[2],
                                    C.c_void_p(out.ptr), C.c_void_p(tri.ptr), flags, sp)
        S._check(rc, "ssdf_mesh_query_grid")

    # one resident handle; every step rebuilds its LBVH and pseudonormals in place from the device
    # mesh arrays (no allocation) and runs the full query
    handle = create()
    brick = [int(v) for v in args.brick.split(",")]
    S._check(L.ssdf_mesh_set_schedule(handle, *brick, -1), "ssdf_mesh_set_schedule")

    def step(flags, qtimes):
        S._check(L.ssdf_mesh_rebuild(handle, 1, sp), "ssdf_mesh_rebuild")
        ev[0].record(stream)
        query(handle, flags)
        ev[1].record(stream)
        stream.synchronize()
        qtimes.append(ev[0].elapsed_ms(ev[1]))

    res = {}
    modes = (("fixup", S.FLAG_FIXUP), ("plain", 0))
    if args.only:
        modes = tuple(md for md in modes if md[0] == args.only)
    idx = _pick(N, args.check_torus, seed=12)
    sx, sy, sz = _grid_coords(idx, n, lo, hi)
    for name, flags in modes:
        qt = []
        for _ in range(args.warmup):
            step(flags, qt)
        qt = []
        stream.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            step(flags, qt)
        stream.synchronize()
        dt = (time.perf_counter() - t0) / args.steps
        res[name] = {"ms_per_step": dt * 1e3, "query_ms": float(np.mean(qt)), "query_ms_min": float(np.min(qt)),
                     "mqps": N / dt / 1e6, "query_mqps": N / (np.mean(qt) * 1e-3) / 1e6}
        log(f"[bench] {name}: step {dt * 1e3:.3f} ms, query {np.mean(qt):.3f} ms")
        if not args.only:
            samples.add(f"torus200k_{name}", m, sx, sy, sz, out.download()[idx], tri.download()[idx])
    if args.only:
        return res
    # fixup statistics of one more fixup query on the resident mesh
    h = handle
    query(h, S.FLAG_FIXUP)
    stream.synchronize()
    q, e = C.c_uint(), C.c_int()
    L.ssdf_mesh_last_stats(h, C.byref(q), C.byref(e))
    info = S._MeshInfo()
    L.ssdf_mesh_info(h, C.byref(info))
    res["fixup_queued"] = int(q.value)
    res["kernel_error"] = int(e.value)
    res["maxdepth"] = int(info.maxdepth)
    res["open_edges"] = int(info.open_edges)
    L.ssdf_mesh_destroy(h)
    # build alone (rebuild in place), for the record
    bt = []
    for _ in range(3):
        h2 = create()
        stream.synchronize()
        t0 = time.perf_counter()
        S._check(L.ssdf_mesh_rebuild(h2, 1, sp), "ssdf_mesh_rebuild")
        stream.synchronize()
        bt.append((time.perf_counter() - t0) * 1e3)
        L.ssdf_mesh_destroy(h2)
    res["build_ms"] = float(np.min(bt))
    # the reference-shaped host call: upload, build, query, download per call (serial/bvh/bvh.cu:80-158)
    outh = np.empty(N, np.float32)
    trih = np.empty(N, np.int32)
    for key, flags in (("plain", 0), ("fixup", S.FLAG_FIXUP)):
        S.sample_sdf_bvh(m.s0, m.s1, m.s2, m.sx, m.sy, m.sz, n[0], n[1], n[2], *lo, *hi, outh, trih, flags=flags)
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            S.sample_sdf_bvh(m.s0, m.s1, m.s2, m.sx, m.sy, m.sz, n[0], n[1], n[2], *lo, *hi, outh, trih,
                             flags=flags)
            ts.append(time.perf_counter() - t0)
        res[f"host_call_{key}_s"] = float(np.min(ts))
        res[f"host_call_{key}_mqps"] = N / res[f"host_call_{key}_s"] / 1e6
        log(f"[bench] host call {key}: {res[f'host_call_{key}_s'] * 1e3:.1f} ms")
    return res


def run_batch(args, samples):
    """64 displaced icospheres (4.1 M triangles), one grid each (84.8 M nodes), signed distance +
    triangle with FP64 fixup: the per-mesh reference-shaped call vs one batch."""
    import ssdf_amd as S
    from ssdf_amd import meshgen as mg
    from ssdf_amd import workloads as W
    meshes, dims, boxes = W.batch64()
    nq = int(dims.prod(1).sum())
    flags = S.FLAG_FIXUP
    res = {"meshes": len(meshes), "triangles": int(sum(m.ntris for m in meshes)), "queries": nq}
    # per-mesh plain path: upload, build, query, download, free per call (serial/bvh/bvh.cu:80-158)
    outs = [np.empty(int(d.prod()), np.float32) for d in dims]
    for m, d, bx, o in zip(meshes[:2], dims, boxes, outs):  # driver and code paths warm
        S.sample_sdf_bvh(m.s0, m.s1, m.s2, m.sx, m.sy, m.sz, int(d[0]), int(d[1]), int(d[2]), *bx.astype(np.float32),
                         o, None, flags=flags)
    t0 = time.perf_counter()
    for m, d, bx, o in zip(meshes, dims, boxes, outs):
        S.sample_sdf_bvh(m.s0, m.s1, m.s2, m.sx, m.sy, m.sz, int(d[0]), int(d[1]), int(d[2]), *bx.astype(np.float32),
                         o, None, flags=flags)
    res["per_mesh_calls_s"] = time.perf_counter() - t0
    res["per_mesh_calls_mqps"] = nq / res["per_mesh_calls_s"] / 1e6
    log(f"[bench] batch64 per-mesh calls: {res['per_mesh_calls_s']:.3f} s")
    # the reference's own per-call GPU path (serial/bvh/bvh.cu:80-158, edf_bvh_cuda): host points in,
    # plain FP32 unsigned distances min-accumulated into a host array, one call per mesh; here
    # ssdf_edf_f over each mesh's grid nodes given as points (the points are the caller's data and
    # are made before the clock starts)
    pts = [mg.grid_points(int(d[0]), int(d[1]), int(d[2]), bx[:3], bx[3:]) for d, bx in zip(dims, boxes)]
    eouts = [np.full(p[0].size, np.inf, np.float32) for p in pts]
    for m, p, o in zip(meshes[:2], pts, eouts):
        S.edf(*p, m.s0, m.s1, m.s2, m.sx, m.sy, m.sz, o)
        o.fill(np.inf)
    t0 = time.perf_counter()
    for m, p, o in zip(meshes, pts, eouts):
        S.edf(*p, m.s0, m.s1, m.s2, m.sx, m.sy, m.sz, o)
    res["per_mesh_edf_calls_s"] = time.perf_counter() - t0
    res["per_mesh_edf_calls_mqps"] = nq / res["per_mesh_edf_calls_s"] / 1e6
    log(f"[bench] batch64 per-mesh edf calls (plain, host points): {res['per_mesh_edf_calls_s']:.3f} s")
    del pts, eouts
    # one batch, host to host: upload + build all, one query launch, download
    off, total = S.Batch.grid_layout(dims)
    st = S.Stream()
    out = S.DeviceArray(total, np.float32)
    tri = S.DeviceArray(total, np.int32)
    host = np.empty(total, np.float32)
    t0 = time.perf_counter()
    b = S.Batch(meshes, stream=st)
    b.query_grids(S.MODE_SDF, dims, boxes, off, out, tri, flags=flags, stream=st)
    st.synchronize()
    host[:] = out.download()
    res["batch_host_to_host_s"] = time.perf_counter() - t0
    res["batch_host_to_host_mqps"] = nq / res["batch_host_to_host_s"] / 1e6
    # the per-mesh calls and the batch agree bit for bit
    res["matches_per_mesh_calls"] = bool(all(np.array_equal(host[off[i]:off[i] + outs[i].size], outs[i])
                                             for i in range(len(meshes))))
    ev = [S.Event(), S.Event()]

    def replay_ms(reps):
        for _ in range(max(1, args.warmup)):
            b.replay(st)
        st.synchronize()
        ev[0].record(st)
        for _ in range(reps):
            b.replay(st)
        ev[1].record(st)
        st.synchronize()
        return ev[0].elapsed_ms(ev[1]) / reps

    # the default schedule, then the autotuned one (per mesh: brick shape and leaf budget; then the
    # launch geometry)
    b.capture(False, S.MODE_SDF, dims, boxes, off, out, tri, flags=flags, stream=st)
    default_ms = replay_ms(args.steps)
    log(f"[bench] batch64 resident_query (default schedule): {default_ms:.3f} ms")
    bpc, ch, tune_ms = b.autotune(S.MODE_SDF, dims, boxes, off, out, tri, flags=flags, stream=st)
    per_mesh = [b.mesh_schedule(i) for i in range(len(meshes))]
    bricks, budgets = {}, {}
    for (bk, hv) in per_mesh:
        key = "x".join(map(str, bk))
        bricks[key] = bricks.get(key, 0) + 1
        hk = str(hv) if hv >= 0 else "process-wide (8192)"
        budgets[hk] = budgets.get(hk, 0) + 1
    res["schedule"] = {"blocks_per_cu": bpc, "units_per_grab": ch, "query_ms": round(tune_ms, 3),
                       "meshes_per_brick": bricks, "meshes_per_leaf_budget": budgets,
                       "resident_query_default_schedule_ms": round(default_ms, 3)}
    for rebuild, key in ((True, "resident_step"), (False, "resident_query")):
        b.capture(rebuild, S.MODE_SDF, dims, boxes, off, out, tri, flags=flags, stream=st)
        ms = replay_ms(args.steps)
        res[key + "_ms"] = ms
        res[key + "_mqps"] = nq / (ms * 1e-3) / 1e6
        log(f"[bench] batch64 {key}: {ms:.3f} ms")
    res["schedule"]["autotune_gain"] = round(default_ms / res["resident_query_ms"], 4)
    q, e = b.last_stats()
    res["fixup_queued"], res["kernel_error"] = q, e
    # the batch query against the FP32 vector peak: canonical per-query work of every mesh's grid
    wc = json.loads((ROOT / "tests" / "golden" / "work_counts.json").read_text())["batch64_icospheres_grids"]
    if wc["queries"] == nq:
        flops = FLOPS_BOX * wc["box_tests"] + FLOPS_TRI * wc["tri_tests"] + FLOPS_SIGN * nq
        ach = flops / (res["resident_query_ms"] * 1e-3) / 1e12
        res["roofline"] = {"bound": "valu", "achieved": round(ach, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
                           "frac": round(ach / PEAK_FP32_TFLOPS, 4), "algorithmic_gflop": round(flops / 1e9, 1),
                           "useful_box_frac": wc["packet"]["useful_box_frac"],
                           "useful_tri_frac": wc["packet"]["useful_tri_frac"]}
    res["speedup_resident_step_vs_per_mesh_calls"] = res["resident_step_mqps"] / res["per_mesh_calls_mqps"]
    res["speedup_resident_step_vs_per_mesh_edf_calls"] = res["resident_step_mqps"] / res["per_mesh_edf_calls_mqps"]
    res["speedup_host_to_host_vs_per_mesh_calls"] = res["batch_host_to_host_mqps"] / res["per_mesh_calls_mqps"]
    ho, ht = out.download(), tri.download()
    for i, m in enumerate(meshes):
        nd = tuple(int(v) for v in dims[i])
        idx = _pick(int(np.prod(nd)), args.check_batch, seed=100 + i)
        x, y, z = _grid_coords(idx, nd, boxes[i][:3], boxes[i][3:])
        samples.add("batch64", m, x, y, z, ho[off[i] + idx], ht[off[i] + idx])
    b.close()
    return {k: (round(v, 4) if isinstance(v, float) else v) for k, v in res.items()}


def _pool_fill(mesh, arrays, npts, seed):
    """Fills three pinned arrays with `npts` query points of `mesh` (workloads.query_cloud) and
    returns how many of them are distinct (a 64-bit hash of the three coordinates' bits; a hash
    collision can only undercount)."""
    from ssdf_amd import workloads as W
    pts = W.query_cloud(mesh, npts, seed=seed)
    for a, p in zip(arrays, pts):
        np.copyto(a.array, p)
    b = [p.view(np.uint32).astype(np.uint64) for p in pts]
    h = (b[0] * np.uint64(0x9E3779B97F4A7C15)) ^ (b[1] * np.uint64(0xC2B2AE3D27D4EB4F)) ^ \
        (b[2] * np.uint64(0x165667B19E3779F9))
    return int(np.unique(h).size)


def _serve_window(srv, S, meshes, pools, outs, times, mesh_ids, nq, offs, t0, on_done=None, tag="serve"):
    """Submits the requests at their scheduled arrivals (t0 + times on the server clock) and waits
    for all; output buffers are a ring, so a slot's earlier request is waited for before reuse
    (latency still counts from the scheduled arrival).  Returns per-request latencies."""
    ring = len(outs)
    slots = [None] * ring
    lat = np.zeros(len(times))
    for k in range(len(times)):
        if k % 1000 == 0:
            log(f"[bench] {tag}: request {k} at t={srv.clock() - t0:.2f} s")
        slot = k % ring
        if slots[slot] is not None:
            kk, tk = slots[slot]
            lat[kk] = srv.wait(tk)
            if on_done:
                on_done(kk, outs[slot])
        arrival = t0 + float(times[k])
        delay = arrival - srv.clock()
        if delay > 0:
            time.sleep(delay)
        i, n, o = int(mesh_ids[k]), int(nq[k]), int(offs[k])
        x, y, z = (p.array[o:o + n] for p in pools[i])
        tk = srv.submit(i, S.MODE_SDF, x, y, z, outs[slot][0].array[:n], outs[slot][1].array[:n],
                        flags=S.FLAG_FIXUP, arrival=arrival)
        slots[slot] = (k, tk)
    log(f"[bench] {tag}: all submitted at t={srv.clock() - t0:.2f} s, draining")
    for s in range(ring):
        if slots[s] is not None:
            kk, tk = slots[s]
            lat[kk] = srv.wait(tk)
            if on_done:
                on_done(kk, outs[s])
    return lat


def _window_stats(times, nq, lat, window):
    done = times + lat
    pend = (times <= window) & (done > window)
    return {"p50_ms": _rnd(np.percentile(lat, 50) * 1e3, 2), "p99_ms": _rnd(np.percentile(lat, 99) * 1e3, 2),
            "max_ms": _rnd(lat.max() * 1e3, 2), "backlog_requests_at_window_end": int(pend.sum()),
            "backlog_queries_at_window_end": int(nq[pend].sum()),
            "drain_s": _rnd(max(0.0, float(done.max()) - window), 3),
            "throughput_mqps": _rnd(nq.sum() / max(float(done.max()), 1e-9) / 1e6, 2)}


def run_serve(args, samples):
    """Open-loop serving: Poisson arrivals over a pool of 32 resident meshes (5 k - 2 M triangles),
    1e4 - 1e7 signed queries per request (half uniform in the padded bbox, half within 0.02 of the
    surface), FP32 + fixup; every request takes its own points (a contiguous slice of its mesh's
    pool of distinct points at a seeded offset); latency from the scheduled arrival to the results in
    host memory."""
    import ssdf_amd as S
    from ssdf_amd import meshgen as mg
    from ssdf_amd import workloads as W
    t_gen = time.perf_counter()
    meshes = W.serve_pool()
    log(f"[bench] serve: mesh pool generated ({time.perf_counter() - t_gen:.1f} s)")
    P = args.serve_pool_points
    times, mesh_ids, nq = W.poisson_requests(args.serve_rate, args.serve_seconds, len(meshes))
    nq = np.minimum(nq, P)
    offs = (mg.SplitMix64(23).uniform(len(times)) * (P - nq + 1)).astype(np.int64)
    pools = [[S.PinnedArray(P, np.float32) for _ in range(3)] for _ in meshes]
    with ThreadPoolExecutor(max_workers=min(HOST_THREADS_MAX, os.cpu_count() or 1)) as ex:
        distinct = list(ex.map(lambda i: _pool_fill(meshes[i], pools[i], P, 100 + i), range(len(meshes))))
    # a request's points are distinct if its pool's are; otherwise at most (P - distinct) repeat
    dup = np.array([P - d for d in distinct], np.int64)
    distinct_q = int(np.maximum(nq - dup[mesh_ids], 0).sum())
    ring = 24
    nmax = int(nq.max())
    outs = [(S.PinnedArray(nmax, np.float32), S.PinnedArray(nmax, np.int32)) for _ in range(ring)]
    gen_s = time.perf_counter() - t_gen
    log(f"[bench] serve: {len(meshes)} pools of {P} distinct points and the output ring pinned ({gen_s:.1f} s)")
    budget = args.serve_budget_gib << 30
    srv = S.Server(budget, max_inflight=16)
    for i, m in enumerate(meshes):
        srv.add_mesh(i, m)
    # warm-up: one small request per mesh loads every mesh into HBM
    warm = [srv.submit(i, S.MODE_SDF, *[p.array[:1024] for p in pools[i]], outs[0][0].array[:1024],
                       outs[0][1].array[:1024], flags=S.FLAG_FIXUP) for i in range(len(meshes))]
    for t in warm:
        srv.wait(t)
    # results of the first request of every mesh are sampled for the FP64 check
    first = {}
    for k, i in enumerate(mesh_ids):
        first.setdefault(int(i), k)
    want = {k: i for i, k in first.items()}

    def keep(k, slot):
        if k in want:
            i = want[k]
            n, o = int(nq[k]), int(offs[k])
            sel = _pick(n, args.check_serve, seed=700 + i)
            samples.add("served", meshes[i], *(pools[i][d].array[o + sel] for d in range(3)), slot[0].array[sel],
                        slot[1].array[sel])

    t0 = srv.clock() + 0.05
    log(f"[bench] serve: {len(meshes)} meshes resident, {len(times)} requests over {args.serve_seconds} s")
    lat = _serve_window(srv, S, meshes, pools, outs, times, mesh_ids, nq, offs, t0, on_done=keep)
    st = srv.stats()
    total_q = int(nq.sum())
    res = {"requests": len(times), "rate_req_s": args.serve_rate, "seconds_offered": args.serve_seconds,
           "meshes": len(meshes), "triangles": int(sum(m.ntris for m in meshes)), "queries": total_q,
           "pool_points_per_mesh": P, "distinct_fraction": _rnd(distinct_q / total_q, 6),
           "offered_mqps": _rnd(total_q / args.serve_seconds / 1e6, 2)}
    res.update(_window_stats(times, nq, lat, args.serve_seconds))
    res.update({"mesh_loads": int(st["mesh_loads"]), "evictions": int(st["evictions"]), "errors": int(st["errors"]),
                "hbm_budget_gib": budget / 2**30, "peak_gib": _rnd(st["peak_bytes"] / 2**30, 3),
                "setup_s": _rnd(gen_s, 1)})
    log(f"[bench] serve: {res}")
    # higher arrival rates, until one is not served with bounded latency (the backlog drains within
    # 1 s of the window's end and p99 stays under 1 s); the highest rate that is, main run included
    bounded = lambda w: w["drain_s"] <= 1.0 and w["p99_ms"] < 1000.0  # noqa: E731
    best = {"rate_req_s": args.serve_rate, "offered_mqps": res["offered_mqps"], "p50_ms": res["p50_ms"],
            "p99_ms": res["p99_ms"], "drain_s": res["drain_s"]} if bounded(res) else None
    if args.sweep_seconds > 0 and best is not None:
        sweep = []
        for f in (1.25, 1.5, 1.75, 2.0, 2.5):
            rate = f * args.serve_rate
            tt, mi, qq = W.poisson_requests(rate, args.sweep_seconds, len(meshes), seed=40 + int(100 * f))
            qq = np.minimum(qq, P)
            oo = (mg.SplitMix64(41 + int(100 * f)).uniform(len(tt)) * (P - qq + 1)).astype(np.int64)
            t1 = srv.clock() + 0.05
            lt = _serve_window(srv, S, meshes, pools, outs, tt, mi, qq, oo, t1, tag=f"serve sweep {rate:.0f}/s")
            w = {"rate_req_s": _rnd(rate, 1), "seconds": args.sweep_seconds,
                 "offered_mqps": _rnd(qq.sum() / args.sweep_seconds / 1e6, 2)}
            w.update(_window_stats(tt, qq, lt, args.sweep_seconds))
            sweep.append(w)
            log(f"[bench] serve sweep: {w}")
            if not bounded(w):
                break
            best = w
        res["sweep"] = sweep
    res["max_sustained"] = best
    srv.close()
    return res


def run_beams(args, samples):
    """200 capped cylinders in the unit cube (4.99 M triangles), intact and with 1% of the triangles
    deleted; signed distance + triangle on the 512^3 grid over the cube (FP32 + fixup; the holed
    mesh signed by winding numbers through SSDF_FLAG_SIGN_AUTO) and at 1e8 uniform scattered points."""
    import ssdf_amd as S
    from ssdf_amd import workloads as W
    t0 = time.perf_counter()
    full = W.beams()
    holed = W.beams(hole_fraction=0.01)
    log(f"[bench] beams: meshes generated ({time.perf_counter() - t0:.1f} s)")
    g = args.beams_grid
    N = g ** 3
    st = S.Stream()
    ev = [S.Event(), S.Event()]
    res = {"triangles": full.ntris, "triangles_holed": holed.ntris, "grid": g}

    def timed(f):
        f()
        st.synchronize()
        ev[0].record(st)
        f()
        ev[1].record(st)
        st.synchronize()
        return ev[0].elapsed_ms(ev[1])

    idx = _pick(N, args.check_beams, seed=31)
    gx, gy, gz = _grid_coords(idx, (g, g, g), (0, 0, 0), (1, 1, 1))
    outs = {}
    for key, m, flags in (("closed", full, S.FLAG_FIXUP), ("holed_gwn", holed, S.FLAG_FIXUP | S.FLAG_SIGN_AUTO)):
        rm = S.ResidentMesh.from_mesh(m, stream=st)
        o = S.DeviceArray(N, np.float32)
        t = S.DeviceArray(N, np.int32)
        ms = timed(lambda: rm.query_grid(S.MODE_SDF, (g, g, g), (0, 0, 0), (1, 1, 1), o, t, flags=flags, stream=st))
        q, e = rm.last_stats()
        res[key] = {"grid_ms": round(ms, 2), "grid_mqps": round(N / ms / 1e3, 1), "fixup_queued": q,
                    "kernel_error": e, "open_edges": rm.info()["open_edges"]}
        outs[key] = o.download()
        ot = t.download()
        if key == "closed":
            samples.add("beams200_closed", m, gx, gy, gz, outs[key][idx], ot[idx])
        else:
            # distances and triangles against the holed mesh, signs against the intact mesh beyond 0.01
            samples.add("beams200_holed_gwn", m, gx, gy, gz, outs[key][idx], ot[idx], sign_mesh=full,
                        sign_band=0.01)
        log(f"[bench] beams {key}: grid {ms:.1f} ms")
        if key == "closed":
            n = args.beams_points
            rng = np.random.default_rng(5)
            dx, dy, dz = (S.DeviceArray.from_host(rng.random(n, dtype=np.float32)) for _ in range(3))
            po = S.DeviceArray(n, np.float32)
            pt = S.DeviceArray(n, np.int32)
            ms = timed(lambda: rm.query_points(S.MODE_SDF, dx, dy, dz, po, pt, flags=flags, stream=st))
            q, e = rm.last_stats()
            res["scattered"] = {"points": n, "ms": round(ms, 2), "mqps": round(n / ms / 1e3, 1), "fixup_queued": q,
                                "kernel_error": e}
            sel = _pick(n, args.check_beams, seed=77)
            samples.add("beams200_scattered", m, dx.download()[sel], dy.download()[sel], dz.download()[sel],
                        po.download()[sel], pt.download()[sel])
            log(f"[bench] beams scattered: {ms:.1f} ms")
            del dx, dy, dz, po, pt
        del rm, o, t
    a, b = outs["closed"], outs["holed_gwn"]
    away = np.abs(a) > 0.01
    res["holed_sign_agreement_beyond_0.01"] = float(np.mean(np.sign(a[away]) == np.sign(b[away])))
    return res


def host_cores():
    """(logical CPUs this process may run on, physical cores among them)."""
    cpus = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else list(range(os.cpu_count() or 1))
    cores = set()
    for c in cpus:
        base = Path(f"/sys/devices/system/cpu/cpu{c}/topology")
        try:
            cores.add((base.joinpath("physical_package_id").read_text().strip(),
                       base.joinpath("core_id").read_text().strip()))
        except OSError:
            cores.add(("cpu", str(c)))
    return len(cpus), len(cores)


def cpu_baseline(args, m, n, lo, hi, samples, legs):
    """The cpu_baseline leg.  Timing: the oracle's restatement of edf_celllist<float>
    (serial/ssdf.hpp:642-963, its static `omp parallel for` schedule) on a seeded uniform subsample
    of the grid, one thread per physical core (at most the box's share), OMP_PROC_BIND=close,
    OMP_PLACES=cores.  Checking: every leg's sampled results against the oracle's exact FP64 argmin
    (its CPU BVH) and the reference's pseudonormal sign rule (bvh.cpp:1734-1792)."""
    nlog, nphys = host_cores()
    threads = max(1, min(nphys, HOST_THREADS_MAX))
    # the OpenMP runtime reads these when the oracle library loads (nothing has loaded it yet)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    os.environ["OMP_PROC_BIND"] = "close"
    os.environ["OMP_PLACES"] = "cores"
    from oracle import oracle as O
    from ssdf_amd import meshgen as mg
    O._load()
    rng = mg.SplitMix64(12345)
    N = n[0] * n[1] * n[2]
    idx = np.unique((rng.uniform(args.cpu_sample) * N).astype(np.int64))
    x, y, z = _grid_coords(idx, n, lo, hi)
    out = np.full(idx.size, np.finfo(np.float32).max, np.float32)
    t0 = time.perf_counter()
    O.celllist(x, y, z, m, out)
    dt = time.perf_counter() - t0
    res = {"value": idx.size / dt / 1e6, "unit": "Mqueries/s", "cores": threads, "kind": "port",
           "threads": O.num_threads(), "host_logical_cpus": nlog, "host_physical_cores": nphys,
           "sample": f"{idx.size} grid nodes drawn uniformly (SplitMix64 seed 12345) from the 256^3 grid; "
                     f"edf_celllist<float> restated in C (static OpenMP schedule as ssdf.hpp:912), "
                     f"{O.num_threads()} threads on {threads} physical cores (OMP_PLACES=cores, "
                     f"OMP_PROC_BIND=close), {dt:.2f} s incl. its cell build"}
    log(f"[bench] cpu_baseline: {res['value']:.3f} Mq/s on {threads} cores ({dt:.1f} s)")
    # ---- accuracy of every leg against the FP64 oracle ----
    t0 = time.perf_counter()
    by_mesh = {}
    for it in samples.items:
        by_mesh.setdefault(id(it["mesh"]), []).append(it)
    acc = {}
    normals = {}

    def normals_of(mesh):
        if id(mesh) not in normals:
            normals[id(mesh)] = O.pseudonormals(mesh)
        return normals[id(mesh)]

    for items in by_mesh.values():
        mesh = items[0]["mesh"]
        X, Y, Z = (np.concatenate([it[c] for it in items]) for c in "xyz")
        d2all, rtri = O.bvh_argmin(X, Y, Z, mesh)
        sg, _, _ = O.sign(X, Y, Z, rtri, mesh, normals_of(mesh))
        s = sg * np.sqrt(d2all)
        lo_m, hi_m = mesh.bbox()
        band = 1e-6 * float(np.linalg.norm(hi_m - lo_m))
        k = 0
        for it in items:
            nn = it["x"].size
            ref = s[k:k + nn]
            rt = rtri[k:k + nn]
            d2r = d2all[k:k + nn]
            k += nn
            if it["sign_mesh"] is not None:
                sref, _, _, _, _ = O.sdf(it["x"], it["y"], it["z"], it["sign_mesh"],
                                         normals=normals_of(it["sign_mesh"]))
                sband = it["sign_band"]
            else:
                sref, sband = ref, band
            d_ref = np.abs(ref)
            got = np.abs(it["d"].astype(np.float64))
            e = got - d_ref
            rel = np.abs(e) / np.maximum(d_ref, 1e-30)
            far = np.abs(sref) > sband
            a = acc.setdefault(it["leg"], {"n": 0, "max_err": 0.0, "sq": 0.0, "max_rel_err": 0.0,
                                           "index_mismatch": 0, "index_mismatch_not_tied": 0,
                                           "sign_mismatch_outside_band": 0,
                                           "sign_band": "1e-6*bbox" if it["sign_mesh"] is None else
                                           f"{sband} (intact mesh's sign)"})
            a["n"] += nn
            a["max_err"] = max(a["max_err"], float(np.abs(e).max()))
            a["sq"] += float((e * e).sum())
            a["max_rel_err"] = max(a["max_rel_err"], float(rel.max()))
            mis = it["tri"] != rt
            a["index_mismatch"] += int(mis.sum())
            if mis.any():
                # a differing triangle at exactly the argmin's FP64 distance is a tie, which the plain
                # FP32 path may break differently (FP32 + fixup takes the lowest index, as the oracle)
                d2_got = O.pair_d2(it["x"][mis], it["y"][mis], it["z"][mis], it["tri"][mis], mesh)
                a["index_mismatch_not_tied"] += int(np.sum(d2_got != d2r[mis]))
            a["sign_mismatch_outside_band"] += int(np.sum((np.sign(it["d"]) != np.sign(sref)) & far))
    for leg, a in acc.items():
        a["rms_err"] = math.sqrt(a.pop("sq") / max(a["n"], 1))
        a["max_err"] = float(f"{a['max_err']:.4g}")
        a["rms_err"] = float(f"{a['rms_err']:.4g}")
        a["max_rel_err"] = float(f"{a['max_rel_err']:.4g}")
        a["reference"] = "oracle FP64 argmin (CPU BVH) + pseudonormal sign"
        legs.setdefault(leg, a)
    log(f"[bench] FP64 checks of {sum(a['n'] for a in acc.values())} sampled results in "
        f"{time.perf_counter() - t0:.1f} s")
    return res


def measured_traffic(kernel):
    """HBM-side bytes per launch of `kernel` from the committed rocprofv3 counter summary
    (FETCH_SIZE and WRITE_SIZE passes, each scaled by what it reads for a copy kernel of known size)."""
    f = ROOT / "profiles" / "k_query_packet_torus200k_grid256_counters.txt"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        parts = line.split()
        if len(parts) == 3 and parts[0] == "traffic_bytes_per_launch" and parts[1] == kernel:
            return int(parts[2])
    return None


def measured_valu_instructions(kernel_prefix):
    """Wave-level VALU instructions per launch of the kernel (SQ_INSTS_VALU from the committed
    rocprofv3 counter summary), or None."""
    f = ROOT / "profiles" / "k_query_packet_torus200k_grid256_counters.txt"
    if not f.exists():
        return None
    for line in f.read_text().splitlines():
        if line.startswith(kernel_prefix):
            parts = line[len(kernel_prefix):].split()
            if len(parts) == 8 and all(p.isdigit() for p in parts):  # the SQ table: SQ_INSTS_VALU is column 2
                return int(parts[1])
    return None


def roofline(query_ms, n_queries, nverts, ntris):
    wc = json.loads((ROOT / "tests" / "golden" / "work_counts.json").read_text())["torus_200k_grid256"]
    assert wc["queries"] == n_queries
    flops = FLOPS_BOX * wc["box_tests"] + FLOPS_TRI * wc["tri_tests"] + FLOPS_SIGN * n_queries
    achieved = flops / (query_ms * 1e-3) / 1e12
    # grid SDF mode: 8 B per query (signed distance + triangle) plus one read of the mesh: vertices
    # 12 B, index triples 12 B, FP32 BVH nodes 32 B x (2T - 1), FP32 triangle corners 36 B
    hbm_bytes = 8 * n_queries + 12 * nverts + 12 * ntris + 32 * (2 * ntris - 1) + 36 * ntris
    hbm = hbm_bytes / (query_ms * 1e-3) / 1e9
    # the same launch against the instruction-issue limit: one wave-level VALU instruction per SIMD
    # per 4 cycles, 1024 SIMDs at 2.4 GHz
    valu = measured_valu_instructions("k_query_packet<64, 1, 1, true>")
    issue = None if valu is None else round(valu * 4 / (1024 * 2.4e9 * query_ms * 1e-3), 4)
    return {"bound": "valu", "achieved": round(achieved, 3), "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s",
            "frac": round(achieved / PEAK_FP32_TFLOPS, 4), "traffic": measured_traffic("k_query_packet_fixup"),
            "kernel": "k_query_packet (FP32+fixup)", "algorithmic_gflop_per_launch": round(flops / 1e9, 2),
            "valu_issue_frac": issue,
            # share of the packet walk's executed per-lane tests that the canonical per-query walk needs
            "useful_box_frac": wc["packet"]["useful_box_frac"], "useful_tri_frac": wc["packet"]["useful_tri_frac"],
            "hbm": {"achieved": round(hbm, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(hbm / PEAK_HBM_GBS, 5),
                    "algorithmic_bytes_per_launch": hbm_bytes}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--brick", default="4,4,4", help="grid nodes per wave of the torus legs (x,y,z; product 64)")
    ap.add_argument("--cpu-sample", type=int, default=1 << 23)
    ap.add_argument("--no-cpu", action="store_true", help="skip the cpu_baseline leg (and the FP64 checks)")
    ap.add_argument("--only", choices=["fixup", "plain", "batch"], default=None,
                    help="time one mode (or the 64-mesh batch leg) only and print nothing else (for profiling)")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-mesh batch leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--no-beams", action="store_true", help="skip the 200-beam stress leg")
    ap.add_argument("--beams-grid", type=int, default=512)
    ap.add_argument("--beams-points", type=int, default=100_000_000)
    ap.add_argument("--serve-seconds", type=float, default=60.0, help="arrival window of the serving leg")
    ap.add_argument("--serve-rate", type=float, default=200.0, help="Poisson arrival rate (requests/s)")
    ap.add_argument("--serve-pool-points", type=int, default=10_000_000, help="distinct points per mesh pool")
    ap.add_argument("--serve-budget-gib", type=int, default=64, help="HBM budget of the server")
    ap.add_argument("--sweep-seconds", type=float, default=8.0, help="window of each lower-rate run (0: none)")
    ap.add_argument("--check-torus", type=int, default=1 << 18, help="sampled nodes checked against FP64")
    ap.add_argument("--check-batch", type=int, default=1024, help="sampled nodes per batch mesh")
    ap.add_argument("--check-serve", type=int, default=512, help="sampled points per mesh (served)")
    ap.add_argument("--check-beams", type=int, default=1 << 16, help="sampled queries per beam query set")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="also run a plain copy kernel of this many MiB (traffic-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (queries shard by "
              f"replicating the mesh and splitting the queries)", file=sys.stderr)
        sys.exit(2)
    samples = Samples()
    if args.only == "batch":
        log(json.dumps(run_batch(args, samples)))
        return
    m, n, lo, hi = make_workload(args)
    N = n[0] * n[1] * n[2]
    log(f"[bench] mesh {m.name}: {m.ntris} triangles; grid {n}; steps {args.steps} warmup {args.warmup}")
    g = run_gpu(args, m, n, lo, hi, samples)
    if args.copy_calibration:
        import ssdf_amd as S
        nbytes = args.copy_calibration << 20
        src, dst = S.DeviceArray(nbytes // 4, np.float32), S.DeviceArray(nbytes // 4, np.float32)
        S._check(S.lib().ssdf_device_copy(dst.ptr, src.ptr, nbytes, None), "ssdf_device_copy")
        S._check(S.lib().ssdf_stream_synchronize(None), "sync")
        log(f"[bench] copy calibration: {nbytes} bytes read + written by k_copy16")
    if args.only:
        log(json.dumps(g))
        return
    errs = {}
    line = {
        "metric": METRIC, "value": round(g["fixup"]["mqps"], 2), "unit": "Mqueries/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(g["fixup"]["ms_per_step"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"noisy torus 200k tris x {n[0]}^3 grid, signed distance + closest triangle, "
                               "FP32 + FP64 fixup; step = in-place LBVH + pseudonormal rebuild + query, inputs and "
                               "outputs resident in HBM (host-to-host rates: host_call_*)",
                   "triangles": m.ntris, "queries": N, "mesh_checksum": m.checksum()},
        "plain": {"value": round(g["plain"]["mqps"], 2), "ms_per_step": round(g["plain"]["ms_per_step"], 4),
                  "query_ms": round(g["plain"]["query_ms"], 4)},
        "fixup_query_only": {"value": round(g["fixup"]["query_mqps"], 2), "query_ms": round(g["fixup"]["query_ms"], 4)},
        "host_call_plain": {"value": round(g["host_call_plain_mqps"], 2), "seconds": round(g["host_call_plain_s"], 4)},
        "host_call_fixup": {"value": round(g["host_call_fixup_mqps"], 2), "seconds": round(g["host_call_fixup_s"], 4)},
        "fixup_queued": g["fixup_queued"], "kernel_error": g["kernel_error"], "lbvh_maxdepth": g["maxdepth"],
        "build_ms": round(g["build_ms"], 3),
    }
    if args.grid == 256:
        line["roofline"] = roofline(g["fixup"]["query_ms"], N, m.nverts, m.ntris)
    if not args.no_batch:
        line["batch64"] = run_batch(args, samples)
    line["served"] = None if args.no_serve else run_serve(args, samples)
    if not args.no_beams:
        line["beams200"] = run_beams(args, samples)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args, m, n, lo, hi, samples, errs)
        line["error_vs_fp64"] = errs
        for key, leg in (("fixup_query_only", "torus200k_fixup"), ("plain", "torus200k_plain")):
            line[key]["error_vs_fp64"] = errs.get(leg)
        if line.get("batch64"):
            line["batch64"]["error_vs_fp64"] = errs.get("batch64")
        if line.get("served"):
            line["served"]["error_vs_fp64"] = errs.get("served")
        if line.get("beams200"):
            line["beams200"]["closed"]["error_vs_fp64"] = errs.get("beams200_closed")
            line["beams200"]["scattered"]["error_vs_fp64"] = errs.get("beams200_scattered")
            line["beams200"]["holed_gwn"]["error_vs_fp64"] = errs.get("beams200_holed_gwn")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
