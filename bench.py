#!/usr/bin/env python3
"""Benchmark of the closest-point / signed-distance hot path on one MI355X.

Workload (default): noisy torus, R=0.6 r=0.2, 500 x 200 segments = 200,000 triangles, every vertex
displaced along its normal by N(0, 0.005^2) (SplitMix64 seed 2); query set: the 256^3 grid on
[-1,1]^3 (16,777,216 FP32 nodes); output: signed distance + closest triangle per node.

A step is one full job on inputs already resident in HBM: LBVH build + FP64 pseudonormals from the
device mesh arrays, then the 256^3 signed query in FP32 + FP64 fixup, outputs written to HBM.
`value` = queries / step time.  Also reported: the same step in plain FP32 (no fixup), the query
alone on a resident BVH, the reference-shaped host call (upload, build, query, download), the
roofline of the query kernel and the CPU baseline (the oracle's restatement of
edf_celllist<float>, timed on this host on a seeded subsample of the same grid).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mqueries/s @256^3×200k tris (FP32+fixup vs plain); served p50/p99 latency"
FLOPS_BOX, FLOPS_TRI, FLOPS_SIGN = 8, 81, 8
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_workload(args):
    from ssdf_amd import meshgen as mg
    m = mg.torus(0.6, 0.2, 500, 200, sigma=0.005, seed=2)
    n = (args.grid,) * 3
    return m, n, (-1.0, -1.0, -1.0), (1.0, 1.0, 1.0)


def run_gpu(args, m, n, lo, hi):
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
    t0 = time.perf_counter()
    S.sample_sdf_bvh(m.s0, m.s1, m.s2, m.sx, m.sy, m.sz, n[0], n[1], n[2], *lo, *hi, outh, None, flags=0)
    res["host_call_plain_s"] = time.perf_counter() - t0
    res["host_call_plain_mqps"] = N / res["host_call_plain_s"] / 1e6
    return res


def run_batch(args):
    """64 displaced icospheres (4.1 M triangles), one grid each (84.8 M nodes), signed distance +
    triangle with FP64 fixup: the per-mesh reference-shaped call vs one batch."""
    import ssdf_amd as S
    from ssdf_amd import workloads as W
    meshes, dims, boxes = W.batch64()
    nq = int(dims.prod(1).sum())
    flags = S.FLAG_FIXUP
    res = {"meshes": len(meshes), "triangles": int(sum(m.ntris for m in meshes)), "queries": nq}
    # per-mesh plain path: upload, build, query, download, free per call (serial/bvh/bvh.cu:80-158)
    outs = [np.empty(int(d.prod()), np.float32) for d in dims]
    t0 = time.perf_counter()
    for m, d, bx, o in zip(meshes, dims, boxes, outs):
        S.sample_sdf_bvh(m.s0, m.s1, m.s2, m.sx, m.sy, m.sz, int(d[0]), int(d[1]), int(d[2]), *bx.astype(np.float32),
                         o, None, flags=flags)
    res["per_mesh_calls_s"] = time.perf_counter() - t0
    res["per_mesh_calls_mqps"] = nq / res["per_mesh_calls_s"] / 1e6
    log(f"[bench] batch64 per-mesh calls: {res['per_mesh_calls_s']:.3f} s")
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
    bpc, ch, tune_ms = b.autotune(S.MODE_SDF, dims, boxes, off, out, tri, flags=flags, stream=st)
    res["schedule"] = {"blocks_per_cu": bpc, "units_per_grab": ch, "query_ms": round(tune_ms, 3)}
    ev = [S.Event(), S.Event()]
    for rebuild, key in ((True, "resident_step"), (False, "resident_query")):
        b.capture(rebuild, S.MODE_SDF, dims, boxes, off, out, tri, flags=flags, stream=st)
        for _ in range(max(1, args.warmup)):
            b.replay(st)
        st.synchronize()
        ev[0].record(st)
        for _ in range(args.steps):
            b.replay(st)
        ev[1].record(st)
        st.synchronize()
        ms = ev[0].elapsed_ms(ev[1]) / args.steps
        res[key + "_ms"] = ms
        res[key + "_mqps"] = nq / (ms * 1e-3) / 1e6
        log(f"[bench] batch64 {key}: {ms:.3f} ms")
    q, e = b.last_stats()
    res["fixup_queued"], res["kernel_error"] = q, e
    res["speedup_resident_step_vs_per_mesh_calls"] = res["resident_step_mqps"] / res["per_mesh_calls_mqps"]
    res["speedup_host_to_host_vs_per_mesh_calls"] = res["batch_host_to_host_mqps"] / res["per_mesh_calls_mqps"]
    b.close()
    return {k: (round(v, 4) if isinstance(v, float) else v) for k, v in res.items()}


def run_serve(args):
    """Open-loop serving: Poisson arrivals (200 requests/s) over a pool of 32 resident meshes
    (5 k - 2 M triangles), 1e4 - 1e7 signed queries per request (half uniform in the padded bbox,
    half within 0.02 of the surface), FP32 + fixup; latency from the scheduled arrival to the
    results in host memory."""
    import ssdf_amd as S
    from ssdf_amd import workloads as W
    t_gen = time.perf_counter()
    meshes = W.serve_pool()
    log(f"[bench] serve: mesh pool generated ({time.perf_counter() - t_gen:.1f} s)")
    times, mesh_ids, nq = W.poisson_requests(200.0, args.serve_seconds, len(meshes))
    nmax = int(nq.max())
    base = 1 << 20
    pools = []
    for i, m in enumerate(meshes):
        pts = W.query_cloud(m, base, seed=100 + i)
        arrs = []
        for v in pts:
            pa = S.PinnedArray(nmax, np.float32)
            reps = -(-nmax // base)
            pa.array[:] = np.tile(v, reps)[:nmax]
            arrs.append(pa)
        pools.append(arrs)
    ring = 24
    outs = [(S.PinnedArray(nmax, np.float32), S.PinnedArray(nmax, np.int32)) for _ in range(ring)]
    gen_s = time.perf_counter() - t_gen
    log(f"[bench] serve: query pools and output ring pinned ({gen_s:.1f} s)")
    budget = 32 << 30
    srv = S.Server(budget, max_inflight=16)
    for i, m in enumerate(meshes):
        srv.add_mesh(i, m)
    # warm-up: one small request per mesh loads every mesh into HBM
    warm = [srv.submit(i, S.MODE_SDF, *[p.array[:1024] for p in pools[i]], outs[0][0].array[:1024],
                       outs[0][1].array[:1024], flags=S.FLAG_FIXUP) for i in range(len(meshes))]
    for t in warm:
        srv.wait(t)
    slots = [None] * ring
    lat = []
    t0 = srv.clock() + 0.05
    log(f"[bench] serve: {len(meshes)} meshes resident, {len(times)} requests over {args.serve_seconds} s")
    for k in range(len(times)):
        if k % 250 == 0:
            log(f"[bench] serve: request {k} at t={srv.clock() - t0:.2f} s")
        slot = k % ring
        if slots[slot] is not None:
            lat.append(srv.wait(slots[slot]))
            slots[slot] = None
        arrival = t0 + float(times[k])
        delay = arrival - srv.clock()
        if delay > 0:
            time.sleep(delay)
        i, n = int(mesh_ids[k]), int(nq[k])
        x, y, z = (p.array[:n] for p in pools[i])
        slots[slot] = srv.submit(i, S.MODE_SDF, x, y, z, outs[slot][0].array[:n], outs[slot][1].array[:n],
                                 flags=S.FLAG_FIXUP, arrival=arrival)
    log(f"[bench] serve: all submitted at t={srv.clock() - t0:.2f} s, draining")
    for tk in slots:
        if tk is not None:
            lat.append(srv.wait(tk))
    t_end = srv.clock()
    st = srv.stats()
    srv.close()
    lat = np.array(lat)
    total_q = int(nq.sum())
    res = {"requests": len(times), "seconds_offered": args.serve_seconds, "meshes": len(meshes),
           "triangles": int(sum(m.ntris for m in meshes)), "queries": total_q,
           "offered_mqps": round(total_q / args.serve_seconds / 1e6, 2),
           "throughput_mqps": round(total_q / (t_end - t0) / 1e6, 2),
           "p50_ms": round(float(np.percentile(lat, 50)) * 1e3, 2),
           "p99_ms": round(float(np.percentile(lat, 99)) * 1e3, 2),
           "max_ms": round(float(lat.max()) * 1e3, 2),
           "mesh_loads": int(st["mesh_loads"]), "evictions": int(st["evictions"]), "errors": int(st["errors"]),
           "hbm_budget_gib": budget / 2**30, "peak_gib": round(st["peak_bytes"] / 2**30, 3),
           "setup_s": round(gen_s, 1)}
    log(f"[bench] serve: {res}")
    return res


def run_beams(args):
    """200 capped cylinders in the unit cube (3.4 M triangles), intact and with 1% of the triangles
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
        log(f"[bench] beams {key}: grid {ms:.1f} ms")
        if key == "closed":
            n = args.beams_points
            P = np.random.default_rng(5).random((3, n), dtype=np.float32)
            dx, dy, dz = (S.DeviceArray.from_host(P[d]) for d in range(3))
            del P
            po = S.DeviceArray(n, np.float32)
            pt = S.DeviceArray(n, np.int32)
            ms = timed(lambda: rm.query_points(S.MODE_SDF, dx, dy, dz, po, pt, flags=flags, stream=st))
            q, e = rm.last_stats()
            res["scattered"] = {"points": n, "ms": round(ms, 2), "mqps": round(n / ms / 1e3, 1), "fixup_queued": q,
                                "kernel_error": e}
            log(f"[bench] beams scattered: {ms:.1f} ms")
            del dx, dy, dz, po, pt
        del rm, o, t
    a, b = outs["closed"], outs["holed_gwn"]
    away = np.abs(a) > 0.01
    res["holed_sign_agreement_beyond_0.01"] = float(np.mean(np.sign(a[away]) == np.sign(b[away])))
    return res


def cpu_baseline(args, m, n, lo, hi):
    """The oracle's edf_celllist<float> restatement (serial/ssdf.hpp:642-963) on a seeded uniform
    subsample of the grid, all host threads."""
    from oracle import oracle as O
    from ssdf_amd import meshgen as mg
    rng = mg.SplitMix64(12345)
    N = n[0] * n[1] * n[2]
    k = args.cpu_sample
    idx = np.unique((rng.uniform(k) * N).astype(np.int64))
    ax = [mg.grid_axis(n[d], lo[d], hi[d]) for d in range(3)]
    i, j, kk = idx % n[0], (idx // n[0]) % n[1], idx // (n[0] * n[1])
    x, y, z = ax[0][i], ax[1][j], ax[2][kk]
    out = np.full(idx.size, np.finfo(np.float32).max, np.float32)
    t0 = time.perf_counter()
    O.celllist(x, y, z, m, out)
    dt = time.perf_counter() - t0
    return {"value": idx.size / dt / 1e6, "unit": "Mqueries/s", "cores": O.num_threads(), "kind": "port",
            "sample": f"{idx.size} grid nodes drawn uniformly (SplitMix64 seed 12345) from the 256^3 grid; "
                      f"edf_celllist<float> restated in C, OpenMP, {dt:.2f} s incl. its cell build"}


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
    wc = json.loads((ROOT / "tests" / "golden" / "work_counts.json").read_text())["c2_torus_200k_grid256"]
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
            "hbm": {"achieved": round(hbm, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(hbm / PEAK_HBM_GBS, 5),
                    "algorithmic_bytes_per_launch": hbm_bytes}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--grid", type=int, default=256)
    ap.add_argument("--cpu-sample", type=int, default=1 << 23)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--only", choices=["fixup", "plain", "batch"], default=None,
                    help="time one mode (or the 64-mesh batch leg) only and print nothing else (for profiling)")
    ap.add_argument("--no-batch", action="store_true", help="skip the 64-mesh batch leg")
    ap.add_argument("--no-serve", action="store_true", help="skip the serving leg")
    ap.add_argument("--no-beams", action="store_true", help="skip the 200-beam stress leg")
    ap.add_argument("--beams-grid", type=int, default=512)
    ap.add_argument("--beams-points", type=int, default=100_000_000)
    ap.add_argument("--serve-seconds", type=float, default=4.0, help="arrival window of the serving leg")
    ap.add_argument("--copy-calibration", type=int, default=0,
                    help="also run a plain copy kernel of this many MiB (traffic-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported (queries shard by "
              f"replicating the mesh and splitting the queries)", file=sys.stderr)
        sys.exit(2)
    if args.only == "batch":
        log(json.dumps(run_batch(args)))
        return
    m, n, lo, hi = make_workload(args)
    N = n[0] * n[1] * n[2]
    log(f"[bench] mesh {m.name}: {m.ntris} triangles; grid {n}; steps {args.steps} warmup {args.warmup}")
    g = run_gpu(args, m, n, lo, hi)
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
    line = {
        "metric": METRIC, "value": round(g["fixup"]["mqps"], 2), "unit": "Mqueries/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(g["fixup"]["ms_per_step"], 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"noisy torus 200k tris x {n[0]}^3 grid, signed distance + closest triangle, "
                               "FP32 + FP64 fixup; step = in-place LBVH + pseudonormal rebuild + query",
                   "triangles": m.ntris, "queries": N, "mesh_checksum": m.checksum()},
        "plain": {"value": round(g["plain"]["mqps"], 2), "ms_per_step": round(g["plain"]["ms_per_step"], 4),
                  "query_ms": round(g["plain"]["query_ms"], 4)},
        "fixup_query_only": {"value": round(g["fixup"]["query_mqps"], 2), "query_ms": round(g["fixup"]["query_ms"], 4)},
        "host_call_plain": {"value": round(g["host_call_plain_mqps"], 2), "seconds": round(g["host_call_plain_s"], 4)},
        "fixup_queued": g["fixup_queued"], "kernel_error": g["kernel_error"], "lbvh_maxdepth": g["maxdepth"],
        "build_ms": round(g["build_ms"], 3),
    }
    if args.grid == 256:
        line["roofline"] = roofline(g["fixup"]["query_ms"], N, m.nverts, m.ntris)
    if not args.no_batch:
        line["batch64"] = run_batch(args)
    line["served"] = None if args.no_serve else run_serve(args)
    if not args.no_beams:
        line["beams200"] = run_beams(args)
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(args, m, n, lo, hi)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
