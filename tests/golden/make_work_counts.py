"""Writes tests/golden/work_counts.json: the canonical algorithmic work of the benchmark workloads,
and the work the GPU's packet walk executes for the same queries.

canonical: the oracle's FP32 LBVH replay (oracle/oracle.c: orc_lbvh_counts — Morton/Karras tree,
  leaf size 1, nearest-child-first, strict '<' pruning from +inf) counts the child-box tests and
  point-triangle tests of every query.  bench.py turns them into the roofline's algorithmic FLOPs:
  8 per box test, 81 per triangle test (serial/ssdf.hpp:50-121 face path), 8 per signed query.
packet: the oracle's replay of the GPU's wave-coherent walk (orc_packet_counts: 64 lanes of a
  4x4x4 brick share one walk, runs of up to 3 leaves) over every `every`-th brick, scaled to all
  bricks: executed box tests = 64 (2 visits + pops), executed triangle tests = 64 triangle
  evaluations; useful fraction = canonical / executed.
Run:  python tests/golden/make_work_counts.py
"""
import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
from oracle import oracle as O  # noqa: E402
from ssdf_amd import meshgen as mg  # noqa: E402
from ssdf_amd import workloads as W  # noqa: E402

OUT = Path(__file__).with_name("work_counts.json")


def canonical(mesh, n, lo, hi):
    x, y, z = mg.grid_points(*n, lo, hi)
    nb, nt, md, _ = O.lbvh_counts(x, y, z, mesh)
    return int(x.size), int(nb.sum()), int(nt.sum()), md


def packet(mesh, n, lo, hi, every):
    r = O.packet_counts(mesh, n, np.asarray(lo, np.float32), np.asarray(hi, np.float32), every)
    nbx = [(v + 3) // 4 for v in n]
    scale = (nbx[0] * nbx[1] * nbx[2]) / r["bricks"]
    return {k: v * scale for k, v in r.items() if k != "bricks"}


def entry(parts, seconds, checksum=None):
    q = sum(p[0] for p in parts)
    nb = sum(p[1] for p in parts)
    nt = sum(p[2] for p in parts)
    e = {"queries": q, "box_tests": nb, "tri_tests": nt, "maxdepth": max(p[3] for p in parts)}
    if checksum:
        e["mesh_checksum"] = checksum
    pk = {k: sum(p[4][k] for p in parts) for k in ("visits", "pops", "leaf_refs", "tri_evals")}
    ex_box = 64 * (2 * pk["visits"] + pk["pops"])
    ex_tri = 64 * pk["tri_evals"]
    e["packet"] = {"every": parts[0][5], **{k: int(round(v)) for k, v in pk.items()},
                   "executed_box_tests": int(round(ex_box)), "executed_tri_tests": int(round(ex_tri)),
                   "useful_box_frac": round(nb / ex_box, 4), "useful_tri_frac": round(nt / ex_tri, 4)}
    e["seconds"] = round(seconds, 1)
    return e


def one(mesh, n, lo, hi, every):
    return (*canonical(mesh, n, lo, hi), packet(mesh, n, lo, hi, every), every)


def main():
    res = {}
    t = time.time()
    m = mg.uv_sphere()
    res["uv_sphere_48x24_grid64"] = entry([one(m, (64, 64, 64), (-1, -1, -1), (1, 1, 1), 1)], time.time() - t,
                                          m.checksum())
    print(res, flush=True)
    t = time.time()
    torus = mg.torus(0.6, 0.2, 500, 200, sigma=0.005, seed=2)
    res["torus_200k_grid256"] = entry([one(torus, (256, 256, 256), (-1, -1, -1), (1, 1, 1), 1)], time.time() - t,
                                      torus.checksum())
    print(res, flush=True)
    t = time.time()
    meshes, dims, boxes = W.batch64()
    parts = []
    for i, (mm, d, bx) in enumerate(zip(meshes, dims, boxes)):
        n = tuple(int(v) for v in d)
        parts.append(one(mm, n, bx[:3].astype(np.float32), bx[3:].astype(np.float32), 4))
        print(i, parts[-1][:3], flush=True)
    res["batch64_icospheres_grids"] = entry(parts, time.time() - t)
    print(res, flush=True)
    OUT.write_text(json.dumps(res, indent=1) + "\n")


if __name__ == "__main__":
    main()
