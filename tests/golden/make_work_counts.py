"""Writes tests/golden/work_counts.json: the canonical algorithmic work of each benchmark workload.

The oracle's FP32 LBVH replay (oracle/oracle.c: orc_lbvh_counts — Morton/Karras tree, leaf size 1,
nearest-child-first, strict '<' pruning from +inf) counts the child-box tests and point-triangle
tests of every query.  bench.py turns them into the roofline's algorithmic FLOPs:
8 per box test, 81 per triangle test (serial/ssdf.hpp:50-121 face path), 8 per signed query.
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

OUT = Path(__file__).with_name("work_counts.json")


def counts(mesh, n, lo, hi):
    x, y, z = mg.grid_points(*n, lo, hi)
    t = time.time()
    nb, nt, md, _ = O.lbvh_counts(x, y, z, mesh)
    return {"queries": int(x.size), "box_tests": int(nb.sum()), "tri_tests": int(nt.sum()), "maxdepth": md,
            "mesh_checksum": mesh.checksum(), "seconds": round(time.time() - t, 2)}


def main():
    res = {}
    res["c1_uv_sphere_48x24_grid64"] = counts(mg.uv_sphere(), (64, 64, 64), (-1, -1, -1), (1, 1, 1))
    print(res)
    torus = mg.torus(0.6, 0.2, 500, 200, sigma=0.005, seed=2)
    res["c2_torus_200k_grid256"] = counts(torus, (256, 256, 256), (-1, -1, -1), (1, 1, 1))
    print(res)
    OUT.write_text(json.dumps(res, indent=1) + "\n")


if __name__ == "__main__":
    main()
