"""The reference's on-disk SoA format (drivers/sdf.exe.cpp:206-211, README.md:163-166): one raw,
native-endian array per file, float32 coordinates x, y, z and int32 triangle corners i0, i1, i2."""
from __future__ import annotations

from pathlib import Path

import numpy as np

LIB = Path(__file__).resolve().parent / "_lib"
SDF_EXE = LIB / "ssdf-sdf"
MESH_TO_SDF_EXE = LIB / "ssdf-mesh_to_sdf"
POINTS_INSIDE_EXE = LIB / "ssdf-points_inside"
SDF_SIGN_EXE = LIB / "ssdf-sdf_sign"
SQEDF_EXE = LIB / "ssdf-sqedf"


def write_surface(folder, m, ext: str = "typed") -> None:
    """Writes mesh m as x/y/z.float32 and i0/i1/i2.int32 (ext="typed") or as <c>.raw (ext="raw")."""
    d = Path(folder)
    d.mkdir(parents=True, exist_ok=True)
    for c, a in (("x", m.sx), ("y", m.sy), ("z", m.sz)):
        np.ascontiguousarray(a, np.float32).tofile(d / (c + (".float32" if ext == "typed" else ".raw")))
    for c, a in (("i0", m.s0), ("i1", m.s1), ("i2", m.s2)):
        np.ascontiguousarray(a, np.int32).tofile(d / (c + (".int32" if ext == "typed" else ".raw")))


def write_points(folder, x, y, z, ext: str = "typed") -> None:
    d = Path(folder)
    d.mkdir(parents=True, exist_ok=True)
    for c, a in (("x", x), ("y", y), ("z", z)):
        np.ascontiguousarray(a, np.float32).tofile(d / (c + (".float32" if ext == "typed" else ".raw")))


def read_field(path) -> np.ndarray:
    """float32[npoints] as written by the sdf command."""
    return np.fromfile(path, dtype=np.float32)


def read_grid(folder):
    """The mesh_to_sdf command's output: (field float32[nz, ny, nx], dims (nx, ny, nz), box
    (xmin, ymin, zmin, xmax, ymax, zmax) as float32)."""
    d = Path(folder)
    meta = dict(line.split() for line in (d / "grid.txt").read_text().splitlines() if line.strip())
    dims = tuple(int(meta[k]) for k in ("nx", "ny", "nz"))
    box = tuple(np.float32(float(meta[k])) for k in ("xmin", "ymin", "zmin", "xmax", "ymax", "zmax"))
    f = np.fromfile(d / "sdf.float32", dtype=np.float32)
    return f.reshape(dims[2], dims[1], dims[0]), dims, box
