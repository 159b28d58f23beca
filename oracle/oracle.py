"""ctypes wrapper of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module; the
product (ssdf_amd) never does.  See oracle.h for what each routine restates.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_build" / "liboracle.so"
REF_LIB = HERE / "_ref" / "libref_edf.so"

_P = C.c_void_p
_PD = C.c_ssize_t
_lib = None
_ref = None


def _load():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise ImportError(f"oracle not built at {LIB}; run `make -C oracle`")
        L = C.CDLL(str(LIB))
        sig = {
            "orc_pt_tri_d2_f": ([_P] * 4, C.c_float), "orc_pt_tri_d2_d": ([_P] * 4, C.c_double),
            "orc_pt_tri_cp_f": ([_P] * 5, None), "orc_pt_tri_cp_d": ([_P] * 5, None),
            "orc_pt_box_d2_f": ([C.c_float] * 9, C.c_float), "orc_pt_box_d2_d": ([C.c_double] * 9, C.c_double),
            "orc_pt_box_sdist_f": ([C.c_float] * 9, C.c_float),
            "orc_pt_box_sdist_d": ([C.c_double] * 9, C.c_double),
            "orc_compute_aabb_f": ([_PD, _P, _P, _P, _P], None), "orc_compute_aabb_d": ([_PD, _P, _P, _P, _P], None),
            "orc_all_points_aabb_sdist_f": ([_PD, _P, _P, _P, _P, _P], None),
            "orc_all_points_aabb_sdist_d": ([_PD, _P, _P, _P, _P, _P], None),
            "orc_edf_f": ([_PD, _P, _P, _P, _PD, _P, _P, _P, _PD, _P, _P, _P, _P], C.c_int),
            "orc_edf_d": ([_PD, _P, _P, _P, _PD, _P, _P, _P, _PD, _P, _P, _P, _P], C.c_int),
            "orc_celllist_f": ([_PD, _P, _P, _P, _PD, _P, _P, _P, _PD, _P, _P, _P, _P], C.c_int),
            "orc_celllist_d": ([_PD, _P, _P, _P, _PD, _P, _P, _P, _PD, _P, _P, _P, _P], C.c_int),
            "orc_brute_argmin": ([_PD, _P, _P, _P, _PD, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P], None),
            "orc_pair_d2": ([_PD, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P], None),
            "orc_bvh_argmin": ([_PD, _P, _P, _P, _PD, _P, _P, _P, _PD, _P, _P, _P, _P, _P], None),
            "orc_pseudonormals": ([_PD, _P, _P, _P, _PD, _P, _P, _P, _P, _P, _P], _PD),
            "orc_sign": ([_PD, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P], None),
            "orc_gwn": ([_PD, _P, _P, _P, _PD, _P, _P, _P, _P, _P, _P, _P], None),
            "orc_lbvh_counts": ([_PD, _P, _P, _P, _PD, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P], None),
            "orc_packet_counts": ([_PD, _P, _P, _P, _P, _P, _P, C.c_int, C.c_int, C.c_int, _P, _P, C.c_int, _P],
                                  None),
            "orc_lbvh_order": ([_PD, _P, _P, _P, _P, _P, _P, _P], None),
            "orc_num_threads": ([], C.c_int),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def ref_lib():
    """The shim over the reference's own serial/edf/edf.hpp, or None if it was not built."""
    global _ref
    if _ref is None and REF_LIB.exists():
        L = C.CDLL(str(REF_LIB))
        for n in ("ref_point_aabb_distance_squared_d", "ref_point_aabb_signed_distance_d"):
            getattr(L, n).argtypes = [C.c_double] * 9
            getattr(L, n).restype = C.c_double
        for n in ("ref_point_aabb_distance_squared_f", "ref_point_aabb_signed_distance_f"):
            getattr(L, n).argtypes = [C.c_float] * 9
            getattr(L, n).restype = C.c_float
        L.ref_point_triangle_barycentric_d.argtypes = [_P] * 5
        L.ref_point_triangle_barycentric_d.restype = C.c_double
        L.ref_point_triangle_barycentric_f.argtypes = [_P] * 5
        L.ref_point_triangle_barycentric_f.restype = C.c_float
        L.ref_all_points_aabb_signed_distance_d.argtypes = [_PD, _P, _P, _P, _P, _P]
        L.ref_all_points_aabb_signed_distance_d.restype = None
        _ref = L
    return _ref


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays must be contiguous"
    return C.c_void_p(a.ctypes.data)


def _f64(*arrs):
    return [np.ascontiguousarray(a, dtype=np.float64) for a in arrs]


def _i32(*arrs):
    return [np.ascontiguousarray(a, dtype=np.int32) for a in arrs]


def num_threads() -> int:
    return int(_load().orc_num_threads())


def pt_tri_d2(p, a, b, c, dtype=np.float64):
    L = _load()
    arrs = [np.ascontiguousarray(v, dtype=dtype) for v in (p, a, b, c)]
    f = L.orc_pt_tri_d2_d if dtype == np.float64 else L.orc_pt_tri_d2_f
    return f(*[_p(v) for v in arrs])


def pt_tri_cp(p, a, b, c, dtype=np.float64):
    L = _load()
    arrs = [np.ascontiguousarray(v, dtype=dtype) for v in (p, a, b, c)]
    out = np.zeros(3, dtype=dtype)
    f = L.orc_pt_tri_cp_d if dtype == np.float64 else L.orc_pt_tri_cp_f
    f(*[_p(v) for v in arrs], _p(out))
    return out


def pt_box_d2(p, lo, hi, dtype=np.float64):
    L = _load()
    f = L.orc_pt_box_d2_d if dtype == np.float64 else L.orc_pt_box_d2_f
    return f(p[0], p[1], p[2], lo[0], hi[0], lo[1], hi[1], lo[2], hi[2])


def pt_box_sdist(p, lo, hi, dtype=np.float64):
    L = _load()
    f = L.orc_pt_box_sdist_d if dtype == np.float64 else L.orc_pt_box_sdist_f
    return f(p[0], p[1], p[2], lo[0], hi[0], lo[1], hi[1], lo[2], hi[2])


def compute_aabb(x, y, z):
    """(xmin, xmax, ymin, ymax, zmin, zmax) in the points' precision (serial/edf/edf.hpp:500-539)."""
    L = _load()
    dt = np.asarray(x).dtype
    x, y, z = (np.ascontiguousarray(v, dtype=dt) for v in (x, y, z))
    box = np.empty(6, dt)
    f = L.orc_compute_aabb_d if dt == np.float64 else L.orc_compute_aabb_f
    f(x.size, _p(x), _p(y), _p(z), _p(box))
    return box


def all_points_aabb_sdist(x, y, z, box6):
    """Signed distance of each point to box6 (serial/edf/edf.hpp:541-557), in the points' precision."""
    L = _load()
    dt = np.asarray(x).dtype
    x, y, z = (np.ascontiguousarray(v, dtype=dt) for v in (x, y, z))
    box = np.ascontiguousarray(box6, dtype=dt)
    out = np.empty(x.size, dt)
    f = L.orc_all_points_aabb_sdist_d if dt == np.float64 else L.orc_all_points_aabb_sdist_f
    f(x.size, _p(x), _p(y), _p(z), _p(box), _p(out))
    return out


def _engine(name, x, y, z, mesh, out):
    L = _load()
    dt = out.dtype
    f = getattr(L, f"orc_{name}_{'f' if dt == np.float32 else 'd'}")
    x, y, z = [np.ascontiguousarray(v, dtype=dt) for v in (x, y, z)]
    sx, sy, sz = [np.ascontiguousarray(v, dtype=dt) for v in (mesh.sx, mesh.sy, mesh.sz)]
    s0, s1, s2 = _i32(mesh.s0, mesh.s1, mesh.s2)
    rc = f(x.size, _p(x), _p(y), _p(z), s0.size, _p(s0), _p(s1), _p(s2), sx.size, _p(sx), _p(sy), _p(sz), _p(out))
    assert rc == 0
    return out


def edf(x, y, z, mesh, out):
    """Restated ssdf::edf (serial/ssdf.hpp:447-619); precision from out.dtype; out in/out."""
    return _engine("edf", x, y, z, mesh, out)


def celllist(x, y, z, mesh, out):
    """Restated ssdf::edf_celllist (serial/ssdf.hpp:642-963); precision from out.dtype; out in/out."""
    return _engine("celllist", x, y, z, mesh, out)


def brute_argmin(x, y, z, mesh, init_d2=None):
    """FP64 brute-force argmin: (d2, tri, ntie, d2_second)."""
    L = _load()
    x, y, z = _f64(x, y, z)
    sx, sy, sz = _f64(mesh.sx, mesh.sy, mesh.sz)
    s0, s1, s2 = _i32(mesh.s0, mesh.s1, mesh.s2)
    n = x.size
    d2 = np.empty(n)
    tri = np.empty(n, np.int32)
    ntie = np.empty(n, np.int32)
    sec = np.empty(n)
    init = None if init_d2 is None else np.ascontiguousarray(init_d2, dtype=np.float64)
    L.orc_brute_argmin(n, _p(x), _p(y), _p(z), s0.size, _p(s0), _p(s1), _p(s2), _p(sx), _p(sy), _p(sz), _p(init),
                       _p(d2), _p(tri), _p(ntie), _p(sec))
    return d2, tri, ntie, sec


def pair_d2(x, y, z, tri, mesh):
    """FP64 squared distance of each point to its given triangle (-1: +inf)."""
    L = _load()
    x, y, z = _f64(x, y, z)
    sx, sy, sz = _f64(mesh.sx, mesh.sy, mesh.sz)
    s0, s1, s2 = _i32(mesh.s0, mesh.s1, mesh.s2)
    t = np.ascontiguousarray(tri, dtype=np.int32)
    out = np.empty(x.size)
    L.orc_pair_d2(x.size, _p(x), _p(y), _p(z), _p(t), _p(s0), _p(s1), _p(s2), _p(sx), _p(sy), _p(sz), _p(out))
    return out


def bvh_argmin(x, y, z, mesh):
    """Exact FP64 argmin through a CPU BVH (same results as brute_argmin)."""
    L = _load()
    x, y, z = _f64(x, y, z)
    sx, sy, sz = _f64(mesh.sx, mesh.sy, mesh.sz)
    s0, s1, s2 = _i32(mesh.s0, mesh.s1, mesh.s2)
    n = x.size
    d2 = np.empty(n)
    tri = np.empty(n, np.int32)
    L.orc_bvh_argmin(n, _p(x), _p(y), _p(z), s0.size, _p(s0), _p(s1), _p(s2), sx.size, _p(sx), _p(sy), _p(sz),
                     _p(d2), _p(tri))
    return d2, tri


def pseudonormals(mesh):
    L = _load()
    sx, sy, sz = _f64(mesh.sx, mesh.sy, mesh.sz)
    s0, s1, s2 = _i32(mesh.s0, mesh.s1, mesh.s2)
    T, V = s0.size, sx.size
    fn, en, vn = np.empty(3 * T), np.empty(9 * T), np.empty(3 * V)
    open_edges = L.orc_pseudonormals(T, _p(s0), _p(s1), _p(s2), V, _p(sx), _p(sy), _p(sz), _p(fn), _p(en), _p(vn))
    return fn, en, vn, int(open_edges)


def sign(x, y, z, tri, mesh, normals=None):
    """Pseudonormal sign at points given their closest triangles: (sign, region, dot)."""
    L = _load()
    if normals is None:
        normals = pseudonormals(mesh)
    fn, en, vn = normals[:3]
    x, y, z = _f64(x, y, z)
    sx, sy, sz = _f64(mesh.sx, mesh.sy, mesh.sz)
    s0, s1, s2 = _i32(mesh.s0, mesh.s1, mesh.s2)
    tri = np.ascontiguousarray(tri, dtype=np.int32)
    n = x.size
    sg = np.empty(n, np.int32)
    rg = np.empty(n, np.int32)
    dot = np.empty(n)
    L.orc_sign(n, _p(x), _p(y), _p(z), _p(tri), _p(s0), _p(s1), _p(s2), _p(sx), _p(sy), _p(sz), _p(fn), _p(en),
               _p(vn), _p(sg), _p(rg), _p(dot))
    return sg, rg, dot


def sdf(x, y, z, mesh, use_bvh=True, normals=None):
    """Signed distance ground truth: exact FP64 argmin + pseudonormal sign.
    Returns (signed distance, tri, sign, region, dot)."""
    if use_bvh:
        d2, tri = bvh_argmin(x, y, z, mesh)
    else:
        d2, tri, _, _ = brute_argmin(x, y, z, mesh)
    sg, rg, dot = sign(x, y, z, tri, mesh, normals)
    return sg * np.sqrt(d2), tri, sg, rg, dot


def gwn(x, y, z, mesh):
    L = _load()
    x, y, z = _f64(x, y, z)
    sx, sy, sz = _f64(mesh.sx, mesh.sy, mesh.sz)
    s0, s1, s2 = _i32(mesh.s0, mesh.s1, mesh.s2)
    w = np.empty(x.size)
    L.orc_gwn(x.size, _p(x), _p(y), _p(z), s0.size, _p(s0), _p(s1), _p(s2), _p(sx), _p(sy), _p(sz), _p(w))
    return w


def lbvh_counts(x, y, z, mesh):
    """Canonical FP32 LBVH replay: per-query (box tests, triangle tests), max depth, FP32 d2."""
    L = _load()
    x, y, z = [np.ascontiguousarray(v, dtype=np.float32) for v in (x, y, z)]
    sx, sy, sz = [np.ascontiguousarray(v, dtype=np.float32) for v in (mesh.sx, mesh.sy, mesh.sz)]
    s0, s1, s2 = _i32(mesh.s0, mesh.s1, mesh.s2)
    n = x.size
    nb = np.empty(n, np.int64)
    nt = np.empty(n, np.int64)
    md = C.c_int(0)
    d2 = np.empty(n, np.float32)
    L.orc_lbvh_counts(n, _p(x), _p(y), _p(z), s0.size, _p(s0), _p(s1), _p(s2), _p(sx), _p(sy), _p(sz), _p(nb),
                      _p(nt), C.byref(md), _p(d2))
    return nb, nt, int(md.value), d2


def packet_counts(mesh, n, lo, hi, every=1):
    """Replay of the GPU's packet walk over the grid's 4x4x4 bricks (every `every`-th):
    dict(bricks, visits, pops, leaf_refs, tri_evals)."""
    L = _load()
    sx, sy, sz = [np.ascontiguousarray(v, dtype=np.float32) for v in (mesh.sx, mesh.sy, mesh.sz)]
    s0, s1, s2 = _i32(mesh.s0, mesh.s1, mesh.s2)
    lo_ = np.ascontiguousarray(lo, dtype=np.float32)
    hi_ = np.ascontiguousarray(hi, dtype=np.float32)
    out = np.zeros(5, np.int64)
    L.orc_packet_counts(s0.size, _p(s0), _p(s1), _p(s2), _p(sx), _p(sy), _p(sz), int(n[0]), int(n[1]), int(n[2]),
                        _p(lo_), _p(hi_), int(every), _p(out))
    return dict(zip(("bricks", "visits", "pops", "leaf_refs", "tri_evals"), (int(v) for v in out)))


def lbvh_order(mesh):
    """Leaf order of the canonical LBVH: triangle index per leaf (stable sort by Morton code)."""
    L = _load()
    sx, sy, sz = [np.ascontiguousarray(v, dtype=np.float32) for v in (mesh.sx, mesh.sy, mesh.sz)]
    s0, s1, s2 = _i32(mesh.s0, mesh.s1, mesh.s2)
    out = np.empty(s0.size, np.int32)
    L.orc_lbvh_order(s0.size, _p(s0), _p(s1), _p(s2), _p(sx), _p(sy), _p(sz), _p(out))
    return out
