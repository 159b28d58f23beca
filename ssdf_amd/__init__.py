"""ssdf_amd — MI355X-native closest-point / signed-distance engine.

Python mirror of zulianp/ssdf's hot-path interface over the C ABI of libssdf_amd.so (ctypes).
The functions keep the reference's names and argument meaning (serial/ssdf.hpp:427-460,
serial/bvh/bvh.hpp:19-227): SoA query coordinates, int32 triangle indices s0, s1, s2, SoA vertex
coordinates, caller-owned outputs, `out` in/out for the unsigned `edf` family.  Floating-point
precision follows the arrays' dtype (float32 -> the FP32 path, float64 -> the exact FP64 path).
Errors raise SsdfError (the reference returns an int; the C ABI here returns its code).

There is no CPU fallback: importing works anywhere, but every call needs the built library and a
GPU and fails loudly otherwise.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

__all__ = [
    "SsdfError", "lib", "available", "edf", "edf_celllist", "edf_select", "edf_bvh", "closest_point_bvh",
    "sample_sdf_bvh", "sdf", "ResidentMesh", "DeviceArray", "Stream", "Event", "FLAG_FIXUP", "FLAG_SIGN_FP64",
    "FLAG_SIGN_GWN", "FLAG_SIGN_AUTO", "points_inside_bvh", "sample_inside_bvh", "compute_aabb",
    "all_points_aabb_signed_distance", "init_volume_box",
    "MODE_EDF", "MODE_SDF", "MODE_CLOSEST", "MODE_SQEDF", "sample_sqedf_bvh", "Batch", "Server", "PinnedArray",
]

FLAG_FIXUP = 1
FLAG_SIGN_FP64 = 2
FLAG_SIGN_GWN = 8
FLAG_SIGN_AUTO = 16
NOALLOC_MAX_POINTS = 1 << 20  # SSDF_NOALLOC_MAX_POINTS
MODE_EDF, MODE_SDF, MODE_CLOSEST, MODE_SQEDF = 0, 1, 2, 3

_LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libssdf_amd.so"
_lib = None

_P = C.c_void_p
_PD = C.c_ssize_t  # ptrdiff_t


class SsdfError(RuntimeError):
    def __init__(self, code: int, where: str, msg: str = ""):
        super().__init__(f"{where} failed with code {code}: {msg}")
        self.code = code


class _ServerStats(C.Structure):
    _fields_ = [("served", C.c_ulonglong), ("queries", C.c_ulonglong), ("errors", C.c_ulonglong),
                ("mesh_loads", C.c_ulonglong), ("evictions", C.c_ulonglong), ("pending", C.c_size_t),
                ("inflight", C.c_size_t), ("bytes_in_use", C.c_size_t), ("peak_bytes", C.c_size_t),
                ("p50_s", C.c_double), ("p99_s", C.c_double), ("max_s", C.c_double), ("throughput_qps", C.c_double)]


class _MeshInfo(C.Structure):
    _fields_ = [("nselements", _PD), ("nspoints", _PD), ("dtype", C.c_int), ("maxdepth", C.c_int),
                ("open_edges", C.c_int), ("flipped_edges", C.c_int), ("bbox", C.c_float * 6),
                ("device_bytes", C.c_size_t)]


_SIGS = {
    "ssdf_edf_f": [_PD, _P, _P, _P, _PD, _P, _P, _P, _PD, _P, _P, _P, _P],
    "ssdf_edf_d": [_PD, _P, _P, _P, _PD, _P, _P, _P, _PD, _P, _P, _P, _P],
    "ssdf_closest_point_f": [_PD, _P, _P, _P, _PD, _P, _P, _P, _PD, _P, _P, _P, _P, _P, _P, _P, _P, C.c_uint],
    "ssdf_closest_point_d": [_PD, _P, _P, _P, _PD, _P, _P, _P, _PD, _P, _P, _P, _P, _P, _P, _P, _P],
    "ssdf_sample_sdf_f": [_PD, _P, _P, _P, _PD, _P, _P, _P, _PD, _PD, _PD, C.c_float, C.c_float, C.c_float,
                          C.c_float, C.c_float, C.c_float, _P, _P, C.c_uint],
    "ssdf_sample_sdf_d": [_PD, _P, _P, _P, _PD, _P, _P, _P, _PD, _PD, _PD, C.c_double, C.c_double, C.c_double,
                          C.c_double, C.c_double, C.c_double, _P, _P],
    "ssdf_sdf_f": [_PD, _P, _P, _P, _PD, _P, _P, _P, _PD, _P, _P, _P, _P, _P, C.c_uint],
    "ssdf_sample_sqedf_f": [_PD, _P, _P, _P, _PD, _P, _P, _P, _PD, _PD, _PD] + [C.c_float] * 6 + [_P, C.c_uint],
    "ssdf_sample_sqedf_d": [_PD, _P, _P, _P, _PD, _P, _P, _P, _PD, _PD, _PD] + [C.c_double] * 6 + [_P],
    "ssdf_points_inside_f": [_PD, _P, _P, _P, _PD, _P, _P, _P, _PD, _P, _P, _P, _P],
    "ssdf_points_inside_d": [_PD, _P, _P, _P, _PD, _P, _P, _P, _PD, _P, _P, _P, _P],
    "ssdf_sample_inside_f": [_PD, _P, _P, _P, _PD, _P, _P, _P, _PD, _PD, _PD] + [C.c_float] * 6 + [_P],
    "ssdf_sample_inside_d": [_PD, _P, _P, _P, _PD, _P, _P, _P, _PD, _PD, _PD] + [C.c_double] * 6 + [_P],
    "ssdf_mesh_winding_number": [_P, _PD, _P, _P, _P, _P, _P],
    "ssdf_mesh_inside_points": [_P, _PD, _P, _P, _P, _P, _P],
    "ssdf_mesh_inside_grid": [_P, _PD, _PD, _PD] + [C.c_double] * 6 + [_P, _P],
    "ssdf_compute_aabb_f": [_PD, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "ssdf_compute_aabb_d": [_PD, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "ssdf_all_points_aabb_signed_distance_f": [_PD, _P, _P, _P] + [C.c_float] * 6 + [_P],
    "ssdf_all_points_aabb_signed_distance_d": [_PD, _P, _P, _P] + [C.c_double] * 6 + [_P],
    "ssdf_init_volume_box": [C.c_int, _PD, _P, _P, _P, _P, _P],
    "ssdf_mesh_create": [C.c_int, _PD, _P, _P, _P, _PD, _P, _P, _P, C.c_int, C.c_int, _P, C.POINTER(_P)],
    "ssdf_mesh_destroy": [_P],
    "ssdf_mesh_rebuild": [_P, C.c_int, _P],
    "ssdf_mesh_update_vertices": [_P, _P, _P, _P, C.c_int, C.c_int, _P],
    "ssdf_mesh_info": [_P, C.POINTER(_MeshInfo)],
    "ssdf_mesh_query_points": [_P, C.c_int, _PD, _P, _P, _P, _P, _P, _P, _P, _P, C.c_uint, _P],
    "ssdf_mesh_query_points_noalloc": [_P, C.c_int, _PD, _P, _P, _P, _P, _P, _P, _P, _P, C.c_uint, _P],
    "ssdf_mesh_query_grid": [_P, C.c_int, _PD, _PD, _PD, C.c_double, C.c_double, C.c_double, C.c_double,
                             C.c_double, C.c_double, _P, _P, C.c_uint, _P],
    "ssdf_mesh_last_stats": [_P, C.POINTER(C.c_uint), C.POINTER(C.c_int)],
    "ssdf_mesh_set_schedule": [_P, C.c_int, C.c_int, C.c_int, C.c_int],
    "ssdf_mesh_brute_points": [_P, _PD, _P, _P, _P, _P, _P, _P],
    "ssdf_mesh_leaf_order": [_P, _P, _P],
    "ssdf_batch_create": [C.c_int, _P, _P, _P, _P, _P, _P, _P, _P, C.c_int, C.c_int, _P, C.POINTER(_P)],
    "ssdf_batch_destroy": [_P],
    "ssdf_batch_rebuild": [_P, C.c_int, _P],
    "ssdf_batch_query_grids": [_P, C.c_int, _P, _P, _P, _P, _P, C.c_uint, _P],
    "ssdf_batch_query_points": [_P, C.c_int, _P, _P, _P, _P, _P, _P, _P, _P, _P, C.c_uint, _P],
    "ssdf_batch_set_schedule": [_P, C.c_int, C.c_int],
    "ssdf_batch_set_mesh_schedule": [_P, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int],
    "ssdf_batch_get_mesh_schedule": [_P, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                     C.POINTER(C.c_int)],
    "ssdf_batch_autotune": [_P, C.c_int, _P, _P, _P, _P, _P, C.c_uint, _P, C.POINTER(C.c_int), C.POINTER(C.c_int),
                            C.POINTER(C.c_float)],
    "ssdf_batch_capture": [_P, C.c_int, C.c_int, _P, _P, _P, _P, _P, C.c_uint, _P],
    "ssdf_batch_replay": [_P, _P],
    "ssdf_batch_last_stats": [_P, C.POINTER(C.c_uint), C.POINTER(C.c_int)],
    "ssdf_batch_info": [_P, C.c_int, C.POINTER(_MeshInfo)],
    "ssdf_server_create": [C.c_size_t, C.c_int, C.POINTER(_P)],
    "ssdf_server_add_mesh": [_P, C.c_int, _PD, _P, _P, _P, _PD, _P, _P, _P],
    "ssdf_server_charge": [_P, C.c_int, _PD, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t)],
    "ssdf_server_clock": [_P, C.POINTER(C.c_double)],
    "ssdf_server_submit": [_P, C.c_int, C.c_int, _PD, _P, _P, _P, _P, _P, C.c_uint, C.c_double,
                           C.POINTER(C.c_uint64)],
    "ssdf_server_wait": [_P, C.c_uint64, C.POINTER(C.c_double)],
    "ssdf_server_poll": [_P, C.c_uint64, C.POINTER(C.c_int)],
    "ssdf_server_stats": [_P, C.c_void_p],
    "ssdf_server_destroy": [_P],
    "ssdf_host_alloc": [C.POINTER(_P), C.c_size_t],
    "ssdf_host_free": [_P],
    "ssdf_device_count": [C.POINTER(C.c_int)],
    "ssdf_device_malloc": [C.POINTER(_P), C.c_size_t],
    "ssdf_device_free": [_P],
    "ssdf_memcpy_h2d": [_P, _P, C.c_size_t],
    "ssdf_memcpy_d2h": [_P, _P, C.c_size_t],
    "ssdf_device_copy": [_P, _P, C.c_size_t, _P],
    "ssdf_stream_create": [C.POINTER(_P)],
    "ssdf_stream_destroy": [_P],
    "ssdf_stream_synchronize": [_P],
    "ssdf_event_create": [C.POINTER(_P)],
    "ssdf_event_destroy": [_P],
    "ssdf_event_record": [_P, _P],
    "ssdf_event_elapsed_ms": [_P, _P, C.POINTER(C.c_float)],
    "ssdf_set_heavy_leaves": [C.c_int],
    "ssdf_last_error": [],
    "ssdf_version": [],
}


def lib() -> C.CDLL:
    """The loaded C library; raises if it has not been built."""
    global _lib
    if _lib is None:
        path = Path(os.environ.get("SSDF_AMD_LIB", _LIB_PATH))
        if not path.exists():
            raise ImportError(f"libssdf_amd.so not built at {path}; run __graft_entry__.build()")
        L = C.CDLL(str(path))
        for name, args in _SIGS.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = C.c_char_p if name in ("ssdf_last_error", "ssdf_version") else C.c_int
        _lib = L
    return _lib


def _check(rc: int, where: str):
    if rc != 0:
        msg = lib().ssdf_last_error()
        raise SsdfError(rc, where, msg.decode() if msg else "")


def available() -> bool:
    """True when the library is built and a GPU is visible."""
    try:
        n = C.c_int(0)
        lib().ssdf_device_count(C.byref(n))
        return n.value > 0
    except (ImportError, OSError):
        return False


def set_heavy_leaves(leaves: int) -> int:
    """Leaf budget of a packet wave before its queries go to the per-query kernel (0: never);
    process-wide, returns the previous value.  A negative value only reads it."""
    return int(lib().ssdf_set_heavy_leaves(int(leaves)))


def _ptr(a) -> C.c_void_p | None:
    if a is None:
        return None
    if isinstance(a, DeviceArray):
        return C.c_void_p(a.ptr)
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("arrays must be C-contiguous")
    return C.c_void_p(a.ctypes.data)


# ------------------------------------------------------------------------------------------------
# device plumbing
# ------------------------------------------------------------------------------------------------
class DeviceArray:
    """A flat device buffer with a numpy dtype (host <-> device through the C ABI)."""

    def __init__(self, n: int, dtype):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        p = C.c_void_p()
        _check(lib().ssdf_device_malloc(C.byref(p), max(1, self.n * self.dtype.itemsize)), "ssdf_device_malloc")
        self.ptr = p.value

    @property
    def size(self) -> int:
        return self.n

    @classmethod
    def from_host(cls, a: np.ndarray) -> "DeviceArray":
        a = np.ascontiguousarray(a)
        d = cls(a.size, a.dtype)
        d.upload(a)
        return d

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        if self.n:
            _check(lib().ssdf_memcpy_h2d(C.c_void_p(self.ptr), C.c_void_p(a.ctypes.data), a.nbytes), "h2d")

    def download(self) -> np.ndarray:
        out = np.empty(self.n, dtype=self.dtype)
        if self.n:
            _check(lib().ssdf_memcpy_d2h(C.c_void_p(out.ctypes.data), C.c_void_p(self.ptr), out.nbytes), "d2h")
        return out

    def free(self):
        if self.ptr:
            lib().ssdf_device_free(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        p = C.c_void_p()
        _check(lib().ssdf_stream_create(C.byref(p)), "ssdf_stream_create")
        self.ptr = p.value

    def synchronize(self):
        _check(lib().ssdf_stream_synchronize(C.c_void_p(self.ptr)), "ssdf_stream_synchronize")

    def __del__(self):
        try:
            if self.ptr:
                lib().ssdf_stream_destroy(C.c_void_p(self.ptr))
        except Exception:
            pass


class Event:
    def __init__(self):
        p = C.c_void_p()
        _check(lib().ssdf_event_create(C.byref(p)), "ssdf_event_create")
        self.ptr = p.value

    def record(self, stream: Stream | None = None):
        _check(lib().ssdf_event_record(C.c_void_p(self.ptr), C.c_void_p(stream.ptr) if stream else None),
               "ssdf_event_record")

    def elapsed_ms(self, later: "Event") -> float:
        ms = C.c_float()
        _check(lib().ssdf_event_elapsed_ms(C.c_void_p(self.ptr), C.c_void_p(later.ptr), C.byref(ms)), "elapsed")
        return float(ms.value)

    def __del__(self):
        try:
            if self.ptr:
                lib().ssdf_event_destroy(C.c_void_p(self.ptr))
        except Exception:
            pass


# ------------------------------------------------------------------------------------------------
# reference-shaped entry points (host arrays, synchronous)
# ------------------------------------------------------------------------------------------------
def _prec(*arrs):
    dt = {np.dtype(a.dtype) for a in arrs if a is not None}
    if dt == {np.dtype(np.float32)}:
        return "f"
    if dt == {np.dtype(np.float64)}:
        return "d"
    raise TypeError(f"coordinates and outputs must all be float32 or all float64, got {dt}")


def _idx(*arrs):
    for a in arrs:
        if a.dtype != np.int32:
            raise TypeError("triangle indices must be int32")


def edf(x, y, z, s0, s1, s2, sx, sy, sz, out) -> int:
    """ssdf::edf (serial/ssdf.hpp:447-460): unsigned distance, min-accumulated into `out`."""
    p = _prec(x, y, z, sx, sy, sz, out)
    _idx(s0, s1, s2)
    fn = lib().ssdf_edf_f if p == "f" else lib().ssdf_edf_d
    rc = fn(x.size, _ptr(x), _ptr(y), _ptr(z), s0.size, _ptr(s0), _ptr(s1), _ptr(s2), sx.size, _ptr(sx), _ptr(sy),
            _ptr(sz), _ptr(out))
    _check(rc, f"ssdf_edf_{p}")
    return 0


# The reference's other engines share edf's contract; on this engine they are the same call.
edf_celllist = edf
edf_select = edf
edf_bvh = edf


def closest_point_bvh(x, y, z, s0, s1, s2, sx, sy, sz, outd, outx, outy, outz, outtri, flags: int = 0) -> int:
    """ssdf::closest_point_bvh (serial/bvh/bvh.hpp:34-51): strict-improvement update of outd, the
    closest point and the closest triangle."""
    p = _prec(x, y, z, sx, sy, sz, outd, outx, outy, outz)
    _idx(s0, s1, s2, outtri)
    if p == "f":
        rc = lib().ssdf_closest_point_f(x.size, _ptr(x), _ptr(y), _ptr(z), s0.size, _ptr(s0), _ptr(s1), _ptr(s2),
                                        sx.size, _ptr(sx), _ptr(sy), _ptr(sz), _ptr(outd), _ptr(outx), _ptr(outy),
                                        _ptr(outz), _ptr(outtri), flags)
    else:
        rc = lib().ssdf_closest_point_d(x.size, _ptr(x), _ptr(y), _ptr(z), s0.size, _ptr(s0), _ptr(s1), _ptr(s2),
                                        sx.size, _ptr(sx), _ptr(sy), _ptr(sz), _ptr(outd), _ptr(outx), _ptr(outy),
                                        _ptr(outz), _ptr(outtri))
    _check(rc, f"ssdf_closest_point_{p}")
    return 0


def sample_sdf_bvh(s0, s1, s2, sx, sy, sz, nx, ny, nz, xmin, ymin, zmin, xmax, ymax, zmax, out, out_tri=None,
                   flags: int = 0) -> int:
    """ssdf::sample_sdf_bvh (serial/bvh/bvh.hpp:209-227): signed distance on grid nodes, x fastest."""
    p = _prec(sx, sy, sz, out)
    _idx(s0, s1, s2)
    if out.size != nx * ny * nz:
        raise ValueError("out must hold nx * ny * nz values")
    if p == "f":
        rc = lib().ssdf_sample_sdf_f(s0.size, _ptr(s0), _ptr(s1), _ptr(s2), sx.size, _ptr(sx), _ptr(sy), _ptr(sz),
                                     nx, ny, nz, xmin, ymin, zmin, xmax, ymax, zmax, _ptr(out), _ptr(out_tri), flags)
    else:
        rc = lib().ssdf_sample_sdf_d(s0.size, _ptr(s0), _ptr(s1), _ptr(s2), sx.size, _ptr(sx), _ptr(sy), _ptr(sz),
                                     nx, ny, nz, xmin, ymin, zmin, xmax, ymax, zmax, _ptr(out), _ptr(out_tri))
    _check(rc, f"ssdf_sample_sdf_{p}")
    return 0


def sdf(x, y, z, s0, s1, s2, sx, sy, sz, out_signed, out_tri=None, flags: int = 0) -> int:
    """Signed distance at points (per-point form of sample_sdf_bvh), FP32."""
    _prec(x, y, z, sx, sy, sz, out_signed)
    _idx(s0, s1, s2)
    rc = lib().ssdf_sdf_f(x.size, _ptr(x), _ptr(y), _ptr(z), s0.size, _ptr(s0), _ptr(s1), _ptr(s2), sx.size,
                          _ptr(sx), _ptr(sy), _ptr(sz), _ptr(out_signed), _ptr(out_tri), flags)
    _check(rc, "ssdf_sdf_f")
    return 0


def sample_sqedf_bvh(s0, s1, s2, sx, sy, sz, nx, ny, nz, xmin, ymin, zmin, xmax, ymax, zmax, out,
                     flags: int = 0) -> int:
    """ssdf::sample_sqedf_bvh (serial/bvh/bvh.hpp:183-201): out (squared distances, in/out) =
    min(out, d^2) on grid nodes, x fastest."""
    p = _prec(sx, sy, sz, out)
    _idx(s0, s1, s2)
    if out.size != nx * ny * nz:
        raise ValueError("out must hold nx * ny * nz values")
    if p == "f":
        rc = lib().ssdf_sample_sqedf_f(s0.size, _ptr(s0), _ptr(s1), _ptr(s2), sx.size, _ptr(sx), _ptr(sy),
                                       _ptr(sz), nx, ny, nz, xmin, ymin, zmin, xmax, ymax, zmax, _ptr(out), flags)
    else:
        rc = lib().ssdf_sample_sqedf_d(s0.size, _ptr(s0), _ptr(s1), _ptr(s2), sx.size, _ptr(sx), _ptr(sy),
                                       _ptr(sz), nx, ny, nz, xmin, ymin, zmin, xmax, ymax, zmax, _ptr(out))
    _check(rc, f"ssdf_sample_sqedf_{p}")
    return 0


def points_inside_bvh(x, y, z, s0, s1, s2, sx, sy, sz, out) -> int:
    """ssdf::points_inside_bvh (serial/bvh/bvh.hpp:53-66): out[i] = 1 inside the surface, else 0
    (uint8), decided by the generalized winding number."""
    p = _prec(x, y, z, sx, sy, sz)
    _idx(s0, s1, s2)
    if out.dtype != np.uint8 or out.size != x.size:
        raise TypeError("out must be a uint8 array with one entry per point")
    fn = lib().ssdf_points_inside_f if p == "f" else lib().ssdf_points_inside_d
    rc = fn(x.size, _ptr(x), _ptr(y), _ptr(z), s0.size, _ptr(s0), _ptr(s1), _ptr(s2), sx.size, _ptr(sx), _ptr(sy),
            _ptr(sz), _ptr(out))
    _check(rc, f"ssdf_points_inside_{p}")
    return 0


def sample_inside_bvh(s0, s1, s2, sx, sy, sz, nx, ny, nz, xmin, ymin, zmin, xmax, ymax, zmax, out) -> int:
    """ssdf::sample_inside_bvh (serial/bvh/bvh.hpp:229-247): inside flags on grid nodes, x fastest."""
    p = _prec(sx, sy, sz)
    _idx(s0, s1, s2)
    if out.dtype != np.uint8 or out.size != nx * ny * nz:
        raise TypeError("out must be a uint8 array of nx * ny * nz entries")
    fn = lib().ssdf_sample_inside_f if p == "f" else lib().ssdf_sample_inside_d
    rc = fn(s0.size, _ptr(s0), _ptr(s1), _ptr(s2), sx.size, _ptr(sx), _ptr(sy), _ptr(sz), nx, ny, nz, xmin, ymin,
            zmin, xmax, ymax, zmax, _ptr(out))
    _check(rc, f"ssdf_sample_inside_{p}")
    return 0


def compute_aabb(x, y, z):
    """ssdf::compute_aabb (serial/edf/edf.hpp:500-539): (xmin, xmax, ymin, ymax, zmin, zmax) of the
    points, in their precision."""
    p = _prec(x, y, z)
    T = C.c_float if p == "f" else C.c_double
    b = [T() for _ in range(6)]
    fn = lib().ssdf_compute_aabb_f if p == "f" else lib().ssdf_compute_aabb_d
    rc = fn(x.size, _ptr(x), _ptr(y), _ptr(z), *[C.byref(v) for v in b])
    _check(rc, f"ssdf_compute_aabb_{p}")
    return tuple(v.value for v in b)


def all_points_aabb_signed_distance(x, y, z, minx, maxx, miny, maxy, minz, maxz, out) -> int:
    """ssdf::all_points_aabb_signed_distance (serial/edf/edf.hpp:541-557): signed distance of every
    point to the box (negative inside), overwriting `out`."""
    p = _prec(x, y, z, out)
    fn = lib().ssdf_all_points_aabb_signed_distance_f if p == "f" else lib().ssdf_all_points_aabb_signed_distance_d
    rc = fn(x.size, _ptr(x), _ptr(y), _ptr(z), minx, maxx, miny, maxy, minz, maxz, _ptr(out))
    _check(rc, f"ssdf_all_points_aabb_signed_distance_{p}")
    return 0


def init_volume_box(x: "DeviceArray", y: "DeviceArray", z: "DeviceArray", out: "DeviceArray",
                    stream: "Stream | None" = None) -> None:
    """The SSDF_INIT_WITH_VOLUME_BOX seed (drivers/sdf.exe.cpp:317-327) on device arrays:
    out = |signed distance to the bounds of all points|, asynchronous on `stream`."""
    dt = {np.dtype(a.dtype) for a in (x, y, z, out)}
    if len(dt) != 1 or dt.pop() not in (np.dtype(np.float32), np.dtype(np.float64)):
        raise TypeError("points and out must share one float dtype")
    if not (x.size == y.size == z.size == out.size):
        raise ValueError("x, y, z and out must have the same length")
    rc = lib().ssdf_init_volume_box(0 if x.dtype == np.float32 else 1, x.size, _ptr(x), _ptr(y), _ptr(z), _ptr(out),
                                    C.c_void_p(stream.ptr) if stream is not None else None)
    _check(rc, "ssdf_init_volume_box")


# ------------------------------------------------------------------------------------------------
# resident meshes (device API)
# ------------------------------------------------------------------------------------------------
class ResidentMesh:
    """A mesh whose LBVH (and pseudonormals) live in HBM; queries take DeviceArrays."""

    def __init__(self, s0, s1, s2, sx, sy, sz, with_normals: bool = True, stream: Stream | None = None):
        dev = isinstance(sx, DeviceArray)
        dtype = (sx.dtype if dev else np.asarray(sx).dtype)
        self.dtype = np.dtype(dtype)
        self.dt_code = 0 if self.dtype == np.float32 else 1
        h = C.c_void_p()
        n = s0.n if dev else s0.size
        nv = sx.n if dev else sx.size
        rc = lib().ssdf_mesh_create(self.dt_code, n, _ptr(s0), _ptr(s1), _ptr(s2), nv, _ptr(sx), _ptr(sy),
                                    _ptr(sz), 1 if dev else 0, 1 if with_normals else 0,
                                    C.c_void_p(stream.ptr) if stream else None, C.byref(h))
        _check(rc, "ssdf_mesh_create")
        self.h = h.value

    @classmethod
    def from_mesh(cls, m, with_normals=True, dtype=np.float32, stream: Stream | None = None):
        return cls(m.s0, m.s1, m.s2, m.sx.astype(dtype), m.sy.astype(dtype), m.sz.astype(dtype), with_normals,
                   stream)

    def info(self) -> dict:
        i = _MeshInfo()
        _check(lib().ssdf_mesh_info(C.c_void_p(self.h), C.byref(i)), "ssdf_mesh_info")
        return dict(nselements=i.nselements, nspoints=i.nspoints, dtype=i.dtype, maxdepth=i.maxdepth,
                    open_edges=i.open_edges, flipped_edges=i.flipped_edges, bbox=list(i.bbox),
                    device_bytes=i.device_bytes)

    def query_points(self, mode, x, y, z, out_d, out_tri=None, cx=None, cy=None, cz=None, flags=0,
                     stream: Stream | None = None):
        rc = lib().ssdf_mesh_query_points(C.c_void_p(self.h), mode, x.n, _ptr(x), _ptr(y), _ptr(z), _ptr(out_d),
                                          _ptr(out_tri), _ptr(cx), _ptr(cy), _ptr(cz), flags,
                                          C.c_void_p(stream.ptr) if stream else None)
        _check(rc, "ssdf_mesh_query_points")

    def query_points_noalloc(self, mode, x, y, z, out_d, out_tri=None, cx=None, cy=None, cz=None, flags=0,
                             stream: Stream | None = None):
        """query_points in the library's static workspace (no allocation; at most NOALLOC_MAX_POINTS
        points; calls that overlap on one GPU race)."""
        rc = lib().ssdf_mesh_query_points_noalloc(C.c_void_p(self.h), mode, x.n, _ptr(x), _ptr(y), _ptr(z),
                                                  _ptr(out_d), _ptr(out_tri), _ptr(cx), _ptr(cy), _ptr(cz), flags,
                                                  C.c_void_p(stream.ptr) if stream else None)
        _check(rc, "ssdf_mesh_query_points_noalloc")

    def query_grid(self, mode, n, lo, hi, out_d, out_tri=None, flags=0, stream: Stream | None = None):
        nx, ny, nz = n
        rc = lib().ssdf_mesh_query_grid(C.c_void_p(self.h), mode, nx, ny, nz, float(lo[0]), float(lo[1]),
                                        float(lo[2]), float(hi[0]), float(hi[1]), float(hi[2]), _ptr(out_d),
                                        _ptr(out_tri), flags, C.c_void_p(stream.ptr) if stream else None)
        _check(rc, "ssdf_mesh_query_grid")

    def winding_number(self, x, y, z, out_w, stream: Stream | None = None):
        rc = lib().ssdf_mesh_winding_number(C.c_void_p(self.h), x.n, _ptr(x), _ptr(y), _ptr(z), _ptr(out_w),
                                            C.c_void_p(stream.ptr) if stream else None)
        _check(rc, "ssdf_mesh_winding_number")

    def inside_points(self, x, y, z, out_inside, stream: Stream | None = None):
        rc = lib().ssdf_mesh_inside_points(C.c_void_p(self.h), x.n, _ptr(x), _ptr(y), _ptr(z), _ptr(out_inside),
                                           C.c_void_p(stream.ptr) if stream else None)
        _check(rc, "ssdf_mesh_inside_points")

    def inside_grid(self, n, lo, hi, out_inside, stream: Stream | None = None):
        nx, ny, nz = n
        rc = lib().ssdf_mesh_inside_grid(C.c_void_p(self.h), nx, ny, nz, float(lo[0]), float(lo[1]), float(lo[2]),
                                         float(hi[0]), float(hi[1]), float(hi[2]), _ptr(out_inside),
                                         C.c_void_p(stream.ptr) if stream else None)
        _check(rc, "ssdf_mesh_inside_grid")

    def brute_points(self, x, y, z, out_d2, out_tri, stream: Stream | None = None):
        rc = lib().ssdf_mesh_brute_points(C.c_void_p(self.h), x.n, _ptr(x), _ptr(y), _ptr(z), _ptr(out_d2),
                                          _ptr(out_tri), C.c_void_p(stream.ptr) if stream else None)
        _check(rc, "ssdf_mesh_brute_points")

    def leaf_order(self, out_tri, stream: Stream | None = None):
        """Original triangle index of each LBVH leaf, in leaf (Morton) order, into a device array."""
        _check(lib().ssdf_mesh_leaf_order(C.c_void_p(self.h), _ptr(out_tri),
                                          C.c_void_p(stream.ptr) if stream else None), "ssdf_mesh_leaf_order")

    def rebuild(self, with_normals: bool = True, stream: Stream | None = None):
        _check(lib().ssdf_mesh_rebuild(C.c_void_p(self.h), 1 if with_normals else 0,
                                       C.c_void_p(stream.ptr) if stream else None), "ssdf_mesh_rebuild")

    def update_vertices(self, sx, sy, sz, with_normals: bool = True, stream: Stream | None = None):
        dev = isinstance(sx, DeviceArray)
        _check(lib().ssdf_mesh_update_vertices(C.c_void_p(self.h), _ptr(sx), _ptr(sy), _ptr(sz), 1 if dev else 0,
                                               1 if with_normals else 0, C.c_void_p(stream.ptr) if stream else None),
               "ssdf_mesh_update_vertices")

    def last_stats(self):
        q, e = C.c_uint(), C.c_int()
        _check(lib().ssdf_mesh_last_stats(C.c_void_p(self.h), C.byref(q), C.byref(e)), "ssdf_mesh_last_stats")
        return int(q.value), int(e.value)

    def set_schedule(self, brick=(4, 4, 4), heavy_leaves: int = -1):
        """Brick of grid nodes per wave (product 64, powers of two) and leaf budget (-1: process-wide)."""
        _check(lib().ssdf_mesh_set_schedule(C.c_void_p(self.h), int(brick[0]), int(brick[1]), int(brick[2]),
                                            int(heavy_leaves)), "ssdf_mesh_set_schedule")

    def close(self):
        if getattr(self, "h", None):
            lib().ssdf_mesh_destroy(C.c_void_p(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------------------------------------
# batches (device API)
# ------------------------------------------------------------------------------------------------
class Batch:
    """Many FP32 meshes resident in shared arenas, answered by one persistent launch per call."""

    def __init__(self, meshes, with_normals: bool = True, stream: Stream | None = None):
        tri_off = np.zeros(len(meshes) + 1, np.int64)
        vert_off = np.zeros(len(meshes) + 1, np.int64)
        for i, m in enumerate(meshes):
            tri_off[i + 1] = tri_off[i] + m.s0.size
            vert_off[i + 1] = vert_off[i] + m.sx.size
        cat = lambda k, dt: np.ascontiguousarray(np.concatenate([getattr(m, k) for m in meshes]), dtype=dt)
        s0, s1, s2 = cat("s0", np.int32), cat("s1", np.int32), cat("s2", np.int32)
        sx, sy, sz = cat("sx", np.float32), cat("sy", np.float32), cat("sz", np.float32)
        self.nmeshes = len(meshes)
        self.ntris = tri_off[1:] - tri_off[:-1]
        h = C.c_void_p()
        rc = lib().ssdf_batch_create(len(meshes), _ptr(tri_off), _ptr(vert_off), _ptr(s0), _ptr(s1), _ptr(s2),
                                     _ptr(sx), _ptr(sy), _ptr(sz), 0, 1 if with_normals else 0,
                                     C.c_void_p(stream.ptr) if stream else None, C.byref(h))
        _check(rc, "ssdf_batch_create")
        self.h = h.value

    @staticmethod
    def grid_layout(dims):
        """out_offsets for per-mesh grids dims[i] = (nx, ny, nz), packed one after another."""
        n = np.array([int(d[0]) * int(d[1]) * int(d[2]) for d in dims], np.int64)
        off = np.zeros(len(dims), np.int64)
        off[1:] = np.cumsum(n)[:-1]
        return off, int(n.sum())

    def query_grids(self, mode, dims, boxes, out_offsets, out_d, out_tri=None, flags=0, stream: Stream | None = None):
        d = np.ascontiguousarray(dims, np.int64).reshape(-1)
        bx = np.ascontiguousarray(boxes, np.float64).reshape(-1)
        off = np.ascontiguousarray(out_offsets, np.int64)
        rc = lib().ssdf_batch_query_grids(C.c_void_p(self.h), mode, _ptr(d), _ptr(bx), _ptr(off), _ptr(out_d),
                                          _ptr(out_tri), flags, C.c_void_p(stream.ptr) if stream else None)
        _check(rc, "ssdf_batch_query_grids")

    def query_points(self, mode, point_offsets, x, y, z, out_d, out_tri=None, cx=None, cy=None, cz=None, flags=0,
                     stream: Stream | None = None):
        off = np.ascontiguousarray(point_offsets, np.int64)
        rc = lib().ssdf_batch_query_points(C.c_void_p(self.h), mode, _ptr(off), _ptr(x), _ptr(y), _ptr(z),
                                           _ptr(out_d), _ptr(out_tri), _ptr(cx), _ptr(cy), _ptr(cz), flags,
                                           C.c_void_p(stream.ptr) if stream else None)
        _check(rc, "ssdf_batch_query_points")

    def rebuild(self, with_normals: bool = True, stream: Stream | None = None):
        _check(lib().ssdf_batch_rebuild(C.c_void_p(self.h), 1 if with_normals else 0,
                                        C.c_void_p(stream.ptr) if stream else None), "ssdf_batch_rebuild")

    def set_schedule(self, blocks_per_cu: int, units_per_grab: int):
        _check(lib().ssdf_batch_set_schedule(C.c_void_p(self.h), blocks_per_cu, units_per_grab),
               "ssdf_batch_set_schedule")

    def set_mesh_schedule(self, mesh: int, brick=(4, 4, 4), heavy_leaves: int = -1):
        """Brick of grid nodes per wave (product 64, powers of two) and leaf budget (-1: process-wide)
        of one mesh's grid queries."""
        _check(lib().ssdf_batch_set_mesh_schedule(C.c_void_p(self.h), mesh, int(brick[0]), int(brick[1]),
                                                  int(brick[2]), int(heavy_leaves)), "ssdf_batch_set_mesh_schedule")

    def mesh_schedule(self, mesh: int):
        """((brick_x, brick_y, brick_z), heavy_leaves) of one mesh."""
        v = [C.c_int() for _ in range(4)]
        _check(lib().ssdf_batch_get_mesh_schedule(C.c_void_p(self.h), mesh, *[C.byref(x) for x in v]),
               "ssdf_batch_get_mesh_schedule")
        return (v[0].value, v[1].value, v[2].value), v[3].value

    def autotune(self, mode, dims, boxes, out_offsets, out_d, out_tri=None, flags=0, stream: Stream | None = None):
        d = np.ascontiguousarray(dims, np.int64).reshape(-1)
        bx = np.ascontiguousarray(boxes, np.float64).reshape(-1)
        off = np.ascontiguousarray(out_offsets, np.int64)
        bb, bc, ms = C.c_int(), C.c_int(), C.c_float()
        rc = lib().ssdf_batch_autotune(C.c_void_p(self.h), mode, _ptr(d), _ptr(bx), _ptr(off), _ptr(out_d),
                                       _ptr(out_tri), flags, C.c_void_p(stream.ptr) if stream else None, C.byref(bb),
                                       C.byref(bc), C.byref(ms))
        _check(rc, "ssdf_batch_autotune")
        return int(bb.value), int(bc.value), float(ms.value)

    def capture(self, rebuild, mode, dims, boxes, out_offsets, out_d, out_tri=None, flags=0,
                stream: Stream | None = None):
        d = np.ascontiguousarray(dims, np.int64).reshape(-1)
        bx = np.ascontiguousarray(boxes, np.float64).reshape(-1)
        off = np.ascontiguousarray(out_offsets, np.int64)
        self._keep = (d, bx, off)
        rc = lib().ssdf_batch_capture(C.c_void_p(self.h), 1 if rebuild else 0, mode, _ptr(d), _ptr(bx), _ptr(off),
                                      _ptr(out_d), _ptr(out_tri), flags, C.c_void_p(stream.ptr) if stream else None)
        _check(rc, "ssdf_batch_capture")

    def replay(self, stream: Stream | None = None):
        _check(lib().ssdf_batch_replay(C.c_void_p(self.h), C.c_void_p(stream.ptr) if stream else None),
               "ssdf_batch_replay")

    def last_stats(self):
        q, e = C.c_uint(), C.c_int()
        _check(lib().ssdf_batch_last_stats(C.c_void_p(self.h), C.byref(q), C.byref(e)), "ssdf_batch_last_stats")
        return int(q.value), int(e.value)

    def info(self, mesh: int = -1) -> dict:
        i = _MeshInfo()
        _check(lib().ssdf_batch_info(C.c_void_p(self.h), mesh, C.byref(i)), "ssdf_batch_info")
        return dict(nselements=i.nselements, nspoints=i.nspoints, maxdepth=i.maxdepth, open_edges=i.open_edges,
                    flipped_edges=i.flipped_edges, bbox=list(i.bbox), device_bytes=i.device_bytes)

    def close(self):
        if getattr(self, "h", None):
            lib().ssdf_batch_destroy(C.c_void_p(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ------------------------------------------------------------------------------------------------
# serving
# ------------------------------------------------------------------------------------------------
class PinnedArray:
    """Pinned host memory viewed as a numpy array (request buffers the server copies asynchronously)."""

    def __init__(self, n: int, dtype):
        self.dtype = np.dtype(dtype)
        p = C.c_void_p()
        _check(lib().ssdf_host_alloc(C.byref(p), max(1, int(n) * self.dtype.itemsize)), "ssdf_host_alloc")
        self.ptr = p.value
        buf = (C.c_char * (max(1, int(n)) * self.dtype.itemsize)).from_address(self.ptr)
        self.array = np.frombuffer(buf, dtype=self.dtype, count=int(n))

    def free(self):
        if self.ptr:
            self.array = None
            lib().ssdf_host_free(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Server:
    """Serving loop over resident meshes (HBM-budget admission, continuous batching, p50/p99)."""

    def __init__(self, hbm_budget_bytes: int, max_inflight: int = 0):
        h = C.c_void_p()
        _check(lib().ssdf_server_create(int(hbm_budget_bytes), int(max_inflight), C.byref(h)), "ssdf_server_create")
        self.h = h.value
        self._keep = {}

    def add_mesh(self, mesh_id: int, m):
        arrs = [np.ascontiguousarray(a) for a in (m.s0, m.s1, m.s2, m.sx, m.sy, m.sz)]
        self._keep[("mesh", mesh_id)] = arrs
        rc = lib().ssdf_server_add_mesh(C.c_void_p(self.h), mesh_id, arrs[0].size, *[_ptr(a) for a in arrs[:3]],
                                        arrs[3].size, *[_ptr(a) for a in arrs[3:]])
        _check(rc, "ssdf_server_add_mesh")

    def charge(self, mesh_id: int, npoints: int) -> tuple[int, int]:
        """(bytes charged for the resident mesh, bytes charged for a request of npoints points)."""
        mb, rb = C.c_size_t(), C.c_size_t()
        _check(lib().ssdf_server_charge(C.c_void_p(self.h), mesh_id, int(npoints), C.byref(mb), C.byref(rb)),
               "ssdf_server_charge")
        return int(mb.value), int(rb.value)

    def clock(self) -> float:
        t = C.c_double()
        _check(lib().ssdf_server_clock(C.c_void_p(self.h), C.byref(t)), "ssdf_server_clock")
        return t.value

    def submit(self, mesh_id, mode, x, y, z, out_d, out_tri=None, flags=0, arrival: float = 0.0) -> int:
        """x, y, z, out_d, out_tri: host arrays (numpy or PinnedArray.array) alive until wait()."""
        t = C.c_uint64()
        rc = lib().ssdf_server_submit(C.c_void_p(self.h), mesh_id, mode, x.size, _ptr(x), _ptr(y), _ptr(z),
                                      _ptr(out_d), _ptr(out_tri), flags, float(arrival), C.byref(t))
        _check(rc, "ssdf_server_submit")
        return int(t.value)

    def wait(self, ticket: int) -> float:
        lat = C.c_double()
        _check(lib().ssdf_server_wait(C.c_void_p(self.h), ticket, C.byref(lat)), "ssdf_server_wait")
        return lat.value

    def poll(self, ticket: int) -> bool:
        d = C.c_int()
        _check(lib().ssdf_server_poll(C.c_void_p(self.h), ticket, C.byref(d)), "ssdf_server_poll")
        return bool(d.value)

    def stats(self) -> dict:
        st = _ServerStats()
        _check(lib().ssdf_server_stats(C.c_void_p(self.h), C.c_void_p(C.addressof(st))), "ssdf_server_stats")
        return {k: getattr(st, k) for k, _ in _ServerStats._fields_}

    def close(self):
        if getattr(self, "h", None):
            lib().ssdf_server_destroy(C.c_void_p(self.h))
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
