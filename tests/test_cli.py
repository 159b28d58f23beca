"""The sdf, mesh_to_sdf, points_inside and sdf_sign commands (drivers/sdf.exe.cpp,
demo/mesh_to_sdf.exe.cpp, drivers/points_inside.exe.cpp, drivers/sdf_sign.exe.cpp contracts):
argument and file checks that exit before any device work, and sdf_sign (a host file transform),
run on the CPU; the GPU runs are in test_gpu_cli.py."""
import subprocess

import numpy as np
import pytest

from ssdf_amd import meshgen as mg
from ssdf_amd import rawio


def _run(*args, env=None):
    if not rawio.SDF_EXE.exists():
        pytest.skip("ssdf-sdf not built")
    return subprocess.run([str(rawio.SDF_EXE), *map(str, args)], capture_output=True, text=True, env=env,
                          timeout=120)


def test_usage_exits_1():
    r = _run()
    assert r.returncode == 1 and "Usage" in r.stderr


def test_missing_and_mismatched_files(tmp_path):
    m = mg.uv_sphere(0.5, 8, 4)
    rawio.write_surface(tmp_path / "s", m)
    rawio.write_points(tmp_path / "p", np.zeros(5), np.zeros(5), np.zeros(5))
    r = _run(tmp_path / "nope", tmp_path / "p", tmp_path / "o.raw")
    assert r.returncode == 1 and "Cannot open" in r.stderr
    np.zeros(3, np.float32).tofile(tmp_path / "s" / "y.float32")
    r = _run(tmp_path / "s", tmp_path / "p", tmp_path / "o.raw")
    assert r.returncode == 1 and "mismatched" in r.stderr
    rawio.write_surface(tmp_path / "s", m)
    np.zeros(4, np.float32).tofile(tmp_path / "p" / "z.float32")
    r = _run(tmp_path / "s", tmp_path / "p", tmp_path / "o.raw")
    assert r.returncode == 1 and "mismatched" in r.stderr


def _run_m2s(*args, env=None):
    if not rawio.MESH_TO_SDF_EXE.exists():
        pytest.skip("ssdf-mesh_to_sdf not built")
    return subprocess.run([str(rawio.MESH_TO_SDF_EXE), *map(str, args)], capture_output=True, text=True, env=env,
                          timeout=120)


def test_mesh_to_sdf_argument_checks(tmp_path):
    import os
    r = _run_m2s()
    assert r.returncode == 1 and "Usage" in r.stderr
    m = mg.uv_sphere(0.5, 8, 4)
    rawio.write_surface(tmp_path / "s", m)
    env = {k: v for k, v in os.environ.items() if not k.startswith("SSDF_")}
    # grids need 2 nodes per axis unless SSDF_HMAX sets the spacing (mesh_to_sdf.exe.cpp:98-101)
    r = _run_m2s(tmp_path / "s", 1, 5, 5, tmp_path / "o", env=env)
    assert r.returncode == 1 and ">= 2" in r.stderr
    r = _run_m2s(tmp_path / "s", "x", 5, 5, tmp_path / "o", env=env)
    assert r.returncode == 1 and ">= 2" in r.stderr
    r = _run_m2s(tmp_path / "nope", 5, 5, 5, tmp_path / "o", env=env)
    assert r.returncode == 1 and "cannot read" in r.stderr
    # block selection belongs to smesh meshes, which a raw folder is not
    r = _run_m2s(tmp_path / "s", 5, 5, 5, tmp_path / "o", env=dict(env, SSDF_BOX_BLOCKS="a,b"))
    assert r.returncode == 1 and "SSDF_BOX_BLOCKS" in r.stderr


def _run_exe(exe, *args):
    if not exe.exists():
        pytest.skip(f"{exe.name} not built")
    return subprocess.run([str(exe), *map(str, args)], capture_output=True, text=True, timeout=120)


def test_sdf_sign(tmp_path):
    """drivers/sdf_sign.exe.cpp:134-139: the distance negated where the inside flag is set."""
    r = _run_exe(rawio.SDF_SIGN_EXE)
    assert r.returncode == 1 and "Usage" in r.stderr
    rng = np.random.default_rng(5)
    udf = rng.uniform(0.0, 2.0, 1001).astype(np.float32)
    udf[:3] = (0.0, np.inf, 1e-30)
    inside = (rng.uniform(size=1001) < 0.4).astype(np.uint8)
    inside[:3] = (1, 1, 1)
    udf.tofile(tmp_path / "u.raw")
    inside.tofile(tmp_path / "in.raw")
    r = _run_exe(rawio.SDF_SIGN_EXE, tmp_path / "u.raw", tmp_path / "in.raw", tmp_path / "s.raw")
    assert r.returncode == 0, r.stderr
    want = np.where(inside != 0, -udf, udf)
    assert np.array_equal(np.fromfile(tmp_path / "s.raw", np.float32).view(np.uint32), want.view(np.uint32))
    inside[:-1].tofile(tmp_path / "short.raw")
    r = _run_exe(rawio.SDF_SIGN_EXE, tmp_path / "u.raw", tmp_path / "short.raw", tmp_path / "s2.raw")
    assert r.returncode == 1 and "mismatch" in r.stderr


def test_points_inside_argument_checks(tmp_path):
    r = _run_exe(rawio.POINTS_INSIDE_EXE)
    assert r.returncode == 1 and "Usage" in r.stderr
    m = mg.uv_sphere(0.5, 8, 4)
    rawio.write_surface(tmp_path / "s", m, ext="raw")
    rawio.write_points(tmp_path / "p", np.zeros(5), np.zeros(5), np.zeros(5), ext="raw")
    r = _run_exe(rawio.POINTS_INSIDE_EXE, tmp_path / "nope", tmp_path / "p", tmp_path / "o.raw")
    assert r.returncode == 1 and "Cannot open" in r.stderr
    np.zeros(3, np.int32).tofile(tmp_path / "s" / "i1.raw")
    r = _run_exe(rawio.POINTS_INSIDE_EXE, tmp_path / "s", tmp_path / "p", tmp_path / "o.raw")
    assert r.returncode == 1 and "mismatched" in r.stderr


def test_sqedf_usage():
    r = _run_exe(rawio.SQEDF_EXE)
    assert r.returncode == 1 and "Usage" in r.stderr and "squared distance" in r.stderr
