"""The FP32 pruning routines of the traversal (branch-free triangle distance, FMA box distance),
built for the host from the same header the kernels use, against the FP64 oracle: the triangle
distance agrees to the FP32 error bound the fixup band assumes (2^-18 of the coordinate scale),
the box distance never exceeds the triangle distance by more than that bound (pruning stays
conservative), and degenerate triangles get their segment distance."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

from tests.test_oracle import _random_cases

ROOT = Path(__file__).resolve().parent.parent
SRC = ROOT / "tests" / "cpp_host" / "fast_routines_test.cpp"
EXE = ROOT / "tests" / "cpp_host" / "_build" / "fast_routines_test"


@pytest.fixture(scope="module")
def exe():
    EXE.parent.mkdir(exist_ok=True)
    cmd = ["/opt/rocm/bin/hipcc", "-O2", "-std=c++17", "-ffp-contract=off", "-I", str(ROOT / "ssdf_amd" / "csrc"),
           "-I", str(ROOT / "include"), str(SRC), "-o", str(EXE)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return EXE


def run(exe, cases):
    lines = [str(len(cases))]
    for cs in cases:
        lines.append(" ".join(float(v).hex() for x in cs for v in x))
    r = subprocess.run([str(exe)], input="\n".join(lines), capture_output=True, text=True, check=True)
    out = [l.split() for l in r.stdout.strip().splitlines()]
    return (np.array([float.fromhex(o[0]) for o in out]), np.array([float.fromhex(o[1]) for o in out]),
            np.array([int(o[2]) for o in out]))


def test_tri_d2_fast_accuracy(exe, oracle):
    cases = _random_cases(4000, 21, np.float32)
    d, bx, fl = run(exe, cases)
    ref = np.array([oracle.pt_tri_d2(*(np.float64(v) for v in cs)) for cs in cases])
    L = np.array([max(1.0, float(np.abs(np.concatenate(cs)).max())) for cs in cases])
    err = np.abs(np.sqrt(d) - np.sqrt(ref))
    assert np.all(err <= 2.0 ** -18 * L), err.max()
    assert np.all(np.sqrt(bx) <= np.sqrt(ref) + 2.0 ** -18 * L)  # conservative pruning
    assert np.all(fl == 0)


def test_tri_d2_fast_degenerate(exe, oracle):
    rng = np.random.default_rng(2)
    cases = []
    for _ in range(200):
        a = rng.uniform(-1, 1, 3)
        b = rng.uniform(-1, 1, 3)
        c = a + rng.uniform(0, 1) * (b - a)  # collinear: a needle
        p = rng.uniform(-1.5, 1.5, 3)
        cases.append(tuple(np.asarray(v, np.float32) for v in (p, a, b, c)))
    cases.append(tuple(np.asarray(v, np.float32) for v in ((0.3, 0.2, 0.1), (0, 0, 0), (0, 0, 0), (0, 0, 0))))
    d, bx, fl = run(exe, cases)
    assert np.all(np.isfinite(d))
    for (p, a, b, c), dd in zip(cases, d):
        p, a, b, c = (x.astype(np.float64) for x in (p, a, b, c))
        seg = []
        for u, v in ((a, b), (b, c), (c, a)):
            e = v - u
            t = 0.0 if e @ e == 0 else np.clip((p - u) @ e / (e @ e), 0, 1)
            seg.append(np.sum((p - u - t * e) ** 2))
        assert abs(np.sqrt(dd) - np.sqrt(min(seg))) <= 2.0 ** -18 * 2
