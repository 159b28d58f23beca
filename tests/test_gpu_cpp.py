"""The C++ drop-in programs (tests/cpp, built by __graft_entry__.build()) run on the GPU."""
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu
BUILD = Path(__file__).resolve().parent / "cpp" / "_build"


@pytest.mark.parametrize("name", ["dropin_edf_test"])
def test_cpp_dropin(gpu_lib, name):
    exe = BUILD / name
    assert exe.exists(), f"{exe} not built; run __graft_entry__.build()"
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "passed" in r.stdout
