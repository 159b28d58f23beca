"""Builds libssdf_amd.so (HIP, gfx950) in-tree with hipcc.

Every .hip file under csrc/ compiles to an object in parallel, then one shared library is linked
into ssdf_amd/_lib/.  Flags:
  -ffp-contract=off                 expressions round as written (FP64 fixup = oracle bits)
  -fno-gpu-flush-denormals-to-zero  subnormals kept, as the reference's HIP build does
  -packed-fp32-ops disabled         no v_pk_*_f32 (unreliable on this pool)
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "_lib"
OBJDIR = PKG / "_build"
LIBNAME = "libssdf_amd.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("SSDF_OFFLOAD_ARCH", "gfx950")

COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-fno-gpu-flush-denormals-to-zero", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-command-line-argument",
    "-Wno-unknown-pragmas",
]


def _compile(src: Path, obj: Path, extra: list[str]) -> tuple[Path, str]:
    cmd = [HIPCC, *COMMON, *extra, "-I", str(CSRC), "-I", str(PKG.parent / "include"), "-c", str(src), "-o",
           str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj, (r.stderr or "")


def build(verbose: bool = False, extra: list[str] | None = None) -> Path:
    extra = list(extra or [])
    OBJDIR.mkdir(exist_ok=True)
    LIBDIR.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip"))
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    objs = []
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        futs = [ex.submit(_compile, s, OBJDIR / (s.stem + ".o"), extra) for s in srcs]
        for f in futs:
            obj, log = f.result()
            objs.append(obj)
            if verbose and log.strip():
                print(log, file=sys.stderr)
    tmp = LIBDIR / (LIBNAME + ".tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    out = LIBDIR / LIBNAME
    os.replace(tmp, out)
    return out


def build_oracle(with_reference: bool | None = None) -> None:
    """Compiles the CPU oracle (test infrastructure) and, when the reference sources are mounted,
    the shim over the reference's own edf.hpp."""
    oracle = PKG.parent / "oracle"
    subprocess.run(["make", "-s", "-C", str(oracle), "all"], check=True)
    ref = Path("/root/reference/serial/edf/edf.hpp")
    if with_reference is None:
        with_reference = ref.exists()
    if with_reference:
        subprocess.run(["make", "-s", "-C", str(oracle), "ref"], check=True)


def build_cpp_tests() -> list[Path]:
    """Compiles the C++ drop-in test programs (tests/cpp/*.cpp) against include/ and the library."""
    root = PKG.parent
    src_dir = root / "tests" / "cpp"
    out_dir = src_dir / "_build"
    out_dir.mkdir(exist_ok=True)
    outs = []
    for src in sorted(src_dir.glob("*.cpp")):
        exe = out_dir / src.stem
        cmd = ["g++", "-std=c++17", "-O1", "-I", str(root / "include"), str(src), "-L", str(LIBDIR), "-lssdf_amd",
               "-Wl,-rpath," + str(LIBDIR), "-Wl,-rpath,$ORIGIN/../../../ssdf_amd/_lib", "-o", str(exe)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"g++ failed for {src.name}:\n{r.stderr}")
        outs.append(exe)
    return outs


def build_tools() -> list[Path]:
    """Compiles the command-line programs (csrc/*_main.cpp, host code over the C ABI) into _lib/,
    next to the library they load (rpath $ORIGIN): sdf_main.cpp -> ssdf-sdf."""
    outs = []
    for src in sorted(CSRC.glob("*_main.cpp")):
        exe = LIBDIR / ("ssdf-" + src.stem[: -len("_main")])
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-I", str(PKG.parent / "include"), str(src), "-L", str(LIBDIR),
               "-lssdf_amd", "-Wl,-rpath,$ORIGIN", "-o", str(exe)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"g++ failed for {src.name}:\n{r.stderr}")
        outs.append(exe)
    return outs


def clean() -> None:
    shutil.rmtree(OBJDIR, ignore_errors=True)


if __name__ == "__main__":
    p = build(verbose="-v" in sys.argv)
    build_tools()
    print(p)
