"""Build recipe for libgpb.so (HIP kernels for gfx950 + C++ host code).

Compiles every ``csrc/*.hip`` with hipcc for gfx950 and every ``csrc/*.cpp``
with g++, then links ``libgpb.so`` next to this file, and ``gpb_cli`` (the
command-line front end, ``apps/gpb_cli.cpp``) against it.  Objects go to
``_build/`` (git-ignored).  Run ``python build.py`` or call :func:`build`.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libgpb.so"
CLI = PKG / "gpb_cli"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics", "-Wno-unused-result",
]
CXX_FLAGS = ["-O3", "-fPIC", "-std=c++17", "-fopenmp", "-ffp-contract=off"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and "not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")


def _compile(src: Path) -> Path:
    obj = BUILD / (src.name + ".o")
    inc = ["-I", str(ROOT / "include"), "-I", str(CSRC)]
    if src.suffix == ".hip":
        cmd = [str(ROCM / "bin" / "hipcc"), *HIP_FLAGS, *inc, "-c", str(src), "-o", str(obj)]
    else:
        cmd = ["g++", *CXX_FLAGS, *inc, f"-I{ROCM / 'include'}", "-c", str(src), "-o", str(obj)]
    _run(cmd)
    return obj


def build(verbose: bool = True) -> Path:
    """Compile all sources and link libgpb.so; returns the library path."""
    BUILD.mkdir(exist_ok=True)
    srcs = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    jobs = min(8, os.cpu_count() or 1, len(srcs))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    _run(["g++", "-shared", "-fopenmp", "-o", str(tmp), *map(str, objs),
          f"-L{ROCM / 'lib'}", "-lamdhip64", f"-Wl,-rpath,{ROCM / 'lib'}"])
    os.replace(tmp, LIB)
    # command-line front end over the reference-named C++ entry points
    cli_tmp = CLI.with_suffix(".tmp")
    _run(["g++", *CXX_FLAGS, "-I", str(ROOT / "include"), str(PKG / "apps" / "gpb_cli.cpp"), "-o", str(cli_tmp),
          f"-L{PKG}", "-lgpb", "-Wl,-rpath,$ORIGIN", f"-L{ROCM / 'lib'}", "-lamdhip64", f"-Wl,-rpath,{ROCM / 'lib'}"])
    os.replace(cli_tmp, CLI)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources", flush=True)
    return LIB


if __name__ == "__main__":
    build()
    sys.exit(0)
