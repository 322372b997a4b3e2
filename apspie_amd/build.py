"""Builds apspie_amd/libapspie.so for gfx950 with hipcc (in-tree, no JIT cache).

    python -m apspie_amd.build [-j N]

Every source under apspie_amd/csrc is compiled to an object in
apspie_amd/_build and linked into one shared library whose exported symbols
are exactly the C ABI of include/apspie.h.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OUT = PKG / "libapspie.so"
BUILD = PKG / "_build"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("APSPIE_ARCH", "gfx950")

COMMON = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
          f"-I{ROOT / 'include'}"]
# Device code: gfx950 only; packed-FP32 instruction forms disabled.
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
          "-packed-fp32-ops", "-munsafe-fp-atomics"]


def sources():
    return sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))


def _compile(src: Path) -> Path:
    obj = BUILD / (src.name + ".o")
    if src.suffix == ".hip":
        cmd = [HIPCC, "-x", "hip", *COMMON, *DEVICE, "-c", str(src), "-o", str(obj)]
    else:
        cmd = [HIPCC, *COMMON, "-pthread", "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src.name}\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs: int | None = None, verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [HIPCC, *DEVICE[:1], "-shared", "-pthread", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    build_driver()
    if verbose:
        print(f"built {OUT} from {len(srcs)} sources, and {DRIVER}")
    return OUT


DRIVER = ROOT / "tests" / "cpp" / "mirror_driver"


def build_driver() -> Path:
    """The C++ driver that exercises include/apspie.hpp (host code only)."""
    src = DRIVER.with_suffix(".cpp")
    cmd = ["g++", "-O2", "-std=c++17", f"-I{ROOT / 'include'}", str(src), "-o", str(DRIVER),
           f"-L{PKG}", "-lapspie", f"-Wl,-rpath,{PKG}", "-Wl,-rpath,$ORIGIN/../../apspie_amd"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"driver build failed\n{' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return DRIVER


def clean():
    shutil.rmtree(BUILD, ignore_errors=True)
    if OUT.exists():
        OUT.unlink()


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("--clean", action="store_true")
    a = ap.parse_args()
    if a.clean:
        clean()
    try:
        build(a.j, verbose=True)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)
