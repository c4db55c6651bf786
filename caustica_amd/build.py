"""Build libcaustica_hip.so (gfx950) from caustica_amd/csrc with hipcc.

Objects go to <repo>/_build/ (one hipcc process per source, in parallel); the shared
library is written next to this file so it travels with the package tree.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
ROOT = PKG.parent
BUILD = ROOT / "_build"
LIB = PKG / "libcaustica_hip.so"

HIPCC_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    # no packed FP32 VALU forms (v_pk_{add,mul,fma}_f32) anywhere in the device code
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
]


def _hipcc() -> str:
    exe = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(exe).exists():
        raise RuntimeError("hipcc not found: the HIP toolchain is required to build caustica_amd")
    return exe


def _compile(src: Path, obj: Path) -> tuple[Path, int, str]:
    cmd = [_hipcc(), *HIPCC_FLAGS, "-I", str(ROOT / "include"), "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(
        ln for ln in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in ln
    )
    return src, r.returncode, out


HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC"]


def _compile_host(src: Path, obj: Path) -> tuple[Path, int, str]:
    """Plain C++ sources (no HIP) build with the host compiler."""
    cxx = shutil.which("g++") or "g++"
    cmd = [cxx, *HOST_FLAGS, "-I", str(ROOT / "include"), "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    return src, r.returncode, r.stdout + r.stderr


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compile every HIP source and link the shared library; returns its path."""
    BUILD.mkdir(exist_ok=True)
    sources = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    jobs = jobs or min(len(sources), max(1, min(16, os.cpu_count() or 1)))
    objs = []
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = []
        for src in sources:
            obj = BUILD / (src.stem + ".o")
            objs.append(obj)
            futs.append(ex.submit(_compile_host if src.suffix == ".cpp" else _compile, src, obj))
        failed = []
        for f in futs:
            src, rc, out = f.result()
            if verbose or rc != 0:
                if out.strip():
                    print(f"[{src.name}]\n{out}", file=sys.stderr)
            if rc != 0:
                failed.append(src.name)
    if failed:
        raise RuntimeError(f"hipcc failed for: {', '.join(failed)}")
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [_hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
