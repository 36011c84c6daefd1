"""Build libopaw_hip.so in-tree with hipcc for gfx950.

Each translation unit under csrc/ is compiled to an object in
opaw-tddft_amd/_build/ (in parallel), then linked into
opaw-tddft_amd/opaw_amd/libopaw_hip.so.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG_DIR = Path(__file__).resolve().parent
ROOT = PKG_DIR.parent                      # opaw-tddft_amd/
REPO = ROOT.parent
CSRC = ROOT / "csrc"
BUILD = ROOT / "_build"
INCLUDE = REPO / "include"
LIB_NAME = "libopaw_hip.so"
LIB_PATH = PKG_DIR / LIB_NAME

ARCH = os.environ.get("OPAW_OFFLOAD_ARCH", "gfx950")
HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # packed FP32 instructions stay off: some of their forms have given
    # silently wrong results on gfx950
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
# the block solver's Rayleigh-Ritz runs rocSOLVER's Hermitian eigensolver and
# rocBLAS products on the device
LINK_LIBS = ["-lrocsolver", "-lrocblas"]


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the OPAW HIP library cannot be built")


def sources():
    return sorted(CSRC.glob("*.hip"))


def _compile(src: Path, verbose: bool) -> Path:
    obj = BUILD / (src.stem + ".o")
    cmd = [hipcc(), *HIPCC_FLAGS, f"-I{INCLUDE}", f"-I{CSRC}", "-c", str(src), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{res.stderr[-6000:]}")
    if verbose:
        print(f"[opaw build] {src.name} -> {obj.name}", flush=True)
    return obj


def build(verbose: bool = True, jobs: int | None = None) -> Path:
    """Compile every HIP source and link the shared library; always rebuilds."""
    BUILD.mkdir(parents=True, exist_ok=True)
    srcs = sources()
    if not srcs:
        raise RuntimeError(f"no HIP sources under {CSRC}")
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), srcs))
    tmp = LIB_PATH.with_suffix(".so.tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), *LINK_LIBS]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed:\n{res.stderr[-4000:]}")
    os.replace(tmp, LIB_PATH)
    if verbose:
        print(f"[opaw build] linked {LIB_PATH}", flush=True)
    return LIB_PATH


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
