"""Build recipe for the gfx950 HIP library ``libnumbagg_amd.so``.

Every ``csrc/*.hip`` translation unit is compiled with hipcc for gfx950 and
linked into one shared library that exposes the C ABI declared in
``include/numbagg_amd.h``. The library is built in-tree so that it travels with
the source tree; nothing is JIT-compiled at import time.
"""

from __future__ import annotations

import concurrent.futures
import os
import pathlib
import shutil
import subprocess

PKG_DIR = pathlib.Path(__file__).resolve().parent
ROOT_DIR = PKG_DIR.parent
CSRC_DIR = PKG_DIR / "csrc"
OBJ_DIR = PKG_DIR / "_build"
LIB_PATH = PKG_DIR / "libnumbagg_amd.so"

ARCH = "gfx950"

# -ffp-contract=off keeps the compensated (TwoSum) arithmetic exact and matches
# the reference's no-contraction numerics; the target-feature flag keeps hipcc
# from emitting packed FP32 instructions.
HIPCC_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang",
    "-target-feature",
    "-Xclang",
    "-packed-fp32-ops",
    "-Wall",
    "-Wno-unused-function",
]


# Extra flags per source. matrix.hip: MFMA results and accumulators in
# ordinary VGPRs (gfx950 accepts either form); with the accumulation registers
# the compiler copied every accumulator in and out of them on each loop
# iteration.
SOURCE_FLAGS = {"matrix.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"]}


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found; a ROCm toolchain is required to build numbagg_amd")
    return path


def _compile(src: pathlib.Path) -> pathlib.Path:
    obj = OBJ_DIR / (src.stem + ".o")
    cmd = [hipcc(), *HIPCC_FLAGS, *SOURCE_FLAGS.get(src.name, []), "-I", str(ROOT_DIR / "include"), "-c", str(src),
           "-o", str(obj)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src.name}:\n{proc.stderr}")
    return obj


def build_library(jobs: int | None = None) -> pathlib.Path:
    """Compile every HIP source for gfx950 and link ``libnumbagg_amd.so``."""
    OBJ_DIR.mkdir(exist_ok=True)
    sources = sorted(CSRC_DIR.glob("*.hip"))
    if not sources:
        raise RuntimeError(f"no HIP sources under {CSRC_DIR}")
    jobs = jobs or min(len(sources), max(1, (os.cpu_count() or 4)), 8)
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as pool:
        objs = list(pool.map(_compile, sources))
    tmp = LIB_PATH.with_name(LIB_PATH.name + ".tmp")
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-o", str(tmp), *map(str, objs)]
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        raise RuntimeError(f"link of {LIB_PATH.name} failed:\n{proc.stderr}")
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


def library_is_stale() -> bool:
    """True when the library is missing or older than a source or header."""
    if not LIB_PATH.exists():
        return True
    built = LIB_PATH.stat().st_mtime
    deps = [*CSRC_DIR.glob("*.hip"), *CSRC_DIR.glob("*.hpp"), *(ROOT_DIR / "include").glob("*.h")]
    return any(p.stat().st_mtime > built for p in deps)
