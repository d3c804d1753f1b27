"""Build libebi_hip.so in-tree with hipcc for gfx950.

Every source is compiled to an object under ebi_amd/_build/ (in parallel),
then linked into ebi_amd/libebi_hip.so.  Run as ``python -m ebi_amd.build``.
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
BUILD = PKG / "_build"
LIB = PKG / "libebi_hip.so"
ARCH = os.environ.get("EBI_HIP_ARCH", "gfx950")

COMMON = [
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Wall",
    "-Wno-unused-function",
    "-Wno-unused-value",
    "-Wno-unused-result",
]
DEVICE = [
    f"--offload-arch={ARCH}",
    # no hand-written or compiler-generated packed FP32 ops on this pool
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found (ROCm toolchain required to build libebi_hip)")


def sources() -> list[Path]:
    srcs = sorted((CSRC / "host").glob("*.cpp")) + sorted((CSRC / "kernels").glob("*.hip"))
    return srcs


def _compile(src: Path, cc: str) -> tuple[Path, str]:
    obj = BUILD / (src.parent.name + "_" + src.stem + ".o")
    flags = COMMON + DEVICE
    if src.suffix == ".cpp":
        flags = COMMON + ["-x", "hip", f"--offload-arch={ARCH}", "-Xclang", "-target-feature",
                          "-Xclang", "-packed-fp32-ops"]
    cmd = [cc, *flags, "-I", str(PKG.parent / "include"), "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    log = "\n".join(
        ln for ln in (r.stdout + r.stderr).splitlines() if "packed-fp32-ops" not in ln
    )
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{log}")
    return obj, log


NATIVE_TESTS = PKG.parent / "tests" / "native"


def build_native_tests(cc: str | None = None) -> list[Path]:
    """Host-run checks of device helpers (tests/native/*.hip), built into
    ebi_amd/_build/."""
    cc = cc or hipcc()
    BUILD.mkdir(exist_ok=True)
    out = []
    for src in sorted(NATIVE_TESTS.glob("*.hip")):
        exe = BUILD / src.stem
        # linked against the library (its host encoder makes the test inputs)
        cmd = [cc, *COMMON, *DEVICE, "-I", str(PKG.parent / "include"), str(src), "-o", str(exe),
               "-L", str(LIB.parent), "-lebi_hip", "-Wl,-rpath,$ORIGIN/.."]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"compile failed: {src}\n{r.stdout}{r.stderr}")
        out.append(exe)
    return out


def build(verbose: bool = False) -> Path:
    cc = hipcc()
    BUILD.mkdir(exist_ok=True)
    srcs = sources()
    workers = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")) or 1)
    with cf.ThreadPoolExecutor(max_workers=workers) as ex:
        results = list(ex.map(lambda s: _compile(s, cc), srcs))
    objs = [o for o, _ in results]
    if verbose:
        for _, log in results:
            if log.strip():
                print(log, file=sys.stderr)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stdout}{r.stderr}")
    os.replace(tmp, LIB)
    build_native_tests(cc)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
