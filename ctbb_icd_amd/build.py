"""Build libctbb_icd.so (HIP kernels + C ABI + host C++) for gfx950, in-tree.

Every HIP translation unit is compiled with -ffp-contract=off (the SSM build
and the slice rotation must round exactly like the reference's FP64/FP32
expressions) and with packed FP32 instructions disabled.  Objects go to
ctbb_icd_amd/_build/, the shared library to ctbb_icd_amd/libctbb_icd.so and
the command-line driver to ctbb_icd_amd/ctbb_icd.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
HOST = PKG / "host"
BUILD = PKG / "_build"
LIB = PKG / "libctbb_icd.so"
DRIVER = PKG / "ctbb_icd"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("CTBB_OFFLOAD_ARCH", "gfx950")

COMMON = [
    "-std=c++17", "-O3", "-fPIC", "-ffp-contract=off", "-fno-fast-math",
    f"-I{ROOT / 'include'}", "-Wno-unused-result",
]
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang",
    "-packed-fp32-ops", "-munsafe-fp-atomics", "-fopenmp",
]


def _sources() -> list[Path]:
    srcs = sorted(CSRC.glob("*.hip")) + sorted(CSRC.glob("*.cpp"))
    srcs += sorted(p for p in HOST.glob("*.cpp") if p.name != "main.cpp")
    return srcs


def _compile(src: Path) -> Path:
    obj = BUILD / (src.stem + ("_hip" if src.suffix == ".hip" else "") + ".o")
    if src.suffix == ".hip":
        cmd = [HIPCC, "-x", "hip", *HIP_FLAGS, *COMMON, "-c", str(src), "-o", str(obj)]
    else:
        cmd = [HIPCC, *COMMON, "-fopenmp", "-c", str(src), "-o", str(obj)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return obj


def build(verbose: bool = True) -> Path:
    BUILD.mkdir(exist_ok=True)
    srcs = _sources()
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-fopenmp",
           *map(str, objs), "-o", str(tmp)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    main = HOST / "main.cpp"
    if main.exists():
        cmd = [HIPCC, *COMMON, "-fopenmp", str(main), "-o", str(DRIVER),
               f"-L{PKG}", "-lctbb_icd", f"-Wl,-rpath,$ORIGIN"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"driver link failed: {' '.join(cmd)}\n{r.stderr}")
    if verbose:
        print(f"built {LIB} ({LIB.stat().st_size // 1024} KiB) from {len(srcs)} sources")
    return LIB


def build_oracle(verbose: bool = True) -> None:
    """Build the CPU oracle (and oracle/_ref where the reference is present)."""
    r = subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")],
                       capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{r.stdout}\n{r.stderr}")
    if verbose and r.stdout.strip():
        print(r.stdout.strip())


if __name__ == "__main__":
    build()
    if "--oracle" in sys.argv:
        build_oracle()
