"""Compile recipe for libzindex_amd.so (HIP for gfx950 + C++ host) and the
`zorindex` CLI. Objects go to zindex_amd/_build/, products next to this file.

Run as `python -m zindex_amd.buildlib` or through __graft_entry__.build().
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
ROCM = Path(os.environ.get("ROCM_PATH", "/opt/rocm"))
HIPCC = str(ROCM / "bin" / "hipcc")
ARCH = "gfx950"

LIB = PKG / "libzindex_amd.so"
CLI = PKG / "zorindex"

HIP_SOURCES = ["zg_kernels.hip", "zg_scan.hip", "zg_sweep.hip", "zg_api.hip", "zg_build.hip",
               "zg_serve.hip", "zg_peel.hip", "zg_noalloc.hip"]
CPP_SOURCES = ["zg_format.cpp", "zg_fastx.cpp"]

# Packed FP32 instructions are disabled for all device code (none of these
# kernels use floating point, the flag keeps the compiler from emitting them).
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]
CPP_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result"]
LINK_LIBS = ["-lz", "-l:libzstd.so.1", "-l:liblzma.so.5", "-l:libbz2.so.1", "-lpthread"]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}")


def _compile(src: str) -> Path:
    obj = BUILD / (src + ".o")
    path = CSRC / src
    if src.endswith(".hip"):
        cmd = [HIPCC, *HIP_FLAGS, "-c", str(path), "-o", str(obj)]
    else:
        cmd = ["g++", *CPP_FLAGS, "-I", str(ROCM / "include"), "-c", str(path), "-o", str(obj)]
    _run(cmd)
    return obj


def build(jobs: int | None = None) -> Path:
    """Compile every source (always, also when products exist) and link."""
    BUILD.mkdir(exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 1)
    with ThreadPoolExecutor(jobs) as pool:
        objs = list(pool.map(_compile, HIP_SOURCES + CPP_SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", str(tmp), *map(str, objs),
          *LINK_LIBS])
    os.replace(tmp, LIB)
    cli_obj = BUILD / "zg_cli.o"
    _run(["g++", *CPP_FLAGS, "-c", str(CSRC / "zg_cli.cpp"), "-o", str(cli_obj)])
    _run(["g++", "-o", str(CLI), str(cli_obj), f"-L{PKG}", "-lzindex_amd",
          "-Wl,-rpath,$ORIGIN", f"-Wl,-rpath,{ROCM / 'lib'}"])
    return LIB


if __name__ == "__main__":
    print(build())
