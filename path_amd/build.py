"""Build the HIP library (gfx950) and the CPU checker library.

    python -m path_amd.build            # both
    python -m path_amd.build --lib      # only path_amd/_lib/libpath_amd.so
    python -m path_amd.build --oracle   # only oracle/_build/liboracle.so

Every source is compiled on each call (no caching), in parallel.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "path_amd" / "csrc"
LIB_DIR = ROOT / "path_amd" / "_lib"
LIB = LIB_DIR / "libpath_amd.so"
ORACLE_DIR = ROOT / "oracle"
ORACLE_LIB = ORACLE_DIR / "_build" / "liboracle.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PT_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off: no implicit fusion, so float results are reproducible on
# the host. The packed-fp32 feature is disabled for all device code.
HIP_FLAGS = [
    "-O3", "-std=c++20", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
    "-Wall", "-Wno-unused-function", "-Wno-unused-result",
]
HOST_FLAGS = ["-O3", "-std=c++20", "-fPIC", "-ffp-contract=off", "-Wall"]

SOURCES = [
    "capi.cpp",
    "host/scene.cpp",
    "host/bvh.cpp",
    "host/settings.cpp",
    "host/renderer.cpp",
    "host/streaming.cpp",
    "host/batch.cpp",
    "host/scene_gen.cpp",
    "host/adjoint.cpp",
    "mlp/mlp.cpp",
    "mlp/nrc.cpp",
    "gpu/megakernel.hip",
    "gpu/wavefront.hip",
    "gpu/adjoint.hip",
    "mlp/mlp_kernels.hip",
    "gpu/nrc.hip",
    "gpu/lbvh.hip",
    "gpu/radix_sort.hip",
    "gpu/noalloc.hip",
]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def build_lib(jobs: int = 8) -> Path:
    obj_dir = ROOT / "_build" / "lib"
    obj_dir.mkdir(parents=True, exist_ok=True)
    LIB_DIR.mkdir(parents=True, exist_ok=True)
    srcs = [s for s in SOURCES if (CSRC / s).exists()]
    if not (CSRC / "host/streaming.cpp").exists():
        srcs.append("host/streaming_stub.cpp")
    objs = []
    cmds = []
    for s in srcs:
        obj = obj_dir / (s.replace("/", "_") + ".o")
        objs.append(obj)
        # Host-only sources target x86-64-v3 (FMA3, AVX2), as the oracle does:
        # for baseline x86-64 every fma_() on the host is a call into libm's
        # fmaf (same correctly rounded result, about 3x slower shading-record
        # packing of a 1024-scene batch).
        flags = HIP_FLAGS if s.endswith(".hip") else [*HIP_FLAGS, "-march=x86-64-v3"]
        lang = ["-x", "hip"] if s.endswith(".hip") else []
        cmds.append([HIPCC, *flags, *lang, "-c", str(CSRC / s), "-o", str(obj)])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-fPIC", "-o", str(tmp),
          *map(str, objs), "-L/opt/rocm/lib", "-lrocrand", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, LIB)
    return LIB


CLI = LIB_DIR / "path"


def build_cli() -> Path:
    """The `path` command line (src/main.cpp flags) linked against the library."""
    cxx = os.environ.get("CXX", "g++")
    tmp = CLI.with_suffix(".tmp")
    _run([cxx, *HOST_FLAGS, "-I/opt/rocm/include", str(CSRC / "cli" / "main.cpp"),
          "-o", str(tmp), f"-L{LIB_DIR}", "-lpath_amd", "-Wl,-rpath,$ORIGIN",
          "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, CLI)
    return CLI


def build_oracle() -> Path:
    (ORACLE_DIR / "_build").mkdir(parents=True, exist_ok=True)
    cxx = os.environ.get("CXX", "g++")
    tmp = ORACLE_LIB.with_suffix(".so.tmp")
    _run([cxx, *HOST_FLAGS, "-march=x86-64-v3", "-fopenmp", "-shared",
          "-I/opt/rocm/include",
          str(ORACLE_DIR / "oracle.cpp"), str(ORACLE_DIR / "indep.cpp"),
          str(CSRC / "host" / "bvh.cpp"),
          "-o", str(tmp), "-L/opt/rocm/lib", "-lrocrand", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, ORACLE_LIB)
    return ORACLE_LIB


def main(argv: list[str] | None = None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true")
    ap.add_argument("--oracle", action="store_true")
    ap.add_argument("-j", type=int, default=8)
    a = ap.parse_args(argv)
    both = not (a.lib or a.oracle)
    if a.lib or both:
        print("built", build_lib(a.j))
        print("built", build_cli())
    if a.oracle or both:
        print("built", build_oracle())
    return 0


if __name__ == "__main__":
    sys.exit(main())
