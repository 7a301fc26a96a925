"""In-tree build of the gfx950 library (hipcc, no JIT cache).

Produces ``pedalsky_amd/lib/libpedalsky_hip.so`` from the sources in
``pedalsky_amd/csrc``. Each translation unit is compiled in parallel, then
linked. Packed FP32 instructions are disabled for all device code.
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
LIB_DIR = PKG / "lib"
LIB = LIB_DIR / "libpedalsky_hip.so"
OBJ_DIR = PKG / "_build"
INCLUDE = PKG.parent / "include"

SOURCES = ["lut_kernels.hip", "render_kernel.hip", "post_kernel.hip", "adjoint_kernel.hip", "atmo_adjoint.hip",
           "noise_kernels.hip", "noalloc.hip", "capi.cpp", "batch.cpp", "host_synthesis.cpp", "astro.cpp"]

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PEDALSKY_OFFLOAD_ARCH", "gfx950")

COMMON = ["-O3", "-std=c++17", "-fPIC", f"-I{INCLUDE}", f"-I{CSRC}", "-Wall", "-Wno-unused-function"]
DEVICE = [f"--offload-arch={ARCH}", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-munsafe-fp-atomics"]


# The march runs FP32 divides and square roots in the loop body; their
# correctly-rounded expansions are ~10 instructions each, the hardware
# approximations (<= 2.5 ulp) one or two. The adjoint recomputes the march
# and must round exactly as the forward does, so it shares the flag. The LUT
# kernels are FP64 and the noise bakes keep the WGSL's IEEE FP32 semantics.
#
# The march kernels are also compiled without machine-level loop-invariant
# code motion: it hoists constant materialisations out of the step loops and
# keeps them live across the whole tile, which costs ~24 VGPRs (and so a wave
# per SIMD, or scratch). Arithmetic is unchanged by it.
# The atmosphere adjoint's image-gradient kernel drops from 175 to 109 VGPRs
# the same way (two waves per SIMD to four).
PER_FILE = {"render_kernel.hip": ["-fno-hip-fp32-correctly-rounded-divide-sqrt", "-mllvm", "-disable-machine-licm"],
            "adjoint_kernel.hip": ["-fno-hip-fp32-correctly-rounded-divide-sqrt", "-mllvm", "-disable-machine-licm"],
            "atmo_adjoint.hip": ["-mllvm", "-disable-machine-licm"]}


def _compile(src: str) -> Path:
    obj = OBJ_DIR / (Path(src).stem + ".o")
    path = CSRC / src
    if src.endswith(".hip"):
        cmd = [HIPCC, *COMMON, *DEVICE, *PER_FILE.get(src, []), "-x", "hip", "-c", str(path), "-o", str(obj)]
    else:
        # host-only translation units still go through hipcc for the HIP runtime headers
        cmd = [HIPCC, *COMMON, "-D__HIP_PLATFORM_AMD__", "-c", str(path), "-o", str(obj)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    return obj


def build(verbose: bool = False) -> Path:
    """Compile every translation unit for gfx950 and link the shared library."""
    OBJ_DIR.mkdir(parents=True, exist_ok=True)
    LIB_DIR.mkdir(parents=True, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


HOST_TEST_SRC = PKG.parent / "tests" / "host" / "subsystem_frame.cpp"
HOST_TEST_BIN = PKG.parent / "tests" / "host" / "_build" / "subsystem_frame"


def build_host_test(verbose: bool = False) -> Path:
    """Compile the driver of the C++ host layer (include/pedalsky_subsystems.hpp)
    that tests/test_host_subsystems.py runs, linked against the library."""
    HOST_TEST_BIN.parent.mkdir(parents=True, exist_ok=True)
    tmp = HOST_TEST_BIN.with_suffix(".tmp")
    cmd = [HIPCC, "-O2", "-std=c++17", "-D__HIP_PLATFORM_AMD__", f"-I{INCLUDE}", "-Wall", str(HOST_TEST_SRC),
           "-o", str(tmp), f"-L{LIB_DIR}", "-lpedalsky_hip", f"-Wl,-rpath,{LIB_DIR}",
           "-Wl,-rpath,$ORIGIN/../../../pedalsky_amd/lib"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"host test build failed: {' '.join(cmd)}\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, HOST_TEST_BIN)
    if verbose:
        print(f"built {HOST_TEST_BIN}")
    return HOST_TEST_BIN


def build_oracle(verbose: bool = False) -> Path:
    """Compile the CPU oracle (test infrastructure) with its Makefile."""
    oracle_dir = PKG.parent / "oracle"
    res = subprocess.run(["make", "-s", "-C", str(oracle_dir), "-B"], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"oracle build failed:\n{res.stdout}\n{res.stderr}")
    out = oracle_dir / "_build" / "liboracle.so"
    if verbose:
        print(f"built {out}")
    return out


if __name__ == "__main__":
    build(verbose=True)
    if "--oracle" in sys.argv:
        build_oracle(verbose=True)
