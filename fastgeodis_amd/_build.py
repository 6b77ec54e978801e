"""In-tree build of the gfx950 library and the torch glue module.

* ``libfastgeodis_hip.so``: ``csrc/fg_kernels.hip`` compiled by hipcc for
  gfx950 with ``-ffp-contract=off`` (no FMA, bit-exact two-rounding relax),
  without fast-math, and with packed FP32 instructions disabled.  The device
  assembly is checked for scalar-cache stores, which this library never uses.
* ``_fastgeodis_hip*.so``: ``csrc/fg_glue.cpp``, the torch extension that
  mirrors the reference's ``FastGeodisCpp`` module (validation, dispatch,
  workspace from the caching allocator) and calls the C ABI above.  It has
  its own module name so it can be loaded beside the reference's module.

Run ``python fastgeodis_amd/_build.py [hip|glue|all]`` or
``__graft_entry__.build()``.  The recipe is loaded by path, not through the
package, because the package refuses to import while its native module is
missing.
"""
from __future__ import annotations

import os
import re
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB_DIR = os.path.join(HERE, "lib")
INCLUDE = os.path.join(ROOT, "include")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off", "-fno-fast-math", "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]



def _run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, **kw)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}")
    return r


def _asm_guard():
    import importlib.util

    spec = importlib.util.spec_from_file_location("fg_asm_guard", os.path.join(HERE, "_asm_guard.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def check_device_asm(src: str) -> None:
    """Compile the device side to assembly and refuse scalar-cache writes."""
    scalar_cache_writes = _asm_guard().scalar_cache_writes

    asm = os.path.join(LIB_DIR, "fg_kernels.s")
    _run([HIPCC, *HIP_FLAGS, "-I", INCLUDE, "--cuda-device-only", "-S", src, "-o", asm])
    with open(asm) as f:
        text = f.read()
    os.remove(asm)
    bad = scalar_cache_writes(text)
    if bad:
        raise RuntimeError(f"scalar-cache write instructions in device code: {bad}")


def build_hip(verbose: bool = False) -> str:
    os.makedirs(LIB_DIR, exist_ok=True)
    src = os.path.join(CSRC, "fg_kernels.hip")
    out = os.path.join(LIB_DIR, "libfastgeodis_hip.so")
    cmd = [HIPCC, *HIP_FLAGS, "-shared", "-I", INCLUDE, src, "-o", out + ".tmp"]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    check_device_asm(src)
    os.replace(out + ".tmp", out)
    return out


def build_glue(verbose: bool = False) -> str:
    import torch
    from torch.utils import cpp_extension as ce

    os.makedirs(LIB_DIR, exist_ok=True)
    src = os.path.join(CSRC, "fg_glue.cpp")
    out = os.path.join(LIB_DIR, "_fastgeodis_hip" + sysconfig.get_config_var("EXT_SUFFIX"))
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    inc = ["-I" + INCLUDE, "-I/opt/rocm/include", "-I" + sysconfig.get_paths()["include"]]
    inc += ["-I" + p for p in ce.include_paths()]
    cmd = [
        "g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-g0", "-w", "-fvisibility=hidden",
        "-D__HIP_PLATFORM_AMD__=1", "-DUSE_ROCM=1",
        "-DTORCH_API_INCLUDE_EXTENSION_H", "-DTORCH_EXTENSION_NAME=_fastgeodis_hip",
        f"-D_GLIBCXX_USE_CXX11_ABI={abi}", *inc, src, "-o", out + ".tmp",
        "-L" + LIB_DIR, "-lfastgeodis_hip", "-Wl,-rpath,$ORIGIN",
    ]
    for d in ce.library_paths():
        cmd += ["-L" + d, "-Wl,-rpath," + d]
    cmd += ["-lc10", "-lc10_hip", "-ltorch", "-ltorch_cpu", "-ltorch_hip", "-ltorch_python",
            "-L/opt/rocm/lib", "-lamdhip64"]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    os.replace(out + ".tmp", out)
    return out


def build(verbose: bool = False) -> None:
    build_hip(verbose)
    build_glue(verbose)


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    if what in ("hip", "all"):
        print(build_hip(verbose=True))
    if what in ("glue", "all"):
        print(build_glue(verbose=True))
