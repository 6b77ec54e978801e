"""Build recipes for the parity oracle (test infrastructure only).

* ``build_oracle()`` compiles ``oracle/fg_oracle.c`` (the C restatement) into
  ``oracle/lib/libfg_oracle.so`` with ``-O3 -fopenmp -ffp-contract=off`` and no
  ``-march``, so there are no fused multiply-adds, like the reference's default
  x86-64 build.
* ``build_ref()`` compiles the reference's own CPU extension straight from its
  sources under ``/root/reference/FastGeodis`` with g++ (not its setup.py) into
  ``oracle/_ref/FastGeodisCpp*.so``.  The flags mirror what the reference's
  setup.py hands to torch's CppExtension (setup.py:62-113): ``-fopenmp``,
  ``-g0``, ``AT_PARALLEL_OPENMP``, no ``WITH_CUDA``.  Nothing from the reference
  is copied into this repository; only the compiled module lands in
  ``oracle/_ref/`` (git-ignored).

Run ``python oracle/build.py [oracle|ref|all]``.
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_DIR = os.path.join(HERE, "lib")
REF_DIR = os.path.join(HERE, "_ref")
REFERENCE_ROOT = "/root/reference"
REF_SOURCES = [
    "fastgeodis.cpp",
    "fastgeodis_cpu.cpp",
    "geodis_toivanen.cpp",
    "geodis_pixelqueue.cpp",
    "geodis_fastmarch.cpp",
]


def _run(cmd):
    subprocess.run(cmd, check=True)


def build_oracle(verbose: bool = False) -> str:
    os.makedirs(LIB_DIR, exist_ok=True)
    out = os.path.join(LIB_DIR, "libfg_oracle.so")
    cmd = [
        "gcc", "-O3", "-fPIC", "-shared", "-fopenmp", "-ffp-contract=off",
        "-fno-fast-math", "-std=c11", "-Wall",
        os.path.join(HERE, "fg_oracle.c"), "-o", out, "-lm",
    ]
    if verbose:
        print(" ".join(cmd))
    _run(cmd)
    return out


def reference_available() -> bool:
    return os.path.isfile(os.path.join(REFERENCE_ROOT, "FastGeodis", "fastgeodis_cpu.cpp"))


def build_ref(verbose: bool = False) -> str | None:
    """Compile the reference CPU extension; returns the module path or None."""
    if not reference_available():
        return None
    import torch
    from torch.utils import cpp_extension as ce

    os.makedirs(REF_DIR, exist_ok=True)
    src_dir = os.path.join(REFERENCE_ROOT, "FastGeodis")
    inc = ["-I" + src_dir, "-I" + os.path.join(REFERENCE_ROOT, "dependency")]
    inc += ["-I" + p for p in ce.include_paths()]
    inc += ["-I" + sysconfig.get_paths()["include"]]
    abi = int(torch._C._GLIBCXX_USE_CXX11_ABI)
    cflags = [
        "-std=c++17", "-O2", "-fwrapv", "-DNDEBUG", "-g0", "-fPIC", "-fopenmp",
        "-DAT_PARALLEL_OPENMP=1", "-DTORCH_API_INCLUDE_EXTENSION_H",
        "-DTORCH_EXTENSION_NAME=FastGeodisCpp", f"-D_GLIBCXX_USE_CXX11_ABI={abi}",
        "-w",
    ]
    objs = []
    cmds = []
    for s in REF_SOURCES:
        o = os.path.join(REF_DIR, s.replace(".cpp", ".o"))
        objs.append(o)
        cmds.append(["g++", *cflags, *inc, "-c", os.path.join(src_dir, s), "-o", o])
    with ThreadPoolExecutor(max_workers=min(5, os.cpu_count() or 1)) as ex:
        list(ex.map(_run, cmds))
    libdirs = ce.library_paths()
    out = os.path.join(REF_DIR, "FastGeodisCpp" + sysconfig.get_config_var("EXT_SUFFIX"))
    link = ["g++", "-shared", "-fopenmp", *objs, "-o", out]
    for d in libdirs:
        link += ["-L" + d, "-Wl,-rpath," + d]
    link += ["-lc10", "-ltorch", "-ltorch_cpu", "-ltorch_python"]
    if verbose:
        print(" ".join(link))
    _run(link)
    for o in objs:
        os.remove(o)
    return out


if __name__ == "__main__":
    what = sys.argv[1] if len(sys.argv) > 1 else "all"
    if what in ("oracle", "all"):
        print(build_oracle(verbose=True))
    if what in ("ref", "all"):
        print(build_ref(verbose=True))
