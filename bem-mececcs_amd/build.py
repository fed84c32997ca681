"""Builds the product in-tree next to the Python package:
  bemx/libbemx.so   HIP kernels + C ABI for gfx950 (hipcc)
  bemx/bemx_main    Fortran driver over the drop-in subroutine surface
                    (amdflang, linked against libbemx.so)
"""
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
FSRC = os.path.join(HERE, "fortran")
OUT = os.path.join(HERE, "bemx", "libbemx.so")
FOUT = os.path.join(HERE, "bemx", "bemx_main")
SOURCES = ["bemx_api.hip", "assembly.hip", "lu.hip", "lu_mixed.hip", "bemx_serve.hip"]
FSOURCES = ["bemx_iface.f90", "bemx_surface.f90", "bemx_main.f90"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FC = os.environ.get("FC", "/opt/rocm/bin/amdflang")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared", "-Werror=return-type",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function"]


def _run(cmd, what, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError("%s failed" % what)


def build_library(verbose: bool = False) -> str:
    _run([HIPCC] + FLAGS + ["-o", OUT + ".tmp"] + [os.path.join(CSRC, s) for s in SOURCES],
         "hipcc (libbemx.so)", verbose)
    os.replace(OUT + ".tmp", OUT)
    return OUT


def build_driver(verbose: bool = False) -> str:
    tmp = tempfile.mkdtemp(prefix="bemx_f_")
    try:
        objs = []
        for s in FSOURCES:
            o = os.path.join(tmp, s.replace(".f90", ".o"))
            _run([FC, "-O2", "-J" + tmp, "-c", os.path.join(FSRC, s), "-o", o],
                 "amdflang (%s)" % s, verbose)
            objs.append(o)
        _run([FC, "-o", FOUT + ".tmp"] + objs +
             ["-L" + os.path.dirname(OUT), "-lbemx", "-Wl,-rpath,$ORIGIN"], "link bemx_main",
             verbose)
        os.replace(FOUT + ".tmp", FOUT)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)
    return FOUT


def build(verbose: bool = False) -> str:
    build_library(verbose)
    build_driver(verbose)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
