"""Build libinca_hip.so (hipcc, gfx950) and the Fortran host `inca` (amdflang).

    python -m inca_parallel_amd.build            # everything
    python -m inca_parallel_amd.build --hip      # the HIP library only

Outputs stay in-tree: inca_parallel_amd/lib/libinca_hip.so and
inca_parallel_amd/bin/inca.  Objects go to inca_parallel_amd/_build/.
"""
from __future__ import annotations

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
FSRC = os.path.join(PKG, "fortran")
BUILD = os.path.join(PKG, "_build")
LIBDIR = os.path.join(PKG, "lib")
BINDIR = os.path.join(PKG, "bin")
LIB = os.path.join(LIBDIR, "libinca_hip.so")
EXE = os.path.join(BINDIR, "inca")

ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = shutil.which("hipcc") or os.path.join(ROCM, "bin", "hipcc")
FC = shutil.which("amdflang") or os.path.join(ROCM, "bin", "amdflang")

HIP_SOURCES = ["util.hip", "grid.hip", "quartet.hip", "group.hip", "eval.hip", "autogrid.hip",
               "batch.hip",
               "mlp.hip", "onee.hip",
               "api.hip"]
# the Fortran modules in dependency order
F_SOURCES = ["mod_inca_hip.f90", "mod_state.f90", "mod_locate.f90",
             "mod_readers.f90", "mod_autogrid.f90", "mod_input.f90",
             "mod_intracule.f90", "mod_cubefile.f90", "inca_main.f90"]

HIP_FLAGS = ["-O3", "--offload-arch=gfx950", "-std=c++17", "-fPIC",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wall", "-Wno-unused-function"]


def _run(cmd, cwd=None):
    r = subprocess.run(cmd, cwd=cwd, stdout=subprocess.PIPE,
                       stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "is not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("command failed: %s\n%s" % (" ".join(cmd), out))
    return out


def build_hip(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)

    def one(src):
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        out = _run([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, src), "-o", obj])
        if verbose and out.strip():
            print(out)
        return obj

    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(one, HIP_SOURCES))
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    return LIB


def build_fortran(verbose=False):
    if not all(os.path.exists(os.path.join(FSRC, f)) for f in F_SOURCES):
        return None
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(BINDIR, exist_ok=True)
    objs = []
    for f in F_SOURCES:
        obj = os.path.join(BUILD, f.replace(".f90", ".o"))
        out = _run([FC, "-O2", "-J", BUILD, "-c", os.path.join(FSRC, f), "-o", obj])
        if verbose and out.strip():
            print(out)
        objs.append(obj)
    tmp = EXE + ".tmp"
    _run([FC, "-o", tmp] + objs + ["-L" + LIBDIR, "-linca_hip",
                                   "-Wl,-rpath,$ORIGIN/../lib"])
    os.replace(tmp, EXE)
    return EXE


def build_oracle():
    subprocess.run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle")], check=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--hip", action="store_true")
    ap.add_argument("--fortran", action="store_true")
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args(argv)
    everything = not (a.hip or a.fortran)
    if a.hip or everything:
        print("built", build_hip(a.v))
    if a.fortran or everything:
        print("built", build_fortran(a.v))
    if everything:
        build_oracle()
        print("built oracle/liborc.so")


if __name__ == "__main__":
    sys.exit(main())
