"""Build libmusica_gpu.so in-tree with hipcc for gfx950.

    python -m musica2_newgrid_amd.build
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libmusica_gpu.so")

SOURCES = ["mgpu_api.hip", "k_stage.hip", "k_visc.hip", "k_filter.hip", "k_iface.hip", "mgpu_serve.hip", "k_diag.hip",
           "k_wall.hip"]

# -Xclang -target-feature -Xclang -packed-fp32-ops: no packed FP32 VALU ops
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


FORTRAN = os.path.join(HERE, "fortran")
BIN = os.path.join(HERE, "bin")


def flang() -> str:
    for cand in (os.environ.get("FC"), "/opt/rocm/bin/amdflang", shutil.which("amdflang"), shutil.which("flang")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("no Fortran compiler (amdflang) found")


def build_fortran(verbose: bool = False) -> None:
    """The Fortran host: binding module, TGV driver and param.ini reader check."""
    os.makedirs(BIN, exist_ok=True)
    mod_dir = os.path.join(HERE, "_build", "fmod")
    os.makedirs(mod_dir, exist_ok=True)
    fc = flang()
    mod = os.path.join(FORTRAN, "mod_musica_gpu.f90")
    progs = {"musica_gpu_tgv": True, "param_check": True}
    for name, needs_lib in progs.items():
        cmd = [fc, "-O2", "-module-dir", mod_dir, mod, os.path.join(FORTRAN, name + ".f90"),
               "-o", os.path.join(BIN, name)]
        if needs_lib:
            cmd += ["-L" + HERE, "-lmusica_gpu", "-Wl,-rpath,$ORIGIN/.."]
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)


def build(verbose: bool = False, extra=()) -> str:
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = OUT + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + list(extra) + ["-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    build_fortran(verbose)
    return OUT


if __name__ == "__main__":
    build(verbose=True, extra=sys.argv[1:])
