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
           "k_wall.hip", "k_cfl.hip"]

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
    """Each source compiles to its own object (in parallel: every kernel is
    launched from its own translation unit, so no relocatable device code is
    needed), then one link."""
    from concurrent.futures import ThreadPoolExecutor
    objdir = os.path.join(HERE, "_build", "obj")
    os.makedirs(objdir, exist_ok=True)
    flags = [f for f in HIPCC_FLAGS if f != "-shared"] + list(extra)
    jobs = []
    for s in SOURCES:
        obj = os.path.join(objdir, os.path.splitext(s)[0] + ".o")
        jobs.append(([hipcc()] + flags + ["-c", "-o", obj, os.path.join(CSRC, s)], obj))

    def run(job):
        cmd, _ = job
        if verbose:
            print(" ".join(cmd), flush=True)
        subprocess.run(cmd, check=True)

    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", "0")) or (os.cpu_count() or 4)))
    with ThreadPoolExecutor(max_workers=workers) as ex:
        list(ex.map(run, jobs))
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + [o for _, o in jobs]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, OUT)
    build_fortran(verbose)
    return OUT


if __name__ == "__main__":
    build(verbose=True, extra=sys.argv[1:])
