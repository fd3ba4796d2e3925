"""The Fortran host (musica2_newgrid_amd/fortran): its positional param.ini
reader, and a TGV run through the iso_c_binding module on the GPU checked
against the oracle."""
import os
import subprocess

import numpy as np
import pytest

from musica2_newgrid_amd import cases
from musica2_newgrid_amd import grid as gr
from tests._util import relerr

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "musica2_newgrid_amd", "bin")
FIXTURE = os.path.join(ROOT, "tests", "golden", "param_tgv.ini")


def parse(path):
    out = subprocess.run([os.path.join(BIN, "param_check"), path], check=True, capture_output=True, text=True)
    d = {}
    for line in out.stdout.splitlines():
        k, v = line.split(None, 1)
        d[k] = v.strip()
    return d


def test_param_reader_on_fixture():
    d = parse(FIXTURE)
    assert d["idepart"] == "1" and d["isolver"] == "1"
    assert d["nrk"] == "4 1" and d["iirs"] == "0" and d["nmax"] == "100"
    assert float(d["CFL"]) == 0.8 and d["stencil"] == "11" and d["is_DRP"] == "F"
    assert d["iorder_visc"] == "4" and d["is_SF"] == "T" and float(d["dissip_coeff"]) == 0.2
    assert d["is_shock"] == "F" and float(d["dissip_shock"]) == 0.1
    assert d["is_ducros"] == "T" and float(d["csens"]) == 0.5


REF_PARAM = "/root/reference/param/param.ini"


@pytest.mark.skipif(not os.path.exists(REF_PARAM), reason="reference tree not mounted")
def test_param_reader_on_the_reference_sample_file():
    """The record layout matches the reference's own sample param.ini."""
    d = parse(REF_PARAM)
    assert d["idepart"] == "1" and d["isolver"] == "2"
    assert d["nrk"] == "4 1" and d["iirs"] == "4" and d["nmax"] == "100000"
    assert float(d["CFL"]) == 8.0 and d["stencil"] == "11"
    assert d["iorder_visc"] == "4" and d["is_SF"] == "T" and float(d["dissip_coeff"]) == 0.2
    assert d["is_shock"] == "F" and d["is_ducros"] == "T" and float(d["csens"]) == 0.5


@pytest.mark.gpu
def test_fortran_tgv_run_matches_oracle(tmp_path):
    n, nsteps = 24, 5
    prefix = str(tmp_path / "tgv")
    print(f"fortran host: TGV {n}^3, {nsteps} steps", flush=True)
    r = subprocess.run([os.path.join(BIN, "musica_gpu_tgv"), FIXTURE, str(n), str(nsteps), prefix],
                       capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    shape = (n + 10, n + 10, n + 10)
    sz = int(np.prod(shape))
    U0 = np.fromfile(prefix + ".init.bin", dtype=np.float64)
    U1 = np.fromfile(prefix + ".final.bin", dtype=np.float64)
    U0 = [U0[q * sz:(q + 1) * sz].reshape(shape).copy() for q in range(5)]
    U1 = [U1[q * sz:(q + 1) * sz].reshape(shape).copy() for q in range(5)]
    info = dict(line.split() for line in open(prefix + ".case.txt"))
    c = cases.tgv(n)
    c.U = U0
    c.dt = float(info["dt"])
    c.diffscale = float(info["diffscale"])
    c.filter = int(info["filter"])
    c.iorder_visc = int(info["iorder_visc"])
    c.chi = float(info["chi"])
    from oracle.oracle import Oracle
    Uc = c.copy_state()
    Oracle(c).step(Uc, nsteps)
    e = relerr(U1, Uc)
    print("fortran-host GPU vs oracle relerr", e)
    assert e < 1e-12
    assert abs(float(np.mean(gr.interior(U1[0]))) / float(np.mean(gr.interior(U0[0]))) - 1.0) < 1e-14


def test_fortran_host_sets_up_the_case_without_crashing(tmp_path):
    """Host-side part of the Fortran driver (param.ini, arrays, initial field,
    case description) runs through; on a host without a GPU it stops cleanly
    at mgpu_create with the library's message."""
    r = subprocess.run([os.path.join(BIN, "musica_gpu_tgv"), FIXTURE, "16", "1", str(tmp_path / "t")],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode >= 0, r.stderr  # no signal (abort, segfault)
    if r.returncode != 0:
        assert "mgpu_create" in r.stdout + r.stderr


@pytest.mark.gpu
def test_fortran_host_artvisc_shock_matches_oracle(tmp_path):
    """is_SF = F with is_shock = T in param.ini: the host selects
    artvisc_o9_shock (assign_procedures.f90:184), checked against the
    oracle."""
    lines = open(FIXTURE).read().splitlines()
    q = lines.index("! selective filter")
    assert lines[q + 1] == "T"
    lines[q + 1] = "F"
    s = lines.index("! shock capturing, low-order coefficient")
    lines[s + 1] = "T 0.1"
    param = tmp_path / "param_avs.ini"
    param.write_text("\n".join(lines) + "\n")
    n, nsteps = 24, 5
    prefix = str(tmp_path / "avs")
    print(f"fortran host: TGV {n}^3 with artvisc_o9_shock, {nsteps} steps", flush=True)
    r = subprocess.run([os.path.join(BIN, "musica_gpu_tgv"), str(param), str(n), str(nsteps), prefix],
                       capture_output=True, text=True, timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    shape = (n + 10, n + 10, n + 10)
    sz = int(np.prod(shape))
    U0 = np.fromfile(prefix + ".init.bin", dtype=np.float64)
    U1 = np.fromfile(prefix + ".final.bin", dtype=np.float64)
    U0 = [U0[k * sz:(k + 1) * sz].reshape(shape).copy() for k in range(5)]
    U1 = [U1[k * sz:(k + 1) * sz].reshape(shape).copy() for k in range(5)]
    info = dict(line.split() for line in open(prefix + ".case.txt"))
    assert int(info["filter"]) == 4
    c = cases.tgv(n)
    c.U = U0
    c.dt = float(info["dt"])
    c.diffscale = float(info["diffscale"])
    c.filter = 4
    c.iorder_visc = int(info["iorder_visc"])
    c.chi = float(info["chi"])
    c.dissip_shock = float(info["dissip_shock"])
    c.csens = float(info["csens"])
    c.ducros = bool(int(info["ducros"]))
    from oracle.oracle import Oracle
    Uc = c.copy_state()
    Oracle(c).step(Uc, nsteps)
    e = relerr(U1, Uc)
    print("fortran-host GPU (artvisc_o9_shock) vs oracle relerr", e)
    assert e < 1e-12
