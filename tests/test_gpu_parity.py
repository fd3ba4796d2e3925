"""Parity of the HIP path (through the C-ABI) with the CPU oracle.

Bar: FP64 GPU within 1e-12 relative of the oracle after 100 steps on the TGV
(64^3, the reference's CPU-runnable configuration), and to round-off on the
other paths (Euler only, selective filter, shock filter + Ducros, curvilinear
metrics).  Relative error: tests._util.relerr (per variable, normalised by the
variable's magnitude; the momentum components by the momentum magnitude).
"""
import numpy as np
import pytest

from musica2_newgrid_amd import cases
from musica2_newgrid_amd import grid as gr
from tests._util import relerr

pytestmark = pytest.mark.gpu


def _solver(case):
    from musica2_newgrid_amd.solver import GpuSolver
    return GpuSolver(case)


def _oracle(case):
    from oracle.oracle import Oracle
    return Oracle(case)


def run_both(case, nsteps):
    print(f"case {case.name} {case.nx}x{case.ny}x{case.nz} visc={case.iorder_visc} "
          f"filter={case.filter} steps={nsteps}", flush=True)
    with _solver(case) as s:
        s.step(nsteps)
        Ug = s.get_state()
    Uc = case.copy_state()
    _oracle(case).step(Uc, nsteps)
    return Ug, Uc


@pytest.mark.parametrize("visc,filt", [(0, 0), (0, 1), (4, 0), (4, 1)])
def test_tgv_one_step(visc, filt):
    case = cases.tgv(24, iorder_visc=visc, filter=filt)
    Ug, Uc = run_both(case, 1)
    e = relerr(Ug, Uc)
    print("relerr", e)
    assert e < 1e-13


def test_tgv_64_100_steps():
    """Configuration 1: TGV 64^3, viscous, SF chi=0.2, RK4, CFL 0.8, 100 steps."""
    case = cases.tgv(64)
    Ug, Uc = run_both(case, 100)
    e = relerr(Ug, Uc)
    print("relerr after 100 steps", e)
    assert e < 1e-12


@pytest.mark.parametrize("filt", [0, 1, 2])
def test_vortex_curvilinear(filt):
    case = cases.vortex_wavy(40, 28, 20, filter=filt)
    Ug, Uc = run_both(case, 3)
    e = relerr(Ug, Uc)
    print("relerr", e)
    assert e < 1e-12


def test_vortex_curvilinear_euler_only():
    case = cases.vortex_wavy(36, 24, 16, filter=0, iorder_visc=0)
    Ug, Uc = run_both(case, 3)
    e = relerr(Ug, Uc)
    print("relerr", e)
    assert e < 1e-12


def test_affine_c3_matches_cartesian():
    """A c3 block with the metrics of an undistorted box is the Cartesian block."""
    cart = cases.tgv(20)
    c3 = cases.tgv(20)
    c3.curvilinear = True
    L = cart.info["L"]
    c3.m, c3.mv = cases.affine_c3_metrics(20, 20, 20, (L, L, L))
    with _solver(cart) as s:
        s.step(2)
        A = s.get_state()
    with _solver(c3) as s:
        s.step(2)
        B = s.get_state()
    e = relerr(B, A)
    print("relerr c3 vs cartesian", e)
    assert e < 1e-12


def test_tgv_mass_conservation():
    """Periodic Cartesian block: the discrete divergence and both filters are
    in conservation form, so total mass is kept to round-off."""
    case = cases.tgv(48)
    m0 = float(np.sum(gr.interior(case.U[0])))
    with _solver(case) as s:
        s.step(20)
        U = s.get_state()
    m1 = float(np.sum(gr.interior(U[0])))
    print("mass drift", (m1 - m0) / m0)
    assert abs(m1 - m0) / m0 < 1e-13


def test_full_size_parity_and_freestream():
    """At the 512x256x128 curvilinear size (shock filter + Ducros, viscous):
    one step matches the oracle, and a uniform flow stays uniform (GCL
    free-stream preservation)."""
    nx, ny, nz = 512, 256, 128
    case = cases.vortex_wavy(nx, ny, nz, filter=2)
    Ug, Uc = run_both(case, 1)
    e = relerr(Ug, Uc)
    print("full-size relerr after 1 step", e)
    assert e < 1e-12
    del Ug, Uc
    U0 = [gr.padded(nx, ny, nz, v) for v in cases.conserved(1.177, 170.0, 10.0, -5.0, 101325.0)]
    case.U = U0
    with _solver(case) as s:
        s.step(2)
        U = s.get_state()
    e = relerr(U, U0)
    print("free-stream relative change", e)
    assert e < 1e-12


@pytest.mark.parametrize("which", ["tgv", "vortex_c3"])
def test_diagnostics_match_oracle(which):
    """mgpu_diagnostics (stats_tgv sums) against the oracle, before and after steps."""
    from musica2_newgrid_amd.solver import GpuSolver
    case = cases.tgv(32) if which == "tgv" else cases.vortex_wavy(40, 28, 20, filter=2)
    print(f"diagnostics {which}", flush=True)
    with GpuSolver(case) as s:
        d0 = s.diagnostics()
        s.step(5)
        d1 = s.diagnostics()
        U1 = s.get_state()
    o = _oracle(case)
    for d, U in ((d0, case.U), (d1, U1)):
        ref = o.diag(U)
        got = np.array([d["mass"], d["kinetic"], d["enstrophy"], d["umax"]])
        e = np.max(np.abs(got - ref) / np.abs(ref))
        print("diag relerr", e, got)
        assert e < 1e-12
        assert d["cells"] == case.ncells


def test_rk6_drp_matches_oracle():
    """nrk = 6 (init_RK's DRP scheme) through the whole path: c3, viscous, shock filter."""
    case = cases.vortex_wavy(40, 28, 20, filter=2)
    case.nrk, case.crk = 6, cases.CRK6_DRP
    case.dt *= 1.5
    Ug, Uc = run_both(case, 10)
    e = relerr(Ug, Uc)
    print("relerr", e)
    assert e < 1e-12


@pytest.mark.parametrize("shape", [(5, 6, 7), (33, 5, 9), (7, 17, 5)])
def test_smallest_and_ragged_blocks(shape):
    """Extents down to the 5 ghost layers and extents that leave partial tiles in
    every direction, c3 with viscous terms and the shock filter."""
    nx, ny, nz = shape
    case = cases.vortex_wavy(nx, ny, nz, filter=2)
    Ug, Uc = run_both(case, 3)
    e = relerr(Ug, Uc)
    print("relerr", e)
    assert e < 1e-12
