"""The shock filter's Edoh switch (is_sw_edoh) on the GPU against the oracle,
through the C-ABI: Cartesian (TGV), c3 (wavy-grid vortex, contravariant
local CFL), the walled channel (the wall rows' j-increment scaled in
k_shock_wall) and the mixed path.  The time steps are cut so that the local
CFL/chi stays mostly below 1, where the switch acts."""
import numpy as np
import pytest

from musica2_newgrid_amd import _lib, cases
from musica2_newgrid_amd import grid as gr
from tests._util import relerr

pytestmark = pytest.mark.gpu


def _gpu(case, nsteps, precision=_lib.PREC_FP64):
    from musica2_newgrid_amd.solver import GpuSolver
    with GpuSolver(case, precision=precision) as s:
        s.step(nsteps)
        return s.get_state()


def _oracle(case, nsteps):
    from oracle.oracle import Oracle
    U = case.copy_state()
    Oracle(case).step(U, nsteps)
    return U


def _edoh_acts(case, nsteps, U_edoh):
    """The switch changes the result by far more than the parity bar."""
    case.sw_edoh = False
    U_plain = _oracle(case, nsteps)
    case.sw_edoh = True
    return relerr(U_plain, U_edoh)


@pytest.mark.parametrize("which", ["tgv", "vortex_c3", "channel"])
def test_edoh_matches_oracle(which):
    if which == "tgv":
        case = cases.tgv(24, cfl=0.1, filter=2)
        nsteps = 10
    elif which == "vortex_c3":
        case = cases.vortex_wavy(40, 28, 20, filter=2)
        case.dt /= 8.0
        nsteps = 5
    else:
        case = cases.channel(48, 40, 16, filter=2)
        case.ducros = True
        case.dt /= 4.0
        nsteps = 10
    case.sw_edoh = True
    print(f"Edoh switch, {which} {case.nx}x{case.ny}x{case.nz}, {nsteps} steps", flush=True)
    Ug = _gpu(case, nsteps)
    Uc = _oracle(case, nsteps)
    e = relerr(Ug, Uc)
    d = _edoh_acts(case, nsteps, Uc)
    print(f"relerr {e:.3e}; with the switch vs without {d:.3e}", flush=True)
    assert e < 1e-12
    assert d > 1e-9


def test_edoh_mixed_precision():
    case = cases.vortex_wavy(32, 24, 16, filter=2)
    case.dt /= 8.0
    case.sw_edoh = True
    nsteps = 200
    print(f"Edoh switch, mixed precision, vortex_c3 32x24x16, {nsteps} steps", flush=True)
    Ug = _gpu(case, nsteps, precision=_lib.PREC_MIXED)
    Uc = _oracle(case, nsteps)
    a, b = gr.interior(Ug[0]), gr.interior(Uc[0])
    linf = float(np.max(np.abs(a - b))) / float(np.max(np.abs(b)))
    print(f"rho Linf {linf:.3e} vs the FP64 oracle", flush=True)
    assert 0.0 < linf <= 1e-5
