"""FP64 parity after 100 steps (the bar the path is held to) on each
geometry and boundary form, not only the periodic TGV: the c3 and 2.5-D
wavy-grid vortices with the shock filter and the Ducros sensor, the walled
channel with the shock filter, and the TGV with the artificial viscosity
with shock capturing; GPU against the oracle within 1e-12 relative."""
import pytest

from musica2_newgrid_amd import cases
from tests._util import relerr
from tests.test_c2_oracle import _add_z_modes

pytestmark = pytest.mark.gpu


def _case(which):
    if which == "vortex_c3":
        return cases.vortex_wavy(32, 24, 16, filter=2)
    if which == "vortex_c2":
        c = cases.vortex_wavy_c2(32, 24, 16, filter=2)
        _add_z_modes(c)
        return c
    if which == "channel":
        c = cases.channel(32, 32, 12, filter=2)
        c.ducros = True
        return c
    c = cases.tgv(24, filter=4)
    c.ducros = True
    return c


@pytest.mark.parametrize("which", ["vortex_c3", "vortex_c2", "channel", "tgv_artvisc_shock"])
def test_100_steps_within_1e12(which):
    from musica2_newgrid_amd.solver import GpuSolver
    from oracle.oracle import Oracle
    c = _case(which)
    print(f"{which} {c.nx}x{c.ny}x{c.nz}, filter {c.filter}, 100 steps", flush=True)
    with GpuSolver(c) as s:
        s.step(100)
        Ug = s.get_state()
    Uc = c.copy_state()
    Oracle(c).step(Uc, 100)
    e = relerr(Ug, Uc)
    moved = relerr(Uc, c.U)
    print(f"relerr {e:.3e} (the state moved by {moved:.3e})", flush=True)
    assert moved > 1e-6
    assert e < 1e-12
