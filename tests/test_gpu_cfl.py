"""mgpu_max_cfl (check_max_cfl on the device) against the oracle on every
geometry: the TGV (Cartesian), the c3 and 2.5-D wavy-grid vortices, the walled
channel (stretched y) and the same channel with its walls on the x-faces (a
relabeled block: the library maps its stored axes back), after a few steps.
Within 1e-13 relative (maxima of the same per-point expression)."""
import numpy as np
import pytest

from musica2_newgrid_amd import cases
from tests.test_gpu_walls_axes import _turned

pytestmark = pytest.mark.gpu


def _case(which):
    if which == "tgv":
        return cases.tgv(24)
    if which == "c3":
        return cases.vortex_wavy(32, 24, 16, filter=2)
    if which == "c2":
        return cases.vortex_wavy_c2(32, 24, 16, filter=2)
    if which == "channel":
        return cases.channel(32, 32, 12)
    return _turned("x")


@pytest.mark.parametrize("which", ["tgv", "c3", "c2", "channel", "channel_x_walls"])
def test_max_cfl_matches_oracle(which):
    from musica2_newgrid_amd.solver import GpuSolver
    from oracle.oracle import Oracle
    c = _case(which)
    print(f"max CFL {which} {c.nx}x{c.ny}x{c.nz}", flush=True)
    with GpuSolver(c) as s:
        s.step(3)
        U = s.get_state()
        got = s.max_cfl()
    want = Oracle(c).max_cfl(U)
    print(f"GPU {got}, oracle {want}", flush=True)
    assert np.all(want > 0)
    assert np.allclose(got, want, rtol=1e-13, atol=0)
