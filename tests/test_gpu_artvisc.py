"""Artificial viscosity (artvisc_o9) on the GPU against the oracle, through
the C-ABI: FP64 within 1e-12 relative, on the TGV and on a Cartesian block
cut into two blocks joined by interfaces (the spectral radii cross the
interface with the viscous exchange)."""
import pytest

from musica2_newgrid_amd import cases, multiblock as mb
from tests._util import relerr

pytestmark = pytest.mark.gpu


def test_tgv_artvisc_matches_oracle():
    from musica2_newgrid_amd.solver import GpuSolver
    from oracle import oracle as orc
    c = cases.tgv(32, filter=3)
    print("TGV 32^3 artificial viscosity, 20 steps", flush=True)
    with GpuSolver(c) as s:
        s.step(20)
        Ug = s.get_state()
    Uc = c.copy_state()
    orc.Oracle(c).step(Uc, 20)
    e = relerr(Ug, Uc)
    print("relerr", e)
    assert e < 1e-12


def test_split_artvisc_matches_oracle():
    from musica2_newgrid_amd.solver import GpuSolver
    from oracle import oracle as orc
    parent = cases.tgv(24, filter=3)
    orients = [((0, 1, 2), (1, 1, 1)), ((0, 1, 2), (-1, 1, 1)), ((0, 1, 2), (1, -1, -1))]
    blocks, bcs = mb.split_oriented(parent, 3, orients)
    print("TGV 24^3 in 3 blocks, artificial viscosity, 10 steps", flush=True)
    with GpuSolver(parent, topology=blocks, block_cases=bcs) as s:
        s.autotune()
        s.step(10)
        Ug = s.get_states()
    Uc = [bc.copy_state() for bc in bcs]
    orc.step_multi(parent, blocks, bcs, Uc, 10)
    e = max(relerr(a, b) for a, b in zip(Ug, Uc))
    print("relerr", e)
    assert e < 1e-12
