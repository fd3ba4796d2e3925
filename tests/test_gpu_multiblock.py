"""Multiblock cases on the GPU: several blocks joined by 1-to-1 interfaces
of every orientation code, advanced in one handle, against the oracle's
multiblock step on the same block table, and against the uncut parent."""
import numpy as np
import pytest

from musica2_newgrid_amd import cases, multiblock as mb
from tests._util import relerr
from tests.test_multiblock_oracle import C3_ORIENTS, C3_PARENT, CART_ORIENTS

pytestmark = pytest.mark.gpu


def _run(parent, orients, nsteps):
    from musica2_newgrid_amd.solver import GpuSolver
    from oracle import oracle as orc
    blocks, bcs = mb.split_oriented(parent, len(orients), orients)
    print(f"multiblock {parent.name} {len(blocks)} blocks visc={parent.iorder_visc} "
          f"filter={parent.filter} steps={nsteps}", flush=True)
    with GpuSolver(parent, topology=blocks, block_cases=bcs) as s:
        s.step(nsteps)
        Ug = s.get_states()
    Uc = [bc.copy_state() for bc in bcs]
    orc.step_multi(parent, blocks, bcs, Uc, nsteps)
    return blocks, bcs, Ug, Uc


@pytest.mark.parametrize("filt", [0, 1, 2, 4])
def test_c3_all_orientations_match_oracle(filt):
    """Filter 4 (artvisc_o9_shock) exchanges the spectral radii along the index
    axes: across an interface that exchanges axes, the radius along this
    block's axis comes from the neighbour's axis it meets."""
    parent = cases.vortex_wavy(*C3_PARENT, filter=filt)
    parent.ducros = True
    _, _, Ug, Uc = _run(parent, C3_ORIENTS, 5)
    e = max(relerr(a, b) for a, b in zip(Ug, Uc))
    print("relerr", e)
    assert e < 1e-12


def test_cartesian_reversed_blocks_match_oracle_and_parent():
    from musica2_newgrid_amd.solver import GpuSolver
    parent = cases.tgv(30)
    _, _, Ug, Uc = _run(parent, CART_ORIENTS, 10)
    e = max(relerr(a, b) for a, b in zip(Ug, Uc))
    print("relerr vs oracle", e)
    assert e < 1e-12
    with GpuSolver(parent) as s:
        s.step(10)
        Up = s.get_state()
    ej = relerr(mb.join_oriented(parent, Ug, CART_ORIENTS), Up)
    print("relerr vs parent block", ej)
    assert ej < 1e-12


def test_uneven_blocks_autotune_and_step():
    """Blocks of different sizes in one handle: tuning per block geometry
    keeps the result."""
    from musica2_newgrid_amd.solver import GpuSolver
    parent = cases.vortex_wavy(*C3_PARENT, filter=1)
    blocks, bcs = mb.split_oriented(parent, len(C3_ORIENTS), C3_ORIENTS)
    with GpuSolver(parent, topology=blocks, block_cases=bcs) as s:
        s.step(3)
        U1 = s.get_states()
    with GpuSolver(parent, topology=blocks, block_cases=bcs) as s:
        s.autotune()
        s.step(3)
        U2 = s.get_states()
    for a, b in zip(U1, U2):
        assert relerr(a, b) < 1e-13


def test_graph_replays_match_eager_launches():
    """The captured step, replayed many times with eager work, host copies and
    synchronisations in between, gives bit-for-bit the eager result; so does a
    case without filter, whose U^n set rotates over three graphs."""
    from musica2_newgrid_amd.solver import GpuSolver
    for filt in (1, 0):
        parent = cases.vortex_wavy(*C3_PARENT, filter=filt)
        blocks, bcs = mb.split_oriented(parent, len(C3_ORIENTS), C3_ORIENTS)
        print(f"graph vs eager filter={filt}", flush=True)
        with GpuSolver(parent, topology=blocks, block_cases=bcs, eager=True) as se, \
                GpuSolver(parent, topology=blocks, block_cases=bcs) as sg:
            for rep in range(4):
                se.step(2)
                sg.step(2)
                ue, ug = se.get_states(), sg.get_states()
                for a, b in zip(ue, ug):
                    for x, y in zip(a, b):
                        assert np.array_equal(x[5:-5, 5:-5, 5:-5], y[5:-5, 5:-5, 5:-5]), rep


def test_uneven_seven_blocks_match_oracle():
    """The uneven 7-block workload, scaled down 8x (48x24x12 ... 12^3)."""
    from musica2_newgrid_amd.solver import GpuSolver
    from oracle import oracle as orc
    flags, blocks, bcs = mb.uneven_islands(scale=8)
    print(f"uneven islands: {len(blocks)} blocks, 10 steps", flush=True)
    with GpuSolver(flags, topology=blocks, block_cases=bcs) as s:
        s.autotune()
        s.step(10)
        Ug = s.get_states()
    Uc = [bc.copy_state() for bc in bcs]
    orc.step_multi(flags, blocks, bcs, Uc, 10)
    e = max(relerr(a, b) for a, b in zip(Ug, Uc))
    print("relerr", e)
    assert e < 1e-12
