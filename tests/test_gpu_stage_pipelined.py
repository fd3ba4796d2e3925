"""The pipelined stage kernel (the halo of plane k+1 copied straight into LDS
while plane k is computed) against the classic one and the oracle, through
the C-ABI.  Both kernels do the same arithmetic in the same order; the
compiler may contract a different set of multiply-adds in each, so they agree
to round-off (1e-14 relative in FP64, 1e-7 in the mixed path) rather than bit
for bit.  The FP64 results must meet the oracle bar (1e-12 relative).  Covered: Cartesian TGV, c3 and 2.5-D
vortices, the walled channel, ragged blocks whose extents are not multiples
of the tile (partial tiles in i and j, short k-chunks), a multiblock case, and
autotuned tiles within the pipelined family."""
import numpy as np
import pytest

from musica2_newgrid_amd import _lib, cases
from tests._util import relerr

pytestmark = pytest.mark.gpu


def _run(case, steps, kernel, precision=_lib.PREC_FP64, tune=False):
    from musica2_newgrid_amd.solver import GpuSolver
    with GpuSolver(case, precision=precision, stage_kernel=kernel) as s:
        if tune:
            s.autotune(force=True)
        assert s.stage_pipelined(0) == (kernel == "pipelined")
        assert s.stage_pipelined(1) == (kernel == "pipelined")
        s.step(steps)
        return s.get_state()


def _case(which):
    if which == "tgv":
        return cases.tgv(24)
    if which == "c3":
        return cases.vortex_wavy(40, 28, 20, filter=2)
    if which == "c2":
        return cases.vortex_wavy_c2(40, 28, 20, filter=2)
    if which == "channel":
        return cases.channel(32, 36, 12)
    if which == "ragged_cart":
        return cases.tgv(13)
    if which == "ragged_c3":
        return cases.vortex_wavy(37, 19, 7, filter=1)
    if which == "smallest_c3":
        return cases.vortex_wavy(5, 6, 7, filter=2)
    raise ValueError(which)


@pytest.mark.parametrize("which", ["tgv", "c3", "c2", "channel", "ragged_cart", "ragged_c3", "smallest_c3"])
def test_pipelined_equals_classic_and_oracle(which):
    from oracle.oracle import Oracle
    c = _case(which)
    steps = 3
    print(f"pipelined vs classic: {which} {c.nx}x{c.ny}x{c.nz}", flush=True)
    Up = _run(c, steps, "pipelined")
    Uc = _run(c, steps, "classic")
    d = relerr(Up, Uc)
    print("relative difference pipelined vs classic", d, flush=True)
    assert d < 1e-14
    Uo = c.copy_state()
    Oracle(c).step(Uo, steps)
    e = relerr(Up, Uo)
    print("relerr vs oracle", e, flush=True)
    assert e < 1e-12


@pytest.mark.parametrize("which", ["tgv", "c3", "c2", "channel"])
def test_pipelined_mixed_equals_classic(which):
    c = _case(which)
    print(f"pipelined vs classic, mixed: {which}", flush=True)
    Up = _run(c, 3, "pipelined", precision=_lib.PREC_MIXED)
    Uc = _run(c, 3, "classic", precision=_lib.PREC_MIXED)
    d = relerr(Up, Uc)
    print("relative difference pipelined vs classic, mixed", d, flush=True)
    # the two kernels round their FP32 sums in different orders
    assert d < 1e-6


@pytest.mark.parametrize("which", ["c3", "channel"])
def test_pipelined_tuned_matches_oracle(which):
    """Autotuned within the pipelined family (tile 64x8 or 32x16, k-chunks
    from 4 planes to the whole block)."""
    from oracle.oracle import Oracle
    c = _case(which)
    print(f"pipelined, tuned: {which}", flush=True)
    Up = _run(c, 4, "pipelined", tune=True)
    Uo = c.copy_state()
    Oracle(c).step(Uo, 4)
    e = relerr(Up, Uo)
    print("relerr vs oracle", e, flush=True)
    assert e < 1e-12


def test_pipelined_multiblock_matches_oracle():
    """Blocks joined by interfaces with reversed and exchanged axes, one
    batched launch per stage for all blocks."""
    from musica2_newgrid_amd import multiblock as mb
    from musica2_newgrid_amd.solver import GpuSolver
    from oracle.oracle import step_multi
    flags, blocks, bcs = mb.uneven_islands(scale=8)
    print(f"pipelined multiblock: {len(blocks)} blocks", flush=True)
    with GpuSolver(flags, topology=blocks, block_cases=bcs, stage_kernel="pipelined") as s:
        s.step(2)
        Ug = s.get_states()
    Us = [bc.copy_state() for bc in bcs]
    step_multi(flags, blocks, bcs, Us, 2)
    for a, b in zip(Ug, Us):
        e = relerr(a, b)
        print("block relerr", e, flush=True)
        assert e < 1e-12
