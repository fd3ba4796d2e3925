"""The artificial viscosity with shock capturing (artvisc_o9_shock, filter 4)
and the 2.5-D artvisc_o9 on the GPU against the oracle, through the C-ABI,
FP64 within 1e-12 relative: the TGV (Ducros sensor off and on), the c3 and
2.5-D wavy-grid vortices, the walled channel (reduced faces next to the
walls in k_shock_wall), the TGV cut into three blocks with reversed axes
(spectral radii, p and the sensor cross the interfaces), graph replays
against eager launches and the autotuner's tiles; the mixed path's density
error against the FP64 oracle."""
import numpy as np
import pytest

from musica2_newgrid_amd import cases, multiblock as mb
from musica2_newgrid_amd import grid as gr
from tests._util import relerr
from tests.test_c2_oracle import _add_z_modes

pytestmark = pytest.mark.gpu


def _gpu(case, nsteps, eager=False, tune=False):
    from musica2_newgrid_amd.solver import GpuSolver
    with GpuSolver(case, eager=eager) as s:
        if tune:
            s.autotune()
        s.step(nsteps)
        return s.get_state()


def _oracle(case, nsteps):
    from oracle.oracle import Oracle
    U = case.copy_state()
    Oracle(case).step(U, nsteps)
    return U


def _case(which):
    if which == "tgv":
        c = cases.tgv(32, filter=4)
        return c, 10
    if which == "tgv_ducros":
        c = cases.tgv(24, filter=4)
        c.ducros = True
        return c, 10
    if which == "vortex_c3":
        return cases.vortex_wavy(40, 28, 20, filter=4), 5
    if which in ("vortex_c2", "vortex_c2_artvisc"):
        c = cases.vortex_wavy_c2(40, 28, 20, filter=4 if which == "vortex_c2" else 3)
        _add_z_modes(c)
        return c, 5
    c = cases.channel(48, 40, 16, filter=4)
    c.ducros = True
    return c, 10


@pytest.mark.parametrize("which", ["tgv", "tgv_ducros", "vortex_c3", "vortex_c2", "vortex_c2_artvisc",
                                   "channel"])
def test_matches_oracle(which):
    c, nsteps = _case(which)
    print(f"filter {c.filter}, {which} {c.nx}x{c.ny}x{c.nz}, {nsteps} steps", flush=True)
    Ug = _gpu(c, nsteps)
    Uc = _oracle(c, nsteps)
    e = relerr(Ug, Uc)
    moved = relerr(Uc, c.U)
    print(f"relerr {e:.3e} (the state moved by {moved:.3e})", flush=True)
    assert moved > 1e-8
    assert e < 1e-12


def test_split_matches_oracle():
    from musica2_newgrid_amd.solver import GpuSolver
    from oracle import oracle as orc
    parent = cases.tgv(24, filter=4)
    parent.ducros = True
    orients = [((0, 1, 2), (1, 1, 1)), ((0, 1, 2), (-1, 1, 1)), ((0, 1, 2), (1, -1, -1))]
    blocks, bcs = mb.split_oriented(parent, 3, orients)
    print("TGV 24^3 in 3 blocks, artvisc_o9_shock, 10 steps", flush=True)
    with GpuSolver(parent, topology=blocks, block_cases=bcs) as s:
        s.step(10)
        Ug = s.get_states()
    Uc = [bc.copy_state() for bc in bcs]
    orc.step_multi(parent, blocks, bcs, Uc, 10)
    e = max(relerr(a, b) for a, b in zip(Ug, Uc))
    print("relerr", e, flush=True)
    assert e < 1e-12


@pytest.mark.parametrize("which", ["tgv", "channel"])
def test_graph_matches_eager_and_tuned(which):
    from musica2_newgrid_amd.solver import GpuSolver
    c, _ = _case(which)
    print(f"filter 4 {which}: graph vs eager (bit for bit), then tuned vs oracle", flush=True)
    with GpuSolver(c, eager=True) as se, GpuSolver(c) as sg:
        for rep in range(3):
            se.step(2)
            sg.step(2)
            for a, b in zip(se.get_state(), sg.get_state()):
                assert np.array_equal(gr.interior(a), gr.interior(b)), rep
    e = relerr(_gpu(c, 4, tune=True), _oracle(c, 4))
    print("tuned relerr", e, flush=True)
    assert e < 1e-12


@pytest.mark.parametrize("which", ["tgv_artvisc", "tgv_ducros", "vortex_c3", "vortex_c2", "channel"])
def test_mixed_precision_error(which):
    """Mixed path (FP32 stage arithmetic and states, FP64 dissipation): density error
    against the FP64 oracle after 200 steps (100 for the channel)."""
    from musica2_newgrid_amd import _lib
    from musica2_newgrid_amd.solver import GpuSolver
    if which == "tgv_artvisc":
        c, nsteps = cases.tgv(24, filter=3), 200
    elif which in ("vortex_c3", "vortex_c2"):
        c = (cases.vortex_wavy(32, 24, 16, filter=4) if which == "vortex_c3"
             else cases.vortex_wavy_c2(32, 24, 16, filter=4))
        nsteps = 200
    else:
        c, _ = _case(which)
        nsteps = 100 if which == "channel" else 200
    print(f"mixed precision, filter {c.filter}, {which} {c.nx}x{c.ny}x{c.nz}, {nsteps} steps", flush=True)
    with GpuSolver(c, precision=_lib.PREC_MIXED) as s:
        s.step(nsteps)
        Ug = s.get_state()
    Uc = _oracle(c, nsteps)
    a, b = gr.interior(Ug[0]), gr.interior(Uc[0])
    linf = float(np.max(np.abs(a - b))) / float(np.max(np.abs(b)))
    print(f"rho Linf {linf:.3e} vs the FP64 oracle", flush=True)
    assert 0.0 < linf <= 1e-5


def test_artvisc_channel_with_walls():
    """artvisc_o9 next to walls runs through the artvisc_o9_shock kernels with
    no sensor term: FP64 parity and the mixed path's density error."""
    from musica2_newgrid_amd import _lib
    from musica2_newgrid_amd.solver import GpuSolver
    c = cases.channel(48, 40, 16, filter=3)
    print("artvisc_o9 channel 48x40x16 with walls, 10 steps", flush=True)
    e = relerr(_gpu(c, 10), _oracle(c, 10))
    print(f"relerr {e:.3e}", flush=True)
    assert e < 1e-12
    with GpuSolver(c, precision=_lib.PREC_MIXED) as s:
        s.step(100)
        Ug = s.get_state()
    Uc = _oracle(c, 100)
    a, b = gr.interior(Ug[0]), gr.interior(Uc[0])
    linf = float(np.max(np.abs(a - b))) / float(np.max(np.abs(b)))
    print(f"mixed, 100 steps: rho Linf {linf:.3e} vs the FP64 oracle", flush=True)
    assert 0.0 < linf <= 1e-6
