"""No-slip walls on the GPU against the oracle, through the C-ABI: a channel
with walls on both j-faces (stretched y-grid, periodic x and z), FP64 within
1e-12 relative.  The cases cover the viscous last stage with the selective
filter, Euler only without filter (the end-of-step wall state then comes from
the last stage's input), the shock filter with and without the Ducros sensor
(reduced faces next to the walls, the wall row's pressure sensor reading p one
row beyond the wall), isothermal walls, ghost layers beyond the walls
filled with NaN (never read), the channel cut into two blocks (the wall-row
face ghosts keep the neighbour's value from before the wall state, as in the
reference), graph replays against eager launches and the autotuner's tiles."""
import numpy as np
import pytest

from musica2_newgrid_amd import cases
from musica2_newgrid_amd import grid as gr
from tests._util import relerr
from tests.test_wall_oracle import split_x_blocks

pytestmark = pytest.mark.gpu

G = gr.G


def _oracle(c, nsteps):
    from oracle import oracle as orc
    U = c.copy_state()
    orc.Oracle(c).step(U, nsteps)
    return U


def _gpu(c, nsteps, eager=False, tune=False):
    from musica2_newgrid_amd.solver import GpuSolver
    with GpuSolver(c, eager=eager) as s:
        if tune:
            s.autotune()
        s.step(nsteps)
        return s.get_state()


def _interior_relerr(A, B):
    return relerr(B, A)


@pytest.mark.parametrize("visc,filt,nrk,ducros", [(4, 1, 4, False), (0, 0, 4, False), (4, 0, 4, False),
                                                   (0, 1, 6, False), (4, 2, 4, False), (4, 2, 4, True)])
def test_channel_matches_oracle(visc, filt, nrk, ducros):
    c = cases.channel(48, 40, 16, filter=filt, iorder_visc=visc)
    c.ducros = ducros
    if nrk == 6:
        c.nrk, c.crk = 6, cases.CRK6_DRP
    print(f"channel 48x40x16 visc={visc} filter={filt} nrk={nrk} ducros={ducros}, 20 steps", flush=True)
    Ug = _gpu(c, 20)
    Uc = _oracle(c, 20)
    e = _interior_relerr(Uc, Ug)
    print("relerr", e)
    assert e < 1e-12


def test_isothermal_channel_matches_oracle():
    c = cases.channel(32, 24, 12, adiabatic=False, t_wall=285.0)
    print("isothermal channel 32x24x12, 20 steps", flush=True)
    e = _interior_relerr(_oracle(c, 20), _gpu(c, 20))
    print("relerr", e)
    assert e < 1e-12


def test_nan_beyond_walls_is_never_read():
    c = cases.channel(32, 24, 12)
    for a in c.U:
        a[:, :G, :] = np.nan
        a[:, -G:, :] = np.nan
    print("channel with NaN beyond the walls, 10 steps", flush=True)
    Ug = _gpu(c, 10)
    for a in Ug:
        assert np.all(np.isfinite(gr.interior(a)))
    c2 = cases.channel(32, 24, 12)
    Uf = _gpu(c2, 10)
    for a, b in zip(Ug, Uf):
        assert np.array_equal(gr.interior(a), gr.interior(b))


@pytest.mark.parametrize("filt", [1, 2])
def test_channel_cut_in_two_blocks_matches_oracle(filt):
    from musica2_newgrid_amd.solver import GpuSolver
    from oracle import oracle as orc
    c = cases.channel(64, 20, 8, filter=filt)
    c.ducros = filt == 2
    blocks, bcs = split_x_blocks(c)
    print(f"channel cut into 2 blocks, 64x20x8, filter={filt}, 10 steps", flush=True)
    with GpuSolver(c, topology=blocks, block_cases=bcs) as s:
        s.step(10)
        Ug = s.get_states()
    Uc = [bc.copy_state() for bc in bcs]
    orc.step_multi(c, blocks, bcs, Uc, 10)
    e = max(relerr(a, b) for a, b in zip(Uc, Ug))
    print("relerr", e)
    assert e < 1e-12


@pytest.mark.parametrize("filt", [1, 2])
def test_walls_graph_matches_eager_and_tuned(filt):
    c = cases.channel(48, 40, 16, filter=filt)
    print(f"channel filter={filt}: graph vs eager vs tuned, 3 x 3 steps", flush=True)
    from musica2_newgrid_amd.solver import GpuSolver
    with GpuSolver(c, eager=True) as se, GpuSolver(c) as sg:
        for rep in range(3):
            se.step(3)
            sg.step(3)
            for a, b in zip(se.get_state(), sg.get_state()):
                assert np.array_equal(gr.interior(a), gr.interior(b)), rep
    Ut = _gpu(c, 9, tune=True)
    e = _interior_relerr(_oracle(c, 9), Ut)
    print("tuned relerr", e)
    assert e < 1e-12


@pytest.mark.parametrize("filt", [1, 2])
def test_mixed_precision_channel(filt):
    """Mixed precision next to walls (the wall state of the FP32 intermediate
    stage sets, the wall variants of the FP32 stage and filter kernels):
    density error against the FP64 oracle after 300 steps."""
    from musica2_newgrid_amd import _lib
    c = cases.channel(48, 40, 16, filter=filt)
    c.ducros = filt == 2
    nsteps = 300
    print(f"mixed-precision channel 48x40x16 filter={filt}, {nsteps} steps", flush=True)
    from musica2_newgrid_amd.solver import GpuSolver
    with GpuSolver(c, precision=_lib.PREC_MIXED) as s:
        s.step(nsteps)
        Ug = s.get_state()
    Uc = _oracle(c, nsteps)
    a, b = gr.interior(Ug[0]), gr.interior(Uc[0])
    linf = float(np.max(np.abs(a - b))) / float(np.max(np.abs(b)))
    e_all = relerr(Ug, Uc)
    print(f"rho Linf {linf:.3e}, all variables {e_all:.3e} vs the FP64 oracle", flush=True)
    assert 0.0 < linf <= 1e-6


def test_step_needs_the_wall_coefficients():
    """A block with walls refuses to step before mgpu_set_wall."""
    import ctypes as C
    from musica2_newgrid_amd import _lib
    from musica2_newgrid_amd.solver import make_case_struct
    c = cases.channel(16, 20, 6)
    cs, keep = make_case_struct(c)
    L = _lib.lib()
    h = C.c_void_p()
    _lib.check(L.mgpu_create(C.byref(cs), C.byref(h)), None, "mgpu_create")
    try:
        arrs = [np.ascontiguousarray(a) for a in c.cart]
        _lib.check(L.mgpu_set_metrics_cart(h, 0, *[_lib.dptr(a) for a in arrs]), h, "metrics")
        U = [np.ascontiguousarray(a) for a in c.U]
        _lib.check(L.mgpu_put_state(h, 0, *[_lib.dptr(a) for a in U]), h, "state")
        assert L.mgpu_step(h, 1) == _lib.ESTATE
        assert "mgpu_set_wall" in L.mgpu_last_error(h).decode()
        d = _lib.MgpuDiag()
        c2 = np.ascontiguousarray(c.cextrp2, dtype=np.float64)
        c3 = np.ascontiguousarray(c.cextrp3, dtype=np.float64)
        _lib.check(L.mgpu_set_wall(h, 0, _lib.dptr(c2), _lib.dptr(c3)), h, "set_wall")
        _lib.check(L.mgpu_diagnostics(h, C.byref(d)), h, "diagnostics")
        _lib.check(L.mgpu_step(h, 1), h, "step")
    finally:
        L.mgpu_destroy(h)


@pytest.mark.parametrize("which", ["y", "x", "z"])
def test_diagnostics_next_to_walls_match_oracle(which):
    """mgpu_diagnostics on walled channels (one-sided velocity gradients on
    the two rows next to each wall, grad_vel with is_boundary): the stats_tgv
    sums and max(|u| + c) against the oracle, before and after steps; walls
    on the x- or z-faces are relabeled blocks."""
    from musica2_newgrid_amd.solver import GpuSolver
    from oracle.oracle import Oracle
    from tests.test_gpu_walls_axes import _turned
    c = cases.channel(32, 32, 12) if which == "y" else _turned(which)
    print(f"diagnostics, walls on the {which}-faces", flush=True)
    with GpuSolver(c) as s:
        d0 = s.diagnostics()
        s.step(4)
        d1 = s.diagnostics()
        U1 = s.get_state()
    o = Oracle(c)
    for d, U in ((d0, c.U), (d1, U1)):
        ref = o.diag(U)
        got = np.array([d["mass"], d["kinetic"], d["enstrophy"], d["umax"]])
        e = float(np.max(np.abs(got - ref) / np.abs(ref)))
        print("diag relerr", e, got)
        assert e < 1e-12
