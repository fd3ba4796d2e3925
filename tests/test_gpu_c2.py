"""The 2.5-D curvilinear path (MGPU_GEOM_C2, the reference's is_curv) on the
GPU against the oracle, through the C-ABI: a wavy z-extruded block carrying
the isentropic vortex plus density and z-momentum modes along z; Euler only,
viscous with no filter, the selective filter, the shock filter with and
without the Edoh switch; FP64 within 1e-12 relative.  Also the block cut in
two along i (interfaces), graph replays against eager launches and the
autotuner's tiles, the diagnostics, and the mixed path's density error."""
import numpy as np
import pytest

from musica2_newgrid_amd import _lib, cases
from musica2_newgrid_amd import grid as gr
from tests._util import relerr
from tests.test_c2_oracle import _add_z_modes

pytestmark = pytest.mark.gpu
G = gr.G


def _case(filt=2, visc=4, n=(40, 28, 20)):
    c = cases.vortex_wavy_c2(*n, filter=filt, iorder_visc=visc)
    _add_z_modes(c)
    return c


def _oracle(c, nsteps):
    from oracle.oracle import Oracle
    U = c.copy_state()
    Oracle(c).step(U, nsteps)
    return U


def _gpu(c, nsteps, eager=False, tune=False, precision=_lib.PREC_FP64):
    from musica2_newgrid_amd.solver import GpuSolver
    with GpuSolver(c, eager=eager, precision=precision) as s:
        if tune:
            s.autotune()
        s.step(nsteps)
        return s.get_state()


@pytest.mark.parametrize("visc,filt,edoh", [(0, 0, False), (4, 0, False), (4, 1, False), (4, 2, False),
                                            (4, 2, True)])
def test_c2_matches_oracle(visc, filt, edoh):
    c = _case(filt, visc)
    c.sw_edoh = edoh
    if edoh:
        c.dt /= 8.0
    print(f"2.5-D vortex 40x28x20 visc={visc} filter={filt} edoh={edoh}, 5 steps", flush=True)
    e = relerr(_gpu(c, 5), _oracle(c, 5))
    print("relerr", e, flush=True)
    assert e < 1e-12


def _split_i(c):
    """c cut at i = nx/2 into two blocks joined by aligned interfaces."""
    h = c.nx // 2
    blocks, bcs = [], []
    for b in range(2):
        sl = slice(b * h, b * h + h + 2 * G)
        sub = cases.Case(f"half{b}", h, c.ny, c.nz, curvilinear=2, iorder_visc=c.iorder_visc,
                         filter=c.filter, ducros=c.ducros, diffscale=c.diffscale, chi=c.chi, dt=c.dt,
                         m2=[np.ascontiguousarray(a[:, sl]) for a in c.m2],
                         m2v=[np.ascontiguousarray(a[:, sl]) for a in c.m2v],
                         cart=[c.cart[0][:h + 2 * G].copy(), c.cart[1], c.cart[2],
                               c.cart[3][:h + 2 * G].copy(), c.cart[4], c.cart[5]])
        sub.U = [np.ascontiguousarray(a[:, :, sl]) for a in c.U]
        bcs.append(sub)
        blocks.append(dict(nx=h, ny=c.ny, nz=c.nz, nbr_block=[1 - b, 1 - b, b, b, b, b],
                           nbr_face=[1, 0, 3, 2, 5, 4], orient=[0] * 6))
    return blocks, bcs


def test_c2_two_blocks_match_oracle():
    from musica2_newgrid_amd.solver import GpuSolver
    from oracle import oracle as orc
    c = _case(2, 4, (48, 24, 16))
    blocks, bcs = _split_i(c)
    print("2.5-D vortex cut into 2 blocks along i, 48x24x16, 5 steps", flush=True)
    with GpuSolver(c, topology=blocks, block_cases=bcs) as s:
        s.step(5)
        Ug = s.get_states()
    Uc = [bc.copy_state() for bc in bcs]
    orc.step_multi(c, blocks, bcs, Uc, 5)
    e = max(relerr(a, b) for a, b in zip(Ug, Uc))
    W = _oracle(c, 5)
    whole = max(float(np.max(np.abs(np.concatenate([gr.interior(Uc[0][v]), gr.interior(Uc[1][v])], axis=2) -
                                    gr.interior(W[v])))) / float(np.max(np.abs(gr.interior(W[v]))))
                for v in (0, 4))
    print("relerr", e, "blocks vs whole (oracle)", whole, flush=True)
    assert e < 1e-12
    assert whole < 1e-13


def test_c2_graph_matches_eager_and_tuned():
    from musica2_newgrid_amd.solver import GpuSolver
    c = _case()
    print("2.5-D: graph vs eager (bit for bit), then tuned vs oracle", flush=True)
    with GpuSolver(c, eager=True) as se, GpuSolver(c) as sg:
        for rep in range(3):
            se.step(2)
            sg.step(2)
            for a, b in zip(se.get_state(), sg.get_state()):
                assert np.array_equal(gr.interior(a), gr.interior(b)), rep
    e = relerr(_gpu(c, 4, tune=True), _oracle(c, 4))
    print("tuned relerr", e, flush=True)
    assert e < 1e-12


def test_c2_diagnostics_match_oracle():
    from musica2_newgrid_amd.solver import GpuSolver
    from oracle.oracle import Oracle
    c = _case()
    with GpuSolver(c) as s:
        d = s.diagnostics()
    ref = Oracle(c).diag(c.U)
    got = [d["mass"], d["kinetic"], d["enstrophy"], d["umax"]]
    print("diag", got, list(ref), flush=True)
    for a, b in zip(got, ref):
        assert abs(a - b) <= 1e-12 * abs(b)


def test_c2_mixed_precision_error():
    c = _case(2, 4, (32, 24, 16))
    nsteps = 200
    print(f"2.5-D mixed precision, 32x24x16, {nsteps} steps", flush=True)
    Ug = _gpu(c, nsteps, precision=_lib.PREC_MIXED)
    Uc = _oracle(c, nsteps)
    a, b = gr.interior(Ug[0]), gr.interior(Uc[0])
    linf = float(np.max(np.abs(a - b))) / float(np.max(np.abs(b)))
    print(f"rho Linf {linf:.3e} vs the FP64 oracle", flush=True)
    assert 0.0 < linf <= 1e-5
