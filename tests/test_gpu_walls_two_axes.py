"""No-slip walls on two axes of one Cartesian block on the GPU against the
oracle, through the C-ABI: a duct walled on its i- and j-faces (clustered at
every wall, periodic z) with no filter, the selective filter and the shock
filter (Ducros), adiabatic and isothermal walls, FP64 within 1e-12 relative;
the edges taking the j-faces' wall state (the one-stage linear-pressure case
of the oracle pin); walls on the i- and k-faces (the block stored with j and
k exchanged); walls on the j- and k-faces (the block stored with its j, k, i
as i, j, k, so the edges take the k-faces' state, last in bc_wall's order),
also cut in two blocks along i (no-slip or advanced walls);
graph replays bit-identical to eager launches; the mixed path against FP64."""
import numpy as np
import pytest

from musica2_newgrid_amd import _lib, cases
from musica2_newgrid_amd import grid as gr
from tests._util import relerr
from tests.test_duct_oracle import cycle_axes as _cycle

pytestmark = pytest.mark.gpu

G = gr.G


def _oracle(c, n):
    from oracle.oracle import Oracle
    U = c.copy_state()
    Oracle(c).step(U, n)
    return U


def _gpu(c, n, eager=False, precision=_lib.PREC_FP64):
    from musica2_newgrid_amd.solver import GpuSolver
    with GpuSolver(c, eager=eager, precision=precision) as s:
        s.step(n)
        return s.get_state()


@pytest.mark.parametrize("filt,ducros,adiabatic", [(2, True, True), (1, False, True), (0, False, True),
                                                   (2, False, False)])
def test_duct_matches_oracle(filt, ducros, adiabatic):
    c = cases.duct(40, 36, 12, filter=filt, adiabatic=adiabatic)
    c.ducros = ducros
    n = 20
    print(f"duct 40x36x12 filter={filt} ducros={ducros} adiabatic={adiabatic}, {n} steps", flush=True)
    Ug = _gpu(c, n)
    Uo = _oracle(c, n)
    e = relerr(Ug, Uo)
    print(f"relerr {e:.3e}", flush=True)
    assert e < 1e-12
    for v in (1, 2, 3):
        w = Ug[v][G:-G, G:-G, G:-G]
        assert np.all(w[:, 0, :] == 0.0) and np.all(w[:, :, 0] == 0.0) and np.all(w[:, :, -1] == 0.0)


def test_edges_take_the_j_faces_wall_state_on_the_gpu():
    """The oracle pin's one-stage linear-pressure duct, GPU against oracle."""
    from tests.test_duct_oracle import linear_duct
    c, _, _ = linear_duct()
    print("linear-pressure duct, one stage", flush=True)
    Ug = _gpu(c, 1)
    Uo = _oracle(c, 1)
    for a, b in zip(Ug, Uo):
        assert np.allclose(gr.interior(a), gr.interior(b), rtol=1e-13, atol=0.0)


def _swap_jk(c):
    """The same duct with y and z exchanged: walls on the i- and k-faces."""
    t = cases.Case(c.name + "_ik", c.nx, c.nz, c.ny, curvilinear=False, iorder_visc=c.iorder_visc,
                   filter=c.filter, ducros=c.ducros, diffscale=c.diffscale, chi=c.chi, dt=c.dt, nrk=c.nrk,
                   crk=c.crk, cart=[c.cart[0], c.cart[2], c.cart[1], c.cart[3], c.cart[5], c.cart[4]],
                   faces=(2, 2, 0, 0, 2, 2), adiabatic=c.adiabatic, t_wall=c.t_wall,
                   cextrp2=(c.cextrp2[0], c.cextrp2[1], 0, 0, c.cextrp2[2], c.cextrp2[3]),
                   cextrp3=(c.cextrp3[0], c.cextrp3[1], 0, 0, c.cextrp3[2], c.cextrp3[3]))
    U = [np.ascontiguousarray(np.swapaxes(a, 0, 1)) for a in c.U]
    U[2], U[3] = U[3], U[2]
    t.U = U
    return t


@pytest.mark.parametrize("filt,ducros,adiabatic", [(2, True, True), (0, False, True), (2, False, False)])
def test_walls_on_the_j_and_k_faces_match_oracle(filt, ducros, adiabatic):
    c = cases.duct(36, 32, 12, filter=filt, adiabatic=adiabatic)
    c.ducros = ducros
    t = _cycle(c)
    n = 10
    print(f"duct with walls on the j- and k-faces {t.nx}x{t.ny}x{t.nz} filter={filt} ducros={ducros} "
          f"adiabatic={adiabatic}, {n} steps", flush=True)
    Ug = _gpu(t, n)
    Uo = _oracle(t, n)
    e = relerr(Ug, Uo)
    print(f"relerr {e:.3e}", flush=True)
    assert e < 1e-12
    # the same flow as the duct itself, axes cycled
    Ud = _oracle(c, n)
    for v, w in zip((0, 1, 2, 3, 4), (0, 3, 1, 2, 4)):
        a = gr.interior(Ug[v])
        b = np.transpose(gr.interior(Ud[w]), (1, 2, 0))
        assert np.allclose(a, b, rtol=1e-10, atol=1e-10 * float(np.max(np.abs(b)) + 1.0))


def test_walls_on_the_j_and_k_faces_graph_and_mixed():
    c = cases.duct(40, 36, 12, filter=2)
    c.ducros = True
    t = _cycle(c)
    print(f"duct with walls on the j- and k-faces {t.nx}x{t.ny}x{t.nz}: graph vs eager (6 steps), "
          f"mixed vs FP64 (30 steps)", flush=True)
    Ua = _gpu(t, 6)
    Ub = _gpu(t, 6, eager=True)
    for a, b in zip(Ua, Ub):
        assert np.array_equal(gr.interior(a), gr.interior(b))
    Um = _gpu(t, 30, precision=_lib.PREC_MIXED)
    Ud = _gpu(t, 30)
    a, b = gr.interior(Um[0]), gr.interior(Ud[0])
    e = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
    print(f"mixed vs FP64 rho Linf {e:.3e}", flush=True)
    assert np.all(np.isfinite(a)) and e < 1e-6


@pytest.mark.parametrize("wall2", [False, True])
def test_j_and_k_walled_duct_in_two_blocks(wall2):
    """The axis-cycled duct cut in two blocks along i: the interfaces sit on
    i-faces, stored as k-faces with their tangential axes reversed; advanced
    walls (no slip on the j-faces, slip on the k-faces) or no-slip walls."""
    from oracle import oracle as orc
    c = cases.duct(40, 36, 16, filter=2)
    c.ducros = True
    if wall2:
        c.wall2 = True
        c.faces = (2, 2, 5, 5, 0, 0)
    t = _cycle(c)
    h = t.nx // 2
    blocks, bcs = [], []
    for b in range(2):
        lo, hi = b * h, b * h + h + 2 * G
        sub = cases.Case(f"half{b}", h, t.ny, t.nz, curvilinear=False, iorder_visc=t.iorder_visc,
                         filter=t.filter, ducros=t.ducros, diffscale=t.diffscale, chi=t.chi, dt=t.dt,
                         cart=[t.cart[0][lo:hi].copy(), t.cart[1], t.cart[2], t.cart[3][lo:hi].copy(), t.cart[4],
                               t.cart[5]],
                         faces=(1, 1) + tuple(t.faces[2:]), adiabatic=t.adiabatic, t_wall=t.t_wall,
                         cextrp2=t.cextrp2, cextrp3=t.cextrp3, wall2=t.wall2)
        sub.U = [np.ascontiguousarray(a[:, :, lo:hi]) for a in t.U]
        bcs.append(sub)
        kinds = [_lib.FACE_INTERFACE] * 2 + [int(k) for k in t.faces[2:]]
        blocks.append(dict(nx=h, ny=t.ny, nz=t.nz, kind=kinds, nbr_block=[1 - b, 1 - b, b, b, b, b],
                           nbr_face=[1, 0, 3, 2, 5, 4], orient=[0] * 6,
                           wall=[0, 0] + [max(0, int(k) - 1) for k in t.faces[2:]]))
    n = 6
    print(f"j- and k-walled duct in 2 blocks along i, {t.nx}x{t.ny}x{t.nz}, wall2={wall2}, {n} steps", flush=True)
    from musica2_newgrid_amd.solver import GpuSolver
    with GpuSolver(t, topology=blocks, block_cases=bcs) as s:
        s.step(n)
        Ug = s.get_states()
    Uc = [bc.copy_state() for bc in bcs]
    orc.step_multi(t, blocks, bcs, Uc, n)
    e = max(relerr(a, b) for a, b in zip(Ug, Uc))
    print(f"relerr {e:.3e}", flush=True)
    assert e < 1e-12


def test_walls_on_the_i_and_k_faces_match_oracle():
    c = cases.duct(36, 32, 12, filter=2)
    c.ducros = True
    t = _swap_jk(c)
    n = 10
    print(f"duct with walls on the i- and k-faces {t.nx}x{t.ny}x{t.nz}, {n} steps", flush=True)
    Ug = _gpu(t, n)
    Uo = _oracle(t, n)
    e = relerr(Ug, Uo)
    print(f"relerr {e:.3e}", flush=True)
    assert e < 1e-12


def test_duct_graph_equals_eager_and_mixed_close_to_fp64():
    c = cases.duct(48, 40, 16, filter=2)
    c.ducros = True
    print("duct 48x40x16: graph vs eager (8 steps), mixed vs FP64 (40 steps)", flush=True)
    Ua = _gpu(c, 8)
    Ub = _gpu(c, 8, eager=True)
    for a, b in zip(Ua, Ub):
        assert np.array_equal(gr.interior(a), gr.interior(b))
    Um = _gpu(c, 40, precision=_lib.PREC_MIXED)
    Ud = _gpu(c, 40)
    a, b = gr.interior(Um[0]), gr.interior(Ud[0])
    e = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
    print(f"mixed vs FP64 rho Linf {e:.3e}", flush=True)
    assert np.all(np.isfinite(a)) and e < 1e-6
