"""Walls on the i- and k-faces on the GPU (the block is stored with that axis
exchanged with j, invisibly to the caller), against the oracle, whose walls
work on any axis: the channel turned so that its walls are x-faces or
z-faces, with the shock filter and Ducros sensor (FP64 within 1e-12, mixed
density error), the GPU result of the turned channel equal to the turned
GPU result of the original (round-off), and a z-walled channel cut in two
blocks along x (interfaces on faces whose tangential axes the relabeling
reverses).  A block walled on axes its neighbour is not walled on is refused
at create (tests/test_topology.py)."""
import numpy as np
import pytest

from musica2_newgrid_amd import _lib, cases
from musica2_newgrid_amd import grid as gr
from tests._util import relerr
from tests.test_wall_oracle import _transpose_case

pytestmark = pytest.mark.gpu
G = gr.G


def _swap_yz(c):
    """The same channel with y and z exchanged: walls on the z-faces."""
    t = cases.Case(c.name + "_z", c.nx, c.nz, c.ny, curvilinear=False, iorder_visc=c.iorder_visc,
                   filter=c.filter, ducros=c.ducros, diffscale=c.diffscale, chi=c.chi, dt=c.dt, nrk=c.nrk,
                   crk=c.crk, cart=[c.cart[0], c.cart[2], c.cart[1], c.cart[3], c.cart[5], c.cart[4]],
                   faces=(0, 0, 0, 0, 2, 2), adiabatic=c.adiabatic, t_wall=c.t_wall,
                   cextrp2=(0, 0, 0, 0, c.cextrp2[2], c.cextrp2[3]),
                   cextrp3=(0, 0, 0, 0, c.cextrp3[2], c.cextrp3[3]))
    U = [np.ascontiguousarray(np.swapaxes(a, 0, 1)) for a in c.U]
    U[2], U[3] = U[3], U[2]
    t.U = U
    return t


def _channel():
    c = cases.channel(40, 32, 12, filter=2)
    c.ducros = True
    return c


def _turned(which):
    c = _channel()
    if which == "x":
        t = _transpose_case(c)
        t.ducros = True
        return t
    return _swap_yz(c)


def _gpu(c, nsteps, precision=_lib.PREC_FP64, topology=None, block_cases=None):
    from musica2_newgrid_amd.solver import GpuSolver
    with GpuSolver(c, precision=precision, topology=topology, block_cases=block_cases) as s:
        s.step(nsteps)
        return s.get_states() if topology else s.get_state()


def _oracle(c, nsteps):
    from oracle.oracle import Oracle
    U = c.copy_state()
    Oracle(c).step(U, nsteps)
    return U


@pytest.mark.parametrize("which", ["x", "z"])
def test_turned_channel_matches_oracle(which):
    t = _turned(which)
    print(f"channel with walls on the {which}-faces, {t.nx}x{t.ny}x{t.nz}, 10 steps", flush=True)
    Ug, Uc = _gpu(t, 10), _oracle(t, 10)
    e = relerr(Ug, Uc)
    print(f"relerr {e:.3e} (moved {relerr(Uc, t.U):.3e})", flush=True)
    assert e < 1e-12


@pytest.mark.parametrize("which", ["x", "z"])
def test_turned_channel_is_the_turned_result(which):
    c = _channel()
    A = _gpu(c, 5)
    B = _gpu(_turned(which), 5)
    if which == "x":
        B = [np.swapaxes(b, 1, 2) for b in B]
        B[1], B[2] = B[2], B[1]
    else:
        B = [np.swapaxes(b, 0, 1) for b in B]
        B[2], B[3] = B[3], B[2]
    e = max(float(np.max(np.abs(gr.interior(a) - gr.interior(b)))) / float(np.max(np.abs(gr.interior(a))))
            for a, b in zip(A, B))
    print(f"walls on {which} vs y, turned back: {e:.3e}", flush=True)
    assert e < 1e-12


@pytest.mark.parametrize("which", ["x", "z"])
def test_turned_channel_mixed_precision(which):
    t = _turned(which)
    Ug = _gpu(t, 100, precision=_lib.PREC_MIXED)
    Uc = _oracle(t, 100)
    a, b = gr.interior(Ug[0]), gr.interior(Uc[0])
    linf = float(np.max(np.abs(a - b))) / float(np.max(np.abs(b)))
    print(f"walls on {which}, mixed, 100 steps: rho Linf {linf:.3e} vs the FP64 oracle", flush=True)
    assert 0.0 < linf <= 1e-6


def test_z_walled_channel_in_two_blocks():
    from oracle import oracle as orc
    t = _turned("z")
    h = t.nx // 2
    blocks, bcs = [], []
    for b in range(2):
        sub = cases.Case(f"half{b}", h, t.ny, t.nz, curvilinear=False, iorder_visc=t.iorder_visc,
                         filter=t.filter, ducros=t.ducros, diffscale=t.diffscale, chi=t.chi, dt=t.dt,
                         cart=[t.cart[0][:h + 2 * G].copy(), t.cart[1], t.cart[2],
                               t.cart[3][:h + 2 * G].copy(), t.cart[4], t.cart[5]],
                         faces=(1, 1, 0, 0, 2, 2), adiabatic=t.adiabatic, t_wall=t.t_wall,
                         cextrp2=t.cextrp2, cextrp3=t.cextrp3)
        sub.U = [np.ascontiguousarray(a[:, :, b * h:b * h + h + 2 * G]) for a in t.U]
        bcs.append(sub)
        blocks.append(dict(nx=h, ny=t.ny, nz=t.nz, kind=[_lib.FACE_INTERFACE] * 4 + [_lib.FACE_WALL] * 2,
                           nbr_block=[1 - b, 1 - b, b, b, b, b], nbr_face=[1, 0, 3, 2, 5, 4], orient=[0] * 6,
                           wall=[0, 0, 0, 0, 1, 1]))
    print(f"z-walled channel in 2 blocks along x, {t.nx}x{t.ny}x{t.nz}, 5 steps", flush=True)
    Ug = _gpu(t, 5, topology=blocks, block_cases=bcs)
    Uc = [bc.copy_state() for bc in bcs]
    orc.step_multi(t, blocks, bcs, Uc, 5)
    e = max(relerr(a, b) for a, b in zip(Ug, Uc))
    print(f"relerr {e:.3e}", flush=True)
    assert e < 1e-12



def _cut_mixed(t, axis, wall_axis):
    """Two blocks cut from case t along `axis` (0 i, 1 j): block 0 keeps the
    walls of t on `wall_axis`, block 1 has periodic faces there, so the two
    blocks are stored with different axis orders and every interface joins
    a relabeled block to one stored as given."""
    n = (t.nx, t.ny, t.nz)
    h = n[axis] // 2
    blocks, bcs = [], []
    for b in range(2):
        ext = list(n)
        ext[axis] = h
        sl = [slice(None)] * 3
        sl[2 - axis] = slice(b * h, b * h + h + 2 * G)
        cart = list(t.cart)
        for q in (axis, axis + 3):
            cart[q] = t.cart[q][b * h:b * h + h + 2 * G].copy()
        faces = [0] * 6
        if b == 0:
            faces[2 * wall_axis] = faces[2 * wall_axis + 1] = 2
        sub = cases.Case(f"part{b}", *ext, curvilinear=False, iorder_visc=t.iorder_visc, filter=t.filter,
                         ducros=t.ducros, diffscale=t.diffscale, chi=t.chi, dt=t.dt, cart=cart, faces=tuple(faces),
                         adiabatic=t.adiabatic, t_wall=t.t_wall, cextrp2=t.cextrp2, cextrp3=t.cextrp3)
        sub.U = [np.ascontiguousarray(a[tuple(sl)]) for a in t.U]
        bcs.append(sub)
        kind = [_lib.FACE_PERIODIC] * 6
        nbr_block = [b] * 6
        for s in range(2):
            kind[2 * axis + s] = _lib.FACE_INTERFACE
            nbr_block[2 * axis + s] = 1 - b
        if b == 0:
            kind[2 * wall_axis] = kind[2 * wall_axis + 1] = _lib.FACE_WALL
        blocks.append(dict(nx=ext[0], ny=ext[1], nz=ext[2], kind=kind, nbr_block=nbr_block,
                           nbr_face=[1, 0, 3, 2, 5, 4], orient=[0] * 6,
                           wall=[int(k == _lib.FACE_WALL) for k in kind]))
    return blocks, bcs
