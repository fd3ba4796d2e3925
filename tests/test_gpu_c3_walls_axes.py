"""Walls on the i- or k-faces of a 3-D curvilinear block on the GPU, through
the C-ABI: the block is stored with that index axis exchanged with j (its
state arrays, metric rows and wall coefficients with it; momentum, tau and
the metric columns stay in the caller's x, y, z).

The reference's c3 wall routines of the i- and k-faces are those of the
j-faces with the index axes exchanged: the wall state extrapolates along the
face normal with the tangential terms of the face's two tangential axes in
order (smod_bc_wall_c3.f90:10-233 imin against :842-1065 jmin), and the
adiabatic wall's temperature gradient takes the same two axes with their unit
tangents (smod_grad_T_c3.f90:220-229 imin, :248-257 jmin, :276-285 kmin).  So
the curvilinear channel turned -- two index axes exchanged and the two
physical axes with them -- is the same flow turned: the GPU on the turned
channel against the oracle on the channel, turned back, FP64 within 1e-12
relative (every filter, slip and no-slip walls), and the mixed path against
FP64 on the turned channel; graph replay and autotuned tiles; the turned
channel in two blocks joined along k against the oracle's two blocks."""
import numpy as np
import pytest

from musica2_newgrid_amd import _lib, cases
from musica2_newgrid_amd import grid as gr
from tests._util import relerr

pytestmark = pytest.mark.gpu

NP_AXIS = (2, 1, 0)  # numpy axis of index axis i, j, k


def _turn(c, a, b):
    """Case c with index axes a, b exchanged and physical axes a, b with them
    (the walls of the j-faces land on the faces of the other axis)."""
    sw = [0, 1, 2]
    sw[a], sw[b] = b, a

    def arr(x):
        return np.ascontiguousarray(np.swapaxes(x, NP_AXIS[a], NP_AXIS[b]))

    def met(m):
        out = [None] * 10
        for r in range(3):
            for q in range(3):
                out[3 * sw[r] + sw[q]] = arr(m[3 * r + q])
        out[9] = arr(m[9])
        return out

    n = [c.nx, c.ny, c.nz]
    n[a], n[b] = n[b], n[a]
    fperm = [2 * sw[f // 2] + f % 2 for f in range(6)]  # face f of c is face fperm[f] of t
    faces = [0] * 6
    for f in range(6):
        faces[fperm[f]] = c.faces[f]
    t = cases.Case(c.name + "_turned", *n, curvilinear=True, iorder_visc=c.iorder_visc, filter=c.filter,
                   ducros=c.ducros, diffscale=c.diffscale, chi=c.chi, dt=c.dt, nrk=c.nrk, crk=c.crk,
                   m=met(c.m), mv=met(c.mv) if c.mv is not None else None, faces=tuple(faces),
                   adiabatic=c.adiabatic, t_wall=c.t_wall, wall2=c.wall2)
    U = [arr(x) for x in c.U]
    mom = [U[1], U[2], U[3]]
    t.U = [U[0]] + [mom[sw[d]] for d in range(3)] + [U[4]]
    wall = {}
    for f, slabs in c.wall_c3.items():
        s = [np.array(x, copy=True) for x in slabs]
        # the unit and dual tangents are physical vectors: their components move
        for base in (2, 5, 8, 11):
            if base + 2 < len(s):
                comp = s[base:base + 3]
                s[base:base + 3] = [comp[sw[d]] for d in range(3)]
        wall[fperm[f]] = s
    t.wall_c3 = wall
    return t


def _unturn(U, a, b):
    """A state of the turned case back in the frame of the original."""
    sw = [0, 1, 2]
    sw[a], sw[b] = b, a
    V = [np.swapaxes(x, NP_AXIS[a], NP_AXIS[b]) for x in U]
    mom = [V[1], V[2], V[3]]
    return [V[0]] + [mom[sw[d]] for d in range(3)] + [V[4]]


def _gpu(c, n, precision=_lib.PREC_FP64):
    from musica2_newgrid_amd.solver import GpuSolver
    with GpuSolver(c, precision=precision) as s:
        s.step(n)
        return s.get_state()


def _oracle(c, n):
    from oracle.oracle import Oracle
    U = c.copy_state()
    Oracle(c).step(U, n)
    return U


@pytest.mark.parametrize("axes,filt,slip", [((0, 1), 2, True), ((0, 1), 2, False), ((1, 2), 2, True),
                                            ((1, 2), 1, False), ((0, 1), 0, True)])
def test_turned_curvilinear_channel_matches_oracle(axes, filt, slip):
    c = cases.vortex_channel(40, 32, 12, filter=filt, slip=slip)
    t = _turn(c, *axes)
    n = 12
    wf = [f for f in range(6) if t.faces[f] in (2, 5)]
    print(f"c3 channel turned {axes} ({t.nx}x{t.ny}x{t.nz}, walls on faces {wf}) filter={filt} slip={slip}, "
          f"{n} steps", flush=True)
    Ug = _unturn(_gpu(t, n), *axes)
    Uo = _oracle(c, n)
    e = relerr(Ug, Uo)
    print(f"relerr {e:.3e}", flush=True)
    assert e < 1e-12


def test_turned_curvilinear_channel_mixed():
    c = cases.vortex_channel(40, 32, 12, slip=True)
    t = _turn(c, 1, 2)
    print(f"c3 channel with walls on the k-faces {t.nx}x{t.ny}x{t.nz}: mixed vs FP64, 30 steps", flush=True)
    Um = _gpu(t, 30, precision=_lib.PREC_MIXED)
    Ud = _gpu(t, 30)
    a, b = gr.interior(Um[0]), gr.interior(Ud[0])
    e = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
    print(f"mixed vs FP64 rho Linf {e:.3e}", flush=True)
    assert np.all(np.isfinite(a)) and e < 1e-6


def test_turned_curvilinear_channel_graph_and_tuned():
    """Walls on the i-faces: graph replay equal to eager launches bit for bit,
    autotuned tiles against the oracle."""
    from musica2_newgrid_amd.solver import GpuSolver
    c = cases.vortex_channel(40, 32, 12, slip=True)
    t = _turn(c, 0, 1)
    n = 8
    print(f"c3 channel with walls on the i-faces {t.nx}x{t.ny}x{t.nz}: graph vs eager, tuned vs oracle, {n} steps",
          flush=True)
    out = []
    for eager, tune in ((False, False), (True, False), (False, True)):
        with GpuSolver(t, eager=eager) as s:
            if tune:
                s.autotune()
            s.step(n)
            out.append(s.get_state())
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(gr.interior(a), gr.interior(b))
    e = relerr(_unturn(out[2], 0, 1), _oracle(c, n))
    print(f"tuned relerr {e:.3e}", flush=True)
    assert e < 1e-12


def _halves(c):
    """c cut in two blocks along k (its periodic k-faces become interfaces)."""
    G = gr.G
    h = c.nz // 2
    subs, blocks = [], []
    for b in range(2):
        lo, hi = b * h, b * h + h + 2 * G
        sub = cases.Case(f"{c.name}_k{b}", c.nx, c.ny, h, curvilinear=True, iorder_visc=c.iorder_visc,
                         filter=c.filter, ducros=c.ducros, diffscale=c.diffscale, chi=c.chi, dt=c.dt, nrk=c.nrk,
                         crk=c.crk, m=[np.ascontiguousarray(x[lo:hi]) for x in c.m],
                         mv=[np.ascontiguousarray(x[lo:hi]) for x in c.mv] if c.mv is not None else None,
                         faces=tuple(c.faces[:4]) + (1, 1), adiabatic=c.adiabatic, t_wall=c.t_wall, wall2=c.wall2)
        sub.U = [np.ascontiguousarray(x[lo:hi]) for x in c.U]
        sub.wall_c3 = {f: [np.ascontiguousarray(x[b * h:(b + 1) * h]) for x in s] for f, s in c.wall_c3.items()}
        subs.append(sub)
        kinds = [int(k) for k in c.faces[:4]] + [_lib.FACE_INTERFACE] * 2
        blocks.append(dict(nx=c.nx, ny=c.ny, nz=h, kind=kinds, nbr_block=[b, b, b, b, 1 - b, 1 - b],
                           nbr_face=[1, 0, 3, 2, 5, 4], orient=[0] * 6,
                           wall=[max(0, k - 1) for k in kinds[:4]] + [0, 0]))
    return subs, blocks


def test_turned_curvilinear_channel_in_two_blocks():
    """Walls on the i-faces of two blocks joined along k, against the oracle's
    two blocks of the channel, turned back."""
    from musica2_newgrid_amd.solver import GpuSolver
    from oracle import oracle as orc
    c = cases.vortex_channel(40, 32, 12, slip=True)
    subs, blocks = _halves(c)
    tsubs = [_turn(s, 0, 1) for s in subs]
    tblocks = []
    for blk, ts in zip(blocks, tsubs):
        kinds = list(blk["kind"])
        tblocks.append(dict(nx=ts.nx, ny=ts.ny, nz=ts.nz, kind=[kinds[2], kinds[3], kinds[0], kinds[1]] + kinds[4:],
                            nbr_block=blk["nbr_block"], nbr_face=blk["nbr_face"], orient=blk["orient"]))
    t = _turn(c, 0, 1)
    n = 8
    print(f"c3 channel with walls on the i-faces in 2 blocks along k ({t.nx}x{t.ny}x{t.nz}), {n} steps", flush=True)
    with GpuSolver(t, topology=tblocks, block_cases=tsubs) as s:
        s.step(n)
        Ug = [_unturn(U, 0, 1) for U in s.get_states()]
    Uo = [bc.copy_state() for bc in subs]
    orc.step_multi(c, blocks, subs, Uo, n)
    e = max(relerr(a, b) for a, b in zip(Ug, Uo))
    print(f"relerr {e:.3e}", flush=True)
    assert e < 1e-12
