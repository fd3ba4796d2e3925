"""Advanced walls (type_wall_bc = 2, mgpu_case.is_wall2) on the GPU against the
oracle, through the C-ABI: the wall rows advanced with the one-sided a02 / a20
derivative across the wall, and the wall states of smod_bc_wall_2*.f90 --
slip walls (MGPU_FACE_SLIP of an adiabatic case, every wall of an inviscid
one), advanced no-slip adiabatic and isothermal walls.  FP64 within 1e-12
relative after 20 steps on a Cartesian channel (every filter), a duct walled
on two axes (slip j-faces meeting no-slip i-faces at the edges, also stored
relabeled), the curvilinear headline channel with slip walls (every filter,
graph, tuned, mixed against FP64), the closed-form linear state of the
oracle pin and the skewed c3 slip wall."""
import numpy as np
import pytest

from musica2_newgrid_amd import _lib, cases
from musica2_newgrid_amd import grid as gr
from tests._util import relerr

pytestmark = pytest.mark.gpu

G = gr.G


def _oracle(c, n):
    from oracle.oracle import Oracle
    U = c.copy_state()
    Oracle(c).step(U, n)
    return U


def _gpu(c, n, eager=False, precision=_lib.PREC_FP64, tune=False):
    from musica2_newgrid_amd.solver import GpuSolver
    with GpuSolver(c, eager=eager, precision=precision) as s:
        if tune:
            s.autotune()
        s.step(n)
        return s.get_state()


@pytest.mark.parametrize("filt,adiabatic,visc,ducros", [(2, True, 4, True), (2, True, 4, False), (1, True, 4, False),
                                                        (0, True, 4, False), (2, False, 4, True),
                                                        (2, False, 4, False), (4, False, 4, True),
                                                        (4, True, 4, True), (1, True, 0, False),
                                                        (0, True, 0, False)])
def test_slip_channel_matches_oracle(filt, adiabatic, visc, ducros):
    """Slip walls (adiabatic, or any inviscid case) or advanced no-slip walls
    (isothermal); the shock filter without the Ducros sensor weights the wall
    rows' dissipation fully, which reads the pressure of the wall rows' face
    ghosts (their own state, from before the wall state was imposed)."""
    c = cases.channel(32, 24, 10, filter=filt, adiabatic=adiabatic, iorder_visc=visc, slip=True)
    c.ducros = ducros
    n = 20
    print(f"advanced-wall channel 32x24x10 filter={filt} adiabatic={adiabatic} visc={visc} ducros={ducros}, "
          f"{n} steps", flush=True)
    Ug = _gpu(c, n)
    Uo = _oracle(c, n)
    e = relerr(Ug, Uo)
    print(f"relerr {e:.3e}", flush=True)
    assert e < 1e-12
    wall = Ug[2][G:-G, G, G:-G]
    assert np.all(wall == 0.0)
    # slip (adiabatic or inviscid) keeps the tangential momentum on the wall row
    slip = adiabatic or visc == 0
    assert (float(np.max(np.abs(Ug[1][G:-G, G, G:-G]))) > 0.0) == slip


def test_linear_state_closed_form_on_the_gpu():
    from tests.test_wall2_oracle import _linear_channel
    c, _, _ = _linear_channel()
    print("linear-flux channel with slip walls, one stage", flush=True)
    Ug = _gpu(c, 1)
    Uo = _oracle(c, 1)
    for a, b in zip(Ug, Uo):
        assert np.allclose(gr.interior(a), gr.interior(b), rtol=1e-13, atol=0.0)


def _duct2():
    """The duct with advanced walls: no slip on the i-faces, slip on the
    j-faces (an adiabatic case), so the edges take both in bc_wall's order."""
    c = cases.duct(40, 36, 12, filter=2)
    c.ducros = True
    c.wall2 = True
    c.faces = (2, 2, 5, 5, 0, 0)
    return c


def test_duct_with_slip_and_no_slip_advanced_walls_matches_oracle():
    c = _duct2()
    n = 20
    print(f"advanced-wall duct 40x36x12 (no-slip i, slip j), {n} steps", flush=True)
    Ug = _gpu(c, n)
    Uo = _oracle(c, n)
    e = relerr(Ug, Uo)
    print(f"relerr {e:.3e}", flush=True)
    assert e < 1e-12
    # edges: the i-face's no slip, then the j-face's slip (nothing left to remove)
    for v in (1, 2, 3):
        assert np.all(Ug[v][G:-G, G:-G, G] == 0.0) and np.all(Ug[v][G:-G, G:-G, G + c.nx - 1] == 0.0)
    assert np.all(Ug[2][G:-G, G, G:-G] == 0.0)


def test_relabeled_advanced_walls_match_oracle():
    """Slip walls on the i- and k-faces (stored with j and k exchanged)."""
    from tests.test_gpu_walls_two_axes import _swap_jk
    c = _duct2()
    t = _swap_jk(c)
    t.wall2 = True
    t.faces = (2, 2, 0, 0, 5, 5)
    n = 10
    print(f"advanced walls on the i- and k-faces {t.nx}x{t.ny}x{t.nz}, {n} steps", flush=True)
    Ug = _gpu(t, n)
    Uo = _oracle(t, n)
    e = relerr(Ug, Uo)
    print(f"relerr {e:.3e}", flush=True)
    assert e < 1e-12


def test_advanced_walls_on_the_j_and_k_faces_match_oracle():
    """No slip on the j-faces, slip on the k-faces (the block stored with its
    j, k, i as i, j, k): the edges take the k-faces' slip last."""
    from tests.test_duct_oracle import cycle_axes
    t = cycle_axes(_duct2())
    n = 10
    print(f"advanced walls on the j- and k-faces {t.nx}x{t.ny}x{t.nz}, {n} steps", flush=True)
    Ug = _gpu(t, n)
    Uo = _oracle(t, n)
    e = relerr(Ug, Uo)
    print(f"relerr {e:.3e}", flush=True)
    assert e < 1e-12


@pytest.mark.parametrize("filt", [2, 1, 0])
def test_slip_curvilinear_channel_matches_oracle(filt):
    c = cases.vortex_channel(48, 40, 12, filter=filt, slip=True)
    n = 20
    print(f"c3 channel with slip walls 48x40x12 filter={filt}, {n} steps", flush=True)
    Ug = _gpu(c, n)
    Uo = _oracle(c, n)
    e = relerr(Ug, Uo)
    print(f"relerr {e:.3e}", flush=True)
    assert e < 1e-12


def test_slip_curvilinear_channel_graph_tuned_and_mixed():
    c = cases.vortex_channel(40, 32, 12, slip=True)
    print("c3 slip channel 40x32x12: graph vs eager, tuned vs oracle (8 steps), mixed vs FP64 (40)", flush=True)
    Ua = _gpu(c, 8)
    Ub = _gpu(c, 8, eager=True)
    for a, b in zip(Ua, Ub):
        assert np.array_equal(gr.interior(a), gr.interior(b))
    Ut = _gpu(c, 8, tune=True)
    e = relerr(Ut, _oracle(c, 8))
    print(f"tuned relerr {e:.3e}", flush=True)
    assert e < 1e-12
    Um = _gpu(c, 40, precision=_lib.PREC_MIXED)
    Ud = _gpu(c, 40)
    a, b = gr.interior(Um[0]), gr.interior(Ud[0])
    e = float(np.max(np.abs(a - b)) / np.max(np.abs(b)))
    print(f"mixed vs FP64 rho Linf {e:.3e}", flush=True)
    assert np.all(np.isfinite(a)) and e < 1e-6


def test_skewed_c3_slip_wall_on_the_gpu():
    """The oracle pin's skewed planar walls with dt = 0: the end-of-step slip
    state alone, GPU against oracle."""
    from tests.test_wall2_oracle import _fill, _skewed
    nx, ny, nz = 12, 12, 10
    m, mv, wall, _ = _skewed(0.7, -0.9)
    rng = np.random.default_rng(3)
    rho = 1.2 + 0.1 * rng.random((nz, ny, nx))
    u, v, w = (30.0 * rng.standard_normal((nz, ny, nx)) for _ in range(3))
    p = 1.0e5 * (1.0 + 0.01 * rng.random((nz, ny, nx)))
    U = _fill(cases.conserved(rho, u, v, w, p), nx, ny, nz, (0, 2))
    c = cases.Case("skew", nx, ny, nz, curvilinear=True, iorder_visc=0, filter=0, nrk=1, crk=(1.0,), m=m, mv=mv,
                   U=U, faces=(0, 0, 5, 5, 0, 0), wall2=True)
    c.wall_c3 = wall
    c.dt = 0.0
    print("skewed c3 slip walls, dt = 0, one step", flush=True)
    Ug = _gpu(c, 1)
    Uo = _oracle(c, 1)
    for a, b in zip(Ug, Uo):
        assert np.allclose(gr.interior(a), gr.interior(b), rtol=1e-14, atol=1e-12)
