"""Characteristic outflow (imax, jmax) and the supersonic inlet (imin) on the
GPU against the oracle, through the C-ABI: bc_charac_imax, bc_charac_jmax and
bc_supersonic_inlet of src/Boundary/mod_charac.f90 (350-653, 816-969,
997-1149) with the operators' closures on those faces, FP64 within 1e-12
relative.

Covered:
* the one-stage cases whose closed forms the oracle is pinned on
  (tests/test_charac_oracle.py): a supersonic imax outflow with an inlet, a
  fluid at rest relaxing through a subsonic imax outflow between a wall and a
  jmax outflow, jmax outflow and inflow; their ghosts beyond the boundary
  faces hold NaN, which nothing may read;
* uniform states kept over RK4 steps with the viscous terms and the shock
  filter;
* the configuration-4 workload (Mach 2 oblique shock on a boundary layer) at
  a reduced size with the shock filter (with and without the Edoh switch), the
  selective filter, no filter, and without viscous terms;
* steps continuing across calls and graph replays bit for bit; an upload of
  the state restarting the run as start_runge_kutta does (the oracle's single
  call from that state);
* the mixed path close to FP64.
"""
import numpy as np
import pytest

from musica2_newgrid_amd import _lib, cases
from musica2_newgrid_amd import grid as gr
from tests._util import relerr
from tests.test_charac_oracle import P0, RHO0, _c, _case

pytestmark = pytest.mark.gpu

G = gr.G


def _oracle(c, n, U=None):
    from oracle.oracle import Oracle
    U = c.copy_state() if U is None else [a.copy() for a in U]
    Oracle(c).step(U, n)
    return U


def _gpu(c, n, eager=False, precision=_lib.PREC_FP64, chunks=None):
    from musica2_newgrid_amd.solver import GpuSolver
    with GpuSolver(c, eager=eager, precision=precision) as s:
        for m in (chunks or [n]):
            s.step(m)
        return s.get_state()


def _err(A, B, vel):
    """max |A - B| over the interior per variable, relative to |rho|, rho
    (|u| + vel) for the momentum (vel: a velocity scale, so that a fluid at rest
    has one) and |rho E| of B."""
    Bi = [gr.interior(b) for b in B]
    rho = float(np.max(np.abs(Bi[0])))
    mom = float(np.max(np.sqrt(Bi[1] ** 2 + Bi[2] ** 2 + Bi[3] ** 2))) + rho * vel
    scale = [rho, mom, mom, mom, float(np.max(np.abs(Bi[4])))]
    return max(float(np.max(np.abs(gr.interior(a) - b))) / s for a, b, s in zip(A, Bi, scale))


def _one_stage(which):
    h = (1.0e-3, 1.0e-3, 1.0e-3)
    c0 = _c(RHO0, P0)
    if which == "imax_supersonic":
        b, s = 0.004 * RHO0 / h[0], 0.006 * P0 / h[0]
        return _case((24, 10, 5), h, (4, 3, 0, 0, 0, 0),
                     lambda X, Y, Z: (RHO0 + b * X, 2.5 * c0 + 0 * X, 0 * X, 0 * X, P0 + s * X))
    if which == "imax_rest":
        return _case((16, 14, 5), h, (4, 3, 2, 3, 0, 0),
                     lambda X, Y, Z: (RHO0 + 0 * X, 0 * X, 0 * X, 0 * X, P0 + 0 * X), p_ref=1.4 * P0)
    b, s = 0.003 * RHO0 / h[1], 0.005 * P0 / h[1]
    v0 = (0.4 if which == "jmax_out" else -0.4) * c0
    return _case((12, 24, 5), h, (0, 0, 2, 3, 0, 0),
                 lambda X, Y, Z: (RHO0 + b * Y, 0 * Y, v0 + 0 * Y, 0 * Y, P0 + s * Y))


@pytest.mark.parametrize("which", ["imax_supersonic", "imax_rest", "jmax_out", "jmax_in"])
def test_one_stage_cases_match_oracle(which):
    c = _one_stage(which)
    print(f"one-stage characteristic case {which} {c.nx}x{c.ny}x{c.nz}, 1 and 3 steps", flush=True)
    c0 = _c(RHO0, P0)
    for n in (1, 3):
        Ug = _gpu(c, n)
        Uo = _oracle(c, n)
        assert all(np.all(np.isfinite(gr.interior(a))) for a in Ug)
        e = _err(Ug, Uo, c0)
        print(f"{n} steps: relerr {e:.3e}", flush=True)
        assert e < 1e-12


@pytest.mark.parametrize("which", ["rest", "supersonic"])
def test_uniform_states_match_oracle(which):
    """RK4 with the viscous terms and the shock filter (Ducros): a fluid at
    rest between a wall and characteristic faces (p_ref = p/0.528) and a
    uniform Mach 2.5 stream from the inlet to the outflow stay uniform, as in
    the oracle."""
    h = (1.0e-3, 1.0e-3, 1.0e-3)
    c0 = _c(RHO0, P0)
    faces, u0 = ((4, 3, 2, 3, 0, 0), 0.0) if which == "rest" else ((4, 3, 0, 0, 0, 0), 2.5 * c0)
    c = _case((20, 14, 5), h, faces, lambda X, Y, Z: (RHO0 + 0 * X, u0 + 0 * X, 0 * X, 0 * X, P0 + 0 * X),
              visc=4, filt=2, nrk=4, crk=cases.CRK4, p_ref=P0 / 0.528, poison=False)
    print(f"uniform {which}: 12 RK4 steps", flush=True)
    Ug = _gpu(c, 12)
    Uo = _oracle(c, 12)
    e = _err(Ug, Uo, c0)
    d = _err(Ug, c.U, c0)
    print(f"relerr vs oracle {e:.3e}, change from the initial state {d:.3e}", flush=True)
    assert e < 1e-12
    assert d < 1e-12


@pytest.mark.parametrize("filt,visc,edoh", [(2, 4, False), (2, 4, True), (1, 4, False), (0, 4, False),
                                            (1, 0, False)])
def test_shock_boundary_layer_matches_oracle(filt, visc, edoh):
    c = cases.shock_bl(48, 40, 8, filter=filt, iorder_visc=visc)
    c.sw_edoh = edoh
    n = 12
    print(f"shock/boundary layer 48x40x8 filter={filt} visc={visc} edoh={edoh}, {n} steps", flush=True)
    Ug = _gpu(c, n)
    Uo = _oracle(c, n)
    e = relerr(Ug, Uo)
    print(f"relerr {e:.3e}", flush=True)
    assert e < 1e-12
    # the inlet row holds the inflow state exactly
    for v in range(5):
        assert np.array_equal(Ug[v][G:-G, G:-G, G], c.U[v][G:-G, G:-G, G])


def test_shock_boundary_layer_steps_continue_and_restart():
    """Steps split over calls and eager launches equal one graph-replayed call
    bit for bit (the characteristic increments, the sound speed and the filter
    increment of the inlet row persist on the device); uploading a state
    restarts the run from it, as the oracle's call from that state does."""
    from musica2_newgrid_amd.solver import GpuSolver
    c = cases.shock_bl(40, 32, 8)
    print("shock/boundary layer 40x32x8: 4+6 steps vs 10, eager vs graph, restart", flush=True)
    U10 = _gpu(c, 10)
    Us = _gpu(c, 10, chunks=[4, 6])
    Ue = _gpu(c, 10, eager=True)
    for a, b, e in zip(U10, Us, Ue):
        assert np.array_equal(gr.interior(a), gr.interior(b))
        assert np.array_equal(gr.interior(a), gr.interior(e))
    with GpuSolver(c) as s:
        s.step(5)
        U5 = s.get_state()
        s.put_state(U5)
        s.step(5)
        Ur = s.get_state()
    # an upload refreshes the ghosts from the neighbours (here the periodic k
    # images), as the reference's start-up exchange does; after a step the
    # k-ghosts of the inlet row hold the state from before the inflow reset
    U5r = [a.copy() for a in U5]
    for a in U5r:
        gr.fill_periodic(a, axes=(0,))
    Uo = _oracle(c, 5, U=U5r)
    e = relerr(Ur, Uo)
    print(f"restart: relerr {e:.3e}", flush=True)
    assert e < 1e-12


def test_shock_boundary_layer_mixed_close_to_fp64():
    c = cases.shock_bl(64, 48, 8)
    n = 20
    print(f"shock/boundary layer 64x48x8 mixed vs FP64, {n} steps", flush=True)
    Um = _gpu(c, n, precision=_lib.PREC_MIXED)
    Ud = _gpu(c, n)
    e = relerr(Um, Ud)
    print(f"mixed vs FP64 relerr {e:.3e}", flush=True)
    assert all(np.all(np.isfinite(gr.interior(a))) for a in Um)
    assert e < 1e-4
    for v in range(5):
        assert np.array_equal(Um[v][G:-G, G:-G, G], c.U[v][G:-G, G:-G, G])
