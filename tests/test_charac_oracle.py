"""Pins of the oracle's characteristic boundaries (configuration 4):
bc_charac_imax, bc_charac_jmax and bc_supersonic_inlet of
src/Boundary/mod_charac.f90 (350-653, 816-969, 997-1149), dispatched after
`primitives` in every stage (runge_kutta.f90:30-37, 64-70, 137-139), with the
operators closed on those faces like on walls (ndx_e = ndx_v = 2,
mpi_index_bounds.f90:20-130).

Each test takes one step of a one-stage scheme (crk = 1) so that the increment
K of every interior point is (U^0 - U^1)/dt.  The LODI amplitudes are
evaluated with one-sided 5-pt derivatives, exact on the linear fields used
here, so closed forms apply:
* at an imax outflow the velocity and pressure parts are the Euler equations
  in characteristic form, K_u = u u_x + p_x/rho and K_p = u p_x + rho c^2 u_x
  for a supersonic point; at a subsonic one the incoming wave is replaced by
  the relaxation L1 = 0.1 (p - 0.528 p_ref);
* at a jmax outflow with 0 < v < c only the outgoing acoustic wave is kept
  (K_v = (v + c) p_y / (2 rho c), K_p = rho c K_v), the tangential velocity is
  advected (K_u = v u_y);
* the density parts are checked in the cases where their as-written form is
  fixed by the other four (the reference's density equations are not the
  textbook LODI ones: rt = -(pt + L2)/c^2 at imax, (pt + L3)/c^2 at jmax);
* the operators' closures on those faces are exact on linear fluxes and never
  read a ghost beyond the faces (the ghosts hold NaN);
* a fluid at rest (with p_ref = p/0.528) and a uniform supersonic stream stay
  unchanged over many full RK4 steps with viscous terms and the shock filter;
* the inlet row keeps the inflow state exactly.
"""
import math

import numpy as np
import pytest

from musica2_newgrid_amd import cases
from musica2_newgrid_amd import grid as gr

G = gr.G
GAM, RG = cases.GAM, cases.RG
CP = GAM * RG / (GAM - 1.0)


def _axis(n, h, lo, hi):
    xe = np.arange(-G, n + G) * h
    if lo or hi:
        ide, idv, c2, c3 = gr.wall_metrics(xe, lo, hi)
    else:
        ide = idv = np.full(n + 2 * G, 1.0 / h)
        c2 = c3 = {}
    return xe, ide, idv, c2, c3


def _case(n, h, faces, fields, visc=0, filt=0, nrk=1, crk=(1.0,), cfl=0.5, p_ref=0.0, poison=True):
    """A Cartesian block with the given face kinds (0 periodic, 2 wall, 3
    characteristic outflow, 4 supersonic inlet) and primitive fields
    rho, u, v, w, p = fields(X, Y, Z) evaluated on the padded nodes."""
    nx, ny, nz = n
    axes = [_axis(n[a], h[a], faces[2 * a] >= 2, faces[2 * a + 1] >= 2) for a in range(3)]
    X = axes[0][0].reshape(1, 1, -1)
    Y = axes[1][0].reshape(1, -1, 1)
    Z = axes[2][0].reshape(-1, 1, 1)
    shape = (nz + 2 * G, ny + 2 * G, nx + 2 * G)
    rho, u, v, w, p = (np.broadcast_to(q, shape).astype(np.float64) for q in fields(X, Y, Z))
    U = [np.ascontiguousarray(a) for a in cases.conserved(rho, u, v, w, p)]
    for a in U:
        per = tuple(ax for ax, (lo, hi) in zip((2, 1, 0), ((faces[0], faces[1]), (faces[2], faces[3]),
                                                          (faces[4], faces[5]))) if lo == 0 and hi == 0)
        if per:
            gr.fill_periodic(a, axes=per)
        if poison:
            if faces[0] >= 2:
                a[:, :, :G] = np.nan
            if faces[1] >= 2:
                a[:, :, G + nx:] = np.nan
            if faces[2] >= 2:
                a[:, :G, :] = np.nan
            if faces[3] >= 2:
                a[:, G + ny:, :] = np.nan
    ext2, ext3 = [0.0] * 6, [0.0] * 6
    for a in range(3):
        c2, c3 = axes[a][3], axes[a][4]
        for s in (0, 1):
            if faces[2 * a + s] == 2:
                ext2[2 * a + s], ext3[2 * a + s] = c2[s], c3[s]
    c = cases.Case("charac", nx, ny, nz, curvilinear=False, iorder_visc=visc, filter=filt, ducros=True, nrk=nrk,
                   crk=crk, cart=[axes[0][1], axes[1][1], axes[2][1], axes[0][2], axes[1][2], axes[2][2]], U=U,
                   faces=tuple(faces), cextrp2=tuple(ext2), cextrp3=tuple(ext3))
    c.p_ref = p_ref
    interior = [gr.interior(a) for a in U]
    c.dt = cases.timestep([np.pad(a, G) for a in interior], min(h), cfl)
    return c


def _step(c, n=1):
    from oracle.oracle import Oracle
    U = c.copy_state()
    Oracle(c).step(U, n)
    return U


def _K(c, U1):
    return [(a - b) / c.dt for a, b in zip(c.U, U1)]


RHO0, P0, T0 = 1.2, 1.0e5, None


def _c(rho, p):
    return math.sqrt(GAM * p / rho)


def test_imax_supersonic_outflow_and_closures():
    """u constant and supersonic, rho and p linear in x: every flux is linear,
    so every closure row (a3..a11, both ends) gives the exact x-derivative;
    the imax row gets K_u = p_x/rho and K_p = u p_x, and
    K_rho = -(pt + u (p_x - c^2 rho_x))/c^2 as written."""
    nx, ny, nz = 24, 6, 5
    h = (1.0e-3, 2.0e-3, 2.0e-3)
    c0 = _c(RHO0, P0)
    U0 = 2.5 * c0
    b, s = 0.004 * RHO0 / h[0], 0.006 * P0 / h[0]  # rho_x, p_x

    def fields(X, Y, Z):
        return RHO0 + b * X, U0 + 0 * X, 0 * X, 0 * X, P0 + s * X

    c = _case((nx, ny, nz), h, (4, 3, 0, 0, 0, 0), fields)
    U1 = _step(c)
    K = _K(c, U1)
    xs = np.arange(nx) * h[0]
    rho, p = RHO0 + b * xs, P0 + s * xs
    # rows 2..nx-1: the exact divergence of the linear fluxes
    want = [U0 * b + 0 * xs, U0 * U0 * b + s + 0 * xs, 0 * xs, 0 * xs,
            U0 * (GAM / (GAM - 1.0) * s + 0.5 * U0 * U0 * b) + 0 * xs]
    for v in range(5):
        got = K[v][G:-G, G:-G, G + 1:G + nx - 1]
        scale = max(abs(want[1][0]), abs(want[4][0]))
        assert np.all(np.isfinite(got))
        assert np.max(np.abs(got - want[v][1:-1])) <= 1e-6 * scale, v
    # the imax row
    r, pp = rho[-1], p[-1]
    cc = _c(r, pp)
    T = pp / (r * RG)
    Kn = [float(np.mean(K[v][G:-G, G:-G, G + nx - 1])) for v in range(5)]
    for v in range(5):
        assert np.ptp(K[v][G:-G, G:-G, G + nx - 1]) <= 1e-7 * max(1.0, abs(Kn[v]))
    Ku = (Kn[1] - U0 * Kn[0]) / r
    assert Ku == pytest.approx(s / r, rel=1e-6)
    H = (pp / (GAM - 1.0) + 0.5 * r * U0 * U0 + pp) / r
    pt = cc * cc * ((Kn[4] - H * Kn[0] - r * U0 * Ku) / (CP * T) + Kn[0])
    assert pt == pytest.approx(U0 * s, rel=1e-6)
    L2 = U0 * (s - cc * cc * b)
    assert Kn[0] == pytest.approx(-(U0 * s + L2) / (cc * cc), rel=1e-6)
    # the inlet row keeps the inflow state
    for v in range(5):
        assert np.array_equal(U1[v][G:-G, G:-G, G], c.U[v][G:-G, G:-G, G])


def test_imax_subsonic_relaxation_at_rest():
    """Fluid at rest with p != 0.528 p_ref: at the imax row only the relaxed
    incoming wave acts, A = L1 = 0.1 (p - 0.528 p_ref): pt = A/2, ut =
    -A/(2 rho c), rt = -pt/c^2, so K = (-A/(2c^2), -A/(2c), 0, 0, A/(2(gam-1)))
    (the wall row j = 1 of the imax column is the wall's; the jmax row gets
    nothing at rest); the rest of the block is unchanged."""
    nx, ny, nz = 16, 14, 5
    h = (1.0e-3, 1.0e-3, 1.0e-3)
    p_ref = 1.4 * P0

    def fields(X, Y, Z):
        return RHO0 + 0 * X, 0 * X, 0 * X, 0 * X, P0 + 0 * X

    c = _case((nx, ny, nz), h, (4, 3, 2, 3, 0, 0), fields, p_ref=p_ref)
    U1 = _step(c)
    K = _K(c, U1)
    A = 0.1 * (P0 - 0.528 * p_ref)
    cc = _c(RHO0, P0)
    want = [-A / (2 * cc * cc), -A / (2 * cc), 0.0, 0.0, A / (2 * (GAM - 1.0))]
    for v in range(5):
        got = K[v][G:-G, G + 1:G + ny, G + nx - 1]
        assert np.allclose(got, want[v], rtol=1e-6, atol=1e-6 * abs(want[4])), v
        # everything else, jmax row included, is at rest
        rest = K[v][G:-G, G + 1:G + ny, G:G + nx - 1]
        assert np.max(np.abs(rest)) <= 1e-6 * abs(want[4]), v


def test_jmax_subsonic_outflow_and_inflow():
    """v constant (0 < v < c, then -c < v < 0), u = w = 0, rho and p linear in
    y: at the jmax row, outflow: K_u = 0, K_v = (v + c) p_y/(2 rho c), K_p =
    (v + c) p_y / 2, K_rho = (pt + v (p_y - c^2 rho_y))/c^2 as written;
    inflow (v <= 0): L2 = L3 = L4 = 0, so K_rho = pt/c^2 with pt =
    (v + c) p_y / 2."""
    nx, ny, nz = 6, 24, 5
    h = (2.0e-3, 1.0e-3, 2.0e-3)
    b, s = 0.003 * RHO0 / h[1], 0.005 * P0 / h[1]
    for sign in (1.0, -1.0):
        c0 = _c(RHO0, P0)
        V0 = sign * 0.4 * c0

        def fields(X, Y, Z):
            return RHO0 + b * Y, 0 * Y, V0 + 0 * Y, 0 * Y, P0 + s * Y

        c = _case((nx, ny, nz), h, (0, 0, 2, 3, 0, 0), fields)
        U1 = _step(c)
        K = _K(c, U1)
        yn = (ny - 1) * h[1]
        r, pp = RHO0 + b * yn, P0 + s * yn
        cc = _c(r, pp)
        T = pp / (r * RG)
        Kn = [float(np.mean(K[v][G:-G, G + ny - 1, G:-G])) for v in range(5)]
        for v in range(5):
            assert np.ptp(K[v][G:-G, G + ny - 1, G:-G]) <= 1e-7 * max(1.0, abs(Kn[v]))
        L5 = (V0 + cc) * s
        pt_want = 0.5 * L5
        assert abs(Kn[1]) <= 1e-7 * abs(Kn[2])
        Kv = (Kn[2] - V0 * Kn[0]) / r
        assert Kv == pytest.approx(L5 / (2 * r * cc), rel=1e-6)
        H = (pp / (GAM - 1.0) + 0.5 * r * V0 * V0 + pp) / r
        pt = cc * cc * ((Kn[4] - H * Kn[0] - r * V0 * Kv) / (CP * T) + Kn[0])
        assert pt == pytest.approx(pt_want, rel=1e-6)
        L3 = V0 * (s - cc * cc * b) if V0 > 0 else 0.0
        assert Kn[0] == pytest.approx((pt_want + L3) / (cc * cc), rel=1e-6)


@pytest.mark.parametrize("which", ["rest", "supersonic"])
def test_uniform_states_are_kept(which):
    """Full RK4 steps with viscous terms and the shock filter (Ducros): a
    fluid at rest between a wall and the characteristic faces (p_ref =
    p/0.528, so the relaxation vanishes), or a uniform Mach 2.5 stream from the
    inlet to the outflow, stays as it is."""
    nx, ny, nz = 20, 14, 5
    h = (1.0e-3, 1.0e-3, 1.0e-3)
    c0 = _c(RHO0, P0)
    if which == "rest":
        faces, U0 = (4, 3, 2, 3, 0, 0), 0.0
    else:
        faces, U0 = (4, 3, 0, 0, 0, 0), 2.5 * c0

    def fields(X, Y, Z):
        return RHO0 + 0 * X, U0 + 0 * X, 0 * X, 0 * X, P0 + 0 * X

    c = _case((nx, ny, nz), h, faces, fields, visc=4, filt=2, nrk=4, crk=cases.CRK4, p_ref=P0 / 0.528,
              poison=False)
    U1 = _step(c, 12)
    # scales: rho, rho (|u| + c) for the momentum, rho E
    scales = [RHO0] + [RHO0 * (abs(U0) + c0)] * 3 + [float(np.max(np.abs(c.U[4][G:-G, G:-G, G:-G])))]
    for v, a in enumerate(U1):
        ref = c.U[v][G:-G, G:-G, G:-G]
        assert np.max(np.abs(a[G:-G, G:-G, G:-G] - ref)) <= 1e-12 * scales[v], v


def test_shock_boundary_layer_case_runs_and_keeps_its_inflow():
    """The configuration-4 workload at a reduced size: finite after 60 RK4
    steps with viscous terms and the shock filter, the inlet row equal to the
    inflow state bit for bit, and a non-trivial flow at the outflow."""
    c = cases.shock_bl(48, 40, 8)
    U1 = _step(c, 60)
    assert all(np.all(np.isfinite(a)) for a in U1)
    for v in range(5):
        assert np.array_equal(U1[v][G:-G, G:-G, G], c.U[v][G:-G, G:-G, G])
    assert np.max(np.abs(U1[0] - c.U[0])[G:-G, G:-G, G:-G]) > 1e-6
