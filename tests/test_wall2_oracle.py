"""Pins of the oracle's advanced walls (is_wall2, type_wall_bc = 2): the wall
rows take an Euler increment with the 3-point one-sided a02 / a20 derivative
across the wall (src/Compute_flow/smod_flux_euler_11pts.f90:42-54, 254-265,
286-297, 584-595; _c3.f90:291-302) and the wall states of
src/Boundary/smod_bc_wall_2.f90 and smod_bc_wall_2_c3.f90, bound as
src/Main/assign_procedures.f90:415-470 binds them.

* A state whose fluxes are linear in y (density linear in y, uniform velocity
  and pressure) between stretched walls: one stage advances every row, the
  wall rows included, by the exact derivative (a02 and the 11-pt scheme are
  exact on linear functions at any spacing, with idy1 = 1/(a02 y) on the wall
  row); the slip walls of the inviscid case then remove the normal momentum
  and keep the other components.  The same on the i-faces.
* An adiabatic no-slip wall with T = T0 + b (y - y1)^2: the extrapolation
  cextrp2 T(2) - cextrp3 T(3) of a zero-slope quadratic is T0 at any spacing,
  so rhoE = rho cv T0 on the wall row whatever the advanced rho.
* A c3 slip wall on a skewed plane (its two unit tangents not orthogonal):
  the momentum keeps its tangential part exactly, m.n = 0 and the change is
  along n, with n the wall normal of the map, computed here on its own.
* The sheared c3 block of the dp/dn = 0 pins: the adiabatic T_wall with its
  tangential term as written.
* Unit-spacing c3 = Cartesian with advanced walls (slip and no slip, every
  filter, viscous).
* The slip flag selects nothing without is_wall2, and an isothermal case binds
  no slip wall.
"""
import math

import numpy as np
import pytest

from musica2_newgrid_amd import cases
from musica2_newgrid_amd import grid as gr

G = gr.G
CV = cases.RG / (cases.GAM - 1.0)


def _oracle(c, n):
    from oracle.oracle import Oracle
    U = c.copy_state()
    Oracle(c).step(U, n)
    return U


def _fill(fields, nx, ny, nz, periodic_axes):
    U = []
    for f in fields:
        a = gr.padded(nx, ny, nz)
        a[G:-G, G:-G, G:-G] = f
        # ghosts beyond the walls repeat the wall rows (nothing on the path reads them)
        a[G:-G, :G, G:-G] = a[G:-G, G:G + 1, G:-G]
        a[G:-G, G + ny:, G:-G] = a[G:-G, G + ny - 1:G + ny, G:-G]
        a[G:-G, :, :G] = a[G:-G, :, G:G + 1]
        a[G:-G, :, G + nx:] = a[G:-G, :, G + nx - 1:G + nx]
        gr.fill_periodic(a, axes=periodic_axes)
        U.append(np.ascontiguousarray(a))
    return U


def _linear_channel(faces=(0, 0, 5, 5, 0, 0), nx=12, ny=24, nz=6, slope=0.3, V0=40.0, U0=25.0, W0=-15.0):
    """Stretched walls on the j-faces, periodic x and z; rho = rho0 (1 + slope
    y), u, v, w, p uniform; inviscid, no filter, one stage of crk = 1."""
    Ly = 1.0
    ye = cases.stretched_wall_coords(ny, Ly, 1.3)
    y = ye[G:G + ny]
    rho0, p0 = 1.2, 1.0e5
    Y = np.broadcast_to(y.reshape(1, -1, 1), (nz, ny, nx))
    rho = rho0 * (1.0 + slope * Y)
    one = np.ones_like(rho)
    fields = cases.conserved(rho, U0 * one, V0 * one, W0 * one, p0 * one)
    U = _fill(fields, nx, ny, nz, (0, 2))
    ide, idv, c2, c3 = gr.wall_metrics(ye, True, True)
    dx = 0.1
    idx = np.full(nx + 2 * G, 1.0 / dx)
    c = cases.Case("lin", nx, ny, nz, curvilinear=False, iorder_visc=0, filter=0, nrk=1, crk=(1.0,),
                   cart=[idx, ide, idx[:nz + 2 * G].copy(), idx.copy(), idv, idx[:nz + 2 * G].copy()], U=U,
                   faces=faces, wall2=True, cextrp2=(0.0, 0.0, c2[0], c2[1], 0.0, 0.0),
                   cextrp3=(0.0, 0.0, c3[0], c3[1], 0.0, 0.0))
    c.dt = 2.0e-6
    return c, Y, rho0


@pytest.mark.parametrize("flag", [2, 5])
def test_advanced_rows_take_the_exact_derivative_and_slip_walls_keep_the_tangential_momentum(flag):
    """Inviscid case with is_wall2: every wall is a slip wall, with or without
    the flag (assign_procedures.f90:416-437)."""
    slope, V0, U0, W0 = 0.3, 40.0, 25.0, -15.0
    c, Y, rho0 = _linear_channel(faces=(0, 0, flag, flag, 0, 0), slope=slope, V0=V0, U0=U0, W0=W0)
    U1 = _oracle(c, 1)
    dt = c.dt
    beta = rho0 * slope  # d rho / dy
    ke = 0.5 * (U0 ** 2 + V0 ** 2 + W0 ** 2)
    # K = d(F_y)/dy with F_y = (rho v, rho u v, rho v^2 + p, rho w v, (rhoE + p) v), linear in y
    want = [rho0 * (1 + slope * Y) - dt * beta * V0,
            rho0 * (1 + slope * Y) * U0 - dt * beta * V0 * U0,
            rho0 * (1 + slope * Y) * V0 - dt * beta * V0 * V0,
            rho0 * (1 + slope * Y) * W0 - dt * beta * V0 * W0,
            gr.interior(c.U[4]) - dt * beta * ke * V0]
    for v in range(5):
        got = gr.interior(U1[v])
        w = want[v].copy()
        if v == 2:
            w[:, 0, :] = 0.0   # slip: no normal momentum on the wall rows
            w[:, -1, :] = 0.0
        scale = float(np.max(np.abs(want[v])))
        assert np.max(np.abs(got - w)) <= 1e-12 * scale, v
    # the wall rows did move: the advanced density differs from the initial one
    assert np.max(np.abs(gr.interior(U1[0])[:, 0, :] - gr.interior(c.U[0])[:, 0, :])) > 1e-7


def test_slip_walls_on_the_i_faces():
    """The same linear state with x and y exchanged: walls on the i-faces."""
    slope, V0, U0, W0 = 0.3, 40.0, 25.0, -15.0
    c, Y, rho0 = _linear_channel(slope=slope, V0=V0, U0=U0, W0=W0)
    t = cases.Case("lin_i", c.ny, c.nx, c.nz, curvilinear=False, iorder_visc=0, filter=0, nrk=1, crk=(1.0,),
                   cart=[c.cart[1], c.cart[0], c.cart[2], c.cart[4], c.cart[3], c.cart[5]],
                   faces=(5, 5, 0, 0, 0, 0), wall2=True, cextrp2=(c.cextrp2[2], c.cextrp2[3], 0, 0, 0, 0),
                   cextrp3=(c.cextrp3[2], c.cextrp3[3], 0, 0, 0, 0))
    U = [np.ascontiguousarray(np.swapaxes(a, 1, 2)) for a in c.U]
    U[1], U[2] = U[2], U[1]
    t.U = U
    t.dt = c.dt
    Ua = _oracle(c, 1)
    Ui = _oracle(t, 1)
    # the i-walls' rows: no momentum along x there
    assert np.all(gr.interior(Ui[1])[:, :, 0] == 0.0) and np.all(gr.interior(Ui[1])[:, :, -1] == 0.0)
    Ub = [np.swapaxes(a, 1, 2) for a in Ui]
    Ub[1], Ub[2] = Ub[2], Ub[1]
    for v in range(5):
        a, b = gr.interior(Ua[v]), gr.interior(Ub[v])
        assert np.max(np.abs(a - b)) <= 1e-12 * float(np.max(np.abs(a))), v


def test_advanced_adiabatic_wall_extrapolates_a_zero_slope_temperature():
    """Viscous adiabatic case with is_wall2 and no slip flag: no-slip advanced
    walls; T = T0 + b (y - y1)^2 gives T_wall = T0 on the jmin wall row."""
    nx, ny, nz = 12, 24, 6
    ye = cases.stretched_wall_coords(ny, 1.0, 1.3)
    y = ye[G:G + ny]
    T0, b, p0 = 300.0, 40.0, 1.0e5
    Y = np.broadcast_to(y.reshape(1, -1, 1), (nz, ny, nx))
    T = T0 + b * (Y - y[0]) ** 2
    rho = p0 / (cases.RG * T)
    u = 20.0 * np.sin(math.pi * Y)
    v = 5.0 * np.sin(math.pi * Y)  # a mass flux through rows 2, 3: the wall row's density moves
    zero = np.zeros_like(rho)
    U = _fill(cases.conserved(rho, u, v, zero, p0 + zero), nx, ny, nz, (0, 2))
    ide, idv, c2, c3 = gr.wall_metrics(ye, True, True)
    idx = np.full(nx + 2 * G, 10.0)
    c = cases.Case("adiab", nx, ny, nz, curvilinear=False, iorder_visc=4, filter=0, nrk=1, crk=(1.0,),
                   cart=[idx, ide, idx[:nz + 2 * G].copy(), idx.copy(), idv, idx[:nz + 2 * G].copy()], U=U,
                   faces=(0, 0, 2, 2, 0, 0), wall2=True, adiabatic=True, diffscale=5.0,
                   cextrp2=(0.0, 0.0, c2[0], c2[1], 0.0, 0.0), cextrp3=(0.0, 0.0, c3[0], c3[1], 0.0, 0.0))
    c.dt = 1.0e-6
    U1 = _oracle(c, 1)
    rho_w = U1[0][G:-G, G, G:-G]
    assert np.max(np.abs(rho_w - rho[:, 0, :])) > 1e-9  # the wall row was advanced
    assert np.allclose(U1[4][G:-G, G, G:-G], rho_w * CV * T0, rtol=1e-13, atol=0.0)
    for v in (1, 2, 3):
        assert np.all(U1[v][G:-G, G, G:-G] == 0.0) and np.all(U1[v][G:-G, G + ny - 1, G:-G] == 0.0)


def _skewed(a, b, nx=12, ny=12, nz=10, dxi=0.5, deta=0.25, dphi=0.4):
    """c3 block x = xi, y = eta + a xi + b zeta, z = zeta (constant metrics),
    walls on the j-faces: planes whose unit tangents along ksi and phi make an
    angle other than 90 degrees when a b != 0."""
    # x_i = (dxi, a dxi, 0), x_j = (0, deta, 0), x_k = (0, b dphi, dphi)
    J = dxi * deta * dphi
    xi_ = np.array([dxi, a * dxi, 0.0])
    xj_ = np.array([0.0, deta, 0.0])
    xk_ = np.array([0.0, b * dphi, dphi])
    # J grad(ksi) = x_j x x_k, J grad(eta) = x_k x x_i, J grad(phi) = x_i x x_j
    gk, ge, gp = np.cross(xj_, xk_), np.cross(xk_, xi_), np.cross(xi_, xj_)
    vals = list(gk) + list(ge) + list(gp) + [1.0 / J]
    m = [gr.padded(nx, ny, nz, float(vals[q])) for q in range(10)]
    mv = [gr.padded(nx, ny, nz, float(vals[q])) for q in range(10)]
    ek = xi_ / np.linalg.norm(xi_)
    ep = xk_ / np.linalg.norm(xk_)
    n2 = float(ge @ ge)
    g1, g2 = -float(gk @ ge) / n2, -float(gp @ ge) / n2
    slab = lambda v: np.full((nz, nx), float(v))  # noqa: E731
    eks, eps = [slab(q) for q in ek], [slab(q) for q in ep]
    wall = [slab(g1), slab(g2)] + eks + eps + gr.dual_tangents(eks, eps)
    return m, mv, {2: wall, 3: [w.copy() for w in wall]}, np.cross(ek, ep)


def test_c3_slip_wall_keeps_exactly_the_tangential_momentum():
    nx, ny, nz = 12, 12, 10
    m, mv, wall, nrm = _skewed(0.7, -0.9)
    nrm = nrm / np.linalg.norm(nrm)
    rng = np.random.default_rng(3)
    rho = 1.2 + 0.1 * rng.random((nz, ny, nx))
    u, v, w = (30.0 * rng.standard_normal((nz, ny, nx)) for _ in range(3))
    p = 1.0e5 * (1.0 + 0.01 * rng.random((nz, ny, nx)))
    U = _fill(cases.conserved(rho, u, v, w, p), nx, ny, nz, (0, 2))
    c = cases.Case("skew", nx, ny, nz, curvilinear=True, iorder_visc=0, filter=0, nrk=1, crk=(1.0,), m=m, mv=mv,
                   U=U, faces=(0, 0, 5, 5, 0, 0), wall2=True)
    c.wall_c3 = wall
    c.dt = 0.0  # the state stays; the end-of-step wall state acts alone
    U1 = _oracle(c, 1)
    assert abs(float(wall[2][2][0, 0] * wall[2][5][0, 0] + wall[2][3][0, 0] * wall[2][6][0, 0] +
                     wall[2][4][0, 0] * wall[2][7][0, 0])) > 0.3  # tangents far from orthogonal
    for j in (0, ny - 1):
        m0 = np.stack([U[q][G:-G, G + j, G:-G] for q in (1, 2, 3)], axis=-1)
        m1 = np.stack([U1[q][G:-G, G + j, G:-G] for q in (1, 2, 3)], axis=-1)
        scale = float(np.max(np.abs(m0)))
        assert np.max(np.abs(m1 @ nrm)) <= 1e-13 * scale                      # no normal momentum
        assert np.max(np.abs(np.cross(m1 - m0, nrm))) <= 1e-13 * scale         # only the normal part removed
        assert np.max(np.abs(m1 - m0)) > 1e-3 * scale
        for q in (0, 4):
            assert np.array_equal(U1[q][G:-G, G + j, G:-G], U[q][G:-G, G + j, G:-G])
    # the interior is untouched with dt = 0
    for q in range(5):
        assert np.array_equal(U1[q][G:-G, G + 1:G + ny - 1, G:-G], U[q][G:-G, G + 1:G + ny - 1, G:-G])


@pytest.mark.parametrize("adiabatic", [True, False])
def test_sheared_advanced_wall_temperature_is_the_closed_form(adiabatic):
    from tests.test_c3_wall_oracle import _sheared
    s, a, b = 0.6, 2.0e3, 40.0
    c, X, p0, T0 = _sheared(s, a, b, adiabatic)
    c.wall2 = True
    c.iorder_visc = 4  # viscous: the no-slip advanced walls of the case's kind
    c.diffscale = 0.0
    U1 = _oracle(c, 1)
    deta = 0.25
    x1 = X[G:-G, G, G:-G]
    T_want = (T0 + b * x1 + 4.0 / 3.0 * b * s * deta) if adiabatic else T0 + 0 * x1
    sl = (slice(None), slice(1, -1))
    rho = U1[0][G:-G, G, G:-G]
    assert np.allclose(U1[4][G:-G, G, G:-G][sl], (rho * CV * T_want)[sl], rtol=1e-12)
    for v in (1, 2, 3):
        assert np.all(U1[v][G:-G, G, G:-G] == 0.0)


@pytest.mark.parametrize("filt,adiabatic,flag", [(2, True, 5), (1, True, 2), (0, False, 5), (2, True, 2)])
def test_unit_c3_advanced_walls_equal_cartesian(filt, adiabatic, flag):
    from tests.test_c3_wall_oracle import _pair
    cc, c3c, _, _ = _pair(20, 16, 10, filt, adiabatic)
    for c in (cc, c3c):
        c.wall2 = True
        c.faces = (0, 0, flag, flag, 0, 0)
    z = np.zeros((10, 20))
    for f in (2, 3):
        c3c.wall_c3[f] = c3c.wall_c3[f] + [z + 1, z, z, z, z, z + 1]
    Ua = _oracle(cc, 6)
    Ub = _oracle(c3c, 6)
    for v in range(5):
        a, b = gr.interior(Ua[v]), gr.interior(Ub[v])
        scale = max(float(np.max(np.abs(gr.interior(cc.U[k])))) for k in ((1, 2, 3) if 1 <= v <= 3 else (v,)))
        assert np.max(np.abs(a - b)) <= 1e-13 * scale, v
    slip = flag == 5 and adiabatic
    # slip: no normal momentum, tangential momentum kept; no slip: no momentum
    assert np.all(Ua[2][G:-G, G, G:-G] == 0.0)
    assert (np.max(np.abs(Ua[1][G:-G, G, G:-G])) > 0.0) == slip


def test_slip_flag_without_wall2_is_a_no_slip_wall():
    c, _, _ = _linear_channel(faces=(0, 0, 5, 5, 0, 0))
    d, _, _ = _linear_channel(faces=(0, 0, 2, 2, 0, 0))
    for x in (c, d):
        x.wall2 = False
    Ua, Ub = _oracle(c, 1), _oracle(d, 1)
    for a, b in zip(Ua, Ub):
        assert np.array_equal(gr.interior(a), gr.interior(b))
    # an isothermal viscous case binds no slip routine either
    e, _, _ = _linear_channel(faces=(0, 0, 5, 5, 0, 0))
    f, _, _ = _linear_channel(faces=(0, 0, 2, 2, 0, 0))
    for x in (e, f):
        x.iorder_visc, x.adiabatic, x.t_wall, x.diffscale = 4, False, 300.0, 1.0
    Ue, Uf = _oracle(e, 1), _oracle(f, 1)
    for a, b in zip(Ue, Uf):
        assert np.array_equal(gr.interior(a), gr.interior(b))
    assert np.all(Ue[1][G:-G, G, G:-G] == 0.0)
