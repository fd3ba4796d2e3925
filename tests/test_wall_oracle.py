"""No-slip walls in the oracle (Cartesian, is_wall2 = F), pinned by what the
mathematics fixes.  The reference ships no tests for its wall closures, so:

* each closure row is exact on the polynomials its order allows and not on the
  next degree (a swapped row or a wrong coefficient fails);
* the ghost layers beyond a wall hold NaN: a closure that reads them fails;
* the wall state after a step is the dp/dn = 0 (and dT/dn = 0) extrapolation of
  the primitives the reference's choreography says (those of U^n, which the
  last stage's primitives_visc writes, viscous or not), exact for quadratics
  with zero slope at the wall;
* a channel symmetric about its mid-plane stays symmetric (the max-side
  closures mirror the min-side ones), walls on x-faces give the transposed
  result of walls on y-faces, and a channel cut into two blocks advances like
  the uncut one."""
import math

import numpy as np
import pytest

from musica2_newgrid_amd import cases
from musica2_newgrid_amd import grid as gr
from oracle import oracle as orc

G = gr.G


def _box(nx, ny, nz, h, fields, visc=4, filt=1):
    """Uniform Cartesian box of spacing h, walls on both y-faces, periodic x/z;
    y-ghosts are NaN."""
    ye = (np.arange(ny + 2 * G) - G) * h - 0.5 * (ny - 1) * h
    ide, idv, c2, c3 = gr.wall_metrics(ye, True, True)
    idx = np.full(nx + 2 * G, 1.0 / h)
    idz = np.full(nz + 2 * G, 1.0 / h)
    U = []
    for f in cases.conserved(*fields):
        a = gr.padded(nx, ny, nz, np.nan)
        a[G:-G, G:-G, G:-G] = f
        gr.fill_periodic(a, axes=(0, 2))
        a[G:-G, :G, :] = np.nan
        a[G:-G, G + ny:, :] = np.nan
        U.append(a)
    c = cases.Case("wallbox", nx, ny, nz, curvilinear=False, iorder_visc=visc, filter=filt, ducros=False,
                   cart=[idx, ide, idz, idx.copy(), idv, idz.copy()], U=U, dt=1e-6,
                   faces=(0, 0, 2, 2, 0, 0),
                   cextrp2=(0, 0, c2[0], c2[1], 0, 0), cextrp3=(0, 0, c3[0], c3[1], 0, 0))
    return c, ye[G:G + ny]


def _fields_y(nx, ny, nz, rho, u, v, w, p):
    sh = (nz, ny, nx)
    return [np.broadcast_to(np.asarray(q, dtype=np.float64).reshape(1, -1, 1) if np.ndim(q) else q, sh).copy()
            for q in (rho, u, v, w, p)]


NX, NY, NZ, H = 6, 24, 6, 0.08


def _rows(K):
    """K[..., j, ...] profile along y (the fields only depend on y)."""
    return gr.interior(K)[2, :, 3]


@pytest.mark.parametrize("m", [2, 4, 6, 8])
def test_euler_wall_rows_exact_on_polynomials(m):
    ye = (np.arange(NY) - 0.5 * (NY - 1)) * H
    v = (ye / ye.max()) ** m
    c, y = _box(NX, NY, NZ, H, _fields_y(NX, NY, NZ, 1.0, 0.0, v, 0.0, 1.0e5))
    K = orc.Oracle(c).euler_rhs(c.U)
    kr = _rows(K[0])
    exact = m * ye ** (m - 1) / ye.max() ** m
    order = np.array([0, 2, 4, 6, 8] + [10] * (NY - 10) + [8, 6, 4, 2, 0])
    scale = np.max(np.abs(exact))
    assert kr[0] == 0.0 and kr[-1] == 0.0, "no Euler increment on the wall rows"
    for r in range(1, NY - 1):
        err = abs(kr[r] - exact[r]) / scale
        if order[r] >= m:
            assert err < 1e-11, (r, order[r], err)
        elif order[r] == m - 2:
            assert err > 1e-6, (r, order[r], err)


def test_viscous_wall_rows_exact_on_quartics():
    ye = (np.arange(NY) - 0.5 * (NY - 1)) * H
    s = ye / ye.max()
    P = 3.0 + 2.0 * s - 1.5 * s ** 2 + 0.7 * s ** 3 - 0.4 * s ** 4
    d2P = (-3.0 + 4.2 * s - 4.8 * s ** 2) / ye.max() ** 2
    c, _ = _box(NX, NY, NZ, H, _fields_y(NX, NY, NZ, 1.0, P, 0.0, 0.0, 1.0e5))
    K = orc.Oracle(c).visc_rhs(c.U)
    T = 1.0e5 / (1.0 * cases.RG)
    mu = cases.C0 * math.sqrt(T ** 3) / (T + cases.S0) * c.diffscale
    kru = _rows(K[1])
    assert np.all(np.isfinite(kru))
    assert np.max(np.abs(kru - (-mu * d2P))) < 1e-10 * np.max(np.abs(mu * d2P))
    assert np.all(_rows(K[0]) == 0.0)


def test_filter_wall_rows():
    chi = 0.2
    ye = (np.arange(NY) - 0.5 * (NY - 1)) * H
    s = ye / ye.max()
    # quartic: the 5-pt filter (4th difference) on row 3 sees it, the wider ones do not
    c, _ = _box(NX, NY, NZ, H, _fields_y(NX, NY, NZ, 1.0 + 0.1 * s ** 4, 0.0, 0.0, 0.0, 1.0e5))
    K = orc.Oracle(c).filter_rhs(c.U)
    kr = _rows(K[0])
    d4 = 0.1 * 24.0 * (H / ye.max()) ** 4        # 4th difference of 0.1 s^4
    expect = np.zeros(NY)
    expect[2] = expect[-3] = -chi / 16.0 * d4
    assert np.max(np.abs(kr - expect)) < 1e-12 * 0.1 + 1e-3 * abs(expect[2])
    assert abs(kr[2] - expect[2]) < 1e-9 * abs(expect[2]) + 1e-15
    # odd-even mode: every filter has unit response at the grid cut-off, rows 1,2 get none
    alt = (-1.0) ** np.arange(NY)
    c, _ = _box(NX, NY, NZ, H, _fields_y(NX, NY, NZ, 1.0 + 0.01 * alt, 0.0, 0.0, 0.0, 1.0e5))
    kr = _rows(orc.Oracle(c).filter_rhs(c.U)[0])
    expect = -chi * 0.01 * alt
    expect[[0, 1, -2, -1]] = 0.0
    assert np.max(np.abs(kr - expect)) < 1e-15


def _channel_quadratic(visc, nrk, adiabatic=True, filt=1):
    c = cases.channel(16, 20, 8, filter=filt, iorder_visc=visc, adiabatic=adiabatic, t_wall=290.0)
    c.nrk = nrk
    c.crk = cases.CRK4 if nrk == 4 else ((0.5, 1.0) if nrk == 2 else (1.0,))
    y = c.info["y"]
    A, B, TA, TB = 1.0e5, 3.0e3, 300.0, 20.0
    rho = gr.interior(c.U[0])
    u, v, w = (gr.interior(c.U[q]) / rho for q in (1, 2, 3))
    p = A + B * ((y - y[0]) / (y[-1] - y[0])) ** 2
    T = TA + TB * ((y - y[0]) / (y[-1] - y[0])) ** 2
    p = np.broadcast_to(p.reshape(1, -1, 1), rho.shape)
    T = np.broadcast_to(T.reshape(1, -1, 1), rho.shape)
    r = p / (cases.RG * T)
    for a, f in zip(c.U, cases.conserved(r, u, v, w, p)):
        gr.interior(a)[...] = f
        a[G:-G, :G, G:-G] = f[:, :1, :]
        a[G:-G, -G:, G:-G] = f[:, -1:, :]
        gr.fill_periodic(a, axes=(0, 2))
    return c, A, TA


@pytest.mark.parametrize("visc,nrk", [(4, 4), (0, 1)])
@pytest.mark.parametrize("adiabatic", [True, False])
def test_wall_state_extrapolates_the_right_primitives(visc, nrk, adiabatic):
    """After a viscous step the end-of-step wall state comes from U^n (the
    primitives primitives_visc wrote last); with one stage from that stage's
    input, which is U^n too: a quadratic with zero slope at the wall is
    extrapolated exactly."""
    c, A, TA = _channel_quadratic(visc, nrk, adiabatic)
    U = c.copy_state()
    orc.Oracle(c).step(U, 1)
    Tw = TA if adiabatic else c.t_wall
    row = [gr.interior(a)[:, 0, :] for a in U]
    rho_w = A / (cases.RG * Tw)
    assert np.all(row[1] == 0.0) and np.all(row[2] == 0.0) and np.all(row[3] == 0.0)
    assert np.max(np.abs(row[0] / rho_w - 1.0)) < 1e-13
    cv = cases.RG / (cases.GAM - 1.0)
    assert np.max(np.abs(row[4] / (rho_w * cv * Tw) - 1.0)) < 1e-13


def test_wall_state_after_an_inviscid_step_comes_from_u_n():
    """nrk = 2, Euler only, no filter: the last stage calls primitives_visc
    (runge_kutta.f90:71) whether or not the case is viscous, so the end-of-step
    wall state is the extrapolation of U^n's primitives, not of the second
    stage's input U1 = U^n - dt crk1 K(U^n)."""
    c, _, _ = _channel_quadratic(0, 2, True, filt=0)
    o = orc.Oracle(c)
    rho0 = gr.interior(c.U[0])
    p0 = (cases.GAM - 1.0) * (gr.interior(c.U[4]) - 0.5 * (gr.interior(c.U[1]) ** 2 + gr.interior(c.U[2]) ** 2
                                                           + gr.interior(c.U[3]) ** 2) / rho0)
    T0 = p0 / (rho0 * cases.RG)
    c2, c3 = c.cextrp2[2], c.cextrp3[2]
    rw0 = (c2 * p0[:, 1, :] - c3 * p0[:, 2, :]) / (cases.RG * (c2 * T0[:, 1, :] - c3 * T0[:, 2, :]))
    U = c.copy_state()
    o.step(U, 1)
    assert np.max(np.abs(gr.interior(U[0])[:, 0, :] / rw0 - 1.0)) < 1e-13
    # the extrapolation of the second stage's input differs: the test sees the choice
    K = o.euler_rhs(c.U)
    r1 = [gr.interior(a - c.dt * c.crk[0] * k) for a, k in zip(c.U, K)]
    q2 = (r1[1] ** 2 + r1[2] ** 2 + r1[3] ** 2) / r1[0]
    p1 = (cases.GAM - 1.0) * (r1[4] - 0.5 * q2)
    T1 = p1 / (r1[0] * cases.RG)
    rw1 = (c2 * p1[:, 1, :] - c3 * p1[:, 2, :]) / (cases.RG * (c2 * T1[:, 1, :] - c3 * T1[:, 2, :]))
    assert np.max(np.abs(gr.interior(U[0])[:, 0, :] / rw1 - 1.0)) > 1e-9


def _mirror_case():
    c = cases.channel(12, 22, 6, symmetric=True)
    return c


def test_symmetric_channel_stays_symmetric():
    c = _mirror_case()
    U = c.copy_state()
    orc.Oracle(c).step(U, 3)
    r = [gr.interior(a) for a in U]
    sgn = [1, 1, -1, 1, 1]
    for a, s in zip(r, sgn):
        scale = np.max(np.abs(a))
        assert np.max(np.abs(a - s * a[:, ::-1, :])) < 1e-12 * scale


def _transpose_case(c):
    """The same channel with x and y exchanged: walls on the x-faces."""
    t = cases.Case(c.name + "_t", c.ny, c.nx, c.nz, curvilinear=False, iorder_visc=c.iorder_visc,
                   filter=c.filter, ducros=False, diffscale=c.diffscale, chi=c.chi, dt=c.dt, nrk=c.nrk,
                   crk=c.crk, cart=[c.cart[1], c.cart[0], c.cart[2], c.cart[4], c.cart[3], c.cart[5]],
                   faces=(2, 2, 0, 0, 0, 0), adiabatic=c.adiabatic, t_wall=c.t_wall,
                   cextrp2=(c.cextrp2[2], c.cextrp2[3], 0, 0, 0, 0),
                   cextrp3=(c.cextrp3[2], c.cextrp3[3], 0, 0, 0, 0))
    U = [np.ascontiguousarray(np.swapaxes(a, 1, 2)) for a in c.U]
    U[1], U[2] = U[2], U[1]
    t.U = U
    return t


def test_walls_on_x_faces_give_the_transposed_result():
    c = cases.channel(12, 20, 6)
    t = _transpose_case(c)
    A = c.copy_state()
    orc.Oracle(c).step(A, 2)
    B = t.copy_state()
    orc.Oracle(t).step(B, 2)
    B = [np.swapaxes(b, 1, 2) for b in B]
    B[1], B[2] = B[2], B[1]
    for a, b in zip(A, B):
        ai, bi = gr.interior(a), gr.interior(b)
        assert np.max(np.abs(ai - bi)) < 1e-12 * np.max(np.abs(ai))


def test_ghosts_beyond_walls_are_never_read():
    c = cases.channel(12, 20, 6)
    A = c.copy_state()
    orc.Oracle(c).step(A, 2)
    B = c.copy_state()
    for a in B:
        a[:, :G, :] = np.nan
        a[:, -G:, :] = np.nan
    orc.Oracle(c).step(B, 2)
    for a, b in zip(A, B):
        assert np.array_equal(gr.interior(a), gr.interior(b))


def split_x_blocks(c):
    """Block table and block cases of c cut at i = nx/2 into two blocks joined
    by aligned interfaces (the cut and the periodic wrap), walls on the j-faces."""
    h = c.nx // 2
    blocks, bcs = [], []
    for b in range(2):
        sub = cases.Case(f"half{b}", h, c.ny, c.nz, curvilinear=False, iorder_visc=c.iorder_visc,
                         filter=c.filter, ducros=c.ducros, diffscale=c.diffscale, chi=c.chi, dt=c.dt,
                         dissip_shock=c.dissip_shock, csens=c.csens,
                         cart=[c.cart[0][:h + 2 * G].copy(), c.cart[1], c.cart[2],
                               c.cart[3][:h + 2 * G].copy(), c.cart[4], c.cart[5]],
                         faces=(1, 1, 2, 2, 1, 1), adiabatic=c.adiabatic, t_wall=c.t_wall,
                         cextrp2=c.cextrp2, cextrp3=c.cextrp3)
        sub.U = [np.ascontiguousarray(a[:, :, b * h:b * h + h + 2 * G]) for a in c.U]
        bcs.append(sub)
        blocks.append(dict(nx=h, ny=c.ny, nz=c.nz, nbr_block=[1 - b, 1 - b, 0, 0, b, b],
                           nbr_face=[1, 0, 0, 0, 5, 4], orient=[0] * 6, wall=[0, 0, 1, 1, 0, 0],
                           kind=[1, 1, 2, 2, 1, 1]))
    return blocks, bcs


def _split_x(c, nsteps):
    """c advanced as two blocks cut at i = nx/2 (aligned interfaces)."""
    blocks, bcs = split_x_blocks(c)
    Us = [bc.copy_state() for bc in bcs]
    orc.step_multi(c, blocks, bcs, Us, nsteps)
    return [np.concatenate([gr.interior(Us[b][v]) for b in range(2)], axis=2) for v in range(5)]


def test_channel_cut_in_two_blocks_euler_equals_parent():
    """Without viscous terms and filter nothing reads a wall row's face
    ghosts, so the cut channel advances exactly like the uncut one."""
    c = cases.channel(16, 20, 6, filter=0, iorder_visc=0)
    A = c.copy_state()
    orc.Oracle(c).step(A, 2)
    B = _split_x(c, 2)
    for v in range(5):
        assert np.array_equal(B[v], gr.interior(A[v]))


def test_channel_cut_differs_only_next_to_the_cut():
    """The exchange runs before bc_wall, so the face ghosts of a wall row hold
    the neighbour's value from before the wall state was imposed, while an
    uncut block reads the imposed value: with viscous terms and the filter,
    which read those ghosts, a cut changes the result from the second step on
    (the initial ghosts are consistent), and only within a few cells of the cut
    (the periodic seam is a cut in both)."""
    c = cases.channel(64, 20, 6)
    A = c.copy_state()
    orc.Oracle(c).step(A, 2)
    B = _split_x(c, 2)
    h = c.nx // 2
    far = np.r_[0:h - 12, h + 12:c.nx]
    differs = False
    for v in range(5):
        ref = gr.interior(A[v])
        assert np.array_equal(B[v][:, :, far], ref[:, :, far])
        differs |= not np.array_equal(B[v], ref)
    assert differs


def test_adiabatic_wall_has_no_normal_heat_flux_on_its_row():
    """Adiabatic walls zero the wall-normal temperature gradient on the wall
    row (smod_grad_T.f90:102-109): on a gas at rest with T linear in y, the
    adiabatic minus the isothermal viscous increment of rho*E is the removed
    wall heat flux kappa*a carried by the divergence stencils that touch the
    wall row: a04(1) on row 1, a13(1) on row 2, -a5(2) on row 3, nothing
    further in."""
    ye = (np.arange(NY) - 0.5 * (NY - 1)) * H
    slope = 40.0
    T = 300.0 + slope * ye
    p0 = 1.0e5
    rho = p0 / (cases.RG * T)
    out = {}
    for adiab in (True, False):
        c, _ = _box(NX, NY, NZ, H, _fields_y(NX, NY, NZ, rho, 0.0, 0.0, 0.0, p0))
        c.adiabatic = adiab
        out[adiab] = _rows(orc.Oracle(c).visc_rhs(c.U)[4])
    T1 = T[0]
    kappa1 = cases.C0 * math.sqrt(T1 ** 3) / (T1 + cases.S0) * c.diffscale * c.cok_ratio
    idv = c.cart[4][G:G + NY]
    coef = np.zeros(NY)
    coef[0], coef[1], coef[2] = gr.A04[0], gr.A13[0], -gr.A5[1]
    expect = coef * kappa1 * slope * idv
    d = out[True] - out[False]
    assert np.max(np.abs(d[:NY // 2] - expect[:NY // 2])) < 1e-9 * np.max(np.abs(expect))
    assert abs(d[0]) > 0.0
