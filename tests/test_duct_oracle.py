"""Pins of the oracle's walls on two axes of one block (a duct walled on its
i- and j-faces): every face's wall state comes from the same primitives, and
on an edge the face applied last wins -- bc_wall applies imin, imax, jmin,
jmax in order (src/Boundary/mod_bc_apply.f90:237-261), each over its whole
row (smod_bc_wall.f90:11-98 and its twins).

A fluid at rest with p = p0 + a x + b y, T uniform, on a uniform grid: one
stage of crk = 1 without viscosity, the wall state after the step from U^n.
With cextrp2 = 4/3 and cextrp3 = 1/3 (uniform rows) a face's extrapolation of
a linear field is the wall value plus 2/3 of one row's increment along its
normal; the edges take the j-faces' value, the rest of each wall row its own
face's.

Walls on the j- and k-faces: the duct with its axes cycled (x, y, z become
y, z, x) is the same flow, its edges taking the k-faces' state (the duct's
j-faces'), within rounding of the reordered sums."""
import numpy as np

from musica2_newgrid_amd import cases
from musica2_newgrid_amd import grid as gr

G = gr.G


NX, NY, NZ = 12, 14, 5
HX, HY, HZ = 0.5, 0.25, 0.4
P0, T0, A, B = 1.0e5, 300.0, 3.0e3, -5.0e3


def linear_duct():
    """The one-stage duct at rest with p = p0 + a x + b y (uniform rows)."""
    nx, ny, nz, hx, hy, hz, p0, a, b = NX, NY, NZ, HX, HY, HZ, P0, A, B
    x = np.arange(-G, nx + G) * hx
    y = np.arange(-G, ny + G) * hy
    X = x.reshape(1, 1, -1) + 0 * y.reshape(1, -1, 1) + np.zeros((nz + 2 * G, 1, 1))
    Y = y.reshape(1, -1, 1) + 0 * X
    p = p0 + a * X + b * Y
    rho = p / (cases.RG * T0)
    zero = np.zeros_like(p)
    U = [np.ascontiguousarray(q) for q in cases.conserved(rho, zero, zero, zero, p)]
    one = lambda n, h: np.full(n + 2 * G, 1.0 / h)  # noqa: E731
    cart = [one(nx, hx), one(ny, hy), one(nz, hz), one(nx, hx), one(ny, hy), one(nz, hz)]
    c = cases.Case("duct_linear", nx, ny, nz, curvilinear=False, iorder_visc=0, filter=0, cart=cart, U=U,
                   faces=(2, 2, 2, 2, 0, 0), adiabatic=False, t_wall=T0, nrk=1, crk=(1.0,),
                   cextrp2=(4 / 3,) * 4 + (0.0, 0.0), cextrp3=(1 / 3,) * 4 + (0.0, 0.0))
    c.dt = 1.0e-9
    return c, x, y


def test_edges_take_the_j_faces_wall_state():
    from oracle.oracle import Oracle
    nx, ny, hx, hy, p0, a, b = NX, NY, HX, HY, P0, A, B
    c, x, y = linear_duct()
    U1 = c.copy_state()
    Oracle(c).step(U1, 1)
    xi = x[G:G + nx].reshape(1, -1)
    yj = y[G:G + ny].reshape(-1, 1)
    pw = {}
    # each face's value: the wall point's own p plus 2/3 of one row's increment along the normal
    pw["jmin"] = p0 + a * xi + b * (yj[0] + 2.0 / 3.0 * hy)
    pw["jmax"] = p0 + a * xi + b * (yj[-1] - 2.0 / 3.0 * hy)
    pw["imin"] = p0 + a * (xi[0, 0] + 2.0 / 3.0 * hx) + b * yj[:, 0]
    pw["imax"] = p0 + a * (xi[0, -1] - 2.0 / 3.0 * hx) + b * yj[:, 0]
    rho1 = U1[0][G:-G, G:-G, G:-G]  # (nz, ny, nx)
    R = cases.RG * T0
    # j rows (edges included)
    assert np.allclose(rho1[:, 0, :], pw["jmin"][0] / R, rtol=1e-13)
    assert np.allclose(rho1[:, -1, :], pw["jmax"][0] / R, rtol=1e-13)
    # i rows without the edges
    assert np.allclose(rho1[:, 1:-1, 0], pw["imin"][1:-1] / R, rtol=1e-13)
    assert np.allclose(rho1[:, 1:-1, -1], pw["imax"][1:-1] / R, rtol=1e-13)
    # the i-face value differs on the edges, so the test tells them apart
    assert abs(pw["imin"][0] - pw["jmin"][0, 0]) > 1e-6 * p0
    for v in (1, 2, 3):
        w = U1[v][G:-G, G:-G, G:-G]
        assert np.all(w[:, 0, :] == 0.0) and np.all(w[:, :, 0] == 0.0) and np.all(w[:, :, -1] == 0.0)


def test_duct_runs_and_keeps_its_walls():
    from oracle.oracle import Oracle
    c = cases.duct(32, 28, 10, filter=2)
    c.ducros = True
    U1 = c.copy_state()
    Oracle(c).step(U1, 20)
    assert all(np.all(np.isfinite(gr.interior(a))) for a in U1)
    for v in (1, 2, 3):
        w = U1[v][G:-G, G:-G, G:-G]
        assert np.all(w[:, 0, :] == 0.0) and np.all(w[:, -1, :] == 0.0)
        assert np.all(w[:, :, 0] == 0.0) and np.all(w[:, :, -1] == 0.0)


def cycle_axes(c):
    """The same duct with its axes cycled (x, y, z become y, z, x): walls on
    the j- and k-faces, periodic along i."""
    t = cases.Case(c.name + "_jk", c.nz, c.nx, c.ny, curvilinear=False, iorder_visc=c.iorder_visc,
                   filter=c.filter, ducros=c.ducros, diffscale=c.diffscale, chi=c.chi, dt=c.dt, nrk=c.nrk,
                   crk=c.crk, cart=[c.cart[2], c.cart[0], c.cart[1], c.cart[5], c.cart[3], c.cart[4]],
                   faces=tuple(c.faces[f] for f in (4, 5, 0, 1, 2, 3)), adiabatic=c.adiabatic,
                   t_wall=c.t_wall, cextrp2=tuple(c.cextrp2[f] for f in (4, 5, 0, 1, 2, 3)),
                   cextrp3=tuple(c.cextrp3[f] for f in (4, 5, 0, 1, 2, 3)), wall2=c.wall2)
    U = [np.ascontiguousarray(np.transpose(a, (1, 2, 0))) for a in c.U]
    t.U = [U[0], U[3], U[1], U[2], U[4]]
    return t


def test_walls_on_the_j_and_k_faces_are_the_duct_cycled():
    from oracle.oracle import Oracle
    for adiabatic in (True, False):
        c = cases.duct(36, 32, 12, filter=2, adiabatic=adiabatic)
        c.ducros = True
        t = cycle_axes(c)
        Ud, Ut = c.copy_state(), t.copy_state()
        Oracle(c).step(Ud, 10)
        Oracle(t).step(Ut, 10)
        for v, w in zip((0, 1, 2, 3, 4), (0, 3, 1, 2, 4)):
            a = gr.interior(Ut[v])
            b = np.transpose(gr.interior(Ud[w]), (1, 2, 0))
            assert np.max(np.abs(a - b)) <= 1e-13 * (np.max(np.abs(b)) + 1.0), (adiabatic, v)
