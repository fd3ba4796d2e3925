"""The oracle's 2.5-D curvilinear geometry (is_curv: flux_euler_11pts_c,
grad_vel_c / grad_T_c, flux_visc_5pts_c; 2-D metrics in (i, j), Cartesian
z), pinned by what the mathematics fixes:

* free stream: a uniform state on a wavy grid has K = 0 to round-off (the
  2-D metrics are centred derivatives of the node coordinates, and the
  centred i- and j-derivatives commute on a periodic grid);
* an undistorted 2.5-D block gives the Cartesian block's step;
* on a z-extruded wavy grid the 2.5-D block and the c3 block with the 3-D
  GCL metrics of the same grid give the same step (the c3 metrics reduce to
  dz times the 2-D ones, 1/J3 to ijacob/dz and phi_z 1/J3 to idz), with the
  viscous terms and the shock filter with and without the Edoh switch (the
  states carry density and z-momentum modes along z, so that the z-fluxes
  and their metric enter)."""
import math

import numpy as np
import pytest

from musica2_newgrid_amd import cases
from musica2_newgrid_amd import grid as gr
from oracle.oracle import Oracle
from tests._util import relerr

G = gr.G


def _add_z_modes(c):
    """Density and z-momentum modes along z (the vortex alone has none), so
    that the z-fluxes and their metric enter."""
    k = np.arange(c.nz + 2 * G) - G
    z = (2.0 * math.pi / c.nz) * k
    fr = (1.0 + 0.02 * np.sin(z))[:, None, None]
    fw = (30.0 * np.cos(z))[:, None, None]
    rho0 = c.U[0].copy()
    c.U[0] = rho0 * fr
    for v in (1, 2, 3):
        c.U[v] = c.U[v] * fr
    c.U[3] = c.U[3] + c.U[0] * fw
    c.U[4] = c.U[4] * fr + 0.5 * c.U[0] * fw ** 2


def _c3_twin(c2):
    """The c3 case of the same z-extruded grid and state."""
    L = c2.info["lengths"]
    f2 = cases.wavy_mapping_2d(L, c2.info["amp"])

    def f3(XI, ET, ZE):
        x, y = f2(XI, ET)
        return x, y, ZE + 0.0 * XI

    m, mv, _ = gr.curvilinear_metrics_periodic(f3, c2.nx, c2.ny, c2.nz, L, viscous=True)
    c3 = cases.Case("c3twin", c2.nx, c2.ny, c2.nz, curvilinear=True, iorder_visc=c2.iorder_visc,
                    filter=c2.filter, ducros=c2.ducros, diffscale=c2.diffscale, m=m, mv=mv,
                    U=[a.copy() for a in c2.U], dt=c2.dt, chi=c2.chi)
    c3.sw_edoh = getattr(c2, "sw_edoh", False)
    return c3


def test_free_stream_is_preserved_on_a_wavy_grid():
    c = cases.vortex_wavy_c2(24, 20, 12, beta=0.0)
    o = Oracle(c)
    K = o.euler_rhs(c.U)
    # scale of one flux derivative term: p |metric| / J
    p = 1.177 * 287.0 * 300.0
    scale = p * max(float(np.max(np.abs(a))) for a in c.m2[:4]) * float(np.max(np.abs(c.m2[4])))
    for v in range(5):
        assert np.max(np.abs(gr.interior(K[v]))) < 1e-11 * scale
    # the same grid with a vortex: the increment is not small
    cv = cases.vortex_wavy_c2(24, 20, 12)
    Kv = Oracle(cv).euler_rhs(cv.U)
    assert np.max(np.abs(gr.interior(Kv[0]))) > 1e-4 * scale / p


def test_undistorted_c2_equals_cartesian():
    c2 = cases.vortex_wavy_c2(24, 20, 12, amp=(0.0, 0.0), filter=1)
    _add_z_modes(c2)
    L = c2.info["lengths"]
    dx, dy, dz = L[0] / 24, L[1] / 20, L[2] / 12
    cc = cases.Case("cart", 24, 20, 12, curvilinear=False, iorder_visc=4, filter=1, ducros=False,
                    diffscale=c2.diffscale, dt=c2.dt, chi=c2.chi, U=[a.copy() for a in c2.U],
                    cart=[np.full(24 + 2 * G, 1 / dx), np.full(20 + 2 * G, 1 / dy), np.full(12 + 2 * G, 1 / dz),
                          np.full(24 + 2 * G, 1 / dx), np.full(20 + 2 * G, 1 / dy), np.full(12 + 2 * G, 1 / dz)])
    A, B = c2.copy_state(), cc.copy_state()
    Oracle(c2).step(A, 3)
    Oracle(cc).step(B, 3)
    assert relerr(A, B) < 1e-13


@pytest.mark.parametrize("filt,edoh", [(1, False), (2, False), (2, True)])
def test_c2_equals_c3_on_an_extruded_grid(filt, edoh):
    c2 = cases.vortex_wavy_c2(24, 20, 12, filter=filt)
    _add_z_modes(c2)
    c2.sw_edoh = edoh
    if edoh:
        c2.dt /= 8.0
    c3 = _c3_twin(c2)
    A, B = c2.copy_state(), c3.copy_state()
    Oracle(c2).step(A, 3)
    Oracle(c3).step(B, 3)
    e = relerr(A, B)
    moved = relerr(A, c2.U)
    assert moved > 1e-6
    assert e < 1e-11
