"""The oracle's maximum CFL numbers (check_max_cfl,
src/Time/timestep_new.f90:660-842), pinned by closed forms:

* a uniform stream on a Cartesian block with a non-uniform spacing along
  each axis: (|u_d| + c) deltat max(idx_d), c = sqrt(gam p / rho), for both
  signs of the velocity (a missing abs() or a wrong component fails);
* a c3 block with constant rotated and stretched metrics, both signs of the
  Jacobian: (|(u xi_x + v xi_y + w xi_z) ijacob3| + c |grad xi| |ijacob3|)
  deltat per direction (the g3 factor and the abs of a negative Jacobian);
* an undistorted c3 block gives the Cartesian numbers;
* a non-uniform stream: the maximum sits where |u| + c is largest.
"""
import math

import numpy as np
import pytest

from musica2_newgrid_amd import cases
from musica2_newgrid_amd import grid as gr
from oracle import oracle as orc

N = 12
RHO0, P0 = 1.2, 1.0e5
C0 = math.sqrt(cases.GAM * P0 / RHO0)


def _stream(vel, n=N, lengths=(1.0, 2.0, 3.0)):
    shape = (n, n, n)
    fields = [np.full(shape, RHO0)] + [np.full(shape, v) for v in vel] + [np.full(shape, P0)]
    c = cases.cartesian_box(n, n, n, lengths, fields, iorder_visc=4)
    c.dt = 1.0e-5
    return c


@pytest.mark.parametrize("vel", [(80.0, -30.0, 10.0), (-80.0, 30.0, -10.0)])
def test_cartesian_uniform_stream(vel):
    c = _stream(vel)
    rng = np.random.default_rng(3)
    for q in range(3):  # non-uniform Euler metrics (the viscous set is not used)
        c.cart[q] = c.cart[q] * rng.uniform(0.5, 1.5, c.cart[q].size)
    got = orc.Oracle(c).max_cfl(c.U)
    for d in range(3):
        want = (abs(vel[d]) + C0) * c.dt * float(np.max(c.cart[d][gr.G:-gr.G]))
        assert abs(got[d] - want) <= 1e-14 * want, (d, got[d], want)


@pytest.mark.parametrize("ij", [0.8, -0.8])
def test_c3_rotated_metrics(ij):
    vel = (60.0, -25.0, 15.0)
    c = _stream(vel)
    c.curvilinear = True
    rows = [(1.3, -0.6, 0.2), (-0.3, 1.7, 0.4), (0.1, 0.4, 2.1)]
    vals = [x for r in rows for x in r] + [ij]
    c.m = [gr.padded(N, N, N, float(v)) for v in vals]
    c.mv = [gr.padded(N, N, N, float(v)) for v in vals]
    got = orc.Oracle(c).max_cfl(c.U)
    for d, (mx, my, mz) in enumerate(rows):
        vc = (vel[0] * mx + vel[1] * my + vel[2] * mz) * ij
        g3 = math.sqrt(mx * mx + my * my + mz * mz) * abs(ij)
        want = (abs(vc) + C0 * g3) * c.dt
        assert abs(got[d] - want) <= 1e-13 * want, (d, got[d], want)


def test_undistorted_c3_equals_cartesian():
    vel = (35.0, -20.0, 10.0)
    L = (1.0, 2.0, 3.0)
    cc = _stream(vel, lengths=L)
    c3 = _stream(vel, lengths=L)
    c3.curvilinear = True
    c3.m, c3.mv = cases.affine_c3_metrics(N, N, N, L)
    a, b = orc.Oracle(cc).max_cfl(cc.U), orc.Oracle(c3).max_cfl(c3.U)
    assert np.allclose(a, b, rtol=1e-13, atol=0)


def test_maximum_where_the_speed_peaks():
    c = cases.tgv(16)
    got = orc.Oracle(c).max_cfl(c.U)
    rho = gr.interior(c.U[0])
    u, v, w = (gr.interior(c.U[q]) / rho for q in (1, 2, 3))
    p = (cases.GAM - 1.0) * (gr.interior(c.U[4]) - 0.5 * rho * (u * u + v * v + w * w))
    cs = np.sqrt(cases.GAM * p / rho)
    h = c.info["dx"]
    for d, vel in enumerate((u, v, w)):
        want = float(np.max((np.abs(vel) + cs) * c.dt / h))
        assert abs(got[d] - want) <= 1e-13 * want
