"""The shock filter's Edoh switch in the oracle (is_sw_edoh,
mod_filtering_shock.f90:620-700, :1059-1140, :1450-1496): each direction's
increment is scaled by min(1, CFL/chi) with the local CFL number of U^n
along that direction.  Pinned by ratios of two oracle runs:

* a uniform stream carrying a small transverse-momentum mode along one axis
  (flat pressure, no Ducros sensor: the plain increment is the 10th-order
  dissipation): K_edoh = min(1, (|u_a| + c) dt / (h chi)) K_plain, for each
  axis a and both signs of u_a; a large dt clamps the factor at 1;
* a c3 block with constant, rotated and stretched metrics (both signs of the
  Jacobian) and the mode along i: the factor uses the contravariant velocity
  |(u xi_x + v xi_y + w xi_z) ijacob3| and c |grad xi| |ijacob3|
  (g3_ksi, smod_grid_metrics3_c3.f90:161);
* an undistorted c3 block gives the Cartesian result."""
import math

import numpy as np
import pytest

from musica2_newgrid_amd import cases
from musica2_newgrid_amd import grid as gr
from oracle import oracle as orc

N = 20
L = 2.0 * math.pi
H = L / N
RHO0, P0 = 1.1, 9.0e4
C0 = math.sqrt(cases.GAM * P0 / RHO0)
CHI = 0.2


def _stream(axis, vel, comp, m=3, amp=0.5):
    """Uniform stream `vel` with amp*cos(m x_axis) added to velocity component
    `comp` (transverse to `axis`)."""
    x = np.arange(N) * H
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    mode = amp * np.cos(m * (X, Y, Z)[axis])
    vs = [np.full_like(X, vel[d]) + (mode if d == comp else 0.0) for d in range(3)]
    fields = [np.full_like(X, RHO0), *vs, np.full_like(X, P0)]
    c = cases.cartesian_box(N, N, N, (L, L, L), fields, iorder_visc=4, filter=2, chi=CHI)
    c.ducros = False
    return c


def _ratio(c):
    c.sw_edoh = False
    K0 = orc.Oracle(c).filter_rhs(c.U)
    c.sw_edoh = True
    K1 = orc.Oracle(c).filter_rhs(c.U)
    return [gr.interior(a) for a in K0], [gr.interior(b) for b in K1]


@pytest.mark.parametrize("axis", [0, 1, 2])
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_cartesian_factor_per_direction(axis, sign):
    vel = [0.0, 0.0, 0.0]
    vel[axis] = sign * 40.0
    vel[(axis + 1) % 3] = 25.0  # a cross-stream component must not enter
    comp = (axis + 2) % 3
    c = _stream(axis, vel, comp)
    c.dt = 0.3 * CHI * H / (abs(vel[axis]) + C0)
    K0, K1 = _ratio(c)
    mul = min(1.0, (abs(vel[axis]) + C0) * c.dt / H / CHI)
    assert 0.25 < mul < 0.35
    scale = max(float(np.max(np.abs(k))) for k in K0)
    assert scale > 0.0
    for a, b in zip(K0, K1):
        assert np.max(np.abs(b - mul * a)) < 1e-12 * scale


def test_large_dt_clamps_the_factor_at_one():
    c = _stream(0, [40.0, 0.0, 0.0], 2)
    c.dt = 5.0 * CHI * H / C0
    K0, K1 = _ratio(c)
    for a, b in zip(K0, K1):
        assert np.array_equal(a, b)


def _c3_box(metric, ij, vel):
    """Stream `vel` with a w-mode along i on a c3 block of constant metrics
    (ksi row `metric`, ksi_z = 0 so the mode leaves u.grad(ksi) unchanged)."""
    c = _stream(0, vel, 2)
    c.curvilinear = True
    vals = [metric[0], metric[1], 0.0, -0.3, 1.7, 0.2, 0.1, 0.4, 2.1, ij]
    c.m = [gr.padded(N, N, N, float(v)) for v in vals]
    c.mv = [gr.padded(N, N, N, float(v)) for v in vals]
    return c


@pytest.mark.parametrize("ij", [0.8, -0.8])
@pytest.mark.parametrize("vel", [(40.0, -15.0, 5.0), (-40.0, 30.0, 0.0)])
def test_c3_factor_uses_the_contravariant_velocity(ij, vel):
    kx, ky = 1.3, -0.6
    c = _c3_box((kx, ky), ij, vel)
    vc = (vel[0] * kx + vel[1] * ky) * ij
    g3 = math.hypot(kx, ky) * abs(ij)
    c.dt = 0.3 * CHI / (abs(vc) + C0 * g3)
    K0, K1 = _ratio(c)
    mul = min(1.0, (abs(vc) + C0 * g3) * c.dt / CHI)
    assert 0.25 < mul < 0.35
    scale = max(float(np.max(np.abs(k))) for k in K0)
    for a, b in zip(K0, K1):
        assert np.max(np.abs(b - mul * a)) < 1e-12 * scale


def test_undistorted_c3_equals_cartesian():
    vel = (35.0, -20.0, 10.0)
    cc = _stream(1, vel, 0)
    cc.dt = 0.4 * CHI * H / (40.0 + C0)
    cc.sw_edoh = True
    c3 = _stream(1, vel, 0)
    c3.dt, c3.sw_edoh, c3.curvilinear = cc.dt, True, True
    c3.m, c3.mv = cases.affine_c3_metrics(N, N, N, (L, L, L))
    A = orc.Oracle(cc).filter_rhs(cc.U)
    B = orc.Oracle(c3).filter_rhs(c3.U)
    scale = max(float(np.max(np.abs(gr.interior(a)))) for a in A)
    for a, b in zip(A, B):
        assert np.max(np.abs(gr.interior(a) - gr.interior(b))) < 1e-12 * scale
