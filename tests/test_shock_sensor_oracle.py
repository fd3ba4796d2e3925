"""Pins of the oracle's shock-capturing sensor (filter_o10_shock,
mod_filtering_shock.f90:223-260 and the interior faces :419-440).

Each pin reads one scalar of the sensor chain off the oracle's filter
increment and compares it with a hand calculation on a field where the
chain reduces to a closed form:

* pressure sensor psi (:246-256, csens = 0.3, so that csens and 1 - csens
  cannot be exchanged unnoticed) on a pressure step and on the kinks of a
  triangle wave; psi = 0 on the linear ramps in between.  With chi = 0 the
  10th-order term is clamped to zero (xk10 = max(0, 0 + xk0/105), xk0 <= 0),
  so the increment of rho*E next to a jump is dissip_shock * psi * Delta(rho E);
* the interior face amplitude xk10 = max(0, xnu10 + xk0/105) (:424-425) with
  chi > 0, unclamped and clamped: the filtered field is a spike placed at the
  far end of the 10-point stencil of exactly one face with an active sensor,
  so the sensor changes the increment by (xk10 - chi/1260) * spike on the
  two points of that face alone;
* the Ducros sensor theta = div^2/(div^2 + |curl u|^2 + 1e-16) (:223-239) of
  the 5-point gradients, on velocity fields with a non-zero divergence where
  each vorticity component enters on its own (theta = 1 under pure
  dilatation).

The sensor is computed from the reference state U^n and applied to U_n, as
in the reference; orc_filter_rhs takes both.
"""
import math

import numpy as np
import pytest

from musica2_newgrid_amd import cases
from musica2_newgrid_amd import grid as gr
from oracle.oracle import Oracle

G = gr.G
GAM = cases.GAM
RHO0 = 1.177
L = 2 * math.pi


def _coords(n):
    return np.arange(n) * (L / n)


def box(fields, nx, ny, nz, chi, ds, csens=0.3, ducros=False):
    c = cases.cartesian_box(nx, ny, nz, (L, L, L), fields, iorder_visc=4, filter=2, chi=chi)
    c.ducros = ducros
    c.dissip_shock = ds
    c.csens = csens
    return c


def uniform_state(nx, ny, nz, rho=RHO0, e=2.5e5):
    """U_n of a uniform gas: every difference vanishes, so the filter adds nothing."""
    out = [gr.padded(nx, ny, nz, v) for v in (rho, 0.0, 0.0, 0.0, e)]
    return out


def step_pressure(nx, ny, nz, m0, pL, pR):
    """p = pL on interior points 0..m0, pR on m0+1..nx-1 (x = axis 2)."""
    p = np.full((nz, ny, nx), pL)
    p[:, :, m0 + 1:] = pR
    return p


def psi_step(pL, pR, cs):
    """The pressure sensor on the point of a step that sits on the pL side:
    |pL - 2 pL + pR| / (cs (pL + 2 pL + pR) + (1 - cs)(|pR - pL| + 0))."""
    d = abs(pR - pL)
    return d / (cs * (3 * pL + pR) + (1 - cs) * d)


@pytest.mark.parametrize("pL,pR", [(1.0e5, 1.2e5), (1.3e5, 0.9e5)])
def test_pressure_sensor_on_a_step(pL, pR):
    """chi = 0, gas at rest: K(rho E) = +-dissip_shock*max(psi_m0, psi_m0+1)*Delta(rho E)
    on the two points of each jump and exactly 0 elsewhere; rho and momentum
    get nothing."""
    nx, ny, nz, m0 = 32, 6, 6, 15
    cs, ds = 0.3, 0.7
    p = step_pressure(nx, ny, nz, m0, pL, pR)
    z = np.zeros_like(p)
    c = box((np.full_like(p, RHO0), z, z, z, p), nx, ny, nz, chi=0.0, ds=ds, csens=cs)
    K = Oracle(c).filter_rhs(c.U)
    KE = gr.interior(K[4])
    dE = (pR - pL) / (GAM - 1.0)
    # the psi of both points of the step: the pL side and the pR side
    psi = max(psi_step(pL, pR, cs), psi_step(pR, pL, cs))
    expect = ds * psi * dE
    assert np.allclose(KE[:, :, m0], expect, rtol=1e-12, atol=0)
    assert np.allclose(KE[:, :, m0 + 1], -expect, rtol=1e-12, atol=0)
    # the periodic wrap: the jump back from pR to pL between points nx-1 and 0
    assert np.allclose(KE[:, :, nx - 1], -expect, rtol=1e-12, atol=0)
    assert np.allclose(KE[:, :, 0], expect, rtol=1e-12, atol=0)
    rest = np.delete(KE, [0, m0, m0 + 1, nx - 1], axis=2)
    assert np.max(np.abs(rest)) <= 1e-12 * abs(expect)
    for q in range(4):
        assert np.max(np.abs(gr.interior(K[q]))) <= 1e-12 * abs(expect)


def test_pressure_sensor_on_kinks_and_zero_on_ramps():
    """Triangle wave p = p0 + a*min(i, n - i): kinks at i = 0 and i = n/2,
    linear ramps between.  psi = 2a/(cs(4 p_k + (p_k-1 - p_k) + (p_k+1 - p_k)) +
    (1 - cs) 2a) at a kink of value p_k, 0 on the ramps, so K(rho E) is
    -ds psi_0 a/(gam-1) at i = 1, +ds psi_n/2 a/(gam-1) at n/2 + 1, the mirror
    images at n - 1 and n/2 - 1, twice the opposite at the kinks themselves,
    and 0 at every other point."""
    nx, ny, nz = 32, 6, 6
    cs, ds, p0, a = 0.3, 0.45, 1.0e5, 900.0
    i = np.arange(nx)
    tri = np.minimum(i, nx - i).astype(float)
    p = np.broadcast_to(p0 + a * tri, (nz, ny, nx)).copy()
    z = np.zeros_like(p)
    c = box((np.full_like(p, RHO0), z, z, z, p), nx, ny, nz, chi=0.0, ds=ds, csens=cs)
    KE = gr.interior(Oracle(c).filter_rhs(c.U)[4])
    # at the minimum (p0, neighbours p0 + a): |2a| / (cs (4 p0 + 2a) + (1 - cs) 2a)
    psi_lo = 2 * a / (cs * (4 * p0 + 2 * a) + (1 - cs) * 2 * a)
    pk = p0 + a * nx / 2
    psi_hi = 2 * a / (cs * (4 * pk - 2 * a) + (1 - cs) * 2 * a)
    e = a / (GAM - 1.0)
    h = nx // 2
    want = np.zeros(nx)
    want[1], want[nx - 1] = -ds * psi_lo * e, -ds * psi_lo * e
    want[h + 1], want[h - 1] = ds * psi_hi * e, ds * psi_hi * e
    # the kink point itself takes both faces: D_{k-1/2} - D_{k+1/2}
    want[0], want[h] = 2 * ds * psi_lo * e, -2 * ds * psi_hi * e
    got = KE[0, 0, :]
    assert np.allclose(KE, got[None, None, :], rtol=0, atol=1e-9)
    scale = ds * psi_lo * e
    assert np.max(np.abs(got - want)) <= 1e-11 * scale, (got, want)
    # csens = 0.5 gives a different value: the test sees the blend
    psi_half = 2 * a / (0.5 * (4 * p0 + 2 * a) + 0.5 * 2 * a)
    assert abs(psi_half - psi_lo) > 1e-3 * psi_lo


@pytest.mark.parametrize("ds,clamped", [(0.05, False), (1.0, True)])
def test_interior_face_amplitude_with_an_active_sensor(ds, clamped):
    """Pressure step between interior points m0 and m0+1 (sensor from U^n),
    filtered field U_n = uniform + eps at q = m0 + 6.  Of the faces with an
    active sensor only face m0+1 (between m0+1 and m0+2) has q in its stencil,
    as its far point m0+1+5, with d10(5) = -1 and no first-difference term,
    so the sensor changes K(rho) by +(xk10 - chi/1260) eps at m0+1 and
    -(xk10 - chi/1260) eps at m0+2 and nowhere else, with
    xk10 = max(0, chi/1260 - ds psi_{m0+1}/105)."""
    nx, ny, nz, m0 = 32, 6, 6, 15
    cs, chi = 0.3, 0.2
    pL, pR = 1.0e5, 1.2e5
    p = step_pressure(nx, ny, nz, m0, pL, pR)
    z = np.zeros_like(p)
    fields = (np.full_like(p, RHO0), z, z, z, p)
    on = box(fields, nx, ny, nz, chi=chi, ds=ds, csens=cs)
    off = box(fields, nx, ny, nz, chi=chi, ds=0.0, csens=cs)
    eps = 1e-3
    Un = uniform_state(nx, ny, nz)
    q = m0 + 6
    Un[0][:, :, G + q] += eps
    for a in Un:
        gr.fill_periodic(a)
    dK = gr.interior(Oracle(on).filter_rhs(Un, on.U)[0]) - gr.interior(Oracle(off).filter_rhs(Un, off.U)[0])
    psi = psi_step(pR, pL, cs)  # point m0+1: pR with neighbours pL, pR
    xk10 = max(0.0, chi / 1260.0 - ds * psi / 105.0)
    assert (xk10 == 0.0) == clamped
    got_plus = dK[:, :, m0 + 1] / eps
    got_minus = dK[:, :, m0 + 2] / eps
    assert np.allclose(got_plus, xk10 - chi / 1260.0, rtol=1e-9, atol=1e-18)
    assert np.allclose(got_minus, -(xk10 - chi / 1260.0), rtol=1e-9, atol=1e-18)
    rest = np.delete(dK, [m0 + 1, m0 + 2], axis=2)
    assert np.max(np.abs(rest)) <= 1e-9 * eps * chi / 1260.0


def _k5(h):
    """5-pt derivative symbol of sin on spacing h: (4/3 sin h - 1/6 sin 2h)/h."""
    return (4.0 / 3.0 * math.sin(h) - 1.0 / 6.0 * math.sin(2 * h)) / h


@pytest.mark.parametrize("a,b,c_,d,e", [
    (30.0, 0.0, 0.0, 0.0, 0.0),    # pure dilatation in y: theta = 1
    (20.0, 25.0, 0.0, 0.0, 0.0),   # dilatation in y and z: theta = 1
    (20.0, 10.0, 15.0, 0.0, 0.0),  # + omega_z = -du/dy
    (20.0, 10.0, 0.0, 15.0, 0.0),  # + omega_x = dw/dy
    (20.0, 10.0, 0.0, 0.0, 15.0),  # + omega_y = du/dz
    (12.0, 7.0, 9.0, 11.0, 13.0),  # all three
])
def test_ducros_sensor_with_divergence(a, b, c_, d, e):
    """u = c sin y + e sin z, v = a sin y, w = b sin z + d sin y over a
    pressure step along x: the 5-pt gradients are k5 cos() times the
    amplitudes, div = k5y a cos y + k5z b cos z, omega = (d k5y cos y,
    e k5z cos z, -c k5y cos y), and K(rho E) at the step is ds theta psi Delta(rho E)
    with theta = div^2/(div^2 + |omega|^2 + 1e-16)."""
    nx, ny, nz, m0 = 32, 16, 12, 15
    cs, ds = 0.3, 0.6
    pL, pR = 1.0e5, 1.15e5
    x, y, zc = _coords(nx), _coords(ny), _coords(nz)
    Z, Y, X = np.meshgrid(zc, y, x, indexing="ij")
    u = c_ * np.sin(Y) + e * np.sin(Z)
    v = a * np.sin(Y)
    w = b * np.sin(Z) + d * np.sin(Y)
    p = step_pressure(nx, ny, nz, m0, pL, pR)
    cs_ = box((np.full_like(p, RHO0), u, v, w, p), nx, ny, nz, chi=0.0, ds=ds, csens=cs, ducros=True)
    KE = gr.interior(Oracle(cs_).filter_rhs(cs_.U)[4])
    k5y, k5z = _k5(L / ny), _k5(L / nz)
    cy, cz = np.cos(y)[None, :, None], np.cos(zc)[:, None, None]
    div = a * k5y * cy + b * k5z * cz
    om2 = (d * k5y * cy) ** 2 + (e * k5z * cz) ** 2 + (c_ * k5y * cy) ** 2
    theta = div * div / (div * div + om2 + 1.0e-16)
    psi = max(psi_step(pL, pR, cs), psi_step(pR, pL, cs))
    expect = ds * theta[:, :, 0] * psi * (pR - pL) / (GAM - 1.0)
    got = KE[:, :, m0]
    scale = ds * psi * (pR - pL) / (GAM - 1.0)
    assert np.max(np.abs(got - expect)) <= 1e-9 * scale, np.max(np.abs(got - expect)) / scale
    if c_ == d == e == 0.0:
        # pure dilatation: theta = 1 wherever the divergence is not ~0
        big = np.abs(div[:, :, 0]) > 1.0
        assert np.allclose(got[big], scale, rtol=1e-12)
    else:
        # the vorticity lowers theta somewhere by a visible amount
        assert np.min(theta) < 0.9
