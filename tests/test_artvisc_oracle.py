"""Artificial viscosity of the oracle (artvisc_o9, Cartesian), pinned by
closed forms: on a uniform stream with a small transverse-momentum Fourier
mode along x, K = -(|u0| + c0) idx (chi/1260) 2^10 sin^10(theta/2) times the
mode (the 10th difference of the 9th-order face dissipation, weighted by the
spectral radius; |u0| is checked with both signs of u0); constants are left
untouched; the increments of a smooth periodic field sum to zero (the face
fluxes telescope only if the face weights of both neighbours agree); and a
step adds deltat*K (update_varn_artvisc), not K."""
import math

import numpy as np
import pytest

from musica2_newgrid_amd import cases
from musica2_newgrid_amd import grid as gr
from oracle import oracle as orc

N = 24


def _stream(u0, mode_amp, m, chi=0.2):
    L = 2.0 * math.pi
    x = np.arange(N) * (L / N)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    rho0, p0 = 1.2, 1.0e5
    w = mode_amp * np.cos(m * X)
    fields = [np.full_like(X, rho0), np.full_like(X, u0), np.zeros_like(X), w, np.full_like(X, p0)]
    c = cases.cartesian_box(N, N, N, (L, L, L), fields, iorder_visc=4, filter=3, chi=chi)
    return c, rho0, p0, L / N


@pytest.mark.parametrize("u0", [50.0, -50.0])
@pytest.mark.parametrize("m", [3, 7, 12])
def test_artvisc_symbol(u0, m):
    c, rho0, p0, dx = _stream(u0, 1e-3, m)
    K = orc.Oracle(c).filter_rhs(c.U)
    theta = 2.0 * math.pi * m / N
    c0 = math.sqrt(cases.GAM * p0 / rho0)
    lam = -(abs(u0) + c0) / dx * (c.chi / 1260.0) * 2.0 ** 10 * math.sin(theta / 2.0) ** 10
    mode = gr.interior(c.U[3])
    got = gr.interior(K[3])
    assert np.max(np.abs(got - lam * mode)) < 1e-9 * abs(lam) * np.max(np.abs(mode))
    for v in (0, 1, 2):
        assert np.max(np.abs(gr.interior(K[v]))) == 0.0


def test_artvisc_leaves_constants_and_conserves():
    c, _, _, _ = _stream(30.0, 0.0, 1)
    K = orc.Oracle(c).filter_rhs(c.U)
    assert all(np.max(np.abs(gr.interior(k))) == 0.0 for k in K)
    t = cases.tgv(N)
    t.filter = 3
    K = orc.Oracle(t).filter_rhs(t.U)
    for k, u in zip(K, t.U):
        ki = gr.interior(k)
        assert abs(float(np.sum(ki))) <= 1e-11 * float(np.sum(np.abs(ki)))


def test_artvisc_step_adds_deltat_times_K():
    """One step of a uniform stream with a mode: Euler and viscous terms of a
    uniform stream with a transverse mode along x are not zero, so compare
    a filter-on step with a filter-off step: their difference is
    deltat*K(U_after_stages) to first order."""
    c, _, _, _ = _stream(50.0, 1e-3, 7)
    c.dt = 1e-6
    A = c.copy_state()
    orc.Oracle(c).step(A, 1)
    c.filter = 0
    B = c.copy_state()
    orc.Oracle(c).step(B, 1)
    c.filter = 3
    K = orc.Oracle(c).filter_rhs(B, c.U)
    d = gr.interior(A[3]) - gr.interior(B[3])
    assert np.max(np.abs(d - c.dt * gr.interior(K[3]))) < 1e-12 * np.max(np.abs(gr.interior(A[3])))
    assert np.max(np.abs(d)) > 1e-3 * c.dt * np.max(np.abs(gr.interior(K[3])))
