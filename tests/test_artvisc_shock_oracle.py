"""The oracle's artificial viscosity with shock capturing (artvisc_o9_shock,
mod_artvisc_shock.f90:113-1287, filter 4) and the 2.5-D form of artvisc_o9
(filter 3), pinned against the operators they share their parts with:

* without the sensor term (dissip_shock = 0) the shock faces reduce to the
  9th-order faces of artvisc_o9 (xk10 = chi/1260, the d10 of the shock
  filter being those of artvisc_o9 with the opposite sign), so filter 4 gives
  the step of filter 3, which the closed-form symbol pins;
* where the spectral radius is one constant c (fluid at rest at one
  temperature, density varying) on a uniform grid of spacing h, each face of
  filter 4 is the shock filter's face times c/h: K4 = (c/h) K2 point by point,
  the sensor included, for the shock filter pinned on its own;
* next to no-slip walls the same holds wherever the sensor clips the reduced
  faces 3 and 4 to the sensor term alone (the two operators scale it by 1, 1
  and by 1/5, 3/70 there), which a sawtooth density along y makes true at
  every wall face: the step's increment of filter 4 is deltat c/h times that
  of filter 2 on rows 2..ny-1, the wall rows included in no j-increment;
* an undistorted 2.5-D block (anisotropic spacings) gives the Cartesian
  block's step for filters 3 and 4; on a z-extruded wavy grid the 2.5-D block
  and the c3 block of the same grid give the same step for filter 4.

Parity unpinned: the constants 1/5 and 3/70 themselves (wall faces 3 and 4
of filter 4 with an unclipped sensor term) are transcribed from the
reference; no closed form tells them apart."""
import math

import numpy as np
import pytest

from musica2_newgrid_amd import cases
from musica2_newgrid_amd import grid as gr
from oracle import oracle as orc
from oracle.oracle import Oracle
from tests._util import relerr
from tests.test_c2_oracle import _add_z_modes, _c3_twin

G = gr.G


def test_without_sensor_term_equals_artvisc():
    t3 = cases.tgv(16)
    t3.filter = 3
    t4 = cases.tgv(16)
    t4.filter = 4
    t4.dissip_shock = 0.0
    A, B = t3.copy_state(), t4.copy_state()
    Oracle(t3).step(A, 2)
    Oracle(t4).step(B, 2)
    assert relerr(A, t3.U) > 1e-6
    assert relerr(A, B) < 1e-13


def _rest_box(n=16, amp=0.05, T=300.0):
    """Fluid at rest at one temperature on a uniform periodic grid, density
    (hence pressure) varying in all three directions."""
    L = 2.0 * math.pi
    x = np.arange(n) * (L / n)
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    rho = 1.2 * (1.0 + amp * np.sin(X + 0.3) * np.cos(2.0 * Y) + amp * np.sin(3.0 * Z) +
                 0.5 * amp * np.tanh(4.0 * np.sin(X)))
    p = rho * cases.RG * T
    zero = np.zeros_like(X)
    c = cases.cartesian_box(n, n, n, (L, L, L), [rho, zero, zero, zero, p], iorder_visc=4, filter=2)
    c.ducros = False
    return c, math.sqrt(cases.GAM * cases.RG * T), L / n


def test_constant_spectral_radius_scales_the_shock_filter():
    c, cs, h = _rest_box()
    K2 = Oracle(c).filter_rhs(c.U)
    c.filter = 4
    K4 = Oracle(c).filter_rhs(c.U)
    for v in (0, 4):
        a, b = gr.interior(K4[v]), (cs / h) * gr.interior(K2[v])
        assert np.max(np.abs(b)) > 0.0
        assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(b))


def _sawtooth_wall_box(nx=8, ny=24, nz=8, h=0.05, T=300.0):
    """Uniform box with walls on both y-faces, fluid at rest at one
    temperature, density a sawtooth along y (with a weak x modulation): the
    pressure sensor is about 0.1 on every row, ghosts beyond the walls
    included."""
    ye = (np.arange(ny + 2 * G) - G) * h
    ide, idv, c2, c3 = gr.wall_metrics(ye, True, True)
    idx = np.full(nx + 2 * G, 1.0 / h)
    idz = np.full(nz + 2 * G, 1.0 / h)
    jj = np.arange(ny + 2 * G).reshape(1, -1, 1)
    ii = np.arange(nx + 2 * G).reshape(1, 1, -1)
    rho = 1.2 * (1.0 + 0.1 * (-1.0) ** jj) * (1.0 + 0.01 * np.sin(2.0 * math.pi * ii / nx))
    rho = np.broadcast_to(rho, (nz + 2 * G, ny + 2 * G, nx + 2 * G)).copy()
    p = rho * cases.RG * T
    zero = np.zeros_like(rho)
    U = [np.ascontiguousarray(a) for a in cases.conserved(rho, zero, zero, zero, p)]
    c = cases.Case("sawtooth", nx, ny, nz, curvilinear=False, iorder_visc=4, filter=2, ducros=False,
                   cart=[idx, ide, idz, idx.copy(), idv, idz.copy()], U=U, dt=1e-7,
                   faces=(0, 0, 2, 2, 0, 0),
                   cextrp2=(0, 0, c2[0], c2[1], 0, 0), cextrp3=(0, 0, c3[0], c3[1], 0, 0))
    c.dissip_shock = 1.0
    return c, math.sqrt(cases.GAM * cases.RG * T), h


def test_walls_with_a_clipping_sensor_scale_the_shock_filter():
    c, cs, h = _sawtooth_wall_box()
    out = {}
    for f in (0, 2, 4):
        c.filter = f
        U = c.copy_state()
        Oracle(c).step(U, 1)
        out[f] = U
    rows = slice(1, c.ny - 1)   # rows 2..ny-1 (the wall rows take the wall state)
    for v in (0, 1, 2, 4):
        d2 = gr.interior(out[2][v])[:, rows, :] - gr.interior(out[0][v])[:, rows, :]
        d4 = gr.interior(out[4][v])[:, rows, :] - gr.interior(out[0][v])[:, rows, :]
        ref = c.dt * cs / h * d2
        scale = np.max(np.abs(ref))
        if v != 2:
            assert scale > 0.0
        assert np.max(np.abs(d4 - ref)) <= 1e-9 * max(scale, 1e-300) + 1e-15 * np.max(np.abs(gr.interior(out[0][v])))


@pytest.mark.parametrize("filt", [3, 4])
def test_undistorted_c2_equals_cartesian(filt):
    h = 10.0 / 24
    c2 = cases.vortex_wavy_c2(24, 20, 12, amp=(0.0, 0.0), filter=filt, lengths=(24 * h, 20 * 1.3 * h, 12 * 0.8 * h))
    _add_z_modes(c2)
    L = c2.info["lengths"]
    dx, dy, dz = L[0] / 24, L[1] / 20, L[2] / 12
    assert len({round(dx, 12), round(dy, 12), round(dz, 12)}) == 3
    cc = cases.Case("cart", 24, 20, 12, curvilinear=False, iorder_visc=4, filter=filt, ducros=c2.ducros,
                    diffscale=c2.diffscale, dt=c2.dt, chi=c2.chi, U=[a.copy() for a in c2.U],
                    cart=[np.full(24 + 2 * G, 1 / dx), np.full(20 + 2 * G, 1 / dy), np.full(12 + 2 * G, 1 / dz),
                          np.full(24 + 2 * G, 1 / dx), np.full(20 + 2 * G, 1 / dy), np.full(12 + 2 * G, 1 / dz)])
    A, B = c2.copy_state(), cc.copy_state()
    Oracle(c2).step(A, 3)
    Oracle(cc).step(B, 3)
    assert relerr(A, c2.U) > 1e-6
    assert relerr(A, B) < 1e-13


def test_c2_equals_c3_on_an_extruded_grid():
    c2 = cases.vortex_wavy_c2(24, 20, 12, filter=4)
    _add_z_modes(c2)
    c3 = _c3_twin(c2)
    A, B = c2.copy_state(), c3.copy_state()
    Oracle(c2).step(A, 3)
    Oracle(c3).step(B, 3)
    assert relerr(A, c2.U) > 1e-6
    assert relerr(A, B) < 1e-11


def test_filter_codes_are_checked():
    c = cases.tgv(16)
    c.filter = 5
    with pytest.raises(RuntimeError):
        Oracle(c).step(c.copy_state(), 1)
    c = cases.vortex_wavy(16, 16, 16, filter=3)
    with pytest.raises(RuntimeError):
        Oracle(c).step(c.copy_state(), 1)


def test_artvisc_at_walls_scales_the_shock_filter_without_sensor():
    """artvisc_o9 next to no-slip walls (faces 1..4 from a wall with d2..d8
    times xnu2..xnu8, nothing from the wall axis on the wall rows,
    mod_artvisc.f90:98-586): with one spectral radius c on a uniform grid its
    step increment is deltat c/h times that of the shock filter without its
    sensor term, whose wall faces are pinned on their own; and it equals
    artvisc_o9_shock without the sensor term."""
    c, cs, h = _sawtooth_wall_box()
    c.dissip_shock = 0.0
    out = {}
    for f in (0, 2, 3, 4):
        c.filter = f
        U = c.copy_state()
        Oracle(c).step(U, 1)
        out[f] = U
    rows = slice(1, c.ny - 1)
    for v in (0, 1, 2, 4):
        base = gr.interior(out[0][v])[:, rows, :]
        d2 = gr.interior(out[2][v])[:, rows, :] - base
        d3 = gr.interior(out[3][v])[:, rows, :] - base
        d4 = gr.interior(out[4][v])[:, rows, :] - base
        ref = c.dt * cs / h * d2
        scale = max(float(np.max(np.abs(ref))), 1e-300)
        if v != 2:
            assert float(np.max(np.abs(ref))) > 0.0
        tol = 1e-9 * scale + 1e-15 * float(np.max(np.abs(base)))
        assert np.max(np.abs(d3 - ref)) <= tol
        assert np.max(np.abs(d3 - d4)) <= tol
