"""The oracle's shock filter next to no-slip j-walls (filter_o10_shock,
mod_filtering_shock.f90:724-1057 and :1128-1140), pinned without retyping
its face formulas:

* at rest with a flat pressure the sensor vanishes and the j-increment of a
  y-profile is the difference of the face dissipations, each of which is a
  binomial forward difference: d_{2h} applied to a face is (-1)^h times the
  (2h-1)-th difference over the 2h rows of its stencil (d2 = -1, d4 = (-3, 1)
  ...), times xnu_2h = 0, chi/12, chi/60, chi/280, chi/1260 for the faces
  1, 2, 3, 4, >= 5 rows from a wall; the wall rows get none;
* with a pressure kink on the row next to each wall the sensor lives on the
  two faces of that row only.  Its strength X is read off the row after
  (chi = 0 leaves one term there), and the row between the kink and the
  wall then fixes that the wall face carries X/10 and the next face X; with
  chi > 0 the amplitude of the reduced face is max(0, chi/12 + X), without
  the 1/105 of the interior faces, clamped once X is large;
* the sensor of the wall row reads p one row beyond the wall, and nothing
  further out is read (those ghost rows hold NaN)."""
import numpy as np
import pytest

from musica2_newgrid_amd import cases
from oracle import oracle as orc
from tests.test_wall_oracle import G, H, NX, NZ, _box, _fields_y, _rows

NY = 24
XNU = {1: 0.0, 2: 1.0 / 12.0, 3: 1.0 / 60.0, 4: 1.0 / 280.0, 5: 1.0 / 1260.0}


def _box_at_rest(rho, p_rows, chi, ds, ducros):
    """Walled box at rest, rho and p given on rows 0..ny+1 (the two ghost rows
    next to the walls included; rows further out are NaN)."""
    ny = len(rho) - 2
    c, _ = _box(NX, ny, NZ, H, _fields_y(NX, ny, NZ, rho[1:-1], 0.0, 0.0, 0.0, p_rows[1:-1]), filt=2)
    c.chi, c.dissip_shock, c.ducros = chi, ds, ducros
    for row, r, pp in ((G - 1, rho[0], p_rows[0]), (G + ny, rho[-1], p_rows[-1])):
        for a, val in zip(c.U, cases.conserved(r, 0.0, 0.0, 0.0, pp)):
            a[G:-G, row, :] = val
    return c


def _half(m, ny):
    return min(5, m, ny - m)


def _dd(r, m):
    """(-1)^h times the (2h-1)-th forward difference of the face m+1/2 of
    half-width h; r[j] is row j (r[0] the ghost row)."""
    h = _half(m, len(r) - 2)
    seg = r[m - h + 1:m + h + 1]
    return (-1) ** h * np.diff(seg, 2 * h - 1)[0]


def _ky(D, ny):
    """j-increment of rows 1..ny from face values D[m], m = 1..ny-1."""
    k = np.zeros(ny)
    for j in range(2, ny):
        k[j - 1] = D[j - 1] - D[j]
    return k


def _rho_profile(seed):
    return 1.0 + 0.3 * np.random.default_rng(seed).random(NY + 2)


def test_flat_pressure_face_differences():
    chi = 0.2
    rho = _rho_profile(1)
    c = _box_at_rest(rho, np.full(NY + 2, 1.0e5), chi, 0.1, True)
    K = orc.Oracle(c).filter_rhs(c.U)
    D = {m: chi * XNU[_half(m, NY)] * _dd(rho, m) for m in range(1, NY)}
    want = _ky(D, NY)
    got = _rows(K[0])
    assert np.max(np.abs(got - want)) < 1e-13 * np.max(np.abs(rho))
    assert got[0] == 0.0 and got[-1] == 0.0
    assert np.max(np.abs(want[1:5])) > 1e-4  # the closures act
    for q in range(1, 5):
        assert np.max(np.abs(K[q][G:-G, G:-G, G:-G])) == 0.0


def _kinked_p(s0, s1, s2):
    """p on rows 0..ny+1, linear with slope changes at rows 2 and ny-1 only."""
    j = np.arange(NY + 2, dtype=np.float64)
    p = 1.0e5 + s1 * (j - 2)
    p[j < 2] = 1.0e5 + s0 * (j[j < 2] - 2)
    top = 1.0e5 + s1 * (NY - 3)
    p[j > NY - 1] = top + s2 * (j[j > NY - 1] - (NY - 1))
    return p


def _kink_strengths(rho, p, ds):
    """X = xnu0 * psens of the kinked rows, read from the chi = 0 increment:
    there K(3) = D(2) = drho(2) X_lo and K(ny-2) = -D(ny-2) = -drho(ny-2) X_hi."""
    c = _box_at_rest(rho, p, 0.0, ds, False)
    k = _rows(orc.Oracle(c).filter_rhs(c.U)[0])
    x_lo = k[2] / (rho[3] - rho[2])
    x_hi = -k[NY - 3] / (rho[NY - 1] - rho[NY - 2])
    return k, x_lo, x_hi


def test_wall_face_carries_a_tenth_of_the_sensor():
    rho = _rho_profile(2)
    p = _kinked_p(-900.0, 1300.0, -700.0)
    k, x_lo, x_hi = _kink_strengths(rho, p, 0.5)
    assert x_lo < -1e-4 and x_hi < -1e-4
    d = np.diff(rho)  # d[m] = rho(m+1) - rho(m)
    tol = 1e-12 * abs(x_lo) * np.max(np.abs(d))
    assert abs(k[1] - (d[1] * x_lo / 10.0 - d[2] * x_lo)) < tol
    assert abs(k[NY - 2] - (d[NY - 2] * x_hi - d[NY - 1] * x_hi / 10.0)) < tol
    assert np.max(np.abs(k[3:NY - 3])) == 0.0
    assert k[0] == 0.0 and k[-1] == 0.0


@pytest.mark.parametrize("ds", [0.5, 20.0])
def test_reduced_faces_with_the_sensor(ds):
    """chi > 0: D = drho xk0 + dd max(0, xnu + xk0) on the reduced faces; the
    larger ds clamps the d4 face next to each kink."""
    chi = 0.2
    rho = _rho_profile(3)
    p = _kinked_p(-900.0, 1300.0, -700.0)
    _, x_lo, x_hi = _kink_strengths(rho, p, ds)
    xk0 = {1: x_lo / 10.0, 2: x_lo, NY - 2: x_hi, NY - 1: x_hi / 10.0}
    clamped = 0
    D = {}
    for m in range(1, NY):
        h = _half(m, NY)
        x0 = xk0.get(m, 0.0)
        amp = chi * XNU[h] + (x0 / 105.0 if h == 5 else x0)
        clamped += amp < 0.0 and h == 2
        D[m] = (rho[m + 1] - rho[m]) * x0 + _dd(rho, m) * max(0.0, amp)
    c = _box_at_rest(rho, p, chi, ds, False)
    got = _rows(orc.Oracle(c).filter_rhs(c.U)[0])
    want = _ky(D, NY)
    assert np.max(np.abs(got - want)) < 1e-12 * np.max(np.abs(want))
    assert clamped == (2 if ds > 1.0 else 0)


def test_wall_row_sensor_reads_p_beyond_the_wall():
    rho = _rho_profile(4)
    p = _kinked_p(-900.0, 1300.0, -700.0)
    c = _box_at_rest(rho, p, 0.2, 0.5, False)
    k0 = _rows(orc.Oracle(c).filter_rhs(c.U)[0])
    p2 = p.copy()
    p2[0] += 5.0e3  # a sensor on row 1 stronger than the kink's
    c = _box_at_rest(rho, p2, 0.2, 0.5, False)
    k1 = _rows(orc.Oracle(c).filter_rhs(c.U)[0])
    assert np.all(np.isfinite(k1))
    assert abs(k1[1] - k0[1]) > 1e-6 * abs(k0[1])  # face 1 takes max(psens(1), psens(2))
    assert np.array_equal(k1[2:], k0[2:])
