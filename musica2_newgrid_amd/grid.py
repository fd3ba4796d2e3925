"""Host-side set-up of block grids: padded array helpers and curvilinear metrics.

Set-up work, done once before the time loop, as in the reference's
grid_finalize (src/Grid/grid_finalize.f90:2-268).  It is not on the per-step
path: the library receives the metrics through mgpu_set_metrics_*.

Arrays are numpy arrays of shape (nz+2G, ny+2G, nx+2G), C order -- the same
bytes as the reference's Fortran arrays (1-G:nx+G, 1-G:ny+G, 1-G:nz+G).
"""
from __future__ import annotations

import numpy as np

G = 5

# 11-pt and 5-pt centred first-derivative coefficients
# (src/Derivatives/mod_coeff_deriv.f90:121-126, 316-325)
A11 = (5.0 / 6.0, -5.0 / 21.0, 5.0 / 84.0, -5.0 / 504.0, 1.0 / 1260.0)
A5 = (2.0 / 3.0, -1.0 / 12.0)


def padded(nx: int, ny: int, nz: int, fill: float = 0.0) -> np.ndarray:
    return np.full((nz + 2 * G, ny + 2 * G, nx + 2 * G), fill, dtype=np.float64)


def interior(a: np.ndarray) -> np.ndarray:
    return a[G:-G, G:-G, G:-G]


def fill_periodic(a: np.ndarray, axes=(0, 1, 2)) -> None:
    """Face ghosts from the periodic image: i-faces over interior (j,k), then
    j-faces over all i, then k-faces over all (i,j) -- mirroring the
    reference's self-neighbour halo (src/Parallel/mod_comm.f90:249-290)."""
    nz, ny, nx = (s - 2 * G for s in a.shape)
    if 2 in axes:  # i
        a[G:G + nz, G:G + ny, :G] = a[G:G + nz, G:G + ny, nx:nx + G]
        a[G:G + nz, G:G + ny, nx + G:] = a[G:G + nz, G:G + ny, G:2 * G]
    if 1 in axes:  # j
        a[G:G + nz, :G, :] = a[G:G + nz, ny:ny + G, :]
        a[G:G + nz, ny + G:, :] = a[G:G + nz, G:2 * G, :]
    if 0 in axes:  # k
        a[:G, :, :] = a[nz:nz + G, :, :]
        a[nz + G:, :, :] = a[G:2 * G, :, :]


def _deriv(f: np.ndarray, axis: int, coeffs) -> np.ndarray:
    """Centred derivative along axis; the result is shorter by 2R there."""
    R = len(coeffs)
    n = f.shape[axis]

    def sl(a, b):
        s = [slice(None)] * f.ndim
        s[axis] = slice(a, b)
        return f[tuple(s)]

    out = None
    for l, a in enumerate(coeffs, start=1):
        t = sl(R + l, n - R + l) - sl(R - l, n - R - l)
        t *= a
        if out is None:
            out = t
        else:
            out += t
    return out


def _crop(f: np.ndarray, e_from: int, e_to: int, axes=(0, 1, 2)) -> np.ndarray:
    """Reduce the extension of the given axes from e_from to e_to layers."""
    d = e_from - e_to
    sl = [slice(None)] * 3
    for ax in axes:
        sl[ax] = slice(d, f.shape[ax] - d)
    return f[tuple(sl)]


def gcl_metrics(xe: np.ndarray, ye: np.ndarray, ze: np.ndarray, ext: int, coeffs):
    """Conservative (Thomas & Lombard) metrics J*dksi/dx ... and the inverse
    Jacobian on the interior, from coordinates given on a box extended by
    `ext` layers (ext >= 2R).  Follows grid_metrics_3d
    (src/Grid/smod_grid_metrics1_c3.f90:152-230): ksi_x = (y_eta z)_phi -
    (y_phi z)_eta and cyclic, and grid_metrics_ijacob_3d
    (src/Grid/smod_grid_metrics3_c3.f90:29-35,131) for 1/J.

    Returns (m, ijac): m = [ksi_x,ksi_y,ksi_z,eta_x,eta_y,eta_z,phi_x,phi_y,phi_z]
    on the interior (shape (nz,ny,nx)), ijac the inverse Jacobian there."""
    R = len(coeffs)
    if ext > 2 * R:  # keep exactly the 2R layers the two derivative levels need
        xe, ye, ze = (_crop(a, ext, 2 * R) for a in (xe, ye, ze))
        ext = 2 * R
    e1 = ext - R  # extension kept after the first derivatives
    # axis 2 = i (ksi), 1 = j (eta), 0 = k (phi)

    def d1(f, axis):
        return _crop(_deriv(f, axis, coeffs), ext, e1, axes=[a for a in (0, 1, 2) if a != axis])

    xk, yk, zk = d1(xe, 2), d1(ye, 2), d1(ze, 2)
    xh, yh, zh = d1(xe, 1), d1(ye, 1), d1(ze, 1)
    xp, yp, zp = d1(xe, 0), d1(ye, 0), d1(ze, 0)
    xc, yc, zc = (_crop(a, ext, e1) for a in (xe, ye, ze))

    def d2(f, axis):
        # derivative along axis, then crop the other axes to the interior
        return _crop(_deriv(f, axis, coeffs), e1, 0, axes=[a for a in (0, 1, 2) if a != axis])

    # ksi_x=(y_eta z)_phi-(y_phi z)_eta ; ksi_y=(z_eta x)_phi-(z_phi x)_eta ; ksi_z=(x_eta y)_phi-(x_phi y)_eta
    ksi_x = d2(yh * zc, 0) - d2(yp * zc, 1)
    ksi_y = d2(zh * xc, 0) - d2(zp * xc, 1)
    ksi_z = d2(xh * yc, 0) - d2(xp * yc, 1)
    # eta_x=(y_phi z)_ksi-(y_ksi z)_phi ...
    eta_x = d2(yp * zc, 2) - d2(yk * zc, 0)
    eta_y = d2(zp * xc, 2) - d2(zk * xc, 0)
    eta_z = d2(xp * yc, 2) - d2(xk * yc, 0)
    # phi_x=(y_ksi z)_eta-(y_eta z)_ksi ...
    phi_x = d2(yk * zc, 1) - d2(yh * zc, 2)
    phi_y = d2(zk * xc, 1) - d2(zh * xc, 2)
    phi_z = d2(xk * yc, 1) - d2(xh * yc, 2)

    c = lambda f: _crop(f, e1, 0)  # noqa: E731
    XK, YK, ZK, XH, YH, ZH, XP, YP, ZP = (c(a) for a in (xk, yk, zk, xh, yh, zh, xp, yp, zp))
    jac = XK * (YH * ZP - YP * ZH) - XH * (YK * ZP - YP * ZK) + XP * (YK * ZH - YH * ZK)
    return [ksi_x, ksi_y, ksi_z, eta_x, eta_y, eta_z, phi_x, phi_y, phi_z], 1.0 / jac


def curvilinear_metrics_periodic(mapping, nx, ny, nz, lengths, viscous=True):
    """Metrics of a block periodic in all three directions.

    mapping(xi, eta, zeta) -> (x, y, z) maps computational coordinates
    (xi = (i-1)*Lx/nx, ...) to physical ones, with x - xi, y - eta, z - zeta
    periodic.  Coordinates are evaluated on a box extended by 10 layers without
    wrapping, metrics are computed on the interior and their ghosts are filled
    from the periodic image (as the reference's metric halo exchange does).

    Returns (m, mv): two lists of 10 padded arrays
    (ksi_x..phi_z, ijacob3), the Euler (11-pt) and viscous (5-pt) sets."""
    ext = 10
    Lx, Ly, Lz = lengths
    i = np.arange(1 - ext, nx + ext + 1, dtype=np.float64)
    j = np.arange(1 - ext, ny + ext + 1, dtype=np.float64)
    k = np.arange(1 - ext, nz + ext + 1, dtype=np.float64)
    XI = (i - 1) * (Lx / nx)
    ET = (j - 1) * (Ly / ny)
    ZE = (k - 1) * (Lz / nz)
    Z3, Y3, X3 = np.meshgrid(ZE, ET, XI, indexing="ij")
    xe, ye, ze = mapping(X3, Y3, Z3)
    out = []
    for coeffs in ((A11, A5) if viscous else (A11,)):
        m9, ij = gcl_metrics(xe, ye, ze, ext, coeffs)
        arrs = []
        for f in m9 + [ij]:
            a = padded(nx, ny, nz)
            a[G:-G, G:-G, G:-G] = f
            fill_periodic(a)
            arrs.append(a)
        out.append(arrs)
    if not viscous:
        out.append(None)
    coords = tuple(_crop(a, ext, 0) for a in (xe, ye, ze))
    return out[0], out[1], coords


def c2_metrics_periodic(mapping, nx, ny, lengths, viscous=True):
    """2-D metrics of a 2.5-D curvilinear block (is_curv), periodic in i and j.

    mapping(xi, eta) -> (x, y) with x - xi, y - eta periodic; z is Cartesian.
    x_ksi, x_eta, y_ksi, y_eta are the centred derivatives of the node
    coordinates (a11 for the Euler set, a5 for the viscous set) and ijacob =
    1/(x_ksi y_eta - x_eta y_ksi) (grid_metrics_ksi/_eta/_ijacob,
    src/Grid/grid_metrics_curv.f90:2-296, 595-630, and the _v twins), on
    the interior, ghosts from the periodic image.

    Returns (m2, m2v, (xc, yc)): two lists of 5 padded 2-D arrays of shape
    (ny+2G, nx+2G), i fastest: x_ksi, x_eta, y_ksi, y_eta, ijacob."""
    ext = 10
    Lx, Ly = lengths[0], lengths[1]
    i = np.arange(1 - ext, nx + ext + 1, dtype=np.float64)
    j = np.arange(1 - ext, ny + ext + 1, dtype=np.float64)
    Y2, X2 = np.meshgrid((j - 1) * (Ly / ny), (i - 1) * (Lx / nx), indexing="ij")
    xe, ye = mapping(X2, Y2)
    out = []
    for coeffs in ((A11, A5) if viscous else (A11,)):
        R = len(coeffs)

        def d(f, axis):
            g = _deriv(f, axis, coeffs)
            other = 1 - axis
            s = [slice(None), slice(None)]
            s[other] = slice(ext, f.shape[other] - ext)
            s[axis] = slice(ext - R, g.shape[axis] - (ext - R))
            return g[tuple(s)]

        xk, xh, yk, yh = d(xe, 1), d(xe, 0), d(ye, 1), d(ye, 0)
        ij = 1.0 / (xk * yh - xh * yk)
        arrs = []
        for f in (xk, xh, yk, yh, ij):
            a = np.zeros((ny + 2 * G, nx + 2 * G))
            a[G:-G, G:-G] = f
            a[G:-G, :G] = a[G:-G, nx:nx + G]
            a[G:-G, nx + G:] = a[G:-G, G:2 * G]
            a[:G, :] = a[ny:ny + G, :]
            a[ny + G:, :] = a[G:2 * G, :]
            arrs.append(a)
        out.append(arrs)
    if not viscous:
        out.append(None)
    return out[0], out[1], (xe[ext:-ext, ext:-ext], ye[ext:-ext, ext:-ext])


# Wall closures of the Cartesian metrics (src/Grid/grid_metrics_cart.f90:2-186,
# 189-436; src/Derivatives/mod_coeff_deriv.f90:100-160, 232-241, 1099-1160).
# a3, a5, a7, a9: centred schemes of half-width 1..4 used on rows 2..5 from a wall.
A_CLOSE = ((0.5,), A5, (3.0 / 4.0, -3.0 / 20.0, 1.0 / 60.0),
           (4.0 / 5.0, -1.0 / 5.0, 4.0 / 105.0, -1.0 / 280.0), A11)
A02 = (-3.0 / 2.0, 2.0, -1.0 / 2.0)
A04 = (-25.0 / 12.0, 4.0, -3.0, 4.0 / 3.0, -1.0 / 4.0)
A13 = (-1.0 / 4.0, -5.0 / 6.0, 3.0 / 2.0, -1.0 / 2.0, 1.0 / 12.0)


def curvilinear_metrics_jwalls(mapping, eta_of, nx, ny, nz, lengths, viscous=True):
    """Metrics of a block periodic in i and k with no-slip walls on its j-faces.

    mapping(xi, eta, zeta) -> (x, y, z), periodic in xi and zeta (periods Lx,
    Lz); eta_of(j) gives the computational coordinate of row j (rows 1 and ny
    are the walls) and must extend smoothly beyond them, so that the
    coordinates can be evaluated on a box extended by 10 layers and the metrics
    computed with the centred 11-pt (5-pt, viscous) stencils everywhere, ghosts
    beyond the walls included (the kernels never read metrics there).  The
    reference computes its metrics from the grid file alone, with one-sided
    stencils next to the walls (smod_grid_metrics1_c3.f90); this synthetic
    set-up uses the analytic map instead.  i- and k-ghosts take the periodic
    image.

    Returns (m, mv, coords, wall): m, mv the 10 padded Euler / viscous arrays
    (ksi_x..phi_z, ijacob3); coords the node coordinates on the interior;
    wall[f] for f = 2 (jmin), 3 (jmax) the 14 per-point coefficients of the
    wall routines (grid_normals.f90:243-300): -grad(ksi).grad(eta)/|grad(eta)|^2,
    -grad(phi).grad(eta)/|grad(eta)|^2 (Euler metrics of the wall row), the
    unit tangents x_ksi/|x_ksi|, x_phi/|x_phi| (11-pt derivatives of the
    coordinates) and their dual basis (dual_tangents, for slip walls), each an
    array (nz, nx)."""
    ext = 10
    Lx, Ly, Lz = lengths
    e2 = ext + G
    i = np.arange(1 - e2, nx + e2 + 1, dtype=np.float64)
    j = np.arange(1 - e2, ny + e2 + 1, dtype=np.float64)
    k = np.arange(1 - e2, nz + e2 + 1, dtype=np.float64)
    XI = (i - 1) * (Lx / nx)
    ET = eta_of(j)
    ZE = (k - 1) * (Lz / nz)
    Z3, Y3, X3 = np.meshgrid(ZE, ET, XI, indexing="ij")
    xe, ye, ze = mapping(X3, Y3, Z3)
    out = []
    for coeffs in ((A11, A5) if viscous else (A11,)):
        # coordinates extended by ext + G layers: the metrics of the padded box
        m9, ij = gcl_metrics(xe, ye, ze, ext, coeffs)
        arrs = []
        for f in m9 + [ij]:
            a = np.ascontiguousarray(f)
            fill_periodic(a, axes=(0, 2))
            arrs.append(a)
        out.append(arrs)
    if not viscous:
        out.append(None)
    m = out[0]
    # unit tangents along ksi and phi of the wall rows
    wall = {}
    for f, row in ((2, 1), (3, ny)):
        jr = G + row - 1 + ext  # row index in the extended box
        line = lambda a: a[:, jr, :]  # noqa: E731
        xk = [_deriv(line(a), 1, A11) for a in (xe, ye, ze)]  # (nz+2e2, nx+2(e2-5))
        xp = [_deriv(line(a), 0, A11) for a in (xe, ye, ze)]
        ck = [c[e2:e2 + nz, e2 - 5:e2 - 5 + nx] for c in xk]
        cp = [c[e2 - 5:e2 - 5 + nz, e2:e2 + nx] for c in xp]
        gk = np.sqrt(ck[0] ** 2 + ck[1] ** 2 + ck[2] ** 2)
        gp = np.sqrt(cp[0] ** 2 + cp[1] ** 2 + cp[2] ** 2)
        w = lambda q: m[q][G:G + nz, G + row - 1, G:G + nx]  # noqa: E731
        n2 = w(3) ** 2 + w(4) ** 2 + w(5) ** 2
        g1 = -(w(0) * w(3) + w(1) * w(4) + w(2) * w(5)) / n2
        g2 = -(w(6) * w(3) + w(7) * w(4) + w(8) * w(5)) / n2
        g1[np.abs(g1) < 1e-16] = 0.0
        g2[np.abs(g2) < 1e-16] = 0.0
        ek = [c / gk for c in ck]
        ep = [c / gp for c in cp]
        wall[f] = [np.ascontiguousarray(a) for a in [g1, g2] + ek + ep + dual_tangents(ek, ep)]
    coords = tuple(a[e2:e2 + nz, e2:e2 + ny, e2:e2 + nx] for a in (xe, ye, ze))
    return out[0], out[1], coords, wall


def dual_tangents(ek, ep):
    """The dual basis (t_ksi, t_phi) of the unit wall tangents ek, ep (3 arrays
    each) in their plane, as grid_normals.f90:279-291 builds it for the slip
    walls: with c = ek.ep, t_ksi = (ek - c ep)/(1 - c^2) and t_phi = (ep - c
    ek)/(1 - c^2), so that m.t_ksi and m.t_phi are the components of m's
    tangential part on ek and ep.  Returns the 6 arrays tx_ksi, ty_ksi, tz_ksi,
    tx_phi, ty_phi, tz_phi."""
    c = ek[0] * ep[0] + ek[1] * ep[1] + ek[2] * ep[2]
    fe = 1.0 / (1.0 - c * c)
    return ([fe * (ek[q] - c * ep[q]) for q in range(3)] +
            [fe * (ep[q] - c * ek[q]) for q in range(3)])


def wall_metrics(ye: np.ndarray, lo_wall: bool, hi_wall: bool):
    """Metrics of one axis whose min and/or max face is a wall, from the node
    coordinates ye of rows 1-G..n+G (padded; the wall rows are rows 1 and n,
    ghost coordinates beyond a wall are never used, those beyond an interface
    are the neighbour's).

    Returns (ide, idv, c2, c3):
      ide  Euler metric as the flux closures use it (length n+2G): 1/(11-pt
           derivative of y) inside, the idy2/idy4/idy6/idy8 of
           grid_metrics_wall_cart (centred a3/a5/a7/a9) on rows 2..5 (n-1..n-4),
           1/(a02 y) on the wall row (unused: wall rows get no Euler increment);
      idv  viscous metric (grid_metrics_cart with ngh=2): 1/(5-pt derivative),
           one-sided a04/a13 (a31/a40) on rows 1,2 (n-1,n);
      c2, c3  extrapolation coefficients cextrp2, cextrp3 of the min (key 0)
           and max (key 1) face (init_coeff_wall_cart).
    Ghost entries of ide/idv repeat the nearest row."""
    n = len(ye) - 2 * G
    y = lambda r: float(ye[G + r - 1])  # row r = 1..n (and ghosts)
    ide = np.zeros(n + 2 * G)
    idv = np.zeros(n + 2 * G)
    for r in range(1, n + 1):
        q = 5
        if lo_wall:
            q = min(q, r - 1)
        if hi_wall:
            q = min(q, n - r)
        if q == 0:
            if r == 1:
                d = A02[0] * y(1) + A02[1] * y(2) + A02[2] * y(3)
            else:
                d = -(A02[0] * y(n) + A02[1] * y(n - 1) + A02[2] * y(n - 2))
        else:
            d = 0.0
            for l, al in enumerate(A_CLOSE[q - 1], start=1):
                d += al * (y(r + l) - y(r - l))
        ide[G + r - 1] = 1.0 / d
        if lo_wall and r == 1:
            dv = sum(c * y(1 + m) for m, c in enumerate(A04))
        elif lo_wall and r == 2:
            dv = sum(c * y(1 + m) for m, c in enumerate(A13))
        elif hi_wall and r == n - 1:
            dv = -sum(c * y(n - m) for m, c in enumerate(A13))
        elif hi_wall and r == n:
            dv = -sum(c * y(n - m) for m, c in enumerate(A04))
        else:
            dv = A5[0] * (y(r + 1) - y(r - 1)) + A5[1] * (y(r + 2) - y(r - 2))
        idv[G + r - 1] = 1.0 / dv
    for a in (ide, idv):
        a[:G] = a[G]
        a[G + n:] = a[G + n - 1]
    c2, c3 = {}, {}
    if lo_wall:
        ry = (y(3) - y(2)) / (y(2) - y(1))
        c2[0] = (1.0 + ry) ** 2 / (ry * (2.0 + ry))
        c3[0] = 1.0 / (ry * (2.0 + ry))
    if hi_wall:
        ry = (y(n - 2) - y(n - 1)) / (y(n - 1) - y(n))
        c2[1] = (1.0 + ry) ** 2 / (ry * (2.0 + ry))
        c3[1] = 1.0 / (ry * (2.0 + ry))
    return ide, idv, c2, c3
