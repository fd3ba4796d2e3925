"""Case descriptions and synthetic workloads (set-up, host side).

A :class:`Case` carries what the reference reads from param.ini /
param_blocks.ini / feos_air.ini for the hot path, the grid metrics, the time
step and the initial conserved state.  Builders:

* :func:`tgv`            -- Taylor-Green vortex on a periodic Cartesian cube
                            (IC: add_tgv, src/Initial_condition/mod_init_flow.f90:577-620;
                            reference quantities: src/Main/setupref.f90:39-69,258-266,357)
* :func:`vortex_wavy`    -- isentropic vortex advected over a curvilinear
                            (wavy, 3-D GCL metrics) block, periodic faces
* :func:`entropy_wave`   -- small-amplitude entropy/shear waves for closed-form tests
* :func:`channel`        -- Cartesian channel: periodic x and z, no-slip walls at
                            y = 0 and y = Ly on a stretched grid
* :func:`shock_tube`     -- periodic shock tube with a wavy front (the shock
                            sensor across a real discontinuity)

The time step follows timestep_new (src/Time/timestep_new.f90:270-288):
deltat = CFL * delta_min / max(|u| + c).
"""
from __future__ import annotations

import dataclasses
import math
from typing import List, Optional

import numpy as np

from . import grid as gr
from .grid import G

# air (harness-authored feos_air.ini) and Sutherland constants of
# init_viscosity (src/Fluid/smod_tranprop_init.f90:33-40)
GAM = 1.4
RG = 287.0
PR = 0.71
C0 = 1.4579326545176254e-06
S0 = 110.4
# low-storage RK4 (init_RK nrk=4, src/Time/mod_time.f90:165-169)
CRK4 = (0.25, 1.0 / 3.0, 0.5, 1.0)
# 6-stage low-storage DRP RK (init_RK nrk=6, src/Time/mod_time.f90:253-294; Bogey & Bailly, JCP 2004):
# the amplification polynomial is sum_m g_m z^m with these g_m; crk(6) = g1, crk(5) = g2,
# crk(4) = g3/crk(5), ... crk(1) = g6/(crk(2) crk(3) crk(4) crk(5))
G_RK6_DRP = (1.0, 0.5, 0.165919771368, 0.040919732041, 0.007555704391, 0.000891421261)


def crk_from_gammas(g):
    """Low-storage stage coefficients whose amplification polynomial is
    1 + sum_m g[m-1] z^m: crk(nrk) = g1 and prod_{j > nrk-m} crk(j) = g_m."""
    n = len(g)
    c = [0.0] * n
    prod = 1.0
    for m in range(1, n + 1):
        c[n - m] = g[m - 1] / prod
        prod *= c[n - m]
    return tuple(c)


CRK6_DRP = crk_from_gammas(G_RK6_DRP)


@dataclasses.dataclass
class Case:
    name: str
    nx: int
    ny: int
    nz: int
    curvilinear: bool
    iorder_visc: int = 4
    filter: int = 1            # 0 none, 1 selective filter, 2 shock filter, 3 artificial viscosity,
    #                            4 artificial viscosity with shock capturing
    ducros: bool = True
    sw_edoh: bool = False      # Edoh switch of the shock filter (is_sw_edoh)
    nrk: int = 4
    crk: tuple = CRK4
    dt: float = 0.0
    gam: float = GAM
    rg: float = RG
    c0: float = C0
    S0: float = S0
    diffscale: float = 1.0
    cok_ratio: float = GAM * RG / (GAM - 1.0) / PR
    chi: float = 0.2           # dissip_coeff
    dissip_shock: float = 0.1
    csens: float = 0.5
    # Cartesian metrics: idx, idy, idz, idx_v, idy_v, idz_v (1-D, length n+2G)
    cart: Optional[List[np.ndarray]] = None
    # curvilinear metrics: 10 padded arrays each (ksi_x..phi_z, ijacob3)
    m: Optional[List[np.ndarray]] = None
    mv: Optional[List[np.ndarray]] = None
    # 2.5-D curvilinear (curvilinear == 2): padded 2-D x_ksi, x_eta, y_ksi, y_eta,
    # ijacob (Euler and viscous sets); cart[2], cart[5] give idz, idz_v
    m2: Optional[List[np.ndarray]] = None
    m2v: Optional[List[np.ndarray]] = None
    # initial conserved state: rho, rhou, rhov, rhow, rhoe (padded)
    U: Optional[List[np.ndarray]] = None
    # face kinds imin..kmax (0 periodic, 2 wall, 3 characteristic outflow,
    # 4 supersonic inlet, 5 wall with the slip flag 's') and the wall condition
    faces: tuple = (0, 0, 0, 0, 0, 0)
    adiabatic: bool = True
    t_wall: float = 0.0
    # type_wall_bc = 2 (is_wall2): wall rows advanced, slip walls possible
    wall2: bool = False
    # wall extrapolation coefficients per face (cextrp2, cextrp3)
    cextrp2: tuple = (0.0,) * 6
    cextrp3: tuple = (0.0,) * 6
    info: dict = dataclasses.field(default_factory=dict)

    @property
    def ncells(self) -> int:
        return self.nx * self.ny * self.nz

    def copy_state(self) -> List[np.ndarray]:
        return [a.copy() for a in self.U]


def sutherland(T):
    return C0 * np.sqrt(T * T * T) / (T + S0)


def conserved(rho, u, v, w, p, gam=GAM):
    return [rho, rho * u, rho * v, rho * w, p / (gam - 1.0) + 0.5 * rho * (u * u + v * v + w * w)]


def _to_padded(fields, nx, ny, nz):
    out = []
    for f in fields:
        a = gr.padded(nx, ny, nz)
        a[G:-G, G:-G, G:-G] = f
        gr.fill_periodic(a)
        out.append(a)
    return out


def timestep(U, delta_min, cfl, gam=GAM, rg=RG):
    """deltat = CFL*delta_min/max(|u|+c) over the interior (timestep_new.f90:256-288)."""
    rho = gr.interior(U[0])
    u, v, w = (gr.interior(U[q]) / rho for q in (1, 2, 3))
    p = (gam - 1.0) * (gr.interior(U[4]) - 0.5 * rho * (u * u + v * v + w * w))
    c = np.sqrt(gam * p / rho)
    umax = float(np.max(np.sqrt(u * u + v * v + w * w) + c))
    return cfl * delta_min / umax


def tgv(n: int = 64, mach: float = 0.1, re: float = 1600.0, cfl: float = 0.8, chi: float = 0.2,
        filter: int = 1, iorder_visc: int = 4, T_ref: float = 300.0, rho_ref: float = 1.177) -> Case:
    """3-D Taylor-Green vortex on [0,2pi)^3, Cartesian, periodic (config: 64^3, M=0.1, Re=1600)."""
    L = 2.0 * math.pi
    dx = L / n
    p_ref = rho_ref * RG * T_ref
    c_ref = math.sqrt(GAM * RG * T_ref)
    u_ref = mach * c_ref
    mu_ref = rho_ref * u_ref * 1.0 / re           # setupref.f90:265-266 (L=1)
    muw_ref = float(sutherland(np.float64(T_ref)))
    x = np.arange(n) * dx
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    u = u_ref * np.sin(X) * np.cos(Y) * np.cos(Z)
    v = -u_ref * np.cos(X) * np.sin(Y) * np.cos(Z)
    w = np.zeros_like(u)
    p = p_ref + rho_ref * u_ref ** 2 / 16.0 * (np.cos(2.0 * Z) + 2.0) * (np.cos(2.0 * X) + np.cos(2.0 * Y))
    T = p / p_ref * T_ref
    rho = p / (RG * T)
    U = _to_padded(conserved(rho, u, v, w, p), n, n, n)
    idx = np.full(n + 2 * G, 1.0 / dx)
    c = Case("tgv", n, n, n, curvilinear=False, iorder_visc=iorder_visc, filter=filter, ducros=False,
             diffscale=mu_ref / muw_ref, chi=chi,
             cart=[idx.copy() for _ in range(6)], U=U)
    c.dt = timestep(U, dx, cfl)
    c.info = dict(mach=mach, re=re, cfl=cfl, u_ref=u_ref, p_ref=p_ref, rho_ref=rho_ref, dx=dx, L=L)
    return c


def wavy_mapping(lengths, amp):
    """x = xi + ax sin(2pi eta/Ly) sin(2pi zeta/Lz), and cyclic: a smooth,
    fully periodic curvilinear distortion of a box."""
    Lx, Ly, Lz = lengths
    ax, ay, az = amp
    tp = 2.0 * math.pi

    def f(XI, ET, ZE):
        x = XI + ax * np.sin(tp * ET / Ly) * np.sin(tp * ZE / Lz)
        y = ET + ay * np.sin(tp * XI / Lx) * np.sin(tp * ZE / Lz)
        z = ZE + az * np.sin(tp * XI / Lx) * np.sin(tp * ET / Ly)
        return x, y, z

    return f


def vortex_wavy(nx: int = 512, ny: int = 256, nz: int = 128, mach: float = 0.5, beta: float = 5.0,
                radius: float = 0.5, re: float = 1.0e4, cfl: float = 0.8, filter: int = 2,
                lengths=None, amp=None, T_inf: float = 300.0, rho_inf: float = 1.177,
                iorder_visc: int = 4) -> Case:
    """Isentropic vortex (strength beta, radius R) advected at Mach `mach` along x
    over a curvilinear block with 3-D GCL metrics and periodic faces."""
    if lengths is None:
        h = 10.0 / nx
        lengths = (nx * h, ny * h, nz * h)
    Lx, Ly, Lz = lengths
    if amp is None:
        amp = (0.02 * Lx / 2, 0.05 * Ly / 2, 0.02 * Lz / 2)
    mapping = wavy_mapping(lengths, amp)
    m, mv, (xc, yc, zc) = gr.curvilinear_metrics_periodic(mapping, nx, ny, nz, lengths,
                                                            viscous=iorder_visc != 0)
    c_inf = math.sqrt(GAM * RG * T_inf)
    u_inf = mach * c_inf
    xv, yv = Lx / 2.0, Ly / 2.0
    r2 = ((xc - xv) ** 2 + (yc - yv) ** 2) / radius ** 2
    e = np.exp(0.5 * (1.0 - r2))
    u = u_inf * (1.0 - beta / (2.0 * math.pi) * (yc - yv) / radius * e)
    v = u_inf * (beta / (2.0 * math.pi) * (xc - xv) / radius * e)
    w = np.zeros_like(u)
    T = T_inf * (1.0 - (GAM - 1.0) * beta ** 2 * mach ** 2 / (8.0 * math.pi ** 2) * e * e)
    rho = rho_inf * (T / T_inf) ** (1.0 / (GAM - 1.0))
    p = rho * RG * T
    U = _to_padded(conserved(rho, u, v, w, p), nx, ny, nz)
    mu_ref = rho_inf * u_inf * radius / re
    diffscale = mu_ref / float(sutherland(np.float64(T_inf)))
    # minimum edge length between neighbouring nodes (timestep_new.f90:40-200)
    dmin = min(float(np.min(np.sqrt(np.diff(xc, axis=a) ** 2 + np.diff(yc, axis=a) ** 2 +
                                    np.diff(zc, axis=a) ** 2))) for a in (0, 1, 2))
    cs = Case("vortex_wavy", nx, ny, nz, curvilinear=True, iorder_visc=iorder_visc, filter=filter,
              ducros=True, diffscale=diffscale, m=m, mv=mv, U=U)
    cs.dt = timestep(U, dmin, cfl)
    cs.info = dict(mach=mach, beta=beta, radius=radius, re=re, cfl=cfl, lengths=lengths, amp=amp,
                   dmin=dmin)
    return cs


def tanh_two_walls(ny: int, Ly: float, ratio: float = 1.05):
    """Computational coordinate eta(j) of rows j (any real j, smooth beyond
    the walls) clustered at both walls j = 1 and j = ny by tanh stretching,
    eta = Ly/2 (1 + tanh(b (2s - 1))/tanh(b)), s = (j-1)/(ny-1), with b
    chosen so that the largest ratio of neighbouring spacings is `ratio`."""
    def nodes(b, s):
        return 0.5 * Ly * (1.0 + np.tanh(b * (2.0 * s - 1.0)) / math.tanh(b))

    s = np.arange(ny) / (ny - 1)

    def max_ratio(b):
        d = np.diff(nodes(b, s))
        return float(np.max(np.maximum(d[1:] / d[:-1], d[:-1] / d[1:])))

    lo, hi = 1e-6, 10.0
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if max_ratio(mid) > ratio:
            hi = mid
        else:
            lo = mid
    b = lo
    return (lambda j: nodes(b, (np.asarray(j, dtype=np.float64) - 1.0) / (ny - 1))), b


def vortex_channel(nx: int = 512, ny: int = 256, nz: int = 128, mach: float = 0.5, beta: float = 5.0,
                   radius: float = 0.5, re: float = 1.0e4, cfl: float = 0.8, filter: int = 2,
                   lengths=None, amp: float = 0.05, ratio: float = 1.05, delta: float = 0.1,
                   adiabatic: bool = True, T_inf: float = 300.0, rho_inf: float = 1.177,
                   iorder_visc: int = 4, slip: bool = False) -> Case:
    """The curvilinear headline block as stated: x uniform and periodic, y
    between walls at j = 1 and j = ny with tanh clustering at both (largest
    spacing ratio `ratio`), the wavy map y' = y + amp sin(2 pi x/Lx) (1 -
    y/Ly) (a wavy lower wall, a flat upper one), z uniform and periodic; 3-D
    GCL metrics; an isentropic vortex (strength beta, radius R) advected at
    Mach `mach` in the middle of the channel.
    slip=False: no-slip walls (is_wall2 = F); the free stream meets them
    through a thin laminar layer u = U tanh(d/delta) of the distance d to the
    nearer wall, zero velocity on the wall rows.
    slip=True: slip walls, type_wall_bc = 2 with the flag 's' on both y-faces
    (an adiabatic case); the uniform free stream runs along them."""
    if slip and iorder_visc and not adiabatic:
        raise ValueError("slip walls need an adiabatic case (an isothermal one binds no slip routine)")
    if lengths is None:
        h = 10.0 / nx
        lengths = (nx * h, ny * h, nz * h)
    Lx, Ly, Lz = lengths
    eta_of, b = tanh_two_walls(ny, Ly, ratio)
    tp = 2.0 * math.pi

    def mapping(XI, ET, ZE):
        return XI, ET + amp * np.sin(tp * XI / Lx) * (1.0 - ET / Ly), ZE

    m, mv, (xc, yc, zc), wall = gr.curvilinear_metrics_jwalls(mapping, eta_of, nx, ny, nz, lengths,
                                                              viscous=iorder_visc != 0)
    c_inf = math.sqrt(GAM * RG * T_inf)
    u_inf = mach * c_inf
    xv, yv = Lx / 2.0, Ly / 2.0
    r2 = ((xc - xv) ** 2 + (yc - yv) ** 2) / radius ** 2
    e = np.exp(0.5 * (1.0 - r2))
    eta = eta_of(np.arange(1, ny + 1)).reshape(1, -1, 1)
    dwall = np.minimum(eta, Ly - eta)
    layer = 1.0 if slip else np.tanh(dwall / delta)
    u = u_inf * (layer - beta / (2.0 * math.pi) * (yc - yv) / radius * e)
    v = u_inf * (beta / (2.0 * math.pi) * (xc - xv) / radius * e)
    w = np.zeros_like(u)
    T = T_inf * (1.0 - (GAM - 1.0) * beta ** 2 * mach ** 2 / (8.0 * math.pi ** 2) * e * e)
    rho = rho_inf * (T / T_inf) ** (1.0 / (GAM - 1.0))
    p = rho * RG * T
    if not slip:
        u[:, 0, :] = v[:, 0, :] = u[:, -1, :] = v[:, -1, :] = 0.0
    U = []
    for f in conserved(rho, u, v, w, p):
        a = gr.padded(nx, ny, nz)
        a[G:-G, G:-G, G:-G] = f
        # ghost rows beyond the walls repeat the wall rows (only the shock
        # filter's pressure sensor of a wall row reads them)
        a[G:-G, :G, G:-G] = a[G:-G, G:G + 1, G:-G]
        a[G:-G, G + ny:, G:-G] = a[G:-G, G + ny - 1:G + ny, G:-G]
        gr.fill_periodic(a, axes=(0, 2))
        U.append(a)
    mu_ref = rho_inf * u_inf * radius / re
    diffscale = mu_ref / float(sutherland(np.float64(T_inf)))
    dmin = min(float(np.min(np.sqrt(np.diff(xc, axis=a) ** 2 + np.diff(yc, axis=a) ** 2 +
                                    np.diff(zc, axis=a) ** 2))) for a in (0, 1, 2))
    cs = Case("vortex_channel", nx, ny, nz, curvilinear=True, iorder_visc=iorder_visc, filter=filter,
              ducros=True, diffscale=diffscale, m=m, mv=mv, U=U, faces=(0, 0, 5, 5, 0, 0) if slip else (0, 0, 2, 2, 0, 0),
              adiabatic=adiabatic, t_wall=T_inf, wall2=slip)
    cs.wall_c3 = wall
    cs.dt = timestep(U, dmin, cfl)
    cs.info = dict(mach=mach, beta=beta, radius=radius, re=re, cfl=cfl, lengths=lengths, amp=amp,
                   ratio=ratio, tanh_b=b, delta=delta, dmin=dmin, slip=slip)
    return cs


def wavy_mapping_2d(lengths, amp):
    """x = xi + ax sin(2pi eta/Ly), y = eta + ay sin(2pi xi/Lx): a smooth,
    periodic distortion in (x, y), extruded along z."""
    Lx, Ly = lengths[0], lengths[1]
    ax, ay = amp
    tp = 2.0 * math.pi

    def f(XI, ET):
        return XI + ax * np.sin(tp * ET / Ly), ET + ay * np.sin(tp * XI / Lx)

    return f


def vortex_wavy_c2(nx: int = 512, ny: int = 256, nz: int = 128, mach: float = 0.5, beta: float = 5.0,
                   radius: float = 0.5, re: float = 1.0e4, cfl: float = 0.8, filter: int = 2,
                   lengths=None, amp=None, T_inf: float = 300.0, rho_inf: float = 1.177,
                   iorder_visc: int = 4) -> Case:
    """Isentropic vortex over a z-extruded wavy block run as 2.5-D curvilinear
    (is_curv: 2-D metrics in (i, j), Cartesian z), periodic faces."""
    if lengths is None:
        h = 10.0 / nx
        lengths = (nx * h, ny * h, nz * h)
    Lx, Ly, Lz = lengths
    if amp is None:
        amp = (0.02 * Lx / 2, 0.05 * Ly / 2)
    m2, m2v, (x2, y2) = gr.c2_metrics_periodic(wavy_mapping_2d(lengths, amp), nx, ny, lengths,
                                               viscous=iorder_visc != 0)
    if m2v is None:
        m2v = [a.copy() for a in m2]
    dz = Lz / nz
    # the flow does not depend on z: fields are built on (j, i) and extruded
    xc, yc = x2[None], y2[None]
    c_inf = math.sqrt(GAM * RG * T_inf)
    u_inf = mach * c_inf
    xv, yv = Lx / 2.0, Ly / 2.0
    r2 = ((xc - xv) ** 2 + (yc - yv) ** 2) / radius ** 2
    e = np.exp(0.5 * (1.0 - r2))
    u = u_inf * (1.0 - beta / (2.0 * math.pi) * (yc - yv) / radius * e)
    v = u_inf * (beta / (2.0 * math.pi) * (xc - xv) / radius * e)
    w = np.zeros_like(u)
    T = T_inf * (1.0 - (GAM - 1.0) * beta ** 2 * mach ** 2 / (8.0 * math.pi ** 2) * e * e)
    rho = rho_inf * (T / T_inf) ** (1.0 / (GAM - 1.0))
    p = rho * RG * T
    U = _to_padded([np.broadcast_to(f, (nz, ny, nx)) for f in conserved(rho, u, v, w, p)], nx, ny, nz)
    mu_ref = rho_inf * u_inf * radius / re
    diffscale = mu_ref / float(sutherland(np.float64(T_inf)))
    dmin = min(float(np.min(np.sqrt(np.diff(x2, axis=a) ** 2 + np.diff(y2, axis=a) ** 2))) for a in (0, 1))
    dmin = min(dmin, dz)
    idz = np.full(nz + 2 * G, 1.0 / dz)
    ones_x, ones_y = np.ones(nx + 2 * G), np.ones(ny + 2 * G)
    cs = Case("vortex_wavy_c2", nx, ny, nz, curvilinear=2, iorder_visc=iorder_visc, filter=filter,
              ducros=True, diffscale=diffscale, m2=m2, m2v=m2v, U=U,
              cart=[ones_x, ones_y, idz, ones_x.copy(), ones_y.copy(), idz.copy()])
    cs.dt = cfl * dmin / float(np.max(np.sqrt(u * u + v * v + w * w) + np.sqrt(GAM * p / rho)))
    cs.info = dict(mach=mach, beta=beta, radius=radius, re=re, cfl=cfl, lengths=lengths, amp=amp,
                   dmin=dmin)
    return cs


def vortex_periodic(n: int = 64, mach: float = 0.5, beta: float = 5.0, radius: float = 0.5,
                    re: float = 1.0e4, cfl: float = 0.8, chi: float = 0.2, filter: int = 1,
                    iorder_visc: int = 4, T_inf: float = 300.0, rho_inf: float = 1.177) -> Case:
    """Isentropic vortex (Shu, strength beta, radius R) advected at Mach `mach`
    along x in a periodic Cartesian cube of edge 10 R."""
    L = 20.0 * radius
    dx = L / n
    x = np.arange(n) * dx
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    c_inf = math.sqrt(GAM * RG * T_inf)
    u_inf = mach * c_inf
    xv = yv = L / 2.0
    r2 = ((X - xv) ** 2 + (Y - yv) ** 2) / radius ** 2
    e = np.exp(0.5 * (1.0 - r2))
    u = u_inf * (1.0 - beta / (2.0 * math.pi) * (Y - yv) / radius * e)
    v = u_inf * (beta / (2.0 * math.pi) * (X - xv) / radius * e)
    w = np.zeros_like(u)
    T = T_inf * (1.0 - (GAM - 1.0) * beta ** 2 * mach ** 2 / (8.0 * math.pi ** 2) * e * e)
    rho = rho_inf * (T / T_inf) ** (1.0 / (GAM - 1.0))
    p = rho * RG * T
    U = _to_padded(conserved(rho, u, v, w, p), n, n, n)
    mu_ref = rho_inf * u_inf * radius / re
    idx = np.full(n + 2 * G, 1.0 / dx)
    c = Case("vortex", n, n, n, curvilinear=False, iorder_visc=iorder_visc, filter=filter, ducros=False,
             diffscale=mu_ref / float(sutherland(np.float64(T_inf))), chi=chi,
             cart=[idx.copy() for _ in range(6)], U=U)
    c.dt = timestep(U, dx, cfl)
    c.info = dict(mach=mach, beta=beta, radius=radius, re=re, cfl=cfl, L=L)
    return c


def cartesian_box(nx, ny, nz, lengths, fields, iorder_visc=0, filter=0, diffscale=1.0, dt=1e-6,
                  chi=0.2) -> Case:
    """Generic periodic Cartesian case from primitive fields (rho,u,v,w,p) on the interior."""
    Lx, Ly, Lz = lengths
    cart = [np.full(nx + 2 * G, nx / Lx), np.full(ny + 2 * G, ny / Ly), np.full(nz + 2 * G, nz / Lz)]
    cart = cart + [a.copy() for a in cart]
    U = _to_padded(conserved(*fields), nx, ny, nz)
    return Case("box", nx, ny, nz, curvilinear=False, iorder_visc=iorder_visc, filter=filter,
                ducros=False, diffscale=diffscale, cart=cart, U=U, dt=dt, chi=chi)


def shock_tube(nx: int = 96, ny: int = 32, nz: int = 8, ducros: bool = True, dissip_shock: float = 0.3,
               csens: float = 0.3, cfl: float = 0.5, filter: int = 2, ratio: float = 2.5,
               noise: float = 5.0, seed: int = 20261203) -> Case:
    """Periodic shock tube with a wavy front: a slab of gas at (2 rho0, ratio p0)
    between fronts x = Lx/2 + 0.1 Lx sin(2 pi y/Ly) and that plus Lx/2, gas at
    (rho0, p0) elsewhere, seeded velocity noise of amplitude `noise` m/s.  The
    jumps cross the i and j directions, so the pressure sensor of both is O(1)
    on the first steps (0.42 at csens = 0.3 for ratio 2.5), and the noise gives
    the Ducros sensor a non-trivial divergence/vorticity blend.  Viscous (the
    sensor reads the gradients of U^n), shock filter."""
    Lx, Ly, Lz = 2.0, 2.0 * ny / nx, 2.0 * nz / nx
    x = (np.arange(nx) + 0.5) * Lx / nx
    y = (np.arange(ny) + 0.5) * Ly / ny
    z = (np.arange(nz) + 0.5) * Lz / nz
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    front = 0.5 * Lx + 0.1 * Lx * np.sin(2.0 * math.pi * Y / Ly)
    inside = np.mod(X - front, Lx) > 0.5 * Lx  # x in (front - Lx/2, front), periodically
    rho0, p0 = 1.177, 1.0e5
    rho = np.where(inside, 2.0 * rho0, rho0)
    p = np.where(inside, ratio * p0, p0)
    rng = np.random.default_rng(seed)
    u, v, w = (noise * rng.uniform(-1.0, 1.0, X.shape) for _ in range(3))
    c = cartesian_box(nx, ny, nz, (Lx, Ly, Lz), (rho, u, v, w, p), iorder_visc=4, filter=filter, chi=0.2)
    c.name = "shock_tube"
    c.ducros = ducros
    c.dissip_shock = dissip_shock
    c.csens = csens
    c.dt = timestep(c.U, Lx / nx, cfl)
    c.info = dict(ratio=ratio, noise=noise, seed=seed, cfl=cfl, lengths=(Lx, Ly, Lz))
    return c


def stretched_wall_coords(n: int, L: float, beta: float) -> np.ndarray:
    """Nodes of rows 1..n on [0, L] clustered at both ends (tanh), padded with
    G ghost coordinates of the end spacings (unused next to walls)."""
    s = np.arange(n) / (n - 1)
    y = 0.5 * L * (1.0 + np.tanh(beta * (2.0 * s - 1.0)) / math.tanh(beta))
    ye = np.empty(n + 2 * G)
    ye[G:G + n] = y
    ye[:G] = y[0] - (y[1] - y[0]) * np.arange(G, 0, -1)
    ye[G + n:] = y[-1] + (y[-1] - y[-2]) * np.arange(1, G + 1)
    return ye


def channel(nx: int = 64, ny: int = 48, nz: int = 16, lengths=(4.0, 2.0, 1.0), beta: float = 1.5,
            mach: float = 0.3, re: float = 2000.0, cfl: float = 0.5, chi: float = 0.2, filter: int = 1,
            iorder_visc: int = 4, adiabatic: bool = True, t_wall: float = 300.0, pulse: float = 1e-2,
            seed: int = 20261202, T_ref: float = 300.0, rho_ref: float = 1.177,
            symmetric: bool = False, wall2: bool = False, slip: bool = False) -> Case:
    """Channel between no-slip walls at y = 0 and y = Ly (rows 1 and ny), periodic in
    x and z, y-nodes clustered at the walls.  Flow: a parabolic streamwise profile
    of Mach mach, 3-D velocity perturbations vanishing at the walls, and a Gaussian
    pressure pulse; temperature T_ref (the isothermal wall temperature t_wall).
    symmetric: the flow is mirror-symmetric about y = Ly/2 (v antisymmetric).
    wall2: advanced walls (type_wall_bc = 2); slip: with the flag 's' on both
    walls (and wall2), and a uniform streamwise 0.3 mach added so that the
    walls carry tangential flow."""
    Lx, Ly, Lz = lengths
    dx, dz = Lx / nx, Lz / nz
    ye = stretched_wall_coords(ny, Ly, beta)
    y = ye[G:G + ny]
    x = np.arange(nx) * dx
    z = np.arange(nz) * dz
    # broadcast (nz,1,1), (1,ny,1), (1,1,nx): the fields come out (nz,ny,nx)
    Z, Y, X = z.reshape(-1, 1, 1), y.reshape(1, -1, 1), x.reshape(1, 1, -1)
    c_ref = math.sqrt(GAM * RG * T_ref)
    u_ref = mach * c_ref
    p_ref = rho_ref * RG * T_ref
    eta = Y / Ly
    env = 4.0 * eta * (1.0 - eta)                      # 0 on both walls
    rng = np.random.default_rng(seed)
    ph = rng.uniform(0.0, 2.0 * math.pi, 6)
    kx, kz = 2.0 * math.pi / Lx, 2.0 * math.pi / Lz
    u = u_ref * (env + 0.05 * env * np.sin(kx * X + ph[0]) * np.cos(kz * Z + ph[1]))
    if slip:
        u = u + 0.3 * u_ref
    v = 0.05 * u_ref * env * env * np.sin(2.0 * kx * X + ph[2]) * np.sin(kz * Z + ph[3])
    if symmetric:
        v = v * (1.0 - 2.0 * eta)
    w = 0.05 * u_ref * env * np.cos(kx * X + ph[4]) * np.sin(2.0 * kz * Z + ph[5])
    yc = 0.5 * Ly if symmetric else 0.4 * Ly
    r2 = ((X - 0.5 * Lx) ** 2 + (Y - yc) ** 2 + (Z - 0.5 * Lz) ** 2) / (0.1 * Ly) ** 2
    p = p_ref * (1.0 + pulse * np.exp(-math.log(2.0) * r2))
    T = T_ref
    rho = p / (RG * T)
    u, v, w = (np.broadcast_to(q, p.shape) for q in (u, v, w))
    U = []
    for f in conserved(rho, u, v, w, p):
        a = gr.padded(nx, ny, nz)
        a[G:-G, G:-G, G:-G] = f
        a[G:-G, :G, G:-G] = f[:, :1, :]                # ghosts beyond the walls: only the shock
        # filter reads them (p one row beyond each wall, in the wall row's pressure sensor)
        a[G:-G, G + ny:, G:-G] = f[:, -1:, :]
        gr.fill_periodic(a, axes=(0, 2))
        U.append(a)
    ide, idv, c2, c3 = gr.wall_metrics(ye, True, True)
    idx = np.full(nx + 2 * G, 1.0 / dx)
    idz = np.full(nz + 2 * G, 1.0 / dz)
    mu_ref = rho_ref * u_ref * (0.5 * Ly) / re
    muw_ref = float(sutherland(np.float64(T_ref)))
    ext2 = [0.0] * 6
    ext3 = [0.0] * 6
    ext2[2], ext2[3], ext3[2], ext3[3] = c2[0], c2[1], c3[0], c3[1]
    c = Case("channel", nx, ny, nz, curvilinear=False, iorder_visc=iorder_visc, filter=filter, ducros=False,
             diffscale=mu_ref / muw_ref, chi=chi, cart=[idx, ide, idz, idx.copy(), idv, idz.copy()], U=U,
             faces=(0, 0, 5, 5, 0, 0) if slip else (0, 0, 2, 2, 0, 0), adiabatic=adiabatic, t_wall=t_wall,
             cextrp2=tuple(ext2), cextrp3=tuple(ext3), wall2=wall2 or slip)
    c.dt = timestep(U, min(dx, dz, float(np.min(np.diff(y)))), cfl)
    c.info = dict(mach=mach, re=re, cfl=cfl, beta=beta, lengths=lengths, y=y, u_ref=u_ref, p_ref=p_ref)
    return c


def duct(nx: int = 48, ny: int = 40, nz: int = 16, lengths=(2.0, 1.6, 2.0), beta: float = 1.2,
         mach: float = 0.3, re: float = 1000.0, cfl: float = 0.5, chi: float = 0.2, filter: int = 1,
         iorder_visc: int = 4, adiabatic: bool = True, t_wall: float = 300.0, pulse: float = 1e-2,
         seed: int = 7, T_ref: float = 300.0, rho_ref: float = 1.177) -> Case:
    """Square-ish duct: no-slip walls on the i-faces (x = 0, Lx) and on the
    j-faces (y = 0, Ly), both axes clustered at their walls, periodic along z.
    Flow: a streamwise (z) profile of Mach mach vanishing on all four walls,
    cross-flow perturbations vanishing there too, and a Gaussian pressure pulse
    off the centre; temperature T_ref (the isothermal wall temperature t_wall)."""
    Lx, Ly, Lz = lengths
    dz = Lz / nz
    xe = stretched_wall_coords(nx, Lx, beta)
    ye = stretched_wall_coords(ny, Ly, beta)
    x, y = xe[G:G + nx], ye[G:G + ny]
    z = np.arange(nz) * dz
    Z, Y, X = z.reshape(-1, 1, 1), y.reshape(1, -1, 1), x.reshape(1, 1, -1)
    c_ref = math.sqrt(GAM * RG * T_ref)
    w_ref = mach * c_ref
    p_ref = rho_ref * RG * T_ref
    ex, ey = 4.0 * (X / Lx) * (1.0 - X / Lx), 4.0 * (Y / Ly) * (1.0 - Y / Ly)
    env = ex * ey                                      # 0 on the four walls
    rng = np.random.default_rng(seed)
    ph = rng.uniform(0.0, 2.0 * math.pi, 6)
    kz = 2.0 * math.pi / Lz
    w = w_ref * env * (1.0 + 0.05 * np.sin(kz * Z + ph[0]))
    u = 0.05 * w_ref * env * np.sin(math.pi * Y / Ly + ph[1]) * np.cos(kz * Z + ph[2])
    v = 0.05 * w_ref * env * np.sin(math.pi * X / Lx + ph[3]) * np.sin(kz * Z + ph[4])
    r2 = ((X - 0.4 * Lx) ** 2 + (Y - 0.45 * Ly) ** 2 + (Z - 0.5 * Lz) ** 2) / (0.1 * Ly) ** 2
    p = p_ref * (1.0 + pulse * np.exp(-math.log(2.0) * r2))
    rho = p / (RG * T_ref)
    u, v, w = (np.broadcast_to(q, p.shape) for q in (u, v, w))
    U = []
    for f in conserved(rho, u, v, w, p):
        a = gr.padded(nx, ny, nz)
        a[G:-G, G:-G, G:-G] = f
        # ghosts beyond the walls repeat the wall rows (only the shock filter's
        # pressure sensor of a wall row reads them)
        a[G:-G, G:-G, :G] = f[:, :, :1]
        a[G:-G, G:-G, G + nx:] = f[:, :, -1:]
        a[G:-G, :G, :] = a[G:-G, G:G + 1, :]
        a[G:-G, G + ny:, :] = a[G:-G, G + ny - 1:G + ny, :]
        gr.fill_periodic(a, axes=(0,))
        U.append(a)
    idxe, idxv, cx2, cx3 = gr.wall_metrics(xe, True, True)
    idye, idyv, cy2, cy3 = gr.wall_metrics(ye, True, True)
    idz = np.full(nz + 2 * G, 1.0 / dz)
    mu_ref = rho_ref * w_ref * (0.5 * Ly) / re
    ext2 = (cx2[0], cx2[1], cy2[0], cy2[1], 0.0, 0.0)
    ext3 = (cx3[0], cx3[1], cy3[0], cy3[1], 0.0, 0.0)
    c = Case("duct", nx, ny, nz, curvilinear=False, iorder_visc=iorder_visc, filter=filter, ducros=False,
             diffscale=mu_ref / float(sutherland(np.float64(T_ref))), chi=chi,
             cart=[idxe, idye, idz, idxv, idyv, idz.copy()], U=U, faces=(2, 2, 2, 2, 0, 0),
             adiabatic=adiabatic, t_wall=t_wall, cextrp2=ext2, cextrp3=ext3)
    c.dt = timestep(U, min(float(np.min(np.diff(x))), float(np.min(np.diff(y))), dz), cfl)
    c.info = dict(mach=mach, re=re, cfl=cfl, beta=beta, lengths=lengths, w_ref=w_ref, p_ref=p_ref)
    return c


def affine_c3_metrics(nx, ny, nz, lengths, viscous=True):
    """c3 metrics of an undistorted box: ksi_x = dy*dz etc., 1/J = 1/(dx dy dz)."""
    dx, dy, dz = (L / n for L, n in zip(lengths, (nx, ny, nz)))
    vals = [dy * dz, 0, 0, 0, dx * dz, 0, 0, 0, dx * dy, 1.0 / (dx * dy * dz)]
    m = [gr.padded(nx, ny, nz, float(v)) for v in vals]
    mv = [gr.padded(nx, ny, nz, float(v)) for v in vals] if viscous else None
    return m, mv


def oblique_shock(mach: float, theta_deg: float, gam: float = GAM):
    """Weak oblique-shock solution of the theta-beta-Mach relation: shock angle
    beta (radians), and the ratios p2/p1, rho2/rho1 and M2 behind it."""
    th = math.radians(theta_deg)

    def f(b):
        return math.tan(th) - 2.0 / math.tan(b) * (mach ** 2 * math.sin(b) ** 2 - 1.0) / (
            mach ** 2 * (gam + math.cos(2.0 * b)) + 2.0)

    lo, hi = math.asin(1.0 / mach) + 1e-9, math.radians(64.0)  # the weak branch lies below ~64 deg at M = 2
    for _ in range(200):
        mid = 0.5 * (lo + hi)
        if f(lo) * f(mid) <= 0.0:
            hi = mid
        else:
            lo = mid
    b = 0.5 * (lo + hi)
    mn1 = mach * math.sin(b)
    p21 = 1.0 + 2.0 * gam / (gam + 1.0) * (mn1 ** 2 - 1.0)
    r21 = (gam + 1.0) * mn1 ** 2 / ((gam - 1.0) * mn1 ** 2 + 2.0)
    mn2 = math.sqrt((1.0 + 0.5 * (gam - 1.0) * mn1 ** 2) / (gam * mn1 ** 2 - 0.5 * (gam - 1.0)))
    return b, p21, r21, mn2 / math.sin(b - th)


def one_sided_coords(n: int, L: float, beta: float) -> np.ndarray:
    """Nodes of rows 1..n on [0, L] clustered at 0 (tanh), padded with G ghost
    coordinates of the end spacings."""
    s = np.arange(n) / (n - 1)
    y = L * (1.0 - np.tanh(beta * (1.0 - s)) / math.tanh(beta)) if beta > 0 else L * s
    ye = np.empty(n + 2 * G)
    ye[G:G + n] = y
    ye[:G] = y[0] - (y[1] - y[0]) * np.arange(G, 0, -1)
    ye[G + n:] = y[-1] + (y[-1] - y[-2]) * np.arange(1, G + 1)
    return ye


def shock_bl(nx: int = 1024, ny: int = 512, nz: int = 8, mach: float = 2.0, wedge_deg: float = 10.0,
             re_delta: float = 1.0e4, lengths=(2.0, 1.0, 0.05), delta: float = 0.05, beta_y: float = 1.2,
             x_shock: float = 0.35, noise: float = 1.0e-4, cfl: float = 0.8, filter: int = 2,
             dissip_shock: float = 0.3, chi: float = 0.2, csens: float = 0.5, adiabatic: bool = True,
             seed: int = 4, T_inf: float = 300.0, rho_inf: float = 1.177, iorder_visc: int = 4) -> Case:
    """Mach 2 oblique shock (wedge angle 10 deg) meeting a laminar boundary layer
    (configuration 4): a 2-D-extruded Cartesian block, supersonic inlet at
    imin (the initial first row), characteristic outflow at imax and jmax
    (bc_charac_imax / bc_charac_jmax), a no-slip wall at jmin, z periodic.
    Initial field: a Pohlhausen (Blasius-like) profile u/U = 2 eta - 2 eta^3 +
    eta^4 of thickness delta with a Crocco-Busemann temperature (recovery factor
    Pr^1/2), uniform pressure; above the shock line, which enters through the
    top at x_shock at the angle beta of the theta-beta-M relation, the
    post-shock state (deflected by the wedge angle), blended over two cells;
    seeded uniform velocity noise of amplitude noise*U (zero on the wall row).
    Re_delta = rho U delta / mu sets the Sutherland scale; shock filter with
    Ducros sensor; y-nodes clustered at the wall."""
    Lx, Ly, Lz = lengths
    dx, dz = Lx / nx, Lz / nz
    xe = np.arange(-G, nx + G) * dx
    ye = one_sided_coords(ny, Ly, beta_y)
    x, y = xe[G:G + nx], ye[G:G + ny]
    z = np.arange(nz) * dz
    Z, Y, X = z.reshape(-1, 1, 1), y.reshape(1, -1, 1), x.reshape(1, 1, -1)
    c_inf = math.sqrt(GAM * RG * T_inf)
    U_inf = mach * c_inf
    p_inf = rho_inf * RG * T_inf
    eta = np.minimum(Y / delta, 1.0)
    fu = 2.0 * eta - 2.0 * eta ** 3 + eta ** 4
    rec = math.sqrt(PR)
    T = T_inf * (1.0 + rec * 0.5 * (GAM - 1.0) * mach ** 2 * (1.0 - fu * fu))
    u = U_inf * fu
    v = np.zeros_like(u)
    p = np.full_like(u, p_inf)
    # the oblique shock: post-shock state above the line y = Ly - (x - x_shock) tan(beta)
    b, p21, r21, m2 = oblique_shock(mach, wedge_deg)
    th = math.radians(wedge_deg)
    T2 = T_inf * p21 / r21
    U2 = m2 * math.sqrt(GAM * RG * T2)
    dist = (Y - (Ly - (X - x_shock) * math.tan(b))) * math.cos(b)  # > 0 above the shock
    hs = 0.5 * (1.0 + np.tanh(dist / (2.0 * max(dx, float(np.min(np.diff(y)))))))
    u = u * (1.0 - hs) + U2 * math.cos(th) * hs
    v = v * (1.0 - hs) - U2 * math.sin(th) * hs
    p = p * (1.0 - hs) + p_inf * p21 * hs
    T = T * (1.0 - hs) + T2 * hs
    shape = (nz, ny, nx)
    u, v, p, T = (np.broadcast_to(q, shape).copy() for q in (u, v, p, T))
    rng = np.random.default_rng(seed)
    w = np.zeros(shape)
    for q in (u, v, w):
        q += noise * U_inf * rng.uniform(-1.0, 1.0, shape)
    u[:, 0, :] = v[:, 0, :] = w[:, 0, :] = 0.0  # no-slip wall row
    rho = p / (RG * T)
    U = []
    for f in conserved(rho, u, v, w, p):
        a = gr.padded(nx, ny, nz)
        a[G:-G, G:-G, G:-G] = f
        # ghosts beyond the boundary faces repeat the boundary row (only the
        # shock filter's pressure sensor of a boundary row reads them)
        a[G:-G, G:-G, :G] = f[:, :, :1]
        a[G:-G, G:-G, G + nx:] = f[:, :, -1:]
        a[G:-G, :G, :] = a[G:-G, G:G + 1, :]
        a[G:-G, G + ny:, :] = a[G:-G, G + ny - 1:G + ny, :]
        gr.fill_periodic(a, axes=(0,))
        U.append(a)
    idxe, idxv, _, _ = gr.wall_metrics(xe, True, True)
    idye, idyv, c2, c3 = gr.wall_metrics(ye, True, True)
    idz = np.full(nz + 2 * G, 1.0 / dz)
    mu_inf = rho_inf * U_inf * delta / re_delta
    ext2, ext3 = [0.0] * 6, [0.0] * 6
    ext2[2], ext3[2] = c2[0], c3[0]
    c = Case("shock_bl", nx, ny, nz, curvilinear=False, iorder_visc=iorder_visc, filter=filter, ducros=True,
             diffscale=mu_inf / float(sutherland(np.float64(T_inf))), chi=chi, dissip_shock=dissip_shock,
             csens=csens, cart=[idxe, idye, idz, idxv, idyv, idz.copy()], U=U, faces=(4, 3, 2, 3, 0, 0),
             adiabatic=adiabatic, t_wall=float(T[0, 0, 0]), cextrp2=tuple(ext2), cextrp3=tuple(ext3))
    c.p_ref = p_inf
    c.dt = timestep(U, min(dx, dz, float(np.min(np.diff(y)))), cfl)
    c.info = dict(mach=mach, wedge_deg=wedge_deg, beta_deg=math.degrees(b), re_delta=re_delta, delta=delta,
                  lengths=lengths, x_shock=x_shock, seed=seed, p_inf=p_inf, U_inf=U_inf)
    return c
