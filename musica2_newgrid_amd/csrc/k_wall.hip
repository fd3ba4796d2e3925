// Boundary faces of Cartesian blocks: no-slip walls on the j-faces, as the
// reference runs them with is_wall2 = F (the default,
// src/File_IO/read_param.f90:75), characteristic outflow at imax / jmax and
// the supersonic inlet at imin.
//
// k_wall        bc_wall_jmin/jmax_adiabatic_dpdn and _isotherm_dpdn
//               (src/Boundary/smod_bc_wall.f90:191-368): zero momentum on the
//               wall row, p (and T, adiabatic) extrapolated with dp/dn = 0 from
//               rows 2,3 of the primitives the reference computed last,
//               rho = p/(R T) (rocalc_pt_pfg, mod_eos_pfg.f90:303-317),
//               rhoE = rho cv T (ecalc_tro_pfg, mod_eos_pfg.f90:170-184).
//
// k_visc_wall   tau on rows 1,2 and n-1,n of every bounded i or j axis, with
//               the 5-pt one-sided gradients a04/a13 (a31/a40) of
//               grad_vel_5pts / grad_T_5pts along each bounded axis
//               (src/Compute_flow/smod_grad_vel.f90:12-100 and its j part,
//               smod_grad_T.f90).  An adiabatic wall has no wall-normal
//               temperature gradient on its row (smod_grad_T.f90:102-109).
//               On a wall row it stores tau/mu and -(cp/Pr) grad T: see
//               k_tau_wallrow.  With the shock filter, also the Ducros sensor
//               of those rows and p one row beyond each bounded face.
//
// k_tau_wallrow after the 2-layer exchange of tau: on the wall row, the
//               reference evaluates the viscous flux at every point, face
//               ghosts included, from that point's own primitives
//               (smod_flux_visc_5pts.f90:27-67) and the exchanged gradients.
//               The exchange runs before bc_wall, so a face ghost of the wall
//               row holds the neighbour's state from before the wall state was
//               imposed, unlike the neighbour's interior point.  With tau/mu
//               and the heat-flux gradient exchanged, this kernel rebuilds the
//               flux of each wall-row point from its own u, v, w and mu(T).
//               Likewise the pressure of the shock sensor and the spectral
//               radii of the artificial viscosities, which primitives_visc
//               computes on the ghosts from their own state
//               (smod_primitive_pfg.f90:84-86), not exchanged: an advanced
//               wall row (is_wall2) keeps the dissipation they weight.
//
// k_vdiv_wall   the viscous divergence on the same rows with one-sided
//               a04/a13 (a31/a40) along each bounded axis
//               (smod_flux_visc_5pts.f90:80-160).
//
// k_charac      the increments of the characteristic rows, bc_charac_imax and
//               bc_charac_jmax (src/Boundary/mod_charac.f90:350-469, 816-969;
//               perfect gas, mod_eos_pfg.f90:38-84), the sound speed the imax
//               routine reads, and the inflow row of bc_supersonic_inlet
//               (:1133-1146; init_BC_ref's state, mod_init_flow.f90:418-440).
//
// The main k_visc / k_vdiv compute the rim rows with the interior stencil
// first; these kernels overwrite them.  All of them touch thin slabs along
// the boundary faces, each thread one point.
#include "mgpu_internal.h"

namespace mgpu {

namespace {

constexpr int kWallThreads = 256;

struct Pw {
  double u, v, w, T, p;
};

__device__ __forceinline__ Pw prim_of(const double U[5], double gam1, double rg) {
  Pw q;
  const double r = U[0];
  const double vi = 1.0 / r;
  q.u = U[1] * vi;
  q.v = U[2] * vi;
  q.w = U[3] * vi;
  q.p = gam1 * (U[4] - 0.5 * r * (q.u * q.u + q.v * q.v + q.w * q.w));
  q.T = q.p * vi / rg;
  return q;
}

__device__ __forceinline__ Pw prim_at(const CState &U, idx_t p, double gam1, double rg) {
  double u[5];
#pragma unroll
  for (int v = 0; v < 5; v++) u[v] = ldg(U.v[v], p);
  return prim_of(u, gam1, rg);
}

// Primitives of a stage set that is stored as FP32 differences with ref.
__device__ __forceinline__ Pw prim_at32(const float *const *U, const double *ref, idx_t p, double gam1, double rg) {
  double u[5];
#pragma unroll
  for (int v = 0; v < 5; v++) u[v] = ref[v] + (double)ldg(U[v], p);
  return prim_of(u, gam1, rg);
}

__device__ __forceinline__ double sutherland(const Gas &gs, double T) {
  return gs.c0 * sqrt(T * T * T) / (T + gs.S0) * gs.diffscale;
}

// Coefficients of the one-sided derivative of row-slot q (0: row 1, 1: row 2,
// 2: row n-1, 3: row n) on the five rows r0..r0+4 of its side.
__device__ __forceinline__ void onesided(int q, double c[5]) {
  const double a04[5] = {A04_1, A04_2, A04_3, A04_4, A04_5};
  const double a13[5] = {A13_1, A13_2, A13_3, A13_4, A13_5};
#pragma unroll
  for (int m = 0; m < 5; m++) {
    if (q == 0) c[m] = a04[m];
    else if (q == 1) c[m] = a13[m];
    else if (q == 2) c[m] = -a13[4 - m];
    else c[m] = -a04[4 - m];
  }
}

// One-sided slot of row r (1..n) along an axis whose bounded faces are bnd:
// 0, 1 for rows 1, 2 next to a bounded min face, 2, 3 for rows n-1, n next to
// a bounded max face, -1 for the centred stencil.
__device__ __forceinline__ int oslot(int bnd, int n, int r) {
  if ((bnd & 1) && r <= 2) return r - 1;
  if ((bnd & 2) && r >= n - 1) return r == n ? 3 : 2;
  return -1;
}

// Point of the rim of a block handled by thread t: section J covers rows 1,2
// (ny-1,ny) of bounded j-faces over every i; section I rows 1,2 (nx-1,nx) of
// bounded i-faces over the j that section J leaves.  False if t maps to no
// point.
__device__ __forceinline__ bool rim_point(const Geom &g, long t, int &i, int &j, int &k) {
  const long nfj = (long)g.nx * g.nz, nfi = (long)g.ny * g.nz;
  if (t < 4 * nfj) {
    const int q = (int)(t / nfj);
    if (!((g.jbnd >> (q >> 1)) & 1)) return false;
    const long r = t - q * nfj;
    i = 1 + (int)(r % g.nx);
    k = 1 + (int)(r / g.nx);
    j = q == 0 ? 1 : (q == 1 ? 2 : (q == 2 ? g.ny - 1 : g.ny));
    return true;
  }
  t -= 4 * nfj;
  if (t >= 4 * nfi) return false;
  const int q = (int)(t / nfi);
  if (!((g.ibnd >> (q >> 1)) & 1)) return false;
  const long r = t - q * nfi;
  j = 1 + (int)(r % g.ny);
  k = 1 + (int)(r / g.ny);
  if (oslot(g.jbnd, g.ny, j) >= 0) return false;  // section J has it
  i = q == 0 ? 1 : (q == 1 ? 2 : (q == 2 ? g.nx - 1 : g.nx));
  return true;
}

// 5-pt first derivatives of u, v, w, T of U^n at (i, j, k) along index axis
// ax (stride st, row r of n): centred, or one-sided in slot q.
__device__ __forceinline__ void grad5(const CState &U, const Geom &g, idx_t p, idx_t st, int q, int r, int n,
                                      double gam1, double rg, double out[4]) {
  if (q < 0) {
    const Pw m2 = prim_at(U, p - 2 * st, gam1, rg), m1 = prim_at(U, p - st, gam1, rg);
    const Pw p1 = prim_at(U, p + st, gam1, rg), p2 = prim_at(U, p + 2 * st, gam1, rg);
    out[0] = A5_1 * (p1.u - m1.u) + A5_2 * (p2.u - m2.u);
    out[1] = A5_1 * (p1.v - m1.v) + A5_2 * (p2.v - m2.v);
    out[2] = A5_1 * (p1.w - m1.w) + A5_2 * (p2.w - m2.w);
    out[3] = A5_1 * (p1.T - m1.T) + A5_2 * (p2.T - m2.T);
    return;
  }
  (void)g;
  double c[5];
  onesided(q, c);
  const int r0 = q < 2 ? 1 : n - 4;
  const idx_t p0 = p - (idx_t)(r - r0) * st;
  out[0] = out[1] = out[2] = out[3] = 0.0;
#pragma unroll
  for (int m = 0; m < 5; m++) {
    const Pw y = prim_at(U, p0 + (idx_t)m * st, gam1, rg);
    out[0] = out[0] + c[m] * y.u;
    out[1] = out[1] + c[m] * y.v;
    out[2] = out[2] + c[m] * y.w;
    out[3] = out[3] + c[m] * y.T;
  }
}

// 5-pt first derivative of f along the index axis of stride st (row r of n):
// centred, or one-sided in slot q.
__device__ __forceinline__ double d5at(const double *f, idx_t p, idx_t st, int q, int r, int n) {
  if (q < 0) return A5_1 * (ldg(f, p + st) - ldg(f, p - st)) + A5_2 * (ldg(f, p + 2 * st) - ldg(f, p - 2 * st));
  double c[5];
  onesided(q, c);
  const int r0 = q < 2 ? 1 : n - 4;
  const idx_t p0 = p - (idx_t)(r - r0) * st;
  double d = 0.0;
#pragma unroll
  for (int m = 0; m < 5; m++) d = d + c[m] * ldg(f, p0 + (idx_t)m * st);
  return d;
}

// Advanced wall point (is_wall2) of the face (jsec ? j : i, side), row point
// r of its section, after the stage's update: smod_bc_wall_2.f90:11-473 and
// smod_bc_wall_2_c3.f90:363-712.  A slip face removes the normal momentum
// (Cartesian: the normal component; c3: m becomes (m.t_ksi) e_ksi +
// (m.t_phi) e_phi); a no-slip face zeroes the momentum and sets rhoE = rho cv
// T_wall with the row's advanced rho, T_wall extrapolated from rows 2, 3 (and
// on c3 walls the row's tangential neighbours) of the primitives of `src`
// (adiabatic) or the isothermal wall's.  At an edge of i- and j-walls the
// j-face's thread applies the i-face first, then its own (bc_wall's order,
// mod_bc_apply.f90:249-255).  With dst = the stash only the extrapolated
// temperatures are kept (j-face, then i-face of an edge); with src = the
// stash they are read back.
__device__ void wall_advanced(const WallArgs &a, bool jsec, int side, long r, int i, int k, idx_t p1, idx_t p2,
                              idx_t p3, idx_t sidx) {
  const Geom &g = a.g;
  const bool from_src = a.src.v[0] || a.src32[0];
  const bool to_stash = !a.dst.v[0] && !a.dst32[0];
  auto prim = [&](idx_t q) {
    return a.src.v[0] ? prim_at(a.src, q, a.gas.gam1, a.gas.rg) : prim_at32(a.src32, a.ref, q, a.gas.gam1, a.gas.rg);
  };
  // the i-face this j-wall point shares (an edge of walls on two axes), -1 if none
  const int es = !jsec ? -1 : (((g.iwall & 1) && i == 1) ? 0 : (((g.iwall & 2) && i == g.nx) ? 1 : -1));
  const bool slip = ((jsec ? a.jslip : a.islip) >> side) & 1;
  const bool slip_e = es >= 0 && ((a.islip >> es) & 1);
  double tw = a.t_wall, tw_e = a.t_wall;
  if (a.adiabatic) {
    if (from_src) {
      if (!slip) {
        const Pw q2 = prim(p2), q3 = prim(p3);
        if (jsec && a.wc3[side]) {
          const long nface = (long)g.nx * g.nz;
          const Pw xm = prim(p1 - 1), xp = prim(p1 + 1), zm = prim(p1 - g.sy), zp = prim(p1 + g.sy);
          const double gk = ldg(a.wc3[side], (idx_t)r), gp = ldg(a.wc3[side], (idx_t)(nface + r));
          tw = (1.0 / 3.0) * (4.0 * q2.T - q3.T + (xp.T - xm.T) * gk + (zp.T - zm.T) * gp);
        } else {
          tw = (jsec ? a.c2[side] : a.ci2[side]) * q2.T - (jsec ? a.c3[side] : a.ci3[side]) * q3.T;
        }
      }
      if (es >= 0 && !slip_e) {
        const idx_t e2 = es ? p1 - 1 : p1 + 1, e3 = es ? p1 - 2 : p1 + 2;
        tw_e = a.ci2[es] * prim(e2).T - a.ci3[es] * prim(e3).T;
      }
    } else {
      tw = ldg(a.stash, sidx);
      tw_e = ldg(a.stash, sidx + 1);
    }
  }
  if (to_stash) {
    stg(a.stash, sidx, tw);
    stg(a.stash, sidx + 1, tw_e);
    return;
  }
  double u[5];
#pragma unroll
  for (int v = 0; v < 5; v++) u[v] = a.dst.v[0] ? ldg(a.dst.v[v], p1) : a.ref[v] + (double)ldg(a.dst32[v], p1);
  bool chg[5] = {false, false, false, false, false};
  // one face's wall state on u: axis 0 (i) or 1 (j)
  auto apply = [&](bool sl, int axis, double t) {
    if (!sl) {
      u[1] = u[2] = u[3] = 0.0;
      u[4] = u[0] * (a.cv * t);
      chg[1] = chg[2] = chg[3] = chg[4] = true;
    } else if (axis == 1 && a.wt3[side]) {
      // c3 slip wall: the tangential part of the momentum
      const long nface = (long)g.nx * g.nz;
      const double *wc = a.wc3[side], *wt = a.wt3[side];
      double vk = 0.0, vp = 0.0;
#pragma unroll
      for (int c = 0; c < 3; c++) {
        vk = vk + u[1 + c] * ldg(wt, (idx_t)(c * nface + r));
        vp = vp + u[1 + c] * ldg(wt, (idx_t)((3 + c) * nface + r));
      }
#pragma unroll
      for (int c = 0; c < 3; c++) {
        u[1 + c] = vk * ldg(wc, (idx_t)((2 + c) * nface + r)) + vp * ldg(wc, (idx_t)((5 + c) * nface + r));
        chg[1 + c] = true;
      }
    } else {
      u[1 + axis] = 0.0;
      chg[1 + axis] = true;
    }
  };
  if (es >= 0) apply(slip_e, 0, tw_e);
  apply(slip, jsec ? 1 : 0, tw);
#pragma unroll
  for (int v = 1; v < 5; v++) {
    if (!chg[v]) continue;
    if (a.dst.v[0]) stg(a.dst.v[v], p1, u[v]);
    else stg(a.dst32[v], p1, (float)(u[v] - a.ref[v]));
  }
}

__global__ __launch_bounds__(kWallThreads) void k_wall(WallArgs a) {
  const Geom &g = a.g;
  const long nface = (long)g.nx * g.nz, nfi = (long)g.ny * g.nz;
  const long t = (long)blockIdx.x * kWallThreads + threadIdx.x;
  if (t >= 2 * nface + 2 * nfi) return;
  // section j (t < 2 nface): the point (i, k) of the jmin / jmax wall row;
  // section i: the point (j, k) of the imin / imax wall row, except the rows
  // of j-walls (their j-face state wins on the edge)
  const bool jsec = t < 2 * nface;
  const long ti = t - 2 * nface;
  const int side = jsec ? (int)(t / nface) : (int)(ti / nfi);
  if (!(((jsec ? g.jwall : g.iwall) >> side) & 1)) return;
  const long r = jsec ? t - side * nface : ti - side * nfi;
  int i, j, k;
  idx_t st;  // from the wall into the block
  if (jsec) {
    i = 1 + (int)(r % g.nx);
    k = 1 + (int)(r / g.nx);
    j = side ? g.ny : 1;
    st = (idx_t)g.sx;
  } else {
    j = 1 + (int)(r % g.ny);
    k = 1 + (int)(r / g.ny);
    if (((g.jwall & 1) && j == 1) || ((g.jwall & 2) && j == g.ny)) return;
    i = side ? g.nx : 1;
    st = 1;
  }
  const idx_t p1 = lin(g, i, j, k);
  const idx_t p2 = side ? p1 - st : p1 + st, p3 = side ? p1 - 2 * st : p1 + 2 * st;
  const idx_t sidx = (idx_t)(2 * t);  // stash slot: the sections follow each other
  if ((jsec ? g.jadv : g.iadv) & (1 << side)) {
    wall_advanced(a, jsec, side, r, i, k, p1, p2, p3, sidx);
    return;
  }
  double rho, rhoe;
  if (a.src.v[0] || a.src32[0]) {
    auto prim = [&](idx_t q) {
      return a.src.v[0] ? prim_at(a.src, q, a.gas.gam1, a.gas.rg) : prim_at32(a.src32, a.ref, q, a.gas.gam1, a.gas.rg);
    };
    const Pw q2 = prim(p2), q3 = prim(p3);
    double tw, pw;
    if (jsec && a.wc3[side]) {
      // c3 wall (bc_wall_jmin/jmax_*_dpdn_c3, smod_bc_wall_c3.f90:842-1671):
      // rows 2, 3 along eta plus the tangential terms with the wall row's own
      // neighbours along ksi and phi (gksigeta3, gphigeta3)
      const Pw xm = prim(p1 - 1), xp = prim(p1 + 1), zm = prim(p1 - g.sy), zp = prim(p1 + g.sy);
      const double gk = ldg(a.wc3[side], (idx_t)r), gp = ldg(a.wc3[side], (idx_t)(nface + r));
      tw = a.adiabatic ? (1.0 / 3.0) * (4.0 * q2.T - q3.T + (xp.T - xm.T) * gk + (zp.T - zm.T) * gp) : a.t_wall;
      pw = (1.0 / 3.0) * (4.0 * q2.p - q3.p + (xp.p - xm.p) * gk + (zp.p - zm.p) * gp);
    } else {
      const double c2 = jsec ? a.c2[side] : a.ci2[side], c3 = jsec ? a.c3[side] : a.ci3[side];
      tw = a.adiabatic ? c2 * q2.T - c3 * q3.T : a.t_wall;
      pw = c2 * q2.p - c3 * q3.p;
    }
    rho = pw / (a.gas.rg * tw);
    rhoe = rho * (a.cv * tw);
  } else {
    rho = ldg(a.stash, sidx);
    rhoe = ldg(a.stash, sidx + 1);
  }
  if (a.dst.v[0]) {
    stg(a.dst.v[0], p1, rho);
    stg(a.dst.v[1], p1, 0.0);
    stg(a.dst.v[2], p1, 0.0);
    stg(a.dst.v[3], p1, 0.0);
    stg(a.dst.v[4], p1, rhoe);
  } else if (a.dst32[0]) {
    const double w[5] = {rho, 0.0, 0.0, 0.0, rhoe};
#pragma unroll
    for (int v = 0; v < 5; v++) stg(a.dst32[v], p1, (float)(w[v] - a.ref[v]));
  } else {
    stg(a.stash, sidx, rho);
    stg(a.stash, sidx + 1, rhoe);
  }
}

__global__ __launch_bounds__(kWallThreads) void k_visc_wall(WallViscArgs a) {
  const Geom &g = a.g;
  const long nrim = 4L * g.nx * g.nz + 4L * g.ny * g.nz;
  long t = (long)blockIdx.x * kWallThreads + threadIdx.x;
  const double gam1 = a.gas.gam1, rg = a.gas.rg;
  if (t >= nrim) {
    // p of U^n one row beyond each bounded face, which the pressure sensor of
    // the boundary row reads: the ghost state, which nothing on the path writes
    if (!a.want_sensor) return;
    t -= nrim;
    const long nfi = (long)g.ny * g.nz, nfj = (long)g.nx * g.nz;
    int i, j, k;
    if (t < 2 * nfi) {
      const int side = (int)(t / nfi);
      if (!((g.ibnd >> side) & 1)) return;
      const long r = t - side * nfi;
      j = 1 + (int)(r % g.ny);
      k = 1 + (int)(r / g.ny);
      i = side ? g.nx + 1 : 0;
    } else {
      t -= 2 * nfi;
      if (t >= 2 * nfj) return;
      const int side = (int)(t / nfj);
      if (!((g.jbnd >> side) & 1)) return;
      const long r = t - side * nfj;
      i = 1 + (int)(r % g.nx);
      k = 1 + (int)(r / g.nx);
      j = side ? g.ny + 1 : 0;
    }
    const idx_t pg = lin(g, i, j, k);
    stg(a.prs, pg, prim_at(a.u0, pg, gam1, rg).p);
    return;
  }
  int i, j, k;
  if (!rim_point(g, t, i, j, k)) return;
  const idx_t p = lin(g, i, j, k);
  const int qi = oslot(g.ibnd, g.nx, i), qj = oslot(g.jbnd, g.ny, j);

  // no-slip wall rows (of a j- or an i-wall)
  const bool jrow = ((g.jwall & 1) && j == 1) || ((g.jwall & 2) && j == g.ny);
  const bool irow = ((g.iwall & 1) && i == 1) || ((g.iwall & 2) && i == g.nx);
  const bool wrow = jrow || irow;
  // gradients gr[u,v,w,T][x,y,z]
  double gr[4][3];
  {
    double gx[4], gy[4], gz[4];
    grad5(a.u0, g, p, 1, qi, i, g.nx, gam1, rg, gx);
    grad5(a.u0, g, p, (idx_t)g.sx, qj, j, g.ny, gam1, rg, gy);
    grad5(a.u0, g, p, g.sy, -1, k, g.nz, gam1, rg, gz);
    if (a.geo == kGeoC3) {
      // index-space derivatives mapped with the viscous metrics
      // (smod_grad_vel_c3.f90:260-300, smod_grad_T_c3.f90:200-213)
      double m[10];
#pragma unroll
      for (int q = 0; q < 10; q++) m[q] = ldg(a.met.mv[q], p);
#pragma unroll
      for (int n = 0; n < 4; n++) {
        const double dvx = gx[n] * m[0] + gy[n] * m[3] + gz[n] * m[6];
        const double dvy = gx[n] * m[1] + gy[n] * m[4] + gz[n] * m[7];
        const double dvz = gx[n] * m[2] + gy[n] * m[5] + gz[n] * m[8];
        gr[n][0] = dvx * m[9];
        gr[n][1] = dvy * m[9];
        gr[n][2] = dvz * m[9];
      }
      if (a.adiabatic && wrow) {
        // adiabatic c3 wall: the index-space derivatives along ksi and phi
        // times the unit tangents, as written (smod_grad_T_c3.f90:249-275)
        const double *wc = a.wc3[j == 1 ? 0 : 1];
        // (c3 blocks have walls on their j-faces only: wrow is a j row)
        const idx_t nf = (idx_t)g.nx * (idx_t)g.nz, q = (idx_t)(i - 1) + (idx_t)g.nx * (idx_t)(k - 1);
#pragma unroll
        for (int c = 0; c < 3; c++)
          gr[3][c] = gx[3] * ldg(wc, (idx_t)(2 + c) * nf + q) + gz[3] * ldg(wc, (idx_t)(5 + c) * nf + q);
      }
    } else {
      const double ix = ldg(a.met.idx_v, (idx_t)(i - 1 + G)), iy = ldg(a.met.idy_v, (idx_t)(j - 1 + G)),
                   iz = ldg(a.met.idz_v, (idx_t)(k - 1 + G));
#pragma unroll
      for (int n = 0; n < 4; n++) {
        gr[n][0] = gx[n] * ix;
        gr[n][1] = gy[n] * iy;
        gr[n][2] = gz[n] * iz;
      }
      // an adiabatic wall has no wall-normal temperature gradient on its row
      // (smod_grad_T.f90:102-109 and its i twin)
      if (a.adiabatic && jrow) gr[3][1] = 0.0;
      if (a.adiabatic && irow) gr[3][0] = 0.0;
    }
  }
  const Pw o = prim_at(a.u0, p, gam1, rg);
  if (a.want_sensor) {
    // the shock filter's Ducros sensor from these gradients
    double s = 1.0;
    if (a.ducros) {
      const double dv = gr[0][0] + gr[1][1] + gr[2][2];
      const double div2 = dv * dv;
      const double o1 = gr[2][1] - gr[1][2];
      const double o2 = gr[0][2] - gr[2][0];
      const double o3 = gr[1][0] - gr[0][1];
      s = div2 / (div2 + o1 * o1 + o2 * o2 + o3 * o3 + 1.0e-16);
    }
    stg(a.sens, p, s);
  }
  // the wall row stores tau/mu and -(cp/Pr) grad T (k_tau_wallrow completes it)
  const double mu_l = wrow ? 1.0 : sutherland(a.gas, o.T);
  const double ck = a.gas.cok_ratio * mu_l;
  double t11 = gr[0][0], t22 = gr[1][1], t33 = gr[2][2];
  double t12 = 0.5 * (gr[0][1] + gr[1][0]);
  double t13 = 0.5 * (gr[0][2] + gr[2][0]);
  double t23 = 0.5 * (gr[1][2] + gr[2][1]);
  const double tr = (1.0 / 3.0) * (t11 + t22 + t33);
  const double mu = -2.0 * mu_l;
  t11 = mu * (t11 - tr);
  t22 = mu * (t22 - tr);
  t33 = mu * (t33 - tr);
  t12 = mu * t12;
  t13 = mu * t13;
  t23 = mu * t23;
  stg(a.tau[0], p, t11);
  stg(a.tau[1], p, t12);
  stg(a.tau[2], p, t13);
  stg(a.tau[3], p, t22);
  stg(a.tau[4], p, t23);
  stg(a.tau[5], p, t33);
  if (wrow) {
    stg(a.tau[6], p, -ck * gr[3][0]);
    stg(a.tau[7], p, -ck * gr[3][1]);
    stg(a.tau[8], p, -ck * gr[3][2]);
  } else {
    stg(a.tau[6], p, (o.u * t11 + o.v * t12 + o.w * t13) - ck * gr[3][0]);
    stg(a.tau[7], p, (o.u * t12 + o.v * t22 + o.w * t23) - ck * gr[3][1]);
    stg(a.tau[8], p, (o.u * t13 + o.v * t23 + o.w * t33) - ck * gr[3][2]);
  }
}

__global__ __launch_bounds__(kWallThreads) void k_tau_wallrow(WallViscArgs a) {
  const Geom &g = a.g;
  const int ex = g.nx + 4, ey = g.ny + 4, ez = g.nz + 4;
  const long nrow = (long)ex * ez, nrowi = (long)ey * ez;
  const long t = (long)blockIdx.x * kWallThreads + threadIdx.x;
  if (t >= 2 * nrow + 2 * nrowi) return;
  int i, j, k;
  if (t < 2 * nrow) {
    // j-wall rows, over i = -1..nx+2 and k = -1..nz+2
    const int side = (int)(t / nrow);
    if (!((g.jwall >> side) & 1)) return;
    const long r = t - side * nrow;
    i = -1 + (int)(r % ex);
    k = -1 + (int)(r / ex);
    const bool in_i = i >= 1 && i <= g.nx, in_k = k >= 1 && k <= g.nz;
    if (!in_i && !in_k) return;  // edge ghosts: never read
    // beyond a bounded i-face: never read (the i-derivatives are one-sided there)
    if (((g.ibnd & 1) && i < 1) || ((g.ibnd & 2) && i > g.nx)) return;
    j = side ? g.ny : 1;
  } else {
    // i-wall rows, over j = -1..ny+2 and k = -1..nz+2, less the points of
    // j-wall rows (done above)
    const long ti = t - 2 * nrow;
    const int side = (int)(ti / nrowi);
    if (!((g.iwall >> side) & 1)) return;
    const long r = ti - side * nrowi;
    j = -1 + (int)(r % ey);
    k = -1 + (int)(r / ey);
    const bool in_j = j >= 1 && j <= g.ny, in_k = k >= 1 && k <= g.nz;
    if (!in_j && !in_k) return;
    if (((g.jbnd & 1) && j < 1) || ((g.jbnd & 2) && j > g.ny)) return;
    if (((g.jwall & 1) && j == 1) || ((g.jwall & 2) && j == g.ny)) return;
    i = side ? g.nx : 1;
  }
  const idx_t p = lin(g, i, j, k);
  const Pw o = prim_at(a.u0, p, a.gas.gam1, a.gas.rg);
  if (a.want_sensor) stg(a.prs, p, o.p);
  if (a.sr[0]) {
    const double c = sqrt(a.gas.gam * a.gas.rg * o.T);
    if (a.geo == kGeoC3) {
      // the contravariant velocity and |grad xi| of the Euler metrics, as k_visc
      const double ij = ldg(a.met.m[9], p);
#pragma unroll
      for (int d = 0; d < 3; d++) {
        const double mx = ldg(a.met.m[3 * d], p), my = ldg(a.met.m[3 * d + 1], p), mz = ldg(a.met.m[3 * d + 2], p);
        const double vc = (o.u * mx + o.v * my + o.w * mz) * ij;
        stg(a.sr[d], p, fabs(vc) + c * (sqrt(mx * mx + my * my + mz * mz) * fabs(ij)));
      }
    } else {
      stg(a.sr[0], p, fabs(o.u) + c);
      stg(a.sr[1], p, fabs(o.v) + c);
      stg(a.sr[2], p, fabs(o.w) + c);
    }
  }
  const double mu = sutherland(a.gas, o.T);
  double tv[6];
#pragma unroll
  for (int c = 0; c < 6; c++) {
    tv[c] = mu * ldg(a.tau[c], p);
    stg(a.tau[c], p, tv[c]);
  }
  stg(a.tau[6], p, (o.u * tv[0] + o.v * tv[1] + o.w * tv[2]) + mu * ldg(a.tau[6], p));
  stg(a.tau[7], p, (o.u * tv[1] + o.v * tv[3] + o.w * tv[4]) + mu * ldg(a.tau[7], p));
  stg(a.tau[8], p, (o.u * tv[2] + o.v * tv[4] + o.w * tv[5]) + mu * ldg(a.tau[8], p));
}

__global__ __launch_bounds__(kWallThreads) void k_vdiv_wall(WallViscArgs a) {
  const Geom &g = a.g;
  const long t = (long)blockIdx.x * kWallThreads + threadIdx.x;
  int i, j, k;
  if (!rim_point(g, t, i, j, k)) return;
  const int qi = oslot(g.ibnd, g.nx, i), qj = oslot(g.jbnd, g.ny, j);
  const idx_t p = lin(g, i, j, k);
  if (a.geo == kGeoC3) {
    // the fluxes of tau projected on the viscous metric row of each direction
    // at every stencil point, their 5-pt differences times ijacob3_v
    // (smod_flux_visc_5pts_c3.f90)
    const idx_t st[3] = {1, (idx_t)g.sx, g.sy};
    const int qd[3] = {qi, qj, -1}, rd[3] = {i, j, k}, nd[3] = {g.nx, g.ny, g.nz};
    double kv[4] = {0.0, 0.0, 0.0, 0.0};
    const double ij = ldg(a.met.mv[9], p);
#pragma unroll
    for (int d = 0; d < 3; d++) {
      // the five stencil points: rows r0..r0+4 (one-sided) or r-2..r+2 (centred)
      const int r0 = qd[d] < 0 ? rd[d] - 2 : (qd[d] < 2 ? 1 : nd[d] - 4);
      double f[5][4];
#pragma unroll
      for (int m = 0; m < 5; m++) {
        const idx_t pm = p + (idx_t)(r0 + m - rd[d]) * st[d];
        double t[NTAU];
#pragma unroll
        for (int c = 0; c < NTAU; c++) t[c] = ldg(a.tau[c], pm);
        const double mx = ldg(a.met.mv[3 * d], pm), my = ldg(a.met.mv[3 * d + 1], pm),
                     mz = ldg(a.met.mv[3 * d + 2], pm);
        f[m][0] = t[0] * mx + t[1] * my + t[2] * mz;
        f[m][1] = t[1] * mx + t[3] * my + t[4] * mz;
        f[m][2] = t[2] * mx + t[4] * my + t[5] * mz;
        f[m][3] = t[6] * mx + t[7] * my + t[8] * mz;
      }
      double c5[5];
      if (qd[d] >= 0) onesided(qd[d], c5);
#pragma unroll
      for (int n = 0; n < 4; n++) {
        double dv;
        if (qd[d] < 0) {
          dv = A5_1 * (f[3][n] - f[1][n]) + A5_2 * (f[4][n] - f[0][n]);
        } else {
          dv = 0.0;
#pragma unroll
          for (int m = 0; m < 5; m++) dv = dv + c5[m] * f[m][n];
        }
        kv[n] = dv * ij + kv[n];
      }
    }
#pragma unroll
    for (int n = 0; n < 4; n++) stg(a.kv[n], p, kv[n]);
    return;
  }
  const double mx = ldg(a.met.idx_v, (idx_t)(i - 1 + G)), my = ldg(a.met.idy_v, (idx_t)(j - 1 + G)),
               mz = ldg(a.met.idz_v, (idx_t)(k - 1 + G));
#pragma unroll
  for (int n = 0; n < 4; n++) {
    const double *fx = a.tau[krow(0, n)], *fy = a.tau[krow(1, n)], *fz = a.tau[krow(2, n)];
    double kv = d5at(fx, p, 1, qi, i, g.nx) * mx;
    kv = d5at(fy, p, (idx_t)g.sx, qj, j, g.ny) * my + kv;
    kv = d5at(fz, p, g.sy, -1, k, g.nz) * mz + kv;
    stg(a.kv[n], p, kv);
  }
}

// One-sided 5-pt derivative a40 at the last row n of an axis (stride st) of
// the values f[0..4] at rows n-4..n, times the metric.
__device__ __forceinline__ double d40(const double f[5]) {
  return (-A04_5) * f[0] + (-A04_4) * f[1] + (-A04_3) * f[2] + (-A04_2) * f[3] + (-A04_1) * f[4];
}

// Conserved variables of the state at element p.
__device__ __forceinline__ void cons_at(const CharArgs &a, idx_t p, double u[5]) {
  if (a.src.v[0]) {
#pragma unroll
    for (int v = 0; v < 5; v++) u[v] = ldg(a.src.v[v], p);
  } else {
#pragma unroll
    for (int v = 0; v < 5; v++) u[v] = a.ref[v] + (double)ldg(a.src32[v], p);
  }
}

// The characteristic increment of a row point from the primitives of the
// rows n-4..n along its normal (LODI, perfect gas: cp/av = cp T).
// Section imax: (j, k) of the imax row; section jmax: (i, k) of the jmax row.
__device__ void charac_point(const CharArgs &a, bool imax, int i, int j, int k, double *kout, idx_t sl, idx_t ss) {
  const Geom &g = a.g;
  const idx_t st = imax ? (idx_t)1 : (idx_t)g.sx;
  const idx_t p = lin(g, i, j, k);
  double fu[5], fv[5], fw[5], fp[5], fr[5], ue[5];
#pragma unroll
  for (int m = 0; m < 5; m++) {
    double u[5];
    cons_at(a, p - (idx_t)(4 - m) * st, u);
    const Pw q = prim_of(u, a.gas.gam1, a.gas.rg);
    fu[m] = q.u;
    fv[m] = q.v;
    fw[m] = q.w;
    fp[m] = q.p;
    fr[m] = u[0];
    if (m == 4) {
#pragma unroll
      for (int v = 0; v < 5; v++) ue[v] = u[v];
    }
  }
  const double met = imax ? ldg(a.idx_v, (idx_t)(i - 1 + G)) : ldg(a.idy_v, (idx_t)(j - 1 + G));
  const double du = d40(fu) * met, dv = d40(fv) * met, dw = d40(fw) * met, dp = d40(fp) * met,
               dr = d40(fr) * met;
  const double rho = ue[0], u = fu[4], v = fv[4], w = fw[4], pr = fp[4];
  const double T = pr * (1.0 / rho) / a.gas.rg;
  double rt, ut, vt, wt, pt, c2;
  if (imax) {
    // bc_charac_imax (mod_charac.f90:350-469): the sound speed that
    // primitives_visc last computed; subsonic outflow relaxes p towards
    // 0.528 p_ref
    const double cj = ldg(a.cst, sl);
    c2 = cj * cj;
    const double upc = u + cj, umc = u - cj;
    const double L1 = umc <= 0.0 ? 0.1 * (pr - 0.528 * a.p_ref) : umc * (dp - rho * cj * du);
    const double L5 = upc <= 0.0 ? 0.0 : upc * (dp + rho * cj * du);
    const double L2 = u * (dp - c2 * dr), L3 = u * dv, L4 = u * dw;
    pt = 0.5 * (L5 + L1);
    rt = -(pt + L2) / c2;
    ut = 0.5 * (L5 - L1) / (rho * cj);
    vt = L3;
    wt = L4;
  } else {
    // bc_charac_jmax (mod_charac.f90:816-969): c of the current primitives
    const double cc = sqrt(a.gas.gam * a.gas.rg * T);
    c2 = cc * cc;
    double L1 = 0.0, L2, L3, L4;
    double L5 = (v + cc) * (dp + rho * cc * dv);
    if (v <= 0.0) {
      L2 = L3 = L4 = 0.0;
      if (v <= -cc) L5 = 0.0;
    } else {
      L2 = v * du;
      L3 = v * (dp - c2 * dr);
      L4 = v * dw;
      if (v >= cc) L1 = (v - cc) * (dp - rho * cc * dv);
    }
    pt = 0.5 * (L5 + L1);
    rt = (pt + L3) / c2;
    ut = L2;
    vt = (L5 - L1) / (2.0 * rho * cc);
    wt = L4;
  }
  const double cpav = a.cp * T;
  stg(kout, sl, rt);
  stg(kout, ss + sl, rho * ut + u * rt);
  stg(kout, 2 * ss + sl, rho * vt + v * rt);
  stg(kout, 3 * ss + sl, rho * wt + w * rt);
  stg(kout, 4 * ss + sl, cpav * (pt / c2 - rt) + (ue[4] + pr) / rho * rt + rho * (u * ut + v * vt + w * wt));
}

// Sound speed sqrt(gam rg T) of the FP64 state U at p (primitives_visc).
__device__ __forceinline__ double sound(const CState &U, idx_t p, const Gas &gs) {
  const Pw q = prim_at(U, p, gs.gam1, gs.rg);
  return sqrt(gs.gam * gs.rg * q.T);
}

__global__ __launch_bounds__(kWallThreads) void k_charac(CharArgs a) {
  const Geom &g = a.g;
  const long nfi = (long)g.ny * g.nz, nfj = (long)g.nx * g.nz;
  long t = (long)blockIdx.x * kWallThreads + threadIdx.x;
  if (t < nfi) {
    // the imax row (every j)
    if (!a.kimax) return;
    const int j = 1 + (int)(t % g.ny), k = 1 + (int)(t / g.ny);
    const idx_t sl = (idx_t)t, p = lin(g, g.nx, j, k);
    if (a.cst_mode == 3) {
      stg(a.cst, sl, sound(a.cnew, p, a.gas));
      return;
    }
    if (a.cst_mode == 2) {
      double u[5];
      cons_at(a, p, u);
      const Pw q = prim_of(u, a.gas.gam1, a.gas.rg);
      stg(a.cst, sl, sqrt(a.gas.gam * a.gas.rg * q.T));
    }
    charac_point(a, true, g.nx, j, k, a.kimax, sl, (idx_t)nfi);
    if (a.cst_mode == 1) stg(a.cst, sl, sound(a.cnew, p, a.gas));
    return;
  }
  t -= nfi;
  if (t < nfj) {
    // the jmax row, i = i0..i1
    if (!a.kjmax || a.cst_mode == 3) return;
    const int i = 1 + (int)(t % g.nx), k = 1 + (int)(t / g.nx);
    if (i < a.i0 || i > a.i1) return;
    charac_point(a, false, i, g.ny, k, a.kjmax, (idx_t)t, (idx_t)nfj);
    return;
  }
  t -= nfj;
  if (t < nfi) {
    // the inlet row i = 1 (every j)
    if (!a.uin || !a.inlet_mode) return;
    const int j = 1 + (int)(t % g.ny), k = 1 + (int)(t / g.ny);
    const idx_t sl = (idx_t)t, ss = (idx_t)nfi, p = lin(g, 1, j, k);
    double u[5];
    cons_at(a, p, u);
#pragma unroll
    for (int v = 0; v < 5; v++) stg(a.ppre, (idx_t)v * ss + sl, u[v]);
    if (a.inlet_mode == 1) {
      // bc_supersonic_inlet (mod_charac.f90:1133-1146): the inflow state
#pragma unroll
      for (int v = 0; v < 5; v++) stg(a.row1.v[v], p, ldg(a.uin, (idx_t)v * ss + sl));
    } else {
#pragma unroll
      for (int v = 0; v < 5; v++) {
        stg(a.uin, (idx_t)v * ss + sl, u[v]);
        stg(a.kf, (idx_t)v * ss + sl, 0.0);
      }
    }
  }
}

__global__ __launch_bounds__(kWallThreads) void k_inlet_kf(Geom g, CState after, CState before, double *kf) {
  const long nfi = (long)g.ny * g.nz;
  const long t = (long)blockIdx.x * kWallThreads + threadIdx.x;
  if (t >= nfi) return;
  const int j = 1 + (int)(t % g.ny), k = 1 + (int)(t / g.ny);
  const idx_t p = lin(g, 1, j, k);
#pragma unroll
  for (int v = 0; v < 5; v++) stg(kf, (idx_t)v * (idx_t)nfi + (idx_t)t, ldg(after.v[v], p) - ldg(before.v[v], p));
}

inline dim3 grid_for(long n) { return dim3((unsigned)((n + kWallThreads - 1) / kWallThreads)); }

}  // namespace

hipError_t launch_wall(const WallArgs &a, hipStream_t s) {
  const long n = 2L * a.g.nx * a.g.nz + 2L * a.g.ny * a.g.nz;
  hipLaunchKernelGGL(k_wall, grid_for(n), dim3(kWallThreads), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_visc_wall(const WallViscArgs &a, hipStream_t s) {
  const long n = 6L * a.g.nx * a.g.nz + 6L * a.g.ny * a.g.nz;
  hipLaunchKernelGGL(k_visc_wall, grid_for(n), dim3(kWallThreads), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_tau_wallrow(const WallViscArgs &a, hipStream_t s) {
  const long n = 2L * (a.g.nx + 4) * (a.g.nz + 4) + 2L * (a.g.ny + 4) * (a.g.nz + 4);
  hipLaunchKernelGGL(k_tau_wallrow, grid_for(n), dim3(kWallThreads), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_vdiv_wall(const WallViscArgs &a, hipStream_t s) {
  const long n = 4L * a.g.nx * a.g.nz + 4L * a.g.ny * a.g.nz;
  hipLaunchKernelGGL(k_vdiv_wall, grid_for(n), dim3(kWallThreads), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_charac(const CharArgs &a, hipStream_t s) {
  const long n = 2L * a.g.ny * a.g.nz + (long)a.g.nx * a.g.nz;
  hipLaunchKernelGGL(k_charac, grid_for(n), dim3(kWallThreads), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_inlet_kf(const Geom &g, const CState &after, const CState &before, double *kf, hipStream_t s) {
  const long n = (long)g.ny * g.nz;
  hipLaunchKernelGGL(k_inlet_kf, grid_for(n), dim3(kWallThreads), 0, s, g, after, before, kf);
  return hipGetLastError();
}

}  // namespace mgpu
