// No-slip walls on the j-faces of Cartesian blocks, as the reference runs them
// with is_wall2 = F (the default, src/File_IO/read_param.f90:75):
//
// k_wall        bc_wall_jmin/jmax_adiabatic_dpdn and _isotherm_dpdn
//               (src/Boundary/smod_bc_wall.f90:191-368): zero momentum on the
//               wall row, p (and T, adiabatic) extrapolated with dp/dn = 0 from
//               rows 2,3 of the primitives the reference computed last,
//               rho = p/(R T) (rocalc_pt_pfg, mod_eos_pfg.f90:303-317),
//               rhoE = rho cv T (ecalc_tro_pfg, mod_eos_pfg.f90:170-184).
//
// k_visc_wall   tau on rows 1,2 (ny-1,ny) with the 5-pt one-sided j-gradients
//               a04/a13 (a31/a40) of grad_vel_5pts / grad_T_5pts
//               (src/Compute_flow/smod_grad_vel.f90:12-100 and its j part,
//               smod_grad_T.f90).  An adiabatic wall has no wall-normal
//               temperature gradient on its row (smod_grad_T.f90:102-109).
//               On the wall row itself it stores tau/mu and -(cp/Pr) grad T:
//               see k_tau_wallrow.
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
//
// k_vdiv_wall   the viscous divergence on rows 1,2 (ny-1,ny) with one-sided
//               a04/a13 (a31/a40) along j (smod_flux_visc_5pts.f90:80-160, j part).
//
// The main k_visc / k_vdiv compute those rows with the interior stencil
// first; these kernels overwrite them.  All of them touch one thin slab per
// wall face, each thread one point.
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

// Coefficients of the one-sided j-derivative of row-slot q (0: row 1, 1: row 2,
// 2: row ny-1, 3: row ny) on the five rows r0..r0+4 of its side.
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

// Row slot q -> (row j, first row r0 of its one-sided stencil); false if
// that side has no wall.
__device__ __forceinline__ bool slot_row(const Geom &g, int q, int &j, int &r0) {
  const int side = q >> 1;
  if (!((g.jwall >> side) & 1)) return false;
  j = q == 0 ? 1 : (q == 1 ? 2 : (q == 2 ? g.ny - 1 : g.ny));
  r0 = side ? g.ny - 4 : 1;
  return true;
}

__global__ __launch_bounds__(kWallThreads) void k_wall(WallArgs a) {
  const Geom &g = a.g;
  const long nface = (long)g.nx * g.nz;
  const long t = (long)blockIdx.x * kWallThreads + threadIdx.x;
  if (t >= 2 * nface) return;
  const int side = (int)(t / nface);
  if (!((g.jwall >> side) & 1)) return;
  const long r = t - side * nface;
  const int i = 1 + (int)(r % g.nx), k = 1 + (int)(r / g.nx);
  const int j1 = side ? g.ny : 1, dj = side ? -1 : 1;
  double rho, rhoe;
  if (a.src.v[0] || a.src32[0]) {
    const idx_t p2 = lin(g, i, j1 + dj, k), p3 = lin(g, i, j1 + 2 * dj, k);
    const Pw q2 = a.src.v[0] ? prim_at(a.src, p2, a.gas.gam1, a.gas.rg)
                             : prim_at32(a.src32, a.ref, p2, a.gas.gam1, a.gas.rg);
    const Pw q3 = a.src.v[0] ? prim_at(a.src, p3, a.gas.gam1, a.gas.rg)
                             : prim_at32(a.src32, a.ref, p3, a.gas.gam1, a.gas.rg);
    const double c2 = a.c2[side], c3 = a.c3[side];
    const double tw = a.adiabatic ? c2 * q2.T - c3 * q3.T : a.t_wall;
    const double pw = c2 * q2.p - c3 * q3.p;
    rho = pw / (a.gas.rg * tw);
    rhoe = rho * (a.cv * tw);
  } else {
    rho = ldg(a.stash, (idx_t)(2 * t));
    rhoe = ldg(a.stash, (idx_t)(2 * t + 1));
  }
  if (a.dst.v[0]) {
    const idx_t p1 = lin(g, i, j1, k);
    stg(a.dst.v[0], p1, rho);
    stg(a.dst.v[1], p1, 0.0);
    stg(a.dst.v[2], p1, 0.0);
    stg(a.dst.v[3], p1, 0.0);
    stg(a.dst.v[4], p1, rhoe);
  } else if (a.dst32[0]) {
    const idx_t p1 = lin(g, i, j1, k);
    const double w[5] = {rho, 0.0, 0.0, 0.0, rhoe};
#pragma unroll
    for (int v = 0; v < 5; v++) stg(a.dst32[v], p1, (float)(w[v] - a.ref[v]));
  } else {
    stg(a.stash, (idx_t)(2 * t), rho);
    stg(a.stash, (idx_t)(2 * t + 1), rhoe);
  }
}

__global__ __launch_bounds__(kWallThreads) void k_visc_wall(WallViscArgs a) {
  const Geom &g = a.g;
  const long nface = (long)g.nx * g.nz;
  const long t = (long)blockIdx.x * kWallThreads + threadIdx.x;
  if (t >= 4 * nface) return;
  const int q = (int)(t / nface);
  int j, r0;
  if (!slot_row(g, q, j, r0)) return;
  const long r = t - q * nface;
  const int i = 1 + (int)(r % g.nx), k = 1 + (int)(r / g.nx);
  const double gam1 = a.gas.gam1, rg = a.gas.rg;

  // gradients gr[u,v,w,T][x,y,z]
  double gr[4][3];
  {
    const Pw xm2 = prim_at(a.u0, lin(g, i - 2, j, k), gam1, rg), xm1 = prim_at(a.u0, lin(g, i - 1, j, k), gam1, rg);
    const Pw xp1 = prim_at(a.u0, lin(g, i + 1, j, k), gam1, rg), xp2 = prim_at(a.u0, lin(g, i + 2, j, k), gam1, rg);
    const Pw zm2 = prim_at(a.u0, lin(g, i, j, k - 2), gam1, rg), zm1 = prim_at(a.u0, lin(g, i, j, k - 1), gam1, rg);
    const Pw zp1 = prim_at(a.u0, lin(g, i, j, k + 1), gam1, rg), zp2 = prim_at(a.u0, lin(g, i, j, k + 2), gam1, rg);
    const double ix = ldg(a.met.idx_v, (idx_t)(i - 1 + G)), iy = ldg(a.met.idy_v, (idx_t)(j - 1 + G)),
                 iz = ldg(a.met.idz_v, (idx_t)(k - 1 + G));
    gr[0][0] = (A5_1 * (xp1.u - xm1.u) + A5_2 * (xp2.u - xm2.u)) * ix;
    gr[1][0] = (A5_1 * (xp1.v - xm1.v) + A5_2 * (xp2.v - xm2.v)) * ix;
    gr[2][0] = (A5_1 * (xp1.w - xm1.w) + A5_2 * (xp2.w - xm2.w)) * ix;
    gr[3][0] = (A5_1 * (xp1.T - xm1.T) + A5_2 * (xp2.T - xm2.T)) * ix;
    gr[0][2] = (A5_1 * (zp1.u - zm1.u) + A5_2 * (zp2.u - zm2.u)) * iz;
    gr[1][2] = (A5_1 * (zp1.v - zm1.v) + A5_2 * (zp2.v - zm2.v)) * iz;
    gr[2][2] = (A5_1 * (zp1.w - zm1.w) + A5_2 * (zp2.w - zm2.w)) * iz;
    gr[3][2] = (A5_1 * (zp1.T - zm1.T) + A5_2 * (zp2.T - zm2.T)) * iz;
    double c[5];
    onesided(q, c);
    double gy[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int m = 0; m < 5; m++) {
      const Pw y = prim_at(a.u0, lin(g, i, r0 + m, k), gam1, rg);
      gy[0] = gy[0] + c[m] * y.u;
      gy[1] = gy[1] + c[m] * y.v;
      gy[2] = gy[2] + c[m] * y.w;
      gy[3] = gy[3] + c[m] * y.T;
    }
#pragma unroll
    for (int n = 0; n < 4; n++) gr[n][1] = gy[n] * iy;
    if (a.adiabatic && (q == 0 || q == 3)) gr[3][1] = 0.0;
  }
  const idx_t p = lin(g, i, j, k);
  const Pw o = prim_at(a.u0, p, gam1, rg);
  const bool wrow = q == 0 || q == 3;
  if (a.want_sensor) {
    // the shock filter's Ducros sensor from these gradients; its pressure
    // sensor on the wall row reads p one row beyond the wall, where the state
    // is the ghost layer that nothing on the path writes
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
    if (wrow) {
      const idx_t pg = lin(g, i, q == 0 ? 0 : g.ny + 1, k);
      stg(a.prs, pg, prim_at(a.u0, pg, gam1, rg).p);
    }
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
  const int ex = g.nx + 4, ez = g.nz + 4;
  const long nrow = (long)ex * ez;
  const long t = (long)blockIdx.x * kWallThreads + threadIdx.x;
  if (t >= 2 * nrow) return;
  const int side = (int)(t / nrow);
  if (!((g.jwall >> side) & 1)) return;
  const long r = t - side * nrow;
  const int i = -1 + (int)(r % ex), k = -1 + (int)(r / ex);
  const bool in_i = i >= 1 && i <= g.nx, in_k = k >= 1 && k <= g.nz;
  if (!in_i && !in_k) return;  // edge ghosts: never read
  const int j = side ? g.ny : 1;
  const idx_t p = lin(g, i, j, k);
  const Pw o = prim_at(a.u0, p, a.gas.gam1, a.gas.rg);
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
  const long nface = (long)g.nx * g.nz;
  const long t = (long)blockIdx.x * kWallThreads + threadIdx.x;
  if (t >= 4 * nface) return;
  const int q = (int)(t / nface);
  int j, r0;
  if (!slot_row(g, q, j, r0)) return;
  const long r = t - q * nface;
  const int i = 1 + (int)(r % g.nx), k = 1 + (int)(r / g.nx);
  const double mx = ldg(a.met.idx_v, (idx_t)(i - 1 + G)), my = ldg(a.met.idy_v, (idx_t)(j - 1 + G)),
               mz = ldg(a.met.idz_v, (idx_t)(k - 1 + G));
  double c[5];
  onesided(q, c);
  const idx_t p = lin(g, i, j, k);
#pragma unroll
  for (int n = 0; n < 4; n++) {
    const double *fx = a.tau[krow(0, n)], *fy = a.tau[krow(1, n)], *fz = a.tau[krow(2, n)];
    double kv = (A5_1 * (ldg(fx, lin(g, i + 1, j, k)) - ldg(fx, lin(g, i - 1, j, k))) +
                 A5_2 * (ldg(fx, lin(g, i + 2, j, k)) - ldg(fx, lin(g, i - 2, j, k)))) *
                mx;
    double dy = 0.0;
#pragma unroll
    for (int m = 0; m < 5; m++) dy = dy + c[m] * ldg(fy, lin(g, i, r0 + m, k));
    kv = dy * my + kv;
    kv = (A5_1 * (ldg(fz, lin(g, i, j, k + 1)) - ldg(fz, lin(g, i, j, k - 1))) +
          A5_2 * (ldg(fz, lin(g, i, j, k + 2)) - ldg(fz, lin(g, i, j, k - 2)))) *
             mz +
         kv;
    stg(a.kv[n], p, kv);
  }
}

inline dim3 grid_for(long n) { return dim3((unsigned)((n + kWallThreads - 1) / kWallThreads)); }

}  // namespace

hipError_t launch_wall(const WallArgs &a, hipStream_t s) {
  const long n = 2L * a.g.nx * a.g.nz;
  hipLaunchKernelGGL(k_wall, grid_for(n), dim3(kWallThreads), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_visc_wall(const WallViscArgs &a, hipStream_t s) {
  const long n = 4L * a.g.nx * a.g.nz;
  hipLaunchKernelGGL(k_visc_wall, grid_for(n), dim3(kWallThreads), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_tau_wallrow(const WallViscArgs &a, hipStream_t s) {
  const long n = 2L * (a.g.nx + 4) * (a.g.nz + 4);
  hipLaunchKernelGGL(k_tau_wallrow, grid_for(n), dim3(kWallThreads), 0, s, a);
  return hipGetLastError();
}

hipError_t launch_vdiv_wall(const WallViscArgs &a, hipStream_t s) {
  const long n = 4L * a.g.nx * a.g.nz;
  hipLaunchKernelGGL(k_vdiv_wall, grid_for(n), dim3(kWallThreads), 0, s, a);
  return hipGetLastError();
}

}  // namespace mgpu
