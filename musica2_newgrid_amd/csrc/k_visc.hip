// Viscous fluxes of U^n for the last RK stage, on the box [-1, n+2]^3 that
// the 5-pt divergence reads.  One pass replaces, for the reference:
//   primitives_visc_pfg  src/Fluid/smod_primitive_pfg.f90:62-136
//   grad_vel_5pts / grad_T_5pts   smod_grad_vel.f90:12-, smod_grad_T.f90:12-
//   grad_*_5pts_c3 (+ mapping to x,y,z)  smod_grad_vel_c3.f90:11-,260-300;
//                                        smod_grad_T_c3.f90:200-213
//   communication_3dv (gradient halo)    src/Parallel/mod_comm.f90:376-402
//   flux_visc_5pts(_c3) flux part        smod_flux_visc_5pts.f90:27-67;
//                                        smod_flux_visc_5pts_c3.f90:27-75
// Gradients in the two ghost layers are evaluated from the state ghosts
// instead of being exchanged: for periodic faces and 1-to-1 interfaces this
// is the same arithmetic on the same values.
// It also emits p(U^n) and the Ducros sensor that the shock filter reads
// (mod_filtering_shock.f90:223-239).
#include "mgpu_internal.h"

namespace mgpu {

namespace {

constexpr int VX = 64, VY = 4;

__device__ __forceinline__ void prim4(const CState &U, long p, double gam1, double rg, double &u,
                                      double &v, double &w, double &T) {
  const double r = U.v[0][p];
  const double vi = 1.0 / r;
  u = U.v[1][p] * vi;
  v = U.v[2][p] * vi;
  w = U.v[3][p] * vi;
  const double pr = gam1 * (U.v[4][p] - 0.5 * r * (u * u + v * v + w * w));
  T = pr * vi / rg;
}

template <bool CURV>
__global__ __launch_bounds__(VX *VY) void k_visc(ViscArgs a) {
  const Geom &g = a.g;
  const int i = blockIdx.x * VX + threadIdx.x - 1;
  const int j = blockIdx.y * VY + threadIdx.y - 1;
  const int k = blockIdx.z - 1;
  if (i > g.nx + 2 || j > g.ny + 2) return;
  const double gam1 = a.gas.gam1, rg = a.gas.rg;
  const long p = lin(g, i, j, k);
  const long st[3] = {1, (long)g.sx, g.sy};

  // gradient of (u,v,w,T) along each computational direction
  double gd[3][4];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    double pm2[4], pm1[4], pp1[4], pp2[4];
    prim4(a.u0, p - 2 * st[d], gam1, rg, pm2[0], pm2[1], pm2[2], pm2[3]);
    prim4(a.u0, p - st[d], gam1, rg, pm1[0], pm1[1], pm1[2], pm1[3]);
    prim4(a.u0, p + st[d], gam1, rg, pp1[0], pp1[1], pp1[2], pp1[3]);
    prim4(a.u0, p + 2 * st[d], gam1, rg, pp2[0], pp2[1], pp2[2], pp2[3]);
#pragma unroll
    for (int q = 0; q < 4; q++) gd[d][q] = A5_1 * (pp1[q] - pm1[q]) + A5_2 * (pp2[q] - pm2[q]);
  }
  // physical gradients g[q][x/y/z]
  double gr[4][3];
  if constexpr (CURV) {
    const double *const *mv = a.met.mv;
    const double ij = mv[9][p];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const double dvx = gd[0][q] * mv[0][p] + gd[1][q] * mv[3][p] + gd[2][q] * mv[6][p];
      const double dvy = gd[0][q] * mv[1][p] + gd[1][q] * mv[4][p] + gd[2][q] * mv[7][p];
      const double dvz = gd[0][q] * mv[2][p] + gd[1][q] * mv[5][p] + gd[2][q] * mv[8][p];
      gr[q][0] = dvx * ij;
      gr[q][1] = dvy * ij;
      gr[q][2] = dvz * ij;
    }
  } else {
    const double ix = a.met.idx_v[i - 1 + G], iy = a.met.idy_v[j - 1 + G], iz = a.met.idz_v[k - 1 + G];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      gr[q][0] = gd[0][q] * ix;
      gr[q][1] = gd[1][q] * iy;
      gr[q][2] = gd[2][q] * iz;
    }
  }

  // primitives, viscosity and conductivity at the point (primitives_visc_pfg)
  const double r = a.u0.v[0][p];
  const double vi = 1.0 / r;
  const double u = a.u0.v[1][p] * vi, v = a.u0.v[2][p] * vi, w = a.u0.v[3][p] * vi;
  const double pr = gam1 * (a.u0.v[4][p] - 0.5 * r * (u * u + v * v + w * w));
  const double T = pr * vi / rg;
  const double mu_l = a.gas.c0 * sqrt(T * T * T) / (T + a.gas.S0) * a.gas.diffscale;
  const double ck = a.gas.cok_ratio * mu_l;

  // stress tensor (flux_visc_5pts: S_ij, then -tau_ij)
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

  if constexpr (CURV) {
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const double mx = a.met.mv[3 * d][p], my = a.met.mv[3 * d + 1][p], mz = a.met.mv[3 * d + 2][p];
      const double fu = t11 * mx + t12 * my + t13 * mz;
      const double fv = t12 * mx + t22 * my + t23 * mz;
      const double fw = t13 * mx + t23 * my + t33 * mz;
      a.fv[4 * d][p] = fu;
      a.fv[4 * d + 1][p] = fv;
      a.fv[4 * d + 2][p] = fw;
      a.fv[4 * d + 3][p] = (u * fu + v * fv + w * fw) - ck * (gr[3][0] * mx + gr[3][1] * my + gr[3][2] * mz);
    }
  } else {
    a.fv[0][p] = t11;
    a.fv[1][p] = t12;
    a.fv[2][p] = t13;
    a.fv[3][p] = (u * t11 + v * t12 + w * t13) - ck * gr[3][0];
    a.fv[4][p] = t12;
    a.fv[5][p] = t22;
    a.fv[6][p] = t23;
    a.fv[7][p] = (u * t12 + v * t22 + w * t23) - ck * gr[3][1];
    a.fv[8][p] = t13;
    a.fv[9][p] = t23;
    a.fv[10][p] = t33;
    a.fv[11][p] = (u * t13 + v * t23 + w * t33) - ck * gr[3][2];
  }

  if (a.want_sensor) {
    a.prs[p] = pr;
    if (i >= 0 && i <= g.nx + 1 && j >= 0 && j <= g.ny + 1 && k >= 0 && k <= g.nz + 1) {
      double s = 1.0;
      if (a.ducros) {
        const double dv = gr[0][0] + gr[1][1] + gr[2][2];
        const double div2 = dv * dv;
        const double o1 = gr[2][1] - gr[1][2];
        const double o2 = gr[0][2] - gr[2][0];
        const double o3 = gr[1][0] - gr[0][1];
        s = div2 / (div2 + o1 * o1 + o2 * o2 + o3 * o3 + 1.0e-16);
      }
      a.sens[p] = s;
    }
  }
}

}  // namespace

hipError_t launch_visc(const ViscArgs &a, bool curv, hipStream_t s) {
  dim3 blk(VX, VY, 1);
  dim3 grd((a.g.nx + 4 + VX - 1) / VX, (a.g.ny + 4 + VY - 1) / VY, a.g.nz + 4);
  if (curv) hipLaunchKernelGGL((k_visc<true>), grd, blk, 0, s, a);
  else hipLaunchKernelGGL((k_visc<false>), grd, blk, 0, s, a);
  return hipGetLastError();
}

}  // namespace mgpu
