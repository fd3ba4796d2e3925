// Viscous terms of U^n for the last RK stage, in two kernels.
//
// k_visc: the physical viscous flux tensor on the interior points.  One pass
// replaces, for the reference:
//   primitives_visc_pfg  src/Fluid/smod_primitive_pfg.f90:62-136
//   grad_vel_5pts / grad_T_5pts   smod_grad_vel.f90:12-, smod_grad_T.f90:12-
//   grad_*_5pts_c3 (+ mapping to x,y,z)  smod_grad_vel_c3.f90:11-,260-300;
//                                        smod_grad_T_c3.f90:200-213
//   flux_visc_5pts(_c3) stress and heat flux  smod_flux_visc_5pts.f90:27-67;
//                                             smod_flux_visc_5pts_c3.f90:27-75
// and, for the artificial viscosity and the shock filter's Edoh switch, the
// spectral radii |u_d| + c of U^n (c_ of primitives_visc,
// smod_primitive_pfg.f90:121; mod_artvisc.f90:278-283), contravariant on c3
// blocks (mod_filtering_shock.f90:627-630).
// It stores the tensor in the physical frame (tau, see mgpu_internal.h), so
// that the 2-layer interface exchange of tau that follows stands for the
// reference's exchange of the gradients (communication_3dv,
// src/Parallel/mod_comm.f90:376-402) and the flux evaluation at the ghosts.
// It also emits p(U^n) and the Ducros sensor that the shock filter reads
// (mod_filtering_shock.f90:223-239), exchanged with tau.
//
// k_vdiv: the 5-pt divergence (flux_visc_5pts(_c3) divergence part,
// smod_flux_visc_5pts.f90:119-, smod_flux_visc_5pts_c3.f90:221-), with the
// projection of tau on the viscous metrics of each stencil point.
//
// 2.5-D layout as in the stage kernel: per k plane the tile + 2-cell halos
// sit in LDS for the i/j derivatives; the thread's own column keeps a 5-deep
// register queue for the k derivative.  All global loads of a plane are
// issued in one batch.
#include "mgpu_internal.h"

namespace mgpu {

namespace {

struct P5 {
  double u, v, w, T, p;
};

__device__ __forceinline__ P5 prim5r(const double U[5], double gam1, double rg) {
  P5 q;
  const double r = U[0];
  const double vi = 1.0 / r;
  q.u = U[1] * vi;
  q.v = U[2] * vi;
  q.w = U[3] * vi;
  q.p = gam1 * (U[4] - 0.5 * r * (q.u * q.u + q.v * q.v + q.w * q.w));
  q.T = q.p * vi / rg;
  return q;
}

__device__ __forceinline__ void ld5(const CState &U, idx_t p, double o[5]) {
#pragma unroll
  for (int v = 0; v < 5; v++) o[v] = ldg(U.v[v], p);
}

template <int TX, int TY, int GEO>
__global__ __launch_bounds__(TX *TY, 2) void k_visc(const Batch<ViscArgs> *__restrict__ tab, int nb, int ntiles) {
  constexpr bool CURV = GEO == kGeoC3;  // 3-D metrics at every point
  constexpr bool C2 = GEO == kGeoC2;    // 2-D metrics of the column
  constexpr int NT = TX * TY;
  constexpr int NHX = 4 * TY;             // i-halo cells per plane
  constexpr int NH = NHX + 4 * TX;        // all halo cells per plane
  constexpr int NS = (NH + NT - 1) / NT;  // halo cells per thread
  __shared__ double spx[4][TY][TX + 4];
  __shared__ double spy[4][TY + 4][TX];

  const int w = xcd_remap(blockIdx.x, ntiles);
  const auto &e = batch_find(tab, nb, w);
  const auto &a = e.a;
  // the block's array pointers, read once from the table into registers
  const CState u0p = a.u0;
  const double *mvp[10];
#pragma unroll
  for (int c = 0; c < 10; c++) mvp[c] = a.met.mv[c];
  double *taup[NTAU];
#pragma unroll
  for (int c = 0; c < NTAU; c++) taup[c] = a.tau[c];
  double *const prsp = a.prs, *const sensp = a.sens;
  double *const sr0 = a.sr[0], *const sr1 = a.sr[1], *const sr2 = a.sr[2];
  const double *const idxv = a.met.idx_v, *const idyv = a.met.idy_v, *const idzv = a.met.idz_v;
  const Geom &g = a.g;
  int tile_x, tile_y, chunk;
  tile_coords(e.t, w, tile_x, tile_y, chunk);
  const int kchunk = e.t.kchunk;
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX, tid = threadIdx.x;
  const int i0 = tile_x * TX + 1, j0 = tile_y * TY + 1;
  const int i = i0 + tx, j = j0 + ty;
  const int kbeg = chunk * kchunk + 1;
  const int kend = min(kbeg + kchunk - 1, g.nz);
  const bool own = (i <= g.nx) && (j <= g.ny);
  const bool needx = (i <= g.nx + 2) && (j <= g.ny);  // point feeds the i-pencil
  const bool needy = (i <= g.nx) && (j <= g.ny + 2);  // point feeds the j-pencil
  const double gam1 = a.gas.gam1, rg = a.gas.rg;
  const idx_t sz = g.sy;

  // halo cells of this thread (LDS slot, offset in plane k=0)
  int hs_dir[NS], hs_r[NS], hs_c[NS];
  idx_t hs_off[NS];
  bool hs_on[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const int h = tid + s * NT;
    hs_on[s] = false;
    hs_dir[s] = hs_r[s] = hs_c[s] = 0;
    hs_off[s] = 0;
    if (h < NHX) {
      const int r = h / 4, c = h % 4;
      const int li = c < 2 ? c : TX + c;
      const int gi = i0 - 2 + li, gj = j0 + r;
      hs_on[s] = gj <= g.ny && gi <= g.nx + 2;
      hs_r[s] = r;
      hs_c[s] = li;
      hs_off[s] = lin(g, gi, gj, 0);
    } else if (h < NH) {
      const int hh = h - NHX;
      const int r = hh / TX, c = hh % TX;
      const int lj = r < 2 ? r : TY + r;
      const int gi = i0 + c, gj = j0 - 2 + lj;
      hs_on[s] = gi <= g.nx && gj <= g.ny + 2;
      hs_dir[s] = 1;
      hs_r[s] = lj;
      hs_c[s] = c;
      hs_off[s] = lin(g, gi, gj, 0);
    }
  }
  const idx_t own_off = lin(g, i, j, 0);
  // 2.5-D: the column's viscous metrics (y_eta, -x_eta, -y_ksi, x_ksi, ijacob),
  // the same in every plane
  double mc2[5] = {0, 0, 0, 0, 0};
  if constexpr (C2) {
    if (own) {
#pragma unroll
      for (int c = 0; c < 5; c++) mc2[c] = ldg(a.met.m2v[c], lin2(g, i, j));
    }
  }

  P5 q[5];  // own column, planes k-2..k+2
  if (own) {
#pragma unroll
    for (int l = 0; l < 4; l++) {
      double U[5];
      ld5(u0p, own_off + (idx_t)(kbeg - 2 + l) * sz, U);
      q[l] = prim5r(U, gam1, rg);
    }
  }

  for (int k = kbeg; k <= kend; k++) {
    const idx_t pk = own_off + (idx_t)k * sz;
    // ---- every global load of the plane ----
    double uq[5], uc[5], uh[NS][5], mvv[10];
    if (own) ld5(u0p, pk + 2 * sz, uq);
    if (!own && (needx || needy)) ld5(u0p, pk, uc);
#pragma unroll
    for (int s = 0; s < NS; s++)
      if (hs_on[s]) ld5(u0p, hs_off[s] + (idx_t)k * sz, uh[s]);
    if (CURV && own) {
#pragma unroll
      for (int c = 0; c < 10; c++) mvv[c] = ldg(mvp[c], pk);
    }

    // ---- primitives into the queue and the LDS pencils ----
    if (own) q[4] = prim5r(uq, gam1, rg);
    if (needx || needy) {
      // own column: from the queue; ghost cells just past a partial tile: loaded
      const P5 c = own ? q[2] : prim5r(uc, gam1, rg);
      if (needx) {
        spx[0][ty][tx + 2] = c.u;
        spx[1][ty][tx + 2] = c.v;
        spx[2][ty][tx + 2] = c.w;
        spx[3][ty][tx + 2] = c.T;
      }
      if (needy) {
        spy[0][ty + 2][tx] = c.u;
        spy[1][ty + 2][tx] = c.v;
        spy[2][ty + 2][tx] = c.w;
        spy[3][ty + 2][tx] = c.T;
      }
    }
#pragma unroll
    for (int s = 0; s < NS; s++) {
      if (hs_on[s]) {
        const P5 h = prim5r(uh[s], gam1, rg);
        if (hs_dir[s] == 0) {
          spx[0][hs_r[s]][hs_c[s]] = h.u;
          spx[1][hs_r[s]][hs_c[s]] = h.v;
          spx[2][hs_r[s]][hs_c[s]] = h.w;
          spx[3][hs_r[s]][hs_c[s]] = h.T;
        } else {
          spy[0][hs_r[s]][hs_c[s]] = h.u;
          spy[1][hs_r[s]][hs_c[s]] = h.v;
          spy[2][hs_r[s]][hs_c[s]] = h.w;
          spy[3][hs_r[s]][hs_c[s]] = h.T;
        }
      }
    }
    __syncthreads();

    if (own) {
      const idx_t p = pk;
      // gradients along the computational directions: gd[dir][u,v,w,T]
      double gd[3][4];
#pragma unroll
      for (int c = 0; c < 4; c++) {
        const double *fx = &spx[c][ty][tx + 2];
        gd[0][c] = A5_1 * (fx[1] - fx[-1]) + A5_2 * (fx[2] - fx[-2]);
        const double *fy = &spy[c][ty + 2][tx];
        gd[1][c] = A5_1 * (fy[TX] - fy[-TX]) + A5_2 * (fy[2 * TX] - fy[-2 * TX]);
      }
      gd[2][0] = A5_1 * (q[3].u - q[1].u) + A5_2 * (q[4].u - q[0].u);
      gd[2][1] = A5_1 * (q[3].v - q[1].v) + A5_2 * (q[4].v - q[0].v);
      gd[2][2] = A5_1 * (q[3].w - q[1].w) + A5_2 * (q[4].w - q[0].w);
      gd[2][3] = A5_1 * (q[3].T - q[1].T) + A5_2 * (q[4].T - q[0].T);

      // physical gradients gr[u,v,w,T][x,y,z]
      double gr[4][3];
      if constexpr (C2) {
        // (f_ksi y_eta - f_eta y_ksi) ijacob, (f_eta x_ksi - f_ksi x_eta) ijacob,
        // f_z idz (smod_grad_vel_c.f90:186-217, :230-300)
        const double iz = ldg(idzv, (idx_t)(k - 1 + G));
#pragma unroll
        for (int c = 0; c < 4; c++) {
          gr[c][0] = (gd[0][c] * mc2[0] + gd[1][c] * mc2[2]) * mc2[4];
          gr[c][1] = (gd[0][c] * mc2[1] + gd[1][c] * mc2[3]) * mc2[4];
          gr[c][2] = gd[2][c] * iz;
        }
      } else if constexpr (CURV) {
        const double ij = mvv[9];
#pragma unroll
        for (int c = 0; c < 4; c++) {
          const double dvx = gd[0][c] * mvv[0] + gd[1][c] * mvv[3] + gd[2][c] * mvv[6];
          const double dvy = gd[0][c] * mvv[1] + gd[1][c] * mvv[4] + gd[2][c] * mvv[7];
          const double dvz = gd[0][c] * mvv[2] + gd[1][c] * mvv[5] + gd[2][c] * mvv[8];
          gr[c][0] = dvx * ij;
          gr[c][1] = dvy * ij;
          gr[c][2] = dvz * ij;
        }
      } else {
        const double ix = ldg(idxv, (idx_t)(i - 1 + G)), iy = ldg(idyv, (idx_t)(j - 1 + G)), iz = ldg(idzv, (idx_t)(k - 1 + G));
#pragma unroll
        for (int c = 0; c < 4; c++) {
          gr[c][0] = gd[0][c] * ix;
          gr[c][1] = gd[1][c] * iy;
          gr[c][2] = gd[2][c] * iz;
        }
      }

      const double u = q[2].u, v = q[2].v, w = q[2].w, T = q[2].T;
      const double mu_l = a.gas.c0 * sqrt(T * T * T) / (T + a.gas.S0) * a.gas.diffscale;
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

      stg(taup[0], p, t11);
      stg(taup[1], p, t12);
      stg(taup[2], p, t13);
      stg(taup[3], p, t22);
      stg(taup[4], p, t23);
      stg(taup[5], p, t33);
      stg(taup[6], p, (u * t11 + v * t12 + w * t13) - ck * gr[3][0]);
      stg(taup[7], p, (u * t12 + v * t22 + w * t23) - ck * gr[3][1]);
      stg(taup[8], p, (u * t13 + v * t23 + w * t33) - ck * gr[3][2]);

      if (a.want_sensor) {
        stg(prsp, p, q[2].p);
        double s = 1.0;
        if (a.ducros) {
          const double dv = gr[0][0] + gr[1][1] + gr[2][2];
          const double div2 = dv * dv;
          const double o1 = gr[2][1] - gr[1][2];
          const double o2 = gr[0][2] - gr[2][0];
          const double o3 = gr[1][0] - gr[0][1];
          s = div2 / (div2 + o1 * o1 + o2 * o2 + o3 * o3 + 1.0e-16);
        }
        stg(sensp, p, s);
      }
      if (a.want_sr) {
        // spectral radii along the three directions, c = sqrt(gam R T)
        // (primitives_visc); on c3 blocks with the contravariant velocity
        // (u xi_x + v xi_y + w xi_z) ijacob3 and c |grad xi| |ijacob3| of the
        // Euler metrics (the Edoh switch, mod_filtering_shock.f90:627-630;
        // g3_ksi, smod_grid_metrics3_c3.f90:161)
        const double c = sqrt(a.gas.gam * rg * T);
        if constexpr (C2) {
          // ksi and eta from the 2-D Euler metrics (mod_filtering_shock.f90:650-652, :1089-1091;
          // g_ksi, g_eta of grid_metrics_curv.f90:709-724), z Cartesian
          const idx_t p2 = lin2(g, i, j);
          const double ij = ldg(a.met.m2[4], p2);
          double *const srp[2] = {sr0, sr1};
#pragma unroll
          for (int d = 0; d < 2; d++) {
            const double mx = ldg(a.met.m2[2 * d], p2), my = ldg(a.met.m2[2 * d + 1], p2);
            const double vc = (u * mx + v * my + w * 0.0) * ij;
            const double g3 = sqrt(mx * mx + my * my + 0.0 * 0.0) * fabs(ij);
            stg(srp[d], p, fabs(vc) + c * g3);
          }
          stg(sr2, p, fabs(w) + c);
        } else if constexpr (CURV) {
          const double ij = ldg(a.met.m[9], p);
          double *const srp[3] = {sr0, sr1, sr2};
#pragma unroll
          for (int d = 0; d < 3; d++) {
            const double mx = ldg(a.met.m[3 * d], p), my = ldg(a.met.m[3 * d + 1], p), mz = ldg(a.met.m[3 * d + 2], p);
            const double vc = (u * mx + v * my + w * mz) * ij;
            const double g3 = sqrt(mx * mx + my * my + mz * mz) * fabs(ij);
            stg(srp[d], p, fabs(vc) + c * g3);
          }
        } else {
          stg(sr0, p, fabs(u) + c);
          stg(sr1, p, fabs(v) + c);
          stg(sr2, p, fabs(w) + c);
        }
      }
    }
    __syncthreads();
#pragma unroll
    for (int l = 0; l < 4; l++) q[l] = q[l + 1];
  }
}

// Tiles of the 32x16 viscous kernels: chunks of >= 16 planes, about 2048
// workgroups per block (4 per CU slot of 256 CUs x 2).
constexpr int VX = 32, VY = 16;
TileRange vtiles(const Geom &g, int kc_req) {
  TileRange t{};
  t.tiles_x = (g.nx + VX - 1) / VX;
  t.tiles_y = (g.ny + VY - 1) / VY;
  const long nxy = (long)t.tiles_x * t.tiles_y;
  const long want = (2048 + nxy - 1) / nxy;
  int kc = kc_req > 0 ? kc_req : (int)((g.nz + want - 1) / want);
  if (kc_req <= 0 && kc < 16) kc = nxy * ((g.nz + 15) / 16) >= 512 ? 16 : (kc < 4 ? 4 : kc);
  if (kc > g.nz) kc = g.nz;
  t.kchunk = kc;
  return t;
}

// Flux along computational direction d at one point: Cartesian, row d of
// tau (t holds those four values); c3, tau projected on the viscous metric
// row (mx,my,mz) of d (t holds all nine values).
template <bool CURV>
__device__ __forceinline__ void vflux(const double *t, double mx, double my, double mz, double f[4]) {
  if constexpr (CURV) {
    f[0] = t[0] * mx + t[1] * my + t[2] * mz;
    f[1] = t[1] * mx + t[3] * my + t[4] * mz;
    f[2] = t[2] * mx + t[4] * my + t[5] * mz;
    f[3] = t[6] * mx + t[7] * my + t[8] * mz;
  } else {
    f[0] = t[0];
    f[1] = t[1];
    f[2] = t[2];
    f[3] = t[3];
  }
}

// 5-pt differences of F^v along i and G^v along j from LDS pencils (tile +
// 2-cell halos), of H^v along k from a 5-deep register queue; times
// idx_v,idy_v,idz_v (Cartesian), ijacob3_v (c3), or ijacob_v along i, j and
// idz_v along k (2.5-D, smod_flux_visc_5pts_c.f90:52-72, :150-).  Interior
// points only.
template <int TX, int TY, int GEO>
__global__ __launch_bounds__(TX *TY, 2) void k_vdiv(const Batch<VdivArgs> *__restrict__ tab, int nb, int ntiles) {
  constexpr bool CURV = GEO == kGeoC3;
  constexpr bool C2 = GEO == kGeoC2;
  constexpr int NT = TX * TY;
  constexpr int NHX = 4 * TY;
  constexpr int NH = NHX + 4 * TX;
  constexpr int NS = (NH + NT - 1) / NT;
  __shared__ double sfx[4][TY][TX + 4];
  __shared__ double sfy[4][TY + 4][TX];

  const int w = xcd_remap(blockIdx.x, ntiles);
  const auto &e = batch_find(tab, nb, w);
  const auto &a = e.a;
  // the block's array pointers, read once from the table into registers
  const double *mvp[10];
#pragma unroll
  for (int c = 0; c < 10; c++) mvp[c] = a.met.mv[c];
  const double *taup[NTAU];
#pragma unroll
  for (int c = 0; c < NTAU; c++) taup[c] = a.tau[c];
  double *kvp[4];
#pragma unroll
  for (int c = 0; c < 4; c++) kvp[c] = a.kv[c];
  const double *const idxv = a.met.idx_v, *const idyv = a.met.idy_v, *const idzv = a.met.idz_v;
  const Geom &g = a.g;
  int tile_x, tile_y, chunk;
  tile_coords(e.t, w, tile_x, tile_y, chunk);
  const int kchunk = e.t.kchunk;
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX, tid = threadIdx.x;
  const int i0 = tile_x * TX + 1, j0 = tile_y * TY + 1;
  const int i = i0 + tx, j = j0 + ty;
  const int kbeg = chunk * kchunk + 1;
  const int kend = min(kbeg + kchunk - 1, g.nz);
  const bool own = i <= g.nx && j <= g.ny;
  const bool needx = i <= g.nx + 2 && j <= g.ny, needy = i <= g.nx && j <= g.ny + 2;
  const idx_t sz = g.sy, own_off = lin(g, i, j, 0);
  const double *const *mv = mvp;

  int hs_r[NS], hs_c[NS];
  idx_t hs_off[NS], hs_off2[NS];
  bool hs_on[NS], hs_x[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const int h = tid + s * NT;
    hs_on[s] = false;
    hs_x[s] = true;
    hs_r[s] = hs_c[s] = 0;
    hs_off[s] = hs_off2[s] = 0;
    if (h < NHX) {
      const int r = h / 4, c = h % 4;
      const int li = c < 2 ? c : TX + c;
      const int gi = i0 - 2 + li, gj = j0 + r;
      hs_on[s] = gj <= g.ny && gi <= g.nx + 2;
      hs_r[s] = r;
      hs_c[s] = li;
      hs_off[s] = lin(g, gi, gj, 0);
      hs_off2[s] = lin2(g, gi, gj);
    } else if (h < NH) {
      const int hh = h - NHX;
      const int r = hh / TX, c = hh % TX;
      const int lj = r < 2 ? r : TY + r;
      const int gi = i0 + c, gj = j0 - 2 + lj;
      hs_on[s] = gi <= g.nx && gj <= g.ny + 2;
      hs_x[s] = false;
      hs_r[s] = lj;
      hs_c[s] = c;
      hs_off[s] = lin(g, gi, gj, 0);
      hs_off2[s] = lin2(g, gi, gj);
    }
  }
  const idx_t own_off2 = lin2(g, i, j);

  // flux of direction d at element p (column p2 of the 2-D arrays) from its
  // nine tau values t9: the projection on the metric row of d (c3; 2.5-D
  // along i, j) or row d of tau (Cartesian; 2.5-D along k)
  auto proj = [&](int d, const double *t9, idx_t p, idx_t p2, double f[4]) {
    if (CURV || (C2 && d < 2)) {
      double mx, my, mz;
      if constexpr (C2) {
        mx = ldg(a.met.m2v[2 * d], p2);
        my = ldg(a.met.m2v[2 * d + 1], p2);
        mz = 0.0;
      } else {
        mx = ldg(mv[3 * d], p);
        my = ldg(mv[3 * d + 1], p);
        mz = ldg(mv[3 * d + 2], p);
      }
      vflux<true>(t9, mx, my, mz, f);
    } else {
      double t[4];
#pragma unroll
      for (int c = 0; c < 4; c++) t[c] = t9[krow(d, c)];
      vflux<false>(t, 0.0, 0.0, 0.0, f);
    }
  };
  // tau values of direction d at element p, projected into f[4]
  auto flux_at = [&](int d, idx_t p, idx_t p2, double f[4]) {
    double t9[NTAU];
    if (CURV || (C2 && d < 2)) {
#pragma unroll
      for (int c = 0; c < NTAU; c++) t9[c] = ldg(taup[c], p);
    } else {
#pragma unroll
      for (int c = 0; c < 4; c++) t9[krow(d, c)] = ldg(taup[krow(d, c)], p);
    }
    proj(d, t9, p, p2, f);
  };
  // the same from tau values already in registers (all nine components)
  auto flux_of = [&](int d, const double *t9, idx_t p, idx_t p2, double f[4]) { proj(d, t9, p, p2, f); };

  // own column: H^v at k-2..k+2, and tau of planes k, k+1, k+2 (each plane's
  // tau is loaded once, at k+2, and reused for the i/j pencils two planes on)
  double q[5][4];
  double tw[3][NTAU];
  if (own) {
#pragma unroll
    for (int l = 0; l < 2; l++) flux_at(2, own_off + (idx_t)(kbeg - 2 + l) * sz, own_off2, q[l]);
#pragma unroll
    for (int l = 0; l < 2; l++) {
      const idx_t p = own_off + (idx_t)(kbeg + l) * sz;
#pragma unroll
      for (int c = 0; c < NTAU; c++) tw[l][c] = ldg(taup[c], p);
      flux_of(2, tw[l], p, own_off2, q[2 + l]);
    }
  }
  for (int k = kbeg; k <= kend; k++) {
    const idx_t pk = own_off + (idx_t)k * sz;
    double fx[4], fy[4], fh[NS][4];
    double ijv = 0.0;
    if (own) {
      const idx_t p2 = pk + 2 * sz;
#pragma unroll
      for (int c = 0; c < NTAU; c++) tw[2][c] = ldg(taup[c], p2);
      flux_of(2, tw[2], p2, own_off2, q[4]);
      if constexpr (CURV) ijv = ldg(mv[9], pk);
      if constexpr (C2) ijv = ldg(a.met.m2v[4], own_off2);
      flux_of(0, tw[0], pk, own_off2, fx);
      flux_of(1, tw[0], pk, own_off2, fy);
    } else {
      if (needx) flux_at(0, pk, own_off2, fx);
      if (needy) flux_at(1, pk, own_off2, fy);
    }
#pragma unroll
    for (int s = 0; s < NS; s++)
      if (hs_on[s]) {
        if (hs_x[s]) flux_at(0, hs_off[s] + (idx_t)k * sz, hs_off2[s], fh[s]);
        else flux_at(1, hs_off[s] + (idx_t)k * sz, hs_off2[s], fh[s]);
      }
    if (needx) {
#pragma unroll
      for (int c = 0; c < 4; c++) sfx[c][ty][tx + 2] = fx[c];
    }
    if (needy) {
#pragma unroll
      for (int c = 0; c < 4; c++) sfy[c][ty + 2][tx] = fy[c];
    }
#pragma unroll
    for (int s = 0; s < NS; s++) {
      if (hs_on[s]) {
#pragma unroll
        for (int c = 0; c < 4; c++) {
          if (hs_x[s]) sfx[c][hs_r[s]][hs_c[s]] = fh[s][c];
          else sfy[c][hs_r[s]][hs_c[s]] = fh[s][c];
        }
      }
    }
    __syncthreads();
    if (own) {
      double mx, my, mz;
      if constexpr (CURV) {
        mx = my = mz = ijv;
      } else if constexpr (C2) {
        mx = my = ijv;
        mz = ldg(idzv, (idx_t)(k - 1 + G));
      } else {
        mx = ldg(idxv, (idx_t)(i - 1 + G));
        my = ldg(idyv, (idx_t)(j - 1 + G));
        mz = ldg(idzv, (idx_t)(k - 1 + G));
      }
#pragma unroll
      for (int c = 0; c < 4; c++) {
        const double *x = &sfx[c][ty][tx + 2];
        const double *y = &sfy[c][ty + 2][tx];
        double kv = (A5_1 * (x[1] - x[-1]) + A5_2 * (x[2] - x[-2])) * mx;
        kv = (A5_1 * (y[TX] - y[-TX]) + A5_2 * (y[2 * TX] - y[-2 * TX])) * my + kv;
        kv = (A5_1 * (q[3][c] - q[1][c]) + A5_2 * (q[4][c] - q[0][c])) * mz + kv;
        stg(kvp[c], pk, kv);
      }
    }
    __syncthreads();
#pragma unroll
    for (int l = 0; l < 4; l++)
#pragma unroll
      for (int c = 0; c < 4; c++) q[l][c] = q[l + 1][c];
#pragma unroll
    for (int l = 0; l < 2; l++)
#pragma unroll
      for (int c = 0; c < NTAU; c++) tw[l][c] = tw[l + 1][c];
  }
}

}  // namespace

TileRange visc_tiles(const Geom &g, int kchunk) { return vtiles(g, kchunk); }
TileRange vdiv_tiles(const Geom &g, int kchunk) { return vtiles(g, kchunk); }

hipError_t launch_visc(const Batch<ViscArgs> *tab, int nb, int nt, int geo, hipStream_t s) {
  if (nt <= 0) return hipSuccess;
  if (geo == kGeoC2) hipLaunchKernelGGL((k_visc<VX, VY, kGeoC2>), dim3(nt), dim3(VX * VY), 0, s, tab, nb, nt);
  else if (geo == kGeoC3) hipLaunchKernelGGL((k_visc<VX, VY, kGeoC3>), dim3(nt), dim3(VX * VY), 0, s, tab, nb, nt);
  else hipLaunchKernelGGL((k_visc<VX, VY, kGeoCart>), dim3(nt), dim3(VX * VY), 0, s, tab, nb, nt);
  return hipGetLastError();
}

hipError_t launch_vdiv(const Batch<VdivArgs> *tab, int nb, int nt, int geo, hipStream_t s) {
  if (nt <= 0) return hipSuccess;
  if (geo == kGeoC2) hipLaunchKernelGGL((k_vdiv<VX, VY, kGeoC2>), dim3(nt), dim3(VX * VY), 0, s, tab, nb, nt);
  else if (geo == kGeoC3) hipLaunchKernelGGL((k_vdiv<VX, VY, kGeoC3>), dim3(nt), dim3(VX * VY), 0, s, tab, nb, nt);
  else hipLaunchKernelGGL((k_vdiv<VX, VY, kGeoCart>), dim3(nt), dim3(VX * VY), 0, s, tab, nb, nt);
  return hipGetLastError();
}

}  // namespace mgpu
