// Maximum CFL number per direction of the current state: check_max_cfl
// (src/Time/timestep_new.f90:660-842).  Not on the time-step path: one
// thread per interior point, each workgroup writes the maximum of its points
// for i, j and k; the host takes the maximum over the workgroups (max is
// order-independent, so the result is exact).
//   Cartesian  (|u_d| + c) deltat idx_d                          (:713-719)
//   c3         (|(u xi_x + v xi_y + w xi_z) ijacob3| + c g3_xi) deltat, with
//              g3_xi = |grad xi| |ijacob3|                        (:679-697)
//   2.5-D      the same with the 2-D rows (y_eta, -x_eta), (-y_ksi, x_ksi) and
//              ijacob in i and j (:700-711), Cartesian in k (:820-826)
// c is the sound speed of primitives_visc, sqrt(gam rg T)
// (src/Fluid/smod_primitive_pfg.f90:121).
#include "mgpu_internal.h"

namespace mgpu {

namespace {

template <int GEO>
__global__ __launch_bounds__(kDiagThreads) void k_cfl(CflArgs a) {
  __shared__ double red[3][kDiagThreads];
  const Geom &g = a.g;
  const long t = (long)blockIdx.x * kDiagThreads + threadIdx.x;
  const long cnt = (long)g.nx * g.ny * g.nz;
  double cm[3] = {0.0, 0.0, 0.0};
  if (t < cnt) {
    const int i = (int)(t % g.nx) + 1;
    const long r = t / g.nx;
    const int j = (int)(r % g.ny) + 1;
    const int k = (int)(r / g.ny) + 1;
    const idx_t p = lin(g, i, j, k);
    const double rho = ldg(a.u.v[0], p), vi = 1.0 / rho;
    const double u[3] = {ldg(a.u.v[1], p) * vi, ldg(a.u.v[2], p) * vi, ldg(a.u.v[3], p) * vi};
    const double pr = a.gam1 * (ldg(a.u.v[4], p) - 0.5 * rho * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
    const double T = pr * vi / a.rg;
    const double cs = sqrt(a.gam * a.rg * T);
    const int rr[3] = {i, j, k};
    const double *cart[3] = {a.met.idx, a.met.idy, a.met.idz};
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const bool curv = GEO == kGeoC3 || (GEO == kGeoC2 && d < 2);
      if (curv) {
        double mx, my, mz, ij;
        if constexpr (GEO == kGeoC3) {
          mx = ldg(a.met.m[3 * d], p);
          my = ldg(a.met.m[3 * d + 1], p);
          mz = ldg(a.met.m[3 * d + 2], p);
          ij = ldg(a.met.m[9], p);
        } else {
          const idx_t p2 = lin2(g, i, j);
          mx = ldg(a.met.m2[2 * d], p2);
          my = ldg(a.met.m2[2 * d + 1], p2);
          mz = 0.0;
          ij = ldg(a.met.m2[4], p2);
        }
        const double val = fabs((u[0] * mx + u[1] * my + u[2] * mz) * ij) + cs * (sqrt(mx * mx + my * my + mz * mz) * fabs(ij));
        cm[d] = val * a.dt;
      } else {
        cm[d] = (fabs(u[d]) + cs) * a.dt * ldg(cart[d], (idx_t)(rr[d] - 1 + G));
      }
    }
  }
#pragma unroll
  for (int d = 0; d < 3; d++) red[d][threadIdx.x] = cm[d];
  __syncthreads();
  for (int s = kDiagThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
#pragma unroll
      for (int d = 0; d < 3; d++) red[d][threadIdx.x] = fmax(red[d][threadIdx.x], red[d][threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 3) a.partial[3 * (long)blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

}  // namespace

hipError_t launch_cfl(const CflArgs &a, int geo, hipStream_t s) {
  const long nb = diag_groups(a.g);
  if (geo == kGeoC2) hipLaunchKernelGGL(k_cfl<kGeoC2>, dim3((unsigned)nb), dim3(kDiagThreads), 0, s, a);
  else if (geo == kGeoC3) hipLaunchKernelGGL(k_cfl<kGeoC3>, dim3((unsigned)nb), dim3(kDiagThreads), 0, s, a);
  else hipLaunchKernelGGL(k_cfl<kGeoCart>, dim3((unsigned)nb), dim3(kDiagThreads), 0, s, a);
  return hipGetLastError();
}

}  // namespace mgpu
