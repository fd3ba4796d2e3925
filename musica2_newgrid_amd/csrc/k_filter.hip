// Numerical dissipation applied once per step after the last RK stage,
// fused with update_varn (src/Time/update.f90:538-565):
//   selective filter   filtering_11pts  src/Num_dissipation/mod_filtering.f90:992-1034,
//                      1192-1233, 1382-1423 (coefficients :63-68, scaled :189-203)
//   shock filter       filter_o10_shock src/Num_dissipation/mod_filtering_shock.f90:
//                      pressure sensor :243-260, face dissipation :419-440,
//                      increment :690-700 (and the j, k repeats)
// Interior rows only (periodic faces and 1-to-1 interfaces).
#include "mgpu_internal.h"

namespace mgpu {

namespace {

constexpr int FX = 64, FY = 4;

__global__ __launch_bounds__(FX *FY) void k_filter_sf(FilterArgs a) {
  const Geom &g = a.g;
  const int i = blockIdx.x * FX + threadIdx.x + 1;
  const int j = blockIdx.y * FY + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  if (i > g.nx || j > g.ny) return;
  const long p = lin(g, i, j, k);
  const long sy = g.sx, sz = g.sy;
  const double *dd = a.dd;
#pragma unroll
  for (int v = 0; v < 5; v++) {
    const double *u = a.uin.v[v];
    double K = dd[0] * u[p] + dd[1] * (u[p + 1] + u[p - 1]) + dd[2] * (u[p + 2] + u[p - 2]) +
               dd[3] * (u[p + 3] + u[p - 3]) + dd[4] * (u[p + 4] + u[p - 4]) + dd[5] * (u[p + 5] + u[p - 5]);
    K = K + dd[0] * u[p] + dd[1] * (u[p + sy] + u[p - sy]) + dd[2] * (u[p + 2 * sy] + u[p - 2 * sy]) +
        dd[3] * (u[p + 3 * sy] + u[p - 3 * sy]) + dd[4] * (u[p + 4 * sy] + u[p - 4 * sy]) +
        dd[5] * (u[p + 5 * sy] + u[p - 5 * sy]);
    K = K + dd[0] * u[p] + dd[1] * (u[p + sz] + u[p - sz]) + dd[2] * (u[p + 2 * sz] + u[p - 2 * sz]) +
        dd[3] * (u[p + 3 * sz] + u[p - 3 * sz]) + dd[4] * (u[p + 4 * sz] + u[p - 4 * sz]) +
        dd[5] * (u[p + 5 * sz] + u[p - 5 * sz]);
    a.uout.v[v][p] = u[p] + K;
  }
}

// pressure sensor psens at padded offset m along stride s
__device__ __forceinline__ double psens(const FilterArgs &a, long m, long s) {
  const double *pr = a.prs;
  const double ptrad = pr[m - s] + 2.0 * pr[m] + pr[m + s];
  const double ptvd = fabs(pr[m + s] - pr[m]) + fabs(pr[m] - pr[m - s]);
  const double sn = a.sens ? a.sens[m] : 1.0;
  return sn * fabs(pr[m - s] - 2.0 * pr[m] + pr[m + s]) / (a.csens * ptrad + (1.0 - a.csens) * ptvd);
}

// face dissipation D_{m+1/2} of variable u along stride s
__device__ __forceinline__ double dface(const double *u, long m, long s, double xk0, double xk10) {
  return (u[m + s] - u[m]) * xk0 +
         (D10_1 * (u[m + s] - u[m]) + D10_2 * (u[m + 2 * s] - u[m - s]) + D10_3 * (u[m + 3 * s] - u[m - 2 * s]) +
          D10_4 * (u[m + 4 * s] - u[m - 3 * s]) + D10_5 * (u[m + 5 * s] - u[m - 4 * s])) *
             xk10;
}

__global__ __launch_bounds__(FX *FY) void k_filter_shock(FilterArgs a) {
  const Geom &g = a.g;
  const int i = blockIdx.x * FX + threadIdx.x + 1;
  const int j = blockIdx.y * FY + threadIdx.y + 1;
  const int k = blockIdx.z + 1;
  if (i > g.nx || j > g.ny) return;
  const long p = lin(g, i, j, k);
  const long st[3] = {1, (long)g.sx, g.sy};
  double K[5];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const long s = st[d];
    const double psm = psens(a, p - s, s), ps0 = psens(a, p, s), psp = psens(a, p + s, s);
    const double xk0m = a.xnu0 * fmax(ps0, psm);
    const double xk10m = fmax(0.0, a.xnu10 + xk0m / 105.0);
    const double xk0p = a.xnu0 * fmax(psp, ps0);
    const double xk10p = fmax(0.0, a.xnu10 + xk0p / 105.0);
#pragma unroll
    for (int v = 0; v < 5; v++) {
      const double *u = a.uin.v[v];
      const double dlo = dface(u, p - s, s, xk0m, xk10m);
      const double dhi = dface(u, p, s, xk0p, xk10p);
      if (d == 0) K[v] = dlo - dhi;
      else K[v] = K[v] + dlo - dhi;
    }
  }
#pragma unroll
  for (int v = 0; v < 5; v++) a.uout.v[v][p] = a.uin.v[v][p] + K[v];
}

}  // namespace

hipError_t launch_filter_sf(const FilterArgs &a, hipStream_t s) {
  dim3 blk(FX, FY, 1);
  dim3 grd((a.g.nx + FX - 1) / FX, (a.g.ny + FY - 1) / FY, a.g.nz);
  hipLaunchKernelGGL(k_filter_sf, grd, blk, 0, s, a);
  return hipGetLastError();
}

hipError_t launch_filter_shock(const FilterArgs &a, hipStream_t s) {
  dim3 blk(FX, FY, 1);
  dim3 grd((a.g.nx + FX - 1) / FX, (a.g.ny + FY - 1) / FY, a.g.nz);
  hipLaunchKernelGGL(k_filter_shock, grd, blk, 0, s, a);
  return hipGetLastError();
}

}  // namespace mgpu
