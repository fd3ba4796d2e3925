// Diagnostics of the state: sums over the interior of rho, 1/2 rho |u|^2 and
// 1/2 rho |curl u|^2, the quantities of stats_tgv
// (src/File_IO/stats.f90:37-112), with the velocity gradients by the 5-pt
// scheme of grad_vel (smod_grad_vel.f90, c3 mapping smod_grad_vel_c3.f90:
// 260-300, 2.5-D mapping smod_grad_vel_c.f90:186-217); and the maximum of |u| + c, the velocity scale of timestep_new
// (src/Time/timestep_new.f90:270-288).  Not on the time-step path: one thread per point, neighbours
// read straight from global memory; each workgroup writes its three partial
// sums, which the host adds in a fixed order (deterministic results).
#include <mutex>

#include "mgpu_internal.h"

namespace mgpu {

namespace {

__device__ __forceinline__ void vel(const CState &U, idx_t p, double o[3]) {
  const double r = ldg(U.v[0], p);
  const double vi = 1.0 / r;
  o[0] = ldg(U.v[1], p) * vi;
  o[1] = ldg(U.v[2], p) * vi;
  o[2] = ldg(U.v[3], p) * vi;
}

template <int GEO>
__global__ __launch_bounds__(kDiagThreads) void k_diag(DiagArgs a) {
  __shared__ double red[4][kDiagThreads];
  const Geom &g = a.g;
  const long t = (long)blockIdx.x * kDiagThreads + threadIdx.x;
  const long cnt = (long)g.nx * g.ny * g.nz;
  double mg = 0.0, ek = 0.0, wr = 0.0, um = 0.0;
  if (t < cnt) {
    const int i = (int)(t % g.nx) + 1;
    const long r = t / g.nx;
    const int j = (int)(r % g.ny) + 1;
    const int k = (int)(r / g.ny) + 1;
    const idx_t p = lin(g, i, j, k);
    const idx_t st[3] = {1, (idx_t)g.sx, g.sy};
    double gd[3][3];  // gd[dir][u,v,w]
#pragma unroll
    for (int d = 0; d < 3; d++) {
      double a1[3], b1[3], a2[3], b2[3];
      vel(a.u, p + st[d], a1);
      vel(a.u, p - st[d], b1);
      vel(a.u, p + 2 * st[d], a2);
      vel(a.u, p - 2 * st[d], b2);
#pragma unroll
      for (int c = 0; c < 3; c++) gd[d][c] = A5_1 * (a1[c] - b1[c]) + A5_2 * (a2[c] - b2[c]);
    }
    // rows 1, 2 (n-1, n) next to a bounded i- or j-face (wall, characteristic
    // outflow, supersonic inlet): the one-sided a04 / a13 (a31 / a40) stencils
    // of grad_vel (smod_grad_vel.f90, is_boundary), which read nothing beyond
    // the face
#pragma unroll
    for (int d = 0; d < 2; d++) {
      const int bnd = d ? g.jbnd : g.ibnd, n = d ? g.ny : g.nx, r = d ? j : i;
      const bool wlo = (bnd & 1) && r <= 2, whi = (bnd & 2) && r >= n - 1;
      if (!(wlo || whi)) continue;
      double f[5][3];  // rows r0..r0+4 of the stencil
      const int r0 = wlo ? 1 : n - 4;
#pragma unroll
      for (int l = 0; l < 5; l++) vel(a.u, d ? lin(g, i, r0 + l, k) : lin(g, r0 + l, j, k), f[l]);
#pragma unroll
      for (int c = 0; c < 3; c++) {
        if (wlo && r == 1)
          gd[d][c] = A04_1 * f[0][c] + A04_2 * f[1][c] + A04_3 * f[2][c] + A04_4 * f[3][c] + A04_5 * f[4][c];
        else if (wlo)
          gd[d][c] = A13_1 * f[0][c] + A13_2 * f[1][c] + A13_3 * f[2][c] + A13_4 * f[3][c] + A13_5 * f[4][c];
        else if (r == n - 1)
          gd[d][c] = (-A13_5) * f[0][c] + (-A13_4) * f[1][c] + (-A13_3) * f[2][c] + (-A13_2) * f[3][c] +
                     (-A13_1) * f[4][c];
        else
          gd[d][c] = (-A04_5) * f[0][c] + (-A04_4) * f[1][c] + (-A04_3) * f[2][c] + (-A04_2) * f[3][c] +
                     (-A04_1) * f[4][c];
      }
    }
    double gr[3][3];  // gr[u,v,w][x,y,z]
    if constexpr (GEO == kGeoC2) {
      // (f_ksi y_eta - f_eta y_ksi) ijacob, (f_eta x_ksi - f_ksi x_eta) ijacob, f_z idz
      const idx_t p2 = lin2(g, i, j);
      double m[5];
#pragma unroll
      for (int q = 0; q < 5; q++) m[q] = ldg(a.met.m2v[q], p2);
      const double iz = ldg(a.met.idz_v, (idx_t)(k - 1 + G));
#pragma unroll
      for (int c = 0; c < 3; c++) {
        gr[c][0] = (gd[0][c] * m[0] + gd[1][c] * m[2]) * m[4];
        gr[c][1] = (gd[0][c] * m[1] + gd[1][c] * m[3]) * m[4];
        gr[c][2] = gd[2][c] * iz;
      }
    } else if constexpr (GEO == kGeoC3) {
      double m[10];
#pragma unroll
      for (int q = 0; q < 10; q++) m[q] = ldg(a.met.mv[q], p);
#pragma unroll
      for (int c = 0; c < 3; c++) {
        const double dvx = gd[0][c] * m[0] + gd[1][c] * m[3] + gd[2][c] * m[6];
        const double dvy = gd[0][c] * m[1] + gd[1][c] * m[4] + gd[2][c] * m[7];
        const double dvz = gd[0][c] * m[2] + gd[1][c] * m[5] + gd[2][c] * m[8];
        gr[c][0] = dvx * m[9];
        gr[c][1] = dvy * m[9];
        gr[c][2] = dvz * m[9];
      }
    } else {
      const double ix = ldg(a.met.idx_v, (idx_t)(i - 1 + G)), iy = ldg(a.met.idy_v, (idx_t)(j - 1 + G)), iz = ldg(a.met.idz_v, (idx_t)(k - 1 + G));
#pragma unroll
      for (int c = 0; c < 3; c++) {
        gr[c][0] = gd[0][c] * ix;
        gr[c][1] = gd[1][c] * iy;
        gr[c][2] = gd[2][c] * iz;
      }
    }
    double u[3];
    vel(a.u, p, u);
    const double rho = ldg(a.u.v[0], p);
    mg = rho;
    ek = 0.5 * rho * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    const double o1 = gr[2][1] - gr[1][2], o2 = gr[0][2] - gr[2][0], o3 = gr[1][0] - gr[0][1];
    wr = 0.5 * rho * (o1 * o1 + o2 * o2 + o3 * o3);
    const double pr = a.gam1 * (ldg(a.u.v[4], p) - 0.5 * rho * (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]));
    um = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) + sqrt(a.gam * pr / rho);
  }
  red[0][threadIdx.x] = mg;
  red[1][threadIdx.x] = ek;
  red[2][threadIdx.x] = wr;
  red[3][threadIdx.x] = um;
  __syncthreads();
  for (int s = kDiagThreads / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
#pragma unroll
      for (int q = 0; q < 3; q++) red[q][threadIdx.x] += red[q][threadIdx.x + s];
      red[3][threadIdx.x] = fmax(red[3][threadIdx.x], red[3][threadIdx.x + s]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 4) a.partial[4 * (long)blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// Totals of the per-workgroup results: sums of the three sums, maximum of the
// maxima.  One workgroup; each thread adds a fixed strided subset, then a
// fixed tree (deterministic).
__global__ __launch_bounds__(kDiagThreads) void k_diag_total(const double *__restrict__ partial, long ng,
                                                             double *__restrict__ total) {
  __shared__ double red[4][kDiagThreads];
  double s[3] = {0.0, 0.0, 0.0}, m = 0.0;
  for (long w = threadIdx.x; w < ng; w += kDiagThreads) {
#pragma unroll
    for (int q = 0; q < 3; q++) s[q] += partial[4 * w + q];
    m = fmax(m, partial[4 * w + 3]);
  }
#pragma unroll
  for (int q = 0; q < 3; q++) red[q][threadIdx.x] = s[q];
  red[3][threadIdx.x] = m;
  __syncthreads();
  for (int h = kDiagThreads / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
#pragma unroll
      for (int q = 0; q < 3; q++) red[q][threadIdx.x] += red[q][threadIdx.x + h];
      red[3][threadIdx.x] = fmax(red[3][threadIdx.x], red[3][threadIdx.x + h]);
    }
    __syncthreads();
  }
  if (threadIdx.x < 4) total[threadIdx.x] = red[threadIdx.x][0];
}

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch[kMaxDevices] = {};  // device address of `scratch`, per device

}  // namespace

// Device memory of the allocation-free diagnostics: the per-workgroup
// results of one block (at most kDiagScratchGroups workgroups) followed by
// its 4 totals (4 MiB).  Each GPU holds its own copy for the life of the
// process.  Host code never passes this name to a kernel or a copy: it uses
// the address diag_scratch_init looked up for the device.
static __device__ __align__(256) double scratch[4 * kDiagScratchGroups + 4];

long diag_groups(const Geom &g) { return ((long)g.nx * g.ny * g.nz + kDiagThreads - 1) / kDiagThreads; }

hipError_t diag_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch[device]) return hipSuccess;
  double *p = nullptr;
  const hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch[device] = p;
  return hipSuccess;
}

hipError_t diag_block_noalloc(DiagArgs a, int geo, int device, hipStream_t s, double tot[4]) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  double *p;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    p = g_scratch[device];
  }
  if (!p) return hipErrorNotInitialized;
  const long ng = diag_groups(a.g);
  if (ng > kDiagScratchGroups) return hipErrorInvalidValue;
  a.partial = p;  // every workgroup writes its 4 results: nothing to clear
  hipError_t e = launch_diag(a, geo, s);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_diag_total, dim3(1), dim3(kDiagThreads), 0, s, p, ng, p + 4 * kDiagScratchGroups);
  if ((e = hipGetLastError()) != hipSuccess) return e;
  if ((e = hipMemcpyAsync(tot, p + 4 * kDiagScratchGroups, 4 * sizeof(double), hipMemcpyDeviceToHost, s)) !=
      hipSuccess)
    return e;
  return hipStreamSynchronize(s);
}

hipError_t launch_diag(const DiagArgs &a, int geo, hipStream_t s) {
  const long nb = diag_groups(a.g);
  if (geo == kGeoC2) hipLaunchKernelGGL(k_diag<kGeoC2>, dim3((unsigned)nb), dim3(kDiagThreads), 0, s, a);
  else if (geo == kGeoC3) hipLaunchKernelGGL(k_diag<kGeoC3>, dim3((unsigned)nb), dim3(kDiagThreads), 0, s, a);
  else hipLaunchKernelGGL(k_diag<kGeoCart>, dim3((unsigned)nb), dim3(kDiagThreads), 0, s, a);
  return hipGetLastError();
}

}  // namespace mgpu
