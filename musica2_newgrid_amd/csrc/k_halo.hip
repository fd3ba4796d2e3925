// Device-side halo fill for periodic faces: the on-GPU form of the
// reference's MPI face exchange with a self neighbour
// (communication3d/commun3d, src/Parallel/mod_comm.f90:249-290; face types
// src/Parallel/mod_mpi_types_two_sided.f90:47-197).
// Three passes: i-faces over interior (j,k); j-faces over all i; k-faces over
// all (i,j), so edge and corner ghosts also hold the periodic image.
#include "mgpu_internal.h"

namespace mgpu {

namespace {

constexpr int HB = 256;

// dir 0: points (l, j, k), l in 1..G, j in 1..ny, k in 1..nz
// dir 1: points (i, l, k), i in 1-G..nx+G, l in 1..G, k in 1..nz
// dir 2: points (i, j, l), i in 1-G..nx+G, j in 1-G..ny+G, l in 1..G
__global__ __launch_bounds__(HB) void k_halo_periodic(HaloArgs a, int dir, int nvar) {
  const Geom &g = a.g;
  long t = (long)blockIdx.x * HB + threadIdx.x;
  int i, j, k, n;
  idx_t s;
  if (dir == 0) {
    const long cnt = (long)G * g.ny * g.nz;
    if (t >= cnt) return;
    const int l = (int)(t % G) + 1;
    t /= G;
    j = (int)(t % g.ny) + 1;
    k = (int)(t / g.ny) + 1;
    i = l; n = g.nx; s = 1;
  } else if (dir == 1) {
    const int wx = g.nx + 2 * G;
    const long cnt = (long)wx * G * g.nz;
    if (t >= cnt) return;
    i = (int)(t % wx) + 1 - G;
    t /= wx;
    const int l = (int)(t % G) + 1;
    k = (int)(t / G) + 1;
    j = l; n = g.ny; s = g.sx;
  } else {
    const int wx = g.nx + 2 * G, wy = g.ny + 2 * G;
    const long cnt = (long)wx * wy * G;
    if (t >= cnt) return;
    i = (int)(t % wx) + 1 - G;
    t /= wx;
    j = (int)(t % wy) + 1 - G;
    const int l = (int)(t / wy) + 1;
    k = l; n = g.nz; s = g.sy;
  }
  // (i,j,k) is interior layer l of the low face along dir.  Ghost n+l takes
  // interior l; ghost 1-l takes interior n+1-l.
  const int l = dir == 0 ? i : (dir == 1 ? j : k);
  const idx_t p_lo_src = lin(g, i, j, k);                      // interior index l
  const idx_t p_hi_dst = p_lo_src + (idx_t)n * s;              // ghost index n+l
  const idx_t p_src2 = p_lo_src + (idx_t)(n + 1 - 2 * l) * s;  // interior index n+1-l
  const idx_t p_dst2 = p_lo_src - (idx_t)(2 * l - 1) * s;      // ghost index 1-l
  for (int v = 0; v < nvar; v++) {
    double *u = a.u.v[v];
    stg(u, p_hi_dst, ldg(u, p_lo_src));
    stg(u, p_dst2, ldg(u, p_src2));
  }
}

}  // namespace

hipError_t launch_halo_periodic(const HaloArgs &a, int nvar, hipStream_t s) {
  const Geom &g = a.g;
  const long cnt[3] = {(long)G * g.ny * g.nz, (long)(g.nx + 2 * G) * G * g.nz,
                       (long)(g.nx + 2 * G) * (g.ny + 2 * G) * G};
  for (int d = 0; d < 3; d++) {
    const unsigned nb = (unsigned)((cnt[d] + HB - 1) / HB);
    hipLaunchKernelGGL(k_halo_periodic, dim3(nb), dim3(HB), 0, s, a, d, nvar);
  }
  return hipGetLastError();
}

}  // namespace mgpu
