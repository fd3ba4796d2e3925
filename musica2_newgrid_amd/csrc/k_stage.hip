// Fused RK stage: primitives -> Euler fluxes -> 11-pt divergence in i, j, k
// (+ 5-pt divergence of the precomputed viscous fluxes in the last stage)
// -> U_out = U^n - alpha*K.
//
// Follows, in one pass per stage, what the reference does in six array
// sweeps plus the update:
//   primitives_pfg          src/Fluid/smod_primitive_pfg.f90:12-59
//   flux_euler_11pts        src/Compute_flow/smod_flux_euler_11pts.f90:25-172,267-,601-
//   flux_euler_11pts_c3     src/Compute_flow/smod_flux_euler_11pts_c3.f90:25-38,139-171,763-773
//   flux_visc_5pts(_c3) divergence part  smod_flux_visc_5pts.f90:119-130, _c3.f90:262-
//   update_var              src/Time/update.f90:32-44
//
// Work decomposition (gfx950): a 64x4 workgroup owns a 64(i) x 4(j) column
// tile and marches a chunk of k planes.  For each plane the i- and j-flux
// pencils (tile + 5-cell halos) are computed into LDS; the k-fluxes of the
// thread's own column live in an 11-deep register queue, so every state
// value is read from HBM once per plane and the k-derivative needs no
// second pass.  i is the fastest index of every array: a wavefront reads
// 64 consecutive doubles (512 B) per variable.
#include "mgpu_internal.h"

namespace mgpu {

namespace {

constexpr int TX = 64;
constexpr int TY = 4;
constexpr int NT = TX * TY;

// Flux along direction DIR at padded offset p of state U.
template <bool CURV, int DIR>
__device__ __forceinline__ void flux(const CState &U, const Metrics &M, long p, double gam1,
                                     double F[5]) {
  const double r = U.v[0][p], ru = U.v[1][p], rv = U.v[2][p], rw = U.v[3][p], re = U.v[4][p];
  const double vi = 1.0 / r;
  const double u = ru * vi, v = rv * vi, w = rw * vi;
  const double pr = gam1 * (re - 0.5 * r * (u * u + v * v + w * w));
  if constexpr (CURV) {
    const double mx = M.m[3 * DIR][p], my = M.m[3 * DIR + 1][p], mz = M.m[3 * DIR + 2][p];
    const double vc = u * mx + v * my + w * mz;
    const double rovc = r * vc;
    F[0] = rovc;
    F[1] = rovc * u + pr * mx;
    F[2] = rovc * v + pr * my;
    F[3] = rovc * w + pr * mz;
    F[4] = (re + pr) * vc;
  } else {
    const double m = DIR == 0 ? ru : (DIR == 1 ? rv : rw);
    const double vel = DIR == 0 ? u : (DIR == 1 ? v : w);
    F[0] = m;
    F[1] = DIR == 0 ? m * u + pr : m * u;
    F[2] = DIR == 1 ? m * v + pr : m * v;
    F[3] = DIR == 2 ? m * w + pr : m * w;
    F[4] = (re + pr) * vel;
  }
}

__device__ __forceinline__ double d11c(const double *f, int s) {
  return A11_1 * (f[s] - f[-s]) + A11_2 * (f[2 * s] - f[-2 * s]) + A11_3 * (f[3 * s] - f[-3 * s]) +
         A11_4 * (f[4 * s] - f[-4 * s]) + A11_5 * (f[5 * s] - f[-5 * s]);
}

__device__ __forceinline__ double d5g(const double *f, long p, long s) {
  return A5_1 * (f[p + s] - f[p - s]) + A5_2 * (f[p + 2 * s] - f[p - 2 * s]);
}

template <bool CURV, bool VISC>
__global__ __launch_bounds__(NT) void k_stage(StageArgs a) {
  __shared__ double sfx[5][TY][TX + 10];
  __shared__ double sfy[5][TY + 10][TX];

  const Geom &g = a.g;
  const int tx = threadIdx.x, ty = threadIdx.y, tid = tx + TX * ty;
  const int i0 = blockIdx.x * TX + 1, j0 = blockIdx.y * TY + 1;
  const int i = i0 + tx, j = j0 + ty;
  const int kbeg = blockIdx.z * a.kchunk + 1;
  const int kend = min(kbeg + a.kchunk - 1, g.nz);
  const bool own = (i <= g.nx) && (j <= g.ny);
  const double gam1 = a.gas.gam1;

  double q[11][5];
  if (own) {
#pragma unroll
    for (int l = 0; l < 10; l++) flux<CURV, 2>(a.uin, a.met, lin(g, i, j, kbeg - 5 + l), gam1, q[l]);
  }

  for (int k = kbeg; k <= kend; k++) {
    if (own) flux<CURV, 2>(a.uin, a.met, lin(g, i, j, k + 5), gam1, q[10]);

    // i-fluxes on the tile + 5-cell halos (interior j only)
    for (int t = tid; t < TY * (TX + 10); t += NT) {
      const int li = t % (TX + 10), lj = t / (TX + 10);
      const int gi = i0 - 5 + li, gj = j0 + lj;
      if (gj <= g.ny && gi <= g.nx + 5) {
        double F[5];
        flux<CURV, 0>(a.uin, a.met, lin(g, gi, gj, k), gam1, F);
#pragma unroll
        for (int v = 0; v < 5; v++) sfx[v][lj][li] = F[v];
      }
    }
    // j-fluxes on the tile + 5-cell halos (interior i only)
    for (int t = tid; t < (TY + 10) * TX; t += NT) {
      const int li = t % TX, lj = t / TX;
      const int gi = i0 + li, gj = j0 - 5 + lj;
      if (gi <= g.nx && gj <= g.ny + 5) {
        double F[5];
        flux<CURV, 1>(a.uin, a.met, lin(g, gi, gj, k), gam1, F);
#pragma unroll
        for (int v = 0; v < 5; v++) sfy[v][lj][li] = F[v];
      }
    }
    __syncthreads();

    if (own) {
      const long p = lin(g, i, j, k);
      double K[5];
#pragma unroll
      for (int v = 0; v < 5; v++) {
        double dx = d11c(&sfx[v][ty][tx + 5], 1);
        double dy = d11c(&sfy[v][ty + 5][tx], TX);
        const double dz = A11_1 * (q[6][v] - q[4][v]) + A11_2 * (q[7][v] - q[3][v]) +
                          A11_3 * (q[8][v] - q[2][v]) + A11_4 * (q[9][v] - q[1][v]) +
                          A11_5 * (q[10][v] - q[0][v]);
        if constexpr (CURV) {
          K[v] = (dz + (dy + dx)) * a.met.m[9][p];
        } else {
          double kx = dx * a.met.idx[i - 1 + G];
          double ky = dy * a.met.idy[j - 1 + G] + kx;
          K[v] = dz * a.met.idz[k - 1 + G] + ky;
        }
      }
      if constexpr (VISC) {
        const long sy = g.sx, sz = g.sy;
        double mx, my, mz;
        if constexpr (CURV) {
          mx = my = mz = a.met.mv[9][p];
        } else {
          mx = a.met.idx_v[i - 1 + G];
          my = a.met.idy_v[j - 1 + G];
          mz = a.met.idz_v[k - 1 + G];
        }
#pragma unroll
        for (int c = 0; c < 4; c++) K[1 + c] = d5g(a.fv[c], p, 1) * mx + K[1 + c];
#pragma unroll
        for (int c = 0; c < 4; c++) K[1 + c] = d5g(a.fv[4 + c], p, sy) * my + K[1 + c];
#pragma unroll
        for (int c = 0; c < 4; c++) K[1 + c] = d5g(a.fv[8 + c], p, sz) * mz + K[1 + c];
      }
#pragma unroll
      for (int v = 0; v < 5; v++) a.uout.v[v][p] = a.u0.v[v][p] - a.alpha * K[v];
    }
    __syncthreads();

#pragma unroll
    for (int l = 0; l < 10; l++)
#pragma unroll
      for (int v = 0; v < 5; v++) q[l][v] = q[l + 1][v];
  }
}

}  // namespace

hipError_t launch_stage(const StageArgs &a, bool curv, bool visc, hipStream_t s) {
  dim3 blk(TX, TY, 1);
  const int nchunk = (a.g.nz + a.kchunk - 1) / a.kchunk;
  dim3 grd((a.g.nx + TX - 1) / TX, (a.g.ny + TY - 1) / TY, nchunk);
  if (curv) {
    if (visc) hipLaunchKernelGGL((k_stage<true, true>), grd, blk, 0, s, a);
    else hipLaunchKernelGGL((k_stage<true, false>), grd, blk, 0, s, a);
  } else {
    if (visc) hipLaunchKernelGGL((k_stage<false, true>), grd, blk, 0, s, a);
    else hipLaunchKernelGGL((k_stage<false, false>), grd, blk, 0, s, a);
  }
  return hipGetLastError();
}

}  // namespace mgpu
