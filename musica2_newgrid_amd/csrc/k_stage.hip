// Fused RK stage: primitives -> Euler fluxes -> 11-pt divergence in i, j, k
// (+ the viscous increment of U^n, precomputed by k_vdiv, in the last stage)
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
// Work decomposition (gfx950): a TX x TY workgroup owns a TX(i) x TY(j)
// column tile and marches a chunk of k planes.  For each plane the i- and
// j-flux pencils (tile + 5-cell halos) are computed into LDS -- each thread
// loads its own point once and produces both its i- and j-flux, the halo
// cells are loaded by the remaining lanes; the k-fluxes of the thread's own
// column live in an 11-deep register queue, so each plane of the state is
// streamed from HBM once and the k-derivative needs no second pass.  i is the
// fastest index of every array: a wavefront reads whole rows.  Tiles are
// numbered XCD-aware so that tiles sharing halos share an L2.
#include "mgpu_internal.h"

namespace mgpu {

namespace {

struct Prim {
  double r, ru, rv, rw, re, u, v, w, pr;
};

__device__ __forceinline__ double d11c(const double *f, int s) {
  return A11_1 * (f[s] - f[-s]) + A11_2 * (f[2 * s] - f[-2 * s]) + A11_3 * (f[3 * s] - f[-3 * s]) +
         A11_4 * (f[4 * s] - f[-4 * s]) + A11_5 * (f[5 * s] - f[-5 * s]);
}

// Raw inputs of one flux evaluation: the 5 conserved variables and the 3
// metric terms of the flux direction (c3).
struct Raw {
  double u[5];
  double m[3];
};

template <bool CURV>
__device__ __forceinline__ void load_raw(const CState &U, const Metrics &M, idx_t p, int dir, Raw &r) {
#pragma unroll
  for (int v = 0; v < 5; v++) r.u[v] = ldg(U.v[v], p);
  if constexpr (CURV) {
#pragma unroll
    for (int c = 0; c < 3; c++) r.m[c] = ldg(M.m[3 * dir + c], p);
  }
}

__device__ __forceinline__ Prim prim_raw(const double u[5], double gam1) {
  Prim q;
  q.r = u[0];
  q.ru = u[1];
  q.rv = u[2];
  q.rw = u[3];
  q.re = u[4];
  const double vi = 1.0 / q.r;
  q.u = q.ru * vi;
  q.v = q.rv * vi;
  q.w = q.rw * vi;
  q.pr = gam1 * (q.re - 0.5 * q.r * (q.u * q.u + q.v * q.v + q.w * q.w));
  return q;
}

// Flux along dir from primitives and the metric terms of dir (c3).
template <bool CURV>
__device__ __forceinline__ void flux_m(const Prim &q, int dir, const double m[3], double F[5]) {
  if constexpr (CURV) {
    const double vc = q.u * m[0] + q.v * m[1] + q.w * m[2];
    const double rovc = q.r * vc;
    F[0] = rovc;
    F[1] = rovc * q.u + q.pr * m[0];
    F[2] = rovc * q.v + q.pr * m[1];
    F[3] = rovc * q.w + q.pr * m[2];
    F[4] = (q.re + q.pr) * vc;
  } else {
    const double vel = dir == 0 ? q.u : (dir == 1 ? q.v : q.w);
    const double mm = dir == 0 ? q.ru : (dir == 1 ? q.rv : q.rw);
    F[0] = mm;
    F[1] = dir == 0 ? mm * q.u + q.pr : mm * q.u;
    F[2] = dir == 1 ? mm * q.v + q.pr : mm * q.v;
    F[3] = dir == 2 ? mm * q.w + q.pr : mm * q.w;
    F[4] = (q.re + q.pr) * vel;
  }
}

template <int TX, int TY, bool CURV, bool VISC>
__global__ __launch_bounds__(TX *TY, 2) void k_stage(const Batch<StageArgs> *__restrict__ tab, int nb, int ntiles) {
  constexpr int NT = TX * TY;
  constexpr int NHX = 10 * TY;            // i-halo cells per plane
  constexpr int NH = NHX + 10 * TX;       // all halo cells per plane
  constexpr int NS = (NH + NT - 1) / NT;  // halo cells per thread
  // flux pencils, double-buffered over k planes so that one barrier per
  // plane suffices (plane k+1 is written while plane k may still be read)
  __shared__ double sfx2[2][5][TY][TX + 10];
  __shared__ double sfy2[2][5][TY + 10][TX];

  const int w = xcd_remap(blockIdx.x, ntiles);
  const Batch<StageArgs> &e = batch_find(tab, nb, w);
  const StageArgs &a = e.a;
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
  const bool needx = (i <= g.nx + 5) && (j <= g.ny);  // own point feeds the i-pencil
  const bool needy = (i <= g.nx) && (j <= g.ny + 5);  // own point feeds the j-pencil
  const double gam1 = a.gas.gam1;

  // halo cells handled by this thread: LDS slot, padded offset in the plane, direction
  int hs_dir[NS], hs_r[NS], hs_c[NS];
  idx_t hs_off[NS];
  bool hs_on[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const int h = tid + s * NT;
    hs_on[s] = false;
    hs_dir[s] = 0;
    hs_r[s] = hs_c[s] = 0;
    hs_off[s] = 0;
    if (h < NHX) {
      const int r = h / 10, c = h % 10;
      const int li = c < 5 ? c : TX + c;
      const int gi = i0 - 5 + li, gj = j0 + r;
      hs_on[s] = gj <= g.ny && gi <= g.nx + 5;
      hs_dir[s] = 0;
      hs_r[s] = r;
      hs_c[s] = li;
      hs_off[s] = lin(g, gi, gj, 0);
    } else if (h < NH) {
      const int hh = h - NHX;
      const int r = hh / TX, c = hh % TX;
      const int lj = r < 5 ? r : TY + r;
      const int gi = i0 + c, gj = j0 - 5 + lj;
      hs_on[s] = gi <= g.nx && gj <= g.ny + 5;
      hs_dir[s] = 1;
      hs_r[s] = lj;
      hs_c[s] = c;
      hs_off[s] = lin(g, gi, gj, 0);
    }
  }
  const idx_t own_off = lin(g, i, j, 0);
  const idx_t sz = g.sy;

  double q[11][5];
  if (own) {
#pragma unroll
    for (int l = 0; l < 10; l++) {
      Raw r;
      load_raw<CURV>(a.uin, a.met, own_off + (idx_t)(kbeg - 5 + l) * sz, 2, r);
      flux_m<CURV>(prim_raw(r.u, gam1), 2, r.m, q[l]);
    }
  }

  for (int k = kbeg; k <= kend; k++) {
    const idx_t pk = own_off + (idx_t)k * sz;
    double (*sfx)[TY][TX + 10] = sfx2[(k - kbeg) & 1];
    double (*sfy)[TY + 10][TX] = sfy2[(k - kbeg) & 1];
    // ---- batch 1: queue point k+5, own point k, U^n and 1/J at k, halo cell 0 ----
    Raw rq, rh0;
    double rxm[3] = {0, 0, 0}, rym[3] = {0, 0, 0}, ru[5] = {0, 0, 0, 0, 0}, u0[5] = {0, 0, 0, 0, 0};
    double ijac = 0.0;
    if (own) load_raw<CURV>(a.uin, a.met, pk + 5 * sz, 2, rq);
    if (needx || needy) {
#pragma unroll
      for (int v = 0; v < 5; v++) ru[v] = ldg(a.uin.v[v], pk);
      if constexpr (CURV) {
#pragma unroll
        for (int c = 0; c < 3; c++) {
          rxm[c] = ldg(a.met.m[c], pk);
          rym[c] = ldg(a.met.m[3 + c], pk);
        }
      }
    }
    if (hs_on[0]) load_raw<CURV>(a.uin, a.met, hs_off[0] + (idx_t)k * sz, hs_dir[0], rh0);
    double kvis[4] = {0, 0, 0, 0};
    if (own) {
#pragma unroll
      for (int v = 0; v < 5; v++) u0[v] = ldg(a.u0.v[v], pk);
      if constexpr (CURV) ijac = ldg(a.met.m[9], pk);
      if constexpr (VISC) {
#pragma unroll
        for (int c = 0; c < 4; c++) kvis[c] = ldg(a.kv[c], pk);
      }
    }
    if (own) flux_m<CURV>(prim_raw(rq.u, gam1), 2, rq.m, q[10]);
    if (needx || needy) {
      const Prim pq = prim_raw(ru, gam1);
      double F[5];
      if (needx) {
        flux_m<CURV>(pq, 0, rxm, F);
#pragma unroll
        for (int v = 0; v < 5; v++) sfx[v][ty][tx + 5] = F[v];
      }
      if (needy) {
        flux_m<CURV>(pq, 1, rym, F);
#pragma unroll
        for (int v = 0; v < 5; v++) sfy[v][ty + 5][tx] = F[v];
      }
    }
    if (hs_on[0]) {
      double F[5];
      flux_m<CURV>(prim_raw(rh0.u, gam1), hs_dir[0], rh0.m, F);
      if (hs_dir[0] == 0) {
#pragma unroll
        for (int v = 0; v < 5; v++) sfx[v][hs_r[0]][hs_c[0]] = F[v];
      } else {
#pragma unroll
        for (int v = 0; v < 5; v++) sfy[v][hs_r[0]][hs_c[0]] = F[v];
      }
    }
    __builtin_amdgcn_sched_barrier(0);
    // ---- batch 2: the remaining halo cells ----
    if constexpr (NS > 1) {
      Raw rh[NS];
#pragma unroll
      for (int s = 1; s < NS; s++)
        if (hs_on[s]) load_raw<CURV>(a.uin, a.met, hs_off[s] + (idx_t)k * sz, hs_dir[s], rh[s]);
#pragma unroll
      for (int s = 1; s < NS; s++) {
        if (hs_on[s]) {
          double F[5];
          flux_m<CURV>(prim_raw(rh[s].u, gam1), hs_dir[s], rh[s].m, F);
          if (hs_dir[s] == 0) {
#pragma unroll
            for (int v = 0; v < 5; v++) sfx[v][hs_r[s]][hs_c[s]] = F[v];
          } else {
#pragma unroll
            for (int v = 0; v < 5; v++) sfy[v][hs_r[s]][hs_c[s]] = F[v];
          }
        }
      }
    }
    __syncthreads();

    if (own) {
      double K[5];
#pragma unroll
      for (int v = 0; v < 5; v++) {
        const double dx = d11c(&sfx[v][ty][tx + 5], 1);
        const double dy = d11c(&sfy[v][ty + 5][tx], TX);
        const double dz = A11_1 * (q[6][v] - q[4][v]) + A11_2 * (q[7][v] - q[3][v]) +
                          A11_3 * (q[8][v] - q[2][v]) + A11_4 * (q[9][v] - q[1][v]) +
                          A11_5 * (q[10][v] - q[0][v]);
        if constexpr (CURV) {
          K[v] = (dz + (dy + dx)) * ijac;
        } else {
          const double kx = dx * a.met.idx[i - 1 + G];
          const double ky = dy * a.met.idy[j - 1 + G] + kx;
          K[v] = dz * a.met.idz[k - 1 + G] + ky;
        }
      }
      if constexpr (VISC) {
#pragma unroll
        for (int c = 0; c < 4; c++) K[1 + c] = K[1 + c] + kvis[c];
      }
#pragma unroll
      for (int v = 0; v < 5; v++) stg_nt(a.uout.v[v], pk, u0[v] - a.alpha * K[v]);
    }

#pragma unroll
    for (int l = 0; l < 10; l++)
#pragma unroll
      for (int v = 0; v < 5; v++) q[l][v] = q[l + 1][v];
  }
}

template <int TX, int TY>
hipError_t launch_t(const Batch<StageArgs> *tab, int nb, int nt, bool curv, bool visc, hipStream_t s) {
  dim3 blk(TX * TY), grd(nt);
  if (curv) {
    if (visc) hipLaunchKernelGGL((k_stage<TX, TY, true, true>), grd, blk, 0, s, tab, nb, nt);
    else hipLaunchKernelGGL((k_stage<TX, TY, true, false>), grd, blk, 0, s, tab, nb, nt);
  } else {
    if (visc) hipLaunchKernelGGL((k_stage<TX, TY, false, true>), grd, blk, 0, s, tab, nb, nt);
    else hipLaunchKernelGGL((k_stage<TX, TY, false, false>), grd, blk, 0, s, tab, nb, nt);
  }
  return hipGetLastError();
}

}  // namespace

// Tile shapes the autotuner may choose from.
const int kStageShapes[][2] = {{64, 4}, {32, 8}, {64, 8}, {16, 16}, {32, 16}};
const int kNumStageShapes = 5;

TileRange stage_tiles(const Geom &g, int shape, int kchunk) {
  const int TX = kStageShapes[shape][0], TY = kStageShapes[shape][1];
  TileRange t{};
  t.tiles_x = (g.nx + TX - 1) / TX;
  t.tiles_y = (g.ny + TY - 1) / TY;
  t.kchunk = kchunk < 1 ? 1 : (kchunk > g.nz ? g.nz : kchunk);
  return t;
}

hipError_t launch_stage(const Batch<StageArgs> *tab, int nb, int nt, bool curv, bool visc, int shape,
                        hipStream_t s) {
  if (nt <= 0) return hipSuccess;
  switch (shape) {
    case 1: return launch_t<32, 8>(tab, nb, nt, curv, visc, s);
    case 2: return launch_t<64, 8>(tab, nb, nt, curv, visc, s);
    case 3: return launch_t<16, 16>(tab, nb, nt, curv, visc, s);
    case 4: return launch_t<32, 16>(tab, nb, nt, curv, visc, s);
    default: return launch_t<64, 4>(tab, nb, nt, curv, visc, s);
  }
}

}  // namespace mgpu
