// Numerical dissipation applied once per step after the last RK stage,
// fused with update_varn (src/Time/update.f90:538-565):
//   selective filter   filtering_11pts  src/Num_dissipation/mod_filtering.f90:992-1034,
//                      1192-1233, 1382-1423 (coefficients :63-68, scaled :189-203)
//   shock filter       filter_o10_shock src/Num_dissipation/mod_filtering_shock.f90:
//                      pressure sensor :243-260, face dissipation :419-440,
//                      increment :690-700 (and the j, k repeats)
//   artificial viscosity artvisc_o9 mod_artvisc.f90:82-874 (coefficients :16-79),
//                      with update_varn_artvisc (update.f90:568-608): the
//                      shock filter's face dissipation without the sensor,
//                      weighted by the spectral radius |u_d| + c of U^n
//                      averaged to the faces, times the metric, U += deltat*K
// Interior rows of periodic faces and 1-to-1 interfaces; next to boundary
// faces on i or j (walls, characteristic outflow, supersonic inlet) the
// selective filter leaves rows 1,2 (n-1,n) unfiltered along that axis and
// uses the 5/7/9-pt filters on rows 3,4,5 (mod_filtering.f90:900-988,
// 1116-1190); along k every row is filtered.
//
// 2.5-D layout: the state of the tile + 5-cell halos of the current k plane
// sits in LDS for the i/j stencils, the own column keeps an 11-deep register
// queue for k.  The shock filter also stages p(U^n) and the Ducros sensor
// and computes each pressure sensor once per plane in LDS.  Halo cells are
// statically assigned to lanes so that all global loads of a plane are
// issued in one batch.
#include <type_traits>

#include "mgpu_internal.h"

namespace mgpu {

namespace {

constexpr int FX = 32, FY = 16, FT = FX * FY;
constexpr int NHX = 10 * FY, NH = NHX + 10 * FX, NS = (NH + FT - 1) / FT;  // state halo cells
constexpr int NPX = FY * (FX + 4), NPY = (FY + 4) * FX;                  // p cells along i / j
constexpr int NSP = (NPX + NPY + FT - 1) / FT;

// Block and tile of this workgroup in a batched launch.
__device__ __forceinline__ const FilterArgs &tile_of(const Batch<FilterArgs> *__restrict__ tab, int nb, int ntiles,
                                                     int &i0, int &j0, int &kbeg, int &kend) {
  const int w = xcd_remap(blockIdx.x, ntiles);
  const Batch<FilterArgs> &e = batch_find(tab, nb, w);
  int tile_x, tile_y, chunk;
  tile_coords(e.t, w, tile_x, tile_y, chunk);
  i0 = tile_x * FX + 1;
  j0 = tile_y * FY + 1;
  kbeg = chunk * e.t.kchunk + 1;
  kend = min(kbeg + e.t.kchunk - 1, e.a.g.nz);
  return e.a;
}

// Static assignment of the 5-cell state halo of a plane to the lanes.
struct Halo {
  idx_t off[NS];
  int r[NS], c[NS];
  bool on[NS], xdir[NS];
};

__device__ __forceinline__ void halo_setup(const Geom &g, int i0, int j0, Halo &h) {
  const int tid = threadIdx.x;
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const int t = tid + s * FT;
    h.on[s] = false;
    h.xdir[s] = true;
    h.r[s] = h.c[s] = 0;
    h.off[s] = 0;
    if (t < NHX) {
      const int r = t / 10, c = t % 10;
      const int li = c < 5 ? c : FX + c;
      const int gi = i0 - 5 + li, gj = j0 + r;
      h.on[s] = gj <= g.ny && gi <= g.nx + 5;
      h.r[s] = r;
      h.c[s] = li;
      h.off[s] = lin(g, gi, gj, 0);
    } else if (t < NH) {
      const int tt = t - NHX;
      const int r = tt / FX, c = tt % FX;
      const int lj = r < 5 ? r : FY + r;
      const int gi = i0 + c, gj = j0 - 5 + lj;
      h.on[s] = gi <= g.nx && gj <= g.ny + 5;
      h.xdir[s] = false;
      h.r[s] = lj;
      h.c[s] = c;
      h.off[s] = lin(g, gi, gj, 0);
    }
  }
}

__device__ __forceinline__ void ld5(const CState &U, idx_t p, double o[5]) {
#pragma unroll
  for (int v = 0; v < 5; v++) o[v] = ldg(U.v[v], p);
}

// State at element p.  The filters run in FP64 in both precisions: they read
// and write FP64 states (the last stage's output and U^{n+1}), so FP32
// arithmetic saved no time (0.93 ms per launch either way on the 512x256x128
// channel), while rounding the state's difference with the reference state
// to FP32 before the stencil was the largest error of the mixed path
// (density 2.4e-6 after 1000 steps there, 4.7e-8 with FP64 filters).
__device__ __forceinline__ void ldr(const CState &U, idx_t p, double o[5]) {
#pragma unroll
  for (int v = 0; v < 5; v++) o[v] = ldg(U.v[v], p);
}

// Filter increment along j of rows next to a wall: nothing on the first two,
// the 5/7/9-pt filters (times -chi) on the next three; y points at the row
// with stride s.  Reads nothing beyond the wall.
template <class R>
__device__ __forceinline__ R fwall(const R *y, int s, int hw, double chi) {
  const R c = R(-chi);
  switch (hw) {
    case 2: return c * (R(D5_0) * y[0] + R(D5_1) * (y[s] + y[-s]) + R(D5_2) * (y[2 * s] + y[-2 * s]));
    case 3:
      return c * (R(D7_0) * y[0] + R(D7_1) * (y[s] + y[-s]) + R(D7_2) * (y[2 * s] + y[-2 * s]) +
                  R(D7_3) * (y[3 * s] + y[-3 * s]));
    case 4:
      return c * (R(D9_0) * y[0] + R(D9_1) * (y[s] + y[-s]) + R(D9_2) * (y[2 * s] + y[-2 * s]) +
                  R(D9_3) * (y[3 * s] + y[-3 * s]) + R(D9_4) * (y[4 * s] + y[-4 * s]));
    default: return R(0);
  }
}

template <bool WALL>
__global__ __launch_bounds__(FT, 2) void k_filter_sf(const Batch<FilterArgs> *__restrict__ tab, int nb, int ntiles) {

  using R = double;
  __shared__ R sux[5][FY][FX + 10];
  __shared__ R suy[5][FY + 10][FX];
  int i0, j0, kbeg, kend;
  const FilterArgs &a = tile_of(tab, nb, ntiles, i0, j0, kbeg, kend);
  // the block's array pointers, read once from the table into registers
  const CState uinp = a.uin;
  const State uoutp = a.uout;
  const Geom &g = a.g;
  const int tx = threadIdx.x % FX, ty = threadIdx.x / FX;
  const int i = i0 + tx, j = j0 + ty;
  const bool own = i <= g.nx && j <= g.ny;
  const bool needx = i <= g.nx + 5 && j <= g.ny, needy = i <= g.nx && j <= g.ny + 5;
  const idx_t sz = g.sy, own_off = lin(g, i, j, 0);
  Halo hl;
  halo_setup(g, i0, j0, hl);
  R dd[6];
#pragma unroll
  for (int l = 0; l < 6; l++) dd[l] = (R)a.dd[l];

  R q[11][5];
  if (own) {
#pragma unroll
    for (int l = 0; l < 10; l++) ldr(uinp, own_off + (idx_t)(kbeg - 5 + l) * sz, q[l]);
  }
  // the state loads of a plane (queue point k+5, pencil cell, halo cells)
  // are issued one plane ahead, during the previous plane's stencils
  R qn[5], uc[5], uh[NS][5];
  {
    const idx_t pk = own_off + (idx_t)kbeg * sz;
    if (own) ldr(uinp, pk + 5 * sz, qn);
    if (!own && (needx || needy)) ldr(uinp, pk, uc);
#pragma unroll
    for (int s = 0; s < NS; s++)
      if (hl.on[s]) ldr(uinp, hl.off[s] + (idx_t)kbeg * sz, uh[s]);
  }
  for (int k = kbeg; k <= kend; k++) {
    const idx_t pk = own_off + (idx_t)k * sz;
    if (own) {
#pragma unroll
      for (int v = 0; v < 5; v++) q[10][v] = qn[v];
    }
    if (needx || needy) {
#pragma unroll
      for (int v = 0; v < 5; v++) {
        const R x = own ? q[5][v] : uc[v];
        if (needx) sux[v][ty][tx + 5] = x;
        if (needy) suy[v][ty + 5][tx] = x;
      }
    }
#pragma unroll
    for (int s = 0; s < NS; s++) {
      if (hl.on[s]) {
#pragma unroll
        for (int v = 0; v < 5; v++) {
          if (hl.xdir[s]) sux[v][hl.r[s]][hl.c[s]] = uh[s][v];
          else suy[v][hl.r[s]][hl.c[s]] = uh[s][v];
        }
      }
    }
    __syncthreads();
    if (k < kend) {
      const idx_t pn = pk + sz;
      if (own) ldr(uinp, pn + 5 * sz, qn);
      if (!own && (needx || needy)) ldr(uinp, pn, uc);
#pragma unroll
      for (int s = 0; s < NS; s++)
        if (hl.on[s]) ldr(uinp, hl.off[s] + (idx_t)(k + 1) * sz, uh[s]);
    }
    if (own) {
      const int hw = WALL ? jhalf(g, j) : 5;
      const int hwi = WALL ? ihalf(g, i) : 5;
#pragma unroll
      for (int v = 0; v < 5; v++) {
        const R *x = &sux[v][ty][tx + 5];
        const R *y = &suy[v][ty + 5][tx];
        R K;
        if (!WALL || hwi == 5) {
          K = dd[0] * x[0] + dd[1] * (x[1] + x[-1]) + dd[2] * (x[2] + x[-2]) + dd[3] * (x[3] + x[-3]) +
              dd[4] * (x[4] + x[-4]) + dd[5] * (x[5] + x[-5]);
        } else {
          K = R(0);
          if constexpr (WALL) K = fwall<R>(x, 1, hwi, a.chi);
        }
        if (!WALL || hw == 5) {
          K = K + dd[0] * y[0] + dd[1] * (y[FX] + y[-FX]) + dd[2] * (y[2 * FX] + y[-2 * FX]) +
              dd[3] * (y[3 * FX] + y[-3 * FX]) + dd[4] * (y[4 * FX] + y[-4 * FX]) +
              dd[5] * (y[5 * FX] + y[-5 * FX]);
        } else {
          if constexpr (WALL) K = K + fwall<R>(y, FX, hw, a.chi);
        }
        K = K + dd[0] * q[5][v] + dd[1] * (q[6][v] + q[4][v]) + dd[2] * (q[7][v] + q[3][v]) +
            dd[3] * (q[8][v] + q[2][v]) + dd[4] * (q[9][v] + q[1][v]) + dd[5] * (q[10][v] + q[0][v]);
        stg_nt(uoutp.v[v], pk, q[5][v] + K);
      }
    }
    __syncthreads();
#pragma unroll
    for (int l = 0; l < 10; l++)
#pragma unroll
      for (int v = 0; v < 5; v++) q[l][v] = q[l + 1][v];
  }
}

// pressure sensor (mod_filtering_shock.f90:252-257) from p at m-1, m, m+1
__device__ __forceinline__ double psens3(double pm, double p0, double pp, double sn, double cs) {
  const double ptrad = pm + 2.0 * p0 + pp;
  const double ptvd = fabs(pp - p0) + fabs(p0 - pm);
  return sn * fabs(pm - 2.0 * p0 + pp) / (cs * ptrad + (1.0 - cs) * ptvd);
}

// D_{m+1/2} from u at m-4..m+5 given as x[-4..5] with stride s
template <class R>
__device__ __forceinline__ R dface(const R *x, int s, R xk0, R xk10) {
  return (x[s] - x[0]) * xk0 +
         (R(D10_1) * (x[s] - x[0]) + R(D10_2) * (x[2 * s] - x[-s]) + R(D10_3) * (x[3 * s] - x[-2 * s]) +
          R(D10_4) * (x[4 * s] - x[-3 * s]) + R(D10_5) * (x[5 * s] - x[-4 * s])) *
             xk10;
}

// 10th difference of the face dissipation (without its amplitude)
template <class R>
__device__ __forceinline__ R d10face(const R *x, int s) {
  return R(D10_1) * (x[s] - x[0]) + R(D10_2) * (x[2 * s] - x[-s]) + R(D10_3) * (x[3 * s] - x[-2 * s]) +
         R(D10_4) * (x[4 * s] - x[-3 * s]) + R(D10_5) * (x[5 * s] - x[-4 * s]);
}

// Half-width of the shock filter's face m+1/2 (between rows m and m+1) along
// an axis of n rows with bounded faces bnd: the d2/d4/d6/d8 faces are the
// first four from a boundary, the boundary face with a tenth of the sensor
// term (filter_o10_shock, mod_filtering_shock.f90:266-466 i, :724-1057 j),
// d10 inside; <= 0: no face (the boundary rows get no increment along that
// axis, :1128-1140 runs over ndy_d..nfy_d).
__device__ __forceinline__ int face_half_n(int bnd, int n, int m) {
  int h = 5;
  if ((bnd & 1) && m < h) h = m;
  if ((bnd & 2) && n - m < h) h = n - m;
  return h;
}
__device__ __forceinline__ int face_half(const Geom &g, int m) { return face_half_n(g.jbnd, g.ny, m); }

// Edoh switch factor of direction d at element p (position r along d):
// min(1, CFL/chi) with CFL = sr_d deltat, times idx_d(r) on Cartesian blocks
// (mod_filtering_shock.f90:668-676 and the j, k repeats; c3 :627-632).
// Metric factor of direction d at position r: idx_d(r) on Cartesian
// directions, 1 where the spectral radius carries the metric (null array).
__device__ __forceinline__ double met_at(const double *met, int r) {
  return met ? ldg(met, (idx_t)(r - 1 + G)) : 1.0;
}

__device__ __forceinline__ double edoh_mul(const FilterArgs &a, int d, idx_t p, int r) {
  const double *met = d == 0 ? a.idx : (d == 1 ? a.idy : a.idz);
  double cfl = ldg(a.sr[d], p) * a.dt;
  if (met) cfl = cfl * ldg(met, (idx_t)(r - 1 + G));
  return fmin(1.0, cfl / a.chi);
}

// ARTV: the artificial viscosity.  The LDS "pressure" pencils then hold the
// spectral radii sr_x (i-pencil) and sr_y (j-pencil), the k queue sr_z.
// AVS: the artificial viscosity with shock capturing (artvisc_o9_shock,
// mod_artvisc_shock.f90:113-1287): the shock filter's faces D, each
// direction's K = (sr_{-1/2} D_{-1/2} - sr_{+1/2} D_{+1/2}) * metric with the
// spectral radius of U^n averaged to the faces (:480-550 i, :866-935 j,
// :1243-1283 k), U_n += deltat*K (update_varn_artvisc).
// WALL: blocks with boundary faces on i or j; rows 1..5 and n-4..n of a
// bounded axis get no increment along it here (k_shock_wall adds it with the
// reduced faces of the boundary closure).
// EDOH: the shock filter's Edoh switch, each direction's increment times
// min(1, CFL_d / chi) with the local CFL_d = sr_d deltat (idx_d on Cartesian
// blocks) (mod_filtering_shock.f90:620-700, :1059-1140, :1450-1496).
template <bool ARTV, bool WALL, bool EDOH, bool AVS>
__global__ __launch_bounds__(FT, 2) void k_filter_shock(const Batch<FilterArgs> *__restrict__ tab, int nb,
                                                        int ntiles) {
  static_assert(!(WALL && ARTV), "walls: not with artvisc_o9");
  static_assert(!(EDOH && (ARTV || AVS)), "the Edoh switch belongs to the shock filter");
  static_assert(!(AVS && ARTV), "one artificial viscosity at a time");
  using R = double;
  __shared__ R sux[5][FY][FX + 10];
  __shared__ R suy[5][FY + 10][FX];
  __shared__ double spx[FY][FX + 4];    // p at i0-2 .. i0+FX+1
  __shared__ double spy[FY + 4][FX];    // p at j0-2 .. j0+FY+1
  __shared__ double ssx[FY][FX + 4];    // sensor at i0-2 .. i0+FX+1 (ends unused)
  __shared__ double ssy[FY + 4][FX];    // sensor at j0-2 .. j0+FY+1 (ends unused)
  __shared__ double psx[FY][FX + 2];    // psens_x at i0-1 .. i0+FX
  __shared__ double psy[FY + 2][FX];    // psens_y at j0-1 .. j0+FY
  // Edoh factors of the plane's own points (AVS: the face spectral radii
  // sr_{-1/2}, sr_{+1/2} of the three directions, then their metrics; ARTV:
  // the metrics)
  __shared__ double smul[EDOH || ARTV ? 3 : (AVS ? 9 : 1)][FY][FX];

  int i0, j0, kbeg, kend;
  const FilterArgs &a = tile_of(tab, nb, ntiles, i0, j0, kbeg, kend);
  // the block's array pointers, read once from the table into registers
  const CState uinp = a.uin;
  const State uoutp = a.uout;
  const double *const prsp = ARTV ? a.sr[2] : a.prs, *const sensp = a.sens;
  const double *const px = ARTV ? a.sr[0] : a.prs, *const py = ARTV ? a.sr[1] : a.prs;
  const Geom &g = a.g;
  const int tid = threadIdx.x, tx = tid % FX, ty = tid / FX;
  const int i = i0 + tx, j = j0 + ty;
  const bool own = i <= g.nx && j <= g.ny;
  const bool needx = i <= g.nx + 5 && j <= g.ny, needy = i <= g.nx && j <= g.ny + 5;
  const bool duc = !ARTV && a.sens != nullptr;
  const double cs = a.csens;
  const idx_t sz = g.sy, own_off = lin(g, i, j, 0);
  Halo hl;
  halo_setup(g, i0, j0, hl);
  // p / sensor cells of this thread: rows of spx (i-direction) then of spy
  idx_t pp_off[NSP];
  int pp_r[NSP], pp_c[NSP];
  bool pp_on[NSP], pp_x[NSP], pp_s[NSP];
#pragma unroll
  for (int s = 0; s < NSP; s++) {
    const int t = tid + s * FT;
    pp_on[s] = false;
    pp_x[s] = true;
    pp_s[s] = false;
    pp_r[s] = pp_c[s] = 0;
    pp_off[s] = 0;
    if (t < NPX) {
      const int r = t / (FX + 4), c = t % (FX + 4);
      const int gi = i0 - 2 + c, gj = j0 + r;
      pp_on[s] = gj <= g.ny && gi <= g.nx + 2;
      pp_s[s] = c >= 1 && c <= FX + 2 && gi <= g.nx + 1;  // sensor needed at i0-1..i0+FX
      pp_r[s] = r;
      pp_c[s] = c;
      pp_off[s] = lin(g, gi, gj, 0);
    } else if (t < NPX + NPY) {
      const int tt = t - NPX;
      const int r = tt / FX, c = tt % FX;
      const int gi = i0 + c, gj = j0 - 2 + r;
      pp_on[s] = gi <= g.nx && gj <= g.ny + 2;
      pp_s[s] = r >= 1 && r <= FY + 2 && gj <= g.ny + 1;
      pp_x[s] = false;
      pp_r[s] = r;
      pp_c[s] = c;
      pp_off[s] = lin(g, gi, gj, 0);
    }
  }

  R q[11][5];        // U at k-5..k+5
  double qp[5];      // p at k-2..k+2 (ARTV: |w|+c at k-1..k+1 in qp[0..2])
  double qs[3];      // sensor at k-1..k+1
  if (own) {
#pragma unroll
    for (int l = 0; l < 10; l++) ldr(uinp, own_off + (idx_t)(kbeg - 5 + l) * sz, q[l]);
    if constexpr (ARTV) {
#pragma unroll
      for (int l = 0; l < 2; l++) qp[l] = ldg(prsp, own_off + (idx_t)(kbeg - 1 + l) * sz);
    } else {
#pragma unroll
      for (int l = 0; l < 4; l++) qp[l] = ldg(prsp, own_off + (idx_t)(kbeg - 2 + l) * sz);
    }
#pragma unroll
    for (int l = 0; l < 2; l++) qs[l] = duc ? ldg(sensp, own_off + (idx_t)(kbeg - 1 + l) * sz) : 1.0;
  }

  // every global load of a plane is issued one plane ahead, during the
  // previous plane's stencils: queue points, pencil cell, halo cells, p/sensor
  R qn[5], uc[5], uh[NS][5];
  double qpn = 0.0, qsn = 1.0, pv[NSP], sv[NSP];
  auto load_plane = [&](int kk) {
    const idx_t pkk = own_off + (idx_t)kk * sz;
    if (own) {
      ldr(uinp, pkk + 5 * sz, qn);
      qpn = ARTV ? ldg(prsp, pkk + sz) : ldg(prsp, pkk + 2 * sz);
      qsn = duc ? ldg(sensp, pkk + sz) : 1.0;
    }
    if (!own && (needx || needy)) ldr(uinp, pkk, uc);
#pragma unroll
    for (int s = 0; s < NS; s++)
      if (hl.on[s]) ldr(uinp, hl.off[s] + (idx_t)kk * sz, uh[s]);
#pragma unroll
    for (int s = 0; s < NSP; s++) {
      if (pp_on[s]) {
        pv[s] = ldg(pp_x[s] ? px : py, pp_off[s] + (idx_t)kk * sz);
        sv[s] = (duc && pp_s[s]) ? ldg(sensp, pp_off[s] + (idx_t)kk * sz) : 1.0;
      }
    }
  };
  load_plane(kbeg);

  for (int k = kbeg; k <= kend; k++) {
    const idx_t pk = own_off + (idx_t)k * sz;
    if (own) {
#pragma unroll
      for (int v = 0; v < 5; v++) q[10][v] = qn[v];
      if constexpr (ARTV) qp[2] = qpn;
      else qp[4] = qpn;
      qs[2] = qsn;
    }
    // ---- stage the plane in LDS ----
    if (needx || needy) {
#pragma unroll
      for (int v = 0; v < 5; v++) {
        const R x = own ? q[5][v] : uc[v];
        if (needx) sux[v][ty][tx + 5] = x;
        if (needy) suy[v][ty + 5][tx] = x;
      }
    }
#pragma unroll
    for (int s = 0; s < NS; s++) {
      if (hl.on[s]) {
#pragma unroll
        for (int v = 0; v < 5; v++) {
          if (hl.xdir[s]) sux[v][hl.r[s]][hl.c[s]] = uh[s][v];
          else suy[v][hl.r[s]][hl.c[s]] = uh[s][v];
        }
      }
    }
#pragma unroll
    for (int s = 0; s < NSP; s++) {
      if (pp_on[s]) {
        if (pp_x[s]) {
          spx[pp_r[s]][pp_c[s]] = pv[s];
          ssx[pp_r[s]][pp_c[s]] = sv[s];
        } else {
          spy[pp_r[s]][pp_c[s]] = pv[s];
          ssy[pp_r[s]][pp_c[s]] = sv[s];
        }
      }
    }
    if constexpr (EDOH) {
      // computed here and kept in LDS, so that nothing of them stays in
      // registers through the stencils
      if (own) {
        smul[0][ty][tx] = edoh_mul(a, 0, pk, i);
        smul[1][ty][tx] = edoh_mul(a, 1, pk, j);
        smul[2][ty][tx] = edoh_mul(a, 2, pk, k);
      }
    }
    if constexpr (AVS) {
      if (own) {
        double s0 = ldg(a.sr[0], pk);
        smul[0][ty][tx] = 0.5 * (ldg(a.sr[0], pk - 1) + s0);
        smul[1][ty][tx] = 0.5 * (ldg(a.sr[0], pk + 1) + s0);
        s0 = ldg(a.sr[1], pk);
        smul[2][ty][tx] = 0.5 * (ldg(a.sr[1], pk - g.sx) + s0);
        smul[3][ty][tx] = 0.5 * (ldg(a.sr[1], pk + g.sx) + s0);
        s0 = ldg(a.sr[2], pk);
        smul[4][ty][tx] = 0.5 * (ldg(a.sr[2], pk - sz) + s0);
        smul[5][ty][tx] = 0.5 * (ldg(a.sr[2], pk + sz) + s0);
        smul[6][ty][tx] = met_at(a.idx, i);
        smul[7][ty][tx] = met_at(a.idy, j);
        smul[8][ty][tx] = met_at(a.idz, k);
      }
    }
    if constexpr (ARTV) {
      if (own) {
        smul[0][ty][tx] = met_at(a.idx, i);
        smul[1][ty][tx] = met_at(a.idy, j);
        smul[2][ty][tx] = met_at(a.idz, k);
      }
    }
    __syncthreads();
    if (k < kend) load_plane(k + 1);
    // pressure sensors of the plane
    for (int t = tid; t < (ARTV ? 0 : FY * (FX + 2)); t += FT) {
      const int r = t / (FX + 2), c = t % (FX + 2);
      const int gi = i0 - 1 + c, gj = j0 + r;
      if (gj <= g.ny && gi <= g.nx + 1)
        psx[r][c] = psens3(spx[r][c], spx[r][c + 1], spx[r][c + 2], ssx[r][c + 1], cs);
    }
    for (int t = tid; t < (ARTV ? 0 : (FY + 2) * FX); t += FT) {
      const int r = t / FX, c = t % FX;
      const int gi = i0 + c, gj = j0 - 1 + r;
      if (gi <= g.nx && gj <= g.ny + 1)
        psy[r][c] = psens3(spy[r][c], spy[r + 1][c], spy[r + 2][c], ssy[r + 1][c], cs);
    }
    __syncthreads();

    if (ARTV && own) {
      // K_d = (sr_{+1/2} D_{+1/2} - sr_{-1/2} D_{-1/2}) * metric with D = -x10 * d10face
      const double x10 = a.xnu10;
      double K[5];
      {
        const double s0 = spx[ty][tx + 2], srp = 0.5 * (spx[ty][tx + 3] + s0), srm = 0.5 * (spx[ty][tx + 1] + s0);
        const double m = smul[0][ty][tx];
#pragma unroll
        for (int v = 0; v < 5; v++) {
          const R *x = &sux[v][ty][tx + 5];
          K[v] = (srm * (x10 * (double)d10face<R>(x - 1, 1)) - srp * (x10 * (double)d10face<R>(x, 1))) * m;
        }
      }
      {
        const double s0 = spy[ty + 2][tx], srp = 0.5 * (spy[ty + 3][tx] + s0), srm = 0.5 * (spy[ty + 1][tx] + s0);
        const double m = smul[1][ty][tx];
#pragma unroll
        for (int v = 0; v < 5; v++) {
          const R *y = &suy[v][ty + 5][tx];
          K[v] = K[v] + (srm * (x10 * (double)d10face<R>(y - FX, FX)) - srp * (x10 * (double)d10face<R>(y, FX))) * m;
        }
      }
      {
        const double s0 = qp[1], srp = 0.5 * (qp[2] + s0), srm = 0.5 * (qp[0] + s0);
        const double m = smul[2][ty][tx];
#pragma unroll
        for (int v = 0; v < 5; v++) {
          const double dlo = R(D10_1) * (q[5][v] - q[4][v]) + R(D10_2) * (q[6][v] - q[3][v]) +
                             R(D10_3) * (q[7][v] - q[2][v]) + R(D10_4) * (q[8][v] - q[1][v]) +
                             R(D10_5) * (q[9][v] - q[0][v]);
          const double dhi = R(D10_1) * (q[6][v] - q[5][v]) + R(D10_2) * (q[7][v] - q[4][v]) +
                             R(D10_3) * (q[8][v] - q[3][v]) + R(D10_4) * (q[9][v] - q[2][v]) +
                             R(D10_5) * (q[10][v] - q[1][v]);
          K[v] = K[v] + (srm * (x10 * dlo) - srp * (x10 * dhi)) * m;
        }
      }
#pragma unroll
      for (int v = 0; v < 5; v++) {
        stg_nt(uoutp.v[v], pk, q[5][v] + a.dt * K[v]);
      }
    }
    if (!ARTV && own) {
      R K[5];
      // i faces i-1/2 and i+1/2 (bounded i: rows 1..5 and nx-4..nx take theirs from k_shock_wall)
      if (WALL && (face_half_n(g.ibnd, g.nx, i - 1) < 5 || face_half_n(g.ibnd, g.nx, i) < 5)) {
#pragma unroll
        for (int v = 0; v < 5; v++) K[v] = R(0);
      } else {
        const double pm = psx[ty][tx], p0 = psx[ty][tx + 1], pp = psx[ty][tx + 2];
        const double xk0m = a.xnu0 * fmax(p0, pm), xk10m = fmax(0.0, a.xnu10 + xk0m / 105.0);
        const double xk0p = a.xnu0 * fmax(pp, p0), xk10p = fmax(0.0, a.xnu10 + xk0p / 105.0);
        const R mul = EDOH ? (R)smul[0][ty][tx] : R(1);
        double srm = 0.0, srp = 0.0, m = 1.0;
        if constexpr (AVS) {
          srm = smul[0][ty][tx];
          srp = smul[1][ty][tx];
          m = smul[6][ty][tx];
        }
#pragma unroll
        for (int v = 0; v < 5; v++) {
          const R *x = &sux[v][ty][tx + 5];
          if constexpr (AVS)
            K[v] = (srm * dface<R>(x - 1, 1, (R)xk0m, (R)xk10m) - srp * dface<R>(x, 1, (R)xk0p, (R)xk10p)) * m;
          else if constexpr (EDOH)
            K[v] = mul * (dface<R>(x - 1, 1, (R)xk0m, (R)xk10m) - dface<R>(x, 1, (R)xk0p, (R)xk10p));
          else K[v] = dface<R>(x - 1, 1, (R)xk0m, (R)xk10m) - dface<R>(x, 1, (R)xk0p, (R)xk10p);
        }
      }
      // j faces (walls: rows 1..5 and ny-4..ny take theirs from k_shock_wall)
      if (!WALL || (face_half(g, j - 1) >= 5 && face_half(g, j) >= 5)) {
        const double pm = psy[ty][tx], p0 = psy[ty + 1][tx], pp = psy[ty + 2][tx];
        const double xk0m = a.xnu0 * fmax(p0, pm), xk10m = fmax(0.0, a.xnu10 + xk0m / 105.0);
        const double xk0p = a.xnu0 * fmax(pp, p0), xk10p = fmax(0.0, a.xnu10 + xk0p / 105.0);
        const R mul = EDOH ? (R)smul[1][ty][tx] : R(1);
        double srm = 0.0, srp = 0.0, m = 1.0;
        if constexpr (AVS) {
          srm = smul[2][ty][tx];
          srp = smul[3][ty][tx];
          m = smul[7][ty][tx];
        }
#pragma unroll
        for (int v = 0; v < 5; v++) {
          const R *y = &suy[v][ty + 5][tx];
          if constexpr (AVS)
            K[v] = K[v] +
                   (srm * dface<R>(y - FX, FX, (R)xk0m, (R)xk10m) - srp * dface<R>(y, FX, (R)xk0p, (R)xk10p)) * m;
          else if constexpr (EDOH)
            K[v] = K[v] + mul * (dface<R>(y - FX, FX, (R)xk0m, (R)xk10m) - dface<R>(y, FX, (R)xk0p, (R)xk10p));
          else K[v] = K[v] + dface<R>(y - FX, FX, (R)xk0m, (R)xk10m) - dface<R>(y, FX, (R)xk0p, (R)xk10p);
        }
      }
      // k faces from the register queues
      {
        const double pm = psens3(qp[0], qp[1], qp[2], qs[0], cs);
        const double p0 = psens3(qp[1], qp[2], qp[3], qs[1], cs);
        const double pp = psens3(qp[2], qp[3], qp[4], qs[2], cs);
        const double xk0m = a.xnu0 * fmax(p0, pm), xk10m = fmax(0.0, a.xnu10 + xk0m / 105.0);
        const double xk0p = a.xnu0 * fmax(pp, p0), xk10p = fmax(0.0, a.xnu10 + xk0p / 105.0);
        const R mul = EDOH ? (R)smul[2][ty][tx] : R(1);
        double srm = 0.0, srp = 0.0, m = 1.0;
        if constexpr (AVS) {
          srm = smul[4][ty][tx];
          srp = smul[5][ty][tx];
          m = smul[8][ty][tx];
        }
#pragma unroll
        for (int v = 0; v < 5; v++) {
          const R dlo = (q[5][v] - q[4][v]) * (R)xk0m +
                        (R(D10_1) * (q[5][v] - q[4][v]) + R(D10_2) * (q[6][v] - q[3][v]) +
                         R(D10_3) * (q[7][v] - q[2][v]) + R(D10_4) * (q[8][v] - q[1][v]) +
                         R(D10_5) * (q[9][v] - q[0][v])) * (R)xk10m;
          const R dhi = (q[6][v] - q[5][v]) * (R)xk0p +
                        (R(D10_1) * (q[6][v] - q[5][v]) + R(D10_2) * (q[7][v] - q[4][v]) +
                         R(D10_3) * (q[8][v] - q[3][v]) + R(D10_4) * (q[9][v] - q[2][v]) +
                         R(D10_5) * (q[10][v] - q[1][v])) * (R)xk10p;
          if constexpr (AVS) K[v] = K[v] + (srm * dlo - srp * dhi) * m;
          else if constexpr (EDOH) K[v] = K[v] + mul * (dlo - dhi);
          else K[v] = K[v] + dlo - dhi;
        }
      }
#pragma unroll
      for (int v = 0; v < 5; v++) {
        if constexpr (AVS) stg_nt(uoutp.v[v], pk, q[5][v] + a.dt * K[v]);
        else stg_nt(uoutp.v[v], pk, q[5][v] + K[v]);
      }
    }
    __syncthreads();
#pragma unroll
    for (int l = 0; l < 10; l++)
#pragma unroll
      for (int v = 0; v < 5; v++) q[l][v] = q[l + 1][v];
    if constexpr (ARTV) {
      qp[0] = qp[1];
      qp[1] = qp[2];
    } else {
#pragma unroll
      for (int l = 0; l < 4; l++) qp[l] = qp[l + 1];
    }
    qs[0] = qs[1];
    qs[1] = qs[2];
  }
}

// Differences d2, d4, d6, d8, d10 of the face dissipation by half-width
// (init_filter_shock with iSF = 0, mod_filtering_shock.f90:75-107).
__constant__ double kShockD[6][5] = {{0.0, 0.0, 0.0, 0.0, 0.0},     {-1.0, 0.0, 0.0, 0.0, 0.0},
                                     {-3.0, 1.0, 0.0, 0.0, 0.0},    {-10.0, 5.0, -1.0, 0.0, 0.0},
                                     {-35.0, 21.0, -7.0, 1.0, 0.0}, {D10_1, D10_2, D10_3, D10_4, D10_5}};

constexpr int kShockWallThreads = 256;

// Shock filter increment along axis ax (0: i, 1: j) of the rows next to the
// bounded faces of that axis, added to the output of k_filter_shock<.., WALL>
// (which leaves rows 1..5 and n-4..n of a bounded axis without one): faces
// 1..4 from a boundary use d2/d4/d6/d8 with amplitude max(0, xnu + xk0), the
// boundary face a tenth of the sensor term (filter_o10_shock,
// mod_filtering_shock.f90:266-466 i, :724-870 jmin, :915-1057 jmax); the
// boundary rows get none (:1128-1140 over ndy_d..nfy_d).  The pressure
// sensor of row 1 (n) reads p one row beyond the face.  One thread per point
// of rows 2..5 and n-4..n-1 of each bounded side; nothing beyond a face is
// read but p.
// a.avs: artvisc_o9_shock (j only), whose faces 3 and 4 from a wall take
// xk0/5 and 3/70 xk0 (mod_artvisc_shock.f90:577-695 jmin, :744-862 jmax) and
// whose increment is deltat (sr_{-1/2} D_{-1/2} - sr_{+1/2} D_{+1/2}) idy(j).
__global__ __launch_bounds__(kShockWallThreads) void k_shock_wall(FilterArgs a, int ax) {
  const Geom &g = a.g;
  const int bnd = ax ? g.jbnd : g.ibnd, n = ax ? g.ny : g.nx;
  const int n1 = ax ? g.nx : g.ny;  // the other in-plane axis
  const long nface = (long)n1 * g.nz;
  const long t = (long)blockIdx.x * kShockWallThreads + threadIdx.x;
  if (t >= 8 * nface) return;
  const int q = (int)(t / nface);
  const int side = q >> 2;
  if (!((bnd >> side) & 1)) return;
  const long r = t - q * nface;
  const int a1 = 1 + (int)(r % n1), k = 1 + (int)(r / n1);
  const int m0 = side ? n - 4 + (q & 3) : 2 + (q & 3);  // row along ax
  const idx_t sj = ax ? (idx_t)g.sx : (idx_t)1;
  const idx_t pj = ax ? lin(g, a1, m0, k) : lin(g, m0, a1, k);
  // pressure sensors of rows m0-1, m0, m0+1 (faces m0-1/2, m0+1/2)
  double ps[3];
#pragma unroll
  for (int l = 0; l < 3; l++) {
    const idx_t pm = pj + (l - 1) * sj;
    const double sn = a.sens ? ldg(a.sens, pm) : 1.0;
    ps[l] = psens3(ldg(a.prs, pm - sj), ldg(a.prs, pm), ldg(a.prs, pm + sj), sn, a.csens);
  }
  double D[2][5];
#pragma unroll
  for (int f = 0; f < 2; f++) {
    const int m = m0 - 1 + f;
    const int hf = face_half_n(bnd, n, m);
    // no sensor term (artvisc_o9): p beyond the wall is not read into it
    double xk0 = a.xnu0 == 0.0 ? 0.0 : a.xnu0 * fmax(ps[f + 1], ps[f]);
    if (hf == 1) xk0 = xk0 / 10.0;
    double xkh;
    switch (hf) {
      case 1: xkh = fmax(0.0, 0.0 + xk0); break;
      case 2: xkh = fmax(0.0, a.chi / 12.0 + xk0); break;
      case 3: xkh = fmax(0.0, a.chi / 60.0 + (a.avs ? xk0 / 5.0 : xk0)); break;
      case 4: xkh = fmax(0.0, a.chi / 280.0 + (a.avs ? 3.0 / 70.0 * xk0 : xk0)); break;
      default: xkh = fmax(0.0, a.xnu10 + xk0 / 105.0);
    }
    const idx_t pf = pj + (f - 1) * sj;  // row m
#pragma unroll
    for (int v = 0; v < 5; v++) {
      const double *u = a.uin.v[v];
      const double d1 = ldg(u, pf + sj) - ldg(u, pf);
      double dd = 0.0;
      for (int l = 1; l <= hf; l++) dd = dd + kShockD[hf][l - 1] * (ldg(u, pf + l * sj) - ldg(u, pf + (1 - l) * sj));
      D[f][v] = d1 * xk0 + dd * xkh;
    }
  }
  if (a.avs) {
    const double s0 = ldg(a.sr[1], pj);
    const double srp = 0.5 * (ldg(a.sr[1], pj + sj) + s0), srm = 0.5 * (ldg(a.sr[1], pj - sj) + s0);
    const double m = met_at(a.idy, m0);
#pragma unroll
    for (int v = 0; v < 5; v++)
      stg(a.uout.v[v], pj, ldg(a.uout.v[v], pj) + a.dt * ((srm * D[0][v] - srp * D[1][v]) * m));
    return;
  }
  const double mul = a.edoh ? edoh_mul(a, ax, pj, m0) : 1.0;
#pragma unroll
  for (int v = 0; v < 5; v++) {
    const double kd = a.edoh ? mul * (D[0][v] - D[1][v]) : D[0][v] - D[1][v];
    stg(a.uout.v[v], pj, ldg(a.uout.v[v], pj) + kd);
  }
}

}  // namespace

hipError_t launch_shock_wall(const FilterArgs &a, int axis, hipStream_t s) {
  if (axis == 0 && a.avs) return hipErrorInvalidValue;  // artvisc_o9_shock: j-walls only
  const long n = 8L * (axis ? a.g.nx : a.g.ny) * a.g.nz;
  hipLaunchKernelGGL(k_shock_wall, dim3((unsigned)((n + kShockWallThreads - 1) / kShockWallThreads)),
                     dim3(kShockWallThreads), 0, s, a, axis);
  return hipGetLastError();
}

// Tiles of the filters: chunks of >= 16 planes, about 1024 workgroups per block.
TileRange filter_tiles(const Geom &g, int kc_req) {
  TileRange t{};
  t.tiles_x = (g.nx + FX - 1) / FX;
  t.tiles_y = (g.ny + FY - 1) / FY;
  const long nxy = (long)t.tiles_x * t.tiles_y;
  const long want = (1024 + nxy - 1) / nxy;
  int kc = kc_req > 0 ? kc_req : (int)((g.nz + want - 1) / want);
  if (kc_req <= 0 && kc < 16) kc = nxy * ((g.nz + 15) / 16) >= 512 ? 16 : (kc < 4 ? 4 : kc);
  if (kc > g.nz) kc = g.nz;
  t.kchunk = kc;
  return t;
}

hipError_t launch_filter_sf(const Batch<FilterArgs> *tab, int nb, int nt, bool wall, hipStream_t s) {
  if (nt <= 0) return hipSuccess;
  if (wall) hipLaunchKernelGGL((k_filter_sf<true>), dim3(nt), dim3(FT), 0, s, tab, nb, nt);
  else hipLaunchKernelGGL((k_filter_sf<false>), dim3(nt), dim3(FT), 0, s, tab, nb, nt);
  return hipGetLastError();
}

hipError_t launch_filter_shock(const Batch<FilterArgs> *tab, int nb, int nt, DissipKind kind, bool wall, bool edoh,
                               hipStream_t s) {
  if (nt <= 0) return hipSuccess;
  if (kind == kDissipArtvisc) {
    if (wall || edoh) return hipErrorInvalidValue;
    hipLaunchKernelGGL((k_filter_shock<true, false, false, false>), dim3(nt), dim3(FT), 0, s, tab, nb, nt);
  } else if (kind == kDissipArtviscShock) {
    if (edoh) return hipErrorInvalidValue;
    if (wall) hipLaunchKernelGGL((k_filter_shock<false, true, false, true>), dim3(nt), dim3(FT), 0, s, tab, nb, nt);
    else hipLaunchKernelGGL((k_filter_shock<false, false, false, true>), dim3(nt), dim3(FT), 0, s, tab, nb, nt);
  } else if (edoh) {
    if (wall) hipLaunchKernelGGL((k_filter_shock<false, true, true, false>), dim3(nt), dim3(FT), 0, s, tab, nb, nt);
    else hipLaunchKernelGGL((k_filter_shock<false, false, true, false>), dim3(nt), dim3(FT), 0, s, tab, nb, nt);
  } else {
    if (wall) hipLaunchKernelGGL((k_filter_shock<false, true, false, false>), dim3(nt), dim3(FT), 0, s, tab, nb, nt);
    else hipLaunchKernelGGL((k_filter_shock<false, false, false, false>), dim3(nt), dim3(FT), 0, s, tab, nb, nt);
  }
  return hipGetLastError();
}

}  // namespace mgpu
