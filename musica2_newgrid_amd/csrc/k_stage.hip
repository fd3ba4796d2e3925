// Fused RK stage: primitives -> Euler fluxes -> 11-pt divergence in i, j, k
// (+ the viscous increment of U^n, precomputed by k_vdiv, in the last stage)
// -> U_out = U^n - alpha*K.
//
// Follows, in one pass per stage, what the reference does in six array
// sweeps plus the update:
//   primitives_pfg          src/Fluid/smod_primitive_pfg.f90:12-59
//   flux_euler_11pts        src/Compute_flow/smod_flux_euler_11pts.f90:25-172,267-,601-
//   flux_euler_11pts_c3     src/Compute_flow/smod_flux_euler_11pts_c3.f90:25-38,139-171,763-773
//   flux_euler_11pts_c      src/Compute_flow/smod_flux_euler_11pts_c.f90:25-37,265-277,527-560
//                           (2.5-D: the 2-D metrics of a column are loaded once
//                           for its whole k-march; z is Cartesian)
//   flux_visc_5pts(_c3) divergence part  smod_flux_visc_5pts.f90:119-130, _c3.f90:262-
//   update_var              src/Time/update.f90:32-44
// With no-slip walls on j-faces (WALL): no Euler increment on the wall rows
// (ndx_e = 2, mpi_index_bounds.f90:20-70) and the centred a3/a5/a7/a9
// closures of rows 2..5 (smod_flux_euler_11pts.f90:267-595 with
// is_bc_1pt), whose metric is the idy entry of the row.
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
#include <type_traits>

#include "mgpu_internal.h"

namespace mgpu {

namespace {

struct Prim {
  double r, ru, rv, rw, re, u, v, w, pr;
};

template <class R>
__device__ __forceinline__ R d11c(const R *f, int s) {
  return R(A11_1) * (f[s] - f[-s]) + R(A11_2) * (f[2 * s] - f[-2 * s]) + R(A11_3) * (f[3 * s] - f[-3 * s]) +
         R(A11_4) * (f[4 * s] - f[-4 * s]) + R(A11_5) * (f[5 * s] - f[-5 * s]);
}

// Coefficients of the centred j-derivative that row j uses next to a wall
// (a3, a5, a7, a9 by half-width, a11 inside), padded with zeros to 5 terms.
// The j-flux pencil holds 0 beyond a wall, so the padded terms add exact
// zeros.
__device__ __forceinline__ void wall_coeffs(int hw, double c[5]) {
  c[0] = hw == 5 ? A11_1 : (hw == 4 ? A9_1 : (hw == 3 ? A7_1 : (hw == 2 ? A5_1 : (hw == 1 ? A3_1 : 0.0))));
  c[1] = hw == 5 ? A11_2 : (hw == 4 ? A9_2 : (hw == 3 ? A7_2 : (hw == 2 ? A5_2 : 0.0)));
  c[2] = hw == 5 ? A11_3 : (hw == 4 ? A9_3 : (hw == 3 ? A7_3 : 0.0));
  c[3] = hw == 5 ? A11_4 : (hw == 4 ? A9_4 : 0.0);
  c[4] = hw == 5 ? A11_5 : 0.0;
}

// Raw inputs of one flux evaluation: the 5 conserved variables and the 3
// metric terms of the flux direction (c3).
struct Raw {
  double u[5];
  double m[3];
};

// Array pointers one stage launch uses for its block.
struct Ptrs {
  const double *uin[5];
  const float *uin32[5];
  const double *u0[5];
  double *uout[5];
  float *uout32[5];
  const double *m[10];
  const double *m2[5];
  const double *kv[4];
  const double *idx, *idy, *idz;
  double ref[5];
};

template <int GEO, bool VISC, bool IN32, bool OUT32>
__device__ __forceinline__ void load_ptrs(const StageArgs &a, Ptrs &P) {
#pragma unroll
  for (int v = 0; v < 5; v++) {
    if constexpr (IN32) P.uin32[v] = a.uin32[v];
    else P.uin[v] = a.uin.v[v];
    if constexpr (OUT32) P.uout32[v] = a.uout32[v];
    else P.uout[v] = a.uout.v[v];
    P.u0[v] = a.u0.v[v];
  }
  if constexpr (GEO == kGeoC3) {
#pragma unroll
    for (int q = 0; q < 10; q++) P.m[q] = a.met.m[q];
  } else if constexpr (GEO == kGeoC2) {
#pragma unroll
    for (int q = 0; q < 5; q++) P.m2[q] = a.met.m2[q];
    P.idz = a.met.idz;
  } else {
    P.idx = a.met.idx;
    P.idy = a.met.idy;
    P.idz = a.met.idz;
  }
  if constexpr (VISC) {
#pragma unroll
    for (int c = 0; c < 4; c++) P.kv[c] = a.kv[c];
  }
  if constexpr (IN32 || OUT32) {
    P.ref[0] = a.ref.r;
    P.ref[1] = a.ref.ru;
    P.ref[2] = a.ref.rv;
    P.ref[3] = a.ref.rw;
    P.ref[4] = a.ref.re;
  }
}

// Stage state at element p: FP64, or (IN32) the reference state plus the
// stored FP32 difference.
template <bool IN32>
__device__ __forceinline__ void load_u(const Ptrs &P, idx_t p, double u[5]) {
  if constexpr (IN32) {
#pragma unroll
    for (int v = 0; v < 5; v++) u[v] = P.ref[v] + (double)ldg(P.uin32[v], p);
  } else {
#pragma unroll
    for (int v = 0; v < 5; v++) u[v] = ldg(P.uin[v], p);
  }
}

template <bool CURV, bool IN32>
__device__ __forceinline__ void load_raw(const Ptrs &P, idx_t p, int dir, Raw &r) {
  load_u<IN32>(P, p, r.u);
  if constexpr (CURV) {
#pragma unroll
    for (int c = 0; c < 3; c++) r.m[c] = ldg(P.m[3 * dir + c], p);
  }
}

// Raw inputs of a halo cell, whose direction is known at run time only:
// branch to a constant direction so that the pointer table stays in registers.
template <bool CURV, bool IN32>
__device__ __forceinline__ void load_raw_halo(const Ptrs &P, idx_t p, int dir, Raw &r) {
  if (dir == 0) load_raw<CURV, IN32>(P, p, 0, r);
  else load_raw<CURV, IN32>(P, p, 1, r);
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

// Flux along dir minus the flux of the reference state, in FP32 (mixed
// precision).  With W = (rho u).m, vc = u.m and primes for differences with
// the reference state (taken in FP64, then rounded):
//   F'0 = W',  F'(1+a) = W' u_a + W_r u'_a + p' m_a,  F'4 = H' vc + H_r vc'
// which equals F(U) - F(U_ref) exactly in real arithmetic.
template <bool CURV>
__device__ __forceinline__ void flux_mix(const Prim &q, int dir, const double m[3], const RefState &rf, float F[5]) {
  const float uu[3] = {(float)q.u, (float)q.v, (float)q.w};
  const float up[3] = {(float)(q.u - rf.u), (float)(q.v - rf.v), (float)(q.w - rf.w)};
  const float wp[3] = {(float)(q.ru - rf.ru), (float)(q.rv - rf.rv), (float)(q.rw - rf.rw)};
  const double pd = q.pr - rf.p;
  const float pp = (float)pd;
  const float hp = (float)((q.re - rf.re) + pd);
  float mm[3];
  if constexpr (CURV) {
    mm[0] = (float)m[0];
    mm[1] = (float)m[1];
    mm[2] = (float)m[2];
  } else {
    mm[0] = dir == 0 ? 1.f : 0.f;
    mm[1] = dir == 1 ? 1.f : 0.f;
    mm[2] = dir == 2 ? 1.f : 0.f;
  }
  const float W1 = wp[0] * mm[0] + wp[1] * mm[1] + wp[2] * mm[2];
  const float Wr = (float)rf.ru * mm[0] + (float)rf.rv * mm[1] + (float)rf.rw * mm[2];
  const float vc = uu[0] * mm[0] + uu[1] * mm[1] + uu[2] * mm[2];
  const float vcp = up[0] * mm[0] + up[1] * mm[1] + up[2] * mm[2];
  F[0] = W1;
#pragma unroll
  for (int c = 0; c < 3; c++) F[1 + c] = W1 * uu[c] + Wr * up[c] + pp * mm[c];
  F[4] = hp * vc + (float)rf.H * vcp;
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

// FP64 flux, or the FP32 flux perturbation of the mixed path.
template <bool CURV, bool MIX, class R>
__device__ __forceinline__ void flux_r(const Prim &q, int dir, const double m[3], const RefState &rf, R F[5]) {
  if constexpr (MIX) flux_mix<CURV>(q, dir, m, rf, F);
  else flux_m<CURV>(q, dir, m, F);
}

// Flux of a halo cell, whose direction (i or j) is only known at run time:
// branch to a constant direction so that the primitives stay in registers.
template <bool CURV, bool MIX, class R>
__device__ __forceinline__ void flux_halo(const Prim &q, int dir, const double m[3], const RefState &rf, R F[5]) {
  if (dir == 0) flux_r<CURV, MIX>(q, 0, m, rf, F);
  else flux_r<CURV, MIX>(q, 1, m, rf, F);
}

template <int TX, int TY, int GEO, bool VISC, int MODE, bool WALL>
__global__ __launch_bounds__(TX *TY, 2) void k_stage(const Batch<StageArgs> *__restrict__ tab, int nb, int ntiles) {
  // CURV: contravariant i/j fluxes (c3 or 2.5-D); C3: metrics at every point
  // (and a contravariant k flux); C2: 2-D metrics per column, Cartesian k
  constexpr bool CURV = GEO != kGeoCart;
  constexpr bool C3 = GEO == kGeoC3;
  constexpr bool C2 = GEO == kGeoC2;
  static_assert(!WALL || !CURV, "walls: Cartesian blocks");
  constexpr bool MIX = MODE != kFP64;
  constexpr bool IN32 = MODE == kMixMiddle || MODE == kMixLast;
  constexpr bool OUT32 = MODE == kMixFirst || MODE == kMixMiddle;
  using R = typename std::conditional<MIX, float, double>::type;
  constexpr int NT = TX * TY;
  constexpr int NHX = 10 * TY;            // i-halo cells per plane
  constexpr int NH = NHX + 10 * TX;       // all halo cells per plane
  constexpr int NS = (NH + NT - 1) / NT;  // halo cells per thread
  // flux pencils, double-buffered over k planes so that one barrier per
  // plane suffices (plane k+1 is written while plane k may still be read)
  __shared__ R sfx2[2][5][TY][TX + 10];
  __shared__ R sfy2[2][5][TY + 10][TX];

  const int w = xcd_remap(blockIdx.x, ntiles);
  const Batch<StageArgs> &e = batch_find(tab, nb, w);
  const StageArgs &a = e.a;
  // the array pointers of the block, read once: they stay in scalar
  // registers for the whole march instead of being re-read from the table
  // in every plane
  Ptrs P;
  load_ptrs<GEO, VISC, IN32, OUT32>(a, P);
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
  // 2.5-D: the metric rows of the halo cells and of the own column, and
  // 1/J, constant along k: loaded once
  double hm[NS][2];
  double rxm[3] = {0, 0, 0}, rym[3] = {0, 0, 0};
  double ijac = 0.0;
  if constexpr (C2) {
#pragma unroll
    for (int s = 0; s < NS; s++) {
      hm[s][0] = hm[s][1] = 0.0;
      if (hs_on[s]) {
        const idx_t p2 = hs_off[s] - (idx_t)(G - 1) * sz;  // the cell's (i, j) column
        // (constant indices only: a run-time index would move the pointer table to scratch)
        const bool xd = hs_dir[s] == 0;
        hm[s][0] = ldg(xd ? P.m2[0] : P.m2[2], p2);
        hm[s][1] = ldg(xd ? P.m2[1] : P.m2[3], p2);
      }
    }
    if (needx || needy) {
      const idx_t p2 = lin2(g, i, j);
      rxm[0] = ldg(P.m2[0], p2);
      rxm[1] = ldg(P.m2[1], p2);
      rym[0] = ldg(P.m2[2], p2);
      rym[1] = ldg(P.m2[3], p2);
      if (own) ijac = ldg(P.m2[4], p2);
    }
  }
  // walls: j-pencil entries beyond a wall hold 0 (their state is never valid),
  // and the row's closure coefficients
  bool hs_beyond[NS];
  bool own_beyond = false;
  double cw[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  int hw = 5;
  if constexpr (WALL) {
#pragma unroll
    for (int s = 0; s < NS; s++) {
      const int gj = j0 - 5 + hs_r[s];
      hs_beyond[s] = hs_dir[s] == 1 && (((g.jwall & 1) && gj < 1) || ((g.jwall & 2) && gj > g.ny));
    }
    own_beyond = (g.jwall & 2) && j > g.ny;
    hw = jhalf(g, j);
    wall_coeffs(hw, cw);
  } else {
#pragma unroll
    for (int s = 0; s < NS; s++) hs_beyond[s] = false;
  }

  const RefState &rf = a.ref;
  const double alpha = a.alpha;
  R q[11][5];
  // raw inputs of the queue point k+5, loaded one plane ahead (after the
  // previous plane's barrier) so that the load is in flight during that
  // plane's derivatives
  Raw rq;
  if (own) {
#pragma unroll
    for (int l = 0; l < 10; l++) {
      Raw r;
      load_raw<C3, IN32>(P, own_off + (idx_t)(kbeg - 5 + l) * sz, 2, r);
      flux_r<C3, MIX>(prim_raw(r.u, gam1), 2, r.m, rf, q[l]);
    }
    load_raw<C3, IN32>(P, own_off + (idx_t)(kbeg + 5) * sz, 2, rq);
  }
  // the own point's state and (c3) i/j metric terms, likewise one plane ahead
  double ru[5] = {0, 0, 0, 0, 0};
  auto load_own = [&](idx_t p) {
    load_u<IN32>(P, p, ru);
    if constexpr (C3) {
#pragma unroll
      for (int c = 0; c < 3; c++) {
        rxm[c] = ldg(P.m[c], p);
        rym[c] = ldg(P.m[3 + c], p);
      }
    }
  };
  if (needx || needy) load_own(own_off + (idx_t)kbeg * sz);

  for (int k = kbeg; k <= kend; k++) {
    const idx_t pk = own_off + (idx_t)k * sz;
    R (*sfx)[TY][TX + 10] = sfx2[(k - kbeg) & 1];
    R (*sfy)[TY + 10][TX] = sfy2[(k - kbeg) & 1];
    // ---- batch 1: U^n and 1/J at k, halo cell 0 (queue point and own point: loaded ahead) ----
    Raw rh0;
    double u0[5] = {0, 0, 0, 0, 0};
    if (hs_on[0]) {
      load_raw_halo<C3, IN32>(P, hs_off[0] + (idx_t)k * sz, hs_dir[0], rh0);
      if constexpr (C2) {
        rh0.m[0] = hm[0][0];
        rh0.m[1] = hm[0][1];
        rh0.m[2] = 0.0;
      }
    }
    double kvis[4] = {0, 0, 0, 0};
    if (own) {
#pragma unroll
      for (int v = 0; v < 5; v++) u0[v] = ldg_nt(P.u0[v], pk);
      if constexpr (C3) ijac = ldg_nt(P.m[9], pk);
      if constexpr (VISC) {
#pragma unroll
        for (int c = 0; c < 4; c++) kvis[c] = ldg_nt(P.kv[c], pk);
      }
    }
    if (own) flux_r<C3, MIX>(prim_raw(rq.u, gam1), 2, rq.m, rf, q[10]);
    if (needx || needy) {
      const Prim pq = prim_raw(ru, gam1);
      R F[5];
      if (needx) {
        flux_r<CURV, MIX>(pq, 0, rxm, rf, F);
#pragma unroll
        for (int v = 0; v < 5; v++) sfx[v][ty][tx + 5] = F[v];
      }
      if (needy) {
        flux_r<CURV, MIX>(pq, 1, rym, rf, F);
#pragma unroll
        for (int v = 0; v < 5; v++) sfy[v][ty + 5][tx] = (WALL && own_beyond) ? R(0) : F[v];
      }
    }
    if (hs_on[0]) {
      R F[5];
      flux_halo<CURV, MIX>(prim_raw(rh0.u, gam1), hs_dir[0], rh0.m, rf, F);
      if (hs_dir[0] == 0) {
#pragma unroll
        for (int v = 0; v < 5; v++) sfx[v][hs_r[0]][hs_c[0]] = F[v];
      } else {
#pragma unroll
        for (int v = 0; v < 5; v++) sfy[v][hs_r[0]][hs_c[0]] = (WALL && hs_beyond[0]) ? R(0) : F[v];
      }
    }
    __builtin_amdgcn_sched_barrier(0);
    // ---- batch 2: the remaining halo cells ----
    if constexpr (NS > 1) {
      Raw rh[NS];
#pragma unroll
      for (int s = 1; s < NS; s++)
        if (hs_on[s]) {
          load_raw_halo<C3, IN32>(P, hs_off[s] + (idx_t)k * sz, hs_dir[s], rh[s]);
          if constexpr (C2) {
            rh[s].m[0] = hm[s][0];
            rh[s].m[1] = hm[s][1];
            rh[s].m[2] = 0.0;
          }
        }
#pragma unroll
      for (int s = 1; s < NS; s++) {
        if (hs_on[s]) {
          R F[5];
          flux_halo<CURV, MIX>(prim_raw(rh[s].u, gam1), hs_dir[s], rh[s].m, rf, F);
          if (hs_dir[s] == 0) {
#pragma unroll
            for (int v = 0; v < 5; v++) sfx[v][hs_r[s]][hs_c[s]] = F[v];
          } else {
#pragma unroll
            for (int v = 0; v < 5; v++) sfy[v][hs_r[s]][hs_c[s]] = (WALL && hs_beyond[s]) ? R(0) : F[v];
          }
        }
      }
    }
    __syncthreads();
    // queue point and own point of the next plane
    if (own && k < kend) load_raw<C3, IN32>(P, pk + 6 * sz, 2, rq);
    if ((needx || needy) && k < kend) load_own(pk + sz);

    if (own) {
      double K[5];
#pragma unroll
      for (int v = 0; v < 5; v++) {
        const R dx = d11c(&sfx[v][ty][tx + 5], 1);
        R dy;
        if constexpr (WALL) {
          const R *f = &sfy[v][ty + 5][tx];
          dy = R(cw[0]) * (f[TX] - f[-TX]) + R(cw[1]) * (f[2 * TX] - f[-2 * TX]) + R(cw[2]) * (f[3 * TX] - f[-3 * TX]) +
               R(cw[3]) * (f[4 * TX] - f[-4 * TX]) + R(cw[4]) * (f[5 * TX] - f[-5 * TX]);
        } else {
          dy = d11c(&sfy[v][ty + 5][tx], TX);
        }
        const R dz = R(A11_1) * (q[6][v] - q[4][v]) + R(A11_2) * (q[7][v] - q[3][v]) +
                     R(A11_3) * (q[8][v] - q[2][v]) + R(A11_4) * (q[9][v] - q[1][v]) +
                     R(A11_5) * (q[10][v] - q[0][v]);
        // metrics and Jacobian multiply in FP64
        if constexpr (C3) {
          K[v] = (double)(dz + (dy + dx)) * ijac;
        } else if constexpr (C2) {
          // (ksi + eta parts) * ijacob, then the Cartesian z part
          K[v] = (double)dz * ldg(P.idz, (idx_t)(k - 1 + G)) + (double)(dy + dx) * ijac;
        } else {
          const double kx = (double)dx * ldg(P.idx, (idx_t)(i - 1 + G));
          const double ky = (double)dy * ldg(P.idy, (idx_t)(j - 1 + G)) + kx;
          K[v] = (double)dz * ldg(P.idz, (idx_t)(k - 1 + G)) + ky;
        }
        // wall rows: no Euler increment (bc_wall zeroes it, flux_euler skips the row)
        if (WALL && hw == 0) K[v] = 0.0;
      }
      if constexpr (VISC) {
#pragma unroll
        for (int c = 0; c < 4; c++) K[1 + c] = K[1 + c] + kvis[c];
      }
      if constexpr (OUT32) {
        const double ref[5] = {rf.r, rf.ru, rf.rv, rf.rw, rf.re};
#pragma unroll
        for (int v = 0; v < 5; v++) stg_nt(P.uout32[v], pk, (float)((u0[v] - alpha * K[v]) - ref[v]));
      } else {
#pragma unroll
        for (int v = 0; v < 5; v++) stg_nt(P.uout[v], pk, u0[v] - alpha * K[v]);
      }
    }

#pragma unroll
    for (int l = 0; l < 10; l++)
#pragma unroll
      for (int v = 0; v < 5; v++) q[l][v] = q[l + 1][v];
  }
}

template <int TX, int TY, int GEO, int MODE>
hipError_t launch_g(const Batch<StageArgs> *tab, int nb, int nt, bool visc, hipStream_t s) {
  dim3 blk(TX * TY), grd(nt);
  if constexpr (MODE == kFP64 || MODE == kMixLast) {
    if (visc) {
      hipLaunchKernelGGL((k_stage<TX, TY, GEO, true, MODE, false>), grd, blk, 0, s, tab, nb, nt);
      return hipGetLastError();
    }
  }
  hipLaunchKernelGGL((k_stage<TX, TY, GEO, false, MODE, false>), grd, blk, 0, s, tab, nb, nt);
  return hipGetLastError();
}

template <int TX, int TY, int MODE>
hipError_t launch_t(const Batch<StageArgs> *tab, int nb, int nt, int geo, bool visc, bool wall, hipStream_t s) {
  dim3 blk(TX * TY), grd(nt);
  // the viscous increment is added in the last stage only
  constexpr bool VOK = MODE == kFP64 || MODE == kMixLast;
  if (visc && !VOK) return hipErrorInvalidValue;
  if (wall) {
    if (geo != kGeoCart) return hipErrorInvalidValue;
    if constexpr (VOK) {
      if (visc) {
        hipLaunchKernelGGL((k_stage<TX, TY, kGeoCart, true, MODE, true>), grd, blk, 0, s, tab, nb, nt);
        return hipGetLastError();
      }
    }
    hipLaunchKernelGGL((k_stage<TX, TY, kGeoCart, false, MODE, true>), grd, blk, 0, s, tab, nb, nt);
    return hipGetLastError();
  }
  if (geo == kGeoC3) return launch_g<TX, TY, kGeoC3, MODE>(tab, nb, nt, visc, s);
  if (geo == kGeoC2) return launch_g<TX, TY, kGeoC2, MODE>(tab, nb, nt, visc, s);
  return launch_g<TX, TY, kGeoCart, MODE>(tab, nb, nt, visc, s);
}

template <int MODE>
hipError_t launch_shape(const Batch<StageArgs> *tab, int nb, int nt, int geo, bool visc, int shape, bool wall,
                        hipStream_t s) {
  switch (shape) {
    case 1: return launch_t<32, 8, MODE>(tab, nb, nt, geo, visc, wall, s);
    case 2: return launch_t<64, 8, MODE>(tab, nb, nt, geo, visc, wall, s);
    case 3: return launch_t<16, 16, MODE>(tab, nb, nt, geo, visc, wall, s);
    case 4: return launch_t<32, 16, MODE>(tab, nb, nt, geo, visc, wall, s);
    default: return launch_t<64, 4, MODE>(tab, nb, nt, geo, visc, wall, s);
  }
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

hipError_t launch_stage(const Batch<StageArgs> *tab, int nb, int nt, int geo, bool visc, int shape, int mode,
                        bool wall, hipStream_t s) {
  if (nt <= 0) return hipSuccess;
  switch (mode) {
    case kMixFirst: return launch_shape<kMixFirst>(tab, nb, nt, geo, visc, shape, wall, s);
    case kMixMiddle: return launch_shape<kMixMiddle>(tab, nb, nt, geo, visc, shape, wall, s);
    case kMixLast: return launch_shape<kMixLast>(tab, nb, nt, geo, visc, shape, wall, s);
    default: return launch_shape<kFP64>(tab, nb, nt, geo, visc, shape, wall, s);
  }
}

}  // namespace mgpu
