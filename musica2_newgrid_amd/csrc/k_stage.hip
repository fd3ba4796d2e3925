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
// With boundary faces on i or j (WALL: no-slip walls, characteristic outflow,
// supersonic inlet): no Euler increment on the boundary rows (ndx_e = 2,
// mpi_index_bounds.f90:20-70) and the centred a3/a5/a7/a9 closures of rows
// 2..5 (smod_flux_euler_11pts.f90:36-133 along i, :267-595 along j, with
// is_bc_1pt), whose metric is the idx / idy entry of the row.  A wall row
// keeps K = 0 plus the viscous increment (bc_wall, smod_bc_wall.f90:226-232);
// the row of a characteristic or inlet face takes the increment its boundary
// routine wrote (bc_charac_imax / _jmax, mod_charac.f90:350-469, 816-969) or,
// at the inlet, the filter increment of the previous step that K still holds
// there.  The i-flux of an inlet row mixes the inflow state's conserved
// variables with the primitives computed before it was imposed
// (bc_supersonic_inlet after `primitives`, runge_kutta.f90:30-37, 64-70).
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

// Value a pencil cell holds for its flux f by code: 0 the flux, 1 zero (a
// cell beyond a bounded face, by a select: its flux may be NaN, from the
// unwritten ghosts of a stage set), 2 three quarters of it (one row beyond an
// advanced face, where the cell's thread evaluated the wall row's flux).
template <class R>
__device__ __forceinline__ R pencil_value(unsigned code, R f) {
  return code == 1u ? R(0) : (code == 2u ? R(-A02_1 / A02_2) * f : f);
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
  const double *kbc[4];
  const double *uin_row, *prim_row;
  const float *m32[9];
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
    if constexpr (IN32 || OUT32) {
#pragma unroll
      for (int q = 0; q < 9; q++) P.m32[q] = a.m32[q];
    }
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

// The flux-direction metric terms are read as FP32 in the mixed path (its
// flux arithmetic rounds them to FP32 anyway: the same values, half the
// bytes), as FP64 otherwise.
template <bool CURV, bool IN32, bool MIX>
__device__ __forceinline__ void load_raw(const Ptrs &P, idx_t p, int dir, Raw &r) {
  load_u<IN32>(P, p, r.u);
  if constexpr (CURV) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
      if constexpr (MIX) r.m[c] = (double)ldg(P.m32[3 * dir + c], p);
      else r.m[c] = ldg(P.m[3 * dir + c], p);
    }
  }
}

// Raw inputs of a halo cell, whose direction is known at run time only:
// branch to a constant direction so that the pointer table stays in registers.
template <bool CURV, bool IN32, bool MIX>
__device__ __forceinline__ void load_raw_halo(const Ptrs &P, idx_t p, int dir, Raw &r) {
  if (dir == 0) load_raw<CURV, IN32, MIX>(P, p, 0, r);
  else load_raw<CURV, IN32, MIX>(P, p, 1, r);
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

// WALL: 0 no bounded face; 1 the closures of bounded faces; 2 also advanced
// wall rows (is_wall2), whose Euler increment takes the one-sided a02 / a20
// derivative across the wall (smod_flux_euler_11pts.f90:42-54, 254-265,
// 286-297, 584-595): -3/2 f0 + 2 f1 - 1/2 f2 at a min face.  The wall row
// evaluates it with the centred form 2 (f1 - f-1) - 1/2 (f2 - f-2) of the
// closure rows: the pencil holds 3/4 of the wall row's flux one row beyond
// an advanced face and 0 further out, cells no other row reads.  The thread
// that fills that cell loads the wall row's point instead of its own and
// scales the flux by 3/4, so the k-march keeps the loop of the other wall
// variants.  A variant of its own all the same, so that they keep their
// registers.
template <int TX, int TY, int GEO, bool VISC, int MODE, int WALL>
__global__ __launch_bounds__(TX *TY, 2) void k_stage(const Batch<StageArgs> *__restrict__ tab, int nb, int ntiles) {
  // CURV: contravariant i/j fluxes (c3 or 2.5-D); C3: metrics at every point
  // (and a contravariant k flux); C2: 2-D metrics per column, Cartesian k
  constexpr bool CURV = GEO != kGeoCart;
  constexpr bool C3 = GEO == kGeoC3;
  constexpr bool C2 = GEO == kGeoC2;
  static_assert(!WALL || !C2, "walls: Cartesian or c3 blocks");
  // bounded i-faces (characteristic outflow, inlet): Cartesian blocks only;
  // c3 blocks take walls on their j-faces
  constexpr bool IBND = WALL && !CURV;
  constexpr bool ADV = WALL == 2;
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
  idx_t own_off = lin(g, i, j, 0);
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
  // bounded faces: pencil entries beyond a boundary face hold 0 (their state
  // is never valid), and the row's closure coefficients along j (cw) and i (ci)
  bool hs_beyond[NS];
  bool own_beyond = false, own_beyond_x = false;
  double cw[5] = {0.0, 0.0, 0.0, 0.0, 0.0}, ci[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  int hw = 5, hwi = 5;
  // ADV: one code per pencil cell in a single word (2 bits each: halo cell s
  // at bit 2s, the own point's j- and i-pencil cells at bits 28 and 30), see
  // pencil_value; cells one row beyond an advanced face take code 2
  unsigned acode = 0u;
  if constexpr (WALL) {
#pragma unroll
    for (int s = 0; s < NS; s++) {
      const int gj = j0 - 5 + hs_r[s], gi = i0 - 5 + hs_c[s];
      hs_beyond[s] = hs_dir[s] == 1 ? (((g.jbnd & 1) && gj < 1) || ((g.jbnd & 2) && gj > g.ny))
                                    : (((g.ibnd & 1) && gi < 1) || ((g.ibnd & 2) && gi > g.nx));
    }
    own_beyond = (g.jbnd & 2) && j > g.ny;
    hw = jhalf(g, j);
    wall_coeffs(hw, cw);
    if constexpr (ADV) {
      if (hw == 0 && (((g.jadv & 1) && j == 1) || ((g.jadv & 2) && j == g.ny))) {
        // the wall row takes an increment: not a row without one (hw 6)
        cw[0] = A02_2;
        cw[1] = A02_3;
        hw = 6;
      }
#pragma unroll
      for (int s = 0; s < NS; s++) {
        const int gj = j0 - 5 + hs_r[s], gi = i0 - 5 + hs_c[s];
        unsigned c = hs_beyond[s] ? 1u : 0u;
        if (hs_dir[s] == 1) {
          if ((g.jadv & 1) && gj == 0) {
            hs_off[s] += (idx_t)g.sx;
            c = 2u;
          } else if ((g.jadv & 2) && gj == g.ny + 1) {
            hs_off[s] -= (idx_t)g.sx;
            c = 2u;
          }
        } else if (IBND) {
          if ((g.iadv & 1) && gi == 0) {
            hs_off[s] += 1;
            c = 2u;
          } else if ((g.iadv & 2) && gi == g.nx + 1) {
            hs_off[s] -= 1;
            c = 2u;
          }
        }
        acode |= c << (2 * s);
      }
      unsigned cy = own_beyond ? 1u : 0u;
      if ((g.jadv & 2) && j == g.ny + 1) {
        // serves the j-pencil only (needx needs j <= ny)
        own_off -= (idx_t)g.sx;
        cy = 2u;
      }
      acode |= cy << 28;
    }
    if constexpr (IBND) {
      own_beyond_x = (g.ibnd & 2) && i > g.nx;
      hwi = ihalf(g, i);
      wall_coeffs(hwi, ci);
      if constexpr (ADV) {
        if (hwi == 0 && (((g.iadv & 1) && i == 1) || ((g.iadv & 2) && i == g.nx))) {
          ci[0] = A02_2;
          ci[1] = A02_3;
          hwi = 6;
        }
        unsigned cx = own_beyond_x ? 1u : 0u;
        if ((g.iadv & 2) && i == g.nx + 1) {
          // serves the i-pencil only (needy needs i <= nx)
          own_off -= 1;
          cx = 2u;
        }
        acode |= cx << 30;
      }
    }
    P.kbc[0] = a.kbc[0];
    P.kbc[1] = a.kbc[1];
    P.kbc[2] = a.kbc[2];
    P.kbc[3] = a.kbc[3];
    P.uin_row = a.uin_row;
    P.prim_row = a.prim_row;
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
      load_raw<C3, IN32, MIX>(P, own_off + (idx_t)(kbeg - 5 + l) * sz, 2, r);
      flux_r<C3, MIX>(prim_raw(r.u, gam1), 2, r.m, rf, q[l]);
    }
    load_raw<C3, IN32, MIX>(P, own_off + (idx_t)(kbeg + 5) * sz, 2, rq);
  }
  // the own point's state and (c3) i/j metric terms, likewise one plane ahead
  double ru[5] = {0, 0, 0, 0, 0};
  auto load_own = [&](idx_t p) {
    load_u<IN32>(P, p, ru);
    if constexpr (C3) {
#pragma unroll
      for (int c = 0; c < 3; c++) {
        rxm[c] = MIX ? (double)ldg(P.m32[c], p) : ldg(P.m[c], p);
        rym[c] = MIX ? (double)ldg(P.m32[3 + c], p) : ldg(P.m[3 + c], p);
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
      load_raw_halo<C3, IN32, MIX>(P, hs_off[0] + (idx_t)k * sz, hs_dir[0], rh0);
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
        if (IBND && i == 1 && P.uin_row) {
          // supersonic inlet: row 1's conserved variables are the inflow
          // state, its primitives those computed before the inflow state was
          // imposed (bc_supersonic_inlet runs after `primitives`)
          const idx_t sl = (idx_t)(j - 1) + (idx_t)g.ny * (idx_t)(k - 1), ss = (idx_t)g.ny * (idx_t)g.nz;
          double uu[5];
#pragma unroll
          for (int v = 0; v < 5; v++) uu[v] = P.prim_row ? ldg(P.prim_row, (idx_t)v * ss + sl) : ru[v];
          Prim pi = prim_raw(uu, gam1);
          pi.r = ldg(P.uin_row, sl);
          pi.ru = ldg(P.uin_row, ss + sl);
          pi.rv = ldg(P.uin_row, 2 * ss + sl);
          pi.rw = ldg(P.uin_row, 3 * ss + sl);
          pi.re = ldg(P.uin_row, 4 * ss + sl);
          flux_r<CURV, MIX>(pi, 0, rxm, rf, F);
        } else {
          flux_r<CURV, MIX>(pq, 0, rxm, rf, F);
        }
#pragma unroll
        for (int v = 0; v < 5; v++)
          sfx[v][ty][tx + 5] = ADV ? pencil_value(acode >> 30, F[v]) : ((IBND && own_beyond_x) ? R(0) : F[v]);
      }
      if (needy) {
        flux_r<CURV, MIX>(pq, 1, rym, rf, F);
#pragma unroll
        for (int v = 0; v < 5; v++)
          sfy[v][ty + 5][tx] = ADV ? pencil_value((acode >> 28) & 3u, F[v]) : ((WALL && own_beyond) ? R(0) : F[v]);
      }
    }
    if (hs_on[0]) {
      R F[5];
      flux_halo<CURV, MIX>(prim_raw(rh0.u, gam1), hs_dir[0], rh0.m, rf, F);
      if (hs_dir[0] == 0) {
#pragma unroll
        for (int v = 0; v < 5; v++)
          sfx[v][hs_r[0]][hs_c[0]] = ADV ? pencil_value(acode & 3u, F[v]) : ((WALL && hs_beyond[0]) ? R(0) : F[v]);
      } else {
#pragma unroll
        for (int v = 0; v < 5; v++)
          sfy[v][hs_r[0]][hs_c[0]] = ADV ? pencil_value(acode & 3u, F[v]) : ((WALL && hs_beyond[0]) ? R(0) : F[v]);
      }
    }
    __builtin_amdgcn_sched_barrier(0);
    // ---- batch 2: the remaining halo cells ----
    if constexpr (NS > 1) {
      Raw rh[NS];
#pragma unroll
      for (int s = 1; s < NS; s++)
        if (hs_on[s]) {
          load_raw_halo<C3, IN32, MIX>(P, hs_off[s] + (idx_t)k * sz, hs_dir[s], rh[s]);
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
            for (int v = 0; v < 5; v++)
              sfx[v][hs_r[s]][hs_c[s]] =
                  ADV ? pencil_value((acode >> (2 * s)) & 3u, F[v]) : ((WALL && hs_beyond[s]) ? R(0) : F[v]);
          } else {
#pragma unroll
            for (int v = 0; v < 5; v++)
              sfy[v][hs_r[s]][hs_c[s]] =
                  ADV ? pencil_value((acode >> (2 * s)) & 3u, F[v]) : ((WALL && hs_beyond[s]) ? R(0) : F[v]);
          }
        }
      }
    }
    __syncthreads();
    // queue point and own point of the next plane
    if (own && k < kend) load_raw<C3, IN32, MIX>(P, pk + 6 * sz, 2, rq);
    if ((needx || needy) && k < kend) load_own(pk + sz);

    if (own) {
      double K[5];
#pragma unroll
      for (int v = 0; v < 5; v++) {
        R dx;
        if constexpr (IBND) {
          const R *f = &sfx[v][ty][tx + 5];
          dx = R(ci[0]) * (f[1] - f[-1]) + R(ci[1]) * (f[2] - f[-2]) + R(ci[2]) * (f[3] - f[-3]) +
               R(ci[3]) * (f[4] - f[-4]) + R(ci[4]) * (f[5] - f[-5]);
        } else {
          dx = d11c(&sfx[v][ty][tx + 5], 1);
        }
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
      }
      // boundary rows: no Euler increment (flux_euler skips them); a wall row
      // keeps 0 (bc_wall zeroes it) plus the viscous increment, a
      // characteristic or inlet row takes its face slab, set by the boundary
      // routines (no viscous increment there either: ndx_v = 2).  An advanced
      // wall row keeps its increment (its half-width reads 6).
      bool kslab = false;
      if constexpr (WALL) {
        if (hwi == 0 || hw == 0) {
          const bool ia = hwi == 0;
          const double *kb = ia ? (i == 1 ? P.kbc[0] : P.kbc[1]) : (j == 1 ? P.kbc[2] : P.kbc[3]);
          if (kb) {
            const idx_t sl = ia ? (idx_t)(j - 1) + (idx_t)g.ny * (idx_t)(k - 1)
                                : (idx_t)(i - 1) + (idx_t)g.nx * (idx_t)(k - 1);
            const idx_t ss = (ia ? (idx_t)g.ny : (idx_t)g.nx) * (idx_t)g.nz;
#pragma unroll
            for (int v = 0; v < 5; v++) K[v] = ldg(kb, (idx_t)v * ss + sl);
            kslab = true;
          } else {
#pragma unroll
            for (int v = 0; v < 5; v++) K[v] = 0.0;
          }
        }
      }
      if constexpr (VISC) {
        if (!kslab) {
#pragma unroll
          for (int c = 0; c < 4; c++) K[1 + c] = K[1 + c] + kvis[c];
        }
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

// ---------------------------------------------------------------------------
// Pipelined variant.  The same arithmetic as k_stage, but the raw inputs of
// the halo cells of plane k+1 (the state and, c3, the flux-direction metrics)
// are copied from HBM straight into LDS by global_load_lds while plane k is
// computed, so they hold no registers and their latency is covered by a whole
// plane of work; the own point and the queue point of plane k+1 are loaded
// into registers after plane k's flux barrier, and U^n, 1/J and K_v of plane
// k at the start of plane k.  Each plane's results are stored at the start of
// the next plane, after its copies are issued, so that the wait at the top of
// a plane does not wait for the stores.  Barriers are raw s_barrier with an
// LDS-only wait: a __syncthreads() fence would also wait for the copies in
// flight.

// Barrier after every wave's LDS accesses (and, with VM, every memory
// access including the LDS copies) of the workgroup have completed.
__device__ __forceinline__ void barrier_lds() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }
__device__ __forceinline__ void barrier_all() {
  asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

typedef __attribute__((address_space(3))) void lds_void_t;

// An empty use of registers, so that the compiler places its wait for their
// loads here.
__device__ __forceinline__ void touch(double x) { asm volatile("" ::"v"(x)); }
template <bool C3>
__device__ __forceinline__ void touch_ready(const double ru[5], const double rxm[3], const double rym[3],
                                            const Raw &rq) {
#pragma unroll
  for (int v = 0; v < 5; v++) {
    touch(ru[v]);
    touch(rq.u[v]);
  }
  if constexpr (C3) {
#pragma unroll
    for (int c = 0; c < 3; c++) {
      touch(rxm[c]);
      touch(rym[c]);
      touch(rq.m[c]);
    }
  }
}

// Halo cells of a TX x TY tile, numbered j-halo first (10 rows of TX cells:
// 5 below the tile, 5 above), then the i-halo (TY rows of 10 cells: 5 left,
// 5 right).  Cell h: its LDS pencil position, its (i, j), and whether it
// lies inside the arrays' valid region (the points some own point reads).
template <int TX, int TY>
struct HaloMap {
  static constexpr int NHY = 10 * TX, NHX = 10 * TY, NH = NHY + NHX;
  __device__ static __forceinline__ void cell(int h, int i0, int j0, const Geom &g, int &gi, int &gj, bool &ok,
                                              bool &jdir, int &pr, int &pc) {
    if (h < NHY) {
      const int r = h / TX, c = h % TX;
      const int lj = r < 5 ? r : TY + r;
      gi = i0 + c;
      gj = j0 - 5 + lj;
      ok = gi <= g.nx && gj <= g.ny + 5;
      jdir = true;
      pr = lj;
      pc = c;
    } else {
      const int hh = h - NHY;
      const int r = hh / 10, c = hh % 10;
      const int li = c < 5 ? c : TX + c;
      gi = i0 - 5 + li;
      gj = j0 + r;
      ok = gj <= g.ny && gi <= g.nx + 5;
      jdir = false;
      pr = r;
      pc = li;
    }
  }
};

// Halo copies of one array of EW-dword elements into LDS, one dword per lane
// and instruction.  The LDS image is the j-halo (10 rows of TX cells, row by
// row) followed by the i-halo (TY rows of 10 cells).  Each j-halo row takes
// R = ceil(TX EW / 64) instructions whose global source is the row's start
// (wave-uniform) plus the lane's dword: one lane offset per part p < R,
// whatever the row.  The i-halo takes NIX instructions with a per-lane offset
// each.  Offsets are computed once per kernel; a cell outside the valid
// region reads the tile's first own point instead (always inside the
// arrays): its LDS slot is never read.
template <int TX, int TY, int EW>
struct DmaPlan {
  using HM = HaloMap<TX, TY>;
  static constexpr int RW = TX * EW;                  // dwords per j-halo row
  static constexpr int R = (RW + 63) / 64;            // instructions per j-halo row
  static constexpr int NDX = HM::NHX * EW;            // i-halo dwords
  static constexpr int NIX = (NDX + 63) / 64;         // i-halo instructions
};

template <int TX, int TY, int EW>
struct DmaLane {
  uint32_t jl[DmaPlan<TX, TY, EW>::R];      // j-halo: byte offset from the row's start
  uint32_t ix[DmaPlan<TX, TY, EW>::NIX];    // i-halo: byte offset in plane k = 0
};

template <int TX, int TY, int EW>
__device__ __forceinline__ void dma_lane(const Geom &g, int i0, int j0, int lane, DmaLane<TX, TY, EW> &L) {
  using D = DmaPlan<TX, TY, EW>;
  using HM = HaloMap<TX, TY>;
  const idx_t org = lin(g, i0, j0, 0);
#pragma unroll
  for (int p = 0; p < D::R; p++) {
    const int dw = 64 * p + lane;  // dword in the row
    const int c = dw / EW;         // cell in the row
    L.jl[p] = (dw < D::RW && i0 + c <= g.nx) ? (uint32_t)(dw * 4) : 0u;
  }
#pragma unroll
  for (int t = 0; t < D::NIX; t++) {
    const int d = 64 * t + lane;
    const int h = HM::NHY + d / EW, part = d % EW;
    idx_t off = org;
    uint32_t extra = 0;
    if (d < D::NDX) {
      int gi, gj, pr, pc;
      bool ok, jd;
      HM::cell(h, i0, j0, g, gi, gj, ok, jd, pr, pc);
      if (ok) {
        off = lin(g, gi, gj, 0);
        extra = (uint32_t)(part * 4);
      }
    }
    L.ix[t] = (uint32_t)(off * (idx_t)(4 * EW)) + extra;
  }
}

// Issue the copies of plane k: bj / bi are the byte bases of plane k of the
// arrays the j-halo / i-halo read (wave-uniform), dst the LDS image.
template <int TX, int TY, int EW>
__device__ __forceinline__ void dma_issue(const MGPU_GLOBAL char *bj, const MGPU_GLOBAL char *bi, void *dst,
                                          const DmaLane<TX, TY, EW> &L, const Geom &g, int i0, int j0, int lane) {
  using D = DmaPlan<TX, TY, EW>;
  uint32_t *ldst = reinterpret_cast<uint32_t *>(dst);
#pragma unroll
  for (int r = 0; r < 10; r++) {
    const int lj = r < 5 ? r : TY + r;
    const int gj = j0 - 5 + lj;
    if (gj > g.ny + 5) continue;  // beyond the valid region: nothing reads this row
    const MGPU_GLOBAL char *row = bj + (uint32_t)(lin(g, i0, gj, 0) * (idx_t)(4 * EW));
#pragma unroll
    for (int p = 0; p < D::R; p++) {
      lds_void_t *l = (lds_void_t *)(ldst + r * D::RW + 64 * p);
      if (64 * p + 64 <= D::RW || 64 * p + lane < D::RW)
        __builtin_amdgcn_global_load_lds((const MGPU_GLOBAL void *)(row + L.jl[p]), l, 4, 0, 0);
    }
  }
#pragma unroll
  for (int t = 0; t < D::NIX; t++) {
    lds_void_t *l = (lds_void_t *)(ldst + 10 * D::RW + 64 * t);
    if (64 * t + 64 <= D::NDX || 64 * t + lane < D::NDX)
      __builtin_amdgcn_global_load_lds((const MGPU_GLOBAL void *)(bi + L.ix[t]), l, 4, 0, 0);
  }
}

template <int TX, int TY, int GEO, bool VISC, int MODE, bool WALL>
__global__ __launch_bounds__(TX *TY, 1) void k_stage_p(const Batch<StageArgs> *__restrict__ tab, int nb,
                                                       int ntiles) {
  constexpr bool CURV = GEO != kGeoCart;
  constexpr bool C3 = GEO == kGeoC3;
  constexpr bool C2 = GEO == kGeoC2;
  static_assert(!WALL || !CURV, "walls: Cartesian blocks");
  constexpr bool MIX = MODE != kFP64;
  constexpr bool IN32 = MODE == kMixMiddle || MODE == kMixLast;
  constexpr bool OUT32 = MODE == kMixFirst || MODE == kMixMiddle;
  using R = typename std::conditional<MIX, float, double>::type;
  using TU = typename std::conditional<IN32, float, double>::type;  // stored stage-input element
  using HM = HaloMap<TX, TY>;
  constexpr int NT = TX * TY;
  constexpr int NW = NT / 64;
  constexpr int NH = HM::NH;
  constexpr int NS = (NH + NT - 1) / NT;  // halo cells per thread
  constexpr int NM = C3 ? 3 : 0;          // metric arrays per halo cell
  static_assert(NT % 64 == 0, "whole waves");
  __shared__ R sfx[5][TY][TX + 10];
  __shared__ R sfy[5][TY + 10][TX];
  __shared__ TU hu[2][5][NH];                          // halo state, two planes
  __shared__ double hmt[2][NM > 0 ? NM : 1][NM > 0 ? NH : 1];  // halo metrics (c3), two planes

  const int w = xcd_remap(blockIdx.x, ntiles);
  const Batch<StageArgs> &e = batch_find(tab, nb, w);
  const StageArgs &a = e.a;
  Ptrs P;
  load_ptrs<GEO, VISC, IN32, OUT32>(a, P);
  const Geom &g = a.g;
  int tile_x, tile_y, chunk;
  tile_coords(e.t, w, tile_x, tile_y, chunk);
  const int kchunk = e.t.kchunk;

  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX, tid = threadIdx.x;
  const int wv = tid / 64, lane = tid % 64;
  const int i0 = tile_x * TX + 1, j0 = tile_y * TY + 1;
  const int i = i0 + tx, j = j0 + ty;
  const int kbeg = chunk * kchunk + 1;
  const int kend = min(kbeg + kchunk - 1, g.nz);
  const bool own = (i <= g.nx) && (j <= g.ny);
  const bool needx = (i <= g.nx + 5) && (j <= g.ny);
  const bool needy = (i <= g.nx) && (j <= g.ny + 5);
  const double gam1 = a.gas.gam1;
  const idx_t sz = g.sy;

  // the thread's halo cells: index, pencil position, direction, validity
  int hs_h[NS], hs_r[NS], hs_c[NS];
  bool hs_on[NS], hs_jd[NS], hs_beyond[NS];
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const int h = tid + s * NT;
    hs_h[s] = h;
    hs_on[s] = false;
    hs_jd[s] = false;
    hs_r[s] = hs_c[s] = 0;
    hs_beyond[s] = false;
    if (h < NH) {
      int gi, gj;
      HM::cell(h, i0, j0, g, gi, gj, hs_on[s], hs_jd[s], hs_r[s], hs_c[s]);
      if constexpr (WALL)
        hs_beyond[s] = hs_jd[s] && (((g.jbnd & 1) && gj < 1) || ((g.jbnd & 2) && gj > g.ny));
    }
  }
  const idx_t own_off = lin(g, i, j, 0);
  // 2.5-D: metric rows of the halo cells and the own column, 1/J (constant in k)
  double hm[NS][2];
  double rxm[3] = {0, 0, 0}, rym[3] = {0, 0, 0};
  double ijac = 0.0;
  if constexpr (C2) {
#pragma unroll
    for (int s = 0; s < NS; s++) {
      hm[s][0] = hm[s][1] = 0.0;
      if (hs_on[s]) {
        int gi, gj, pr, pc;
        bool ok, jd;
        HM::cell(hs_h[s], i0, j0, g, gi, gj, ok, jd, pr, pc);
        const idx_t p2 = lin2(g, gi, gj);
        hm[s][0] = ldg(jd ? P.m2[2] : P.m2[0], p2);
        hm[s][1] = ldg(jd ? P.m2[3] : P.m2[1], p2);
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
  bool own_beyond = false;
  double cw[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  int hw = 5;
  if constexpr (WALL) {
    own_beyond = (g.jbnd & 2) && j > g.ny;
    hw = jhalf(g, j);
    wall_coeffs(hw, cw);
  }

  // this wave's share of the halo copies of plane k into LDS buffer b:
  // wave q copies array q (the 5 states, then the NM metric components; c3:
  // the i-halo reads the xi metrics, the j-halo the eta metrics)
  static_assert(NW >= 5 + NM, "one array per wave");
  constexpr int EWU = IN32 ? 1 : 2;
  DmaLane<TX, TY, EWU> dlu;
  DmaLane<TX, TY, 2> dlm;
  const MGPU_GLOBAL char *dbj = nullptr, *dbi = nullptr;  // plane-0 bases of this wave's arrays
  const bool dstate = wv < 5, dmet = C3 && wv >= 5 && wv < 5 + NM;
  if (dstate) {
    const TU *src = nullptr;
#pragma unroll
    for (int q = 0; q < 5; q++)
      if (q == wv) {
        if constexpr (IN32) src = P.uin32[q];
        else src = P.uin[q];
      }
    dbj = dbi = reinterpret_cast<const MGPU_GLOBAL char *>((const MGPU_GLOBAL TU *)src);
    dma_lane<TX, TY, EWU>(g, i0, j0, lane, dlu);
  } else if (dmet) {
    if constexpr (C3) {
#pragma unroll
      for (int c = 0; c < NM; c++)
        if (c == wv - 5) {
          dbi = reinterpret_cast<const MGPU_GLOBAL char *>((const MGPU_GLOBAL double *)P.m[c]);
          dbj = reinterpret_cast<const MGPU_GLOBAL char *>((const MGPU_GLOBAL double *)P.m[3 + c]);
        }
      dma_lane<TX, TY, 2>(g, i0, j0, lane, dlm);
    }
  }
  auto dma = [&](int k, int b) {
    if (dstate) {
      const uint32_t kb = (uint32_t)((idx_t)k * sz * (idx_t)sizeof(TU));
#pragma unroll
      for (int q = 0; q < 5; q++)
        if (q == wv) dma_issue<TX, TY, EWU>(dbj + kb, dbi + kb, &hu[b][q][0], dlu, g, i0, j0, lane);
    } else if (dmet) {
      if constexpr (C3) {
        const uint32_t kb = (uint32_t)((idx_t)k * sz * (idx_t)sizeof(double));
#pragma unroll
        for (int c = 0; c < NM; c++)
          if (c == wv - 5) dma_issue<TX, TY, 2>(dbj + kb, dbi + kb, &hmt[b][c][0], dlm, g, i0, j0, lane);
      }
    }
  };

  // the mixed path's reference state, read once (not in every plane)
  const RefState rf = a.ref;
  const double alpha = a.alpha;
  R q[11][5];
  Raw rq;
  if (own) {
#pragma unroll
    for (int l = 0; l < 10; l++) {
      Raw r;
      load_raw<C3, IN32, MIX>(P, own_off + (idx_t)(kbeg - 5 + l) * sz, 2, r);
      flux_r<C3, MIX>(prim_raw(r.u, gam1), 2, r.m, rf, q[l]);
    }
    load_raw<C3, IN32, MIX>(P, own_off + (idx_t)(kbeg + 5) * sz, 2, rq);
  }
  double ru[5] = {0, 0, 0, 0, 0};
  auto load_own = [&](idx_t p) {
    load_u<IN32>(P, p, ru);
    if constexpr (C3) {
#pragma unroll
      for (int c = 0; c < 3; c++) {
        rxm[c] = MIX ? (double)ldg(P.m32[c], p) : ldg(P.m[c], p);
        rym[c] = MIX ? (double)ldg(P.m32[3 + c], p) : ldg(P.m[3 + c], p);
      }
    }
  };
  if (needx || needy) load_own(own_off + (idx_t)kbeg * sz);
  // U^n, 1/J (c3) and K_v of the own point, one plane ahead like the rest
  // (Cartesian and 2.5-D: idz of the plane; the own column's idx and idy
  // are constant along the march)
  double u0n[5] = {0, 0, 0, 0, 0}, kvisn[4] = {0, 0, 0, 0}, ijacn = 0.0, idzn = 0.0;
  double idxi = 0.0, idyj = 0.0;
  if constexpr (GEO == kGeoCart) {
    if (own) {
      idxi = ldg(P.idx, (idx_t)(i - 1 + G));
      idyj = ldg(P.idy, (idx_t)(j - 1 + G));
    }
  }
  auto load_u0 = [&](idx_t p, int kk) {
#pragma unroll
    for (int v = 0; v < 5; v++) u0n[v] = ldg_nt(P.u0[v], p);
    if constexpr (C3) ijacn = ldg_nt(P.m[9], p);
    else idzn = ldg(P.idz, (idx_t)(kk - 1 + G));
    if constexpr (VISC) {
#pragma unroll
      for (int c = 0; c < 4; c++) kvisn[c] = ldg_nt(P.kv[c], p);
    }
  };
  if (own) load_u0(own_off + (idx_t)kbeg * sz, kbeg);
  dma(kbeg, 0);

  // the previous plane's result, stored at the start of the next plane
  double outv[5] = {0, 0, 0, 0, 0};
  idx_t out_p = 0;
  bool out_pending = false;
  auto store_out = [&]() {
    if constexpr (OUT32) {
#pragma unroll
      for (int v = 0; v < 5; v++) stg_nt(P.uout32[v], out_p, (float)outv[v]);
    } else {
#pragma unroll
      for (int v = 0; v < 5; v++) stg_nt(P.uout[v], out_p, outv[v]);
    }
  };

  for (int k = kbeg; k <= kend; k++) {
    const int b = (k - kbeg) & 1;
    const idx_t pk = own_off + (idx_t)k * sz;
    // plane k's halo copies and own-point loads have landed in every wave;
    // the pencils are free (every wave is past plane k-1's derivatives)
    barrier_all();
    // the own and queue point's registers loaded during plane k-1 are ready
    // (waited for above); say so to the compiler here, before the copies are
    // issued -- otherwise its wait before their first use also waits for the
    // copies just issued
    touch_ready<C3>(ru, rxm, rym, rq);
    double u0[5], kvis[4];
#pragma unroll
    for (int v = 0; v < 5; v++) {
      touch(u0n[v]);
      u0[v] = u0n[v];
    }
    double idzk = 0.0;
    if constexpr (C3) {
      touch(ijacn);
      ijac = ijacn;
    } else {
      touch(idzn);
      idzk = idzn;
    }
    if constexpr (GEO == kGeoCart) {
      touch(idxi);
      touch(idyj);
    }
#pragma unroll
    for (int c = 0; c < 4; c++) {
      if constexpr (VISC) touch(kvisn[c]);
      kvis[c] = kvisn[c];
    }
    if (k < kend) dma(k + 1, b ^ 1);
    if (out_pending) store_out();
    out_pending = false;
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
#pragma unroll
    for (int s = 0; s < NS; s++) {
      if (hs_on[s]) {
        const int h = hs_h[s];
        Raw r;
        if constexpr (IN32) {
#pragma unroll
          for (int v = 0; v < 5; v++) r.u[v] = P.ref[v] + (double)hu[b][v][h];
        } else {
#pragma unroll
          for (int v = 0; v < 5; v++) r.u[v] = hu[b][v][h];
        }
        if constexpr (C3) {
#pragma unroll
          for (int c = 0; c < 3; c++) r.m[c] = hmt[b][c < NM ? c : 0][h];
        } else if constexpr (C2) {
          r.m[0] = hm[s][0];
          r.m[1] = hm[s][1];
          r.m[2] = 0.0;
        }
        R F[5];
        flux_halo<CURV, MIX>(prim_raw(r.u, gam1), hs_jd[s] ? 1 : 0, r.m, rf, F);
        if (!hs_jd[s]) {
#pragma unroll
          for (int v = 0; v < 5; v++) sfx[v][hs_r[s]][hs_c[s]] = F[v];
        } else {
#pragma unroll
          for (int v = 0; v < 5; v++) sfy[v][hs_r[s]][hs_c[s]] = (WALL && hs_beyond[s]) ? R(0) : F[v];
        }
      }
    }
    barrier_lds();
    // own point and queue point of plane k+1, in flight during the derivatives
    if (own && k < kend) {
      load_raw<C3, IN32, MIX>(P, pk + 6 * sz, 2, rq);
      load_u0(pk + sz, k + 1);
    }
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
        if constexpr (C3) {
          K[v] = (double)(dz + (dy + dx)) * ijac;
        } else if constexpr (C2) {
          K[v] = (double)dz * idzk + (double)(dy + dx) * ijac;
        } else {
          const double kx = (double)dx * idxi;
          const double ky = (double)dy * idyj + kx;
          K[v] = (double)dz * idzk + ky;
        }
        if (WALL && hw == 0) K[v] = 0.0;
      }
      if constexpr (VISC) {
#pragma unroll
        for (int c = 0; c < 4; c++) K[1 + c] = K[1 + c] + kvis[c];
      }
      if constexpr (OUT32) {
        const double ref[5] = {rf.r, rf.ru, rf.rv, rf.rw, rf.re};
#pragma unroll
        for (int v = 0; v < 5; v++) outv[v] = (u0[v] - alpha * K[v]) - ref[v];
      } else {
#pragma unroll
        for (int v = 0; v < 5; v++) outv[v] = u0[v] - alpha * K[v];
      }
      out_p = pk;
      out_pending = true;
    }

#pragma unroll
    for (int l = 0; l < 10; l++)
#pragma unroll
      for (int v = 0; v < 5; v++) q[l][v] = q[l + 1][v];
  }
  if (out_pending) store_out();
}

template <int TX, int TY, int GEO, int MODE>
hipError_t launch_gp(const Batch<StageArgs> *tab, int nb, int nt, bool visc, hipStream_t s) {
  dim3 blk(TX * TY), grd(nt);
  if constexpr (MODE == kFP64 || MODE == kMixLast) {
    if (visc) {
      hipLaunchKernelGGL((k_stage_p<TX, TY, GEO, true, MODE, false>), grd, blk, 0, s, tab, nb, nt);
      return hipGetLastError();
    }
  }
  hipLaunchKernelGGL((k_stage_p<TX, TY, GEO, false, MODE, false>), grd, blk, 0, s, tab, nb, nt);
  return hipGetLastError();
}

template <int TX, int TY, int MODE>
hipError_t launch_tp(const Batch<StageArgs> *tab, int nb, int nt, int geo, bool visc, bool wall, hipStream_t s) {
  dim3 blk(TX * TY), grd(nt);
  constexpr bool VOK = MODE == kFP64 || MODE == kMixLast;
  if (visc && !VOK) return hipErrorInvalidValue;
  if (wall) {
    if (geo != kGeoCart) return hipErrorInvalidValue;
    if constexpr (VOK) {
      if (visc) {
        hipLaunchKernelGGL((k_stage_p<TX, TY, kGeoCart, true, MODE, true>), grd, blk, 0, s, tab, nb, nt);
        return hipGetLastError();
      }
    }
    hipLaunchKernelGGL((k_stage_p<TX, TY, kGeoCart, false, MODE, true>), grd, blk, 0, s, tab, nb, nt);
    return hipGetLastError();
  }
  if (geo == kGeoC3) return launch_gp<TX, TY, kGeoC3, MODE>(tab, nb, nt, visc, s);
  if (geo == kGeoC2) return launch_gp<TX, TY, kGeoC2, MODE>(tab, nb, nt, visc, s);
  return launch_gp<TX, TY, kGeoCart, MODE>(tab, nb, nt, visc, s);
}

template <int TX, int TY, int GEO, int MODE>
hipError_t launch_g(const Batch<StageArgs> *tab, int nb, int nt, bool visc, hipStream_t s) {
  dim3 blk(TX * TY), grd(nt);
  if constexpr (MODE == kFP64 || MODE == kMixLast) {
    if (visc) {
      hipLaunchKernelGGL((k_stage<TX, TY, GEO, true, MODE, 0>), grd, blk, 0, s, tab, nb, nt);
      return hipGetLastError();
    }
  }
  hipLaunchKernelGGL((k_stage<TX, TY, GEO, false, MODE, 0>), grd, blk, 0, s, tab, nb, nt);
  return hipGetLastError();
}

template <int TX, int TY, int MODE, int W>
hipError_t launch_w(const Batch<StageArgs> *tab, int nb, int nt, int geo, bool visc, hipStream_t s) {
  dim3 blk(TX * TY), grd(nt);
  constexpr bool VOK = MODE == kFP64 || MODE == kMixLast;
  if (geo == kGeoC2) return hipErrorInvalidValue;
  if (geo == kGeoC3) {
    if constexpr (VOK) {
      if (visc) {
        hipLaunchKernelGGL((k_stage<TX, TY, kGeoC3, true, MODE, W>), grd, blk, 0, s, tab, nb, nt);
        return hipGetLastError();
      }
    }
    hipLaunchKernelGGL((k_stage<TX, TY, kGeoC3, false, MODE, W>), grd, blk, 0, s, tab, nb, nt);
    return hipGetLastError();
  }
  if constexpr (VOK) {
    if (visc) {
      hipLaunchKernelGGL((k_stage<TX, TY, kGeoCart, true, MODE, W>), grd, blk, 0, s, tab, nb, nt);
      return hipGetLastError();
    }
  }
  hipLaunchKernelGGL((k_stage<TX, TY, kGeoCart, false, MODE, W>), grd, blk, 0, s, tab, nb, nt);
  return hipGetLastError();
}

// wall: 0 no bounded face, 1 bounded faces, 2 also advanced wall rows
template <int TX, int TY, int MODE>
hipError_t launch_t(const Batch<StageArgs> *tab, int nb, int nt, int geo, bool visc, int wall, hipStream_t s) {
  // the viscous increment is added in the last stage only
  constexpr bool VOK = MODE == kFP64 || MODE == kMixLast;
  if (visc && !VOK) return hipErrorInvalidValue;
  if (wall == 2) return launch_w<TX, TY, MODE, 2>(tab, nb, nt, geo, visc, s);
  if (wall) return launch_w<TX, TY, MODE, 1>(tab, nb, nt, geo, visc, s);
  if (geo == kGeoC3) return launch_g<TX, TY, kGeoC3, MODE>(tab, nb, nt, visc, s);
  if (geo == kGeoC2) return launch_g<TX, TY, kGeoC2, MODE>(tab, nb, nt, visc, s);
  return launch_g<TX, TY, kGeoCart, MODE>(tab, nb, nt, visc, s);
}

template <int MODE>
hipError_t launch_shape(const Batch<StageArgs> *tab, int nb, int nt, int geo, bool visc, int shape, int wall,
                        hipStream_t s) {
  switch (shape) {
    case 1: return launch_t<32, 8, MODE>(tab, nb, nt, geo, visc, wall, s);
    case 2: return launch_t<64, 8, MODE>(tab, nb, nt, geo, visc, wall, s);
    case 3: return launch_t<16, 16, MODE>(tab, nb, nt, geo, visc, wall, s);
    case 4: return launch_t<32, 16, MODE>(tab, nb, nt, geo, visc, wall, s);
    case 5: return wall == 2 ? hipErrorInvalidValue : launch_tp<64, 8, MODE>(tab, nb, nt, geo, visc, wall != 0, s);
    case 6: return wall == 2 ? hipErrorInvalidValue : launch_tp<32, 16, MODE>(tab, nb, nt, geo, visc, wall != 0, s);
    default: return launch_t<64, 4, MODE>(tab, nb, nt, geo, visc, wall, s);
  }
}

}  // namespace

// Tile shapes the autotuner may choose from.
// Shapes 5 and 6 are the pipelined kernel (halo copies straight into LDS).
const int kStageShapes[][2] = {{64, 4}, {32, 8}, {64, 8}, {16, 16}, {32, 16}, {64, 8}, {32, 16}};
const int kNumStageShapes = 7;
const int kFirstPipelinedShape = 5;

TileRange stage_tiles(const Geom &g, int shape, int kchunk) {
  const int TX = kStageShapes[shape][0], TY = kStageShapes[shape][1];
  TileRange t{};
  t.tiles_x = (g.nx + TX - 1) / TX;
  t.tiles_y = (g.ny + TY - 1) / TY;
  t.kchunk = kchunk < 1 ? 1 : (kchunk > g.nz ? g.nz : kchunk);
  return t;
}

hipError_t launch_stage(const Batch<StageArgs> *tab, int nb, int nt, int geo, bool visc, int shape, int mode,
                        int wall, hipStream_t s) {
  if (nt <= 0) return hipSuccess;
  switch (mode) {
    case kMixFirst: return launch_shape<kMixFirst>(tab, nb, nt, geo, visc, shape, wall, s);
    case kMixMiddle: return launch_shape<kMixMiddle>(tab, nb, nt, geo, visc, shape, wall, s);
    case kMixLast: return launch_shape<kMixLast>(tab, nb, nt, geo, visc, shape, wall, s);
    default: return launch_shape<kFP64>(tab, nb, nt, geo, visc, shape, wall, s);
  }
}

}  // namespace mgpu
