/*
 * musica_oracle.c -- CPU restatement of MUSICAA's explicit RK time step.
 *
 * TEST INFRASTRUCTURE ONLY (what the tests compare the HIP product path
 * with; see the header).  Plain C99 + OpenMP over k-planes, FP64 throughout, no FMA
 * contraction (built with -ffp-contract=off), evaluation order of every sum
 * as in the reference's Fortran so that results track the reference to
 * round-off.
 *
 * Reference files followed (LouenasZm/musica2_newGrid):
 *   runge_kutta.f90:2-152        step choreography            orc_step
 *   update.f90:32-44, 538-565    update_var / update_varn      orc_step
 *   mod_time.f90:165-190         crk for nrk=4                 (caller)
 *   smod_primitive_pfg.f90:12-136 primitives / primitives_visc prim_*
 *   smod_flux_euler_11pts.f90    Cartesian Euler divergence    euler_cart
 *   smod_flux_euler_11pts_c3.f90 curvilinear Euler divergence  euler_c3
 *   mod_coeff_deriv.f90:108-325  a3,a5,a7,a9,a11 coefficients  A11,A5
 *   smod_grad_vel*.f90, smod_grad_T*.f90   5-pt gradients      grad_*
 *   smod_flux_visc_5pts*.f90     viscous flux divergence       visc_*
 *   mod_filtering.f90:18-205,894-1492  selective filter        filter_sf
 *   mod_filtering_shock.f90:22-1512    shock filter            filter_shock
 *   mod_artvisc.f90:16-79, 82-874      artificial viscosity    artvisc
 *   update.f90:568-608            update_varn_artvisc
 *   mod_comm.f90:249-290         halo exchange (periodic self) orc_fill_periodic
 *   mpi_index_bounds.f90         loop bounds for periodic faces and walls
 *   smod_bc_wall.f90:11-98       no-slip wall, dp/dn = 0         bc_wall
 *   smod_bc_wall_2*.f90          advanced walls (is_wall2)       bc_wall2_face
 *   mod_coeff_deriv.f90:100-160, 1099-1160  closure and extrapolation coefficients
 */
#include "musica_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NG ORC_NG

/* Planted mistakes, for the test that shows the pins fail on them
 * (tests/test_planted_mistakes.py builds copies with -DORC_PLANT=n into
 * _build/plant_n.so).  0 -- the default -- in every other build:
 *   1  pressure sensor: 4 p_i instead of 2 p_i in the cs-weighted denominator
 *   2  interior face amplitude: xk0/10.5 instead of xk0/105
 *   3  Ducros sensor: omega_z missing from the denominator
 *   4  bc_charac_imax: the textbook sign of the density equation
 *   5  bc_charac_imax: 0.5 for the 0.528 of the subsonic relaxation
 *   6  bc_charac_jmax: v for v + c in the outgoing wave
 *   7  bc_charac_imax: the a04 rows for the a40 ones
 *   8  c3 wall state: the tangential term with the opposite sign
 *   9  c3 adiabatic wall gradient: the two tangents exchanged
 *   10 advanced wall row at the max face: a02 for a20 (no sign change)
 *   11 c3 slip wall: the unit tangents for their dual basis */
#ifndef ORC_PLANT
#define ORC_PLANT 0
#endif

typedef struct {
  int nx, ny, nz;
  long sx, sy, sz;
  size_t n;
} dims_t;

static dims_t mkdims(const orc_cfg *c) {
  dims_t d;
  d.nx = c->nx; d.ny = c->ny; d.nz = c->nz;
  d.sx = c->nx + 2 * NG; d.sy = c->ny + 2 * NG; d.sz = c->nz + 2 * NG;
  d.n = (size_t)d.sx * d.sy * d.sz;
  return d;
}

/* Fortran index (i,j,k), interior 1..n, to linear offset. */
#define ID(d, i, j, k) \
  ((size_t)((i) + NG - 1) + (size_t)(d).sx * ((size_t)((j) + NG - 1) + (size_t)(d).sy * (size_t)((k) + NG - 1)))

/* 11-pt centred 10th-order first derivative, mod_coeff_deriv.f90:316-325. */
static const double A11[6] = {0.0, 5.0 / 6.0, -5.0 / 21.0, 5.0 / 84.0, -5.0 / 504.0, 1.0 / 1260.0};
/* 5-pt centred 4th-order first derivative, mod_coeff_deriv.f90:121-126. */
static const double A5[3] = {0.0, 2.0 / 3.0, -1.0 / 12.0};
/* 11-pt standard selective filter, mod_filtering.f90:63-68 (before -chi scaling). */
static const double D11[6] = {252.0 / 1024.0, -210.0 / 1024.0, 120.0 / 1024.0,
                              -45.0 / 1024.0, 10.0 / 1024.0, -1.0 / 1024.0};
/* 9th-order conservative dissipation, mod_filtering_shock.f90:95-100. */
/* Centred closures next to a wall by half-width q = 1..4: a3, a5, a7, a9
 * (mod_coeff_deriv.f90:102-104, 109-113, 125-131, 232-241; stencil 11 keeps
 * the standard a7/a9). */
static const double ACL[5][5] = {{0.0, 0.0, 0.0, 0.0, 0.0},
                                 {0.0, 0.5, 0.0, 0.0, 0.0},
                                 {0.0, 2.0 / 3.0, -1.0 / 12.0, 0.0, 0.0},
                                 {0.0, 3.0 / 4.0, -3.0 / 20.0, 1.0 / 60.0, 0.0},
                                 {0.0, 4.0 / 5.0, -1.0 / 5.0, 4.0 / 105.0, -1.0 / 280.0}};
/* 3-pt one-sided first derivative of an advanced wall row (is_wall2),
 * mod_coeff_deriv.f90:113-117; a20 = -a02 at the max face. */
static const double A02[3] = {-3.0 / 2.0, 2.0, -1.0 / 2.0};
/* 5-pt one-sided first derivatives at the first two rows (mod_coeff_deriv.f90:114-124);
 * a31 = -a13, a40 = -a04 at the last two. */
static const double A04[6] = {0.0, -25.0 / 12.0, 4.0, -3.0, 4.0 / 3.0, -1.0 / 4.0};
static const double A13[6] = {0.0, -1.0 / 4.0, -5.0 / 6.0, 3.0 / 2.0, -1.0 / 2.0, 1.0 / 12.0};
/* Standard filters on 5, 7, 9 points used on rows 3, 4, 5 from a wall
 * (mod_filtering.f90:24-38), before the -chi scaling. */
static const double DCL[6][6] = {{0.0}, {0.0},
                                 {6.0 / 16.0, -4.0 / 16.0, 1.0 / 16.0, 0.0, 0.0, 0.0},
                                 {20.0 / 64.0, -15.0 / 64.0, 6.0 / 64.0, -1.0 / 64.0, 0.0, 0.0},
                                 {70.0 / 256.0, -56.0 / 256.0, 28.0 / 256.0, -8.0 / 256.0, 1.0 / 256.0, 0.0},
                                 {252.0 / 1024.0, -210.0 / 1024.0, 120.0 / 1024.0,
                                  -45.0 / 1024.0, 10.0 / 1024.0, -1.0 / 1024.0}};

/* Wall faces of a block: lo[a] / hi[a] for the min / max face of axis a. */
typedef struct {
  int lo[3], hi[3], any;
} walls_t;

static walls_t walls_of(const orc_cfg *c) {
  walls_t W;
  W.any = 0;
  for (int a = 0; a < 3; a++) {
    W.lo[a] = c->face[a][0] == ORC_FACE_WALL || c->face[a][0] == ORC_FACE_SLIP;
    W.hi[a] = c->face[a][1] == ORC_FACE_WALL || c->face[a][1] == ORC_FACE_SLIP;
    W.any |= W.lo[a] | W.hi[a];
  }
  return W;
}

/* Boundary faces of any kind (no-slip wall, characteristic outflow,
 * supersonic inlet): the operators close them alike (is_boundary / is_bc_1pt
 * of mpi_index_bounds.f90:20-70: ndx_e = 2 for sort 0 and sort <= -3; the
 * one-sided gradients; the filters' closures, mod_filtering.f90:900-988). */
static walls_t bounds_of(const orc_cfg *c) {
  walls_t W;
  W.any = 0;
  for (int a = 0; a < 3; a++) {
    W.lo[a] = c->face[a][0] >= ORC_FACE_WALL;
    W.hi[a] = c->face[a][1] >= ORC_FACE_WALL;
    W.any |= W.lo[a] | W.hi[a];
  }
  return W;
}

/* Characteristic faces (sort -3: outflow or supersonic inlet): their row
 * gets no Euler and no viscous increment at all (ndx_e = ndx_v = 2); the
 * boundary routines set it. */
static walls_t chars_of(const orc_cfg *c) {
  walls_t W;
  W.any = 0;
  for (int a = 0; a < 3; a++) {
    W.lo[a] = c->face[a][0] == ORC_FACE_CHARAC || c->face[a][0] == ORC_FACE_INLET;
    W.hi[a] = c->face[a][1] == ORC_FACE_CHARAC || c->face[a][1] == ORC_FACE_INLET;
    W.any |= W.lo[a] | W.hi[a];
  }
  return W;
}

/* Slip wall state on face f (0..5): the bindings of assign_procedures.f90:
 * 415-470 -- with is_wall2, every wall of an inviscid case and the walls with
 * the slip flag of an adiabatic one; isothermal cases bind no slip routine,
 * and without is_wall2 the flag selects nothing on this path. */
static int slip_face(const orc_cfg *c, int f) {
  const int k = c->face[f / 2][f % 2];
  if (!c->wall2 || (k != ORC_FACE_WALL && k != ORC_FACE_SLIP)) return 0;
  if (!c->visc) return 1;
  return c->adiabatic && k == ORC_FACE_SLIP;
}

/* Half-width of the centred stencil at row r (1..n) of an axis: 5 inside,
 * r-1 (n-r) near a min (max) wall, 0 on the wall row itself. */
static inline int halfw(int r, int n, int lo, int hi) {
  int q = 5;
  if (lo && r - 1 < q) q = r - 1;
  if (hi && n - r < q) q = n - r;
  return q;
}

/* 5-pt first derivative at row r with the one-sided rows of a wall axis
 * (smod_grad_vel.f90:12-100: a04 at 1, a13 at 2, a31 at n-1, a40 at n). */
static inline double g5w(const double *f, long p, long s, int r, int n, int lo, int hi) {
  if (lo && r == 1)
    return A04[1] * f[p] + A04[2] * f[p + s] + A04[3] * f[p + 2 * s] + A04[4] * f[p + 3 * s] +
           A04[5] * f[p + 4 * s];
  if (lo && r == 2)
    return A13[1] * f[p - s] + A13[2] * f[p] + A13[3] * f[p + s] + A13[4] * f[p + 2 * s] +
           A13[5] * f[p + 3 * s];
  if (hi && r == n - 1)
    return (-A13[5]) * f[p - 3 * s] + (-A13[4]) * f[p - 2 * s] + (-A13[3]) * f[p - s] + (-A13[2]) * f[p] +
           (-A13[1]) * f[p + s];
  if (hi && r == n)
    return (-A04[5]) * f[p - 4 * s] + (-A04[4]) * f[p - 3 * s] + (-A04[3]) * f[p - 2 * s] +
           (-A04[2]) * f[p - s] + (-A04[1]) * f[p];
  return A5[1] * (f[p + s] - f[p - s]) + A5[2] * (f[p + 2 * s] - f[p - 2 * s]);
}

int orc_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_num_threads(int n) {
#ifdef _OPENMP
  if (n > 0) omp_set_num_threads(n);
#else
  (void)n;
#endif
}

static double *alloc_arr(const dims_t *d) {
  double *p = (double *)calloc(d->n, sizeof(double));
  return p;
}

/* ------------------------------------------------------------------------ */
/* Halo: periodic self-neighbour faces (mod_comm.f90:249-290).  The x pass   */
/* covers interior j,k; the y pass all i; the z pass all i,j -- so edge and  */
/* corner ghosts also hold the periodic image (the cross stencils never read */
/* them, but primitives are evaluated on the whole padded box).              */
/* ------------------------------------------------------------------------ */
void orc_fill_periodic(const orc_cfg *c, double *a) {
  dims_t d = mkdims(c);
  if (c->face[0][0] == ORC_FACE_PERIODIC) {
#pragma omp parallel for collapse(2)
    for (int k = 1; k <= d.nz; k++)
      for (int j = 1; j <= d.ny; j++)
        for (int l = 1; l <= NG; l++) {
          a[ID(d, 1 - l, j, k)] = a[ID(d, d.nx + 1 - l, j, k)];
          a[ID(d, d.nx + l, j, k)] = a[ID(d, l, j, k)];
        }
  }
  if (c->face[1][0] == ORC_FACE_PERIODIC) {
#pragma omp parallel for
    for (int k = 1; k <= d.nz; k++)
      for (int l = 1; l <= NG; l++)
        for (int i = 1 - NG; i <= d.nx + NG; i++) {
          a[ID(d, i, 1 - l, k)] = a[ID(d, i, d.ny + 1 - l, k)];
          a[ID(d, i, d.ny + l, k)] = a[ID(d, i, l, k)];
        }
  }
  if (c->face[2][0] == ORC_FACE_PERIODIC) {
#pragma omp parallel for
    for (int l = 1; l <= NG; l++)
      for (int j = 1 - NG; j <= d.ny + NG; j++)
        for (int i = 1 - NG; i <= d.nx + NG; i++) {
          a[ID(d, i, j, 1 - l)] = a[ID(d, i, j, d.nz + 1 - l)];
          a[ID(d, i, j, d.nz + l)] = a[ID(d, i, j, l)];
        }
  }
}

/* ------------------------------------------------------------------------ */
/* Primitive variables (smod_primitive_pfg.f90:12-59, 62-136), whole box.    */
/* ------------------------------------------------------------------------ */
typedef struct {
  double *uu, *vv, *ww, *prs, *tmp; /* primitives                          */
  double *cs, *visc, *cok;          /* sound speed, viscosity, conductivity */
  double *grad[12];                 /* dux,duy,duz,dvx,dvy,dvz,dwx,dwy,dwz,dTx,dTy,dTz */
} work_t;

enum { DUX, DUY, DUZ, DVX, DVY, DVZ, DWX, DWY, DWZ, DTX, DTY, DTZ };

static void prim_euler(const orc_cfg *c, const dims_t *d, double *const U[5], work_t *w) {
  const double gam1 = c->gam - 1.0, rg = c->rg;
  const long n = (long)d->n;
#pragma omp parallel for
  for (long p = 0; p < n; p++) {
    double rro = U[0][p];
    if (rro == 0.0) { w->uu[p] = w->vv[p] = w->ww[p] = w->prs[p] = w->tmp[p] = 0.0; continue; }
    double v = 1.0 / rro;
    double u1 = U[1][p] * v, u2 = U[2][p] * v, u3 = U[3][p] * v;
    w->uu[p] = u1; w->vv[p] = u2; w->ww[p] = u3;
    double pr = gam1 * (U[4][p] - 0.5 * rro * (u1 * u1 + u2 * u2 + u3 * u3));
    w->prs[p] = pr;
    w->tmp[p] = pr * v / rg;
  }
}

static void prim_visc(const orc_cfg *c, const dims_t *d, double *const U[5], work_t *w) {
  const double gam1 = c->gam - 1.0, rg = c->rg, gam = c->gam;
  const long n = (long)d->n;
#pragma omp parallel for
  for (long p = 0; p < n; p++) {
    double rro = U[0][p];
    if (rro == 0.0) {
      w->uu[p] = w->vv[p] = w->ww[p] = w->prs[p] = w->tmp[p] = 0.0;
      w->cs[p] = w->visc[p] = w->cok[p] = 0.0;
      continue;
    }
    double v = 1.0 / rro;
    double u1 = U[1][p] * v, u2 = U[2][p] * v, u3 = U[3][p] * v;
    w->uu[p] = u1; w->vv[p] = u2; w->ww[p] = u3;
    double pr = gam1 * (U[4][p] - 0.5 * rro * (u1 * u1 + u2 * u2 + u3 * u3));
    double t = pr * v / rg;
    w->prs[p] = pr;
    w->tmp[p] = t;
    w->cs[p] = sqrt(gam * rg * t);
    double mu = c->c0 * sqrt(t * t * t) / (t + c->S0) * c->diffscale;
    w->visc[p] = mu;
    w->cok[p] = c->cok_ratio * mu;
  }
}

/* ------------------------------------------------------------------------ */
/* Euler flux divergence.                                                    */
/* Flux along one direction on the whole line (incl. 5 ghosts), then the     */
/* interior 11-pt difference.  Periodic faces only: loop bounds ndxt..nfxt = */
/* 1-ng..n+ng and ndx..nfx = 1..n (mpi_index_bounds.f90:82-107).             */
/* ------------------------------------------------------------------------ */
static inline double d11(const double *f, long p, long s) {
  return A11[1] * (f[p + s] - f[p - s]) + A11[2] * (f[p + 2 * s] - f[p - 2 * s]) +
         A11[3] * (f[p + 3 * s] - f[p - 3 * s]) + A11[4] * (f[p + 4 * s] - f[p - 4 * s]) +
         A11[5] * (f[p + 5 * s] - f[p - 5 * s]);
}

/* Metric term q of the c3 set at element p (ksi_x,ksi_y,ksi_z,eta_x,..,
 * phi_z, ijacob3 = 0..9), or on a 2.5-D block (is_curv) its counterpart from
 * the 2-D arrays m2 = x_ksi, x_eta, y_ksi, y_eta, ijacob: the ksi-flux
 * projects on (y_eta, -x_eta, 0) and the eta-flux on (-y_ksi, x_ksi, 0)
 * (flux_euler_11pts_c, smod_flux_euler_11pts_c.f90:25-37, :265-277; the
 * gradients smod_grad_vel_c.f90:186-217), 1/J = ijacob.  The phi rows are
 * not used there: z is Cartesian. */
static inline double mtr(const orc_cfg *c, const double *const m3[10], const double *const m2[5], const dims_t *d,
                         int q, long p) {
  if (c->curv != 2) return m3[q][p];
  const long p2 = p % ((long)d->sx * d->sy);
  switch (q) {
    case 0: return m2[3][p2];
    case 1: return -m2[1][p2];
    case 3: return -m2[2][p2];
    case 4: return m2[0][p2];
    case 9: return m2[4][p2];
    default: return 0.0;
  }
}

/* Physical (Cartesian) flux along direction dir at linear offset p. */
static inline void flux_cart(int dir, double *const U[5], const work_t *w, long p, double F[5]) {
  double pr = w->prs[p];
  double m = U[1 + dir][p]; /* rhou, rhov or rhow */
  double vel = dir == 0 ? w->uu[p] : (dir == 1 ? w->vv[p] : w->ww[p]);
  F[0] = m;
  F[1] = m * w->uu[p] + (dir == 0 ? pr : 0.0);
  F[2] = m * w->vv[p] + (dir == 1 ? pr : 0.0);
  F[3] = m * w->ww[p] + (dir == 2 ? pr : 0.0);
  F[4] = (U[4][p] + pr) * vel;
}

/* Contravariant flux along direction dir (smod_flux_euler_11pts_c3.f90:25-38;
 * 2.5-D: smod_flux_euler_11pts_c.f90:25-37, :265-277). */
static inline void flux_c3(const orc_cfg *c, const dims_t *d, int dir, double *const U[5], const work_t *w,
                           const orc_grid *g, long p, double F[5]) {
  const double mx = mtr(c, g->m, g->m2, d, 3 * dir, p), my = mtr(c, g->m, g->m2, d, 3 * dir + 1, p),
               mz = mtr(c, g->m, g->m2, d, 3 * dir + 2, p);
  double u = w->uu[p], v = w->vv[p], ww = w->ww[p], pr = w->prs[p];
  double vc = u * mx + v * my + ww * mz;
  double rovc = U[0][p] * vc;
  F[0] = rovc;
  F[1] = rovc * u + pr * mx;
  F[2] = rovc * v + pr * my;
  F[3] = rovc * ww + pr * mz;
  F[4] = (U[4][p] + pr) * vc;
}

static void euler_rhs_w(const orc_cfg *c, const orc_grid *g, const dims_t *d, double *const U[5],
                        const work_t *w, double *K[5]) {
  const long sx = 1, sy = d->sx, sz = d->sx * d->sy;
  const long strides[3] = {sx, sy, sz};
  const int nn[3] = {d->nx, d->ny, d->nz};
  const walls_t W = bounds_of(c), Wl = walls_of(c);
  /* rows with an Euler increment along each axis: 2..n-1 next to boundaries
   * (ndx_e, nfx_e); K on the boundary rows is left as it is (bc_wall zeroes it
   * on wall rows, the characteristic routines set it on theirs).  With
   * is_wall2 the wall rows are advanced too (mpi_index_bounds.f90:28-70). */
  int lo_e[3], hi_e[3], adv_lo[3], adv_hi[3];
  for (int a = 0; a < 3; a++) {
    adv_lo[a] = c->wall2 && Wl.lo[a];
    adv_hi[a] = c->wall2 && Wl.hi[a];
    lo_e[a] = (W.lo[a] && !adv_lo[a]) ? 2 : 1;
    hi_e[a] = nn[a] - ((W.hi[a] && !adv_hi[a]) ? 1 : 0);
  }
  /* line buffers: fluxes along a line with ghosts */
  for (int dir = 0; dir < 3; dir++) {
    const long s = strides[dir];
    const int n = nn[dir];
    /* iterate over the 2 transverse indices with an increment */
    int o1 = dir == 0 ? 1 : 0, o2 = dir == 2 ? 1 : 2;
#pragma omp parallel
    {
      double *fl = (double *)malloc(sizeof(double) * 5 * (size_t)(n + 2 * NG));
#pragma omp for collapse(2)
      for (int b = lo_e[o2]; b <= hi_e[o2]; b++)
        for (int a = lo_e[o1]; a <= hi_e[o1]; a++) {
          int idx3[3];
          idx3[o1] = a; idx3[o2] = b; idx3[dir] = 1 - NG;
          long p0 = (long)ID(*d, idx3[0], idx3[1], idx3[2]);
          /* fluxes on 1-ng..n+ng */
          for (int l = 0; l < n + 2 * NG; l++) {
            double F[5];
            long p = p0 + l * s;
            if (c->curv == 1 || (c->curv == 2 && dir < 2)) flux_c3(c, d, dir, U, w, g, p, F);
            else flux_cart(dir, U, w, p, F);
            for (int v = 0; v < 5; v++) fl[v * (n + 2 * NG) + l] = F[v];
          }
          for (int i = lo_e[dir]; i <= hi_e[dir]; i++) {
            long p = p0 + (long)(i - 1 + NG) * s;
            long q = i - 1 + NG; /* position in line buffer */
            const int hw = halfw(i, n, W.lo[dir], W.hi[dir]);
            double metric = 1.0;
            const int cart_dir = !c->curv || (c->curv == 2 && dir == 2);
            if (cart_dir) metric = dir == 0 ? g->idx[i - 1 + NG] : (dir == 1 ? g->idy[i - 1 + NG] : g->idz[i - 1 + NG]);
            for (int v = 0; v < 5; v++) {
              const double *f = fl + v * (n + 2 * NG);
              double dv;
              if (hw == 5) {
                dv = d11(f, q, 1);
              } else if (hw == 0) {
                /* advanced wall row: a02 at the min face, a20 at the max face
                 * (smod_flux_euler_11pts.f90:42-54, 254-265; _c3.f90:291-302) */
                dv = i == 1 ? A02[0] * f[q] + A02[1] * f[q + 1] + A02[2] * f[q + 2]
                            : (ORC_PLANT == 10 ? 1.0 : -1.0) * (A02[0] * f[q] + A02[1] * f[q - 1] + A02[2] * f[q - 2]);
              } else {
                /* rows 2..5 from a wall: a3, a5, a7, a9 (smod_flux_euler_11pts.f90:36-133) */
                dv = 0.0;
                for (int l = 1; l <= hw; l++) dv = dv + ACL[hw][l] * (f[q + l] - f[q - l]);
              }
              if (cart_dir) dv = dv * metric;
              if (dir == 0) K[v][p] = dv;
              else K[v][p] = dv + K[v][p];
            }
          }
        }
      free(fl);
    }
    if (c->curv == 2 && dir == 1) {
      /* 2.5-D: the ksi and eta parts times ijacob, then the Cartesian z part
       * (smod_flux_euler_11pts_c.f90:527-540) */
#pragma omp parallel for collapse(2)
      for (int k = 1; k <= d->nz; k++)
        for (int j = 1; j <= d->ny; j++)
          for (int i = 1; i <= d->nx; i++) {
            long p = (long)ID(*d, i, j, k);
            double ij = mtr(c, g->m, g->m2, d, 9, p);
            for (int v = 0; v < 5; v++) K[v][p] = K[v][p] * ij;
          }
    }
  }
  if (c->curv == 1) {
    /* K *= ijacob3 (smod_flux_euler_11pts_c3.f90:763-773) */
#pragma omp parallel for collapse(2)
    for (int k = 1; k <= d->nz; k++)
      for (int j = 1; j <= d->ny; j++)
        for (int i = 1; i <= d->nx; i++) {
          long p = (long)ID(*d, i, j, k);
          double ij = g->m[9][p];
          for (int v = 0; v < 5; v++) K[v][p] = K[v][p] * ij;
        }
  }
}

/* ------------------------------------------------------------------------ */
/* Viscous terms: 5-pt gradients of u,v,w,T on the interior                  */
/* (smod_grad_vel.f90:12-, smod_grad_T.f90:12-; c3: smod_grad_vel_c3.f90:11- */
/* with the Cartesian mapping at :260-300), halo of the 12 gradients         */
/* (communication_3dv, mod_comm.f90:376-402), then the viscous flux and its  */
/* 5-pt divergence (smod_flux_visc_5pts.f90:27-, smod_flux_visc_5pts_c3.f90).*/
/* ------------------------------------------------------------------------ */
static inline double d5(const double *f, long p, long s) {
  return A5[1] * (f[p + s] - f[p - s]) + A5[2] * (f[p + 2 * s] - f[p - 2 * s]);
}

static void gradients(const orc_cfg *c, const orc_grid *g, const dims_t *d, work_t *w) {
  const long sy = d->sx, sz = d->sx * d->sy;
  double *src[4] = {w->uu, w->vv, w->ww, w->tmp};
  const walls_t W = bounds_of(c), Wall = walls_of(c);
#pragma omp parallel for collapse(2)
  for (int k = 1; k <= d->nz; k++)
    for (int j = 1; j <= d->ny; j++)
      for (int i = 1; i <= d->nx; i++) {
        long p = (long)ID(*d, i, j, k);
        for (int q = 0; q < 4; q++) {
          double gx = g5w(src[q], p, 1, i, d->nx, W.lo[0], W.hi[0]);
          double gy = g5w(src[q], p, sy, j, d->ny, W.lo[1], W.hi[1]);
          double gz = g5w(src[q], p, sz, k, d->nz, W.lo[2], W.hi[2]);
          if (!c->curv) {
            gx = gx * g->idx_v[i - 1 + NG];
            gy = gy * g->idy_v[j - 1 + NG];
            gz = gz * g->idz_v[k - 1 + NG];
          } else if (c->curv == 2) {
            /* (f_ksi y_eta_v - f_eta y_ksi_v) ijacob_v, (f_eta x_ksi_v - f_ksi x_eta_v)
             * ijacob_v, f_z idz_v (smod_grad_vel_c.f90:186-217, :230-300) */
            const double ijv = mtr(c, g->mv, g->m2v, d, 9, p);
            double dvx = gx * mtr(c, g->mv, g->m2v, d, 0, p) + gy * mtr(c, g->mv, g->m2v, d, 3, p);
            double dvy = gx * mtr(c, g->mv, g->m2v, d, 1, p) + gy * mtr(c, g->mv, g->m2v, d, 4, p);
            gx = dvx * ijv;
            gy = dvy * ijv;
            gz = gz * g->idz_v[k - 1 + NG];
          } else {
            const double *const *mv = g->mv;
            double dvx = gx * mv[0][p] + gy * mv[3][p] + gz * mv[6][p];
            double dvy = gx * mv[1][p] + gy * mv[4][p] + gz * mv[7][p];
            double dvz = gx * mv[2][p] + gy * mv[5][p] + gz * mv[8][p];
            gx = dvx * mv[9][p];
            gy = dvy * mv[9][p];
            gz = dvz * mv[9][p];
          }
          w->grad[3 * q][p] = gx;
          w->grad[3 * q + 1][p] = gy;
          w->grad[3 * q + 2][p] = gz;
        }
        /* adiabatic c3 j-walls: on the wall row the temperature gradient is
         * the index-space derivatives along ksi and phi times the unit tangents
         * x_ksi/|x_ksi|, x_phi/|x_phi| (smod_grad_T_c3.f90:80-85, 194-197,
         * 249-275 as written: the derivatives are not divided by the tangents'
         * lengths) */
        if (c->curv == 1 && c->adiabatic && ((Wall.lo[1] && j == 1) || (Wall.hi[1] && j == d->ny))) {
          const int f = j == 1 ? 2 : 3;
          const long q = (long)(i - 1) + (long)d->nx * (k - 1);
          const double dk = g5w(src[3], p, 1, i, d->nx, W.lo[0], W.hi[0]);
          const double dp = g5w(src[3], p, sz, k, d->nz, W.lo[2], W.hi[2]);
          for (int a = 0; a < 3; a++)
            w->grad[DTX + a][p] = ORC_PLANT == 9 ? dk * g->wc3[f][5 + a][q] + dp * g->wc3[f][2 + a][q]
                                                 : dk * g->wc3[f][2 + a][q] + dp * g->wc3[f][5 + a][q];
        }
        /* adiabatic walls: no wall-normal temperature gradient on the wall row
         * (smod_grad_T.f90:102-109 and its twins) */
        if (c->curv != 1 && c->adiabatic && Wall.any) {
          const int rr[3] = {i, j, k}, nn[3] = {d->nx, d->ny, d->nz};
          for (int a = 0; a < 3; a++)
            if ((Wall.lo[a] && rr[a] == 1) || (Wall.hi[a] && rr[a] == nn[a])) w->grad[DTX + a][p] = 0.0;
        }
      }
}

static void visc_rhs_w(const orc_cfg *c, const orc_grid *g, const dims_t *d, work_t *w, double *K[5]) {
  const long sy = d->sx, sz = d->sx * d->sy;
  /* viscous fluxes on ndxt_v..nfxt_v = -1..n+2 (periodic: ndx_vi=1, ngh_v=2) */
  double *F[12];
  for (int q = 0; q < 12; q++) F[q] = alloc_arr(d); /* Frhou..Frhoe, Grhou.., Hrhou.. */
  const double one_third = 1.0 / 3.0;
#pragma omp parallel for collapse(2)
  for (int k = -1; k <= d->nz + 2; k++)
    for (int j = -1; j <= d->ny + 2; j++)
      for (int i = -1; i <= d->nx + 2; i++) {
        long p = (long)ID(*d, i, j, k);
        double *const *gr = w->grad;
        double t11 = gr[DUX][p], t22 = gr[DVY][p], t33 = gr[DWZ][p];
        double t12 = 0.5 * (gr[DUY][p] + gr[DVX][p]);
        double t13 = 0.5 * (gr[DUZ][p] + gr[DWX][p]);
        double t23 = 0.5 * (gr[DVZ][p] + gr[DWY][p]);
        double tr = one_third * (t11 + t22 + t33);
        double mu = -2.0 * w->visc[p];
        t11 = mu * (t11 - tr);
        t22 = mu * (t22 - tr);
        t33 = mu * (t33 - tr);
        t12 = mu * t12;
        t13 = mu * t13;
        t23 = mu * t23;
        double u = w->uu[p], v = w->vv[p], ww = w->ww[p], ck = w->cok[p];
        if (!c->curv) {
          F[0][p] = t11; F[1][p] = t12; F[2][p] = t13;
          F[3][p] = (u * F[0][p] + v * F[1][p] + ww * F[2][p]) - ck * gr[DTX][p];
          F[4][p] = t12; F[5][p] = t22; F[6][p] = t23;
          F[7][p] = (u * F[4][p] + v * F[5][p] + ww * F[6][p]) - ck * gr[DTY][p];
          F[8][p] = t13; F[9][p] = t23; F[10][p] = t33;
          F[11][p] = (u * F[8][p] + v * F[9][p] + ww * F[10][p]) - ck * gr[DTZ][p];
        } else {
          for (int dir = 0; dir < 3; dir++) {
            if (c->curv == 2 && dir == 2) {
              /* 2.5-D: Cartesian z flux (smod_flux_visc_5pts_c.f90:67-72) */
              F[8][p] = t13; F[9][p] = t23; F[10][p] = t33;
              F[11][p] = (u * F[8][p] + v * F[9][p] + ww * F[10][p]) - ck * gr[DTZ][p];
              continue;
            }
            double mx = mtr(c, g->mv, g->m2v, d, 3 * dir, p), my = mtr(c, g->mv, g->m2v, d, 3 * dir + 1, p),
                   mz = mtr(c, g->mv, g->m2v, d, 3 * dir + 2, p);
            double fu = t11 * mx + t12 * my + t13 * mz;
            double fv = t12 * mx + t22 * my + t23 * mz;
            double fw = t13 * mx + t23 * my + t33 * mz;
            F[4 * dir][p] = fu;
            F[4 * dir + 1][p] = fv;
            F[4 * dir + 2][p] = fw;
            F[4 * dir + 3][p] = (u * fu + v * fv + ww * fw) -
                                ck * (gr[DTX][p] * mx + gr[DTY][p] * my + gr[DTZ][p] * mz);
          }
        }
      }
  /* divergence, x then y then z, each ( ... )*metric + K; one-sided rows at walls
   * (smod_flux_visc_5pts.f90:80-160) */
  const walls_t W = bounds_of(c), Wc = chars_of(c);
  const int nn[3] = {d->nx, d->ny, d->nz};
#pragma omp parallel for collapse(2)
  for (int k = 1; k <= d->nz; k++)
    for (int j = 1; j <= d->ny; j++)
      for (int i = 1; i <= d->nx; i++) {
        long p = (long)ID(*d, i, j, k);
        const long st[3] = {1, sy, sz};
        const int rr[3] = {i, j, k};
        /* the row of a characteristic face gets no viscous increment along any
         * axis (ndx_v = 2 there: smod_flux_visc_5pts.f90:80-175 loops) */
        int onrow = 0;
        for (int a = 0; a < 3; a++) onrow |= (Wc.lo[a] && rr[a] == 1) || (Wc.hi[a] && rr[a] == nn[a]);
        if (onrow) continue;
        for (int dir = 0; dir < 3; dir++) {
          double metric;
          if (c->curv == 1 || (c->curv == 2 && dir < 2)) metric = mtr(c, g->mv, g->m2v, d, 9, p);
          else metric = dir == 0 ? g->idx_v[i - 1 + NG] : (dir == 1 ? g->idy_v[j - 1 + NG] : g->idz_v[k - 1 + NG]);
          for (int v = 0; v < 4; v++)
            K[1 + v][p] = g5w(F[4 * dir + v], p, st[dir], rr[dir], nn[dir], W.lo[dir], W.hi[dir]) * metric +
                          K[1 + v][p];
        }
      }
  for (int q = 0; q < 12; q++) free(F[q]);
}

/* ------------------------------------------------------------------------ */
/* Selective filter, interior 11-pt stencil (mod_filtering.f90:992-1034,     */
/* 1192-1233, 1382-1423); coefficients d = -chi*d11 (init_filtering :189-203)*/
/* ------------------------------------------------------------------------ */
static void filter_sf(const orc_cfg *c, const dims_t *d, double *const Un[5], double *K[5]) {
  double dd[6];
  for (int l = 0; l < 6; l++) dd[l] = -D11[l] * c->chi;
  const long st[3] = {1, d->sx, d->sx * d->sy};
  const walls_t W = bounds_of(c);
  if (W.any) {
    /* rows 1,2 from a boundary get nothing along its axis, rows 3,4,5 the 5/7/9-pt
     * filters (mod_filtering.f90:900-988, 1116-1190) */
    const int nn[3] = {d->nx, d->ny, d->nz};
    double dc[6][6];
    for (int q = 2; q <= 5; q++)
      for (int l = 0; l < 6; l++) dc[q][l] = -DCL[q][l] * c->chi;
#pragma omp parallel for collapse(2)
    for (int k = 1; k <= d->nz; k++)
      for (int j = 1; j <= d->ny; j++)
        for (int i = 1; i <= d->nx; i++) {
          long p = (long)ID(*d, i, j, k);
          const int rr[3] = {i, j, k};
          int hw[3];
          for (int a = 0; a < 3; a++) hw[a] = halfw(rr[a], nn[a], W.lo[a], W.hi[a]);
          for (int v = 0; v < 5; v++) {
            const double *u = Un[v];
            double acc = 0.0;
            for (int dir = 0; dir < 3; dir++) {
              const int q = hw[dir];
              if (q < 2) continue;
              const long s = st[dir];
              double t = dc[q][0] * u[p];
              for (int l = 1; l <= q; l++) t = t + dc[q][l] * (u[p + l * s] + u[p - l * s]);
              acc = acc + t;
            }
            K[v][p] = acc;
          }
        }
    return;
  }
#pragma omp parallel for collapse(2)
  for (int k = 1; k <= d->nz; k++)
    for (int j = 1; j <= d->ny; j++)
      for (int i = 1; i <= d->nx; i++) {
        long p = (long)ID(*d, i, j, k);
        for (int v = 0; v < 5; v++) {
          const double *u = Un[v];
          /* x sets K (:993), y and z accumulate as K + d0*u + ... (:1192, :1382) */
          double acc = dd[0] * u[p] + dd[1] * (u[p + 1] + u[p - 1]) + dd[2] * (u[p + 2] + u[p - 2]) +
                       dd[3] * (u[p + 3] + u[p - 3]) + dd[4] * (u[p + 4] + u[p - 4]) +
                       dd[5] * (u[p + 5] + u[p - 5]);
          for (int dir = 1; dir < 3; dir++) {
            long s = st[dir];
            acc = acc + dd[0] * u[p] + dd[1] * (u[p + s] + u[p - s]) + dd[2] * (u[p + 2 * s] + u[p - 2 * s]) +
                  dd[3] * (u[p + 3 * s] + u[p - 3 * s]) + dd[4] * (u[p + 4 * s] + u[p - 4 * s]) +
                  dd[5] * (u[p + 5 * s] + u[p - 5 * s]);
          }
          K[v][p] = acc;
        }
      }
}

/* ------------------------------------------------------------------------ */
/* Shock filter (filter_o10_shock, mod_filtering_shock.f90:204-1512), and   */
/* with filter 4 the artificial viscosity with shock capturing that shares  */
/* its sensor and faces (artvisc_o9_shock, mod_artvisc_shock.f90:113-1287): */
/* Ducros sensor (:223-239), pressure sensor per direction (:243-260), face  */
/* dissipation D_{i+1/2} on 0..n (:419-440) and K_i = D_{i-1/2} - D_{i+1/2}  */
/* accumulated over directions (:690-700).  Along an axis with no-slip walls */
/* the sensor runs over 1..n (row 1 reads p one row beyond the wall, i.e.    */
/* the pressure of the ghost state, which nothing on the path changes), the  */
/* faces over 1..n-1 and K over 2..n-1.                                      */
/* ------------------------------------------------------------------------ */
/* Spectral radius of U^n along dir at p and the metric its face differences
 * are multiplied by: Cartesian |u_dir| + c and idx_dir(r); curvilinear (c3,
 * and the i, j directions of a 2.5-D block) |contravariant velocity| +
 * c |grad xi_dir| |1/J| and 1 (artvisc_o9 mod_artvisc.f90:324-372;
 * artvisc_o9_shock mod_artvisc_shock.f90:431-470; g_ksi grid_metrics_curv.f90:709-724,
 * g3_ksi smod_grid_metrics3_c3.f90:161). */
static inline double spec_radius(const orc_cfg *c, const orc_grid *g, const dims_t *d, const work_t *w, int dir,
                                 long p, int r, double *met) {
  if (c->curv == 1 || (c->curv == 2 && dir < 2)) {
    const double mx = mtr(c, g->m, g->m2, d, 3 * dir, p), my = mtr(c, g->m, g->m2, d, 3 * dir + 1, p),
                 mz = mtr(c, g->m, g->m2, d, 3 * dir + 2, p), ij = mtr(c, g->m, g->m2, d, 9, p);
    const double uc = (w->uu[p] * mx + w->vv[p] * my + w->ww[p] * mz) * ij;
    *met = 1.0;
    return fabs(uc) + w->cs[p] * (sqrt(mx * mx + my * my + mz * mz) * fabs(ij));
  }
  const double *vel = dir == 0 ? w->uu : (dir == 1 ? w->vv : w->ww);
  const double *m = dir == 0 ? g->idx : (dir == 1 ? g->idy : g->idz);
  *met = m[r - 1 + NG];
  return fabs(vel[p]) + w->cs[p];
}

static void filter_shock(const orc_cfg *c, const orc_grid *g, const dims_t *d, double *const Un[5],
                         const work_t *w, double *K[5]) {
  const double xnu0 = -c->dissip_shock, cs = c->csens;
  /* face amplitudes and differences by half-width hf = 1..5: d2..d10 with
   * xnu2 = 0, xnu4 = chi/12, xnu6 = chi/60, xnu8 = chi/280, xnu10 = chi/1260
   * (init_filter_shock with iSF = 0, mod_filtering_shock.f90:75-107); next to
   * a wall the faces 1..4 use d2..d8, face 1 a tenth of the sensor term and
   * xk = max(0, xnu + xk0) without the 1/105 of d10 (filter_o10_shock
   * :724-1057 j, :266-466 i), and the wall rows get no increment along that
   * axis (:1128-1140 over ndy_d..nfy_d) */
  const double xnu[6] = {0.0, 0.0, c->chi / 12.0, c->chi / 60.0, c->chi / 280.0, c->chi / 1260.0};
  static const double DF[6][6] = {{0}, {0, -1.0}, {0, -3.0, 1.0}, {0, -10.0, 5.0, -1.0},
                                  {0, -35.0, 21.0, -7.0, 1.0}, {0, -126.0, 84.0, -36.0, 9.0, -1.0}};
  const walls_t W = bounds_of(c);
  double *sens = alloc_arr(d);
  /* Ducros sensor on 0..n+1 (1..n along a wall axis: ndx_d1..nfx_d1) */
#pragma omp parallel for collapse(2)
  for (int k = 0; k <= d->nz + 1; k++)
    for (int j = 0; j <= d->ny + 1; j++)
      for (int i = 0; i <= d->nx + 1; i++) {
        long p = (long)ID(*d, i, j, k);
        if (c->ducros) {
          double *const *gr = w->grad;
          double dv = gr[DUX][p] + gr[DVY][p] + gr[DWZ][p];
          double div2 = dv * dv;
          double o1 = gr[DWY][p] - gr[DVZ][p];
          double o2 = gr[DUZ][p] - gr[DWX][p];
          double o3 = ORC_PLANT == 3 ? 0.0 : gr[DVX][p] - gr[DUY][p];
          sens[p] = div2 / (div2 + o1 * o1 + o2 * o2 + o3 * o3 + 1.0e-16);
        } else {
          sens[p] = 1.0;
        }
      }
  const long st[3] = {1, d->sx, d->sx * d->sy};
  const int nn[3] = {d->nx, d->ny, d->nz};
  const double *prs = w->prs;
  for (int dir = 0; dir < 3; dir++) {
    const long s = st[dir];
    const int n = nn[dir];
    const int lo = W.lo[dir], hi = W.hi[dir];
    int o1 = dir == 0 ? 1 : 0, o2 = dir == 2 ? 1 : 2;
    int n1 = nn[o1], n2 = nn[o2];
#pragma omp parallel
    {
      double *ps = (double *)malloc(sizeof(double) * (size_t)(n + 2));
      double *D = (double *)malloc(sizeof(double) * 5 * (size_t)(n + 1));
#pragma omp for collapse(2)
      for (int b = 1; b <= n2; b++)
        for (int a = 1; a <= n1; a++) {
          int idx3[3];
          idx3[o1] = a; idx3[o2] = b; idx3[dir] = 0;
          long p0 = (long)ID(*d, idx3[0], idx3[1], idx3[2]); /* offset of index 0 */
          /* pressure sensor on 0..n+1 (1..n next to walls; row 1 reads p one row beyond
           * the wall, the pressure of the ghost state, which nothing changes) */
          for (int i = lo ? 1 : 0; i <= (hi ? n : n + 1); i++) {
            long p = p0 + i * s;
            double ptrad = prs[p - s] + (ORC_PLANT == 1 ? 4.0 : 2.0) * prs[p] + prs[p + s];
            double ptvd = fabs(prs[p + s] - prs[p]) + fabs(prs[p] - prs[p - s]);
            ps[i] = sens[p] * fabs(prs[p - s] - 2.0 * prs[p] + prs[p + s]) / (cs * ptrad + (1.0 - cs) * ptvd);
          }
          /* face dissipation D_{m+1/2} on 0..n (1..n-1 next to walls) */
          for (int m = lo ? 1 : 0; m <= (hi ? n - 1 : n); m++) {
            int hf = 5;
            if (lo && m < hf) hf = m;
            if (hi && n - m < hf) hf = n - m;
            long p = p0 + m * s;
            double xk0 = xnu0 * fmax(ps[m + 1], ps[m]);
            if (hf == 1) xk0 = xk0 / 10.0;
            /* artvisc_o9_shock scales the sensor term of faces 3 and 4 from a wall
             * by 1/5 and 3/70 (mod_artvisc_shock.f90:577-695, :744-862) */
            double xks = hf == 5 ? xk0 / (ORC_PLANT == 2 ? 10.5 : 105.0) : xk0;
            if (c->filter == 4 && hf == 3) xks = xk0 / 5.0;
            if (c->filter == 4 && hf == 4) xks = 3.0 / 70.0 * xk0;
            const double xkh = fmax(0.0, xnu[hf] + xks);
            for (int v = 0; v < 5; v++) {
              const double *u = Un[v];
              double dd = 0.0;
              for (int l = 1; l <= hf; l++) dd = dd + DF[hf][l] * (u[p + l * s] - u[p + (1 - l) * s]);
              D[v * (n + 1) + m] = (u[p + s] - u[p]) * xk0 + dd * xkh;
            }
          }
          /* K_i = D_{i-1/2} - D_{i+1/2} on 1..n (2..n-1 next to walls); with the
           * Edoh switch times min(1, cfl/chi), cfl the local CFL number of U^n
           * along dir: (|u_dir| + c) dt idx_dir(i) (Cartesian), (|vc| + c g3) dt
           * with the contravariant velocity vc = (u xi_x + v xi_y + w xi_z) ijacob3
           * and g3 = |grad xi| |ijacob3| (c3) (mod_filtering_shock.f90:620-700,
           * :1059-1140, :1450-1496; g3_ksi smod_grid_metrics3_c3.f90:161) */
          for (int i = 1; i <= n; i++) {
            long p = p0 + i * s;
            const int edge = (lo && i == 1) || (hi && i == n);
            double mul = 1.0;
            if (c->sw_edoh) {
              double cfl;
              if (c->curv == 1 || (c->curv == 2 && dir < 2)) {
                /* 2.5-D: vc = (u y_eta - v x_eta) ijacob, g_ksi = |grad ksi| |ijacob|
                 * (mod_filtering_shock.f90:650-652, :1089-1091; grid_metrics_curv.f90:709-724) */
                const double mx = mtr(c, g->m, g->m2, d, 3 * dir, p), my = mtr(c, g->m, g->m2, d, 3 * dir + 1, p),
                             mz = mtr(c, g->m, g->m2, d, 3 * dir + 2, p);
                const double ij = mtr(c, g->m, g->m2, d, 9, p);
                const double vc = (w->uu[p] * mx + w->vv[p] * my + w->ww[p] * mz) * ij;
                const double g3 = sqrt(mx * mx + my * my + mz * mz) * fabs(ij);
                cfl = (fabs(vc) + w->cs[p] * g3) * c->dt;
              } else {
                const double *vel = dir == 0 ? w->uu : (dir == 1 ? w->vv : w->ww);
                const double *met = dir == 0 ? g->idx : (dir == 1 ? g->idy : g->idz);
                cfl = (fabs(vel[p]) + w->cs[p]) * c->dt * met[i - 1 + NG];
              }
              mul = fmin(1.0, cfl / c->chi);
            }
            /* artvisc_o9_shock: the faces weighted by the spectral radius of U^n
             * averaged to them, times the metric (mod_artvisc_shock.f90:480-550 i,
             * :866-935 j, :1243-1283 k); the wall rows get no increment along
             * the wall axis (K over ndy_d..nfy_d) */
            double srm = 0.0, srp = 0.0, met = 1.0;
            if (c->filter == 4) {
              double m1, m2;
              const double sr = spec_radius(c, g, d, w, dir, p, i, &met);
              srp = 0.5 * (spec_radius(c, g, d, w, dir, p + s, i + 1, &m1) + sr);
              srm = 0.5 * (spec_radius(c, g, d, w, dir, p - s, i - 1, &m2) + sr);
            }
            for (int v = 0; v < 5; v++) {
              const double *Dv = D + v * (n + 1);
              double kd = edge ? 0.0 : (c->sw_edoh ? mul * (Dv[i - 1] - Dv[i]) : Dv[i - 1] - Dv[i]);
              if (c->filter == 4) kd = edge ? 0.0 : (srm * Dv[i - 1] - srp * Dv[i]) * met;
              if (dir == 0) K[v][p] = kd;
              else K[v][p] = K[v][p] + kd;
            }
          }
        }
      free(ps);
      free(D);
    }
  }
  free(sens);
}

/* ------------------------------------------------------------------------ */
/* Artificial viscosity (artvisc_o9, mod_artvisc.f90:82-874; coefficients   */
/* init_artvisc :16-79), periodic faces: per direction the face dissipation */
/* D_{i+1/2} = sum_l d10(l) (u(i+l) - u(i+1-l)) on 0..n with                */
/* d10 = (126,-84,36,-9,1) chi/1260, weighted by the spectral radius of U^n */
/* averaged to the faces, K = (sr_{i+1/2} D_{i+1/2} - sr_{i-1/2} D_{i-1/2}) */
/* * metric, x setting K and y, z adding to it: Cartesian |u_d| + c and     */
/* idx_d(i); 2.5-D (is_curv, :324-372) i, j with the contravariant velocity */
/* and g_ksi / g_eta and metric 1, z Cartesian (spec_radius).               */
/* w holds primitives_visc of U^n (velocity and sound speed).               */
/* ------------------------------------------------------------------------ */
static void artvisc(const orc_cfg *c, const orc_grid *g, const dims_t *d, double *const Un[5], const work_t *w,
                    double *K[5]) {
  /* coefficients times their amplitude (init_artvisc :30-77): d2, d4, d6, d8
   * for the faces 1..4 from a wall (xnu2 = 0), d10 inside */
  const double xnu[6] = {0.0, 0.0, c->chi / 12.0, c->chi / 60.0, c->chi / 280.0, c->chi / 1260.0};
  static const double DP[6][6] = {{0}, {0, 1.0}, {0, 3.0, -1.0}, {0, 10.0, -5.0, 1.0},
                                  {0, 35.0, -21.0, 7.0, -1.0}, {0, 126.0, -84.0, 36.0, -9.0, 1.0}};
  double dk[6][6];
  for (int h = 0; h < 6; h++)
    for (int l = 0; l < 6; l++) dk[h][l] = DP[h][l] * xnu[h];
  const double *d10 = dk[5];
  const walls_t W = bounds_of(c);
  const long st[3] = {1, d->sx, d->sx * d->sy};
  const int nn[3] = {d->nx, d->ny, d->nz};
  for (int dir = 0; dir < 3; dir++) {
    const long s = st[dir];
    const int n = nn[dir];
    const int lo = W.lo[dir], hi = W.hi[dir];
    const int o1 = dir == 0 ? 1 : 0, o2 = dir == 2 ? 1 : 2;
#pragma omp parallel
    {
      double *D = (double *)malloc(sizeof(double) * 5 * (size_t)(n + 1));
#pragma omp for collapse(2)
      for (int b = 1; b <= nn[o2]; b++)
        for (int a = 1; a <= nn[o1]; a++) {
          int idx3[3];
          idx3[o1] = a; idx3[o2] = b; idx3[dir] = 0;
          const long p0 = (long)ID(*d, idx3[0], idx3[1], idx3[2]);
          /* faces 0..n (1..n-1 along a wall axis: mod_artvisc.f90:98-376 i, :380-586 j) */
          for (int i = lo ? 1 : 0; i <= (hi ? n - 1 : n); i++) {
            const long p = p0 + i * s;
            int hf = 5;
            if (lo && i < hf) hf = i;
            if (hi && n - i < hf) hf = n - i;
            for (int v = 0; v < 5; v++) {
              const double *u = Un[v];
              if (hf == 5) {
                D[v * (n + 1) + i] = d10[1] * (u[p + s] - u[p]) + d10[2] * (u[p + 2 * s] - u[p - s]) +
                                     d10[3] * (u[p + 3 * s] - u[p - 2 * s]) + d10[4] * (u[p + 4 * s] - u[p - 3 * s]) +
                                     d10[5] * (u[p + 5 * s] - u[p - 4 * s]);
              } else {
                double dd = 0.0;
                for (int l = 1; l <= hf; l++) dd = dd + dk[hf][l] * (u[p + l * s] - u[p + (1 - l) * s]);
                D[v * (n + 1) + i] = dd;
              }
            }
          }
          /* K on 1..n; the wall rows get nothing along the wall axis (ndy_d..nfy_d) */
          for (int i = 1; i <= n; i++) {
            const long p = p0 + i * s;
            if ((lo && i == 1) || (hi && i == n)) {
              if (dir == 0)
                for (int v = 0; v < 5; v++) K[v][p] = 0.0;
              continue;
            }
            double m, m1, m2;
            const double sr = spec_radius(c, g, d, w, dir, p, i, &m);
            const double srp = 0.5 * (spec_radius(c, g, d, w, dir, p + s, i + 1, &m1) + sr);
            const double srm = 0.5 * (spec_radius(c, g, d, w, dir, p - s, i - 1, &m2) + sr);
            for (int v = 0; v < 5; v++) {
              const double *Dv = D + v * (n + 1);
              const double kd = (srp * Dv[i] - srm * Dv[i - 1]) * m;
              if (dir == 0) K[v][p] = kd;
              else K[v][p] = K[v][p] + kd;
            }
          }
        }
      free(D);
    }
  }
}

/* ------------------------------------------------------------------------ */
/* No-slip wall with dp/dn = 0 (bc_wall_*_adiabatic_dpdn / _isotherm_dpdn,   */
/* smod_bc_wall.f90:11-98 and its imax..kmax twins), faces in the order of   */
/* bc_wall (mod_bc_apply.f90:237-261), so a later face wins on an edge.  T   */
/* and p are the primitives last computed (w), not those of Un.              */
/* ------------------------------------------------------------------------ */
/* Advanced wall (is_wall2) on face f, over every point of its row
 * (smod_bc_wall_2.f90:11-473, smod_bc_wall_2_c3.f90:363-712):
 *   slip: the wall-normal momentum component = 0 on Cartesian blocks
 *         (bc_wall_jmin_slip :165-185); on the j-faces of c3 blocks the
 *         momentum m becomes (m.t_ksi) e_ksi + (m.t_phi) e_phi with the unit
 *         tangents e and their dual basis t (bc_wall_jmin_slip_c3 :363-393);
 *   no slip: zero momentum and rhoE = rho cv T_wall with the advanced rho,
 *         T_wall = cextrp2 T(2) - cextrp3 T(3) (adiabatic, :188-214), on c3
 *         blocks (4 T(2) - T(3) + the tangential terms)/3 (_c3.f90:396-536),
 *         or the isothermal wall's T_wall (:217-241), T the primitives last
 *         computed. */
static void bc_wall2_face(const orc_cfg *c, const orc_grid *g, const dims_t *d, double *const Un[5],
                          const work_t *w, int f) {
  const int a = f / 2, side = f % 2;
  const int nn[3] = {d->nx, d->ny, d->nz};
  const long st[3] = {1, d->sx, d->sx * d->sy};
  const int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
  const long s = side ? -st[a] : st[a]; /* from the wall into the block */
  const double cv = 1.0 / (c->gam - 1.0) * c->rg;
  const int slip = slip_face(c, f);
  const double c2 = g->cextrp2[f], c3 = g->cextrp3[f];
  const double *const *wc = g->wc3[f];
#pragma omp parallel for
  for (int q2 = 1; q2 <= nn[t2]; q2++)
    for (int q1 = 1; q1 <= nn[t1]; q1++) {
      int idx3[3];
      idx3[t1] = q1; idx3[t2] = q2; idx3[a] = side ? nn[a] : 1;
      const long p1 = (long)ID(*d, idx3[0], idx3[1], idx3[2]);
      const long q = (long)(q1 - 1) + (long)nn[t1] * (q2 - 1); /* c3 j-face: (i-1) + nx (k-1) */
      if (slip) {
        if (c->curv == 1) {
          const double m[3] = {Un[1][p1], Un[2][p1], Un[3][p1]};
          const int o = ORC_PLANT == 11 ? 2 : 8; /* planted: the unit tangents for their dual basis */
          const double vk = m[0] * wc[o][q] + m[1] * wc[o + 1][q] + m[2] * wc[o + 2][q];
          const double vp = m[0] * wc[o + 3][q] + m[1] * wc[o + 4][q] + m[2] * wc[o + 5][q];
          for (int x = 0; x < 3; x++) Un[1 + x][p1] = vk * wc[2 + x][q] + vp * wc[5 + x][q];
        } else {
          Un[1 + a][p1] = 0.0;
        }
        continue;
      }
      const long p2 = p1 + s, p3 = p1 + 2 * s;
      double t_wall = c->t_wall;
      if (c->adiabatic) {
        if (c->curv == 1) {
          const long si = 1, sk = st[2];
          t_wall = (1.0 / 3.0) * (4.0 * w->tmp[p2] - w->tmp[p3] + (w->tmp[p1 + si] - w->tmp[p1 - si]) * wc[0][q] +
                                  (w->tmp[p1 + sk] - w->tmp[p1 - sk]) * wc[1][q]);
        } else {
          t_wall = c2 * w->tmp[p2] - c3 * w->tmp[p3];
        }
      }
      Un[1][p1] = 0.0;
      Un[2][p1] = 0.0;
      Un[3][p1] = 0.0;
      Un[4][p1] = Un[0][p1] * (cv * t_wall); /* rho_n*ecalc_tro_pfg */
    }
}

static void bc_wall(const orc_cfg *c, const orc_grid *g, const dims_t *d, double *const Un[5], const work_t *w,
                    double *K[5]) {
  const walls_t W = walls_of(c);
  if (!W.any) return;
  if (c->wall2) {
    /* faces in the order of bc_wall (mod_bc_apply.f90:237-261); K stays */
    for (int f = 0; f < 6; f++)
      if (f % 2 ? W.hi[f / 2] : W.lo[f / 2]) bc_wall2_face(c, g, d, Un, w, f);
    return;
  }
  const walls_t B = bounds_of(c);
  const int nn[3] = {d->nx, d->ny, d->nz};
  /* K = 0 on the wall rows over the tangential rows with an Euler increment
   * (ndx_e..nfx_e, ndz_e..nfz_e: smod_bc_wall.f90:226-232); the rows of a
   * characteristic face keep the increment their routine gives them */
  if (K)
    for (int f = 0; f < 6; f++) {
      const int a = f / 2, side = f % 2;
      if (!(side ? W.hi[a] : W.lo[a])) continue;
      const int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
      for (int q2 = B.lo[t2] ? 2 : 1; q2 <= nn[t2] - (B.hi[t2] ? 1 : 0); q2++)
        for (int q1 = B.lo[t1] ? 2 : 1; q1 <= nn[t1] - (B.hi[t1] ? 1 : 0); q1++) {
          int idx3[3];
          idx3[t1] = q1; idx3[t2] = q2; idx3[a] = side ? nn[a] : 1;
          const long p1 = (long)ID(*d, idx3[0], idx3[1], idx3[2]);
          for (int v = 0; v < 5; v++) K[v][p1] = 0.0;
        }
    }
  const long st[3] = {1, d->sx, d->sx * d->sy};
  const double cv = 1.0 / (c->gam - 1.0) * c->rg; /* cvfg, mod_eos_pfg.f90:33 */
  for (int f = 0; f < 6; f++) {
    const int a = f / 2, side = f % 2;
    if (!(side ? W.hi[a] : W.lo[a])) continue;
    const int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
    const long s = side ? -st[a] : st[a]; /* from the wall into the block */
    if (c->curv == 1) {
      /* c3 j-wall (bc_wall_jmin/jmax_adiabatic_dpdn_c3, _isotherm_dpdn_c3,
       * smod_bc_wall_c3.f90:842-1065, 1068-1255, 1258-1671): p (and T) from
       * rows 2, 3 along eta plus the tangential terms
       *   v_wall = (4 v2 - v3 + (v(i+1,1) - v(i-1,1)) gksigeta3 + (v(i,1,k+1) - v(i,1,k-1)) gphigeta3)/3
       * with the wall row's own neighbours (i and k run over interfaces or
       * periodic faces here, so no edge rows) */
      const double *g1 = g->wc3[f][0], *g2 = g->wc3[f][1];
#pragma omp parallel for
      for (int k = 1; k <= d->nz; k++)
        for (int i = 1; i <= d->nx; i++) {
          const int j1 = side ? d->ny : 1;
          const long p1 = (long)ID(*d, i, j1, k);
          const long p2 = p1 + s, p3 = p1 + 2 * s;
          const long q = (long)(i - 1) + (long)d->nx * (k - 1);
          const long si = 1, sk = st[2];
          const double gk = (ORC_PLANT == 8 ? -1.0 : 1.0) * g1[q], gp = g2[q];
          const double t_wall =
              c->adiabatic ? (1.0 / 3.0) * (4.0 * w->tmp[p2] - w->tmp[p3] + (w->tmp[p1 + si] - w->tmp[p1 - si]) * gk +
                                             (w->tmp[p1 + sk] - w->tmp[p1 - sk]) * gp)
                           : c->t_wall;
          const double p_wall = (1.0 / 3.0) * (4.0 * w->prs[p2] - w->prs[p3] + (w->prs[p1 + si] - w->prs[p1 - si]) * gk +
                                                (w->prs[p1 + sk] - w->prs[p1 - sk]) * gp);
          const double ro = p_wall / (c->rg * t_wall);
          Un[1][p1] = 0.0;
          Un[2][p1] = 0.0;
          Un[3][p1] = 0.0;
          Un[0][p1] = ro;
          Un[4][p1] = ro * (cv * t_wall);
        }
      continue;
    }
    const double c2 = g->cextrp2[f], c3 = g->cextrp3[f];
#pragma omp parallel for
    for (int q2 = 1; q2 <= nn[t2]; q2++)
      for (int q1 = 1; q1 <= nn[t1]; q1++) {
        int idx3[3];
        idx3[t1] = q1; idx3[t2] = q2; idx3[a] = side ? nn[a] : 1;
        const long p1 = (long)ID(*d, idx3[0], idx3[1], idx3[2]);
        const long p2 = p1 + s, p3 = p1 + 2 * s;
        const double t_wall = c->adiabatic ? c2 * w->tmp[p2] - c3 * w->tmp[p3] : c->t_wall;
        const double p_wall = c2 * w->prs[p2] - c3 * w->prs[p3];
        const double ro = p_wall / (c->rg * t_wall); /* rocalc_pt_pfg */
        Un[1][p1] = 0.0;
        Un[2][p1] = 0.0;
        Un[3][p1] = 0.0;
        Un[0][p1] = ro;
        Un[4][p1] = ro * (cv * t_wall); /* ro*ecalc_tro_pfg */
      }
  }
}

/* ------------------------------------------------------------------------ */
/* Characteristic (LODI) outflow of a Cartesian block: the increment K of the */
/* boundary row from one-sided 5-pt derivatives (a40, times idx_v / idy_v)   */
/* of the primitives of the stage state Un, as bc_charac_imax and            */
/* bc_charac_jmax write it (src/Boundary/mod_charac.f90:350-469, 816-969),   */
/* for a perfect gas (cp/av = cp T, mod_eos_pfg.f90:38-84):                  */
/*   imax: L1 = (u-c)(px - rho c ux), or 0.1 (p - 0.528 p_ref) if u-c <= 0;  */
/*         L5 = (u+c)(px + rho c ux), 0 if u+c <= 0; L2 = u (px - c^2 rhox); */
/*         L3 = u vx; L4 = u wx; pt = (L5+L1)/2; rt = -(pt+L2)/c^2;          */
/*         ut = (L5-L1)/(2 rho c); vt = L3; wt = L4, with c the sound speed  */
/*         that primitives_visc last computed (cstale);                      */
/*   jmax: L1 = 0, L5 = (v+c)(py + rho c vy); v <= 0: L2 = L3 = L4 = 0 (and  */
/*         L5 = 0 if v <= -c); else L2 = v uy, L3 = v (py - c^2 rhoy),       */
/*         L4 = v wy (L1 = (v-c)(py - rho c vy) if v >= c); pt = (L5+L1)/2;  */
/*         rt = (pt+L3)/c^2; ut = L2; vt = (L5-L1)/(2 rho c); wt = L4, with  */
/*         c = sqrt(gam rg T) of the current primitives, on i = ndx_c..nfx_c;*/
/*   K = (rt, rho ut + u rt, rho vt + v rt, rho wt + w rt,                   */
/*        cp T (pt/c^2 - rt) + (rhoE + p)/rho rt + rho (u ut + v vt + w wt)).*/
/* The density equations are restated as written (they are not the textbook */
/* LODI ones); see the notes where the tests pin the other four.             */
/* ------------------------------------------------------------------------ */
static void bc_charac(const orc_cfg *c, const orc_grid *g, const dims_t *d, double *const Un[5],
                      const work_t *w, const double *cstale, double *K[5]) {
  const double cp = c->gam * c->rg / (c->gam - 1.0); /* cpfg */
  const long sy = d->sx;
  const walls_t B = bounds_of(c);
  if (c->face[0][1] == ORC_FACE_CHARAC) {
    const int i = d->nx;
    const double ix = g->idx_v[i - 1 + NG];
#pragma omp parallel for collapse(2)
    for (int k = 1; k <= d->nz; k++)
      for (int j = 1; j <= d->ny; j++) {
        const long p = (long)ID(*d, i, j, k);
        double dd[5];
        const double *f[5] = {w->uu, w->vv, w->ww, w->prs, Un[0]};
        for (int q = 0; q < 5; q++)
          dd[q] = (ORC_PLANT == 7 ? -1.0 : 1.0) *
                  ((-A04[5]) * f[q][p - 4] + (-A04[4]) * f[q][p - 3] + (-A04[3]) * f[q][p - 2] +
                   (-A04[2]) * f[q][p - 1] + (-A04[1]) * f[q][p]) * ix;
        const double dux = dd[0], dvx = dd[1], dwx = dd[2], dpx = dd[3], drx = dd[4];
        const double rho = Un[0][p], u = w->uu[p], v = w->vv[p], ww = w->ww[p], pr = w->prs[p];
        const double cj = cstale[p], cj2 = cj * cj, upc = u + cj, umc = u - cj;
        const double L1 = umc <= 0.0 ? 0.1 * (pr - (ORC_PLANT == 5 ? 0.5 : 0.528) * c->p_ref)
                                     : umc * (dpx - rho * cj * dux);
        const double L5 = upc <= 0.0 ? 0.0 : upc * (dpx + rho * cj * dux);
        const double L2 = u * (dpx - cj2 * drx), L3 = u * dvx, L4 = u * dwx;
        const double pt = 0.5 * (L5 + L1);
        const double rt = ORC_PLANT == 4 ? (pt - L2) / cj2 : -(pt + L2) / cj2;
        const double ut = 0.5 * (L5 - L1) / (rho * cj), vt = L3, wt = L4;
        const double cpav = cp * w->tmp[p];
        K[0][p] = rt;
        K[1][p] = rho * ut + u * rt;
        K[2][p] = rho * vt + v * rt;
        K[3][p] = rho * wt + ww * rt;
        K[4][p] = cpav * (pt / (cj * cj) - rt) + (Un[4][p] + pr) / rho * rt + rho * (u * ut + v * vt + ww * wt);
      }
  }
  if (c->face[1][1] == ORC_FACE_CHARAC) {
    const int j = d->ny;
    const double iy = g->idy_v[j - 1 + NG];
    /* i = ndx_c..nfx_c: one row in from a characteristic or inlet i-face
     * (mpi_index_bounds.f90:298-360) */
    const walls_t Wc = chars_of(c);
    const int i0 = Wc.lo[0] ? 2 : 1, i1 = d->nx - (Wc.hi[0] ? 1 : 0);
    (void)B;
#pragma omp parallel for collapse(2)
    for (int k = 1; k <= d->nz; k++)
      for (int i = i0; i <= i1; i++) {
        const long p = (long)ID(*d, i, j, k);
        double dd[5];
        const double *f[5] = {w->uu, w->vv, w->ww, w->prs, Un[0]};
        for (int q = 0; q < 5; q++)
          dd[q] = ((-A04[5]) * f[q][p - 4 * sy] + (-A04[4]) * f[q][p - 3 * sy] + (-A04[3]) * f[q][p - 2 * sy] +
                   (-A04[2]) * f[q][p - sy] + (-A04[1]) * f[q][p]) * iy;
        const double duy = dd[0], dvy = dd[1], dwy = dd[2], dpy = dd[3], dry = dd[4];
        const double rho = Un[0][p], u = w->uu[p], v = w->vv[p], ww = w->ww[p], pr = w->prs[p];
        const double cc = sqrt(c->gam * c->rg * w->tmp[p]); /* c2calc_tro */
        double L1 = 0.0, L2, L3, L4;
        double L5 = (ORC_PLANT == 6 ? v : v + cc) * (dpy + rho * cc * dvy);
        if (v <= 0.0) {
          L2 = L3 = L4 = 0.0;
          if (v <= -cc) L5 = 0.0;
        } else {
          L2 = v * duy;
          L3 = v * (dpy - (cc * cc) * dry);
          L4 = v * dwy;
          if (v >= cc) L1 = (v - cc) * (dpy - rho * cc * dvy);
        }
        const double pt = 0.5 * (L5 + L1);
        const double rt = (pt + L3) / (cc * cc);
        const double ut = L2, vt = (L5 - L1) / (2.0 * rho * cc), wt = L4;
        const double cpav = cp * w->tmp[p];
        K[0][p] = rt;
        K[1][p] = rho * ut + u * rt;
        K[2][p] = rho * vt + v * rt;
        K[3][p] = rho * wt + ww * rt;
        K[4][p] = cpav * (pt / (cc * cc) - rt) + (Un[4][p] + pr) / rho * rt + rho * (u * ut + v * vt + ww * wt);
      }
  }
}

/* Supersonic inlet (bc_supersonic_inlet without eigenmodes or RFM,
 * mod_charac.f90:1133-1146): the conserved state of row i = 1 is the inflow
 * state Uin (the row of the initial field, BC_face(1,1)%Uref(:,:,7:10) of
 * init_BC_ref with is_BCref_init, mod_init_flow.f90:418-440).  The primitives
 * the caller computed before keep their values. */
static void bc_inlet(const orc_cfg *c, const dims_t *d, double *const Un[5], double *const Uin[5]) {
  if (c->face[0][0] != ORC_FACE_INLET) return;
  for (int k = 1; k <= d->nz; k++)
    for (int j = 1; j <= d->ny; j++) {
      const long p = (long)ID(*d, 1, j, k);
      for (int v = 0; v < 5; v++) Un[v][p] = Uin[v][p];
    }
}

/* ------------------------------------------------------------------------ */
/* Public entry points                                                       */
/* ------------------------------------------------------------------------ */
/* walls_ok: the routine reads the state one block at a time with its ghosts
 * as given (diagnostics, CFL): wall faces are allowed; the time step itself
 * runs walled blocks through orc_step_multi. */
static int check_cfg_faces(const orc_cfg *c, int walls_ok) {
  if (c->nx < NG || c->ny < NG || c->nz < NG) return -1;
  for (int dir = 0; dir < 3; dir++)
    for (int s = 0; s < 2; s++)
      if (c->face[dir][s] != ORC_FACE_PERIODIC &&
          !(walls_ok && (c->face[dir][s] == ORC_FACE_WALL || c->face[dir][s] == ORC_FACE_SLIP)))
        return -2;
  if (c->filter >= 2 && !c->visc) return -3;              /* sensor / sound speed of primitives_visc */
  if (c->filter == 3 && c->curv == 1) return -8;          /* artvisc_o9: Cartesian or 2.5-D */
  if (c->filter < 0 || c->filter > 4) return -11;
  if (c->curv < 0 || c->curv > 2) return -9;
  if (c->nrk < 1 || c->nrk > 8) return -4;
  return 0;
}

static int check_cfg(const orc_cfg *c) { return check_cfg_faces(c, 0); }

static int work_alloc(const dims_t *d, work_t *w, int visc) {
  memset(w, 0, sizeof(*w));
  w->uu = alloc_arr(d); w->vv = alloc_arr(d); w->ww = alloc_arr(d);
  w->prs = alloc_arr(d); w->tmp = alloc_arr(d);
  if (!w->uu || !w->vv || !w->ww || !w->prs || !w->tmp) return -1;
  if (visc) {
    w->cs = alloc_arr(d); w->visc = alloc_arr(d); w->cok = alloc_arr(d);
    if (!w->cs || !w->visc || !w->cok) return -1;
    for (int q = 0; q < 12; q++) {
      w->grad[q] = alloc_arr(d);
      if (!w->grad[q]) return -1;
    }
  }
  return 0;
}

static void work_free(work_t *w) {
  free(w->uu); free(w->vv); free(w->ww); free(w->prs); free(w->tmp);
  free(w->cs); free(w->visc); free(w->cok);
  for (int q = 0; q < 12; q++) free(w->grad[q]);
}

void orc_euler_rhs(const orc_cfg *c, const orc_grid *g, double *const U[5], double *K[5]) {
  dims_t d = mkdims(c);
  work_t w;
  if (work_alloc(&d, &w, 0)) { work_free(&w); return; }
  prim_euler(c, &d, U, &w);
  euler_rhs_w(c, g, &d, U, &w, K);
  work_free(&w);
}

void orc_visc_rhs(const orc_cfg *c, const orc_grid *g, double *const Uref[5], double *K[5]) {
  dims_t d = mkdims(c);
  work_t w;
  if (work_alloc(&d, &w, 1)) { work_free(&w); return; }
  prim_visc(c, &d, Uref, &w);
  gradients(c, g, &d, &w);
  for (int q = 0; q < 12; q++) orc_fill_periodic(c, w.grad[q]);
  visc_rhs_w(c, g, &d, &w, K);
  work_free(&w);
}

void orc_filter_rhs(const orc_cfg *c, const orc_grid *g, double *const Un[5], double *const Uref[5],
                    double *K[5]) {
  (void)g;
  dims_t d = mkdims(c);
  if (c->filter == 1) {
    filter_sf(c, &d, Un, K);
  } else if (c->filter == 2 || c->filter == 4) {
    work_t w;
    if (work_alloc(&d, &w, 1)) { work_free(&w); return; }
    prim_visc(c, &d, Uref, &w);
    gradients(c, g, &d, &w);
    for (int q = 0; q < 12; q++) orc_fill_periodic(c, w.grad[q]);
    filter_shock(c, g, &d, Un, &w, K);
    work_free(&w);
  } else if (c->filter == 3) {
    work_t w;
    if (work_alloc(&d, &w, 1)) { work_free(&w); return; }
    prim_visc(c, &d, Uref, &w);
    artvisc(c, g, &d, Un, &w, K);
    work_free(&w);
  }
}

int orc_step(const orc_cfg *c, const orc_grid *g, double *U[5], int nsteps) {
  int rc = check_cfg(c);
  if (rc) return rc;
  dims_t d = mkdims(c);
  work_t w;
  double *Un[5], *K[5];
  for (int v = 0; v < 5; v++) { Un[v] = alloc_arr(&d); K[v] = alloc_arr(&d); }
  if (work_alloc(&d, &w, c->visc)) rc = -10;
  for (int v = 0; v < 5; v++) if (!Un[v] || !K[v]) rc = -10;
  if (rc) goto out;

  for (int step = 0; step < nsteps; step++) {
    /* U_n = U (solver.f90:199-203 / end of previous step runge_kutta.f90:146-150) */
    for (int v = 0; v < 5; v++) memcpy(Un[v], U[v], d.n * sizeof(double));
    for (int irk = 1; irk <= c->nrk; irk++) {
      prim_euler(c, &d, Un, &w);                      /* primitives(U_n) */
      euler_rhs_w(c, g, &d, Un, &w, K);               /* flux_euler       */
      if (irk == c->nrk && c->visc) {
        prim_visc(c, &d, U, &w);                      /* primitives_visc(U^n) */
        gradients(c, g, &d, &w);                      /* grad_vel, grad_T */
        for (int q = 0; q < 12; q++) orc_fill_periodic(c, w.grad[q]); /* communication_v */
        visc_rhs_w(c, g, &d, &w, K);                  /* flux_visc        */
      }
      const double alpha = c->dt * c->crk[irk - 1];   /* update_var, update.f90:32-44 */
#pragma omp parallel for collapse(2)
      for (int k = 1; k <= d.nz; k++)
        for (int j = 1; j <= d.ny; j++)
          for (int i = 1; i <= d.nx; i++) {
            long p = (long)ID(d, i, j, k);
            for (int v = 0; v < 5; v++) Un[v][p] = U[v][p] - alpha * K[v][p];
          }
      for (int v = 0; v < 5; v++) orc_fill_periodic(c, Un[v]); /* communication */
    }
    if (c->filter) {
      /* num_dissip then update_varn (runge_kutta.f90:113-126; update.f90:538-565) */
      if (c->filter == 1) filter_sf(c, &d, Un, K);
      else if (c->filter == 2 || c->filter == 4) filter_shock(c, g, &d, Un, &w, K); /* sensor on p, grads of U^n */
      else artvisc(c, g, &d, Un, &w, K);               /* |u|+c of U^n */
      /* update_varn_filter adds K, update_varn_artvisc deltat*K (update.f90:538-608) */
      const int av = c->filter >= 3;
      const double fk = c->dt;
#pragma omp parallel for collapse(2)
      for (int k = 1; k <= d.nz; k++)
        for (int j = 1; j <= d.ny; j++)
          for (int i = 1; i <= d.nx; i++) {
            long p = (long)ID(d, i, j, k);
            for (int v = 0; v < 5; v++) Un[v][p] = av ? Un[v][p] + fk * K[v][p] : Un[v][p] + K[v][p];
          }
      for (int v = 0; v < 5; v++) orc_fill_periodic(c, Un[v]);
    }
    for (int v = 0; v < 5; v++) memcpy(U[v], Un[v], d.n * sizeof(double)); /* U = U_n */
  }
out:
  for (int v = 0; v < 5; v++) { free(Un[v]); free(K[v]); }
  work_free(&w);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* Multiblock: interface copies with orientation, and the step over blocks.  */
/* ------------------------------------------------------------------------ */
static void face_axes(int f, int *a, int *t1, int *t2) {
  *a = f / 2;
  *t1 = *a == 0 ? 1 : 0;
  *t2 = *a == 2 ? 1 : 2;
}

void orc_fill_interfaces(int nblocks, const orc_block *bl, double *const *arrs) {
  for (int b = 0; b < nblocks; b++) {
    const orc_block *A = &bl[b];
    const int na[3] = {A->nx, A->ny, A->nz};
    const long sA[3] = {1, A->nx + 2 * NG, (long)(A->nx + 2 * NG) * (A->ny + 2 * NG)};
    for (int f = 0; f < 6; f++) {
      if (A->wall[f]) continue;
      const orc_block *B = &bl[A->nbr_block[f]];
      const int fb = A->nbr_face[f], o = A->orient[f];
      const int nb[3] = {B->nx, B->ny, B->nz};
      const long sB[3] = {1, B->nx + 2 * NG, (long)(B->nx + 2 * NG) * (B->ny + 2 * NG)};
      int a, t1, t2, bb, u1, u2;
      face_axes(f, &a, &t1, &t2);
      face_axes(fb, &bb, &u1, &u2);
      const int side = f % 2, sideb = fb % 2;
      double *dst = arrs[b];
      const double *src = arrs[A->nbr_block[f]];
#pragma omp parallel for collapse(2)
      for (int qq = 1; qq <= na[t2]; qq++)
        for (int pp = 1; pp <= na[t1]; pp++) {
          int x1 = (o & 1) ? qq : pp, x2 = (o & 1) ? pp : qq;
          if (o & 2) x1 = nb[u1] + 1 - x1;
          if (o & 4) x2 = nb[u2] + 1 - x2;
          for (int l = 1; l <= NG; l++) {
            const int ia = side == 0 ? 1 - l : na[a] + l;
            const int ib = sideb == 0 ? l : nb[bb] + 1 - l;
            const long pa = (long)(ia + NG - 1) * sA[a] + (long)(pp + NG - 1) * sA[t1] + (long)(qq + NG - 1) * sA[t2];
            const long pb = (long)(ib + NG - 1) * sB[bb] + (long)(x1 + NG - 1) * sB[u1] + (long)(x2 + NG - 1) * sB[u2];
            dst[pa] = src[pb];
          }
        }
    }
  }
}

int orc_step_multi(const orc_cfg *c0, int nb, const orc_block *bl, const orc_grid *g, double *const *U, int nsteps) {
  if (nb < 1 || nb > 64) return -1;
  if (c0->filter >= 2 && !c0->visc) return -3;
  if (c0->filter == 3 && c0->curv == 1) return -8;
  if (c0->filter < 0 || c0->filter > 4) return -11;
  for (int b = 0; b < nb; b++)
    for (int f = 0; f < 6; f++)
      if (bl[b].wall[f]) {
        const int ext = f < 2 ? bl[b].nx : (f < 4 ? bl[b].ny : bl[b].nz);
        /* boundaries: Cartesian blocks; c3 blocks take no-slip walls on their
         * j-faces, with interfaces or periodic faces on i and k */
        if (c0->curv == 2) return -5;
        const int wk = bl[b].wall[f];
        const int is_w = wk == 1 || wk == 4;
        if (c0->curv == 1 && (!is_w || f < 2 || f > 3 || bl[b].wall[0] || bl[b].wall[1] ||
                              bl[b].wall[4] || bl[b].wall[5]))
          return -5;
        if (ext < 10) return -7;        /* closures of the two sides must not overlap */
        if (wk < 1 || wk > 4) return -12;
        /* characteristic outflow at imax or jmax, supersonic inlet at imin
         * (the faces config. 4 uses); walls on the i-faces, or advanced walls
         * (is_wall2), together with a characteristic face are not covered */
        if (wk == 2 && f != 1 && f != 3) return -12;
        if (wk == 3 && f != 0) return -12;
        if ((wk == 2 || wk == 3) &&
            (c0->wall2 || bl[b].wall[0] == 1 || bl[b].wall[1] == 1 || bl[b].wall[0] == 4 || bl[b].wall[1] == 4))
          return -12;
      }
  orc_cfg cb[64];
  dims_t d[64];
  work_t w[64];
  double *Un[64][5], *K[64][5];
  /* characteristic faces: the module arrays the reference keeps between calls
   * of its boundary routines -- the sound speed primitives_visc last computed
   * (c_, read by bc_charac_imax), the primitives of row i = 1 from the last
   * `primitives` before the inlet reset (the flux at i = 1 of stage 1 reads
   * them), the inflow state -- and K itself, whose rows of characteristic and
   * inlet faces only the boundary routines and the filter write */
  double *cst[64], *pst[64][5], *Uin[64][5];
  int chb[64];
  int rc = 0;
  for (int b = 0; b < nb; b++) {
    cb[b] = *c0;
    cb[b].nx = bl[b].nx;
    cb[b].ny = bl[b].ny;
    cb[b].nz = bl[b].nz;
    static const int kinds[5] = {ORC_FACE_INTERFACE, ORC_FACE_WALL, ORC_FACE_CHARAC, ORC_FACE_INLET, ORC_FACE_SLIP};
    for (int f = 0; f < 6; f++) cb[b].face[f / 2][f % 2] = kinds[bl[b].wall[f]];
    if (bl[b].nx < NG || bl[b].ny < NG || bl[b].nz < NG) return -1;
    /* c3 walls need their coefficient slabs, slip ones also the dual tangents */
    if (cb[b].curv == 1)
      for (int f = 2; f < 4; f++)
        if (bl[b].wall[f] && (!g[b].wc3[f][0] || (slip_face(&cb[b], f) && !g[b].wc3[f][8]))) return -13;
    d[b] = mkdims(&cb[b]);
    chb[b] = chars_of(&cb[b]).any;
    cst[b] = NULL;
    for (int v = 0; v < 5; v++) {
      Un[b][v] = alloc_arr(&d[b]);
      K[b][v] = alloc_arr(&d[b]);
      pst[b][v] = chb[b] ? alloc_arr(&d[b]) : NULL;
      Uin[b][v] = chb[b] ? alloc_arr(&d[b]) : NULL;
      if (!Un[b][v] || !K[b][v] || (chb[b] && (!pst[b][v] || !Uin[b][v]))) rc = -10;
    }
    if (chb[b] && !(cst[b] = alloc_arr(&d[b]))) rc = -10;
    if (work_alloc(&d[b], &w[b], 1)) rc = -10;
  }
  double *tab[64];
  if (rc) goto out;
  /* start_runge_kutta (runge_kutta.f90:316-342) on blocks with characteristic
   * faces: the increments of their rows from the primitives of U, the sound
   * speed of U (the one primitives_visc gave at set-up), the inflow state =
   * the first row of U */
  for (int b = 0; b < nb; b++) {
    if (!chb[b]) continue;
    prim_visc(&cb[b], &d[b], &U[5 * b], &w[b]);
    memcpy(cst[b], w[b].cs, d[b].n * sizeof(double));
    for (int v = 0; v < 5; v++) memcpy(Uin[b][v], U[5 * b + v], d[b].n * sizeof(double));
    prim_euler(&cb[b], &d[b], &U[5 * b], &w[b]);
    bc_charac(&cb[b], &g[b], &d[b], &U[5 * b], &w[b], cst[b], K[b]);
    double *pw[5] = {w[b].uu, w[b].vv, w[b].ww, w[b].prs, w[b].tmp};
    for (int q = 0; q < 5; q++) memcpy(pst[b][q], pw[q], d[b].n * sizeof(double));
  }
  for (int step = 0; step < nsteps; step++) {
    for (int b = 0; b < nb; b++)
      for (int v = 0; v < 5; v++) memcpy(Un[b][v], U[5 * b + v], d[b].n * sizeof(double));
    for (int irk = 1; irk <= c0->nrk; irk++) {
      const int last = irk == c0->nrk;
      /* bc_wall at the start of stages 2..nrk, from the previous stage's primitives */
      if (irk > 1)
        for (int b = 0; b < nb; b++) bc_wall(&cb[b], &g[b], &d[b], Un[b], &w[b], K[b]);
      if (last && c0->visc) {
        for (int b = 0; b < nb; b++) {
          prim_visc(&cb[b], &d[b], &U[5 * b], &w[b]);
          gradients(&cb[b], &g[b], &d[b], &w[b]);
        }
        for (int q = 0; q < 12; q++) {
          for (int b = 0; b < nb; b++) tab[b] = w[b].grad[q];
          orc_fill_interfaces(nb, bl, tab);
        }
      }
      for (int b = 0; b < nb; b++) {
        /* the primitives of the stage state (`primitives`); stage 1 keeps the
         * end-of-step ones, which differ from those of U^n only on an inlet row */
        work_t we = w[b];
        double *tmp5[5] = {NULL, NULL, NULL, NULL, NULL};
        if (last && c0->visc) {
          /* primitives of U_n for the Euler part, primitives_visc of U^n kept for the sensor */
          for (int q = 0; q < 5; q++) tmp5[q] = alloc_arr(&d[b]);
          we.uu = tmp5[0]; we.vv = tmp5[1]; we.ww = tmp5[2]; we.prs = tmp5[3]; we.tmp = tmp5[4];
        }
        prim_euler(&cb[b], &d[b], Un[b], &we);
        if (chb[b]) {
          double *pw[5] = {we.uu, we.vv, we.ww, we.prs, we.tmp};
          if (irk == 1 && cb[b].face[0][0] == ORC_FACE_INLET)
            for (int k = 1; k <= d[b].nz; k++)
              for (int j = 1; j <= d[b].ny; j++) {
                const long p = (long)ID(d[b], 1, j, k);
                for (int q = 0; q < 5; q++) pw[q][p] = pst[b][q][p];
              }
          /* bc_non_reflecting after `primitives` (runge_kutta.f90:30-37, 64-70) */
          if (irk > 1) {
            bc_charac(&cb[b], &g[b], &d[b], Un[b], &we, cst[b], K[b]);
            bc_inlet(&cb[b], &d[b], Un[b], Uin[b]);
          }
        }
        euler_rhs_w(&cb[b], &g[b], &d[b], Un[b], &we, K[b]);
        for (int q = 0; q < 5; q++) free(tmp5[q]);
        if (last && c0->visc) {
          visc_rhs_w(&cb[b], &g[b], &d[b], &w[b], K[b]);
        } else if (last) {
          /* the last stage calls primitives_visc(U^n) whether or not the case is
           * viscous (runge_kutta.f90:71; only grad_vel and flux_visc are void
           * without viscosity): the primitives bc_wall finds at the end of the
           * step are those of U^n */
          prim_visc(&cb[b], &d[b], &U[5 * b], &w[b]);
        }
        /* primitives_visc's sound speed of U^n: bc_charac_imax reads it until the
         * next step's last stage */
        if (last && chb[b]) memcpy(cst[b], w[b].cs, d[b].n * sizeof(double));
        const double alpha = c0->dt * c0->crk[irk - 1];
        const dims_t dd = d[b];
#pragma omp parallel for collapse(2)
        for (int k = 1; k <= dd.nz; k++)
          for (int j = 1; j <= dd.ny; j++)
            for (int i = 1; i <= dd.nx; i++) {
              long p = (long)ID(dd, i, j, k);
              for (int v = 0; v < 5; v++) Un[b][v][p] = U[5 * b + v][p] - alpha * K[b][v][p];
            }
      }
      for (int v = 0; v < 5; v++) {
        for (int b = 0; b < nb; b++) tab[b] = Un[b][v];
        orc_fill_interfaces(nb, bl, tab);
      }
    }
    if (c0->filter) {
      for (int b = 0; b < nb; b++) {
        if (c0->filter == 1) filter_sf(&cb[b], &d[b], Un[b], K[b]);
        else if (c0->filter == 2 || c0->filter == 4) filter_shock(&cb[b], &g[b], &d[b], Un[b], &w[b], K[b]);
        else artvisc(&cb[b], &g[b], &d[b], Un[b], &w[b], K[b]);
        const dims_t dd = d[b];
        const int av = c0->filter >= 3;
        const double fk = c0->dt;
#pragma omp parallel for collapse(2)
        for (int k = 1; k <= dd.nz; k++)
          for (int j = 1; j <= dd.ny; j++)
            for (int i = 1; i <= dd.nx; i++) {
              long p = (long)ID(dd, i, j, k);
              for (int v = 0; v < 5; v++)
                Un[b][v][p] = av ? Un[b][v][p] + fk * K[b][v][p] : Un[b][v][p] + K[b][v][p];
            }
      }
      for (int v = 0; v < 5; v++) {
        for (int b = 0; b < nb; b++) tab[b] = Un[b][v];
        orc_fill_interfaces(nb, bl, tab);
      }
    }
    /* bc_wall at the end of the step (runge_kutta.f90:130-136): from the primitives
     * of U^n, which the last stage's primitives_visc wrote */
    for (int b = 0; b < nb; b++) bc_wall(&cb[b], &g[b], &d[b], Un[b], &w[b], K[b]);
    /* then `primitives` and bc_non_reflecting (:137-139): the increments of the
     * characteristic rows for the next step's first stage, the inlet reset */
    for (int b = 0; b < nb; b++) {
      if (!chb[b]) continue;
      prim_euler(&cb[b], &d[b], Un[b], &w[b]);
      bc_charac(&cb[b], &g[b], &d[b], Un[b], &w[b], cst[b], K[b]);
      double *pw[5] = {w[b].uu, w[b].vv, w[b].ww, w[b].prs, w[b].tmp};
      for (int q = 0; q < 5; q++) memcpy(pst[b][q], pw[q], d[b].n * sizeof(double));
      bc_inlet(&cb[b], &d[b], Un[b], Uin[b]);
    }
    for (int b = 0; b < nb; b++)
      for (int v = 0; v < 5; v++) memcpy(U[5 * b + v], Un[b][v], d[b].n * sizeof(double));
  }
out:
  for (int b = 0; b < nb; b++) {
    for (int v = 0; v < 5; v++) {
      free(Un[b][v]);
      free(K[b][v]);
      free(pst[b][v]);
      free(Uin[b][v]);
    }
    free(cst[b]);
    work_free(&w[b]);
  }
  return rc;
}

/* ------------------------------------------------------------------------ */
/* TGV statistics of the state U (stats_tgv, src/File_IO/stats.f90:37-112): */
/* sums over the interior of rho, 1/2 rho |u|^2 and 1/2 rho |curl u|^2, the */
/* velocity gradients by the 5-pt scheme of grad_vel (with the viscous      */
/* metrics).  The reference sums gradients its last stage computed; here    */
/* they are those of U itself, whose ghosts must be valid.                  */
/* ------------------------------------------------------------------------ */
int orc_diag(const orc_cfg *c, const orc_grid *g, double *const U[5], double out[4]) {
  int rc = check_cfg_faces(c, 1);
  if (rc) return rc;
  dims_t d = mkdims(c);
  work_t w;
  memset(&w, 0, sizeof w);
  if (work_alloc(&d, &w, 1)) return -10;
  prim_euler(c, &d, U, &w);
  gradients(c, g, &d, &w);
  double mg = 0.0, ek = 0.0, wr = 0.0, um = 0.0;
  for (int k = 1; k <= d.nz; k++)
    for (int j = 1; j <= d.ny; j++)
      for (int i = 1; i <= d.nx; i++) {
        long p = (long)ID(d, i, j, k);
        double r = U[0][p];
        double *const *gr = w.grad;
        mg += r;
        ek += 0.5 * r * (w.uu[p] * w.uu[p] + w.vv[p] * w.vv[p] + w.ww[p] * w.ww[p]);
        double o1 = gr[DWY][p] - gr[DVZ][p], o2 = gr[DUZ][p] - gr[DWX][p], o3 = gr[DVX][p] - gr[DUY][p];
        wr += 0.5 * r * (o1 * o1 + o2 * o2 + o3 * o3);
        /* u_max of timestep_new.f90:270-288: |u| + c */
        double q2 = w.uu[p] * w.uu[p] + w.vv[p] * w.vv[p] + w.ww[p] * w.ww[p];
        double cs = sqrt(c->gam * w.prs[p] / r);
        double s = sqrt(q2) + cs;
        if (s > um) um = s;
      }
  out[0] = mg;
  out[1] = ek;
  out[2] = wr;
  out[3] = um;
  work_free(&w);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Maximum CFL number per direction (check_max_cfl,                          */
/* src/Time/timestep_new.f90:660-842): over the interior, the local           */
/* characteristic speed times deltat -- Cartesian (|u_d| + c) deltat idx_d,   */
/* c3 (|u.grad xi_d| |1/J| ... as |(u xi_x + v xi_y + w xi_z) ijacob3| +      */
/* c g3_xi with g3_xi = |grad xi| |ijacob3|) times deltat, and on a 2.5-D     */
/* block the same with the 2-D metrics in i and j and Cartesian k.  c is the  */
/* sound speed of primitives_visc, sqrt(gam rg T).                             */
/* ------------------------------------------------------------------------ */
int orc_max_cfl(const orc_cfg *c, const orc_grid *g, double *const U[5], double out[3]) {
  int rc = check_cfg_faces(c, 1);
  if (rc) return rc;
  dims_t d = mkdims(c);
  work_t w;
  memset(&w, 0, sizeof w);
  if (work_alloc(&d, &w, 1)) return -10;
  prim_visc(c, &d, U, &w);
  for (int dir = 0; dir < 3; dir++) {
    const double *vel = dir == 0 ? w.uu : (dir == 1 ? w.vv : w.ww);
    const double *met = dir == 0 ? g->idx : (dir == 1 ? g->idy : g->idz);
    const int curv = c->curv == 1 || (c->curv == 2 && dir < 2);
    double cm = 0.0;
    for (int k = 1; k <= d.nz; k++)
      for (int j = 1; j <= d.ny; j++)
        for (int i = 1; i <= d.nx; i++) {
          const long p = (long)ID(d, i, j, k);
          double cfl;
          if (curv) {
            const double mx = mtr(c, g->m, g->m2, &d, 3 * dir, p), my = mtr(c, g->m, g->m2, &d, 3 * dir + 1, p),
                         mz = mtr(c, g->m, g->m2, &d, 3 * dir + 2, p), ij = mtr(c, g->m, g->m2, &d, 9, p);
            const double val = fabs((w.uu[p] * mx + w.vv[p] * my + w.ww[p] * mz) * ij) +
                               w.cs[p] * (sqrt(mx * mx + my * my + mz * mz) * fabs(ij));
            cfl = val * c->dt;
          } else {
            const int r = dir == 0 ? i : (dir == 1 ? j : k);
            cfl = (fabs(vel[p]) + w.cs[p]) * c->dt * met[r - 1 + NG];
          }
          cm = fmax(cm, cfl);
        }
    out[dir] = cm;
  }
  work_free(&w);
  return 0;
}
