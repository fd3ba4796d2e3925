/*
 * musica_oracle.h -- CPU restatement of MUSICAA's explicit RK time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP
 * product path: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  Nothing in musica2_newgrid_amd/ links or
 * calls it.
 *
 * Every function restates the reference's algorithm, citing the file:line
 * of LouenasZm/musica2_newGrid (src/...) that it follows.  The reference is
 * Fortran+MPI; its hot-path subset cannot be built in this image without
 * stand-ins (an `mpi` module for flang), so there is no oracle/_ref build:
 * these routines are pinned by closed forms and invariants in tests/.
 *
 * Data layout (same as the reference's module arrays, src/Main/alloc.f90:116):
 * one contiguous array per variable, Fortran order (i fastest), extents
 * (1-ng:nx+ng, 1-ng:ny+ng, 1-ng:nz+ng) with ng = 5 ghost layers.
 */
#ifndef MUSICA_ORACLE_H
#define MUSICA_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NG 5

/* Boundary kind of a block face (restates BC_face(dir,side)%sort,
 * src/Boundary/mod_bc.f90:7-61 for the kinds this path covers). */
enum orc_face_kind {
  ORC_FACE_PERIODIC = 0,   /* neighbour block is the block itself          */
  ORC_FACE_INTERFACE = 1,  /* 1-to-1 interface with another block          */
  ORC_FACE_WALL = 2        /* slip wall, sort=0 with flag 's'              */
};

typedef struct orc_cfg {
  int nx, ny, nz;          /* interior cells                                */
  int curv;                /* 0 Cartesian (idx 1-D metrics), 1 3-D curvilinear (c3) */
  int visc;                /* 1: 5-pt viscous terms in the last RK stage   */
  int filter;              /* 0 none, 1 selective filter, 2 shock filter   */
  int ducros;              /* Ducros sensor in the shock filter            */
  int nrk;                 /* number of RK stages                          */
  int face[3][2];          /* enum orc_face_kind per (direction, side)      */
  int pad_;
  double crk[8];           /* RK stage coefficients                         */
  double dt;               /* time step                                     */
  double gam, rg;          /* ratio of specific heats, gas constant         */
  double c0, S0, diffscale;/* Sutherland law: mu = c0 T^1.5/(T+S0)*diffscale */
  double cok_ratio;        /* cp/Pr (thermal conductivity = cok_ratio*mu)   */
  double chi;              /* filter amplitude (dissip_coeff)               */
  double dissip_shock;     /* shock-capturing amplitude                     */
  double csens;            /* pressure sensor blend                         */
} orc_cfg;

typedef struct orc_grid {
  /* Cartesian: inverse spacings with ng ghosts, length n+2ng (Euler, viscous) */
  const double *idx, *idy, *idz;
  const double *idx_v, *idy_v, *idz_v;
  /* c3: padded 3-D arrays: ksi_x,ksi_y,ksi_z,eta_x,eta_y,eta_z,phi_x,phi_y,phi_z,ijacob3 */
  const double *m[10];
  const double *mv[10];    /* the viscous set (*_v arrays)                 */
} orc_grid;

/* Fill the face ghost layers of a padded array from the periodic image.
 * Restates communication3d for self-neighbour faces
 * (src/Parallel/mod_comm.f90:249-290). */
void orc_fill_periodic(const orc_cfg *c, double *a);

/* K = flux_euler(U): Euler flux divergence of the state U (which carries valid
 * ghosts).  src/Compute_flow/smod_flux_euler_11pts.f90:12-869 (Cartesian),
 * smod_flux_euler_11pts_c3.f90:12-776 (c3).  K is a padded array set. */
void orc_euler_rhs(const orc_cfg *c, const orc_grid *g, double *const U[5], double *K[5]);

/* K += viscous flux divergence of Uref (grad_vel/grad_T + flux_visc). */
void orc_visc_rhs(const orc_cfg *c, const orc_grid *g, double *const Uref[5], double *K[5]);

/* K = numerical dissipation increment of Un (selective filter or shock
 * filter; the shock filter takes its sensor from Uref). */
void orc_filter_rhs(const orc_cfg *c, const orc_grid *g, double *const Un[5],
                    double *const Uref[5], double *K[5]);

/* Advance U (interior + valid ghosts) by nsteps full RK steps
 * (src/Compute_flow/runge_kutta.f90:2-152).  Returns 0 on success. */
int orc_step(const orc_cfg *c, const orc_grid *g, double *U[5], int nsteps);

/* Number of OpenMP threads the oracle uses. */
int orc_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
