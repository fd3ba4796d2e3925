/*
 * musica_gpu.h -- C-ABI of the MI355X (gfx950) time-advance library
 * libmusica_gpu.so.
 *
 * The reference (LouenasZm/musica2_newGrid, Fortran) advances one explicit
 * low-storage Runge-Kutta step through a table of no-argument procedure
 * pointers bound at start-up:
 *   - pointer table   src/Main/mod_interface.f90:10-13,63-166
 *   - binding         src/Main/assign_procedures.f90:109-411
 *   - the step        src/Compute_flow/runge_kutta.f90:2-152, called from
 *                     src/Compute_flow/solver.f90:415-417
 * All data flows through module globals (src/Compute_flow/mod_flow.f90:13-41,
 * src/Grid/mod_grid.f90:131-145).  This header replaces that seam with one
 * handle that owns the device copy of those globals and advances whole steps,
 * so that the Euler, viscous, filter and halo passes can be fused on the GPU.
 *
 * Conventions
 *   - plain pointers and sizes only; every host array is borrowed for the
 *     duration of the call;
 *   - host arrays are the reference's Fortran module arrays: one contiguous
 *     array per variable, i fastest, extents (1-5:nx+5, 1-5:ny+5, 1-5:nz+5)
 *     (src/Main/alloc.f90:116-120, ngh=5 from src/File_IO/read_param.f90:495);
 *   - return 0 on success, a negative MGPU_E* code on error; the message is
 *     available from mgpu_last_error().  The Fortran host maps an error to
 *     mpistop(msg,1) (src/Utils/warnstop.f90:20-90);
 *   - one host thread per handle; handles are independent.
 */
#ifndef MUSICA_GPU_H
#define MUSICA_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MGPU_NGH 5 /* ghost layers of state and metric arrays */

/* error codes */
#define MGPU_OK 0
#define MGPU_EINVAL -1   /* invalid argument or case description        */
#define MGPU_EDEVICE -2  /* HIP runtime error                           */
#define MGPU_ENOMEM -3   /* device allocation failed                    */
#define MGPU_EUNSUP -4   /* feature combination not supported           */
#define MGPU_ESTATE -5   /* call out of order (e.g. step before metrics) */

/* Face kinds: restate BC_face(dir,side)%sort (src/Boundary/mod_bc.f90:7-61)
 * for the kinds this library advances.  A periodic face is the reference's
 * neighbour-is-itself interface (src/Parallel/mod_mpi_part.f90:333-). */
/* Geometry of a case (mgpu_case.curvilinear) */
#define MGPU_GEOM_CART 0 /* Cartesian, 1-D metrics idx, idy, idz                     */
#define MGPU_GEOM_C3 1   /* is_curv3: 3-D metrics at every point (flux_euler_11pts_c3) */
#define MGPU_GEOM_C2 2   /* is_curv: 2-D metrics in (i, j), Cartesian z (flux_euler_11pts_c, */
                         /* src/Compute_flow/smod_flux_euler_11pts_c.f90)             */

#define MGPU_FACE_PERIODIC 0
#define MGPU_FACE_INTERFACE 1
/* Wall, BC_face%sort = 0.  With is_wall2 = F (the reference's default,
 * src/File_IO/read_param.f90:75) a no-slip wall: bc_wall_*_adiabatic_dpdn
 * or _isotherm_dpdn (src/Boundary/smod_bc_wall.f90, smod_bc_wall_c3.f90)
 * with the boundary-row closures of the Euler, viscous and filter operators.
 * With mgpu_case.is_wall2 (type_wall_bc = 2) the wall rows are advanced
 * (mpi_index_bounds.f90:28-70: the interior derivatives along the wall, the
 * 3-point one-sided a02 / a20 across it, smod_flux_euler_11pts.f90:42-54,
 * 286-297) and take the states of smod_bc_wall_2*.f90: no slip (zero
 * momentum, rhoE = rho cv T_wall with the advanced rho), or slip for every
 * wall of an inviscid case (see MGPU_FACE_SLIP).  On the faces of one axis
 * per block (i, j or k; a block with walls on its i- or k-faces is stored
 * with that axis exchanged with j, invisibly to the caller) or of two axes
 * of Cartesian blocks (walls on the j- and k-faces: the block stored with its
 * j, k, i as i, j, k, the k-faces' state winning on the edges as in bc_wall's
 * face order), or on the faces of one axis of 3-D curvilinear blocks (walls
 * on their i- or k-faces: the block stored with that index axis exchanged
 * with j, its metric rows with it; momentum and tau stay in x, y, z)
 * (FP64 or mixed, any filter), at least 10 cells between the walls;
 * the block's extrapolation coefficients come from mgpu_set_wall
 * (Cartesian) or mgpu_set_wall_c3 (curvilinear).
 * With a shock sensor (filters 2, 4) the ghost row beyond a wall must hold a
 * valid state:
 * the wall row's pressure sensor reads p there (mod_filtering_shock.f90:
 * 243-260 with ndy_d1 = 1), and nothing on the path writes it. */
#define MGPU_FACE_WALL 2
/* Characteristic (LODI) outflow, BC_face%sort = -3 at imax or jmax:
 * bc_charac_imax / bc_charac_jmax (src/Boundary/mod_charac.f90:350-653,
 * 816-969) for a perfect gas.  The boundary row gets no Euler or viscous
 * increment; its increment comes from one-sided 5-pt derivatives of the
 * stage state's primitives, computed after `primitives` in stages 2..nrk and
 * at the end of the step (runge_kutta.f90:30-37, 64-70, 137-139), and kept
 * for the next stage.  At imax the sound speed is the one primitives_visc
 * last computed (U^n of the previous step's last stage), and a subsonic
 * outflow relaxes p towards 0.528*p_ref (mgpu_case.p_ref). */
#define MGPU_FACE_CHARAC 3
/* Supersonic inlet, BC_face%sort = -3 at imin: bc_supersonic_inlet without
 * eigenmodes (mod_charac.f90:1133-1146).  Row 1 is held at the inflow state,
 * which is row 1 of the state given by mgpu_put_state (init_BC_ref with
 * is_BCref_init, mod_init_flow.f90:418-440).
 * Characteristic and inlet faces: Cartesian cases, blocks whose other
 * boundary faces are j-walls, filter NONE, SF or SHOCK, the classic stage
 * kernel, >= 10 cells along a bounded axis.  mgpu_put_state
 * starts a run as start_runge_kutta does (runge_kutta.f90:316-342): the first
 * mgpu_step after it derives the characteristic increments, the sound speed
 * and the inflow state from the uploaded state.  Like walls, the ghost row
 * beyond such a face must hold a valid state when the shock filter runs. */
#define MGPU_FACE_INLET 4
/* Wall with the slip flag 's' (bl%flag = 's', is_slip, src/Boundary/mod_bc.f90:
 * 127-182): with is_wall2 and an adiabatic (or inviscid) case a slip wall,
 * bc_wall_*_slip (smod_bc_wall_2.f90:11-30, 165-185): the advanced wall row
 * loses its wall-normal momentum component and keeps the rest; on a 3-D
 * curvilinear block its momentum becomes its projection on the wall's
 * tangent plane (bc_wall_jmin_slip_c3, smod_bc_wall_2_c3.f90:363-393), which
 * needs mgpu_set_wall_c3_slip.  Otherwise (is_wall2 = F, or an isothermal
 * viscous case) exactly an MGPU_FACE_WALL: the binding of
 * src/Main/assign_procedures.f90:415-530 selects no slip routine there. */
#define MGPU_FACE_SLIP 5

/* Filter selection: is_SF / is_shock (src/Main/assign_procedures.f90:161-219) */
#define MGPU_FILTER_NONE 0
#define MGPU_FILTER_SF 1    /* filtering_11pts, mod_filtering.f90:894      */
#define MGPU_FILTER_SHOCK 2 /* filter_o10_shock, mod_filtering_shock.f90:204 */
#define MGPU_FILTER_ARTVISC 3 /* artvisc_o9 (is_SF = F), mod_artvisc.f90:82: Cartesian or 2.5-D,
                               * iorder_visc = 4 (it reads the sound speed of primitives_visc) */
#define MGPU_FILTER_ARTVISC_SHOCK 4 /* artvisc_o9_shock (is_SF = F, is_shock = T),
                                     * mod_artvisc_shock.f90:113: the shock filter's sensor-weighted
                                     * faces times the spectral radius, U += deltat*K; every geometry,
                                     * iorder_visc = 4 */

/* Arithmetic: FP64 everywhere, or FP32 flux and derivative arithmetic and
 * FP32 intermediate stage states, with FP64 RK accumulation, metrics, viscous
 * terms, shock sensor and filters (the filters read and write FP64 states). */
#define MGPU_PREC_FP64 0
#define MGPU_PREC_MIXED 1

/* Launch flags.  By default every step is replayed from a HIP graph captured
 * on first use (one per U^n buffer), with one batched launch per kernel for
 * all blocks; EAGER launches the same kernels one by one instead. */
#define MGPU_LAUNCH_EAGER 1
/* Stage-kernel family.  The stage kernel exists in two forms with the same
 * arithmetic: the classic one (each halo cell loaded into registers in the
 * plane that uses it) and the pipelined one (the halo of plane k+1 copied
 * straight into LDS while plane k is computed).  By default the autotuner
 * picks among both; these flags restrict it, and the untuned default, to
 * one family. */
#define MGPU_LAUNCH_STAGE_CLASSIC 2
#define MGPU_LAUNCH_STAGE_PIPELINED 4

/* Faces are numbered imin,imax,jmin,jmax,kmin,kmax = 0..5.  An interface
 * face f of block b takes its ghosts from face nbr_face[f] of block
 * nbr_block[f].  The tangential axes of a face are its two other axes in
 * increasing order; orient bit 0 swaps them, bit 1 reverses the neighbour's
 * first tangential axis, bit 2 its second.  Both sides must declare the
 * interface, with inverse orientations (swap'=swap, rev1'=swap?rev2:rev1,
 * rev2'=swap?rev1:rev2). */
typedef struct mgpu_block {
  int32_t nx, ny, nz;     /* interior cells of the block                     */
  int32_t face_kind[6];   /* MGPU_FACE_*                                     */
  int32_t nbr_block[6];   /* neighbour block of an interface face            */
  int32_t nbr_face[6];    /* neighbour face (0..5) of an interface face       */
  int32_t orient[6];      /* interface orientation code (0 = aligned)         */
} mgpu_block;

/* Case description: the param.ini flags of the hot path plus the block table
 * (param_blocks.ini, src/Main/mod_block.f90:109-229). */
typedef struct mgpu_case {
  int32_t nblocks;
  int32_t curvilinear;    /* MGPU_GEOM_*: Cartesian 1-D metrics (idx,idy,idz),  */
                          /* full 3-D (is_curv3) or 2.5-D (is_curv)            */
  int32_t iorder_visc;    /* 0 (Euler) or 4 (5-pt viscous terms)               */
  int32_t filter;         /* MGPU_FILTER_*                                     */
  int32_t is_ducros;      /* Ducros sensor in the shock filter                 */
  int32_t nrk;            /* RK stages (mod_time.f90:121)                       */
  int32_t precision;      /* MGPU_PREC_*                                       */
  int32_t launch_flags;   /* MGPU_LAUNCH_* (0: default)                        */
  const mgpu_block *blocks;
  double crk[8];          /* RK stage coefficients crk(1..nrk)                 */
  double deltat;          /* time step                                         */
  double gam, rg;         /* gamma, gas constant                               */
  double c0, S0;          /* Sutherland: mu = c0*T^1.5/(T+S0)*diffscale        */
  double diffscale;       /* viscosity rescaling (setupref.f90:357)            */
  double cok_ratio;       /* cp/Pr: conductivity = cok_ratio*mu                */
  double dissip_coeff;    /* filter amplitude chi                              */
  double dissip_shock;    /* shock-capturing amplitude                         */
  double csens;           /* pressure-sensor blend                             */
  int32_t is_adiab;       /* walls: 1 adiabatic (dT/dn = 0), 0 isothermal (is_adiab) */
  int32_t is_sw_edoh;     /* shock filter: Edoh et al. switch, each direction's  */
                          /* increment times min(1, CFL_d/chi) with CFL_d the   */
                          /* local (|u_d|+c) dt / dx_d of U^n (contravariant on */
                          /* c3 blocks) (mod_filtering_shock.f90:620-700)       */
  double t_wall;          /* walls: temperature of an isothermal wall (T_wall) */
  double p_ref;           /* characteristic imax outflow: reference pressure of */
                          /* its subsonic relaxation (p_ref, mod_charac.f90:426) */
  int32_t is_wall2;       /* type_wall_bc = 2 (is_wall2, read_param.f90:373):  */
                          /* walls advanced, smod_bc_wall_2*.f90 wall states;   */
                          /* not with characteristic or inlet faces            */
  int32_t reserved0;      /* 0                                                 */
} mgpu_case;

typedef struct mgpu_handle mgpu_handle;

/* Create a handle for one case on the current HIP device: allocates the
 * device state (3 state sets: U^n and two stage buffers), metrics and
 * scratch.  Replaces alloc_tab + assign_procedures for the hot path
 * (src/Main/alloc.f90:2-295, src/Main/assign_procedures.f90:2-979). */
int mgpu_create(const mgpu_case *c, mgpu_handle **out);

/* Cartesian metrics of a block: idx, idy, idz (Euler, 11-pt) and idx_v,
 * idy_v, idz_v (viscous, 5-pt); each of length n+2*MGPU_NGH, ghosts
 * included (src/Grid/mod_grid.f90:60-64, grid_metrics_cart.f90:2-). */
int mgpu_set_metrics_cart(mgpu_handle *h, int block, const double *idx, const double *idy,
                          const double *idz, const double *idx_v, const double *idy_v,
                          const double *idz_v);

/* Curvilinear metrics of a block: m[0..9] = ksi_x,ksi_y,ksi_z,eta_x,eta_y,
 * eta_z,phi_x,phi_y,phi_z,ijacob3 and mv[0..9] the same for the viscous set
 * (*_v), each a padded 3-D array with ghosts
 * (src/Grid/mod_grid.f90:131-145, smod_grid_metrics1_c3.f90:8-484). */
int mgpu_set_metrics_c3(mgpu_handle *h, int block, const double *const *m, const double *const *mv);

/* 2.5-D curvilinear metrics of a block (MGPU_GEOM_C2, the reference's
 * is_curv): m2[0..4] = x_ksi, x_eta, y_ksi, y_eta, ijacob (Euler set, 11-pt)
 * and m2v[0..4] the viscous set (*_v, 5-pt), each a padded 2-D array of
 * (nx+2*MGPU_NGH)*(ny+2*MGPU_NGH) doubles, i fastest; idz, idz_v as in
 * mgpu_set_metrics_cart (src/Grid/mod_grid.f90, grid_metrics_curv.f90). */
int mgpu_set_metrics_c2(mgpu_handle *h, int block, const double *const *m2, const double *const *m2v,
                        const double *idz, const double *idz_v);

/* Wall extrapolation coefficients of a block with MGPU_FACE_WALL faces:
 * cextrp2[f], cextrp3[f] for faces f = 0..5 (entries of faces without a
 * wall are ignored), as init_coeff_wall_cart computes them from the grid
 * (src/Derivatives/mod_coeff_deriv.f90:1099-1160): p_wall = cextrp2*p(2) -
 * cextrp3*p(3) along the wall normal.  Required before the first step of
 * such a block.  The Euler metric (idy) of a wall block holds, on rows 2..5
 * (ny-1..ny-4), the idy2/idy4/idy6/idy8 of grid_metrics_wall_cart
 * (src/Grid/grid_metrics_cart.f90:190-436), which the closure rows use, and
 * on the wall row idy1_jmin = 1/(a02 y) (idy1_jmax), which an advanced wall
 * row (is_wall2) uses (:245-256). */
int mgpu_set_wall(mgpu_handle *h, int block, const double *cextrp2, const double *cextrp3);

/* Wall coefficients of a wall face of a 3-D curvilinear block
 * (MGPU_GEOM_C3), as grid_normals.f90:232-300 computes them for
 * bc_wall_jmin/jmax_*_dpdn_c3 (src/Boundary/smod_bc_wall_c3.f90:842-1671)
 * and the adiabatic wall's tangential temperature gradient
 * (src/Compute_flow/smod_grad_T_c3.f90:249-275): for a j-face coef[0..7] are
 * arrays of nx*nz doubles (i fastest) over the wall points: gksigeta3 =
 * -grad(ksi).grad(eta)/|grad(eta)|^2, gphigeta3 (phi for ksi), the unit
 * tangent x_ksi/|x_ksi| (x, y, z: txn_ksi, tyn_ksi, tzn_ksi) and x_phi/|x_phi|
 * (txn_phi, tyn_phi, tzn_phi).  An i- or k-face takes the same with its own
 * two tangential axes in order (i-face: eta, phi over ny*nz points, j
 * fastest, getagksi3, gphigksi3, :10-839; k-face: ksi, eta over nx*ny, i
 * fastest, :1674-2505).  Required for every wall face of such a block before
 * its first step. */
int mgpu_set_wall_c3(mgpu_handle *h, int block, int face, const double *const *coef);

/* Dual tangents of a slip j-face (face 2 or 3) of a 3-D curvilinear block,
 * as grid_normals.f90:279-291 computes them for bc_wall_jmin/jmax_slip_c3
 * (src/Boundary/smod_bc_wall_2_c3.f90:363-393, 539-569): t[0..5] are arrays
 * of nx*nz doubles (i fastest): tx_ksi, ty_ksi, tz_ksi, tx_phi, ty_phi,
 * tz_phi, the dual basis of the unit tangents of mgpu_set_wall_c3 in the
 * tangent plane ((e_ksi - c e_phi)/(1 - c^2) and its twin, c = e_ksi.e_phi),
 * so that the wall momentum m becomes (m.t_ksi) e_ksi + (m.t_phi) e_phi.
 * Required, with mgpu_set_wall_c3, for every face of such a block that is a
 * slip wall (MGPU_FACE_SLIP with is_wall2 and an adiabatic case, or any wall
 * of an inviscid is_wall2 case) before its first step. */
int mgpu_set_wall_c3_slip(mgpu_handle *h, int block, int face, const double *const *t);

/* Upload the conserved variables rho,rhou,rhov,rhow,rhoe of a block (padded
 * arrays; ghosts are refreshed on the device). */
int mgpu_put_state(mgpu_handle *h, int block, const double *rho, const double *rhou,
                   const double *rhov, const double *rhow, const double *rhoe);

/* Download the conserved variables of a block (padded arrays, ghosts valid). */
int mgpu_get_state(mgpu_handle *h, int block, double *rho, double *rhou, double *rhov,
                   double *rhow, double *rhoe);

/* Advance nsteps full time steps: nrk stages (viscous terms in the last),
 * then the filter, with device-side halo/interface copies in between
 * (runge_kutta, src/Compute_flow/runge_kutta.f90:2-152).  Asynchronous on
 * the handle's stream. */
int mgpu_step(mgpu_handle *h, int nsteps);

/* Tile autotuner of the stage kernel.  For every block (and kernel variant:
 * Euler stages, last stage with the viscous divergence) it times each tile
 * shape (i x j workgroup tile) and k-chunk length on the block's own data
 * and keeps the fastest; results are cached per process by block geometry.
 * flags bit 0: re-tune even when cached.  Synchronous.  Needs metrics and
 * state set.  No counterpart in the reference (its loops have no tiles). */
int mgpu_autotune(mgpu_handle *h, int flags);

/* Launch parameters in use for a block: variant 0 = Euler stages, 1 = last
 * stage with viscous divergence, 2 = viscous fluxes, 3 = viscous divergence,
 * 4 = filter (tile extents and planes per workgroup). */
int mgpu_stage_config(mgpu_handle *h, int block, int variant, int *shape_tx, int *shape_ty, int *kchunk);

/* 1 if stage variant 0 (Euler stages) or 1 (the last stage) runs the
 * pipelined stage kernel, 0 if the classic one, < 0 on error. */
int mgpu_stage_pipelined(mgpu_handle *h, int variant);

/* Wait for all work queued on the handle. */
int mgpu_sync(mgpu_handle *h);

/* Per-kernel timing with HIP events on the handle's stream.  enable=1 starts
 * recording (and clears the totals); mgpu_kernel_time returns, for kernel
 * class kind (MGPU_K_*), the number of launches and their summed duration
 * in milliseconds. */
#define MGPU_K_STAGE 0      /* Euler (+viscous divergence) stage + RK update */
#define MGPU_K_VISC 1       /* viscous fluxes of U^n                         */
#define MGPU_K_FILTER 2     /* filter + update_varn                          */
#define MGPU_K_HALO 3       /* halo / interface copies                       */
#define MGPU_K_COUNT 4
int mgpu_profile(mgpu_handle *h, int enable);
int mgpu_kernel_time(mgpu_handle *h, int kind, int64_t *launches, double *ms);

/* Message of the last error on h (or of the last failed mgpu_create when h
 * is NULL). */
const char *mgpu_last_error(const mgpu_handle *h);

/* Statistics of the current state U^n over the interior cells of all
 * blocks: mass sum(rho), kinetic energy sum(1/2 rho |u|^2) and enstrophy
 * sum(1/2 rho |curl u|^2) with the 5-pt velocity gradients -- the sums of
 * stats_tgv (src/File_IO/stats.f90:37-112), before its normalisation and
 * MPI_REDUCE (stats_tgv takes the gradients its last stage computed; these
 * are the current state's) -- and umax = max(|u| + c), the velocity scale
 * of timestep_new (src/Time/timestep_new.f90:270-288: deltat =
 * CFL*delta_min/umax).  Deterministic: fixed summation order.  Next to a
 * no-slip wall the velocity gradients take the one-sided 5-pt stencils of
 * grad_vel (rows 1, 2 and n-1, n). */
typedef struct mgpu_diag {
  double mass, kinetic, enstrophy, umax;
  int64_t cells;
} mgpu_diag;
int mgpu_diagnostics(mgpu_handle *h, mgpu_diag *out);

/* mgpu_diagnostics without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The per-workgroup partial results and their totals go to a
 * static device array of the library (4 MiB per GPU, held until the process
 * ends), so a block may hold at most 2^17 * 256 = 33 554 432 interior cells
 * (MGPU_EINVAL beyond).  It needs the handle's exchange tables, which
 * mgpu_create builds (MGPU_ESTATE if they are missing).
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (other handles on other threads or streams) race and must be serialised
 * by the caller.  Synchronous; the same results as mgpu_diagnostics up to
 * the order of the floating-point additions. */
int mgpu_diagnostics_noalloc(mgpu_handle *h, mgpu_diag *out);

/* Maximum CFL number along i, j and k over the interior cells of all blocks
 * of the current state U^n, as check_max_cfl prints it
 * (src/Time/timestep_new.f90:660-842): Cartesian (|u_d| + c) deltat idx_d;
 * 3-D curvilinear (|(u xi_x + v xi_y + w xi_z) ijacob3| + c |grad xi|
 * |ijacob3|) deltat; 2.5-D the same with the 2-D metrics in i and j,
 * Cartesian in k; c = sqrt(gam rg T).  Synchronous. */
int mgpu_max_cfl(mgpu_handle *h, double cfl[3]);

/* Release every device resource of the handle. */
void mgpu_destroy(mgpu_handle *h);

/* Ghost-fill plan of the interface exchange of a block table, computed on the
 * host (no device, no allocation on it): what mgpu_step's halo copies will
 * write, in the caller's frame.  Entry e of map (6 int64 each, at most cap
 * entries written, *count = the total) reads: destination block, element of
 * the destination's padded array (i fastest), destination array, source
 * block, source element, source array.  set 0: the 5 conserved variables
 * rho..rhoe over 5 ghost layers (communication3d, src/Parallel/
 * mod_comm.f90:249-290); set 1: the viscous set over 2 layers
 * (communication_3dv :376-402): 0..5 the stress components xx,xy,xz,yy,yz,zz,
 * 6..8 the energy fluxes along x,y,z, 9 p, 10 the Ducros sensor, 11..13 the
 * spectral radii along i,j,k.  Blocks with walls on their i- or k-faces are
 * stored with that axis exchanged with j, blocks with walls on their j- and
 * k-faces with their axes cycled; the plan undoes that, so it can be
 * checked against the reference's face types (src/Parallel/
 * mod_mpi_types_two_sided.f90:47-197) on the caller's arrays. */
int mgpu_exchange_plan(const mgpu_case *c, int set, int64_t *map, int64_t cap, int64_t *count);

/* ------------------------------------------------------------------------
 * Serving loop.  No counterpart in the reference, which runs one case per
 * MPI job (src/Main/musica_main.f90): this is for callers that advance many
 * independent cases of different sizes on one GPU.  Each request is a
 * one-block case with periodic faces; requests are admitted in submission
 * order while their device memory fits the budget, stay resident, and
 * advance together -- the RK stages of all resident requests of the same
 * kernel variant run as one batched launch per kernel.  Their steps are
 * exactly those of mgpu_step.
 * ------------------------------------------------------------------------ */
typedef struct mgpu_server mgpu_server;

int mgpu_server_create(int64_t hbm_budget_bytes, mgpu_server **out);

/* Queue a request.  c->nblocks must be 1 with periodic faces; metrics: the 6
 * Cartesian arrays idx,idy,idz,idx_v,idy_v,idz_v, or (MGPU_GEOM_C3) the 10
 * c3 arrays followed by the 10 viscous ones, or (MGPU_GEOM_C2) the 5 2-D
 * arrays of mgpu_set_metrics_c2, its 5 viscous ones, idz and idz_v (the
 * viscous entries may be null for inviscid cases); state: rho..rhoe padded.
 * The inputs are copied.  *id receives the request id. */
int mgpu_server_submit(mgpu_server *s, const mgpu_case *c, const double *const *metrics,
                       const double *const *state, int nsteps, int64_t *id);

/* One scheduling round: retire finished requests (result copied to host
 * memory held by the server, device memory released), admit queued
 * requests while the budget allows, then advance every group of resident
 * requests by up to max_steps steps (never past a member's last step).
 * Reports the resident and queued counts. */
int mgpu_server_tick(mgpu_server *s, int max_steps, int *n_resident, int *n_queued);

/* state: 0 queued, 1 resident, 2 done (result waiting in mgpu_server_take). */
int mgpu_server_status(mgpu_server *s, int64_t id, int *state, int *steps_done);

/* Copy a finished request's state (padded arrays) and forget the request. */
int mgpu_server_take(mgpu_server *s, int64_t id, double *rho, double *rhou, double *rhov, double *rhow,
                     double *rhoe);

/* Device bytes held by resident requests. */
int64_t mgpu_server_memory(const mgpu_server *s);

/* Co-batching record since mgpu_server_create: *group_steps = time steps
 * launched for groups (one batched launch per kernel and stage each),
 * *request_steps = request time steps they advanced (their ratio is the mean
 * number of requests per launch), *max_members = the most requests one group
 * step advanced. */
int mgpu_server_batching(const mgpu_server *s, int64_t *group_steps, int64_t *request_steps, int32_t *max_members);

const char *mgpu_server_last_error(const mgpu_server *s);
void mgpu_server_destroy(mgpu_server *s);

#ifdef __cplusplus
}
#endif
#endif
