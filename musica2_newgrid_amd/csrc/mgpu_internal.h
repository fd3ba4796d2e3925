// Internal definitions shared by the HIP kernels and the C-ABI of
// libmusica_gpu.so.  gfx950 only (64-wide wavefronts).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/musica_gpu.h"

namespace mgpu {

constexpr int G = MGPU_NGH;  // ghost layers (ngh=5, read_param.f90:495)

// 11-pt centred 10th-order first derivative (mod_coeff_deriv.f90:316-325).
constexpr double A11_1 = 5.0 / 6.0, A11_2 = -5.0 / 21.0, A11_3 = 5.0 / 84.0,
                 A11_4 = -5.0 / 504.0, A11_5 = 1.0 / 1260.0;
// 5-pt centred 4th-order first derivative (mod_coeff_deriv.f90:121-126).
constexpr double A5_1 = 2.0 / 3.0, A5_2 = -1.0 / 12.0;
// 11-pt selective filter (mod_filtering.f90:63-68), scaled by -chi at run time.
constexpr double D11_0 = 252.0 / 1024.0, D11_1 = -210.0 / 1024.0, D11_2 = 120.0 / 1024.0,
                 D11_3 = -45.0 / 1024.0, D11_4 = 10.0 / 1024.0, D11_5 = -1.0 / 1024.0;
// 9th-order conservative dissipation (mod_filtering_shock.f90:95-100).
constexpr double D10_1 = -126.0, D10_2 = 84.0, D10_3 = -36.0, D10_4 = 9.0, D10_5 = -1.0;
// Centred first derivatives of half-width 1, 3, 4 used on rows 2, 4, 5 from a
// wall (a3, a7, a9; row 3 uses a5): mod_coeff_deriv.f90:102-104, 125-131, 232-241.
constexpr double A3_1 = 0.5;
constexpr double A7_1 = 3.0 / 4.0, A7_2 = -3.0 / 20.0, A7_3 = 1.0 / 60.0;
constexpr double A9_1 = 4.0 / 5.0, A9_2 = -1.0 / 5.0, A9_3 = 4.0 / 105.0, A9_4 = -1.0 / 280.0;
// 3-pt one-sided first derivative of an advanced wall row (is_wall2),
// a20 = -a02 at a max face (mod_coeff_deriv.f90:113-117).
constexpr double A02_1 = -3.0 / 2.0, A02_2 = 2.0, A02_3 = -1.0 / 2.0;
// 5-pt one-sided first derivatives on the wall row (a04) and the next (a13);
// a40 = -a04 and a31 = -a13 mirror them at a max face (mod_coeff_deriv.f90:114-124).
constexpr double A04_1 = -25.0 / 12.0, A04_2 = 4.0, A04_3 = -3.0, A04_4 = 4.0 / 3.0, A04_5 = -1.0 / 4.0;
constexpr double A13_1 = -1.0 / 4.0, A13_2 = -5.0 / 6.0, A13_3 = 3.0 / 2.0, A13_4 = -1.0 / 2.0, A13_5 = 1.0 / 12.0;
// Standard 5-, 7- and 9-pt filters on rows 3, 4, 5 from a wall
// (mod_filtering.f90:24-38), scaled by -chi at run time.
constexpr double D5_0 = 6.0 / 16.0, D5_1 = -4.0 / 16.0, D5_2 = 1.0 / 16.0;
constexpr double D7_0 = 20.0 / 64.0, D7_1 = -15.0 / 64.0, D7_2 = 6.0 / 64.0, D7_3 = -1.0 / 64.0;
constexpr double D9_0 = 70.0 / 256.0, D9_1 = -56.0 / 256.0, D9_2 = 28.0 / 256.0, D9_3 = -8.0 / 256.0,
                 D9_4 = 1.0 / 256.0;

// Element offset inside one padded array.  32-bit: a block's padded array
// holds at most 2^28 points (checked at mgpu_create), so byte offsets fit in
// 32 bits and loads use the SGPR-base + 32-bit VGPR-offset form.
using idx_t = uint32_t;
constexpr size_t kMaxPadded = size_t(1) << 28;

// Geometry of one block's padded arrays (i fastest).
struct Geom {
  int nx, ny, nz;
  int sx;     // nx + 2G
  idx_t sy;   // sx*(ny+2G)
  int jwall;  // no-slip walls: bit 0 on the jmin face, bit 1 on the jmax face
  int iwall;  // no-slip walls on the i-faces (a block walled on two axes), bits as jwall
  // boundary faces of any kind (no-slip wall, characteristic outflow,
  // supersonic inlet), at which the operators take their closures: bit 0 on
  // the min face, bit 1 on the max face of axis i (ibnd) and j (jbnd)
  int ibnd, jbnd;
  // advanced walls (is_wall2): the wall rows of these faces take an Euler
  // increment with the one-sided a02 / a20 derivative across the wall; bits
  // as jwall (jadv) and iwall (iadv)
  int iadv, jadv;
  long n;     // total padded points
};

// Half-width of the centred stencil that row r (1..n) of an axis may use: 5
// inside, r-1 (n-r) near a bounded min (max) face, 0 on the boundary row.
__host__ __device__ inline int bhalf(int bnd, int n, int r) {
  int q = 5;
  if ((bnd & 1) && r - 1 < q) q = r - 1;
  if ((bnd & 2) && n - r < q) q = n - r;
  return q;
}
__host__ __device__ inline int jhalf(const Geom &g, int j) { return bhalf(g.jbnd, g.ny, j); }
__host__ __device__ inline int ihalf(const Geom &g, int i) { return bhalf(g.ibnd, g.nx, i); }

// Face slabs of the boundary rows: one value per point of the face's row,
// the i-faces indexed (j-1) + ny*(k-1), the j-faces (i-1) + nx*(k-1); a slab
// of 5 variables stores them one after the other.
__host__ __device__ inline long slab_size(const Geom &g, int axis) {
  return axis == 0 ? (long)g.ny * g.nz : (long)g.nx * g.nz;
}

__host__ __device__ inline idx_t lin(const Geom &g, int i, int j, int k) {
  return (idx_t)(i + G - 1) + (idx_t)g.sx * (idx_t)(j + G - 1) + g.sy * (idx_t)(k + G - 1);
}

// Load/store at element offset i through a 32-bit byte offset.  The array
// pointers reach the kernels through device-memory tables, so the compiler
// cannot infer their address space: cast them to the global one, so that the
// accesses are global (SGPR base + 32-bit VGPR offset, counted by vmcnt only)
// instead of flat ones, which also count against lgkmcnt and make every LDS
// wait a wait for the outstanding memory loads.
#define MGPU_GLOBAL __attribute__((address_space(1)))
template <class T>
__device__ __forceinline__ const MGPU_GLOBAL T *gaddr(const T *b, idx_t i) {
  return reinterpret_cast<const MGPU_GLOBAL T *>(reinterpret_cast<const MGPU_GLOBAL char *>(
                                                     (const MGPU_GLOBAL T *)b) +
                                                 (uint32_t)(i * (idx_t)sizeof(T)));
}
template <class T>
__device__ __forceinline__ MGPU_GLOBAL T *gaddr(T *b, idx_t i) {
  return reinterpret_cast<MGPU_GLOBAL T *>(reinterpret_cast<MGPU_GLOBAL char *>((MGPU_GLOBAL T *)b) +
                                           (uint32_t)(i * (idx_t)sizeof(T)));
}
__device__ __forceinline__ double ldg(const double *b, idx_t i) { return *gaddr(b, i); }
__device__ __forceinline__ void stg(double *b, idx_t i, double v) { *gaddr(b, i) = v; }
// Streaming load for data only the loading point reads (no neighbour re-reads
// it from L2), so that it does not evict halo data other workgroups re-read.
__device__ __forceinline__ double ldg_nt(const double *b, idx_t i) { return __builtin_nontemporal_load(gaddr(b, i)); }
// Streaming store for results no kernel of the same launch reads back.
__device__ __forceinline__ void stg_nt(double *b, idx_t i, double v) { __builtin_nontemporal_store(v, gaddr(b, i)); }
// The same for FP32 arrays.
__device__ __forceinline__ float ldg(const float *b, idx_t i) { return *gaddr(b, i); }
__device__ __forceinline__ void stg(float *b, idx_t i, float v) { *gaddr(b, i) = v; }
__device__ __forceinline__ void stg_nt(float *b, idx_t i, float v) { __builtin_nontemporal_store(v, gaddr(b, i)); }

// XCD-aware workgroup order.  gfx950 dispatches workgroup w to XCD w % 8;
// each XCD has its own L2.  Renumber so that XCD x owns one contiguous range
// of logical tiles: neighbouring tiles (which re-read each other's halos)
// then share an L2.  Bijective for any count n.
constexpr int NXCD = 8;
__device__ __forceinline__ int xcd_remap(int w, int n) {
  const int x = w % NXCD, l = w / NXCD;
  const int q = n / NXCD, r = n % NXCD;
  return x * q + (x < r ? x : r) + l;
}

// Tiles of one block in a batched launch: logical tiles [first, first +
// tiles_x*tiles_y*ceil(nz/kchunk)) belong to it.  One launch covers every
// block of a case; each workgroup finds its block by binary search over
// `first` and reads that block's arguments from a device table.
struct TileRange {
  int first, tiles_x, tiles_y, kchunk;
  // tile numbering inside the block: 0 i fastest, 1 j fastest.  With j
  // fastest the consecutive logical tiles that xcd_remap gives one XCD form
  // columns of tiles along j, whose 5-row j-halos are their neighbours' own
  // rows: those re-reads can then hit the XCD's L2.
  int jfirst;
};
template <class A>
struct Batch {
  A a;
  TileRange t;
};

__host__ __device__ inline int tile_count(const TileRange &t, int nz) {
  return t.tiles_x * t.tiles_y * ((nz + t.kchunk - 1) / t.kchunk);
}

// Entry of the block that owns logical tile w (entries sorted by first).
template <class E>
__device__ __forceinline__ const E &batch_find(const E *__restrict__ tab, int nb, int w) {
  int lo = 0, hi = nb - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (tab[mid].t.first <= w) lo = mid;
    else hi = mid - 1;
  }
  return tab[lo];
}

// Tile coordinates of logical tile w inside its block.
__device__ __forceinline__ void tile_coords(const TileRange &t, int w, int &tile_x, int &tile_y, int &chunk) {
  const int l = w - t.first;
  if (t.jfirst) {
    tile_y = l % t.tiles_y;
    const int r = l / t.tiles_y;
    tile_x = r % t.tiles_x;
    chunk = r / t.tiles_x;
  } else {
    tile_x = l % t.tiles_x;
    const int r = l / t.tiles_x;
    tile_y = r % t.tiles_y;
    chunk = r / t.tiles_y;
  }
}

// Thermodynamics / transport constants of a case.
struct Gas {
  double gam1, rg, gam;
  double c0, S0, diffscale, cok_ratio;
};

// Five conserved variables as separate arrays.
struct State {
  double *v[5];
};
struct CState {
  const double *v[5];
};

// Metrics of a block on the device.
struct Metrics {
  // Cartesian (length n+2G each)
  const double *idx, *idy, *idz, *idx_v, *idy_v, *idz_v;
  // c3 padded 3-D arrays
  const double *m[10];
  const double *mv[10];
  // 2.5-D padded 2-D arrays (index lin2): the ksi-flux metric row (y_eta,
  // -x_eta), the eta-flux row (-y_ksi, x_ksi), ijacob; Euler and viscous
  // sets.  z uses idz / idz_v.
  const double *m2[5];
  const double *m2v[5];
};

// Offset of column (i, j) in a padded 2-D (i, j) array.
__host__ __device__ inline idx_t lin2(const Geom &g, int i, int j) {
  return (idx_t)(i + G - 1) + (idx_t)g.sx * (idx_t)(j + G - 1);
}

// Reference state of the mixed-precision path: the FP32 arithmetic works on
// fluxes and states minus their values at this constant state.  The 11-pt
// divergence and the filters annihilate constants (and, on c3 metrics, the
// constant-state contravariant flux by the discrete GCL), so the results are
// those of the full quantities, with FP32 rounding relative to the
// perturbation instead of the mean.
struct RefState {
  double r, ru, rv, rw, re;  // conserved
  double u, v, w, p, H;      // primitives, H = rho E + p
};

// Arguments of the Euler (+viscous divergence) stage kernel.
struct StageArgs {
  Geom g;
  Gas gas;
  CState uin;    // stage input U_n
  CState u0;     // step-start U^n
  State uout;    // stage output
  Metrics met;
  const double *kv[4];   // viscous increment of rhou,rhov,rhow,rhoe (last stage)
  double alpha;          // deltat*crk(irk)
  RefState ref;          // mixed precision only
  // mixed precision: intermediate stage states as FP32 differences with ref
  const float *uin32[5];
  float *uout32[5];
  // boundary rows, which get no Euler increment: their increment K per face
  // imin, imax, jmin, jmax, as a face slab of 5 variables (the characteristic
  // increment, or at a supersonic inlet the filter increment of the previous
  // step); null: K = 0 and the viscous increment (no-slip walls).  A row of
  // an i-face takes its face's slab also where it meets a j-face.
  const double *kbc[4];
  // supersonic inlet at imin: the i-flux of row 1 takes its conserved
  // variables from uin_row (face slab) and its primitives from prim_row
  // (face slab; first stage) or from the stage state (later stages)
  const double *uin_row, *prim_row;
  // mixed precision, 3-D curvilinear: the 9 flux metric terms as FP32
  // (ksi_x..phi_z rounded; the FP32 flux arithmetic rounds them so anyway)
  const float *m32[9];
};

// Stage kernel variants: FP64, or mixed precision with the stage state read
// and written as FP64 (U^n, the last stage's output) or as FP32 differences
// with the reference state (the intermediate stages).
enum StageMode { kFP64 = 0, kMixFirst = 1, kMixMiddle = 2, kMixLast = 3 };

// Geometry of a case (mgpu_case.curvilinear): Cartesian 1-D metrics, full
// 3-D curvilinear metrics (is_curv3), or 2.5-D curvilinear (is_curv: 2-D
// metrics in i, j and a Cartesian z).
enum Geometry { kGeoCart = 0, kGeoC3 = 1, kGeoC2 = 2 };

// Viscous flux tensor of U^n in the physical frame, per interior point:
// tau[0..5] = -tau_xx,-tau_xy,-tau_xz,-tau_yy,-tau_yz,-tau_zz (the
// reference's t11..t33 after scaling by -2mu), tau[6..8] = energy fluxes
// e_d = u t_d1 + v t_d2 + w t_d3 - kappa dT/dx_d.  Frame-independent, so
// that an interface copy of them equals the reference's copy of the
// gradients followed by the flux evaluation at the ghost point.
constexpr int NTAU = 9;
struct ViscArgs {
  Geom g;
  Gas gas;
  CState u0;          // U^n
  Metrics met;
  double *tau[NTAU];  // out: interior points
  double *prs;        // out: p(U^n), interior points (shock filter)
  double *sens;       // out: Ducros sensor, interior points (shock filter)
  double *sr[3];      // out: |u|+c, |v|+c, |w|+c of U^n, interior points (artificial viscosity,
                      // Edoh switch); c3: |contravariant velocity| + c |grad xi_d| per direction
  int want_sensor;    // 1: write prs and sens
  int want_sr;        // 1: write sr
  int ducros;
};

// Divergence of the viscous fluxes: K_v = sum_d D5_d(F^v_d) * metric, with
// F^v_d the projection of tau on the viscous metrics of direction d (c3) or
// its row d (Cartesian), at every stencil point.
struct VdivArgs {
  Geom g;
  Metrics met;
  const double *tau[NTAU];  // interior + 2 face-ghost layers
  double *kv[4];            // out, interior points
};

struct FilterArgs {
  Geom g;
  CState uin;   // U_n after the last stage
  State uout;   // U_n + K
  const double *prs;   // shock filter: p(U^n)
  const double *sens;  // shock filter: Ducros sensor
  double dd[6];        // -chi*d11 (selective filter)
  double chi;          // filter amplitude (the 5/7/9-pt filters next to walls)
  double xnu0, xnu10, csens;
  const double *sr[3];  // artificial viscosity, Edoh switch: spectral radii of U^n (see ViscArgs)
  // artificial viscosity, Edoh switch: Cartesian metrics (null where the
  // spectral radius carries its metric: all directions of c3 blocks, i and j
  // of 2.5-D blocks)
  const double *idx, *idy, *idz;
  double dt;            // artificial viscosity (U += deltat*K), Edoh switch (local CFL)
  int edoh;             // shock filter: Edoh switch
  int avs;              // artificial viscosity with shock capturing (k_shock_wall)
};

// No-slip wall state (bc_wall with dp/dn = 0) of the j-faces (and i-faces)
// of one block: rho and rhoE from p and T extrapolated from rows 2,3 of `src`
// (or read from `stash`), zero momentum, written to the wall rows of `dst` (or
// to `stash`: two values per wall point, jmin, jmax, then imin, imax).  On an
// edge the j-face's state wins (bc_wall applies imin..jmax in order,
// mod_bc_apply.f90:237-261, each face from the same primitives).
// Advanced walls (g.jadv / g.iadv) keep the wall row of `dst` and change it
// in place: slip faces remove its normal momentum, the others zero it and set
// rhoE = rho cv T_wall, T_wall extrapolated from `src` (or the stash, which
// then holds T_wall alone); on an edge the j-face acts last.
struct WallArgs {
  Geom g;
  Gas gas;
  CState src;
  State dst;
  // mixed precision: an intermediate stage set stored as FP32 differences
  // with the reference state ref (used when the FP64 pointer is null)
  const float *src32[5];
  float *dst32[5];
  double ref[5];
  double *stash;
  double c2[2], c3[2];  // extrapolation coefficients cextrp2, cextrp3 (jmin, jmax)
  double ci2[2], ci3[2];  // the same of the imin, imax faces
  double cv;            // cv = rg/(gam-1)
  double t_wall;        // isothermal wall temperature
  int adiabatic;
  // c3 blocks: per side (jmin, jmax) the 8 coefficient slabs of the wall
  // points (nx*nz each, i fastest): gksigeta3, gphigeta3, then the unit
  // tangents x_ksi/|x_ksi| and x_phi/|x_phi| (x, y, z components); null on
  // Cartesian blocks, whose extrapolation uses c2, c3
  const double *wc3[2];
  // slip walls among the advanced ones: bits as g.jwall (jslip), g.iwall
  // (islip); a c3 slip face's dual tangents t_ksi, t_phi (6 slabs)
  int jslip, islip;
  const double *wt3[2];
};

// The viscous rows next to the bounded faces of one block (rows 1,2 and
// n-1,n of a bounded i or j axis): tau with the one-sided gradients along
// the bounded axes, the wall-row rescaling after the exchange, and the
// viscous divergence on those rows.
struct WallViscArgs {
  Geom g;
  Gas gas;
  CState u0;  // U^n
  Metrics met;
  double *tau[NTAU];
  double *kv[4];
  double *prs, *sens;  // shock filter: p and Ducros sensor of U^n
  int adiabatic;       // no wall-normal temperature gradient on the wall row
  int want_sensor;     // write the sensor on those rows and p beyond the bounded faces
  int ducros;          // Ducros sensor (else 1)
  int geo;             // kGeoCart or kGeoC3 (viscous metrics mapping the index-space derivatives)
  const double *wc3[2];  // c3 walls: the coefficient slabs of WallArgs::wc3 (the unit tangents)
  double *sr[3];       // artificial viscosities: spectral radii of U^n (null without)
};

// Characteristic (LODI) boundaries and the supersonic inlet of one block:
// the increments of the characteristic rows from the primitives of a state
// set, and the bookkeeping of the inflow row.
struct CharArgs {
  Geom g;
  Gas gas;
  double cp, p_ref;       // cp = gam rg/(gam-1); reference pressure of the imax relaxation
  CState src;             // the state (FP64), or with src.v[0] null
  const float *src32[5];  // FP32 differences with ref
  double ref[5];
  const double *idx_v, *idy_v;
  double *kimax, *kjmax;  // out: increment slabs of the imax / jmax row (null: no such face)
  double *cst;            // imax: sound speed primitives_visc last computed (face slab)
  // cst_mode 1: after the increments, cst := c of state cnew (U^n, FP64);
  // 2: before them, cst := c of src; 3: only cst := c of cnew, no increments
  int cst_mode;
  CState cnew;
  // supersonic inlet at imin (null: none): inflow state, state whose
  // primitives row 1 takes in the first stage, filter increment of row 1
  double *uin, *ppre, *kf;
  // inlet_mode 1 (end of a step): ppre := src row 1, then row 1 of row1 (the
  // src set) := uin; 2 (start of a run): uin := ppre := src row 1, kf := 0
  int inlet_mode;
  State row1;
  int i0, i1;             // jmax row: i = i0..i1
};

// Component c of tau in the flux along direction d of a Cartesian block (row
// d of the stress tensor, then the energy flux e_d); constant-folds for
// constant d and c.
__host__ __device__ constexpr int krow(int d, int c) {
  return c == 3 ? 6 + d : (d == 0 ? c : (d == 1 ? (c == 0 ? 1 : 2 + c) : (c == 0 ? 2 : (c == 1 ? 4 : 5))));
}

// One face of one destination block in the interface exchange: the ghost
// layers of block `dst` on that face take the interior layers of block
// `src` on the neighbour face.  Point (a0,a1,a2) of the box n[0]xn[1]xn[2]
// writes element dst_base + sum a_d*dst_s[d] of array v of the dst block and
// reads src_base + sum a_d*src_s[d] of array vmap[v] of the src block: the
// two blocks may store their axes in different orders (blocks with walls on
// their i- or k-faces), and vector and tensor components follow the axes.
constexpr int kIfaceMaxVar = 16;
struct FaceTask {
  int dst, src;
  int n[3];
  int dst_base, src_base;
  int dst_s[3], src_s[3];
  int8_t vmap[kIfaceMaxVar];
};
struct IfaceChunk {
  int task, start;
};
constexpr int kIfaceThreads = 256;
template <class T>
struct IfaceArgsT {
  const FaceTask *tasks;
  const IfaceChunk *chunks;
  T *const *ptrs;  // ptrs[block*nvar + v]: array v of block
  int nvar;
};
using IfaceArgs = IfaceArgsT<double>;
using IfaceArgs32 = IfaceArgsT<float>;

// Diagnostics of one block: per-workgroup partial results (3 sums and a max).
constexpr int kDiagThreads = 256;
struct DiagArgs {
  Geom g;
  CState u;
  Metrics met;
  double gam, gam1;
  double *partial;
};

// Maximum CFL number per direction of one block: per-workgroup maxima for i,
// j, k (3 per workgroup, diag_groups workgroups).
struct CflArgs {
  Geom g;
  CState u;
  Metrics met;
  double gam, gam1, rg, dt;
  double *partial;
};

}  // namespace mgpu

// Kernel launchers (defined in the .hip files).  Each takes a device table
// of nb Batch entries (sorted by first tile) and the launch's tile count.
namespace mgpu {
extern const int kStageShapes[][2];
extern const int kNumStageShapes;
extern const int kFirstPipelinedShape;  // shapes from here on: the pipelined stage kernel
TileRange stage_tiles(const Geom &g, int shape, int kchunk);
// wall: 0 no bounded face, 1 bounded faces (their closures), 2 also advanced
// wall rows (is_wall2; classic shapes only)
hipError_t launch_stage(const Batch<StageArgs> *tab, int nb, int ntiles, int geo, bool visc, int shape,
                        int mode, int wall, hipStream_t s);
TileRange visc_tiles(const Geom &g, int kchunk);  // kchunk 0: default
hipError_t launch_visc(const Batch<ViscArgs> *tab, int nb, int ntiles, int geo, hipStream_t s);
TileRange vdiv_tiles(const Geom &g, int kchunk);
hipError_t launch_vdiv(const Batch<VdivArgs> *tab, int nb, int ntiles, int geo, hipStream_t s);
TileRange filter_tiles(const Geom &g, int kchunk);
// The filters run in FP64 in both precisions (they read and write FP64
// states); `wall`: the closures of blocks with bounded faces.
hipError_t launch_filter_sf(const Batch<FilterArgs> *tab, int nb, int ntiles, bool wall, hipStream_t s);
// The operators sharing the shock filter's face-dissipation structure.
enum DissipKind {
  kDissipShock = 0,         // filter_o10_shock: U_n += K
  kDissipArtvisc = 1,       // artvisc_o9: U_n += deltat*K, spectral-radius weighted d10 faces
  kDissipArtviscShock = 2,  // artvisc_o9_shock: U_n += deltat*K, spectral-radius weighted shock faces
};
// `wall`: the closures of blocks with bounded faces (not artvisc_o9); `edoh`:
// the shock filter's Edoh switch.
hipError_t launch_filter_shock(const Batch<FilterArgs> *tab, int nb, int ntiles, DissipKind kind, bool wall,
                               bool edoh, hipStream_t s);
// The increment along axis (0: i, 1: j) of the rows next to the bounded
// faces of that axis of one block, after launch_filter_shock(.., wall), in
// FP64 (the filter's input is the last stage's FP64 output): the shock
// filter (a.avs = 0) or artvisc_o9_shock (a.avs = 1, j only).
hipError_t launch_shock_wall(const FilterArgs &a, int axis, hipStream_t s);
hipError_t launch_iface(const IfaceArgs &a, int nchunks, hipStream_t s);
hipError_t launch_iface(const IfaceArgs32 &a, int nchunks, hipStream_t s);
long diag_groups(const Geom &g);
hipError_t launch_cfl(const CflArgs &a, int geo, hipStream_t s);
hipError_t launch_wall(const WallArgs &a, hipStream_t s);
hipError_t launch_visc_wall(const WallViscArgs &a, hipStream_t s);
hipError_t launch_tau_wallrow(const WallViscArgs &a, hipStream_t s);
hipError_t launch_vdiv_wall(const WallViscArgs &a, hipStream_t s);
hipError_t launch_charac(const CharArgs &a, hipStream_t s);
// Filter increment of the inlet row: kf = after - before on row 1 (face slab).
hipError_t launch_inlet_kf(const Geom &g, const CState &after, const CState &before, double *kf, hipStream_t s);
hipError_t launch_diag(const DiagArgs &a, int geo, hipStream_t s);
// Allocation-free diagnostics of one block (mgpu_diagnostics_noalloc): the
// per-workgroup results and their totals go to a static device array of
// 4 * kDiagScratchGroups + 4 doubles; diag_scratch_init looks its address up
// once per device (that device current).  diag_block_noalloc returns the
// block's 3 sums and maximum in tot, synchronously; hipErrorInvalidValue if
// the block needs more than kDiagScratchGroups workgroups.
constexpr long kDiagScratchGroups = 1L << 17;
hipError_t diag_scratch_init(int device);
hipError_t diag_block_noalloc(DiagArgs a, int geo, int device, hipStream_t s, double tot[4]);
}  // namespace mgpu
