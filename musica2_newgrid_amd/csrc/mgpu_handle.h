// Host-side state of a libmusica_gpu.so handle, shared by the C-ABI
// (mgpu_api.hip) and the serving loop (mgpu_serve.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "mgpu_internal.h"

namespace mgpu_host {

using namespace mgpu;

struct BlockData {
  Geom g{};
  mgpu_case p{};          // scalars of this block's case (time step, gas, filter amplitudes)
  double *st[3][5] = {};  // three state sets: st[cur] = U^n, the others stage buffers
  float *sf[2][5] = {};   // mixed precision: intermediate stage states, FP32 differences with ref
  double *idx[6] = {};    // Cartesian metrics idx,idy,idz,idx_v,idy_v,idz_v
  double *m[10] = {};
  float *m32[9] = {};     // mixed precision, c3: the flux metric terms of m as FP32
  double *mv[10] = {};
  double *m2[5] = {};     // 2.5-D: 2-D metric rows (y_eta, -x_eta, -y_ksi, x_ksi, ijacob), Euler set
  double *m2v[5] = {};    // 2.5-D: the viscous set
  double *tau[NTAU] = {}; // physical viscous flux tensor of U^n (last stage)
  double *kv[4] = {};     // viscous increment of rhou..rhoe (last stage)
  double *prs = nullptr, *sens = nullptr;
  double *sr[3] = {};     // artificial viscosity: |u|+c, |v|+c, |w|+c of U^n (last stage)
  RefState ref{};         // mixed precision: interior mean of the state at put_state
  double *wstash = nullptr;  // walls: rho, rhoE of the end-of-step wall state (2 per wall point)
  double c2[2] = {0, 0}, c3[2] = {0, 0};  // walls: cextrp2, cextrp3 of the jmin, jmax faces
  double ci2[2] = {0, 0}, ci3[2] = {0, 0};  // and of the imin, imax faces (walls on two axes)
  // c3 walls: per j-face, the 8 coefficient slabs of mgpu_set_wall_c3 (8*nx*nz doubles)
  double *wc3[2] = {nullptr, nullptr};
  bool wc3_set[2] = {false, false};
  // slip walls (bits as g.jwall, g.iwall) and the dual tangents of a c3 slip
  // face (6 slabs of nx*nz)
  int jslip = 0, islip = 0;
  double *wt3[2] = {nullptr, nullptr};
  bool wt3_set[2] = {false, false};
  // characteristic outflow (imax, jmax) and supersonic inlet (imin): face
  // slabs (see slab_size) of what the reference keeps between the calls of
  // its boundary routines -- the increments of the characteristic rows, the
  // sound speed bc_charac_imax reads, and at the inlet the inflow state, the
  // state whose primitives row 1 takes in the first stage and the filter
  // increment of row 1 (K there is written by the filter only)
  double *kimax = nullptr, *kjmax = nullptr, *cst = nullptr;
  double *uin = nullptr, *ppre = nullptr, *kf = nullptr;
  bool bc_start = false;   // start_runge_kutta pending (state uploaded)
  size_t bytes = 0;       // device memory held by the block
  bool have_metrics = false;
  bool have_state = false;
  bool have_wall = false;  // wall coefficients set (blocks with walls)
  // Axis relabeling: the kernels close walls on the j-faces only, so a block
  // with walls on its i-faces (perm 1) is stored with i and j exchanged, one
  // with walls on its k-faces (perm 2) with j and k exchanged, one with walls
  // on its j- and k-faces (perm 3) with its j, k, i stored as i, j, k; the state,
  // Cartesian metrics and wall coefficients are exchanged with it at the API.
  int perm = 0;
  // the relabeling of the physical axes (momentum, tau): perm on a Cartesian
  // case, none on a curvilinear one (only its index axes and metric rows move)
  int vperm = 0;
};

struct ProfRec {
  int kind;
  hipEvent_t a, b;
};

}  // namespace mgpu_host

struct mgpu_handle {
  mgpu_case c{};                    // flags (uniform over the blocks) and defaults
  std::vector<mgpu_block> blocks;   // topology, c.blocks points here
  std::vector<mgpu_host::BlockData> b;
  int cur = 0;  // index of the set holding U^n
  hipStream_t stream = nullptr;
  int device = 0;
  std::string err;
  bool profiling = false;
  std::vector<mgpu_host::ProfRec> recs;
  std::vector<hipEvent_t> ev_pool;
  int64_t kcount[MGPU_K_COUNT] = {};
  double kms[MGPU_K_COUNT] = {};
  // interface exchange: task and chunk tables (5 state layers, 2 viscous
  // layers) and per-array-set pointer tables, all in device memory
  mgpu::IfaceArgs xs[3] = {};  // state set s
  mgpu::IfaceArgs xv = {};     // tau (+ prs, sens)
  mgpu::IfaceArgs32 xf[2] = {};  // mixed precision: FP32 stage sets
  int nchunks_s = 0, nchunks_v = 0;
  std::vector<void *> xtables;  // device memory of the exchange tables
  bool exchange_ready = false;
  bool ghosts_dirty = false;
  // stage kernel launch parameters: [0] Euler stages, [1] last stage with the
  // viscous divergence (chosen by mgpu_autotune, else a default); kchunk 0
  // means the per-block default
  int shape[2] = {2, 2};
  int stage_family = 0;  // 0: the tuner may pick either stage kernel, 1: classic only, 2: pipelined only
  int kchunk[2] = {0, 0};
  // planes per workgroup of k_visc, k_vdiv and the filter (0: per-block default)
  int kc_aux[3] = {0, 0, 0};
  // tile numbering (TileRange::jfirst) of the stage variants and of k_visc,
  // k_vdiv and the filter
  int jfirst[2] = {0, 0};
  int jf_aux[3] = {0, 0, 0};
  // batched launch tables in device memory, one entry per block
  mgpu::Batch<mgpu::StageArgs> *d_stage = nullptr;  // [cur][irk-1][block]
  mgpu::Batch<mgpu::ViscArgs> *d_visc = nullptr;    // [cur][block]
  mgpu::Batch<mgpu::VdivArgs> *d_vdiv = nullptr;    // [block]
  mgpu::Batch<mgpu::FilterArgs> *d_filt = nullptr;  // [cur][block]
  mgpu::Batch<mgpu::StageArgs> *d_tune = nullptr;   // [block], autotuner
  mgpu::Batch<mgpu::FilterArgs> *d_tune_f = nullptr;  // [block], autotuner (writes a stage buffer)
  int tab_cap = 0;                                  // blocks the tables hold
  int nt_stage[2] = {0, 0}, nt_visc = 0, nt_vdiv = 0, nt_filt = 0;
  bool tables_ready = false;
  // one captured step per U^n set (a case without filter rotates the sets)
  bool use_graph = true;
  hipGraphExec_t gexec[3] = {};
  bool mixed = false;  // FP32 flux/derivative arithmetic and intermediate stage states
  bool wall = false;   // some block has no-slip walls
  bool bnd = false;    // some block has boundary faces (closure variants of the stage and filter kernels)
  bool adv = false;    // some block has advanced wall rows (is_wall2: their stage-kernel variant)
  bool charac = false; // some block has characteristic or inlet faces
  bool shared_ref = false;  // one mixed-precision reference state for all blocks (set from block 0)
  double *diag_buf = nullptr;  // per-workgroup partial sums of mgpu_diagnostics
  long diag_cap = 0;
  double *cfl_buf = nullptr;   // per-workgroup maxima of mgpu_max_cfl
  long cfl_cap = 0;
};

namespace mgpu_host {

extern thread_local std::string g_create_error;

// Records an error message on the handle (or the thread's create error) and
// returns `code`.
int fail(mgpu_handle *h, int code, const char *fmt, ...);

// Validates the flags of a case (not its blocks).  Empty string if valid.
std::string check_flags(const mgpu_case &c);
// Empty handle with the flags of c on the current device.
int handle_new(const mgpu_case &c, mgpu_handle **out);
// Appends a block (extents of bl, scalars of p) and allocates its arrays.
int handle_add_block(mgpu_handle *h, const mgpu_block &bl, const mgpu_case &p);
// Frees block i and removes it from the handle (the other blocks keep their order).
int handle_remove_block(mgpu_handle *h, int i);
// Device bytes a block of extents nx,ny,nz needs under the flags of c.
size_t block_bytes(const mgpu_case &c, int nx, int ny, int nz);
// Uploads metrics / state of block i (host arrays in the padded layout).
int put_metrics_cart(mgpu_handle *h, int i, const double *const *a6);
int put_metrics_c3(mgpu_handle *h, int i, const double *const *m, const double *const *mv);
int put_metrics_c2(mgpu_handle *h, int i, const double *const *m2, const double *const *m2v, const double *idz,
                   const double *idz_v);
int put_state(mgpu_handle *h, int i, const double *const *u5);
int get_state(mgpu_handle *h, int i, double *const *u5);
// Advances every block of the handle by nsteps (asynchronous on h->stream).
int step(mgpu_handle *h, int nsteps);
void destroy(mgpu_handle *h);

}  // namespace mgpu_host
