// C-ABI of libmusica_gpu.so: handle, device memory, and the RK step
// choreography of runge_kutta (src/Compute_flow/runge_kutta.f90:2-152).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <string>
#include <vector>

#include "mgpu_handle.h"

using namespace mgpu;

namespace mgpu_host {

thread_local std::string g_create_error;

int fail(mgpu_handle *h, int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (h) h->err = buf;
  else g_create_error = buf;
  return code;
}

#define HIPCK(h, call)                                                                        \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess)                                                                     \
      return fail((h), MGPU_EDEVICE, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_),   \
                  __FILE__, __LINE__);                                                        \
  } while (0)

hipEvent_t get_event(mgpu_handle *h) {
  if (!h->ev_pool.empty()) {
    hipEvent_t e = h->ev_pool.back();
    h->ev_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Brackets one launch with events when profiling is on.
struct ProfScope {
  mgpu_handle *h;
  int kind;
  hipEvent_t a = nullptr, b = nullptr;
  ProfScope(mgpu_handle *h_, int k) : h(h_), kind(k) {
    if (h->profiling) {
      a = get_event(h);
      b = get_event(h);
      if (a) (void)hipEventRecord(a, h->stream);
    }
  }
  ~ProfScope() {
    if (h->profiling && a && b) {
      (void)hipEventRecord(b, h->stream);
      h->recs.push_back({kind, a, b});
    }
  }
};

int drain_profile(mgpu_handle *h) {
  if (h->recs.empty()) return 0;
  HIPCK(h, hipStreamSynchronize(h->stream));
  for (auto &r : h->recs) {
    float ms = 0.f;
    HIPCK(h, hipEventElapsedTime(&ms, r.a, r.b));
    h->kcount[r.kind] += 1;
    h->kms[r.kind] += ms;
    h->ev_pool.push_back(r.a);
    h->ev_pool.push_back(r.b);
  }
  h->recs.clear();
  return 0;
}

Gas gas_of(const mgpu_case &c) {
  Gas gs;
  gs.gam = c.gam;
  gs.gam1 = c.gam - 1.0;
  gs.rg = c.rg;
  gs.c0 = c.c0;
  gs.S0 = c.S0;
  gs.diffscale = c.diffscale;
  gs.cok_ratio = c.cok_ratio;
  return gs;
}

// The shock filter's Edoh switch is on (is_sw_edoh acts in filter_o10_shock only).
bool edoh_on(const mgpu_case &c) { return c.filter == MGPU_FILTER_SHOCK && c.is_sw_edoh != 0; }
// The dissipation kernels read p and the Ducros sensor of U^n (artvisc_o9:
// only through the kernels of artvisc_o9_shock it shares on blocks with
// walls, with no sensor term; kept for every artvisc_o9 case, so that every
// block of a case holds the same arrays).
bool has_sensor(const mgpu_case &c) {
  return c.filter == MGPU_FILTER_SHOCK || c.filter == MGPU_FILTER_ARTVISC_SHOCK ||
         c.filter == MGPU_FILTER_ARTVISC;
}
// The artificial viscosities: spectral radii of U^n at the faces' two sides.
bool is_artvisc(const mgpu_case &c) {
  return c.filter == MGPU_FILTER_ARTVISC || c.filter == MGPU_FILTER_ARTVISC_SHOCK;
}
// The spectral radii of U^n are stored (artificial viscosities, Edoh switch).
bool has_sr(const mgpu_case &c) { return is_artvisc(c) || edoh_on(c); }
// artvisc_o9 next to walls is artvisc_o9_shock without its sensor term: the
// faces 1..4 from a wall with d2..d8 times xnu2..xnu8 (mod_artvisc.f90:98-586)
// are its reduced faces with xk0 = 0.
DissipKind dissip_kind(const mgpu_case &c, bool wall) {
  if (c.filter == MGPU_FILTER_ARTVISC) return wall ? kDissipArtviscShock : kDissipArtvisc;
  if (c.filter == MGPU_FILTER_ARTVISC_SHOCK) return kDissipArtviscShock;
  return kDissipShock;
}

// Walls, with or without the slip flag.
bool is_wall(int kind) { return kind == MGPU_FACE_WALL || kind == MGPU_FACE_SLIP; }
// Boundary faces: no neighbour; the operators take their closures there.
bool is_bound(int kind) { return is_wall(kind) || kind == MGPU_FACE_CHARAC || kind == MGPU_FACE_INLET; }
// A slip wall: the bindings of assign_procedures.f90:415-470 -- with is_wall2,
// every wall of an inviscid case and the walls with the slip flag of an
// adiabatic one; an isothermal case binds no slip routine, and without
// is_wall2 the flag selects nothing.
bool is_slip(const mgpu_case &c, int kind) {
  if (!c.is_wall2 || !is_wall(kind)) return false;
  if (c.iorder_visc == 0) return true;
  return c.is_adiab && kind == MGPU_FACE_SLIP;
}
bool is_charac(int kind) { return kind == MGPU_FACE_CHARAC || kind == MGPU_FACE_INLET; }

// Characteristic outflow and supersonic inlet, checked on the caller's block
// table: bc_charac_imax / bc_charac_jmax at imax / jmax and the inlet at imin
// of Cartesian blocks whose other boundary faces are j-walls (so that no
// axis relabeling moves them), with the flags their kernels cover.  "" or an
// error.
std::string check_charac(const mgpu_case &c) {
  char buf[256];
  for (int b = 0; b < c.nblocks; b++) {
    const mgpu_block &A = c.blocks[b];
    bool any = false;
    for (int f = 0; f < 6; f++) {
      const int kind = A.face_kind[f];
      if (!is_charac(kind)) continue;
      any = true;
      if (kind == MGPU_FACE_CHARAC && f != 1 && f != 3) {
        snprintf(buf, sizeof buf, "block %d face %d: characteristic outflow is supported at imax and jmax (faces 1, 3)", b,
                 f);
        return buf;
      }
      if (kind == MGPU_FACE_INLET && f != 0) {
        snprintf(buf, sizeof buf, "block %d face %d: the supersonic inlet is supported at imin (face 0)", b, f);
        return buf;
      }
    }
    if (!any) continue;
    for (int f : {0, 1, 4, 5})
      if (is_wall(A.face_kind[f])) {
        snprintf(buf, sizeof buf, "block %d: walls with characteristic faces are supported on the j-faces only", b);
        return buf;
      }
    if (c.is_wall2) {
      snprintf(buf, sizeof buf, "block %d: characteristic faces with advanced walls (is_wall2) are not supported", b);
      return buf;
    }
    if (c.curvilinear) {
      snprintf(buf, sizeof buf, "block %d: characteristic faces need a Cartesian case", b);
      return buf;
    }
    if (c.filter != MGPU_FILTER_NONE && c.filter != MGPU_FILTER_SF && c.filter != MGPU_FILTER_SHOCK) {
      snprintf(buf, sizeof buf, "block %d: characteristic faces with filter %d are not supported", b, c.filter);
      return buf;
    }
    if (c.launch_flags & MGPU_LAUNCH_STAGE_PIPELINED) {
      snprintf(buf, sizeof buf, "block %d: characteristic faces run the classic stage kernel only", b);
      return buf;
    }
    if ((is_bound(A.face_kind[0]) || is_bound(A.face_kind[1])) && A.nx < 10) {
      snprintf(buf, sizeof buf, "block %d: %d cells between its i boundary faces, >= 10 needed", b, A.nx);
      return buf;
    }
    if ((is_bound(A.face_kind[2]) || is_bound(A.face_kind[3])) && A.ny < 10) {
      snprintf(buf, sizeof buf, "block %d: %d cells between its j boundary faces, >= 10 needed", b, A.ny);
      return buf;
    }
  }
  return std::string();
}

// Walls of 3-D curvilinear blocks: on their j-faces only (the axis
// relabeling that turns i- or k-walls into j-walls covers Cartesian metrics
// alone), checked on the caller's block table.  "" or an error.
std::string check_c3_walls(const mgpu_case &c) {
  char buf[256];
  if (c.curvilinear != MGPU_GEOM_C3) return std::string();
  for (int b = 0; b < c.nblocks; b++) {
    int axes = 0;
    for (int a = 0; a < 3; a++)
      if (is_wall(c.blocks[b].face_kind[2 * a]) || is_wall(c.blocks[b].face_kind[2 * a + 1])) axes++;
    if (axes > 1) {
      snprintf(buf, sizeof buf, "block %d: 3-D curvilinear blocks take walls on the faces of one axis", b);
      return buf;
    }
  }
  return std::string();
}

Metrics metrics_of(const BlockData &bd) {
  Metrics m{};
  m.idx = bd.idx[0];
  m.idy = bd.idx[1];
  m.idz = bd.idx[2];
  m.idx_v = bd.idx[3];
  m.idy_v = bd.idx[4];
  m.idz_v = bd.idx[5];
  for (int q = 0; q < 10; q++) {
    m.m[q] = bd.m[q];
    m.mv[q] = bd.mv[q];
  }
  for (int q = 0; q < 5; q++) {
    m.m2[q] = bd.m2[q];
    m.m2v[q] = bd.m2v[q];
  }
  return m;
}

// Ghost layers of state set `set` of every block from the neighbours.
int exchange_state(mgpu_handle *h, int set) {
  ProfScope ps(h, MGPU_K_HALO);
  HIPCK(h, launch_iface(h->xs[set], h->nchunks_s, h->stream));
  return 0;
}

// Neighbour of face f of block b: (block, face, orientation code).
void face_nbr(const mgpu_block &bl, int b, int f, int &nb, int &nf, int &o) {
  if (bl.face_kind[f] == MGPU_FACE_PERIODIC) {
    nb = b;
    nf = f ^ 1;
    o = 0;
  } else {
    nb = bl.nbr_block[f];
    nf = bl.nbr_face[f];
    o = bl.orient[f];
  }
}

// Axes of face f: normal axis, then the two tangential axes in increasing order.
void face_axes(int f, int &a, int &t1, int &t2) {
  a = f / 2;
  t1 = a == 0 ? 1 : 0;
  t2 = a == 2 ? 1 : 2;
}

// Checks the block table: neighbours exist, faces match in extent under the
// orientation, and every interface is declared from both sides with the
// inverse orientation.  Returns an error message or an empty string.
// ---- axis relabeling of blocks with walls on the i- or k-faces ----
// face f of the caller's block is face kFaceMap[perm][f] of the stored block;
// the caller's axis a is the stored axis kAxisMap[perm][a]; momentum v of the
// caller is stored as kVarMap[perm][v].  Perms 1 and 2 exchange two axes (each
// its own inverse); perm 3 cycles them (the caller's j, k, i are stored as i,
// j, k), so the stored i-face before the stored j-face keeps bc_wall's face
// order at the edges; caller_axis / caller_face / caller_var invert the maps.
constexpr int kFaceMap[4][6] = {{0, 1, 2, 3, 4, 5}, {2, 3, 0, 1, 4, 5}, {0, 1, 4, 5, 2, 3}, {4, 5, 0, 1, 2, 3}};
constexpr int kAxisMap[4][3] = {{0, 1, 2}, {1, 0, 2}, {0, 2, 1}, {2, 0, 1}};
constexpr int kVarMap[4][5] = {{0, 1, 2, 3, 4}, {0, 2, 1, 3, 4}, {0, 1, 3, 2, 4}, {0, 3, 1, 2, 4}};

// The caller's axis / face / variable stored as s.
int caller_axis(int perm, int s) {
  for (int a = 0; a < 3; a++)
    if (kAxisMap[perm][a] == s) return a;
  return -1;
}
int caller_face(int perm, int s) {
  for (int f = 0; f < 6; f++)
    if (kFaceMap[perm][f] == s) return f;
  return -1;
}
int caller_var(int perm, int s) {
  for (int v = 0; v < 5; v++)
    if (kVarMap[perm][v] == s) return v;
  return -1;
}

// 0: walls (if any) on the j-faces, or on the i- and j-faces; 1: on the
// i-faces alone; 2: on the k-faces alone, or on the i- and k-faces (stored as
// i and j); 3: on the j- and k-faces (stored as i and j); -1: walls on all
// three axes.
int wall_perm(const mgpu_block &b) {
  bool w[3];
  for (int a = 0; a < 3; a++) w[a] = is_wall(b.face_kind[2 * a]) || is_wall(b.face_kind[2 * a + 1]);
  if (w[0] && w[1] && w[2]) return -1;
  if (w[1] && w[2]) return 3;
  if (w[0] && !w[1] && !w[2]) return 1;
  if (w[2]) return 2;
  return 0;
}

// The stored block's tangential axes of a face (the two other axes in
// increasing order) come out in the caller's reverse order: the caller's
// k-faces under perm 1, i-faces under perm 2, j- and k-faces under perm 3.
bool tangential_flipped(int perm, int f) {
  const int a = f / 2, p = a == 0 ? 1 : 0, q = a == 2 ? 1 : 2;
  return kAxisMap[perm][p] > kAxisMap[perm][q];
}

// Orientation code of an interface after relabeling: the caller's code maps
// this face's tangential coordinates (p, q) to the neighbour's (s1, s2) by a
// signed permutation; the stored faces see (p, q) and (s1, s2) in swapped
// order where flipped.  The composed signed permutation is re-encoded.
int relabel_orient(int code, bool flip_self, bool flip_nbr) {
  // M maps (p, q) to (s1, s2): rows of a signed permutation matrix
  int M[2][2] = {{0, 0}, {0, 0}};
  const bool sw = code & 1;
  M[0][sw ? 1 : 0] = (code & 2) ? -1 : 1;
  M[1][sw ? 0 : 1] = (code & 4) ? -1 : 1;
  int C[2][2];
  for (int r = 0; r < 2; r++)
    for (int col = 0; col < 2; col++) {
      const int rr = flip_nbr ? 1 - r : r, cc = flip_self ? 1 - col : col;
      C[r][col] = M[rr][cc];
    }
  const bool sw2 = C[0][0] == 0;
  const int s1 = sw2 ? C[0][1] : C[0][0], s2 = sw2 ? C[1][0] : C[1][1];
  return (sw2 ? 1 : 0) | (s1 < 0 ? 2 : 0) | (s2 < 0 ? 4 : 0);
}

// The stored block table of a case (see BlockData::perm); "" or an error.
std::string relabel_blocks(const mgpu_case &c, std::vector<mgpu_block> &out, std::vector<int> &perm) {
  char buf[256];
  const int nb = c.nblocks;
  out.assign(c.blocks, c.blocks + nb);
  perm.assign(nb, 0);
  for (int b = 0; b < nb; b++) {
    perm[b] = wall_perm(c.blocks[b]);
    if (perm[b] < 0) {
      snprintf(buf, sizeof buf,
               "block %d: walls on all three axes are not supported", b);
      return buf;
    }
  }
  for (int b = 0; b < nb; b++) {
    const mgpu_block &U = c.blocks[b];
    mgpu_block &S = out[b];
    const int ext[3] = {U.nx, U.ny, U.nz};
    int se[3];
    for (int a = 0; a < 3; a++) se[kAxisMap[perm[b]][a]] = ext[a];
    S.nx = se[0];
    S.ny = se[1];
    S.nz = se[2];
    for (int fu = 0; fu < 6; fu++) {
      const int fs = kFaceMap[perm[b]][fu];
      S.face_kind[fs] = U.face_kind[fu];
      S.nbr_block[fs] = U.nbr_block[fu];
      S.nbr_face[fs] = U.nbr_face[fu];
      S.orient[fs] = U.orient[fu];
      if (U.face_kind[fu] != MGPU_FACE_INTERFACE) continue;
      const int n = U.nbr_block[fu], g = U.nbr_face[fu];
      if (n < 0 || n >= nb || g < 0 || g > 5) continue;  // reported by check_topology
      S.nbr_face[fs] = kFaceMap[perm[n]][g];
      S.orient[fs] = relabel_orient(U.orient[fu], tangential_flipped(perm[b], fu), tangential_flipped(perm[n], g));
    }
  }
  return "";
}

// On a Cartesian case each block's index axes are the physical axes (the
// fluxes tie rho u to i through idx, etc.), so an interface can only join
// faces of the same normal axis without exchanging the tangential axes;
// reversed axes are fine (the metrics of a reversed block are negative).
// Checked on the caller's block table.  "" or an error.
std::string check_cartesian_frames(const mgpu_case &c) {
  char buf[256];
  if (c.curvilinear) return "";
  for (int b = 0; b < c.nblocks; b++) {
    const mgpu_block &A = c.blocks[b];
    for (int f = 0; f < 6; f++) {
      if (A.face_kind[f] != MGPU_FACE_INTERFACE) continue;
      const int nf = A.nbr_face[f], o = A.orient[f];
      if (nf < 0 || nf > 5) continue;  // reported by check_topology
      if (nf / 2 != f / 2 || (o & 1)) {
        snprintf(buf, sizeof buf,
                 "block %d face %d: on a Cartesian case an interface must join faces of the same axis without "
                 "exchanging the tangential axes (neighbour face %d, orient %d)",
                 b, f, nf, o);
        return buf;
      }
    }
  }
  return "";
}

// A caller's padded array of block extents (as the caller sees them) to the
// stored layout (to_stored) or back.
void relabel_array(const Geom &g, int perm, const double *src, double *dst, bool to_stored) {
  const long Q[3] = {g.nx + 2 * G, g.ny + 2 * G, g.nz + 2 * G};
  const int *m = kAxisMap[perm];
  const long P0 = Q[m[0]], P1 = Q[m[1]];  // the caller's padded extents along its i and j
  long x[3];
  for (x[2] = 0; x[2] < Q[2]; x[2]++)
    for (x[1] = 0; x[1] < Q[1]; x[1]++)
      for (x[0] = 0; x[0] < Q[0]; x[0]++) {
        const long s = x[0] + Q[0] * (x[1] + Q[1] * x[2]), u = x[m[0]] + P0 * (x[m[1]] + P1 * x[m[2]]);
        if (to_stored) dst[s] = src[u];
        else dst[u] = src[s];
      }
}

std::string check_topology(const mgpu_case &c) {
  char buf[256];
  for (int b = 0; b < c.nblocks; b++) {
    const mgpu_block &A = c.blocks[b];
    const int na[3] = {A.nx, A.ny, A.nz};
    for (int f = 0; f < 6; f++) {
      const int kind = A.face_kind[f];
      if (kind != MGPU_FACE_PERIODIC && kind != MGPU_FACE_INTERFACE && !is_bound(kind)) {
        snprintf(buf, sizeof buf, "block %d face %d: face kind %d not supported", b, f, kind);
        return buf;
      }
      if (is_charac(kind)) continue;  // check_charac
      if (is_wall(kind)) {
        // the stored table: walls on the j-faces, or on the i- and j-faces
        if (f > 3 || (f < 2 && !(is_wall(A.face_kind[2]) || is_wall(A.face_kind[3])))) {
          snprintf(buf, sizeof buf, "block %d face %d: walls are supported on the j-faces (2, 3), or on the "
                   "i- and j-faces together", b, f);
          return buf;
        }
        if (f < 2 && c.curvilinear) {
          snprintf(buf, sizeof buf, "block %d face %d: walls on two axes need a Cartesian case", b, f);
          return buf;
        }
        if (f < 2 && (c.filter == MGPU_FILTER_ARTVISC || c.filter == MGPU_FILTER_ARTVISC_SHOCK)) {
          snprintf(buf, sizeof buf, "block %d: walls on two axes with the artificial viscosities are not supported",
                   b);
          return buf;
        }
        if (f < 2 && A.nx < 10) {
          snprintf(buf, sizeof buf, "block %d: %d cells between its i-walls, a block with walls needs >= 10", b,
                   A.nx);
          return buf;
        }
        if (c.curvilinear == MGPU_GEOM_C2) {
          snprintf(buf, sizeof buf, "block %d face %d: walls need a Cartesian or a 3-D curvilinear case", b, f);
          return buf;
        }
        if (A.ny < 10) {
          snprintf(buf, sizeof buf, "block %d: %d cells between its walls, a block with walls needs >= 10", b,
                   A.ny);
          return buf;
        }
        continue;
      }
      if (kind == MGPU_FACE_PERIODIC) {
        if (A.face_kind[f ^ 1] != MGPU_FACE_PERIODIC) {
          snprintf(buf, sizeof buf, "block %d face %d: periodic face without a periodic opposite face", b, f);
          return buf;
        }
        continue;
      }
      int nb, nf, o;
      face_nbr(A, b, f, nb, nf, o);
      if (nb < 0 || nb >= c.nblocks || nf < 0 || nf > 5 || o < 0 || o > 7) {
        snprintf(buf, sizeof buf, "block %d face %d: neighbour (%d, face %d, orient %d) out of range", b, f, nb,
                 nf, o);
        return buf;
      }
      const mgpu_block &B = c.blocks[nb];
      const int nbx[3] = {B.nx, B.ny, B.nz};
      int a, t1, t2, bb, u1, u2;
      face_axes(f, a, t1, t2);
      face_axes(nf, bb, u1, u2);
      const int m1 = (o & 1) ? u2 : u1, m2 = (o & 1) ? u1 : u2;
      if (na[t1] != nbx[m1] || na[t2] != nbx[m2]) {
        snprintf(buf, sizeof buf, "block %d face %d: extents %dx%d do not match face %d of block %d under orient %d",
                 b, f, na[t1], na[t2], nf, nb, o);
        return buf;
      }
      if (B.face_kind[nf] != MGPU_FACE_INTERFACE || B.nbr_block[nf] != b || B.nbr_face[nf] != f) {
        snprintf(buf, sizeof buf, "block %d face %d: face %d of block %d does not point back", b, f, nf, nb);
        return buf;
      }
      const int swap = o & 1, r1 = (o >> 1) & 1, r2 = (o >> 2) & 1;
      const int inv = swap | ((swap ? r2 : r1) << 1) | ((swap ? r1 : r2) << 2);
      if (B.orient[nf] != inv) {
        snprintf(buf, sizeof buf, "block %d face %d: orient %d, but the neighbour declares %d (expected %d)", b, f,
                 o, B.orient[nf], inv);
        return buf;
      }
    }
  }
  return std::string();
}

// One affine coordinate: c0 + cl*l + cp*p + cq*q (l, p, q from 0).
struct Aff {
  long c0, cl, cp, cq;
};

// Ghost point (l, p, q) of face f (layer l from the face, p and q along the
// face's first and second tangential axes, all from 0) as block coordinates
// d[axis] of the destination block, and the neighbour point it copies as
// coordinates sr[axis] of the neighbour (face nf, orientation o), each an
// affine function c0 + cl l + cp p + cq q.  na / nbx: the two blocks' extents.
void face_affine(int f, int nf, int o, const int na[3], const int nbx[3], Aff d[3], Aff sr[3]) {
  int a, t1, t2, bb, u1, u2;
  face_axes(f, a, t1, t2);
  face_axes(nf, bb, u1, u2);
  d[a] = (f & 1) ? Aff{na[a] + 1, 1, 0, 0} : Aff{0, -1, 0, 0};
  d[t1] = Aff{1, 0, 1, 0};
  d[t2] = Aff{1, 0, 0, 1};
  sr[bb] = (nf & 1) ? Aff{nbx[bb], -1, 0, 0} : Aff{1, 1, 0, 0};
  // neighbour tangential coordinates: s1 = swap ? q : p, s2 = swap ? p : q
  Aff s1 = (o & 1) ? Aff{0, 0, 0, 1} : Aff{0, 0, 1, 0};
  Aff s2 = (o & 1) ? Aff{0, 0, 1, 0} : Aff{0, 0, 0, 1};
  sr[u1] = (o & 2) ? Aff{nbx[u1] - s1.c0, -s1.cl, -s1.cp, -s1.cq} : Aff{1 + s1.c0, s1.cl, s1.cp, s1.cq};
  sr[u2] = (o & 4) ? Aff{nbx[u2] - s2.c0, -s2.cl, -s2.cp, -s2.cq} : Aff{1 + s2.c0, s2.cl, s2.cp, s2.cq};
}

// Interfaces must join wall rows to wall rows.  On the row of a no-slip wall
// the viscous kernels exchange tau/mu and the heat-flux gradient, and every
// receiving point of that row rebuilds its flux from its own primitives
// (the reference evaluates the flux at a face ghost from the ghost's own
// primitives, which hold the neighbour's state from before the wall state was
// imposed).  A wall row facing a row that is not a wall row on the other
// side (a wall that ends at the interface, the neighbour being periodic
// there) would mix the two forms, so such a table is refused.  c: the blocks
// as stored (walls on the j-faces, or on the i- and j-faces).
std::string check_wall_rows(const mgpu_case &c) {
  char buf[320];
  // wall bits of a block along axis ax (0: i, 1: j)
  auto wbits = [](const mgpu_block &X, int ax) {
    return (is_wall(X.face_kind[2 * ax]) ? 1 : 0) | (is_wall(X.face_kind[2 * ax + 1]) ? 2 : 0);
  };
  for (int b = 0; b < c.nblocks; b++) {
    const mgpu_block &A = c.blocks[b];
    const int na[3] = {A.nx, A.ny, A.nz};
    for (int f = 0; f < 6; f++) {
      if (is_bound(A.face_kind[f])) continue;
      int nb, nf, o;
      face_nbr(A, b, f, nb, nf, o);
      const mgpu_block &B = c.blocks[nb];
      const int nbx[3] = {B.nx, B.ny, B.nz};
      int a, t1, t2;
      face_axes(f, a, t1, t2);
      for (int ax = 0; ax < 2; ax++) {
        if (ax == a) continue;  // the face's own axis: its ghost layers are no rows of it
        const int wa = wbits(A, ax), wb = wbits(B, ax);
        if (!wa && !wb) continue;
        Aff d[3] = {}, sr[3] = {};
        face_affine(f, nf, o, na, nbx, d, sr);
        for (int p = 0; p < na[t1]; p++)
          for (int q = 0; q < na[t2]; q++) {
            const long dj = d[ax].c0 + d[ax].cp * p + d[ax].cq * q;  // layer 0
            const long sj = sr[ax].c0 + sr[ax].cp * p + sr[ax].cq * q;
            const bool dw = ((wa & 1) && dj == 1) || ((wa & 2) && dj == na[ax]);
            const bool sw = ((wb & 1) && sj == 1) || ((wb & 2) && sj == nbx[ax]);
            if (dw != sw) {
              snprintf(buf, sizeof buf,
                       "block %d face %d: a no-slip wall row of block %d meets a row of block %d that is not a wall "
                       "row (walls must continue across an interface; not supported)",
                       b, f, dw ? b : nb, dw ? nb : b);
              return buf;
            }
          }
      }
    }
  }
  return std::string();
}

// Every element a task touches lies inside the padded arrays: the offsets
// are affine in the box indices, so their extremes sit at the box corners.
bool task_in_bounds(const FaceTask &t, long ndst, long nsrc) {
  for (int c = 0; c < 8; c++) {
    long pd = t.dst_base, ps = t.src_base;
    for (int x = 0; x < 3; x++) {
      const long a = ((c >> x) & 1) ? t.n[x] - 1 : 0;
      pd += a * t.dst_s[x];
      ps += a * t.src_s[x];
    }
    if (pd < 0 || pd >= ndst || ps < 0 || ps >= nsrc) return false;
  }
  return true;
}

// Arrays an exchange moves: the 5 conserved variables, or the viscous set
// (the NTAU components of tau, then p and the Ducros sensor when `sensor`,
// then the 3 spectral radii when `radii`).
struct XSet {
  bool visc = false, sensor = false, radii = false;
  int nvar() const { return visc ? NTAU + (sensor ? 2 : 0) + (radii ? 3 : 0) : 5; }
};

// Stored axis of block n holding the caller's axis that block b stores as s.
int axis_between(int pb, int pn, int s) { return kAxisMap[pn][caller_axis(pb, s)]; }

// Index of the symmetric tensor component (s, t) in tau[0..5] (xx, xy, xz, yy, yz, zz).
int tau_index(int s, int t) {
  static const int I[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
  return I[s][t];
}
const int kTauPair[6][2] = {{0, 0}, {0, 1}, {0, 2}, {1, 1}, {1, 2}, {2, 2}};

// Source array of each destination array of a face from block b (perm pb,
// face axes a,t1,t2) to block n (perm pn; the axes bb, m1, m2 of its face
// that meet a, t1, t2): the conserved variables and tau follow the caller's
// frame through both blocks' axis relabelings; the spectral radius along an
// index axis is that of the neighbour's index axis it meets (on c3 and 2.5-D
// blocks a swapping interface exchanges them; the radius has no sign).
// vb, vn: the blocks' relabelings of the physical axes (their perms on a
// Cartesian case, none on a curvilinear one, whose momentum and tau stay in
// the caller's x, y, z while only the index axes move).
void face_vmap(const XSet &xs, int vb, int vn, const int dst_ax[3], const int src_ax[3], int8_t vmap[kIfaceMaxVar]) {
  for (int v = 0; v < kIfaceMaxVar; v++) vmap[v] = (int8_t)v;
  if (!xs.visc) {
    for (int v = 0; v < 5; v++) vmap[v] = (int8_t)kVarMap[vn][caller_var(vb, v)];
    return;
  }
  for (int q = 0; q < 6; q++)
    vmap[q] = (int8_t)tau_index(axis_between(vb, vn, kTauPair[q][0]), axis_between(vb, vn, kTauPair[q][1]));
  for (int d = 0; d < 3; d++) vmap[6 + d] = (int8_t)(6 + axis_between(vb, vn, d));
  if (xs.radii) {
    const int off = NTAU + (xs.sensor ? 2 : 0);
    int meet[3];
    for (int x = 0; x < 3; x++) meet[dst_ax[x]] = src_ax[x];
    for (int d = 0; d < 3; d++) vmap[off + d] = (int8_t)(off + meet[d]);
  }
}

// Face tasks of every face of every block for `ng` ghost layers, and the
// chunk table that maps workgroups onto them.  False if a task would reach
// outside its arrays (a block table that passed check_topology never does)
// or if, on a Cartesian case, a face's orientation does not meet the
// neighbour's axes as the two blocks' axis relabelings say it must.
bool build_tasks(const mgpu_case &c, const std::vector<BlockData> &bds, int ng, const XSet &xs,
                 std::vector<FaceTask> &tasks, std::vector<IfaceChunk> &chunks) {
  tasks.clear();
  chunks.clear();
  for (int b = 0; b < c.nblocks; b++) {
    const mgpu_block &A = c.blocks[b];
    const Geom &ga = bds[b].g;
    const int na[3] = {ga.nx, ga.ny, ga.nz};
    for (int f = 0; f < 6; f++) {
      if (is_bound(A.face_kind[f])) continue;
      int nb, nf, o;
      face_nbr(A, b, f, nb, nf, o);
      const Geom &gb = bds[nb].g;
      const int nbx[3] = {gb.nx, gb.ny, gb.nz};
      int a, t1, t2, bb, u1, u2;
      face_axes(f, a, t1, t2);
      face_axes(nf, bb, u1, u2);
      Aff d[3] = {}, sr[3] = {};
      face_affine(f, nf, o, na, nbx, d, sr);
      const long sta[3] = {1, ga.sx, (long)ga.sy}, stb[3] = {1, gb.sx, (long)gb.sy};
      long db = 0, sb = 0, dv[3] = {0, 0, 0}, sv[3] = {0, 0, 0};
      for (int x = 0; x < 3; x++) {
        db += (d[x].c0 - 1 + G) * sta[x];
        sb += (sr[x].c0 - 1 + G) * stb[x];
        dv[0] += d[x].cl * sta[x];
        dv[1] += d[x].cp * sta[x];
        dv[2] += d[x].cq * sta[x];
        sv[0] += sr[x].cl * stb[x];
        sv[1] += sr[x].cp * stb[x];
        sv[2] += sr[x].cq * stb[x];
      }
      const int ext[3] = {ng, na[t1], na[t2]};
      // box axis 0 = the variable with the smallest destination stride
      int ord[3] = {0, 1, 2};
      for (int x = 0; x < 3; x++)
        for (int y = x + 1; y < 3; y++)
          if (labs(dv[ord[y]]) < labs(dv[ord[x]])) std::swap(ord[x], ord[y]);
      FaceTask t{};
      t.dst = b;
      t.src = nb;
      t.dst_base = (int)db;
      t.src_base = (int)sb;
      const int dst_ax[3] = {a, t1, t2};
      const int src_ax[3] = {bb, (o & 1) ? u2 : u1, (o & 1) ? u1 : u2};
      if (!c.curvilinear)
        for (int x = 0; x < 3; x++)
          if (src_ax[x] != axis_between(bds[b].perm, bds[nb].perm, dst_ax[x])) return false;
      face_vmap(xs, bds[b].vperm, bds[nb].vperm, dst_ax, src_ax, t.vmap);
      for (int x = 0; x < 3; x++) {
        t.n[x] = ext[ord[x]];
        t.dst_s[x] = (int)dv[ord[x]];
        t.src_s[x] = (int)sv[ord[x]];
      }
      if (!task_in_bounds(t, ga.n, gb.n)) return false;
      const int cnt = t.n[0] * t.n[1] * t.n[2];
      const int id = (int)tasks.size();
      tasks.push_back(t);
      for (int st = 0; st < cnt; st += kIfaceThreads) chunks.push_back({id, st});
    }
  }
  return true;
}

template <class T>
int upload(mgpu_handle *h, const std::vector<T> &v, T **out) {
  *out = nullptr;
  if (v.empty()) return 0;
  void *p = nullptr;
  HIPCK(h, hipMalloc(&p, v.size() * sizeof(T)));
  h->xtables.push_back(p);
  HIPCK(h, hipMemcpy(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  *out = (T *)p;
  return 0;
}

void free_exchange(mgpu_handle *h) {
  for (void *p : h->xtables) (void)hipFree(p);
  h->xtables.clear();
  for (auto &x : h->xs) x = IfaceArgs{};
  for (auto &x : h->xf) x = IfaceArgs32{};
  h->xv = IfaceArgs{};
  h->nchunks_s = h->nchunks_v = 0;
  h->exchange_ready = false;
}

// Task, chunk and pointer tables of the interface exchange (rebuilt when the
// set of blocks changes; the stream is idle then).
int setup_exchange(mgpu_handle *h) {
  if (h->exchange_ready) return 0;
  HIPCK(h, hipStreamSynchronize(h->stream));
  free_exchange(h);
  h->c.blocks = h->blocks.data();
  h->c.nblocks = (int)h->blocks.size();
  if (h->b.empty()) {
    h->exchange_ready = true;
    return 0;
  }
  std::vector<FaceTask> tasks;
  std::vector<IfaceChunk> chunks;
  const int nb = (int)h->b.size();
  int rc;
  // state: 5 layers, 5 variables, one pointer table per state set
  if (!build_tasks(h->c, h->b, G, XSet{}, tasks, chunks))
    return fail(h, MGPU_EINVAL,
                "interface exchange would reach outside a block's arrays, or a face meets its neighbour's axes "
                "otherwise than the blocks' axis relabelings require");
  FaceTask *dt = nullptr;
  IfaceChunk *dc = nullptr;
  if ((rc = upload(h, tasks, &dt)) || (rc = upload(h, chunks, &dc))) return rc;
  h->nchunks_s = (int)chunks.size();
  for (int s = 0; s < 3; s++) {
    std::vector<double *> ptrs;
    for (int b = 0; b < nb; b++)
      for (int v = 0; v < 5; v++) ptrs.push_back(h->b[b].st[s][v]);
    double **dp = nullptr;
    if ((rc = upload(h, ptrs, &dp))) return rc;
    h->xs[s] = IfaceArgs{dt, dc, dp, 5};
  }
  if (h->mixed) {
    for (int s = 0; s < 2; s++) {
      std::vector<float *> ptrs;
      for (int b = 0; b < nb; b++)
        for (int v = 0; v < 5; v++) ptrs.push_back(h->b[b].sf[s][v]);
      float **dp = nullptr;
      if ((rc = upload(h, ptrs, &dp))) return rc;
      h->xf[s] = IfaceArgs32{dt, dc, dp, 5};
    }
  }
  // viscous: 2 layers of tau (+ p and the sensor for the shock filter)
  if (h->c.iorder_visc) {
    const bool shock = has_sensor(h->c), artv = is_artvisc(h->c);
    XSet xv;
    xv.visc = true;
    xv.sensor = shock;
    xv.radii = artv;
    const int nv = xv.nvar();
    if (!build_tasks(h->c, h->b, 2, xv, tasks, chunks))
      return fail(h, MGPU_EINVAL, "viscous interface exchange would reach outside a block's arrays");
    if ((rc = upload(h, tasks, &dt)) || (rc = upload(h, chunks, &dc))) return rc;
    h->nchunks_v = (int)chunks.size();
    std::vector<double *> ptrs;
    for (int b = 0; b < nb; b++) {
      for (int q = 0; q < NTAU; q++) ptrs.push_back(h->b[b].tau[q]);
      if (shock) {
        ptrs.push_back(h->b[b].prs);
        ptrs.push_back(h->b[b].sens);
      }
      if (artv)
        for (int q = 0; q < 3; q++) ptrs.push_back(h->b[b].sr[q]);
    }
    double **dp = nullptr;
    if ((rc = upload(h, ptrs, &dp))) return rc;
    h->xv = IfaceArgs{dt, dc, dp, nv};
  }
  h->exchange_ready = true;
  return 0;
}

// Default k-chunk: aim at >= 2048 workgroups, at least 32 planes per chunk
// (the 10-plane queue warm-up is paid once per chunk), but at least 8 planes
// only on blocks too small to give every CU a workgroup otherwise.
int kchunk_for(const Geom &g, int shape) {
  const int TX = kStageShapes[shape][0], TY = kStageShapes[shape][1];
  const long nxy = (long)((g.nx + TX - 1) / TX) * ((g.ny + TY - 1) / TY);
  long want = (2048 + nxy - 1) / nxy;
  if (want < 1) want = 1;
  long kc = (g.nz + want - 1) / want;
  const long kmin = nxy * ((g.nz + 31) / 32) >= 512 ? 32 : 8;
  if (kc < kmin) kc = kmin;
  if (kc > g.nz) kc = g.nz;
  return (int)kc;
}

// Process-wide cache of tuned stage parameters, keyed by block geometry and
// kernel variant, so that equal blocks (and later handles) are tuned once.
struct TuneKey {
  int dev, curv, visc, family;
  std::vector<int> dims;  // nx,ny,nz of every block of the batch
  bool operator<(const TuneKey &o) const {
    return std::tie(dev, curv, visc, family, dims) < std::tie(o.dev, o.curv, o.visc, o.family, o.dims);
  }
};
std::mutex g_tune_mu;
std::map<TuneKey, std::pair<int, int>> g_tune_cache;

void free_block(BlockData &bd) {
  for (auto &s : bd.st)
    for (auto &p : s) { if (p) (void)hipFree(p); p = nullptr; }
  for (auto &s : bd.sf)
    for (auto &p : s) { if (p) (void)hipFree(p); p = nullptr; }
  for (auto &p : bd.idx) { if (p) (void)hipFree(p); p = nullptr; }
  for (auto &p : bd.m) { if (p) (void)hipFree(p); p = nullptr; }
  for (auto &p : bd.m32) { if (p) (void)hipFree(p); p = nullptr; }
  for (auto &p : bd.mv) { if (p) (void)hipFree(p); p = nullptr; }
  for (auto &p : bd.m2) { if (p) (void)hipFree(p); p = nullptr; }
  for (auto &p : bd.m2v) { if (p) (void)hipFree(p); p = nullptr; }
  for (auto &p : bd.tau) { if (p) (void)hipFree(p); p = nullptr; }
  for (auto &p : bd.kv) { if (p) (void)hipFree(p); p = nullptr; }
  if (bd.prs) (void)hipFree(bd.prs);
  if (bd.sens) (void)hipFree(bd.sens);
  if (bd.wstash) (void)hipFree(bd.wstash);
  bd.prs = bd.sens = bd.wstash = nullptr;
  for (double **q : {&bd.kimax, &bd.kjmax, &bd.cst, &bd.uin, &bd.ppre, &bd.kf, &bd.wc3[0], &bd.wc3[1], &bd.wt3[0],
                     &bd.wt3[1]}) {
    if (*q) (void)hipFree(*q);
    *q = nullptr;
  }
  for (auto &p : bd.sr) { if (p) (void)hipFree(p); p = nullptr; }
}

int dalloc(mgpu_handle *h, double **p, size_t count) {
  hipError_t e = hipMalloc((void **)p, count * sizeof(double));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(h, MGPU_ENOMEM, "hipMalloc of %zu doubles failed: %s", count, hipGetErrorString(e));
  }
  e = hipMemset(*p, 0, count * sizeof(double));
  if (e != hipSuccess) return fail(h, MGPU_EDEVICE, "hipMemset failed: %s", hipGetErrorString(e));
  return 0;
}

int dalloc32(mgpu_handle *h, float **p, size_t count) {
  hipError_t e = hipMalloc((void **)p, count * sizeof(float));
  if (e != hipSuccess) {
    *p = nullptr;
    return fail(h, MGPU_ENOMEM, "hipMalloc of %zu floats failed: %s", count, hipGetErrorString(e));
  }
  e = hipMemset(*p, 0, count * sizeof(float));
  if (e != hipSuccess) return fail(h, MGPU_EDEVICE, "hipMemset failed: %s", hipGetErrorString(e));
  return 0;
}

// Stage kernel variant of stage irk.
int stage_mode(const mgpu_handle *h, int irk) {
  if (!h->mixed) return kFP64;
  return irk == 1 ? kMixFirst : (irk == h->c.nrk ? kMixLast : kMixMiddle);
}

// FP32 set written by intermediate stage irk (mixed precision).
int fset_of(int irk) { return (irk % 2 == 1) ? 0 : 1; }

StageArgs stage_args(const mgpu_handle *h, const BlockData &bd, int in, int s0, int out, int irk) {
  (void)h;
  StageArgs sa{};
  sa.g = bd.g;
  sa.gas = gas_of(bd.p);
  for (int v = 0; v < 5; v++) {
    sa.uin.v[v] = bd.st[in][v];
    sa.u0.v[v] = bd.st[s0][v];
    sa.uout.v[v] = bd.st[out][v];
  }
  sa.met = metrics_of(bd);
  for (int q = 0; q < 4; q++) sa.kv[q] = bd.kv[q];
  sa.alpha = bd.p.deltat * bd.p.crk[irk - 1];
  sa.ref = bd.ref;
  // boundary rows: the inlet row keeps the filter increment, the
  // characteristic rows their routines' increments; walls K = 0
  sa.kbc[0] = bd.kf;
  sa.kbc[1] = bd.kimax;
  sa.kbc[2] = nullptr;
  sa.kbc[3] = bd.kjmax;
  sa.uin_row = bd.uin;
  sa.prim_row = irk == 1 ? bd.ppre : nullptr;
  for (int q = 0; q < 9; q++) sa.m32[q] = bd.m32[q];
  if (h->mixed) {
    const int mode = stage_mode(h, irk);
    for (int v = 0; v < 5; v++) {
      if (mode == kMixMiddle || mode == kMixLast) sa.uin32[v] = bd.sf[fset_of(irk - 1)][v];
      if (mode == kMixFirst || mode == kMixMiddle) sa.uout32[v] = bd.sf[fset_of(irk)][v];
    }
  }
  return sa;
}

// Stage tiles of every block for variant v, numbered consecutively; returns
// the launch's tile count.
int stage_ranges(const mgpu_handle *h, int v, std::vector<TileRange> &tr) {
  tr.clear();
  int first = 0;
  for (const auto &bd : h->b) {
    const int kc = h->kchunk[v] > 0 ? h->kchunk[v] : kchunk_for(bd.g, h->shape[v]);
    TileRange t = stage_tiles(bd.g, h->shape[v], kc);
    t.jfirst = h->jfirst[v];
    t.first = first;
    first += tile_count(t, bd.g.nz);
    tr.push_back(t);
  }
  return first;
}

template <class T>
int dalloc_table(mgpu_handle *h, T **p, size_t count) {
  if (*p) (void)hipFree(*p);
  *p = nullptr;
  void *q = nullptr;
  HIPCK(h, hipMalloc(&q, count * sizeof(T)));
  *p = (T *)q;
  return 0;
}

void free_tables(mgpu_handle *h) {
  if (h->d_stage) (void)hipFree(h->d_stage);
  if (h->d_visc) (void)hipFree(h->d_visc);
  if (h->d_vdiv) (void)hipFree(h->d_vdiv);
  if (h->d_filt) (void)hipFree(h->d_filt);
  if (h->d_tune) (void)hipFree(h->d_tune);
  if (h->d_tune_f) (void)hipFree(h->d_tune_f);
  h->d_tune_f = nullptr;
  h->d_stage = nullptr;
  h->d_visc = nullptr;
  h->d_vdiv = nullptr;
  h->d_filt = nullptr;
  h->d_tune = nullptr;
  h->tab_cap = 0;
}

void drop_graphs(mgpu_handle *h) {
  for (auto &g : h->gexec) {
    if (g) (void)hipGraphExecDestroy(g);
    g = nullptr;
  }
}

FilterArgs filter_args(const mgpu_handle *h, const BlockData &bd, int in, int out) {
  const double d11[6] = {D11_0, D11_1, D11_2, D11_3, D11_4, D11_5};
  FilterArgs fa{};
  fa.g = bd.g;
  for (int v = 0; v < 5; v++) {
    fa.uin.v[v] = bd.st[in][v];
    fa.uout.v[v] = bd.st[out][v];
  }
  for (int l = 0; l < 6; l++) fa.dd[l] = -d11[l] * bd.p.dissip_coeff;
  fa.chi = bd.p.dissip_coeff;
  fa.prs = bd.prs;
  fa.sens = h->c.is_ducros ? bd.sens : nullptr;
  fa.xnu0 = h->c.filter == MGPU_FILTER_ARTVISC ? 0.0 : -bd.p.dissip_shock;
  fa.xnu10 = bd.p.dissip_coeff / 1260.0;
  fa.csens = bd.p.csens;
  for (int q = 0; q < 3; q++) fa.sr[q] = bd.sr[q];
  // Cartesian metrics of the Edoh factor and the artificial viscosity (on a
  // 2.5-D block the i, j spectral radii carry their metric already)
  const bool c2 = h->c.curvilinear == MGPU_GEOM_C2;
  fa.idx = c2 ? nullptr : bd.idx[0];
  fa.idy = c2 ? nullptr : bd.idx[1];
  fa.idz = bd.idx[2];
  fa.dt = bd.p.deltat;
  fa.edoh = edoh_on(h->c);
  fa.avs = h->c.filter == MGPU_FILTER_ARTVISC_SHOCK || (h->c.filter == MGPU_FILTER_ARTVISC && h->wall);
  return fa;
}

// Fills the batched launch tables (after create, and after the autotuner
// changed the tiles).  The stream is idle when they are rewritten.
int build_tables(mgpu_handle *h) {
  if (h->tables_ready) return 0;
  const int nb = (int)h->b.size(), nrk = h->c.nrk;
  const mgpu_case &c = h->c;
  int rc;
  HIPCK(h, hipStreamSynchronize(h->stream));
  drop_graphs(h);
  if (nb > h->tab_cap) {
    const int cap = nb < 8 ? 8 : nb;
    if ((rc = dalloc_table(h, &h->d_stage, (size_t)3 * nrk * cap)) || (rc = dalloc_table(h, &h->d_visc, 3 * cap)) ||
        (rc = dalloc_table(h, &h->d_vdiv, cap)) || (rc = dalloc_table(h, &h->d_filt, 3 * cap)) ||
        (rc = dalloc_table(h, &h->d_tune, cap)) || (rc = dalloc_table(h, &h->d_tune_f, cap)))
      return rc;
    h->tab_cap = cap;
  }
  if (nb == 0) {
    h->nt_stage[0] = h->nt_stage[1] = h->nt_visc = h->nt_vdiv = h->nt_filt = 0;
    h->tables_ready = true;
    return 0;
  }
  std::vector<TileRange> tr[2];
  h->nt_stage[0] = stage_ranges(h, 0, tr[0]);
  h->nt_stage[1] = stage_ranges(h, 1, tr[1]);
  std::vector<Batch<StageArgs>> st((size_t)3 * nrk * nb);
  std::vector<Batch<ViscArgs>> vi(3 * nb);
  std::vector<Batch<VdivArgs>> vd(nb);
  std::vector<Batch<FilterArgs>> fi(3 * nb);
  const bool visc = c.iorder_visc != 0;
  int fv = 0, fd = 0, ff = 0;
  for (int b = 0; b < nb; b++) {
    const BlockData &bd = h->b[b];
    TileRange tv = visc_tiles(bd.g, h->kc_aux[0]), td = vdiv_tiles(bd.g, h->kc_aux[1]),
              tf = filter_tiles(bd.g, h->kc_aux[2]);
    tv.jfirst = h->jf_aux[0];
    td.jfirst = h->jf_aux[1];
    tf.jfirst = h->jf_aux[2];
    tv.first = fv;
    td.first = fd;
    tf.first = ff;
    fv += tile_count(tv, bd.g.nz);
    fd += tile_count(td, bd.g.nz);
    ff += tile_count(tf, bd.g.nz);
    for (int s0 = 0; s0 < 3; s0++) {
      const int sa = (s0 + 1) % 3, sb = (s0 + 2) % 3;
      int in = s0;
      for (int irk = 1; irk <= nrk; irk++) {
        const int out = (irk % 2 == 1) ? sa : sb;
        const bool vl = visc && irk == nrk;
        Batch<StageArgs> &e = st[((size_t)s0 * nrk + (irk - 1)) * nb + b];
        e.a = stage_args(h, bd, in, s0, out, irk);
        e.t = tr[vl ? 1 : 0][b];
        in = out;
      }
      ViscArgs &va = vi[s0 * nb + b].a;
      va.g = bd.g;
      va.gas = gas_of(bd.p);
      for (int v = 0; v < 5; v++) va.u0.v[v] = bd.st[s0][v];
      va.met = metrics_of(bd);
      for (int q = 0; q < NTAU; q++) va.tau[q] = bd.tau[q];
      va.prs = bd.prs;
      va.sens = bd.sens;
      for (int q = 0; q < 3; q++) va.sr[q] = bd.sr[q];
      va.want_sensor = has_sensor(c);
      va.want_sr = has_sr(c);
      va.ducros = c.is_ducros;
      vi[s0 * nb + b].t = tv;
      fi[s0 * nb + b].a = filter_args(h, bd, in, s0);  // reads the last stage's output
      fi[s0 * nb + b].t = tf;
    }
    VdivArgs &va = vd[b].a;
    va.g = bd.g;
    va.met = metrics_of(bd);
    for (int q = 0; q < NTAU; q++) va.tau[q] = bd.tau[q];
    for (int q = 0; q < 4; q++) va.kv[q] = bd.kv[q];
    vd[b].t = td;
  }
  h->nt_visc = fv;
  h->nt_vdiv = fd;
  h->nt_filt = ff;
  HIPCK(h, hipMemcpy(h->d_stage, st.data(), st.size() * sizeof(st[0]), hipMemcpyHostToDevice));
  HIPCK(h, hipMemcpy(h->d_visc, vi.data(), vi.size() * sizeof(vi[0]), hipMemcpyHostToDevice));
  HIPCK(h, hipMemcpy(h->d_vdiv, vd.data(), vd.size() * sizeof(vd[0]), hipMemcpyHostToDevice));
  HIPCK(h, hipMemcpy(h->d_filt, fi.data(), fi.size() * sizeof(fi[0]), hipMemcpyHostToDevice));
  h->tables_ready = true;
  return 0;
}

// Wall state of the j-walls of every block with walls: from rows 2,3 of set
// `src` (-1: from the block's stash) into the wall rows of set `dst` (-1:
// into the stash).  A set >= 0 is an FP64 state set, a set <= -2 the
// mixed-precision FP32 stage set -2 - set.
// The stage kernel's wall variant: 0 no bounded face, 1 bounded faces, 2
// also advanced wall rows.
int stage_wall(const mgpu_handle *h) { return h->bnd ? (h->adv ? 2 : 1) : 0; }

int enqueue_walls(mgpu_handle *h, int src, int dst) {
  for (const auto &bd : h->b) {
    if (!bd.g.jwall && !bd.g.iwall) continue;
    WallArgs wa{};
    wa.g = bd.g;
    wa.gas = gas_of(bd.p);
    for (int v = 0; v < 5; v++) {
      wa.src.v[v] = src >= 0 ? bd.st[src][v] : nullptr;
      wa.dst.v[v] = dst >= 0 ? bd.st[dst][v] : nullptr;
      wa.src32[v] = src <= -2 ? bd.sf[-2 - src][v] : nullptr;
      wa.dst32[v] = dst <= -2 ? bd.sf[-2 - dst][v] : nullptr;
    }
    wa.ref[0] = bd.ref.r;
    wa.ref[1] = bd.ref.ru;
    wa.ref[2] = bd.ref.rv;
    wa.ref[3] = bd.ref.rw;
    wa.ref[4] = bd.ref.re;
    wa.stash = bd.wstash;
    for (int s = 0; s < 2; s++) {
      wa.c2[s] = bd.c2[s];
      wa.c3[s] = bd.c3[s];
      wa.ci2[s] = bd.ci2[s];
      wa.ci3[s] = bd.ci3[s];
    }
    wa.cv = 1.0 / (bd.p.gam - 1.0) * bd.p.rg;
    wa.t_wall = bd.p.t_wall;
    wa.adiabatic = bd.p.is_adiab;
    wa.wc3[0] = bd.wc3[0];
    wa.wc3[1] = bd.wc3[1];
    wa.jslip = bd.jslip;
    wa.islip = bd.islip;
    wa.wt3[0] = bd.wt3[0];
    wa.wt3[1] = bd.wt3[1];
    HIPCK(h, launch_wall(wa, h->stream));
  }
  return 0;
}

WallViscArgs wall_visc_args(const mgpu_handle *h, const BlockData &bd, int s0) {
  WallViscArgs a{};
  a.g = bd.g;
  a.gas = gas_of(bd.p);
  for (int v = 0; v < 5; v++) a.u0.v[v] = bd.st[s0][v];
  a.met = metrics_of(bd);
  for (int q = 0; q < NTAU; q++) a.tau[q] = bd.tau[q];
  for (int q = 0; q < 4; q++) a.kv[q] = bd.kv[q];
  a.adiabatic = bd.p.is_adiab;
  a.prs = bd.prs;
  a.sens = bd.sens;
  for (int q = 0; q < 3; q++) a.sr[q] = bd.sr[q];
  a.want_sensor = bd.prs != nullptr;
  a.ducros = h->c.is_ducros;
  a.geo = h->c.curvilinear;
  a.wc3[0] = bd.wc3[0];
  a.wc3[1] = bd.wc3[1];
  return a;
}

// Characteristic rows and inlet of every block that has them, from set `src`
// (>= 0 an FP64 set, <= -2 the mixed-precision FP32 stage set -2 - src).
// cst_mode / inlet_mode as in CharArgs; cnew: the FP64 set of U^n (cst_mode 1).
int enqueue_charac(mgpu_handle *h, int src, int cst_mode, int cnew, int inlet_mode) {
  for (auto &bd : h->b) {
    if (!bd.kimax && !bd.kjmax && !bd.uin) continue;
    CharArgs a{};
    a.g = bd.g;
    a.gas = gas_of(bd.p);
    a.cp = bd.p.gam * bd.p.rg / (bd.p.gam - 1.0);
    a.p_ref = bd.p.p_ref;
    for (int v = 0; v < 5; v++) {
      a.src.v[v] = src >= 0 ? bd.st[src][v] : nullptr;
      a.src32[v] = src <= -2 ? bd.sf[-2 - src][v] : nullptr;
      a.cnew.v[v] = cnew >= 0 ? bd.st[cnew][v] : nullptr;
      a.row1.v[v] = src >= 0 ? bd.st[src][v] : nullptr;
    }
    a.ref[0] = bd.ref.r;
    a.ref[1] = bd.ref.ru;
    a.ref[2] = bd.ref.rv;
    a.ref[3] = bd.ref.rw;
    a.ref[4] = bd.ref.re;
    a.idx_v = bd.idx[3];
    a.idy_v = bd.idx[4];
    a.kimax = bd.kimax;
    a.kjmax = bd.kjmax;
    a.cst = bd.cst;
    a.cst_mode = cst_mode;
    a.uin = bd.uin;
    a.ppre = bd.ppre;
    a.kf = bd.kf;
    a.inlet_mode = inlet_mode;
    // bc_charac_jmax runs over i = ndx_c..nfx_c: one row in from a
    // characteristic or inlet i-face (mpi_index_bounds.f90:298-360)
    a.i0 = bd.uin ? 2 : 1;
    a.i1 = bd.kimax ? bd.g.nx - 1 : bd.g.nx;
    if (inlet_mode == 1 && src < 0) return fail(h, MGPU_EINVAL, "inlet reset needs an FP64 set");
    if (cst_mode == 3 && !bd.kimax) continue;
    HIPCK(h, launch_charac(a, h->stream));
  }
  return 0;
}

// start_runge_kutta (runge_kutta.f90:316-342) for blocks whose state was just
// uploaded: the sound speed of U (primitives_visc at set-up), the increments
// of the characteristic rows from U, and the inflow state = row 1 of U, with
// no filter increment yet on the inlet row.
int start_charac(mgpu_handle *h) {
  bool any = false;
  for (auto &bd : h->b) any |= bd.bc_start;
  if (!any) return 0;
  for (auto &bd : h->b) {
    if (!bd.bc_start) continue;
    CharArgs a{};
    a.g = bd.g;
    a.gas = gas_of(bd.p);
    a.cp = bd.p.gam * bd.p.rg / (bd.p.gam - 1.0);
    a.p_ref = bd.p.p_ref;
    for (int v = 0; v < 5; v++) a.src.v[v] = bd.st[h->cur][v];
    a.idx_v = bd.idx[3];
    a.idy_v = bd.idx[4];
    a.kimax = bd.kimax;
    a.kjmax = bd.kjmax;
    a.cst = bd.cst;
    a.cst_mode = 2;
    a.uin = bd.uin;
    a.ppre = bd.ppre;
    a.kf = bd.kf;
    a.inlet_mode = 2;
    a.i0 = bd.uin ? 2 : 1;
    a.i1 = bd.kimax ? bd.g.nx - 1 : bd.g.nx;
    HIPCK(h, launch_charac(a, h->stream));
    bd.bc_start = false;
  }
  return 0;
}

// Enqueues one step with U^n in set s0 (no host synchronisation: the same
// sequence is captured into a graph).  runge_kutta, runge_kutta.f90:2-152.
// With walls, bc_wall runs after each stage's exchange, from the stage's
// input (the primitives the reference computed last), except after the last
// stage: there the wall state (from U^n, whose primitives the last stage's
// primitives_visc wrote, viscous or not, runge_kutta.f90:71) is kept aside
// and imposed after the filter and its exchange (runge_kutta.f90:20-24,
// 42-45, 130-136).
// With characteristic faces, bc_non_reflecting runs after `primitives` in
// stages 2..nrk (runge_kutta.f90:30-37, 64-70: after that stage's wall
// state, before its fluxes) and at the end of the step (:137-139, after the
// final wall state), where the inlet row also takes the inflow state; the
// last stage then records the sound speed of U^n that bc_charac_imax reads
// until the next step's last stage (primitives_visc, :71).  The filter
// increment of the inlet row is kept: K there is not written again before
// the next step's stages use it.
int enqueue_step(mgpu_handle *h, int s0) {
  const mgpu_case &c = h->c;
  const int geo = c.curvilinear;
  const bool visc = c.iorder_visc != 0;
  const int nb = (int)h->b.size(), nrk = c.nrk;
  const int sa = (s0 + 1) % 3, sb = (s0 + 2) % 3;
  int rc, in = s0;
  for (int irk = 1; irk <= nrk; irk++) {
    const bool last = irk == nrk;
    const int out = (irk % 2 == 1) ? sa : sb;
    if (last && visc) {
      ProfScope ps(h, MGPU_K_VISC);
      HIPCK(h, launch_visc(h->d_visc + s0 * nb, nb, h->nt_visc, geo, h->stream));
      if (h->bnd)
        for (const auto &bd : h->b)
          if (bd.g.ibnd || bd.g.jbnd) HIPCK(h, launch_visc_wall(wall_visc_args(h, bd, s0), h->stream));
      HIPCK(h, launch_iface(h->xv, h->nchunks_v, h->stream));
      if (h->wall)
        for (const auto &bd : h->b)
          if (bd.g.jwall || bd.g.iwall) HIPCK(h, launch_tau_wallrow(wall_visc_args(h, bd, s0), h->stream));
      HIPCK(h, launch_vdiv(h->d_vdiv, nb, h->nt_vdiv, geo, h->stream));
      if (h->bnd)
        for (const auto &bd : h->b)
          if (bd.g.ibnd || bd.g.jbnd) HIPCK(h, launch_vdiv_wall(wall_visc_args(h, bd, s0), h->stream));
    }
    if (h->charac && irk > 1) {
      ProfScope ps(h, MGPU_K_HALO);
      // the stage input (mixed precision: an FP32 stage set)
      const int in_d = h->mixed ? -2 - fset_of(irk - 1) : in;
      if ((rc = enqueue_charac(h, in_d, last ? 1 : 0, s0, 0))) return rc;
    } else if (h->charac && last) {
      // one stage: no characteristic rows before it, only the sound speed of U^n
      ProfScope ps(h, MGPU_K_HALO);
      if ((rc = enqueue_charac(h, -1, 3, s0, 0))) return rc;
    }
    {
      ProfScope ps(h, MGPU_K_STAGE);
      const int v = (last && visc) ? 1 : 0;
      HIPCK(h, launch_stage(h->d_stage + ((size_t)s0 * nrk + (irk - 1)) * nb, nb, h->nt_stage[v], geo,
                            last && visc, h->shape[v], stage_mode(h, irk), stage_wall(h), h->stream));
    }
    if (h->mixed && !last) {
      ProfScope ps(h, MGPU_K_HALO);
      HIPCK(h, launch_iface(h->xf[fset_of(irk)], h->nchunks_s, h->stream));
    } else if ((rc = exchange_state(h, out))) {
      return rc;
    }
    if (h->wall) {
      ProfScope ps(h, MGPU_K_HALO);
      // the stage's input and output sets (mixed precision: the intermediate
      // stage states are FP32 sets)
      const int in_d = (h->mixed && irk > 1) ? -2 - fset_of(irk - 1) : in;
      const int out_d = (h->mixed && !last) ? -2 - fset_of(irk) : out;
      // after the last stage from U^n: its primitives_visc runs viscous or not
      if (!last) rc = enqueue_walls(h, in_d, out_d);
      else if (c.filter != MGPU_FILTER_NONE) rc = enqueue_walls(h, s0, -1);
      else rc = enqueue_walls(h, s0, out);
      if (rc) return rc;
    }
    in = out;
  }
  if (c.filter != MGPU_FILTER_NONE) {
    {
      ProfScope ps(h, MGPU_K_FILTER);
      if (c.filter == MGPU_FILTER_SF)
        HIPCK(h, launch_filter_sf(h->d_filt + s0 * nb, nb, h->nt_filt, h->bnd, h->stream));
      else
        HIPCK(h, launch_filter_shock(h->d_filt + s0 * nb, nb, h->nt_filt, dissip_kind(c, h->wall), h->bnd,
                                     edoh_on(c), h->stream));
      if (h->bnd && has_sensor(c))
        for (const auto &bd : h->b) {
          if (bd.g.jbnd) HIPCK(h, launch_shock_wall(filter_args(h, bd, in, s0), 1, h->stream));
          if (bd.g.ibnd) HIPCK(h, launch_shock_wall(filter_args(h, bd, in, s0), 0, h->stream));
        }
      // the filter increment of the inlet row, before the wall state changes
      // its corner
      for (const auto &bd : h->b)
        if (bd.kf) {
          CState after{}, before{};
          for (int v = 0; v < 5; v++) {
            after.v[v] = bd.st[s0][v];
            before.v[v] = bd.st[in][v];
          }
          HIPCK(h, launch_inlet_kf(bd.g, after, before, bd.kf, h->stream));
        }
    }
    if ((rc = exchange_state(h, s0))) return rc;
    if (h->wall) {
      ProfScope ps(h, MGPU_K_HALO);
      if ((rc = enqueue_walls(h, -1, s0))) return rc;
    }
  }
  if (h->charac) {
    ProfScope ps(h, MGPU_K_HALO);
    if ((rc = enqueue_charac(h, c.filter != MGPU_FILTER_NONE ? s0 : in, 0, -1, 1))) return rc;
  }
  return 0;
}

// Set holding U^{n+1} after a step from set s0.
int next_set(const mgpu_case &c, int s0) {
  if (c.filter != MGPU_FILTER_NONE) return s0;
  return (c.nrk % 2 == 1) ? (s0 + 1) % 3 : (s0 + 2) % 3;
}

// One step through the captured graph of set h->cur (captured on first use).
int graph_step(mgpu_handle *h) {
  const int s0 = h->cur;
  if (!h->gexec[s0]) {
    hipGraph_t g = nullptr;
    HIPCK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_step(h, s0);
    const hipError_t e = hipStreamEndCapture(h->stream, &g);
    if (rc) {
      if (e == hipSuccess && g) (void)hipGraphDestroy(g);
      return rc;
    }
    HIPCK(h, e);
    const hipError_t ei = hipGraphInstantiate(&h->gexec[s0], g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) h->gexec[s0] = nullptr;
    HIPCK(h, ei);
  }
  HIPCK(h, hipGraphLaunch(h->gexec[s0], h->stream));
  return 0;
}

// Exchange the ghosts of U^n after put_state, once every block that has
// state is in place.
int refresh_ghosts(mgpu_handle *h) {
  int rc = setup_exchange(h);
  if (rc || !h->ghosts_dirty) return rc;
  if ((rc = exchange_state(h, h->cur))) return rc;
  h->ghosts_dirty = false;
  return 0;
}

std::string check_flags(const mgpu_case &c) {
  char buf[256];
  buf[0] = 0;
  if (c.nrk < 1 || c.nrk > 8) snprintf(buf, sizeof buf, "nrk=%d out of range 1..8", c.nrk);
  else if (c.iorder_visc != 0 && c.iorder_visc != 4)
    snprintf(buf, sizeof buf, "iorder_visc=%d: only 0 and 4 are supported", c.iorder_visc);
  else if (c.filter < 0 || c.filter > 4) snprintf(buf, sizeof buf, "filter=%d", c.filter);
  else if (c.curvilinear < MGPU_GEOM_CART || c.curvilinear > MGPU_GEOM_C2)
    snprintf(buf, sizeof buf, "curvilinear=%d: 0 Cartesian, 1 3-D, 2 2.5-D", c.curvilinear);
  else if (c.filter == MGPU_FILTER_SHOCK && c.iorder_visc == 0)
    snprintf(buf, sizeof buf, "shock filter needs iorder_visc=4 (its sensor reads the viscous gradients of U^n)");
  else if (is_artvisc(c) && c.iorder_visc == 0)
    snprintf(buf, sizeof buf, "artificial viscosity needs iorder_visc=4 (it reads the sound speed of U^n)");
  else if (c.filter == MGPU_FILTER_ARTVISC && c.curvilinear == MGPU_GEOM_C3)
    snprintf(buf, sizeof buf, "artificial viscosity: Cartesian or 2.5-D cases (artvisc_o9 has no c3 form)");
  else if (c.precision != MGPU_PREC_FP64 && c.precision != MGPU_PREC_MIXED)
    snprintf(buf, sizeof buf, "precision=%d", c.precision);
  else if (c.is_wall2 != 0 && c.is_wall2 != 1)
    snprintf(buf, sizeof buf, "is_wall2=%d: 0 or 1", c.is_wall2);
  else if (c.precision == MGPU_PREC_MIXED && c.nrk < 2)
    snprintf(buf, sizeof buf, "mixed precision needs nrk >= 2");
  else if ((c.launch_flags & MGPU_LAUNCH_STAGE_CLASSIC) && (c.launch_flags & MGPU_LAUNCH_STAGE_PIPELINED))
    snprintf(buf, sizeof buf, "launch_flags: MGPU_LAUNCH_STAGE_CLASSIC and MGPU_LAUNCH_STAGE_PIPELINED together");
  return std::string(buf);
}

int handle_new(const mgpu_case &c, mgpu_handle **out) {
  *out = nullptr;
  mgpu_handle *h = new mgpu_handle();
  h->c = c;
  h->c.nblocks = 0;
  h->c.blocks = nullptr;
  h->use_graph = (c.launch_flags & MGPU_LAUNCH_EAGER) == 0;
  h->stage_family = (c.launch_flags & MGPU_LAUNCH_STAGE_PIPELINED) ? 2
                    : ((c.launch_flags & MGPU_LAUNCH_STAGE_CLASSIC) ? 1 : 0);
  if (h->stage_family == 2) h->shape[0] = h->shape[1] = kFirstPipelinedShape;
  h->mixed = c.precision == MGPU_PREC_MIXED;
  if (hipGetDevice(&h->device) != hipSuccess) {
    delete h;
    return fail(nullptr, MGPU_EDEVICE, "no HIP device");
  }
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return fail(nullptr, MGPU_EDEVICE, "hipStreamCreate failed");
  }
  // device address of the allocation-free diagnostics' static array, looked
  // up once per device (this device is current)
  if (diag_scratch_init(h->device) != hipSuccess) {
    (void)hipStreamDestroy(h->stream);
    delete h;
    return fail(nullptr, MGPU_EDEVICE, "device address of the diagnostics scratch array not found");
  }
  *out = h;
  return 0;
}

size_t block_bytes(const mgpu_case &c, int nx, int ny, int nz) {
  const size_t n = (size_t)(nx + 2 * G) * (ny + 2 * G) * (nz + 2 * G);
  size_t arrays = 15;  // three state sets
  if (c.curvilinear == MGPU_GEOM_C3) arrays += c.iorder_visc ? 20 : 10;
  if (c.iorder_visc)
    arrays += NTAU + 4 + (has_sensor(c) ? 2 : 0) + (has_sr(c) ? 3 : 0);
  size_t bytes = arrays * n * sizeof(double);
  if (c.precision == MGPU_PREC_MIXED) bytes += 10 * n * sizeof(float);
  if (c.precision == MGPU_PREC_MIXED && c.curvilinear == MGPU_GEOM_C3) bytes += 9 * n * sizeof(float);
  if (c.curvilinear != MGPU_GEOM_C3) bytes += 6 * (size_t)(nx + ny + nz + 6 * G) * sizeof(double);
  if (c.curvilinear == MGPU_GEOM_C2) bytes += 10 * (size_t)(nx + 2 * G) * (ny + 2 * G) * sizeof(double);
  return bytes;  // (+ 4*nx*nz doubles of wall state for a block with walls)
}

int handle_add_block(mgpu_handle *h, const mgpu_block &bl, const mgpu_case &p) {
  const mgpu_case &c = h->c;
  if (bl.nx < G || bl.ny < G || bl.nz < G)
    return fail(h, MGPU_EINVAL, "block extents %dx%dx%d: every extent must be >= %d", bl.nx, bl.ny, bl.nz, G);
  const size_t padded = (size_t)(bl.nx + 2 * G) * (bl.ny + 2 * G) * (bl.nz + 2 * G);
  if (padded > kMaxPadded)
    return fail(h, MGPU_EINVAL, "%zu padded points exceed the 2^28 limit of one block", padded);
  HIPCK(h, hipStreamSynchronize(h->stream));
  h->b.emplace_back();
  h->blocks.push_back(bl);
  BlockData &bd = h->b.back();
  bd.p = p;
  bd.g.nx = bl.nx;
  bd.g.ny = bl.ny;
  bd.g.nz = bl.nz;
  bd.g.sx = bl.nx + 2 * G;
  bd.g.sy = (idx_t)bd.g.sx * (idx_t)(bl.ny + 2 * G);
  bd.g.n = (long)bd.g.sy * (bl.nz + 2 * G);
  bd.g.jwall = (is_wall(bl.face_kind[2]) ? 1 : 0) | (is_wall(bl.face_kind[3]) ? 2 : 0);
  bd.g.iwall = (is_wall(bl.face_kind[0]) ? 1 : 0) | (is_wall(bl.face_kind[1]) ? 2 : 0);
  bd.g.jadv = p.is_wall2 ? bd.g.jwall : 0;
  bd.g.iadv = p.is_wall2 ? bd.g.iwall : 0;
  bd.jslip = (is_slip(p, bl.face_kind[2]) ? 1 : 0) | (is_slip(p, bl.face_kind[3]) ? 2 : 0);
  bd.islip = (is_slip(p, bl.face_kind[0]) ? 1 : 0) | (is_slip(p, bl.face_kind[1]) ? 2 : 0);
  bd.g.ibnd = (is_bound(bl.face_kind[0]) ? 1 : 0) | (is_bound(bl.face_kind[1]) ? 2 : 0);
  bd.g.jbnd = (is_bound(bl.face_kind[2]) ? 1 : 0) | (is_bound(bl.face_kind[3]) ? 2 : 0);
  bd.bytes = block_bytes(c, bl.nx, bl.ny, bl.nz);
  const size_t n = (size_t)bd.g.n;
  int rc = 0;
  for (int s = 0; s < 3 && !rc; s++)
    for (int v = 0; v < 5 && !rc; v++) rc = dalloc(h, &bd.st[s][v], n);
  if (h->mixed)
    for (int s = 0; s < 2 && !rc; s++)
      for (int v = 0; v < 5 && !rc; v++) rc = dalloc32(h, &bd.sf[s][v], n);
  if (c.curvilinear == MGPU_GEOM_C3) {
    for (int q = 0; q < 10 && !rc; q++) rc = dalloc(h, &bd.m[q], n);
    if (h->mixed)
      for (int q = 0; q < 9 && !rc; q++) rc = dalloc32(h, &bd.m32[q], n);
    if (c.iorder_visc)
      for (int q = 0; q < 10 && !rc; q++) rc = dalloc(h, &bd.mv[q], n);
  } else {
    // Cartesian metrics; a 2.5-D block uses idz, idz_v of them and 2-D arrays
    const int len[3] = {bl.nx + 2 * G, bl.ny + 2 * G, bl.nz + 2 * G};
    for (int q = 0; q < 6 && !rc; q++) rc = dalloc(h, &bd.idx[q], (size_t)len[q % 3]);
    if (c.curvilinear == MGPU_GEOM_C2) {
      const size_t n2 = (size_t)bd.g.sx * (bl.ny + 2 * G);
      for (int q = 0; q < 5 && !rc; q++) rc = dalloc(h, &bd.m2[q], n2);
      for (int q = 0; q < 5 && !rc; q++) rc = dalloc(h, &bd.m2v[q], n2);
    }
  }
  if ((bd.g.jwall || bd.g.iwall) && !rc)
    rc = dalloc(h, &bd.wstash, (size_t)4 * bl.nx * bl.nz + (size_t)4 * bl.ny * bl.nz);
  if (c.curvilinear == MGPU_GEOM_C3)
    for (int s = 0; s < 2 && !rc; s++)
      if ((bd.g.jwall >> s) & 1) {
        rc = dalloc(h, &bd.wc3[s], (size_t)8 * bl.nx * bl.nz);
        if (!rc && ((bd.jslip >> s) & 1)) rc = dalloc(h, &bd.wt3[s], (size_t)6 * bl.nx * bl.nz);
      }
  {
    const size_t fi = (size_t)bl.ny * bl.nz, fj = (size_t)bl.nx * bl.nz;
    if (bl.face_kind[1] == MGPU_FACE_CHARAC) {
      if (!rc) rc = dalloc(h, &bd.kimax, 5 * fi);
      if (!rc) rc = dalloc(h, &bd.cst, fi);
    }
    if (bl.face_kind[3] == MGPU_FACE_CHARAC && !rc) rc = dalloc(h, &bd.kjmax, 5 * fj);
    if (bl.face_kind[0] == MGPU_FACE_INLET) {
      if (!rc) rc = dalloc(h, &bd.uin, 5 * fi);
      if (!rc) rc = dalloc(h, &bd.ppre, 5 * fi);
      if (!rc) rc = dalloc(h, &bd.kf, 5 * fi);
    }
  }
  if (c.iorder_visc) {
    for (int q = 0; q < NTAU && !rc; q++) rc = dalloc(h, &bd.tau[q], n);
    for (int q = 0; q < 4 && !rc; q++) rc = dalloc(h, &bd.kv[q], n);
    if (has_sensor(c)) {
      if (!rc) rc = dalloc(h, &bd.prs, n);
      if (!rc) rc = dalloc(h, &bd.sens, n);
    }
    if (has_sr(c))
      for (int q = 0; q < 3 && !rc; q++) rc = dalloc(h, &bd.sr[q], n);
  }
  if (rc) {
    free_block(bd);
    h->b.pop_back();
    h->blocks.pop_back();
    return rc;
  }
  if (bd.g.jwall || bd.g.iwall) h->wall = true;
  if (bd.g.ibnd || bd.g.jbnd) h->bnd = true;
  if (bd.g.ibnd) h->stage_family = 1;  // the classic stage kernel closes i-faces
  if (bd.kimax || bd.kjmax || bd.uin) {
    h->charac = true;
    h->stage_family = 1;  // the classic stage kernel closes i-faces
  }
  if (bd.wc3[0] || bd.wc3[1]) h->stage_family = 1;  // and c3 walls
  if (bd.g.jadv || bd.g.iadv) {
    h->adv = true;
    h->stage_family = 1;  // and advanced wall rows
  }
  h->exchange_ready = false;
  h->tables_ready = false;
  return 0;
}

int handle_remove_block(mgpu_handle *h, int i) {
  if (i < 0 || i >= (int)h->b.size()) return fail(h, MGPU_EINVAL, "block %d out of range", i);
  HIPCK(h, hipStreamSynchronize(h->stream));
  drop_graphs(h);
  free_block(h->b[i]);
  h->b.erase(h->b.begin() + i);
  h->blocks.erase(h->blocks.begin() + i);
  h->exchange_ready = false;
  h->tables_ready = false;
  return 0;
}

int put_metrics_cart(mgpu_handle *h, int block, const double *const *src_in) {
  if (block < 0 || block >= (int)h->b.size()) return fail(h, MGPU_EINVAL, "block %d out of range", block);
  if (h->c.curvilinear) return fail(h, MGPU_ESTATE, "case is curvilinear: use mgpu_set_metrics_c3 or _c2");
  BlockData &bd = h->b[block];
  const int len[3] = {bd.g.nx + 2 * G, bd.g.ny + 2 * G, bd.g.nz + 2 * G};
  const double *sp[6];  // in the stored axis order
  for (int q = 0; q < 6; q++) sp[3 * (q / 3) + kAxisMap[bd.perm][q % 3]] = src_in[q];
  const double *const *src = sp;
  for (int q = 0; q < 6; q++) {
    if (!src[q]) {
      // the characteristic rows take their one-sided derivatives with the
      // viscous metrics idx_v, idy_v (mod_charac.f90:386-404 and the jmax twin)
      const bool charac = bd.kimax || bd.kjmax;
      if (q >= 3 && h->c.iorder_visc == 0 && !(charac && q < 5)) continue;
      return fail(h, MGPU_EINVAL, "metric array %d is null", q);
    }
    HIPCK(h, hipMemcpy(bd.idx[q], src[q], sizeof(double) * len[q % 3], hipMemcpyHostToDevice));
  }
  bd.have_metrics = true;
  return 0;
}

int put_metrics_c3(mgpu_handle *h, int block, const double *const *m, const double *const *mv) {
  if (block < 0 || block >= (int)h->b.size()) return fail(h, MGPU_EINVAL, "block %d out of range", block);
  if (h->c.curvilinear != MGPU_GEOM_C3)
    return fail(h, MGPU_ESTATE, "case is not 3-D curvilinear: use mgpu_set_metrics_cart or _c2");
  if (!m) return fail(h, MGPU_EINVAL, "null metric table");
  BlockData &bd = h->b[block];
  const size_t bytes = sizeof(double) * (size_t)bd.g.n;
  // a relabeled block (walls on its i- or k-faces): the rows of its stored
  // index axes, each array in the stored axis order; the physical components
  // x, y, z and 1/J keep their places
  std::vector<double> tmp(bd.perm ? (size_t)bd.g.n : 0);
  auto stored = [&](const double *const *src, int q) -> const double * {
    const double *a = q < 9 ? src[3 * caller_axis(bd.perm, q / 3) + q % 3] : src[q];
    if (!bd.perm) return a;
    relabel_array(bd.g, bd.perm, a, tmp.data(), true);
    return tmp.data();
  };
  for (int q = 0; q < 10; q++)
    if (!m[q]) return fail(h, MGPU_EINVAL, "metric %d is null", q);
  std::vector<float> f(bd.m32[0] ? (size_t)bd.g.n : 0);
  for (int q = 0; q < 10; q++) {
    const double *a = stored(m, q);
    HIPCK(h, hipMemcpy(bd.m[q], a, bytes, hipMemcpyHostToDevice));
    if (bd.m32[0] && q < 9) {
      // the mixed path's FP32 copies of the flux metric terms
      for (size_t e = 0; e < f.size(); e++) f[e] = (float)a[e];
      HIPCK(h, hipMemcpy(bd.m32[q], f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice));
    }
  }
  if (h->c.iorder_visc) {
    if (!mv) return fail(h, MGPU_EINVAL, "viscous metric table is null");
    for (int q = 0; q < 10; q++)
      if (!mv[q]) return fail(h, MGPU_EINVAL, "viscous metric %d is null", q);
    for (int q = 0; q < 10; q++) HIPCK(h, hipMemcpy(bd.mv[q], stored(mv, q), bytes, hipMemcpyHostToDevice));
  }
  bd.have_metrics = true;
  return 0;
}

int put_metrics_c2(mgpu_handle *h, int block, const double *const *m2, const double *const *m2v, const double *idz,
                   const double *idz_v) {
  if (block < 0 || block >= (int)h->b.size()) return fail(h, MGPU_EINVAL, "block %d out of range", block);
  if (h->c.curvilinear != MGPU_GEOM_C2)
    return fail(h, MGPU_ESTATE, "case is not 2.5-D curvilinear: use mgpu_set_metrics_cart or _c3");
  if (!m2 || !idz) return fail(h, MGPU_EINVAL, "null metric table");
  BlockData &bd = h->b[block];
  const size_t n2 = (size_t)bd.g.sx * (bd.g.ny + 2 * G);
  // stored as the flux metric rows: ksi (y_eta, -x_eta), eta (-y_ksi, x_ksi), ijacob
  auto up = [&](const double *const *src, double *const *dst) -> int {
    for (int q = 0; q < 5; q++)
      if (!src[q]) return fail(h, MGPU_EINVAL, "2-D metric %d is null", q);
    std::vector<double> row(n2);
    const int from[5] = {3, 1, 2, 0, 4};
    const double sign[5] = {1.0, -1.0, -1.0, 1.0, 1.0};
    for (int q = 0; q < 5; q++) {
      for (size_t e = 0; e < n2; e++) row[e] = sign[q] * src[from[q]][e];
      HIPCK(h, hipMemcpy(dst[q], row.data(), n2 * sizeof(double), hipMemcpyHostToDevice));
    }
    return 0;
  };
  int rc = up(m2, bd.m2);
  if (rc) return rc;
  const size_t nz = (size_t)bd.g.nz + 2 * G;
  HIPCK(h, hipMemcpy(bd.idx[2], idz, nz * sizeof(double), hipMemcpyHostToDevice));
  if (h->c.iorder_visc) {
    if (!m2v || !idz_v) return fail(h, MGPU_EINVAL, "viscous metric table is null");
    if ((rc = up(m2v, bd.m2v))) return rc;
    HIPCK(h, hipMemcpy(bd.idx[5], idz_v, nz * sizeof(double), hipMemcpyHostToDevice));
  }
  bd.have_metrics = true;
  return 0;
}

int put_state(mgpu_handle *h, int block, const double *const *src_in) {
  if (block < 0 || block >= (int)h->b.size()) return fail(h, MGPU_EINVAL, "block %d out of range", block);
  BlockData &bd = h->b[block];
  const size_t bytes = sizeof(double) * (size_t)bd.g.n;
  for (int v = 0; v < 5; v++)
    if (!src_in[v]) return fail(h, MGPU_EINVAL, "state array %d is null", v);
  // a relabeled block: the caller's arrays in the stored axis order
  std::vector<double> tmp[5];
  const double *sp[5];
  for (int v = 0; v < 5; v++) {
    const double *u = src_in[caller_var(bd.vperm, v)];
    if (bd.perm) {
      tmp[v].resize((size_t)bd.g.n);
      relabel_array(bd.g, bd.perm, u, tmp[v].data(), true);
      u = tmp[v].data();
    }
    sp[v] = u;
  }
  const double *const *src = sp;
  HIPCK(h, hipStreamSynchronize(h->stream));
  for (int v = 0; v < 5; v++) HIPCK(h, hipMemcpy(bd.st[h->cur][v], src[v], bytes, hipMemcpyHostToDevice));
  bd.have_state = true;
  h->ghosts_dirty = true;  // the exchange runs before the next use of the ghosts
  bd.bc_start = bd.kimax || bd.kjmax || bd.uin;
  if (h->mixed && (block == 0 || !h->shared_ref)) {
    // reference state of the mixed path: interior mean of this block's state;
    // blocks joined by interfaces exchange FP32 differences with it, so they
    // all take block 0's (requests of the serving loop keep their own)
    double mean[5] = {0, 0, 0, 0, 0};
    const Geom &g = bd.g;
    for (int v = 0; v < 5; v++) {
      double acc = 0.0;
      for (int k = 1; k <= g.nz; k++)
        for (int j = 1; j <= g.ny; j++) {
          const double *row = src[v] + lin(g, 1, j, k);
          for (int i = 0; i < g.nx; i++) acc += row[i];
        }
      mean[v] = acc / ((double)g.nx * g.ny * g.nz);
    }
    // the mean in the caller's frame, then in each block's stored axis order
    double mc[5];
    for (int v = 0; v < 5; v++) mc[caller_var(bd.vperm, v)] = mean[v];
    auto ref_of = [&](int perm) {
      RefState r{};
      r.r = mc[caller_var(perm, 0)];
      r.ru = mc[caller_var(perm, 1)];
      r.rv = mc[caller_var(perm, 2)];
      r.rw = mc[caller_var(perm, 3)];
      r.re = mc[caller_var(perm, 4)];
      r.u = r.ru / r.r;
      r.v = r.rv / r.r;
      r.w = r.rw / r.r;
      r.p = (bd.p.gam - 1.0) * (r.re - 0.5 * r.r * (r.u * r.u + r.v * r.v + r.w * r.w));
      r.H = r.re + r.p;
      return r;
    };
    if (h->shared_ref)
      for (auto &b : h->b) b.ref = ref_of(b.vperm);
    else
      bd.ref = ref_of(bd.vperm);
    h->tables_ready = false;
  }
  return 0;
}

int get_state(mgpu_handle *h, int block, double *const *dst) {
  if (block < 0 || block >= (int)h->b.size()) return fail(h, MGPU_EINVAL, "block %d out of range", block);
  BlockData &bd = h->b[block];
  const size_t bytes = sizeof(double) * (size_t)bd.g.n;
  for (int v = 0; v < 5; v++)
    if (!dst[v]) return fail(h, MGPU_EINVAL, "state array %d is null", v);
  int rc = refresh_ghosts(h);
  if (rc) return rc;
  HIPCK(h, hipStreamSynchronize(h->stream));
  if (!bd.perm) {
    for (int v = 0; v < 5; v++) HIPCK(h, hipMemcpy(dst[v], bd.st[h->cur][v], bytes, hipMemcpyDeviceToHost));
    return 0;
  }
  std::vector<double> tmp((size_t)bd.g.n);
  for (int v = 0; v < 5; v++) {
    HIPCK(h, hipMemcpy(tmp.data(), bd.st[h->cur][v], bytes, hipMemcpyDeviceToHost));
    relabel_array(bd.g, bd.perm, tmp.data(), dst[caller_var(bd.vperm, v)], false);
  }
  return 0;
}

int step(mgpu_handle *h, int nsteps) {
  if (nsteps < 0) return fail(h, MGPU_EINVAL, "nsteps=%d", nsteps);
  for (size_t bI = 0; bI < h->b.size(); bI++) {
    if (!h->b[bI].have_metrics) return fail(h, MGPU_ESTATE, "block %zu: metrics not set", bI);
    if (!h->b[bI].have_state) return fail(h, MGPU_ESTATE, "block %zu: state not set", bI);
    if ((h->b[bI].g.jwall || h->b[bI].g.iwall) && !h->b[bI].have_wall)
      return fail(h, MGPU_ESTATE, "block %zu: wall coefficients not set (mgpu_set_wall)", bI);
  }
  if (h->b.empty()) return 0;
  int rc = refresh_ghosts(h);
  if (!rc) rc = build_tables(h);
  if (!rc) rc = start_charac(h);
  if (rc) return rc;
  const bool graph = h->use_graph && !h->profiling;
  for (int s = 0; s < nsteps; s++) {
    rc = graph ? graph_step(h) : enqueue_step(h, h->cur);
    if (rc) return rc;
    h->cur = next_set(h->c, h->cur);
  }
  return 0;
}

void destroy(mgpu_handle *h) {
  if (!h) return;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  for (auto &r : h->recs) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (auto e : h->ev_pool) (void)hipEventDestroy(e);
  drop_graphs(h);
  for (auto &bd : h->b) free_block(bd);
  free_exchange(h);
  free_tables(h);
  if (h->diag_buf) (void)hipFree(h->diag_buf);
  if (h->cfl_buf) (void)hipFree(h->cfl_buf);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

int diagnostics(mgpu_handle *h, mgpu_diag *out) {
  for (size_t bI = 0; bI < h->b.size(); bI++) {
    if (!h->b[bI].have_metrics) return fail(h, MGPU_ESTATE, "block %zu: metrics not set", bI);
    if (!h->b[bI].have_state) return fail(h, MGPU_ESTATE, "block %zu: state not set", bI);
  }
  int rc = refresh_ghosts(h);
  if (rc) return rc;
  double sum[3] = {0.0, 0.0, 0.0}, umax = 0.0;
  int64_t cells = 0;
  std::vector<double> part;
  for (auto &bd : h->b) {
    const long ng = diag_groups(bd.g);
    if (4 * ng > h->diag_cap) {
      HIPCK(h, hipStreamSynchronize(h->stream));
      if (h->diag_buf) (void)hipFree(h->diag_buf);
      h->diag_buf = nullptr;
      h->diag_cap = 0;
      HIPCK(h, hipMalloc((void **)&h->diag_buf, sizeof(double) * 4 * ng));
      h->diag_cap = 4 * ng;
    }
    DiagArgs da{};
    da.g = bd.g;
    for (int v = 0; v < 5; v++) da.u.v[v] = bd.st[h->cur][v];
    da.met = metrics_of(bd);
    da.gam = bd.p.gam;
    da.gam1 = bd.p.gam - 1.0;
    da.partial = h->diag_buf;
    HIPCK(h, launch_diag(da, h->c.curvilinear, h->stream));
    HIPCK(h, hipStreamSynchronize(h->stream));
    part.resize(4 * ng);
    HIPCK(h, hipMemcpy(part.data(), h->diag_buf, sizeof(double) * 4 * ng, hipMemcpyDeviceToHost));
    for (long w = 0; w < ng; w++) {
      for (int q = 0; q < 3; q++) sum[q] += part[4 * w + q];
      umax = part[4 * w + 3] > umax ? part[4 * w + 3] : umax;
    }
    cells += (int64_t)bd.g.nx * bd.g.ny * bd.g.nz;
  }
  out->mass = sum[0];
  out->kinetic = sum[1];
  out->enstrophy = sum[2];
  out->umax = umax;
  out->cells = cells;
  return 0;
}

// mgpu_diagnostics without any device allocation: the per-workgroup results
// go to the static array of k_diag.hip, block by block; the block totals are
// added on the host in block order.  The exchange tables must exist already
// (mgpu_create builds them), since building them allocates; stale ghosts are
// refreshed with the exchange kernel alone.
int diagnostics_noalloc(mgpu_handle *h, mgpu_diag *out) {
  for (size_t bI = 0; bI < h->b.size(); bI++) {
    if (!h->b[bI].have_metrics) return fail(h, MGPU_ESTATE, "block %zu: metrics not set", bI);
    if (!h->b[bI].have_state) return fail(h, MGPU_ESTATE, "block %zu: state not set", bI);
    const long ng = diag_groups(h->b[bI].g);
    if (ng > kDiagScratchGroups)
      return fail(h, MGPU_EINVAL,
                  "block %zu: %ld interior cells; the allocation-free diagnostics take at most %ld per block",
                  bI, (long)h->b[bI].g.nx * h->b[bI].g.ny * h->b[bI].g.nz, kDiagScratchGroups * kDiagThreads);
  }
  if (!h->exchange_ready)
    return fail(h, MGPU_ESTATE,
                "exchange tables not built (building them allocates); mgpu_create builds them");
  if (h->ghosts_dirty) {
    HIPCK(h, launch_iface(h->xs[h->cur], h->nchunks_s, h->stream));
    h->ghosts_dirty = false;
  }
  double sum[3] = {0.0, 0.0, 0.0}, umax = 0.0;
  int64_t cells = 0;
  for (auto &bd : h->b) {
    DiagArgs da{};
    da.g = bd.g;
    for (int v = 0; v < 5; v++) da.u.v[v] = bd.st[h->cur][v];
    da.met = metrics_of(bd);
    da.gam = bd.p.gam;
    da.gam1 = bd.p.gam - 1.0;
    double tot[4];
    HIPCK(h, diag_block_noalloc(da, h->c.curvilinear, h->device, h->stream, tot));
    for (int q = 0; q < 3; q++) sum[q] += tot[q];
    umax = tot[3] > umax ? tot[3] : umax;
    cells += (int64_t)bd.g.nx * bd.g.ny * bd.g.nz;
  }
  out->mass = sum[0];
  out->kinetic = sum[1];
  out->enstrophy = sum[2];
  out->umax = umax;
  out->cells = cells;
  return 0;
}

// Maximum CFL number along the caller's i, j, k over every block
// (check_max_cfl): per-workgroup maxima of each block, maximum on the host;
// a relabeled block's stored axes are mapped back to the caller's.
int max_cfl(mgpu_handle *h, double out[3]) {
  for (size_t bI = 0; bI < h->b.size(); bI++) {
    if (!h->b[bI].have_metrics) return fail(h, MGPU_ESTATE, "block %zu: metrics not set", bI);
    if (!h->b[bI].have_state) return fail(h, MGPU_ESTATE, "block %zu: state not set", bI);
  }
  double cm[3] = {0.0, 0.0, 0.0};
  std::vector<double> part;
  for (auto &bd : h->b) {
    const long ng = diag_groups(bd.g);
    if (3 * ng > h->cfl_cap) {
      HIPCK(h, hipStreamSynchronize(h->stream));
      if (h->cfl_buf) (void)hipFree(h->cfl_buf);
      h->cfl_buf = nullptr;
      h->cfl_cap = 0;
      HIPCK(h, hipMalloc((void **)&h->cfl_buf, sizeof(double) * 3 * ng));
      h->cfl_cap = 3 * ng;
    }
    CflArgs ca{};
    ca.g = bd.g;
    for (int v = 0; v < 5; v++) ca.u.v[v] = bd.st[h->cur][v];
    ca.met = metrics_of(bd);
    ca.gam = bd.p.gam;
    ca.gam1 = bd.p.gam - 1.0;
    ca.rg = bd.p.rg;
    ca.dt = bd.p.deltat;
    ca.partial = h->cfl_buf;
    HIPCK(h, launch_cfl(ca, h->c.curvilinear, h->stream));
    part.resize(3 * ng);
    HIPCK(h, hipMemcpyAsync(part.data(), h->cfl_buf, sizeof(double) * 3 * ng, hipMemcpyDeviceToHost, h->stream));
    HIPCK(h, hipStreamSynchronize(h->stream));
    for (int a = 0; a < 3; a++) {
      const int s = kAxisMap[bd.perm][a];  // the caller's axis a is the block's stored axis s
      for (long w = 0; w < ng; w++) cm[a] = part[3 * w + s] > cm[a] ? part[3 * w + s] : cm[a];
    }
  }
  for (int a = 0; a < 3; a++) out[a] = cm[a];
  return 0;
}

// k-chunks of k_visc, k_vdiv and the filter: every candidate timed on three
// launches of the kernel on the case's own data (the filter writes into a
// stage buffer, never into U^n).
int tune_aux(mgpu_handle *h, bool force, hipEvent_t e0, hipEvent_t e1) {
  const int nb = (int)h->b.size();
  const int geo = h->c.curvilinear;
  const int s0 = h->cur, sa = (h->cur + 1) % 3;
  int maxnz = 0;
  for (const auto &bd : h->b) maxnz = bd.g.nz > maxnz ? bd.g.nz : maxnz;
  const int kcs[] = {0, 4, 8, 16, 32, 64, 128, 256};
  int rc = 0;
  for (int kind = 0; kind < 3 && !rc; kind++) {
    if (kind < 2 && !h->c.iorder_visc) continue;
    if (kind == 2 && h->c.filter == MGPU_FILTER_NONE) continue;
    TuneKey key{h->device, h->c.curvilinear, 10 + 3 * h->c.filter + kind + (h->mixed ? 100 : 0), {}};
    for (const auto &bd : h->b) key.dims.insert(key.dims.end(), {bd.g.nx, bd.g.ny, bd.g.nz});
    if (!force) {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      auto it = g_tune_cache.find(key);
      if (it != g_tune_cache.end()) {
        h->jf_aux[kind] = it->second.first;
        h->kc_aux[kind] = it->second.second;
        h->tables_ready = false;
        continue;
      }
    }
    float best = 1e30f;
    int best_kc = 0, best_jf = 0;
    for (int cand = 0; cand < 2 * (int)(sizeof kcs / sizeof kcs[0]) && !rc; cand++) {
      const int jf = cand & 1, kc = kcs[cand >> 1];
      if (kc > 16 && kc / 2 >= maxnz) break;
      h->kc_aux[kind] = kc;
      h->jf_aux[kind] = jf;
      h->tables_ready = false;
      if ((rc = build_tables(h))) break;
      if (kind == 2) {
        std::vector<Batch<FilterArgs>> tab(nb);
        int first = 0;
        for (int b = 0; b < nb; b++) {
          tab[b].a = filter_args(h, h->b[b], s0, sa);
          tab[b].t = filter_tiles(h->b[b].g, kc);
          tab[b].t.jfirst = jf;
          tab[b].t.first = first;
          first += tile_count(tab[b].t, h->b[b].g.nz);
        }
        HIPCK(h, hipMemcpy(h->d_tune_f, tab.data(), nb * sizeof(tab[0]), hipMemcpyHostToDevice));
      }
      auto launch = [&]() -> hipError_t {
        if (kind == 0) return launch_visc(h->d_visc + s0 * nb, nb, h->nt_visc, geo, h->stream);
        if (kind == 1) return launch_vdiv(h->d_vdiv, nb, h->nt_vdiv, geo, h->stream);
        if (h->c.filter == MGPU_FILTER_SF)
          return launch_filter_sf(h->d_tune_f, nb, h->nt_filt, h->bnd, h->stream);
        return launch_filter_shock(h->d_tune_f, nb, h->nt_filt, dissip_kind(h->c, h->wall), h->bnd,
                                   edoh_on(h->c), h->stream);
      };
      hipError_t e = launch();
      if (e == hipSuccess) e = hipEventRecord(e0, h->stream);
      for (int r = 0; r < 3 && e == hipSuccess; r++) e = launch();
      if (e == hipSuccess) e = hipEventRecord(e1, h->stream);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      float ms = 0.f;
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      if (e != hipSuccess) {
        rc = fail(h, MGPU_EDEVICE, "autotune launch failed: %s", hipGetErrorString(e));
        break;
      }
      if (ms < best) {
        best = ms;
        best_kc = kc;
        best_jf = jf;
      }
    }
    h->kc_aux[kind] = best_kc;
    h->jf_aux[kind] = best_jf;
    h->tables_ready = false;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[key] = {best_jf, best_kc};
  }
  return rc;
}

}  // namespace mgpu_host

using namespace mgpu_host;

extern "C" {

int mgpu_create(const mgpu_case *c, mgpu_handle **out) {
  if (!c || !out) return fail(nullptr, MGPU_EINVAL, "null argument");
  *out = nullptr;
  if (c->nblocks < 1 || !c->blocks) return fail(nullptr, MGPU_EINVAL, "nblocks must be >= 1");
  {
    const std::string msg = check_flags(*c);
    if (!msg.empty()) return fail(nullptr, MGPU_EINVAL, "%s", msg.c_str());
  }
  for (int bI = 0; bI < c->nblocks; bI++) {
    const mgpu_block &bl = c->blocks[bI];
    if (bl.nx < G || bl.ny < G || bl.nz < G)
      return fail(nullptr, MGPU_EINVAL, "block %d: every extent must be >= %d", bI, G);
    const size_t padded = (size_t)(bl.nx + 2 * G) * (bl.ny + 2 * G) * (bl.nz + 2 * G);
    if (padded > kMaxPadded)
      return fail(nullptr, MGPU_EINVAL, "block %d: %zu padded points exceed the 2^28 limit of one block", bI,
                  padded);
  }
  std::vector<mgpu_block> stored;
  std::vector<int> perm;
  {
    std::string msg = check_cartesian_frames(*c);
    if (msg.empty()) msg = check_charac(*c);
    if (msg.empty()) msg = check_c3_walls(*c);
    if (msg.empty()) msg = relabel_blocks(*c, stored, perm);
    mgpu_case cs = *c;
    cs.blocks = stored.data();
    if (msg.empty()) msg = check_topology(cs);
    if (msg.empty()) msg = check_wall_rows(cs);
    if (!msg.empty()) return fail(nullptr, MGPU_EINVAL, "%s", msg.c_str());
  }
  mgpu_handle *h = nullptr;
  int rc = handle_new(*c, &h);
  if (rc) return rc;
  for (int bI = 0; bI < c->nblocks && !rc; bI++) {
    rc = handle_add_block(h, stored[bI], *c);
    if (!rc) {
      h->b[bI].perm = perm[bI];
      h->b[bI].vperm = c->curvilinear ? 0 : perm[bI];
    }
  }
  if (!rc) rc = setup_exchange(h);
  if (rc) {
    g_create_error = h->err;
    destroy(h);
    return rc;
  }
  h->shared_ref = true;
  *out = h;
  return 0;
}

int mgpu_set_metrics_cart(mgpu_handle *h, int block, const double *idx, const double *idy,
                          const double *idz, const double *idx_v, const double *idy_v,
                          const double *idz_v) {
  if (!h) return MGPU_EINVAL;
  const double *src[6] = {idx, idy, idz, idx_v, idy_v, idz_v};
  return put_metrics_cart(h, block, src);
}

int mgpu_set_metrics_c3(mgpu_handle *h, int block, const double *const *m, const double *const *mv) {
  if (!h) return MGPU_EINVAL;
  return put_metrics_c3(h, block, m, mv);
}

int mgpu_set_metrics_c2(mgpu_handle *h, int block, const double *const *m2, const double *const *m2v,
                        const double *idz, const double *idz_v) {
  if (!h) return MGPU_EINVAL;
  return put_metrics_c2(h, block, m2, m2v, idz, idz_v);
}

int mgpu_set_wall(mgpu_handle *h, int block, const double *cextrp2, const double *cextrp3) {
  if (!h) return MGPU_EINVAL;
  if (block < 0 || block >= (int)h->b.size()) return fail(h, MGPU_EINVAL, "block %d out of range", block);
  if (!cextrp2 || !cextrp3) return fail(h, MGPU_EINVAL, "null argument");
  BlockData &bd = h->b[block];
  if (!bd.g.jwall && !bd.g.iwall) return fail(h, MGPU_ESTATE, "block %d has no wall face", block);
  if (h->c.curvilinear == MGPU_GEOM_C3)
    return fail(h, MGPU_ESTATE, "block %d: curvilinear walls take their coefficients from mgpu_set_wall_c3", block);
  for (int s = 0; s < 2; s++) {
    if (!((bd.g.jwall >> s) & 1)) continue;
    const int fu = caller_face(bd.perm, 2 + s);  // the caller's face stored as j-face 2 + s
    const double a = cextrp2[fu], b = cextrp3[fu];
    if (!std::isfinite(a) || !std::isfinite(b))
      return fail(h, MGPU_EINVAL, "block %d face %d: extrapolation coefficients not finite", block, fu);
    bd.c2[s] = a;
    bd.c3[s] = b;
  }
  for (int s = 0; s < 2; s++) {
    if (!((bd.g.iwall >> s) & 1)) continue;
    const int fu = caller_face(bd.perm, s);  // the caller's face stored as i-face s
    const double a = cextrp2[fu], b = cextrp3[fu];
    if (!std::isfinite(a) || !std::isfinite(b))
      return fail(h, MGPU_EINVAL, "block %d face %d: extrapolation coefficients not finite", block, fu);
    bd.ci2[s] = a;
    bd.ci3[s] = b;
  }
  HIPCK(h, hipStreamSynchronize(h->stream));
  drop_graphs(h);  // the captured steps hold the coefficients
  bd.have_wall = true;
  return 0;
}

// A c3 block's walls are ready once every wall face has its coefficients and
// every slip face its dual tangents.
bool c3_walls_ready(const BlockData &bd) {
  for (int q = 0; q < 2; q++) {
    if (!((bd.g.jwall >> q) & 1)) continue;
    if (!bd.wc3_set[q] || (((bd.jslip >> q) & 1) && !bd.wt3_set[q])) return false;
  }
  return true;
}

int mgpu_set_wall_c3(mgpu_handle *h, int block, int face, const double *const *coef) {
  if (!h) return MGPU_EINVAL;
  if (block < 0 || block >= (int)h->b.size()) return fail(h, MGPU_EINVAL, "block %d out of range", block);
  if (!coef) return fail(h, MGPU_EINVAL, "null argument");
  BlockData &bd = h->b[block];
  if (h->c.curvilinear != MGPU_GEOM_C3) return fail(h, MGPU_ESTATE, "case is not 3-D curvilinear: use mgpu_set_wall");
  if (face < 0 || face > 5) return fail(h, MGPU_EINVAL, "face %d out of range", face);
  // the face as stored (a j-face: walls on the i- or k-faces are stored
  // there, with the face's two tangential axes in the same order)
  const int fs = kFaceMap[bd.perm][face];
  if (fs != 2 && fs != 3) return fail(h, MGPU_ESTATE, "block %d face %d is no wall", block, face);
  const int s = fs - 2;
  if (!bd.wc3[s]) return fail(h, MGPU_ESTATE, "block %d face %d is no wall", block, face);
  const size_t nf = (size_t)bd.g.nx * bd.g.nz;
  std::vector<double> tmp(8 * nf);
  for (int q = 0; q < 8; q++) {
    if (!coef[q]) return fail(h, MGPU_EINVAL, "coefficient array %d is null", q);
    for (size_t e = 0; e < nf; e++) {
      if (!std::isfinite(coef[q][e]))
        return fail(h, MGPU_EINVAL, "block %d face %d: coefficient %d not finite at %zu", block, face, q, e);
      tmp[q * nf + e] = coef[q][e];
    }
  }
  HIPCK(h, hipStreamSynchronize(h->stream));
  HIPCK(h, hipMemcpy(bd.wc3[s], tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice));
  bd.wc3_set[s] = true;
  bd.have_wall = c3_walls_ready(bd);
  return 0;
}

int mgpu_set_wall_c3_slip(mgpu_handle *h, int block, int face, const double *const *t) {
  if (!h) return MGPU_EINVAL;
  if (block < 0 || block >= (int)h->b.size()) return fail(h, MGPU_EINVAL, "block %d out of range", block);
  if (!t) return fail(h, MGPU_EINVAL, "null argument");
  BlockData &bd = h->b[block];
  if (h->c.curvilinear != MGPU_GEOM_C3) return fail(h, MGPU_ESTATE, "case is not 3-D curvilinear");
  if (face < 0 || face > 5) return fail(h, MGPU_EINVAL, "face %d out of range", face);
  const int fs = kFaceMap[bd.perm][face];
  if (fs != 2 && fs != 3) return fail(h, MGPU_ESTATE, "block %d face %d is no slip wall", block, face);
  const int s = fs - 2;
  if (!bd.wt3[s]) return fail(h, MGPU_ESTATE, "block %d face %d is no slip wall", block, face);
  const size_t nf = (size_t)bd.g.nx * bd.g.nz;
  std::vector<double> tmp(6 * nf);
  for (int q = 0; q < 6; q++) {
    if (!t[q]) return fail(h, MGPU_EINVAL, "tangent array %d is null", q);
    for (size_t e = 0; e < nf; e++) {
      if (!std::isfinite(t[q][e]))
        return fail(h, MGPU_EINVAL, "block %d face %d: tangent %d not finite at %zu", block, face, q, e);
      tmp[q * nf + e] = t[q][e];
    }
  }
  HIPCK(h, hipStreamSynchronize(h->stream));
  HIPCK(h, hipMemcpy(bd.wt3[s], tmp.data(), tmp.size() * sizeof(double), hipMemcpyHostToDevice));
  bd.wt3_set[s] = true;
  bd.have_wall = c3_walls_ready(bd);
  return 0;
}

int mgpu_put_state(mgpu_handle *h, int block, const double *rho, const double *rhou, const double *rhov,
                   const double *rhow, const double *rhoe) {
  if (!h) return MGPU_EINVAL;
  const double *src[5] = {rho, rhou, rhov, rhow, rhoe};
  return put_state(h, block, src);
}

int mgpu_get_state(mgpu_handle *h, int block, double *rho, double *rhou, double *rhov, double *rhow,
                   double *rhoe) {
  if (!h) return MGPU_EINVAL;
  double *dst[5] = {rho, rhou, rhov, rhow, rhoe};
  return get_state(h, block, dst);
}

int mgpu_step(mgpu_handle *h, int nsteps) {
  if (!h) return MGPU_EINVAL;
  return step(h, nsteps);
}

int mgpu_autotune(mgpu_handle *h, int flags) {
  if (!h) return MGPU_EINVAL;
  for (size_t bI = 0; bI < h->b.size(); bI++) {
    if (!h->b[bI].have_metrics) return fail(h, MGPU_ESTATE, "block %zu: metrics not set", bI);
    if (!h->b[bI].have_state) return fail(h, MGPU_ESTATE, "block %zu: state not set", bI);
  }
  const bool force = (flags & 1) != 0;
  int rc = build_tables(h);
  if (!rc) rc = refresh_ghosts(h);
  if (rc) return rc;
  const int nb = (int)h->b.size();
  const int s0 = h->cur, sa = (h->cur + 1) % 3;
  const int geo = h->c.curvilinear;
  hipEvent_t e0 = get_event(h), e1 = get_event(h);
  if (!e0 || !e1) return fail(h, MGPU_EDEVICE, "hipEventCreate failed");
  for (int v = 0; v < (h->c.iorder_visc ? 2 : 1) && !rc; v++) {
    TuneKey key{h->device, h->c.curvilinear, v + (h->mixed ? 2 : 0), h->stage_family, {}};
    for (const auto &bd : h->b) key.dims.insert(key.dims.end(), {bd.g.nx, bd.g.ny, bd.g.nz});
    if (!force) {
      std::lock_guard<std::mutex> lk(g_tune_mu);
      auto it = g_tune_cache.find(key);
      if (it != g_tune_cache.end()) {
        h->shape[v] = it->second.first & 0xff;
        h->jfirst[v] = it->second.first >> 8;
        h->kchunk[v] = it->second.second;
        continue;
      }
    }
    // candidates: every tile shape x k-chunks from 16 planes to the whole
    // block (0 = the per-block default), timed on one batched launch of
    // the stage (U^n in, a stage buffer out)
    float best = 1e30f;
    int best_shape = h->shape[v], best_kc = h->kchunk[v], best_jf = h->jfirst[v];
    const int kcs[] = {0, 4, 8, 16, 32, 64, 128, 256};
    int maxnz = 0;
    for (const auto &bd : h->b) maxnz = bd.g.nz > maxnz ? bd.g.nz : maxnz;
    std::vector<Batch<StageArgs>> tab(nb);
    for (int sj = 0; sj < 2 * kNumStageShapes && !rc; sj++) {
      const int sh = sj >> 1, jf = sj & 1;
      if ((h->stage_family == 1 && sh >= kFirstPipelinedShape) || (h->stage_family == 2 && sh < kFirstPipelinedShape))
        continue;
      for (int kc : kcs) {
        if (kc > 16 && kc / 2 >= maxnz) break;  // same as the previous candidate for every block
        h->shape[v] = sh;
        h->kchunk[v] = kc;
        h->jfirst[v] = jf;
        std::vector<TileRange> tr;
        const int nt = stage_ranges(h, v, tr);
        // Euler variant timed as an intermediate stage, viscous variant as the last one
        const int irk = v == 1 ? h->c.nrk : (h->c.nrk > 2 ? 2 : 1);
        for (int b = 0; b < nb; b++) {
          tab[b].a = stage_args(h, h->b[b], s0, s0, sa, irk);
          tab[b].t = tr[b];
        }
        const int mode = stage_mode(h, irk);
        hipError_t e = hipMemcpy(h->d_tune, tab.data(), nb * sizeof(tab[0]), hipMemcpyHostToDevice);
        // one warm-up launch, then three timed ones
        if (e == hipSuccess) e = launch_stage(h->d_tune, nb, nt, geo, v == 1, sh, mode, stage_wall(h), h->stream);
        if (e == hipSuccess) e = hipEventRecord(e0, h->stream);
        for (int r = 0; r < 3 && e == hipSuccess; r++)
          e = launch_stage(h->d_tune, nb, nt, geo, v == 1, sh, mode, stage_wall(h), h->stream);
        if (e == hipSuccess) e = hipEventRecord(e1, h->stream);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        if (e != hipSuccess) {
          rc = fail(h, MGPU_EDEVICE, "autotune launch failed: %s", hipGetErrorString(e));
          break;
        }
        if (ms < best) {
          best = ms;
          best_shape = sh;
          best_kc = kc;
          best_jf = jf;
        }
      }
    }
    h->shape[v] = best_shape;
    h->kchunk[v] = best_kc;
    h->jfirst[v] = best_jf;
    std::lock_guard<std::mutex> lk(g_tune_mu);
    g_tune_cache[key] = {best_shape | (best_jf << 8), best_kc};
  }
  if (!rc) rc = tune_aux(h, force, e0, e1);
  h->ev_pool.push_back(e0);
  h->ev_pool.push_back(e1);
  h->tables_ready = false;  // rebuilt with the chosen tiles before the next step
  return rc;
}

int mgpu_stage_config(mgpu_handle *h, int block, int variant, int *shape_tx, int *shape_ty, int *kchunk) {
  if (!h) return MGPU_EINVAL;
  if (block < 0 || block >= (int)h->b.size()) return fail(h, MGPU_EINVAL, "block %d out of range", block);
  if (variant < 0 || variant > 4) return fail(h, MGPU_EINVAL, "variant %d", variant);
  const BlockData &bd = h->b[block];
  if (variant >= 2) {
    const int kind = variant - 2;
    const TileRange t = kind == 0 ? visc_tiles(bd.g, h->kc_aux[0])
                                  : (kind == 1 ? vdiv_tiles(bd.g, h->kc_aux[1]) : filter_tiles(bd.g, h->kc_aux[2]));
    if (shape_tx) *shape_tx = 32;  // the viscous kernels and the filters use 32x16 tiles
    if (shape_ty) *shape_ty = 16;
    if (kchunk) *kchunk = t.kchunk;
    return 0;
  }
  if (shape_tx) *shape_tx = kStageShapes[h->shape[variant]][0];
  if (shape_ty) *shape_ty = kStageShapes[h->shape[variant]][1];
  if (kchunk) {
    const int kc = h->kchunk[variant] > 0 ? h->kchunk[variant] : kchunk_for(bd.g, h->shape[variant]);
    *kchunk = stage_tiles(bd.g, h->shape[variant], kc).kchunk;
  }
  return 0;
}

int mgpu_stage_pipelined(mgpu_handle *h, int variant) {
  if (!h) return MGPU_EINVAL;
  if (variant < 0 || variant > 1) return fail(h, MGPU_EINVAL, "variant %d: 0 Euler stages, 1 last stage", variant);
  return h->shape[variant] >= kFirstPipelinedShape ? 1 : 0;
}

int mgpu_sync(mgpu_handle *h) {
  if (!h) return MGPU_EINVAL;
  HIPCK(h, hipStreamSynchronize(h->stream));
  return 0;
}

int mgpu_profile(mgpu_handle *h, int enable) {
  if (!h) return MGPU_EINVAL;
  int rc = drain_profile(h);
  if (rc) return rc;
  for (int k = 0; k < MGPU_K_COUNT; k++) {
    h->kcount[k] = 0;
    h->kms[k] = 0.0;
  }
  h->profiling = enable != 0;
  return 0;
}

int mgpu_kernel_time(mgpu_handle *h, int kind, int64_t *launches, double *ms) {
  if (!h) return MGPU_EINVAL;
  if (kind < 0 || kind >= MGPU_K_COUNT) return fail(h, MGPU_EINVAL, "kernel kind %d", kind);
  int rc = drain_profile(h);
  if (rc) return rc;
  if (launches) *launches = h->kcount[kind];
  if (ms) *ms = h->kms[kind];
  return 0;
}

const char *mgpu_last_error(const mgpu_handle *h) {
  if (!h) return g_create_error.c_str();
  return h->err.c_str();
}

int mgpu_max_cfl(mgpu_handle *h, double cfl[3]) {
  if (!h) return MGPU_EINVAL;
  if (!cfl) return fail(h, MGPU_EINVAL, "null argument");
  return max_cfl(h, cfl);
}

int mgpu_exchange_plan(const mgpu_case *c, int set, int64_t *map, int64_t cap, int64_t *count) {
  if (!c || !count || (cap > 0 && !map)) return fail(nullptr, MGPU_EINVAL, "null argument");
  if (set != 0 && set != 1) return fail(nullptr, MGPU_EINVAL, "set=%d: 0 state, 1 viscous", set);
  if (c->nblocks < 1 || !c->blocks) return fail(nullptr, MGPU_EINVAL, "nblocks must be >= 1");
  std::vector<mgpu_block> stored;
  std::vector<int> perm;
  std::string msg = check_cartesian_frames(*c);
  if (msg.empty()) msg = check_charac(*c);
  if (msg.empty()) msg = check_c3_walls(*c);
  if (msg.empty()) msg = relabel_blocks(*c, stored, perm);
  mgpu_case cs = *c;
  cs.blocks = stored.data();
  if (msg.empty()) msg = check_topology(cs);
  if (!msg.empty()) return fail(nullptr, MGPU_EINVAL, "%s", msg.c_str());
  std::vector<BlockData> bds(c->nblocks);
  for (int b = 0; b < c->nblocks; b++) {
    const mgpu_block &bl = stored[b];
    Geom &g = bds[b].g;
    g.nx = bl.nx;
    g.ny = bl.ny;
    g.nz = bl.nz;
    g.sx = bl.nx + 2 * G;
    g.sy = (idx_t)g.sx * (idx_t)(bl.ny + 2 * G);
    g.n = (long)g.sy * (bl.nz + 2 * G);
    bds[b].perm = perm[b];
    bds[b].vperm = c->curvilinear ? 0 : perm[b];
  }
  XSet xs;
  xs.visc = set == 1;
  xs.sensor = xs.radii = set == 1;
  std::vector<FaceTask> tasks;
  std::vector<IfaceChunk> chunks;
  if (!build_tasks(cs, bds, set == 1 ? 2 : G, xs, tasks, chunks))
    return fail(nullptr, MGPU_EINVAL,
                "interface exchange would reach outside a block's arrays, or a face meets its neighbour's axes "
                "otherwise than the blocks' axis relabelings require");
  // stored element and array of block b -> the caller's element and array
  auto caller_elem = [&](int b, long off) -> long {
    const Geom &g = bds[b].g;
    const long Q0 = g.sx, Q1 = g.ny + 2 * G, Q2 = g.nz + 2 * G;
    const long x[3] = {off % Q0, (off / Q0) % Q1, off / (Q0 * Q1)}, Q[3] = {Q0, Q1, Q2};
    const int *m = kAxisMap[perm[b]];
    return x[m[0]] + Q[m[0]] * (x[m[1]] + Q[m[1]] * x[m[2]]);
  };
  auto plan_var = [&](int b, int v) -> int {
    const int p = perm[b], vp = bds[b].vperm;
    if (set == 0) return caller_var(vp, v);
    if (v < 6) return tau_index(caller_axis(vp, kTauPair[v][0]), caller_axis(vp, kTauPair[v][1]));
    if (v < NTAU) return 6 + caller_axis(vp, v - 6);
    if (v < NTAU + 2) return v;
    return NTAU + 2 + caller_axis(p, v - NTAU - 2);
  };
  const int nv = xs.nvar();
  int64_t n = 0;
  for (const FaceTask &t : tasks)
    for (int a2 = 0; a2 < t.n[2]; a2++)
      for (int a1 = 0; a1 < t.n[1]; a1++)
        for (int a0 = 0; a0 < t.n[0]; a0++) {
          const long pd = t.dst_base + (long)a0 * t.dst_s[0] + (long)a1 * t.dst_s[1] + (long)a2 * t.dst_s[2];
          const long ps = t.src_base + (long)a0 * t.src_s[0] + (long)a1 * t.src_s[1] + (long)a2 * t.src_s[2];
          for (int v = 0; v < nv; v++, n++) {
            if (n >= cap) continue;
            int64_t *e = map + 6 * n;
            e[0] = t.dst;
            e[1] = caller_elem(t.dst, pd);
            e[2] = plan_var(t.dst, v);
            e[3] = t.src;
            e[4] = caller_elem(t.src, ps);
            e[5] = plan_var(t.src, t.vmap[v]);
          }
        }
  *count = n;
  return 0;
}

int mgpu_diagnostics(mgpu_handle *h, mgpu_diag *out) {
  if (!h) return MGPU_EINVAL;
  if (!out) return fail(h, MGPU_EINVAL, "null argument");
  return diagnostics(h, out);
}

int mgpu_diagnostics_noalloc(mgpu_handle *h, mgpu_diag *out) {
  if (!h) return MGPU_EINVAL;
  if (!out) return fail(h, MGPU_EINVAL, "null argument");
  return diagnostics_noalloc(h, out);
}

void mgpu_destroy(mgpu_handle *h) { destroy(h); }

}  // extern "C"
