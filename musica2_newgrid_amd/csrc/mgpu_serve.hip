// Serving loop: many independent one-block cases ("requests") resident on
// one GPU at once.  The reference runs one case per MPI job and has no
// counterpart; the time step of every request is the reference's
// runge_kutta (src/Compute_flow/runge_kutta.f90:2-152) as mgpu_step runs it.
//
// Requests are admitted in submission order while their device memory fits
// the budget, keep their state resident, and advance together: resident
// requests with the same kernel variant (geometry kind, viscous terms,
// filter, RK stages, precision) form a group that owns one handle whose
// blocks are those requests, so each RK stage of the whole group is one
// batched launch per kernel (and, while the group's membership is stable,
// one replay of a captured graph per step).  Each group has its own stream,
// so groups also overlap on the device.  A request that finishes is copied
// back to host memory and its device memory is released at the next round.
#include <chrono>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <tuple>
#include <vector>

#include "mgpu_handle.h"

using namespace mgpu;
using namespace mgpu_host;

namespace {

struct Request {
  int64_t id = 0;
  mgpu_case c{};
  mgpu_block blk{};
  int nsteps = 0, done_steps = 0;
  int state = 0;  // 0 queued, 1 resident, 2 done
  size_t dev_bytes = 0;
  std::vector<double> metrics;  // host copies (6 1-D arrays, 20 padded arrays, or 10 2-D + 2 1-D)
  std::vector<double> u;        // 5 padded arrays: initial state, then the result
  size_t n = 0;                 // padded points
};

using Key = std::tuple<int, int, int, int, int, int>;

Key key_of(const mgpu_case &c) {
  return Key(c.curvilinear, c.iorder_visc, c.filter, c.is_ducros, c.nrk, c.precision);
}

struct Group {
  mgpu_handle *h = nullptr;
  std::vector<Request *> members;  // in the order of h's blocks
};

}  // namespace

struct mgpu_server {
  int64_t budget = 0, used = 0;
  int64_t next_id = 1;
  // co-batching record: group steps launched, request steps they advanced,
  // and the most requests one group step has advanced
  int64_t group_steps = 0, request_steps = 0;
  int max_members = 0;
  std::map<int64_t, std::unique_ptr<Request>> reqs;
  std::vector<Request *> queue;  // submission order
  std::map<Key, Group> groups;
  std::string err;
};

namespace {

int sfail(mgpu_server *s, int code, const std::string &msg) {
  if (s) s->err = msg;
  else g_create_error = msg;
  return code;
}

int retire(mgpu_server *s, Group &g, int i) {
  Request *r = g.members[i];
  double *dst[5];
  for (int v = 0; v < 5; v++) dst[v] = r->u.data() + v * r->n;
  int rc = get_state(g.h, i, dst);
  if (!rc) rc = handle_remove_block(g.h, i);
  if (rc) return sfail(s, rc, g.h->err);
  g.members.erase(g.members.begin() + i);
  r->state = 2;
  s->used -= (int64_t)r->dev_bytes;
  r->metrics.clear();
  r->metrics.shrink_to_fit();
  return 0;
}

int admit(mgpu_server *s, Request *r) {
  Group &g = s->groups[key_of(r->c)];
  int rc = 0;
  if (!g.h) {
    mgpu_case flags = r->c;
    rc = handle_new(flags, &g.h);
    if (rc) return sfail(s, rc, g_create_error);
  }
  rc = handle_add_block(g.h, r->blk, r->c);
  if (rc) return sfail(s, rc, g.h->err);
  const int b = (int)g.h->b.size() - 1;
  if (r->c.curvilinear == MGPU_GEOM_C2) {
    const size_t n2 = (size_t)(r->blk.nx + 2 * G) * (r->blk.ny + 2 * G);
    const double *m2[5], *m2v[5];
    for (int q = 0; q < 5; q++) {
      m2[q] = r->metrics.data() + q * n2;
      m2v[q] = r->metrics.data() + (5 + q) * n2;
    }
    const double *idz = r->metrics.data() + 10 * n2;
    const double *idz_v = idz + (r->blk.nz + 2 * G);
    const bool visc = r->c.iorder_visc != 0;
    rc = put_metrics_c2(g.h, b, m2, visc ? m2v : nullptr, idz, visc ? idz_v : nullptr);
  } else if (r->c.curvilinear) {
    const size_t n = r->n;
    const double *m[10], *mv[10];
    for (int q = 0; q < 10; q++) {
      m[q] = r->metrics.data() + q * n;
      mv[q] = r->metrics.data() + (10 + q) * n;
    }
    rc = put_metrics_c3(g.h, b, m, r->c.iorder_visc ? mv : nullptr);
  } else {
    const int len[3] = {r->blk.nx + 2 * G, r->blk.ny + 2 * G, r->blk.nz + 2 * G};
    const double *a[6];
    size_t off = 0;
    for (int q = 0; q < 6; q++) {
      a[q] = r->metrics.data() + off;
      off += len[q % 3];
    }
    rc = put_metrics_cart(g.h, b, a);
  }
  const double *u[5];
  for (int v = 0; v < 5; v++) u[v] = r->u.data() + v * r->n;
  if (!rc) rc = put_state(g.h, b, u);
  if (rc) {
    (void)handle_remove_block(g.h, b);
    return sfail(s, rc, g.h->err);
  }
  g.members.push_back(r);
  r->state = 1;
  s->used += (int64_t)r->dev_bytes;
  return 0;
}

}  // namespace

extern "C" {

int mgpu_server_create(int64_t hbm_budget_bytes, mgpu_server **out) {
  if (!out) return sfail(nullptr, MGPU_EINVAL, "null argument");
  *out = nullptr;
  if (hbm_budget_bytes <= 0) return sfail(nullptr, MGPU_EINVAL, "budget must be positive");
  mgpu_server *s = new mgpu_server();
  s->budget = hbm_budget_bytes;
  *out = s;
  return 0;
}

int mgpu_server_submit(mgpu_server *s, const mgpu_case *c, const double *const *metrics,
                       const double *const *state, int nsteps, int64_t *id) {
  if (!s || !c || !metrics || !state || !id) return sfail(s, MGPU_EINVAL, "null argument");
  if (c->nblocks != 1 || !c->blocks) return sfail(s, MGPU_EINVAL, "a request is one block (nblocks = 1)");
  const std::string msg = check_flags(*c);
  if (!msg.empty()) return sfail(s, MGPU_EINVAL, msg);
  const mgpu_block &bl = c->blocks[0];
  for (int f = 0; f < 6; f++)
    if (bl.face_kind[f] != MGPU_FACE_PERIODIC)
      return sfail(s, MGPU_EUNSUP, "a request's faces must be periodic");
  if (bl.nx < G || bl.ny < G || bl.nz < G) return sfail(s, MGPU_EINVAL, "block extents below 5");
  const size_t n = (size_t)(bl.nx + 2 * G) * (bl.ny + 2 * G) * (bl.nz + 2 * G);
  if (n > kMaxPadded) return sfail(s, MGPU_EINVAL, "block exceeds 2^28 padded points");
  if (nsteps < 0) return sfail(s, MGPU_EINVAL, "nsteps < 0");
  auto r = std::make_unique<Request>();
  r->id = s->next_id++;
  r->c = *c;
  r->c.blocks = nullptr;
  r->blk = bl;
  r->nsteps = nsteps;
  r->n = n;
  r->dev_bytes = block_bytes(*c, bl.nx, bl.ny, bl.nz);
  if ((int64_t)r->dev_bytes > s->budget) return sfail(s, MGPU_EINVAL, "request larger than the whole budget");
  if (c->curvilinear == MGPU_GEOM_C2) {
    const size_t n2 = (size_t)(bl.nx + 2 * G) * (bl.ny + 2 * G), nz = (size_t)bl.nz + 2 * G;
    const bool visc = c->iorder_visc != 0;
    r->metrics.resize(10 * n2 + 2 * nz, 0.0);
    for (int q = 0; q < (visc ? 10 : 5); q++) {
      if (!metrics[q]) return sfail(s, MGPU_EINVAL, "metric array is null");
      memcpy(r->metrics.data() + q * n2, metrics[q], n2 * sizeof(double));
    }
    for (int q = 0; q < (visc ? 2 : 1); q++) {
      if (!metrics[10 + q]) return sfail(s, MGPU_EINVAL, "metric array is null");
      memcpy(r->metrics.data() + 10 * n2 + q * nz, metrics[10 + q], nz * sizeof(double));
    }
  } else if (c->curvilinear) {
    const int na = c->iorder_visc ? 20 : 10;
    r->metrics.resize(20 * n, 0.0);
    for (int q = 0; q < na; q++) {
      if (!metrics[q]) return sfail(s, MGPU_EINVAL, "metric array is null");
      memcpy(r->metrics.data() + q * n, metrics[q], n * sizeof(double));
    }
  } else {
    const int len[3] = {bl.nx + 2 * G, bl.ny + 2 * G, bl.nz + 2 * G};
    r->metrics.resize(2 * (size_t)(len[0] + len[1] + len[2]), 0.0);
    size_t off = 0;
    for (int q = 0; q < 6; q++) {
      if (metrics[q]) memcpy(r->metrics.data() + off, metrics[q], len[q % 3] * sizeof(double));
      else if (q < 3 || c->iorder_visc) return sfail(s, MGPU_EINVAL, "metric array is null");
      off += len[q % 3];
    }
  }
  r->u.resize(5 * n);
  for (int v = 0; v < 5; v++) {
    if (!state[v]) return sfail(s, MGPU_EINVAL, "state array is null");
    memcpy(r->u.data() + v * n, state[v], n * sizeof(double));
  }
  *id = r->id;
  s->queue.push_back(r.get());
  s->reqs[r->id] = std::move(r);
  return 0;
}

int mgpu_server_tick(mgpu_server *s, int max_steps, int *n_resident, int *n_queued) {
  if (!s) return MGPU_EINVAL;
  if (max_steps < 1) return sfail(s, MGPU_EINVAL, "max_steps < 1");
  int rc = 0;
  // retire finished requests
  for (auto &kv : s->groups) {
    Group &g = kv.second;
    for (int i = (int)g.members.size() - 1; i >= 0; i--)
      if (g.members[i]->done_steps >= g.members[i]->nsteps && (rc = retire(s, g, i))) return rc;
  }
  // admit in submission order while the budget allows
  while (!s->queue.empty()) {
    Request *r = s->queue.front();
    if (s->used + (int64_t)r->dev_bytes > s->budget) break;
    if ((rc = admit(s, r))) return rc;
    s->queue.erase(s->queue.begin());
  }
  // advance every group: as many steps as its nearest finishing member needs (at most max_steps)
  int resident = 0;
  for (auto &kv : s->groups) {
    Group &g = kv.second;
    if (g.members.empty()) continue;
    int k = max_steps;
    for (Request *r : g.members) {
      const int left = r->nsteps - r->done_steps;
      if (left < k) k = left;
    }
    if (k > 0) {
      rc = step(g.h, k);
      if (rc) return sfail(s, rc, g.h->err);
      for (Request *r : g.members) r->done_steps += k;
      s->group_steps += k;
      s->request_steps += (int64_t)k * (int64_t)g.members.size();
      if ((int)g.members.size() > s->max_members) s->max_members = (int)g.members.size();
    }
    resident += (int)g.members.size();
  }
  if (n_resident) *n_resident = resident;
  if (n_queued) *n_queued = (int)s->queue.size();
  return 0;
}

int mgpu_server_status(mgpu_server *s, int64_t id, int *state, int *steps_done) {
  if (!s) return MGPU_EINVAL;
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return sfail(s, MGPU_EINVAL, "unknown request");
  if (state) *state = it->second->state;
  if (steps_done) *steps_done = it->second->done_steps;
  return 0;
}

int mgpu_server_take(mgpu_server *s, int64_t id, double *rho, double *rhou, double *rhov, double *rhow,
                     double *rhoe) {
  if (!s) return MGPU_EINVAL;
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return sfail(s, MGPU_EINVAL, "unknown request");
  Request *r = it->second.get();
  if (r->state != 2) return sfail(s, MGPU_ESTATE, "request not finished");
  double *dst[5] = {rho, rhou, rhov, rhow, rhoe};
  for (int v = 0; v < 5; v++) {
    if (!dst[v]) return sfail(s, MGPU_EINVAL, "state array is null");
    memcpy(dst[v], r->u.data() + v * r->n, r->n * sizeof(double));
  }
  s->reqs.erase(it);
  return 0;
}

int64_t mgpu_server_memory(const mgpu_server *s) { return s ? s->used : 0; }

int mgpu_server_batching(const mgpu_server *s, int64_t *group_steps, int64_t *request_steps, int32_t *max_members) {
  if (!s) return MGPU_EINVAL;
  if (group_steps) *group_steps = s->group_steps;
  if (request_steps) *request_steps = s->request_steps;
  if (max_members) *max_members = s->max_members;
  return 0;
}

const char *mgpu_server_last_error(const mgpu_server *s) {
  if (!s) return g_create_error.c_str();
  return s->err.c_str();
}

void mgpu_server_destroy(mgpu_server *s) {
  if (!s) return;
  for (auto &kv : s->groups) destroy(kv.second.h);
  delete s;
}

}  // extern "C"
