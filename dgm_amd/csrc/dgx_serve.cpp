// Serving loop: continuous batching of independent DG forward solves on one
// GPU (include/dgx.h, dgx_server_*).  No counterpart in the reference; it
// drives the step seam (dgx_advance) for optimisation and UQ ensembles.
//
// Requests are admitted FIFO against a slot count and a device-memory budget
// (dgx_domain_bytes).  Every admitted request owns a domain and a HIP stream
// of default priority, so the slices of different requests overlap on the
// device.  A slice is `steps_per_slice` steps: the first runs directly; the
// second is captured into a hipGraph (no memset nodes: the stepping kernels
// clear nothing), which later full slices replay.
//
// With batching, 2D Euler and shallow-water LSRK14 requests of the same order
// and precision (FP64, or DGX_MIXED: FP32 contractions) join a group instead: each stage of a group is two launches (face
// traces; residual with the fused register update) over the elements of all
// its members, on fixed grids of workgroups that walk the members' elements.
// Members step in lockstep slices of min(steps_per_slice, fewest remaining
// steps); they join and leave at slice boundaries.  With `persistent`, a
// group's whole slice is one cooperative launch of resident workgroups that
// step every stage with grid-wide barriers between them; all groups' slices
// then share one stream, so that two such launches never wait on each other.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/dgx.h"
#include "dgx_domain.hpp"
#include "dgx_kernels.hpp"

namespace {

// Failures keep their message for dgx_last_error(); s_err keeps it for the
// request that failed.
thread_local std::string s_err;

int sfail(int code, const std::string &msg) {
  s_err = msg;
  return dgx::set_error(code, msg);
}

struct Request {
  long long id = 0;
  dgx_domain_desc desc{};
  std::vector<double> vert, xnodes, U0, result;
  std::vector<int> con;
  int layout = DGX_NODAL, inttype = DGX_LSRK14, nsteps = 0;
  double dt = 0.0;
  int state = DGX_QUEUED;
  int done_steps = 0, inflight_steps = 0, slices = 0;
  bool inflight = false;
  size_t bytes = 0, state_len = 0;
  dgx_handle h = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  double t_submit = 0.0, t_admit = 0.0, t_done = 0.0;
  std::string msg;
  struct Group *g = nullptr;   // batched group, if any
};

// Batched requests of one (physics, order): members, the device copy of their
// descriptors, and the stream and event of the group's slices.
struct Group {
  int eqn = 0, p = 0;
  bool mixed = false;                // FP32 LDS tensors and contractions (DGX_MIXED members)
  std::vector<Request *> members;
  std::vector<Request *> stepping;   // the members of the slice in flight
  std::vector<dgx::Batch2> hd;
  dgx::Batch2 *dd = nullptr;
  size_t cap = 0;
  int nwork = 0, k = 0;
  bool dirty = true, inflight = false;
  bool shared_stream = false;        // persistent slices: the server's stream
  hipStream_t stream = nullptr;
  hipEvent_t ev = nullptr;
};

}  // namespace

struct dgx_server {
  dgx_server_config cfg{};
  std::chrono::steady_clock::time_point t0;
  long long next_id = 1;
  std::deque<Request *> queue;
  std::vector<Request *> running;
  std::map<long long, std::unique_ptr<Request>> all;
  std::vector<std::unique_ptr<Group>> groups;
  size_t used = 0, peak_bytes = 0;
  int peak_active = 0;
  hipStream_t pstream = nullptr;   // persistent slices of every group, one after another

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
};

namespace {

void release(Request *r) {
  if (r->exec) (void)hipGraphExecDestroy(r->exec);
  if (r->graph) (void)hipGraphDestroy(r->graph);
  if (r->h) (void)dgx_domain_destroy(r->h);
  if (r->ev) (void)hipEventDestroy(r->ev);
  if (r->stream) (void)hipStreamDestroy(r->stream);
  r->exec = nullptr; r->graph = nullptr; r->h = nullptr; r->ev = nullptr; r->stream = nullptr;
  std::vector<double>().swap(r->U0);
}

void fail_request(dgx_server *s, Request *r, const std::string &msg) {
  r->state = DGX_FAILED;
  r->msg = msg;
  r->t_done = s->now();
  release(r);
}

// Admit: own stream (default priority), domain on it, initial state uploaded.
int admit(dgx_server *s, Request *r) {
  if (hipStreamCreate(&r->stream) != hipSuccess) return sfail(DGX_ERR_DEVICE, "hipStreamCreate failed");
  if (hipEventCreateWithFlags(&r->ev, hipEventDisableTiming) != hipSuccess)
    return sfail(DGX_ERR_DEVICE, "hipEventCreate failed");
  int rc = dgx_domain_create(&r->desc, r->stream, &r->h);
  if (rc) return sfail(rc, dgx_last_error());
  int qtot, nm, nf, ne;
  if ((rc = dgx_sizes(r->h, &qtot, &nm, &nf, &ne))) return sfail(rc, dgx_last_error());
  r->state_len = (size_t)nf * ne * (r->layout == DGX_NODAL ? qtot : nm);
  if (r->U0.size() != r->state_len) return sfail(DGX_ERR_ARGUMENT, "initial state size does not match the domain");
  if ((rc = dgx_state_upload(r->h, r->U0.data(), r->layout, 0))) return sfail(rc, dgx_last_error());
  return DGX_OK;
}

// Enqueue the next slice of r on its stream.
int enqueue_slice(dgx_server *s, Request *r) {
  const int k = std::min(s->cfg.steps_per_slice, r->nsteps - r->done_steps);
  int rc;
  const bool graph = s->cfg.use_graphs && k == s->cfg.steps_per_slice && r->slices >= 1;
  if (graph && !r->exec) {
    if (hipStreamBeginCapture(r->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
      return sfail(DGX_ERR_DEVICE, "hipStreamBeginCapture failed");
    rc = dgx_advance(r->h, r->inttype, r->dt, k);
    hipError_t e = hipStreamEndCapture(r->stream, &r->graph);
    if (rc) return sfail(rc, dgx_last_error());
    if (e != hipSuccess) return sfail(DGX_ERR_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    if (hipGraphInstantiate(&r->exec, r->graph, nullptr, nullptr, 0) != hipSuccess)
      return sfail(DGX_ERR_DEVICE, "hipGraphInstantiate failed");
  }
  if (graph) {
    if (hipGraphLaunch(r->exec, r->stream) != hipSuccess) return sfail(DGX_ERR_DEVICE, "hipGraphLaunch failed");
  } else if ((rc = dgx_advance(r->h, r->inttype, r->dt, k))) {
    return sfail(rc, dgx_last_error());
  }
  if (hipEventRecord(r->ev, r->stream) != hipSuccess) return sfail(DGX_ERR_DEVICE, "hipEventRecord failed");
  r->inflight = true;
  r->inflight_steps = k;
  ++r->slices;
  return DGX_OK;
}

// Affine quads only: the batched kernels use constant metrics, so domains on
// the curved path (xnodes, non-parallelogram quads) step on their own.
bool batchable(const dgx_server *s, const Request *r) {
  const dgx_domain_desc &d = r->desc;
  return s->cfg.batch && d.nsd == 2 && !d.xnodes && r->h && !r->h->curved && !r->h->varp &&
         (d.eqntype == DGX_EULER || d.eqntype == DGX_SHALLOW) &&
         (d.precision == DGX_FP64 || d.precision == DGX_MIXED) &&
         r->inttype == DGX_LSRK14 && d.p >= 1 && d.p <= 10;
}

// Join r (admitted, state uploaded) to the group of its physics and order.
int join_group(dgx_server *s, Request *r) {
  Group *g = nullptr;
  const bool mixed = r->desc.precision == DGX_MIXED;
  for (auto &x : s->groups)
    if (x->eqn == r->desc.eqntype && x->p == r->desc.p && x->mixed == mixed) g = x.get();
  if (!g) {
    auto ng = std::make_unique<Group>();
    ng->eqn = r->desc.eqntype; ng->p = r->desc.p; ng->mixed = mixed;
    if (s->cfg.persistent) {
      // cooperative launches must not run side by side (each waits for all its
      // workgroups to be resident): every group's slices share one stream
      if (!s->pstream && hipStreamCreate(&s->pstream) != hipSuccess)
        return sfail(DGX_ERR_DEVICE, "hipStreamCreate failed");
      ng->stream = s->pstream;
      ng->shared_stream = true;
    } else if (hipStreamCreate(&ng->stream) != hipSuccess) {
      return sfail(DGX_ERR_DEVICE, "hipStreamCreate failed");
    }
    if (hipEventCreateWithFlags(&ng->ev, hipEventDisableTiming) != hipSuccess)
      return sfail(DGX_ERR_DEVICE, "hipEventCreate failed");
    g = ng.get();
    s->groups.push_back(std::move(ng));
  }
  int rc;
  // the balance term of shallow water is the first residual: store it now, as
  // the first stage would (its residual is then R(U0) - R(U0) = 0 there too)
  if (r->h->bal == 1 && (rc = dgx_rhs(r->h))) return sfail(rc, dgx_last_error());
  if (hipStreamSynchronize(r->stream) != hipSuccess) return sfail(DGX_ERR_DEVICE, "hipStreamSynchronize failed");
  r->g = g;
  g->members.push_back(r);
  g->dirty = true;
  return DGX_OK;
}

void leave_group(Request *r) {
  Group *g = r->g;
  if (!g) return;
  g->members.erase(std::find(g->members.begin(), g->members.end(), r));
  g->dirty = true;
  r->g = nullptr;
}

// Enqueue the group's next slice: k LSRK14 steps of every member.
int enqueue_group(dgx_server *s, Group *g) {
  int k = s->cfg.steps_per_slice;
  for (Request *r : g->members) k = std::min(k, r->nsteps - r->done_steps);
  if (k <= 0) return DGX_OK;
  if (g->dirty) {
    g->hd.resize(g->members.size());
    int off = 0;
    for (size_t i = 0; i < g->members.size(); ++i) {
      const dgx_domain *d = g->members[i]->h;
      dgx::Batch2 &b = g->hd[i];
      if (!d->T2) return sfail(DGX_ERR_ARGUMENT, "batched slice: a member has no second trace buffer");
      b.U = d->U; b.Ut = d->Ut; b.T = d->T; b.T2 = d->T2; b.Bal = d->Bal; b.flag = d->d_flag;
      b.geo = d->geo; b.phys = d->phys; b.ne = d->ne; b.off = off; b.bal = d->bal;
      b.dt = g->members[i]->dt;
      off += d->ne;
    }
    g->nwork = off;
    if (g->hd.size() > g->cap) {
      if (g->dd) (void)hipFree(g->dd);
      g->cap = std::max<size_t>(g->hd.size(), 2 * g->cap);
      if (hipMalloc((void **)&g->dd, g->cap * sizeof(dgx::Batch2)) != hipSuccess) {
        g->dd = nullptr; g->cap = 0;
        return sfail(DGX_ERR_MEMORY, "hipMalloc of the batch descriptors failed");
      }
    }
    if (hipMemcpyAsync(g->dd, g->hd.data(), g->hd.size() * sizeof(dgx::Batch2), hipMemcpyHostToDevice,
                       g->stream) != hipSuccess)
      return sfail(DGX_ERR_DEVICE, "descriptor upload failed");
    g->dirty = false;
  }
  dgx::BatchLaunch a;
  a.p = g->p; a.nreq = (int)g->hd.size(); a.nwork = g->nwork; a.shallow = g->eqn == DGX_SHALLOW;
  a.mixed = g->mixed; a.b = g->dd; a.stream = g->stream;
  if (s->cfg.persistent) {   // the whole slice in one cooperative launch
    a.op = 2; a.nsteps = k;
    std::copy(dgx::kLsrkA, dgx::kLsrkA + 14, a.lsrkA);
    std::copy(dgx::kLsrkB, dgx::kLsrkB + 14, a.lsrkB);
    const hipError_t e = dgx::launch_batch2(a);
    if (e != hipSuccess) return sfail(DGX_ERR_DEVICE, std::string("persistent slice: ") + hipGetErrorString(e));
  } else
  for (int step = 0; step < k; ++step)
    for (int st = 0; st < 14; ++st) {
      // the slice's first stage evaluates the traces; every stage then writes
      // the next stage's into the other buffer (T, T2 alternate; after the
      // even number of stages of a step the current ones are in T again)
      hipError_t e = hipSuccess;
      if (step == 0 && st == 0) {
        a.op = 0;
        e = dgx::launch_batch2(a);
      }
      if (e == hipSuccess) {
        a.op = 1; a.A = dgx::kLsrkA[st]; a.B = dgx::kLsrkB[st];
        a.odd = st & 1; a.next = !(step == k - 1 && st == 13);
        e = dgx::launch_batch2(a);
      }
      if (e != hipSuccess) return sfail(DGX_ERR_DEVICE, std::string("batched launch: ") + hipGetErrorString(e));
    }
  if (hipEventRecord(g->ev, g->stream) != hipSuccess) return sfail(DGX_ERR_DEVICE, "hipEventRecord failed");
  g->k = k;
  g->stepping = g->members;
  g->inflight = true;
  return DGX_OK;
}

// Final state to host memory, flag check, resources freed.
int finish(dgx_server *s, Request *r) {
  int rc = dgx_synchronize(r->h);
  if (rc) return sfail(rc, dgx_last_error());
  r->result.assign(r->state_len, 0.0);
  if ((rc = dgx_state_download(r->h, r->result.data(), r->layout, 0))) return sfail(rc, dgx_last_error());
  if ((rc = dgx_synchronize(r->h))) return sfail(rc, dgx_last_error());
  r->state = DGX_DONE;
  r->t_done = s->now();
  release(r);
  return DGX_OK;
}

}  // namespace

extern "C" {

int dgx_server_create(const dgx_server_config *cfg, dgx_server_handle *out) {
  if (!cfg || !out) return sfail(DGX_ERR_ARGUMENT, "null argument");
  *out = nullptr;
  if (cfg->hbm_budget == 0 || cfg->max_active < 1 || cfg->steps_per_slice < 1)
    return sfail(DGX_ERR_ARGUMENT, "hbm_budget, max_active and steps_per_slice must be positive");
  dgx_server *s = new dgx_server();
  s->cfg = *cfg;
  s->t0 = std::chrono::steady_clock::now();
  *out = s;
  return DGX_OK;
}

int dgx_server_destroy(dgx_server_handle s) {
  if (!s) return DGX_OK;
  for (auto &g : s->groups)
    if (g->stream) (void)hipStreamSynchronize(g->stream);
  for (Request *r : s->running) {
    if (r->stream) (void)hipStreamSynchronize(r->stream);
    release(r);
  }
  for (auto &g : s->groups) {
    if (g->dd) (void)hipFree(g->dd);
    if (g->ev) (void)hipEventDestroy(g->ev);
    if (g->stream && !g->shared_stream) (void)hipStreamDestroy(g->stream);
  }
  if (s->pstream) (void)hipStreamDestroy(s->pstream);
  delete s;
  return DGX_OK;
}

int dgx_server_submit(dgx_server_handle s, const dgx_domain_desc *desc, const double *U0, int layout,
                      int inttype, double dt, int nsteps, long long *id) {
  if (!s || !desc || !U0 || !id) return sfail(DGX_ERR_ARGUMENT, "null argument");
  if (layout != DGX_NODAL && layout != DGX_MODAL) return sfail(DGX_ERR_ARGUMENT, "bad layout");
  if (inttype != DGX_TVD && inttype != DGX_RK4 && inttype != DGX_LSRK14)
    return sfail(DGX_ERR_ARGUMENT, "unsupported inttype");
  if (nsteps < 1 || !(dt > 0.0)) return sfail(DGX_ERR_ARGUMENT, "nsteps and dt must be positive");
  if (!desc->vert || !desc->con || desc->ne <= 0 || desc->ncon <= 0)
    return sfail(DGX_ERR_ARGUMENT, "empty mesh");
  if (desc->bc_function || desc->source)
    return sfail(DGX_ERR_ARGUMENT, "boundary and source functions are not served; create the domain directly");
  auto r = std::make_unique<Request>();
  int rc = dgx_domain_bytes(desc, &r->bytes);
  if (rc) return sfail(rc, dgx_last_error());
  if (r->bytes > s->cfg.hbm_budget) return sfail(DGX_ERR_ARGUMENT, "request exceeds the memory budget on its own");
  const size_t nv = (size_t)desc->ne * desc->nsd * (desc->nsd == 2 ? 4 : 8);
  r->vert.assign(desc->vert, desc->vert + nv);
  r->con.assign(desc->con, desc->con + 4 * (size_t)desc->ncon);
  r->desc = *desc;
  r->desc.vert = r->vert.data();
  r->desc.con = r->con.data();
  const size_t L = desc->p + 1, Q = desc->p + 2, nf = desc->eqntype == DGX_SHALLOW ? 3 : desc->nsd + 2;
  if (desc->xnodes) {   // curved elements: the nodes are kept with the request like the mesh
    const size_t nx = desc->nsd == 2 ? 2 * (size_t)desc->ne * Q * Q : 3 * (size_t)desc->ne * Q * Q * Q;
    r->xnodes.assign(desc->xnodes, desc->xnodes + nx);
    r->desc.xnodes = r->xnodes.data();
  }
  const size_t per = desc->nsd == 2 ? (layout == DGX_NODAL ? Q * Q : L * L)
                                    : (layout == DGX_NODAL ? Q * Q * Q : L * L * L);
  r->U0.assign(U0, U0 + nf * desc->ne * per);
  r->layout = layout;
  r->inttype = inttype;
  r->dt = dt;
  r->nsteps = nsteps;
  r->id = s->next_id++;
  r->t_submit = s->now();
  *id = r->id;
  s->queue.push_back(r.get());
  s->all[r->id] = std::move(r);
  return DGX_OK;
}

int dgx_server_progress(dgx_server_handle s, int *pending) {
  if (!s) return sfail(DGX_ERR_ARGUMENT, "null handle");
  // finished group slices
  for (auto &g : s->groups) {
    if (!g->inflight) continue;
    hipError_t e = hipEventQuery(g->ev);
    if (e == hipErrorNotReady) continue;
    g->inflight = false;
    for (Request *r : g->stepping) {   // members that joined during the slice did not step
      if (e == hipSuccess) r->done_steps += g->k;
      else r->msg = std::string("batched slice failed: ") + hipGetErrorString(e);
    }
    g->stepping.clear();
  }
  // retire finished slices and requests
  for (size_t i = 0; i < s->running.size();) {
    Request *r = s->running[i];
    bool retire = false;
    if (r->g) {
      if (!r->msg.empty()) {
        leave_group(r);
        fail_request(s, r, r->msg);
        retire = true;
      }
    } else if (r->inflight) {
      hipError_t e = hipEventQuery(r->ev);
      if (e == hipSuccess) {
        r->inflight = false;
        r->done_steps += r->inflight_steps;
      } else if (e != hipErrorNotReady) {
        fail_request(s, r, std::string("slice failed: ") + hipGetErrorString(e));
        retire = true;
      }
    }
    const bool busy = r->g ? r->g->inflight : r->inflight;
    if (!retire && !busy && r->done_steps >= r->nsteps) {
      leave_group(r);
      if (finish(s, r)) fail_request(s, r, s_err);
      retire = true;
    }
    if (retire) {
      s->used -= r->bytes;
      s->running.erase(s->running.begin() + i);
    } else {
      ++i;
    }
  }
  // admit in arrival order while a slot and the budget allow
  while (!s->queue.empty()) {
    Request *r = s->queue.front();
    if ((int)s->running.size() >= s->cfg.max_active || s->used + r->bytes > s->cfg.hbm_budget) break;
    s->queue.pop_front();
    r->t_admit = s->now();
    if (admit(s, r) || (batchable(s, r) && join_group(s, r))) {
      leave_group(r);
      fail_request(s, r, s_err);
      continue;
    }
    r->state = DGX_RUNNING;
    s->used += r->bytes;
    s->running.push_back(r);
    if (s->used > s->peak_bytes) s->peak_bytes = s->used;
    if ((int)s->running.size() > s->peak_active) s->peak_active = (int)s->running.size();
  }
  // next slice of every group and every unbatched request with none in flight
  for (auto &g : s->groups) {
    if (g->inflight || g->members.empty()) continue;
    if (enqueue_group(s, g.get())) {
      const std::string msg = s_err;
      for (Request *r : std::vector<Request *>(g->members)) {
        leave_group(r);
        fail_request(s, r, msg);
        s->used -= r->bytes;
        s->running.erase(std::find(s->running.begin(), s->running.end(), r));
      }
    }
  }
  for (size_t i = 0; i < s->running.size();) {
    Request *r = s->running[i];
    if (!r->g && !r->inflight && r->done_steps < r->nsteps && enqueue_slice(s, r)) {
      fail_request(s, r, s_err);
      s->used -= r->bytes;
      s->running.erase(s->running.begin() + i);
      continue;
    }
    ++i;
  }
  if (pending) *pending = (int)(s->queue.size() + s->running.size());
  return DGX_OK;
}

int dgx_server_status(dgx_server_handle s, long long id, int *state) {
  if (!s || !state) return sfail(DGX_ERR_ARGUMENT, "null argument");
  auto it = s->all.find(id);
  if (it == s->all.end()) return sfail(DGX_ERR_ARGUMENT, "unknown request id");
  *state = it->second->state;
  return DGX_OK;
}

int dgx_server_result(dgx_server_handle s, long long id, double *U, int layout, double *times) {
  if (!s) return sfail(DGX_ERR_ARGUMENT, "null handle");
  auto it = s->all.find(id);
  if (it == s->all.end()) return sfail(DGX_ERR_ARGUMENT, "unknown request id");
  Request *r = it->second.get();
  if (r->state == DGX_FAILED) return sfail(DGX_ERR_DEVICE, "request failed: " + r->msg);
  if (r->state != DGX_DONE) return sfail(DGX_ERR_ARGUMENT, "request not finished");
  if (U) {
    if (layout != r->layout) return sfail(DGX_ERR_ARGUMENT, "results are kept in the submission layout");
    std::copy(r->result.begin(), r->result.end(), U);
  }
  if (times) { times[0] = r->t_submit; times[1] = r->t_admit; times[2] = r->t_done; }
  return DGX_OK;
}

int dgx_server_peak(dgx_server_handle s, size_t *bytes, int *active) {
  if (!s) return sfail(DGX_ERR_ARGUMENT, "null handle");
  if (bytes) *bytes = s->peak_bytes;
  if (active) *active = s->peak_active;
  return DGX_OK;
}

}  // extern "C"
