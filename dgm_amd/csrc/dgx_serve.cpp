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
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <deque>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/dgx.h"
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
};

}  // namespace

struct dgx_server {
  dgx_server_config cfg{};
  std::chrono::steady_clock::time_point t0;
  long long next_id = 1;
  std::deque<Request *> queue;
  std::vector<Request *> running;
  std::map<long long, std::unique_ptr<Request>> all;
  size_t used = 0, peak_bytes = 0;
  int peak_active = 0;

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
  for (Request *r : s->running) {
    if (r->stream) (void)hipStreamSynchronize(r->stream);
    release(r);
  }
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
  if (desc->xnodes) {   // curved hexes: the nodes are kept with the request like the mesh
    r->xnodes.assign(desc->xnodes, desc->xnodes + 3 * (size_t)desc->ne * Q * Q * Q);
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
  // retire finished slices and requests
  for (size_t i = 0; i < s->running.size();) {
    Request *r = s->running[i];
    bool retire = false;
    if (r->inflight) {
      hipError_t e = hipEventQuery(r->ev);
      if (e == hipSuccess) {
        r->inflight = false;
        r->done_steps += r->inflight_steps;
      } else if (e != hipErrorNotReady) {
        fail_request(s, r, std::string("slice failed: ") + hipGetErrorString(e));
        retire = true;
      }
    }
    if (!retire && !r->inflight && r->done_steps >= r->nsteps) {
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
    if (admit(s, r)) {
      fail_request(s, r, s_err);
      continue;
    }
    r->state = DGX_RUNNING;
    s->used += r->bytes;
    s->running.push_back(r);
    if (s->used > s->peak_bytes) s->peak_bytes = s->used;
    if ((int)s->running.size() > s->peak_active) s->peak_active = (int)s->running.size();
  }
  // next slice of every admitted request with none in flight
  for (size_t i = 0; i < s->running.size();) {
    Request *r = s->running[i];
    if (!r->inflight && r->done_steps < r->nsteps && enqueue_slice(s, r)) {
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
