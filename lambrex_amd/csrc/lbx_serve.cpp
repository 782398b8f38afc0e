// lbx_serve.cpp -- batched serving of independent periodic LBM runs
// (the lbx_serve_* entry points of include/lbx.h).
//
// Requests queue on the host.  A request is admitted when the device arena
// (one allocation made at server creation, carved first-fit) has room for its
// two population buffers and fewer than kServeMaxRuns runs are active.  The
// default policy admits in arrival order and stops at the first request that
// does not fit; LBX_SERVE_SHORTEST_FIRST admits the queued requests in order
// of their work (cells x steps, ties by arrival), skipping any that do not
// fit yet.  All
// active runs advance together: one patch-list kernel launch per global step
// walks the patch list of every active run.  At most kInFlight steps are
// queued ahead of the GPU, so admission and completion are observed within a
// few steps.  A run finishes with its stream-only pass; once the GPU has
// executed it, its FP64 sums (reduced on the stream right behind its final
// pass and copied to pinned host memory) and any requested fields are
// handed to the caller and its arena space is released; nothing in the
// steady state waits for the GPU except the in-flight bound.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <deque>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "../../include/lbx.h"
#include "lbx_internal.h"

using lbx::ServeBatch;
using lbx::ServeRun;

namespace {

constexpr int kInFlight = 4;
constexpr size_t kAlign = 256;

double now_s() {
  using namespace std::chrono;
  return duration<double>(steady_clock::now().time_since_epoch()).count();
}

// First-fit arena over one device allocation.
struct Arena {
  char* base = nullptr;
  size_t size = 0;
  std::map<size_t, size_t> free_blocks;  // offset -> length

  void reset(char* b, size_t n) {
    base = b;
    size = n;
    free_blocks.clear();
    if (n) free_blocks[0] = n;
  }
  bool alloc(size_t n, size_t* off) {
    n = (n + kAlign - 1) / kAlign * kAlign;
    for (auto it = free_blocks.begin(); it != free_blocks.end(); ++it)
      if (it->second >= n) {
        *off = it->first;
        const size_t rest = it->second - n, at = it->first + n;
        free_blocks.erase(it);
        if (rest) free_blocks[at] = rest;
        return true;
      }
    return false;
  }
  void release(size_t off, size_t n) {
    n = (n + kAlign - 1) / kAlign * kAlign;
    auto it = free_blocks.emplace(off, n).first;
    auto next = std::next(it);
    if (next != free_blocks.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_blocks.erase(next);
    }
    if (it != free_blocks.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_blocks.erase(it);
      }
    }
  }
  size_t largest() const {
    size_t m = 0;
    for (auto& b : free_blocks) m = std::max(m, b.second);
    return m;
  }
};

struct Request {
  uint64_t id;
  lbx_request req;
  std::vector<double> rho, u;  // copies of caller fields (LBX_INIT_FIELDS)
  double t_submit;
};

struct Run {
  Request r;
  size_t off = 0, bytes = 0;
  lbx::Lattice lat{};    // the run's level-array geometry
  long long start = 0;  // global step of pass 0
  double t_admit = 0.0;
  bool launched_last = false;
  long long last_step = 0;
  int slot = -1;          // per-run result slot (device sums, pinned copy, event)
  bool finish_queued = false;
};

}  // namespace

struct lbx_server {
  lbx_ctx* ctx = nullptr;
  int device = 0;
  hipStream_t stream = nullptr;
  int precision = LBX_FP32_SHIFTED;
  double rho0 = 1.0;
  int max_active = lbx::kServeMaxRuns;
  int policy = LBX_SERVE_FIFO;
  double max_wait_s = 0.0;  // shortest-first: runs queued longer go first (0 = no limit)
  int grid = 2048;
  Arena arena;
  void* arena_mem = nullptr;
  double* staging = nullptr;  // device: hydro and packed populations of a finishing run
  size_t staging_bytes = 0;
  double* sums = nullptr;     // device: conservation partials + per-slot results
  double* pinned = nullptr;   // host pinned: per-slot results
  hipEvent_t slot_ev[lbx::kServeMaxRuns] = {};
  std::vector<int> free_slots;
  std::deque<Request> queue;
  std::vector<Run> active;
  std::vector<lbx_result> done;
  uint64_t next_id = 1;
  long long step = 0;  // next global step to launch
  hipEvent_t ev[kInFlight] = {};
  long long ev_step[kInFlight] = {-1, -1, -1, -1};
  std::string err;
  int64_t lattice_updates = 0;
};

namespace {

int sfail(lbx_server* s, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  s->err = buf;
  return code;
}

#define SRV_HIP(s, call)                                                                     \
  do {                                                                                       \
    hipError_t e_ = (call);                                                                  \
    if (e_ != hipSuccess) return sfail(s, LBX_E_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

size_t esize(int precision) { return precision == LBX_FP64 ? 8 : 4; }

long long cells_of(const lbx_request& q) { return (long long)q.dims[0] * q.dims[1] * q.dims[2]; }

// Admission: move queued requests into the active set while they fit, in
// arrival order (FIFO, stopping at the first that does not fit) or in order
// of work (shortest first, skipping those that do not fit yet).
int admit(lbx_server* s) {
  // admitted entries leave a tombstone (id 0) until the scan is over; clear
  // any left by a scan that stopped on an error
  auto sweep = [s]() {
    s->queue.erase(std::remove_if(s->queue.begin(), s->queue.end(), [](const Request& q) { return q.id == 0; }),
                   s->queue.end());
  };
  sweep();
  std::vector<size_t> order;  // queue positions to try
  if (s->policy == LBX_SERVE_SHORTEST_FIRST) {
    // runs that have waited longer than max_wait_s first, in arrival order;
    // then the rest by work (cells x steps), ties by arrival
    const double now = now_s();
    auto overdue = [&](size_t i) { return s->max_wait_s > 0.0 && now - s->queue[i].t_submit > s->max_wait_s; };
    order.resize(s->queue.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t a, size_t b) {
      const bool oa = overdue(a), ob = overdue(b);
      if (oa != ob) return oa;
      if (oa) return false;  // both overdue: keep arrival order
      return cells_of(s->queue[a].req) * s->queue[a].req.steps < cells_of(s->queue[b].req) * s->queue[b].req.steps;
    });
  }
  for (size_t k = 0; (int)s->active.size() < s->max_active && !s->free_slots.empty(); ++k) {
    const bool fifo = s->policy != LBX_SERVE_SHORTEST_FIRST;
    if (fifo ? s->queue.empty() : k >= order.size()) break;
    Request& q = fifo ? s->queue.front() : s->queue[order[k]];
    const lbx::Lattice L = lbx::make_lattice(q.req.dims[0], q.req.dims[1], q.req.dims[2]);
    const size_t bytes = (size_t)2 * L.elems() * esize(s->precision);
    size_t off = 0;
    if (!s->arena.alloc(bytes, &off)) {
      if (fifo) break;  // head of line waits
      continue;         // a smaller one later in the order may fit
    }
    Run r;
    r.slot = s->free_slots.back();
    s->free_slots.pop_back();
    r.r = std::move(q);
    if (fifo)
      s->queue.pop_front();
    else
      q.id = 0;
    r.off = off;
    r.bytes = bytes;
    r.lat = L;
    r.start = s->step;
    r.t_admit = now_s();
    char* b0 = s->arena.base + off;
    const lbx_request& rq = r.r.req;
    if (rq.init_kind == LBX_INIT_PLANE_PULSE) {
      SRV_HIP(s, lbx::launch_serve_init_pulse(s->precision, b0, L, rq.axis, rq.plane, rq.amplitude,
                                              s->rho0, s->stream));
    } else {
      const long long n = cells_of(rq);
      lbx::Region R{{0, 0, 0}, {rq.dims[0], rq.dims[1], rq.dims[2]}};
      // staging holds rho then u for the upload
      SRV_HIP(s, hipMemcpyAsync(s->staging, r.r.rho.data(), n * sizeof(double), hipMemcpyHostToDevice,
                                s->stream));
      SRV_HIP(s, hipMemcpyAsync(s->staging + n, r.r.u.data(), 3 * n * sizeof(double),
                                hipMemcpyHostToDevice, s->stream));
      SRV_HIP(s, lbx::launch_equilibrium(s->precision, b0, L, R, s->staging, s->staging + n, s->rho0,
                                         s->stream));
      SRV_HIP(s, hipStreamSynchronize(s->stream));  // staging is reused
      r.r.rho.clear();
      r.r.u.clear();
    }
    s->active.push_back(std::move(r));
  }
  sweep();
  return LBX_OK;
}

// Enqueue a finished run's FP64 sums right behind its final pass: reduce
// into the run's slot, copy to pinned host memory, record the slot event.
int queue_finish(lbx_server* s, Run& r) {
  const lbx_request& rq = r.r.req;
  const lbx::Lattice& L = r.lat;
  char* fin = s->arena.base + r.off + (size_t)((rq.steps + 1) & 1) * L.elems() * esize(s->precision);
  double* out = s->sums + 1024 + 4 * r.slot;
  SRV_HIP(s, lbx::launch_conservation(s->precision, false, fin, L, s->rho0, s->sums, 256, out, s->stream));
  SRV_HIP(s, hipMemcpyAsync(s->pinned + 4 * r.slot, out, 4 * sizeof(double), hipMemcpyDeviceToHost,
                            s->stream));
  SRV_HIP(s, hipEventRecord(s->slot_ev[r.slot], s->stream));
  r.finish_queued = true;
  return LBX_OK;
}

// Deliver runs whose sums have reached the host; optional field outputs are
// copied synchronously (the run's arena space is still intact), then the
// arena space and the slot are released.
int collect(lbx_server* s, bool wait) {
  for (size_t i = 0; i < s->active.size();) {
    Run& r = s->active[i];
    if (!r.finish_queued) {
      ++i;
      continue;
    }
    hipError_t q = wait ? hipEventSynchronize(s->slot_ev[r.slot]) : hipEventQuery(s->slot_ev[r.slot]);
    if (q == hipErrorNotReady) {
      ++i;
      continue;
    }
    if (q != hipSuccess) return sfail(s, LBX_E_HIP, "run completion: %s", hipGetErrorString(q));
    const lbx_request& rq = r.r.req;
    const long long n = cells_of(rq);
    const lbx::Lattice& L = r.lat;
    lbx::Region R{{0, 0, 0}, {rq.dims[0], rq.dims[1], rq.dims[2]}};
    char* fin = s->arena.base + r.off + (size_t)((rq.steps + 1) & 1) * L.elems() * esize(s->precision);
    lbx_result res{};
    res.id = r.r.id;
    res.status = LBX_OK;
    res.steps = rq.steps;
    res.cells = n;
    res.mass = s->pinned[4 * r.slot];
    for (int a = 0; a < 3; ++a) res.momentum[a] = s->pinned[4 * r.slot + 1 + a];
    res.t_submit = r.r.t_submit;
    res.t_admit = r.t_admit;
    res.t_done = now_s();
    if (rq.out_rho || rq.out_u) {
      SRV_HIP(s, lbx::launch_hydro(s->precision, false, fin, L, R, s->staging, s->staging + n, s->rho0, s->stream));
      if (rq.out_rho)
        SRV_HIP(s, hipMemcpyAsync(rq.out_rho, s->staging, n * sizeof(double), hipMemcpyDeviceToHost, s->stream));
      if (rq.out_u)
        SRV_HIP(s, hipMemcpyAsync(rq.out_u, s->staging + n, 3 * n * sizeof(double), hipMemcpyDeviceToHost,
                                  s->stream));
      SRV_HIP(s, hipStreamSynchronize(s->stream));
    }
    if (rq.out_f) {
      SRV_HIP(s, lbx::launch_pack_f(s->precision, false, fin, L, R, s->staging, s->rho0, s->stream));
      SRV_HIP(s, hipMemcpyAsync(rq.out_f, s->staging, lbx::NV_TOTAL * n * sizeof(double),
                                hipMemcpyDeviceToHost, s->stream));
      SRV_HIP(s, hipStreamSynchronize(s->stream));
    }
    s->lattice_updates += n * (int64_t)rq.steps;
    s->arena.release(r.off, r.bytes);
    s->free_slots.push_back(r.slot);
    s->done.push_back(res);
    s->active.erase(s->active.begin() + i);
  }
  return LBX_OK;
}

int launch_step(lbx_server* s) {
  ServeBatch b{};
  int items = 0;
  for (Run& r : s->active) {
    if (r.launched_last) continue;
    ServeRun& d = b.run[b.nruns++];
    d.buf[0] = s->arena.base + r.off;
    d.buf[1] = s->arena.base + r.off + (size_t)r.lat.elems() * esize(s->precision);
    for (int k = 0; k < 3; ++k) d.n[k] = r.r.req.dims[k];
    d.item_begin = items;
    d.steps = r.r.req.steps;
    d.start = r.start;
    d.ps = r.lat.ps;
    d.omega_s = 1.0 / (r.r.req.tau_s + 0.5);
    d.omega_b = 1.0 / (r.r.req.tau_b + 0.5);
    d.rho0 = s->rho0;
    items += (int)((r.lat.nx * (long long)r.lat.ny * r.lat.nz + 255) / 256);
    if (s->step - r.start == r.r.req.steps) {
      r.launched_last = true;
      r.last_step = s->step;
    }
  }
  b.total_items = items;
  if (b.nruns == 0) return LBX_OK;
  // reuse the event slot of step - kInFlight once the GPU has passed it
  const int k = int(s->step % kInFlight);
  if (s->ev_step[k] >= 0) SRV_HIP(s, hipEventSynchronize(s->ev[k]));
  SRV_HIP(s, lbx::launch_serve_step(s->precision, b, s->step, s->grid, s->stream));
  SRV_HIP(s, hipEventRecord(s->ev[k], s->stream));
  s->ev_step[k] = s->step;
  for (Run& r : s->active)
    if (r.launched_last && !r.finish_queued && r.last_step == s->step)
      if (int rc = queue_finish(s, r)) return rc;
  ++s->step;
  return LBX_OK;
}

}  // namespace

extern "C" {

int lbx_serve_create(lbx_ctx* ctx, int precision, double rho0, size_t arena_bytes, int max_active,
                     lbx_server** out) {
  if (!ctx || !out) return LBX_E_ARG;
  *out = nullptr;
  if (precision != LBX_FP64 && precision != LBX_FP32_SHIFTED) return LBX_E_ARG;
  if (max_active < 1) max_active = lbx::kServeMaxRuns;
  lbx_server* s = new (std::nothrow) lbx_server;
  if (!s) return LBX_E_NOMEM;
  s->ctx = ctx;
  s->precision = precision;
  s->rho0 = rho0;
  s->max_active = std::min(max_active, lbx::kServeMaxRuns);
  s->device = lbx::ctx_device(ctx);
  hipError_t e = hipSetDevice(s->device);
  int cus = 256;
  if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s->device);
  // One workgroup per work item (the hardware refills CUs as items finish;
  // a persistent grid of 4 per CU walking the list measured 16 % slower),
  // bounded for very long lists.
  s->grid = std::max(1, cus) * 4096;
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e == hipSuccess && arena_bytes == 0) {
    size_t fr = 0, tot = 0;
    e = hipMemGetInfo(&fr, &tot);
    arena_bytes = fr / 2;
  }
  // staging holds the largest run's packed FP64 populations (128^3 x 15 doubles
  // fits 256 MB); larger runs are refused at submit
  s->staging_bytes = (size_t)256 << 20;
  if (e == hipSuccess) e = hipMalloc(&s->arena_mem, arena_bytes);
  if (e == hipSuccess) e = hipMalloc(&s->staging, s->staging_bytes);
  if (e == hipSuccess) e = hipMalloc(&s->sums, (1024 + 4 * lbx::kServeMaxRuns) * sizeof(double));
  if (e == hipSuccess) e = hipHostMalloc(&s->pinned, 4 * lbx::kServeMaxRuns * sizeof(double));
  for (int k = 0; k < lbx::kServeMaxRuns && e == hipSuccess; ++k)
    e = hipEventCreateWithFlags(&s->slot_ev[k], hipEventDisableTiming);
  for (int k = lbx::kServeMaxRuns - 1; k >= 0; --k) s->free_slots.push_back(k);
  for (int k = 0; k < kInFlight && e == hipSuccess; ++k) e = hipEventCreateWithFlags(&s->ev[k], hipEventDisableTiming);
  if (e != hipSuccess) {
    lbx_serve_destroy(s);
    return LBX_E_NOMEM;
  }
  s->arena.reset((char*)s->arena_mem, arena_bytes);
  *out = s;
  return LBX_OK;
}

int lbx_serve_set_policy(lbx_server* s, int policy, double max_wait_s) {
  if (!s) return LBX_E_ARG;
  if (policy != LBX_SERVE_FIFO && policy != LBX_SERVE_SHORTEST_FIRST)
    return sfail(s, LBX_E_ARG, "unknown admission policy %d", policy);
  if (!(max_wait_s >= 0.0) || !std::isfinite(max_wait_s))
    return sfail(s, LBX_E_ARG, "max_wait_s must be finite and >= 0");
  s->policy = policy;
  s->max_wait_s = max_wait_s;
  return LBX_OK;
}

int lbx_serve_destroy(lbx_server* s) {
  if (!s) return LBX_OK;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  for (int k = 0; k < kInFlight; ++k)
    if (s->ev[k]) (void)hipEventDestroy(s->ev[k]);
  if (s->arena_mem) (void)hipFree(s->arena_mem);
  if (s->staging) (void)hipFree(s->staging);
  if (s->sums) (void)hipFree(s->sums);
  if (s->pinned) (void)hipHostFree(s->pinned);
  for (int k = 0; k < lbx::kServeMaxRuns; ++k)
    if (s->slot_ev[k]) (void)hipEventDestroy(s->slot_ev[k]);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
  return LBX_OK;
}

const char* lbx_serve_last_error(const lbx_server* s) { return s ? s->err.c_str() : ""; }

int lbx_serve_submit(lbx_server* s, const lbx_request* req, uint64_t* id) {
  if (!s || !req) return LBX_E_ARG;
  const long long max_cells = (long long)(s->staging_bytes / (lbx::NV_TOTAL * sizeof(double)));
  long long n = 1;
  for (int d = 0; d < 3; ++d) {
    if (req->dims[d] < 1) return sfail(s, LBX_E_ARG, "dims must be positive");
    n *= req->dims[d];  // <= max_cells after every factor, so no overflow
    if (n > max_cells) return sfail(s, LBX_E_ARG, "run exceeds the serving staging size (%lld cells)", max_cells);
  }
  if (req->steps < 1) return sfail(s, LBX_E_ARG, "steps must be >= 1");
  if (!(req->tau_s > -0.5) || !(req->tau_b > -0.5) || !std::isfinite(req->tau_s) || !std::isfinite(req->tau_b))
    return sfail(s, LBX_E_ARG, "tau must be finite and > -1/2");
  const lbx::Lattice L = lbx::make_lattice(req->dims[0], req->dims[1], req->dims[2]);
  if ((size_t)2 * L.elems() * esize(s->precision) > s->arena.size)
    return sfail(s, LBX_E_NOMEM, "run does not fit the serving arena");
  if (req->init_kind != LBX_INIT_PLANE_PULSE && req->init_kind != LBX_INIT_FIELDS)
    return sfail(s, LBX_E_ARG, "unknown init kind %d", req->init_kind);
  if (req->init_kind == LBX_INIT_FIELDS && (!req->rho || !req->u))
    return sfail(s, LBX_E_ARG, "LBX_INIT_FIELDS needs rho and u");
  if (req->init_kind == LBX_INIT_PLANE_PULSE && (req->axis < 0 || req->axis > 2))
    return sfail(s, LBX_E_ARG, "pulse axis must be 0, 1 or 2");
  if (req->init_kind == LBX_INIT_PLANE_PULSE && (req->plane < 0 || req->plane >= req->dims[req->axis]))
    return sfail(s, LBX_E_ARG, "pulse plane %d outside the run", req->plane);
  Request q;
  q.id = s->next_id++;
  q.req = *req;
  if (req->init_kind == LBX_INIT_FIELDS) {
    q.rho.assign(req->rho, req->rho + n);
    q.u.assign(req->u, req->u + 3 * n);
  }
  q.req.rho = nullptr;
  q.req.u = nullptr;
  q.t_submit = now_s();
  if (id) *id = q.id;
  s->queue.push_back(std::move(q));
  return LBX_OK;
}

int lbx_serve_progress(lbx_server* s, int max_steps, int* active, int* queued) {
  if (!s) return LBX_E_ARG;
  SRV_HIP(s, hipSetDevice(s->device));
  for (int it = 0; it < std::max(max_steps, 1); ++it) {
    if (int rc = collect(s, false)) return rc;
    if (int rc = admit(s)) return rc;
    bool pending = false;
    for (const Run& r : s->active) pending |= !r.launched_last;
    if (!pending) break;
    if (int rc = launch_step(s)) return rc;
  }
  if (int rc = collect(s, false)) return rc;
  if (active) *active = (int)s->active.size();
  if (queued) *queued = (int)s->queue.size();
  return LBX_OK;
}

int lbx_serve_drain(lbx_server* s) {
  if (!s) return LBX_E_ARG;
  while (!s->queue.empty() || !s->active.empty()) {
    int a = 0, q = 0;
    if (int rc = lbx_serve_progress(s, 16, &a, &q)) return rc;
    bool pending = false;
    for (const Run& r : s->active) pending |= !r.launched_last;
    if (!pending && !s->active.empty())
      if (int rc = collect(s, true)) return rc;
  }
  return LBX_OK;
}

int lbx_serve_poll(lbx_server* s, lbx_result* out, int max, int* n) {
  if (!s || !n || (max > 0 && !out)) return LBX_E_ARG;
  const int k = std::min<int>(max, (int)s->done.size());
  for (int i = 0; i < k; ++i) out[i] = s->done[i];
  s->done.erase(s->done.begin(), s->done.begin() + k);
  *n = k;
  return LBX_OK;
}

int lbx_serve_stats(lbx_server* s, int64_t* lattice_updates, int64_t* steps_launched,
                    size_t* arena_free_largest) {
  if (!s) return LBX_E_ARG;
  if (lattice_updates) *lattice_updates = s->lattice_updates;
  if (steps_launched) *steps_launched = s->step;
  if (arena_free_largest) *arena_free_largest = s->arena.largest();
  return LBX_OK;
}

}  // extern "C"
