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
//
// A request with fine boxes is a refined run: a two-level hierarchy (ratio
// 2, the fine-grid mask) held in a simulation of its own, made through the
// C ABI on the server's context, whose level-0 steps (one captured Rohde
// cycle each) go to the context's stream while the patch-list launches of
// the single-level runs go to the server's stream, so the two overlap on the
// device.  It takes one coarse step per global step, is admitted against
// the device's free memory (its levels are allocated when it is admitted
// and released when it is delivered), and its result is what lbx_iterate
// computes for the same hierarchy alone, bit for bit.
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
// device memory left free when a refined run is admitted
constexpr size_t kRefinedReserve = (size_t)512 << 20;

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
  std::vector<int32_t> boxes, fine;  // copies of the level-0 / level-1 boxes
  double t_submit;
};

bool refined(const lbx_request& q) { return q.n_fine_boxes > 0; }

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
  lbx_sim* sim = nullptr;  // refined run: its own hierarchy
  int steps_done = 0;      // refined run: level-0 steps enqueued
  size_t device_bytes = 0;  // refined run: device memory it was admitted against
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
  hipEvent_t rev[kInFlight] = {};  // refined runs' steps on the context's stream
  long long rev_step[kInFlight] = {-1, -1, -1, -1};
  std::string err;
  int64_t lattice_updates = 0;
  size_t refined_bytes = 0;  // device memory held by active refined runs
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

// Device memory a refined run's simulation allocates: two level-0 arrays
// and their flag bytes, four buffers of grown fine patches (2 ghost layers,
// population stride rounded to 64) with their fill table, mask bytes and
// block averages, and the parent's touched-cell slots; 1/8 more for the
// smaller tables.
size_t refined_bytes_estimate(const Request& q, int precision) {
  const lbx::Lattice L = lbx::make_lattice(q.req.dims[0], q.req.dims[1], q.req.dims[2], (int)esize(precision));
  const long long coarse_cells = (long long)L.nx * L.ny * L.nz;
  size_t bytes = 2 * (size_t)L.elems() * esize(precision) + (size_t)coarse_cells * 5;
  for (size_t b = 0; b + 5 < q.fine.size(); b += 6) {
    long long grown = 1, blocks = 1;
    for (int d = 0; d < 3; ++d) {
      const long long g = q.fine[b + 3 + d] - q.fine[b + d] + 1 + 4;
      grown *= g;
      blocks *= g / 2;
    }
    const long long qs = (grown + 63) / 64 * 64;
    bytes += (size_t)(4 * qs * lbx::NV_TOTAL) * esize(precision) + (size_t)grown * 9 +
             (size_t)blocks * lbx::NV_TOTAL * sizeof(double);
  }
  return bytes + bytes / 8;
}

// Level-0 fields of box [lo, hi] of a refined run, x fastest, from the
// request: the plane pulse (rho = 1 + amplitude on the plane, u = 0) or the
// caller's whole-domain fields.
void box_fields(const Request& q, const int32_t* lohi, std::vector<double>& rho, std::vector<double>& u) {
  const lbx_request& rq = q.req;
  const int n0 = lohi[3] - lohi[0] + 1, n1 = lohi[4] - lohi[1] + 1, n2 = lohi[5] - lohi[2] + 1;
  const long long n = (long long)n0 * n1 * n2, dom = (long long)rq.dims[0] * rq.dims[1] * rq.dims[2];
  rho.assign((size_t)n, 1.0);
  u.assign((size_t)(3 * n), 0.0);
  for (int z = 0; z < n2; ++z)
    for (int y = 0; y < n1; ++y)
      for (int x = 0; x < n0; ++x) {
        const int p[3] = {lohi[0] + x, lohi[1] + y, lohi[2] + z};
        const long long i = ((long long)z * n1 + y) * n0 + x;
        const long long g = ((long long)p[2] * rq.dims[1] + p[1]) * rq.dims[0] + p[0];
        if (rq.init_kind == LBX_INIT_PLANE_PULSE) {
          if (p[rq.axis] == rq.plane) rho[i] = 1.0 + rq.amplitude;
        } else {
          rho[i] = q.rho[g];
          for (int a = 0; a < 3; ++a) u[a * n + i] = q.u[a * dom + g];
        }
      }
}

// The refined run's simulation on the server's context: level 0 from its
// boxes (one box over the domain if none were given) at equilibrium, level 1
// at the request's ratio filled from it by piecewise-constant interpolation
// with tau rescaled as ComputeDt does (src/AmrSim.cpp:297-322), and level 0
// masked under level 1.
int start_refined(lbx_server* s, Run& r) {
  const Request& q = r.r;
  const lbx_request& rq = q.req;
  const int32_t per[3] = {1, 1, 1};
  auto abi = [s](int rc) {
    if (rc != LBX_OK) sfail(s, rc, "refined run: %s", lbx_last_error(s->ctx));
    return rc;
  };
  if (int rc = abi(lbx_sim_create(s->ctx, rq.dims, per, s->precision, s->rho0, &r.sim))) return rc;
  std::vector<int32_t> boxes = q.boxes;
  if (boxes.empty()) boxes = {0, 0, 0, rq.dims[0] - 1, rq.dims[1] - 1, rq.dims[2] - 1};
  const int nb = (int)boxes.size() / 6, nf = (int)q.fine.size() / 6;
  if (int rc = abi(lbx_level_define(r.sim, 0, nb, boxes.data()))) return rc;
  if (int rc = abi(lbx_set_relax(r.sim, 0, 1.0 / (rq.tau_s + 0.5), 1.0 / (rq.tau_b + 0.5)))) return rc;
  std::vector<double> rho, u;
  for (int b = 0; b < nb; ++b) {
    box_fields(q, &boxes[6 * b], rho, u);
    if (int rc = abi(lbx_init_equilibrium(r.sim, 0, b, rho.data(), u.data()))) return rc;
  }
  const int ratio = rq.ref_ratio ? rq.ref_ratio : 2;
  if (ratio != 2)
    if (int rc = abi(lbx_set_ref_ratio(r.sim, ratio))) return rc;
  const double tf_s = ratio * (rq.tau_s - 0.5) + 0.5, tf_b = ratio * (rq.tau_b - 0.5) + 0.5;
  if (int rc = abi(lbx_level_define(r.sim, 1, nf, q.fine.data()))) return rc;
  if (int rc = abi(lbx_set_relax(r.sim, 1, 1.0 / (tf_s + 0.5), 1.0 / (tf_b + 0.5)))) return rc;
  if (int rc = abi(lbx_fill_from_coarse(r.sim, 1))) return rc;
  return abi(lbx_set_fine_mask(r.sim, nf, q.fine.data()));
}

// A refined run that could not be set up is delivered at once with the
// error status; nothing of it stays on the device.
void reject_refined(lbx_server* s, Run& r, int rc) {
  if (r.sim) lbx_sim_destroy(r.sim);
  r.sim = nullptr;
  lbx_result res{};
  res.id = r.r.id;
  res.status = rc;
  res.t_submit = r.r.t_submit;
  res.t_admit = res.t_done = now_s();
  s->done.push_back(res);
  s->free_slots.push_back(r.slot);
}

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
    if (refined(q.req)) {
      const size_t need = refined_bytes_estimate(q, s->precision);
      size_t fr = 0, tot = 0;
      SRV_HIP(s, hipMemGetInfo(&fr, &tot));
      if (need + kRefinedReserve > fr) {
        if (fifo) break;
        continue;
      }
      Run r;
      r.slot = s->free_slots.back();
      s->free_slots.pop_back();
      r.r = std::move(q);
      if (fifo)
        s->queue.pop_front();
      else
        q.id = 0;
      r.t_admit = now_s();
      if (int rc = start_refined(s, r)) {
        reject_refined(s, r, rc);
        continue;
      }
      r.device_bytes = need;
      s->refined_bytes += need;
      s->active.push_back(std::move(r));
      continue;
    }
    const lbx::Lattice L = lbx::make_lattice(q.req.dims[0], q.req.dims[1], q.req.dims[2], (int)esize(s->precision));
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

// A finished refined run: level-0 sums and fields (optional outputs are
// level 0's, over the whole domain like a single-level run's), then its
// simulation is released.
int deliver_refined(lbx_server* s, Run& r) {
  const lbx_request& rq = r.r.req;
  auto abi = [s](int rc) {
    if (rc != LBX_OK) sfail(s, rc, "refined run: %s", lbx_last_error(s->ctx));
    return rc;
  };
  lbx_result res{};
  res.id = r.r.id;
  res.status = LBX_OK;
  res.steps = rq.steps;
  const long long n = cells_of(rq);
  res.cells = n;
  double sums[4];
  if (int rc = abi(lbx_conservation(r.sim, 0, sums))) return rc;
  res.mass = sums[0];
  for (int a = 0; a < 3; ++a) res.momentum[a] = sums[1 + a];
  std::vector<int32_t> boxes = r.r.boxes;
  if (boxes.empty()) boxes = {0, 0, 0, rq.dims[0] - 1, rq.dims[1] - 1, rq.dims[2] - 1};
  const int nb = (int)boxes.size() / 6;
  // box b's cell (x, y, z) of its [n2][n1][n0] block -> whole-domain index
  auto scatter = [&](const int32_t* lohi, const double* src, double* dst, int ncomp, long long box_n) {
    const int n0 = lohi[3] - lohi[0] + 1, n1 = lohi[4] - lohi[1] + 1, n2 = lohi[5] - lohi[2] + 1;
    for (int c = 0; c < ncomp; ++c)
      for (int z = 0; z < n2; ++z)
        for (int y = 0; y < n1; ++y)
          for (int x = 0; x < n0; ++x)
            dst[c * n + ((long long)(lohi[2] + z) * rq.dims[1] + lohi[1] + y) * rq.dims[0] + lohi[0] + x] =
                src[c * box_n + ((long long)z * n1 + y) * n0 + x];
  };
  if (rq.out_rho || rq.out_u) {
    std::vector<double> rho((size_t)n), u((size_t)(3 * n));
    if (int rc = abi(lbx_hydro_level(r.sim, 0, rho.data(), u.data()))) return rc;
    long long off = 0;
    for (int b = 0; b < nb; ++b) {
      const int32_t* lohi = &boxes[6 * b];
      const long long bn = (long long)(lohi[3] - lohi[0] + 1) * (lohi[4] - lohi[1] + 1) * (lohi[5] - lohi[2] + 1);
      if (rq.out_rho) scatter(lohi, rho.data() + off, rq.out_rho, 1, bn);
      if (rq.out_u) scatter(lohi, u.data() + 3 * off, rq.out_u, 3, bn);
      off += bn;
    }
  }
  if (rq.out_f) {
    std::vector<double> f;
    for (int b = 0; b < nb; ++b) {
      const int32_t* lohi = &boxes[6 * b];
      const long long bn = (long long)(lohi[3] - lohi[0] + 1) * (lohi[4] - lohi[1] + 1) * (lohi[5] - lohi[2] + 1);
      f.resize((size_t)(lbx::NV_TOTAL * bn));
      if (int rc = abi(lbx_download_f(r.sim, 0, b, f.data()))) return rc;
      scatter(lohi, f.data(), rq.out_f, lbx::NV_TOTAL, bn);
    }
  }
  long long fine_cells = 0;
  for (size_t b = 0; b + 5 < r.r.fine.size(); b += 6)
    fine_cells += (long long)(r.r.fine[b + 3] - r.r.fine[b] + 1) * (r.r.fine[b + 4] - r.r.fine[b + 1] + 1) *
                  (r.r.fine[b + 5] - r.r.fine[b + 2] + 1);
  const int ratio = rq.ref_ratio ? rq.ref_ratio : 2;  // fine substeps per coarse step
  s->lattice_updates += (n + ratio * fine_cells) * (int64_t)rq.steps;
  lbx_sim_destroy(r.sim);
  r.sim = nullptr;
  s->refined_bytes -= r.device_bytes;
  res.t_submit = r.r.t_submit;
  res.t_admit = r.t_admit;
  res.t_done = now_s();
  s->done.push_back(res);
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
    if (r.sim) {
      if (int rc = deliver_refined(s, r)) return rc;
      s->free_slots.push_back(r.slot);
      s->active.erase(s->active.begin() + i);
      continue;
    }
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
  // refined runs: one level-0 step each on the context's stream, at most
  // kInFlight steps ahead of the GPU
  bool stepped = false;
  for (Run& r : s->active) {
    if (!r.sim || r.launched_last) continue;
    if (!stepped) {
      const int k = int(s->step % kInFlight);
      if (s->rev_step[k] >= 0) SRV_HIP(s, hipEventSynchronize(s->rev[k]));
    }
    if (int rc = lbx_iterate(r.sim, 1)) return sfail(s, rc, "refined run: %s", lbx_last_error(s->ctx));
    stepped = true;
    if (++r.steps_done == r.r.req.steps) {
      r.launched_last = true;
      r.last_step = s->step;
      SRV_HIP(s, hipEventRecord(s->slot_ev[r.slot], lbx::ctx_stream(s->ctx)));
      r.finish_queued = true;
    }
  }
  if (stepped) {
    const int k = int(s->step % kInFlight);
    SRV_HIP(s, hipEventRecord(s->rev[k], lbx::ctx_stream(s->ctx)));
    s->rev_step[k] = s->step;
  }
  ServeBatch b{};
  int items = 0;
  for (Run& r : s->active) {
    if (r.sim || r.launched_last) continue;
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
  if (b.nruns == 0) {
    if (stepped) ++s->step;
    return LBX_OK;
  }
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
  for (int k = 0; k < kInFlight && e == hipSuccess; ++k) e = hipEventCreateWithFlags(&s->rev[k], hipEventDisableTiming);
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
  for (Run& r : s->active)
    if (r.sim) lbx_sim_destroy(r.sim);
  for (int k = 0; k < kInFlight; ++k) {
    if (s->ev[k]) (void)hipEventDestroy(s->ev[k]);
    if (s->rev[k]) (void)hipEventDestroy(s->rev[k]);
  }
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
  const bool ref = req->n_fine_boxes > 0;
  // single-level runs pass their fields through the staging buffer
  const long long max_cells =
      ref ? (1ll << 31) : (long long)(s->staging_bytes / (lbx::NV_TOTAL * sizeof(double)));
  long long n = 1;
  for (int d = 0; d < 3; ++d) {
    if (req->dims[d] < 1) return sfail(s, LBX_E_ARG, "dims must be positive");
    n *= req->dims[d];  // <= max_cells after every factor, so no overflow
    if (n > max_cells) return sfail(s, LBX_E_ARG, "run exceeds the serving staging size (%lld cells)", max_cells);
  }
  if (req->n_fine_boxes < 0 || req->n_boxes < 0) return sfail(s, LBX_E_ARG, "negative box count");
  if ((ref && !req->fine_lohi) || (req->n_boxes > 0 && !req->box_lohi))
    return sfail(s, LBX_E_ARG, "box count without a box list");
  if (req->n_boxes > 0 && !ref)
    return sfail(s, LBX_E_ARG, "level-0 boxes are taken for refined runs only (single-level runs are one lattice)");
  if (req->steps < 1) return sfail(s, LBX_E_ARG, "steps must be >= 1");
  if (req->ref_ratio != 0 && (req->ref_ratio < 2 || req->ref_ratio > 4))
    return sfail(s, LBX_E_ARG, "refinement ratio %d is not 2, 3 or 4", req->ref_ratio);
  if (!(req->tau_s > -0.5) || !(req->tau_b > -0.5) || !std::isfinite(req->tau_s) || !std::isfinite(req->tau_b))
    return sfail(s, LBX_E_ARG, "tau must be finite and > -1/2");
  const lbx::Lattice L = lbx::make_lattice(req->dims[0], req->dims[1], req->dims[2], (int)esize(s->precision));
  if (!ref && (size_t)2 * L.elems() * esize(s->precision) > s->arena.size)
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
  if (req->n_boxes > 0) q.boxes.assign(req->box_lohi, req->box_lohi + 6 * (size_t)req->n_boxes);
  if (ref) q.fine.assign(req->fine_lohi, req->fine_lohi + 6 * (size_t)req->n_fine_boxes);
  q.req.box_lohi = nullptr;
  q.req.fine_lohi = nullptr;
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
