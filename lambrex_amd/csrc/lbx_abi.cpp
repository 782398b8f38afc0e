// lbx_abi.cpp -- the C ABI declared in include/lbx.h.
//
// Owns device memory and the per-context HIP stream; never aborts: every
// failure becomes a negative status plus a message in lbx_last_error().
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/lbx.h"
#include "lbx_internal.h"

using lbx::Lattice;
using lbx::LaunchShape;
using lbx::Region;
using lbx::Relax;

namespace {

thread_local std::string g_global_error;

struct Box {
  int lo[3], hi[3];
  long long cells() const {
    return (long long)(hi[0] - lo[0] + 1) * (hi[1] - lo[1] + 1) * (hi[2] - lo[2] + 1);
  }
};

struct Level {
  bool defined = false;
  std::vector<Box> boxes;
  Lattice lat{};
  void* buf[2] = {nullptr, nullptr};  // state lives in buf[cur]
  int cur = 0;
  double omega_s = 0.0, omega_b = 0.0;
  bool relax_set = false;
};

constexpr int kPartials = 1024;  // blocks of the conservation reduction

}  // namespace

struct lbx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  double* scratch = nullptr;  // kPartials*4 partials + 4 results
  // profiling of stream-collide launches
  bool prof = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  double prof_ms = 0.0;
  int64_t prof_launches = 0;
};

struct lbx_sim {
  lbx_ctx* ctx = nullptr;
  int dims[3] = {0, 0, 0};
  int precision = LBX_FP64;
  double rho0 = 1.0;
  Level lev[2];
};

namespace {

int fail(lbx_ctx* ctx, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (ctx)
    ctx->err = buf;
  else
    g_global_error = buf;
  return code;
}

int hip_fail(lbx_ctx* ctx, hipError_t e, const char* what) {
  return fail(ctx, LBX_E_HIP, "%s: %s", what, hipGetErrorString(e));
}

#define LBX_HIP(ctx, call)                                \
  do {                                                    \
    hipError_t e_ = (call);                               \
    if (e_ != hipSuccess) return hip_fail(ctx, e_, #call); \
  } while (0)

size_t elem_size(int precision) { return precision == LBX_FP64 ? 8 : 4; }

bool box_ok(const Box& b, const int* dom) {
  for (int d = 0; d < 3; ++d)
    if (b.lo[d] < 0 || b.hi[d] >= dom[d] || b.hi[d] < b.lo[d]) return false;
  return true;
}

Region region_of(const Box& b) {
  Region r;
  for (int d = 0; d < 3; ++d) {
    r.lo[d] = b.lo[d];
    r.n[d] = b.hi[d] - b.lo[d] + 1;
  }
  return r;
}

// Boxes of level 0 must tile [0,dims) exactly: disjoint and covering.
bool tiles_domain(const std::vector<Box>& boxes, const int* dims) {
  long long total = 0;
  for (const Box& b : boxes) total += b.cells();
  if (total != (long long)dims[0] * dims[1] * dims[2]) return false;
  for (size_t a = 0; a < boxes.size(); ++a)
    for (size_t c = a + 1; c < boxes.size(); ++c) {
      bool overlap = true;
      for (int d = 0; d < 3; ++d)
        if (boxes[a].hi[d] < boxes[c].lo[d] || boxes[c].hi[d] < boxes[a].lo[d]) overlap = false;
      if (overlap) return false;
    }
  return true;
}

void free_level(Level& L) {
  for (void*& p : L.buf)
    if (p) {
      (void)hipFree(p);
      p = nullptr;
    }
  L = Level{};
}

int check_sim(lbx_sim* sim) {
  if (!sim || !sim->ctx) return fail(nullptr, LBX_E_ARG, "null simulation handle");
  return LBX_OK;
}

int check_box(lbx_sim* sim, int level, int box) {
  if (level < 0 || level > 1) return fail(sim->ctx, LBX_E_ARG, "level %d out of range", level);
  Level& L = sim->lev[level];
  if (!L.defined) return fail(sim->ctx, LBX_E_STATE, "level %d is not defined", level);
  if (box < 0 || box >= (int)L.boxes.size())
    return fail(sim->ctx, LBX_E_ARG, "box %d out of range (level %d has %zu)", box, level,
                L.boxes.size());
  return LBX_OK;
}

// Scratch device staging buffer for one box transfer (FP64 host layout).
struct DeviceTemp {
  void* p = nullptr;
  ~DeviceTemp() {
    if (p) (void)hipFree(p);
  }
};

hipEvent_t take_event(lbx_ctx* ctx) {
  if (ctx->ev_used == ctx->ev_pool.size()) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    ctx->ev_pool.push_back(e);
  }
  return ctx->ev_pool[ctx->ev_used++];
}

// Fold recorded event pairs into the running total.
int drain_profile(lbx_ctx* ctx) {
  if (ctx->ev_used == 0) return LBX_OK;
  LBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  for (size_t i = 0; i + 1 < ctx->ev_used; i += 2) {
    float ms = 0.0f;
    LBX_HIP(ctx, hipEventElapsedTime(&ms, ctx->ev_pool[i], ctx->ev_pool[i + 1]));
    ctx->prof_ms += ms;
    ctx->prof_launches += 1;
  }
  ctx->ev_used = 0;
  return LBX_OK;
}

int launch_step(lbx_sim* sim, Level& L, lbx::StepKind kind) {
  lbx_ctx* ctx = sim->ctx;
  Relax r{L.omega_s, L.omega_b, sim->rho0};
  const LaunchShape shape = lbx::default_shape(L.lat);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (ctx->prof && kind == lbx::STEP_FUSED) {
    if (ctx->ev_used + 2 > 4096) {
      int rc = drain_profile(ctx);
      if (rc) return rc;
    }
    e0 = take_event(ctx);
    e1 = take_event(ctx);
    if (!e0 || !e1) return fail(ctx, LBX_E_HIP, "hipEventCreate failed");
    LBX_HIP(ctx, hipEventRecord(e0, ctx->stream));
  }
  hipError_t e;
  void* src = L.buf[L.cur];
  void* dst = L.buf[1 - L.cur];
  if (sim->precision == LBX_FP64)
    e = lbx::launch_step_f64(kind, (const double*)src, (double*)dst, L.lat, r, shape, ctx->stream);
  else
    e = lbx::launch_step_f32s(kind, (const float*)src, (float*)dst, L.lat, r, shape, ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "stream-collide launch");
  if (e1) LBX_HIP(ctx, hipEventRecord(e1, ctx->stream));
  L.cur = 1 - L.cur;
  return LBX_OK;
}

}  // namespace

extern "C" {

const char* lbx_last_error(const lbx_ctx* ctx) {
  return ctx ? ctx->err.c_str() : g_global_error.c_str();
}

int lbx_ctx_create(int device, lbx_ctx** out) {
  if (!out) return fail(nullptr, LBX_E_ARG, "out is null");
  *out = nullptr;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e != hipSuccess) return hip_fail(nullptr, e, "hipGetDeviceCount");
  if (device < 0 || device >= n)
    return fail(nullptr, LBX_E_ARG, "device %d not present (%d visible)", device, n);
  e = hipSetDevice(device);
  if (e != hipSuccess) return hip_fail(nullptr, e, "hipSetDevice");
  lbx_ctx* ctx = new (std::nothrow) lbx_ctx;
  if (!ctx) return fail(nullptr, LBX_E_NOMEM, "out of host memory");
  ctx->device = device;
  e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc(&ctx->scratch, (kPartials * 4 + 4) * sizeof(double));
  if (e != hipSuccess) {
    hip_fail(nullptr, e, "context set-up");
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return LBX_E_HIP;
  }
  *out = ctx;
  return LBX_OK;
}

int lbx_ctx_destroy(lbx_ctx* ctx) {
  if (!ctx) return LBX_OK;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  for (hipEvent_t e : ctx->ev_pool) (void)hipEventDestroy(e);
  if (ctx->scratch) (void)hipFree(ctx->scratch);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return LBX_OK;
}

int lbx_sync(lbx_ctx* ctx) {
  if (!ctx) return fail(nullptr, LBX_E_ARG, "null context");
  LBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LBX_OK;
}

int lbx_mem_info(lbx_ctx* ctx, size_t* free_bytes, size_t* total_bytes) {
  if (!ctx || !free_bytes || !total_bytes) return fail(ctx, LBX_E_ARG, "null argument");
  LBX_HIP(ctx, hipSetDevice(ctx->device));
  LBX_HIP(ctx, hipMemGetInfo(free_bytes, total_bytes));
  return LBX_OK;
}

int lbx_sim_create(lbx_ctx* ctx, const int32_t dims[3], const int32_t periodic[3], int precision,
                   double rho0, lbx_sim** out) {
  if (!ctx || !dims || !periodic || !out) return fail(ctx, LBX_E_ARG, "null argument");
  *out = nullptr;
  for (int d = 0; d < 3; ++d) {
    if (dims[d] < 1) return fail(ctx, LBX_E_ARG, "domain extent %d in direction %d", dims[d], d);
    if (!periodic[d])
      return fail(ctx, LBX_E_UNSUPPORTED, "Currently only periodic boundary conditions allowed.");
  }
  if ((long long)dims[0] * dims[1] * dims[2] >= (1ll << 31))
    return fail(ctx, LBX_E_ARG, "domain of %lld cells exceeds 2^31",
                (long long)dims[0] * dims[1] * dims[2]);
  if (precision != LBX_FP64 && precision != LBX_FP32_SHIFTED)
    return fail(ctx, LBX_E_ARG, "unknown precision %d", precision);
  lbx_sim* sim = new (std::nothrow) lbx_sim;
  if (!sim) return fail(ctx, LBX_E_NOMEM, "out of host memory");
  sim->ctx = ctx;
  for (int d = 0; d < 3; ++d) sim->dims[d] = dims[d];
  sim->precision = precision;
  sim->rho0 = rho0;
  *out = sim;
  return LBX_OK;
}

int lbx_sim_destroy(lbx_sim* sim) {
  if (!sim) return LBX_OK;
  (void)hipSetDevice(sim->ctx->device);
  (void)hipStreamSynchronize(sim->ctx->stream);
  for (Level& L : sim->lev) free_level(L);
  delete sim;
  return LBX_OK;
}

int lbx_level_define(lbx_sim* sim, int level, int n_boxes, const int32_t* lohi) {
  if (int rc = check_sim(sim)) return rc;
  lbx_ctx* ctx = sim->ctx;
  if (level != 0)
    return fail(ctx, LBX_E_UNSUPPORTED, "only level 0 is supported by this build");
  if (n_boxes < 1 || !lohi) return fail(ctx, LBX_E_ARG, "need at least one box");
  std::vector<Box> boxes(n_boxes);
  for (int b = 0; b < n_boxes; ++b) {
    for (int d = 0; d < 3; ++d) {
      boxes[b].lo[d] = lohi[6 * b + d];
      boxes[b].hi[d] = lohi[6 * b + 3 + d];
    }
    if (!box_ok(boxes[b], sim->dims))
      return fail(ctx, LBX_E_ARG, "box %d lies outside the domain or is empty", b);
  }
  if (!tiles_domain(boxes, sim->dims))
    return fail(ctx, LBX_E_ARG, "level-0 boxes must tile the periodic domain exactly");
  LBX_HIP(ctx, hipSetDevice(ctx->device));
  LBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  Level& L = sim->lev[level];
  free_level(L);
  L.boxes = boxes;
  L.lat.nx = sim->dims[0];
  L.lat.ny = sim->dims[1];
  L.lat.nz = sim->dims[2];
  const long long n = (long long)L.lat.nx * L.lat.ny * L.lat.nz;
  L.lat.qs = (n + 63) / 64 * 64;
  const size_t bytes = (size_t)L.lat.qs * lbx::NV_TOTAL * elem_size(sim->precision);
  for (void*& p : L.buf) {
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
      free_level(L);
      return fail(ctx, LBX_E_NOMEM, "hipMalloc of %zu bytes for level %d failed: %s", bytes, level,
                  hipGetErrorString(e));
    }
    e = hipMemsetAsync(p, 0, bytes, ctx->stream);
    if (e != hipSuccess) {
      free_level(L);
      return hip_fail(ctx, e, "hipMemsetAsync");
    }
  }
  L.defined = true;
  return LBX_OK;
}

int lbx_level_clear(lbx_sim* sim, int level) {
  if (int rc = check_sim(sim)) return rc;
  if (level < 0 || level > 1) return fail(sim->ctx, LBX_E_ARG, "level %d out of range", level);
  (void)hipStreamSynchronize(sim->ctx->stream);
  free_level(sim->lev[level]);
  return LBX_OK;
}

int lbx_set_relax(lbx_sim* sim, int level, double omega_s, double omega_b) {
  if (int rc = check_sim(sim)) return rc;
  if (level < 0 || level > 1) return fail(sim->ctx, LBX_E_ARG, "level %d out of range", level);
  sim->lev[level].omega_s = omega_s;
  sim->lev[level].omega_b = omega_b;
  sim->lev[level].relax_set = true;
  return LBX_OK;
}

int lbx_init_equilibrium(lbx_sim* sim, int level, int box, const double* rho, const double* u) {
  if (int rc = check_sim(sim)) return rc;
  if (int rc = check_box(sim, level, box)) return rc;
  if (!rho || !u) return fail(sim->ctx, LBX_E_ARG, "null field buffer");
  lbx_ctx* ctx = sim->ctx;
  Level& L = sim->lev[level];
  const Region R = region_of(L.boxes[box]);
  const size_t n = (size_t)L.boxes[box].cells();
  DeviceTemp t;
  LBX_HIP(ctx, hipMalloc(&t.p, n * 4 * sizeof(double)));
  double* drho = (double*)t.p;
  double* du = drho + n;
  LBX_HIP(ctx, hipMemcpyAsync(drho, rho, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  LBX_HIP(ctx, hipMemcpyAsync(du, u, 3 * n * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  LBX_HIP(ctx, lbx::launch_equilibrium(sim->precision, L.buf[L.cur], L.lat, R, drho, du, sim->rho0,
                                       ctx->stream));
  LBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LBX_OK;
}

int lbx_upload_f(lbx_sim* sim, int level, int box, const double* f) {
  if (int rc = check_sim(sim)) return rc;
  if (int rc = check_box(sim, level, box)) return rc;
  if (!f) return fail(sim->ctx, LBX_E_ARG, "null buffer");
  lbx_ctx* ctx = sim->ctx;
  Level& L = sim->lev[level];
  const size_t n = (size_t)L.boxes[box].cells();
  DeviceTemp t;
  LBX_HIP(ctx, hipMalloc(&t.p, n * lbx::NV_TOTAL * sizeof(double)));
  LBX_HIP(ctx, hipMemcpyAsync(t.p, f, n * lbx::NV_TOTAL * sizeof(double), hipMemcpyHostToDevice,
                              ctx->stream));
  LBX_HIP(ctx, lbx::launch_unpack_f(sim->precision, L.buf[L.cur], L.lat, region_of(L.boxes[box]),
                                    (const double*)t.p, sim->rho0, ctx->stream));
  LBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LBX_OK;
}

int lbx_download_f(lbx_sim* sim, int level, int box, double* f) {
  if (int rc = check_sim(sim)) return rc;
  if (int rc = check_box(sim, level, box)) return rc;
  if (!f) return fail(sim->ctx, LBX_E_ARG, "null buffer");
  lbx_ctx* ctx = sim->ctx;
  Level& L = sim->lev[level];
  const size_t n = (size_t)L.boxes[box].cells();
  DeviceTemp t;
  LBX_HIP(ctx, hipMalloc(&t.p, n * lbx::NV_TOTAL * sizeof(double)));
  LBX_HIP(ctx, lbx::launch_pack_f(sim->precision, L.buf[L.cur], L.lat, region_of(L.boxes[box]),
                                  (double*)t.p, sim->rho0, ctx->stream));
  LBX_HIP(ctx, hipMemcpyAsync(f, t.p, n * lbx::NV_TOTAL * sizeof(double), hipMemcpyDeviceToHost,
                              ctx->stream));
  LBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LBX_OK;
}

int lbx_iterate(lbx_sim* sim, int nsteps) {
  if (int rc = check_sim(sim)) return rc;
  lbx_ctx* ctx = sim->ctx;
  if (nsteps < 0) return fail(ctx, LBX_E_ARG, "negative step count %d", nsteps);
  Level& L = sim->lev[0];
  if (!L.defined) return fail(ctx, LBX_E_STATE, "level 0 is not defined");
  if (!L.relax_set) return fail(ctx, LBX_E_STATE, "relaxation rates of level 0 are not set");
  if (sim->lev[1].defined)
    return fail(ctx, LBX_E_UNSUPPORTED, "two-level iteration is not available in this build");
  if (nsteps == 0) return LBX_OK;
  LBX_HIP(ctx, hipSetDevice(ctx->device));
  // Reference order per step is f <- S(C(f)).  Over n steps:
  //   f_n = S (C S)^(n-1) C f_0
  // so: one collide-only pass, n-1 fused pull+collide passes, one
  // stream-only pass.  Per-cell arithmetic is unchanged, so results are
  // bit-identical to the reference order.
  if (int rc = launch_step(sim, L, lbx::STEP_COLLIDE)) return rc;
  for (int s = 1; s < nsteps; ++s)
    if (int rc = launch_step(sim, L, lbx::STEP_FUSED)) return rc;
  if (int rc = launch_step(sim, L, lbx::STEP_STREAM)) return rc;
  return LBX_OK;
}

int lbx_hydro(lbx_sim* sim, int level, int box, double* rho, double* u) {
  if (int rc = check_sim(sim)) return rc;
  if (int rc = check_box(sim, level, box)) return rc;
  if (!rho || !u) return fail(sim->ctx, LBX_E_ARG, "null buffer");
  lbx_ctx* ctx = sim->ctx;
  Level& L = sim->lev[level];
  const size_t n = (size_t)L.boxes[box].cells();
  DeviceTemp t;
  LBX_HIP(ctx, hipMalloc(&t.p, n * 4 * sizeof(double)));
  double* drho = (double*)t.p;
  double* du = drho + n;
  LBX_HIP(ctx, lbx::launch_hydro(sim->precision, L.buf[L.cur], L.lat, region_of(L.boxes[box]), drho,
                                 du, sim->rho0, ctx->stream));
  LBX_HIP(ctx, hipMemcpyAsync(rho, drho, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  LBX_HIP(ctx, hipMemcpyAsync(u, du, 3 * n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  LBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LBX_OK;
}

int lbx_conservation(lbx_sim* sim, int level, double out[4]) {
  if (int rc = check_sim(sim)) return rc;
  if (!out) return fail(sim->ctx, LBX_E_ARG, "null buffer");
  if (level < 0 || level > 1 || !sim->lev[level].defined)
    return fail(sim->ctx, LBX_E_STATE, "level %d is not defined", level);
  lbx_ctx* ctx = sim->ctx;
  Level& L = sim->lev[level];
  double* partials = ctx->scratch;
  double* res = ctx->scratch + kPartials * 4;
  LBX_HIP(ctx, lbx::launch_conservation(sim->precision, L.buf[L.cur], L.lat, sim->rho0, partials,
                                        kPartials, res, ctx->stream));
  LBX_HIP(ctx, hipMemcpyAsync(out, res, 4 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  LBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LBX_OK;
}

int lbx_profile_enable(lbx_ctx* ctx, int enable) {
  if (!ctx) return fail(nullptr, LBX_E_ARG, "null context");
  if (int rc = drain_profile(ctx)) return rc;
  ctx->prof = enable != 0;
  ctx->prof_ms = 0.0;
  ctx->prof_launches = 0;
  return LBX_OK;
}

int lbx_profile_read(lbx_ctx* ctx, double* ms, int64_t* launches) {
  if (!ctx || !ms || !launches) return fail(ctx, LBX_E_ARG, "null argument");
  if (int rc = drain_profile(ctx)) return rc;
  *ms = ctx->prof_ms;
  *launches = ctx->prof_launches;
  return LBX_OK;
}

}  // extern "C"
