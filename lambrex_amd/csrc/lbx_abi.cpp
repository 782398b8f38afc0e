// lbx_abi.cpp -- the C ABI declared in include/lbx.h.
//
// Owns device memory and the per-context HIP stream; never aborts: every
// failure becomes a negative status plus a message in lbx_last_error().
//
// Level 0 is a periodic level array (z-plane major; its boxes are index
// ranges into it).  Every finer level is a set of patches, one per box, each
// grown by 2 ghost layers; the tables that drive its ghost fill (from its own
// patches or its parent) and its average-down into the parent are built here
// on the host once per level definition.  One level-0 step of a hierarchy is
// the recursive Rohde cycle, captured as one HIP graph per buffer parity.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <map>
#include <mutex>
#include <tuple>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "../../include/lbx.h"
#include "lbx_internal.h"

using lbx::Lattice;
using lbx::LaunchShape;
using lbx::PatchDesc;
using lbx::Region;
using lbx::Relax;

namespace {

thread_local std::string g_global_error;

struct Box {
  int lo[3], hi[3];
  int n(int d) const { return hi[d] - lo[d] + 1; }
  long long cells() const { return (long long)n(0) * n(1) * n(2); }
  bool contains(int x, int y, int z) const {
    return x >= lo[0] && x <= hi[0] && y >= lo[1] && y <= hi[1] && z >= lo[2] && z <= hi[2];
  }
};

struct Level {
  bool defined = false;
  bool patched = false;  // per-box ghosted storage (levels >= 1)
  int ratio = 2;         // refinement ratio to the parent (patched levels)
  std::vector<Box> boxes;
  Lattice lat{};         // level-array geometry (level 0)
  void* buf[3] = {nullptr, nullptr, nullptr};  // state in buf[cur]; buf[2] scratch (patched)
  int cur = 0;
  // Level 0 of a single-level simulation between Iterate calls: buf[cur]
  // holds the post-collision state g of the last step, and the reference's
  // pre-collision populations are f = S(g) (one pull).  Iterate then runs
  // only fused passes, and readers gather f through the pull (lazy
  // materialisation); anything else that needs f in place first runs the
  // stream-only pass (materialise()).  False: buf[cur] holds f itself.
  bool post = false;
  double omega_s = 0.0, omega_b = 0.0;
  bool relax_set = false;
  // device box/patch descriptors
  std::vector<PatchDesc> desc;
  PatchDesc* d_desc = nullptr;
  int max_cells = 0;
  // level 0 with a finer level: per-cell box-face and fine-mask flags of the
  // coarse push (launch_coarse_push), and its launch shape
  unsigned char* d_mask = nullptr;
  LaunchShape push_shape{};
  // patched levels: ghost fill sources and the average-down contributions
  // into the parent; d_mask is then one byte per grown cell (the parent's
  // view of this level's own finer level)
  long long* d_fill = nullptr;
  long long grown_cells = 0;        // length of the fill table
  long long* d_sum_cell = nullptr;  // level-array parent: its element per touched cell
  int* d_sum_start = nullptr;
  int* d_sum_contrib = nullptr;  // index into the block averages, per contribution
  int sum_ncell = 0;
  // patched parent: touched-cell index per grown cell of the parent (-1 if
  // none), for the parent's fused stream + sum
  int* d_sum_slot = nullptr;
  // Patches are stored in the buffers in `gdesc` order (box corners, x
  // fastest), whatever order the caller gave the boxes in, and the finest
  // level's substep pair runs over `groups` of consecutive ones there (about
  // kFineGroupBytes of one buffer each), so a group is a compact set of
  // neighbouring boxes.  gdesc holds copies of desc's entries in that order.
  // With `regions` the array between the two substeps lives in one region of
  // buf[2] per stream (region_span elements).
  std::vector<PatchDesc> gdesc;
  PatchDesc* d_gdesc = nullptr;
  std::vector<std::pair<int, int>> groups;  // (first position in gdesc, count)
  long long region_span = 0;
  bool regions = false;
  double* d_avg = nullptr;       // 15 planes of avg_stride 2x2x2 block averages
  long long avg_stride = 0;
  int max_blocks = 0;            // largest (g/2)^3 of a patch
};

// One captured level-0 step of a hierarchy per (level-0 parity, finest-level
// parity).
// Captured level-0 steps, one per combination of the levels' buffer
// parities at the start of the step (bit l = level l's `cur`).
struct StepGraphs {
  hipGraphExec_t exec[1 << LBX_MAX_LEVELS] = {};
  void clear() {
    for (auto& e : exec)
      if (e) {
        (void)hipGraphExecDestroy(e);
        e = nullptr;
      }
  }
};

// Conservation sums: a level array is reduced by kLatticePartials
// workgroups; a patched level by patch_chunks(n) workgroups per patch.  The
// context's scratch holds kPartials partials (4 doubles each) + 4 results.
constexpr int kLatticePartials = 1024;
constexpr int kPartials = 65536;
int patch_chunks(int npatch) { return std::max(1, std::min(64, kLatticePartials / std::max(1, npatch))); }

}  // namespace

struct lbx_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  // second stream (default priority) and the events that fork it from and
  // join it back to `stream` for independent patch groups
  hipStream_t side = nullptr;
  hipEvent_t fork_ev = nullptr, join_ev = nullptr;
  // third stream (default priority): a hierarchy step's level-0 push, which
  // only reads level 0's current state, runs beside the finer levels' work
  hipStream_t push = nullptr;
  hipEvent_t push_fork_ev = nullptr, push_join_ev = nullptr;
  std::string err;
  double* sums = nullptr;  // conservation sums: kPartials*4 partials + 4 results
  // staging buffer of host transfers (grows on demand, kept until destroy;
  // every call that uses it synchronises before returning)
  void* stage = nullptr;
  size_t stage_bytes = 0;
  // profiling of stream-collide launches
  bool prof = false;
  std::vector<hipEvent_t> ev_pool;
  size_t ev_used = 0;
  double prof_ms = 0.0;
  int64_t prof_launches = 0;
  // autotuned workgroup shapes of the fused kernel, per (nx, ny, precision)
  bool autotune = false;
  std::map<std::tuple<int, int, int>, LaunchShape> shapes;
  std::map<std::tuple<int, int, int>, std::vector<std::pair<LaunchShape, float>>> trials;
};

struct lbx_sim {
  lbx_ctx* ctx = nullptr;
  int dims[3] = {0, 0, 0};
  unsigned walls = 0;  // bit d: bounce-back walls normal to axis d
  int precision = LBX_FP64;
  double rho0 = 1.0;
  // refinement ratio between consecutive levels (2, 3 or 4; the same on
  // every level)
  int ratio = 2;
  Level lev[LBX_MAX_LEVELS];
  StepGraphs graphs;
  bool use_graphs = true;
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

#define LBX_HIP(ctx, call)                                 \
  do {                                                     \
    hipError_t e_ = (call);                                \
    if (e_ != hipSuccess) return hip_fail(ctx, e_, #call); \
  } while (0)

size_t elem_size(int precision) { return precision == LBX_FP64 ? 8 : 4; }

bool box_ok(const Box& b, const int* dom) {
  for (int d = 0; d < 3; ++d)
    if (b.lo[d] < 0 || b.hi[d] >= dom[d] || b.hi[d] < b.lo[d]) return false;
  return true;
}

bool overlap(const Box& a, const Box& b) {
  for (int d = 0; d < 3; ++d)
    if (a.hi[d] < b.lo[d] || b.hi[d] < a.lo[d]) return false;
  return true;
}

bool any_overlap(const std::vector<Box>& boxes) {
  for (size_t a = 0; a < boxes.size(); ++a)
    for (size_t c = a + 1; c < boxes.size(); ++c)
      if (overlap(boxes[a], boxes[c])) return true;
  return false;
}

Region region_of(const Box& b) {
  Region r;
  for (int d = 0; d < 3; ++d) {
    r.lo[d] = b.lo[d];
    r.n[d] = b.n(d);
  }
  return r;
}

int wrap(int v, int n) {
  const int r = v % n;
  return r < 0 ? r + n : r;
}
int floor_div(int v, int r) { return v >= 0 ? v / r : -((-v + r - 1) / r); }

template <typename T>
hipError_t upload(T** dptr, const std::vector<T>& h, hipStream_t st) {
  if (h.empty()) return hipSuccess;
  hipError_t e = hipMalloc(dptr, h.size() * sizeof(T));
  if (e != hipSuccess) return e;
  e = hipMemcpyAsync(*dptr, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, st);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(st);
}

void free_level(Level& L) {
  for (void*& p : L.buf)
    if (p) (void)hipFree(p);
  if (L.d_desc) (void)hipFree(L.d_desc);
  if (L.d_mask) (void)hipFree(L.d_mask);
  if (L.d_fill) (void)hipFree(L.d_fill);
  if (L.d_sum_cell) (void)hipFree(L.d_sum_cell);
  if (L.d_sum_start) (void)hipFree(L.d_sum_start);
  if (L.d_sum_contrib) (void)hipFree(L.d_sum_contrib);
  if (L.d_sum_slot) (void)hipFree(L.d_sum_slot);
  if (L.d_gdesc) (void)hipFree(L.d_gdesc);
  if (L.d_avg) (void)hipFree(L.d_avg);
  L = Level{};
}

int check_sim(lbx_sim* sim) {
  if (!sim || !sim->ctx) return fail(nullptr, LBX_E_ARG, "null simulation handle");
  return LBX_OK;
}

int check_box(lbx_sim* sim, int level, int box) {
  if (level < 0 || level >= LBX_MAX_LEVELS) return fail(sim->ctx, LBX_E_ARG, "level %d out of range", level);
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

// The context's staging buffer, at least `bytes` long.
hipError_t stage_buffer(lbx_ctx* ctx, size_t bytes, void** out) {
  if (bytes > ctx->stage_bytes) {
    if (ctx->stage) {
      const hipError_t e = hipFree(ctx->stage);
      ctx->stage = nullptr;
      ctx->stage_bytes = 0;
      if (e != hipSuccess) return e;
    }
    const size_t want = std::max(bytes, (size_t)1 << 20);
    const hipError_t e = hipMalloc(&ctx->stage, want);
    if (e != hipSuccess) return e;
    ctx->stage_bytes = want;
  }
  *out = ctx->stage;
  return hipSuccess;
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

// Event pair around the timed region of one step (when profiling is on).
struct ProfileScope {
  lbx_ctx* ctx;
  hipEvent_t e1 = nullptr;
  int rc = LBX_OK;
  ProfileScope(lbx_ctx* c, bool timed) : ctx(c) {
    if (!ctx->prof || !timed) return;
    if (ctx->ev_used + 2 > 4096 && (rc = drain_profile(ctx)) != LBX_OK) return;
    hipEvent_t e0 = take_event(ctx);
    e1 = take_event(ctx);
    if (!e0 || !e1) {
      rc = fail(ctx, LBX_E_HIP, "hipEventCreate failed");
      e1 = nullptr;
      return;
    }
    if (hipEventRecord(e0, ctx->stream) != hipSuccess) rc = fail(ctx, LBX_E_HIP, "hipEventRecord");
  }
  int close() {
    if (e1 && rc == LBX_OK && hipEventRecord(e1, ctx->stream) != hipSuccess)
      rc = fail(ctx, LBX_E_HIP, "hipEventRecord");
    return rc;
  }
};

// Autotuner: every candidate workgroup shape of the fused kernel is timed on
// a proxy lattice with the same rows (nx, ny) and a few z-planes; the
// fastest is kept for that (nx, ny, precision).  Shapes change only which
// lanes compute which cells, never the arithmetic, so results do not depend
// on the choice.
std::vector<LaunchShape> shape_candidates(const Lattice& L) {
  std::vector<LaunchShape> c;
  for (int lanes : {256, 512}) {
    for (int cap : {0, 1, 2, 3, 4, 6}) {
      if (cap == 1 && lanes == 256) continue;  // 4 wavefronts per CU
      if (cap > 3 && lanes == 512) continue;
      const int lds = lbx::lds_bytes_for_cap(cap);
      if (lds < 0) continue;
      for (int bx : {256, 128, 64, 32}) {
        const int x = std::min(bx, L.nx);
        LaunchShape s{x, std::max(1, std::min(lanes / x, L.ny)), cap, lds};
        bool seen = false;
        for (const LaunchShape& o : c) seen |= (o.bx == s.bx && o.by == s.by && o.cap == s.cap);
        if (!seen) c.push_back(s);
      }
    }
  }
  return c;
}

int tune_shape(lbx_ctx* ctx, const Lattice& L, int precision, LaunchShape* out) {
  const auto key = std::make_tuple(L.nx, L.ny, precision);
  auto it = ctx->shapes.find(key);
  if (it != ctx->shapes.end()) {
    *out = it->second;
    return LBX_OK;
  }
  const Lattice P =
      lbx::make_lattice(L.nx, L.ny, std::min(L.nz, std::max(4, (1 << 24) / std::max(1, L.nx * L.ny))),
                        (int)elem_size(precision));
  const size_t bytes = (size_t)P.elems() * elem_size(precision);
  DeviceTemp a, b;
  LBX_HIP(ctx, hipMalloc(&a.p, bytes));
  LBX_HIP(ctx, hipMalloc(&b.p, bytes));
  // a positive finite value in every element (0x3F800000 pairs / floats)
  LBX_HIP(ctx, hipMemsetD32Async((hipDeviceptr_t)a.p, 0x3F800000, bytes / 4, ctx->stream));
  LBX_HIP(ctx, hipMemsetD32Async((hipDeviceptr_t)b.p, 0x3F800000, bytes / 4, ctx->stream));
  const Relax r{1.0 / 1.1, 1.0 / 1.1, 1.0};
  hipEvent_t e0 = nullptr, e1 = nullptr;
  LBX_HIP(ctx, hipEventCreate(&e0));
  if (hipEventCreate(&e1) != hipSuccess) {
    (void)hipEventDestroy(e0);
    return fail(ctx, LBX_E_HIP, "hipEventCreate");
  }
  std::vector<std::pair<LaunchShape, float>> trial;
  const LaunchShape dflt = lbx::default_shape(L, precision);
  LaunchShape best = dflt;
  float best_ms = 1e30f, dflt_ms = 1e30f;
  hipError_t e = hipSuccess;
  for (const LaunchShape& sh : shape_candidates(P)) {
    auto launch = [&](int k) {
      hipError_t ee = hipSuccess;
      for (int i = 0; i < k && ee == hipSuccess; ++i)
        ee = precision == LBX_FP64
                 ? lbx::launch_step_f64(lbx::STEP_FUSED, (const double*)a.p, (double*)b.p, P, r, sh, ctx->stream)
                 : lbx::launch_step_f32s(lbx::STEP_FUSED, (const float*)a.p, (float*)b.p, P, r, sh, ctx->stream);
      return ee;
    };
    e = launch(2);
    if (e == hipSuccess) e = hipEventRecord(e0, ctx->stream);
    if (e == hipSuccess) e = launch(10);
    if (e == hipSuccess) e = hipEventRecord(e1, ctx->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e != hipSuccess) break;
    ms /= 10;
    trial.push_back({sh, ms});
    if (sh.bx == dflt.bx && sh.by == dflt.by && sh.cap == dflt.cap) dflt_ms = ms;
    if (ms < best_ms) {
      best_ms = ms;
      best = sh;
    }
  }
  // keep the default unless a candidate is clearly faster (timing noise is ~1%)
  if (dflt_ms < 1e30f && best_ms > 0.99f * dflt_ms) best = dflt;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (e != hipSuccess) return hip_fail(ctx, e, "autotune");
  ctx->shapes[key] = best;
  ctx->trials[key] = trial;
  *out = best;
  return LBX_OK;
}

int launch_step(lbx_sim* sim, Level& L, lbx::StepKind kind) {
  lbx_ctx* ctx = sim->ctx;
  Relax r{L.omega_s, L.omega_b, sim->rho0};
  LaunchShape shape = lbx::default_shape(L.lat, sim->precision);
  if (ctx->autotune)
    if (int rc = tune_shape(ctx, L.lat, sim->precision, &shape)) return rc;
  ProfileScope ps(ctx, kind == lbx::STEP_FUSED);
  if (ps.rc) return ps.rc;
  hipError_t e;
  void* src = L.buf[L.cur];
  void* dst = L.buf[1 - L.cur];
  if (sim->precision == LBX_FP64)
    e = lbx::launch_step_f64(kind, (const double*)src, (double*)dst, L.lat, r, shape, ctx->stream);
  else
    e = lbx::launch_step_f32s(kind, (const float*)src, (float*)dst, L.lat, r, shape, ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "stream-collide launch");
  L.cur = 1 - L.cur;
  return ps.close();
}

// Bring level 0 back to the pre-collision populations f in place (the
// stream-only pass) if it holds the post-collision state (Level::post).
int materialise(lbx_sim* sim) {
  Level& L = sim->lev[0];
  if (!L.defined || !L.post) return LBX_OK;
  if (int rc = launch_step(sim, L, lbx::STEP_STREAM)) return rc;
  L.post = false;
  return LBX_OK;
}

// The finest defined level (levels are defined contiguously from 0).
int finest_level(const lbx_sim* sim) {
  int f = 0;
  while (f + 1 < LBX_MAX_LEVELS && sim->lev[f + 1].defined) ++f;
  return f;
}

void free_levels_from(lbx_sim* sim, int level) {
  for (int l = level; l < LBX_MAX_LEVELS; ++l) free_level(sim->lev[l]);
}

// The kernels of one level-0 step of the Rohde cycle (src/AmrSim.cpp:430-469)
// for the current buffer parities.
// The finest level's patches run in groups of about kFineGroupBytes of state
// (one buffer), alternating over two streams: two groups' intermediate
// arrays (≈ 2 x 2 x 48 MB) then stay in the 256 MB MALL.
constexpr long long kFineGroupBytes = 48ll << 20;

// The finest level's substeps (two at ratio 2, r at ratio r), from its
// parent's current state (level 0's `now`, or a middle level's state, which
// its own step rewrites only after this).  Its block averages land in
// F.d_avg.  par[l] is the buffer holding level l's state while the step is
// enqueued; the finest level's flips.
hipError_t enqueue_finest_pair(lbx_sim* sim, int f, int* par, hipStream_t st) {
  Level& F = sim->lev[f];
  const Level& C = sim->lev[f - 1];
  const int prec = sim->precision;
  const Relax rf{F.omega_s, F.omega_b, sim->rho0};
  int& fpar = par[f];
  const void* cnow = C.buf[par[f - 1]];
  void* fnow = F.buf[fpar];
  void* fnext = F.buf[1 - fpar];
  void* ftmp = F.buf[2];
  const PatchDesc* cpatches = C.patched ? C.d_desc : nullptr;
  // The fine patches' two substeps depend on no other patch's substeps (the
  // ghosts are filled once, from fnow and cnow), so they run in the level's
  // groups of neighbouring patches, whose arrays stay in the MALL between
  // the passes.  Groups write disjoint patches (and block sums), so
  // consecutive groups alternate between the context's stream and its side
  // stream, forked from and joined back to `st` around the loop; one group's
  // launch ramp then overlaps the other's tail.  The array between the
  // substeps is dead once its group ends: with F.regions each stream
  // rewrites one region of buf[2] for it (the group's patches are
  // contiguous, patch p at p.base - base(first)) while its lines are still
  // in the MALL, instead of writing each group's back to HBM.  The groups
  // run in storage order when fpar is 0 and in reverse order when it is 1,
  // so each pair starts with the groups whose state the previous pair wrote
  // last, while it may still be in the MALL.
  lbx_ctx* ctx = sim->ctx;
  const auto& groups = F.groups;
  const size_t ng = groups.size();
  const bool fork = ng > 1, rev = fpar == 1;
  hipError_t e = hipSuccess;
  if (fork) {
    e = hipEventRecord(ctx->fork_ev, st);
    if (e == hipSuccess) e = hipStreamWaitEvent(ctx->side, ctx->fork_ev, 0);
  }
  for (size_t k = 0; k < ng && e == hipSuccess; ++k) {
    const auto& grp = groups[rev ? ng - 1 - k : k];
    const int first = grp.first, n = grp.second;
    hipStream_t gs = (k & 1) ? ctx->side : st;
    void* mid = ftmp;
    long long shift = 0;
    if (F.regions) {
      mid = static_cast<char*>(ftmp) + (k & 1) * F.region_span * elem_size(prec);
      shift = F.gdesc[first].base;
    }
    if (F.ratio == 2) {
      e = lbx::launch_fine_fill_collide(prec, fnow, fnext, cnow, C.lat, cpatches, F.d_desc, F.d_gdesc, first, n,
                                        F.max_cells, F.d_fill, nullptr, rf, gs);
      if (e == hipSuccess)
        e = lbx::launch_fine_stream_collide(prec, fnext, mid, shift, F.d_gdesc, first, n, F.max_cells, rf, gs);
      if (e == hipSuccess)
        e = lbx::launch_fine_stream_avg(prec, mid, shift, fnext, F.d_avg, F.avg_stride, F.d_gdesc, first, n,
                                        F.max_blocks, sim->rho0, gs);
    } else {
      // ratio r > 2: r substeps, each a collide then a stream, chained
      // through buf[2] (whole, no regions) and fnext so that the last one
      // reads buf[2]: the fill-collide writes stage 1, the k-th
      // stream-collide reads stage k and writes stage k+1, stage k in buf[2]
      // when r - k is even; then the last stream into fnext and the
      // average-down of its valid cells.
      const int r = F.ratio;
      auto stage = [&](int k) { return ((r - k) % 2 == 0) ? mid : fnext; };
      e = lbx::launch_fine_fill_collide(prec, fnow, stage(1), cnow, C.lat, cpatches, F.d_desc, F.d_gdesc, first, n,
                                        F.max_cells, F.d_fill, nullptr, rf, gs);
      for (int k = 1; k < r && e == hipSuccess; ++k)
        e = lbx::launch_fine_stream_collide(prec, stage(k), stage(k + 1), 0, F.d_gdesc, first, n, F.max_cells, rf, gs,
                                            k > 1);
      if (e == hipSuccess)
        e = lbx::launch_fine_stream_valid(prec, mid, fnext, F.d_gdesc, first, n, F.max_cells, sim->rho0, gs);
      if (e == hipSuccess)
        e = lbx::launch_fine_avg_valid(prec, fnext, F.d_avg, F.avg_stride, F.d_gdesc, first, n, F.max_blocks, r,
                                       sim->rho0, gs);
    }
  }
  if (fork) {  // join even after a failed launch, so a capture can end
    const hipError_t ej = hipEventRecord(ctx->join_ev, ctx->side);
    const hipError_t ew = ej == hipSuccess ? hipStreamWaitEvent(st, ctx->join_ev, 0) : ej;
    if (e == hipSuccess) e = ew;
  }
  fpar = 1 - fpar;
  return e;
}

hipError_t enqueue_children(lbx_sim* sim, int l, int finest, int* par, hipStream_t st);

// One step of a middle level m (0 < m < finest), RohdeCycle(m): the finer
// levels' steps, which read m's state; then, per group of m's patches, the
// two-level fill from the parent with comp 0 zeroed on the outer layer and
// the masked collide, and Stream + ZeroInvalidComponents with SumFromFine
// added into the valid cells (one fused kernel) into m's other buffer,
// which becomes its state.  The reference collides before the finer levels'
// steps; the collide reads only m's and its parent's current states, which
// those steps do not change, so it can follow them, and m's passes then run
// in the same MALL-sized groups as the finest level's (the array between
// the two in a group-sized region of buf[2] per stream, consecutive groups
// on two streams): the fill of a group reads the neighbours' current state,
// which no group overwrites.  In the last of its parent's r steps (`last`)
// the block averages of the result for the parent are formed too: in the
// same pass at ratio 2, where they include 2 fine ghost layers, or by a pass
// over the valid cells above it (nGrow = 2 / r = 0).
hipError_t enqueue_middle_step(lbx_sim* sim, int m, int finest, int* par, hipStream_t st, bool last) {
  Level& L = sim->lev[m];
  const Level& P = sim->lev[m - 1];
  const Level& F = sim->lev[m + 1];
  const int prec = sim->precision;
  const Relax r{L.omega_s, L.omega_b, sim->rho0};
  hipError_t e = enqueue_children(sim, m, finest, par, st);
  if (e != hipSuccess) return e;
  void* now = L.buf[par[m]];
  void* next = L.buf[1 - par[m]];
  const void* pnow = P.buf[par[m - 1]];
  lbx_ctx* ctx = sim->ctx;
  const auto& groups = L.groups;
  const bool fork = groups.size() > 1;
  if (fork) {
    e = hipEventRecord(ctx->fork_ev, st);
    if (e == hipSuccess) e = hipStreamWaitEvent(ctx->side, ctx->fork_ev, 0);
  }
  for (size_t k = 0; k < groups.size() && e == hipSuccess; ++k) {
    const int first = groups[k].first, n = groups[k].second;
    hipStream_t gs = (k & 1) ? ctx->side : st;
    void* mid = L.buf[2];
    long long shift = 0;
    if (L.regions) {
      mid = static_cast<char*>(L.buf[2]) + (k & 1) * L.region_span * elem_size(prec);
      shift = L.gdesc[first].base;
    }
    e = lbx::launch_fine_fill_collide(prec, now, mid, pnow, P.lat, P.patched ? P.d_desc : nullptr, L.d_desc,
                                      L.d_gdesc, first, n, L.max_cells, L.d_fill, L.d_mask, r, gs, shift);
    if (e == hipSuccess && last && L.ratio != 2) {
      e = lbx::launch_fine_stream_zero(prec, mid, next, L.d_gdesc, first, n, L.max_cells, sim->rho0, F.d_sum_slot,
                                       F.d_avg, F.avg_stride, F.d_sum_start, F.d_sum_contrib, gs, shift);
      if (e == hipSuccess)
        e = lbx::launch_fine_avg_valid(prec, next, L.d_avg, L.avg_stride, L.d_gdesc, first, n, L.max_blocks,
                                       L.ratio, sim->rho0, gs);
    } else if (e == hipSuccess && last) {  // the state the parent averages down: its block sums in the same pass
      e = lbx::launch_middle_stream_avg(prec, mid, shift, next, L.d_avg, L.avg_stride, L.d_gdesc, first, n,
                                        L.max_blocks, sim->rho0, F.d_sum_slot, F.d_avg, F.avg_stride, F.d_sum_start,
                                        F.d_sum_contrib, gs);
    } else if (e == hipSuccess) {
      e = lbx::launch_fine_stream_zero(prec, mid, next, L.d_gdesc, first, n, L.max_cells, sim->rho0, F.d_sum_slot,
                                       F.d_avg, F.avg_stride, F.d_sum_start, F.d_sum_contrib, gs, shift);
    }
  }
  if (fork) {  // join even after a failed launch, so a capture can end
    const hipError_t ej = hipEventRecord(ctx->join_ev, ctx->side);
    const hipError_t ew = ej == hipSuccess ? hipStreamWaitEvent(st, ctx->join_ev, 0) : ej;
    if (e == hipSuccess) e = ew;
  }
  par[m] = 1 - par[m];
  return e;
}

// What level l's step needs of level l + 1: the finest level's substeps,
// or refRatio middle steps, the last followed by the block averages of the
// result (src/AmrSim.cpp:445-455).
hipError_t enqueue_children(lbx_sim* sim, int l, int finest, int* par, hipStream_t st) {
  const int c = l + 1;
  if (c == finest) return enqueue_finest_pair(sim, c, par, st);
  const int r = sim->lev[c].ratio;
  hipError_t e = hipSuccess;
  for (int k = 0; k < r && e == hipSuccess; ++k) e = enqueue_middle_step(sim, c, finest, par, st, k == r - 1);
  return e;
}

hipError_t enqueue_hierarchy_step(lbx_sim* sim, hipStream_t st) {
  Level& C = sim->lev[0];
  const Level& F = sim->lev[1];
  const int finest = finest_level(sim);
  const Relax rc{C.omega_s, C.omega_b, sim->rho0};
  const void* cnow = C.buf[C.cur];
  void* cnext = C.buf[1 - C.cur];
  int par[LBX_MAX_LEVELS];
  for (int l = 0; l < LBX_MAX_LEVELS; ++l) par[l] = sim->lev[l].cur;
  // CoarseCollide + Stream(0) read level 0's `now` and write its `next`;
  // the finer levels only read `now` (their parent fills), so the push runs
  // on its own stream beside them, and SumFromFine(0) into `next` waits for
  // both.
  lbx_ctx* ctx = sim->ctx;
  hipError_t e = hipEventRecord(ctx->push_fork_ev, st);
  if (e == hipSuccess) e = hipStreamWaitEvent(ctx->push, ctx->push_fork_ev, 0);
  if (e != hipSuccess) return e;
  const hipError_t ep = lbx::launch_coarse_push(sim->precision, cnow, cnext, C.lat, C.d_mask, rc, C.push_shape,
                                                ctx->push);
  e = enqueue_children(sim, 0, finest, par, st);
  // join even after a failed launch, so a capture can end
  const hipError_t ej = hipEventRecord(ctx->push_join_ev, ctx->push);
  const hipError_t ew = ej == hipSuccess ? hipStreamWaitEvent(st, ctx->push_join_ev, 0) : ej;
  if (e == hipSuccess) e = ep;
  if (e == hipSuccess) e = ew;
  if (e == hipSuccess)
    e = lbx::launch_sum_apply(sim->precision, F.d_avg, F.avg_stride, cnext, C.lat, F.d_sum_cell,
                              F.d_sum_start, F.d_sum_contrib, F.sum_ncell, sim->rho0, st);
  return e;
}

// One level-0 step: replayed from a captured graph, or launched eagerly
// while profiling (event records cannot sit inside the capture).
int launch_hierarchy_step(lbx_sim* sim) {
  lbx_ctx* ctx = sim->ctx;
  Level& C = sim->lev[0];
  const int finest = finest_level(sim);
  Level& F = sim->lev[finest];
  if (sim->use_graphs && !ctx->prof) {
    unsigned key = 0;
    for (int l = 0; l <= finest; ++l) key |= (unsigned)sim->lev[l].cur << l;
    hipGraphExec_t& ex = sim->graphs.exec[key];
    if (!ex) {
      hipGraph_t g = nullptr;
      LBX_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
      const hipError_t ek = enqueue_hierarchy_step(sim, ctx->stream);
      const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
      if (ek != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return hip_fail(ctx, ek, "hierarchy step capture");
      }
      if (ec != hipSuccess) return hip_fail(ctx, ec, "hipStreamEndCapture");
      const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
      (void)hipGraphDestroy(g);
      if (ei != hipSuccess) {
        ex = nullptr;
        return hip_fail(ctx, ei, "hipGraphInstantiate");
      }
    }
    LBX_HIP(ctx, hipGraphLaunch(ex, ctx->stream));
  } else {
    ProfileScope ps(ctx, true);
    if (ps.rc) return ps.rc;
    LBX_HIP(ctx, enqueue_hierarchy_step(sim, ctx->stream));
    if (int rc = ps.close()) return rc;
  }
  // every step of a middle level flips its buffer, and level l takes the
  // product of the ratios of levels 1..l steps per level-0 step; the finest
  // level's substeps run (and flip its buffer) once per step of its parent
  long long steps = 1;
  for (int l = 1; l < finest; ++l) {
    steps *= sim->lev[l].ratio;
    if (steps % 2) sim->lev[l].cur = 1 - sim->lev[l].cur;
  }
  if (steps % 2) F.cur = 1 - F.cur;
  C.cur = 1 - C.cur;
  return LBX_OK;
}

std::vector<Box> read_boxes(int n, const int32_t* lohi) {
  std::vector<Box> boxes(n);
  for (int b = 0; b < n; ++b)
    for (int d = 0; d < 3; ++d) {
      boxes[b].lo[d] = lohi[6 * b + d];
      boxes[b].hi[d] = lohi[6 * b + 3 + d];
    }
  return boxes;
}

int define_level0(lbx_sim* sim, const std::vector<Box>& boxes) {
  lbx_ctx* ctx = sim->ctx;
  long long total = 0;
  for (size_t b = 0; b < boxes.size(); ++b) {
    if (!box_ok(boxes[b], sim->dims))
      return fail(ctx, LBX_E_ARG, "box %zu lies outside the domain or is empty", b);
    total += boxes[b].cells();
  }
  if (total != (long long)sim->dims[0] * sim->dims[1] * sim->dims[2] || any_overlap(boxes))
    return fail(ctx, LBX_E_ARG, "level-0 boxes must tile the periodic domain exactly");
  Level& L = sim->lev[0];
  free_levels_from(sim, 0);
  L.boxes = boxes;
  L.lat = lbx::make_lattice(sim->dims[0], sim->dims[1], sim->dims[2], (int)elem_size(sim->precision));
  L.lat.walls = sim->walls;
  const size_t bytes = (size_t)L.lat.elems() * elem_size(sim->precision);
  for (int i = 0; i < 2; ++i) {
    hipError_t e = hipMalloc(&L.buf[i], bytes);
    if (e == hipSuccess) e = hipMemsetAsync(L.buf[i], 0, bytes, ctx->stream);
    if (e != hipSuccess) {
      free_level(L);
      return fail(ctx, LBX_E_NOMEM, "allocation of %zu bytes for level 0 failed: %s", bytes,
                  hipGetErrorString(e));
    }
  }
  for (const Box& b : boxes) {
    PatchDesc d{};
    for (int k = 0; k < 3; ++k) {
      d.lo[k] = b.lo[k];
      d.hi[k] = b.hi[k];
      d.g[k] = b.n(k);
    }
    d.cells = (int)b.cells();
    L.max_cells = std::max(L.max_cells, d.cells);
    L.desc.push_back(d);
  }
  hipError_t e = upload(&L.d_desc, L.desc, ctx->stream);
  if (e != hipSuccess) {
    free_level(L);
    return hip_fail(ctx, e, "level-0 box table");
  }
  L.defined = true;
  return LBX_OK;
}

// Per coarse cell (x fastest): bits 0-5 = moving by +x, -x, +y, -y, +z, -z
// leaves the cell's level-0 box (the coarse push then streams the
// pre-collision value, as the reference's pull from stale ghost copies
// does); bit 7 = the cell lies under one of the mask boxes (fine index
// space, coarsened by the refinement ratio r) and is zeroed instead of
// collided.
std::vector<unsigned char> coarse_flags(const Level& C, const std::vector<Box>& mask, int r) {
  const Lattice& L = C.lat;
  std::vector<unsigned char> fl((size_t)L.nx * L.ny * L.nz, 0);
  auto at = [&](int x, int y, int z) -> unsigned char& { return fl[((size_t)z * L.ny + y) * L.nx + x]; };
  for (const Box& b : C.boxes)
    for (int z = b.lo[2]; z <= b.hi[2]; ++z)
      for (int y = b.lo[1]; y <= b.hi[1]; ++y)
        for (int x = b.lo[0]; x <= b.hi[0]; ++x)
          at(x, y, z) = (unsigned char)((x == b.hi[0] ? 1 : 0) | (x == b.lo[0] ? 2 : 0) | (y == b.hi[1] ? 4 : 0) |
                                        (y == b.lo[1] ? 8 : 0) | (z == b.hi[2] ? 16 : 0) | (z == b.lo[2] ? 32 : 0));
  for (const Box& b : mask)
    for (int z = std::max(0, floor_div(b.lo[2], r)); z <= std::min(L.nz - 1, floor_div(b.hi[2], r)); ++z)
      for (int y = std::max(0, floor_div(b.lo[1], r)); y <= std::min(L.ny - 1, floor_div(b.hi[1], r)); ++y)
        for (int x = std::max(0, floor_div(b.lo[0], r)); x <= std::min(L.nx - 1, floor_div(b.hi[0], r)); ++x)
          at(x, y, z) |= 0x80;
  return fl;
}

// Grown index of absolute position (x, y, z) in patch P.
long long grown_index(const PatchDesc& P, int x, int y, int z) {
  return ((long long)(z - P.lo[2] + 2) * P.g[1] + (y - P.lo[1] + 2)) * P.g[0] + (x - P.lo[0] + 2);
}

// Where a level keeps the valid cell at (x, y, z) of its own index space
// (already wrapped into the domain): the element of population 0 in a level
// array, or (patch << 32) | grown index on a patched level.  False if no
// valid cell of the level holds it.  `hint` is the box found last.
bool cell_ref(const Level& C, int x, int y, int z, long long* ref, int& hint) {
  if (!C.patched) {
    *ref = C.lat.at(0, y * C.lat.nx + x, z);
    return true;
  }
  const int nb = (int)C.boxes.size();
  int src = -1;
  if (hint < nb && C.boxes[hint].contains(x, y, z)) {
    src = hint;
  } else {
    for (int s = 0; s < nb; ++s)
      if (C.boxes[s].contains(x, y, z)) {
        src = s;
        break;
      }
  }
  if (src < 0) return false;
  hint = src;
  *ref = ((long long)src << 32) | grown_index(C.desc[src], x, y, z);
  return true;
}

// Mask bytes of a patched level, one per grown cell: 1 where a valid cell
// lies in coarsen(mask box, r) for some mask box (finer index space).
std::vector<unsigned char> patch_mask(const Level& C, const std::vector<Box>& mask, int r) {
  std::vector<unsigned char> m((size_t)C.grown_cells, 0);
  for (size_t b = 0; b < C.boxes.size(); ++b) {
    const PatchDesc& P = C.desc[b];
    for (const Box& mb : mask) {
      int lo[3], hi[3];
      bool empty = false;
      for (int d = 0; d < 3; ++d) {
        lo[d] = std::max(P.lo[d], floor_div(mb.lo[d], r));
        hi[d] = std::min(P.hi[d], floor_div(mb.hi[d], r));
        empty |= lo[d] > hi[d];
      }
      if (empty) continue;
      for (int z = lo[2]; z <= hi[2]; ++z)
        for (int y = lo[1]; y <= hi[1]; ++y)
          for (int x = lo[0]; x <= hi[0]; ++x) m[P.fill_base + grown_index(P, x, y, z)] = 1;
    }
  }
  return m;
}

// Level l >= 1: patches grown by 2; ghost-fill sources; average-down table
// into level l - 1.  Builds the level over `boxes` into F (a fresh Level); on
// failure F holds nothing.
int build_level(lbx_sim* sim, int level, const std::vector<Box>& boxes, Level& F) {
  lbx_ctx* ctx = sim->ctx;
  Level& C = sim->lev[level - 1];
  if (!C.defined) return fail(ctx, LBX_E_STATE, "define level %d before level %d", level - 1, level);
  if (sim->walls) return fail(ctx, LBX_E_UNSUPPORTED, "refined levels need a fully periodic domain");
  const int r = sim->ratio;
  long long scale = 1;
  for (int l = 0; l < level; ++l) scale *= r;
  int fdom[3];
  for (int d = 0; d < 3; ++d) {
    if ((long long)sim->dims[d] * scale > (1ll << 30))
      return fail(ctx, LBX_E_ARG, "level %d's index space exceeds 2^30 cells per direction", level);
    fdom[d] = (int)(sim->dims[d] * scale);
  }
  for (size_t b = 0; b < boxes.size(); ++b) {
    if (!box_ok(boxes[b], fdom))
      return fail(ctx, LBX_E_ARG, "level-%d box %zu lies outside the refined domain or is empty", level, b);
    for (int d = 0; d < 3; ++d)
      if (boxes[b].lo[d] % r != 0 || (boxes[b].hi[d] + 1) % r != 0)
        return fail(ctx, LBX_E_ARG, "level-%d box %zu is not aligned to refinement ratio %d", level, b, r);
  }
  if (any_overlap(boxes)) return fail(ctx, LBX_E_ARG, "level-%d boxes overlap", level);
  if (boxes.size() > 65535)  // one grid row of workgroups per patch
    return fail(ctx, LBX_E_ARG, "level %d has more than 65535 boxes", level);
  // Proper nesting: the parent of every valid cell lies in the parent
  // level's valid region (the parent of a ghost is checked with the fill).
  int chint = 0;
  if (C.patched)
    for (size_t b = 0; b < boxes.size(); ++b)
      for (int z = boxes[b].lo[2] / r; z <= boxes[b].hi[2] / r; ++z)
        for (int y = boxes[b].lo[1] / r; y <= boxes[b].hi[1] / r; ++y)
          for (int x = boxes[b].lo[0] / r; x <= boxes[b].hi[0] / r; ++x) {
            long long ref;
            if (!cell_ref(C, x, y, z, &ref, chint))
              return fail(ctx, LBX_E_ARG, "level-%d box %zu is not properly nested in level %d", level, b,
                          level - 1);
          }
  F = Level{};
  F.boxes = boxes;
  F.patched = true;
  F.ratio = r;
  // Average-down blocks per patch: the box coarsened by r and grown by
  // ng = 2 / r coarse cells (AMReX's sum_fine_to_coarse: the fine ghosts
  // over the ratio), c[d] blocks per direction.
  const int ng = 2 / r;
  long long base = 0, fill_base = 0, avg_base = 0;
  for (const Box& b : boxes) {
    PatchDesc d{};
    for (int k = 0; k < 3; ++k) {
      d.lo[k] = b.lo[k];
      d.hi[k] = b.hi[k];
      d.g[k] = b.n(k) + 4;
    }
    const long long cells = (long long)d.g[0] * d.g[1] * d.g[2];
    if (cells >= (1ll << 31) || fill_base + cells >= (1ll << 32)) {
      F = Level{};
      return fail(ctx, LBX_E_ARG, "level-%d box too large", level);
    }
    d.cells = (int)cells;
    d.qs = (cells + 63) / 64 * 64;
    d.fill_base = fill_base;
    d.avg_base = avg_base;
    const int blocks = (b.n(0) / r + 2 * ng) * (b.n(1) / r + 2 * ng) * (b.n(2) / r + 2 * ng);
    fill_base += cells;
    avg_base += blocks;
    F.max_cells = std::max(F.max_cells, d.cells);
    F.max_blocks = std::max(F.max_blocks, blocks);
    F.desc.push_back(d);
  }
  if (avg_base >= (1ll << 31)) {
    F = Level{};
    return fail(ctx, LBX_E_ARG, "level %d too large", level);
  }
  F.grown_cells = fill_base;
  // Storage order and patch groups (see Level::gdesc).
  {
    std::vector<int> order(boxes.size());
    for (size_t b = 0; b < boxes.size(); ++b) order[b] = (int)b;
    std::stable_sort(order.begin(), order.end(), [&](int a, int c) {
      const Box& A = boxes[a];
      const Box& B = boxes[c];
      return std::make_tuple(A.lo[2], A.lo[1], A.lo[0]) < std::make_tuple(B.lo[2], B.lo[1], B.lo[0]);
    });
    for (int b : order) {
      F.desc[b].base = base;
      base += F.desc[b].qs * lbx::NV_TOTAL;
    }
    for (int b : order) F.gdesc.push_back(F.desc[b]);
    const long long esz = (long long)elem_size(sim->precision);
    for (size_t k = 0; k < F.gdesc.size();) {
      long long elems = 0;
      size_t n = 0;
      while (k + n < F.gdesc.size()) {
        const long long pe = F.gdesc[k + n].qs * lbx::NV_TOTAL;
        if (n > 0 && (elems + pe) * esz > kFineGroupBytes) break;
        elems += pe;
        ++n;
      }
      F.groups.push_back({(int)k, (int)n});
      F.region_span = std::max(F.region_span, elems);
      k += n;
    }
    // per-stream regions of buf[2] for the array between the two substeps;
    // at a ratio above 2 the substeps chain through the whole buf[2]
    F.regions = r == 2 && F.groups.size() > 1 && 2 * F.region_span <= base;  // base: one level buffer's elements
  }
  // Ghost fill sources (FillPatchTwoLevels with piecewise-constant
  // interpolation, src/AmrSim.cpp:373-388): the valid cell of this level
  // holding the periodic image if a box has it, else the coarse parent.
  std::vector<long long> fill((size_t)fill_base, 0);
  int hint = 0;
  for (size_t b = 0; b < boxes.size(); ++b) {
    const PatchDesc& P = F.desc[b];
    for (int gz = 0; gz < P.g[2]; ++gz)
      for (int gy = 0; gy < P.g[1]; ++gy)
        for (int gx = 0; gx < P.g[0]; ++gx) {
          const int x = P.lo[0] - 2 + gx, y = P.lo[1] - 2 + gy, z = P.lo[2] - 2 + gz;
          if (boxes[b].contains(x, y, z)) continue;
          const int wx = wrap(x, fdom[0]), wy = wrap(y, fdom[1]), wz = wrap(z, fdom[2]);
          int src = -1;
          if (boxes[hint].contains(wx, wy, wz)) {
            src = hint;
          } else {
            for (size_t s = 0; s < boxes.size(); ++s)
              if (boxes[s].contains(wx, wy, wz)) {
                src = (int)s;
                break;
              }
          }
          long long d;
          if (src >= 0) {
            hint = src;
            d = ((long long)src << 32) | grown_index(F.desc[src], wx, wy, wz);
          } else {
            long long ref;
            if (!cell_ref(C, wx / r, wy / r, wz / r, &ref, chint)) {
              F = Level{};
              return fail(ctx, LBX_E_ARG, "level-%d box %zu is not properly nested in level %d", level, b,
                          level - 1);
            }
            d = -1 - ref;
          }
          fill[P.fill_base + ((long long)gz * P.g[1] + gy) * P.g[0] + gx] = d;
        }
  }
  // Average-down contributions (sum_fine_to_coarse, src/AmrSim.cpp:592-602):
  // each box's coarsened box grown by one coarse cell, periodic image in the
  // parent level, added only where a valid parent cell holds it; per parent
  // cell, contributions in box order and, within a box, in (i, j, k) loop
  // order.  A contribution is the average of one 2x2x2 block of a grown
  // patch, stored at avg_base + block index.
  struct Contrib {
    long long cell;  // parent cell reference
    int avg;         // block-average index
  };
  const int cdom[3] = {fdom[0] / r, fdom[1] / r, fdom[2] / r};
  std::vector<Contrib> list;
  for (size_t b = 0; b < boxes.size(); ++b) {
    const Box& B = boxes[b];
    const PatchDesc& P = F.desc[b];
    const int c0 = B.n(0) / r + 2 * ng, c1 = B.n(1) / r + 2 * ng;
    const int i0 = floor_div(B.lo[0], r) - ng, j0 = floor_div(B.lo[1], r) - ng, k0 = floor_div(B.lo[2], r) - ng;
    for (int ci = i0; ci <= floor_div(B.hi[0], r) + ng; ++ci)
      for (int cj = j0; cj <= floor_div(B.hi[1], r) + ng; ++cj)
        for (int ck = k0; ck <= floor_div(B.hi[2], r) + ng; ++ck) {
          long long ref;
          if (!cell_ref(C, wrap(ci, cdom[0]), wrap(cj, cdom[1]), wrap(ck, cdom[2]), &ref, chint)) continue;
          const long long blk = ((long long)(ck - k0) * c1 + (cj - j0)) * c0 + (ci - i0);
          list.push_back({ref, (int)(P.avg_base + blk)});
        }
  }
  std::stable_sort(list.begin(), list.end(), [](const Contrib& a, const Contrib& b) { return a.cell < b.cell; });
  std::vector<long long> cell;
  std::vector<int> start;
  std::vector<int> contrib;
  for (size_t j = 0; j < list.size(); ++j) {
    if (j == 0 || list[j].cell != list[j - 1].cell) {
      cell.push_back(list[j].cell);
      start.push_back((int)j);
    }
    contrib.push_back(list[j].avg);
  }
  start.push_back((int)list.size());
  F.sum_ncell = (int)cell.size();
  std::vector<int> slot;
  if (C.patched) {
    slot.assign((size_t)C.grown_cells, -1);
    for (size_t u = 0; u < cell.size(); ++u)
      slot[C.desc[cell[u] >> 32].fill_base + (cell[u] & 0xffffffffll)] = (int)u;
  }
  F.avg_stride = (avg_base + 63) / 64 * 64;
  const size_t bytes = (size_t)base * elem_size(sim->precision);
  {
    const size_t abytes = (size_t)F.avg_stride * lbx::NV_TOTAL * sizeof(double);
    hipError_t e = hipMalloc(&F.d_avg, abytes);
    if (e != hipSuccess) {
      free_level(F);
      return fail(ctx, LBX_E_NOMEM, "allocation of %zu bytes for level %d failed: %s", abytes, level,
                  hipGetErrorString(e));
    }
  }
  for (int i = 0; i < 3; ++i) {
    hipError_t e = hipMalloc(&F.buf[i], bytes);
    if (e == hipSuccess) e = hipMemsetAsync(F.buf[i], 0, bytes, ctx->stream);
    if (e != hipSuccess) {
      free_level(F);
      return fail(ctx, LBX_E_NOMEM, "allocation of %zu bytes for level %d failed: %s", bytes, level,
                  hipGetErrorString(e));
    }
  }
  hipError_t e = upload(&F.d_desc, F.desc, ctx->stream);
  if (e == hipSuccess) e = upload(&F.d_gdesc, F.gdesc, ctx->stream);
  if (e == hipSuccess) e = upload(&F.d_fill, fill, ctx->stream);
  if (e == hipSuccess && !C.patched) e = upload(&F.d_sum_cell, cell, ctx->stream);
  if (e == hipSuccess) e = upload(&F.d_sum_start, start, ctx->stream);
  if (e == hipSuccess) e = upload(&F.d_sum_contrib, contrib, ctx->stream);
  if (e == hipSuccess && C.patched) e = upload(&F.d_sum_slot, slot, ctx->stream);
  if (!C.patched) C.push_shape = lbx::default_shape(C.lat, sim->precision);
  if (e != hipSuccess) {
    free_level(F);
    return hip_fail(ctx, e, "level tables");
  }
  F.defined = true;
  return LBX_OK;
}

// Level `level`'s mask bytes with no cell masked (a level starts unmasked;
// on a level array they still carry the box-face bits of the coarse push).
// The device buffer is kept if there is one, so captured steps stay valid.
int unmask_level(lbx_sim* sim, int level) {
  lbx_ctx* ctx = sim->ctx;
  Level& C = sim->lev[level];
  if (!C.defined) return LBX_OK;
  const std::vector<unsigned char> bytes = C.patched ? patch_mask(C, {}, sim->ratio) : coarse_flags(C, {}, sim->ratio);
  if (!C.d_mask) LBX_HIP(ctx, hipMalloc(&C.d_mask, bytes.size()));
  LBX_HIP(ctx, hipMemcpyAsync(C.d_mask, bytes.data(), bytes.size(), hipMemcpyHostToDevice, ctx->stream));
  LBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LBX_OK;
}

// A level replaces any level at its index and discards the finer ones; its
// parent starts unmasked.
int define_level(lbx_sim* sim, int level, const std::vector<Box>& boxes) {
  Level nf;
  if (int rc = build_level(sim, level, boxes, nf)) return rc;
  free_levels_from(sim, level);
  sim->lev[level] = nf;
  return unmask_level(sim, level - 1);
}

// RemakeLevel (src/AmrSim.cpp:715-744) of the finest level: a new BoxArray
// whose valid cells take the old level's values where an old box covers
// them and the parent's otherwise (DistFnFillPatch: FillPatchTwoLevels with
// piecewise-constant interpolation, src/AmrSim.cpp:359-391).  Relaxation
// rates carry over.
int remake_level(lbx_sim* sim, int level, const std::vector<Box>& boxes) {
  lbx_ctx* ctx = sim->ctx;
  Level& C = sim->lev[level - 1];
  Level& F = sim->lev[level];
  if (!F.defined) return fail(ctx, LBX_E_STATE, "level %d is not defined", level);
  if (finest_level(sim) != level)
    return fail(ctx, LBX_E_STATE, "level %d has finer levels; clear them before remaking it", level);
  Level nf;
  if (int rc = build_level(sim, level, boxes, nf)) return rc;
  // Source of every new valid cell: (old patch << 32) | old grown index, or
  // -1 - parent reference.
  std::vector<long long> src((size_t)nf.grown_cells, 0);
  int chint = 0;
  for (size_t b = 0; b < boxes.size(); ++b) {
    const PatchDesc& P = nf.desc[b];
    for (int z = P.lo[2]; z <= P.hi[2]; ++z)
      for (int y = P.lo[1]; y <= P.hi[1]; ++y)
        for (int x = P.lo[0]; x <= P.hi[0]; ++x) {
          long long ref = 0;
          (void)cell_ref(C, x / nf.ratio, y / nf.ratio, z / nf.ratio, &ref, chint);  // nesting checked by build_level
          src[P.fill_base + grown_index(P, x, y, z)] = -1 - ref;
        }
    for (size_t o = 0; o < F.boxes.size(); ++o) {
      const PatchDesc& O = F.desc[o];
      int lo[3], hi[3];
      bool empty = false;
      for (int d = 0; d < 3; ++d) {
        lo[d] = std::max(P.lo[d], O.lo[d]);
        hi[d] = std::min(P.hi[d], O.hi[d]);
        empty |= lo[d] > hi[d];
      }
      if (empty) continue;
      for (int z = lo[2]; z <= hi[2]; ++z)
        for (int y = lo[1]; y <= hi[1]; ++y)
          for (int x = lo[0]; x <= hi[0]; ++x)
            src[P.fill_base + grown_index(P, x, y, z)] = ((long long)o << 32) | grown_index(O, x, y, z);
    }
  }
  DeviceTemp t;
  hipError_t e = hipMalloc(&t.p, src.size() * sizeof(long long));
  if (e == hipSuccess)
    e = hipMemcpyAsync(t.p, src.data(), src.size() * sizeof(long long), hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess)
    e = lbx::launch_fine_remake(sim->precision, nf.buf[0], F.buf[F.cur], C.buf[C.cur], C.lat,
                                C.patched ? C.d_desc : nullptr, nf.d_desc, (int)nf.desc.size(), nf.max_cells,
                                F.d_desc, (const long long*)t.p, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) {
    free_level(nf);
    return hip_fail(ctx, e, "level remake");
  }
  nf.omega_s = F.omega_s;
  nf.omega_b = F.omega_b;
  nf.relax_set = F.relax_set;
  free_level(F);
  F = nf;
  return LBX_OK;
}

// DistFnFillFromCoarse (src/AmrSim.cpp:393-411) for a level whose parent is
// patched: every valid cell copies its parent through the remake kernel.
int fill_from_patched(lbx_sim* sim, int level) {
  lbx_ctx* ctx = sim->ctx;
  const Level& C = sim->lev[level - 1];
  Level& F = sim->lev[level];
  std::vector<long long> src((size_t)F.grown_cells, 0);
  int chint = 0;
  for (size_t b = 0; b < F.boxes.size(); ++b) {
    const PatchDesc& P = F.desc[b];
    for (int z = P.lo[2]; z <= P.hi[2]; ++z)
      for (int y = P.lo[1]; y <= P.hi[1]; ++y)
        for (int x = P.lo[0]; x <= P.hi[0]; ++x) {
          long long ref;
          if (!cell_ref(C, x / F.ratio, y / F.ratio, z / F.ratio, &ref, chint))
            return fail(ctx, LBX_E_STATE, "level %d is not nested in level %d", level, level - 1);
          src[P.fill_base + grown_index(P, x, y, z)] = -1 - ref;
        }
  }
  DeviceTemp t;
  LBX_HIP(ctx, hipMalloc(&t.p, src.size() * sizeof(long long)));
  LBX_HIP(ctx, hipMemcpyAsync(t.p, src.data(), src.size() * sizeof(long long), hipMemcpyHostToDevice, ctx->stream));
  LBX_HIP(ctx, lbx::launch_fine_remake(sim->precision, F.buf[F.cur], nullptr, C.buf[C.cur], C.lat, C.d_desc,
                                       F.d_desc, (int)F.boxes.size(), F.max_cells, F.d_desc, (const long long*)t.p,
                                       ctx->stream));
  LBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LBX_OK;
}

// FP64 sums over a level's valid cells, enqueued on the context's stream
// with the reduction's partials in `partials` (room for max_partials
// partials of 4 doubles) and the 4 results written to `res` (device
// memory): a level array in kLatticePartials workgroups, gathering through
// the pull when it holds the post-collision state; a patched level in one
// launch over all patches.  The partial layout depends only on the level, so
// any partials buffer gives the same bits.
int enqueue_conservation(lbx_sim* sim, int level, double* partials, long long max_partials, double* res) {
  lbx_ctx* ctx = sim->ctx;
  Level& L = sim->lev[level];
  if (!L.patched) {
    if (kLatticePartials > max_partials)
      return fail(ctx, LBX_E_ARG, "conservation needs %d partial sums, the buffer holds %lld", kLatticePartials,
                  max_partials);
    LBX_HIP(ctx, lbx::launch_conservation(sim->precision, L.post, L.buf[L.cur], L.lat, sim->rho0, partials,
                                          kLatticePartials, res, ctx->stream));
  } else {
    const int np = (int)L.boxes.size();
    const int chunks = patch_chunks(np);
    if ((long long)np * chunks > max_partials)
      return fail(ctx, LBX_E_ARG, "conservation of level %d (%d boxes) needs %lld partial sums, the buffer holds %lld",
                  level, np, (long long)np * chunks, max_partials);
    LBX_HIP(ctx, lbx::launch_patch_conservation(sim->precision, L.buf[L.cur], L.d_desc, np, chunks, sim->rho0,
                                                partials, res, ctx->stream));
  }
  return LBX_OK;
}

// lbx_conservation: partials and results in the context's own buffer, then
// a synchronous copy of the results to the host.
int conservation_sums(lbx_sim* sim, int level, double* buf, long long max_partials, double out[4]) {
  lbx_ctx* ctx = sim->ctx;
  double* res = buf + 4 * max_partials;
  if (int rc = enqueue_conservation(sim, level, buf, max_partials, res)) return rc;
  LBX_HIP(ctx, hipMemcpyAsync(out, res, 4 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  LBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LBX_OK;
}

// Device memory of lbx_conservation_noalloc: the reduction's partials live in
// this array, one copy per GPU, held until the process ends.  Host code never
// passes `scratch` itself to a kernel (in host code the name is a stand-in,
// not a device address); it uses the address hipGetSymbolAddress returns for
// the device, looked up once per device at context creation.
constexpr long long kNoAllocPartials = 65536;  // 2 MiB of partial sums
constexpr int kMaxDevices = 64;
static __device__ __align__(256) double scratch[kNoAllocPartials * 4];
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// The calling thread's current device must be `device`.
hipError_t scratch_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return hipSuccess;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e == hipSuccess) g_scratch_ptr[device] = (double*)p;
  return e;
}

}  // namespace

int lbx::ctx_device(const lbx_ctx* ctx) { return ctx->device; }
hipStream_t lbx::ctx_stream(const lbx_ctx* ctx) { return ctx->stream; }

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
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->side, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->fork_ev, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->join_ev, hipEventDisableTiming);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->push, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->push_fork_ev, hipEventDisableTiming);
  if (e == hipSuccess) e = hipEventCreateWithFlags(&ctx->push_join_ev, hipEventDisableTiming);
  if (e == hipSuccess) e = hipMalloc(&ctx->sums, (kPartials * 4 + 4) * sizeof(double));
  if (e == hipSuccess) e = scratch_setup(device);
  if (e != hipSuccess) {
    hip_fail(nullptr, e, "context set-up");
    if (ctx->sums) (void)hipFree(ctx->sums);
    if (ctx->push_join_ev) (void)hipEventDestroy(ctx->push_join_ev);
    if (ctx->push_fork_ev) (void)hipEventDestroy(ctx->push_fork_ev);
    if (ctx->push) (void)hipStreamDestroy(ctx->push);
    if (ctx->join_ev) (void)hipEventDestroy(ctx->join_ev);
    if (ctx->fork_ev) (void)hipEventDestroy(ctx->fork_ev);
    if (ctx->side) (void)hipStreamDestroy(ctx->side);
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
  (void)hipStreamSynchronize(ctx->side);
  (void)hipStreamSynchronize(ctx->push);
  if (ctx->sums) (void)hipFree(ctx->sums);
  if (ctx->stage) (void)hipFree(ctx->stage);
  (void)hipEventDestroy(ctx->join_ev);
  (void)hipEventDestroy(ctx->fork_ev);
  (void)hipStreamDestroy(ctx->side);
  (void)hipEventDestroy(ctx->push_join_ev);
  (void)hipEventDestroy(ctx->push_fork_ev);
  (void)hipStreamDestroy(ctx->push);
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
  long long cells = 1;
  for (int d = 0; d < 3; ++d) {
    if (dims[d] < 1 || dims[d] > (1 << 28))
      return fail(ctx, LBX_E_ARG, "domain extent %d in direction %d", dims[d], d);
    cells *= dims[d];  // < 2^31 after every factor, so no overflow
    if (cells >= (1ll << 31)) return fail(ctx, LBX_E_ARG, "domain exceeds 2^31 cells");
  }
  if (precision != LBX_FP64 && precision != LBX_FP32_SHIFTED)
    return fail(ctx, LBX_E_ARG, "unknown precision %d", precision);
  if (!std::isfinite(rho0)) return fail(ctx, LBX_E_ARG, "rho0 must be finite");
  lbx_sim* sim = new (std::nothrow) lbx_sim;
  if (!sim) return fail(ctx, LBX_E_NOMEM, "out of host memory");
  sim->ctx = ctx;
  for (int d = 0; d < 3; ++d) {
    sim->dims[d] = dims[d];
    if (!periodic[d]) sim->walls |= 1u << d;
  }
  sim->precision = precision;
  sim->rho0 = rho0;
  *out = sim;
  return LBX_OK;
}

int lbx_sim_destroy(lbx_sim* sim) {
  if (!sim) return LBX_OK;
  (void)hipSetDevice(sim->ctx->device);
  (void)hipStreamSynchronize(sim->ctx->stream);
  sim->graphs.clear();
  for (Level& L : sim->lev) free_level(L);
  delete sim;
  return LBX_OK;
}

int lbx_level_define(lbx_sim* sim, int level, int n_boxes, const int32_t* lohi) {
  if (int rc = check_sim(sim)) return rc;
  lbx_ctx* ctx = sim->ctx;
  if (level < 0 || level >= LBX_MAX_LEVELS)
    return fail(ctx, LBX_E_UNSUPPORTED, "levels 0 to %d are supported", LBX_MAX_LEVELS - 1);
  if (n_boxes < 1 || !lohi) return fail(ctx, LBX_E_ARG, "need at least one box");
  LBX_HIP(ctx, hipSetDevice(ctx->device));
  LBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const std::vector<Box> boxes = read_boxes(n_boxes, lohi);
  sim->graphs.clear();
  if (level == 0) return define_level0(sim, boxes);
  if (int rc = materialise(sim)) return rc;  // the Rohde cycle keeps level 0 pre-collision
  return define_level(sim, level, boxes);
}

int lbx_level_remake(lbx_sim* sim, int level, int n_boxes, const int32_t* lohi) {
  if (int rc = check_sim(sim)) return rc;
  lbx_ctx* ctx = sim->ctx;
  if (level < 1 || level >= LBX_MAX_LEVELS) return fail(ctx, LBX_E_UNSUPPORTED, "level %d cannot be remade", level);
  if (n_boxes < 1 || !lohi) return fail(ctx, LBX_E_ARG, "need at least one box");
  LBX_HIP(ctx, hipSetDevice(ctx->device));
  LBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const std::vector<Box> boxes = read_boxes(n_boxes, lohi);
  sim->graphs.clear();
  if (int rc = materialise(sim)) return rc;
  return remake_level(sim, level, boxes);
}

int lbx_tag_cells(lbx_sim* sim, int level, int criterion, double threshold, uint8_t* tags) {
  if (int rc = check_sim(sim)) return rc;
  lbx_ctx* ctx = sim->ctx;
  if (level != 0) return fail(ctx, LBX_E_UNSUPPORTED, "cells of level 0 are tagged");
  if (!tags) return fail(ctx, LBX_E_ARG, "null tag buffer");
  if (criterion != LBX_TAG_DENSITY_GRADIENT && criterion != LBX_TAG_SPEED)
    return fail(ctx, LBX_E_ARG, "unknown tagging criterion %d", criterion);
  if (!std::isfinite(threshold) || threshold < 0.0) return fail(ctx, LBX_E_ARG, "threshold must be finite, >= 0");
  const Level& L = sim->lev[0];
  if (!L.defined) return fail(ctx, LBX_E_STATE, "level 0 is not defined");
  LBX_HIP(ctx, hipSetDevice(ctx->device));
  const size_t n = (size_t)L.lat.nx * L.lat.ny * L.lat.nz;
  void* t = nullptr;
  LBX_HIP(ctx, stage_buffer(ctx, n, &t));
  LBX_HIP(ctx, lbx::launch_tag(sim->precision, L.post, criterion, L.buf[L.cur], L.lat, sim->rho0,
                               threshold * threshold, (unsigned char*)t, ctx->stream));
  LBX_HIP(ctx, hipMemcpyAsync(tags, t, n, hipMemcpyDeviceToHost, ctx->stream));
  LBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LBX_OK;
}

int lbx_level_clear(lbx_sim* sim, int level) {
  if (int rc = check_sim(sim)) return rc;
  if (level < 0 || level >= LBX_MAX_LEVELS) return fail(sim->ctx, LBX_E_ARG, "level %d out of range", level);
  LBX_HIP(sim->ctx, hipSetDevice(sim->ctx->device));
  (void)hipStreamSynchronize(sim->ctx->stream);
  sim->graphs.clear();
  free_levels_from(sim, level);
  // the new finest level has no finer level to mask for
  return level > 0 ? unmask_level(sim, level - 1) : LBX_OK;
}

int lbx_set_ref_ratio(lbx_sim* sim, int ratio) {
  if (int rc = check_sim(sim)) return rc;
  if (ratio < 2 || ratio > 4) return fail(sim->ctx, LBX_E_ARG, "refinement ratio %d is not 2, 3 or 4", ratio);
  if (sim->lev[1].defined)
    return fail(sim->ctx, LBX_E_STATE, "the refinement ratio is set before level 1 is defined");
  sim->ratio = ratio;
  return LBX_OK;
}

int lbx_set_relax(lbx_sim* sim, int level, double omega_s, double omega_b) {
  if (int rc = check_sim(sim)) return rc;
  if (level < 0 || level >= LBX_MAX_LEVELS) return fail(sim->ctx, LBX_E_ARG, "level %d out of range", level);
  if (!std::isfinite(omega_s) || !std::isfinite(omega_b))
    return fail(sim->ctx, LBX_E_ARG, "relaxation rates must be finite");
  sim->lev[level].omega_s = omega_s;
  sim->lev[level].omega_b = omega_b;
  sim->lev[level].relax_set = true;
  sim->graphs.clear();
  return LBX_OK;
}

int lbx_set_level_mask(lbx_sim* sim, int level, int n_boxes, const int32_t* lohi_fine) {
  if (int rc = check_sim(sim)) return rc;
  lbx_ctx* ctx = sim->ctx;
  if (level < 0 || level >= LBX_MAX_LEVELS) return fail(ctx, LBX_E_ARG, "level %d out of range", level);
  Level& C = sim->lev[level];
  if (!C.defined) return fail(ctx, LBX_E_STATE, "level %d is not defined", level);
  if (n_boxes < 0 || (n_boxes > 0 && !lohi_fine)) return fail(ctx, LBX_E_ARG, "bad mask boxes");
  const std::vector<Box> mask = read_boxes(n_boxes, lohi_fine);
  const std::vector<unsigned char> bytes = C.patched ? patch_mask(C, mask, sim->ratio) : coarse_flags(C, mask, sim->ratio);
  LBX_HIP(ctx, hipSetDevice(ctx->device));
  LBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (!C.d_mask) {
    sim->graphs.clear();  // captured steps hold the mask's address
    LBX_HIP(ctx, hipMalloc(&C.d_mask, bytes.size()));
  }
  LBX_HIP(ctx, hipMemcpyAsync(C.d_mask, bytes.data(), bytes.size(), hipMemcpyHostToDevice, ctx->stream));
  LBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LBX_OK;
}

int lbx_set_fine_mask(lbx_sim* sim, int n_boxes, const int32_t* lohi_fine) {
  return lbx_set_level_mask(sim, 0, n_boxes, lohi_fine);
}

int lbx_fill_from_coarse(lbx_sim* sim, int level) {
  if (int rc = check_sim(sim)) return rc;
  lbx_ctx* ctx = sim->ctx;
  if (level < 1 || level >= LBX_MAX_LEVELS)
    return fail(ctx, LBX_E_ARG, "level %d has no coarser level to fill from", level);
  Level& C = sim->lev[level - 1];
  Level& F = sim->lev[level];
  if (!C.defined || !F.defined) return fail(ctx, LBX_E_STATE, "levels %d and %d must be defined", level - 1, level);
  LBX_HIP(ctx, hipSetDevice(ctx->device));
  if (int rc = materialise(sim)) return rc;
  if (C.patched) return fill_from_patched(sim, level);
  LBX_HIP(ctx, lbx::launch_fine_from_coarse(sim->precision, F.buf[F.cur], C.buf[C.cur], C.lat, F.d_desc,
                                            (int)F.boxes.size(), F.max_cells, sim->rho0, F.ratio, ctx->stream));
  LBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LBX_OK;
}

int lbx_init_equilibrium(lbx_sim* sim, int level, int box, const double* rho, const double* u) {
  if (int rc = check_sim(sim)) return rc;
  if (int rc = check_box(sim, level, box)) return rc;
  if (!rho || !u) return fail(sim->ctx, LBX_E_ARG, "null field buffer");
  lbx_ctx* ctx = sim->ctx;
  Level& L = sim->lev[level];
  if (L.patched)
    return fail(ctx, LBX_E_UNSUPPORTED, "level %d is filled from the coarser level", level);
  const Region R = region_of(L.boxes[box]);
  const size_t n = (size_t)L.boxes[box].cells();
  LBX_HIP(ctx, hipSetDevice(ctx->device));
  if (int rc = materialise(sim)) return rc;
  void* t = nullptr;
  LBX_HIP(ctx, stage_buffer(ctx, n * 4 * sizeof(double), &t));
  double* drho = (double*)t;
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
  LBX_HIP(ctx, hipSetDevice(ctx->device));
  if (level == 0)
    if (int rc = materialise(sim)) return rc;
  void* t = nullptr;
  LBX_HIP(ctx, stage_buffer(ctx, n * lbx::NV_TOTAL * sizeof(double), &t));
  LBX_HIP(ctx, hipMemcpyAsync(t, f, n * lbx::NV_TOTAL * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
  if (L.patched)
    LBX_HIP(ctx, lbx::launch_patch_io(sim->precision, 1, L.buf[L.cur], L.d_desc, box, (int)n, (double*)t,
                                      nullptr, sim->rho0, ctx->stream));
  else
    LBX_HIP(ctx, lbx::launch_unpack_f(sim->precision, L.buf[L.cur], L.lat, region_of(L.boxes[box]),
                                      (const double*)t, sim->rho0, ctx->stream));
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
  LBX_HIP(ctx, hipSetDevice(ctx->device));
  void* t = nullptr;
  LBX_HIP(ctx, stage_buffer(ctx, n * lbx::NV_TOTAL * sizeof(double), &t));
  if (L.patched)
    LBX_HIP(ctx, lbx::launch_patch_io(sim->precision, 0, L.buf[L.cur], L.d_desc, box, (int)n, (double*)t,
                                      nullptr, sim->rho0, ctx->stream));
  else
    LBX_HIP(ctx, lbx::launch_pack_f(sim->precision, L.post, L.buf[L.cur], L.lat, region_of(L.boxes[box]),
                                    (double*)t, sim->rho0, ctx->stream));
  LBX_HIP(ctx, hipMemcpyAsync(f, t, n * lbx::NV_TOTAL * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
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
  if (nsteps == 0) return LBX_OK;
  LBX_HIP(ctx, hipSetDevice(ctx->device));
  if (const int finest = finest_level(sim)) {
    for (int l = 1; l <= finest; ++l)
      if (!sim->lev[l].relax_set) return fail(ctx, LBX_E_STATE, "relaxation rates of level %d are not set", l);
    if (int rc = materialise(sim)) return rc;
    for (int s = 0; s < nsteps; ++s)
      if (int rc = launch_hierarchy_step(sim)) return rc;
    return LBX_OK;
  }
  // Reference order per step is f <- S(C(f)).  Over n steps:
  //   f_n = S (C S)^(n-1) C f_0 = S g_n,  g_n = (C S)^(n-1) C f_0
  // so from f in place: one collide-only pass and n-1 fused pull+collide
  // passes leave g_n (Level::post); from g: n fused passes.  The last S is
  // applied only when f is observed (a pull inside the readers, or the
  // stream-only pass of materialise()).  Per-cell arithmetic is unchanged,
  // so results are bit-identical to the reference order, and Iterate(a)
  // then Iterate(b) is Iterate(a + b) pass for pass.
  int s = 0;
  if (!L.post) {
    if (int rc = launch_step(sim, L, lbx::STEP_COLLIDE)) return rc;
    L.post = true;
    s = 1;
  }
  for (; s < nsteps; ++s)
    if (int rc = launch_step(sim, L, lbx::STEP_FUSED)) return rc;
  return LBX_OK;
}

int lbx_hydro(lbx_sim* sim, int level, int box, double* rho, double* u) {
  if (int rc = check_sim(sim)) return rc;
  if (int rc = check_box(sim, level, box)) return rc;
  if (!rho || !u) return fail(sim->ctx, LBX_E_ARG, "null buffer");
  lbx_ctx* ctx = sim->ctx;
  Level& L = sim->lev[level];
  const size_t n = (size_t)L.boxes[box].cells();
  LBX_HIP(ctx, hipSetDevice(ctx->device));
  void* t = nullptr;
  LBX_HIP(ctx, stage_buffer(ctx, n * 4 * sizeof(double), &t));
  double* drho = (double*)t;
  double* du = drho + n;
  if (L.patched)
    LBX_HIP(ctx, lbx::launch_patch_io(sim->precision, 2, L.buf[L.cur], L.d_desc, box, (int)n, drho, du,
                                      sim->rho0, ctx->stream));
  else
    LBX_HIP(ctx, lbx::launch_hydro(sim->precision, L.post, L.buf[L.cur], L.lat, region_of(L.boxes[box]), drho,
                                   du, sim->rho0, ctx->stream));
  LBX_HIP(ctx, hipMemcpyAsync(rho, drho, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  LBX_HIP(ctx, hipMemcpyAsync(u, du, 3 * n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  LBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LBX_OK;
}

int lbx_hydro_level(lbx_sim* sim, int level, double* rho, double* u) {
  if (int rc = check_sim(sim)) return rc;
  lbx_ctx* ctx = sim->ctx;
  if (level < 0 || level >= LBX_MAX_LEVELS) return fail(ctx, LBX_E_ARG, "level %d out of range", level);
  Level& L = sim->lev[level];
  if (!L.defined) return fail(ctx, LBX_E_STATE, "level %d is not defined", level);
  if (!rho || !u) return fail(ctx, LBX_E_ARG, "null buffer");
  LBX_HIP(ctx, hipSetDevice(ctx->device));
  size_t total = 0;
  for (const Box& b : L.boxes) total += (size_t)b.cells();
  void* t = nullptr;
  LBX_HIP(ctx, stage_buffer(ctx, total * 4 * sizeof(double), &t));
  double* drho = (double*)t;
  double* du = drho + total;
  size_t off = 0;
  for (size_t b = 0; b < L.boxes.size(); ++b) {
    const size_t n = (size_t)L.boxes[b].cells();
    if (L.patched)
      LBX_HIP(ctx, lbx::launch_patch_io(sim->precision, 2, L.buf[L.cur], L.d_desc, (int)b, (int)n, drho + off,
                                        du + 3 * off, sim->rho0, ctx->stream));
    else
      LBX_HIP(ctx, lbx::launch_hydro(sim->precision, L.post, L.buf[L.cur], L.lat, region_of(L.boxes[b]),
                                     drho + off, du + 3 * off, sim->rho0, ctx->stream));
    off += n;
  }
  LBX_HIP(ctx, hipMemcpyAsync(rho, drho, total * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  LBX_HIP(ctx, hipMemcpyAsync(u, du, 3 * total * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  LBX_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return LBX_OK;
}

int lbx_conservation(lbx_sim* sim, int level, double out[4]) {
  if (int rc = check_sim(sim)) return rc;
  if (!out) return fail(sim->ctx, LBX_E_ARG, "null buffer");
  if (level < 0 || level >= LBX_MAX_LEVELS || !sim->lev[level].defined)
    return fail(sim->ctx, LBX_E_STATE, "level %d is not defined", level);
  LBX_HIP(sim->ctx, hipSetDevice(sim->ctx->device));
  return conservation_sums(sim, level, sim->ctx->sums, kPartials, out);
}

int lbx_conservation_noalloc(lbx_sim* sim, int level, double* d_out) {
  if (int rc = check_sim(sim)) return rc;
  lbx_ctx* ctx = sim->ctx;
  if (!d_out) return fail(ctx, LBX_E_ARG, "null output");
  if (level < 0 || level >= LBX_MAX_LEVELS || !sim->lev[level].defined)
    return fail(ctx, LBX_E_STATE, "level %d is not defined", level);
  LBX_HIP(ctx, hipSetDevice(ctx->device));
  double* partials = nullptr;
  if (ctx->device >= 0 && ctx->device < kMaxDevices) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    partials = g_scratch_ptr[ctx->device];
  }
  if (!partials) return fail(ctx, LBX_E_STATE, "no scratch address for device %d", ctx->device);
  return enqueue_conservation(sim, level, partials, kNoAllocPartials, d_out);
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

int lbx_set_autotune(lbx_ctx* ctx, int enable) {
  if (!ctx) return fail(nullptr, LBX_E_ARG, "null context");
  ctx->autotune = enable != 0;
  return LBX_OK;
}

int lbx_tuned_shape(lbx_ctx* ctx, int nx, int ny, int precision, int32_t* bx, int32_t* by,
                    int32_t* wg_per_cu, float* ms_per_launch) {
  if (!ctx || !bx || !by) return fail(ctx, LBX_E_ARG, "null argument");
  if (nx < 1 || ny < 1 || (precision != LBX_FP64 && precision != LBX_FP32_SHIFTED))
    return fail(ctx, LBX_E_ARG, "bad geometry or precision");
  auto it = ctx->shapes.find(std::make_tuple(nx, ny, precision));
  if (it == ctx->shapes.end()) {  // never tuned: launches use the default shape
    LBX_HIP(ctx, hipSetDevice(ctx->device));
    const LaunchShape d = lbx::default_shape(lbx::make_lattice(nx, ny, 1, (int)elem_size(precision)), precision);
    *bx = d.bx;
    *by = d.by;
    if (wg_per_cu) *wg_per_cu = d.cap;
    if (ms_per_launch) *ms_per_launch = 0.0f;
    return LBX_OK;
  }
  *bx = it->second.bx;
  *by = it->second.by;
  if (wg_per_cu) *wg_per_cu = it->second.cap;
  if (ms_per_launch) {
    *ms_per_launch = 0.0f;
    for (auto& t : ctx->trials[it->first])
      if (t.first.bx == it->second.bx && t.first.by == it->second.by && t.first.cap == it->second.cap)
        *ms_per_launch = t.second;
  }
  return LBX_OK;
}

int lbx_bench_copy(lbx_ctx* ctx, size_t bytes, int reps, double* gbs) {
  if (!ctx || !gbs || reps < 1 || bytes < 8) return fail(ctx, LBX_E_ARG, "bad copy benchmark arguments");
  LBX_HIP(ctx, hipSetDevice(ctx->device));
  const long long n = (long long)(bytes / 16) * 2;  // whole 16-byte elements
  DeviceTemp a, b;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  LBX_HIP(ctx, hipMalloc(&a.p, n * 8));
  LBX_HIP(ctx, hipMalloc(&b.p, n * 8));
  LBX_HIP(ctx, hipMemsetAsync(a.p, 0, n * 8, ctx->stream));
  LBX_HIP(ctx, lbx::launch_copy((const double*)a.p, (double*)b.p, n, ctx->stream));  // warm-up
  LBX_HIP(ctx, hipEventCreate(&e0));
  hipError_t e = hipEventCreate(&e1);
  if (e != hipSuccess) {
    (void)hipEventDestroy(e0);
    return hip_fail(ctx, e, "hipEventCreate");
  }
  e = hipEventRecord(e0, ctx->stream);
  for (int r = 0; r < reps && e == hipSuccess; ++r)
    e = lbx::launch_copy((const double*)a.p, (double*)b.p, n, ctx->stream);
  if (e == hipSuccess) e = hipEventRecord(e1, ctx->stream);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (e != hipSuccess) return hip_fail(ctx, e, "copy benchmark");
  *gbs = 2.0 * (double)n * 8.0 * reps / (ms * 1e-3) / 1e9;
  return LBX_OK;
}

}  // extern "C"
