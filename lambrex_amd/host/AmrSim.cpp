// AmrSim.cpp -- the reference's AmrSim API (src/AmrSim.cpp) as a facade over
// the C ABI; all distribution-function work happens on the GPU.
#include "AmrSim.h"

#include <sys/stat.h>

#include <algorithm>
#include <cerrno>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <limits>
#include <sstream>
#include <string>

namespace {
// ratio 2 on every level (src/AmrSim.cpp:766) unless SetRefRatio sets
// another one for every level
constexpr int kRefRatio = 2;
}

void AmrSim::SetRefRatio(int ratio) {
  if (sim_) throw lambrex::Error("SetRefRatio must be called before InitFromScratch");
  if (ratio < 2 || ratio > 4) throw lambrex::Error("SetRefRatio: the ratio must be 2, 3 or 4");
  ref_ratio_ = ratio;
}

// Ratio between level `level` and level `level` + 1.
int AmrSim::Ratio(int level) const {
  (void)level;
  return ref_ratio_;
}

AmrSim::AmrSim(int nx, int ny, int nz, int max_ref_level,
               const std::array<int, NDIMS>& periodicity, double tau_s, double tau_b)
    : NX(nx), NY(ny), NZ(nz), NUMEL(nx * ny * nz), PERIODICITY(periodicity),
      max_level_(max_ref_level) {
  if (nx < 1 || ny < 1 || nz < 1) throw lambrex::Error("domain extents must be positive");
  if (max_ref_level < 0) throw lambrex::Error("max_ref_level must be >= 0");
  if (max_ref_level >= LBX_MAX_LEVELS)
    throw lambrex::Error("max_ref_level must be below " + std::to_string(LBX_MAX_LEVELS));
  // The reference aborts on a non-periodic direction (src/AmrSim.cpp:788-797);
  // here such a direction gets halfway bounce-back walls (single level only).
  const int num_levels = max_level_ + 1;
  tau_s_.assign(num_levels, 0.0);
  tau_b_.assign(num_levels, 0.0);
  mass_.assign(num_levels, 0.0);
  levels_.resize(num_levels);
  static_tags_.assign(num_levels, lambrex::IBox{});
  tau_s_.at(0) = tau_s;
  tau_b_.at(0) = tau_b;
}

AmrSim::~AmrSim() {
  if (sim_) lbx_sim_destroy(sim_);
}

void AmrSim::Check(int rc) const {
  if (rc != LBX_OK) {
    const char* msg = lbx_last_error(lambrex::Context());
    throw lambrex::Error(std::string("lbx: ") + (msg && *msg ? msg : "error") + " (code " +
                         std::to_string(rc) + ")");
  }
}

void AmrSim::SetMaxGridSize(int max_grid_size) {
  if (max_grid_size < 1) throw lambrex::Error("max_grid_size must be >= 1");
  max_grid_size_ = max_grid_size;
}

void AmrSim::SetPrecision(int precision, double rho0) {
  if (sim_) throw lambrex::Error("SetPrecision must be called before InitFromScratch");
  if (precision != LBX_FP64 && precision != LBX_FP32_SHIFTED)
    throw lambrex::Error("unknown precision");
  precision_ = precision;
  rho0_ = rho0;
}

// src/AmrSim.cpp:804-822
void AmrSim::SetInitialDensity(double rho_init) { initial_density.assign(NUMEL, rho_init); }
void AmrSim::SetInitialDensity(std::vector<double> rho_init) { initial_density = std::move(rho_init); }
void AmrSim::SetInitialVelocity(double u_init) { initial_velocity.assign(3 * NUMEL, u_init); }
void AmrSim::SetInitialVelocity(std::vector<double> u_init) { initial_velocity = std::move(u_init); }

std::vector<std::array<int, NDIMS>> AmrSim::refRatio() const {
  std::vector<std::array<int, NDIMS>> out;
  for (int l = 0; l <= max_level_; ++l) out.push_back({Ratio(l), Ratio(l), Ratio(l)});
  return out;
}

std::array<int, NDIMS> AmrSim::refRatio(int level) const {
  if (level < 0 || level > max_level_) throw lambrex::Error("refRatio: level out of range");
  return {Ratio(level), Ratio(level), Ratio(level)};
}

// src/AmrSim.cpp:297-322
void AmrSim::ComputeDt(int level) {
  LevelHost& L = levels_.at(level);
  if (level) {
    const LevelHost& C = levels_.at(level - 1);
    const int r = Ratio(level - 1);
    L.dt = C.dt / r;
    mass_.at(level) = mass_.at(level - 1) / r;
    tau_s_.at(level) = r * (tau_s_.at(level - 1) - 0.5) + 0.5;
    tau_b_.at(level) = r * (tau_b_.at(level - 1) - 0.5) + 0.5;
  } else {
    L.dt = 1.0;
    mass_.at(level) = 1.0;
  }
}

long long AmrSim::LocalIndex(const lambrex::IBox& b, int i, int j, int k) const {
  return ((long long)(k - b.lo[2]) * b.length(1) + (j - b.lo[1])) * b.length(0) + (i - b.lo[0]);
}

// The reference searches the FABs linearly for every point
// (src/AmrSim.cpp:824-843); the box of the previous query is tried first, so
// scans over neighbouring points cost O(1).
int AmrSim::BoxOf(int level, int i, int j, int k) const {
  const LevelHost& L = levels_.at(level);
  if (box_hint_.size() != levels_.size()) box_hint_.assign(levels_.size(), 0);
  int& h = box_hint_[level];
  if (h < (int)L.boxes.size() && L.boxes[h].contains(i, j, k)) return h;
  for (size_t b = 0; b < L.boxes.size(); ++b)
    if (L.boxes[b].contains(i, j, k)) {
      h = (int)b;
      return h;
    }
  return -1;
}

// src/AmrSim.cpp:138-215: C-order density into the level's boxes.
void AmrSim::InitDensity(int level) {
  if (level) throw lambrex::Error("Only level 0 should be initialised from scratch currently.");
  LevelHost& L = levels_.at(level);
  for (size_t b = 0; b < L.boxes.size(); ++b) {
    const lambrex::IBox& B = L.boxes[b];
    auto& rho = L.rho[L.cur][b];
    for (int k = B.lo[2]; k <= B.hi[2]; ++k)
      for (int j = B.lo[1]; j <= B.hi[1]; ++j)
        for (int i = B.lo[0]; i <= B.hi[0]; ++i)
          rho[LocalIndex(B, i, j, k)] = initial_density.at((size_t)i * NY * NZ + (size_t)j * NZ + k);
  }
}

// src/AmrSim.cpp:217-295: C-order velocity (component fastest).
void AmrSim::InitVelocity(int level) {
  if (level) throw lambrex::Error("Only LEVEL 0 should be initialised from scratch currently.");
  LevelHost& L = levels_.at(level);
  for (size_t b = 0; b < L.boxes.size(); ++b) {
    const lambrex::IBox& B = L.boxes[b];
    auto& u = L.vel[b];
    const long long n = B.numPts();
    for (int k = B.lo[2]; k <= B.hi[2]; ++k)
      for (int j = B.lo[1]; j <= B.hi[1]; ++j)
        for (int i = B.lo[0]; i <= B.hi[0]; ++i)
          for (int c = 0; c < NDIMS; ++c)
            u[c * n + LocalIndex(B, i, j, k)] =
                initial_velocity.at(((size_t)i * NY * NZ + (size_t)j * NZ + k) * NDIMS + c);
  }
}

// src/AmrSim.cpp:665-687
void AmrSim::MakeNewLevelFromScratch(int level, double time, const std::vector<lambrex::IBox>& ba) {
  LevelHost& L = levels_.at(level);
  L.boxes = ba;
  const double nan = std::numeric_limits<double>::quiet_NaN();
  for (int s = 0; s < 2; ++s) {
    L.rho[s].assign(ba.size(), {});
    for (size_t b = 0; b < ba.size(); ++b) L.rho[s][b].assign(ba[b].numPts(), nan);
  }
  L.vel.assign(ba.size(), {});
  for (size_t b = 0; b < ba.size(); ++b) L.vel[b].assign(3 * ba[b].numPts(), nan);
  L.cur = 0;
  const std::vector<int32_t> lohi = lambrex::to_lohi(ba);
  Check(lbx_level_define(sim_, level, (int)ba.size(), lohi.data()));
  L.time = time;
  ComputeDt(level);
  L.step = 0;
  Check(lbx_set_relax(sim_, level, 1.0 / (tau_s_.at(level) + 0.5), 1.0 / (tau_b_.at(level) + 0.5)));
  if (!level) {
    InitDensity(level);
    InitVelocity(level);
    CalcEquilibriumDist(level);
  }
}

// AmrCore::InitFromScratch: base grids then level 0 from scratch.
void AmrSim::InitFromScratch(double time) {
  if (!sim_) {
    const int32_t dims[3] = {NX, NY, NZ};
    const int32_t per[3] = {PERIODICITY[0], PERIODICITY[1], PERIODICITY[2]};
    Check(lbx_sim_create(lambrex::Context(), dims, per, precision_, rho0_, &sim_));
    if (ref_ratio_ != kRefRatio) Check(lbx_set_ref_ratio(sim_, ref_ratio_));
  }
  lambrex::IBox domain;
  domain.lo = {0, 0, 0};
  domain.hi = {NX - 1, NY - 1, NZ - 1};
  MakeNewLevelFromScratch(0, time, lambrex::max_size_chop(domain, max_grid_size_));
  finest_level_ = 0;
}

// src/AmrSim.cpp:746-751
void AmrSim::ClearLevel(int level) {
  Check(lbx_level_clear(sim_, level));  // and every finer level
  for (int l = level; l <= max_level_; ++l) levels_.at(l) = LevelHost{};
  if (level <= finest_level_) finest_level_ = level > 0 ? level - 1 : 0;
}

// src/AmrSim.cpp:845-936 (the arithmetic runs on the GPU in the same order).
void AmrSim::CalcEquilibriumDist(int level) {
  LevelHost& L = levels_.at(level);
  for (size_t b = 0; b < L.boxes.size(); ++b)
    Check(lbx_init_equilibrium(sim_, level, (int)b, L.rho[L.cur][b].data(), L.vel[b].data()));
}

// src/AmrSim.cpp:938-979, as one bulk readback of the level.
void AmrSim::CalcHydroVars(int level) {
  LevelHost& L = levels_.at(level);
  size_t total = 0;
  for (const lambrex::IBox& b : L.boxes) total += (size_t)b.numPts();
  std::vector<double> rho(total), u(3 * total);
  Check(lbx_hydro_level(sim_, level, rho.data(), u.data()));
  size_t off = 0;
  for (size_t b = 0; b < L.boxes.size(); ++b) {
    const size_t n = (size_t)L.boxes[b].numPts();
    std::copy(rho.begin() + off, rho.begin() + off + n, L.rho[L.cur][b].begin());
    std::copy(u.begin() + 3 * off, u.begin() + 3 * (off + n), L.vel[b].begin());
    off += n;
  }
}

// src/AmrSim.cpp:824-843
double AmrSim::GetDensity(int i, int j, int k, int level) const {
  const LevelHost& L = levels_.at(level);
  const int b = BoxOf(level, i, j, k);
  if (b < 0 || L.rho[L.cur].empty()) return NL_DENSITY;
  return L.rho[L.cur][b][LocalIndex(L.boxes[b], i, j, k)];
}

double AmrSim::GetVelocity(int i, int j, int k, int n, int level) const {
  const LevelHost& L = levels_.at(level);
  const int b = BoxOf(level, i, j, k);
  if (b < 0 || L.vel.empty()) return NL_VELOCITY;
  return L.vel[b][n * L.boxes[b].numPts() + LocalIndex(L.boxes[b], i, j, k)];
}

// src/AmrSim.cpp:981-993 with IterateLevel (:324-333) per step, or one
// RohdeCycle(0) per step (:430-469) when a fine level exists.
void AmrSim::Iterate(int nsteps) {
  if (!sim_) throw lambrex::Error("Iterate before InitFromScratch");
  if (nsteps <= 0) return;
  Check(lbx_iterate(sim_, nsteps));
  LevelHost& C = levels_.at(0);
  if (finest_level_ == 0) {
    for (int t = 0; t < nsteps; ++t) {
      C.time += C.dt;
      ++C.step;
    }
    if (nsteps % 2) C.cur = 1 - C.cur;  // UpdateNow swaps Density too
  } else {
    // UpdateDistribution (:619-631): a level between 0 and the finest takes
    // one step per RohdeCycle, r^l cycles per level-0 step (the product of
    // the ratios above it); the finest level takes r steps per cycle of its
    // parent (two at ratio 2, as the reference writes out).
    for (int t = 0; t < nsteps; ++t)
      for (int l = 0, cycles = 1; l <= finest_level_; cycles *= Ratio(l), ++l) {
        LevelHost& L = levels_.at(l);
        if (l > 0 && l == finest_level_) {
          const int r = Ratio(l - 1);  // substeps per cycle of its parent
          for (int c = 0; c < cycles / r; ++c) {
            L.time += r * L.dt;
            L.step += r;
          }
        } else {
          for (int c = 0; c < cycles; ++c) {
            L.time += L.dt;
            ++L.step;
          }
        }
      }
  }
}

// src/AmrSim.cpp:689-713: the fine level from the coarse one.
void AmrSim::MakeNewLevelFromCoarse(int level, double time, const std::vector<lambrex::IBox>& ba) {
  if (!level) throw lambrex::Error("Cannot construct level 0 from a coarser level.");
  if (level > max_level_ || level != finest_level_ + 1)
    throw lambrex::Error("level " + std::to_string(level) + " cannot be made from level " +
                         std::to_string(level - 1));
  LevelHost& L = levels_.at(level);
  L.boxes = ba;
  const double nan = std::numeric_limits<double>::quiet_NaN();
  for (int s = 0; s < 2; ++s) {
    L.rho[s].assign(ba.size(), {});
    for (size_t b = 0; b < ba.size(); ++b) L.rho[s][b].assign(ba[b].numPts(), nan);
  }
  L.vel.assign(ba.size(), {});
  for (size_t b = 0; b < ba.size(); ++b) L.vel[b].assign(3 * ba[b].numPts(), nan);
  L.cur = 0;
  const std::vector<int32_t> lohi = lambrex::to_lohi(ba);
  Check(lbx_level_define(sim_, level, (int)ba.size(), lohi.data()));
  L.time = time;
  ComputeDt(level);
  L.step = 0;
  Check(lbx_set_relax(sim_, level, 1.0 / (tau_s_.at(level) + 0.5), 1.0 / (tau_b_.at(level) + 0.5)));
  Check(lbx_fill_from_coarse(sim_, level));  // DistFnFillFromCoarse
  finest_level_ = level;
  CalcHydroVars(level);
  MakeFineMask(level - 1);
}

// ErrorEst from the state + clustering + the level update of
// AmrCore::regrid, for every level below max_level.  No reference
// counterpart: the reference tags only its static boxes (TagCell,
// src/AmrSim.cpp:413-417).
void AmrSim::Regrid(int criterion, double threshold, int blocking, int buffer) {
  if (!sim_) throw lambrex::Error("Regrid before InitFromScratch");
  if (max_level_ < 1) throw lambrex::Error("Regrid: max_level must be >= 1");
  for (auto& t : static_tags_) t = lambrex::IBox{};
  for (int l = 0; l < max_level_ && l <= finest_level_; ++l) {
    const std::vector<lambrex::IBox> fba = RefineFromState(l, criterion, threshold, blocking, buffer);
    if (fba.empty()) {
      if (finest_level_ > l) ClearLevel(l + 1);
      return;
    }
    if (finest_level_ < l + 1) {
      MakeNewLevelFromCoarse(l + 1, GetTime(l), fba);
    } else if (!(fba == levels_.at(l + 1).boxes)) {
      if (finest_level_ > l + 1) ClearLevel(l + 2);  // level l+1 is remade as the finest level
      RemakeLevel(l + 1, GetTime(l), fba);
    }
    MakeFineMask(l);
  }
}

// Boxes of level `level`+1 from the tags of level `level`.  Level 0 is
// tagged on the device (lbx_tag_cells) over the whole domain.  A finer level
// is tagged here from its CalcHydroVars fields with the same arithmetic
// (central differences, (rho+ - rho-) * 0.5 per axis, or |u|^2), a neighbour
// outside the level standing in as the cell's own value; only cells whose
// 3x3x3 neighbourhood lies in the level may be refined (proper nesting).  The
// level's bounding box is padded by buffer + blocking cells in the directions
// it does not span, so the clustering's periodic growth cannot wrap across it.
std::vector<lambrex::IBox> AmrSim::RefineFromState(int level, int criterion, double threshold, int blocking,
                                                   int buffer) {
  std::vector<uint8_t> tags, allowed;
  int32_t ext[3];
  int org[3] = {0, 0, 0};
  if (level == 0) {
    ext[0] = NX;
    ext[1] = NY;
    ext[2] = NZ;
    tags.resize((size_t)NUMEL);
    Check(lbx_tag_cells(sim_, 0, criterion, threshold, tags.data()));
  } else {
    if (criterion != LBX_TAG_DENSITY_GRADIENT && criterion != LBX_TAG_SPEED)
      throw lambrex::Error("Regrid: unknown tagging criterion");
    if (!std::isfinite(threshold) || threshold < 0.0) throw lambrex::Error("Regrid: threshold must be finite, >= 0");
    if (blocking < 1 || buffer < 0) throw lambrex::Error("Regrid: blocking must be >= 1 and buffer >= 0");
    CalcHydroVars(level);
    const LevelHost& L = levels_.at(level);
    const lambrex::IBox B = lambrex::minimal_box(L.boxes);
    int scale = 1;  // level's index space: the domain refined by every ratio below it
    for (int l = 0; l < level; ++l) scale *= Ratio(l);
    const int dom[3] = {NX * scale, NY * scale, NZ * scale};
    bool wrap[3];
    for (int d = 0; d < 3; ++d) {
      wrap[d] = B.lo[d] == 0 && B.hi[d] == dom[d] - 1;
      const int pad = wrap[d] ? 0 : buffer + blocking;
      org[d] = B.lo[d] - pad;
      ext[d] = B.length(d) + 2 * pad;
    }
    const size_t ncell = (size_t)ext[0] * ext[1] * ext[2];
    std::vector<double> rho(ncell, 0.0), usq(ncell, 0.0);
    std::vector<uint8_t> valid(ncell, 0);
    auto idx = [&](int x, int y, int z) { return ((size_t)z * ext[1] + y) * ext[0] + x; };
    for (size_t b = 0; b < L.boxes.size(); ++b) {
      const lambrex::IBox& Bx = L.boxes[b];
      const long long n = Bx.numPts();
      for (int k = Bx.lo[2]; k <= Bx.hi[2]; ++k)
        for (int j = Bx.lo[1]; j <= Bx.hi[1]; ++j)
          for (int i = Bx.lo[0]; i <= Bx.hi[0]; ++i) {
            const long long li = LocalIndex(Bx, i, j, k);
            const size_t c = idx(i - org[0], j - org[1], k - org[2]);
            rho[c] = L.rho[L.cur][b][li];
            const double u0 = L.vel[b][li], u1 = L.vel[b][n + li], u2 = L.vel[b][2 * n + li];
            usq[c] = u0 * u0 + u1 * u1 + u2 * u2;
            valid[c] = 1;
          }
    }
    // neighbour of (x, y, z) along axis d by s, or -1 outside the padded box
    auto nb = [&](int x, int y, int z, int d, int s) -> long long {
      int p[3] = {x, y, z};
      p[d] += s;
      if (wrap[d])
        p[d] = (p[d] + ext[d]) % ext[d];
      else if (p[d] < 0 || p[d] >= ext[d])
        return -1;
      return (long long)idx(p[0], p[1], p[2]);
    };
    const double thr2 = threshold * threshold;
    tags.assign(ncell, 0);
    allowed.assign(ncell, 0);
    for (int z = 0; z < ext[2]; ++z)
      for (int y = 0; y < ext[1]; ++y)
        for (int x = 0; x < ext[0]; ++x) {
          const size_t c = idx(x, y, z);
          if (!valid[c]) continue;
          double s = 0.0;
          if (criterion == LBX_TAG_DENSITY_GRADIENT) {
            for (int d = 0; d < 3; ++d) {
              const long long ip = nb(x, y, z, d, 1), im = nb(x, y, z, d, -1);
              const double rp = (ip >= 0 && valid[ip]) ? rho[ip] : rho[c];
              const double rm = (im >= 0 && valid[im]) ? rho[im] : rho[c];
              const double g = (rp - rm) * 0.5;
              s = d == 0 ? g * g : s + g * g;
            }
          } else {
            s = usq[c];
          }
          tags[c] = s > thr2 ? 1 : 0;
          bool inner = true;
          for (int dz = -1; dz <= 1 && inner; ++dz)
            for (int dy = -1; dy <= 1 && inner; ++dy)
              for (int dx = -1; dx <= 1 && inner; ++dx) {
                long long q = (long long)c;
                const int sh[3] = {dx, dy, dz};
                int p[3] = {x, y, z};
                for (int d = 0; d < 3 && q >= 0; ++d) {
                  q = nb(p[0], p[1], p[2], d, sh[d]);
                  if (q >= 0) {
                    p[d] += sh[d];
                    if (wrap[d]) p[d] = (p[d] + ext[d]) % ext[d];
                  }
                }
                inner = q >= 0 && valid[q];
              }
          allowed[c] = inner ? 1 : 0;
        }
  }
  // The clustering refines by 2 and chops at max_grid_size.  At another
  // ratio r it runs without the chop (a max_grid_size above twice every
  // extent), its boxes are taken back to this level's index space, chopped
  // there at max_grid_size / r and refined by r, so every fine box is
  // aligned to the ratio (as static refinement does).
  const int r = Ratio(level);
  const int mgs = r == kRefRatio ? max_grid_size_ : 2 * std::max({ext[0], ext[1], ext[2]}) + 2;
  int32_t n = 0;
  const uint8_t* al = allowed.empty() ? nullptr : allowed.data();
  Check(lbx_cluster_tags_allowed(ext, tags.data(), al, blocking, buffer, mgs, 0, nullptr, &n));
  std::vector<int32_t> lohi(6 * (size_t)std::max(n, 1));
  if (n > 0) Check(lbx_cluster_tags_allowed(ext, tags.data(), al, blocking, buffer, mgs, n, lohi.data(), &n));
  std::vector<lambrex::IBox> fba;
  for (int b = 0; b < n; ++b) {
    lambrex::IBox c;  // refined by 2 relative to org
    for (int d = 0; d < NDIMS; ++d) {
      c.lo[d] = lohi[6 * b + d];
      c.hi[d] = lohi[6 * b + 3 + d];
    }
    if (r == kRefRatio) {
      for (int d = 0; d < NDIMS; ++d) {
        c.lo[d] += 2 * org[d];
        c.hi[d] += 2 * org[d];
      }
      fba.push_back(c);
      continue;
    }
    for (int d = 0; d < NDIMS; ++d) {
      c.lo[d] = c.lo[d] / 2 + org[d];
      c.hi[d] = (c.hi[d] + 1) / 2 - 1 + org[d];
    }
    for (const lambrex::IBox& piece : lambrex::max_size_chop(c, std::max(1, max_grid_size_ / r)))
      fba.push_back(lambrex::refine(piece, r));
  }
  return fba;
}

namespace {
std::string box_text(const lambrex::IBox& b) {
  std::ostringstream os;
  os << "((" << b.lo[0] << ',' << b.lo[1] << ',' << b.lo[2] << ") (" << b.hi[0] << ',' << b.hi[1] << ','
     << b.hi[2] << ") (0,0,0))";
  return os.str();
}

void make_dir(const std::string& d) {
  if (mkdir(d.c_str(), 0755) != 0 && errno != EEXIST) throw lambrex::Error("cannot create directory " + d);
}
}  // namespace

void AmrSim::WritePlotfile(const std::string& dir) {
  if (!sim_) throw lambrex::Error("WritePlotfile before InitFromScratch");
  const int nlev = finest_level_ + 1;
  for (int l = 0; l < nlev; ++l) CalcHydroVars(l);
  make_dir(dir);
  const double time = GetTime(0);
  std::ofstream h(dir + "/Header");
  h.precision(17);
  h << "HyperCLaw-V1.1\n4\ndensity\nx_velocity\ny_velocity\nz_velocity\n3\n" << time << '\n'
    << finest_level_ << '\n';
  h << "0 0 0 \n1 1 1 \n";
  for (int l = 0; l < finest_level_; ++l) h << Ratio(l) << ' ';
  h << '\n';
  for (int l = 0, r = 1; l < nlev; r *= Ratio(l), ++l) {
    lambrex::IBox d;
    d.lo = {0, 0, 0};
    d.hi = {r * NX - 1, r * NY - 1, r * NZ - 1};
    h << box_text(d) << ' ';
  }
  h << '\n';
  for (int l = 0; l < nlev; ++l) h << levels_.at(l).step << ' ';
  h << '\n';
  for (int l = 0, r = 1; l < nlev; r *= Ratio(l), ++l)
    h << 1.0 / (r * NX) << ' ' << 1.0 / (r * NY) << ' ' << 1.0 / (r * NZ) << " \n";
  h << "0\n0\n";
  for (int l = 0, r = 1; l < nlev; r *= Ratio(l), ++l) {
    const LevelHost& L = levels_.at(l);
    const int ext[3] = {r * NX, r * NY, r * NZ};
    h << l << ' ' << L.boxes.size() << ' ' << time << '\n' << L.step << '\n';
    for (const lambrex::IBox& b : L.boxes)
      for (int d = 0; d < 3; ++d) h << double(b.lo[d]) / ext[d] << ' ' << double(b.hi[d] + 1) / ext[d] << '\n';
    h << "Level_" << l << "/Cell\n";
    // one FAB per box: "FAB <real descriptor><box> <ncomp>\n" then the data
    const std::string ldir = dir + "/Level_" + std::to_string(l);
    make_dir(ldir);
    std::ofstream dfile(ldir + "/Cell_D_00000", std::ios::binary);
    std::vector<long long> offsets;
    std::vector<std::array<double, 4>> mins, maxs;
    for (size_t b = 0; b < L.boxes.size(); ++b) {
      offsets.push_back((long long)dfile.tellp());
      dfile << "FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (8 7 6 5 4 3 2 1)))" << box_text(L.boxes[b]) << " 4\n";
      const size_t n = (size_t)L.boxes[b].numPts();
      const double* comp[4] = {L.rho[L.cur][b].data(), L.vel[b].data(), L.vel[b].data() + n,
                               L.vel[b].data() + 2 * n};
      std::array<double, 4> mn, mx;
      for (int c = 0; c < 4; ++c) {
        dfile.write(reinterpret_cast<const char*>(comp[c]), (std::streamsize)(n * sizeof(double)));
        // NaN entries (u where rho = 0) are skipped; all-NaN gives NaN
        mn[c] = mx[c] = std::numeric_limits<double>::quiet_NaN();
        for (size_t i = 0; i < n; ++i) {
          mn[c] = std::fmin(mn[c], comp[c][i]);
          mx[c] = std::fmax(mx[c], comp[c][i]);
        }
      }
      mins.push_back(mn);
      maxs.push_back(mx);
    }
    if (!dfile) throw lambrex::Error("cannot write " + ldir + "/Cell_D_00000");
    std::ofstream ch(ldir + "/Cell_H");
    ch.precision(17);
    ch << "1\n1\n4\n0\n(" << L.boxes.size() << " 0\n";
    for (const lambrex::IBox& b : L.boxes) ch << box_text(b) << '\n';
    ch << ")\n" << L.boxes.size() << '\n';
    for (long long o : offsets) ch << "FabOnDisk: Cell_D_00000 " << o << '\n';
    ch << '\n' << L.boxes.size() << ",4\n";
    for (const auto& m : mins) ch << m[0] << ',' << m[1] << ',' << m[2] << ',' << m[3] << ",\n";
    ch << '\n' << L.boxes.size() << ",4\n";
    for (const auto& m : maxs) ch << m[0] << ',' << m[1] << ',' << m[2] << ',' << m[3] << ",\n";
    if (!ch) throw lambrex::Error("cannot write " + ldir + "/Cell_H");
  }
  if (!h) throw lambrex::Error("cannot write " + dir + "/Header");
}

// src/AmrSim.cpp:715-744: new boxes for the existing fine level; each new
// valid cell keeps the old fine populations where an old fine box covered
// it and takes its coarse parent's otherwise (DistFnFillPatch, :359-391).
void AmrSim::RemakeLevel(int level, double time, const std::vector<lambrex::IBox>& ba) {
  if (level < 1 || level != finest_level_)
    throw lambrex::Error("RemakeLevel: level " + std::to_string(level) + " is not the finest level");
  LevelHost& L = levels_.at(level);
  const std::vector<int32_t> lohi = lambrex::to_lohi(ba);
  Check(lbx_level_remake(sim_, level, (int)ba.size(), lohi.data()));
  L.boxes = ba;
  const double nan = std::numeric_limits<double>::quiet_NaN();
  for (int s = 0; s < 2; ++s) {
    L.rho[s].assign(ba.size(), {});
    for (size_t b = 0; b < ba.size(); ++b) L.rho[s][b].assign(ba[b].numPts(), nan);
  }
  L.vel.assign(ba.size(), {});
  for (size_t b = 0; b < ba.size(); ++b) L.vel[b].assign(3 * ba[b].numPts(), nan);
  L.cur = 0;
  L.time = time;
  CalcHydroVars(level);
}

// src/AmrSim.cpp:419-428.  The reference builds the mask from the coarse
// level's own BoxArray (its fba argument), i.e. the coarse cells covered by
// coarsen(coarse boxes, 2); SetFineMaskFromFineGrids(true) uses the fine
// BoxArray instead.
void AmrSim::MakeFineMask(int coarse_level) {
  if (coarse_level < 0 || coarse_level >= finest_level_) return;
  const std::vector<lambrex::IBox>& fba =
      mask_from_fine_grids_ ? levels_.at(coarse_level + 1).boxes : levels_.at(coarse_level).boxes;
  const std::vector<int32_t> lohi = lambrex::to_lohi(fba);
  Check(lbx_set_level_mask(sim_, coarse_level, (int)fba.size(), lohi.data()));
}

void AmrSim::SetFineMaskFromFineGrids(bool on) {
  mask_from_fine_grids_ = on;
  for (int l = 0; l < finest_level_; ++l) MakeFineMask(l);
}

// src/AmrSim.cpp:995-1017.  AMReX's regrid (tagging, error buffer,
// clustering) is not reproduced: level l+1 is exactly the refined tag box of
// level l, clipped to level l's extent (for l >= 1 shrunk by one cell, so
// the finer level's ghost cells keep parents in level l: proper nesting) and
// cut at max_grid_size.
void AmrSim::SetStaticRefinement(int level, const std::array<int, NDIMS>& lo_corner,
                                 const std::array<int, NDIMS>& hi_corner) {
  if (!sim_) throw lambrex::Error("SetStaticRefinement before InitFromScratch");
  if (level < 0 || level >= max_level_ || level > finest_level_)
    throw lambrex::Error("SetStaticRefinement: level " + std::to_string(level) +
                         " cannot be refined (it must exist and be below max_level)");
  lambrex::IBox tag;
  for (int d = 0; d < NDIMS; ++d) {
    tag.lo[d] = lo_corner[d];
    tag.hi[d] = hi_corner[d];
  }
  if (!tag.ok()) throw lambrex::Error("SetStaticRefinement: empty refinement region");
  static_tags_.at(level) = tag;
  RegridFrom(level);
}

void AmrSim::RegridFrom(int base) {
  for (int l = base; l < max_level_ && l <= finest_level_; ++l) {
    lambrex::IBox tag = static_tags_.at(l);
    if (tag.ok()) {
      const lambrex::IBox ext = lambrex::minimal_box(levels_.at(l).boxes);
      const int shrink = l > 0 ? 1 : 0;
      for (int d = 0; d < NDIMS; ++d) {
        tag.lo[d] = std::max(tag.lo[d], ext.lo[d] + shrink);
        tag.hi[d] = std::min(tag.hi[d], ext.hi[d] - shrink);
      }
    }
    if (!tag.ok()) {  // nothing to refine: no finer levels
      if (finest_level_ > l) ClearLevel(l + 1);
      return;
    }
    std::vector<lambrex::IBox> fba;
    if (Ratio(l) == kRefRatio) {
      fba = lambrex::max_size_chop(lambrex::refine(tag, kRefRatio), max_grid_size_);
    } else {  // boxes aligned to the ratio: chop in the coarse index space, then refine
      const int r = Ratio(l);
      for (const lambrex::IBox& b : lambrex::max_size_chop(tag, std::max(1, max_grid_size_ / r)))
        fba.push_back(lambrex::refine(b, r));
    }
    // AmrCore::regrid: a new level is made from the coarse one, an existing
    // level with a different BoxArray is remade (keeping its data; its finer
    // levels are rebuilt above it), an unchanged one is left alone.
    if (finest_level_ < l + 1) {
      MakeNewLevelFromCoarse(l + 1, GetTime(l), fba);
    } else if (!(fba == levels_.at(l + 1).boxes)) {
      if (finest_level_ > l + 1) ClearLevel(l + 2);
      RemakeLevel(l + 1, GetTime(l), fba);
    }
    MakeFineMask(l);
  }
}

void AmrSim::UnsetStaticRefinement(int level) {
  if (level < 0 || level >= max_level_) throw lambrex::Error("UnsetStaticRefinement: level out of range");
  static_tags_.at(level) = lambrex::IBox{};
  if (finest_level_ > level) ClearLevel(level + 1);
}

// src/AmrSim.cpp:1019-1030
std::pair<std::array<int, NDIMS>, std::array<int, NDIMS>> AmrSim::GetExtent(int level) const {
  const lambrex::IBox m = lambrex::minimal_box(levels_.at(level).boxes);
  return {m.lo, m.hi};
}

std::vector<double> AmrSim::DistFn(int level, int box) const {
  const LevelHost& L = levels_.at(level);
  std::vector<double> f((size_t)NMODES * L.boxes.at(box).numPts());
  Check(lbx_download_f(sim_, level, box, f.data()));
  return f;
}

void AmrSim::SetDistFn(int level, int box, const std::vector<double>& f) {
  const LevelHost& L = levels_.at(level);
  if (f.size() != (size_t)NMODES * L.boxes.at(box).numPts())
    throw lambrex::Error("SetDistFn: wrong buffer size");
  Check(lbx_upload_f(sim_, level, box, f.data()));
}

std::array<double, 4> AmrSim::Conservation(int level) const {
  std::array<double, 4> out{};
  Check(lbx_conservation(sim_, level, out.data()));
  return out;
}

namespace {
constexpr const char* kCheckpointTag = "LAMBReX-MI355X-Checkpoint-V1";

std::string hexd(double v) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%a", v);
  return buf;
}

// Whitespace-separated tokens of a checkpoint Header.
struct HeaderReader {
  std::istringstream in;
  std::string path;
  std::string word() {
    std::string w;
    if (!(in >> w)) throw lambrex::Error("checkpoint " + path + ": truncated header");
    return w;
  }
  void expect(const char* key) {
    const std::string w = word();
    if (w != key) throw lambrex::Error("checkpoint " + path + ": expected '" + key + "', found '" + w + "'");
  }
  long long integer() {
    const std::string w = word();
    char* end = nullptr;
    const long long v = std::strtoll(w.c_str(), &end, 10);
    if (end == w.c_str() || *end) throw lambrex::Error("checkpoint " + path + ": bad integer '" + w + "'");
    return v;
  }
  double real() {
    const std::string w = word();
    char* end = nullptr;
    const double v = std::strtod(w.c_str(), &end);
    if (end == w.c_str() || *end) throw lambrex::Error("checkpoint " + path + ": bad number '" + w + "'");
    return v;
  }
  lambrex::IBox box() {
    lambrex::IBox b;
    for (int d = 0; d < 3; ++d) b.lo[d] = (int)integer();
    for (int d = 0; d < 3; ++d) b.hi[d] = (int)integer();
    return b;
  }
};

void write_doubles(std::ofstream& f, const double* p, size_t n) {
  f.write(reinterpret_cast<const char*>(p), (std::streamsize)(n * sizeof(double)));
}

void read_doubles(std::ifstream& f, double* p, size_t n, const std::string& path) {
  f.read(reinterpret_cast<char*>(p), (std::streamsize)(n * sizeof(double)));
  if (!f) throw lambrex::Error("checkpoint " + path + ": file too short");
}
}  // namespace

void AmrSim::WriteCheckpoint(const std::string& dir) const {
  if (!sim_) throw lambrex::Error("WriteCheckpoint before InitFromScratch");
  make_dir(dir);
  std::ostringstream h;
  h << kCheckpointTag << '\n'
    << "dims " << NX << ' ' << NY << ' ' << NZ << '\n'
    << "periodicity " << PERIODICITY[0] << ' ' << PERIODICITY[1] << ' ' << PERIODICITY[2] << '\n'
    << "max_level " << max_level_ << '\n'
    << "finest_level " << finest_level_ << '\n'
    << "precision " << precision_ << ' ' << hexd(rho0_) << '\n'
    << "max_grid_size " << max_grid_size_ << '\n'
    << "mask_from_fine_grids " << (mask_from_fine_grids_ ? 1 : 0) << '\n'
    << "ref_ratio " << ref_ratio_ << '\n';
  for (int l = 0; l <= max_level_; ++l) {
    const lambrex::IBox& t = static_tags_.at(l);
    h << "static_tag " << l << ' ' << t.lo[0] << ' ' << t.lo[1] << ' ' << t.lo[2] << ' ' << t.hi[0] << ' '
      << t.hi[1] << ' ' << t.hi[2] << '\n';
  }
  for (int l = 0; l <= max_level_; ++l)
    h << "tau " << l << ' ' << hexd(tau_s_.at(l)) << ' ' << hexd(tau_b_.at(l)) << ' ' << hexd(mass_.at(l)) << '\n';
  for (int l = 0; l <= finest_level_; ++l) {
    const LevelHost& L = levels_.at(l);
    h << "level " << l << " boxes " << L.boxes.size() << " time " << hexd(L.time) << " dt " << hexd(L.dt)
      << " step " << L.step << " cur " << L.cur << '\n';
    for (const lambrex::IBox& b : L.boxes)
      h << "box " << b.lo[0] << ' ' << b.lo[1] << ' ' << b.lo[2] << ' ' << b.hi[0] << ' ' << b.hi[1] << ' '
        << b.hi[2] << '\n';
  }
  h << "end\n";
  for (int l = 0; l <= finest_level_; ++l) {
    const LevelHost& L = levels_.at(l);
    const std::string ldir = dir + "/Level_" + std::to_string(l);
    make_dir(ldir);
    std::ofstream df(ldir + "/DistFn", std::ios::binary);
    std::ofstream hf(ldir + "/Hydro", std::ios::binary);
    for (size_t b = 0; b < L.boxes.size(); ++b) {
      const std::vector<double> f = DistFn(l, (int)b);
      write_doubles(df, f.data(), f.size());
      const size_t n = (size_t)L.boxes[b].numPts();
      for (int s = 0; s < 2; ++s) {
        if (L.rho[s].size() == L.boxes.size() && L.rho[s][b].size() == n)
          write_doubles(hf, L.rho[s][b].data(), n);
        else
          write_doubles(hf, std::vector<double>(n, std::numeric_limits<double>::quiet_NaN()).data(), n);
      }
      write_doubles(hf, L.vel[b].data(), 3 * n);
    }
    if (!df || !hf) throw lambrex::Error("cannot write the data of " + ldir);
  }
  std::ofstream hout(dir + "/Header");
  hout << h.str();
  if (!hout) throw lambrex::Error("cannot write " + dir + "/Header");
}

void AmrSim::ReadCheckpoint(const std::string& dir) {
  HeaderReader r;
  r.path = dir + "/Header";
  {
    std::ifstream hin(r.path);
    if (!hin) throw lambrex::Error("cannot read " + r.path);
    std::ostringstream all;
    all << hin.rdbuf();
    r.in.str(all.str());
  }
  if (r.word() != kCheckpointTag) throw lambrex::Error("checkpoint " + r.path + ": not a checkpoint of this format");
  r.expect("dims");
  const long long nx = r.integer(), ny = r.integer(), nz = r.integer();
  if (nx != NX || ny != NY || nz != NZ)
    throw lambrex::Error("ReadCheckpoint: the checkpoint's domain differs from this AmrSim's");
  r.expect("periodicity");
  for (int d = 0; d < NDIMS; ++d)
    if (r.integer() != PERIODICITY[d])
      throw lambrex::Error("ReadCheckpoint: the checkpoint's periodicity differs from this AmrSim's");
  r.expect("max_level");
  if (r.integer() != max_level_) throw lambrex::Error("ReadCheckpoint: the checkpoint's max_level differs");
  r.expect("finest_level");
  const long long finest = r.integer();
  if (finest < 0 || finest > max_level_) throw lambrex::Error("ReadCheckpoint: bad finest level");
  r.expect("precision");
  const long long precision = r.integer();
  const double rho0 = r.real();
  if (precision != LBX_FP64 && precision != LBX_FP32_SHIFTED)
    throw lambrex::Error("ReadCheckpoint: unknown precision");
  r.expect("max_grid_size");
  const long long mgs = r.integer();
  r.expect("mask_from_fine_grids");
  const bool mask_fine = r.integer() != 0;
  r.expect("ref_ratio");
  const long long ratio = r.integer();
  if (ratio < 2 || ratio > 4) throw lambrex::Error("ReadCheckpoint: bad refinement ratio");
  std::vector<lambrex::IBox> tags(max_level_ + 1);
  for (int l = 0; l <= max_level_; ++l) {
    r.expect("static_tag");
    if (r.integer() != l) throw lambrex::Error("checkpoint " + r.path + ": static tags out of order");
    tags[l] = r.box();
  }
  std::vector<double> ts(max_level_ + 1), tb(max_level_ + 1), ms(max_level_ + 1);
  for (int l = 0; l <= max_level_; ++l) {
    r.expect("tau");
    if (r.integer() != l) throw lambrex::Error("checkpoint " + r.path + ": tau lines out of order");
    ts[l] = r.real();
    tb[l] = r.real();
    ms[l] = r.real();
  }
  std::vector<LevelHost> lv(max_level_ + 1);
  for (int l = 0; l <= finest; ++l) {
    r.expect("level");
    if (r.integer() != l) throw lambrex::Error("checkpoint " + r.path + ": levels out of order");
    r.expect("boxes");
    const long long nb = r.integer();
    if (nb < 1 || nb > (1 << 24)) throw lambrex::Error("checkpoint " + r.path + ": bad box count");
    LevelHost& L = lv[l];
    r.expect("time");
    L.time = r.real();
    r.expect("dt");
    L.dt = r.real();
    r.expect("step");
    L.step = (int)r.integer();
    r.expect("cur");
    L.cur = r.integer() ? 1 : 0;
    for (long long b = 0; b < nb; ++b) {
      r.expect("box");
      L.boxes.push_back(r.box());
      if (!L.boxes.back().ok()) throw lambrex::Error("checkpoint " + r.path + ": empty box");
    }
  }
  r.expect("end");

  // A fresh device simulation in the checkpoint's precision, every level
  // defined before any data goes up (defining a level resets its parent's
  // mask), then the masks, the populations and the host fields.
  if (sim_) {
    Check(lbx_sim_destroy(sim_));
    sim_ = nullptr;
  }
  precision_ = (int)precision;
  rho0_ = rho0;
  ref_ratio_ = (int)ratio;
  max_grid_size_ = (int)mgs;
  mask_from_fine_grids_ = mask_fine;
  static_tags_ = tags;
  tau_s_ = ts;
  tau_b_ = tb;
  mass_ = ms;
  for (auto& L : levels_) L = LevelHost{};
  finest_level_ = 0;
  const int32_t dims[3] = {NX, NY, NZ};
  const int32_t per[3] = {PERIODICITY[0], PERIODICITY[1], PERIODICITY[2]};
  Check(lbx_sim_create(lambrex::Context(), dims, per, precision_, rho0_, &sim_));
  if (ref_ratio_ != kRefRatio) Check(lbx_set_ref_ratio(sim_, ref_ratio_));
  for (int l = 0; l <= finest; ++l) {
    const std::vector<int32_t> lohi = lambrex::to_lohi(lv[l].boxes);
    Check(lbx_level_define(sim_, l, (int)lv[l].boxes.size(), lohi.data()));
    Check(lbx_set_relax(sim_, l, 1.0 / (tau_s_.at(l) + 0.5), 1.0 / (tau_b_.at(l) + 0.5)));
    levels_.at(l) = lv[l];
    finest_level_ = l;
  }
  for (int l = 0; l < finest_level_; ++l) MakeFineMask(l);
  for (int l = 0; l <= finest_level_; ++l) {
    LevelHost& L = levels_.at(l);
    const std::string ldir = dir + "/Level_" + std::to_string(l);
    std::ifstream df(ldir + "/DistFn", std::ios::binary);
    std::ifstream hf(ldir + "/Hydro", std::ios::binary);
    if (!df || !hf) throw lambrex::Error("cannot read the data of " + ldir);
    for (int s = 0; s < 2; ++s) L.rho[s].assign(L.boxes.size(), {});
    L.vel.assign(L.boxes.size(), {});
    for (size_t b = 0; b < L.boxes.size(); ++b) {
      const size_t n = (size_t)L.boxes[b].numPts();
      std::vector<double> f((size_t)NMODES * n);
      read_doubles(df, f.data(), f.size(), ldir + "/DistFn");
      Check(lbx_upload_f(sim_, l, (int)b, f.data()));
      for (int s = 0; s < 2; ++s) {
        L.rho[s][b].resize(n);
        read_doubles(hf, L.rho[s][b].data(), n, ldir + "/Hydro");
      }
      L.vel[b].resize(3 * n);
      read_doubles(hf, L.vel[b].data(), 3 * n, ldir + "/Hydro");
    }
  }
  box_hint_.clear();
}
