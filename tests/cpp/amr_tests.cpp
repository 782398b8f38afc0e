// amr_tests.cpp -- the reference's two-level AMR tests
// (tests/catch2AMRTests.cpp:118-423) against this build's AmrSim facade.
// Sections that inspect AMReX TagBoxArrays (ErrorEst) have no counterpart
// here: tagging is reduced to the static refinement box.
#include <algorithm>
#include <cmath>
#include <vector>

#include "check.h"
#include "lambrex.h"

namespace {
const int NX = 48, NY = 24, NZ = 12, MAX_LEVEL = 1;
const double DENSITY = 0.8, VELOCITY = 0.1, TAU = 0.01;

AmrSim* make() {
  AmrSim* sim = new AmrSim(NX, NY, NZ, MAX_LEVEL, {1, 1, 1}, TAU, TAU);
  sim->SetInitialDensity(DENSITY);
  sim->SetInitialVelocity(VELOCITY);
  sim->InitFromScratch(0.0);
  return sim;
}
}  // namespace

TEST_CASE("TwoLevel Initialisation") {
  std::unique_ptr<AmrSim> sim(make());
  REQUIRE(sim->refRatio().size() == (size_t)MAX_LEVEL + 1);
  for (int d = 0; d < NDIMS; ++d) REQUIRE(sim->refRatio(0)[d] == 2);
  REQUIRE(sim->maxLevel() == MAX_LEVEL);
  REQUIRE(sim->finestLevel() == 0);
  REQUIRE_FALSE(sim->boxArray(0).empty());
  REQUIRE(sim->boxArray(1).empty());
  // DistributionMap beside boxArray, as tests/catch2AMRTests.cpp:27 builds a
  // TagBoxArray from the pair: one owner per box, all on this GPU (rank 0)
  const lambrex::DistributionMapping dm = sim->DistributionMap(0);
  REQUIRE(dm.size() == (int)sim->boxArray(0).size());
  for (int b = 0; b < dm.size(); ++b) REQUIRE(dm[b] == 0);
  REQUIRE(sim->DistributionMap(1).size() == 0);
}

TEST_CASE("TwoLevel FineGrids") {
  std::unique_ptr<AmrSim> sim(make());
  const std::array<int, NDIMS> REF_LO{NX / 4, NY / 4, NZ / 4};
  const std::array<int, NDIMS> REF_HI{3 * NX / 4, 3 * NY / 4, 3 * NZ / 4};
  sim->SetStaticRefinement(0, REF_LO, REF_HI);
  REQUIRE(sim->finestLevel() == 1);
  long long num_pts = 1;
  for (int d = 0; d < NDIMS; ++d) num_pts *= REF_HI[d] - REF_LO[d];
  int ratio = 1;
  for (int level = 0; level <= MAX_LEVEL; ++level) {
    auto ext = sim->GetExtent(level);
    lambrex::IBox m;
    m.lo = ext.first;
    m.hi = ext.second;
    REQUIRE(m.contains(ratio * REF_LO[0], ratio * REF_LO[1], ratio * REF_LO[2]));
    REQUIRE(m.contains(ratio * REF_HI[0], ratio * REF_HI[1], ratio * REF_HI[2]));
    REQUIRE(m.numPts() >= (long long)ratio * ratio * ratio * num_pts);
    ratio *= 2;
  }
  // ComputeDt (tests/catch2AMRTests.cpp:348-353)
  REQUIRE(sim->GetDt(1) == 0.5);
  REQUIRE(sim->GetMass(1) == 0.5);
  REQUIRE(sim->GetTauS(1) == Approx(2 * (sim->GetTauS(0) - 0.5) + 0.5));
  REQUIRE(sim->GetTauB(1) == Approx(2 * (sim->GetTauB(0) - 0.5) + 0.5));
  REQUIRE(sim->GetTime(1) == 0.0);
  REQUIRE(sim->GetTimeStep(1) == 0);
}

TEST_CASE("TwoLevel MakeNewLevelFromCoarse keeps uniform fields") {
  std::unique_ptr<AmrSim> sim(make());
  sim->SetStaticRefinement(0, {0, 0, 0}, {NX - 1, NY - 1, NZ - 1});
  for (int level = 0; level <= MAX_LEVEL; ++level) {
    const int r = level ? 2 : 1;
    for (int k = 0; k < r * NZ; k += 3)
      for (int j = 0; j < r * NY; j += 2)
        for (int i = 0; i < r * NX; ++i) {
          REQUIRE(sim->GetDensity(i, j, k, level) == Approx(DENSITY));
          for (int d = 0; d < NDIMS; ++d) REQUIRE(sim->GetVelocity(i, j, k, d, level) == Approx(VELOCITY));
        }
  }
}

TEST_CASE("TwoLevel Iterate advances both levels") {
  std::unique_ptr<AmrSim> sim(make());
  sim->SetStaticRefinement(0, {8, 4, 2}, {23, 15, 9});
  sim->Iterate(3);
  REQUIRE(sim->GetTime(0) == 3.0);
  REQUIRE(sim->GetTimeStep(0) == 3);
  REQUIRE(sim->GetTime(1) == 3.0);  // 2 fine steps of dt 1/2 per coarse step
  REQUIRE(sim->GetTimeStep(1) == 6);
  sim->CalcHydroVars(0);
  sim->CalcHydroVars(1);
  REQUIRE(sim->OnProcessDensity(sim->GetDensity(16, 8, 4, 1)));
  REQUIRE(!sim->OnProcessDensity(sim->GetDensity(0, 0, 0, 1)));  // not refined there
}

// A second SetStaticRefinement with another region remakes the fine level
// (AmrCore::regrid -> RemakeLevel, src/AmrSim.cpp:715-744): cells refined
// before and after keep their fine data, newly refined cells take the
// coarse parent's state.
TEST_CASE("TwoLevel re-refinement keeps fine data") {
  std::unique_ptr<AmrSim> sim(make());
  sim->SetStaticRefinement(0, {8, 4, 2}, {23, 15, 9});
  sim->Iterate(3);
  sim->CalcHydroVars(0);
  sim->CalcHydroVars(1);
  std::vector<double> before;
  for (int k = 8; k <= 19; ++k)
    for (int j = 8; j <= 31; ++j)
      for (int i = 24; i <= 47; ++i) before.push_back(sim->GetDensity(i, j, k, 1));
  const auto fine_before = sim->boxArray(1);
  sim->SetStaticRefinement(0, {12, 4, 4}, {31, 15, 9});
  REQUIRE(sim->finestLevel() == 1);
  REQUIRE(!(sim->boxArray(1) == fine_before));
  REQUIRE(sim->GetTimeStep(1) == 6);
  size_t n = 0;
  bool all_same = true;
  for (int k = 8; k <= 19; ++k)
    for (int j = 8; j <= 31; ++j)
      for (int i = 24; i <= 47; ++i) all_same &= sim->GetDensity(i, j, k, 1) == before[n++];
  REQUIRE(all_same);  // refined before and after: the old fine values
  bool parent = true;
  for (int k = 8; k <= 19; ++k)
    for (int j = 8; j <= 31; ++j)
      for (int i = 48; i <= 63; ++i) parent &= sim->GetDensity(i, j, k, 1) == sim->GetDensity(i / 2, j / 2, k / 2, 0);
  REQUIRE(parent);  // newly refined: the coarse parent's density
  REQUIRE(!sim->OnProcessDensity(sim->GetDensity(16, 8, 4, 1)));  // no longer refined
  sim->Iterate(1);
  REQUIRE(sim->GetTimeStep(1) == 8);
}

// Regrid from the state: a density pulse on one plane is refined, the fine
// level starts from its coarse parents, a higher threshold remakes a
// smaller level, an unreachable one removes it.
TEST_CASE("TwoLevel Regrid from a density gradient") {
  AmrSim sim(NX, NY, NZ, MAX_LEVEL, {1, 1, 1}, 0.8, 0.8);
  std::vector<double> rho((size_t)NX * NY * NZ, 1.0);
  for (int i = 0; i < NX; ++i)
    for (int j = 0; j < NY; ++j)
      for (int k = 0; k < NZ; ++k)
        if (i == 20) rho[((size_t)i * NY + j) * NZ + k] = 1.01;  // C order, i slowest
  sim.SetInitialDensity(rho);
  sim.SetInitialVelocity(0.0);
  sim.InitFromScratch(0.0);
  sim.Regrid(LBX_TAG_DENSITY_GRADIENT, 1e-3);
  REQUIRE(sim.finestLevel() == 1);
  // the tagged cells are x = 19 and 21 (central differences); tiles of 4 with
  // a buffer of 1 refine x in [16, 24) at every y and z
  auto ext = sim.GetExtent(1);
  REQUIRE(ext.first[0] == 32);
  REQUIRE(ext.second[0] == 47);
  REQUIRE(ext.first[1] == 0);
  REQUIRE(ext.second[1] == 2 * NY - 1);
  sim.CalcHydroVars(0);
  bool parent = true;
  for (int i = 32; i <= 47; ++i) parent &= sim.GetDensity(i, 5, 3, 1) == sim.GetDensity(i / 2, 2, 1, 0);
  REQUIRE(parent);
  sim.Iterate(4);
  REQUIRE(sim.GetTimeStep(1) == 8);
  sim.Regrid(LBX_TAG_DENSITY_GRADIENT, 1e9);
  REQUIRE(sim.finestLevel() == 0);
  REQUIRE(sim.boxArray(1).empty());
  sim.Iterate(1);
}

TEST_CASE("TwoLevel UnsetStaticRefinement") {
  std::unique_ptr<AmrSim> sim(make());
  sim->SetStaticRefinement(0, {8, 4, 2}, {23, 15, 9});
  REQUIRE(sim->finestLevel() == 1);
  sim->UnsetStaticRefinement(0);
  REQUIRE(sim->finestLevel() == 0);
  REQUIRE(sim->boxArray(1).empty());
  sim->Iterate(2);
  REQUIRE(sim->GetTimeStep(0) == 2);
}

TEST_CASE("Refinement needs max_level >= 1") {
  AmrSim sim(8, 8, 8, 0, {1, 1, 1}, 0.5, 0.5);
  sim.SetInitialDensity(1.0);
  sim.SetInitialVelocity(0.0);
  sim.InitFromScratch(0.0);
  REQUIRE_THROWS(sim.SetStaticRefinement(0, {0, 0, 0}, {3, 3, 3}));
}

// Three levels (RohdeCycle's recursion, src/AmrSim.cpp:453-455): static
// boxes on levels 0 and 1, the dt / tau / step bookkeeping of ComputeDt and
// UpdateDistribution, re-refinement of level 0 under an existing level 2
// (level 1 is remade, level 2 rebuilt from its static box), and removal.
TEST_CASE("ThreeLevel static refinement") {
  AmrSim sim(32, 16, 16, 2, {1, 1, 1}, 0.8, 0.8);
  sim.SetInitialDensity(1.0);
  sim.SetInitialVelocity(0.01);
  sim.InitFromScratch(0.0);
  REQUIRE_THROWS(sim.SetStaticRefinement(1, {20, 10, 10}, {43, 21, 21}));  // level 1 does not exist yet
  sim.SetStaticRefinement(0, {8, 4, 4}, {23, 11, 11});
  sim.SetStaticRefinement(1, {20, 10, 10}, {43, 21, 21});
  REQUIRE(sim.finestLevel() == 2);
  auto e2 = sim.GetExtent(2);
  REQUIRE(e2.first[0] == 40);
  REQUIRE(e2.second[0] == 87);
  REQUIRE(e2.first[1] == 20);
  REQUIRE(e2.second[2] == 43);
  REQUIRE(sim.GetDt(2) == 0.25);
  REQUIRE(sim.GetTauS(2) == Approx(2 * (sim.GetTauS(1) - 0.5) + 0.5));
  REQUIRE(sim.GetTauS(1) == Approx(2 * (sim.GetTauS(0) - 0.5) + 0.5));
  sim.Iterate(3);
  for (int l = 0; l <= 2; ++l) {
    REQUIRE(sim.GetTimeStep(l) == 3 << l);
    REQUIRE(sim.GetTime(l) == 3.0);
  }
  sim.CalcHydroVars(2);
  REQUIRE(sim.OnProcessDensity(sim.GetDensity(60, 30, 30, 2)));
  REQUIRE(std::isfinite(sim.GetDensity(60, 30, 30, 2)));
  REQUIRE(!sim.OnProcessDensity(sim.GetDensity(10, 10, 10, 2)));
  // a level-1 static box reaching past level 1 is clipped one cell inside it
  sim.SetStaticRefinement(1, {0, 0, 0}, {63, 31, 31});
  auto e2b = sim.GetExtent(2);
  REQUIRE(e2b.first[0] == 2 * 17);
  REQUIRE(e2b.second[0] == 2 * 46 + 1);
  // moving level 0's box remakes level 1 and rebuilds level 2 inside it
  sim.SetStaticRefinement(0, {6, 4, 4}, {21, 11, 11});
  REQUIRE(sim.finestLevel() == 2);
  auto e1 = sim.GetExtent(1);
  REQUIRE(e1.first[0] == 12);
  REQUIRE(sim.GetExtent(2).first[0] == 2 * 13);
  sim.Iterate(1);
  REQUIRE(sim.GetTimeStep(2) == 4);      // a level made from coarse restarts at step 0 (src/AmrSim.cpp:700)
  REQUIRE(sim.GetTimeStep(1) == 6 + 2);  // a remade level keeps its count (:715-744)
  sim.UnsetStaticRefinement(1);
  REQUIRE(sim.finestLevel() == 1);
  REQUIRE(sim.boxArray(2).empty());
  sim.Iterate(1);
  REQUIRE(sim.GetTimeStep(1) == 8 + 2);
  REQUIRE_THROWS(sim.SetStaticRefinement(2, {0, 0, 0}, {3, 3, 3}));  // max_level is 2
  REQUIRE_THROWS(AmrSim(8, 8, 8, LBX_MAX_LEVELS, {1, 1, 1}, 0.8, 0.8));
}

// Regrid on three levels: level 0's pulse tags make level 1, level 1's own
// gradient tags (from its fields) make level 2 inside it, properly nested;
// an unreachable threshold removes both.
TEST_CASE("ThreeLevel Regrid from a density gradient") {
  const int nx = 32, ny = 16, nz = 16;
  AmrSim sim(nx, ny, nz, 2, {1, 1, 1}, 0.8, 0.8);
  std::vector<double> rho((size_t)nx * ny * nz, 1.0);
  for (int j = 0; j < ny; ++j)
    for (int k = 0; k < nz; ++k) rho[((size_t)16 * ny + j) * nz + k] = 1.01;  // plane x = 16, C order
  sim.SetInitialDensity(rho);
  sim.SetInitialVelocity(0.0);
  sim.InitFromScratch(0.0);
  sim.Regrid(LBX_TAG_DENSITY_GRADIENT, 1e-3);
  REQUIRE(sim.finestLevel() == 2);
  auto e1 = sim.GetExtent(1);
  REQUIRE(e1.first[0] == 24);  // tags at coarse x 15, 17 grown by 1: tiles 12..19
  REQUIRE(e1.second[0] == 39);
  REQUIRE(e1.second[1] == 2 * ny - 1);
  auto e2 = sim.GetExtent(2);
  REQUIRE(e2.first[0] == 54);  // level-1 tags at x 31..34 grown by 1, tiles of 4 from x 19
  REQUIRE(e2.second[0] == 77);
  REQUIRE(e2.first[1] == 0);
  REQUIRE(e2.second[2] == 4 * nz - 1);
  sim.CalcHydroVars(1);
  sim.CalcHydroVars(2);
  bool parent = true;
  for (int i = 54; i <= 77; ++i) parent &= sim.GetDensity(i, 9, 5, 2) == sim.GetDensity(i / 2, 4, 2, 1);
  REQUIRE(parent);
  sim.CalcHydroVars(0);
  REQUIRE(sim.GetDensity(64, 0, 0, 2) == sim.GetDensity(16, 0, 0, 0));  // the pulse plane, copied down
  REQUIRE(sim.GetDensity(64, 0, 0, 2) > sim.GetDensity(60, 0, 0, 2));
  sim.Iterate(2);
  REQUIRE(sim.GetTimeStep(2) == 8);
  sim.CalcHydroVars(2);
  REQUIRE(std::isfinite(sim.GetDensity(64, 0, 0, 2)));
  sim.Regrid(LBX_TAG_DENSITY_GRADIENT, 1e9);
  REQUIRE(sim.finestLevel() == 0);
  REQUIRE(sim.boxArray(1).empty());
  REQUIRE(sim.boxArray(2).empty());
  sim.Iterate(1);
}

// Regrid at refinement ratio 3: the same pulse plane tags the same coarse
// tiles, which are refined by 3 and chopped in the coarse index space, so
// every fine box is aligned to the ratio; level 1's own tags then make a
// level 2 aligned to 3 and properly nested inside it.
TEST_CASE("Regrid at ratio 3") {
  AmrSim sim(NX, NY, NZ, 2, {1, 1, 1}, 0.8, 0.8);
  sim.SetRefRatio(3);
  std::vector<double> rho((size_t)NX * NY * NZ, 1.0);
  for (int j = 0; j < NY; ++j)
    for (int k = 0; k < NZ; ++k) rho[((size_t)20 * NY + j) * NZ + k] = 1.01;  // plane x = 20, C order
  sim.SetInitialDensity(rho);
  sim.SetInitialVelocity(0.0);
  sim.InitFromScratch(0.0);
  sim.Regrid(LBX_TAG_DENSITY_GRADIENT, 1e-3);
  REQUIRE(sim.finestLevel() >= 1);
  // tags at x = 19 and 21, grown by 1: tiles of 4 cover coarse x in [16, 24)
  auto ext = sim.GetExtent(1);
  REQUIRE(ext.first[0] == 3 * 16);
  REQUIRE(ext.second[0] == 3 * 24 - 1);
  REQUIRE(ext.first[1] == 0);
  REQUIRE(ext.second[1] == 3 * NY - 1);
  REQUIRE(ext.second[2] == 3 * NZ - 1);
  for (int l = 1; l <= sim.finestLevel(); ++l)
    for (const lambrex::IBox& b : sim.boxArray(l))
      for (int d = 0; d < NDIMS; ++d) {
        REQUIRE(b.lo[d] % 3 == 0);
        REQUIRE((b.hi[d] + 1) % 3 == 0);
        REQUIRE(b.length(d) <= 32);
      }
  if (sim.finestLevel() == 2) {  // level 2 inside level 1, one level-1 cell from its x faces
    auto e2 = sim.GetExtent(2);
    REQUIRE(e2.first[0] >= 3 * (ext.first[0] + 1));
    REQUIRE(e2.second[0] <= 3 * ext.second[0] - 1);
  }
  sim.CalcHydroVars(0);
  sim.CalcHydroVars(1);
  bool parent = true;
  for (int i = ext.first[0]; i <= ext.second[0]; ++i)
    parent &= sim.GetDensity(i, 7, 4, 1) == sim.GetDensity(i / 3, 7 / 3, 4 / 3, 0);
  REQUIRE(parent);
  sim.Iterate(2);
  REQUIRE(sim.GetTimeStep(1) == 6);
  sim.CalcHydroVars(1);
  REQUIRE(std::isfinite(sim.GetDensity(3 * 20, 0, 0, 1)));
  sim.Regrid(LBX_TAG_DENSITY_GRADIENT, 1e9);
  REQUIRE(sim.finestLevel() == 0);
  REQUIRE(sim.boxArray(1).empty());
  sim.Iterate(1);
}

// SetPrecision(LBX_FP32_SHIFTED, 1.0) through the facade: a two-level run
// tracks the same run in FP64 to 1e-6 in density on both levels, and
// SetPrecision after InitFromScratch is refused.
TEST_CASE("FP32-shifted facade run tracks FP64") {
  const int nx = 24, ny = 16, nz = 12;
  std::vector<double> rho((size_t)nx * ny * nz), u(3 * (size_t)nx * ny * nz, 0.0);
  for (size_t c = 0; c < rho.size(); ++c) {
    rho[c] = 1.0 + 1e-3 * (double)((c * 2654435761u) % 1000) / 1000.0;
    u[3 * c] = 1e-3 * (double)((c * 40503u) % 100) / 100.0;
  }
  double max_err = 0.0;
  std::vector<double> dens[2][2];  // [precision][level]
  for (int p = 0; p < 2; ++p) {
    AmrSim sim(nx, ny, nz, 1, {1, 1, 1}, 0.8, 0.8);
    if (p) sim.SetPrecision(LBX_FP32_SHIFTED, 1.0);
    sim.SetInitialDensity(rho);
    sim.SetInitialVelocity(u);
    sim.SetMaxGridSize(8);
    sim.InitFromScratch(0.0);
    REQUIRE_THROWS(sim.SetPrecision(LBX_FP64, 1.0));
    sim.SetFineMaskFromFineGrids(true);
    sim.SetStaticRefinement(0, {6, 0, 0}, {17, ny - 1, nz - 1});
    sim.Iterate(10);
    for (int l = 0; l < 2; ++l) {
      sim.CalcHydroVars(l);
      auto e = sim.GetExtent(l);
      for (int i = e.first[0]; i <= e.second[0]; ++i)
        for (int j = e.first[1]; j <= e.second[1]; ++j)
          for (int k = e.first[2]; k <= e.second[2]; ++k) dens[p][l].push_back(sim.GetDensity(i, j, k, l));
    }
  }
  for (int l = 0; l < 2; ++l) {
    REQUIRE(dens[0][l].size() == dens[1][l].size());
    for (size_t c = 0; c < dens[0][l].size(); ++c)
      max_err = std::max(max_err, std::fabs(dens[1][l][c] - dens[0][l][c]) / dens[0][l][c]);
  }
  REQUIRE(max_err > 0.0);  // the two storages really differ
  REQUIRE(max_err <= 1e-6);
}

// Refinement ratio 3 between levels 0 and 1 (SetRefRatio; the reference
// fixes 2): ComputeDt at that ratio, boxes aligned to it, three fine steps
// per coarse step, conservation of the fine level's arithmetic (finite
// state, masses of both levels reported).
TEST_CASE("TwoLevel Ratio3") {
  AmrSim sim(24, 24, 12, 1, {1, 1, 1}, 0.7, 0.7);
  REQUIRE_THROWS(sim.SetRefRatio(5));
  sim.SetRefRatio(3);
  sim.SetInitialDensity(1.0);
  sim.SetInitialVelocity(0.01);
  sim.InitFromScratch(0.0);
  REQUIRE_THROWS(sim.SetRefRatio(2));  // after InitFromScratch
  REQUIRE(sim.refRatio(0)[0] == 3);
  sim.SetStaticRefinement(0, {6, 6, 3}, {14, 14, 8});
  REQUIRE(sim.finestLevel() == 1);
  for (const lambrex::IBox& b : sim.boxArray(1))
    for (int d = 0; d < NDIMS; ++d) {
      REQUIRE(b.lo[d] % 3 == 0);
      REQUIRE((b.hi[d] + 1) % 3 == 0);
    }
  const auto ext = sim.GetExtent(1);
  REQUIRE(ext.first[0] == 18);
  REQUIRE(ext.second[0] == 44);
  REQUIRE(sim.GetDt(1) == 1.0 / 3);
  REQUIRE(sim.GetTauS(1) == 3 * (0.7 - 0.5) + 0.5);
  sim.Iterate(2);
  REQUIRE(sim.GetTimeStep(0) == 2);
  REQUIRE(sim.GetTimeStep(1) == 6);
  REQUIRE(sim.GetTime(1) == Approx(2.0));
  sim.CalcHydroVars(1);
  const double rho = sim.GetDensity(30, 30, 15, 1);
  REQUIRE(std::isfinite(rho));
  REQUIRE(rho > 0.5);
  REQUIRE(rho < 1.5);
}

// Ratio 3 on every level of a three-level hierarchy: ComputeDt per level
// (dt 3^-l, tau_l = 3 (tau_{l-1} - 1/2) + 1/2), boxes aligned to 3 in each
// level's index space 3^l * dims, 3^l steps of level l per level-0 step.
TEST_CASE("ThreeLevel Ratio3") {
  AmrSim sim(24, 24, 12, 2, {1, 1, 1}, 0.7, 0.7);
  sim.SetRefRatio(3);
  sim.SetMaxGridSize(9);
  sim.SetInitialDensity(1.0);
  sim.SetInitialVelocity(0.01);
  sim.InitFromScratch(0.0);
  sim.SetFineMaskFromFineGrids(true);
  sim.SetStaticRefinement(0, {6, 6, 3}, {14, 14, 8});
  sim.SetStaticRefinement(1, {24, 24, 12}, {35, 35, 20});
  REQUIRE(sim.finestLevel() == 2);
  for (int l = 1; l <= 2; ++l) {
    REQUIRE(sim.refRatio(l - 1)[0] == 3);
    for (const lambrex::IBox& b : sim.boxArray(l))
      for (int d = 0; d < NDIMS; ++d) {
        REQUIRE(b.lo[d] % 3 == 0);
        REQUIRE((b.hi[d] + 1) % 3 == 0);
      }
  }
  const auto ext = sim.GetExtent(2);
  REQUIRE(ext.first[0] == 72);
  REQUIRE(ext.second[0] == 107);
  REQUIRE(sim.GetDt(2) == sim.GetDt(1) / 3);
  const double tau1 = 3 * (0.7 - 0.5) + 0.5;
  REQUIRE(sim.GetTauS(2) == 3 * (tau1 - 0.5) + 0.5);
  sim.Iterate(2);
  REQUIRE(sim.GetTimeStep(0) == 2);
  REQUIRE(sim.GetTimeStep(1) == 6);
  REQUIRE(sim.GetTimeStep(2) == 18);
  REQUIRE(sim.GetTime(2) == Approx(2.0));
  // The cycle as coded is not conservative (the masked coarse cells take
  // streamed values as well as the fine averages), and at ratio 3 the
  // density moves fast: the oracle gives 0.86, 0.61 and 0.46 at these three
  // centres after 2 level-0 steps; the bitwise comparison with the oracle
  // is in tests/test_gpu_multilevel.py.
  for (int l = 0; l <= 2; ++l) {
    sim.CalcHydroVars(l);
    const auto e = sim.GetExtent(l);
    const double rho = sim.GetDensity((e.first[0] + e.second[0]) / 2, (e.first[1] + e.second[1]) / 2,
                                      (e.first[2] + e.second[2]) / 2, l);
    REQUIRE(std::isfinite(rho));
    REQUIRE(rho > 0.2);
    REQUIRE(rho < 1.5);
  }
}

int main() {
  lambrexInit();
  const int rc = check::run_all();
  lambrexFinalise();
  return rc;
}
