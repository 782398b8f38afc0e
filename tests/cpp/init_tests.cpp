// init_tests.cpp -- the reference's initialisation and single-level AMR
// tests (tests/catch2InitTests.cpp:5-109, tests/catch2AMRTests.cpp:8-116)
// against this build's AmrSim facade.
#include <cmath>

#include "check.h"
#include "lambrex.h"

TEST_CASE("Single Value Initialisation") {
  const double TEST_DENSITY = 0.63, TEST_VELOCITY = 0.23;
  const int TEST_NX = 11, TEST_NY = 12, TEST_NZ = 13, LEVEL = 0;
  AmrSim sim(TEST_NX, TEST_NY, TEST_NZ, LEVEL, {1, 1, 1}, 0.01, 0.01);
  sim.SetInitialDensity(TEST_DENSITY);
  sim.SetInitialVelocity(TEST_VELOCITY);
  sim.InitFromScratch(0.0);
  std::array<int, NDIMS> dims = sim.GetDims();
  REQUIRE(dims[0] == TEST_NX);
  REQUIRE(dims[1] == TEST_NY);
  REQUIRE(dims[2] == TEST_NZ);
  for (int k = 0; k < TEST_NZ; ++k)
    for (int j = 0; j < TEST_NY; ++j)
      for (int i = 0; i < TEST_NX; ++i) {
        REQUIRE(sim.GetDensity(i, j, k, LEVEL) == Approx(TEST_DENSITY));
        for (int d = 0; d < NDIMS; ++d) REQUIRE(sim.GetVelocity(i, j, k, d, LEVEL) == Approx(TEST_VELOCITY));
      }
  // the same values recomputed from the distribution on the GPU
  sim.CalcHydroVars(LEVEL);
  for (int k = 0; k < TEST_NZ; ++k)
    for (int j = 0; j < TEST_NY; ++j)
      for (int i = 0; i < TEST_NX; ++i) {
        REQUIRE(sim.GetDensity(i, j, k, LEVEL) == Approx(TEST_DENSITY));
        for (int d = 0; d < NDIMS; ++d) REQUIRE(sim.GetVelocity(i, j, k, d, LEVEL) == Approx(TEST_VELOCITY));
      }
}

TEST_CASE("Element-wise Initialisation") {
  const int TEST_NX = 16, TEST_NY = 9, TEST_NZ = 8, LEVEL = 0;
  const int NUMEL = TEST_NX * TEST_NY * TEST_NZ;
  std::vector<double> init_density(NUMEL), init_velocity(3 * NUMEL);
  double val = 0.0;
  for (double& e : init_density) e = val++;
  val = 0.0;
  for (double& e : init_velocity) e = val++;
  AmrSim sim(TEST_NX, TEST_NY, TEST_NZ, LEVEL, {1, 1, 1}, 0.01, 0.01);
  sim.SetInitialDensity(init_density);
  sim.SetInitialVelocity(init_velocity);
  sim.InitFromScratch(0.0);
  val = 0.0;
  for (int i = 0; i < TEST_NX; ++i)  // C-like indexing
    for (int j = 0; j < TEST_NY; ++j)
      for (int k = 0; k < TEST_NZ; ++k) REQUIRE(sim.GetDensity(i, j, k, LEVEL) == Approx(val++));
  val = 0.0;
  for (int i = 0; i < TEST_NX; ++i)
    for (int j = 0; j < TEST_NY; ++j)
      for (int k = 0; k < TEST_NZ; ++k)
        for (int d = 0; d < NDIMS; ++d) REQUIRE(sim.GetVelocity(i, j, k, d, LEVEL) == Approx(val++));
}

TEST_CASE("OneLevel") {
  const int NX = 10, NY = 10, NZ = 50, LEVEL = 0;
  const double DENSITY = 0.5, VELOCITY = 0.2, TAU = 0.01;
  AmrSim sim(NX, NY, NZ, LEVEL, {1, 1, 1}, TAU, TAU);
  sim.SetInitialDensity(DENSITY);
  sim.SetInitialVelocity(VELOCITY);
  sim.InitFromScratch(0.0);
  // base grids: domain chopped at max_grid_size 32 -> z in [24,49] and [0,23]
  REQUIRE(sim.boxArray(0).size() == 2);
  REQUIRE(sim.GetTime(LEVEL) == 0.0);
  REQUIRE(sim.GetTimeStep(LEVEL) == 0);
  REQUIRE(sim.GetTauS(LEVEL) == Approx(TAU));
  REQUIRE(sim.GetTauB(LEVEL) == Approx(TAU));
  REQUIRE(sim.GetDt(LEVEL) == 1.0);
  REQUIRE(sim.GetMass(LEVEL) == 1.0);
  auto ext = sim.GetExtent(LEVEL);
  REQUIRE(ext.first[0] == 0);
  REQUIRE(ext.second[2] == NZ - 1);
  // a uniform state is a fixed point of collide + stream
  sim.Iterate(3);
  REQUIRE(sim.GetTime(LEVEL) == 3.0);
  REQUIRE(sim.GetTimeStep(LEVEL) == 3);
  sim.CalcHydroVars(LEVEL);
  for (int k = 0; k < NZ; ++k)
    for (int j = 0; j < NY; ++j)
      for (int i = 0; i < NX; ++i) {
        REQUIRE(sim.GetDensity(i, j, k, LEVEL) == Approx(DENSITY));
        for (int d = 0; d < NDIMS; ++d) REQUIRE(sim.GetVelocity(i, j, k, d, LEVEL) == Approx(VELOCITY));
      }
  // off-domain queries return the sentinels
  REQUIRE(!sim.OnProcessDensity(sim.GetDensity(NX, 0, 0, LEVEL)));
  REQUIRE(!sim.OnProcessVelocity(sim.GetVelocity(0, NY, 0, 0, LEVEL)));
}

// The reference aborts on a non-periodic direction (src/AmrSim.cpp:788-797);
// here it gets bounce-back walls: a closed box keeps its mass, and a fine
// level is refused.
TEST_CASE("Non-periodic directions get bounce-back walls") {
  AmrSim sim(8, 6, 4, 1, {1, 0, 1}, 0.7, 0.7);
  sim.SetInitialDensity(1.0);
  sim.SetInitialVelocity(0.02);
  sim.InitFromScratch(0.0);
  const double m0 = sim.Conservation(0)[0];
  sim.Iterate(40);
  REQUIRE(std::abs(sim.Conservation(0)[0] - m0) <= 1e-12 * m0);
  sim.CalcHydroVars(0);
  REQUIRE(std::abs(sim.GetVelocity(3, 0, 1, 1, 0)) < 0.02);  // the wall stops the normal flow
  REQUIRE_THROWS(sim.SetStaticRefinement(0, {0, 0, 0}, {3, 3, 3}));
}

TEST_CASE("Conservation reductions") {
  const int N = 16;
  AmrSim sim(N, N, N, 0, {1, 1, 1}, 0.3, 0.3);
  std::vector<double> rho(N * N * N), u(3 * N * N * N);
  for (size_t c = 0; c < rho.size(); ++c) rho[c] = 1.0 + 1e-3 * std::sin(0.1 * c);
  for (size_t c = 0; c < u.size(); ++c) u[c] = 1e-3 * std::cos(0.37 * c);
  sim.SetInitialDensity(rho);
  sim.SetInitialVelocity(u);
  sim.SetMaxGridSize(8);
  sim.InitFromScratch(0.0);
  const auto c0 = sim.Conservation(0);
  sim.Iterate(25);
  const auto c1 = sim.Conservation(0);
  REQUIRE(std::fabs(c1[0] - c0[0]) <= 1e-12 * c0[0]);
  for (int a = 1; a < 4; ++a) REQUIRE(std::fabs(c1[a] - c0[a]) <= 1e-12 * c0[0]);
}

int main() {
  lambrexInit();
  const int rc = check::run_all();
  lambrexFinalise();
  return rc;
}
