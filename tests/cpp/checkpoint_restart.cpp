// checkpoint_restart.cpp -- WriteCheckpoint / ReadCheckpoint of the AmrSim
// facade: Iterate(a), a checkpoint, a fresh AmrSim restarted from it and
// Iterate(b) give the same populations, bit for bit, on every box of every
// level as one AmrSim that runs Iterate(a + b), with the same time, step
// count and (after CalcHydroVars) density.  Single level, two levels with the
// reference's mask (src/AmrSim.cpp:419-428) and with the fine-grid mask,
// three levels (at ratios 2 and 3), and FP32-shifted storage.  The
// reference has no I/O; the stepping it continues is its Iterate
// (src/AmrSim.cpp:981-993).
#include <unistd.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <memory>
#include <random>
#include <string>
#include <vector>

#include "check.h"
#include "lambrex.h"

namespace {
const int NX = 24, NY = 20, NZ = 16;

std::string scratch_dir(const char* name) {
  const char* base = std::getenv("LBX_CHECKPOINT_DIR");
  std::string d = base && *base ? base : "/tmp";
  d += std::string("/lbx_ckpt_") + name + "_" + std::to_string(getpid());
  return d;
}

struct Setup {
  int max_level = 0;
  int precision = LBX_FP64;
  bool fine_mask = false;
  int ratio = 2;
  std::array<int, NDIMS> periodic{1, 1, 1};
};

std::unique_ptr<AmrSim> make(const Setup& s) {
  std::unique_ptr<AmrSim> sim(new AmrSim(NX, NY, NZ, s.max_level, s.periodic, 0.7, 0.7));
  sim->SetMaxGridSize(s.ratio == 2 ? 8 : 9);
  if (s.precision != LBX_FP64) sim->SetPrecision(s.precision, 1.0);
  if (s.ratio != 2) sim->SetRefRatio(s.ratio);
  std::mt19937_64 rng(12345);
  std::normal_distribution<double> n(0.0, 1e-3);
  std::uniform_real_distribution<double> uu(-0.01, 0.01);
  std::vector<double> rho((size_t)NX * NY * NZ), u(3 * rho.size());
  for (double& r : rho) r = 1.0 + n(rng);
  for (double& v : u) v = uu(rng);
  sim->SetInitialDensity(rho);
  sim->SetInitialVelocity(u);
  sim->InitFromScratch(0.0);
  if (s.max_level >= 1) {
    sim->SetFineMaskFromFineGrids(s.fine_mask);
    sim->SetStaticRefinement(0, {6, 4, 2}, {17, 15, 13});
  }
  if (s.max_level >= 2) sim->SetStaticRefinement(1, {16, 12, 8}, {31, 27, 21});
  return sim;
}

bool same_bits(const std::vector<double>& a, const std::vector<double>& b) {
  return a.size() == b.size() && std::memcmp(a.data(), b.data(), a.size() * sizeof(double)) == 0;
}

// Iterate(a) + checkpoint + restart + Iterate(b) against Iterate(a + b).
void restart_matches(const Setup& s, const char* name, int a, int b) {
  const std::string dir = scratch_dir(name);
  std::unique_ptr<AmrSim> whole = make(s);
  whole->Iterate(a + b);
  {
    std::unique_ptr<AmrSim> first = make(s);
    first->Iterate(a);
    first->WriteCheckpoint(dir);
  }
  std::unique_ptr<AmrSim> resumed(new AmrSim(NX, NY, NZ, s.max_level, s.periodic, 0.7, 0.7));
  resumed->ReadCheckpoint(dir);
  REQUIRE(resumed->finestLevel() == whole->finestLevel());
  resumed->Iterate(b);
  for (int l = 0; l <= whole->finestLevel(); ++l) {
    REQUIRE(resumed->boxArray(l) == whole->boxArray(l));
    REQUIRE(resumed->GetTime(l) == whole->GetTime(l));
    REQUIRE(resumed->GetTimeStep(l) == whole->GetTimeStep(l));
    REQUIRE(resumed->GetTauS(l) == whole->GetTauS(l));
    int differ = 0;
    for (int bx = 0; bx < (int)whole->boxArray(l).size(); ++bx)
      if (!same_bits(resumed->DistFn(l, bx), whole->DistFn(l, bx))) ++differ;
    if (differ) std::printf("  level %d: %d boxes differ\n", l, differ);
    REQUIRE(differ == 0);
    whole->CalcHydroVars(l);
    resumed->CalcHydroVars(l);
    const auto e = whole->GetExtent(l);
    for (int i = e.first[0]; i <= e.second[0]; i += 3)
      for (int j = e.first[1]; j <= e.second[1]; j += 2)
        for (int k = e.first[2]; k <= e.second[2]; ++k) {
          const double x = whole->GetDensity(i, j, k, l), y = resumed->GetDensity(i, j, k, l);
          REQUIRE(std::memcmp(&x, &y, sizeof x) == 0);
        }
  }
  std::string cmd = "rm -rf '" + dir + "'";
  REQUIRE(std::system(cmd.c_str()) == 0);
}
}  // namespace

TEST_CASE("Checkpoint single level FP64") {
  Setup s;
  restart_matches(s, "single", 5, 4);
}

TEST_CASE("Checkpoint single level with bounce-back walls") {
  Setup s;
  s.periodic = {1, 0, 1};
  restart_matches(s, "walls", 3, 3);
}

TEST_CASE("Checkpoint two levels, reference mask") {
  Setup s;
  s.max_level = 1;
  restart_matches(s, "two_ref", 3, 2);
}

TEST_CASE("Checkpoint two levels, fine-grid mask") {
  Setup s;
  s.max_level = 1;
  s.fine_mask = true;
  restart_matches(s, "two_fine", 2, 3);
}

TEST_CASE("Checkpoint three levels") {
  Setup s;
  s.max_level = 2;
  s.fine_mask = true;
  restart_matches(s, "three", 2, 2);
}

TEST_CASE("Checkpoint two levels at refinement ratio 3") {
  Setup s;
  s.max_level = 1;
  s.fine_mask = true;
  s.ratio = 3;
  restart_matches(s, "ratio3", 2, 2);
}

TEST_CASE("Checkpoint three levels at refinement ratio 3") {
  Setup s;
  s.max_level = 2;
  s.fine_mask = true;
  s.ratio = 3;
  restart_matches(s, "ratio3_three", 1, 2);
}

TEST_CASE("Checkpoint FP32-shifted two levels") {
  Setup s;
  s.max_level = 1;
  s.fine_mask = true;
  s.precision = LBX_FP32_SHIFTED;
  restart_matches(s, "fp32", 3, 3);
}

TEST_CASE("Checkpoint refuses another domain or a missing directory") {
  Setup s;
  const std::string dir = scratch_dir("refuse");
  {
    std::unique_ptr<AmrSim> sim = make(s);
    sim->WriteCheckpoint(dir);
  }
  std::unique_ptr<AmrSim> other(new AmrSim(NX, NY, NZ + 2, 0, {1, 1, 1}, 0.7, 0.7));
  REQUIRE_THROWS(other->ReadCheckpoint(dir));
  std::unique_ptr<AmrSim> deeper(new AmrSim(NX, NY, NZ, 1, {1, 1, 1}, 0.7, 0.7));
  REQUIRE_THROWS(deeper->ReadCheckpoint(dir));
  std::unique_ptr<AmrSim> none(new AmrSim(NX, NY, NZ, 0, {1, 1, 1}, 0.7, 0.7));
  REQUIRE_THROWS(none->ReadCheckpoint(dir + "_missing"));
  std::string cmd = "rm -rf '" + dir + "'";
  REQUIRE(std::system(cmd.c_str()) == 0);
}

int main() {
  lambrexInit();
  const int rc = check::run_all();
  lambrexFinalise();
  return rc;
}
