// ml_pulse_dump.cpp -- the reference's "ml_pulse Regression" case
// (tests/catch2RegressionTests.cpp:95-196) through this build's AmrSim: the
// 10x10x50 pulse of amr_pulse with max_level 1 and the whole domain refined.
// Prints, for the points the reference test samples, density and velocity
// with 17 digits: both levels at t = 0 (level 1 at ratio * (i, j, k)) and
// level 0 at t = 100; tests/test_gpu_cpp.py compares them with the goldens.
#include <cstdio>
#include <vector>

#include "lambrex.h"

static void dump(const AmrSim& sim, int t, int level, int ratio, int nx, int ny, int nz) {
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny; ++j)
      for (int i = 0; i < nx; ++i) {
        const int I = ratio * i, J = ratio * j, K = ratio * k;
        std::printf("P %d %d %d %.17g %.17g %.17g %.17g\n", t, level, k, sim.GetDensity(I, J, K, level),
                    sim.GetVelocity(I, J, K, 0, level), sim.GetVelocity(I, J, K, 1, level),
                    sim.GetVelocity(I, J, K, 2, level));
      }
}

int main() {
  lambrexInit();
  int rc = 0;
  try {
    const int NX = 10, NY = 10, NZ = 50, NUMEL = NX * NY * NZ;
    const double TAU = 0.5, AMPLITUDE = 0.01;
    std::vector<double> rho_0(NUMEL, 1.0);
    for (int i = 0; i < NX; ++i)
      for (int j = 0; j < NY; ++j) rho_0[i * NY * NZ + j * NZ + NZ / 2 - 1] += AMPLITUDE;
    double z_mean = 0.0;
    for (int k = 0; k < NZ; ++k) z_mean += rho_0[(NUMEL + NY * NZ) / 2 + k];
    z_mean /= NZ;
    for (double& r : rho_0) r /= z_mean;
    AmrSim sim(NX, NY, NZ, 1, {1, 1, 1}, TAU, TAU);
    sim.SetInitialDensity(rho_0);
    sim.SetInitialVelocity(0.0);
    sim.InitFromScratch(0.0);
    sim.SetStaticRefinement(0, {0, 0, 0}, {NX - 1, NY - 1, NZ - 1});
    dump(sim, 0, 0, 1, NX, NY, NZ);
    dump(sim, 0, 1, 2, NX, NY, NZ);
    sim.Iterate(100);
    sim.CalcHydroVars(0);
    dump(sim, 100, 0, 1, NX, NY, NZ);
  } catch (const std::exception& e) {
    std::printf("error: %s\n", e.what());
    rc = 1;
  }
  lambrexFinalise();
  return rc;
}
