// amr_pulse.cpp -- the reference's example driver (examples/amr_pulse.cpp:7-58)
// against this build's AmrSim: a 10x10x50 periodic density pulse, tau 0.5,
// density and velocity printed at t = 0, 100 and 200 with precision 6 in the
// reference's print format (examples/amr_pulse.cpp:60-108), which is the
// format of the reference's golden arrays.
#include <iostream>

#include "lambrex.h"

static void printDensity(const AmrSim& sim) {
  const std::array<int, NDIMS> DIMS = sim.GetDims();
  const int NUMEL = DIMS[0] * DIMS[1] * DIMS[2];
  const int LEVEL = 0;
  std::cout.precision(6);
  std::cout << "Time: " << sim.GetTime(LEVEL) << std::endl;
  std::cout << "Density:\n{ ";
  for (int k = 0; k < DIMS[2]; ++k)
    for (int j = 0; j < DIMS[1]; ++j)
      for (int i = 0; i < DIMS[0]; ++i) {
        const bool last = !(k * DIMS[0] * DIMS[1] + j * DIMS[0] + i < NUMEL - 1);
        std::cout << sim.GetDensity(i, j, k, LEVEL) << (last ? " };" : ", ");
      }
  std::cout << std::endl;
}

static void printVelocity(const AmrSim& sim) {
  const std::array<int, NDIMS> DIMS = sim.GetDims();
  const int NUMEL = DIMS[0] * DIMS[1] * DIMS[2] * NDIMS;
  const int LEVEL = 0;
  std::cout.precision(6);
  std::cout << "Time step: " << sim.GetTimeStep(LEVEL) << std::endl;
  std::cout << "Velocity:\n{ ";
  for (int k = 0; k < DIMS[2]; ++k)
    for (int j = 0; j < DIMS[1]; ++j)
      for (int i = 0; i < DIMS[0]; ++i)
        for (int n = 0; n < NDIMS; ++n) {
          const bool last =
              !(n * DIMS[0] * DIMS[1] * DIMS[2] + k * DIMS[0] * DIMS[1] + j * DIMS[0] + i < NUMEL - 1);
          std::cout << sim.GetVelocity(i, j, k, n, LEVEL) << (last ? " };" : ", ");
        }
  std::cout << std::endl;
}

int main() {
  const int nx = 10, ny = 10, nz = 50;
  const double tau = 0.5;
  const double amplitude = 0.01;
  const std::array<int, 3> periodicity = {1, 1, 1};
  const int max_level = 0;
  const int numel = nx * ny * nz;

  std::vector<double> rho_0(numel, 1.0);
  int k = nz / 2;
  for (int i = 0; i < nx; ++i)
    for (int j = 0; j < ny; ++j) rho_0[i * ny * nz + j * nz + k - 1] += amplitude;
  double z_mean = 0.0;
  for (k = 0; k < nz; ++k) z_mean += rho_0[(numel + ny * nz) / 2 + k];
  z_mean /= nz;
  for (int i = 0; i < numel; ++i) rho_0[i] /= z_mean;

  lambrexInit();
  {
    AmrSim lbrx(nx, ny, nz, max_level, periodicity, tau, tau);
    lbrx.SetInitialDensity(rho_0);
    lbrx.SetInitialVelocity(0.0);
    lbrx.InitFromScratch(0.0);
    printDensity(lbrx);
    printVelocity(lbrx);
    lbrx.Iterate(100);
    lbrx.CalcHydroVars(0);
    printDensity(lbrx);
    printVelocity(lbrx);
    lbrx.Iterate(100);
    lbrx.CalcHydroVars(0);
    printDensity(lbrx);
    printVelocity(lbrx);
  }
  lambrexFinalise();
  return 0;
}
