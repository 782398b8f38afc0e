// three_level_dump.cpp -- a three-level run through the AmrSim facade only
// (static refinement of levels 0 and 1, the reference's own fine masks),
// written with AmrSim::WritePlotfile to argv[1]; tests/test_gpu_cpp.py
// replays the same run on the oracle from the boxes in the plotfile and
// compares every level's density and velocity bitwise.
#include <cstdio>
#include <vector>

#include "lambrex.h"

int main(int argc, char** argv) {
  if (argc != 2) {
    std::fprintf(stderr, "usage: three_level_dump <dir>\n");
    return 2;
  }
  lambrexInit();
  int rc = 0;
  try {
    const int NX = 16, NY = 12, NZ = 8;
    AmrSim sim(NX, NY, NZ, 2, {1, 1, 1}, 0.7, 0.9);
    std::vector<double> rho((size_t)NX * NY * NZ), u(3 * (size_t)NX * NY * NZ);
    for (int i = 0; i < NX; ++i)
      for (int j = 0; j < NY; ++j)
        for (int k = 0; k < NZ; ++k) {
          const size_t c = ((size_t)i * NY + j) * NZ + k;  // C order, i slowest
          rho[c] = 1.0 + 0.01 * (i == 7) + 0.001 * ((j + 2 * k) % 5);
          u[3 * c + 0] = 0.002 * ((i + k) % 3);
          u[3 * c + 1] = -0.001 * (j % 2);
          u[3 * c + 2] = 0.0005 * ((i * j) % 4);
        }
    sim.SetInitialDensity(rho);
    sim.SetInitialVelocity(u);
    sim.SetMaxGridSize(8);
    sim.InitFromScratch(0.0);
    sim.SetStaticRefinement(0, {4, 2, 0}, {11, 9, 7});
    sim.SetStaticRefinement(1, {10, 6, 2}, {21, 15, 13});
    sim.Iterate(2);
    sim.WritePlotfile(argv[1]);
    std::printf("FINEST %d\n", sim.finestLevel());
  } catch (const std::exception& e) {
    std::printf("error: %s\n", e.what());
    rc = 1;
  }
  lambrexFinalise();
  return rc;
}
