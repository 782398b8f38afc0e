// plotfile_dump.cpp -- writes a two-level state with AmrSim::WritePlotfile
// to argv[1] and prints sampled GetDensity / GetVelocity values, which
// tests/test_gpu_cpp.py compares with what it reads back from the files.
#include <cstdio>
#include <vector>

#include "lambrex.h"

int main(int argc, char** argv) {
  if (argc != 2) {
    std::fprintf(stderr, "usage: plotfile_dump <dir>\n");
    return 2;
  }
  lambrexInit();
  int rc = 0;
  try {
    const int NX = 24, NY = 16, NZ = 12;
    AmrSim sim(NX, NY, NZ, 1, {1, 1, 1}, 0.8, 0.8);
    std::vector<double> rho((size_t)NX * NY * NZ, 1.0);
    for (int i = 0; i < NX; ++i)
      for (int j = 0; j < NY; ++j)
        for (int k = 0; k < NZ; ++k) rho[((size_t)i * NY + j) * NZ + k] = 1.0 + 0.01 * (i == 9) + 0.001 * (j % 3);
    sim.SetInitialDensity(rho);
    sim.SetInitialVelocity(0.0);
    sim.SetMaxGridSize(16);
    sim.InitFromScratch(0.0);
    sim.SetFineMaskFromFineGrids(true);  // the reference's own mask zeroes the coarse low octant
    sim.SetStaticRefinement(0, {6, 0, 0}, {13, NY - 1, NZ - 1});
    sim.Iterate(3);
    sim.WritePlotfile(argv[1]);
    // samples: level i j k rho ux uy uz
    const int pts[][4] = {{0, 0, 0, 0}, {0, 9, 5, 3}, {0, 23, 15, 11}, {1, 12, 0, 0}, {1, 18, 17, 9}, {1, 27, 31, 23}};
    for (const auto& p : pts)
      std::printf("SAMPLE %d %d %d %d %.17g %.17g %.17g %.17g\n", p[0], p[1], p[2], p[3],
                  sim.GetDensity(p[1], p[2], p[3], p[0]), sim.GetVelocity(p[1], p[2], p[3], 0, p[0]),
                  sim.GetVelocity(p[1], p[2], p[3], 1, p[0]), sim.GetVelocity(p[1], p[2], p[3], 2, p[0]));
  } catch (const std::exception& e) {
    std::printf("error: %s\n", e.what());
    rc = 1;
  }
  lambrexFinalise();
  return rc;
}
