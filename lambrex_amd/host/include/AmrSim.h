// AmrSim.h -- LAMBReX's AmrSim API (include/AmrSim.h:23-155 in the
// reference) over the MI355X C ABI (include/lbx.h).
//
// Same public names, argument meaning and state semantics as the reference:
//   - domain [0,nx)x[0,ny)x[0,nz), periodic in each direction whose
//     periodicity flag is set; where the reference aborts on a non-periodic
//     direction (src/AmrSim.cpp:788-797), this build puts halfway
//     bounce-back walls on its faces (single-level simulations only);
//   - up to LBX_MAX_LEVELS levels (max_ref_level < LBX_MAX_LEVELS), ratio 2
//     (or 3 or 4 on every level, SetRefRatio), advanced by the reference's
//     recursive Rohde cycle;
//   - initial density / velocity vectors in C order, i slowest, velocity
//     component fastest (include/AmrSim.h:79-83);
//   - omega = 1/(tau + 1/2) (src/AmrSim.cpp:125-126);
//   - GetDensity returns the density array of the current state, which
//     Iterate swaps every single-level step (include/multilevel.h:75-79), so
//     callers run CalcHydroVars after Iterate as the reference's tests do;
//     GetVelocity returns the last computed velocity;
//   - off-domain queries return the NL_DENSITY / NL_VELOCITY sentinels.
// Differences, all documented where they matter:
//   - errors throw lambrex::Error (carrying the reference's message where it
//     has one) instead of amrex::Abort ending the process;
//   - AMReX's AmrCore is not available: the members callers used from it
//     (InitFromScratch, boxArray, DistributionMap, refRatio, maxLevel,
//     finestLevel) are provided here, with base grids chopped like AMReX's
//     BoxList::maxSize (max_grid_size 32 by default);
//   - the distribution function lives on the GPU; DistFn()/SetDistFn() and
//     Device() expose it for tests and drivers.
#pragma once

#include <array>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "boxes.h"
#include "lbx.h"

#define NDIMS 3
#define NMODES 15

namespace lambrex {

struct Error : std::runtime_error {
  using std::runtime_error::runtime_error;
};

// The process-wide device context created by lambrexInit().
lbx_ctx* Context();

}  // namespace lambrex

class AmrSim {
 public:
  AmrSim(int nx, int ny, int nz, int max_ref_level, const std::array<int, NDIMS>& periodicity,
         double tau_s, double tau_b);
  virtual ~AmrSim();
  AmrSim(const AmrSim&) = delete;
  AmrSim& operator=(const AmrSim&) = delete;

  // ---- public API of the reference (include/AmrSim.h:128-154)
  double GetTime(int level) const { return levels_.at(level).time; }
  int GetTimeStep(int level) const { return levels_.at(level).step; }
  std::array<int, NDIMS> GetDims() const { return {NX, NY, NZ}; }
  bool OnProcessDensity(double rho) const { return rho != NL_DENSITY; }
  bool OnProcessVelocity(double u) const { return u != NL_VELOCITY; }
  void SetInitialDensity(double rho_init);
  void SetInitialDensity(std::vector<double> rho_init);
  void SetInitialVelocity(double u_init);
  void SetInitialVelocity(std::vector<double> u_init);
  double GetDensity(int i, int j, int k, int level) const;
  double GetVelocity(int i, int j, int k, int n, int level) const;
  void CalcEquilibriumDist(int level);
  void CalcHydroVars(int level);
  void Iterate(int nsteps);
  void SetStaticRefinement(int level, const std::array<int, NDIMS>& lo_corner,
                           const std::array<int, NDIMS>& hi_corner);
  void UnsetStaticRefinement(int level);
  std::pair<std::array<int, NDIMS>, std::array<int, NDIMS>> GetExtent(int level) const;

  // ---- members callers used from amrex::AmrCore
  void InitFromScratch(double time);
  const std::vector<lambrex::IBox>& boxArray(int level) const { return levels_.at(level).boxes; }
  // Owner rank of each box of the level: all on this process's GPU (rank 0).
  lambrex::DistributionMapping DistributionMap(int level) const {
    return {std::vector<int>(levels_.at(level).boxes.size(), 0)};
  }
  int maxLevel() const { return max_level_; }
  int finestLevel() const { return finest_level_; }
  std::vector<std::array<int, NDIMS>> refRatio() const;
  std::array<int, NDIMS> refRatio(int level) const;

  // ---- extensions (no reference counterpart)
  // Largest box edge used when the base grids are chopped (AMReX default 32).
  void SetMaxGridSize(int max_grid_size);
  // LBX_FP64 (default) or LBX_FP32_SHIFTED with reference density rho0.
  void SetPrecision(int precision, double rho0);
  // Refinement ratio between consecutive levels, the same on every level: 2
  // (the reference's, fixed in its constructor, src/AmrSim.cpp:766), 3 or 4.
  // Before InitFromScratch.  A level then takes `ratio` steps per step of
  // its parent (dt / ratio, tau rescaled by ComputeDt at that ratio), and
  // static refinement and Regrid make fine boxes aligned to the ratio.
  void SetRefRatio(int ratio);
  // Mask the coarse collide with the fine BoxArray instead of the coarse one
  // the reference passes to makeFineMask (src/AmrSim.cpp:419-428).
  void SetFineMaskFromFineGrids(bool on);
  // Dynamic regrid on device data, level by level from 0 to max_level-1:
  // tag the cells where `criterion` (LBX_TAG_DENSITY_GRADIENT or
  // LBX_TAG_SPEED) exceeds `threshold`, grow the tags by `buffer` cells,
  // refine the tiles of `blocking` cells that hold one (on levels >= 1 only
  // tiles whose cells have their whole 3x3x3 neighbourhood in the level, so
  // the next level is properly nested), and make, remake (keeping fine data)
  // or clear the next level like AmrCore::regrid.  Replaces every static
  // refinement box.
  void Regrid(int criterion, double threshold, int blocking = 4, int buffer = 1);
  // Bulk output: density and velocity of every level (CalcHydroVars first)
  // written to `dir` in AMReX's plotfile layout: a HyperCLaw-V1.1 Header,
  // Level_<l>/Cell_H (VisMF header) and one Level_<l>/Cell_D_00000 holding
  // every box as a FAB of 4 components (density, x/y/z velocity),
  // little-endian IEEE doubles, x fastest.  The problem domain is the
  // reference's RealBox [0,1]^3.
  void WritePlotfile(const std::string& dir);
  // Checkpoint and restart (the reference has no I/O).  WriteCheckpoint
  // writes to `dir`: a text Header (format tag, domain, periodicity,
  // max/finest level, precision and rho0, max_grid_size, mask choice,
  // static refinement boxes, and per level its boxes, time, dt, step,
  // tau_s, tau_b, mass and density-swap parity, doubles as C hexfloats; the
  // refinement ratio), and
  // per level Level_<l>/DistFn (every box's populations of the current
  // state, [15][nz][ny][nx] little-endian FP64, boxes in order) and
  // Level_<l>/Hydro (both density arrays and the velocity of every box).
  // ReadCheckpoint takes the place of InitFromScratch on an AmrSim made with
  // the same domain and periodicity: it rebuilds every level and its masks
  // and uploads the populations, after which Iterate continues as the
  // original would have, bit for bit (FP64; FP32-shifted storage restores
  // the stored deviations exactly wherever f - w rho0 is representable).
  void WriteCheckpoint(const std::string& dir) const;
  void ReadCheckpoint(const std::string& dir);
  lbx_sim* Device() const { return sim_; }
  // Populations of the current state of one box, [15][nz][ny][nx] FP64.
  std::vector<double> DistFn(int level, int box) const;
  void SetDistFn(int level, int box, const std::vector<double>& f);
  // {sum rho, sum rho u_x, sum rho u_y, sum rho u_z} over the level, FP64.
  std::array<double, 4> Conservation(int level) const;
  double GetTauS(int level) const { return tau_s_.at(level); }
  double GetTauB(int level) const { return tau_b_.at(level); }
  double GetDt(int level) const { return levels_.at(level).dt; }
  double GetMass(int level) const { return mass_.at(level); }

  static constexpr double CS2 = 1.0 / 3.0;
  static constexpr double NL_DENSITY = -1.0;
  static constexpr double NL_VELOCITY = -3E8;

 protected:
  struct LevelHost {
    std::vector<lambrex::IBox> boxes;
    // Host copies of the derived fields, per box, x fastest.  Density has
    // two arrays because the reference swaps it with the state
    // (include/multilevel.h:75-79); rho[cur] is the current one.
    std::vector<std::vector<double>> rho[2];
    std::vector<std::vector<double>> vel;  // [3][n] per box
    int cur = 0;
    double time = 0.0, dt = 0.0;
    int step = 0;
  };

  void ComputeDt(int level);
  int Ratio(int level) const;
  void MakeNewLevelFromScratch(int level, double time, const std::vector<lambrex::IBox>& ba);
  void MakeNewLevelFromCoarse(int level, double time, const std::vector<lambrex::IBox>& ba);
  void RemakeLevel(int level, double time, const std::vector<lambrex::IBox>& ba);
  void MakeFineMask(int coarse_level);
  void ClearLevel(int level);
  // AmrCore::regrid(base): levels base+1 .. max from the static boxes of
  // levels base .. max-1.
  void RegridFrom(int base);
  std::vector<lambrex::IBox> RefineFromState(int level, int criterion, double threshold, int blocking, int buffer);
  void InitDensity(int level);
  void InitVelocity(int level);
  void Check(int rc) const;
  int BoxOf(int level, int i, int j, int k) const;
  mutable std::vector<int> box_hint_;  // last box BoxOf found, per level
  long long LocalIndex(const lambrex::IBox& b, int i, int j, int k) const;

  const int NX, NY, NZ, NUMEL;
  std::array<int, NDIMS> PERIODICITY;
  int max_level_;
  int finest_level_ = 0;
  int max_grid_size_ = 32;
  int ref_ratio_ = 2;
  int precision_ = LBX_FP64;
  double rho0_ = 1.0;
  bool mask_from_fine_grids_ = false;
  std::vector<lambrex::IBox> static_tags_;  // per level (TagCell, src/AmrSim.cpp:413-417)
  std::vector<double> initial_density;
  std::vector<double> initial_velocity;
  std::vector<double> tau_s_, tau_b_, mass_;
  std::vector<LevelHost> levels_;
  lbx_sim* sim_ = nullptr;
};
