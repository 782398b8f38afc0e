// boxes.h -- the small host-side box bookkeeping the AmrSim facade needs in
// place of AMReX's Box / BoxArray (AMReX is not part of this build).
#pragma once

#include <array>
#include <cstdint>
#include <vector>

namespace lambrex {

struct IBox {
  std::array<int, 3> lo{0, 0, 0};
  std::array<int, 3> hi{-1, -1, -1};

  bool ok() const { return hi[0] >= lo[0] && hi[1] >= lo[1] && hi[2] >= lo[2]; }
  int length(int d) const { return hi[d] - lo[d] + 1; }
  long long numPts() const {
    return ok() ? (long long)length(0) * length(1) * length(2) : 0;
  }
  bool contains(int i, int j, int k) const {
    return i >= lo[0] && i <= hi[0] && j >= lo[1] && j <= hi[1] && k >= lo[2] && k <= hi[2];
  }
  bool contains(const IBox& b) const {
    return contains(b.lo[0], b.lo[1], b.lo[2]) && contains(b.hi[0], b.hi[1], b.hi[2]);
  }
  bool operator==(const IBox& o) const { return lo == o.lo && hi == o.hi; }
};

// AMReX's DistributionMapping as far as the reference's callers use it
// (tests/catch2AMRTests.cpp:27, 70, 97 pass it on beside boxArray(l)): the
// rank owning each box of a level.  Every box lives on the one GPU of the
// process, rank 0.
struct DistributionMapping {
  std::vector<int> owner;
  int size() const { return (int)owner.size(); }
  int operator[](int box) const { return owner.at(box); }
  const std::vector<int>& ProcessorMap() const { return owner; }
};

// Cells of the coarse level covering a fine box (ratio r), and back.
IBox coarsen(const IBox& b, int r);
IBox refine(const IBox& b, int r);

// Split a box into pieces no longer than max_grid_size in any direction.
// Same cut positions as AMReX's BoxList::maxSize, which the reference gets
// through AmrCore::MakeBaseGrids (blocking factor 1, one rank): per
// direction, common powers of two are factored out, the rest is split into
// near-equal blocks (larger ones first) cut from the top.
std::vector<IBox> max_size_chop(const IBox& domain, int max_grid_size);

// Smallest box containing all boxes (BoxArray::minimalBox).
IBox minimal_box(const std::vector<IBox>& boxes);

// 6 int32 per box (lo x,y,z, hi x,y,z) as the C ABI takes them.
std::vector<int32_t> to_lohi(const std::vector<IBox>& boxes);

}  // namespace lambrex
