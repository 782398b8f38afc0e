// boxes.cpp -- host box bookkeeping for the AmrSim facade.
#include "boxes.h"

#include <algorithm>
#include <climits>

namespace lambrex {

namespace {
int floor_div(int a, int r) { return a >= 0 ? a / r : -((-a + r - 1) / r); }
}  // namespace

IBox coarsen(const IBox& b, int r) {
  IBox c;
  for (int d = 0; d < 3; ++d) {
    c.lo[d] = floor_div(b.lo[d], r);
    c.hi[d] = floor_div(b.hi[d], r);
  }
  return c;
}

IBox refine(const IBox& b, int r) {
  IBox f;
  for (int d = 0; d < 3; ++d) {
    f.lo[d] = b.lo[d] * r;
    f.hi[d] = (b.hi[d] + 1) * r - 1;
  }
  return f;
}

std::vector<IBox> max_size_chop(const IBox& domain, int max_grid_size) {
  std::vector<IBox> boxes{domain};
  if (max_grid_size < 1) return boxes;
  for (int d = 0; d < 3; ++d) {
    std::vector<IBox> out;
    for (IBox bx : boxes) {
      const int len = bx.length(d);
      if (len > max_grid_size) {
        int ratio = 1, bs = max_grid_size, nlen = len;
        while (bs % 2 == 0 && nlen % 2 == 0) {
          ratio *= 2;
          bs /= 2;
          nlen /= 2;
        }
        const int numblk = nlen / bs + (nlen % bs ? 1 : 0);
        const int sz = nlen / numblk;
        const int extra = nlen % numblk;
        for (int k = 0; k < numblk - 1; ++k) {
          const int ksize = (k < extra ? sz + 1 : sz) * ratio;
          const int pos = bx.hi[d] - ksize + 1;
          IBox top = bx;
          top.lo[d] = pos;
          out.push_back(top);
          bx.hi[d] = pos - 1;
        }
      }
      out.push_back(bx);
    }
    boxes.swap(out);
  }
  return boxes;
}

IBox minimal_box(const std::vector<IBox>& boxes) {
  IBox m;
  if (boxes.empty()) return m;
  m.lo = {INT_MAX, INT_MAX, INT_MAX};
  m.hi = {INT_MIN, INT_MIN, INT_MIN};
  for (const IBox& b : boxes)
    for (int d = 0; d < 3; ++d) {
      m.lo[d] = std::min(m.lo[d], b.lo[d]);
      m.hi[d] = std::max(m.hi[d], b.hi[d]);
    }
  return m;
}

std::vector<int32_t> to_lohi(const std::vector<IBox>& boxes) {
  std::vector<int32_t> v;
  v.reserve(boxes.size() * 6);
  for (const IBox& b : boxes) {
    for (int d = 0; d < 3; ++d) v.push_back(b.lo[d]);
    for (int d = 0; d < 3; ++d) v.push_back(b.hi[d]);
  }
  return v;
}

}  // namespace lambrex
