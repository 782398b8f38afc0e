// lbx_regrid.cpp -- clustering of refinement tags into fine boxes
// (lbx_cluster_tags of include/lbx.h).  Host code only.
//
// The reference regrids through AMReX (AmrCore::regrid: ErrorEst tags,
// an error buffer, Berger-Rigoutsos clustering under a blocking factor,
// max_grid_size chopping).  Here the clustering is deliberately simple and
// deterministic: tags are grown by the buffer (periodic), the coarse domain
// is cut into tiles of `blocking` cells, a tile is refined if it holds a
// grown tag, and tagged tiles are merged into boxes -- runs along x, then
// equal runs along y, then equal rectangles along z.  The boxes do not
// overlap and cover every grown tag; they are refined by 2 and chopped at
// max_grid_size with AMReX's BoxList::maxSize rule.  With an `allowed` mask,
// a tile holding any cell that is not allowed is not refined (proper nesting
// of a level inside a patched parent).
#include <algorithm>
#include <cstdint>
#include <vector>

#include "../../include/lbx.h"

namespace {

struct B3 {
  int lo[3], hi[3];
};

// AMReX BoxList::maxSize: pieces of at most mgs cells, sizes as equal as the
// largest common power of two of length and mgs allows, cut from the high
// end.
void chop_dir(std::vector<B3>& boxes, int d, int mgs) {
  std::vector<B3> out;
  for (B3 bx : boxes) {
    const int len = bx.hi[d] - bx.lo[d] + 1;
    if (len > mgs) {
      int ratio = 1, bs = mgs, nlen = len;
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
        B3 top = bx;
        top.lo[d] = bx.hi[d] - ksize + 1;
        out.push_back(top);
        bx.hi[d] = top.lo[d] - 1;
      }
    }
    out.push_back(bx);
  }
  boxes.swap(out);
}

}  // namespace

extern "C" int lbx_cluster_tags_allowed(const int32_t dims[3], const uint8_t* tags, const uint8_t* allowed,
                                        int32_t blocking, int32_t buffer, int32_t max_grid_size, int32_t max_boxes,
                                        int32_t* lohi, int32_t* n_boxes) {
  if (!dims || !tags || !n_boxes || blocking < 1 || buffer < 0 || max_grid_size < 2 || max_grid_size % 2 ||
      max_boxes < 0 || (max_boxes > 0 && !lohi))
    return LBX_E_ARG;
  const int nx = dims[0], ny = dims[1], nz = dims[2];
  if (nx < 1 || ny < 1 || nz < 1 || (long long)nx * ny * nz >= (1ll << 31)) return LBX_E_ARG;
  auto at = [&](int x, int y, int z) { return ((long long)z * ny + y) * nx + x; };
  // tiles holding a tag grown by `buffer` cells (periodic)
  const int tx = (nx + blocking - 1) / blocking, ty = (ny + blocking - 1) / blocking,
            tz = (nz + blocking - 1) / blocking;
  std::vector<uint8_t> tile((size_t)tx * ty * tz, 0);
  auto wrap = [](int v, int n) { return ((v % n) + n) % n; };
  for (int z = 0; z < nz; ++z)
    for (int y = 0; y < ny; ++y)
      for (int x = 0; x < nx; ++x) {
        if (!tags[at(x, y, z)]) continue;
        for (int dz = -buffer; dz <= buffer; ++dz)
          for (int dy = -buffer; dy <= buffer; ++dy)
            for (int dx = -buffer; dx <= buffer; ++dx) {
              const int gx = wrap(x + dx, nx), gy = wrap(y + dy, ny), gz = wrap(z + dz, nz);
              tile[((size_t)(gz / blocking) * ty + gy / blocking) * tx + gx / blocking] = 1;
            }
      }
  // a tile holding a cell that is not allowed is never refined
  if (allowed)
    for (int z = 0; z < nz; ++z)
      for (int y = 0; y < ny; ++y)
        for (int x = 0; x < nx; ++x)
          if (!allowed[at(x, y, z)]) tile[((size_t)(z / blocking) * ty + y / blocking) * tx + x / blocking] = 0;
  // runs along x, merged along y, then along z
  struct Rect {
    int x0, x1, y0, y1;
  };
  std::vector<B3> coarse;
  std::vector<std::pair<Rect, int>> open;  // rectangles of the previous z-tile and their start z
  auto emit = [&](const Rect& r, int z0, int z1) {
    B3 b;
    b.lo[0] = r.x0 * blocking;
    b.hi[0] = std::min(nx, (r.x1 + 1) * blocking) - 1;
    b.lo[1] = r.y0 * blocking;
    b.hi[1] = std::min(ny, (r.y1 + 1) * blocking) - 1;
    b.lo[2] = z0 * blocking;
    b.hi[2] = std::min(nz, (z1 + 1) * blocking) - 1;
    coarse.push_back(b);
  };
  for (int k = 0; k < tz; ++k) {
    std::vector<Rect> rects;
    std::vector<std::pair<int, int>> prev_runs;
    std::vector<int> prev_start;
    for (int j = 0; j < ty; ++j) {
      std::vector<std::pair<int, int>> runs;
      for (int i = 0; i < tx;) {
        if (!tile[((size_t)k * ty + j) * tx + i]) {
          ++i;
          continue;
        }
        int e = i;
        while (e + 1 < tx && tile[((size_t)k * ty + j) * tx + e + 1]) ++e;
        runs.push_back({i, e});
        i = e + 1;
      }
      // a run continues a rectangle if the row above had the same run
      std::vector<int> start(runs.size());
      for (size_t r = 0; r < runs.size(); ++r) {
        start[r] = j;
        for (size_t p = 0; p < prev_runs.size(); ++p)
          if (prev_runs[p] == runs[r]) start[r] = prev_start[p];
      }
      for (size_t p = 0; p < prev_runs.size(); ++p)
        if (std::find(runs.begin(), runs.end(), prev_runs[p]) == runs.end())
          rects.push_back({prev_runs[p].first, prev_runs[p].second, prev_start[p], j - 1});
      prev_runs = runs;
      prev_start = start;
    }
    for (size_t p = 0; p < prev_runs.size(); ++p)
      rects.push_back({prev_runs[p].first, prev_runs[p].second, prev_start[p], ty - 1});
    // a rectangle continues a box if the previous z-tile had the same one
    std::vector<std::pair<Rect, int>> next;
    for (const Rect& r : rects) {
      int z0 = k;
      for (const auto& o : open)
        if (o.first.x0 == r.x0 && o.first.x1 == r.x1 && o.first.y0 == r.y0 && o.first.y1 == r.y1) z0 = o.second;
      next.push_back({r, z0});
    }
    for (const auto& o : open) {
      bool continued = false;
      for (const Rect& r : rects)
        continued |= o.first.x0 == r.x0 && o.first.x1 == r.x1 && o.first.y0 == r.y0 && o.first.y1 == r.y1;
      if (!continued) emit(o.first, o.second, k - 1);
    }
    open.swap(next);
  }
  for (const auto& o : open) emit(o.first, o.second, tz - 1);
  // refine by 2, chop at max_grid_size
  std::vector<B3> fine;
  for (const B3& c : coarse) {
    B3 f;
    for (int d = 0; d < 3; ++d) {
      f.lo[d] = 2 * c.lo[d];
      f.hi[d] = 2 * c.hi[d] + 1;
    }
    fine.push_back(f);
  }
  for (int d = 0; d < 3; ++d) chop_dir(fine, d, max_grid_size);
  *n_boxes = (int32_t)fine.size();
  if ((int)fine.size() > max_boxes) return max_boxes == 0 ? LBX_OK : LBX_E_ARG;
  for (size_t b = 0; b < fine.size(); ++b)
    for (int d = 0; d < 3; ++d) {
      lohi[6 * b + d] = fine[b].lo[d];
      lohi[6 * b + 3 + d] = fine[b].hi[d];
    }
  return LBX_OK;
}

extern "C" int lbx_cluster_tags(const int32_t dims[3], const uint8_t* tags, int32_t blocking, int32_t buffer,
                                int32_t max_grid_size, int32_t max_boxes, int32_t* lohi, int32_t* n_boxes) {
  return lbx_cluster_tags_allowed(dims, tags, nullptr, blocking, buffer, max_grid_size, max_boxes, lohi, n_boxes);
}
