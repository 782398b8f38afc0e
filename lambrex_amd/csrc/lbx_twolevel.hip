// lbx_twolevel.hip -- device kernels of the two-level Rohde cycle
// (src/AmrSim.cpp:430-631) on gfx950, FP64.
//
// Storage.  The coarse level is a periodic level array (lbx_lattice.hip).
// Each fine box is a patch of its own, grown by 2 ghost layers like the
// reference's FABs (population-major, x fastest, per-patch population
// stride).  The fine ghosts must be separate per box because the reference
// lets them go stale between the two fine substeps (they are filled once per
// coarse step and never collided), so a cell seen as a ghost by one box and
// as valid by its neighbour evolves differently in the two places.
//
// One coarse step, in the reference's order of effects:
//   k_fine_fill_collide   FillPatchTwoLevels(fine next <- fine now | coarse
//                         now, piecewise constant) + zero comp 0 on the outer
//                         ghost layer + collide valid fine cells
//   k_fine_stream<0>      Stream(1) (valid + 1 ghost layer; outer layer of the
//                         fresh array defined as 0) + collide valid cells
//   k_fine_stream<1>      Stream(1) + ZeroInvalidComponents(1)
//   k_coarse_push         CoarseCollide(0) (masked cells -> 0) + Stream(0) in
//                         push form; pulls across a box face read the
//                         pre-collision ghost copies, as the reference's
//                         RohdeCycle does (no FillBoundary between collide and
//                         stream)
//   k_sum_fine_to_coarse  SumFromFine(0): ADD the 2x2x2 average of fine
//                         (valid + ghosts) into coarse next, contributions to
//                         one coarse cell in a fixed order
#include <hip/hip_runtime.h>

#include "d3q15.h"
#include "lbx_internal.h"

#pragma clang fp contract(off)

namespace lbx {
namespace {

__device__ __forceinline__ void patch_coords(const PatchDesc& P, int i, int& gx, int& gy, int& gz) {
  gx = i % P.g[0];
  const int t = i / P.g[0];
  gy = t % P.g[1];
  gz = t / P.g[1];
}

__device__ __forceinline__ bool patch_valid(const PatchDesc& P, int gx, int gy, int gz) {
  return gx >= 2 && gx < P.g[0] - 2 && gy >= 2 && gy < P.g[1] - 2 && gz >= 2 && gz < P.g[2] - 2;
}

__device__ __forceinline__ bool patch_outer(const PatchDesc& P, int gx, int gy, int gz) {
  return gx == 0 || gy == 0 || gz == 0 || gx == P.g[0] - 1 || gy == P.g[1] - 1 || gz == P.g[2] - 1;
}

// valid-region membership of absolute position p for patch P
__device__ __forceinline__ bool in_valid(const PatchDesc& P, int x, int y, int z) {
  return x >= P.lo[0] && x <= P.hi[0] && y >= P.lo[1] && y <= P.hi[1] && z >= P.lo[2] && z <= P.hi[2];
}

__global__ __launch_bounds__(256) void k_fine_fill_collide(
    const double* __restrict__ fnow, double* __restrict__ fnext, const double* __restrict__ cnow,
    Lattice clat, const PatchDesc* __restrict__ patches, const long long* __restrict__ fill,
    Relax r) {
  const PatchDesc P = patches[blockIdx.y];
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= P.cells) return;
  int gx, gy, gz;
  patch_coords(P, i, gx, gy, gz);
  double f[NV];
  if (patch_valid(P, gx, gy, gz)) {
#pragma unroll
    for (int q = 0; q < NV; ++q) f[q] = fnow[P.base + q * P.qs + i];
    collide_f64(f, r.omega_s, r.omega_b);
  } else {
    const long long d = fill[P.fill_base + i];
    if (d >= 0) {
      const PatchDesc S = patches[d >> 32];
      const int li = int(d & 0xffffffffll);
#pragma unroll
      for (int q = 0; q < NV; ++q) f[q] = fnow[S.base + q * S.qs + li];
    } else {
      const long long c = -1 - d;  // coarse element of population 0
#pragma unroll
      for (int q = 0; q < NV; ++q) f[q] = cnow[c + q * clat.ps];
    }
    if (patch_outer(P, gx, gy, gz)) f[0] = 0.0;
  }
#pragma unroll
  for (int q = 0; q < NV; ++q) fnext[P.base + q * P.qs + i] = f[q];
}

// MODE 0: pull over valid + 1 ghost layer, collide valid cells.
// MODE 1: pull over valid + 1 ghost layer, then keep ghost population m only
//         if pos - 2 c_m lies in the valid box.
// The outer ghost layer is written as zero in both modes.
template <int MODE>
__global__ __launch_bounds__(256) void k_fine_stream(const double* __restrict__ in,
                                                     double* __restrict__ out,
                                                     const PatchDesc* __restrict__ patches,
                                                     Relax r) {
  const PatchDesc P = patches[blockIdx.y];
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= P.cells) return;
  int gx, gy, gz;
  patch_coords(P, i, gx, gy, gz);
  double f[NV];
  if (patch_outer(P, gx, gy, gz)) {
#pragma unroll
    for (int q = 0; q < NV; ++q) f[q] = 0.0;
  } else {
    const int sx = 1, sy = P.g[0], sz = P.g[0] * P.g[1];
#pragma unroll
    for (int q = 0; q < NV; ++q)
      f[q] = in[P.base + q * P.qs + i - kC[q][0] * sx - kC[q][1] * sy - kC[q][2] * sz];
    const bool valid = patch_valid(P, gx, gy, gz);
    if (MODE == 0 && valid) collide_f64(f, r.omega_s, r.omega_b);
    if (MODE == 1 && !valid) {
      const int x = P.lo[0] - 2 + gx, y = P.lo[1] - 2 + gy, z = P.lo[2] - 2 + gz;
#pragma unroll
      for (int q = 0; q < NV; ++q)
        if (!in_valid(P, x - 2 * kC[q][0], y - 2 * kC[q][1], z - 2 * kC[q][2])) f[q] = 0.0;
    }
  }
#pragma unroll
  for (int q = 0; q < NV; ++q) out[P.base + q * P.qs + i] = f[q];
}

__device__ __forceinline__ int wrap(int v, int n) { return v < 0 ? v + n : (v >= n ? v - n : v); }

// One workgroup row per coarse box: blockIdx.y = box.
__global__ __launch_bounds__(256) void k_coarse_push(const double* __restrict__ cnow,
                                                     double* __restrict__ cnext, Lattice L,
                                                     const PatchDesc* __restrict__ boxes,
                                                     const unsigned char* __restrict__ mask,
                                                     Relax r) {
  const PatchDesc B = boxes[blockIdx.y];
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= B.cells) return;
  const int nx = B.hi[0] - B.lo[0] + 1, ny = B.hi[1] - B.lo[1] + 1;
  const int x = B.lo[0] + i % nx;
  const int t = i / nx;
  const int y = B.lo[1] + t % ny;
  const int z = B.lo[2] + t / ny;
  const long long self = L.at(0, y * L.nx + x, z);
  double pre[NV], post[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) pre[q] = post[q] = cnow[self + q * L.ps];
  if (mask[((long long)z * L.ny + y) * L.nx + x]) {
#pragma unroll
    for (int q = 0; q < NV; ++q) post[q] = 0.0;
  } else {
    collide_f64(post, r.omega_s, r.omega_b);
  }
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    const int dx = x + kC[q][0], dy = y + kC[q][1], dz = z + kC[q][2];
    const bool same_box = in_valid(B, dx, dy, dz);
    const long long dest = L.at(q, wrap(dy, L.ny) * L.nx + wrap(dx, L.nx), wrap(dz, L.nz));
    cnext[dest] = same_box ? post[q] : pre[q];
  }
}

// Thread per touched coarse cell; contributions [start[u], start[u+1]) in
// the order the reference's loops would add them.
__global__ __launch_bounds__(256) void k_sum_fine_to_coarse(
    const double* __restrict__ fnow, double* __restrict__ cnext, Lattice L,
    const PatchDesc* __restrict__ patches, const long long* __restrict__ cell,
    const int* __restrict__ start, const int4* __restrict__ contrib, int ncell) {
  const int u = blockIdx.x * 256 + threadIdx.x;
  if (u >= ncell) return;
  const long long c = cell[u];  // coarse element of population 0
  double acc[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) acc[q] = cnext[c + q * L.ps];
  for (int j = start[u]; j < start[u + 1]; ++j) {
    const int4 k = contrib[j];  // box, unwrapped coarse position
    const PatchDesc P = patches[k.x];
    // fine child (2ci + di, ...) -> grown-patch local index
    const int gx0 = 2 * k.y - (P.lo[0] - 2), gy0 = 2 * k.z - (P.lo[1] - 2), gz0 = 2 * k.w - (P.lo[2] - 2);
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      double s = 0.0;
#pragma unroll
      for (int dk = 0; dk < 2; ++dk)
#pragma unroll
        for (int dj = 0; dj < 2; ++dj)
#pragma unroll
          for (int di = 0; di < 2; ++di)
            s += fnow[P.base + q * P.qs + ((gz0 + dk) * P.g[1] + (gy0 + dj)) * P.g[0] + (gx0 + di)];
      s *= 0.125;
      acc[q] += s;
    }
  }
#pragma unroll
  for (int q = 0; q < NV; ++q) cnext[c + q * L.ps] = acc[q];
}

// DistFnFillFromCoarse (src/AmrSim.cpp:393-411): valid fine cells take the
// coarse value of their parent; ghosts are zeroed (they are rewritten before
// they are read).
__global__ __launch_bounds__(256) void k_fine_from_coarse(double* __restrict__ fnow,
                                                          const double* __restrict__ cnow,
                                                          Lattice clat,
                                                          const PatchDesc* __restrict__ patches) {
  const PatchDesc P = patches[blockIdx.y];
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= P.cells) return;
  int gx, gy, gz;
  patch_coords(P, i, gx, gy, gz);
  if (patch_valid(P, gx, gy, gz)) {
    const int x = (P.lo[0] - 2 + gx) >> 1, y = (P.lo[1] - 2 + gy) >> 1, z = (P.lo[2] - 2 + gz) >> 1;
    const long long c = clat.at(0, y * clat.nx + x, z);
#pragma unroll
    for (int q = 0; q < NV; ++q) fnow[P.base + q * P.qs + i] = cnow[c + q * clat.ps];
  } else {
#pragma unroll
    for (int q = 0; q < NV; ++q) fnow[P.base + q * P.qs + i] = 0.0;
  }
}

// Valid cells of one patch <-> packed FP64 [15][nz][ny][nx]; hydro of one patch.
template <int OP>  // 0 pack, 1 unpack, 2 hydro
__global__ __launch_bounds__(256) void k_patch_io(double* __restrict__ f, const PatchDesc* __restrict__ patches,
                                                  int b, double* __restrict__ a, double* __restrict__ u) {
  const PatchDesc P = patches[b];
  const int nx = P.hi[0] - P.lo[0] + 1, ny = P.hi[1] - P.lo[1] + 1, nz = P.hi[2] - P.lo[2] + 1;
  const int n = nx * ny * nz;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int x = i % nx, t = i / nx, y = t % ny, z = t / ny;
  const long long li = ((long long)(z + 2) * P.g[1] + (y + 2)) * P.g[0] + (x + 2);
  if (OP == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) a[(long long)q * n + i] = f[P.base + q * P.qs + li];
  } else if (OP == 1) {
#pragma unroll
    for (int q = 0; q < NV; ++q) f[P.base + q * P.qs + li] = a[(long long)q * n + i];
  } else {
    double g[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) g[q] = f[P.base + q * P.qs + li];
    double rr, u0, u1, u2;
    hydro_f64(g, rr, u0, u1, u2);
    a[i] = rr;
    u[i] = u0;
    u[n + i] = u1;
    u[2 * n + i] = u2;
  }
}

}  // namespace

hipError_t launch_fine_fill_collide(const double* fnow, double* fnext, const double* cnow,
                                    const Lattice& clat, const PatchDesc* patches, int npatch,
                                    int max_cells, const long long* fill, const Relax& r,
                                    hipStream_t st) {
  k_fine_fill_collide<<<dim3((max_cells + 255) / 256, npatch), 256, 0, st>>>(fnow, fnext, cnow, clat,
                                                                            patches, fill, r);
  return hipGetLastError();
}

hipError_t launch_fine_stream(int mode, const double* in, double* out, const PatchDesc* patches,
                              int npatch, int max_cells, const Relax& r, hipStream_t st) {
  const dim3 grid((max_cells + 255) / 256, npatch);
  if (mode == 0)
    k_fine_stream<0><<<grid, 256, 0, st>>>(in, out, patches, r);
  else
    k_fine_stream<1><<<grid, 256, 0, st>>>(in, out, patches, r);
  return hipGetLastError();
}

hipError_t launch_coarse_push(const double* cnow, double* cnext, const Lattice& L,
                              const PatchDesc* boxes, int nbox, int max_cells,
                              const unsigned char* mask, const Relax& r, hipStream_t st) {
  k_coarse_push<<<dim3((max_cells + 255) / 256, nbox), 256, 0, st>>>(cnow, cnext, L, boxes, mask, r);
  return hipGetLastError();
}

hipError_t launch_sum_fine_to_coarse(const double* fnow, double* cnext, const Lattice& L,
                                     const PatchDesc* patches, const long long* cell, const int* start,
                                     const int4* contrib, int ncell, hipStream_t st) {
  if (ncell > 0)
    k_sum_fine_to_coarse<<<(ncell + 255) / 256, 256, 0, st>>>(fnow, cnext, L, patches, cell, start,
                                                              contrib, ncell);
  return hipGetLastError();
}

hipError_t launch_fine_from_coarse(double* fnow, const double* cnow, const Lattice& clat,
                                   const PatchDesc* patches, int npatch, int max_cells,
                                   hipStream_t st) {
  k_fine_from_coarse<<<dim3((max_cells + 255) / 256, npatch), 256, 0, st>>>(fnow, cnow, clat, patches);
  return hipGetLastError();
}

hipError_t launch_patch_io(int op, double* f, const PatchDesc* patches, int b, int valid_cells,
                           double* a, double* u, hipStream_t st) {
  const int blocks = (valid_cells + 255) / 256;
  if (op == 0)
    k_patch_io<0><<<blocks, 256, 0, st>>>(f, patches, b, a, u);
  else if (op == 1)
    k_patch_io<1><<<blocks, 256, 0, st>>>(f, patches, b, a, u);
  else
    k_patch_io<2><<<blocks, 256, 0, st>>>(f, patches, b, a, u);
  return hipGetLastError();
}

}  // namespace lbx
