// lbx_twolevel.hip -- device kernels of the two-level Rohde cycle
// (src/AmrSim.cpp:430-631) on gfx950, FP64 or FP32-shifted storage.
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
//   k_fine_stream_collide Stream(1) (valid + 1 ghost layer; outer layer of the
//                         fresh array defined as 0, not stored) + collide
//                         valid cells
//   k_fine_stream_avg     Stream(1) + ZeroInvalidComponents(1), and the 1/8
//                         weighted sum of every 2x2x2 block of the result;
//                         valid cells stored (the next fill rewrites ghosts)
//   k_coarse_push         CoarseCollide(0) (masked cells -> 0) + Stream(0) in
//                         push form; pulls across a box face read the
//                         pre-collision ghost copies, as the reference's
//                         RohdeCycle does (no FillBoundary between collide and
//                         stream)
//   k_sum_apply           SumFromFine(0): ADD those block averages (fine valid
//                         + ghosts) into coarse next, contributions to one
//                         coarse cell in a fixed order
//
// Deeper hierarchies (RohdeCycle's recursion, src/AmrSim.cpp:453-455) reuse
// these kernels on patched "middle" levels, which have both a parent and a
// child: the fill-collide takes its coarse values from patches (PC) and
// zeroes the cells under the child's mask (MASK); k_fine_stream_zero is the
// middle level's stream of its valid cells with the child's block averages
// added into them; and k_fine_stream_avg<S, true> is the same plus
// ZeroInvalidComponents for the middle level's last step of its parent's
// step, with the block averages of its result for the parent.  Every pass
// stores valid cells only.
//
// Storage type S is double (FP64, bit-exact with the oracle) or float (the
// shifted deviation d = f - w rho0 of the FP32 path).  Either way the block
// sums and the coarse additions of SumFromFine are done in FP64 on
// reconstructed f, and a population set to zero means f = 0 (d = -w rho0).
#include <hip/hip_runtime.h>

#include "d3q15.h"
#include "lbx_internal.h"

#pragma clang fp contract(off)

namespace lbx {
namespace {

// f of a stored value, and the stored value of f, for population q.
template <typename S>
__device__ __forceinline__ double to_f64(S v, int q, double rho0) {
  if constexpr (sizeof(S) == 8)
    return v;
  else
    return kT.W[q] * rho0 + double(v);
}

template <typename S>
__device__ __forceinline__ S from_f64(double f, int q, double rho0) {
  if constexpr (sizeof(S) == 8)
    return f;
  else
    return float(f - kT.W[q] * rho0);
}

// The stored value of f_q = 0.
template <typename S>
__device__ __forceinline__ S zero_pop(int q, double rho0) {
  return from_f64<S>(0.0, q, rho0);
}

template <typename S>
__device__ __forceinline__ void collide_s(S* f, const Relax& r) {
  if constexpr (sizeof(S) == 8)
    collide_f64(f, r.omega_s, r.omega_b);
  else
    collide_f32_shifted(f, float(r.rho0), float(r.omega_s), float(r.omega_b));
}

// n / d for 0 <= n < 2^22 and d >= 1: the float estimate from the hardware
// reciprocal (relative error below 2^-22 with the product's rounding) is
// off by less than one, and one correction makes it exact.
__device__ __forceinline__ int div_small(int n, int d) {
  int q = (int)((float)n * __builtin_amdgcn_rcpf((float)d));
  const int r = n - q * d;
  q += (r >= d) - (r < 0);
  return q;
}

// Chunk of its patch handled by this workgroup.  Workgroups go to the 8
// XCDs round-robin by linear id; with gridDim.x a multiple of 8, XCD k takes
// chunks [k G/8, (k+1) G/8) of every patch (G = gridDim.x): a contiguous
// slab, so neighbouring chunks, and the same rows of neighbouring patches,
// share that XCD's L2.  The identity for any other grid.
__device__ __forceinline__ int xcd_chunk() {
  const int G = gridDim.x, c = blockIdx.x;
  return (G & 7) ? c : (c & 7) * (G >> 3) + (c >> 3);
}

// (x, y, z) of index i < n of a box with rows of n0 and planes of n1 rows;
// n is uniform over the workgroup, so the branch is too.
__device__ __forceinline__ void split3(int i, int n, int n0, int n1, int& x, int& y, int& z) {
  if (n <= (1 << 22)) {
    const int t = div_small(i, n0);
    x = i - t * n0;
    z = div_small(t, n1);
    y = t - z * n1;
  } else {
    x = i % n0;
    const int t = i / n0;
    y = t % n1;
    z = t / n1;
  }
}

__device__ __forceinline__ void patch_coords(const PatchDesc& P, int i, int& gx, int& gy, int& gz) {
  split3(i, P.cells, P.g[0], P.g[1], gx, gy, gz);
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

// Every lane first picks where its 15 populations come from (its own valid
// cell, the fine valid cell holding its periodic image, or the coarse
// parent) so the loads of valid and ghost lanes issue together; valid cells
// are then collided.  Coarse sources are -1 - (element of population 0) in
// a level array, or with PC -1 - ((coarse patch << 32) | grown index).  With
// MASK, a valid cell whose mask byte is set is zeroed instead of collided
// (CoarseCollide on a level that has a finer one, src/AmrSim.cpp:499-500).
// The patch is own[first + blockIdx.y] (`own` lists the level's patches in
// launch order); `patches` is the level's table by box index, which the
// fill sources refer to.
template <typename S, bool PC, bool MASK>
__global__ __launch_bounds__(256) void k_fine_fill_collide(
    const S* __restrict__ fnow, S* __restrict__ fnext, const S* __restrict__ cnow, Lattice clat,
    const PatchDesc* __restrict__ cpatches, const PatchDesc* __restrict__ patches,
    const PatchDesc* __restrict__ own, int first, const long long* __restrict__ fill,
    const unsigned char* __restrict__ mask, Relax r, long long oshift) {
  // P is read in place: with a by-value copy the PC variants merged the
  // loads of base / qs from P, patches[s] and cpatches[] into one load through
  // a selected pointer, which kept the copy in 80 B of scratch per lane.
  const PatchDesc& P = own[first + blockIdx.y];
  const int i = xcd_chunk() * 256 + threadIdx.x;
  if (i >= P.cells) return;
  int gx, gy, gz;
  patch_coords(P, i, gx, gy, gz);
  const bool valid = patch_valid(P, gx, gy, gz);
  const S* src;
  long long stride;
  if (valid) {
    src = fnow + P.base + i;
    stride = P.qs;
  } else {
    const long long d = fill[P.fill_base + i];
    if (d >= 0) {
      const int s = int(d >> 32);
      src = fnow + patches[s].base + (d & 0xffffffffll);
      stride = patches[s].qs;
    } else if constexpr (PC) {
      const long long e = -1 - d;
      const PatchDesc& CP = cpatches[e >> 32];
      src = cnow + CP.base + (e & 0xffffffffll);
      stride = CP.qs;
    } else {
      src = cnow + (-1 - d);  // coarse element of population 0
      stride = clat.ps;
    }
  }
  S f[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) f[q] = src[q * stride];
  if (valid) {
    if (MASK && mask[P.fill_base + i]) {
#pragma unroll
      for (int q = 0; q < NV; ++q) f[q] = zero_pop<S>(q, r.rho0);
    } else {
      collide_s(f, r);
    }
  } else if (patch_outer(P, gx, gy, gz)) {
    f[0] = zero_pop<S>(0, r.rho0);
  }
  const long long ob = P.base - oshift;
#pragma unroll
  for (int q = 0; q < NV; ++q) fnext[ob + q * P.qs + i] = f[q];
}

// Pull of grown cell i over valid + 1 ghost layer, from the patch stored at
// element `base` of `in`; the outer ghost layer of the fresh array is
// defined as zero.  Returns whether i is a valid cell.
template <typename S>
__device__ __forceinline__ bool fine_pull(const PatchDesc& P, const S* __restrict__ in, long long base, int i,
                                          int gx, int gy, int gz, double rho0, S* f) {
  if (patch_outer(P, gx, gy, gz)) {
#pragma unroll
    for (int q = 0; q < NV; ++q) f[q] = zero_pop<S>(q, rho0);
    return false;
  }
  const int sx = 1, sy = P.g[0], sz = P.g[0] * P.g[1];
#pragma unroll
  for (int q = 0; q < NV; ++q) f[q] = in[base + q * P.qs + i - kC[q][0] * sx - kC[q][1] * sy - kC[q][2] * sz];
  return patch_valid(P, gx, gy, gz);
}

// fine_pull for an input whose outer ghost layer is not stored (the first
// substep's output, k_fine_stream_collide): a population whose upstream cell
// lies on the outer layer takes that layer's defined value, zero, without
// reading that layer.
template <typename S>
__device__ __forceinline__ bool fine_pull_inner(const PatchDesc& P, const S* __restrict__ in, long long base,
                                                int i, int gx, int gy, int gz, double rho0, S* f) {
  if (patch_outer(P, gx, gy, gz)) {
#pragma unroll
    for (int q = 0; q < NV; ++q) f[q] = zero_pop<S>(q, rho0);
    return false;
  }
  const int sx = 1, sy = P.g[0], sz = P.g[0] * P.g[1];
  // One load per population for every lane (a wave holds valid and ghost
  // cells alike): a lane whose upstream cell is on the outer layer (never
  // the case for a valid cell) loads its own cell instead and keeps zero.
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    const bool outer = patch_outer(P, gx - kC[q][0], gy - kC[q][1], gz - kC[q][2]);
    const int off = outer ? 0 : kC[q][0] * sx + kC[q][1] * sy + kC[q][2] * sz;
    const S v = in[base + q * P.qs + i - off];
    f[q] = outer ? zero_pop<S>(q, rho0) : v;
  }
  return patch_valid(P, gx, gy, gz);
}

// First fine substep's stream + the second substep's collision: pull, then
// collide valid cells.  `out` holds the patch at element P.base - oshift (a
// group-sized scratch array when oshift is the group's first base).  The
// outer ghost layer of the result is zero by definition (Stream's fresh
// array) and is not stored: the next pass reads it as zero.  With INNER the
// input is itself such a result (a later substep at a refinement ratio above
// 2), whose outer layer is read as zero.
template <typename S, bool INNER = false>
__global__ __launch_bounds__(256) void k_fine_stream_collide(const S* __restrict__ in, S* __restrict__ out,
                                                             const PatchDesc* __restrict__ patches, int first,
                                                             Relax r, long long oshift) {
  const PatchDesc P = patches[first + blockIdx.y];
  const int i = xcd_chunk() * 256 + threadIdx.x;
  if (i >= P.cells) return;
  int gx, gy, gz;
  patch_coords(P, i, gx, gy, gz);
  if (patch_outer(P, gx, gy, gz)) return;
  S f[NV];
  const bool valid = INNER ? fine_pull_inner(P, in, P.base, i, gx, gy, gz, r.rho0, f)
                           : fine_pull(P, in, P.base, i, gx, gy, gz, r.rho0, f);
  if (valid) collide_s(f, r);
  const long long ob = P.base - oshift;
#pragma unroll
  for (int q = 0; q < NV; ++q) out[ob + q * P.qs + i] = f[q];
}

// ZeroInvalidComponents on a ghost cell at absolute (x, y, z): population m
// survives only if pos - 2 c_m lies in the patch's valid box.
template <typename S>
__device__ __forceinline__ void zero_invalid(const PatchDesc& P, int gx, int gy, int gz, double rho0, S* f) {
  const int x = P.lo[0] - 2 + gx, y = P.lo[1] - 2 + gy, z = P.lo[2] - 2 + gz;
#pragma unroll
  for (int q = 0; q < NV; ++q)
    if (!in_valid(P, x - 2 * kC[q][0], y - 2 * kC[q][1], z - 2 * kC[q][2])) f[q] = zero_pop<S>(q, rho0);
}

// SumFromFine's addition into one valid cell of a patched level: the block
// averages avg[q * avg_stride + contrib[j]], j in [start[u], start[u+1]),
// added in FP64 in that order to the cell's streamed value.
template <typename S>
__device__ __forceinline__ void add_block_sums(S* f, int u, const double* __restrict__ avg, long long avg_stride,
                                               const int* __restrict__ start, const int* __restrict__ contrib,
                                               double rho0) {
  double acc[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) acc[q] = to_f64(f[q], q, rho0);
  for (int j = start[u]; j < start[u + 1]; ++j) {
    const long long a = contrib[j];
#pragma unroll
    for (int q = 0; q < NV; ++q) acc[q] += avg[q * avg_stride + a];
  }
#pragma unroll
  for (int q = 0; q < NV; ++q) f[q] = from_f64<S>(acc[q], q, rho0);
}

// A middle level's Stream + ZeroInvalidComponents (src/AmrSim.cpp:459-463)
// in a step whose result needs no block averages: the pull of the valid
// cells.  The ghost shell the reference computes (populations that came
// from valid cells) is not formed: nothing reads a patched level's ghost
// cells from its state (its next fill rewrites them, a finer level's fill
// and every read-back take valid cells only).  Fused with the second half
// of SumFromFine into this level (src/AmrSim.cpp:464-466, 592-602): a valid
// cell whose slot (per grown cell, -1 if none) names a touched cell u adds
// the child's block averages [start[u], start[u+1]) to its streamed value in
// FP64, in the same order as k_sum_apply.  `in` holds each patch at element
// P.base - ishift (a group-sized region when ishift is the group's first
// base); `out` at P.base.
template <typename S>
__global__ __launch_bounds__(256) void k_fine_stream_zero(const S* __restrict__ in, S* __restrict__ out,
                                                          const PatchDesc* __restrict__ patches, int first,
                                                          double rho0, const int* __restrict__ slot,
                                                          const double* __restrict__ avg, long long avg_stride,
                                                          const int* __restrict__ start,
                                                          const int* __restrict__ contrib, long long ishift) {
  const PatchDesc P = patches[first + blockIdx.y];
  const int i = xcd_chunk() * 256 + threadIdx.x;
  if (i >= P.cells) return;
  int gx, gy, gz;
  patch_coords(P, i, gx, gy, gz);
  if (!patch_valid(P, gx, gy, gz)) return;
  S f[NV];
  fine_pull(P, in, P.base - ishift, i, gx, gy, gz, rho0, f);
  const int u = slot[P.fill_base + i];
  if (u >= 0) add_block_sums(f, u, avg, avg_stride, start, contrib, rho0);
#pragma unroll
  for (int q = 0; q < NV; ++q) out[P.base + q * P.qs + i] = f[q];
}

// Second fine substep's stream + ZeroInvalidComponents(1), fused with the
// first half of SumFromFine: the 1/8-weighted sum of each 2x2x2 block of the
// grown patch (= one coarse cell of the coarsened box grown by 1), children
// added in (k, j, i) order in FP64, to avg[q * avg_stride + avg_base + block].
// A workgroup takes B = avg_blocks<S>() consecutive blocks; thread t
// computes child t / B of block t % B, so each wavefront reads whole x-pairs
// of its blocks.  The children meet in LDS for the ordered sums, as stored
// values (S), and each child is converted to FP64 f as it is added (the
// same bits as converting first).
// With SUM (a middle level's
// last step of its parent's step) the stream is k_fine_stream_zero's, the
// child's block averages (slot, cavg, start, contrib) added into valid
// cells, and the block sums are those of the level's resulting state, ghost
// shell included.  Without SUM (the finest level's second substep) the
// input's outer ghost layer is not stored and reads as zero.  Either way
// only valid cells are stored: the ghost cells of a level's state are
// rewritten by its next fill-collide before anything reads them.
// Blocks per workgroup (8 threads each): 32 in FP64; 16 in FP32, whose
// passes are shorter, so that a barrier spans two waves instead of four.
template <typename S>
constexpr int avg_blocks() { return sizeof(S) == 4 ? 16 : 32; }

template <typename S, bool SUM = false>
__global__ __launch_bounds__(256) void k_fine_stream_avg(const S* __restrict__ in, S* __restrict__ out,
                                                         double* __restrict__ avg, long long avg_stride,
                                                         const PatchDesc* __restrict__ patches, int first,
                                                         double rho0, long long ishift,
                                                         const int* __restrict__ slot = nullptr,
                                                         const double* __restrict__ cavg = nullptr,
                                                         long long cavg_stride = 0,
                                                         const int* __restrict__ start = nullptr,
                                                         const int* __restrict__ contrib = nullptr) {
  constexpr int kAvgBlocks = avg_blocks<S>();
  __shared__ S part[NV][8][kAvgBlocks];
  const PatchDesc P = patches[first + blockIdx.y];
  const int c0 = P.g[0] >> 1, c1 = P.g[1] >> 1, c2 = P.g[2] >> 1;
  const int nblk = c0 * c1 * c2;
  const int b0 = xcd_chunk() * kAvgBlocks;  // first block of this workgroup
  if (b0 >= nblk) return;                  // uniform over the workgroup
  const int child = threadIdx.x / kAvgBlocks, lb = threadIdx.x % kAvgBlocks;
  const int u = b0 + lb;
  if (u < nblk) {
    const int di = child & 1, dj = (child >> 1) & 1, dk = child >> 2;
    int cx, cy, cz;
    split3(u, nblk, c0, c1, cx, cy, cz);
    const int gx = 2 * cx + di, gy = 2 * cy + dj, gz = 2 * cz + dk;
    const int i = (gz * P.g[1] + gy) * P.g[0] + gx;
    S f[NV];
    const bool valid = SUM ? fine_pull(P, in, P.base - ishift, i, gx, gy, gz, rho0, f)
                           : fine_pull_inner(P, in, P.base - ishift, i, gx, gy, gz, rho0, f);
    if (valid) {
      if constexpr (SUM) {
        const int su = slot[P.fill_base + i];
        if (su >= 0) add_block_sums(f, su, cavg, cavg_stride, start, contrib, rho0);
      }
    } else if (!patch_outer(P, gx, gy, gz)) {
      zero_invalid(P, gx, gy, gz, rho0, f);
    }
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      if (valid) out[P.base + q * P.qs + i] = f[q];
      part[q][child][lb] = f[q];
    }
  }
  __syncthreads();
  for (int task = threadIdx.x; task < NV * kAvgBlocks; task += 8 * kAvgBlocks) {
    const int q = task / kAvgBlocks, b = task % kAvgBlocks;
    if (b0 + b >= nblk) continue;
    double sum = 0.0;
#pragma unroll
    for (int c = 0; c < 8; ++c) sum += to_f64(part[q][c][b], q, rho0);
    avg[q * avg_stride + P.avg_base + b0 + b] = sum * 0.125;
  }
}

// Coarse cell flags: bit 0/1 = the cell is on its box's high/low x face
// (x+1 / x-1 leaves the box), bits 2/3 the same in y, 4/5 in z; bit 7 = the
// cell lies under the fine mask.
constexpr unsigned kMasked = 0x80u;

__device__ constexpr unsigned leave_bits(int q) {
  return (kC[q][0] > 0 ? 1u : 0u) | (kC[q][0] < 0 ? 2u : 0u) | (kC[q][1] > 0 ? 4u : 0u) |
         (kC[q][1] < 0 ? 8u : 0u) | (kC[q][2] > 0 ? 16u : 0u) | (kC[q][2] < 0 ? 32u : 0u);
}

// Over the whole coarse level array, shaped like the stream-collide launch.
template <typename S>
__global__ __launch_bounds__(512) void k_coarse_push(const S* __restrict__ cnow, S* __restrict__ cnext,
                                                     Lattice L, const unsigned char* __restrict__ flags,
                                                     Relax r) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int z = blockIdx.z;
  if (x >= L.nx || y >= L.ny) return;
  const long long self = L.at(0, y * L.nx + x, z);
  const unsigned fl = flags[((long long)z * L.ny + y) * L.nx + x];
  S pre[NV], post[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) pre[q] = post[q] = __builtin_nontemporal_load(cnow + self + q * L.ps);
  if (fl & kMasked) {
#pragma unroll
    for (int q = 0; q < NV; ++q) post[q] = zero_pop<S>(q, r.rho0);
  } else {
    collide_s(post, r);
  }
  const int xs[3] = {(x == 0) ? L.nx - 1 : x - 1, x, (x == L.nx - 1) ? 0 : x + 1};
  const int ys[3] = {((y == 0) ? L.ny - 1 : y - 1) * L.nx, y * L.nx, ((y == L.ny - 1) ? 0 : y + 1) * L.nx};
  const int zs[3] = {(z == 0) ? L.nz - 1 : z - 1, z, (z == L.nz - 1) ? 0 : z + 1};
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    const long long dest = L.at(q, ys[1 + kC[q][1]] + xs[1 + kC[q][0]], zs[1 + kC[q][2]]);
    __builtin_nontemporal_store((fl & leave_bits(q)) ? pre[q] : post[q], cnext + dest);
  }
}

// Second half of SumFromFine: thread per touched coarse cell; its
// contributions [start[u], start[u+1]) (indices into avg, in the order the
// reference's loops add them) are added in FP64 to the streamed coarse value.
// A coarse cell is the element of population 0 in the level array.
template <typename S>
__global__ __launch_bounds__(256) void k_sum_apply(const double* __restrict__ avg, long long avg_stride,
                                                   S* __restrict__ cnext, Lattice L,
                                                   const long long* __restrict__ cell,
                                                   const int* __restrict__ start,
                                                   const int* __restrict__ contrib, int ncell, double rho0) {
  const int u = xcd_chunk() * 256 + threadIdx.x;
  if (u >= ncell) return;
  const long long c = cell[u], stride = L.ps;
  double acc[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) acc[q] = to_f64(cnext[c + q * stride], q, rho0);
  for (int j = start[u]; j < start[u + 1]; ++j) {
    const long long a = contrib[j];
#pragma unroll
    for (int q = 0; q < NV; ++q) acc[q] += avg[q * avg_stride + a];
  }
#pragma unroll
  for (int q = 0; q < NV; ++q) cnext[c + q * stride] = from_f64<S>(acc[q], q, rho0);
}

// The finest level's last substep at a refinement ratio r above 2: the
// stream of its valid cells (the input's outer layer read as zero) into
// `out`.  Its ghost cells need no ZeroInvalidComponents: at such a ratio the
// average-down reads valid cells only (nGrow = 2 / r = 0) and the next
// fill-collide rewrites the ghosts.
template <typename S>
__global__ __launch_bounds__(256) void k_fine_stream_valid(const S* __restrict__ in, S* __restrict__ out,
                                                           const PatchDesc* __restrict__ patches, int first,
                                                           double rho0) {
  const PatchDesc& P = patches[first + blockIdx.y];
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= P.cells) return;
  int gx, gy, gz;
  patch_coords(P, i, gx, gy, gz);
  if (!patch_valid(P, gx, gy, gz)) return;
  S f[NV];
  fine_pull_inner(P, in, P.base, i, gx, gy, gz, rho0, f);
#pragma unroll
  for (int q = 0; q < NV; ++q) out[P.base + q * P.qs + i] = f[q];
}

// The first half of SumFromFine at a refinement ratio r above 2
// (amrex_avgdown over the valid box only, nGrow = 2 / r = 0): per r x r x r
// block of valid cells, the FP64 sum of its children in (k, j, i) order
// times 1/r^3, to avg[q * avg_stride + P.avg_base + block] (blocks x
// fastest over the coarsened valid box).
template <typename S>
__global__ __launch_bounds__(256) void k_fine_avg_valid(const S* __restrict__ f, double* __restrict__ avg,
                                                        long long avg_stride, const PatchDesc* __restrict__ patches,
                                                        int first, int r, double rho0) {
  const PatchDesc& P = patches[first + blockIdx.y];
  const int c0 = (P.hi[0] - P.lo[0] + 1) / r, c1 = (P.hi[1] - P.lo[1] + 1) / r, c2 = (P.hi[2] - P.lo[2] + 1) / r;
  const int u = blockIdx.x * 256 + threadIdx.x;
  const int nblk = c0 * c1 * c2;
  if (u >= nblk) return;
  int cx, cy, cz;
  split3(u, nblk, c0, c1, cx, cy, cz);
  const double volfrac = 1.0 / (double)(r * r * r);
  const int sy = P.g[0], sz = P.g[0] * P.g[1];
  const long long corner = (long long)(2 + r * cz) * sz + (long long)(2 + r * cy) * sy + (2 + r * cx);
  for (int q = 0; q < NV; ++q) {
    double sum = 0.0;
    for (int dk = 0; dk < r; ++dk)
      for (int dj = 0; dj < r; ++dj)
        for (int di = 0; di < r; ++di)
          sum += to_f64(f[P.base + q * P.qs + corner + (long long)dk * sz + dj * sy + di], q, rho0);
    avg[q * avg_stride + P.avg_base + u] = sum * volfrac;
  }
}

// DistFnFillFromCoarse (src/AmrSim.cpp:393-411): valid fine cells take the
// coarse value of their parent; ghosts are zeroed (they are rewritten before
// they are read).
template <typename S>
__global__ __launch_bounds__(256) void k_fine_from_coarse(S* __restrict__ fnow, const S* __restrict__ cnow,
                                                          Lattice clat, const PatchDesc* __restrict__ patches,
                                                          double rho0, int ratio) {
  const PatchDesc P = patches[blockIdx.y];
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= P.cells) return;
  int gx, gy, gz;
  patch_coords(P, i, gx, gy, gz);
  if (patch_valid(P, gx, gy, gz)) {
    const int x = (P.lo[0] - 2 + gx) / ratio, y = (P.lo[1] - 2 + gy) / ratio, z = (P.lo[2] - 2 + gz) / ratio;
    const long long c = clat.at(0, y * clat.nx + x, z);
#pragma unroll
    for (int q = 0; q < NV; ++q) fnow[P.base + q * P.qs + i] = cnow[c + q * clat.ps];
  } else {
#pragma unroll
    for (int q = 0; q < NV; ++q) fnow[P.base + q * P.qs + i] = zero_pop<S>(q, rho0);
  }
}

// RemakeLevel's fill (src/AmrSim.cpp:715-744 via DistFnFillPatch): each
// valid cell of the new patches copies its populations from an old fine
// patch (src >= 0: (old patch << 32) | old grown index) or from its coarse
// parent (src < 0: -1 - coarse element of population 0, or with PC
// -1 - ((coarse patch << 32) | grown index)).  With every source a parent it
// is DistFnFillFromCoarse on a level whose parent is patched.
template <typename S, bool PC>
__global__ __launch_bounds__(256) void k_fine_remake(S* __restrict__ fnew, const S* __restrict__ fold,
                                                     const S* __restrict__ cnow, Lattice clat,
                                                     const PatchDesc* __restrict__ cpatches,
                                                     const PatchDesc* __restrict__ patches,
                                                     const PatchDesc* __restrict__ old_patches,
                                                     const long long* __restrict__ src) {
  const PatchDesc P = patches[blockIdx.y];
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= P.cells) return;
  int gx, gy, gz;
  patch_coords(P, i, gx, gy, gz);
  if (!patch_valid(P, gx, gy, gz)) return;
  const long long d = src[P.fill_base + i];
  const S* from;
  long long stride;
  if (d >= 0) {
    const PatchDesc& O = old_patches[d >> 32];
    from = fold + O.base + (d & 0xffffffffll);
    stride = O.qs;
  } else if constexpr (PC) {
    const long long e = -1 - d;
    const PatchDesc& CP = cpatches[e >> 32];
    from = cnow + CP.base + (e & 0xffffffffll);
    stride = CP.qs;
  } else {
    from = cnow + (-1 - d);
    stride = clat.ps;
  }
#pragma unroll
  for (int q = 0; q < NV; ++q) fnew[P.base + q * P.qs + i] = from[q * stride];
}

// Valid cells of one patch <-> packed FP64 f [15][nz][ny][nx]; hydro of one
// patch.
template <int OP, typename S>  // 0 pack, 1 unpack, 2 hydro
__global__ __launch_bounds__(256) void k_patch_io(S* __restrict__ f, const PatchDesc* __restrict__ patches, int b,
                                                  double* __restrict__ a, double* __restrict__ u, double rho0) {
  const PatchDesc P = patches[b];
  const int nx = P.hi[0] - P.lo[0] + 1, ny = P.hi[1] - P.lo[1] + 1, nz = P.hi[2] - P.lo[2] + 1;
  const int n = nx * ny * nz;
  const int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int x = i % nx, t = i / nx, y = t % ny, z = t / ny;
  const long long li = ((long long)(z + 2) * P.g[1] + (y + 2)) * P.g[0] + (x + 2);
  if (OP == 0) {
#pragma unroll
    for (int q = 0; q < NV; ++q) a[(long long)q * n + i] = to_f64(f[P.base + q * P.qs + li], q, rho0);
  } else if (OP == 1) {
#pragma unroll
    for (int q = 0; q < NV; ++q) f[P.base + q * P.qs + li] = from_f64<S>(a[(long long)q * n + i], q, rho0);
  } else {
    double g[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) g[q] = to_f64(f[P.base + q * P.qs + li], q, rho0);
    double rr, u0, u1, u2;
    hydro_f64(g, rr, u0, u1, u2);
    a[i] = rr;
    u[i] = u0;
    u[n + i] = u1;
    u[2 * n + i] = u2;
  }
}

template <typename S>
hipError_t patch_io(int op, S* f, const PatchDesc* patches, int b, int valid_cells, double* a, double* u,
                    double rho0, hipStream_t st) {
  const int blocks = (valid_cells + 255) / 256;
  if (op == 0)
    k_patch_io<0, S><<<blocks, 256, 0, st>>>(f, patches, b, a, u, rho0);
  else if (op == 1)
    k_patch_io<1, S><<<blocks, 256, 0, st>>>(f, patches, b, a, u, rho0);
  else
    k_patch_io<2, S><<<blocks, 256, 0, st>>>(f, patches, b, a, u, rho0);
  return hipGetLastError();
}

}  // namespace

template <typename S>
hipError_t fill_collide(const S* fnow, S* fnext, const S* cnow, const Lattice& clat, const PatchDesc* cpatches,
                        const PatchDesc* patches, const PatchDesc* own, int first, const long long* fill,
                        const unsigned char* mask, const Relax& r, long long oshift, dim3 grid, hipStream_t st) {
  if (cpatches && mask)
    k_fine_fill_collide<S, true, true><<<grid, 256, 0, st>>>(fnow, fnext, cnow, clat, cpatches, patches, own, first,
                                                             fill, mask, r, oshift);
  else if (cpatches)
    k_fine_fill_collide<S, true, false><<<grid, 256, 0, st>>>(fnow, fnext, cnow, clat, cpatches, patches, own,
                                                              first, fill, mask, r, oshift);
  else if (mask)
    k_fine_fill_collide<S, false, true><<<grid, 256, 0, st>>>(fnow, fnext, cnow, clat, cpatches, patches, own,
                                                              first, fill, mask, r, oshift);
  else
    k_fine_fill_collide<S, false, false><<<grid, 256, 0, st>>>(fnow, fnext, cnow, clat, cpatches, patches, own,
                                                               first, fill, mask, r, oshift);
  return hipGetLastError();
}

hipError_t launch_fine_fill_collide(int precision, const void* fnow, void* fnext, const void* cnow,
                                    const Lattice& clat, const PatchDesc* cpatches, const PatchDesc* patches,
                                    const PatchDesc* own, int first, int npatch, int max_cells,
                                    const long long* fill, const unsigned char* mask, const Relax& r,
                                    hipStream_t st, long long oshift) {
  // A multiple of 8 workgroups per patch, so that xcd_chunk() gives each XCD
  // the same slab of every patch: the rows whose first or last two cells an
  // x-neighbour's ghost fill reads (and the rows a y-neighbour's reads) are
  // then fetched into that XCD's L2 by both patches' workgroups at about the
  // same time.  The extra workgroups exit at once.
  const dim3 grid(((max_cells + 255) / 256 + 7) / 8 * 8, npatch);
  if (precision == 0)
    return fill_collide((const double*)fnow, (double*)fnext, (const double*)cnow, clat, cpatches, patches, own,
                        first, fill, mask, r, oshift, grid, st);
  return fill_collide((const float*)fnow, (float*)fnext, (const float*)cnow, clat, cpatches, patches, own, first,
                      fill, mask, r, oshift, grid, st);
}

hipError_t launch_fine_stream_zero(int precision, const void* in, void* out, const PatchDesc* patches, int first,
                                   int npatch, int max_cells, double rho0, const int* slot, const double* avg,
                                   long long avg_stride, const int* start, const int* contrib, hipStream_t st,
                                   long long ishift) {
  const dim3 grid(((max_cells + 255) / 256 + 7) / 8 * 8, npatch);  // for xcd_chunk()
  if (precision == 0)
    k_fine_stream_zero<double><<<grid, 256, 0, st>>>((const double*)in, (double*)out, patches, first, rho0, slot,
                                                     avg, avg_stride, start, contrib, ishift);
  else
    k_fine_stream_zero<float><<<grid, 256, 0, st>>>((const float*)in, (float*)out, patches, first, rho0, slot, avg,
                                                    avg_stride, start, contrib, ishift);
  return hipGetLastError();
}

// Grid of k_fine_stream_avg<S>: a multiple of 8 workgroups per patch, for
// xcd_chunk(); each workgroup has 8 avg_blocks<S>() threads.
template <typename S>
dim3 avg_grid(int max_blocks, int npatch) {
  constexpr int ab = avg_blocks<S>();
  return dim3(((max_blocks + ab - 1) / ab + 7) / 8 * 8, npatch);
}

hipError_t launch_middle_stream_avg(int precision, const void* in, long long ishift, void* out, double* avg,
                                   long long avg_stride, const PatchDesc* patches, int first, int npatch,
                                   int max_blocks, double rho0, const int* slot, const double* cavg,
                                   long long cavg_stride, const int* start, const int* contrib, hipStream_t st) {
  if (precision == 0)
    k_fine_stream_avg<double, true><<<avg_grid<double>(max_blocks, npatch), 8 * avg_blocks<double>(), 0, st>>>(
        (const double*)in, (double*)out, avg, avg_stride, patches, first, rho0, ishift, slot, cavg, cavg_stride,
        start, contrib);
  else
    k_fine_stream_avg<float, true><<<avg_grid<float>(max_blocks, npatch), 8 * avg_blocks<float>(), 0, st>>>(
        (const float*)in, (float*)out, avg, avg_stride, patches, first, rho0, ishift, slot, cavg, cavg_stride, start,
        contrib);
  return hipGetLastError();
}

hipError_t launch_fine_stream_collide(int precision, const void* in, void* out, long long oshift,
                                      const PatchDesc* patches, int first, int npatch, int max_cells, const Relax& r,
                                      hipStream_t st, bool inner) {
  const dim3 grid(((max_cells + 255) / 256 + 7) / 8 * 8, npatch);  // for xcd_chunk()
  if (precision == 0 && inner)
    k_fine_stream_collide<double, true><<<grid, 256, 0, st>>>((const double*)in, (double*)out, patches, first, r,
                                                              oshift);
  else if (precision == 0)
    k_fine_stream_collide<double><<<grid, 256, 0, st>>>((const double*)in, (double*)out, patches, first, r, oshift);
  else if (inner)
    k_fine_stream_collide<float, true><<<grid, 256, 0, st>>>((const float*)in, (float*)out, patches, first, r,
                                                             oshift);
  else
    k_fine_stream_collide<float><<<grid, 256, 0, st>>>((const float*)in, (float*)out, patches, first, r, oshift);
  return hipGetLastError();
}

hipError_t launch_fine_stream_avg(int precision, const void* in, long long ishift, void* out, double* avg,
                                  long long avg_stride, const PatchDesc* patches, int first, int npatch,
                                  int max_blocks, double rho0, hipStream_t st) {
  if (precision == 0)
    k_fine_stream_avg<double><<<avg_grid<double>(max_blocks, npatch), 8 * avg_blocks<double>(), 0, st>>>(
        (const double*)in, (double*)out, avg, avg_stride, patches, first, rho0, ishift);
  else
    k_fine_stream_avg<float><<<avg_grid<float>(max_blocks, npatch), 8 * avg_blocks<float>(), 0, st>>>(
        (const float*)in, (float*)out, avg, avg_stride, patches, first, rho0, ishift);
  return hipGetLastError();
}

hipError_t launch_coarse_push(int precision, const void* cnow, void* cnext, const Lattice& L,
                              const unsigned char* flags, const Relax& r, LaunchShape sh, hipStream_t st) {
  const dim3 grid((L.nx + sh.bx - 1) / sh.bx, (L.ny + sh.by - 1) / sh.by, L.nz);
  const dim3 block(sh.bx, sh.by, 1);
  if (precision == 0)
    k_coarse_push<double><<<grid, block, sh.lds_bytes, st>>>((const double*)cnow, (double*)cnext, L, flags, r);
  else
    k_coarse_push<float><<<grid, block, sh.lds_bytes, st>>>((const float*)cnow, (float*)cnext, L, flags, r);
  return hipGetLastError();
}

hipError_t launch_sum_apply(int precision, const double* avg, long long avg_stride, void* cnext, const Lattice& L,
                            const long long* cell, const int* start, const int* contrib, int ncell, double rho0,
                            hipStream_t st) {
  if (ncell <= 0) return hipSuccess;
  const int blocks = ((ncell + 255) / 256 + 7) / 8 * 8;  // for xcd_chunk()
  if (precision == 0)
    k_sum_apply<double><<<blocks, 256, 0, st>>>(avg, avg_stride, (double*)cnext, L, cell, start, contrib, ncell,
                                                rho0);
  else
    k_sum_apply<float><<<blocks, 256, 0, st>>>(avg, avg_stride, (float*)cnext, L, cell, start, contrib, ncell,
                                               rho0);
  return hipGetLastError();
}

hipError_t launch_fine_from_coarse(int precision, void* fnow, const void* cnow, const Lattice& clat,
                                   const PatchDesc* patches, int npatch, int max_cells, double rho0, int ratio,
                                   hipStream_t st) {
  const dim3 grid((max_cells + 255) / 256, npatch);
  if (precision == 0)
    k_fine_from_coarse<double><<<grid, 256, 0, st>>>((double*)fnow, (const double*)cnow, clat, patches, rho0, ratio);
  else
    k_fine_from_coarse<float><<<grid, 256, 0, st>>>((float*)fnow, (const float*)cnow, clat, patches, rho0, ratio);
  return hipGetLastError();
}

hipError_t launch_fine_stream_valid(int precision, const void* in, void* out, const PatchDesc* patches, int first,
                                    int npatch, int max_cells, double rho0, hipStream_t st) {
  const dim3 grid((max_cells + 255) / 256, npatch);
  if (precision == 0)
    k_fine_stream_valid<double><<<grid, 256, 0, st>>>((const double*)in, (double*)out, patches, first, rho0);
  else
    k_fine_stream_valid<float><<<grid, 256, 0, st>>>((const float*)in, (float*)out, patches, first, rho0);
  return hipGetLastError();
}

hipError_t launch_fine_avg_valid(int precision, const void* f, double* avg, long long avg_stride,
                                 const PatchDesc* patches, int first, int npatch, int max_blocks, int ratio,
                                 double rho0, hipStream_t st) {
  const dim3 grid((max_blocks + 255) / 256, npatch);
  if (precision == 0)
    k_fine_avg_valid<double><<<grid, 256, 0, st>>>((const double*)f, avg, avg_stride, patches, first, ratio, rho0);
  else
    k_fine_avg_valid<float><<<grid, 256, 0, st>>>((const float*)f, avg, avg_stride, patches, first, ratio, rho0);
  return hipGetLastError();
}

hipError_t launch_fine_remake(int precision, void* fnew, const void* fold, const void* cnow, const Lattice& clat,
                              const PatchDesc* cpatches, const PatchDesc* patches, int npatch, int max_cells,
                              const PatchDesc* old_patches, const long long* src, hipStream_t st) {
  const dim3 grid((max_cells + 255) / 256, npatch);
  if (precision == 0 && cpatches)
    k_fine_remake<double, true><<<grid, 256, 0, st>>>((double*)fnew, (const double*)fold, (const double*)cnow, clat,
                                                      cpatches, patches, old_patches, src);
  else if (precision == 0)
    k_fine_remake<double, false><<<grid, 256, 0, st>>>((double*)fnew, (const double*)fold, (const double*)cnow, clat,
                                                       cpatches, patches, old_patches, src);
  else if (cpatches)
    k_fine_remake<float, true><<<grid, 256, 0, st>>>((float*)fnew, (const float*)fold, (const float*)cnow, clat,
                                                     cpatches, patches, old_patches, src);
  else
    k_fine_remake<float, false><<<grid, 256, 0, st>>>((float*)fnew, (const float*)fold, (const float*)cnow, clat,
                                                      cpatches, patches, old_patches, src);
  return hipGetLastError();
}

hipError_t launch_patch_io(int precision, int op, void* f, const PatchDesc* patches, int b, int valid_cells,
                           double* a, double* u, double rho0, hipStream_t st) {
  if (precision == 0) return patch_io(op, (double*)f, patches, b, valid_cells, a, u, rho0, st);
  return patch_io(op, (float*)f, patches, b, valid_cells, a, u, rho0, st);
}

}  // namespace lbx
