// lbx_serve.hip -- kernels of the batched serving path.
//
// Every admitted run is an independent periodic lattice (nx, ny, nz) in the
// serving arena.  One launch advances every active run by one pass over the
// patch list (the work items of all runs: one item = 256 consecutive cells,
// x fastest, of one run, so every lane has a cell whatever the run's shape).
// The launch normally has one workgroup per item, with resident workgroups
// per CU capped as for the stream-collide kernels; the loop below walks
// further items with a static stride when the grid is smaller than the list,
// so every workgroup ends after a bounded number of items.  Each run's pass kind
// follows from the global step: pass p = step - start is collide-only for
// p = 0, fused pull + collide for 0 < p < steps, stream-only for p = steps
// (the same S (C S)^(n-1) C association as lbx_iterate), reading buffer
// p & 1 and writing the other one.  The per-cell arithmetic does not depend
// on which other runs share the launch, so a run's result is the same alone
// and in any mix, bit for bit.
#include <hip/hip_runtime.h>

#include "d3q15.h"
#include "lbx_internal.h"

#pragma clang fp contract(off)

namespace lbx {
namespace {

template <typename S>
__device__ __forceinline__ void serve_cell(const ServeRun& R, int kind, const S* __restrict__ src,
                                           S* __restrict__ dst, int x, int y, int z) {
  const int nx = R.n[0], ny = R.n[1], nz = R.n[2];
  const long long zst = NV * R.ps;
  const long long selfz = z * zst;
  const int selfp = y * nx + x;
  int xs[3], ys[3];
  long long zs[3];
  xs[0] = (x == nx - 1) ? 0 : x + 1;
  xs[1] = x;
  xs[2] = (x == 0) ? nx - 1 : x - 1;
  ys[0] = ((y == ny - 1) ? 0 : y + 1) * nx;
  ys[1] = y * nx;
  ys[2] = ((y == 0) ? ny - 1 : y - 1) * nx;
  zs[0] = ((z == nz - 1) ? 0 : z + 1) * zst;
  zs[1] = selfz;
  zs[2] = ((z == 0) ? nz - 1 : z - 1) * zst;
  S f[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    const long long off =
        (kind == STEP_COLLIDE) ? selfz + selfp : (zs[1 + kC[q][2]] + ys[1 + kC[q][1]] + xs[1 + kC[q][0]]);
    f[q] = src[off + q * R.ps];  // cached: small runs stay in the 256 MB MALL between steps
  }
  if (kind != STEP_STREAM) {
    if constexpr (sizeof(S) == 8)
      collide_f64(f, R.omega_s, R.omega_b);
    else
      collide_f32_shifted(f, float(R.rho0), float(R.omega_s), float(R.omega_b));
  }
#pragma unroll
  for (int q = 0; q < NV; ++q) __builtin_nontemporal_store(f[q], dst + selfz + q * R.ps + selfp);
}

template <typename S>
__global__ __launch_bounds__(256) void k_serve_step(ServeBatch B, long long step) {
  for (int item = blockIdx.x; item < B.total_items; item += gridDim.x) {
    // run holding this item: last r with item_begin <= item
    int lo = 0, hi = B.nruns - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (B.run[mid].item_begin <= item) lo = mid; else hi = mid - 1;
    }
    const ServeRun& R = B.run[lo];
    const long long p = step - R.start;
    if (p < 0 || p > R.steps) continue;  // not active at this step
    const int kind = (p == 0) ? STEP_COLLIDE : (p == R.steps ? STEP_STREAM : STEP_FUSED);
    const S* src = (const S*)R.buf[p & 1];
    S* dst = (S*)R.buf[(p + 1) & 1];
    const int plane = R.n[0] * R.n[1];
    const long long c = (long long)(item - R.item_begin) * 256 + threadIdx.x;
    if (c < (long long)plane * R.n[2]) {
      const int z = int(c / plane), r = int(c - (long long)z * plane);
      const int y = r / R.n[0], x = r - y * R.n[0];
      serve_cell<S>(R, kind, src, dst, x, y, z);
    }
  }
}

// Equilibrium of a plane pulse (rho = 1 + amplitude on one plane normal to
// `axis`, 1 elsewhere, u = 0), the reference's operation order, into the
// run's buffer 0.
template <typename S>
__global__ __launch_bounds__(256) void k_serve_init_pulse(S* __restrict__ f, Lattice L, int axis, int plane,
                                                         double amplitude, double rho0) {
  const int nx = L.nx, ny = L.ny;
  const long long n = (long long)nx * ny * L.nz;
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  const int x = int(i % nx), y = int((i / nx) % ny), z = int(i / ((long long)nx * ny));
  const long long self = L.at(0, y * nx + x, z);
  const int c = axis == 0 ? x : (axis == 1 ? y : z);
  const double rho = (c == plane) ? 1.0 + amplitude : 1.0;
  double fe[NV];
  equilibrium_f64(rho, 0.0, 0.0, 0.0, fe);
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    if constexpr (sizeof(S) == 8)
      f[self + q * L.ps] = fe[q];
    else
      f[self + q * L.ps] = float(fe[q] - kT.W[q] * rho0);
  }
}

}  // namespace

hipError_t launch_serve_step(int precision, const ServeBatch& b, long long step, int grid,
                             hipStream_t st) {
  if (b.total_items <= 0) return hipSuccess;
  const int blocks = b.total_items < grid ? b.total_items : grid;
  // resident workgroups per CU capped like the stream-collide launches
  const int lds = lds_bytes_for_cap(precision == 0 ? 3 : 4);
  if (precision == 0)
    k_serve_step<double><<<blocks, 256, lds > 0 ? lds : 0, st>>>(b, step);
  else
    k_serve_step<float><<<blocks, 256, lds > 0 ? lds : 0, st>>>(b, step);
  return hipGetLastError();
}

hipError_t launch_serve_init_pulse(int precision, void* f, const Lattice& L, int axis, int plane,
                                   double amplitude, double rho0, hipStream_t st) {
  const long long cells = (long long)L.nx * L.ny * L.nz;
  const int blocks = int((cells + 255) / 256);
  if (precision == 0)
    k_serve_init_pulse<double><<<blocks, 256, 0, st>>>((double*)f, L, axis, plane, amplitude, rho0);
  else
    k_serve_init_pulse<float><<<blocks, 256, 0, st>>>((float*)f, L, axis, plane, amplitude, rho0);
  return hipGetLastError();
}

}  // namespace lbx
