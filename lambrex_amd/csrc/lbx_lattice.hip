// lbx_lattice.hip -- stream-collide and field kernels on periodic lattices
// (gfx950).
//
// The reference step (src/AmrSim.cpp:124-135, 109-122; include/AmrSim.h:89-94)
// is: copy now -> next with ghosts, collide valid cells, refill ghosts,
// pull-stream valid+1 ghost layer into a fresh array, swap.  That is four
// passes over 15 populations per step.  Here one pass does pull + collide:
// h'(x) = C(S(h))(x), reading population q of the upstream cell x - c_q and
// writing the collided populations of x.  Each z-plane stores its 15
// population planes in turn (x fastest), so every load and store of a
// wavefront is a contiguous 64-element run (shifted by at most one element
// for c_x = +-1).
//
// Bandwidth bound: 15 loads + 15 stores per lattice update and ~430 FP64
// add/mul (+5 divisions).  No MFMA: there is no dense contraction here.
#include <hip/hip_runtime.h>

#include "d3q15.h"
#include "lbx_internal.h"

#pragma clang fp contract(off)

namespace lbx {
namespace {

// Source offsets of the 15 populations of cell (x, y, z) on a periodic
// lattice.  xs/ys/zo are indexed by 1 + c: the upstream coordinate of
// x is x - c, so index 0 (c = -1) holds x+1, 1 holds x, 2 (c = +1) holds x-1.
struct Upstream {
  int xs[3], ys[3];
  long long zo[3];
};

__device__ __forceinline__ Upstream upstream(int x, int y, int z, const Lattice& L) {
  Upstream u;
  const int row = L.nx;
  u.xs[0] = (x == L.nx - 1) ? 0 : x + 1;
  u.xs[1] = x;
  u.xs[2] = (x == 0) ? L.nx - 1 : x - 1;
  u.ys[0] = ((y == L.ny - 1) ? 0 : y + 1) * row;
  u.ys[1] = y * row;
  u.ys[2] = ((y == 0) ? L.ny - 1 : y - 1) * row;
  u.zo[0] = ((z == L.nz - 1) ? 0 : z + 1) * L.zs;
  u.zo[1] = z * L.zs;
  u.zo[2] = ((z == 0) ? L.nz - 1 : z - 1) * L.zs;
  return u;
}

// self: (z*zs, y*nx + x) of the cell itself
template <typename S, bool STREAM>
__device__ __forceinline__ void load_cell(const S* __restrict__ src, const Lattice& L, long long selfz,
                                          int selfp, const Upstream& u, S* f) {
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    const long long off = STREAM ? (u.zo[1 + kC[q][2]] + u.ys[1 + kC[q][1]] + u.xs[1 + kC[q][0]])
                                 : selfz + selfp;
    f[q] = __builtin_nontemporal_load(src + off + q * L.ps);
  }
}

// Index of the population opposite to q (c_opp = -c_q).
__device__ constexpr int opposite(int q) {
  int o = 0;
  for (int p = 0; p < NV; ++p)
    if (kC[p][0] == -kC[q][0] && kC[p][1] == -kC[q][1] && kC[p][2] == -kC[q][2]) o = p;
  return o;
}

// Pull with halfway bounce-back: population q of cell (x, y, z) comes from
// x - c_q, unless that crosses a wall face (an axis with its bit in
// L.walls), where the cell's own opposite population, which left towards
// the wall, returns instead.  Periodic axes wrap as in load_cell.
template <typename S>
__device__ __forceinline__ void load_cell_walls(const S* __restrict__ src, const Lattice& L, int x, int y,
                                                int z, long long selfz, int selfp, const Upstream& u, S* f) {
  const bool wx = L.walls & 1u, wy = L.walls & 2u, wz = L.walls & 4u;
  const bool lo_x = wx && x == 0, hi_x = wx && x == L.nx - 1;
  const bool lo_y = wy && y == 0, hi_y = wy && y == L.ny - 1;
  const bool lo_z = wz && z == 0, hi_z = wz && z == L.nz - 1;
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    const bool bounce = (kC[q][0] > 0 && lo_x) || (kC[q][0] < 0 && hi_x) || (kC[q][1] > 0 && lo_y) ||
                        (kC[q][1] < 0 && hi_y) || (kC[q][2] > 0 && lo_z) || (kC[q][2] < 0 && hi_z);
    const long long off = bounce ? selfz + opposite(q) * L.ps + selfp
                                 : u.zo[1 + kC[q][2]] + q * L.ps + u.ys[1 + kC[q][1]] + u.xs[1 + kC[q][0]];
    f[q] = __builtin_nontemporal_load(src + off);
  }
}

template <typename S>
__device__ __forceinline__ void store_cell(S* __restrict__ dst, const Lattice& L, long long selfz, int selfp,
                                           const S* f) {
#pragma unroll
  for (int q = 0; q < NV; ++q) __builtin_nontemporal_store(f[q], dst + selfz + q * L.ps + selfp);
}

// Resident workgroups per CU are capped at launch time by reserving
// dynamic LDS (LaunchShape::lds_bytes); the kernels do not use it.  With 30
// streams per cell, fewer resident wavefronts contend less for DRAM: at
// 256^3 one 512-lane workgroup per CU (FP64) and four 256-lane workgroups
// per CU (FP32) run ~8% and ~7% faster than uncapped launches.  Every
// element is loaded once per pass (the x+-1 neighbours of a wavefront are
// in the lines its own row fetches), so loads are non-temporal too.
template <int KIND, bool WALLS>
__global__ __launch_bounds__(512) void k_step_f64(const double* __restrict__ src,
                                                  double* __restrict__ dst, Lattice L, Relax r) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int z = blockIdx.z;
  if (x >= L.nx || y >= L.ny) return;
  const long long selfz = (long long)z * L.zs;
  const int selfp = y * L.nx + x;
  double f[NV];
  if (KIND == STEP_COLLIDE) {
    load_cell<double, false>(src, L, selfz, selfp, Upstream{}, f);
  } else if (WALLS) {
    load_cell_walls<double>(src, L, x, y, z, selfz, selfp, upstream(x, y, z, L), f);
  } else {
    const Upstream u = upstream(x, y, z, L);
    load_cell<double, true>(src, L, selfz, selfp, u, f);
  }
  if (KIND != STEP_STREAM) collide_f64(f, r.omega_s, r.omega_b);
  store_cell<double>(dst, L, selfz, selfp, f);
}

template <int KIND, bool WALLS>
__global__ __launch_bounds__(512) void k_step_f32s(const float* __restrict__ src,
                                                   float* __restrict__ dst, Lattice L, Relax r) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int y = blockIdx.y * blockDim.y + threadIdx.y;
  const int z = blockIdx.z;
  if (x >= L.nx || y >= L.ny) return;
  const long long selfz = (long long)z * L.zs;
  const int selfp = y * L.nx + x;
  float f[NV];
  if (KIND == STEP_COLLIDE) {
    load_cell<float, false>(src, L, selfz, selfp, Upstream{}, f);
  } else if (WALLS) {
    load_cell_walls<float>(src, L, x, y, z, selfz, selfp, upstream(x, y, z, L), f);
  } else {
    const Upstream u = upstream(x, y, z, L);
    load_cell<float, true>(src, L, selfz, selfp, u, f);
  }
  if (KIND != STEP_STREAM)
    collide_f32_shifted(f, float(r.rho0), float(r.omega_s), float(r.omega_b));
  store_cell<float>(dst, L, selfz, selfp, f);
}

dim3 grid_of(const Lattice& L, LaunchShape s) {
  return dim3((L.nx + s.bx - 1) / s.bx, (L.ny + s.by - 1) / s.by, L.nz);
}

// ------------------------------------------------------------ field kernels

__device__ __forceinline__ int region_cell(const Region& R, int i, int& x, int& y, int& z) {
  const int n0 = R.n[0], n1 = R.n[1];
  const int a = i % n0;
  const int t = i / n0;
  const int b = t % n1;
  const int c = t / n1;
  x = R.lo[0] + a;
  y = R.lo[1] + b;
  z = R.lo[2] + c;
  return i;
}

template <typename S>
__global__ __launch_bounds__(256) void k_equilibrium(S* __restrict__ f, Lattice L, Region R,
                                                     const double* __restrict__ rho,
                                                     const double* __restrict__ u, double rho0) {
  const int n = R.n[0] * R.n[1] * R.n[2];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int x, y, z;
  region_cell(R, i, x, y, z);
  const long long selfz = (long long)z * L.zs;
  const int selfp = y * L.nx + x;
  double fe[NV];
  equilibrium_f64(rho[i], u[i], u[n + i], u[2 * n + i], fe);
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    if constexpr (sizeof(S) == 8)
      f[selfz + q * L.ps + selfp] = fe[q];
    else
      f[selfz + q * L.ps + selfp] = float(fe[q] - kT.W[q] * rho0);
  }
}

// FP64 populations of cell (x, y, z).  Without PULL they are stored at the
// cell.  With PULL the lattice holds the post-collision state g of the
// previous step and f = S(g) is gathered from the upstream cells exactly as
// the stream-only pass would (periodic wrap or bounce-back walls), so a
// reader sees the same bits as after that pass.
template <typename S, bool PULL>
__device__ __forceinline__ void cell_at(const S* __restrict__ f, const Lattice& L, int x, int y, int z,
                                        double rho0, double* g) {
  const long long selfz = (long long)z * L.zs;
  const int selfp = y * L.nx + x;
  S v[NV];
  if (!PULL)
    load_cell<S, false>(f, L, selfz, selfp, Upstream{}, v);
  else if (L.walls)
    load_cell_walls<S>(f, L, x, y, z, selfz, selfp, upstream(x, y, z, L), v);
  else
    load_cell<S, true>(f, L, selfz, selfp, upstream(x, y, z, L), v);
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    if constexpr (sizeof(S) == 8)
      g[q] = v[q];
    else
      g[q] = kT.W[q] * rho0 + double(v[q]);
  }
}

template <typename S, bool PULL>
__global__ __launch_bounds__(256) void k_pack(const S* __restrict__ f, Lattice L, Region R,
                                              double* __restrict__ out, double rho0) {
  const int n = R.n[0] * R.n[1] * R.n[2];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int x, y, z;
  region_cell(R, i, x, y, z);
  double g[NV];
  cell_at<S, PULL>(f, L, x, y, z, rho0, g);
#pragma unroll
  for (int q = 0; q < NV; ++q) out[q * (long long)n + i] = g[q];
}

template <typename S>
__global__ __launch_bounds__(256) void k_unpack(S* __restrict__ f, Lattice L, Region R,
                                                const double* __restrict__ in, double rho0) {
  const int n = R.n[0] * R.n[1] * R.n[2];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int x, y, z;
  region_cell(R, i, x, y, z);
  const long long self = L.at(0, y * L.nx + x, z);
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    if constexpr (sizeof(S) == 8)
      f[self + q * L.ps] = in[q * (long long)n + i];
    else
      f[self + q * L.ps] = float(in[q * (long long)n + i] - kT.W[q] * rho0);
  }
}

template <typename S, bool PULL>
__global__ __launch_bounds__(256) void k_hydro(const S* __restrict__ f, Lattice L, Region R,
                                               double* __restrict__ rho, double* __restrict__ u,
                                               double rho0) {
  const int n = R.n[0] * R.n[1] * R.n[2];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int x, y, z;
  region_cell(R, i, x, y, z);
  double g[NV];
  cell_at<S, PULL>(f, L, x, y, z, rho0, g);
  double r, u0, u1, u2;
  hydro_f64(g, r, u0, u1, u2);
  rho[i] = r;
  u[i] = u0;
  u[n + i] = u1;
  u[2 * n + i] = u2;
}

// ErrorEst on device data: tag[c] = 1 (c = (z*ny + y)*nx + x) where
// CRIT 0: |grad rho|^2 > thr2, central differences with periodic wrap;
// CRIT 1: |u|^2 > thr2.  rho and u are exactly what k_hydro returns.
template <typename S, int CRIT, bool PULL>
__global__ __launch_bounds__(256) void k_tag(const S* __restrict__ f, Lattice L, double rho0, double thr2,
                                             unsigned char* __restrict__ tag) {
  const long long n = (long long)L.nx * L.ny * L.nz;
  const long long c = (long long)blockIdx.x * 256 + threadIdx.x;
  if (c >= n) return;
  const int x = int(c % L.nx), y = int((c / L.nx) % L.ny), z = int(c / ((long long)L.nx * L.ny));
  auto hydro_at = [&](int xx, int yy, int zz, double& r, double& u0, double& u1, double& u2) {
    double g[NV];
    cell_at<S, PULL>(f, L, xx, yy, zz, rho0, g);
    hydro_f64(g, r, u0, u1, u2);
  };
  double s;
  if (CRIT == 0) {
    double rp, rm, u0, u1, u2;
    hydro_at(x == L.nx - 1 ? 0 : x + 1, y, z, rp, u0, u1, u2);
    hydro_at(x == 0 ? L.nx - 1 : x - 1, y, z, rm, u0, u1, u2);
    const double gx = (rp - rm) * 0.5;
    hydro_at(x, y == L.ny - 1 ? 0 : y + 1, z, rp, u0, u1, u2);
    hydro_at(x, y == 0 ? L.ny - 1 : y - 1, z, rm, u0, u1, u2);
    const double gy = (rp - rm) * 0.5;
    hydro_at(x, y, z == L.nz - 1 ? 0 : z + 1, rp, u0, u1, u2);
    hydro_at(x, y, z == 0 ? L.nz - 1 : z - 1, rm, u0, u1, u2);
    const double gz = (rp - rm) * 0.5;
    s = gx * gx + gy * gy + gz * gz;
  } else {
    double r, u0, u1, u2;
    hydro_at(x, y, z, r, u0, u1, u2);
    s = u0 * u0 + u1 * u1 + u2 * u2;
  }
  tag[c] = s > thr2 ? 1 : 0;
}

// Block sum of 4 doubles in a fixed order: wavefront butterfly, then the
// wavefront partials in index order.
__device__ __forceinline__ void block_sum4(double v[4], double* out) {
  __shared__ double part[4][4];  // [wave][component], 256 threads = 4 waves
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    double s = v[c];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    v[c] = s;
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0)
    for (int c = 0; c < 4; ++c) part[wave][c] = v[c];
  __syncthreads();
  if (threadIdx.x == 0)
    for (int c = 0; c < 4; ++c) out[c] = ((part[0][c] + part[1][c]) + part[2][c]) + part[3][c];
}

template <typename S, bool PULL>
__global__ __launch_bounds__(256) void k_conservation(const S* __restrict__ f, Lattice L,
                                                      double rho0, double* __restrict__ partials) {
  const long long n = (long long)L.nx * L.ny * L.nz;
  const long long plane = (long long)L.nx * L.ny;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (long long i = blockIdx.x * 256ll + threadIdx.x; i < n; i += (long long)gridDim.x * 256) {
    double g[NV];
    const long long pc = i % plane;
    cell_at<S, PULL>(f, L, int(pc % L.nx), int(pc / L.nx), int(i / plane), rho0, g);
    acc[0] += mode_of<0>(g, std::make_integer_sequence<int, NV>{});
    acc[1] += mode_of<1>(g, std::make_integer_sequence<int, NV>{});
    acc[2] += mode_of<2>(g, std::make_integer_sequence<int, NV>{});
    acc[3] += mode_of<3>(g, std::make_integer_sequence<int, NV>{});
  }
  block_sum4(acc, partials + 4 * blockIdx.x);
}

// The same sums over the valid cells of patched boxes (2 ghost layers, see
// PatchDesc), every patch in one launch: workgroup (c, b) of the grid
// (chunks, npatch) adds the c-th strided share of patch b's valid cells, x
// fastest, into partials[4 * (b * chunks + c)].
template <typename S>
__global__ __launch_bounds__(256) void k_patch_conservation(const S* __restrict__ f,
                                                            const PatchDesc* __restrict__ patches, double rho0,
                                                            double* __restrict__ partials) {
  const PatchDesc& P = patches[blockIdx.y];
  const int nx = P.hi[0] - P.lo[0] + 1, ny = P.hi[1] - P.lo[1] + 1, nz = P.hi[2] - P.lo[2] + 1;
  const int n = nx * ny * nz;
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    const int x = i % nx, t = i / nx, y = t % ny, z = t / ny;
    const long long li = P.base + ((long long)(z + 2) * P.g[1] + (y + 2)) * P.g[0] + (x + 2);
    double g[NV];
#pragma unroll
    for (int q = 0; q < NV; ++q) {
      if constexpr (sizeof(S) == 8)
        g[q] = f[li + q * P.qs];
      else
        g[q] = kT.W[q] * rho0 + double(f[li + q * P.qs]);
    }
    acc[0] += mode_of<0>(g, std::make_integer_sequence<int, NV>{});
    acc[1] += mode_of<1>(g, std::make_integer_sequence<int, NV>{});
    acc[2] += mode_of<2>(g, std::make_integer_sequence<int, NV>{});
    acc[3] += mode_of<3>(g, std::make_integer_sequence<int, NV>{});
  }
  block_sum4(acc, partials + 4 * ((long long)blockIdx.y * gridDim.x + blockIdx.x));
}

__global__ __launch_bounds__(256) void k_reduce_partials(const double* __restrict__ partials,
                                                         int nblocks, double* __restrict__ out) {
  double acc[4] = {0.0, 0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nblocks; b += 256)
    for (int c = 0; c < 4; ++c) acc[c] += partials[4 * b + c];
  __shared__ double res[4];
  block_sum4(acc, res);
  __syncthreads();
  if (threadIdx.x < 4) out[threadIdx.x] = res[threadIdx.x];
}

int blocks_for(long long n) { return int((n + 255) / 256); }

// Plain streaming copy (one read + one write per element): the achievable
// read+write bandwidth ceiling, and the calibration kernel for counter-based
// traffic measurements.  16-byte elements, one per lane, non-temporal on both
// sides: the fastest plain copy shape measured on gfx950.
typedef double copy_vec __attribute__((ext_vector_type(2)));
__global__ __launch_bounds__(256) void k_copy(const copy_vec* __restrict__ a, copy_vec* __restrict__ b,
                                              long long n) {
  const long long i = (long long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) __builtin_nontemporal_store(__builtin_nontemporal_load(a + i), b + i);
}

}  // namespace

int lds_bytes_for_cap(int cap) {
  static int per_cu = -1, per_block = -1;
  if (per_cu < 0) {
    int dev = 0;
    per_cu = 64 * 1024;
    per_block = 64 * 1024;
    if (hipGetDevice(&dev) == hipSuccess) {
      (void)hipDeviceGetAttribute(&per_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev);
      (void)hipDeviceGetAttribute(&per_block, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    }
  }
  if (cap <= 0) return 0;
  const int bytes = per_cu / cap - 512;
  return (bytes > 0 && bytes <= per_block) ? bytes : -1;
}

LaunchShape default_shape(const Lattice& L, int precision) {
  LaunchShape s;
  const int lanes = precision == 0 ? 512 : 256;
  s.bx = L.nx < 256 ? L.nx : 256;
  s.by = lanes / s.bx;
  if (s.by > L.ny) s.by = L.ny;
  if (s.by < 1) s.by = 1;
  s.cap = precision == 0 ? 1 : 4;
  s.lds_bytes = lds_bytes_for_cap(s.cap);
  if (s.lds_bytes < 0) {
    s.cap = 0;
    s.lds_bytes = 0;
  }
  return s;
}

#define LBX_STEP_LAUNCH(KERNEL, SRC, DST)                                                          \
  do {                                                                                          \
    const dim3 grid = grid_of(L, s), block(s.bx, s.by, 1);                                      \
    const bool w = L.walls != 0;                                                                \
    switch (kind) {                                                                             \
      case STEP_FUSED:                                                                          \
        if (w) KERNEL<STEP_FUSED, true><<<grid, block, s.lds_bytes, st>>>(SRC, DST, L, r);      \
        else KERNEL<STEP_FUSED, false><<<grid, block, s.lds_bytes, st>>>(SRC, DST, L, r);       \
        break;                                                                                  \
      case STEP_COLLIDE:                                                                        \
        KERNEL<STEP_COLLIDE, false><<<grid, block, s.lds_bytes, st>>>(SRC, DST, L, r);          \
        break;                                                                                  \
      case STEP_STREAM:                                                                         \
        if (w) KERNEL<STEP_STREAM, true><<<grid, block, s.lds_bytes, st>>>(SRC, DST, L, r);     \
        else KERNEL<STEP_STREAM, false><<<grid, block, s.lds_bytes, st>>>(SRC, DST, L, r);      \
        break;                                                                                  \
    }                                                                                           \
    return hipGetLastError();                                                                   \
  } while (0)

hipError_t launch_step_f64(StepKind kind, const double* src, double* dst, const Lattice& L,
                           const Relax& r, LaunchShape s, hipStream_t st) {
  LBX_STEP_LAUNCH(k_step_f64, src, dst);
}

hipError_t launch_step_f32s(StepKind kind, const float* src, float* dst, const Lattice& L,
                            const Relax& r, LaunchShape s, hipStream_t st) {
  LBX_STEP_LAUNCH(k_step_f32s, src, dst);
}

hipError_t launch_equilibrium(int precision, void* f, const Lattice& L, const Region& R,
                              const double* rho, const double* u, double rho0, hipStream_t st) {
  const long long n = (long long)R.n[0] * R.n[1] * R.n[2];
  if (precision == 0)
    k_equilibrium<double><<<blocks_for(n), 256, 0, st>>>((double*)f, L, R, rho, u, rho0);
  else
    k_equilibrium<float><<<blocks_for(n), 256, 0, st>>>((float*)f, L, R, rho, u, rho0);
  return hipGetLastError();
}

// Launch KERNEL<S, ..., PULL> for the storage precision and the pull flag.
#define LBX_PULL_LAUNCH(KERNEL, GRID, ...)                                                              \
  do {                                                                                                \
    if (precision == 0 && pull)                                                                       \
      KERNEL<double, true><<<GRID, 256, 0, st>>>((const double*)f, __VA_ARGS__);                      \
    else if (precision == 0)                                                                          \
      KERNEL<double, false><<<GRID, 256, 0, st>>>((const double*)f, __VA_ARGS__);                     \
    else if (pull)                                                                                    \
      KERNEL<float, true><<<GRID, 256, 0, st>>>((const float*)f, __VA_ARGS__);                        \
    else                                                                                              \
      KERNEL<float, false><<<GRID, 256, 0, st>>>((const float*)f, __VA_ARGS__);                       \
  } while (0)

hipError_t launch_pack_f(int precision, bool pull, const void* f, const Lattice& L, const Region& R,
                         double* out, double rho0, hipStream_t st) {
  const long long n = (long long)R.n[0] * R.n[1] * R.n[2];
  LBX_PULL_LAUNCH(k_pack, blocks_for(n), L, R, out, rho0);
  return hipGetLastError();
}

hipError_t launch_unpack_f(int precision, void* f, const Lattice& L, const Region& R,
                           const double* in, double rho0, hipStream_t st) {
  const long long n = (long long)R.n[0] * R.n[1] * R.n[2];
  if (precision == 0)
    k_unpack<double><<<blocks_for(n), 256, 0, st>>>((double*)f, L, R, in, rho0);
  else
    k_unpack<float><<<blocks_for(n), 256, 0, st>>>((float*)f, L, R, in, rho0);
  return hipGetLastError();
}

hipError_t launch_hydro(int precision, bool pull, const void* f, const Lattice& L, const Region& R,
                        double* rho, double* u, double rho0, hipStream_t st) {
  const long long n = (long long)R.n[0] * R.n[1] * R.n[2];
  LBX_PULL_LAUNCH(k_hydro, blocks_for(n), L, R, rho, u, rho0);
  return hipGetLastError();
}

template <typename S, bool PULL>
static void tag_launch(int criterion, const S* f, const Lattice& L, double rho0, double thr2, unsigned char* tag,
                int blocks, hipStream_t st) {
  if (criterion == 0)
    k_tag<S, 0, PULL><<<blocks, 256, 0, st>>>(f, L, rho0, thr2, tag);
  else
    k_tag<S, 1, PULL><<<blocks, 256, 0, st>>>(f, L, rho0, thr2, tag);
}

hipError_t launch_tag(int precision, bool pull, int criterion, const void* f, const Lattice& L, double rho0,
                      double thr2, unsigned char* tag, hipStream_t st) {
  const long long n = (long long)L.nx * L.ny * L.nz;
  const int blocks = int((n + 255) / 256);
  if (precision == 0 && pull)
    tag_launch<double, true>(criterion, (const double*)f, L, rho0, thr2, tag, blocks, st);
  else if (precision == 0)
    tag_launch<double, false>(criterion, (const double*)f, L, rho0, thr2, tag, blocks, st);
  else if (pull)
    tag_launch<float, true>(criterion, (const float*)f, L, rho0, thr2, tag, blocks, st);
  else
    tag_launch<float, false>(criterion, (const float*)f, L, rho0, thr2, tag, blocks, st);
  return hipGetLastError();
}

hipError_t launch_conservation(int precision, bool pull, const void* f, const Lattice& L, double rho0,
                               double* partials, int nblocks, double* out, hipStream_t st) {
  LBX_PULL_LAUNCH(k_conservation, nblocks, L, rho0, partials);
  k_reduce_partials<<<1, 256, 0, st>>>(partials, nblocks, out);
  return hipGetLastError();
}

hipError_t launch_patch_conservation(int precision, const void* f, const PatchDesc* patches, int npatch,
                                     int chunks, double rho0, double* partials, double* out, hipStream_t st) {
  const dim3 grid(chunks, npatch);
  if (precision == 0)
    k_patch_conservation<double><<<grid, 256, 0, st>>>((const double*)f, patches, rho0, partials);
  else
    k_patch_conservation<float><<<grid, 256, 0, st>>>((const float*)f, patches, rho0, partials);
  k_reduce_partials<<<1, 256, 0, st>>>(partials, chunks * npatch, out);
  return hipGetLastError();
}

hipError_t launch_copy(const double* a, double* b, long long n, hipStream_t st) {
  const long long nv = n / 2;  // whole 16-byte elements; callers pass even n
  k_copy<<<(unsigned)((nv + 255) / 256), 256, 0, st>>>((const copy_vec*)a, (copy_vec*)b, nv);
  return hipGetLastError();
}

}  // namespace lbx
