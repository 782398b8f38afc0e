// lbx_internal.h -- declarations shared by the C-ABI implementation
// (lbx_abi.cpp) and the HIP kernels (lbx_lattice.hip).  Not installed.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

struct lbx_ctx;

namespace lbx {

constexpr int NV_TOTAL = 15;

// Device ordinal and HIP stream of a context (lbx_abi.cpp).
int ctx_device(const lbx_ctx* ctx);
hipStream_t ctx_stream(const lbx_ctx* ctx);

#if defined(__HIPCC__)
#define LBX_IHD __host__ __device__ __forceinline__
#else
#define LBX_IHD inline
#endif

// A periodic lattice: one level stored as a single array over the whole
// domain, z-plane major: each z-plane holds its 15 population planes one
// after the other, x fastest.  Population q of cell (x,y,z) lives at
// z*zs + q*ps + y*nx + x.  Keeping a plane's populations together keeps the
// 30 streams of a stream-collide pass inside a few MB of address space
// (measured ~4% faster than population-major at 256^3).  Boxes of the level
// are index ranges into it, so neighbouring boxes exchange nothing: the pull
// reads the neighbour box's cells directly (periodic wrap at the faces).
struct Lattice {
  int nx, ny, nz;
  long long ps;  // population-plane stride: nx*ny rounded up to 64 elements
  long long zs;  // z-plane stride: 15 * ps
  // bit d set: halfway bounce-back walls on both faces normal to axis d
  // instead of the periodic wrap (single-level lattices only)
  unsigned walls;
  LBX_IHD long long at(int q, long long plane_cell, long long z) const { return z * zs + q * ps + plane_cell; }
  // linear cell index (z*ny + y)*nx + x
  LBX_IHD long long at_cell(int q, long long c) const {
    const long long plane = (long long)nx * ny;
    return at(q, c % plane, c / plane);
  }
  LBX_IHD long long elems() const { return nz * zs; }
};

// elem_bytes: 8 (FP64) or 4 (FP32-shifted storage).  A population plane
// whose size is a multiple of 512 KB (256^2 FP64, 512^2 either precision) is
// padded by 1 KB: at such strides the 15 population streams of a pass start
// at the same offset modulo 512 KB and share HBM channels.  Measured on
// MI355X with the fused FP64 pass: 256^3 648 -> 640 us, 512x512x64 724 ->
// 660 us; unpadded 128^2 and 192^2 planes were fastest as they are.
inline Lattice make_lattice(int nx, int ny, int nz, int elem_bytes) {
  Lattice L;
  L.nx = nx;
  L.ny = ny;
  L.nz = nz;
  L.ps = ((long long)nx * ny + 63) / 64 * 64;
  if ((L.ps * elem_bytes) % (512 * 1024) == 0) L.ps += 1024 / elem_bytes;
  L.zs = NV_TOTAL * L.ps;
  L.walls = 0;
  return L;
}

struct Relax {
  double omega_s, omega_b;
  double rho0;  // FP32-shifted storage reference density
};

// Region of a lattice (inclusive lo, extent n) for box-wise transfers.
struct Region {
  int lo[3];
  int n[3];
};

enum StepKind { STEP_FUSED = 0, STEP_COLLIDE = 1, STEP_STREAM = 2 };

// Launch configuration of the lattice stream-collide kernels.
struct LaunchShape {
  int bx, by;     // block extent in x and y (bx*by <= 512)
  int cap;        // resident workgroups per CU (0 = no cap)
  int lds_bytes;  // dynamic LDS reserved per workgroup to enforce `cap`
};

// Workgroups over whole x-rows: FP64 (precision 0) 512 lanes, at most one
// resident per CU; FP32 256 lanes, at most four per CU (the measured optima
// at 256^3).
LaunchShape default_shape(const Lattice& L, int precision);
// Dynamic LDS bytes that limit residency to `cap` workgroups per CU on the
// current device (0 for no cap, -1 if the device cannot reserve that much).
int lds_bytes_for_cap(int cap);

// FP64 storage
hipError_t launch_step_f64(StepKind kind, const double* src, double* dst, const Lattice& L,
                           const Relax& r, LaunchShape s, hipStream_t st);
// FP32 shifted storage
hipError_t launch_step_f32s(StepKind kind, const float* src, float* dst, const Lattice& L,
                            const Relax& r, LaunchShape s, hipStream_t st);

// Equilibrium of a region from device rho[nz][ny][nx], u[3][nz][ny][nx].
hipError_t launch_equilibrium(int precision, void* f, const Lattice& L, const Region& R,
                              const double* rho, const double* u, double rho0, hipStream_t st);
// Readers of a level array take `pull`: the array holds the post-collision
// state g of the previous step and the populations are f = S(g), gathered
// from the upstream cells as the stream-only pass would (see Level::post in
// lbx_abi.cpp); without it they are read in place.
// Copy a region between the lattice and a packed FP64 [15][n2][n1][n0] buffer.
hipError_t launch_pack_f(int precision, bool pull, const void* f, const Lattice& L, const Region& R,
                         double* out, double rho0, hipStream_t st);
hipError_t launch_unpack_f(int precision, void* f, const Lattice& L, const Region& R,
                           const double* in, double rho0, hipStream_t st);
// Density and velocity of a region into packed FP64 buffers.
hipError_t launch_hydro(int precision, bool pull, const void* f, const Lattice& L, const Region& R,
                        double* rho, double* u, double rho0, hipStream_t st);
// FP64 sums over the whole lattice: partials[nblocks*4] then out[4].
// tag[(z*ny + y)*nx + x] = 1 where criterion 0 (|grad rho|^2) or 1 (|u|^2)
// exceeds thr2.
hipError_t launch_tag(int precision, bool pull, int criterion, const void* f, const Lattice& L, double rho0,
                      double thr2, unsigned char* tag, hipStream_t st);
hipError_t launch_conservation(int precision, bool pull, const void* f, const Lattice& L, double rho0,
                               double* partials, int nblocks, double* out, hipStream_t st);

// A box stored on its own with 2 ghost layers (fine level), or a box of a
// level array (coarse push: only lo/hi/cells used).  Population q of grown
// local cell i is at base + q*qs + i.
struct PatchDesc {
  int lo[3], hi[3];  // valid box, level index space (inclusive)
  int g[3];          // grown extents (n + 4)
  int cells;         // grown cells (or valid cells for level-array boxes)
  long long base;    // element offset of the patch in the level buffer
  long long qs;      // population stride (cells rounded up to 64)
  long long fill_base;  // offset of this patch's rows in the fill table
  long long avg_base;   // offset of this patch's 2x2x2 block averages ((g/2)^3 per population)
};

// FP64 sums over the valid cells of patches [0, npatch) in one launch:
// chunks workgroups per patch write partials[4 * (patch * chunks + chunk)],
// then out[4] = their sum in a fixed order.
hipError_t launch_patch_conservation(int precision, const void* f, const PatchDesc* patches, int npatch,
                                     int chunks, double rho0, double* partials, double* out, hipStream_t st);

// Level-hierarchy kernels (lbx_twolevel.hip); precision 0 = FP64 storage,
// else FP32-shifted (rho0 in Relax or the argument).  The patch kernels run
// over patches [first, first + npatch).  A coarse level is either a level
// array (clat, cpatches null) or patched (cpatches its descriptors); coarse
// references in the tables are then -1 - element or
// -1 - ((patch << 32) | grown index).
// mask (null = none): per grown cell of the level (fill_base + i), nonzero =
// the valid cell lies under a finer level's mask and is zeroed, not collided.
// own: the level's patches in launch order (patches [first, first + npatch)
// of it run); patches: the same table by box index, for the fill sources.
hipError_t launch_fine_fill_collide(int precision, const void* fnow, void* fnext, const void* cnow,
                                    const Lattice& clat, const PatchDesc* cpatches, const PatchDesc* patches,
                                    const PatchDesc* own, int first, int npatch, int max_cells,
                                    const long long* fill, const unsigned char* mask, const Relax& r,
                                    hipStream_t st, long long oshift = 0);
// out holds each patch at element P.base - oshift (0: a level buffer; the
// group's first base: a group-sized scratch array); an input the same way
// at P.base - ishift.
// inner: `in` is itself such an output (outer layer not stored, read as
// zero), for the later substeps at a refinement ratio above 2.
hipError_t launch_fine_stream_collide(int precision, const void* in, void* out, long long oshift,
                                      const PatchDesc* patches, int first, int npatch, int max_cells, const Relax& r,
                                      hipStream_t st, bool inner = false);
// The finest level's last substep at a refinement ratio above 2: stream of
// the valid cells (input outer layer read as zero) into out ...
hipError_t launch_fine_stream_valid(int precision, const void* in, void* out, const PatchDesc* patches, int first,
                                    int npatch, int max_cells, double rho0, hipStream_t st);
// ... and the average-down of the result over ratio^3 blocks of valid cells
// (children in (k, j, i) order, times 1/ratio^3) into avg.
hipError_t launch_fine_avg_valid(int precision, const void* f, double* avg, long long avg_stride,
                                 const PatchDesc* patches, int first, int npatch, int max_blocks, int ratio,
                                 double rho0, hipStream_t st);
// Stream + ZeroInvalidComponents of a middle level (no collide), with the
// child's block averages added into valid cells: slot[fill_base + i] is the
// touched-cell index of grown cell i (-1 if none), start/contrib as in
// launch_sum_apply.
hipError_t launch_fine_stream_zero(int precision, const void* in, void* out, const PatchDesc* patches, int first,
                                   int npatch, int max_cells, double rho0, const int* slot, const double* avg,
                                   long long avg_stride, const int* start, const int* contrib, hipStream_t st,
                                   long long ishift = 0);
// avg holds 15 FP64 planes of avg_stride block averages; max_blocks = max (g/2)^3;
// in holds each patch at element P.base - ishift.
hipError_t launch_fine_stream_avg(int precision, const void* in, long long ishift, void* out, double* avg,
                                  long long avg_stride, const PatchDesc* patches, int first, int npatch,
                                  int max_blocks, double rho0, hipStream_t st);
// A middle level's last step of its parent's step, all patches: the stream of
// launch_fine_stream_zero (the child's block averages cavg added through
// slot / start / contrib) fused with the level's own block averages of the
// result into avg.
hipError_t launch_middle_stream_avg(int precision, const void* in, long long ishift, void* out, double* avg,
                                   long long avg_stride, const PatchDesc* patches, int first, int npatch,
                                   int max_blocks, double rho0, const int* slot, const double* cavg,
                                   long long cavg_stride, const int* start, const int* contrib, hipStream_t st);
// flags: per coarse cell (x fastest), bits 0-5 = x+1, x-1, y+1, y-1, z+1,
// z-1 leave the cell's box; bit 7 = masked by the fine level.
hipError_t launch_coarse_push(int precision, const void* cnow, void* cnext, const Lattice& L,
                              const unsigned char* flags, const Relax& r, LaunchShape shape, hipStream_t st);
// Block averages into a level-array parent: cell[u] is the element of
// population 0 of touched cell u, its contributions avg indices
// contrib[start[u] .. start[u+1]).
hipError_t launch_sum_apply(int precision, const double* avg, long long avg_stride, void* cnext, const Lattice& L,
                            const long long* cell, const int* start, const int* contrib, int ncell, double rho0,
                            hipStream_t st);
hipError_t launch_fine_from_coarse(int precision, void* fnow, const void* cnow, const Lattice& clat,
                                   const PatchDesc* patches, int npatch, int max_cells, double rho0, int ratio,
                                   hipStream_t st);
// Valid cells of new patches from old patches or the coarse parent
// (src per new grown cell: (old patch << 32) | old grown index, or a coarse
// reference as above).
hipError_t launch_fine_remake(int precision, void* fnew, const void* fold, const void* cnow, const Lattice& clat,
                              const PatchDesc* cpatches, const PatchDesc* patches, int npatch, int max_cells,
                              const PatchDesc* old_patches, const long long* src, hipStream_t st);
// op 0: pack valid cells of patch b into FP64 f a[15][n]; 1: unpack; 2: hydro
// into a (rho), u.
hipError_t launch_patch_io(int precision, int op, void* f, const PatchDesc* patches, int b, int valid_cells,
                           double* a, double* u, double rho0, hipStream_t st);

// Serving: one active run of the batch, passed by value in the kernel
// arguments (kServeMaxRuns * sizeof(ServeRun) stays under 4 KB).
struct ServeRun {
  void* buf[2];      // state buffers (pass p reads buf[p & 1])
  int n[3];          // lattice extents
  int item_begin;    // first work item (256 cells, x fastest) of this run in the batch
  int steps;         // LB steps of the run (passes 0..steps)
  long long start;   // global step of pass 0
  long long ps;      // population-plane stride of the run's lattice (zs = 15 ps)
  double omega_s, omega_b, rho0;
};
constexpr int kServeMaxRuns = 44;
struct ServeBatch {
  ServeRun run[kServeMaxRuns];
  int nruns;
  int total_items;
};
hipError_t launch_serve_step(int precision, const ServeBatch& b, long long step, int grid,
                             hipStream_t st);
hipError_t launch_serve_init_pulse(int precision, void* f, const Lattice& L, int axis, int plane,
                                   double amplitude, double rho0, hipStream_t st);

// Streaming copy of n doubles (bandwidth ceiling / counter calibration).
hipError_t launch_copy(const double* a, double* b, long long n, hipStream_t st);

}  // namespace lbx
