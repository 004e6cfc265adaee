// Device-side data structures shared by the HIP kernels and the host solver.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace msc {

// Planes of z-tiles: each wavefront covers 64 consecutive k, lane 0 being
// the k0-1 halo, so a tile owns 63 cells along z.
constexpr int kLanes = 64;
constexpr int kMaxLkDeg = 24;
constexpr int kLkPolyDeg = 12;  // highest degree the in-kernel Horner takes
constexpr int kNtffBlock = 256;  // surface points per NTFF workgroup

// current models selected per launch
enum { kCurDI = 0, kCurLKPoly = 1, kCurLKClosed = 2 };

// CPML slab axes a fused-step variant carries code for: a tile whose owned
// cells meet the slabs of one axis only takes that axis's variant, a tile
// meeting several takes the full one
enum { kPmlX = 1, kPmlY = 2, kPmlZ = 4, kPmlAll = 7 };
// the tile lists of one solve, in this order, and the variant of each
enum { kListInterior = 0, kListX, kListY, kListZ, kListAll, kNumLists };
inline int list_pm(int l) {
  return l == kListX ? kPmlX : l == kListY ? kPmlY : l == kListZ ? kPmlZ : l == kListAll ? kPmlAll : 0;
}

// packed tile descriptor: z-tile (8 bits) | y-tile (12 bits) | x-chunk (11 bits)
// | kTileNoCurrent, indices into StepArgs::zstart / ystart / xstart
__host__ __device__ inline uint32_t pack_tile(int tz, int ty, int ch) {
  return (uint32_t)tz | ((uint32_t)ty << 8) | ((uint32_t)ch << 20);
}
constexpr int kMaxChunks = 2047;
// Set on tiles where eps0 wp^2 is zero (in the field precision) on every
// cell the tile updates J on, halo included.  There J starts at zero and
// J' = A J + eps0 wp^2 K E stays zero, so the kernels skip J and the media
// (neither read nor written: all three J sets hold zero there).
constexpr uint32_t kTileNoCurrent = 1u << 31;

// Everything one fused time step of one solve needs.  Kept in device memory
// (one element of an array indexed by solve id); it does not change from
// step to step (the step index lives in device memory at `tdev`, and the
// ping-pong buffers are selected by its parity), so a launch sequence can be
// replayed or captured in a hipGraph unchanged.
template <typename T>
struct StepArgs {
  int nx, ny, nz, L;
  int xs, ys, zs;                       // nx-2L-1, ny-2L-1, nz-2L-1
  int bx1, bx2, by1, by2, bz1, bz2;     // TF/SF box
  int ssx, ssy, ssz;                    // slice planes
  int dopml, doplate, doB0;
  int jcol;                             // Lee-Kalluri J driven by collocated E (else node average)
  int plate_lo, plate_up;
  int chunk, tiles_y, tiles_z, nchunks; // launch decomposition
  const int* xstart;  // [nchunks+1] x-chunk starts (aligned to the CPML slab edges)
  const int* ystart;  // [tiles_y+1] row-tile starts of interior tiles (ty rows)
  const int* ystart_b;  // row-tile starts of CPML-boundary tiles (ty_b rows)
  const int* zstart;  // [tiles_z+1] z-tile starts
  int cadence;                          // tsteps/numfiles
  long long sy, sx;                     // strides: sy = nz, sx = ny*nz
  T c2h, c2e0;                          // dt/mu0, dt/eps0 (sigma = 0)
  T inv_d;                              // 1/dx (dx == dy == dz)
  T c1s, c2s;                           // sphere (sigma = 1e7) coefficients
  double x0, y0, z0, dd, rsph;          // coordinates for the sphere test
  double dt;
  // Lee-Kalluri constants: R = Ce1 w w^T + cos I + Se1 [w x]
  double wx, wy, wz, w0, cth, sth, Se1, Ce1;
  double R[9];
  // Lee-Kalluri per-cell coefficients as polynomials in x = nu dt:
  // lkc[0] -> e^{-x}, lkc[1] -> kI, lkc[2] -> kX, lkc[3] -> kP (degree lkdeg)
  int lkdeg;
  double lkc[4][kMaxLkDeg + 1];
  // device pointers
  // Field sets rotate through three buffers: after step t, E and H are in
  // set t % 3 and the (deferred) J^(t-1) in set (t - 1) % 3; step t reads
  // set t % 3 and writes set (t + 1) % 3, so a kernel advancing two steps
  // at once can write step t + 2 while step t is still being read.
  T* e[3][3];        // E: [set][x,y,z]
  T* h[3][3];        // H
  T* j[3][3];        // J
  const T* wpe2;     // eps0 * wpe^2 per cell
  const T* nue;      // collision frequency per cell
  T* psih[2][6];     // ping-pong psi for H: hxy hxz hyx hyz hzx hzy
  T* psie[6];        // psi for E (owner-only, in place): exy exz eyx eyz ezx ezy
  const T* bh; const T* ah; const T* be; const T* ae;  // 2L each
  const T* inc;      // [tsteps][4][nx]: c2h Ezinc/dx, c2h Eyinc/dx, c2e0 Hyinc/dx, c2e0 Hzinc/dx
  T* eslice;         // staging of E slices: [2][nx][ny] [2][nx][nz] [2][ny][nz]
  const int* tdev;   // step index
  // Two-step tiles: the region [x2lo,x2hi) x [y2lo,y2hi) x [z2lo,z2hi), at
  // least 6 cells clear of the grid faces on every side (so clear of the
  // CPML slabs, the NTFF planes and their H averages), advanced two steps
  // per launch by fdtd_pair; its x-chunk, row and z starts.
  int x2lo, x2hi, y2lo, y2hi, z2lo, z2hi;
  const int* xstart2;
  const int* ystart2;
  const int* zstart2;
  // probes, whose neighbourhoods need step t+1 values in pair mode
  int nprobes;
  const int* px; const int* py; const int* pz;
};

// Rows and z lanes of a two-step workgroup (2 halo rows and lanes on each
// side) and the z cells a tile owns.  FP64 rows are 32 lanes wide: its
// 512-thread workgroups may use twice the registers of 1024-thread ones.
constexpr int kPairRows = 16;
template <typename T>
constexpr int kPairLanes = sizeof(T) == 4 ? 64 : 32;
template <typename T>
constexpr int kPairOwnZ = kPairLanes<T> - 4;

// Near-to-far-field surface DFT.  Every step writes the 24 surface
// sample arrays into slot t % K of a K-step buffer; every K steps (and at
// the last step) the buffer is contracted with the sin/cos table into the
// FP64 accumulators: acc[q][2m+c][pt] += sum_k s_k[q][pt] * sc[t0+k][m][c].
// A flush contracts the slots of the block so far at any step and zeroes
// them, so the block's own contraction adds only the later slots.
template <typename T>
struct NtffArgs {
  int nx, ny, nz, M, b;
  int K, T_steps;           // buffer depth, tsteps
  long long sy, sx;
  T* e[3][3];
  T* h[3][3];
  double* acc;              // 24 arrays, each [M][2][n1*n2]
  long long off[24];        // offsets of the 24 arrays in acc
  long long uoff[25];       // offsets of the 24 arrays in one sample slot (uoff[24] = U)
  long long wofs[25];       // first 16-point wavefront of each array in the MFMA contraction
  T* sbuf;                  // [K][U] surface samples
  const double* sincos;     // [tsteps][M][2]
  const int* tdev;
  // compensated FP32 sums (acc32 = 1): each 8-byte slot of acc holds the
  // pair (sum, compensation) as two floats, the contraction runs on FP32
  // matrix tiles with the FP32 copy of the phase table
  int acc32;
  const float* sincos32;    // [tsteps][M][2] when acc32
  int* health;              // set to 1 once a surface sample is NaN or infinite
};

template <typename T>
struct AuxArgs {
  int nx, ny, nz, nprobes, tsteps, cadence, nframes;
  long long sy, sx;
  T* e[3][3];
  T* h[3][3];
  const int* px; const int* py; const int* pz;
  double* probes;           // [6][tsteps][nprobes]
  int ssx, ssy, ssz;
  const T* eslice;          // staging written by the step kernel
  double* eframes;          // [nframes][2(nx ny + nx nz + ny nz)]
  double* hframes;
  unsigned long long* emax; // [nframes] bit patterns of max|Ex| (as double)
  int* tdev;
};

template <typename T>
hipError_t launch_step(const StepArgs<T>* a, const uint16_t* sids, int ty, int cur, int pm, bool slice, int zl,
                       const uint32_t* tiles, int n, hipStream_t s);
template <typename T>
hipError_t launch_pair(const StepArgs<T>* a, const uint16_t* sids, int cur, const uint32_t* tiles, int n,
                       hipStream_t s);
// The kNumLists one-step lists (list l = tiles[first[l], first[l] +
// count[l]), all of `ty` rows; tile b of solve sids[b], or solve 0 when sids
// is null) in one launch, each tile running its list's CPML variant; the
// same results as one launch_step per list.
template <typename T>
hipError_t launch_steps(const StepArgs<T>* a, const uint16_t* sids, int ty, int cur, bool slice, int zl,
                        const uint32_t* tiles, const int* first, const int* count, hipStream_t s);
template <typename T>
hipError_t launch_pair_pattern(const StepArgs<T>* a, const uint32_t* tiles, int n, hipStream_t s);
template <typename T>
hipError_t launch_ntff(const NtffArgs<T>* a, const uint2* blocks, int nblocks, hipStream_t s);
template <typename T>
hipError_t launch_ntff_contract(const NtffArgs<T>* a, long long U, int M, long long nwaves, hipStream_t s,
                                int flush = 0, int acc32 = 0);
template <typename T>
hipError_t launch_frames(const AuxArgs<T>* a, hipStream_t s);
template <typename T>
hipError_t launch_probes(const AuxArgs<T>* a, int nsolves, hipStream_t s);
template <typename T>
hipError_t launch_catchup(const StepArgs<T>* a, int cur, T* ox, T* oy, T* oz, hipStream_t s);
// Component `comp` (MSC_EX..MSC_JZ, J caught up) of cells [c0, c0 + nc) as
// FP64 into dst (device memory of nc doubles).
template <typename T>
hipError_t launch_field_slab(const StepArgs<T>* a, int cur, int comp, long long c0, long long nc, double* dst,
                             hipStream_t s);
// Grid-stride 16-byte copy of `bytes` with `blocks` workgroups of 256 threads.
hipError_t launch_stream_copy(const void* src, void* dst, size_t bytes, int blocks, hipStream_t s);

}  // namespace msc
