// HIP kernels (gfx950) of the meteor-trail FDTD time step.
//
// Reference time loop: source/meteor-fork.cpp:683-1667.  One launch of
// fdtd_step advances the whole grid by one step and fuses, per cell:
//   - the CPML psi_H updates and H corrections  (:686-791, :824-968)
//   - the H Yee update and TF/SF H corrections  (:795-934)
//   - the plasma current J, deferred by one step (:1523-1597): J^n is formed
//     from J^(n-1) and the node-averaged E^n at the start of step n instead
//     of at the end of step n-1, which is the same arithmetic on the same
//     inputs but lets E^(n+1) consume J^n in the same pass
//   - the CPML psi_E updates and E corrections  (:1049-1156, :1190-1216,
//     :1245-1275, :1373-1405)
//   - the E Yee update, TF/SF E corrections, plate and sphere (:1160-1351)
//   - the E slice capture before the Ez CPML correction (:1353-1371)
// so the interior streams E, H, J and the two media scalars once per step.
//
// E and H are ping-pong buffered: each workgroup recomputes the H values of
// its one-cell low halo (row j0-1, column k0-1) from read-only step-n data,
// which removes any ordering between workgroups.
//
// Every kernel reads its arguments from device memory (an array of per-solve
// argument blocks indexed by a per-workgroup solve id), so one launch can
// advance the tiles of many independent solves together.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "msc_kernels.h"

namespace msc {

// Argument blocks are read through the constant address space: the compiler
// then treats them as invariant for the whole kernel and keeps them in
// scalar registers / reloads them with scalar loads, as it does for kernel
// arguments, even across the kernel's own global stores.
template <typename A>
using ConstArgs = __attribute__((address_space(4))) const A;
template <typename A>
__device__ __forceinline__ ConstArgs<A>& as_const_space(const A* p) {
  return *(ConstArgs<A>*)p;
}

// Spread consecutive logical tiles over one XCD: hardware dispatch sends
// workgroup b to XCD b % 8, so logical tile ids are laid out XCD-major and
// neighbouring tiles share that XCD's L2.
__device__ __forceinline__ int xcd_remap(int b, int n) {
  constexpr int X = 8;
  const int per = n / X, rem = n % X;
  const int xcd = b % X, idx = b / X;
  return (xcd < rem ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per) + idx;
}

// CPML slab index of an H-grid position p (meteor-fork.cpp:690-700):
// [0,L) low slab, [n-L-1, n-1) high slab mapped to [L, 2L).
__device__ __forceinline__ int slab_h(int p, int n, int L) {
  if (p < L) return p;
  if (p >= n - L - 1 && p < n - 1) return p - (n - 2 * L - 1);
  return -1;
}
// E-grid slab index (meteor-fork.cpp:1053-1064): [1, L+1) -> p-1.
__device__ __forceinline__ int slab_e(int p, int n, int L) {
  if (p >= 1 && p < L + 1) return p - 1;
  if (p >= n - L - 1 && p < n - 1) return p - (n - 2 * L - 1);
  return -1;
}

// Plasma current update for one cell (meteor-fork.cpp:1523-1597):
//   CUR == kCurDI:       Young's direct integration (doB0 = 0, :1525-1536)
//   CUR == kCurLKPoly:   Lee & Kalluri (doB0 = 1, :1546-1595) with the
//                        per-cell coefficients e^{-x}, kI, kX, kP evaluated in
//                        FP64 as Horner polynomials in x = nu dt whose FP64
//                        coefficients the host derives from the closed-form
//                        integrals (exact to FP64 over the solve's nu range)
//   CUR == kCurLKClosed: Lee & Kalluri with the reference's closed form in
//                        FP64 (finite limits at nu == 0 and |B| == 0)
// J^(n+1) = e^{-x} R J^n + w (kI E + kX (w_b x E) + kP w_b (w_b . E)), w = eps0 wp^2,
// with E the node-averaged field for Lee & Kalluri and the collocated one for DI.
template <typename T, int CUR, typename A>
__device__ __forceinline__ void current_update(const A& a, T nu_t, T w_t, T& jx, T& jy, T& jz,
                                               T ex, T ey, T ez) {
  const double nu = nu_t, w = w_t;
  if constexpr (CUR == kCurDI) {
    const double g = 2 + nu * a.dt;
    const double f = (2 - nu * a.dt) / g;
    const double s = 2 * ((w * a.dt) / g);
    jx = (T)(f * (double)jx + s * (double)ex);
    jy = (T)(f * (double)jy + s * (double)ey);
    jz = (T)(f * (double)jz + s * (double)ez);
    return;
  } else {
    double Ee1, kP, kI, kX;
    if constexpr (CUR == kCurLKPoly) {
      // Horner from degree lkdeg (a loop: unrolling it would hoist every
      // coefficient of every degree into registers across the x march)
      const double x = nu * a.dt;
      const int d = a.lkdeg;
      double pa = a.lkc[0][d], pi = a.lkc[1][d], px = a.lkc[2][d], pp = a.lkc[3][d];
      for (int n = d - 1; n >= 0; --n) {
        pa = fma(pa, x, a.lkc[0][n]);
        pi = fma(pi, x, a.lkc[1][n]);
        px = fma(px, x, a.lkc[2][n]);
        pp = fma(pp, x, a.lkc[3][n]);
      }
      Ee1 = pa; kI = pi; kX = px; kP = pp;
    } else {
      if (nu > 0 && a.w0 > 0) {
        Ee1 = exp(-nu * a.dt);
        const double Ee2 = 1 / (a.w0 * a.w0 + nu * nu);
        const double Ce2 = (1 - Ee1) / nu - Ee1 * nu * a.Ce1 - Ee1 * a.Se1;
        const double Ce3 = nu * (1 - Ee1 * a.cth) + Ee1 * a.w0 * a.sth;
        const double Ce4 = 1 - Ee1 * a.cth - Ee1 * nu * a.Se1;
        kP = Ee2 * Ce2;
        kI = Ee2 * Ce3;
        kX = Ee2 * Ce4;
      } else if (a.w0 > 0) {
        Ee1 = 1;
        kI = a.Se1;
        kX = a.Ce1;
        kP = (a.dt - a.Se1) / (a.w0 * a.w0);
      } else {
        Ee1 = exp(-nu * a.dt);
        kI = nu > 0 ? -expm1(-nu * a.dt) / nu : a.dt;
        kX = 0;
        kP = 0;
      }
    }
    // per-cell coefficient products in FP64, then the update in T
    const T ca = (T)Ee1, ci = (T)(w * kI), cx = (T)(w * kX), cp = (T)(w * kP);
    const T wx = (T)a.wx, wy = (T)a.wy, wz = (T)a.wz;
    const T rx = (T)a.R[0] * jx + (T)a.R[1] * jy + (T)a.R[2] * jz;
    const T ry = (T)a.R[3] * jx + (T)a.R[4] * jy + (T)a.R[5] * jz;
    const T rz = (T)a.R[6] * jx + (T)a.R[7] * jy + (T)a.R[8] * jz;
    const T wE = cp * (wx * ex + wy * ey + wz * ez);
    jx = ca * rx + ci * ex + cx * (wy * ez - wz * ey) + wE * wx;
    jy = ca * ry + ci * ey + cx * (wz * ex - wx * ez) + wE * wy;
    jz = ca * rz + ci * ez + cx * (wx * ey - wy * ex) + wE * wz;
  }
}

// Global access through a uniform base pointer and a 32-bit byte offset, so
// the compiler can use the scalar-base + 32-bit vector-offset addressing
// form and keep one VGPR per address instead of two.  The host only takes
// this path when every array is below 4 GiB.
template <typename T>
__device__ __forceinline__ T ldo(const T* __restrict__ base, uint32_t boff) {
  return *reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + boff);
}
template <typename T>
__device__ __forceinline__ void sto(T* __restrict__ base, uint32_t boff, T v) {
  *reinterpret_cast<T*>(reinterpret_cast<char*>(base) + boff) = v;
}

// Workgroup barrier that orders LDS only.  __syncthreads() is a
// workgroup-scope fence and so also waits for every outstanding global load,
// which would drain the x-direction prefetch pipeline twice per plane.  The
// barriers in fdtd_step only publish LDS data: no thread of a workgroup reads
// global memory another thread of it writes during the step (E and H are
// ping-pong buffers, J and psi_E are owner-only).
__device__ __forceinline__ void lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory");
}

// Whether positions [lo, hi) include f - 1 or f for a TF/SF face f = f1 or
// f2: the only positions where a lane's face signs can be nonzero.  A tile
// whose rows and lanes hold none of them weights the per-plane incident
// terms by zero on every lane, so it skips their loads and corrections (the
// field is then unchanged, where fma(0, term, field) could turn a field of
// exactly -0 into +0).
__device__ __forceinline__ bool holds_face(int lo, int hi, int f1, int f2) {
  return (f1 >= lo && f1 <= hi) || (f2 >= lo && f2 <= hi);
}

template <typename A>
__device__ __forceinline__ bool in_sphere(const A& a, int i, int j, int k) {
  // meteor-fork.cpp:203-209, coordinates as :137-141 (one loop per axis)
  const double x = __dadd_rn(a.x0, __dmul_rn(a.dd, (double)i));
  const double y = __dadd_rn(a.y0, __dmul_rn(a.dd, (double)j));
  const double z = __dadd_rn(a.z0, __dmul_rn(a.dd, (double)k));
  return sqrt(__dadd_rn(__dadd_rn(__dmul_rn(x, x), __dmul_rn(y, y)), __dmul_rn(z, z))) < a.rsph;
}

// LDS of a one-step workgroup.  sE[b]: one x-plane of E^n with a +1 row
// and +1 column halo.  Three buffers rotate: plane i (read), plane i+1
// (staged at the start of plane i), and plane i-1, which slower threads may
// still read in phase C of plane i-1 while plane i+1 is being staged.
template <typename T, int TY, int ZL>
struct StepLds {
  T sE[3][3][TY + 2][ZL + 1];
  T sH[3][TY + 1][ZL];
};

// One one-step tile (tile `blk` of the list `tiles`) with the CPML code of
// slab axes PM.  The body of fdtd_step, and of fdtd_steps for each variant.
template <typename T, int TY, int CUR, int PM, bool SLICE, int ZL>
__device__ __forceinline__ void step_tile(const StepArgs<T>* __restrict__ args, const uint32_t* __restrict__ tiles,
                                          const uint16_t* __restrict__ sids, int blk, StepLds<T, TY, ZL>& lds) {
  constexpr int R = TY + 1;  // rows including the j0-1 halo row
  constexpr uint32_t W = sizeof(T);
  auto& sE = lds.sE;
  auto& sH = lds.sH;

  const int lane = threadIdx.x & (ZL - 1);
  const int r = threadIdx.x / ZL;
  const uint32_t tile = tiles[blk];
  const ConstArgs<StepArgs<T>>& a = as_const_space(args + (sids ? sids[blk] : 0));
  const int tz = tile & 0xff;
  const int ty = (tile >> 8) & 0xfff;
  const int chunk = (tile >> 20) & 0x7ff;
  const bool cj = (tile & kTileNoCurrent) == 0;  // the tile carries a current
  const int k0 = a.zstart[tz], k1 = a.zstart[tz + 1];
  const int* __restrict__ yst = PM ? a.ystart_b : a.ystart;
  const int j0 = yst[ty], j1 = yst[ty + 1];
  const int ib = a.xstart[chunk], ie = a.xstart[chunk + 1];

  const int nx = a.nx, ny = a.ny, nz = a.nz, L = a.L;
  const int j = j0 - 1 + r;
  const int k = k0 - 1 + lane;
  const bool valid = j >= 0 && j < ny && k >= 0 && k < nz;
  const bool own = valid && r >= 1 && lane >= 1 && j < j1 && k < k1;
  const int istart = ib > 0 ? ib - 1 : 0;
  // owned cells inside the current's domain [1,N-1)^3 (the x test is per plane)
  const bool jdom = own && j >= 1 && j < ny - 1 && k >= 1 && k < nz - 1;
  // TF/SF face signs of this lane's (j, k), fixed along the x march
  // (meteor-fork.cpp:808-822, 868-875, 927-934, 1175-1188, 1236-1243,
  // 1295-1302); the x conditions are uniform per plane.  Each correction is
  // fma(sign, term, field): exactly field +/- term, or field for sign 0.
  const bool jin = j >= a.by1 && j < a.by2, kin = k >= a.bz1 && k < a.bz2;
  const T sHx0 = kin ? (T)((int)(j == a.by1 - 1) - (int)(j == a.by2 - 1)) : (T)0;
  const T sHx1 = jin ? (T)((int)(k == a.bz2 - 1) - (int)(k == a.bz1 - 1)) : (T)0;
  const T sEx2 = jin ? (T)((int)(k == a.bz1) - (int)(k == a.bz2)) : (T)0;
  const T sEx3 = kin ? (T)((int)(j == a.by2) - (int)(j == a.by1)) : (T)0;
  const T mJK = (jin && kin) ? (T)1 : (T)0;
  const bool tfl = holds_face(j0 - 1, j0 - 1 + R, a.by1, a.by2) || holds_face(k0 - 1, k0 - 1 + ZL, a.bz1, a.bz2);

  const int t = *a.tdev;
  const int par = t & 1;  // psi_H ping-pong
  const int si = t % 3, so = (t + 1) % 3, sj = (t + 2) % 3;  // field sets: read, write, J^(t-1)
  const T* __restrict__ Exo = a.e[si][0];
  const T* __restrict__ Eyo = a.e[si][1];
  const T* __restrict__ Ezo = a.e[si][2];
  const T* __restrict__ Hxo = a.h[si][0];
  const T* __restrict__ Hyo = a.h[si][1];
  const T* __restrict__ Hzo = a.h[si][2];
  T* __restrict__ Exn = a.e[so][0];
  T* __restrict__ Eyn = a.e[so][1];
  T* __restrict__ Ezn = a.e[so][2];
  T* __restrict__ Hxn = a.h[so][0];
  T* __restrict__ Hyn = a.h[so][1];
  T* __restrict__ Hzn = a.h[so][2];
  const T* __restrict__ Jxo = a.j[sj][0];
  const T* __restrict__ Jyo = a.j[sj][1];
  const T* __restrict__ Jzo = a.j[sj][2];
  T* __restrict__ Jxn = a.j[si][0];
  T* __restrict__ Jyn = a.j[si][1];
  T* __restrict__ Jzn = a.j[si][2];
  const T* __restrict__ inc = a.inc + (size_t)t * 4 * nx;
  const T c2h = a.c2h, c2e0 = a.c2e0, inv = a.inv_d;
  // TF/SF values at fixed planes (meteor-fork.cpp:870-872, 929-931, 1238-1240, 1297-1299)
  const T tfEzB1 = inc[a.bx1], tfEzB2 = inc[a.bx2 - 1];
  const T tfEyB1 = inc[nx + a.bx1], tfEyB2 = inc[nx + a.bx2 - 1];
  const T tfHyB1 = inc[2 * nx + a.bx1 - 1], tfHyB2 = inc[2 * nx + a.bx2 - 1];
  const T tfHzB1 = inc[3 * nx + a.bx1 - 1], tfHzB2 = inc[3 * nx + a.bx2 - 1];
  // byte strides (every array is < 4 GiB on this path)
  const uint32_t sxb = (uint32_t)a.sx * W, syb = (uint32_t)a.sy * W;

  const bool xrow = (r == R - 1);
  const bool xcol = (lane == ZL - 1);
  const bool rowok = xrow && (j + 1) < ny && k >= 0 && k < nz;
  const bool colok = xcol && j >= 0 && j < ny && (k + 1) < nz;
  // byte offset of (0, j, k); only dereferenced when valid / rowok / colok
  const uint32_t ob = valid || rowok || colok ? ((uint32_t)j * (uint32_t)a.sy + (uint32_t)k) * W : 0u;

  // Pipeline along x: one batch of loads per plane, issued in phase A of
  // plane i and consumed in plane i+1, so the only wait on global loads is at
  // the start of the next plane, one full plane of work later.  The batch
  // holds E^n(i+2) (staged into LDS at the start of plane i+1), and
  // H^(n-1/2), J^(n-1), the media and the incident-line terms of plane i+1.
  T px = 0, py = 0, pz = 0, prx = 0, prz = 0, pcx = 0, pcy = 0;  // E^n staging registers
  T nhx = 0, nhy = 0, nhz = 0;                                   // H^(n-1/2), next plane
  T nqx = 0, nqy = 0, nqz = 0, nqw = 0, nqn = 0;                 // J^(n-1), eps0 wp^2, nu, next plane
  T nt0 = 0, nt1 = 0, nt2 = 0, nt3 = 0;                          // incident-line terms, next plane
  {
    const uint32_t c = (uint32_t)istart * sxb + ob;
    T ex = 0, ey = 0, ez = 0, rx = 0, rz = 0, cx = 0, cy = 0;
    if (valid) {
      ex = ldo(Exo, c); ey = ldo(Eyo, c); ez = ldo(Ezo, c);
      nhx = ldo(Hxo, c); nhy = ldo(Hyo, c); nhz = ldo(Hzo, c);
    }
    if (rowok) { rx = ldo(Exo, c + syb); rz = ldo(Ezo, c + syb); }
    if (colok) { cx = ldo(Exo, c + W); cy = ldo(Eyo, c + W); }
    sE[0][0][r][lane] = ex; sE[0][1][r][lane] = ey; sE[0][2][r][lane] = ez;
    if (xrow) { sE[0][0][R][lane] = rx; sE[0][2][R][lane] = rz; }
    if (xcol) { sE[0][0][r][ZL] = cx; sE[0][1][r][ZL] = cy; }
    if (istart + 1 < nx) {
      const uint32_t cn = c + sxb;
      if (valid) { px = ldo(Exo, cn); py = ldo(Eyo, cn); pz = ldo(Ezo, cn); }
      if (rowok) { prx = ldo(Exo, cn + syb); prz = ldo(Ezo, cn + syb); }
      if (colok) { pcx = ldo(Exo, cn + W); pcy = ldo(Eyo, cn + W); }
    }
    // (no J or media for plane istart: it is either the warm-up plane ib-1,
    // which only recomputes H, or plane 0, outside the current's domain)
    if (tfl) { nt0 = inc[istart]; nt1 = inc[nx + istart]; nt2 = inc[2 * nx + istart]; nt3 = inc[3 * nx + istart]; }
  }
  T exprev = 0, hyprev = 0, hzprev = 0;
  int buf = 0;

  for (int i = istart; i < ie; ++i, buf = (buf == 2 ? 0 : buf + 1)) {
    const uint32_t c = (uint32_t)i * sxb + ob;
    const int nb = buf == 2 ? 0 : buf + 1;
    // ---- phase A: take the previous batch, stage E^n(i+1), issue the next batch
    const T hx0 = nhx, hy0 = nhy, hz0 = nhz;
    const T qx = nqx, qy = nqy, qz = nqz, qw = nqw, qn = nqn;
    const T tf0 = nt0, tf1 = nt1, tf2 = nt2, tf3 = nt3;
    if (i + 1 < nx) {
      sE[nb][0][r][lane] = px; sE[nb][1][r][lane] = py; sE[nb][2][r][lane] = pz;
      if (xrow) { sE[nb][0][R][lane] = prx; sE[nb][2][R][lane] = prz; }
      if (xcol) { sE[nb][0][r][ZL] = pcx; sE[nb][1][r][ZL] = pcy; }
    }
    if (i + 1 < ie) {
      const uint32_t cn = c + sxb;
      if (i + 2 < nx) {
        const uint32_t c2 = cn + sxb;
        if (valid) { px = ldo(Exo, c2); py = ldo(Eyo, c2); pz = ldo(Ezo, c2); }
        if (rowok) { prx = ldo(Exo, c2 + syb); prz = ldo(Ezo, c2 + syb); }
        if (colok) { pcx = ldo(Exo, c2 + W); pcy = ldo(Eyo, c2 + W); }
      }
      if (valid) { nhx = ldo(Hxo, cn); nhy = ldo(Hyo, cn); nhz = ldo(Hzo, cn); }
      if (cj && jdom && i + 1 >= 1 && i + 1 < nx - 1) {
        nqx = ldo(Jxo, cn); nqy = ldo(Jyo, cn); nqz = ldo(Jzo, cn);
        nqw = ldo(a.wpe2, cn); nqn = ldo(a.nue, cn);
      }
      if (tfl) { nt0 = inc[i + 1]; nt1 = inc[nx + i + 1]; nt2 = inc[2 * nx + i + 1]; nt3 = inc[3 * nx + i + 1]; }
    }
    lds_barrier();

    // ---- phase B: H^(n+1/2) at plane i (owned cells and the low halo)
    const bool wr = own && i >= ib;
    T hx = hx0, hy = hy0, hz = hz0;
    const T ecx = sE[buf][0][r][lane], ecy = sE[buf][1][r][lane], ecz = sE[buf][2][r][lane];
    if (valid) {
      const T ezjp = sE[buf][2][r + 1][lane];
      const T eykp = sE[buf][1][r][lane + 1];
      const T exkp = sE[buf][0][r][lane + 1];
      const T exjp = sE[buf][0][r + 1][lane];
      const T eny = sE[nb][1][r][lane], enz = sE[nb][2][r][lane];
      // Hx: i in [0,nx), j in [0,ny-1), k in [0,nz-1)   (:797-806)
      if (j < ny - 1 && k < nz - 1) {
        const T dEz = (ezjp - ecz) * inv;
        const T dEy = (eykp - ecy) * inv;
        hx = hx - c2h * (dEz - dEy);
        if (tfl && i >= a.bx1 && i < a.bx2) {  // TF/SF (:808-822)
          hx = fma(sHx0, tf0, hx);
          hx = fma(sHx1, tf1, hx);
        }
        if constexpr (PM != 0) {  // (:690-714, :824-850)
          if constexpr ((PM & 2) != 0) {
            const int jj = slab_h(j, ny, L);
            if (jj >= 0) {
              const uint32_t q = (((uint32_t)i * 2 * L + jj) * nz + k) * W;
              const T p = a.bh[jj] * ldo(a.psih[par][0], q) + a.ah[jj] * dEz;
              if (wr) sto(a.psih[par ^ 1][0], q, p);
              hx = hx - c2h * p;
            }
          }
          if constexpr ((PM & 4) != 0) {
            const int kk = slab_h(k, nz, L);
            if (kk >= 0) {
              const uint32_t q = (((uint32_t)i * ny + j) * 2 * L + kk) * W;
              const T p = a.bh[kk] * ldo(a.psih[par][1], q) + a.ah[kk] * dEy;
              if (wr) sto(a.psih[par ^ 1][1], q, p);
              hx = hx + c2h * p;
            }
          }
        }
        if (wr) sto(Hxn, c, hx);
      }
      // Hy: i in [0,nx-1), j in [0,ny), k in [0,nz-1)   (:854-865)
      if (i < nx - 1 && k < nz - 1) {
        const T dEx = (exkp - ecx) * inv;
        const T dEz = (enz - ecz) * inv;
        hy = hy - c2h * (dEx - dEz);
        if (i == a.bx1 - 1) hy = fma(-mJK, tfEzB1, hy);  // (:868-875)
        if (i == a.bx2 - 1) hy = fma(mJK, tfEzB2, hy);
        if constexpr (PM != 0) {  // (:718-744, :877-910)
          if constexpr ((PM & 1) != 0) {
            const int ii = slab_h(i, nx, L);
            if (ii >= 0) {
              const uint32_t q = (((uint32_t)ii * ny + j) * nz + k) * W;
              const T p = a.bh[ii] * ldo(a.psih[par][2], q) + a.ah[ii] * dEz;
              if (wr) sto(a.psih[par ^ 1][2], q, p);
              hy = hy + c2h * p;
            }
          }
          if constexpr ((PM & 4) != 0) {
            const int kk = slab_h(k, nz, L);
            if (kk >= 0) {
              const uint32_t q = (((uint32_t)i * ny + j) * 2 * L + kk) * W;
              const T p = a.bh[kk] * ldo(a.psih[par][3], q) + a.ah[kk] * dEx;
              if (wr) sto(a.psih[par ^ 1][3], q, p);
              hy = hy - c2h * p;
            }
          }
        }
        if (wr) sto(Hyn, c, hy);
      }
      // Hz: i in [0,nx-1), j in [0,ny-1), k in [0,nz)   (:913-924)
      if (i < nx - 1 && j < ny - 1) {
        const T dEy = (eny - ecy) * inv;
        const T dEx = (exjp - ecx) * inv;
        hz = hz - c2h * (dEy - dEx);
        if (i == a.bx1 - 1) hz = fma(mJK, tfEyB1, hz);  // (:927-934)
        if (i == a.bx2 - 1) hz = fma(-mJK, tfEyB2, hz);
        if constexpr (PM != 0) {  // (:747-791, :936-968)
          if constexpr ((PM & 1) != 0) {
            const int ii = slab_h(i, nx, L);
            if (ii >= 0) {
              const uint32_t q = (((uint32_t)ii * ny + j) * nz + k) * W;
              const T p = a.bh[ii] * ldo(a.psih[par][4], q) + a.ah[ii] * dEy;
              if (wr) sto(a.psih[par ^ 1][4], q, p);
              hz = hz - c2h * p;
            }
          }
          if constexpr ((PM & 2) != 0) {
            const int jj = slab_h(j, ny, L);
            if (jj >= 0) {
              const uint32_t q = (((uint32_t)i * 2 * L + jj) * nz + k) * W;
              const T p = a.bh[jj] * ldo(a.psih[par][5], q) + a.ah[jj] * dEx;
              if (wr) sto(a.psih[par ^ 1][5], q, p);
              hz = hz + c2h * p;
            }
          }
        }
        if (wr) sto(Hzn, c, hz);
      }
    }
    sH[0][r][lane] = hx;
    sH[1][r][lane] = hy;
    sH[2][r][lane] = hz;
    lds_barrier();

    // ---- phase C: J^n (deferred) and E^(n+1) on owned cells
    if (wr) {
      // J^n from J^(n-1) and E^n; domain [1,N-1)^3   (:1523-1597)
      T jx = 0, jy = 0, jz = 0;
      if (cj && jdom && i >= 1 && i < nx - 1) {
        jx = qx;
        jy = qy;
        jz = qz;
        if constexpr (CUR == kCurDI) {
          current_update<T, CUR>(a, qn, qw, jx, jy, jz, ecx, ecy, ecz);
        } else {
          if (a.jcol) {
            current_update<T, CUR>(a, qn, qw, jx, jy, jz, ecx, ecy, ecz);
          } else {
            const T mx = (ecx + exprev) / 2;
            const T my = (ecy + sE[buf][1][r - 1][lane]) / 2;
            const T mz = (ecz + sE[buf][2][r][lane - 1]) / 2;
            current_update<T, CUR>(a, qn, qw, jx, jy, jz, mx, my, mz);
          }
        }
        sto(Jxn, c, jx);
        sto(Jyn, c, jy);
        sto(Jzn, c, jz);
      }
      T c1 = 1, c2 = c2e0;
      if (a.doplate == 2 && in_sphere(a, i, j, k)) { c1 = a.c1s; c2 = a.c2s; }
      // Ex: i in [0,nx-1), j in [1,ny-1), k in [1,nz-1)   (:1161-1216)
      if (i < nx - 1 && j >= 1 && j < ny - 1 && k >= 1 && k < nz - 1) {
        const T dHz = (hz - sH[2][r - 1][lane]) * inv;
        const T dHy = (hy - sH[1][r][lane - 1]) * inv;
        T ex = c1 * ecx + c2 * (dHz - dHy) - c2 * jx;
        if (tfl && i >= a.bx1 && i < a.bx2 - 1) {
          ex = fma(sEx2, tf2, ex);
          ex = fma(sEx3, tf3, ex);
        }
        if constexpr (PM != 0) {
          if constexpr ((PM & 2) != 0) {
            const int jj = slab_e(j, ny, L);
            if (jj >= 0) {
              const uint32_t q = (((uint32_t)i * 2 * L + jj) * nz + k) * W;
              const T p = a.be[jj] * ldo(a.psie[0], q) + a.ae[jj] * dHz;
              sto(a.psie[0], q, p);
              ex = ex + c2e0 * p;
            }
          }
          if constexpr ((PM & 4) != 0) {
            const int kk = slab_e(k, nz, L);
            if (kk >= 0) {
              const uint32_t q = (((uint32_t)i * ny + j) * 2 * L + kk) * W;
              const T p = a.be[kk] * ldo(a.psie[1], q) + a.ae[kk] * dHy;
              sto(a.psie[1], q, p);
              ex = ex - c2e0 * p;
            }
          }
        }
        sto(Exn, c, ex);
      }
      // Ey: i in [1,nx-1), j in [0,ny), k in [1,nz)   (:1221-1275)
      T ey = ecy;
      if (i >= 1 && i < nx - 1 && k >= 1) {
        const T dHx = (hx - sH[0][r][lane - 1]) * inv;
        const T dHz = (hz - hzprev) * inv;
        ey = c1 * ecy + c2 * (dHx - dHz) - c2 * jy;
        if (i == a.bx1) ey = fma(mJK, tfHzB1, ey);
        if (i == a.bx2 - 1) ey = fma(-mJK, tfHzB2, ey);
        if constexpr (PM != 0) {
          if constexpr ((PM & 1) != 0) {
            const int ii = slab_e(i, nx, L);
            if (ii >= 0 && j < ny - 1 && k < nz - 1) {
              const uint32_t q = (((uint32_t)ii * ny + j) * nz + k) * W;
              const T p = a.be[ii] * ldo(a.psie[2], q) + a.ae[ii] * dHz;
              sto(a.psie[2], q, p);
              ey = ey - c2e0 * p;
            }
          }
          if constexpr ((PM & 4) != 0) {
            const int kk = slab_e(k, nz, L);
            if (kk >= 0 && j < ny - 1) {
              const uint32_t q = (((uint32_t)i * ny + j) * 2 * L + kk) * W;
              const T p = a.be[kk] * ldo(a.psie[3], q) + a.ae[kk] * dHx;
              sto(a.psie[3], q, p);
              ey = ey + c2e0 * p;
            }
          }
        }
        if (a.doplate == 1 && i == a.ssx && j >= a.plate_lo && j < a.plate_up - 1 && k >= a.plate_lo && k < a.plate_up)
          ey = 0;  // (:1337-1343)
        sto(Eyn, c, ey);
      }
      // Ez: i in [1,nx-1), j in [1,ny), k in [0,nz)   (:1279-1302, :1373-1405)
      T ez = ecz, ezpre = ecz;
      if (i >= 1 && i < nx - 1 && j >= 1) {
        const T dHy = (hy - hyprev) * inv;
        const T dHx = (hx - sH[0][r - 1][lane]) * inv;
        ez = c1 * ecz + c2 * (dHy - dHx) - c2 * jz;
        if (i == a.bx1) ez = fma(-mJK, tfHyB1, ez);
        if (i == a.bx2 - 1) ez = fma(mJK, tfHyB2, ez);
        if (a.doplate == 1 && i == a.ssx && j >= a.plate_lo && j < a.plate_up && k >= a.plate_lo && k < a.plate_up - 1)
          ez = 0;  // (:1344-1350)
        ezpre = ez;
        if constexpr (PM != 0) {
          if constexpr ((PM & 1) != 0) {
            const int ii = slab_e(i, nx, L);
            if (ii >= 0 && j < ny - 1 && k < nz - 1) {
              const uint32_t q = (((uint32_t)ii * ny + j) * nz + k) * W;
              const T p = a.be[ii] * ldo(a.psie[4], q) + a.ae[ii] * dHy;
              sto(a.psie[4], q, p);
              ez = ez + c2e0 * p;
            }
          }
          if constexpr ((PM & 2) != 0) {
            const int jj = slab_e(j, ny, L);
            if (jj >= 0 && k < nz - 1) {
              const uint32_t q = (((uint32_t)i * 2 * L + jj) * nz + k) * W;
              const T p = a.be[jj] * ldo(a.psie[5], q) + a.ae[jj] * dHx;
              sto(a.psie[5], q, p);
              ez = ez - c2e0 * p;
            }
          }
        }
        sto(Ezn, c, ez);
      }
      if (SLICE) {  // E slices before the Ez CPML correction (:1353-1371)
        T* s = a.eslice;
        const size_t zxy = (size_t)nx * ny, zxz = (size_t)nx * nz;
        if (k == a.ssz) { s[(size_t)i * ny + j] = ey; s[zxy + (size_t)i * ny + j] = ezpre; }
        s += 2 * zxy;
        if (j == a.ssy) { s[(size_t)i * nz + k] = ey; s[zxz + (size_t)i * nz + k] = ezpre; }
        s += 2 * zxz;
        if (i == a.ssx) { s[(size_t)j * nz + k] = ey; s[(size_t)ny * nz + (size_t)j * nz + k] = ezpre; }
      }
    }
    exprev = ecx;
    hyprev = hy;
    hzprev = hz;
  }
}

template <typename T, int TY, int CUR, int PM, bool SLICE, int ZL>
__global__ void __launch_bounds__((TY + 1) * ZL)
fdtd_step(const StepArgs<T>* __restrict__ args, const uint32_t* __restrict__ tiles,
          const uint16_t* __restrict__ sids) {
  __shared__ StepLds<T, TY, ZL> lds;
  step_tile<T, TY, CUR, PM, SLICE, ZL>(args, tiles, sids, xcd_remap(blockIdx.x, gridDim.x), lds);
}

// The five one-step lists of a step in one launch, so a small grid's lists
// (or a small gang's) fill the device together instead of one launch each.
// List l (interior, x, y, z and all slab axes, whose variants step_tile
// carries) is tiles[first[l] .. first[l] + n_l), and its workgroups are
// [end[l-1], end[l]) (end[-1] = 0).  Every tile runs its list's code;
// interior and CPML tiles must have the same rows (TY).
struct ListSpans {
  int first[kNumLists];
  int end[kNumLists];
};
template <typename T, int TY, int CUR, bool SLICE, int ZL>
__global__ void __launch_bounds__((TY + 1) * ZL)
fdtd_steps(const StepArgs<T>* __restrict__ args, const uint32_t* __restrict__ tiles,
           const uint16_t* __restrict__ sids, ListSpans sp) {
  __shared__ StepLds<T, TY, ZL> lds;
  const int blk = xcd_remap(blockIdx.x, gridDim.x);
  if (blk < sp.end[kListInterior])
    step_tile<T, TY, CUR, 0, SLICE, ZL>(args, tiles, sids, sp.first[kListInterior] + blk, lds);
  else if (blk < sp.end[kListX])
    step_tile<T, TY, CUR, kPmlX, SLICE, ZL>(args, tiles, sids, sp.first[kListX] + blk - sp.end[kListInterior], lds);
  else if (blk < sp.end[kListY])
    step_tile<T, TY, CUR, kPmlY, SLICE, ZL>(args, tiles, sids, sp.first[kListY] + blk - sp.end[kListX], lds);
  else if (blk < sp.end[kListZ])
    step_tile<T, TY, CUR, kPmlZ, SLICE, ZL>(args, tiles, sids, sp.first[kListZ] + blk - sp.end[kListY], lds);
  else
    step_tile<T, TY, CUR, kPmlAll, SLICE, ZL>(args, tiles, sids, sp.first[kListAll] + blk - sp.end[kListZ], lds);
}

// Two time steps per launch on the tiles of the two-step region (no CPML
// terms, no slice capture: the region is clear of both).  Step t is
// computed on the owned cells plus a halo of two cells on every side in y
// and z, step t+1 on the owned cells, so every value step t+1 needs is
// produced inside the workgroup and tiles stay independent.  The march
// along x is skewed by one plane: iteration i computes step t at plane i
// and step t+1 at plane i-1, with the step-t planes i-1 and i of E and the
// step-t plane i of H held in LDS.  Per two steps each of E, H, J and the
// media is read once (set t % 3, J^(t-1)) and E, H, J written once (set
// (t+2) % 3, J^(t+1)).  Step-t values are written (set (t+1) % 3) only for
// owned cells the one-step launches around the region read as halo (the
// two outermost layers of the region) and for probe neighbourhoods.  The
// arithmetic of every cell and step is the one-step kernel's, term by term.
template <typename T, int CUR, int PR, int PL>
__global__ void __launch_bounds__(PR * PL)
fdtd_pair(const StepArgs<T>* __restrict__ args, const uint32_t* __restrict__ tiles,
          const uint16_t* __restrict__ sids) {
  constexpr int R = PR;  // rows; PL z lanes per row
  constexpr uint32_t W = sizeof(T);
  __shared__ T s0[2][3][R][PL];   // E^t of planes i and i+1 (by parity)
  __shared__ T s1h[3][R][PL];     // H^(t+1/2) of plane i
  __shared__ T s1e[2][3][R][PL];  // E^(t+1) of planes i-1 and i (by parity)
  __shared__ T s2h[3][R][PL];     // H^(t+3/2) of plane i-1

  const int lane = threadIdx.x & (PL - 1);
  const int r = threadIdx.x / PL;
  const int blk = xcd_remap(blockIdx.x, gridDim.x);
  const uint32_t tile = tiles[blk];
  const ConstArgs<StepArgs<T>>& a = as_const_space(args + (sids ? sids[blk] : 0));
  const int k0 = a.zstart2[tile & 0xff], k1 = a.zstart2[(tile & 0xff) + 1];
  const int j0 = a.ystart2[(tile >> 8) & 0xfff], j1 = a.ystart2[((tile >> 8) & 0xfff) + 1];
  const int ib = a.xstart2[(tile >> 20) & 0x7ff], ie = a.xstart2[((tile >> 20) & 0x7ff) + 1];
  const bool cj = (tile & kTileNoCurrent) == 0;  // the tile carries a current

  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const int j = j0 - 2 + r;
  const int k = k0 - 2 + lane;
  // positions inside the grid away from its faces: every domain condition
  // of the one-step update holds there
  const bool inside = j >= 1 && j < ny - 1 && k >= 1 && k < nz - 1;
  const bool valid = j >= 0 && j < ny && k >= 0 && k < nz;
  const bool h1ok = inside && r < R - 1 && lane < PL - 1;  // H^(t+1/2): needs the +1 neighbours
  const bool e1ok = h1ok && r >= 1 && lane >= 1;                // E^(t+1): needs the -1 H neighbours
  const bool h2ok = e1ok && r < R - 2 && lane < PL - 2;     // H^(t+3/2)
  const bool own = inside && r >= 2 && r < R - 2 && j < j1 && lane >= 2 && lane < PL - 2 && k < k1;
  // owned cells whose step-t values the one-step launches or the probes
  // read: the two outer layers of the region and the 3x3x3 box of each probe
  // (the x planes of up to two probes whose (j, k) box holds the cell are
  // kept; a cell in the boxes of more writes back on every plane, which
  // stores only exact step-t+1 values)
  const bool rimjk = j < a.y2lo + 2 || j >= a.y2hi - 2 || k < a.z2lo + 2 || k >= a.z2hi - 2;
  int pxa = -8, pxb = -8;
  bool pall = false;
  for (int p = 0; p < a.nprobes; ++p)
    if (abs(j - a.py[p]) <= 1 && abs(k - a.pz[p]) <= 1) {
      if (pxa == -8) pxa = a.px[p];
      else if (pxb == -8) pxb = a.px[p];
      else pall = true;
    }

  const bool jin = j >= a.by1 && j < a.by2, kin = k >= a.bz1 && k < a.bz2;
  const T sHx0 = kin ? (T)((int)(j == a.by1 - 1) - (int)(j == a.by2 - 1)) : (T)0;
  const T sHx1 = jin ? (T)((int)(k == a.bz2 - 1) - (int)(k == a.bz1 - 1)) : (T)0;
  const T sEx2 = jin ? (T)((int)(k == a.bz1) - (int)(k == a.bz2)) : (T)0;
  const T sEx3 = kin ? (T)((int)(j == a.by2) - (int)(j == a.by1)) : (T)0;
  const T mJK = (jin && kin) ? (T)1 : (T)0;
  const bool tfl = holds_face(j0 - 2, j0 - 2 + R, a.by1, a.by2) || holds_face(k0 - 2, k0 - 2 + PL, a.bz1, a.bz2);

  const int t = *a.tdev;
  const int s0i = t % 3, s1o = (t + 1) % 3, s2o = (t + 2) % 3;  // step t, t+1, t+2 field sets
  const T* __restrict__ Exo = a.e[s0i][0];
  const T* __restrict__ Eyo = a.e[s0i][1];
  const T* __restrict__ Ezo = a.e[s0i][2];
  const T* __restrict__ Hxo = a.h[s0i][0];
  const T* __restrict__ Hyo = a.h[s0i][1];
  const T* __restrict__ Hzo = a.h[s0i][2];
  const T* __restrict__ Jxo = a.j[s2o][0];  // J^(t-1)
  const T* __restrict__ Jyo = a.j[s2o][1];
  const T* __restrict__ Jzo = a.j[s2o][2];
  // incident terms, prefetched one plane ahead on the tiles that hold a TF/SF
  // face (tfl; at 400x400x800 one tile in fifteen).  FP64 and the closed-form
  // current read them with scalar loads (they are uniform; waited for at the
  // next barrier), which frees VGPRs there (FP64 Lee-Kalluri: 31, and its
  // launch goes from 5.61 to 4.98 ms at 400x400x800; the FP32 closed form
  // stops spilling); the other FP32 variants keep vector loads, 0.7% faster
  // for them.
  using IncT = std::conditional_t<sizeof(T) == 8 || CUR == kCurLKClosed, __attribute__((address_space(4))) const T,
                                  const T>;
  IncT* inc1 = (IncT*)(a.inc + (size_t)t * 4 * nx);        // step t
  IncT* inc2 = (IncT*)(a.inc + (size_t)(t + 1) * 4 * nx);  // step t+1
  const T c2h = a.c2h, c2e0 = a.c2e0, inv = a.inv_d;
  const T tfEzB1 = inc1[a.bx1], tfEzB2 = inc1[a.bx2 - 1];
  const T tfEyB1 = inc1[nx + a.bx1], tfEyB2 = inc1[nx + a.bx2 - 1];
  const T tfHyB1 = inc1[2 * nx + a.bx1 - 1], tfHyB2 = inc1[2 * nx + a.bx2 - 1];
  const T tfHzB1 = inc1[3 * nx + a.bx1 - 1], tfHzB2 = inc1[3 * nx + a.bx2 - 1];
  const T ufEzB1 = inc2[a.bx1], ufEzB2 = inc2[a.bx2 - 1];
  const T ufEyB1 = inc2[nx + a.bx1], ufEyB2 = inc2[nx + a.bx2 - 1];
  const T ufHyB1 = inc2[2 * nx + a.bx1 - 1], ufHyB2 = inc2[2 * nx + a.bx2 - 1];
  const T ufHzB1 = inc2[3 * nx + a.bx1 - 1], ufHzB2 = inc2[3 * nx + a.bx2 - 1];
  const uint32_t sxb = (uint32_t)a.sx * W;
  const uint32_t ob = valid ? ((uint32_t)j * (uint32_t)a.sy + (uint32_t)k) * W : 0u;

  // batch of plane i+1 (issued at plane i): E^t(i+2), H^(t-1/2)(i+1),
  // J^(t-1)(i+1), media(i+1), incident terms of step t at i+1 and step t+1 at i
  T px = 0, py = 0, pz = 0;
  T nhx = 0, nhy = 0, nhz = 0;
  T nqx = 0, nqy = 0, nqz = 0, nqw = 0, nqn = 0;
  T nt0 = 0, nt1 = 0, nt2 = 0, nt3 = 0, mt0 = 0, mt1 = 0, mt2 = 0, mt3 = 0;
  const int i0 = ib - 2;
  {
    const uint32_t c = (uint32_t)i0 * sxb + ob;
    T ex = 0, ey = 0, ez = 0;
    if (valid) {
      ex = ldo(Exo, c); ey = ldo(Eyo, c); ez = ldo(Ezo, c);
      nhx = ldo(Hxo, c); nhy = ldo(Hyo, c); nhz = ldo(Hzo, c);
      px = ldo(Exo, c + sxb); py = ldo(Eyo, c + sxb); pz = ldo(Ezo, c + sxb);
    }
    s0[i0 & 1][0][r][lane] = ex; s0[i0 & 1][1][r][lane] = ey; s0[i0 & 1][2][r][lane] = ez;
    if (tfl) { nt0 = inc1[i0]; nt1 = inc1[nx + i0]; nt2 = inc1[2 * nx + i0]; nt3 = inc1[3 * nx + i0]; }
  }
  // carried along x: step-t H and E, step-t+1 E, J^t and media of plane i-1,
  // step-t+1 E x of plane i-2, step-t+1 H of plane i-2
  T h1px = 0, h1py = 0, h1pz = 0, e0xp = 0;
  T e1px = 0, e1py = 0, e1pz = 0, e1xpp = 0;
  T j1px = 0, j1py = 0, j1pz = 0, w1p = 0, n1p = 0;
  T h2py = 0, h2pz = 0;

  for (int i = i0; i <= ie; ++i) {
    const uint32_t c = (uint32_t)i * sxb + ob;
    const int b0 = i & 1, bn = b0 ^ 1;
    // ---- A: take the batch of plane i, stage E^t(i+1), issue the batch of plane i+1
    const T hx0 = nhx, hy0 = nhy, hz0 = nhz;
    const T qx = nqx, qy = nqy, qz = nqz, qw = nqw, qn = nqn;
    const T tf0 = nt0, tf1 = nt1, tf2 = nt2, tf3 = nt3;  // step t, plane i
    const T uf0 = mt0, uf1 = mt1, uf2 = mt2, uf3 = mt3;  // step t+1, plane i-1
    s0[bn][0][r][lane] = px; s0[bn][1][r][lane] = py; s0[bn][2][r][lane] = pz;
    if (i + 1 <= ie) {
      const uint32_t cn = c + sxb;
      if (valid) {
        px = ldo(Exo, cn + sxb); py = ldo(Eyo, cn + sxb); pz = ldo(Ezo, cn + sxb);
        nhx = ldo(Hxo, cn); nhy = ldo(Hyo, cn); nhz = ldo(Hzo, cn);
      }
      if (e1ok && cj) {
        nqx = ldo(Jxo, cn); nqy = ldo(Jyo, cn); nqz = ldo(Jzo, cn);
        nqw = ldo(a.wpe2, cn); nqn = ldo(a.nue, cn);
      }
      if (tfl) {
        nt0 = inc1[i + 1]; nt1 = inc1[nx + i + 1]; nt2 = inc1[2 * nx + i + 1]; nt3 = inc1[3 * nx + i + 1];
        mt0 = inc2[i]; mt1 = inc2[nx + i]; mt2 = inc2[2 * nx + i]; mt3 = inc2[3 * nx + i];
      }
    }
    const bool rim = own && i >= ib && i < ie &&
                     (rimjk || i < a.x2lo + 2 || i >= a.x2hi - 2 || pall || abs(i - pxa) <= 1 ||
                      abs(i - pxb) <= 1);
    lds_barrier();

    // ---- B1: H^(t+1/2) at plane i   (:797-806, :854-865, :913-924 and TF/SF)
    T h1x = hx0, h1y = hy0, h1z = hz0;
    const T ecx = s0[b0][0][r][lane], ecy = s0[b0][1][r][lane], ecz = s0[b0][2][r][lane];
    if (h1ok) {
      const T ezjp = s0[b0][2][r + 1][lane];
      const T eykp = s0[b0][1][r][lane + 1];
      const T exkp = s0[b0][0][r][lane + 1];
      const T exjp = s0[b0][0][r + 1][lane];
      const T eny = s0[bn][1][r][lane], enz = s0[bn][2][r][lane];
      {
        const T dEz = (ezjp - ecz) * inv;
        const T dEy = (eykp - ecy) * inv;
        h1x = h1x - c2h * (dEz - dEy);
        if (tfl && i >= a.bx1 && i < a.bx2) {
          h1x = fma(sHx0, tf0, h1x);
          h1x = fma(sHx1, tf1, h1x);
        }
      }
      {
        const T dEx = (exkp - ecx) * inv;
        const T dEz = (enz - ecz) * inv;
        h1y = h1y - c2h * (dEx - dEz);
        if (i == a.bx1 - 1) h1y = fma(-mJK, tfEzB1, h1y);
        if (i == a.bx2 - 1) h1y = fma(mJK, tfEzB2, h1y);
      }
      {
        const T dEy = (eny - ecy) * inv;
        const T dEx = (exjp - ecx) * inv;
        h1z = h1z - c2h * (dEy - dEx);
        if (i == a.bx1 - 1) h1z = fma(mJK, tfEyB1, h1z);
        if (i == a.bx2 - 1) h1z = fma(-mJK, tfEyB2, h1z);
      }
      if (rim) {
        sto(a.h[s1o][0], c, h1x);
        sto(a.h[s1o][1], c, h1y);
        sto(a.h[s1o][2], c, h1z);
      }
    }
    s1h[0][r][lane] = h1x;
    s1h[1][r][lane] = h1y;
    s1h[2][r][lane] = h1z;
    lds_barrier();

    // ---- C1: J^t (deferred) and E^(t+1) at plane i
    T e1x = 0, e1y = 0, e1z = 0, j1x = 0, j1y = 0, j1z = 0;
    if (e1ok && i >= ib - 1) {
      if (cj) {
        j1x = qx;
        j1y = qy;
        j1z = qz;
        if constexpr (CUR == kCurDI) {
          current_update<T, CUR>(a, qn, qw, j1x, j1y, j1z, ecx, ecy, ecz);
        } else {
          if (a.jcol) {
            current_update<T, CUR>(a, qn, qw, j1x, j1y, j1z, ecx, ecy, ecz);
          } else {
            const T mx = (ecx + e0xp) / 2;
            const T my = (ecy + s0[b0][1][r - 1][lane]) / 2;
            const T mz = (ecz + s0[b0][2][r][lane - 1]) / 2;
            current_update<T, CUR>(a, qn, qw, j1x, j1y, j1z, mx, my, mz);
          }
        }
      }
      T c1 = 1, c2 = c2e0;
      if (a.doplate == 2 && in_sphere(a, i, j, k)) { c1 = a.c1s; c2 = a.c2s; }
      {
        const T dHz = (h1z - s1h[2][r - 1][lane]) * inv;
        const T dHy = (h1y - s1h[1][r][lane - 1]) * inv;
        e1x = c1 * ecx + c2 * (dHz - dHy) - c2 * j1x;
        if (tfl && i >= a.bx1 && i < a.bx2 - 1) {
          e1x = fma(sEx2, tf2, e1x);
          e1x = fma(sEx3, tf3, e1x);
        }
      }
      {
        const T dHx = (h1x - s1h[0][r][lane - 1]) * inv;
        const T dHz = (h1z - h1pz) * inv;
        e1y = c1 * ecy + c2 * (dHx - dHz) - c2 * j1y;
        if (i == a.bx1) e1y = fma(mJK, tfHzB1, e1y);
        if (i == a.bx2 - 1) e1y = fma(-mJK, tfHzB2, e1y);
        if (a.doplate == 1 && i == a.ssx && j >= a.plate_lo && j < a.plate_up - 1 && k >= a.plate_lo && k < a.plate_up)
          e1y = 0;
      }
      {
        const T dHy = (h1y - h1py) * inv;
        const T dHx = (h1x - s1h[0][r - 1][lane]) * inv;
        e1z = c1 * ecz + c2 * (dHy - dHx) - c2 * j1z;
        if (i == a.bx1) e1z = fma(-mJK, tfHyB1, e1z);
        if (i == a.bx2 - 1) e1z = fma(mJK, tfHyB2, e1z);
        if (a.doplate == 1 && i == a.ssx && j >= a.plate_lo && j < a.plate_up && k >= a.plate_lo && k < a.plate_up - 1)
          e1z = 0;
      }
      if (rim) {
        sto(a.e[s1o][0], c, e1x);
        sto(a.e[s1o][1], c, e1y);
        sto(a.e[s1o][2], c, e1z);
      }
    }
    s1e[b0][0][r][lane] = e1x;
    s1e[b0][1][r][lane] = e1y;
    s1e[b0][2][r][lane] = e1z;
    lds_barrier();

    // ---- B2: H^(t+3/2) at plane i-1
    const int ip = i - 1;
    const uint32_t cp = c - sxb;
    T h2x = h1px, h2y = h1py, h2z = h1pz;
    if (h2ok && i >= ib) {
      const T ezjp = s1e[bn][2][r + 1][lane];
      const T eykp = s1e[bn][1][r][lane + 1];
      const T exkp = s1e[bn][0][r][lane + 1];
      const T exjp = s1e[bn][0][r + 1][lane];
      {
        const T dEz = (ezjp - e1pz) * inv;
        const T dEy = (eykp - e1py) * inv;
        h2x = h2x - c2h * (dEz - dEy);
        if (tfl && ip >= a.bx1 && ip < a.bx2) {
          h2x = fma(sHx0, uf0, h2x);
          h2x = fma(sHx1, uf1, h2x);
        }
      }
      {
        const T dEx = (exkp - e1px) * inv;
        const T dEz = (e1z - e1pz) * inv;
        h2y = h2y - c2h * (dEx - dEz);
        if (ip == a.bx1 - 1) h2y = fma(-mJK, ufEzB1, h2y);
        if (ip == a.bx2 - 1) h2y = fma(mJK, ufEzB2, h2y);
      }
      {
        const T dEy = (e1y - e1py) * inv;
        const T dEx = (exjp - e1px) * inv;
        h2z = h2z - c2h * (dEy - dEx);
        if (ip == a.bx1 - 1) h2z = fma(mJK, ufEyB1, h2z);
        if (ip == a.bx2 - 1) h2z = fma(-mJK, ufEyB2, h2z);
      }
      if (own && ip >= ib) {
        sto(a.h[s2o][0], cp, h2x);
        sto(a.h[s2o][1], cp, h2y);
        sto(a.h[s2o][2], cp, h2z);
      }
    }
    s2h[0][r][lane] = h2x;
    s2h[1][r][lane] = h2y;
    s2h[2][r][lane] = h2z;
    lds_barrier();

    // ---- C2: J^(t+1) and E^(t+2) at plane i-1 (owned cells)
    if (own && ip >= ib) {
      T j2x = 0, j2y = 0, j2z = 0;
      if (cj) {
        j2x = j1px;
        j2y = j1py;
        j2z = j1pz;
        if constexpr (CUR == kCurDI) {
          current_update<T, CUR>(a, n1p, w1p, j2x, j2y, j2z, e1px, e1py, e1pz);
        } else {
          if (a.jcol) {
            current_update<T, CUR>(a, n1p, w1p, j2x, j2y, j2z, e1px, e1py, e1pz);
          } else {
            const T mx = (e1px + e1xpp) / 2;
            const T my = (e1py + s1e[bn][1][r - 1][lane]) / 2;
            const T mz = (e1pz + s1e[bn][2][r][lane - 1]) / 2;
            current_update<T, CUR>(a, n1p, w1p, j2x, j2y, j2z, mx, my, mz);
          }
        }
        sto(a.j[s1o][0], cp, j2x);
        sto(a.j[s1o][1], cp, j2y);
        sto(a.j[s1o][2], cp, j2z);
      }
      T c1 = 1, c2 = c2e0;
      if (a.doplate == 2 && in_sphere(a, ip, j, k)) { c1 = a.c1s; c2 = a.c2s; }
      {
        const T dHz = (h2z - s2h[2][r - 1][lane]) * inv;
        const T dHy = (h2y - s2h[1][r][lane - 1]) * inv;
        T ex = c1 * e1px + c2 * (dHz - dHy) - c2 * j2x;
        if (tfl && ip >= a.bx1 && ip < a.bx2 - 1) {
          ex = fma(sEx2, uf2, ex);
          ex = fma(sEx3, uf3, ex);
        }
        sto(a.e[s2o][0], cp, ex);
      }
      {
        const T dHx = (h2x - s2h[0][r][lane - 1]) * inv;
        const T dHz = (h2z - h2pz) * inv;
        T ey = c1 * e1py + c2 * (dHx - dHz) - c2 * j2y;
        if (ip == a.bx1) ey = fma(mJK, ufHzB1, ey);
        if (ip == a.bx2 - 1) ey = fma(-mJK, ufHzB2, ey);
        if (a.doplate == 1 && ip == a.ssx && j >= a.plate_lo && j < a.plate_up - 1 && k >= a.plate_lo && k < a.plate_up)
          ey = 0;
        sto(a.e[s2o][1], cp, ey);
      }
      {
        const T dHy = (h2y - h2py) * inv;
        const T dHx = (h2x - s2h[0][r - 1][lane]) * inv;
        T ez = c1 * e1pz + c2 * (dHy - dHx) - c2 * j2z;
        if (ip == a.bx1) ez = fma(-mJK, ufHyB1, ez);
        if (ip == a.bx2 - 1) ez = fma(mJK, ufHyB2, ez);
        if (a.doplate == 1 && ip == a.ssx && j >= a.plate_lo && j < a.plate_up && k >= a.plate_lo && k < a.plate_up - 1)
          ez = 0;
        sto(a.e[s2o][2], cp, ez);
      }
    }
    // ---- carry along x
    e1xpp = e1px;
    e1px = e1x; e1py = e1y; e1pz = e1z;
    h1px = h1x; h1py = h1y; h1pz = h1z;
    e0xp = ecx;
    j1px = j1x; j1py = j1y; j1pz = j1z;
    w1p = qw; n1p = qn;
    h2py = h2y; h2pz = h2z;
  }
}

// fdtd_pair's memory traffic with no arithmetic: the same tiles in the same
// order, the same loads one plane ahead (E and H of the tile and its 2-cell
// halo; J and the media on tiles that carry J, halo included) and one store
// per owned value and array, of the value just read, to where it was read,
// so the solve's state is unchanged.  Timed beside fdtd_pair it measures
// what this access pattern reaches on the device: an upper bound for
// fdtd_pair, which moves the same bytes and also computes.
template <typename T, int PR, int PL>
__global__ void __launch_bounds__(PR * PL) pair_pattern(const StepArgs<T>* __restrict__ args,
                                                         const uint32_t* __restrict__ tiles) {
  constexpr uint32_t W = sizeof(T);
  const int lane = threadIdx.x & (PL - 1);
  const int r = threadIdx.x / PL;
  const uint32_t tile = tiles[xcd_remap(blockIdx.x, gridDim.x)];
  const ConstArgs<StepArgs<T>>& a = as_const_space(args);
  const int k0 = a.zstart2[tile & 0xff], k1 = a.zstart2[(tile & 0xff) + 1];
  const int j0 = a.ystart2[(tile >> 8) & 0xfff], j1 = a.ystart2[((tile >> 8) & 0xfff) + 1];
  const int ib = a.xstart2[(tile >> 20) & 0x7ff], ie = a.xstart2[((tile >> 20) & 0x7ff) + 1];
  const bool cj = (tile & kTileNoCurrent) == 0;
  const int j = j0 - 2 + r, k = k0 - 2 + lane;
  const bool valid = j >= 0 && j < a.ny && k >= 0 && k < a.nz;
  const bool own = valid && r >= 2 && r < PR - 2 && j < j1 && lane >= 2 && lane < PL - 2 && k < k1;
  const bool jl = cj && valid && r >= 1 && r < PR - 1 && lane >= 1 && lane < PL - 1;  // fdtd_pair's J lanes
  const int t = *a.tdev;
  const int s0i = t % 3, sj = (t + 2) % 3;
  T* const ex = a.e[s0i][0]; T* const ey = a.e[s0i][1]; T* const ez = a.e[s0i][2];
  T* const hx = a.h[s0i][0]; T* const hy = a.h[s0i][1]; T* const hz = a.h[s0i][2];
  T* const jx = a.j[sj][0]; T* const jy = a.j[sj][1]; T* const jz = a.j[sj][2];
  const uint32_t sxb = (uint32_t)a.sx * W;
  const uint32_t ob = valid ? ((uint32_t)j * (uint32_t)a.sy + (uint32_t)k) * W : 0u;
  T v[11] = {};
  auto load = [&](int i) {
    const uint32_t c = (uint32_t)i * sxb + ob;
    if (valid) { v[0] = ldo(ex, c); v[1] = ldo(ey, c); v[2] = ldo(ez, c); v[3] = ldo(hx, c); v[4] = ldo(hy, c); v[5] = ldo(hz, c); }
    if (jl) { v[6] = ldo(jx, c); v[7] = ldo(jy, c); v[8] = ldo(jz, c); v[9] = ldo(a.wpe2, c); v[10] = ldo(a.nue, c); }
  };
  load(ib - 2);
  for (int i = ib - 2; i <= ie; ++i) {
    T c[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) c[q] = v[q];
    if (i + 1 <= ie) load(i + 1);
    lds_barrier();  // paces the waves as fdtd_pair's barriers do
    if (own && i >= ib && i < ie) {
      const uint32_t cc = (uint32_t)i * sxb + ob;
      sto(ex, cc, c[0]); sto(ey, cc, c[1]); sto(ez, cc, c[2]);
      sto(hx, cc, c[3]); sto(hy, cc, c[4]); sto(hz, cc, c[5]);
      if (cj) { sto(jx, cc, c[6]); sto(jy, cc, c[7]); sto(jz, cc, c[8]); }
    }
  }
}

// J^(t+1) from J^t and E^(t+1) into `out` (what the reference holds after
// the same number of steps); used only to read J back.
template <typename T, int CUR>
__global__ void current_catchup(const StepArgs<T>* __restrict__ ap, T* ox, T* oy, T* oz) {
  const ConstArgs<StepArgs<T>>& a = as_const_space(ap);
  const long long n = (long long)a.nx * a.ny * a.nz;
  const int t = *a.tdev;
  const T* Ex = a.e[t % 3][0];
  const T* Ey = a.e[t % 3][1];
  const T* Ez = a.e[t % 3][2];
  const int sj = (t + 2) % 3;  // J^(t-1)
  for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < n; c += (long long)gridDim.x * blockDim.x) {
    const int k = (int)(c % a.nz);
    const int j = (int)((c / a.nz) % a.ny);
    const int i = (int)(c / a.sx);
    T jx = a.j[sj][0][c], jy = a.j[sj][1][c], jz = a.j[sj][2][c];
    if (i >= 1 && i < a.nx - 1 && j >= 1 && j < a.ny - 1 && k >= 1 && k < a.nz - 1) {
      if (CUR == kCurDI || a.jcol) {
        current_update<T, CUR>(a, a.nue[c], a.wpe2[c], jx, jy, jz, Ex[c], Ey[c], Ez[c]);
      } else {
        current_update<T, CUR>(a, a.nue[c], a.wpe2[c], jx, jy, jz, (Ex[c] + Ex[c - a.sx]) / 2,
                               (Ey[c] + Ey[c - a.sy]) / 2, (Ez[c] + Ez[c - 1]) / 2);
      }
    }
    ox[c] = jx;
    oy[c] = jy;
    oz[c] = jz;
  }
}

// One field component of cells [c0, c0 + nc) (flat [nx][ny][nz] index),
// as FP64, into dst[0, nc): E or H of the current set, or J caught up as
// current_catchup does (all three components are updated, one is kept, so
// the arithmetic is current_catchup's).  Reads the step from the device
// counter, so it follows the solver's stream.
template <typename T, int CUR>
__global__ void __launch_bounds__(256) field_slab(const StepArgs<T>* __restrict__ ap, int comp, long long c0,
                                                  long long nc, double* __restrict__ dst) {
  const ConstArgs<StepArgs<T>>& a = as_const_space(ap);
  const int t = *a.tdev;
  const T* Ex = a.e[t % 3][0];
  const T* Ey = a.e[t % 3][1];
  const T* Ez = a.e[t % 3][2];
  const int sj = (t + 2) % 3;  // J^(t-1)
  for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < nc; q += (long long)gridDim.x * blockDim.x) {
    const long long c = c0 + q;
    T v;
    if (comp < 3) {
      v = a.e[t % 3][comp][c];
    } else if (comp < 6) {
      v = a.h[t % 3][comp - 3][c];
    } else {
      const int k = (int)(c % a.nz);
      const int j = (int)((c / a.nz) % a.ny);
      const int i = (int)(c / a.sx);
      T jx = a.j[sj][0][c], jy = a.j[sj][1][c], jz = a.j[sj][2][c];
      if (i >= 1 && i < a.nx - 1 && j >= 1 && j < a.ny - 1 && k >= 1 && k < a.nz - 1) {
        if (CUR == kCurDI || a.jcol) {
          current_update<T, CUR>(a, a.nue[c], a.wpe2[c], jx, jy, jz, Ex[c], Ey[c], Ez[c]);
        } else {
          current_update<T, CUR>(a, a.nue[c], a.wpe2[c], jx, jy, jz, (Ex[c] + Ex[c - a.sx]) / 2,
                                 (Ey[c] + Ey[c - a.sy]) / 2, (Ez[c] + Ez[c - 1]) / 2);
        }
      }
      v = comp == 6 ? jx : comp == 7 ? jy : jz;
    }
    dst[q] = (double)v;
  }
}

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// The same contraction on the matrix cores, for 2M >= 16 frequency rows:
// per sample array q it is the product acc[q] (2M x n) += Phi (2M x kn) S
// (kn x n), with S[k][p] the buffered samples and Phi[f][k] =
// sincos[t0+k][f] (f = 2m + c).  One wavefront per 16 points of one array
// runs FP64 16x16x4 MFMA tiles over the 2M rows, up to 8 tiles (128 rows)
// per pass, so each sample is read once per 128 rows instead of once per 8.
// The slots go in blocks of 16: the workgroup stages a block of Phi (16
// slots x 128 rows, shared by its four wavefronts) in LDS while it runs the
// previous block's MFMAs, and each wavefront loads the next block's samples
// meanwhile, so the loads overlap the matrix work (one barrier per block).
// Loads use clamped indices and are masked where used, so nothing waits on
// a load before its value is needed.  Lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16] and holds D[4 v + l / 16][l % 16] in element v (checked
// on gfx950).  FP64 products and sums, in a different order than
// ntff_contract.
constexpr int kPhiBlock = 16;  // slots per staged block of Phi

// Neumaier's compensated sum of (s, c) and x: the new pair.
__device__ __forceinline__ float2 compensated_sum(float2 e, float x) {
  const float t = e.x + x;
  const float c = e.y + (fabsf(e.x) >= fabsf(x) ? (e.x - t) + x : (x - t) + e.x);
  return make_float2(t, c);
}

// One pass of the matrix-core contraction over rows [f0, f0 + nf), nf <=
// 16 NT.  C32 = false: FP64 Phi, samples and 16x16x4 FP64 tiles, added to
// the FP64 sums (lane l holds D[4 v + l / 16][l % 16] in element v).  C32 =
// true: FP32 Phi, samples and 16x16x4 FP32 tiles summed over the slots in
// FP32, then added to the compensated (sum, compensation) pairs (lane l
// holds D[4 (l / 16) + v][l % 16] in element v).  Both layouts were checked
// on gfx950.
template <int NT, bool C32, typename T>
__device__ __forceinline__ void mfma_pass(std::conditional_t<C32, float, double> (*stab)[kPhiBlock][128],
                                          const std::conditional_t<C32, float, double>* __restrict__ tab,
                                          const T* __restrict__ sb, double* acc, long long U, long long np, int f0,
                                          int nf, int M2, int kn, int t0, bool wv, bool pv, int col, int kr) {
  using P = std::conditional_t<C32, float, double>;
  using V = std::conditional_t<C32, v4f, v4d>;
  constexpr int KB = kPhiBlock, PER = KB * 128 / 256;
  const int nblk = (kn + KB - 1) / KB;
  P rphi[PER];
  T bc[KB / 4], bn[KB / 4];
  // this thread's share of a Phi block: elements e = threadIdx.x + 256 m of
  // the 16 x 128 block (slot e / 128, row e % 128), clamped into the table
  auto load_phi = [&](int k0) {
#pragma unroll
    for (int m = 0; m < PER; ++m) {
      const int e = threadIdx.x + 256 * m;
      const int k = min(k0 + (e >> 7), kn - 1), fe = min(e & 127, nf - 1);
      rphi[m] = ldo(tab, (uint32_t)((t0 + k) * M2 + f0 + fe) * (uint32_t)sizeof(P));
    }
  };
  auto store_phi = [&](int buf, int k0) {
#pragma unroll
    for (int m = 0; m < PER; ++m) {
      const int e = threadIdx.x + 256 * m;
      stab[buf][e >> 7][e & 127] = (k0 + (e >> 7) < kn && (e & 127) < nf) ? rphi[m] : (P)0;
    }
  };
  // this lane's samples of a block: slots k0 + 4 s + kr (clamped; masked at use)
  auto load_s = [&](int k0, T* bs) {
#pragma unroll
    for (int s4 = 0; s4 < KB / 4; ++s4)
      bs[s4] = ldo(sb, (uint32_t)(min(k0 + 4 * s4 + kr, kn - 1) * U) * (uint32_t)sizeof(T));
  };
  load_phi(0);
  load_s(0, bc);
  store_phi(0, 0);
  lds_barrier();
  V d[NT];
#pragma unroll
  for (int tt = 0; tt < NT; ++tt) d[tt] = V{0, 0, 0, 0};
  for (int bk = 0; bk < nblk; ++bk) {
    const int buf = bk & 1, k0 = bk * KB;
    const bool more = bk + 1 < nblk;  // uniform
    if (more) {
      load_phi(k0 + KB);
      load_s(k0 + KB, bn);
    }
    if (wv) {
#pragma unroll
      for (int s4 = 0; s4 < KB / 4; ++s4) {
        const P b = (pv && k0 + 4 * s4 + kr < kn) ? (P)bc[s4] : (P)0;
#pragma unroll
        for (int tt = 0; tt < NT; ++tt) {
          if constexpr (C32)
            d[tt] = __builtin_amdgcn_mfma_f32_16x16x4f32(stab[buf][4 * s4 + kr][tt * 16 + col], b, d[tt], 0, 0, 0);
          else
            d[tt] = __builtin_amdgcn_mfma_f64_16x16x4f64(stab[buf][4 * s4 + kr][tt * 16 + col], b, d[tt], 0, 0, 0);
        }
      }
    }
    if (more) {
      store_phi(buf ^ 1, k0 + KB);
#pragma unroll
      for (int s4 = 0; s4 < KB / 4; ++s4) bc[s4] = bn[s4];
    }
    lds_barrier();
  }
  // the running sums, four tiles at a time: their rows read first, then
  // written (as one read-modify-write per row, each load would wait for the
  // store before it)
  if (pv) {
    constexpr int G = NT < 4 ? NT : 4;
    auto row = [&](int tt, int v) { return C32 ? tt * 16 + 4 * kr + v : tt * 16 + 4 * v + kr; };
#pragma unroll
    for (int g0 = 0; g0 < NT; g0 += G) {
      double old[G][4];
#pragma unroll
      for (int tt = 0; tt < G; ++tt)
#pragma unroll
        for (int v = 0; v < 4; ++v) old[tt][v] = acc[(size_t)(f0 + min(row(g0 + tt, v), nf - 1)) * np];
#pragma unroll
      for (int tt = 0; tt < G; ++tt)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int f = row(g0 + tt, v);
          if (f >= nf) continue;
          if constexpr (C32) {
            const float2 e = compensated_sum(__builtin_bit_cast(float2, old[tt][v]), d[g0 + tt][v]);
            acc[(size_t)(f0 + f) * np] = __builtin_bit_cast(double, e);
          } else {
            acc[(size_t)(f0 + f) * np] = old[tt][v] + d[g0 + tt][v];
          }
        }
    }
  }
}

template <typename T>
__global__ void __launch_bounds__(256) ntff_contract_mfma(const NtffArgs<T>* __restrict__ args, int flush) {
  __shared__ double stab[2][kPhiBlock][128];  // Phi blocks, double-buffered
  const ConstArgs<NtffArgs<T>>& a = as_const_space(args);
  const int t = *a.tdev - 1;
  if (t < 0 || (!flush && (t + 1) % a.K != 0 && t + 1 != a.T_steps)) return;  // uniform
  const long long w = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / kLanes;
  // wavefronts past the last one stage Phi and meet the barriers, nothing else
  const bool wv = w < a.wofs[24];  // uniform per wavefront
  int q = 0;
  if (wv)
    while (q < 23 && w >= a.wofs[q + 1]) ++q;
  const int lane = threadIdx.x & (kLanes - 1);
  const int col = lane & 15, kr = lane >> 4;
  const long long np = a.uoff[q + 1] - a.uoff[q];
  const long long p = wv ? (w - a.wofs[q]) * 16 + col : 0;
  const bool pv = wv && p < np;
  const long long U = a.uoff[24];
  const int kn = t % a.K + 1, t0 = t + 1 - kn, M2 = 2 * a.M;
  const T* __restrict__ sb = a.sbuf + a.uoff[q] + (pv ? p : 0);
  double* acc = a.acc + a.off[q] + (pv ? p : 0);
  for (int f0 = 0; f0 < M2; f0 += 128) {
    const int nf = min(128, M2 - f0);  // uniform
    if (nf > 64) mfma_pass<8, false>(stab, a.sincos, sb, acc, U, np, f0, nf, M2, kn, t0, wv, pv, col, kr);
    else if (nf > 32) mfma_pass<4, false>(stab, a.sincos, sb, acc, U, np, f0, nf, M2, kn, t0, wv, pv, col, kr);
    else if (nf > 16) mfma_pass<2, false>(stab, a.sincos, sb, acc, U, np, f0, nf, M2, kn, t0, wv, pv, col, kr);
    else mfma_pass<1, false>(stab, a.sincos, sb, acc, U, np, f0, nf, M2, kn, t0, wv, pv, col, kr);
  }
  // a flush consumes the block's samples so far: zeroed, they add nothing
  // when the block's own contraction sums all of its slots
  if (flush && pv) {
    T* __restrict__ sw = a.sbuf + a.uoff[q] + p;
    for (int k = kr; k < kn; k += 4) sw[(size_t)k * U] = (T)0;
  }
}


// The contraction with FP32 inputs and FP32 accumulation (compensated
// FP32 sums): FP32 samples and the FP32 phase table on FP32 16x16x4 matrix
// tiles, summed over the block in FP32; each block sum is added to the
// running sum of its entry with Neumaier's compensation, the pair (sum,
// compensation) held as two floats in the entry's 8 bytes.  The structure
// is ntff_contract_mfma's (mfma_pass).  Every 2M (from 2 rows up) takes this
// path in the compensated mode.
template <typename T>
__global__ void __launch_bounds__(256) ntff_contract_mfma32(const NtffArgs<T>* __restrict__ args, int flush) {
  __shared__ float stab[2][kPhiBlock][128];  // Phi blocks, double-buffered
  const ConstArgs<NtffArgs<T>>& a = as_const_space(args);
  const int t = *a.tdev - 1;
  if (t < 0 || (!flush && (t + 1) % a.K != 0 && t + 1 != a.T_steps)) return;  // uniform
  const long long w = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / kLanes;
  // wavefronts past the last one stage Phi and meet the barriers, nothing else
  const bool wv = w < a.wofs[24];  // uniform per wavefront
  int q = 0;
  if (wv)
    while (q < 23 && w >= a.wofs[q + 1]) ++q;
  const int lane = threadIdx.x & (kLanes - 1);
  const int col = lane & 15, kr = lane >> 4;
  const long long np = a.uoff[q + 1] - a.uoff[q];
  const long long p = wv ? (w - a.wofs[q]) * 16 + col : 0;
  const bool pv = wv && p < np;
  const long long U = a.uoff[24];
  const int kn = t % a.K + 1, t0 = t + 1 - kn, M2 = 2 * a.M;
  const T* __restrict__ sb = a.sbuf + a.uoff[q] + (pv ? p : 0);
  double* acc = a.acc + a.off[q] + (pv ? p : 0);
  for (int f0 = 0; f0 < M2; f0 += 128) {
    const int nf = min(128, M2 - f0);  // uniform
    if (nf > 64) mfma_pass<8, true>(stab, a.sincos32, sb, acc, U, np, f0, nf, M2, kn, t0, wv, pv, col, kr);
    else if (nf > 32) mfma_pass<4, true>(stab, a.sincos32, sb, acc, U, np, f0, nf, M2, kn, t0, wv, pv, col, kr);
    else if (nf > 16) mfma_pass<2, true>(stab, a.sincos32, sb, acc, U, np, f0, nf, M2, kn, t0, wv, pv, col, kr);
    else mfma_pass<1, true>(stab, a.sincos32, sb, acc, U, np, f0, nf, M2, kn, t0, wv, pv, col, kr);
  }
  if (flush && pv) {
    T* __restrict__ sw = a.sbuf + a.uoff[q] + p;
    for (int k = kr; k < kn; k += 4) sw[(size_t)k * U] = (T)0;
  }
}

// Surface samples of the running DFT (meteor-fork.cpp:971-1023 for the
// face-averaged H, :1417-1470 for E) into slot t % K of the sample buffer.
template <typename T>
__global__ void __launch_bounds__(256) ntff_sample(const NtffArgs<T>* __restrict__ args,
                                                        const uint2* __restrict__ blocks) {
  // blocks[b] = {solve id, face << 29 | block index within the face}
  const uint2 d = blocks[blockIdx.x];
  const ConstArgs<NtffArgs<T>>& a = as_const_space(args + d.x);
  const int face = (int)(d.y >> 29);
  const int nx = a.nx, ny = a.ny, nz = a.nz, b = a.b;
  int n2;
  long long npts;
  if (face < 2) { n2 = nz; npts = (long long)ny * nz; }
  else if (face < 4) { n2 = nz; npts = (long long)nx * nz; }
  else { n2 = ny; npts = (long long)nx * ny; }
  const long long pt = (long long)(d.y & 0x1fffffffu) * blockDim.x + threadIdx.x;
  if (pt >= npts) return;
  const int u = (int)(pt / n2), v = (int)(pt % n2);
  const int t = *a.tdev;
  const int par = (t + 1) % 3;  // the field set this step wrote
  const T* Ex = a.e[par][0];
  const T* Ey = a.e[par][1];
  const T* Ez = a.e[par][2];
  const T* Hx = a.h[par][0];
  const T* Hy = a.h[par][1];
  const T* Hz = a.h[par][2];
  double q0, q1, q2, q3;
  long long c;
  switch (face) {
    case 0:  // front, x = b: My=-Ez, Mz=Ey, Jy=+avg Hz, Jz=-avg Hy
      c = (long long)b * a.sx + (long long)u * a.sy + v;
      q0 = -(double)Ez[c]; q1 = Ey[c];
      q2 = 0.5 * ((double)Hz[c] + (double)Hz[c - a.sx]);
      q3 = -0.5 * ((double)Hy[c] + (double)Hy[c - a.sx]);
      break;
    case 1:  // back, x = nx-b
      c = (long long)(nx - b) * a.sx + (long long)u * a.sy + v;
      q0 = Ez[c]; q1 = -(double)Ey[c];
      q2 = -0.5 * ((double)Hz[c] + (double)Hz[c - a.sx]);
      q3 = 0.5 * ((double)Hy[c] + (double)Hy[c - a.sx]);
      break;
    case 2:  // top, y = ny-b: Mx=-Ez, Mz=Ex, Jx=+avg Hz, Jz=-avg Hx
      c = (long long)u * a.sx + (long long)(ny - b) * a.sy + v;
      q0 = -(double)Ez[c]; q1 = Ex[c];
      q2 = 0.5 * ((double)Hz[c] + (double)Hz[c - a.sy]);
      q3 = -0.5 * ((double)Hx[c] + (double)Hx[c - a.sy]);
      break;
    case 3:  // bottom, y = b
      c = (long long)u * a.sx + (long long)b * a.sy + v;
      q0 = Ez[c]; q1 = -(double)Ex[c];
      q2 = -0.5 * ((double)Hz[c] + (double)Hz[c - a.sy]);
      q3 = 0.5 * ((double)Hx[c] + (double)Hx[c - a.sy]);
      break;
    case 4:  // left, z = b: Mx=-Ey, My=Ex, Jx=+avg Hy, Jy=-avg Hx
      c = (long long)u * a.sx + (long long)v * a.sy + b;
      q0 = -(double)Ey[c]; q1 = Ex[c];
      q2 = 0.5 * ((double)Hy[c] + (double)Hy[c - 1]);
      q3 = -0.5 * ((double)Hx[c] + (double)Hx[c - 1]);
      break;
    default:  // right, z = nz-b
      c = (long long)u * a.sx + (long long)v * a.sy + (nz - b);
      q0 = Ey[c]; q1 = -(double)Ex[c];
      q2 = -0.5 * ((double)Hy[c] + (double)Hy[c - 1]);
      q3 = 0.5 * ((double)Hx[c] + (double)Hx[c - 1]);
      break;
  }
  // watchdog: a non-finite field anywhere reaches the NTFF surface within
  // the steps the stencil needs to carry it there
  if (!(isfinite(q0) && isfinite(q1) && isfinite(q2) && isfinite(q3))) atomicOr(a.health, 1);
  T* slot = a.sbuf + (size_t)(t % a.K) * (size_t)a.uoff[24];
  slot[a.uoff[4 * face + 0] + pt] = (T)q0;
  slot[a.uoff[4 * face + 1] + pt] = (T)q1;
  slot[a.uoff[4 * face + 2] + pt] = (T)q2;
  slot[a.uoff[4 * face + 3] + pt] = (T)q3;
}

// FP64 read through the constant address space (uniform table entries go
// through scalar loads).
__device__ __forceinline__ double ldc(const double* p, size_t i) {
  return ((const __attribute__((address_space(4))) double*)p)[i];
}

// Contraction of the buffered samples at the end of a K-step block (the
// DFT of meteor-fork.cpp:977-1022, 1417-1470 summed over the block, phase
// t*dt for both E and H, as the reference's integer (t+1/2) gives): one
// thread per sample point, frequencies in groups of 8, FP64 products and
// sums.  Runs after the step counter moved past step t (so it can follow a
// captured step sequence); returns at once unless step t closes a block or
// `flush` asks for the block so far (its slots are then zeroed).
template <typename T>
__global__ void __launch_bounds__(256) ntff_contract(const NtffArgs<T>* __restrict__ args, int flush) {
  const ConstArgs<NtffArgs<T>>& a = as_const_space(args);
  const int t = *a.tdev - 1;
  if (t < 0 || (!flush && (t + 1) % a.K != 0 && t + 1 != a.T_steps)) return;
  const long long U = a.uoff[24];
  const long long u = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= U) return;
  int q = 0;
  while (q < 23 && u >= a.uoff[q + 1]) ++q;
  const long long np = a.uoff[q + 1] - a.uoff[q], pt = u - a.uoff[q];
  const int kn = t % a.K + 1, t0 = t + 1 - kn, M = a.M;
  const T* __restrict__ sb = a.sbuf;
  double* acc = a.acc + a.off[q] + pt;
  for (int m0 = 0; m0 < M; m0 += 8) {
    double ss[8], cc[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) ss[j] = cc[j] = 0;
    for (int k = 0; k < kn; ++k) {
      const double x = (double)sb[(size_t)k * U + u];
      const size_t row = (size_t)(t0 + k) * M * 2;
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        if (m0 + j < M) {
          ss[j] = fma(x, ldc(a.sincos, row + 2 * (m0 + j)), ss[j]);
          cc[j] = fma(x, ldc(a.sincos, row + 2 * (m0 + j) + 1), cc[j]);
        }
      }
    }
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      if (m0 + j < M) {
        acc[(size_t)(2 * (m0 + j)) * np] += ss[j];
        acc[(size_t)(2 * (m0 + j) + 1) * np] += cc[j];
      }
    }
  }
  if (flush) {
    T* __restrict__ sw = a.sbuf;
    for (int k = 0; k < kn; ++k) sw[(size_t)k * U + u] = (T)0;
  }
}

// Output cadence (t % cadence == 0): H slices (meteor-fork.cpp:1025-1045),
// staged E slices, and max|Ex| (:1621-1630) into the frame store.
template <typename T>
__global__ void __launch_bounds__(256) cadence_frames(const AuxArgs<T>* __restrict__ ap) {
  const ConstArgs<AuxArgs<T>>& a = as_const_space(ap);
  const int t = *a.tdev;
  if (t % a.cadence) return;
  const int f = t / a.cadence;
  if (f >= a.nframes) return;
  const int par = (t + 1) % 3;  // the field set this step wrote
  const int nx = a.nx, ny = a.ny, nz = a.nz;
  const long long gsz = (long long)gridDim.x * blockDim.x;
  const long long tid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  // max|Ex|
  {
    const T* Ex = a.e[par][0];
    const long long n = (long long)nx * ny * nz;
    double m = 0;
    for (long long c = tid; c < n; c += gsz) m = fmax(m, fabs((double)Ex[c]));
    for (int s = 32; s > 0; s >>= 1) m = fmax(m, __shfl_xor(m, s));
    if ((threadIdx.x & 63) == 0) atomicMax(a.emax + f, (unsigned long long)__double_as_longlong(m));
  }
  if (!a.hframes) return;
  const long long fsz = 2LL * ((long long)nx * ny + (long long)nx * nz + (long long)ny * nz);
  double* hf = a.hframes + (size_t)f * fsz;
  double* ef = a.eframes + (size_t)f * fsz;
  const T* Hy = a.h[par][1];
  const T* Hz = a.h[par][2];
  for (long long q = tid; q < fsz; q += gsz) {
    ef[q] = (double)a.eslice[q];
    long long rem = q;
    double v = 0;
    const long long zxy = (long long)nx * ny, zxz = (long long)nx * nz, zyz = (long long)ny * nz;
    if (rem < 2 * zxy) {
      const int comp = (int)(rem / zxy);
      const long long e = rem % zxy;
      const int i = (int)(e / ny), jj = (int)(e % ny);
      if (jj < ny - 1) {
        const long long c = (long long)i * a.sx + (long long)jj * a.sy + a.ssz;
        v = comp ? (double)Hz[c] : (double)Hy[c];
      }
    } else if ((rem -= 2 * zxy) < 2 * zxz) {
      const int comp = (int)(rem / zxz);
      const long long e = rem % zxz;
      const int i = (int)(e / nz), kk = (int)(e % nz);
      if (kk < nz - 1) {
        const long long c = (long long)i * a.sx + (long long)a.ssy * a.sy + kk;
        v = comp ? (double)Hz[c] : (double)Hy[c];
      }
    } else {
      rem -= 2 * zxz;
      const int comp = (int)(rem / zyz);
      const long long e = rem % zyz;
      const int jj = (int)(e / nz), kk = (int)(e % nz);
      if (jj < ny - 1 && kk < nz - 1) {
        const long long c = (long long)a.ssx * a.sx + (long long)jj * a.sy + kk;
        v = comp ? (double)Hz[c] : (double)Hy[c];
      }
    }
    hf[q] = v;
  }
}

// Probes (meteor-fork.cpp:1474-1486) and the step counter; one workgroup per solve.
template <typename T>
__global__ void __launch_bounds__(64) probes_and_tick(const AuxArgs<T>* __restrict__ args) {
  const ConstArgs<AuxArgs<T>>& a = as_const_space(args + blockIdx.x);
  const int t = *a.tdev;
  const int par = (t + 1) % 3;  // the field set this step wrote
  const T* Ex = a.e[par][0];
  const T* Ey = a.e[par][1];
  const T* Ez = a.e[par][2];
  const T* Hx = a.h[par][0];
  const T* Hy = a.h[par][1];
  const T* Hz = a.h[par][2];
  const long long sx = a.sx, sy = a.sy;
  for (int n = threadIdx.x; n < a.nprobes; n += blockDim.x) {
    const long long c = (long long)a.px[n] * sx + (long long)a.py[n] * sy + a.pz[n];
    const size_t e = (size_t)t * a.nprobes + n, T6 = (size_t)a.tsteps * a.nprobes;
    a.probes[0 * T6 + e] = 0.5 * ((double)Ex[c] + (double)Ex[c - sx]);
    a.probes[1 * T6 + e] = 0.5 * ((double)Ey[c] + (double)Ey[c - sy]);
    a.probes[2 * T6 + e] = 0.5 * ((double)Ez[c] + (double)Ez[c - 1]);
    a.probes[3 * T6 + e] = 0.25 * ((double)Hx[c] + (double)Hx[c - sy] + (double)Hx[c - 1] + (double)Hx[c - sy - 1]);
    a.probes[4 * T6 + e] = 0.25 * ((double)Hy[c] + (double)Hy[c - sx] + (double)Hy[c - 1] + (double)Hy[c - sx - 1]);
    a.probes[5 * T6 + e] = 0.25 * ((double)Hz[c] + (double)Hz[c - sy] + (double)Hz[c - sx] + (double)Hz[c - sx - sy]);
  }
  __syncthreads();
  if (threadIdx.x == 0) *a.tdev = t + 1;
}

// Plain streaming copy (16 B per lane per access): the bandwidth reference
// for the roofline and the known-size calibration of the HBM byte counters.
__global__ void __launch_bounds__(256) stream_copy(const float4* __restrict__ src, float4* __restrict__ dst, size_t n) {
  for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x)
    dst[q] = src[q];
}

hipError_t launch_stream_copy(const void* src, void* dst, size_t bytes, int blocks, hipStream_t s) {
  const size_t n = bytes / 16;
  hipLaunchKernelGGL(stream_copy, dim3(blocks), dim3(256), 0, s, (const float4*)src, (float4*)dst, n);
  return hipGetLastError();
}

// ---------------------------------------------------------------- launchers

template <typename T, int TY, int CUR, int PM, bool SLICE, int ZL>
static hipError_t launch_v(const StepArgs<T>* a, const uint16_t* sids, const uint32_t* tiles, int n, hipStream_t s) {
  hipLaunchKernelGGL((fdtd_step<T, TY, CUR, PM, SLICE, ZL>), dim3(n), dim3((TY + 1) * ZL), 0, s, a, tiles, sids);
  return hipGetLastError();
}

// 32 z lanes with or without slice capture; 64 z lanes without it
template <typename T, int TY, int CUR, int PM>
static hipError_t launch_s(const StepArgs<T>* a, const uint16_t* sids, bool slice, int zl, const uint32_t* tiles,
                           int n, hipStream_t s) {
  if (zl == 64) return slice ? hipErrorInvalidValue : launch_v<T, TY, CUR, PM, false, 64>(a, sids, tiles, n, s);
  if (zl != 32) return hipErrorInvalidValue;
  return slice ? launch_v<T, TY, CUR, PM, true, 32>(a, sids, tiles, n, s)
               : launch_v<T, TY, CUR, PM, false, 32>(a, sids, tiles, n, s);
}

template <typename T, int TY, int CUR>
static hipError_t launch_ps(const StepArgs<T>* a, const uint16_t* sids, int pm, bool slice, int zl,
                            const uint32_t* tiles, int n, hipStream_t s) {
  switch (pm) {
    case 0: return launch_s<T, TY, CUR, 0>(a, sids, slice, zl, tiles, n, s);
    case kPmlX: return launch_s<T, TY, CUR, kPmlX>(a, sids, slice, zl, tiles, n, s);
    case kPmlY: return launch_s<T, TY, CUR, kPmlY>(a, sids, slice, zl, tiles, n, s);
    case kPmlZ: return launch_s<T, TY, CUR, kPmlZ>(a, sids, slice, zl, tiles, n, s);
    case kPmlAll: return launch_s<T, TY, CUR, kPmlAll>(a, sids, slice, zl, tiles, n, s);
  }
  return hipErrorInvalidValue;
}

template <typename T, int TY>
static hipError_t launch_c(const StepArgs<T>* a, const uint16_t* sids, int cur, int pm, bool slice, int zl,
                           const uint32_t* tiles, int n, hipStream_t s) {
  switch (cur) {
    case kCurDI: return launch_ps<T, TY, kCurDI>(a, sids, pm, slice, zl, tiles, n, s);
    case kCurLKPoly: return launch_ps<T, TY, kCurLKPoly>(a, sids, pm, slice, zl, tiles, n, s);
    case kCurLKClosed: return launch_ps<T, TY, kCurLKClosed>(a, sids, pm, slice, zl, tiles, n, s);
  }
  return hipErrorInvalidValue;
}

// One launch of the fused step over `n` tiles of the list `tiles`, tile b
// belonging to solve sids[b] of the argument array `a` (sids == nullptr:
// every tile is solve 0); `pm` is the set of CPML slab axes the variant
// carries code for (kPmlX | kPmlY | kPmlZ, 0 = interior), `slice` selects
// the variant capturing E slices, `zl` the z lanes of a workgroup row (32
// or 64; tiles own zl - 1 z cells).
template <typename T>
hipError_t launch_step(const StepArgs<T>* a, const uint16_t* sids, int ty, int cur, int pm, bool slice, int zl,
                       const uint32_t* tiles, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  switch (ty) {
    case 3: return launch_c<T, 3>(a, sids, cur, pm, slice, zl, tiles, n, s);
    case 7: return launch_c<T, 7>(a, sids, cur, pm, slice, zl, tiles, n, s);
    case 15: return launch_c<T, 15>(a, sids, cur, pm, slice, zl, tiles, n, s);
  }
  return hipErrorInvalidValue;
}

template <typename T, int TY, int CUR>
static hipError_t launch_fc(const StepArgs<T>* a, const uint16_t* sids, bool slice, int zl, const uint32_t* tiles,
                            const ListSpans& sp, int n, hipStream_t s) {
  if (zl == 64) {
    if (slice) return hipErrorInvalidValue;
    hipLaunchKernelGGL((fdtd_steps<T, TY, CUR, false, 64>), dim3(n), dim3((TY + 1) * 64), 0, s, a, tiles, sids, sp);
  } else if (zl == 32) {
    if (slice)
      hipLaunchKernelGGL((fdtd_steps<T, TY, CUR, true, 32>), dim3(n), dim3((TY + 1) * 32), 0, s, a, tiles, sids, sp);
    else
      hipLaunchKernelGGL((fdtd_steps<T, TY, CUR, false, 32>), dim3(n), dim3((TY + 1) * 32), 0, s, a, tiles, sids, sp);
  } else {
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

template <typename T, int TY>
static hipError_t launch_fty(const StepArgs<T>* a, const uint16_t* sids, int cur, bool slice, int zl,
                             const uint32_t* tiles, const ListSpans& sp, int n, hipStream_t s) {
  switch (cur) {
    case kCurDI: return launch_fc<T, TY, kCurDI>(a, sids, slice, zl, tiles, sp, n, s);
    case kCurLKPoly: return launch_fc<T, TY, kCurLKPoly>(a, sids, slice, zl, tiles, sp, n, s);
    case kCurLKClosed: return launch_fc<T, TY, kCurLKClosed>(a, sids, slice, zl, tiles, sp, n, s);
  }
  return hipErrorInvalidValue;
}

// The kNumLists one-step lists in one launch (fdtd_steps): list l is
// tiles[first[l], first[l] + count[l]) (solve ids sids[...] alike, or every
// tile solve 0 when sids is null); every list has tiles of `ty` rows.
template <typename T>
hipError_t launch_steps(const StepArgs<T>* a, const uint16_t* sids, int ty, int cur, bool slice, int zl,
                        const uint32_t* tiles, const int* first, const int* count, hipStream_t s) {
  ListSpans sp;
  int n = 0;
  for (int l = 0; l < kNumLists; l++) {
    if (count[l] < 0) return hipErrorInvalidValue;
    sp.first[l] = first[l];
    n += count[l];
    sp.end[l] = n;
  }
  if (n <= 0) return hipSuccess;
  switch (ty) {
    case 3: return launch_fty<T, 3>(a, sids, cur, slice, zl, tiles, sp, n, s);
    case 7: return launch_fty<T, 7>(a, sids, cur, slice, zl, tiles, sp, n, s);
    case 15: return launch_fty<T, 15>(a, sids, cur, slice, zl, tiles, sp, n, s);
  }
  return hipErrorInvalidValue;
}

// Two steps of the tiles of the two-step region (fdtd_pair): workgroups of
// kPairRows rows of kPairLanes<T> z lanes (64 in FP32; 32 in FP64, whose
// 512-thread workgroups leave room for its registers).
template <typename T>
hipError_t launch_pair(const StepArgs<T>* a, const uint16_t* sids, int cur, const uint32_t* tiles, int n,
                       hipStream_t s) {
  if (n <= 0) return hipSuccess;
  constexpr int PL = kPairLanes<T>;
  const dim3 g(n), b(kPairRows * PL);
  switch (cur) {
    case kCurDI: hipLaunchKernelGGL((fdtd_pair<T, kCurDI, kPairRows, PL>), g, b, 0, s, a, tiles, sids); break;
    case kCurLKPoly: hipLaunchKernelGGL((fdtd_pair<T, kCurLKPoly, kPairRows, PL>), g, b, 0, s, a, tiles, sids); break;
    default: hipLaunchKernelGGL((fdtd_pair<T, kCurLKClosed, kPairRows, PL>), g, b, 0, s, a, tiles, sids); break;
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_pair_pattern(const StepArgs<T>* a, const uint32_t* tiles, int n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  constexpr int PL = kPairLanes<T>;
  hipLaunchKernelGGL((pair_pattern<T, kPairRows, PL>), dim3(n), dim3(kPairRows * PL), 0, s, a, tiles);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_ntff(const NtffArgs<T>* a, const uint2* blocks, int nblocks, hipStream_t s) {
  if (nblocks <= 0) return hipSuccess;
  hipLaunchKernelGGL((ntff_sample<T>), dim3(nblocks), dim3(kNtffBlock), 0, s, a, blocks);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_ntff_contract(const NtffArgs<T>* a, long long U, int M, long long nwaves, hipStream_t s,
                                int flush, int acc32) {
  if (U <= 0) return hipSuccess;
  if (acc32)  // compensated FP32 sums on FP32 matrix tiles, every M
    hipLaunchKernelGGL((ntff_contract_mfma32<T>), dim3((unsigned)((nwaves + 3) / 4)), dim3(256), 0, s, a, flush);
  else if (2 * M >= 16)  // matrix cores: 4 wavefronts of 16 points per workgroup
    hipLaunchKernelGGL((ntff_contract_mfma<T>), dim3((unsigned)((nwaves + 3) / 4)), dim3(256), 0, s, a, flush);
  else
    hipLaunchKernelGGL((ntff_contract<T>), dim3((unsigned)((U + 255) / 256)), dim3(256), 0, s, a, flush);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_frames(const AuxArgs<T>* a, hipStream_t s) {
  hipLaunchKernelGGL((cadence_frames<T>), dim3(1024), dim3(256), 0, s, a);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_probes(const AuxArgs<T>* a, int nsolves, hipStream_t s) {
  if (nsolves <= 0) return hipSuccess;
  hipLaunchKernelGGL((probes_and_tick<T>), dim3(nsolves), dim3(64), 0, s, a);
  return hipGetLastError();
}

template <typename T>
hipError_t launch_catchup(const StepArgs<T>* a, int cur, T* ox, T* oy, T* oz, hipStream_t s) {
  switch (cur) {
    case kCurDI: hipLaunchKernelGGL((current_catchup<T, kCurDI>), dim3(2048), dim3(256), 0, s, a, ox, oy, oz); break;
    case kCurLKPoly: hipLaunchKernelGGL((current_catchup<T, kCurLKPoly>), dim3(2048), dim3(256), 0, s, a, ox, oy, oz); break;
    default: hipLaunchKernelGGL((current_catchup<T, kCurLKClosed>), dim3(2048), dim3(256), 0, s, a, ox, oy, oz); break;
  }
  return hipGetLastError();
}

template <typename T>
hipError_t launch_field_slab(const StepArgs<T>* a, int cur, int comp, long long c0, long long nc, double* dst,
                             hipStream_t s) {
  if (nc <= 0) return hipSuccess;
  const unsigned blocks = (unsigned)std::min<long long>((nc + 255) / 256, 4096);
  switch (cur) {
    case kCurDI: hipLaunchKernelGGL((field_slab<T, kCurDI>), dim3(blocks), dim3(256), 0, s, a, comp, c0, nc, dst); break;
    case kCurLKPoly:
      hipLaunchKernelGGL((field_slab<T, kCurLKPoly>), dim3(blocks), dim3(256), 0, s, a, comp, c0, nc, dst);
      break;
    default:
      hipLaunchKernelGGL((field_slab<T, kCurLKClosed>), dim3(blocks), dim3(256), 0, s, a, comp, c0, nc, dst);
      break;
  }
  return hipGetLastError();
}

#define MSC_INST(T)                                                                                                \
  template hipError_t launch_step<T>(const StepArgs<T>*, const uint16_t*, int, int, int, bool, int, const uint32_t*, int, \
                                     hipStream_t);                                                                 \
  template hipError_t launch_steps<T>(const StepArgs<T>*, const uint16_t*, int, int, bool, int, const uint32_t*,     \
                                      const int*, const int*, hipStream_t);                                      \
  template hipError_t launch_pair<T>(const StepArgs<T>*, const uint16_t*, int, const uint32_t*, int, hipStream_t); \
  template hipError_t launch_ntff<T>(const NtffArgs<T>*, const uint2*, int, hipStream_t);                          \
  template hipError_t launch_pair_pattern<T>(const StepArgs<T>*, const uint32_t*, int, hipStream_t);           \
  template hipError_t launch_ntff_contract<T>(const NtffArgs<T>*, long long, int, long long, hipStream_t, int, int); \
  template hipError_t launch_frames<T>(const AuxArgs<T>*, hipStream_t);                                            \
  template hipError_t launch_probes<T>(const AuxArgs<T>*, int, hipStream_t);                                       \
  template hipError_t launch_catchup<T>(const StepArgs<T>*, int, T*, T*, T*, hipStream_t);                     \
  template hipError_t launch_field_slab<T>(const StepArgs<T>*, int, int, long long, long long, double*, hipStream_t);
MSC_INST(float)
MSC_INST(double)
#undef MSC_INST

}  // namespace msc
