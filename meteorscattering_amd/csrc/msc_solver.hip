// Solver handle: FP64 host precompute, device state, step sequencing and
// read-back.  The set-up mirrors meteor-fork.cpp:119-646; the per-step work
// is the kernel sequence of msc_kernels.hip.
#include "msc_solver.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <initializer_list>
#include <memory>
#include <mutex>
#include <string>
#include <type_traits>
#include <vector>

namespace msc {

// ------------------------------------------------- allocation-free read-back
// Device scratch of msc_solver_get_field_noalloc: one fixed array of
// kScratchDoubles FP64 values (8 MiB) on each device the library runs on,
// shared by every solver there and held until the process ends.  In host
// code `scratch` names a host-side stand-in, never the device memory: the
// device address is looked up once per device, with that device current,
// at solver creation (scratch_on_current_device), and only that address is
// handed to kernels and copies.
constexpr size_t kScratchDoubles = size_t(1) << 20;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
double* scratch_addr[kMaxDevices] = {};
}  // namespace

// The scratch array's address on the current device (looked up once per
// device).
static int scratch_on_current_device(int* dev_out, double** out) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(MSC_EINVAL, "device %d beyond the %d supported", dev, kMaxDevices);
  std::lock_guard<std::mutex> lk(scratch_mu);
  if (!scratch_addr[dev]) {
    void* p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) return fail(MSC_EHIP, "no device address for the read-back scratch on device %d", dev);
    scratch_addr[dev] = static_cast<double*>(p);
  }
  *dev_out = dev;
  *out = scratch_addr[dev];
  return MSC_OK;
}

// ------------------------------------------------------------ host precompute


// Source waveforms, meteor-fork.cpp:153-179 (+ extension type 4).
static void plan_sources(HostPlan& h) {
  const msc_params& p = h.p;
  const double f0 = p.f0, fsig = p.fsig, dt = h.dt;
  const double rdelayA = 1.6 * std::sqrt(6) / kPi / f0;
  const double rdelayB = (1.6 + 1 / std::sqrt(3)) * std::sqrt(6) / kPi / f0;
  const double tau = p.src_tau > 0 ? p.src_tau : 1.0 / (kPi * f0);
  h.srcA.resize(h.T);
  h.srcB.resize(h.T);
  for (int t = 0; t < h.T; t++) {
    double a, b;
    if (p.sourcetype == 2) {
      a = std::sin(2 * kPi * f0 * (t - 3 * fsig) * dt) * std::exp(-(t - 3 * fsig) * (t - 3 * fsig) / (2 * fsig * fsig));
      b = std::cos(2 * kPi * f0 * (t - 3 * fsig) * dt) * std::exp(-(t - 3 * fsig) * (t - 3 * fsig) / (2 * fsig * fsig));
    } else if (p.sourcetype == 3) {
      a = (1 - 2 * kPi * kPi * f0 * f0 * (t * dt - rdelayA) * (t * dt - rdelayA)) *
          std::exp(-kPi * kPi * f0 * f0 * (t * dt - rdelayA) * (t * dt - rdelayA));
      b = (1 - 2 * kPi * kPi * f0 * f0 * (t * dt - rdelayB) * (t * dt - rdelayB)) *
          std::exp(-kPi * kPi * f0 * f0 * (t * dt - rdelayB) * (t * dt - rdelayB));
    } else if (p.sourcetype == 4) {
      const double u = (t * dt - 4 * tau) / tau, v = (t * dt - 5 * tau) / tau;
      a = -u * std::exp(0.5 - 0.5 * u * u);
      b = -v * std::exp(0.5 - 0.5 * v * v);
    } else {
      a = std::sin(2 * kPi * f0 * t * dt);
      b = std::cos(2 * kPi * f0 * t * dt);
    }
    h.srcA[t] = a;
    h.srcB[t] = b;
  }
}

// CPML tables (meteor-fork.cpp:444-561, whose stretching kappa is 1 and
// frequency shift alpha 0, so its recursion coefficients reduce exactly to
// b = exp(-sigma dt / eps0) and a = b - 1).  Slab cell n of the high side
// (table entry L + n) and its mirror on the low side (entry L - 1 - n) take
// the polynomially graded conductivity sigma_max ((n + 1/2) / L)^4 on the
// E grid and sigma_max ((n + 1) / L)^4 on the H grid, sigma_max = 1.5 (4 + 1)
// / (150 pi dx).  One table per grid since dx == dy == dz.
static void plan_cpml(HostPlan& h) {
  const int L = h.L;
  h.be.assign(2 * L, 0); h.ae.assign(2 * L, 0); h.bh.assign(2 * L, 0); h.ah.assign(2 * L, 0);
  if (!h.p.dopml) return;
  const double grade = 4;
  const double sigma_max = 1.5 * (grade + 1) / (150 * kPi * h.dx);
  for (int n = 0; n < L; n++) {
    const double sig_e = sigma_max * std::pow((n + 0.5) / L, grade);
    const double sig_h = sigma_max * std::pow((n + 1.0) / L, grade);
    const double be = std::exp(-sig_e * h.dt / kE0), bh = std::exp(-sig_h * h.dt / kE0);
    for (const int m : {L + n, L - 1 - n}) {
      h.be[m] = be;
      h.ae[m] = be - 1;
      h.bh[m] = bh;
      h.ah[m] = bh - 1;
    }
  }
}

// The 1-D incident line (Eyinc, Ezinc, Hyinc, Hzinc with its +x CPML and the
// hard source at i=0) does not depend on the 3-D grid, so its whole history
// is computed here in FP64 (meteor-fork.cpp:736,772,864,894,923,952,1101,
// 1137,1232,1261,1290,1308-1335,1389), giving per step the TF/SF correction
// values, the source DFT (:1407-1414) and the lines written with each frame.
static void plan_incident(HostPlan& h) {
  const int nx = h.nx, L = h.L, T = h.T;
  const int xs = nx - 2 * L - 1;
  const double dx = h.dx, c1h = h.c1h, c2h = h.c2h, c1e0 = h.c1e0, c2e0 = h.c2e0;
  std::vector<double> eyi(nx, 0), ezi(nx, 0), hyi(nx, 0), hzi(nx, 0);
  std::vector<double> phy(2 * L, 0), phz(2 * L, 0), pey(2 * L, 0), pez(2 * L, 0);
  const bool pml = h.p.dopml != 0;
  h.inc.assign((size_t)T * 4 * nx, 0);
  h.eline.assign((size_t)h.nframes * 2 * nx, 0);
  h.hline.assign((size_t)h.nframes * 2 * nx, 0);
  for (int t = 0; t < T; t++) {
    double* row = &h.inc[(size_t)t * 4 * nx];
    for (int i = 0; i < nx; i++) {
      row[i] = c2h * ezi[i] / dx;
      row[nx + i] = c2h * eyi[i] / dx;
    }
    if (pml)
      for (int i = nx - L - 1; i < nx - 1; i++) {
        phy[i - xs] = h.bh[i - xs] * phy[i - xs] + h.ah[i - xs] * (ezi[i + 1] - ezi[i]) / dx;
        phz[i - xs] = h.bh[i - xs] * phz[i - xs] + h.ah[i - xs] * (eyi[i + 1] - eyi[i]) / dx;
      }
    for (int i = 0; i < nx - 1; i++) hyi[i] = c1h * hyi[i] + c2h * (ezi[i + 1] - ezi[i]) / dx;
    if (pml)
      for (int i = nx - L - 1; i < nx - 1; i++) hyi[i] = hyi[i] + c2h * phy[i - xs];
    for (int i = 0; i < nx - 1; i++) hzi[i] = c1h * hzi[i] - c2h * (eyi[i + 1] - eyi[i]) / dx;
    if (pml)
      for (int i = nx - L - 1; i < nx - 1; i++) hzi[i] = hzi[i] - c2h * phz[i - xs];
    for (int i = 0; i < nx; i++) {
      row[2 * nx + i] = c2e0 * hyi[i] / dx;
      row[3 * nx + i] = c2e0 * hzi[i] / dx;
    }
    if (pml)
      for (int i = nx - L - 1; i < nx - 1; i++) {
        pey[i - xs] = h.be[i - xs] * pey[i - xs] + h.ae[i - xs] * (hzi[i] - hzi[i - 1]) / dx;
        pez[i - xs] = h.be[i - xs] * pez[i - xs] + h.ae[i - xs] * (hyi[i] - hyi[i - 1]) / dx;
      }
    for (int i = 1; i < nx - 1; i++) eyi[i] = c1e0 * eyi[i] - c2e0 * (hzi[i] - hzi[i - 1]) / dx;
    if (pml)
      for (int i = nx - L - 1; i < nx - 1; i++) eyi[i] = eyi[i] - c2e0 * pey[i - xs];
    for (int i = 1; i < nx - 1; i++) ezi[i] = c1e0 * ezi[i] + c2e0 * (hyi[i] - hyi[i - 1]) / dx;
    const double a = h.srcA[t], b = h.srcB[t];
    switch (h.p.sourcepol) {  // meteor-fork.cpp:1308-1335
      case 2: ezi[0] = 0; eyi[0] = a; break;
      case 3: ezi[0] = a; eyi[0] = a; break;
      case 4: ezi[0] = a; eyi[0] = b; break;
      case 5: ezi[0] = b; eyi[0] = a; break;
      default: ezi[0] = a; eyi[0] = 0; break;
    }
    if (pml)
      for (int i = nx - L - 1; i < nx - 1; i++) ezi[i] = ezi[i] + c2e0 * pez[i - xs];
    if (t % h.cadence == 0) {
      const int f = t / h.cadence;
      std::copy(eyi.begin(), eyi.end(), h.eline.begin() + (size_t)f * 2 * nx);
      std::copy(ezi.begin(), ezi.end(), h.eline.begin() + (size_t)f * 2 * nx + nx);
      std::copy(hyi.begin(), hyi.end(), h.hline.begin() + (size_t)f * 2 * nx);
      std::copy(hzi.begin(), hzi.end(), h.hline.begin() + (size_t)f * 2 * nx + nx);
    }
  }
}

// Every media value finite (the reference reads wp.dat/nu.dat unchecked,
// meteor-fork.cpp:340-347; a NaN or infinity there would poison every
// field it reaches).
int check_media(const msc_params* p, const double* wp, const double* nu) {
  const size_t n = (size_t)p->nx * p->ny * (p->media3d ? (size_t)p->nz : 1);
  for (size_t c = 0; c < n; c++)
    if (!std::isfinite(wp[c]) || !std::isfinite(nu[c]))
      return fail(MSC_EINVAL, "media value %zu is not finite (wp %g, nu %g)", c, wp[c], nu[c]);
  return MSC_OK;
}

// Source DFT (meteor-fork.cpp:1407-1414) through `steps` steps: the hard
// source values of the incident line's first cell (:1308-1335) times the
// phase table, summed in step order as the reference does.
void source_dft(const HostPlan& h, int steps, double* ey, double* ez) {
  const int M = h.M;
  std::fill(ey, ey + 2 * M, 0.0);
  std::fill(ez, ez + 2 * M, 0.0);
  for (int t = 0; t < std::min(steps, h.T); t++) {
    const double a = h.srcA[t], b = h.srcB[t];
    double e0z, e0y;
    switch (h.p.sourcepol) {
      case 2: e0z = 0; e0y = a; break;
      case 3: e0z = a; e0y = a; break;
      case 4: e0z = a; e0y = b; break;
      case 5: e0z = b; e0y = a; break;
      default: e0z = a; e0y = 0; break;
    }
    for (int m = 0; m < M; m++) {
      const double sn = h.sincos[((size_t)t * M + m) * 2], cs = h.sincos[((size_t)t * M + m) * 2 + 1];
      ez[2 * m] += e0z * sn;
      ez[2 * m + 1] += e0z * cs;
      ey[2 * m] += e0y * sn;
      ey[2 * m + 1] += e0y * cs;
    }
  }
}

// Lee-Kalluri coefficients as polynomials in x = nu dt (FP64 Taylor
// coefficients of the closed-form integrals, meteor-fork.cpp:1546-1573):
//   e^{-x}                               = sum (-x)^n / n!
//   kI = int_0^dt e^{-nu s} cos(w0 s) ds  = dt   sum (-x)^n/n! int_0^1 u^n cos(th u) du
//   kX = int e^{-nu s} sin(w0 s)/w0 ds    = dt^2 sum (-x)^n/n! int_0^1 u^n sin(th u)/th du
//   kP = int e^{-nu s}(1-cos w0 s)/w0^2 ds = dt^3 sum (-x)^n/n! int_0^1 u^n (1-cos th u)/th^2 du
// with th = w0 dt and the u-moments summed as series in th.  Returns false
// (closed form instead) when the series would not be exact to FP64 over
// [0, xmax].
static bool plan_lk_poly(const HostPlan& h, double xmax, int& deg, double c[4][kMaxLkDeg + 1]) {
  const double th = h.w0 * h.dt, dt = h.dt;
  if (!(xmax >= 0) || xmax > 2.0 || th > 1.0) return false;
  deg = -1;
  double term = 1;  // xmax^(d+1)/(d+1)!
  for (int d = 0; d <= kMaxLkDeg; ++d) {
    term *= xmax / (d + 1);
    if (term < 1e-18) { deg = d; break; }
  }
  if (deg < 0 || deg > kLkPolyDeg) return false;
  if (deg < 1) deg = 1;
  double fact = 1;  // n!
  for (int n = 0; n <= kMaxLkDeg; ++n) {
    if (n > 0) fact *= n;
    double mc = 0, ms = 0, mp = 0, th2m = 1, f2m = 1;  // th^(2m), (2m)!
    for (int m = 0; m < 40; ++m) {
      if (m > 0) { th2m *= th * th; f2m *= (2.0 * m - 1) * (2.0 * m); }
      const double sg = (m & 1) ? -1.0 : 1.0;
      const double tc = sg * th2m / f2m / (n + 2 * m + 1);
      const double ts = sg * th2m / (f2m * (2 * m + 1)) / (n + 2 * m + 2);
      const double tp = sg * th2m / (f2m * (2 * m + 1) * (2 * m + 2)) / (n + 2 * m + 3);
      mc += tc; ms += ts; mp += tp;
      if (std::fabs(tc) < 1e-22 && m > 0) break;
    }
    const double sn = ((n & 1) ? -1.0 : 1.0) / fact;
    c[0][n] = sn;
    c[1][n] = dt * sn * mc;
    c[2][n] = dt * dt * sn * ms;
    c[3][n] = dt * dt * dt * sn * mp;
  }
  return true;
}

int make_plan(const msc_params* p, HostPlan& h) {
  int rc = msc_params_validate(p);
  if (rc) return rc;
  h.p = *p;
  h.nx = p->nx; h.ny = p->ny; h.nz = p->nz;
  h.L = pml_len(p);
  h.M = p->nntf;
  h.T = p->tsteps;
  h.cadence = p->tsteps / p->numfiles;
  h.nframes = (h.T - 1) / h.cadence + 1;
  h.dt = p->dt;
  const double lambda = kC / p->f0;
  h.dx = lambda / p->gridfactor;  // meteor-fork.cpp:127-132
  h.x0 = -(h.nx - 1) / 2 * h.dx;  // :137-141 (integer division), one loop per axis
  h.y0 = -(h.ny - 1) / 2 * h.dx;
  h.z0 = -(h.nz - 1) / 2 * h.dx;
  h.zcoord.resize(h.nz);
  for (int k = 0; k < h.nz; k++) h.zcoord[k] = -(h.nz - 1) / 2 * h.dx + h.dx * k;
  const double sigm = 0, sig = 0, dt = h.dt;  // :181-213
  h.c1h = (2 * kU0 - sigm * dt) / (2 * kU0 + sigm * dt);
  h.c2h = 2 * dt / (2 * kU0 + sigm * dt);
  h.c1e0 = (2 * kE0 - sig * dt) / (2 * kE0 + sig * dt);
  h.c2e0 = 2 * dt / (2 * kE0 + sig * dt);
  h.c1s = (2 * kE0 - 1e7 * dt) / (2 * kE0 + 1e7 * dt);
  h.c2s = 2 * dt / (2 * kE0 + 1e7 * dt);
  for (int q = 0; q < 3; q++) h.wce[q] = -kQE * p->B0[q] / kME;  // :228-232
  h.w0 = std::sqrt(h.wce[0] * h.wce[0] + h.wce[1] * h.wce[1] + h.wce[2] * h.wce[2]);
  if (h.w0 > 0) {
    h.cth = std::cos(h.w0 * dt);
    h.sth = std::sin(h.w0 * dt);
    h.Se1 = std::sin(h.w0 * dt) / h.w0;
    h.Ce1 = (1 - std::cos(h.w0 * dt)) / (h.w0 * h.w0);
  } else {
    h.cth = 1; h.sth = 0; h.Se1 = dt; h.Ce1 = 0.5 * dt * dt;
  }
  const double* w = h.wce;
  // R = Ce1 w w^T + cos I + Se1 [w x]   (A = Ee1 R, meteor-fork.cpp:1554-1562)
  h.R[0] = h.Ce1 * w[0] * w[0] + h.cth; h.R[1] = h.Ce1 * w[0] * w[1] - h.Se1 * w[2]; h.R[2] = h.Ce1 * w[0] * w[2] + h.Se1 * w[1];
  h.R[3] = h.Ce1 * w[1] * w[0] + h.Se1 * w[2]; h.R[4] = h.Ce1 * w[1] * w[1] + h.cth; h.R[5] = h.Ce1 * w[1] * w[2] - h.Se1 * w[0];
  h.R[6] = h.Ce1 * w[2] * w[0] - h.Se1 * w[1]; h.R[7] = h.Ce1 * w[2] * w[1] + h.Se1 * w[0]; h.R[8] = h.Ce1 * w[2] * w[2] + h.cth;
  h.sincos.resize((size_t)h.T * h.M * 2);
  for (int t = 0; t < h.T; t++)
    for (int m = 0; m < h.M; m++) {
      h.sincos[((size_t)t * h.M + m) * 2] = std::sin(2 * kPi * p->fntf[m] * t * dt);  // :977, :1410
      h.sincos[((size_t)t * h.M + m) * 2 + 1] = std::cos(2 * kPi * p->fntf[m] * t * dt);
    }
  plan_sources(h);
  plan_cpml(h);
  plan_incident(h);
  return MSC_OK;
}

// ---------------------------------------------------------------- arena

Arena::~Arena() {
  if (base_) (void)hipFree(base_);
}

int Arena::init(size_t bytes) {
  bytes &= ~(size_t)255;
  if (bytes == 0) return fail(MSC_EINVAL, "arena of 0 bytes");
  void* q = nullptr;
  hipError_t e = hipMalloc(&q, bytes);
  if (e != hipSuccess) return fail(MSC_ENOMEM, "hipMalloc of a %zu-byte arena: %s", bytes, hipGetErrorString(e));
  base_ = (char*)q;
  cap_ = bytes;
  free_[0] = bytes;
  return MSC_OK;
}

void* Arena::alloc(size_t bytes) {
  std::lock_guard<std::mutex> g(mu_);
  bytes = (std::max<size_t>(bytes, 1) + 255) & ~(size_t)255;
  for (auto it = free_.begin(); it != free_.end(); ++it) {
    if (it->second < bytes) continue;
    const size_t off = it->first, len = it->second;
    free_.erase(it);
    if (len > bytes) free_[off + bytes] = len - bytes;
    live_[off] = bytes;
    used_ += bytes;
    return base_ + off;
  }
  return nullptr;
}

void Arena::release(void* ptr) {
  if (!ptr) return;
  std::lock_guard<std::mutex> g(mu_);
  const size_t off = (size_t)((char*)ptr - base_);
  auto lv = live_.find(off);
  if (lv == live_.end()) return;
  size_t o = off, len = lv->second;
  live_.erase(lv);
  used_ -= len;
  auto nx = free_.lower_bound(o);
  if (nx != free_.end() && nx->first == o + len) {  // merge with the following range
    len += nx->second;
    nx = free_.erase(nx);
  }
  if (nx != free_.begin()) {  // and with the preceding one
    auto pv = std::prev(nx);
    if (pv->first + pv->second == o) {
      o = pv->first;
      len += pv->second;
      free_.erase(pv);
    }
  }
  free_[o] = len;
}

size_t Arena::used() const {
  std::lock_guard<std::mutex> g(mu_);
  return used_;
}

size_t Arena::largest_free() const {
  std::lock_guard<std::mutex> g(mu_);
  size_t m = 0;
  for (const auto& f : free_) m = std::max(m, f.second);
  return m;
}

// ---------------------------------------------------------------- solver

SolverBase::~SolverBase() {
  if (stream) (void)hipStreamSynchronize(stream);
  for (int q = 0; q < kMaxAux; q++) {
    if (aux[q]) {
      (void)hipStreamSynchronize(aux[q]);
      (void)hipStreamDestroy(aux[q]);
    }
    if (join_ev[q]) (void)hipEventDestroy(join_ev[q]);
  }
  if (fork_ev) (void)hipEventDestroy(fork_ev);
  if (blk) {
    if (arena) arena->release(blk);
    else (void)hipFree(blk);
  }
  for (hipEvent_t e : ev) (void)hipEventDestroy(e);
  for (hipEvent_t e : span)
    if (e) (void)hipEventDestroy(e);
  if (stream && own_stream) (void)hipStreamDestroy(stream);
}

int SolverBase::sync() {
  HIPCHK(hipStreamSynchronize(stream));
  if (timing && !ev.empty() && timed_steps > 0) {
    double tot = 0, tp = 0;
    int np = 0;
    for (size_t q = 0; q + 1 < ev.size(); q += 2) {
      float ms = 0;
      HIPCHK(hipEventElapsedTime(&ms, ev[q], ev[q + 1]));
      tot += ms;
      if (ev_pair[q / 2]) { tp += ms; np++; }
    }
    kernel_ms = tot / (double)timed_steps;
    pair_launches = np;
    pair_ms = np ? tp / np : 0.0;
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, span[0], span[1]));
    span_ms = ms / timed_steps;
  }
  return MSC_OK;
}

// The layout's tiles, decoded as the kernels decode them (fdtd_step:
// tile & 0xff -> zstart, bits 8-19 -> ystart (ystart_b for the CPML
// variants), bits 20-30 -> xstart; a workgroup owns at most its rows and
// zlanes - 1 lanes; fdtd_pair the same fields over xstart2 / ystart2 /
// zstart2, at most kPairRows - 4 rows and kPairOwnZ lanes), checked for
// bounds, for the CPML variant each tile needs, and for owning every cell
// exactly once.
template <typename T>
int Solver<T>::check_layout(msc_layout_report* rep) const {
  const HostPlan& h = plan;
  const int nx = h.nx, ny = h.ny, nz = h.nz, L = h.L;
  const bool pml = h.p.dopml != 0;
  std::vector<uint8_t> own((size_t)nx * ny * nz);
  auto zone = [&](int lo, int hi, int n) { return pml && (lo <= L || hi - 1 >= n - L - 1); };
  // mark [x0,x1) x [y0,y1) x [z0,z1); false if a cell is owned twice
  auto mark = [&](int x0, int x1, int y0, int y1, int z0, int z1) {
    for (int x = x0; x < x1; x++)
      for (int y = y0; y < y1; y++) {
        uint8_t* row = own.data() + ((size_t)x * ny + y) * nz;
        for (int z = z0; z < z1; z++) {
          if (row[z]) return false;
          row[z] = 1;
        }
      }
    return true;
  };
  auto all_owned = [&](const char* what) {
    for (size_t c = 0; c < own.size(); c++)
      if (!own[c])
        return fail(MSC_EINVAL, "layout: %s leaves cell (%zu, %zu, %zu) unowned", what, c / ((size_t)ny * nz),
                    c / nz % ny, c % nz);
    return MSC_OK;
  };
  // one one-step tile of list l; inregion: whether it may meet the region
  auto one_step = [&](uint32_t w, int l, bool outer) {
    const int tz = w & 0xff, ty_ = (w >> 8) & 0xfff, ch = (w >> 20) & 0x7ff;
    const int pm = list_pm(l);
    const std::vector<int>& yt = pm ? ysb_ : ys_;
    const int rows = l == kListInterior ? ty : ty_b;
    if (tz + 1 >= (int)zs_.size() || ty_ + 1 >= (int)yt.size() || ch + 1 >= (int)xs_.size())
      return fail(MSC_EINVAL, "layout: list %d tile %08x indexes past its tables (%zu, %zu, %zu)", l, w, zs_.size(),
                  yt.size(), xs_.size());
    const int x0 = xs_[ch], x1 = xs_[ch + 1], y0 = yt[ty_], y1 = yt[ty_ + 1], z0 = zs_[tz], z1 = zs_[tz + 1];
    if (!(0 <= x0 && x0 < x1 && x1 <= nx && 0 <= y0 && y0 < y1 && y1 <= ny && 0 <= z0 && z0 < z1 && z1 <= nz))
      return fail(MSC_EINVAL, "layout: list %d tile [%d,%d)x[%d,%d)x[%d,%d) is empty or leaves the grid", l, x0, x1,
                  y0, y1, z0, z1);
    if (y1 - y0 > rows || z1 - z0 > zlanes - 1)
      return fail(MSC_EINVAL, "layout: list %d tile owns %d rows x %d lanes, its workgroup %d x %d", l, y1 - y0,
                  z1 - z0, rows, zlanes - 1);
    const int need = (zone(x0, x1, nx) ? kPmlX : 0) | (zone(y0, y1, ny) ? kPmlY : 0) | (zone(z0, z1, nz) ? kPmlZ : 0);
    if ((need & ~pm) != 0)
      return fail(MSC_EINVAL, "layout: list %d tile [%d,%d)x[%d,%d)x[%d,%d) meets CPML axes %d, its variant %d", l, x0,
                  x1, y0, y1, z0, z1, need, pm);
    if (outer && x0 < x2hi && x1 > x2lo && y0 < y2hi && y1 > y2lo && z0 < z2hi && z1 > z2lo)
      return fail(MSC_EINVAL, "layout: outer tile [%d,%d)x[%d,%d)x[%d,%d) meets the two-step region", x0, x1, y0, y1,
                  z0, z1);
    if (!mark(x0, x1, y0, y1, z0, z1))
      return fail(MSC_EINVAL, "layout: list %d tile [%d,%d)x[%d,%d)x[%d,%d) owns a cell another tile owns", l, x0, x1,
                  y0, y1, z0, z1);
    return MSC_OK;
  };
  msc_layout_report r{};
  r.cells = (long long)nx * ny * nz;
  int rc;
  for (int l = 0; l < kNumLists; l++) {
    r.max_list_tiles = std::max<long long>(r.max_list_tiles, list_off[l + 1] - list_off[l]);
    for (int q = list_off[l]; q < list_off[l + 1]; q++)
      if ((rc = one_step(tiles_host[q], l, false))) return rc;
  }
  r.one_step_tiles = (long long)tiles_host.size();
  if ((rc = all_owned("the one-step tiling"))) return rc;
  if (!pair_host.empty()) {
    if (x2lo < L + 6 || y2lo < L + 6 || z2lo < L + 6 || x2hi > nx - L - 6 || y2hi > ny - L - 6 || z2hi > nz - L - 6)
      return fail(MSC_EINVAL, "layout: two-step region [%d,%d)x[%d,%d)x[%d,%d) is not 6 cells clear of the faces",
                  x2lo, x2hi, y2lo, y2hi, z2lo, z2hi);
    std::fill(own.begin(), own.end(), 0);
    for (int l = 0; l < kNumLists; l++)
      for (int q = outer_off[l]; q < outer_off[l + 1]; q++)
        if ((rc = one_step(pair_host[q], l, true))) return rc;
    for (size_t q = outer_off[kNumLists]; q < pair_host.size(); q++) {
      const uint32_t w = pair_host[q];
      const int tz = w & 0xff, ty_ = (w >> 8) & 0xfff, ch = (w >> 20) & 0x7ff;
      if (tz + 1 >= (int)zs2_.size() || ty_ + 1 >= (int)ys2_.size() || ch + 1 >= (int)xs2_.size())
        return fail(MSC_EINVAL, "layout: two-step tile %08x indexes past its tables", w);
      const int x0 = xs2_[ch], x1 = xs2_[ch + 1], y0 = ys2_[ty_], y1 = ys2_[ty_ + 1], z0 = zs2_[tz], z1 = zs2_[tz + 1];
      if (!(x2lo <= x0 && x0 < x1 && x1 <= x2hi && y2lo <= y0 && y0 < y1 && y1 <= y2hi && z2lo <= z0 && z0 < z1 &&
            z1 <= z2hi))
        return fail(MSC_EINVAL, "layout: two-step tile [%d,%d)x[%d,%d)x[%d,%d) is empty or leaves the region", x0, x1,
                    y0, y1, z0, z1);
      if (y1 - y0 > kPairRows - 4 || z1 - z0 > kPairOwnZ<T>)
        return fail(MSC_EINVAL, "layout: two-step tile owns %d rows x %d lanes, its workgroup %d x %d", y1 - y0,
                    z1 - z0, kPairRows - 4, kPairOwnZ<T>);
      if (!mark(x0, x1, y0, y1, z0, z1))
        return fail(MSC_EINVAL, "layout: two-step tile [%d,%d)x[%d,%d)x[%d,%d) owns a cell another tile owns", x0,
                    x1, y0, y1, z0, z1);
    }
    if ((rc = all_owned("the outer tiles with the two-step tiles"))) return rc;
    r.outer_tiles = outer_off[kNumLists];
    r.pair_tiles = (long long)pair_host.size() - outer_off[kNumLists];
    r.region_cells = pair_cells;
  }
  if (rep) *rep = r;
  return MSC_OK;
}

template <typename T>
int Solver<T>::pattern_probe(int reps, double* ms) {
  const int npair = pair_host.empty() ? 0 : (int)pair_host.size() - outer_off[kNumLists];
  if (npair <= 0) return fail(MSC_EINVAL, "the solve has no two-step region");
  if (reps < 1) return fail(MSC_EINVAL, "reps must be >= 1");
  hipEvent_t e0 = nullptr, e1 = nullptr;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  hipError_t e = launch_pair_pattern<T>(sa_dev, tiles2 + outer_off[kNumLists], npair, stream);  // warm
  if (e == hipSuccess) e = hipEventRecord(e0, stream);
  for (int q = 0; q < reps && e == hipSuccess; ++q)
    e = launch_pair_pattern<T>(sa_dev, tiles2 + outer_off[kNumLists], npair, stream);
  if (e == hipSuccess) e = hipEventRecord(e1, stream);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float t = 0;
  if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  if (e != hipSuccess) return fail(MSC_EHIP, "pattern probe: %s", hipGetErrorString(e));
  *ms = t / reps;
  return MSC_OK;
}

size_t SolverBase::ntff_doubles() const {
  const size_t M2 = 2 * (size_t)plan.M;
  return 2 * M2 + 8 * M2 * ((size_t)plan.ny * plan.nz + (size_t)plan.nx * plan.nz + (size_t)plan.nx * plan.ny);
}

// Tile lists and NTFF workgroup list, then the block layout.
template <typename T>
size_t Solver<T>::layout() {
  const HostPlan& h = plan;
  const int nx = h.nx, ny = h.ny, nz = h.nz, L = h.L;
  const int ownz = zlanes - 1;  // z cells of a one-step tile (zlanes is 32 or 64)
  // Two-step region: 6 cells clear of every face (L+6: clear of the CPML
  // slabs [0, L] and [n-L-1, n), of the NTFF planes at L+5 and n-L-5 with
  // their one-cell H averages, and far enough in that a one-step tile's
  // one-cell halo and the two rim layers never reach a slab).
  x2lo = x2hi = y2lo = y2hi = z2lo = z2hi = 0;
  if (pair) {
    const int m = L + 6;
    if (nx - m > m && ny - m > m) {
      // z bounds on the one-step z tiles: the z-band tiles below and above
      // the region stay full width (a narrow tile costs as much as a full one)
      const int zlo = (m + ownz - 1) / ownz * ownz, zhi = (nz - m) / ownz * ownz;
      // the region's tile count, as the even split below makes it
      auto pieces = [](int n, int piece) { return (long long)((n + piece - 1) / piece); };
      const long long nt2 = zhi > zlo ? pieces(nx - 2 * m, chunk2) * pieces(ny - 2 * m, kPairRows - 4) *
                                            pieces(zhi - zlo, kPairOwnZ<T>)
                                      : 0;
      if (zhi > zlo && nt2 >= min_pair_tiles) {
        x2lo = m; x2hi = nx - m; y2lo = m; y2hi = ny - m; z2lo = zlo; z2hi = zhi;
      }
    }
  }
  const bool region = x2hi > x2lo;
  // Tiles: x and y are cut at the CPML slab edges (L+1 and n-L-1), so a tile
  // either meets a slab on some axis (the variant carrying the CPML code) or
  // meets none (the interior variant); z keeps full ownz-wide tiles from 0,
  // the first and last of which hold the z slabs.  A tile's one-cell low
  // halo (row/column/plane L when the tile starts at L+1) needs no CPML
  // term: slab_h(L) is outside both H slabs, and no E is computed there.
  // (all three axes are also cut at the two-step region's bounds, so a
  // one-step tile is inside the region or outside it)
  auto cut = [&](int n, int piece, bool align, int rlo, int rhi) {
    std::vector<int> br;
    if (align && h.p.dopml) br = {L + 1, n - L - 1};
    if (rhi > rlo) { br.push_back(rlo); br.push_back(rhi); }
    std::sort(br.begin(), br.end());
    br.push_back(n);
    std::vector<int> st;
    int lo = 0;
    for (int b : br) {
      for (int q = lo; q < b; q += piece) st.push_back(q);
      lo = std::max(lo, b);
    }
    st.push_back(n);
    return st;
  };
  xs_ = cut(nx, chunk, true, x2lo, x2hi);
  // interior and boundary tiles cut y differently; both cuttings align at
  // L+1 and n-L-1, so a row is in a slab zone in both or in neither
  ys_ = cut(ny, ty, true, y2lo, y2hi);
  ysb_ = cut(ny, h.p.dopml ? ty_b : ty, true, 0, 0);
  zs_ = cut(nz, ownz, false, z2lo, z2hi);
  // fdtd_pair tiles: pieces of near-equal length
  auto even = [](int lo, int hi, int piece) {
    std::vector<int> st;
    const int np = hi > lo ? (hi - lo + piece - 1) / piece : 0;
    for (int q = 0; q <= np; q++) st.push_back(lo + (int)((long long)(hi - lo) * q / std::max(np, 1)));
    return st;
  };
  xs2_ = even(x2lo, x2hi, chunk2);
  ys2_ = even(y2lo, y2hi, kPairRows - 4);
  zs2_ = even(z2lo, z2hi, kPairOwnZ<T>);
  const int nch = (int)xs_.size() - 1, tiles_y = (int)ys_.size() - 1, tiles_z = (int)zs_.size() - 1;
  const int tiles_yb = (int)ysb_.size() - 1;
  auto zone = [&](int lo, int hi, int n) {  // owned range [lo, hi) meets p <= L or p >= n-L-1
    return lo < hi && (lo <= L || hi - 1 >= n - L - 1);
  };
  // interior tiles (ty rows) and CPML tiles (ty_b rows), the latter by the
  // set of slab axes their owned cells meet
  std::vector<uint32_t> lists[kNumLists];
  std::vector<uint32_t> outer;  // interior tiles outside the two-step region
  tiles_host.clear();
  pair_host.clear();
  pair_cells = 0;
  if (tiles_z <= 255 && tiles_y <= 4095 && tiles_yb <= 4095 && nch <= kMaxChunks) {
    for (int c = 0; c < nch; c++) {
      const bool xz = zone(xs_[c], xs_[c + 1], nx);
      for (int yy = 0; yy < tiles_y; yy++)
        for (int zz = 0; zz < tiles_z; zz++) {
          const bool b = h.p.dopml && (xz || zone(ys_[yy], ys_[yy + 1], ny) || zone(zs_[zz], zs_[zz + 1], nz));
          if (!b) {
            lists[kListInterior].push_back(pack_tile(zz, yy, c));
            const bool in = region && xs_[c] >= x2lo && xs_[c + 1] <= x2hi && ys_[yy] >= y2lo &&
                            ys_[yy + 1] <= y2hi && zs_[zz] >= z2lo && zs_[zz + 1] <= z2hi;
            if (!in) outer.push_back(pack_tile(zz, yy, c));
          }
        }
      if (!h.p.dopml) continue;
      for (int yy = 0; yy < tiles_yb; yy++)
        for (int zz = 0; zz < tiles_z; zz++) {
          const int m = (xz ? kPmlX : 0) | (zone(ysb_[yy], ysb_[yy + 1], ny) ? kPmlY : 0) |
                        (zone(zs_[zz], zs_[zz + 1], nz) ? kPmlZ : 0);
          if (!m) continue;
          const int l = m == kPmlX ? kListX : m == kPmlY ? kListY : m == kPmlZ ? kListZ : kListAll;
          lists[l].push_back(pack_tile(zz, yy, c));
        }
    }
  }
  for (int l = 0; l < kNumLists; l++) {
    list_off[l] = (int)tiles_host.size();
    tiles_host.insert(tiles_host.end(), lists[l].begin(), lists[l].end());
  }
  list_off[kNumLists] = (int)tiles_host.size();
  if (region && (int)zs2_.size() - 1 <= 255 && (int)ys2_.size() - 1 <= 4095 && (int)xs2_.size() - 1 <= kMaxChunks) {
    // the CPML lists are outside the region as they are
    for (int l = 0; l < kNumLists; l++) {
      outer_off[l] = (int)pair_host.size();
      const auto& src = l == kListInterior ? outer : lists[l];
      pair_host.insert(pair_host.end(), src.begin(), src.end());
    }
    outer_off[kNumLists] = (int)pair_host.size();
    pair_cells = (long long)(x2hi - x2lo) * (y2hi - y2lo) * (z2hi - z2lo);
    for (int c = 0; c + 1 < (int)xs2_.size(); c++)
      for (int yy = 0; yy + 1 < (int)ys2_.size(); yy++)
        for (int zz = 0; zz + 1 < (int)zs2_.size(); zz++) pair_host.push_back(pack_tile(zz, yy, c));
  }
  // NTFF: faces x (ny*nz points), y (nx*nz), z (nx*ny), two each
  ntff_host.clear();
  for (int face = 0; face < 6; face++) {
    const long long npts = face < 2 ? (long long)ny * nz : face < 4 ? (long long)nx * nz : (long long)nx * ny;
    const long long nb = (npts + kNtffBlock - 1) / kNtffBlock;
    for (long long b = 0; b < nb; b++) ntff_host.push_back(make_uint2(0u, ((uint32_t)face << 29) | (uint32_t)b));
  }
  const size_t M2 = 2 * (size_t)h.M;
  const size_t fb = (size_t)ny * nz, tb = (size_t)nx * nz, lr = (size_t)nx * ny;
  size_t off = 0, uo = 0;
  for (int q = 0; q < 24; q++) {
    const size_t np = q < 8 ? fb : q < 16 ? tb : lr;
    na.off[q] = (long long)off;
    na.uoff[q] = (long long)uo;
    off += M2 * np;
    uo += np;
  }
  na.uoff[24] = (long long)uo;
  na.wofs[0] = 0;
  for (int q = 0; q < 24; q++) na.wofs[q + 1] = na.wofs[q] + (na.uoff[q + 1] - na.uoff[q] + 15) / 16;
  acc_n = off;
  // sample-buffer depth: the accumulators' read-modify-write per step
  // (2 x 2M x 8 B per sample point, once per K steps) stays near 2.5% of the
  // field step's 80 B per cell, and at least 32 steps share one contraction
  // launch (a latency-bound pass: 80 us per call at 400x400x800); at most
  // 256 steps (96x96x192 with 64 frequencies: K = 213, a 314 MB FP32
  // sample buffer against 377 MB of FP64 sums)
  const double kk = std::ceil(16.0 * h.M * (double)uo / ((double)nx * ny * nz));
  na.K = (int)std::min(256.0, std::max(32.0, kk));
  // the contraction addresses the sample buffer with 32-bit byte offsets
  const double kmax = std::floor(4294967295.0 / ((double)uo * sizeof(T)));
  na.K = (int)std::max(1.0, std::min((double)na.K, kmax));
  na.T_steps = h.T;
  return carve(nullptr);
}

// Assigns every device pointer inside the block at `base` (nullptr: sizes
// only) and returns the block's length.  Arrays are 256-byte aligned.
template <typename T>
size_t Solver<T>::carve(char* base) {
  const HostPlan& h = plan;
  const int nx = h.nx, ny = h.ny, nz = h.nz, L = h.L;
  const size_t N = (size_t)nx * ny * nz;
  size_t total = 0;
  auto take = [&](auto& ptr, size_t n) {
    using P = std::remove_reference_t<decltype(ptr)>;
    using U = std::remove_const_t<std::remove_pointer_t<P>>;
    const size_t o = (total + 255) & ~(size_t)255;
    total = o + std::max<size_t>(n, 1) * sizeof(U);
    ptr = base ? reinterpret_cast<P>(base + o) : nullptr;
  };
  for (int q = 0; q < 3; q++)
    for (int c = 0; c < 3; c++) {
      take(sa.e[q][c], N);
      take(sa.h[q][c], N);
      take(sa.j[q][c], N);
    }
  take(sa.wpe2, N);
  take(sa.nue, N);
  // psi slabs: H ping-pong, E in place
  const size_t sxy = (size_t)nx * 2 * L * nz, sxz = (size_t)nx * ny * 2 * L, syx = (size_t)2 * L * ny * nz;
  const size_t psz[6] = {sxy, sxz, syx, sxz, syx, sxy};  // hxy hxz hyx hyz hzx hzy (and exy ... ezy)
  for (int q = 0; q < 6; q++) {
    if (h.p.dopml) {
      take(sa.psih[0][q], psz[q]);
      take(sa.psih[1][q], psz[q]);
      take(sa.psie[q], psz[q]);
    } else {
      sa.psih[0][q] = sa.psih[1][q] = sa.psie[q] = nullptr;
    }
  }
  take(sa.bh, 2 * L);
  take(sa.ah, 2 * L);
  take(sa.be, 2 * L);
  take(sa.ae, 2 * L);
  take(sa.inc, (size_t)h.T * 4 * nx);
  take(tdev, 1);
  take(health, 1);
  const size_t fsz = frame_doubles();
  if (h.p.savefields) {
    take(sa.eslice, fsz);
    take(eframes, fsz * h.nframes);
    take(hframes, fsz * h.nframes);
  } else {
    sa.eslice = nullptr;
    eframes = hframes = nullptr;
  }
  take(emax, h.nframes);
  take(acc, acc_n);
  take(na.sbuf, (size_t)na.K * (size_t)na.uoff[24]);
  take(na.sincos, (size_t)h.T * h.M * 2);
  if (acc32) take(na.sincos32, (size_t)h.T * h.M * 2);
  else na.sincos32 = nullptr;
  take(probes, (size_t)6 * h.T * std::max(1, h.p.nprobes));
  take(xa.px, MSC_MAX_PROBES);
  take(xa.py, MSC_MAX_PROBES);
  take(xa.pz, MSC_MAX_PROBES);
  take(tiles, tiles_host.size());
  take(tiles2, pair_host.size());
  take(sa.xstart2, xs2_.size());
  take(sa.ystart2, ys2_.size());
  take(sa.zstart2, zs2_.size());
  take(sa.xstart, xs_.size());
  take(sa.ystart, ys_.size());
  take(sa.ystart_b, ysb_.size());
  take(sa.zstart, zs_.size());
  take(ntff_blocks, ntff_host.size());
  take(sa_dev, 1);
  take(na_dev, 1);
  take(xa_dev, 1);
  return (total + 255) & ~(size_t)255;
}

template <typename T>
int Solver<T>::init(const double* wp, const double* nu) {
  const HostPlan& h = plan;
  const int nx = h.nx, ny = h.ny, nz = h.nz, L = h.L;
  const size_t N = (size_t)nx * ny * nz;
  if (tiling_from_env) {
    if (const char* s = std::getenv("MSC_TY")) ty = std::atoi(s);
    if (const char* s = std::getenv("MSC_TY_B")) ty_b = std::atoi(s);
    // one-step x chunks of 4 planes below 24 M cells: there the tiles are
    // few and their marches latency-bound, and shorter marches pay (3-27%
    // faster on 8 grids of 1.3-22 M cells, 3% slower at 33 M and 128 M)
    if (N < 24000000 && nx / 4 + 8 <= kMaxChunks) chunk = 4;
    // shorter still on the smallest grids, whose few tiles run as one
    // launch (FP32 ms per step, chunks of 2 / 3 / 4 planes: 48x48x96
    // 0.027 / 0.031 / 0.035, 64x64x128 0.039 / 0.035 / 0.039, 96x96x128
    // 0.059 / 0.057 / 0.063, 96x96x192 0.075 / 0.072 / 0.071)
    if (N < 1500000) chunk = N < 350000 ? 2 : 3;
    if (const char* s = std::getenv("MSC_CHUNK")) chunk = std::max(1, std::atoi(s));
    if (const char* s = std::getenv("MSC_PAIR")) pair = std::atoi(s) != 0;
    if (const char* s = std::getenv("MSC_CHUNK2")) chunk2 = std::max(1, std::atoi(s));
    if (const char* s = std::getenv("MSC_ZLANES")) zlanes = std::atoi(s);
  }
  if (const char* s = std::getenv("MSC_NTFF_ACC")) {
    const int v = std::atoi(s);
    if (v != 32 && v != 64) return fail(MSC_EINVAL, "MSC_NTFF_ACC must be 32 or 64 (got %s)", s);
    acc32 = v == 32;
  }
  auto tyok = [](int v) { return v == 3 || v == 7 || v == 15; };
  if (!tyok(ty) || !tyok(ty_b)) return fail(MSC_EINVAL, "MSC_TY and MSC_TY_B must be 3, 7 or 15");
  if (chunk < 1 || chunk2 < 1) return fail(MSC_EINVAL, "x chunks must be >= 1 plane");
  if (zlanes != 32 && zlanes != 64) return fail(MSC_EINVAL, "MSC_ZLANES must be 32 or 64");
  if (zlanes == 64 && plan.p.savefields)
    return fail(MSC_EINVAL, "64-lane one-step tiles carry no slice capture: savefields needs 32 lanes");
  if ((double)N * sizeof(T) >= 4294967296.0)
    return fail(MSC_EINVAL, "grid of %zu cells exceeds the 4 GiB-per-array limit of the 32-bit offset kernel", N);
  if ((double)plan.T * 2.0 * plan.M * sizeof(double) >= 4294967296.0)
    return fail(MSC_EINVAL, "%d steps x %d frequencies: the sin/cos table exceeds the 4 GiB the contraction addresses",
                plan.T, plan.M);
  const size_t bytes = layout();
  if ((int)zs_.size() - 1 > 255 || (int)ys_.size() - 1 > 4095 || (int)ysb_.size() - 1 > 4095 ||
      (int)xs_.size() - 1 > kMaxChunks)
    return fail(MSC_EINVAL, "grid too large for the tile descriptor");
  if (!stream) {
    HIPCHK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    own_stream = true;
  }
  if (const char* s = std::getenv("MSC_LIST_STREAMS")) list_streams = std::max(0, std::min(kMaxAux, std::atoi(s)));
  // one launch for the one-step lists below 16 MB per field array (4 M
  // cells in FP32, 2 M in FP64), where launch and fork/join gaps dominate a
  // step (FP32: 2.1x faster at 0.5 M cells, 1.45x at 1.8 M, 1.17x at 2.8 M;
  // FP64: 1.49x at 0.5 M, 1.08x at 1.8 M); above, the per-list launches
  // keep the interior tiles at their own occupancy (7 waves per SIMD
  // against the fused kernel's 5 in FP32) and are 2-9% faster
  fuse_lists = (double)N * sizeof(T) < 16e6;
  if (const char* s = std::getenv("MSC_FUSE_LISTS")) fuse_lists = std::atoi(s) != 0;

  if (arena) {
    blk = (char*)arena->alloc(bytes);
    if (!blk) return fail(MSC_EBUSY, "arena has no free range of %zu bytes", bytes);
  } else {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, bytes);
    if (e != hipSuccess) return fail(MSC_ENOMEM, "hipMalloc(%zu bytes): %s", bytes, hipGetErrorString(e));
    blk = (char*)q;
  }
  dev_bytes = bytes;
  carve(blk);
  HIPCHK(hipMemsetAsync(blk, 0, bytes, stream));
  // host buffers stay alive until the final synchronisation
  std::vector<std::vector<T>> keep;
  auto upload = [&](const std::vector<double>& src, const T* dst) -> int {
    keep.emplace_back(src.begin(), src.end());
    HIPCHK(hipMemcpyAsync((void*)dst, keep.back().data(), keep.back().size() * sizeof(T), hipMemcpyHostToDevice, stream));
    return MSC_OK;
  };
  int rc;
  if ((rc = upload(h.bh, sa.bh)) || (rc = upload(h.ah, sa.ah)) || (rc = upload(h.be, sa.be)) ||
      (rc = upload(h.ae, sa.ae)) || (rc = upload(h.inc, sa.inc)))
    return rc;
  // Tiles that carry a current: those whose J cells (owned cells, plus a
  // halo of 1 for one-step tiles and 2 for two-step tiles, which also
  // update J on their halo) meet a nonzero eps0 wp^2; the rest get
  // kTileNoCurrent below.  Boxes are inclusive, clipped to the grid.
  struct JBox {
    int x0, x1, y0, y1, z0, z1;
    uint32_t* word;
    long long own;
    bool hit;
  };
  std::vector<JBox> jb;
  auto add_boxes = [&](std::vector<uint32_t>& v, size_t lo, size_t hi, const std::vector<int>& xs,
                       const std::vector<int>& ys, const std::vector<int>& zs, int halo) {
    for (size_t q = lo; q < hi; q++) {
      const uint32_t w = v[q];
      const int tz = w & 0xff, tyy = (w >> 8) & 0xfff, ch = (w >> 20) & 0x7ff;
      const long long own = (long long)(xs[ch + 1] - xs[ch]) * (ys[tyy + 1] - ys[tyy]) * (zs[tz + 1] - zs[tz]);
      jb.push_back({std::max(0, xs[ch] - halo), std::min(nx - 1, xs[ch + 1] - 1 + halo), std::max(0, ys[tyy] - halo),
                    std::min(ny - 1, ys[tyy + 1] - 1 + halo), std::max(0, zs[tz] - halo),
                    std::min(nz - 1, zs[tz + 1] - 1 + halo), &v[q], own, false});
    }
  };
  for (int l = 0; l < kNumLists; l++) {
    const auto& ys = l == kListInterior ? ys_ : ysb_;
    add_boxes(tiles_host, list_off[l], list_off[l + 1], xs_, ys, zs_, 1);
    if (!pair_host.empty()) add_boxes(pair_host, outer_off[l], outer_off[l + 1], xs_, ys, zs_, 1);
  }
  const size_t jb_pair = jb.size();
  if (!pair_host.empty()) add_boxes(pair_host, outer_off[kNumLists], pair_host.size(), xs2_, ys2_, zs2_, 2);
  std::vector<int> sat((size_t)(ny + 1) * (nz + 1), 0);  // per-plane summed-area table of wpe2 != 0
  // media: eps0 wpe^2 and nu per cell, computed in FP64 as meteor-fork.cpp:336-366
  double numax = 0;
  {
    std::vector<T> pw((size_t)ny * nz), pn((size_t)ny * nz);
    std::vector<double> g(nz), gn(nz);
    const double msig0 = h.p.metsize[0];
    for (int k = 0; k < nz; k++) {
      g[k] = std::exp(-std::pow((h.zcoord[k] - h.p.metloc[2]), 2) / (msig0 * msig0));
      gn[k] = h.p.nu_scale_height > 0 ? std::exp(-(h.zcoord[k] - h.p.metloc[2]) / h.p.nu_scale_height) : 1.0;
    }
    for (int i = 0; i < nx; i++) {
      for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++) {
          double wpe = 0, nue = 0;
          if (!h.p.doplate) {
            if (h.p.media3d) {
              const size_t c = ((size_t)i * ny + j) * nz + k;
              wpe = wp[c];
              nue = nu[c];
            } else {
              wpe = wp[(size_t)i * ny + j] * g[k];
              nue = nu[(size_t)i * ny + j] * g[k];
            }
            if (h.p.nu_scale_height > 0) nue *= gn[k];
          }
          pw[(size_t)j * nz + k] = (T)(kE0 * wpe * wpe);
          pn[(size_t)j * nz + k] = (T)nue;
          numax = nue > numax ? nue : numax;
        }
      for (int j = 0; j < ny; j++)
        for (int k = 0; k < nz; k++)
          sat[(size_t)(j + 1) * (nz + 1) + k + 1] = (pw[(size_t)j * nz + k] != (T)0) + sat[(size_t)j * (nz + 1) + k + 1] +
                                                    sat[(size_t)(j + 1) * (nz + 1) + k] - sat[(size_t)j * (nz + 1) + k];
      for (JBox& b : jb) {
        if (b.hit || i < b.x0 || i > b.x1) continue;
        auto at = [&](int j, int k) { return sat[(size_t)j * (nz + 1) + k]; };
        b.hit = at(b.y1 + 1, b.z1 + 1) - at(b.y0, b.z1 + 1) - at(b.y1 + 1, b.z0) + at(b.y0, b.z0) > 0;
      }
      HIPCHK(hipMemcpyAsync((void*)(sa.wpe2 + (size_t)i * ny * nz), pw.data(), pw.size() * sizeof(T),
                            hipMemcpyHostToDevice, stream));
      HIPCHK(hipMemcpyAsync((void*)(sa.nue + (size_t)i * ny * nz), pn.data(), pn.size() * sizeof(T),
                            hipMemcpyHostToDevice, stream));
      HIPCHK(hipStreamSynchronize(stream));
    }
  }
  jc_single = jc_outer = jc_pair = 0;
  for (size_t q = 0; q < jb.size(); q++) {
    JBox& b = jb[q];
    if (!b.hit) *b.word |= kTileNoCurrent;
    else if (q >= jb_pair) jc_pair += b.own;
    else if (b.word >= tiles_host.data() && b.word < tiles_host.data() + tiles_host.size()) jc_single += b.own;
    else jc_outer += b.own;
  }
  if (std::getenv("MSC_CURRENT_EVERYWHERE")) {  // every tile updates J (no skipping)
    for (uint32_t& w : tiles_host) w &= ~kTileNoCurrent;
    for (uint32_t& w : pair_host) w &= ~kTileNoCurrent;
  }
  // current model: DI, or Lee-Kalluri with polynomial (default) / closed-form coefficients
  cur = kCurDI;
  if (h.p.doB0) {
    const char* force = std::getenv("MSC_LK_CLOSED");
    const bool want_closed = force && std::atoi(force) != 0;
    // the FP32 media rounding can lift nu slightly above the FP64 maximum
    if (!want_closed && plan_lk_poly(h, numax * 1.001 * h.dt, sa.lkdeg, sa.lkc)) cur = kCurLKPoly;
    else cur = kCurLKClosed;
  }
  HIPCHK(hipMemcpyAsync((void*)na.sincos, h.sincos.data(), h.sincos.size() * 8, hipMemcpyHostToDevice, stream));
  std::vector<float> sincos32;
  if (acc32) {
    sincos32.assign(h.sincos.begin(), h.sincos.end());  // FP64 phases rounded once
    HIPCHK(hipMemcpyAsync((void*)na.sincos32, sincos32.data(), sincos32.size() * 4, hipMemcpyHostToDevice, stream));
  }
  HIPCHK(hipMemcpyAsync((void*)xa.px, h.p.probex, 4 * MSC_MAX_PROBES, hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemcpyAsync((void*)xa.py, h.p.probey, 4 * MSC_MAX_PROBES, hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemcpyAsync((void*)xa.pz, h.p.probez, 4 * MSC_MAX_PROBES, hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemcpyAsync(tiles, tiles_host.data(), tiles_host.size() * 4, hipMemcpyHostToDevice, stream));
  if (!pair_host.empty()) {
    HIPCHK(hipMemcpyAsync(tiles2, pair_host.data(), pair_host.size() * 4, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync((void*)sa.xstart2, xs2_.data(), xs2_.size() * 4, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync((void*)sa.ystart2, ys2_.data(), ys2_.size() * 4, hipMemcpyHostToDevice, stream));
    HIPCHK(hipMemcpyAsync((void*)sa.zstart2, zs2_.data(), zs2_.size() * 4, hipMemcpyHostToDevice, stream));
  }
  HIPCHK(hipMemcpyAsync((void*)sa.xstart, xs_.data(), xs_.size() * 4, hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemcpyAsync((void*)sa.ystart, ys_.data(), ys_.size() * 4, hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemcpyAsync((void*)sa.ystart_b, ysb_.data(), ysb_.size() * 4, hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemcpyAsync((void*)sa.zstart, zs_.data(), zs_.size() * 4, hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemcpyAsync(ntff_blocks, ntff_host.data(), ntff_host.size() * sizeof(uint2), hipMemcpyHostToDevice, stream));

  // step arguments
  sa.nx = nx; sa.ny = ny; sa.nz = nz; sa.L = L;
  sa.xs = nx - 2 * L - 1; sa.ys = ny - 2 * L - 1; sa.zs = nz - 2 * L - 1;
  sa.bx1 = L + 10; sa.bx2 = nx - L - 10; sa.by1 = L + 10; sa.by2 = ny - L - 10; sa.bz1 = L + 10; sa.bz2 = nz - L - 10;
  sa.ssx = (nx + 1) / 2; sa.ssy = (ny + 1) / 2; sa.ssz = (nz + 1) / 2;
  sa.dopml = h.p.dopml; sa.doplate = h.p.doplate; sa.doB0 = h.p.doB0;
  sa.jcol = h.p.j_collocated;
  sa.plate_lo = (int)(sa.ssy - h.p.platesize / 2 / h.dx);  // meteor-fork.cpp:1339-1340
  sa.plate_up = (int)(sa.ssy + h.p.platesize / 2 / h.dx);
  sa.chunk = chunk;
  sa.tiles_y = (int)ys_.size() - 1;
  sa.tiles_z = (int)zs_.size() - 1;
  sa.nchunks = (int)xs_.size() - 1;
  sa.cadence = h.cadence;
  sa.sy = nz; sa.sx = (long long)ny * nz;
  sa.c2h = (T)h.c2h; sa.c2e0 = (T)h.c2e0; sa.inv_d = (T)(1.0 / h.dx);
  sa.c1s = (T)h.c1s; sa.c2s = (T)h.c2s;
  sa.x0 = h.x0; sa.y0 = h.y0; sa.z0 = h.z0; sa.dd = h.dx; sa.rsph = h.p.platesize;
  sa.dt = h.dt;
  sa.wx = h.wce[0]; sa.wy = h.wce[1]; sa.wz = h.wce[2]; sa.w0 = h.w0;
  sa.cth = h.cth; sa.sth = h.sth; sa.Se1 = h.Se1; sa.Ce1 = h.Ce1;
  for (int q = 0; q < 9; q++) sa.R[q] = h.R[q];
  sa.tdev = tdev;
  sa.x2lo = x2lo; sa.x2hi = x2hi; sa.y2lo = y2lo; sa.y2hi = y2hi; sa.z2lo = z2lo; sa.z2hi = z2hi;
  sa.nprobes = h.p.nprobes;
  sa.px = xa.px; sa.py = xa.py; sa.pz = xa.pz;

  na.nx = nx; na.ny = ny; na.nz = nz; na.M = h.M; na.b = L + 5;
  na.sy = sa.sy; na.sx = sa.sx;
  for (int q = 0; q < 3; q++)
    for (int c = 0; c < 3; c++) { na.e[q][c] = sa.e[q][c]; na.h[q][c] = sa.h[q][c]; }
  na.acc = acc; na.tdev = tdev;
  na.acc32 = acc32 ? 1 : 0;
  na.health = health;

  xa.nx = nx; xa.ny = ny; xa.nz = nz; xa.nprobes = h.p.nprobes; xa.tsteps = h.T;
  xa.cadence = h.cadence; xa.nframes = h.nframes;
  xa.sy = sa.sy; xa.sx = sa.sx;
  for (int q = 0; q < 3; q++)
    for (int c = 0; c < 3; c++) { xa.e[q][c] = sa.e[q][c]; xa.h[q][c] = sa.h[q][c]; }
  xa.probes = probes;
  xa.ssx = sa.ssx; xa.ssy = sa.ssy; xa.ssz = sa.ssz;
  xa.eslice = sa.eslice; xa.eframes = eframes; xa.hframes = hframes; xa.emax = emax;
  xa.tdev = tdev;
  HIPCHK(hipMemcpyAsync(sa_dev, &sa, sizeof(sa), hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemcpyAsync(na_dev, &na, sizeof(na), hipMemcpyHostToDevice, stream));
  HIPCHK(hipMemcpyAsync(xa_dev, &xa, sizeof(xa), hipMemcpyHostToDevice, stream));
  HIPCHK(hipStreamSynchronize(stream));
  return MSC_OK;
}

template <typename T>
int Solver<T>::step(int n) {
  if (n < 0) return fail(MSC_EINVAL, "nsteps < 0");
  if (t_host + n > plan.T) return fail(MSC_ESTATE, "stepping past tsteps (%d + %d > %d)", t_host, n, plan.T);
  for (hipEvent_t e : ev) (void)hipEventDestroy(e);
  ev.clear();
  ev_pair.clear();
  timed_steps = timing ? n : 0;
  if (timing) {
    for (hipEvent_t& e : span)
      if (!e) HIPCHK(hipEventCreate(&e));
    HIPCHK(hipEventRecord(span[0], stream));
  }
  // timed brackets: tick(kind) opens one on the stream, tock() closes it
  auto tick = [&](bool is_pair) -> int {
    if (!timing) return MSC_OK;
    hipEvent_t a = nullptr, b = nullptr;
    HIPCHK(hipEventCreate(&a));
    HIPCHK(hipEventCreate(&b));
    ev.push_back(a);
    ev.push_back(b);
    ev_pair.push_back(is_pair ? 1 : 0);
    HIPCHK(hipEventRecord(a, stream));
    return MSC_OK;
  };
  auto tock = [&]() -> int {
    if (timing) HIPCHK(hipEventRecord(ev.back(), stream));
    return MSC_OK;
  };
  // everything after the field update of step t_host: NTFF samples, frames,
  // probes and the step tick, the contraction at the end of a K-step block
  auto after = [&](bool cad) -> int {
    HIPCHK(launch_ntff<T>(na_dev, ntff_blocks, (int)ntff_host.size(), stream));
    if (cad) HIPCHK(launch_frames<T>(xa_dev, stream));
    HIPCHK(launch_probes<T>(xa_dev, 1, stream));
    if ((t_host + 1) % na.K == 0 || t_host + 1 == plan.T)
      HIPCHK(launch_ntff_contract<T>(na_dev, na.uoff[24], na.M, na.wofs[24], stream, 0, na.acc32));
    ++t_host;
    return MSC_OK;
  };
  // One step's one-step lists.  With auxiliary streams, the lists go to the
  // least loaded of `stream` and the aux streams, largest first (load = tiles
  // x tile rows, CPML tiles counted 1.5 times), between a fork and a join on
  // `stream`.
  if (list_streams > 0 && !fork_ev) {
    HIPCHK(hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming));
    for (int q = 0; q < std::min(list_streams, kMaxAux); q++) {
      HIPCHK(hipStreamCreateWithFlags(&aux[q], hipStreamNonBlocking));
      HIPCHK(hipEventCreateWithFlags(&join_ev[q], hipEventDisableTiming));
    }
  }
  int naux = 0;
  while (naux < std::min(list_streams, kMaxAux) && aux[naux]) naux++;
  auto lists = [&](const uint32_t* base, const int* off, bool slice) -> int {
    if (fuse_lists && ty == ty_b) {
      int cnt[kNumLists];
      for (int l = 0; l < kNumLists; l++) cnt[l] = off[l + 1] - off[l];
      HIPCHK(launch_steps<T>(sa_dev, nullptr, ty, cur, slice, zlanes, base, off, cnt, stream));
      return MSC_OK;
    }
    if (naux == 0) {
      for (int l = 0; l < kNumLists; l++)
        HIPCHK(launch_step<T>(sa_dev, nullptr, l == kListInterior ? ty : ty_b, cur, list_pm(l), slice, zlanes,
                              base + off[l], off[l + 1] - off[l], stream));
      return MSC_OK;
    }
    double w[kNumLists];
    int order[kNumLists];
    for (int l = 0; l < kNumLists; l++) {
      w[l] = (double)(off[l + 1] - off[l]) * (l == kListInterior ? ty : 1.5 * ty_b);
      order[l] = l;
    }
    std::sort(order, order + kNumLists, [&](int a, int b) { return w[a] > w[b]; });
    hipStream_t st[kMaxAux + 1] = {stream};
    for (int q = 0; q < naux; q++) st[q + 1] = aux[q];
    double load[kMaxAux + 1] = {};
    bool used[kMaxAux + 1] = {};
    HIPCHK(hipEventRecord(fork_ev, stream));
    for (int l : order) {
      const int n = off[l + 1] - off[l];
      if (n == 0) continue;
      int q = 0;
      for (int k = 1; k <= naux; k++)
        if (load[k] < load[q]) q = k;
      load[q] += w[l];
      if (q > 0 && !used[q]) HIPCHK(hipStreamWaitEvent(st[q], fork_ev, 0));
      used[q] = true;
      HIPCHK(launch_step<T>(sa_dev, nullptr, l == kListInterior ? ty : ty_b, cur, list_pm(l), slice, zlanes,
                            base + off[l], n, st[q]));
    }
    for (int q = 1; q <= naux; q++) {
      if (!used[q]) continue;
      HIPCHK(hipEventRecord(join_ev[q - 1], st[q]));
      HIPCHK(hipStreamWaitEvent(stream, join_ev[q - 1], 0));
    }
    return MSC_OK;
  };
  const int npair = pair_host.empty() ? 0 : (int)pair_host.size() - outer_off[kNumLists];
  int rc;
  for (int s = 0; s < n;) {
    const bool cad = (t_host % plan.cadence) == 0;
    const bool slice = cad && plan.p.savefields;
    // Two steps at once when nothing reads the whole grid between them:
    // no frame or slice after step t, no slice in step t+1.
    const bool cad1 = ((t_host + 1) % plan.cadence) == 0;
    if (npair > 0 && s + 1 < n && !cad && !(cad1 && plan.p.savefields)) {
      // step t: fdtd_pair over the region (steps t and t+1) and the
      // one-step launches outside it; step t+1 outside the region, which
      // reads the region's rim at step t+1
      if ((rc = tick(false)) || (rc = lists(tiles2, outer_off, false)) || (rc = tock())) return rc;
      if ((rc = tick(true))) return rc;
      HIPCHK(launch_pair<T>(sa_dev, nullptr, cur, tiles2 + outer_off[kNumLists], npair, stream));
      if ((rc = tock()) || (rc = after(false))) return rc;
      if ((rc = tick(false)) || (rc = lists(tiles2, outer_off, false)) || (rc = tock())) return rc;
      if ((rc = after(cad1))) return rc;
      s += 2;
      continue;
    }
    if ((rc = tick(false))) return rc;
    if ((rc = lists(tiles, list_off, slice))) return rc;
    if ((rc = tock())) return rc;
    if ((rc = after(cad))) return rc;
    s++;
  }
  if (timing) HIPCHK(hipEventRecord(span[1], stream));
  return MSC_OK;
}

template <typename T>
int Solver<T>::get_field(int comp, double* out) {
  if (comp < 0 || comp > 8) return fail(MSC_EINVAL, "component %d out of range", comp);
  const size_t N = (size_t)plan.nx * plan.ny * plan.nz;
  std::vector<T> tmp(N);
  const int par = t_host % 3;
  const T* src;
  T* scratch[3] = {nullptr, nullptr, nullptr};
  if (comp < 3) src = sa.e[par][comp];
  else if (comp < 6) src = sa.h[par][comp - 3];
  else {
    // the device J lags one step (deferred update); catch it up into scratch
    for (int c = 0; c < 3; c++) HIPCHK(hipMalloc(&scratch[c], N * sizeof(T)));
    hipError_t e = launch_catchup<T>(sa_dev, cur, scratch[0], scratch[1], scratch[2], stream);
    if (e != hipSuccess) {
      for (int c = 0; c < 3; c++) (void)hipFree(scratch[c]);
      return fail(MSC_EHIP, "catch-up kernel: %s", hipGetErrorString(e));
    }
    src = scratch[comp - 6];
  }
  hipError_t e = hipMemcpyAsync(tmp.data(), src, N * sizeof(T), hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  for (int c = 0; c < 3; c++)
    if (scratch[c]) (void)hipFree(scratch[c]);
  if (e != hipSuccess) return fail(MSC_EHIP, "field read-back: %s", hipGetErrorString(e));
  for (size_t c = 0; c < N; c++) out[c] = (double)tmp[c];
  return MSC_OK;
}

// msc_solver_get_field without allocating: the grid in passes of at most
// kScratchDoubles cells through the device scratch (field_slab writes the
// component as FP64, J caught up as current_catchup does), each pass copied
// straight into `out`.  Kernel and copy of a pass share the solver's stream,
// so the next pass's kernel starts after the copy has read the scratch.
template <typename T>
int Solver<T>::get_field_noalloc(int comp, double* out) {
  if (comp < 0 || comp > 8) return fail(MSC_EINVAL, "component %d out of range", comp);
  if (!scratch_dev) return fail(MSC_ESTATE, "no read-back scratch on this solver's device");
  const long long N = (long long)plan.nx * plan.ny * plan.nz, per = (long long)kScratchDoubles;
  for (long long c0 = 0; c0 < N; c0 += per) {
    const long long nc = std::min(per, N - c0);
    HIPCHK(launch_field_slab<T>(sa_dev, cur, comp, c0, nc, scratch_dev, stream));
    HIPCHK(hipMemcpyAsync(out + c0, scratch_dev, (size_t)nc * sizeof(double), hipMemcpyDeviceToHost, stream));
  }
  HIPCHK(hipStreamSynchronize(stream));
  return MSC_OK;
}

template <typename T>
int Solver<T>::get_ntff(double* out, int steps) {
  const size_t M2 = 2 * (size_t)plan.M;
  std::vector<double> a(acc_n);
  HIPCHK(hipMemcpyAsync(a.data(), acc, acc_n * 8, hipMemcpyDeviceToHost, stream));
  HIPCHK(hipStreamSynchronize(stream));
  source_dft(plan, steps, out, out + M2);
  if (acc32)  // (sum, compensation) float pairs: the value is their sum
    for (size_t e = 0; e < acc_n; e++) {
      float f[2];
      std::memcpy(f, &a[e], sizeof f);
      a[e] = (double)f[0] + (double)f[1];
    }
  double* o = out + 2 * M2;
  const size_t fb = (size_t)plan.ny * plan.nz, tb = (size_t)plan.nx * plan.nz, lr = (size_t)plan.nx * plan.ny;
  for (int q = 0; q < 24; q++) {
    const size_t np = q < 8 ? fb : q < 16 ? tb : lr;
    const double* src = a.data() + na.off[q];
    for (size_t pt = 0; pt < np; pt++)
      for (size_t m = 0; m < M2; m++) o[pt * M2 + m] = src[m * np + pt];
    o += np * M2;
  }
  return MSC_OK;
}

template <typename T>
int Solver<T>::flush_ntff() {
  if (t_host == 0 || t_host % na.K == 0 || t_host >= plan.T) return MSC_OK;
  HIPCHK(launch_ntff_contract<T>(na_dev, na.uoff[24], na.M, na.wofs[24], stream, 1, na.acc32));
  return MSC_OK;
}

template <typename T>
int Solver<T>::nonfinite(int* flag) {
  int v = 0;
  HIPCHK(hipMemcpyAsync(&v, health, sizeof v, hipMemcpyDeviceToHost, stream));
  HIPCHK(hipStreamSynchronize(stream));
  *flag = v != 0;
  return MSC_OK;
}

template <typename T>
int Solver<T>::get_probes(double* out) {
  const size_t n = (size_t)6 * plan.T * plan.p.nprobes;
  if (!n) return MSC_OK;
  HIPCHK(hipMemcpyAsync(out, probes, n * 8, hipMemcpyDeviceToHost, stream));
  HIPCHK(hipStreamSynchronize(stream));
  return MSC_OK;
}

template <typename T>
int Solver<T>::get_frame(int f, double* e, double* hh, double* em) {
  const int done = t_host == 0 ? 0 : (t_host - 1) / plan.cadence + 1;
  if (f < 0 || f >= done) return fail(MSC_EINVAL, "frame %d not captured (have %d)", f, done);
  const size_t fsz = frame_doubles(), nx = plan.nx;
  if ((e || hh) && !plan.p.savefields) return fail(MSC_ESTATE, "savefields=0: no slice frames");
  if (e) {
    HIPCHK(hipMemcpyAsync(e, eframes + (size_t)f * fsz, fsz * 8, hipMemcpyDeviceToHost, stream));
  }
  if (hh) {
    HIPCHK(hipMemcpyAsync(hh, hframes + (size_t)f * fsz, fsz * 8, hipMemcpyDeviceToHost, stream));
  }
  unsigned long long bits = 0;
  if (em) HIPCHK(hipMemcpyAsync(&bits, emax + f, 8, hipMemcpyDeviceToHost, stream));
  HIPCHK(hipStreamSynchronize(stream));
  if (e) std::memcpy(e + fsz, plan.eline.data() + (size_t)f * 2 * nx, 2 * nx * 8);
  if (hh) std::memcpy(hh + fsz, plan.hline.data() + (size_t)f * 2 * nx, 2 * nx * 8);
  if (em) std::memcpy(em, &bits, 8);
  return MSC_OK;
}

template struct Solver<float>;
template struct Solver<double>;

size_t solve_footprint(const msc_params* p, int precision) {
  std::unique_ptr<SolverBase> s;
  if (precision == 32) s.reset(new Solver<float>());
  else if (precision == 64) s.reset(new Solver<double>());
  else return 0;
  // the server's settings: its tiling, two-step mode, its one-step tile width
  s->tiling_from_env = false;
  s->pair = true;
  s->zlanes = kServerZLanes;
  if (const char* e = std::getenv("MSC_NTFF_ACC")) s->acc32 = std::atoi(e) == 32;
  if (make_plan(p, s->plan) != MSC_OK) return 0;
  return s->layout();
}

}  // namespace msc

using namespace msc;

struct msc_solver {
  std::unique_ptr<SolverBase> impl;
};

#define SOLVER_OK(s) \
  if (!(s) || !(s)->impl) return fail(MSC_EINVAL, "NULL solver")

namespace {
int check_tiling(const msc_tiling* t) {
  auto tyok = [](int v) { return v == 3 || v == 7 || v == 15; };
  if (!tyok(t->ty) || !tyok(t->ty_b)) return fail(MSC_EINVAL, "tiling rows must be 3, 7 or 15 (got %d, %d)", t->ty, t->ty_b);
  if (t->chunk < 1 || t->chunk2 < 1) return fail(MSC_EINVAL, "tiling chunks must be >= 1 (got %d, %d)", t->chunk, t->chunk2);
  return MSC_OK;
}

// a solver with an explicit tiling, set up and synchronised
// Two-step mode of a solve with an explicit tiling: on, unless MSC_PAIR=0.
bool tiled_pair(int) {
  if (const char* e = std::getenv("MSC_PAIR")) return std::atoi(e) != 0;
  return true;
}

// A single solve uses its two-step region only when the region gives at
// least 1.5 fdtd_pair workgroups (one per CU at a time) per CU of the
// current device: with fewer, most CUs idle through the long two-step
// marches and the one-step tiling is faster (on 48 random trails of 1-22 M cells
// the break-even lies near 400 tiles on 256 CUs).  MSC_PAIR=1 uses the
// region whatever its size.
void set_min_pair_tiles(SolverBase& s) {
  s.min_pair_tiles = 0;
  if (const char* e = std::getenv("MSC_PAIR"))
    if (std::atoi(e) != 0) return;
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess)
    s.min_pair_tiles = 3LL * cus / 2;
}

int make_tiled(const msc_params* p, const double* wp, const double* nu, int precision, const msc_tiling* t,
               std::unique_ptr<SolverBase>& s) {
  if (precision == 32) s.reset(new Solver<float>());
  else s.reset(new Solver<double>());
  s->tiling_from_env = false;
  s->ty = t->ty;
  s->ty_b = t->ty_b;
  s->chunk = t->chunk;
  s->chunk2 = t->chunk2;
  s->pair = tiled_pair(precision);
  set_min_pair_tiles(*s);
  int rc = make_plan(p, s->plan);
  if (rc) return rc;
  if ((rc = s->init(wp, nu))) return rc;
  if ((rc = scratch_on_current_device(&s->device, &s->scratch_dev))) return rc;
  return s->sync();
}
}  // namespace

extern "C" {

int msc_solver_create(const msc_params* p, const double* wp, const double* nu, int precision, msc_solver** out) {
  if (!p || !wp || !nu || !out) return fail(MSC_EINVAL, "NULL argument");
  *out = nullptr;
  if (precision != 32 && precision != 64) return fail(MSC_EINVAL, "precision must be 32 or 64 (got %d)", precision);
  int rc0 = msc_params_validate(p);
  if (rc0 || (rc0 = check_media(p, wp, nu))) return rc0;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(MSC_EHIP, "no HIP device visible");
  std::unique_ptr<SolverBase> s;
  if (precision == 32) s.reset(new Solver<float>());
  else s.reset(new Solver<double>());
  set_min_pair_tiles(*s);
  int rc = make_plan(p, s->plan);
  if (rc) return rc;
  rc = s->init(wp, nu);
  if (rc) return rc;
  if ((rc = scratch_on_current_device(&s->device, &s->scratch_dev))) return rc;
  rc = s->sync();
  if (rc) return rc;
  msc_solver* h = new msc_solver;
  h->impl = std::move(s);
  *out = h;
  return MSC_OK;
}

int msc_layout_check(const msc_params* p, int precision, int zlanes, const msc_tiling* tiling, msc_layout_report* rep) {
  if (!p) return fail(MSC_EINVAL, "NULL parameters");
  if (precision != 32 && precision != 64) return fail(MSC_EINVAL, "precision must be 32 or 64 (got %d)", precision);
  if (zlanes != 32 && zlanes != 64) return fail(MSC_EINVAL, "zlanes must be 32 or 64 (got %d)", zlanes);
  int rc = msc_params_validate(p);
  if (rc) return rc;
  msc_tiling t;
  msc_tiling_default(&t);
  if (tiling) {
    if ((rc = check_tiling(tiling))) return rc;
    t = *tiling;
  }
  std::unique_ptr<SolverBase> s;
  if (precision == 32) s.reset(new Solver<float>());
  else s.reset(new Solver<double>());
  // the server's settings: its tiling, two-step mode whatever the region's size
  s->tiling_from_env = false;
  s->pair = true;
  s->zlanes = zlanes;
  s->ty = t.ty;
  s->ty_b = t.ty_b;
  s->chunk = t.chunk;
  s->chunk2 = t.chunk2;
  if ((rc = make_plan(p, s->plan))) return rc;
  s->layout();
  if (s->tiles_host.empty()) return fail(MSC_EINVAL, "layout: no one-step tiles (a descriptor field would overflow)");
  return s->check_layout(rep);
}

void msc_tiling_default(msc_tiling* t) {
  if (!t) return;
  t->ty = 7;
  t->ty_b = 7;
  t->chunk = 8;
  t->chunk2 = 32;
}


int msc_solver_create_tiled(const msc_params* p, const double* wp, const double* nu, int precision,
                            const msc_tiling* tiling, msc_solver** out) {
  if (!p || !wp || !nu || !tiling || !out) return fail(MSC_EINVAL, "NULL argument");
  *out = nullptr;
  if (precision != 32 && precision != 64) return fail(MSC_EINVAL, "precision must be 32 or 64 (got %d)", precision);
  int rc = check_tiling(tiling);
  if (rc || (rc = msc_params_validate(p)) || (rc = check_media(p, wp, nu))) return rc;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(MSC_EHIP, "no HIP device visible");
  std::unique_ptr<SolverBase> s;
  if ((rc = make_tiled(p, wp, nu, precision, tiling, s))) return rc;
  msc_solver* h = new msc_solver;
  h->impl = std::move(s);
  *out = h;
  return MSC_OK;
}

int msc_autotune(const msc_params* p, const double* wp, const double* nu, int precision, int steps,
                 msc_tiling* best, double* best_ms) {
  if (!p || !wp || !nu || !best) return fail(MSC_EINVAL, "NULL argument");
  if (precision != 32 && precision != 64) return fail(MSC_EINVAL, "precision must be 32 or 64 (got %d)", precision);
  if (p->tsteps < 4) return fail(MSC_EINVAL, "autotuning needs tsteps >= 4");
  const int n = std::max(2, std::min(steps, p->tsteps - 2));
  // elapsed GPU time per step of one candidate
  auto time_one = [&](const msc_tiling& t, double* ms) -> int {
    std::unique_ptr<SolverBase> s;
    int rc = make_tiled(p, wp, nu, precision, &t, s);
    if (!rc) rc = s->step(2);
    if (!rc) rc = s->sync();
    if (rc) return rc;
    s->timing = true;
    if ((rc = s->step(n)) || (rc = s->sync())) return rc;
    *ms = s->span_ms;
    return MSC_OK;
  };
  msc_tiling cur;
  msc_tiling_default(&cur);
  double cur_ms = 0;
  int rc = time_one(cur, &cur_ms);
  if (rc) return rc;
  const std::vector<int> chunk2s = {16, 24, 32, 48, 64}, chunks = {4, 6, 8, 12, 16}, tys = {3, 7, 15};
  auto search = [&](int32_t msc_tiling::*field, const std::vector<int>& values) -> int {
    const int start = cur.*field;
    for (int v : values) {
      if (v == start) continue;
      msc_tiling c = cur;
      c.*field = v;
      double ms = 0;
      int r = time_one(c, &ms);
      if (r) return r;
      if (ms < cur_ms * 0.995) { cur = c; cur_ms = ms; }
    }
    return MSC_OK;
  };
  // chunk2 shapes only two-step tiles: searched only when the candidates run them
  if ((tiled_pair(precision) && (rc = search(&msc_tiling::chunk2, chunk2s))) ||
      (rc = search(&msc_tiling::chunk, chunks)) || (rc = search(&msc_tiling::ty, tys)) ||
      (rc = search(&msc_tiling::ty_b, tys)))
    return rc;
  *best = cur;
  if (best_ms) *best_ms = cur_ms;
  return MSC_OK;
}

int msc_solver_pattern_probe(msc_solver* s, int reps, double* ms_per_launch) {
  SOLVER_OK(s);
  if (!ms_per_launch) return fail(MSC_EINVAL, "NULL argument");
  int rc = s->impl->sync();
  if (rc) return rc;
  return s->impl->pattern_probe(reps, ms_per_launch);
}

int msc_solver_step(msc_solver* s, int nsteps) {
  SOLVER_OK(s);
  return s->impl->step(nsteps);
}

int msc_solver_sync(msc_solver* s) {
  SOLVER_OK(s);
  return s->impl->sync();
}

int msc_solver_time(const msc_solver* s) { return (s && s->impl) ? s->impl->t_host : -1; }

int msc_solver_get_field(msc_solver* s, int comp, double* out) {
  SOLVER_OK(s);
  if (!out) return fail(MSC_EINVAL, "NULL output");
  return s->impl->get_field(comp, out);
}

int msc_solver_get_field_noalloc(msc_solver* s, int comp, double* out, size_t n) {
  SOLVER_OK(s);
  const HostPlan& h = s->impl->plan;
  const size_t N = (size_t)h.nx * h.ny * h.nz;
  if (!out || n < N) return fail(MSC_EINVAL, "output holds %zu values, the grid has %zu", out ? n : 0, N);
  return s->impl->get_field_noalloc(comp, out);
}

size_t msc_solver_ntff_count(const msc_solver* s) {
  if (!s || !s->impl) return 0;
  const HostPlan& h = s->impl->plan;
  const size_t M2 = 2 * (size_t)h.M;
  return 2 * M2 + 8 * M2 * ((size_t)h.ny * h.nz + (size_t)h.nx * h.nz + (size_t)h.nx * h.ny);
}

int msc_solver_get_ntff(msc_solver* s, double* out, size_t n) {
  SOLVER_OK(s);
  if (!out || n < msc_solver_ntff_count(s)) return fail(MSC_EINVAL, "output buffer too small");
  int rc = s->impl->flush_ntff();
  return rc ? rc : s->impl->get_ntff(out, s->impl->t_host);
}

int msc_solver_flush_ntff(msc_solver* s) {
  SOLVER_OK(s);
  return s->impl->flush_ntff();
}

int msc_solver_nonfinite(msc_solver* s, int* flag) {
  SOLVER_OK(s);
  if (!flag) return fail(MSC_EINVAL, "NULL argument");
  return s->impl->nonfinite(flag);
}

int msc_solver_get_probes(msc_solver* s, double* out, size_t n) {
  SOLVER_OK(s);
  const HostPlan& h = s->impl->plan;
  if (n < (size_t)6 * h.T * h.p.nprobes || (!out && h.p.nprobes)) return fail(MSC_EINVAL, "output buffer too small");
  return s->impl->get_probes(out);
}

size_t msc_solver_frame_count(const msc_solver* s) {
  if (!s || !s->impl) return 0;
  return s->impl->frame_doubles() + 2 * (size_t)s->impl->plan.nx;
}

int msc_solver_nframes(const msc_solver* s) {
  if (!s || !s->impl) return -1;
  const int t = s->impl->t_host;
  return t == 0 ? 0 : (t - 1) / s->impl->plan.cadence + 1;
}

int msc_solver_get_frame(msc_solver* s, int f, double* e, double* h, double* emax) {
  SOLVER_OK(s);
  return s->impl->get_frame(f, e, h, emax);
}

int msc_solver_max_abs_ex(msc_solver* s, double* out) {
  SOLVER_OK(s);
  const HostPlan& h = s->impl->plan;
  std::vector<double> ex((size_t)h.nx * h.ny * h.nz);
  int rc = s->impl->get_field(MSC_EX, ex.data());
  if (rc) return rc;
  double m = 0;
  for (double v : ex) m = std::fabs(v) > m ? std::fabs(v) : m;
  *out = m;
  return MSC_OK;
}

size_t msc_solver_device_bytes(const msc_solver* s) { return (s && s->impl) ? s->impl->dev_bytes : 0; }

int msc_solver_set_timing(msc_solver* s, int on) {
  SOLVER_OK(s);
  s->impl->timing = on != 0;
  return MSC_OK;
}

double msc_solver_kernel_ms(const msc_solver* s) { return (s && s->impl) ? s->impl->kernel_ms : 0.0; }

int msc_solver_get_timing(const msc_solver* s, msc_timing* out) {
  SOLVER_OK(s);
  if (!out) return fail(MSC_EINVAL, "NULL argument");
  const SolverBase& b = *s->impl;
  out->steps = b.timed_steps;
  out->pair_launches = b.pair_launches;
  out->step_ms = b.kernel_ms;
  out->pair_ms = b.pair_ms;
  out->pair_cells = b.pair_cells;
  out->span_ms = b.span_ms;
  out->pair_j_cells = b.jc_pair;
  out->step_j_cells = b.pair_cells > 0 ? b.jc_pair + b.jc_outer : b.jc_single;
  return MSC_OK;
}

void msc_solver_destroy(msc_solver* s) { delete s; }

int msc_stream_copy(size_t bytes, int iters, double* gbps) {
  if (bytes < 16 || iters < 1 || !gbps) return fail(MSC_EINVAL, "bad arguments");
  bytes &= ~(size_t)15;
  void *a = nullptr, *b = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  int rc = MSC_OK;
  float ms = 0;
  if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
    rc = fail(MSC_ENOMEM, "hipMalloc of 2 x %zu bytes failed", bytes);
  } else if (hipStreamCreate(&st) != hipSuccess || hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    rc = fail(MSC_EHIP, "stream/event creation failed");
  } else {
    // the best of several grid sizes (1 to 16 workgroups per CU on 256 CUs)
    hipError_t e = hipMemsetAsync(a, 0, bytes, st);
    double best = 0;
    for (int blocks : {256, 512, 1024, 2048, 4096}) {
      if (e == hipSuccess) e = launch_stream_copy(a, b, bytes, blocks, st);  // warm
      if (e == hipSuccess) e = hipEventRecord(e0, st);
      for (int q = 0; q < iters && e == hipSuccess; ++q) e = launch_stream_copy(a, b, bytes, blocks, st);
      if (e == hipSuccess) e = hipEventRecord(e1, st);
      if (e == hipSuccess) e = hipEventSynchronize(e1);
      if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
      if (e == hipSuccess) best = std::max(best, 2.0 * (double)bytes * iters / (ms * 1e-3) / 1e9);
    }
    if (e != hipSuccess) rc = fail(MSC_EHIP, "stream copy: %s", hipGetErrorString(e));
    else *gbps = best;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  return rc;
}

}  // extern "C"
