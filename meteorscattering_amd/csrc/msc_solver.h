// Host plan and per-solve device state, shared by the single-solve handle
// (msc_solver.hip) and the multi-solve server (msc_server.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <mutex>
#include <vector>

#include "msc_common.h"
#include "msc_kernels.h"

namespace msc {

#define HIPCHK(expr)                                                                  \
  do {                                                                                \
    hipError_t e_ = (expr);                                                           \
    if (e_ != hipSuccess) return fail(MSC_EHIP, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

// FP64 host precompute of one solve (meteor-fork.cpp:119-646).
struct HostPlan {
  msc_params p;
  int nx, ny, nz, L, M, T, cadence, nframes;
  double dx, dt, x0, y0, z0;
  double c1h, c2h, c1e0, c2e0, c1s, c2s;
  double wce[3], w0, cth, sth, Se1, Ce1, R[9];
  std::vector<double> be, ae, bh, ah;  // 2L
  std::vector<double> srcA, srcB;      // T
  std::vector<double> inc;             // [T][4][nx]
  std::vector<double> eline, hline;    // [nframes][2][nx] incident lines at frame steps
  std::vector<double> sincos;          // [T][M][2]
  std::vector<double> zcoord;          // nz
};

int make_plan(const msc_params* p, HostPlan& h);
// Every media value (wp, nu maps as msc_read_run_dir returns them) finite.
int check_media(const msc_params* p, const double* wp, const double* nu);
// The source DFT sums Eydft[2M], Ezdft[2M] through `steps` time steps.
void source_dft(const HostPlan& h, int steps, double* ey, double* ez);

// z lanes of the one-step tiles of the server's gang launches: 32, so the
// z bands around a solve's two-step region are 31 planes deep, not 63
constexpr int kServerZLanes = 32;

// Device bytes of one solve (its single block) with the server's settings,
// without allocating.
size_t solve_footprint(const msc_params* p, int precision);

// One device allocation carved into solve blocks: first fit over a free
// list, 256-byte aligned, coalescing on release.  Thread-safe.
class Arena {
 public:
  ~Arena();
  int init(size_t bytes);
  void* alloc(size_t bytes);  // nullptr when no free range fits
  void release(void* ptr);
  size_t capacity() const { return cap_; }
  size_t used() const;
  size_t largest_free() const;

 private:
  char* base_ = nullptr;
  size_t cap_ = 0, used_ = 0;
  std::map<size_t, size_t> free_;  // offset -> length
  std::map<size_t, size_t> live_;  // offset -> length
  mutable std::mutex mu_;
};

// Device state of one solve.  All of it lives in ONE block (hipMalloc, or a
// range of a server's Arena), laid out array after array.
struct SolverBase {
  HostPlan plan;
  hipStream_t stream = nullptr;
  bool own_stream = true;
  Arena* arena = nullptr;
  char* blk = nullptr;
  size_t dev_bytes = 0;
  int t_host = 0;
  int ty = 7, ty_b = 7, chunk = 8;  // interior / CPML-boundary tile rows, x-chunk planes
  // z lanes of a one-step workgroup row (32: two rows per wave; 64); a
  // one-step tile owns zlanes - 1 z cells, and the two-step region's z
  // bounds are multiples of that, so the z bands around it stay full tiles
  int zlanes = 32;
  // NTFF sums as compensated FP32 pairs on FP32 matrix tiles (MSC_NTFF_ACC=32)
  // instead of FP64 sums (the default, MSC_NTFF_ACC=64)
  bool acc32 = false;
  bool tiling_from_env = true;      // MSC_TY / MSC_CHUNK may override ty / chunk
  int cur = 0;                      // kCurDI / kCurLKPoly / kCurLKClosed
  std::vector<uint32_t> tiles_host; // the kNumLists tile lists, one after another
  int list_off[kNumLists + 1] = {};  // list l is tiles_host[list_off[l], list_off[l+1])
  std::vector<uint2> ntff_host;     // NTFF workgroup descriptors (solve id 0)
  // Two-step mode: steps t, t+1 advance the two-step region with fdtd_pair
  // and the rest of the grid with the one-step lists restricted to tiles
  // outside it.  pair_host holds those outer lists (offsets outer_off) and
  // then the fdtd_pair tiles (from outer_off[kNumLists] to the end).
  bool pair = true;
  int chunk2 = 32;                  // x planes per fdtd_pair tile
  // fewest fdtd_pair tiles for which a solve uses the two-step region (0:
  // any); single solves need enough to fill the GPU, server gangs pool them
  long long min_pair_tiles = 0;
  long long pair_cells = 0;         // cells of the two-step region (0: none)
  // owned cells of tiles that carry a current (see kTileNoCurrent): of the
  // one-step tiling, of the one-step tiles outside the region, of the
  // two-step tiles
  long long jc_single = 0, jc_outer = 0, jc_pair = 0;
  std::vector<uint32_t> pair_host;
  int outer_off[kNumLists + 1] = {};
  // The one-step lists of one step write disjoint cells of the next set and
  // read only the current one, so they may run at once: up to kMaxAux
  // auxiliary streams take lists beside `stream` (fork and join events keep
  // every other launch ordered on `stream`).
  static constexpr int kMaxAux = 3;
  int list_streams = 2;             // auxiliary streams used (0: every list on `stream`)
  // 1: the one-step lists of a step run as one launch_steps on `stream`
  // (when interior and CPML tiles have the same rows) instead of one launch
  // per list
  int fuse_lists = 0;
  hipStream_t aux[kMaxAux] = {};
  hipEvent_t fork_ev = nullptr, join_ev[kMaxAux] = {};
  bool timing = false;
  std::vector<hipEvent_t> ev;       // pairs around the step kernels of the last step() call
  std::vector<char> ev_pair;        // per pair: 1 if it brackets an fdtd_pair launch
  int timed_steps = 0;
  hipEvent_t span[2] = {nullptr, nullptr};  // first and last work of the last step() call
  double span_ms = 0;               // all GPU work of that call, per step
  double kernel_ms = 0;             // all field-update kernels, per step
  double pair_ms = 0;               // mean fdtd_pair launch
  int pair_launches = 0;
  virtual ~SolverBase();
  // Plan the device block: builds the tile and NTFF lists and returns the
  // block's size (no device call).
  virtual size_t layout() = 0;
  // Decode every tile of the layout and check it (msc_layout_check).
  virtual int check_layout(msc_layout_report* rep) const = 0;
  // Mean duration of `reps` launches of pair_pattern over the region's tiles
  virtual int pattern_probe(int reps, double* ms) = 0;
  // Allocate (arena or hipMalloc), clear and fill the block on `stream`.
  virtual int init(const double* wp, const double* nu) = 0;
  virtual int step(int n) = 0;
  virtual int get_field(int comp, double* out) = 0;
  // get_field through the device's fixed read-back scratch (no allocation)
  virtual int get_field_noalloc(int comp, double* out) = 0;
  int device = 0;                   // the device the solver was created on
  double* scratch_dev = nullptr;    // that device's read-back scratch
  // The NTF payload (source DFT, then the 24 face arrays) after `steps`
  // steps: the source DFT is summed through that many steps, the face sums
  // are what the device holds.
  virtual int get_ntff(double* out, int steps) = 0;
  // Contract the NTFF samples of the current partial K-step block now
  // (asynchronous, on `stream`; nothing to do when the last step closed a
  // block or the run is complete)
  virtual int flush_ntff() = 0;
  virtual int get_probes(double* out) = 0;
  // The non-finite watchdog flag (1: a NaN or infinite value reached the
  // NTFF surface), read after the work queued so far; and its device word.
  virtual int nonfinite(int* flag) = 0;
  virtual int* health_word() = 0;
  virtual int get_frame(int f, double* e, double* h, double* emax) = 0;
  int sync();
  size_t frame_doubles() const {
    return 2 * ((size_t)plan.nx * plan.ny + (size_t)plan.nx * plan.nz + (size_t)plan.ny * plan.nz);
  }
  size_t ntff_doubles() const;
};

template <typename T>
struct Solver : SolverBase {
  // host copies of the device-resident argument blocks
  StepArgs<T> sa{};
  NtffArgs<T> na{};
  AuxArgs<T> xa{};
  StepArgs<T>* sa_dev = nullptr;
  NtffArgs<T>* na_dev = nullptr;
  AuxArgs<T>* xa_dev = nullptr;
  uint32_t* tiles = nullptr;
  uint32_t* tiles2 = nullptr;  // pair_host on the device
  uint2* ntff_blocks = nullptr;
  size_t acc_n = 0;
  double* acc = nullptr;
  double* probes = nullptr;
  double* eframes = nullptr;
  double* hframes = nullptr;
  unsigned long long* emax = nullptr;
  int* tdev = nullptr;
  int* health = nullptr;  // non-finite watchdog flag (ntff_sample)

  size_t layout() override;
  int check_layout(msc_layout_report* rep) const override;
  int pattern_probe(int reps, double* ms) override;
  int init(const double* wp, const double* nu) override;
  int step(int n) override;
  int get_field(int comp, double* out) override;
  int get_field_noalloc(int comp, double* out) override;
  int get_ntff(double* out, int steps) override;
  int flush_ntff() override;
  int get_probes(double* out) override;
  int get_frame(int f, double* e, double* h, double* emax) override;
  int nonfinite(int* flag) override;
  int* health_word() override { return health; }

 private:
  // CPML-aligned tile starts per axis, computed by layout()
  std::vector<int> xs_, ys_, ysb_, zs_;
  // two-step region bounds and fdtd_pair tile starts
  int x2lo = 0, x2hi = 0, y2lo = 0, y2hi = 0, z2lo = 0, z2hi = 0;
  std::vector<int> xs2_, ys2_, zs2_;
  size_t carve(char* base);
};

extern template struct Solver<float>;
extern template struct Solver<double>;

}  // namespace msc
