// Multi-solve server: many independent solves resident on one GPU, advanced
// together by ganged launches.
//
// Reference counterpart: setup/meteorbatch.m and setup/meteorrandom.m submit
// one PBS job per run (setup/meteorrun.m:154-183), each a separate process
// running the whole time loop of meteor-fork.cpp on its own.  Here the runs
// share one device:
//   - every solve owns one block of a fixed device arena (no hipMalloc or
//     hipFree while serving); requests are admitted first come, first served
//     when their block fits;
//   - a worker thread builds each request's FP64 plan and fills its block on
//     a set-up stream while the resident solves keep stepping;
//   - one gang iteration advances every resident solve by one time step with
//     one fused-step launch per current model and CPML variant, one NTFF
//     launch and one probe launch, whatever the number of solves; the tile
//     lists carry a solve id per workgroup, so small grids fill the GPU
//     together instead of one after another;
//   - a solve leaves the gang after its last step; its results are copied
//     out once that step has completed and its block returns to the arena.
// Each solve's arithmetic is exactly its single-solve arithmetic (the same
// kernels on the same per-solve argument block), so results do not depend
// on what else is resident.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "msc_solver.h"

namespace msc {

struct ServerBase {
  virtual ~ServerBase() = default;
  virtual int submit(const msc_params* p, const double* wp, const double* nu, int64_t* id) = 0;
  virtual int advance(int max_iter, int* ndone) = 0;
  virtual int info(int64_t id, msc_request_info* out) = 0;
  virtual int get_ntff(int64_t id, double* out, size_t n) = 0;
  virtual int get_probes(int64_t id, double* out, size_t n) = 0;
  virtual int release(int64_t id) = 0;
  virtual int stats(msc_server_stats* out) = 0;
  virtual double clock() const = 0;
  virtual int configure(int fill_tiles, int use_graphs) = 0;
  virtual int set_tilings(const msc_tiling* tilings, const int64_t* max_cells, int n) = 0;
  virtual int cancel(int64_t id) = 0;
  virtual int set_time_limit(double seconds) = 0;
};

template <typename T>
struct Server : ServerBase {
  struct Req {
    int64_t id = 0;
    msc_params p{};
    const double* wp = nullptr;  // borrowed from the caller until READY
    const double* nu = nullptr;
    std::unique_ptr<Solver<T>> s;
    int state = MSC_REQ_QUEUED;
    std::atomic<int> t{0};  // steps launched
    size_t bytes = 0;
    hipEvent_t done = nullptr;
    double t_submit = 0, t_start = 0, t_done = 0;
    std::vector<double> ntff, probes;
    std::string err;
    // evict at the next chance (cancelled, over the time limit or flagged
    // non-finite), ending in end_state with `err` as the reason
    bool cancel = false;
    int end_state = MSC_REQ_CANCELLED;
  };
  // device copies of the gang's argument blocks and launch lists; two sets
  // so a rebuild never overwrites lists that queued launches still read
  struct Gang {
    StepArgs<T>* sa = nullptr;
    NtffArgs<T>* na = nullptr;
    AuxArgs<T>* xa = nullptr;
    uint32_t* tiles = nullptr;
    uint16_t* sids = nullptr;
    uint2* nb = nullptr;
    size_t cap_solves = 0, cap_tiles = 0, cap_nb = 0;
    // host staging (pinned)
    char* host = nullptr;
    size_t host_bytes = 0;
    hipEvent_t last_use = nullptr;
    bool used = false;
    // one captured gang iteration: [0] step t of every solve (one-step
    // tiles, two-step tiles, NTFF samples, probes/tick), [1] step t+1 of
    // the two-step solves outside their regions
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
  };

  int device = 0;
  // Every gang launch goes to `run`, one after another (a captured graph is
  // a single chain).  Two streams in all: `run` and the worker's `setup`.
  hipStream_t run = nullptr, setup = nullptr;
  Arena arena;
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

  std::mutex mu;
  std::condition_variable cv_work, cv_ready;
  std::deque<Req*> queue;    // submitted, not yet prepared (FIFO)
  std::vector<Req*> ready;   // prepared, waiting to join the gang
  std::vector<Req*> active;  // resident and not finished
  std::vector<Req*> sel;     // the subset the current gang lists step
  // one-step tile rows a gang launch can carry (the kernel's TY variants)
  static constexpr int kTys[3] = {3, 7, 15};
  static int ty_index(int ty) { return ty == 3 ? 0 : ty == 7 ? 1 : 2; }
  // tilings by size class: (max cells, tiling); empty: the defaults
  std::vector<std::pair<long long, msc_tiling>> tilings;
  int fill_tiles = 32768;   // ~40 rounds of 3 workgroups on each of the 256 CUs
  bool use_graphs = true;     // replay each gang half as a captured graph
  std::vector<Req*> draining;  // all steps launched, waiting for completion
  std::vector<Req*> evicted;   // left the gang unfinished, waiting for its queued launches
  double time_limit = 0;       // seconds a request may run once in the gang (0: no limit)
  // Non-finite watchdog: every kHealthEvery iterations the flag words of
  // the running solves (set by ntff_sample) are copied to pinned host slots
  // on the run stream; once that copy has completed, flagged solves are
  // evicted as failed.
  static constexpr int kHealthEvery = 16;
  static constexpr size_t kHealthSlots = 65536;
  int* health_host = nullptr;
  hipEvent_t health_ev = nullptr;
  std::vector<int64_t> health_ids;
  bool health_pending = false;
  std::map<int64_t, std::unique_ptr<Req>> reqs;
  int64_t next_id = 1;
  bool stop = false;
  bool worker_busy = false;
  std::thread worker;

  Gang gang[2];
  int cur_gang = 0;
  bool dirty = true;
  // launch lists of the current gang.  The two-step solves come first in
  // `sel` (n_pair of them): per current model and one-step variant, the
  // list holds their tiles outside the two-step region and then the other
  // solves' tiles (n_tiles), so its first n_tiles_b entries are step t+1 of
  // the two-step solves; pair tiles per current model; NTFF blocks of the
  // two-step solves first (n_nb_b of them).
  int n_tiles[3][kNumLists][3] = {};    // [current model][variant][tile rows 3, 7, 15]
  int n_tiles_b[3][kNumLists][3] = {};
  size_t tile_off[3][kNumLists][3] = {};
  int n_ptiles[3] = {};
  size_t ptile_off[3] = {};
  int n_nb = 0, n_nb_b = 0;
  size_t n_pair = 0;
  // The gang's one-step lists run as one launch per current model and tile
  // rows (fdtd_steps) when its one-step cells fill less than 64 MB of one
  // field array (16 M cells in FP32).  Gang launches run one after another,
  // so the lists gain more from one launch than a single solve's, whose
  // lists overlap on three streams: gangs of 0.5-33 M cells ran 1.1-2.3x
  // faster, 65-134 M within 1%, 180 M 4% and 422 M (the 48-trail sweep) 9%
  // slower.
  // MSC_SERVER_FUSE=0/1 decides instead.
  int fuse_env = -1;
  bool fuse_gang = false;
  // The gang runs two-step tiles only when its solves' regions give at
  // least pair_min fdtd_pair workgroups together (1.5 per CU, the single-
  // solve rule): a gang of a few small solves otherwise spends each
  // iteration in a handful of long two-step marches.  MSC_SERVER_PAIR_MIN
  // sets the count (0: a gang pairs whenever a solve has a region).
  long long pair_min = 0;
  bool gang_pairs = true;
  // MSC_SERVER_TIMING=1: launches go out directly (no graphs) with HIP
  // events between the parts of each half, every iteration waits for its
  // events, and the destructor prints where the GPU time went (to stderr).
  // Serialises host and device: for diagnosis, not for serving.
  bool timing = false;
  hipEvent_t tev[2][5] = {};
  double tsum[2][4] = {};  // [half][one-step lists, two-step launch, NTFF samples + probes, contractions]
  double tcells[2][2] = {};  // [half][cell-updates of the one-step lists, of the two-step launch]
  long long thalves[2] = {};
  double gang_cells[3] = {};  // the current gang's one-step cells of half 0, of half 1, and region cells
  // run-ahead limit: at most kDepth iterations queued on the run stream
  static constexpr int kDepth = 4;
  hipEvent_t iter_ev[kDepth] = {};
  long long iters = 0;
  long long steps_done = 0;  // solve-steps launched

  double now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  double clock() const override { return now(); }
  int configure(int fill, int graphs) override {
    fill_tiles = fill;
    use_graphs = graphs != 0 && !timing;  // timed launches go out directly
    dirty = true;
    return MSC_OK;
  }
  int set_tilings(const msc_tiling* t, const int64_t* max_cells, int n) override {
    if (n < 0 || (n > 0 && (!t || !max_cells))) return fail(MSC_EINVAL, "bad tiling table");
    std::vector<std::pair<long long, msc_tiling>> v;
    for (int c = 0; c < n; c++) {
      const msc_tiling& x = t[c];
      auto tyok = [](int r) { return r == 3 || r == 7 || r == 15; };
      if (!tyok(x.ty) || !tyok(x.ty_b) || x.chunk < 1 || x.chunk2 < 1)
        return fail(MSC_EINVAL, "tiling class %d: rows must be 3, 7 or 15 and chunks >= 1", c);
      v.push_back({(long long)max_cells[c], x});
    }
    std::lock_guard<std::mutex> g(mu);
    if (!queue.empty() || !ready.empty() || !active.empty() || !draining.empty() || worker_busy)
      return fail(MSC_ESTATE, "tilings change only while the server holds no request");
    tilings = std::move(v);
    return MSC_OK;
  }

  int init(size_t budget) {
    HIPCHK(hipGetDevice(&device));
    HIPCHK(hipStreamCreateWithFlags(&run, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&setup, hipStreamNonBlocking));
    for (auto& e : iter_ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&health_ev, hipEventDisableTiming));
    HIPCHK(hipHostMalloc((void**)&health_host, kHealthSlots * sizeof(int)));
    for (auto& g : gang) HIPCHK(hipEventCreateWithFlags(&g.last_use, hipEventDisableTiming));
    if (budget == 0) {
      size_t fr = 0, tot = 0;
      HIPCHK(hipMemGetInfo(&fr, &tot));
      budget = fr / 10 * 9;
    }
    int rc = arena.init(budget);
    if (rc) return rc;
    if (const char* f = std::getenv("MSC_FILL_TILES")) fill_tiles = std::atoi(f);
    if (const char* f = std::getenv("MSC_GRAPHS")) use_graphs = std::atoi(f) != 0;
    if (const char* f = std::getenv("MSC_SERVER_FUSE")) fuse_env = std::atoi(f) != 0;
    {
      int cus = 0;
      HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
      pair_min = 3LL * cus / 2;
    }
    if (const char* f = std::getenv("MSC_SERVER_PAIR_MIN")) pair_min = std::max(0LL, std::atoll(f));
    if (const char* f = std::getenv("MSC_SERVER_TIMING")) timing = std::atoi(f) != 0;
    if (timing) {
      use_graphs = false;
      for (auto& hv : tev)
        for (auto& e : hv) HIPCHK(hipEventCreate(&e));
    }
    worker = std::thread([this] { work(); });
    return MSC_OK;
  }

  ~Server() override {
    if (timing && thalves[0] > 0) {
      static const char* kPart[4] = {"one-step lists", "two-step launch", "NTFF samples + probes", "contractions"};
      double tot = 0;
      for (auto& hv : tsum)
        for (double v : hv) tot += v;
      std::fprintf(stderr, "server timing: %lld iterations (%lld with a second half), %.1f ms of GPU time\n",
                   thalves[0], thalves[1], tot);
      for (int h = 0; h < 2; h++)
        for (int k = 0; k < 4; k++)
          if (tsum[h][k] > 0) {
            std::fprintf(stderr, "  half %d %-22s %10.1f ms  %5.1f%%  %8.1f us per half", h, kPart[k], tsum[h][k],
                         100 * tsum[h][k] / tot, 1e3 * tsum[h][k] / thalves[h]);
            if (k < 2 && tcells[h][k] > 0)
              std::fprintf(stderr, "  %6.1f Gcell-updates/s", tcells[h][k] / (tsum[h][k] * 1e6));
            std::fprintf(stderr, "\n");
          }
    }
    for (auto& hv : tev)
      for (auto& e : hv)
        if (e) hipEventDestroy(e);
    {
      std::lock_guard<std::mutex> g(mu);
      stop = true;
    }
    cv_work.notify_all();
    if (worker.joinable()) worker.join();
    if (run) hipStreamSynchronize(run);
    if (setup) hipStreamSynchronize(setup);
    for (auto& kv : reqs) {
      if (kv.second->done) hipEventDestroy(kv.second->done);
      kv.second->s.reset();
    }
    for (auto& g : gang) {
      hipFree(g.sa); hipFree(g.na); hipFree(g.xa); hipFree(g.tiles); hipFree(g.sids); hipFree(g.nb);
      if (g.host) hipHostFree(g.host);
      if (g.last_use) hipEventDestroy(g.last_use);
      for (int h = 0; h < 2; h++) {
        if (g.exec[h]) hipGraphExecDestroy(g.exec[h]);
        if (g.graph[h]) hipGraphDestroy(g.graph[h]);
      }
    }
    for (auto& e : iter_ev)
      if (e) hipEventDestroy(e);
    if (health_ev) hipEventDestroy(health_ev);
    if (health_host) hipHostFree(health_host);
    if (run) hipStreamDestroy(run);
    if (setup) hipStreamDestroy(setup);
  }

  // ---- worker: FIFO preparation of queued requests
  void work() {
    hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      cv_work.wait(lk, [&] { return stop || !queue.empty(); });
      if (stop) return;
      Req* r = queue.front();
      if (r->cancel) {  // cancelled while queued (also while waiting for space)
        queue.pop_front();
        r->s.reset();
        end_unrun(r);
        continue;
      }
      worker_busy = true;
      lk.unlock();
      int rc = prepare(r);
      lk.lock();
      worker_busy = false;
      if (rc == MSC_EBUSY) {
        // no free range yet: wait for a retirement (or shutdown), keep FIFO order
        cv_work.wait_for(lk, std::chrono::milliseconds(50));
        continue;
      }
      queue.pop_front();
      if (rc == MSC_OK && r->cancel) {
        r->s.reset();  // its block back (its set-up has completed)
        end_unrun(r);
      } else if (rc == MSC_OK) {
        r->state = MSC_REQ_READY;
        ready.push_back(r);
      } else {
        r->state = MSC_REQ_FAILED;
        r->err = msc_last_error();
        r->t_done = now();
        r->wp = r->nu = nullptr;
      }
      cv_ready.notify_all();
    }
  }

  // A request that never joined the gang ends in its end state (mu held).
  void end_unrun(Req* r) {
    r->state = r->end_state;
    r->t_done = now();
    r->wp = r->nu = nullptr;
  }

  int cancel(int64_t id) override {
    std::lock_guard<std::mutex> g(mu);
    Req* r = find(id);
    if (!r) return fail(MSC_EINVAL, "unknown request %lld", (long long)id);
    if (r->state >= MSC_REQ_DRAINING)
      return fail(MSC_ESTATE, "request %lld has no steps left to cancel", (long long)id);
    if (!r->cancel) {
      r->cancel = true;
      r->end_state = MSC_REQ_CANCELLED;
      r->err = "cancelled";
    }
    cv_work.notify_all();
    return MSC_OK;
  }

  int set_time_limit(double seconds) override {
    if (!(seconds >= 0)) return fail(MSC_EINVAL, "time limit must be >= 0 s (got %g)", seconds);
    std::lock_guard<std::mutex> g(mu);
    time_limit = seconds;
    return MSC_OK;
  }

  int prepare(Req* r) {
    if (!r->s) {
      r->s.reset(new Solver<T>());
      r->s->tiling_from_env = false;  // the size class's tiling (the defaults without a table)
      r->s->pair = true;              // two-step gangs (FP64 tiles are 16 x 32)
      r->s->zlanes = kServerZLanes;
      msc_tiling t;
      msc_tiling_default(&t);
      if (!tilings.empty()) {
        const long long n = (long long)r->p.nx * r->p.ny * r->p.nz;
        t = tilings.back().second;
        for (const auto& c : tilings)
          if (n <= c.first) {
            t = c.second;
            break;
          }
      }
      r->s->ty = t.ty;
      r->s->ty_b = t.ty_b;
      r->s->chunk = t.chunk;
      r->s->chunk2 = t.chunk2;
      int rc = make_plan(&r->p, r->s->plan);
      if (rc) return rc;
      r->bytes = r->s->layout();
      if (r->bytes > arena.capacity())
        return fail(MSC_ENOMEM, "solve needs %zu device bytes, server arena holds %zu", r->bytes, arena.capacity());
      r->s->arena = &arena;
      r->s->stream = setup;
      r->s->own_stream = false;
    }
    int rc = r->s->init(r->wp, r->nu);
    if (rc == MSC_EBUSY) return rc;
    if (rc) {
      r->s.reset();
      return rc;
    }
    r->wp = r->nu = nullptr;
    return MSC_OK;
  }

  int submit(const msc_params* p, const double* wp, const double* nu, int64_t* id) override {
    int rc = msc_params_validate(p);
    if (rc) return rc;
    if (p->savefields) return fail(MSC_EINVAL, "server solves produce NTFF and probe results only: set savefields = 0");
    if (!wp || !nu) return fail(MSC_EINVAL, "NULL media");
    if ((rc = check_media(p, wp, nu))) return rc;
    std::unique_ptr<Req> r(new Req());
    r->p = *p;
    r->wp = wp;
    r->nu = nu;
    r->t_submit = now();
    std::lock_guard<std::mutex> g(mu);
    r->id = next_id++;
    *id = r->id;
    queue.push_back(r.get());
    reqs[r->id] = std::move(r);
    cv_work.notify_all();
    return MSC_OK;
  }

  // ---- gang lists
  // grow a device list (rare: capacities double); the set is idle here
  template <typename U>
  int regrow(U*& ptr, size_t n) {
    if (ptr) hipFree(ptr);
    ptr = nullptr;
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, n * sizeof(U));
    if (e != hipSuccess) return fail(MSC_ENOMEM, "gang list of %zu entries: %s", n, hipGetErrorString(e));
    ptr = (U*)q;
    return MSC_OK;
  }

  int ensure(Gang& g, size_t nsol, size_t ntiles, size_t nnb) {
    int rc;
    if (nsol > g.cap_solves) {
      const size_t c = std::max(nsol, 2 * g.cap_solves);
      if ((rc = regrow(g.sa, c)) || (rc = regrow(g.na, c)) || (rc = regrow(g.xa, c))) return rc;
      g.cap_solves = c;
    }
    if (ntiles > g.cap_tiles) {
      const size_t c = std::max(ntiles, 2 * g.cap_tiles);
      if ((rc = regrow(g.tiles, c)) || (rc = regrow(g.sids, c))) return rc;
      g.cap_tiles = c;
    }
    if (nnb > g.cap_nb) {
      const size_t c = std::max(nnb, 2 * g.cap_nb);
      if ((rc = regrow(g.nb, c))) return rc;
      g.cap_nb = c;
    }
    const size_t hb = nsol * (sizeof(StepArgs<T>) + sizeof(NtffArgs<T>) + sizeof(AuxArgs<T>)) + ntiles * 6 +
                      nnb * sizeof(uint2) + 1024;
    if (hb > g.host_bytes) {
      if (g.host) hipHostFree(g.host);
      g.host = nullptr;
      void* q = nullptr;
      HIPCHK(hipHostMalloc(&q, hb * 2));
      g.host = (char*)q;
      g.host_bytes = hb * 2;
    }
    return MSC_OK;
  }

  // Which resident solves the gang steps: the oldest always (no request
  // starves), then the others by least remaining work (cells x steps left),
  // until their tiles fill the device; the rest wait resident.  The GPU
  // stays full either way, and short requests are not slowed to the pace
  // of the largest grid in the mix.  fill_tiles <= 0 steps every solve.
  void select() {
    sel = active;
    if (fill_tiles <= 0 || sel.size() <= 1) return;
    auto left = [](const Req* r) {
      const HostPlan& h = r->s->plan;
      return (double)h.nx * h.ny * h.nz * (double)(h.T - r->t);
    };
    auto oldest = std::min_element(sel.begin(), sel.end(), [](const Req* a, const Req* b) { return a->id < b->id; });
    std::iter_swap(sel.begin(), oldest);
    std::sort(sel.begin() + 1, sel.end(), [&](const Req* a, const Req* b) {
      const double la = left(a), lb = left(b);
      return la != lb ? la < lb : a->id < b->id;
    });
    // a solve's workgroups in one iteration: its two-step launch list when
    // it runs two-step tiles (outer one-step tiles, then the region's), the
    // one-step tiling otherwise (the count rebuild() sizes the gang by)
    size_t tiles = 0, q = 0;
    for (; q < sel.size() && (q == 0 || tiles < (size_t)fill_tiles); q++)
      tiles += pairs(sel[q]) ? sel[q]->s->pair_host.size() : sel[q]->s->tiles_host.size();
    sel.resize(q);
  }

  int rebuild() {
    const int gi = cur_gang ^ 1;  // the set the queued launches are not reading
    Gang& g = gang[gi];
    if (g.used) HIPCHK(hipEventSynchronize(g.last_use));
    gang_pairs = true;
    select();
    long long region_tiles = 0;
    for (const Req* r : sel)
      if (has_pair(r)) region_tiles += (long long)r->s->pair_host.size() - r->s->outer_off[kNumLists];
    gang_pairs = region_tiles >= pair_min;
    // two-step solves (a region, two or more steps left) first
    std::stable_partition(sel.begin(), sel.end(), [this](const Req* r) { return pairs(r); });
    n_pair = (size_t)std::count_if(sel.begin(), sel.end(), [this](const Req* r) { return pairs(r); });
    const std::vector<Req*>& active = sel;
    const size_t ns = active.size();
    if (ns > 65535) return fail(MSC_EINVAL, "more than 65535 resident solves");
    size_t nt = 0, nnb = 0;
    double cells1 = 0;  // cells the one-step lists of half 0 update
    gang_cells[1] = gang_cells[2] = 0;
    for (Req* r : active) {
      nt += pairs(r) ? r->s->pair_host.size() : r->s->tiles_host.size();
      nnb += r->s->ntff_host.size();
      const HostPlan& hp = r->s->plan;
      const double n = (double)hp.nx * hp.ny * hp.nz, nr = pairs(r) ? (double)r->s->pair_cells : 0.0;
      cells1 += n - nr;
      if (pairs(r)) {
        gang_cells[1] += n - nr;
        gang_cells[2] += nr;
      }
    }
    gang_cells[0] = cells1;
    fuse_gang = (fuse_env >= 0 ? fuse_env != 0 : cells1 * sizeof(T) < 64e6) && nt < (size_t)INT32_MAX;
    int rc = ensure(g, std::max<size_t>(ns, 1), std::max<size_t>(nt, 1), std::max<size_t>(nnb, 1));
    if (rc) return rc;
    // host image: args, then tiles/sids grouped by (current model, variant), then NTFF blocks
    char* h = g.host;
    StepArgs<T>* hsa = (StepArgs<T>*)h;
    NtffArgs<T>* hna = (NtffArgs<T>*)(hsa + ns);
    AuxArgs<T>* hxa = (AuxArgs<T>*)(hna + ns);
    uint32_t* htiles = (uint32_t*)(hxa + ns);
    uint16_t* hsids = (uint16_t*)(htiles + nt);
    uint2* hnb = (uint2*)(((uintptr_t)(hsids + nt) + 7) & ~(uintptr_t)7);
    for (size_t q = 0; q < ns; q++) {
      hsa[q] = active[q]->s->sa;
      hna[q] = active[q]->s->na;
      hxa[q] = active[q]->s->xa;
    }
    size_t o = 0;
    auto put = [&](const std::vector<uint32_t>& v, size_t b, size_t e, size_t q) {
      for (size_t k = b; k < e; k++) {
        htiles[o] = v[k];
        hsids[o] = (uint16_t)q;
        ++o;
      }
    };
    auto rows = [](const Solver<T>& s, int v) { return ty_index(v == kListInterior ? s.ty : s.ty_b); };
    for (int c = 0; c < 3; c++) {
      for (int v = 0; v < kNumLists; v++)
        for (int y = 0; y < 3; y++) {
          tile_off[c][v][y] = o;
          for (size_t q = 0; q < n_pair; q++) {  // two-step solves: one-step tiles outside the region
            const Solver<T>& s = *active[q]->s;
            if (s.cur == c && rows(s, v) == y) put(s.pair_host, s.outer_off[v], s.outer_off[v + 1], q);
          }
          n_tiles_b[c][v][y] = (int)(o - tile_off[c][v][y]);
          for (size_t q = n_pair; q < ns; q++) {  // the others: every tile
            const Solver<T>& s = *active[q]->s;
            if (s.cur == c && rows(s, v) == y) put(s.tiles_host, s.list_off[v], s.list_off[v + 1], q);
          }
          n_tiles[c][v][y] = (int)(o - tile_off[c][v][y]);
        }
      ptile_off[c] = o;
      for (size_t q = 0; q < n_pair; q++) {
        const Solver<T>& s = *active[q]->s;
        if (s.cur == c) put(s.pair_host, s.outer_off[kNumLists], s.pair_host.size(), q);
      }
      n_ptiles[c] = (int)(o - ptile_off[c]);
    }
    size_t m = 0;
    for (size_t q = 0; q < ns; q++) {
      for (const uint2& d : active[q]->s->ntff_host) hnb[m++] = make_uint2((uint32_t)q, d.y);
      if (q + 1 == n_pair) n_nb_b = (int)m;
    }
    if (n_pair == 0) n_nb_b = 0;
    n_nb = (int)m;
    HIPCHK(hipMemcpyAsync(g.sa, hsa, ns * sizeof(StepArgs<T>), hipMemcpyHostToDevice, run));
    HIPCHK(hipMemcpyAsync(g.na, hna, ns * sizeof(NtffArgs<T>), hipMemcpyHostToDevice, run));
    HIPCHK(hipMemcpyAsync(g.xa, hxa, ns * sizeof(AuxArgs<T>), hipMemcpyHostToDevice, run));
    if (nt) {
      HIPCHK(hipMemcpyAsync(g.tiles, htiles, nt * 4, hipMemcpyHostToDevice, run));
      HIPCHK(hipMemcpyAsync(g.sids, hsids, nt * 2, hipMemcpyHostToDevice, run));
    }
    if (m) HIPCHK(hipMemcpyAsync(g.nb, hnb, m * sizeof(uint2), hipMemcpyHostToDevice, run));
    cur_gang = gi;
    g.used = true;
    dirty = false;
    return use_graphs ? capture(g) : MSC_OK;
  }

  // A solve steps two at a time when it has a two-step region and at least
  // two steps left (the server's solves write no frames or slices).
  static bool has_pair(const Req* r) { return !r->s->pair_host.empty() && r->t + 2 <= r->s->plan.T; }
  bool pairs(const Req* r) const { return gang_pairs && has_pair(r); }

  // The fixed part of an iteration, in two halves.  Half 0: step t of every
  // solve (the one-step launches per current model and variant, the
  // two-step launches, the NTFF samples, the probe/tick launch).  Half 1:
  // step t+1 of the two-step solves outside their regions, their NTFF
  // samples and probes/tick.  The kernels read each solve's step from
  // device memory, so the same launch sequences (or their captured graphs)
  // serve every iteration until the gang changes.
  hipError_t launch_fixed(const Gang& g, int half) {
    hipError_t e = hipSuccess;
    hipEvent_t* te = timing ? tev[half] : nullptr;
    if (te) e = hipEventRecord(te[0], run);
    for (int c = 0; c < 3 && e == hipSuccess; c++) {
      if (fuse_gang) {
        for (int y = 0; y < 3 && e == hipSuccess; y++) {
          int first[kNumLists], count[kNumLists];
          for (int v = 0; v < kNumLists; v++) {
            first[v] = (int)tile_off[c][v][y];
            count[v] = half ? n_tiles_b[c][v][y] : n_tiles[c][v][y];
          }
          e = launch_steps<T>(g.sa, g.sids, kTys[y], c, false, kServerZLanes, g.tiles, first, count, run);
        }
      } else {
        for (int v = 0; v < kNumLists && e == hipSuccess; v++)
          for (int y = 0; y < 3 && e == hipSuccess; y++) {
            const int n = half ? n_tiles_b[c][v][y] : n_tiles[c][v][y];
            if (!n) continue;
            const size_t o = tile_off[c][v][y];
            e = launch_step<T>(g.sa, g.sids + o, kTys[y], c, list_pm(v), false, kServerZLanes, g.tiles + o, n, run);
          }
      }
    }
    if (te && e == hipSuccess) e = hipEventRecord(te[1], run);
    for (int c = 0; c < 3 && e == hipSuccess; c++)
      if (!half && n_ptiles[c])
        e = launch_pair<T>(g.sa, g.sids + ptile_off[c], c, g.tiles + ptile_off[c], n_ptiles[c], run);
    if (te && e == hipSuccess) e = hipEventRecord(te[2], run);
    if (e == hipSuccess) e = launch_ntff<T>(g.na, g.nb, half ? n_nb_b : n_nb, run);
    if (e == hipSuccess) e = launch_probes<T>(g.xa, half ? (int)n_pair : (int)sel.size(), run);
    if (te && e == hipSuccess) e = hipEventRecord(te[3], run);
    return e;
  }

  int capture(Gang& g) {
    for (int h = 0; h < 2; h++) {
      if (g.exec[h]) (void)hipGraphExecDestroy(g.exec[h]);
      if (g.graph[h]) (void)hipGraphDestroy(g.graph[h]);
      g.exec[h] = nullptr;
      g.graph[h] = nullptr;
    }
    for (int h = 0; h < (n_pair ? 2 : 1); h++) {
      HIPCHK(hipStreamBeginCapture(run, hipStreamCaptureModeThreadLocal));
      const hipError_t e = launch_fixed(g, h);
      hipGraph_t gr = nullptr;
      const hipError_t e2 = hipStreamEndCapture(run, &gr);
      if (e != hipSuccess || e2 != hipSuccess) {
        if (gr) (void)hipGraphDestroy(gr);
        return fail(MSC_EHIP, "gang capture: %s", hipGetErrorString(e != hipSuccess ? e : e2));
      }
      g.graph[h] = gr;
      HIPCHK(hipGraphInstantiate(&g.exec[h], gr, nullptr, nullptr, 0));
    }
    return MSC_OK;
  }

  // completed solves: copy results out, free their blocks
  int retire(int* ndone) {
    for (size_t q = 0; q < draining.size();) {
      Req* r = draining[q];
      hipError_t e = hipEventQuery(r->done);
      if (e == hipErrorNotReady) {
        ++q;
        continue;
      }
      if (e != hipSuccess) return fail(MSC_EHIP, "solve %lld: %s", (long long)r->id, hipGetErrorString(e));
      r->ntff.resize(r->s->ntff_doubles());
      int rc = r->s->get_ntff(r->ntff.data(), r->t);
      if (rc) return rc;
      r->probes.resize((size_t)6 * r->s->plan.T * r->p.nprobes);
      if (!r->probes.empty() && (rc = r->s->get_probes(r->probes.data()))) return rc;
      int bad = 0;
      if ((rc = r->s->nonfinite(&bad))) return rc;
      r->s.reset();  // block back to the arena
      hipEventDestroy(r->done);
      r->done = nullptr;
      {
        std::lock_guard<std::mutex> g(mu);
        r->state = bad ? MSC_REQ_FAILED : MSC_REQ_DONE;
        if (bad) r->err = "non-finite fields (NaN or infinity) reached the NTFF surface";
        r->t_done = now();
      }
      cv_work.notify_all();
      draining.erase(draining.begin() + q);
      if (ndone) ++*ndone;
    }
    for (size_t q = 0; q < evicted.size();) {
      Req* r = evicted[q];
      hipError_t e = hipEventQuery(r->done);
      if (e == hipErrorNotReady) {
        ++q;
        continue;
      }
      if (e != hipSuccess) return fail(MSC_EHIP, "solve %lld: %s", (long long)r->id, hipGetErrorString(e));
      r->s.reset();  // its queued launches are complete: block back to the arena
      hipEventDestroy(r->done);
      r->done = nullptr;
      {
        std::lock_guard<std::mutex> g(mu);
        r->state = r->end_state;
        r->t_done = now();
      }
      cv_work.notify_all();
      evicted.erase(evicted.begin() + q);
    }
    return MSC_OK;
  }

  // Running solves that were cancelled, ran past the time limit or were
  // flagged non-finite leave the gang; their blocks return in retire() once
  // the launches already queued for them have completed.
  int evict() {
    const double tnow = now();
    for (size_t q = 0; q < active.size();) {
      Req* r = active[q];
      {
        std::lock_guard<std::mutex> lk(mu);
        if (!r->cancel && time_limit > 0 && tnow - r->t_start > time_limit) {
          r->cancel = true;
          r->end_state = MSC_REQ_FAILED;
          char buf[128];
          std::snprintf(buf, sizeof buf, "time limit of %g s exceeded at step %d", time_limit, r->t.load());
          r->err = buf;
        }
        if (!r->cancel) {
          ++q;
          continue;
        }
      }
      HIPCHK(hipEventCreateWithFlags(&r->done, hipEventDisableTiming));
      HIPCHK(hipEventRecord(r->done, run));
      evicted.push_back(r);
      active.erase(active.begin() + q);
      auto it = std::find(sel.begin(), sel.end(), r);
      if (it != sel.end()) sel.erase(it);
      dirty = true;
    }
    return MSC_OK;
  }

  // Watchdog: queue copies of the running solves' flag words, or read the
  // copies queued earlier once they have completed.
  int watch_health() {
    if (health_pending) {
      hipError_t e = hipEventQuery(health_ev);
      if (e == hipErrorNotReady) return MSC_OK;
      if (e != hipSuccess) return fail(MSC_EHIP, "watchdog copy: %s", hipGetErrorString(e));
      std::lock_guard<std::mutex> lk(mu);
      for (size_t q = 0; q < health_ids.size(); q++) {
        Req* r = find(health_ids[q]);
        if (!health_host[q] || !r || r->state != MSC_REQ_RUNNING || r->cancel) continue;
        r->cancel = true;
        r->end_state = MSC_REQ_FAILED;
        char buf[160];
        std::snprintf(buf, sizeof buf, "non-finite fields (NaN or infinity) reached the NTFF surface by step %d",
                      r->t.load());
        r->err = buf;
      }
      health_pending = false;
      return MSC_OK;
    }
    if (iters % kHealthEvery != 0 || active.empty()) return MSC_OK;
    health_ids.clear();
    for (Req* r : active) {
      if (health_ids.size() == kHealthSlots) break;
      HIPCHK(hipMemcpyAsync(&health_host[health_ids.size()], r->s->health_word(), sizeof(int),
                            hipMemcpyDeviceToHost, run));
      health_ids.push_back(r->id);
    }
    HIPCHK(hipEventRecord(health_ev, run));
    health_pending = true;
    return MSC_OK;
  }

  int launch_iteration() {
    const Gang& g = gang[cur_gang];
    const int slot = (int)(iters % kDepth);
    if (iters >= kDepth) HIPCHK(hipEventSynchronize(iter_ev[slot]));
    for (int h = 0; h < (n_pair ? 2 : 1); h++) {
      if (g.exec[h]) HIPCHK(hipGraphLaunch(g.exec[h], run));
      else HIPCHK(launch_fixed(g, h));
      // solves whose step (t, or t+1 in half 1) closes a sample block
      for (size_t q = 0; q < (h ? n_pair : sel.size()); q++) {
        const Solver<T>& s = *sel[q]->s;
        const int t = sel[q]->t + h;
        if ((t + 1) % s.na.K == 0 || t + 1 == s.plan.T)
          HIPCHK(launch_ntff_contract<T>(g.na + q, s.na.uoff[24], s.na.M, s.na.wofs[24], run, 0, s.na.acc32));
      }
      if (timing) {
        hipEvent_t* te = tev[h];
        HIPCHK(hipEventRecord(te[4], run));
        HIPCHK(hipEventSynchronize(te[4]));
        for (int k = 0; k < 4; k++) {
          float ms = 0;
          HIPCHK(hipEventElapsedTime(&ms, te[k], te[k + 1]));
          tsum[h][k] += ms;
        }
        tcells[h][0] += gang_cells[h];
        if (h == 0) tcells[h][1] += 2 * gang_cells[2];
        ++thalves[h];
      }
    }
    HIPCHK(hipEventRecord(iter_ev[slot], run));
    HIPCHK(hipEventRecord(gang[cur_gang].last_use, run));
    ++iters;
    steps_done += (long long)(sel.size() + n_pair);
    const std::vector<Req*> stepped = sel;  // positions [0, n_pair) stepped twice
    for (size_t q = 0; q < stepped.size(); q++) {
      Req* r = stepped[q];
      const bool two = q < n_pair;
      r->t += two ? 2 : 1;
      if (r->t < r->s->plan.T) {
        // a two-step solve with one step left leaves the two-step group
        if (two && r->t + 2 > r->s->plan.T) dirty = true;
        continue;
      }
      HIPCHK(hipEventCreateWithFlags(&r->done, hipEventDisableTiming));
      HIPCHK(hipEventRecord(r->done, run));
      {
        std::lock_guard<std::mutex> lk(mu);
        r->state = MSC_REQ_DRAINING;
      }
      draining.push_back(r);
      sel.erase(std::find(sel.begin(), sel.end(), r));
      active.erase(std::find(active.begin(), active.end(), r));
      dirty = true;
    }
    return MSC_OK;
  }

  int advance(int max_iter, int* ndone) override {
    if (ndone) *ndone = 0;
    hipSetDevice(device);
    int it = 0;
    while (it < max_iter) {
      int rc = retire(ndone);
      if (rc) return rc;
      {
        std::unique_lock<std::mutex> lk(mu);
        if (!ready.empty()) {
          for (Req* r : ready) {
            if (r->cancel) {  // cancelled before it joined: its set-up has completed
              r->s.reset();
              end_unrun(r);
              cv_work.notify_all();
              continue;
            }
            r->state = MSC_REQ_RUNNING;
            r->t_start = now();
            active.push_back(r);
          }
          ready.clear();
          dirty = true;
        }
        if (active.empty() && draining.empty() && evicted.empty()) {
          // idle, or only preparation pending: give the worker a moment and
          // return so the caller can submit or poll meanwhile
          if (!queue.empty()) cv_ready.wait_for(lk, std::chrono::milliseconds(2));
          if (ready.empty()) break;
          continue;
        }
      }
      if ((rc = evict())) return rc;
      if (active.empty()) {
        // wait for the draining and evicted solves' work
        HIPCHK(hipStreamSynchronize(run));
        continue;
      }
      if (dirty && (rc = rebuild())) return rc;
      if ((rc = launch_iteration())) return rc;
      if ((rc = watch_health())) return rc;
      ++it;
    }
    return retire(ndone);
  }

  Req* find(int64_t id) {
    auto it = reqs.find(id);
    return it == reqs.end() ? nullptr : it->second.get();
  }

  int info(int64_t id, msc_request_info* out) override {
    std::lock_guard<std::mutex> g(mu);
    Req* r = find(id);
    if (!r) return fail(MSC_EINVAL, "unknown request %lld", (long long)id);
    out->state = r->state;
    out->t = r->t;
    out->tsteps = r->p.tsteps;
    out->device_bytes = r->bytes;
    out->t_submit = r->t_submit;
    out->t_start = r->t_start;
    out->t_done = r->t_done;
    if (r->state == MSC_REQ_FAILED || r->state == MSC_REQ_CANCELLED) fail(MSC_EINVAL, "%s", r->err.c_str());
    return MSC_OK;
  }

  int get_ntff(int64_t id, double* out, size_t n) override {
    std::lock_guard<std::mutex> g(mu);
    Req* r = find(id);
    if (!r) return fail(MSC_EINVAL, "unknown request %lld", (long long)id);
    if (r->state != MSC_REQ_DONE) return fail(MSC_ESTATE, "request %lld not done", (long long)id);
    if (!out || n < r->ntff.size()) return fail(MSC_EINVAL, "output buffer too small (%zu < %zu)", n, r->ntff.size());
    std::memcpy(out, r->ntff.data(), r->ntff.size() * 8);
    return MSC_OK;
  }

  int get_probes(int64_t id, double* out, size_t n) override {
    std::lock_guard<std::mutex> g(mu);
    Req* r = find(id);
    if (!r) return fail(MSC_EINVAL, "unknown request %lld", (long long)id);
    if (r->state != MSC_REQ_DONE) return fail(MSC_ESTATE, "request %lld not done", (long long)id);
    if (n < r->probes.size() || (!out && !r->probes.empty())) return fail(MSC_EINVAL, "output buffer too small");
    if (!r->probes.empty()) std::memcpy(out, r->probes.data(), r->probes.size() * 8);
    return MSC_OK;
  }

  int release(int64_t id) override {
    std::lock_guard<std::mutex> g(mu);
    Req* r = find(id);
    if (!r) return fail(MSC_EINVAL, "unknown request %lld", (long long)id);
    if (r->state != MSC_REQ_DONE && r->state != MSC_REQ_FAILED && r->state != MSC_REQ_CANCELLED)
      return fail(MSC_ESTATE, "request %lld still in flight", (long long)id);
    reqs.erase(id);
    return MSC_OK;
  }

  int stats(msc_server_stats* out) override {
    std::lock_guard<std::mutex> g(mu);
    out->queued = (int)queue.size();
    out->ready = (int)ready.size();
    out->running = (int)active.size();
    out->draining = (int)(draining.size() + evicted.size());
    out->arena_bytes = arena.capacity();
    out->arena_used = arena.used();
    out->iterations = iters;
    out->solve_steps = steps_done;
    return MSC_OK;
  }
};

}  // namespace msc

using namespace msc;

struct msc_server {
  std::unique_ptr<ServerBase> impl;
};

#define SERVER_OK(s) \
  if (!(s) || !(s)->impl) return fail(MSC_EINVAL, "NULL server")

extern "C" {

size_t msc_solve_footprint(const msc_params* p, int precision) {
  if (!p || msc_params_validate(p) != MSC_OK) return 0;
  return solve_footprint(p, precision);
}

int msc_server_create(int precision, size_t budget_bytes, msc_server** out) {
  if (!out) return fail(MSC_EINVAL, "NULL argument");
  *out = nullptr;
  if (precision != 32 && precision != 64) return fail(MSC_EINVAL, "precision must be 32 or 64 (got %d)", precision);
  // Gang launches spread over auxiliary streams (parallel branches in the
  // captured graphs) were removed: two large-gang batch runs with them
  // stopped making progress, with no cause found.  The setting that chose
  // them is refused rather than ignored.
  if (const char* f = std::getenv("MSC_LIST_STREAMS"))
    if (std::atoi(f) != 0)
      return fail(MSC_EINVAL, "MSC_LIST_STREAMS=%s: the server launches on one stream only (unset it or set 0)", f);
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(MSC_EHIP, "no HIP device visible");
  std::unique_ptr<ServerBase> s;
  int rc;
  if (precision == 32) {
    auto* q = new Server<float>();
    s.reset(q);
    rc = q->init(budget_bytes);
  } else {
    auto* q = new Server<double>();
    s.reset(q);
    rc = q->init(budget_bytes);
  }
  if (rc) return rc;
  msc_server* h = new msc_server;
  h->impl = std::move(s);
  *out = h;
  return MSC_OK;
}

int msc_server_submit(msc_server* s, const msc_params* p, const double* wp, const double* nu, int64_t* id) {
  SERVER_OK(s);
  if (!p || !wp || !nu || !id) return fail(MSC_EINVAL, "NULL argument");
  return s->impl->submit(p, wp, nu, id);
}

int msc_server_advance(msc_server* s, int max_iterations, int* ndone) {
  SERVER_OK(s);
  if (max_iterations < 0) return fail(MSC_EINVAL, "max_iterations < 0");
  return s->impl->advance(max_iterations, ndone);
}

int msc_server_info(msc_server* s, int64_t id, msc_request_info* info) {
  SERVER_OK(s);
  if (!info) return fail(MSC_EINVAL, "NULL argument");
  return s->impl->info(id, info);
}

int msc_server_get_ntff(msc_server* s, int64_t id, double* out, size_t n) {
  SERVER_OK(s);
  return s->impl->get_ntff(id, out, n);
}

int msc_server_get_probes(msc_server* s, int64_t id, double* out, size_t n) {
  SERVER_OK(s);
  return s->impl->get_probes(id, out, n);
}

int msc_server_release(msc_server* s, int64_t id) {
  SERVER_OK(s);
  return s->impl->release(id);
}

int msc_server_get_stats(msc_server* s, msc_server_stats* out) {
  SERVER_OK(s);
  if (!out) return fail(MSC_EINVAL, "NULL argument");
  return s->impl->stats(out);
}

double msc_server_clock(const msc_server* s) { return (s && s->impl) ? s->impl->clock() : -1.0; }

int msc_server_cancel(msc_server* s, int64_t id) {
  SERVER_OK(s);
  return s->impl->cancel(id);
}

int msc_server_set_time_limit(msc_server* s, double seconds) {
  SERVER_OK(s);
  return s->impl->set_time_limit(seconds);
}

int msc_server_configure(msc_server* s, int fill_tiles, int use_graphs) {
  SERVER_OK(s);
  return s->impl->configure(fill_tiles, use_graphs);
}

int msc_server_set_tilings(msc_server* s, const msc_tiling* tilings, const int64_t* max_cells, int nclasses) {
  SERVER_OK(s);
  return s->impl->set_tilings(tilings, max_cells, nclasses);
}

void msc_server_destroy(msc_server* s) { delete s; }

}  // extern "C"
