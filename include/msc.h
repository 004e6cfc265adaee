/*
 * msc.h - C ABI of the MI355X meteor-trail FDTD time-stepping core.
 *
 * The reference solver (source/meteor-fork.cpp, built by source/makefile:6-9)
 * is one monolithic main() that reads four binary files from the working
 * directory, runs the time loop (meteor-fork.cpp:683-1667) and writes five
 * binary files plus log.txt.  This header splits that process into the entry
 * points a host program binds: read the run directory, create a solver (the
 * set-up part of main), advance it, and fetch the same payloads the reference
 * writes.  The `meteor` executable built from this library keeps the
 * reference's process contract (no argv, files in the cwd).
 *
 * Conventions: every function returns 0 on success and a negative MSC_E*
 * code on failure; msc_last_error() returns a thread-local message for the
 * last failure on the calling thread.  No C++ exception crosses this ABI.
 * Host buffers are owned by the caller; device memory is owned by the
 * solver.  A solver handle is used by one host thread at a time; distinct
 * handles may run concurrently (each owns its own HIP stream).
 *
 * Array layout: every 3-D array is C row-major [nx][ny][nz] (z fastest),
 * the reference's own layout (meteor-fork.cpp:248-283).
 */
#ifndef MSC_H
#define MSC_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MSC_MAX_PROBES 64
#define MSC_MAX_NTF 256

/* status codes */
#define MSC_OK 0
#define MSC_EINVAL -1   /* bad argument or inconsistent parameters */
#define MSC_EIO -2      /* missing or truncated file */
#define MSC_ENOMEM -3   /* device or host allocation failed */
#define MSC_EHIP -4     /* HIP runtime error */
#define MSC_ESTATE -5   /* call not valid in the handle's state */
#define MSC_EBUSY -6    /* server: no capacity right now */

/* field component ids for msc_solver_get_field */
enum {
  MSC_EX = 0, MSC_EY, MSC_EZ, MSC_HX, MSC_HY, MSC_HZ, MSC_JX, MSC_JY, MSC_JZ
};

/* source types: meteor-fork.cpp:160-179; 4 is an extension */
enum {
  MSC_SRC_SIN = 1, MSC_SRC_MODGAUSS = 2, MSC_SRC_RICKER = 3, MSC_SRC_GAUSSDERIV = 4
};

/*
 * Run parameters.  The first block is inputs.dat in file order
 * (reader meteor-fork.cpp:84-115, writer setup/meteorrun.m:45-69), then
 * B0.dat (meteor-fork.cpp:221-226).  The last block is this build's
 * optional extension record (absent from the reference; zero/default values
 * reproduce the reference exactly).
 */
typedef struct msc_params {
  int32_t dopml;
  int32_t savefields;
  double f0;
  double fsig;
  int32_t sourcetype;
  int32_t sourcepol;
  double gridfactor;
  double dt;
  int32_t tsteps;
  int32_t nx, ny, nz;
  double metsize[2];
  double wpmax;
  double numax;
  double metloc[3];
  double metangle;
  int32_t mettype;
  int32_t numfiles;
  int32_t nprobes;
  int32_t probex[MSC_MAX_PROBES];
  int32_t probey[MSC_MAX_PROBES];
  int32_t probez[MSC_MAX_PROBES];
  int32_t nntf;
  double fntf[MSC_MAX_NTF];
  int32_t doplate;
  double platesize;
  int32_t doB0;        /* fork only; meteorrun.m never writes it (EOF -> 0) */
  double B0[3];        /* tesla, B0.dat */
  /* ---- extension record ("MSCX" block after doB0 in inputs.dat) ---- */
  int32_t pmllen;      /* CPML thickness in cells; 0 -> 10 (meteor-fork.cpp:374) */
  int32_t media3d;     /* 1: wp.dat/nu.dat hold [nx][ny][nz], no z-taper */
  double nu_scale_height; /* >0: nu *= exp(-(z - metloc[2]) / H) */
  double src_tau;      /* sourcetype 4 width (s); 0 -> 1/(pi f0) */
  int32_t j_collocated; /* 1: the Lee-Kalluri current is driven by E at its own
                          edge; 0 (reference): by the node average of the two
                          neighbouring edges (meteor-fork.cpp:1577-1579), a
                          half-cell mismatch that makes the magnetized update
                          grow without bound */
} msc_params;

/* Fill `p` with the defaults of setup/meteordefaults.m (plus doB0=0). */
void msc_params_default(msc_params* p);

/* Validate a parameter set (sizes, cadence, probe indices, ...). */
int msc_params_validate(const msc_params* p);

/* Thread-local message for the last failing call on this thread. */
const char* msc_last_error(void);

/* Library version string. */
const char* msc_version(void);

/* ------------------------------------------------------------------ */
/* Run-directory I/O (the reference's file contract, Appendix layouts) */
/* ------------------------------------------------------------------ */

/*
 * Read inputs.dat, B0.dat, wp.dat and nu.dat from `dir`
 * (replaces meteor-fork.cpp:49-117, 215-226, 336-347).  On success *wp and
 * *nu point to malloc'd arrays of nx*ny doubles ([nx][ny], row-major, the
 * byte order the reference's fread gives) or nx*ny*nz when p->media3d;
 * release them with msc_free.  A missing trailing doB0 means 0 (the
 * reference leaves it indeterminate).
 */
int msc_read_run_dir(const char* dir, msc_params* p, double** wp, double** nu);

/* Write inputs.dat (with doB0 and, when any extension field differs from its
 * default, the extension record), B0.dat, wp.dat and nu.dat. */
int msc_write_run_dir(const char* dir, const msc_params* p, const double* wp, const double* nu);

void msc_free(void* ptr);

/* ------------------------------------------------------------------ */
/* Single solve                                                        */
/* ------------------------------------------------------------------ */

typedef struct msc_solver msc_solver;

/*
 * Create a solver on the current HIP device: host FP64 precompute of the
 * media, CPML tables, source table and incident line (meteor-fork.cpp:
 * 119-646), device state allocated and zeroed.  `precision` is 32 (fields,
 * J and CPML in FP32; plasma coefficients FP64; NTFF sums FP64) or 64.
 * `wp`/`nu` as returned by msc_read_run_dir.  The tiling is the defaults
 * of msc_tiling_default with one change: one-step x chunks of 4 planes
 * below 24 M cells.  The two-step region is used when it gives at least
 * 1.5 two-step tiles per CU.
 */
int msc_solver_create(const msc_params* p, const double* wp, const double* nu,
                      int precision, msc_solver** out);

/* Tile shapes of one solve.  No counterpart in the reference (its loops are
 * untiled).  Results do not depend on them: every cell is computed by the
 * same code whatever the tiling.  Speed does. */
typedef struct msc_tiling {
  int32_t ty;      /* rows of the one-step interior tiles: 3, 7 or 15 */
  int32_t ty_b;    /* rows of the one-step CPML tiles: 3, 7 or 15 */
  int32_t chunk;   /* x planes per one-step tile (>= 1) */
  int32_t chunk2;  /* x planes per two-step tile (>= 1) */
} msc_tiling;

/* The defaults: 7, 7, 8, 32. */
void msc_tiling_default(msc_tiling* t);

/* msc_solver_create with the given tiling instead of the defaults (and of
 * the MSC_TY / MSC_TY_B / MSC_CHUNK / MSC_CHUNK2 environment settings). */
int msc_solver_create_tiled(const msc_params* p, const double* wp, const double* nu, int precision,
                            const msc_tiling* tiling, msc_solver** out);

/* Autotuner: times tilings on this solve and returns the fastest.  A
 * coordinate search from the defaults over chunk2 (unless MSC_PAIR=0
 * turns two-step tiles off), chunk, ty and ty_b, each
 * candidate a fresh solver stepped 2 steps and then `steps` timed steps
 * (at most tsteps-2); a candidate replaces the current one when it is
 * faster by more than 0.5%.  *best_ms = its elapsed GPU time per step.
 * No counterpart in the reference. */
int msc_autotune(const msc_params* p, const double* wp, const double* nu, int precision, int steps,
                 msc_tiling* best, double* best_ms);

/* Enqueue `nsteps` time steps (meteor-fork.cpp:683-1667) on the solver's
 * stream.  Asynchronous: returns before the GPU finishes. */
int msc_solver_step(msc_solver* s, int nsteps);

/* Wait for all enqueued work. */
int msc_solver_sync(msc_solver* s);

/* Steps completed so far (enqueued). */
int msc_solver_time(const msc_solver* s);

/* Copy one field component to host as FP64, [nx][ny][nz].  The J
 * components are those the reference holds after the same number of steps. */
int msc_solver_get_field(msc_solver* s, int comp, double* out);

/* msc_solver_get_field without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (serving
 * loops and harnesses that forbid hipMalloc).  Same values as
 * msc_solver_get_field, bit for bit; `n` (the capacity of `out`) must be at
 * least nx*ny*nz.  The component is converted to FP64 (and J caught up by
 * the deferred current update) in a fixed device scratch array of 2^20
 * doubles (8 MiB) that the library holds on every device it runs on until
 * the process ends, in passes of at most 2^20 cells copied straight into
 * `out`, so any grid size fits.  Every solver on a device shares that
 * device's scratch: calls that overlap on one device race (make them from
 * one host thread per device at a time).  Synchronous. */
int msc_solver_get_field_noalloc(msc_solver* s, int comp, double* out, size_t n);

/* Number of doubles in the NTF payload: Eydft[2M], Ezdft[2M], then the 24
 * face arrays of output_NTF.dat in file order (meteor-fork.cpp:1677-1714).
 * msc_solver_get_ntff returns the running DFT through the last step
 * enqueued, as the reference's arrays hold it after that step (it flushes
 * the samples buffered since the last K-step contraction first). */
size_t msc_solver_ntff_count(const msc_solver* s);
int msc_solver_get_ntff(msc_solver* s, double* out, size_t n);

/* Contract the NTFF surface samples buffered since the last K-step block
 * into the FP64 sums now, on the solver's stream (asynchronous).  The
 * reference adds every step's sample to its sums at once
 * (meteor-fork.cpp:971-1023, 1417-1470); this build buffers K steps and
 * contracts them together, so a caller that times or reads a partial block
 * flushes it.  Results are the same sums in another order (FP64). */
int msc_solver_flush_ntff(msc_solver* s);

/* Non-finite watchdog: *flag = 1 once a NaN or infinite value has reached
 * the NTFF surface (every step's surface samples are checked on the
 * device; a non-finite value anywhere gets there within the steps the
 * stencil needs to carry it), else 0.  Waits for the work enqueued so far.
 * The reference has no such check: its only health signal is the max|Ex|
 * it logs at each output frame (meteor-fork.cpp:1616-1632). */
int msc_solver_nonfinite(msc_solver* s, int* flag);

/* Probe histories (Probe.dat body after the index header): 6 arrays
 * [tsteps][nprobes] (Ex,Ey,Ez,Hx,Hy,Hz), meteor-fork.cpp:1719-1729. */
int msc_solver_get_probes(msc_solver* s, double* out, size_t n);

/* Output frames, captured on the device every tsteps/numfiles steps
 * (meteor-fork.cpp:1616-1660).  msc_solver_frame_count gives the doubles in
 * one E or H frame: slices z, y, x (2 components each) then the two
 * incident lines, 2(nx*ny + nx*nz + ny*nz) + 2nx.  msc_solver_nframes gives
 * the frames captured so far; msc_solver_get_frame copies frame f (E, H and
 * the max|Ex| the reference logs at that step; any pointer may be NULL). */
size_t msc_solver_frame_count(const msc_solver* s);
int msc_solver_nframes(const msc_solver* s);
int msc_solver_get_frame(msc_solver* s, int f, double* eframe, double* hframe, double* emax);

/* max |Ex| over the grid (meteor-fork.cpp:1621-1630). */
int msc_solver_max_abs_ex(msc_solver* s, double* out);

/* Device bytes held by the solver. */
size_t msc_solver_device_bytes(const msc_solver* s);

/* Average duration (ms) of the fused step kernel over the last
 * msc_solver_step call, measured with HIP events on the solver's stream;
 * 0 when timing is off.  msc_solver_set_timing(s, 1) turns it on. */
int msc_solver_set_timing(msc_solver* s, int on);
double msc_solver_kernel_ms(const msc_solver* s);

/* Timing of the last msc_solver_step call (timing on), split by launch
 * kind.  In two-step mode (the default; MSC_PAIR=0 turns it off) a pair of
 * steps t, t+1 runs the two-step kernel once over the region pair_cells
 * (6 cells clear of every grid face) and the one-step kernels twice over
 * the rest; steps that write a frame or a slice run one-step everywhere.
 * No counterpart in the reference (its loop is untimed). */
typedef struct msc_timing {
  int32_t steps;          /* steps of that call */
  int32_t pair_launches;  /* two-step launches among them */
  double step_ms;         /* all field-update kernels, per step */
  double pair_ms;         /* mean duration of one two-step launch */
  int64_t pair_cells;     /* cells a two-step launch advances (0: mode off) */
  double span_ms;         /* all GPU work of that call, first to last, per step */
  int64_t pair_j_cells;   /* of pair_cells, those in tiles that update J */
  int64_t step_j_cells;   /* cells per step in tiles that update J (tiles where
                             eps0 wp^2 is zero everywhere skip J: it stays 0) */
} msc_timing;
int msc_solver_get_timing(const msc_solver* s, msc_timing* out);

void msc_solver_destroy(msc_solver* s);

/* The memory traffic of the two-step kernel over this solve's region with no
 * arithmetic: the same tiles, the same loads (each tile and its 2-cell halo,
 * one x plane ahead) and one store per owned value, of the value just read,
 * to where it was read, so the solve's state is unchanged.  Waits for the
 * solver's queued work, then times `reps` launches; *ms_per_launch = their
 * mean.  No counterpart in the reference: it measures the bandwidth the
 * two-step kernel's access pattern reaches on the device (MSC_EINVAL when
 * the solve has no two-step region). */
int msc_solver_pattern_probe(msc_solver* s, int reps, double* ms_per_launch);

/* Streaming copy of `bytes` between two fresh device buffers, `iters` times
 * at each of several grid sizes; *gbps = bytes read + written per second of
 * the fastest.  No counterpart in the
 * reference: it measures the device's attainable HBM bandwidth next to the
 * solver's roofline and serves as the known-size calibration of the HBM byte
 * counters. */
int msc_stream_copy(size_t bytes, int iters, double* gbps);

/*
 * Multi-solve server.  Replaces the batch layer of the reference: the loops
 * of setup/meteorbatch.m:26-40 and setup/meteorrandom.m:30-44 that submit
 * one PBS job per run (setup/meteorrun.m:154-183), each job one process
 * running the whole time loop.  Here the runs share one GPU:
 *  - msc_server_create reserves one device arena of `budget_bytes`
 *    (0: 90% of the free device memory) and starts a worker thread;
 *  - msc_server_submit queues a solve (savefields must be 0: the server
 *    returns the NTFF payload and the probes).  The parameters are copied;
 *    wp and nu are read in place by the worker, so they must stay valid
 *    until the request's state is past MSC_REQ_QUEUED;
 *  - the worker prepares queued solves in submission order, each as soon
 *    as its block fits in the arena (msc_solve_footprint gives its size);
 *  - msc_server_advance runs up to `max_iterations` gang iterations, each
 *    advancing every resident solve by one time step, and returns early
 *    when nothing is resident; *ndone counts the solves completed in the
 *    call.  Call it from one host thread;
 *  - msc_server_info / _get_ntff / _get_probes read a request's state,
 *    timestamps and results; msc_server_release forgets a finished request.
 * A solve's results are the same bits as the same solve run alone through
 * msc_solver_* (same kernels on the same per-solve data).
 */
typedef struct msc_server msc_server;

enum {
  MSC_REQ_QUEUED = 0,   /* submitted, waiting for preparation / arena space */
  MSC_REQ_READY = 1,    /* resident, joins the gang at the next iteration */
  MSC_REQ_RUNNING = 2,  /* stepping */
  MSC_REQ_DRAINING = 3, /* all steps launched, completing */
  MSC_REQ_DONE = 4,     /* results available */
  MSC_REQ_FAILED = 5,   /* rejected, over the time limit, or its fields became
                           non-finite: msc_server_info sets msc_last_error */
  MSC_REQ_CANCELLED = 6 /* msc_server_cancel before its last step */
};

typedef struct msc_request_info {
  int32_t state;
  int32_t t;            /* time steps launched */
  int32_t tsteps;
  size_t device_bytes;  /* its arena block */
  double t_submit, t_start, t_done;  /* seconds on the server clock; 0 = not yet */
} msc_request_info;

typedef struct msc_server_stats {
  int32_t queued, ready, running, draining;  /* draining: completing or being evicted */
  size_t arena_bytes, arena_used;
  long long iterations;   /* gang iterations launched */
  long long solve_steps;  /* solve time steps launched */
} msc_server_stats;

/* Device bytes one solve occupies in a server (0 for invalid parameters). */
size_t msc_solve_footprint(const msc_params* p, int precision);

/* Host-only check of the tile layout a server gang would launch for one
 * solve (two-step mode, one-step tiles of `zlanes` z lanes, 32 or 64, the
 * given tiling or the defaults when NULL).  Every tile descriptor is decoded
 * as the kernels decode it: its indices must lie inside the cut tables, its
 * owned range inside the grid and inside what its workgroup covers, its
 * CPML variant must carry every slab axis its cells meet, and the one-step
 * tiling, and the outer tiles with the two-step tiles, must each own every
 * cell exactly once.  No device call.  No counterpart in the reference. */
typedef struct msc_layout_report {
  long long cells;           /* nx ny nz */
  long long one_step_tiles;  /* workgroups of one one-step step over the grid */
  long long outer_tiles;     /* one-step workgroups outside the two-step region */
  long long pair_tiles;      /* fdtd_pair workgroups (0: no two-step region) */
  long long region_cells;    /* cells of the two-step region */
  long long max_list_tiles;  /* largest single one-step list */
} msc_layout_report;
int msc_layout_check(const msc_params* p, int precision, int zlanes, const msc_tiling* tiling, msc_layout_report* rep);

int msc_server_create(int precision, size_t budget_bytes, msc_server** out);
int msc_server_submit(msc_server* s, const msc_params* p, const double* wp, const double* nu, int64_t* id);
int msc_server_advance(msc_server* s, int max_iterations, int* ndone);
int msc_server_info(msc_server* s, int64_t id, msc_request_info* info);
int msc_server_get_ntff(msc_server* s, int64_t id, double* out, size_t n);
int msc_server_get_probes(msc_server* s, int64_t id, double* out, size_t n);
int msc_server_release(msc_server* s, int64_t id);
/* Failure handling, none of which the reference has (a PBS job runs until
 * its walltime; nothing checks its fields): msc_server_cancel stops a
 * request that is queued, ready or running (MSC_ESTATE once all its steps
 * are launched); it ends MSC_REQ_CANCELLED at a later msc_server_advance,
 * and its arena block returns once the launches already queued for it have
 * completed.  Other solves' results do not change.  msc_server_set_time_limit
 * ends a request that has been in the gang longer than `seconds` (0: no
 * limit, the default) as MSC_REQ_FAILED.  A request whose fields become
 * non-finite (the watchdog of msc_solver_nonfinite, read every 16 gang
 * iterations and at completion) ends MSC_REQ_FAILED as well; a failed
 * request that completed keeps its results for inspection.  Media with
 * non-finite values are refused at submission. */
int msc_server_cancel(msc_server* s, int64_t id);
int msc_server_set_time_limit(msc_server* s, double seconds);
int msc_server_get_stats(msc_server* s, msc_server_stats* out);
double msc_server_clock(const msc_server* s);
/* Scheduling knobs.  fill_tiles: each gang iteration steps the oldest
 * resident solve, then the others by least remaining work (cells x steps
 * left), until their workgroup tiles reach fill_tiles (<= 0: every resident
 * solve, the batch setting); default 32768.  use_graphs: replay each gang's
 * fixed launch sequence as a captured hipGraph (default 1). */
int msc_server_configure(msc_server* s, int fill_tiles, int use_graphs);
/* Tilings by size class (e.g. from msc_autotune on one solve per class): a
 * solve of n cells takes tilings[c] for the first class c with
 * n <= max_cells[c], the last class if none.  nclasses = 0 restores the
 * defaults.  Only while no request is queued, resident or draining. */
int msc_server_set_tilings(msc_server* s, const msc_tiling* tilings, const int64_t* max_cells, int nclasses);
void msc_server_destroy(msc_server* s);

/*
 * Whole drop-in run in `dir`: read the files, solve, write output_E.dat,
 * output_H.dat, output_J.dat (when savefields), output_NTF.dat, Probe.dat
 * and log.txt with the reference's layouts.  This is main() of
 * meteor-fork.cpp:42-1742.  `precision` as in msc_solver_create.
 */
int msc_run_dir(const char* dir, int precision);

#ifdef __cplusplus
}
#endif

#endif /* MSC_H */
