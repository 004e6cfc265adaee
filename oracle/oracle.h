/*
 * oracle.h - FP64 CPU restatement of the reference meteor FDTD solver.
 *
 * TEST INFRASTRUCTURE ONLY.  This library checks the HIP path's results:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and never as the thing measured or shipped.
 *
 * It restates source/meteor-fork.cpp:42-1742 (the makefile's solver) as
 * separate functions, keeping the reference's loop order and arithmetic
 * order so that it reproduces the reference bit for bit on the inputs where
 * the reference is well defined (cubic grids, nu>0 and B!=0 when doB0=1).
 * Pinned against: golden fixtures written by the reference compiled from its
 * own source (oracle/Makefile -> oracle/_ref/), closed forms (Lee-Kalluri
 * A,K versus a matrix exponential, CPML tables, source waveforms, DFT of a
 * tone) and invariants (vacuum TF/SF leakage, energy, plate/sphere RCS).
 */
#ifndef MSC_ORACLE_H
#define MSC_ORACLE_H

#include <stddef.h>
#include "../include/msc.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc;

/* Set-up part of main() (meteor-fork.cpp:49-646).  Returns NULL and fills
 * err on failure. wp/nu: [nx][ny] (or [nx][ny][nz] when p->media3d). */
orc* orc_create(const msc_params* p, const double* wp, const double* nu, char* err, int errlen);
/* nsteps iterations of the time loop (meteor-fork.cpp:683-1667). */
void orc_step(orc* o, int nsteps);
int orc_time(const orc* o);
/* comp: MSC_EX..MSC_JZ; out: nx*ny*nz doubles. */
void orc_get_field(const orc* o, int comp, double* out);
size_t orc_ntff_count(const orc* o);
void orc_get_ntff(const orc* o, double* out);
size_t orc_probe_count(const orc* o);
void orc_get_probes(const orc* o, double* out);
size_t orc_frame_count(const orc* o);
void orc_get_frames(const orc* o, double* eframe, double* hframe);
double orc_max_abs_ex(const orc* o);
/* The extruded media the time loop uses (meteor-fork.cpp:336-366 and the
 * media3d / nu scale-height extensions): wpe = plasma frequency (rad/s),
 * nue = collision frequency (1/s), each nx*ny*nz doubles. */
void orc_get_media(const orc* o, double* wpe, double* nue);
/* Lee-Kalluri 3x3 matrices for one cell (meteor-fork.cpp:1546-1573). */
void orc_lk_matrices(double nu, double dt, const double wce[3], double A[9], double K[9]);
/* CPML tables: b[2L], a[2L] for the E (half-offset) and H grids. */
void orc_cpml_tables(const orc* o, double* be, double* ae, double* bh, double* ah);
/* Source table EsourceA/B (meteor-fork.cpp:153-179). */
void orc_source_tables(const orc* o, double* a, double* b);
void orc_destroy(orc* o);

#ifdef __cplusplus
}
#endif

#endif
