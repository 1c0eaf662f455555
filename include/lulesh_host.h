/* lulesh_host.h -- C API of the host-side Domain (lulesh_amd/csrc/host).
 *
 * A flat C view of the C++ host Domain and its driver loop, for callers that
 * are not C++ (the Python package, benchmarks).  Each function names the
 * reference piece it stands for.  Real fields use the lgpu_field numbering of
 * lulesh_gpu.h.
 */
#ifndef LULESH_HOST_H
#define LULESH_HOST_H

#include <stdint.h>

#include "lulesh_gpu.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lh_domain lh_domain;

/* new Domain(1, 0, 0, 0, s, 1, nr, balance, cost) (lulesh.cc:2715-2716),
 * optionally with the deposited energy scaled by emult. */
lh_domain *lh_create(int32_t s, int32_t nr, int32_t balance, int32_t cost, double emult);
void lh_destroy(lh_domain *d);

/* Creates the device context for this domain and uploads its state. */
int lh_attach_device(lh_domain *d, int32_t precision);

/* As lh_attach_device with lgpu_mesh_desc.flags (e.g. LGPU_FLAG_GENERIC). */
int lh_attach_device_flags(lh_domain *d, int32_t precision, int32_t flags);
lgpu_ctx *lh_device(lh_domain *d);

/* One iteration of the main loop: TimeIncrement + LagrangeLeapFrog
 * (lulesh.cc:2747-2748).  Returns 0 or the error code. */
int lh_step(lh_domain *d);

/* The main loop (lulesh.cc:2745-2757) until time >= stoptime or cycle >= its.
 * device_loop = 1 runs TimeIncrement on the device (lgpu_run).  Returns 0 or
 * the error code. */
int lh_run(lh_domain *d, int32_t its, int32_t device_loop);

/* Copies the fields a cycle writes (x, y, z, xd, yd, zd, e, p, q, v, ss,
 * vnew, delv, vdov, arealg) from the device into the host Domain; nodalMass,
 * volo and elemMass keep the host's FP64 values. */
int lh_sync_from_device(lh_domain *d);

int32_t lh_num_elem(lh_domain *d);
int32_t lh_num_node(lh_domain *d);
int32_t lh_num_reg(lh_domain *d);
double *lh_real(lh_domain *d, int32_t field);

/* Integer arrays: 0 nodelist, 1..6 lxim..lzetap, 7 elemBC, 8 nodeElemStart,
 * 9 nodeElemCornerList, 10..12 symmX..Z, 13 regNumList, 14 regElemSize,
 * 15 regElemlist[region]. */
int32_t *lh_int(lh_domain *d, int32_t which, int32_t region, int64_t *len);

/* Scalars: 0 time, 1 deltatime, 2 dtcourant, 3 dthydro, 4 cycle,
 * 5 stoptime, 6 dtfixed, 7 dtmax, 8 deltatimemultlb, 9 deltatimemultub. */
double lh_get(lh_domain *d, int32_t which);
void lh_set(lh_domain *d, int32_t which, double value);

/* VerifyAndWriteFinalOutput's plane-0 symmetry figures
 * (lulesh-util.cc:197-212) on the host copy of e. */
void lh_symmetry(lh_domain *d, double out[3]);

/* The end-of-run visualization dump (DumpToVisit, lulesh-viz.cc:279-364) as a
 * legacy binary VTK file: mesh, regions, e, p, v, q per element and speed,
 * xd, yd, zd per node, from the host copies (call lh_sync_from_device
 * first).  0 on success, -1 if the file cannot be written. */
int lh_write_vtk(lh_domain *d, const char *path);

/* The first n values of glibc's rand() after srand(seed), from the host's
 * own generator (the source of the region layout, lulesh-init.cc:406-480). */
void lh_glibc_rand(uint32_t seed, int32_t n, int32_t *out);

/* The region layout of an s^3 problem with nr regions and the given balance
 * (CreateRegionIndexSets, lulesh-init.cc:401-510) as runs of consecutive
 * elements, the input of lgpu_create_sedov.  Returns the number of runs;
 * writes min(that, capacity) of them to reg (1-based region numbers) and len.
 * Returns -1 on bad arguments. */
int32_t lh_region_runs(int32_t s, int32_t nr, int32_t balance, int32_t *reg, int32_t *len,
                       int32_t capacity);

/* ------------------------------------------------------------------------
 * Ensembles of independent Sedov problems (device side: lgpu_ensemble).
 * No counterpart in the reference (one Domain per process there).
 * ---------------------------------------------------------------------- */
typedef struct lh_ensemble lh_ensemble;

/* Member parameters drawn from std::mt19937_64(seed), four raw 64-bit
 * outputs per member, in order: edge s = 16 + x0 mod 49 (16..64); energy
 * multiplier exp(ln 0.5 + ln 4 * u), u = (x1 >> 11) * 2^-53 (log-uniform in
 * [0.5, 2)); balance = 1 + x2 mod 2; cost = 1 + x3 mod 2. */
void lh_ensemble_draw(uint64_t seed, int32_t n, int32_t *s, double *emult, int32_t *balance,
                      int32_t *cost);

/* Builds n host Domains (nr[i] regions each) and their device ensemble. */
lh_ensemble *lh_ensemble_create(int32_t n, const int32_t *s, const double *emult, const int32_t *nr,
                                const int32_t *balance, const int32_t *cost, int32_t precision);
/* The same with lgpu_mesh_desc.flags for the device ensemble
 * (LGPU_FLAG_GENERIC: CSR-gather kernels instead of the brick path). */
lh_ensemble *lh_ensemble_create_flags(int32_t n, const int32_t *s, const double *emult,
                                      const int32_t *nr, const int32_t *balance, const int32_t *cost,
                                      int32_t precision, int32_t flags);
void lh_ensemble_destroy(lh_ensemble *e);

/* Runs every member's main loop to cycle its (or t_stop); status optional
 * (n entries).  Returns 0 or the first member error. */
int lh_ensemble_run(lh_ensemble *e, int32_t its, int32_t *status);

/* Copies the fields a cycle writes (as lh_sync_from_device) of every member
 * back into its host Domain. */
int lh_ensemble_sync(lh_ensemble *e);

/* Host Domain of member m, owned by the ensemble (do not lh_destroy it). */
lh_domain *lh_ensemble_member(lh_ensemble *e, int32_t m);
lgpu_ensemble *lh_ensemble_device(lh_ensemble *e);
int32_t lh_ensemble_size(lh_ensemble *e);

#ifdef __cplusplus
}
#endif

#endif
