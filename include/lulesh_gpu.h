/* lulesh_gpu.h -- C ABI of the MI355X (gfx950) LULESH Lagrange leapfrog step.
 *
 * The reference has no plugin or FFI interface: its hot path is the in-file
 * call structure LagrangeLeapFrog(Domain&) -> LagrangeNodal(Domain&),
 * LagrangeElements(Domain&, Index_t), CalcTimeConstraintsForElems(Domain&)
 * (lulesh.cc:2600-2645), all of it operating on the Domain data contract
 * (lulesh.h:148-595).  This header is the seam that replaces those calls: a
 * host Domain hands its mesh, connectivity, region sets and state to an opaque
 * device context once, then calls the three phases per cycle.
 *
 * Conventions
 *   - Plain pointers and sizes only.  Host arrays are read during the call
 *     and not retained.
 *   - Real fields are FP64 on the host side of the ABI in every precision
 *     mode; the mixed mode stores some of them in FP32 on the device and
 *     converts on upload/download.
 *   - lgpu_lagrange_nodal and lgpu_lagrange_elements only enqueue work on the
 *     context's stream.  lgpu_time_constraints is the per-cycle sync point: it
 *     returns the new dt constraints together with the status of every check
 *     the reference makes with exit(): LGPU_VOLUME_ERROR (-1) or
 *     LGPU_QSTOP_ERROR (-2), as VolumeError/QStopError in lulesh.h:42, and in
 *     the order the reference would have hit them.
 *   - One context per host thread; calls on one context are not re-entrant.
 *   - Return codes: 0 ok, -1/-2 as above, <= -100 usage or device errors.
 */
#ifndef LULESH_GPU_H
#define LULESH_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lgpu_ctx lgpu_ctx;

enum lgpu_status {
   LGPU_OK = 0,
   LGPU_VOLUME_ERROR = -1, /* VolumeError, lulesh.h:42 */
   LGPU_QSTOP_ERROR = -2,  /* QStopError,  lulesh.h:42 */
   LGPU_EINVAL = -100,     /* bad argument or mesh description */
   LGPU_EDEVICE = -101,    /* HIP runtime error */
   LGPU_ENOMEM = -102      /* device allocation failed */
};

enum lgpu_precision {
   LGPU_FP64 = 0,  /* every quantity in FP64, arithmetic as the reference */
   LGPU_MIXED = 1, /* FP32 arithmetic: stress and hourglass corner forces
                      and their node sums, Q gradients and limiter, EOS
                      pressure and sound speed; FP32 storage: p, q, ss and
                      the Q gradients.  FP64: node accelerations, velocities
                      and positions, element volume, vnew, delv, strain rate,
                      characteristic length, the energy update, the dt
                      candidates and reductions */
   LGPU_FP64_FMA = 2 /* every quantity in FP64, the reference's association
                        order, with each multiply-add fused (one rounding
                        instead of two); results agree with the reference's
                        to roundoff rather than bit for bit */
};

/* Real fields (Domain accessors of lulesh.h:266-367).  Node fields first. */
enum lgpu_field {
   LGPU_X = 0, LGPU_Y, LGPU_Z, LGPU_XD, LGPU_YD, LGPU_ZD, LGPU_XDD, LGPU_YDD, LGPU_ZDD,
   LGPU_FX, LGPU_FY, LGPU_FZ, LGPU_NODALMASS,
   LGPU_E, LGPU_P, LGPU_Q, LGPU_QL, LGPU_QQ, LGPU_V, LGPU_VOLO, LGPU_VNEW, LGPU_DELV,
   LGPU_VDOV, LGPU_AREALG, LGPU_SS, LGPU_ELEMMASS,
   LGPU_DELX_XI, LGPU_DELX_ETA, LGPU_DELX_ZETA, LGPU_DELV_XI, LGPU_DELV_ETA, LGPU_DELV_ZETA,
   LGPU_NUM_FIELDS
};

/* Physics constants (Domain constructor, lulesh-init.cc:20-38). */
typedef struct lgpu_constants {
   double e_cut, p_cut, q_cut, v_cut, u_cut;
   double hgcoef, ss4o3, qstop, monoq_max_slope, monoq_limiter_mult;
   double qlc_monoq, qqc_monoq, qqc, eosvmax, eosvmin, pmin, emin, dvovmax, refdens;
} lgpu_constants;

/* Mesh, connectivity and region description: the integer part of Domain
 * (lulesh.h:478-500, 587-588) plus the EOS cost knob. */
typedef struct lgpu_mesh_desc {
   int32_t numElem, numNode;
   int32_t sizeX, sizeY, sizeZ;               /* elements per edge           */
   const int32_t *nodelist;                   /* 8*numElem                   */
   const int32_t *lxim, *lxip, *letam, *letap, *lzetam, *lzetap; /* numElem  */
   const int32_t *elemBC;                     /* numElem                     */
   const int32_t *nodeElemStart;              /* numNode+1                   */
   const int32_t *nodeElemCornerList;         /* nodeElemStart[numNode]      */
   const int32_t *symmX, *symmY, *symmZ;      /* node sets, may be NULL      */
   int32_t numSymmX, numSymmY, numSymmZ;
   int32_t numReg;
   const int32_t *regElemSize;                /* numReg                      */
   const int32_t *const *regElemlist;         /* numReg lists                */
   int32_t cost;                              /* EOS repeat rule, lulesh.cc:2390-2400 */
   int32_t precision;                         /* enum lgpu_precision         */
   int32_t flags;                             /* LGPU_FLAG_* below           */
   lgpu_constants k;
} lgpu_mesh_desc;

/* lgpu_mesh_desc.flags */
enum lgpu_flags {
   /* Use the general CSR-gather kernels even when the mesh is the
    * reference's structured brick (lulesh-init.cc:217-337), for which the
    * fused brick kernel is chosen by default. */
   LGPU_FLAG_GENERIC = 1
};

/* Fills *k with the reference's constants (lulesh-init.cc:20-38). */
void lgpu_default_constants(lgpu_constants *k);

/* Builds the device context: validates the description, allocates one device
 * arena for every array the step needs (the reference mallocs its scratch
 * per cycle, lulesh.cc:514-518, 737-741, 1001-1006, 2222-2235), uploads the
 * connectivity and picks the kernel path.  Replaces the device-relevant part
 * of Domain::Domain (lulesh-init.cc:16-194).  State fields start at zero
 * except v = 1; upload the rest with lgpu_upload. */
int lgpu_create(const lgpu_mesh_desc *desc, lgpu_ctx **out);

/* The reference's own problem -- the Sedov blast on an s^3 brick -- described
 * by its command-line parameters, for set-up on the device. */
typedef struct lgpu_sedov_desc {
   int32_t size;                  /* elements per edge (-s)                   */
   int32_t numReg, balance, cost; /* -r, -b, -c                               */
   double emult;                  /* initial energy multiplier (1 = -s run)   */
   /* The region layout as runs of consecutive elements, in element order:
    * run i gives runLen[i] elements to region runReg[i] (1-based).  These
    * are the draws of CreateRegionIndexSets (lulesh-init.cc:401-510), a
    * sequential glibc rand() sequence the host makes (lh_region_runs). */
   int32_t numRuns;
   const int32_t *runReg, *runLen;
   int32_t precision;             /* enum lgpu_precision                      */
   int32_t flags;                 /* LGPU_FLAG_*                              */
   lgpu_constants k;
} lgpu_sedov_desc;

/* Frees the context (Domain::~Domain, lulesh-init.cc:198-213). */
void lgpu_destroy(lgpu_ctx *ctx);

/* Synchronous host -> device copy of one real field (count = numNode or
 * numElem values). */
int lgpu_upload(lgpu_ctx *ctx, int field, const double *host);

/* Synchronous device -> host copy of one real field.  Fields the step does
 * not keep (fx.., xdd.., ql, qq, delv) hold their last value only while
 * lgpu_set_option(LGPU_OPT_KEEP_DIAGNOSTICS) is on. */
int lgpu_download(lgpu_ctx *ctx, int field, double *host);

/* LagrangeNodal (lulesh.cc:1223-1269): nodal forces (stress integration and
 * Flanagan-Belytschko hourglass control, summed per node in
 * nodeElemCornerList order), acceleration, symmetry-plane BCs, velocity and
 * position, with dt = domain.deltatime().  Enqueues only. */
int lgpu_lagrange_nodal(lgpu_ctx *ctx, double dt);

/* LagrangeElements (lulesh.cc:2431-2443): kinematics, monotonic Q, the
 * region-by-region EOS with its repeat counts, and the volume update.
 * Enqueues only. */
int lgpu_lagrange_elements(lgpu_ctx *ctx, double dt);

/* CalcTimeConstraintsForElems (lulesh.cc:2576-2596): Courant and hydro dt
 * minima over all regions.  Waits for the cycle to finish and returns its
 * status (0, -1 or -2); on 0 writes *dtcourant and *dthydro. */
int lgpu_time_constraints(lgpu_ctx *ctx, double *dtcourant, double *dthydro);

/* CalcTimeConstraintsForElems without a context and without any device
 * allocation.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  It computes what lgpu_time_constraints
 * returns -- the Courant (lulesh.cc:2447-2512) and hydro (lulesh.cc:2517-2572)
 * minima, from 1.0e+20, over elements with vdov != 0 -- from device arrays
 * the caller holds:
 *   ss, vdov, arealg   FP64 element fields (device pointers)
 *   elemlist           numElem element indices (device), the union of the
 *                      region lists; NULL = elements 0..numElem-1
 *   qqc, dvovmax       Domain constants (lulesh-init.cc:32, 37)
 *   dtOut              device pointer to 2 doubles: [0] dtcourant, [1] dthydro
 *   stream             hipStream_t to enqueue on (NULL: the null stream)
 * Enqueues only; the caller synchronises the stream before reading dtOut.
 * Workgroup partials go to a fixed per-GPU array of the library (32 KiB,
 * 2048 partial pairs; a grid-stride loop covers any numElem >= 0), so calls
 * that overlap on one GPU race.  Returns 0, LGPU_EINVAL (null array or
 * numElem < 0) or LGPU_EDEVICE. */
int lgpu_time_constraints_noalloc(const double *ss, const double *vdov, const double *arealg,
                                  const int32_t *elemlist, int32_t numElem, double qqc,
                                  double dvovmax, double *dtOut, void *stream);

/* Diagnostics; no counterpart in the reference.  Evaluates the FP64
 * reciprocal and square root that k_eos uses inside its repeats on n device
 * values x, writing 1/x[i] to recip and sqrt(x[i]) to root (device arrays;
 * either may be NULL), enqueued on stream.  They are correctly rounded for
 * 2^-400 <= |x| <= 2^400 (reciprocal) and 2^-767 <= x < +inf (square root);
 * outside those ranges k_eos uses the general routines and so may this
 * call's results differ from 1/x and sqrt(x).  Returns 0, LGPU_EINVAL (null
 * x or n < 0) or LGPU_EDEVICE. */
int lgpu_fp64_routines(const double *x, int64_t n, double *recip, double *root, void *stream);

/* Time-step controls of Domain (lulesh.h:557-567) used by the device-side
 * loop below. */
typedef struct lgpu_time_state {
   double time, deltatime, dtcourant, dthydro, stoptime, dtfixed, dtmax;
   double deltatimemultlb, deltatimemultub;
   int32_t cycle;
} lgpu_time_state;

/* The reference's main loop (lulesh.cc:2745-2748) with TimeIncrement
 * (lulesh.cc:166-222) evaluated on the device: runs while time < stoptime and
 * cycle < max_cycle without a host round trip per cycle, then writes the
 * time state back to *ts.  Returns the status as lgpu_time_constraints. */
int lgpu_run(lgpu_ctx *ctx, lgpu_time_state *ts, int32_t max_cycle);

/* Domain::Domain (lulesh-init.cc:16-194) on the device: builds the context
 * and, with set-up kernels instead of host loops and uploads, the mesh
 * (BuildMesh :218-268), the node-to-corner lists
 * (SetupThreadSupportStructures :272-337), the region work list from the
 * runs (CreateRegionIndexSets :401-510), symmetry planes, face neighbours and
 * boundary masks (:514-673), reference volumes and masses (:158-178) and the
 * energy deposit (:180-192).  The result is bit-identical to lgpu_create +
 * lgpu_upload of a host-built Domain.  *ts receives the initial time state
 * (lulesh-init.cc:146-156, deltatime from :192), ready for lgpu_run. */
int lgpu_create_sedov(const lgpu_sedov_desc *desc, lgpu_ctx **out, lgpu_time_state *ts);

/* Options. */
enum lgpu_option {
   LGPU_OPT_KEEP_DIAGNOSTICS = 1, /* also store fx..,xdd..,ql,qq,delv      */
   LGPU_OPT_KERNEL_TIMING = 2,    /* time every kernel with HIP events       */
   LGPU_OPT_BLOCK_SIZE = 3,       /* threads per block of the element/node
                                     kernels of the generic path: 64, 128
                                     or 256                                  */
   LGPU_OPT_BRICK_CHUNK = 4,      /* node planes per workgroup of the fused
                                     brick kernel (1..mesh edge)             */
   LGPU_OPT_XCD_REMAP = 5,        /* 1 (default): contiguous block ranges per
                                     XCD; 0: hardware order                 */
   LGPU_OPT_KERNEL_TIMING_MASK = 6, /* bit k: time kernel k while timing is on
                                     (default: all)                          */
   LGPU_OPT_KIN_CHUNK = 7,        /* element planes per workgroup of the
                                     brick kinematics kernel (1..mesh edge)  */
   LGPU_OPT_BRICK_ROWS = 8,       /* thread rows of a fused brick workgroup:
                                     16 (16x16 element columns, default) or 8
                                     (16x8)                                  */
   LGPU_OPT_EOS_INTERLEAVE = 9    /* k_eos dispatch order: 0 work-list order,
                                     largest EOS repeat count first; k > 0
                                     (default 4) one workgroup of the largest
                                     count every k+1 dispatches (an odd k is
                                     raised by one: the 8 XCDs are dealt
                                     workgroups in turn)                    */
};
int lgpu_set_option(lgpu_ctx *ctx, int option, int64_t value);
/* Current value of an option (the brick chunks report the defaults chosen at
 * lgpu_create: from the CU count and the kernels' occupancy). */
int lgpu_get_option(lgpu_ctx *ctx, int option, int64_t *value);

/* Kernel timing (LGPU_OPT_KERNEL_TIMING).  Kernels are numbered
 * 0..lgpu_num_kernels()-1. */
int lgpu_num_kernels(void);
const char *lgpu_kernel_name(int kernel);
int lgpu_kernel_stats(lgpu_ctx *ctx, int kernel, int64_t *launches, double *total_ms);
int lgpu_reset_kernel_stats(lgpu_ctx *ctx);

/* Context facts for reports: path (0 generic CSR gather, 1 structured
 * brick), device arena bytes, precision; brick_faces = 1 when the face
 * neighbours and boundary masks are those of the brick with the reference's
 * boundary conditions, so that the EOS pass computes them instead of reading
 * them; brick_blocks_per_cu = resident workgroups of the fused nodal kernel
 * per CU, as used for the default chunk. */
typedef struct lgpu_info {
   int32_t path, precision, numElem, numNode, numReg, block;
   int64_t arena_bytes;
   int32_t brick_faces, brick_blocks_per_cu;
} lgpu_info;
int lgpu_get_info(lgpu_ctx *ctx, lgpu_info *info);

/* Last HIP error text of this context (empty if none). */
const char *lgpu_last_error(lgpu_ctx *ctx);

/* ------------------------------------------------------------------------
 * Ensemble: a batch of independent problems advanced together.
 *
 * No counterpart in the reference, which runs one Domain per process; this
 * is the reference's main loop (lulesh.cc:2745-2748) applied to many
 * Domains at once.  Each member is described as for lgpu_create (all members
 * share precision and constants) and keeps its own time state: TimeIncrement
 * (lulesh.cc:166-222) runs per member on the device, so members of different
 * sizes advance with their own dt.  One cycle of every member is captured as
 * one hipGraph and replayed.  Results of each member equal a single-problem
 * run of that member.  When every member is the reference's brick (and
 * members[0].flags lacks LGPU_FLAG_GENERIC) the fused brick kernels run over
 * all members' workgroups in one launch each; otherwise the CSR-gather
 * kernels are used.  Ensembles keep no diagnostics: the fields a single
 * context keeps only with LGPU_OPT_KEEP_DIAGNOSTICS (fx.., xdd.., ql, qq,
 * delv) are not written.
 * ---------------------------------------------------------------------- */
typedef struct lgpu_ensemble lgpu_ensemble;

int lgpu_ensemble_create(const lgpu_mesh_desc *members, int32_t nMembers, lgpu_ensemble **out);
void lgpu_ensemble_destroy(lgpu_ensemble *ens);
int lgpu_ensemble_upload(lgpu_ensemble *ens, int32_t member, int field, const double *host);
int lgpu_ensemble_download(lgpu_ensemble *ens, int32_t member, int field, double *host);

/* Runs every member until its own loop test fails (time >= stoptime or
 * cycle >= max_cycle) or its step fails.  ts holds nMembers time states and
 * is updated; status (optional, nMembers) receives each member's status.
 * Returns the first member error, or 0. */
int lgpu_ensemble_run(lgpu_ensemble *ens, lgpu_time_state *ts, int32_t max_cycle, int32_t *status);

enum lgpu_ensemble_option {
   LGPU_ENS_OPT_GRAPH = 1,        /* 1 (default): replay a captured cycle graph */
   LGPU_ENS_OPT_TIMING = 2,       /* time every kernel (no graph while on)      */
   LGPU_ENS_OPT_BLOCK_FORCE = 3,  /* threads per block: 64, 128 or 256          */
   LGPU_ENS_OPT_BLOCK_NODE = 4,
   LGPU_ENS_OPT_BLOCK_KIN = 5,
   LGPU_ENS_OPT_XCD_REMAP = 6,
   LGPU_ENS_OPT_BRICK_CHUNK = 7,  /* brick path: node planes per workgroup of
                                     the fused nodal kernel, every member      */
   LGPU_ENS_OPT_KIN_CHUNK = 8,    /* brick path: element planes per
                                     workgroup of the kinematics kernel       */
   LGPU_ENS_OPT_PATH = 9          /* read only: 1 brick path, 0 CSR gathers   */
};
int lgpu_ensemble_set_option(lgpu_ensemble *ens, int option, int64_t value);
int lgpu_ensemble_get_option(lgpu_ensemble *ens, int option, int64_t *value);

/* Autotuner: times trial cycles of this ensemble for each candidate block
 * shape of the force, node and kinematics kernels and for both XCD block
 * mappings, keeps the fastest, and restores the state afterwards. */
int lgpu_ensemble_autotune(lgpu_ensemble *ens, const lgpu_time_state *ts, int32_t trial_cycles);

int lgpu_ensemble_num_kernels(void);
const char *lgpu_ensemble_kernel_name(int kernel);
int lgpu_ensemble_kernel_stats(lgpu_ensemble *ens, int kernel, int64_t *launches, double *total_ms);
int64_t lgpu_ensemble_num_elem(lgpu_ensemble *ens);
const char *lgpu_ensemble_last_error(lgpu_ensemble *ens);

#ifdef __cplusplus
}
#endif

#endif
