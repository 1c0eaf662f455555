/* lulesh_oracle.h -- CPU restatement of the LULESH 2.x Lagrange leapfrog step.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the device
 * path is compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product never does.
 *
 * It follows the reference's arithmetic operation by operation (same
 * association order, no FMA contraction) and sums node forces in the order of
 * the reference's multi-threaded build: per node, over nodeElemCornerList in
 * ascending corner order (lulesh.cc:562-586, 967-990).  A scatter mode
 * reproduces the single-threaded build's order (lulesh.cc:548-558, 932-963).
 *
 * Errors are returned instead of calling exit(): VolumeError = -1,
 * QStopError = -2 (lulesh.h:42).
 */
#ifndef LULESH_ORACLE_H
#define LULESH_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_domain orc_domain;

/* Real-valued fields.  Numbering is shared with include/lulesh_gpu.h. */
enum orc_real_field {
   ORC_X = 0, ORC_Y, ORC_Z, ORC_XD, ORC_YD, ORC_ZD, ORC_XDD, ORC_YDD, ORC_ZDD,
   ORC_FX, ORC_FY, ORC_FZ, ORC_NODALMASS,
   ORC_E, ORC_P, ORC_Q, ORC_QL, ORC_QQ, ORC_V, ORC_VOLO, ORC_VNEW, ORC_DELV,
   ORC_VDOV, ORC_AREALG, ORC_SS, ORC_ELEMMASS,
   ORC_DELX_XI, ORC_DELX_ETA, ORC_DELX_ZETA, ORC_DELV_XI, ORC_DELV_ETA, ORC_DELV_ZETA,
   ORC_NUM_REAL_FIELDS
};

/* Integer arrays. */
enum orc_int_field {
   ORC_NODELIST = 0, ORC_LXIM, ORC_LXIP, ORC_LETAM, ORC_LETAP, ORC_LZETAM, ORC_LZETAP,
   ORC_ELEMBC, ORC_NODEELEMSTART, ORC_NODEELEMCORNERLIST, ORC_SYMMX, ORC_SYMMY, ORC_SYMMZ,
   ORC_REGNUMLIST, ORC_REGELEMSIZE, ORC_REGELEMLIST /* + region index */,
   ORC_NUM_INT_FIELDS
};

/* Scalars. */
enum orc_scalar {
   ORC_TIME = 0, ORC_DELTATIME, ORC_DTCOURANT, ORC_DTHYDRO, ORC_CYCLE, ORC_STOPTIME,
   ORC_DTFIXED, ORC_DTMAX, ORC_DELTATIMEMULTLB, ORC_DELTATIMEMULTUB,
   ORC_NUM_SCALARS
};

/* Builds the Sedov problem of edge s exactly as the reference's Domain
 * constructor (lulesh-init.cc:16-194) with one rank.  emult multiplies the
 * deposited energy (1.0 = the reference problem). */
orc_domain *orc_create(int s, int nr, int balance, int cost, double emult);
void orc_destroy(orc_domain *d);

int orc_num_elem(const orc_domain *d);
int orc_num_node(const orc_domain *d);
int orc_num_reg(const orc_domain *d);
int orc_region_rep(const orc_domain *d, int r);

double *orc_real(orc_domain *d, int field);
int32_t *orc_int(orc_domain *d, int field, int region);
int64_t orc_int_len(const orc_domain *d, int field, int region);
double orc_get(const orc_domain *d, int which);
void orc_set(orc_domain *d, int which, double value);

/* 0 = gather in nodeElemCornerList order (multi-threaded reference, default),
 * 1 = scatter in element order (single-threaded reference). */
void orc_set_force_order(orc_domain *d, int scatter);

/* Phases, named after the reference functions they restate. */
void orc_time_increment(orc_domain *d);                 /* lulesh.cc:166-222   */
int orc_calc_force_for_nodes(orc_domain *d);            /* lulesh.cc:1104-1135 */
void orc_node_update(orc_domain *d);                    /* lulesh.cc:1139-1219 */
int orc_lagrange_nodal(orc_domain *d);                  /* lulesh.cc:1223-1269 */
int orc_calc_lagrange_elements(orc_domain *d);          /* lulesh.cc:1572-1609 */
void orc_calc_monotonic_q_gradients(orc_domain *d);     /* lulesh.cc:1613-1757 */
int orc_calc_q(orc_domain *d);                          /* lulesh.cc:1945-2010 */
int orc_apply_material_properties(orc_domain *d);       /* lulesh.cc:2328-2406 */
void orc_update_volumes(orc_domain *d);                 /* lulesh.cc:2410-2427 */
int orc_lagrange_elements(orc_domain *d);               /* lulesh.cc:2431-2443 */
void orc_calc_time_constraints(orc_domain *d);          /* lulesh.cc:2576-2596 */
int orc_lagrange_leapfrog(orc_domain *d);               /* lulesh.cc:2600-2645 */

/* Main loop (lulesh.cc:2745-2757): TimeIncrement + LagrangeLeapFrog until
 * time >= stoptime or cycle >= its.  Returns 0 or the first error code. */
int orc_run(orc_domain *d, int its);

/* Plane-0 symmetry figures (lulesh-util.cc:197-212): MaxAbsDiff,
 * TotalAbsDiff, MaxRelDiff. */
void orc_symmetry(const orc_domain *d, double out[3]);

/* Element-level kernels exposed for closed-form tests. */
double orc_elem_volume(const double x[8], const double y[8], const double z[8]);
void orc_elem_node_normals(const double x[8], const double y[8], const double z[8],
                           double nx[8], double ny[8], double nz[8]);
void orc_elem_volume_derivative(const double x[8], const double y[8], const double z[8],
                                double dx[8], double dy[8], double dz[8]);
double orc_elem_char_length(const double x[8], const double y[8], const double z[8],
                            double volume);

#ifdef __cplusplus
}
#endif

#endif
