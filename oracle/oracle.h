/*
 * oracle.h -- FP64 CPU restatement of ChaNGa's gravity force-evaluation path.
 *
 * TEST INFRASTRUCTURE ONLY. Nothing in the product (changa_amd/, include/)
 * links, loads or calls this code. It is used by tests/, by
 * __graft_entry__.smoke() as the checker, and by bench.py's cpu_baseline leg.
 *
 * Every function states the reference file:line it follows. The formulation
 * is physics-first (symmetric trace-free tensor algebra, the Ewald
 * g-recursion, the Hernquist-Katz spline) rather than a transcription of the
 * reference's hand-scheduled arithmetic; the functions are pinned against
 * moments.c compiled from the reference (oracle/_ref) and against closed forms
 * (see tests/test_oracle_*.py).
 *
 * Moment storage: 22 doubles in the FMOMR component order of moments.h:64-68
 *   m, xx,yy,xy,xz,yz, xxx,xyy,xxy,yyy,xxz,yyz,xyz,
 *   xxxx,xyyy,xxxy,yyyy,xxxz,yyyz,xxyy,xxyz,xyyz
 * Components of order n are the symmetric trace-free (STF) tensor components
 * of sum_i m_i x_i^n about the centre of mass, divided by u^n (u = the cell
 * radius, "radius-scaled reduced moments").
 */
#ifndef CHANGA_ORACLE_H
#define CHANGA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_NMOM 22

/* ---- spline softening (gravity.h:147-182) ---------------------------- */
void orc_spline(double r2, double twoh, double *a, double *b);

/* ---- multipole algebra (moments.c) ------------------------------------ */
/* momMakeFmomr (moments.c:772-846): STF moments of a point mass m at x/u.
   Returns |x/u|^2. */
double orc_mom_make(double q[ORC_NMOM], double m, double u, double x, double y, double z);
/* momShiftFmomr (moments.c:1002-1058): shift moments (about the old centre,
   scale u) by d = old centre - new centre; the dipole created by the shift is
   dropped, exactly as the reference does. */
void orc_mom_shift(double q[ORC_NMOM], double u, double dx, double dy, double dz);
/* momRescaleFmomr (moments.c:238-268). */
void orc_mom_rescale(double q[ORC_NMOM], double unew, double uold);
/* momScaledAddFmomr (moments.c:202-235): qr(ur) += qa(ua). */
void orc_mom_scaled_add(double qr[ORC_NMOM], double ur, const double qa[ORC_NMOM], double ua);
/* momFmomr2Momc (moments.c:1563-1600): the 32-component complete STF form
   (MOMC order of moments.h:42-50). */
void orc_mom_to_momc(const double q[ORC_NMOM], double mc[32]);
/* momEvalFmomrcm (moments.c:1469-1525): potential and acceleration at
   r = (x,y,z) (target - cell centre) from radius-scaled moments q with scale u.
   Adds into *pot and a[0..2]. */
void orc_mom_eval(const double q[ORC_NMOM], double u, double x, double y, double z,
                  double *pot, double a[3]);

/* ---- tree moments build ------------------------------------------------
 * Follows GenericTreeNode::makeBucket (GenericTreeNode.h:219-257),
 * MultipoleMoments::operator+= (MultipoleMoments.h:207-352),
 * calculateRadius{Box,FarthestParticle,FarthestCorner}
 * (MultipoleMoments.h:460-510) and LocalTreeBuilder::doneChildren
 * (Compute.cpp:2534-2580).
 *
 * Nodes are numbered so that every child index is larger than its parent's.
 * Leaves (child0 < 0) are buckets. Output per node:
 *   geo[n*12 + 0..11] = cm.xyz, radius, soft, mass, bbox_lo.xyz, bbox_hi.xyz
 *   mom[n*22 + ...]   = radius-scaled reduced moments.
 */
void orc_build_moments(int n_nodes, const int32_t *node_first, const int32_t *node_count,
                       const int32_t *node_child, const double *node_box_radius,
                       const double *pos, const double *mass, const double *soft,
                       double *geo, double *mom);

/* ---- forces over interaction lists ------------------------------------
 * nodeBucketForce (gravity.h:398-475) with the openSoftening monopole
 * fallback (gravity.h:251-260, 414-421), partBucketForce (gravity.h:266-303).
 * Lists use ChaNGa's wire layout (cuda_typedef.h:208-235):
 *   cells: int32 pairs {index, offsetID}; parts: int32 triples {index, off, num}
 *   cell_marks[nb+1], part_marks[nb+1] index the two lists per bucket.
 * offsetID bits 22-30 carry the periodic replica (TreePiece.cpp:3631-3645);
 * bit 31 (the softened-cell hint of the GPU list) is ignored here -- the
 * oracle re-derives the softening decision itself.
 * Accumulates into acc[3N], pot[N], dtgrav[N] (max) for particles with
 * rung >= active_rung. Returns the number of (target, source) interactions
 * as counts[0] = PC (hexadecapole), counts[1] = softened-cell monopoles,
 * counts[2] = PP.
 */
void orc_forces_lists(int n_buckets, const int32_t *bucket_node,
                      const int32_t *node_first, const int32_t *node_count,
                      const double *geo, const double *mom,
                      const double *pos, const double *mass, const double *soft,
                      const int32_t *rung, int active_rung,
                      const int32_t *cells, const int64_t *cell_marks,
                      const int32_t *parts, const int64_t *part_marks,
                      double period, int soft_flags, int n_threads,
                      double *acc, double *pot, double *dtgrav, int64_t counts[3]);
/* The same with a period per axis (period[3]: x, y, z). */
void orc_forces_lists3(int n_buckets, const int32_t *bucket_node,
                       const int32_t *node_first, const int32_t *node_count,
                       const double *geo, const double *mom,
                       const double *pos, const double *mass, const double *soft,
                       const int32_t *rung, int active_rung,
                       const int32_t *cells, const int64_t *cell_marks,
                       const int32_t *parts, const int64_t *part_marks,
                       const double period[3], int soft_flags, int n_threads,
                       double *acc, double *pot, double *dtgrav, int64_t counts[3]);

/* openSoftening (gravity.h:251-260) for one (cell, bucket, offset) triple. */
int orc_open_softening(const double *cell_geo, const double *bucket_geo,
                       double ox, double oy, double oz);

/* ---- Ewald (Ewald.cpp) ------------------------------------------------- */
/* EwaldInit h-loop table (Ewald.cpp:285-383). q_root/u_root: root moments.
   ewt[i*5 + 0..4] = kx, ky, kz, hCfac, hSfac. Returns the number of entries
   written (at most max_n). */
int orc_ewald_table(const double q_root[ORC_NMOM], double u_root, double L, double hcut,
                    double *ewt, int max_n);
/* BucketEwald (Ewald.cpp:72-281) for the listed target particles. */
void orc_bucket_ewald(int n_targets, const int32_t *targets, const double *pos,
                      const double root_cm[3], double root_mass,
                      const double q_root[ORC_NMOM], double u_root,
                      const double *ewt, int n_ewt, double L, int n_reps, double ew_cut,
                      int n_threads, double *acc, double *pot, int64_t counts[2]);

/* ---- FP64 accuracy yardsticks (no reference counterpart) --------------- */
/* Direct summation with the same spline softening, onto the listed targets. */
void orc_direct_sum(int n, const double *pos, const double *mass, const double *soft,
                    int n_targets, const int32_t *targets, int n_threads,
                    double *acc, double *pot);
/* Fully periodic (cubic box L) softened force by an Ewald sum with per-pair
   real space (cut rcut < L/2) and structure-factor k space (|h| <= hmax). */
void orc_ewald_exact(int n, const double *pos, const double *mass, const double *soft,
                     int n_targets, const int32_t *targets, double L, double alpha,
                     double rcut, int hmax, int n_threads, double *acc, double *pot);

/* Adjoint of orc_forces_lists (softened cells as flagged in the lists) for
   L = sum_i w_i . a_i (w: 3 per particle): grad[4k..4k+2] += dL/dx_k,
   grad[4k+3] += dL/dm_k. Isolated or periodic tree part (no Ewald term). */
void orc_adjoint_lists(int n_nodes, const int32_t *node_child, int n_buckets,
                       const int32_t *bucket_node, const int32_t *node_first,
                       const int32_t *node_count, const double *geo, const double *mom,
                       const double *pos, const double *mass, const double *soft,
                       const int32_t *cells, const int64_t *cell_marks,
                       const int32_t *parts, const int64_t *part_marks, double period,
                       const double *w, const double *root_extra, int n_threads, double *grad);
/* Adjoint of orc_bucket_ewald over all n particles (w: 3 per particle, zero
   for non-targets): grad[4t..4t+2] += the target-side terms; hroot[35]
   receives the root's polynomial (monomial coefficients about the root's
   centre of mass), to be passed as root_extra to orc_adjoint_lists. */
void orc_adjoint_ewald(int n, const double *pos, const double *w, const double root_cm[3],
                       double root_mass, const double q_root[ORC_NMOM], double u_root,
                       const double *ewt, int n_ewt, double L, int n_reps, double ew_cut,
                       int n_threads, double *grad, double hroot[35]);
/* Exact gradient of L for the softened direct sum over all n particles. */
void orc_direct_grad(int n, const double *pos, const double *mass, const double *soft,
                     const double *w, int n_threads, double *grad);
/* d^(a+b+c)(1/|r|)/dx^a dy^b dz^c into out[a*49 + b*7 + c], a+b+c <= nmax <= 6. */
void orc_rinv_derivs(const double r[3], int nmax, double *out);

#ifdef __cplusplus
}
#endif
#endif
