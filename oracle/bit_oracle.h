/*
 * bit_oracle.h -- CPU restatement of the particle-cluster treecode of ongbw/bit.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker for the HIP product
 * path: tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it; nothing in the product (bit_amd/) links, loads or calls it.
 *
 * Everything is FP64 and follows the reference function by function; each
 * function cites the reference file:line it restates (paths relative to the
 * reference root).  Build flags keep -ffp-contract=off so that the box, MAC and
 * moment arithmetic round exactly like the reference's g++ -O2 x86-64 build.
 *
 * Pinning (what checks each function; see tests/test_oracle_*.py):
 *   ora_tree_*        bit-exact against fixtures written by the compiled
 *                     reference (oracle/_ref, fixtures under tests/golden);
 *   ora_moments       brute-force sum of q*dx^i*dy^j*dz^k in long double and
 *                     the reference's own moments (fixtures);
 *   ora_taylor_coeffs independent multivariate power-series expansion of
 *                     exp(-kappa R)/R, R^2=|x-y|^2+delta^2 (tests) and the
 *                     reference recurrence through end-to-end potentials;
 *   ora_potential     reference potentials (fixtures; Coulomb via
 *                     treepotential(p=-1) and screened via shielded_treepot)
 *                     and convergence to ora_direct as nterms grows;
 *   ora_direct        closed form in long double on tiny sets.
 */
#ifndef BIT_ORACLE_H
#define BIT_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One cluster of the bisection tree; field meaning as TREE in tree3d.h:51-74. */
typedef struct ora_node {
  int nummemb;
  double x0, x1, y0, y1, z0, z1;
  double xmid, ymid, zmid;
  double aspectxy, aspectxz, aspectyz;
  double sqradius;
  int children[8];
  int *members; /* ascending particle indices (stable partitions of 0..n-1) */
} ora_node;

typedef struct ora_tree ora_tree;

/* Kernel / MAC parameters of one evaluation.
 *   G(r)   = exp(-kappa R)/R,  R = sqrt(r^2 + delta2)        (near and far)
 *   accept = sqradius < acc * ((r^2 + delta2) - mac_sub)     (MAC)
 * The reference's functions are special cases:
 *   shielded_treepot (tree3d.cpp:3448)  delta2 = 0,    mac_sub = DEL = 0.01
 *   treepotential p=-1 (tree3d.cpp:4679) kappa = 0, delta2 = mac_sub = DEL. */
typedef struct ora_params {
  int nterms;     /* Taylor order p = nterms-1 (degrees 0..nterms-1) */
  double acc;     /* MAC parameter theta, multiplies the squared distance */
  double kappa;   /* screening */
  double delta2;  /* regularization delta^2 */
  double mac_sub; /* value subtracted from (r^2+delta2) in the MAC */
  int kernel_order; /* field only: regularized family G_L (problem.h KERNEL_ORDER) */
  int root;         /* node the traversals start at: the reference's `ind`
                       (tree3d.cpp:851, 3448, 4679); 0 = the whole tree */
} ora_params;

/* ---- tree (tree3d.cpp:990-1551) ---- */
/* Builds the tree; returns NULL on failure (node cap exceeded / child overflow). */
ora_tree *ora_tree_build(int n, const double *x, const double *y, const double *z,
                         int maxmemb);
void ora_tree_free(ora_tree *t);
int ora_tree_numtree(const ora_tree *t);
/* Number of z-split checks (tree3d.cpp:1458-1477) that addressed a node other
 * than the upper half of the preceding y-split: inputs with hits > 0 exercise
 * the reference's indexing quirk. */
int ora_tree_quirk_hits(const ora_tree *t);
const ora_node *ora_tree_node(const ora_tree *t, int i);
/* Flat export: boxes[13*numtree] = x0,x1,y0,y1,z0,z1,xmid,ymid,zmid,axy,axz,ayz,sqr;
 * children[8*numtree]; nummemb[numtree]; members concatenated in node order. */
void ora_tree_export(const ora_tree *t, double *boxes, int *children, int *nummemb,
                     int *members);
int64_t ora_tree_total_members(const ora_tree *t);

/* ---- moments (tree3d.cpp:2631-2711): out[numtree][nterms][nterms][nterms] ---- */
void ora_moments(const ora_tree *t, const double *x, const double *y, const double *z,
                 const double *q, int nterms, double *out);

/* ---- Taylor coefficients of exp(-kappa R)/R about a cluster centre
 * (tree3d.cpp:3004-3227): a[(i*(nterms+1)+j)*(nterms+1)+k] for i+j+k < nterms,
 * with (dx,dy,dz) = target - centre and R2 = dx^2+dy^2+dz^2+delta2. ---- */
void ora_taylor_coeffs(int nterms, double dx, double dy, double dz, double R2,
                       double kappa, double *a);

/* ---- per-target evaluation (tree3d.cpp:3448-3627, 4679-4778) ----
 * Potential phi_i = sum_{j != i} q_j G(|x_i - x_j|) via the treecode, for the
 * particle targets tidx[0..nt-1].  Optional counters (may be NULL):
 * n_near[t] = source pairs summed directly, n_far[t] = clusters accepted. */
void ora_potential(const ora_tree *t, const double *x, const double *y, const double *z,
                   const double *q, const double *moments, ora_params prm, int nt,
                   const int *tidx, double *pot, int64_t *n_near, int64_t *n_far,
                   int nthreads);

/* Potential at arbitrary points (p = -1 of the reference traversals,
 * tree3d.cpp:3538-3551, 4744-4757): all sources, pairs with R^2 <= 1e-14
 * skipped. */
void ora_potential_points(const ora_tree *t, const double *x, const double *y,
                          const double *z, const double *q, const double *moments,
                          ora_params prm, int m, const double *px, const double *py,
                          const double *pz, double *pot, int nthreads);
/* Field at arbitrary points (treeforce p = -1, tree3d.cpp:937-950). */
void ora_treeforce_points(const ora_tree *t, const double *x, const double *y,
                          const double *z, const double *q, const double *moments,
                          ora_params prm, int m, const double *px, const double *py,
                          const double *pz, double *ex, double *ey, double *ez, int nthreads);

/* Per-target interaction lists in reference DFS order: far clusters and near
 * leaves of target tidx; returns counts, writes up to cap entries each. */
void ora_lists(const ora_tree *t, const double *x, const double *y, const double *z,
               ora_params prm, int ti, int *far, int *nfar, int *near, int *nnear, int cap);

/* ---- direct sum (tree3d.cpp:1966-2016): phi_i = sum_{j != i} q_j G ---- */
void ora_direct(int n, const double *x, const double *y, const double *z, const double *q,
                double kappa, double delta2, int nt, const int *tidx, double *pot,
                int nthreads);

/* ---- field, regularized Coulomb (tree3d.cpp:851-971, 456-810, 73-97) ----
 * E_i += sum q_j grad G  with the reference's sign convention (xfor of treeforce),
 * delta2 = DEL semantics: rs = r^2 + delta2, MAC acc*(rs - mac_sub). */
void ora_treeforce(const ora_tree *t, const double *x, const double *y, const double *z,
                   const double *q, const double *moments, ora_params prm, int nt,
                   const int *tidx, double *ex, double *ey, double *ez, int nthreads);
void ora_direct_force(int n, const double *x, const double *y, const double *z,
                      const double *q, double delta2, int kernel_order, int nt, const int *tidx,
                      double *ex, double *ey, double *ez, int nthreads);
/* Taylor coefficients, in the source offset y, of R^-(2l+1) with
 * R^2 = |x - y|^2 + delta2 (x = (dx,dy,dz), R2 = |x|^2 + delta2), degrees
 * < nterms, layout [(i*S + j)*S + k], S = nterms + 1.  Term l of the
 * regularized kernel family (tree3d.cpp:614-791). */
void ora_powinv_coeffs(int nterms, int l, double dx, double dy, double dz, double R2, double *a);

#ifdef __cplusplus
}
#endif
#endif
