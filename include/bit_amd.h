/*
 * bit_amd.h -- C-ABI of the MI355X particle-cluster treecode.
 *
 * This is the drop-in boundary under the reference's C++ library API
 * (src/tree3d.h of ongbw/bit).  Every entry point takes plain pointers and
 * sizes, returns an int status (BIT_OK = 0, negative BIT_E_* on failure, the
 * message in bit_last_error()), never throws, and issues its device work on
 * the context's stream.  One host thread owns a context at a time.
 *
 * Kernel evaluated (potential at particle targets, self excluded):
 *     phi_i = sum_{j != i} q_j exp(-kappa R_ij) / R_ij,   R_ij = sqrt(|x_i-x_j|^2 + delta2)
 * Multipole acceptance (reference convention, tree3d.cpp:3473 / 4691):
 *     accept(cluster, target) = sqradius < acc * ((r^2 + delta2) - mac_sub),
 * r = distance from the target to the cluster's box midpoint.
 * The reference's functions are parameter choices of this one evaluation:
 *     shielded_treepot (tree3d.cpp:3448):   delta2 = 0, mac_sub = 0.01 (DEL), result -phi
 *     treepotential p=-1 (tree3d.cpp:4679): kappa = 0, delta2 = mac_sub = 0.01
 */
#ifndef BIT_AMD_H
#define BIT_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define BIT_OK 0
#define BIT_E_INVALID (-1)   /* bad argument (size, NaN/Inf input, unsupported nterms) */
#define BIT_E_OOM (-2)       /* device allocation failed or HBM budget exceeded */
#define BIT_E_HIP (-3)       /* HIP runtime error */
#define BIT_E_TREE (-4)      /* tree could not be built (inseparable particles) */
#define BIT_E_STATE (-5)     /* call out of order (e.g. evaluate before moments) */

/* precision modes */
#define BIT_FP64 0 /* moments, far and near field in FP64 */
#define BIT_FP32 1 /* FP64 moments and lists; far/near in FP32, Kahan-compensated */

/* largest supported Taylor order: nterms <= BIT_MAX_NTERMS (p = nterms-1) */
#define BIT_MAX_NTERMS 13
/* largest order of the regularized field family (bit_params.kernel_order) */
#define BIT_MAX_KERNEL_ORDER 8

typedef struct bit_ctx bit_ctx;
typedef struct bit_system bit_system;
typedef struct bit_tree bit_tree;

/* Evaluation parameters.  Defaults (bit_params_default) follow problem.h:11-13:
 * acc = ACC = 0.5, delta2 = 0, kappa = 0, mac_sub = delta2. */
typedef struct bit_params {
  int nterms;       /* Taylor terms: degrees 0..nterms-1 (reference `nterms`) */
  double acc;       /* MAC parameter (reference `acc`, ACC in problem.h:11) */
  double kappa;     /* screening (reference `alpha` of shielded_treepot) */
  double delta2;    /* regularization delta^2 (reference DEL, problem.h:12) */
  double mac_sub;   /* subtracted in the MAC; reference: DEL */
  int precision;    /* BIT_FP64 or BIT_FP32 */
  int kernel_order; /* field only: regularized family G_L = sum_{l<=L} w_l delta2^l
                       R^-(2l+1), w_l = prod_{j<=l} (j-1/2)/j (reference KERNEL_ORDER,
                       problem.h:13; tree3d.cpp:88, 614-791); 0 = 1/R.  Potential
                       entry points require 0. */
  int root;         /* node the traversal starts at (reference `ind` of treeforce,
                       shielded_treepot and treepotential, tree3d.cpp:851, 3448,
                       4679): the result is the contribution of that node's subtree,
                       the node's own MAC tested first.  0 = the whole tree; must be
                       in [0, numtree). */
} bit_params;

/* Interaction statistics of the last evaluation of a system (exact counts). */
typedef struct bit_stats {
  int64_t near_pairs;  /* source-target pairs summed directly */
  int64_t far_pairs;   /* cluster-target Taylor evaluations */
  int64_t far_items;   /* batch list entries (far), incl. masked lanes */
  int64_t near_items;  /* batch list entries (near) */
  int64_t lane_far;    /* lane slots of the FP32 far kernel: 64 x its sweeps (two
                        * consecutive items of at most 64 targets share one) */
  int64_t lane_near_pairs; /* lane-pair slots of the FP32 near kernel: 64 x the
                            * leaf size per target-parallel item, 4 per 64 sources
                            * per four targets for items of at most 40 targets */
  double ms_tree, ms_upload, ms_moments, ms_lists, ms_far, ms_near; /* last timings */
  /* tree traversals of the last list build: 1 when every batch fit the
   * per-batch regions sized from the previous build, else 2 (count, fill) */
  int64_t list_passes;
} bit_stats;

void bit_params_default(bit_params *p);

/* Thread-local message of the last failing call on this thread. */
const char *bit_last_error(void);

/* Context: one device, one stream (default priority), an HBM budget for
 * everything the context allocates (0 = no limit). */
int bit_ctx_create(int device, size_t hbm_budget_bytes, bit_ctx **out);
/* Fails with BIT_E_STATE while systems of the context are alive; frees the
 * context's bit_device_alloc buffers. */
int bit_ctx_destroy(bit_ctx *ctx);
/* Use an existing hipStream_t (e.g. torch's current stream) instead of the
 * context's own; NULL restores the context stream. */
int bit_ctx_set_stream(bit_ctx *ctx, void *hip_stream);
int bit_ctx_sync(bit_ctx *ctx);
/* Host threads the tree build of this context's systems may use (0 = all);
 * a request server building many trees at once gives each builder a few. */
int bit_ctx_set_host_threads(bit_ctx *ctx, int threads);
/* Bytes the context currently holds on the device. */
size_t bit_ctx_bytes_in_use(const bit_ctx *ctx);
/* Test hook: make the next k device allocations of this context fail with
 * BIT_E_OOM (k = 0 disables). */
int bit_ctx_inject_alloc_failures(bit_ctx *ctx, int k);

/* Device buffers for callers without a GPU framework (no counterpart in the
 * reference: plumbing so that inputs and outputs can stay resident in HBM).
 * Allocations count against the context's HBM budget. */
int bit_device_alloc(bit_ctx *ctx, size_t bytes, void **out);
int bit_device_free(bit_ctx *ctx, void *ptr);
int bit_copy_to_device(bit_ctx *ctx, void *dst, const void *src, size_t bytes);
int bit_copy_to_host(bit_ctx *ctx, void *dst, const void *src, size_t bytes);
int bit_copy_on_device(bit_ctx *ctx, void *dst, const void *src, size_t bytes);

/* Builds the tree and uploads the sorted sources.  Replaces treemake
 * (tree3d.cpp:1345-1551; tree3d.h:115): same bisection rules, node numbering,
 * children order and member lists (leaf size maxmemb = reference `maxmemb`).
 * x,y,z,q are HOST arrays of n doubles (q = PARTICLE::tot_charge). */
int bit_system_create(bit_ctx *ctx, int n, const double *x, const double *y,
                      const double *z, const double *q, int maxmemb, bit_system **out);
/* Same, from DEVICE arrays of n doubles already resident in HBM. */
int bit_system_create_device(bit_ctx *ctx, int n, const double *d_x, const double *d_y,
                             const double *d_z, const double *d_q, int maxmemb,
                             bit_system **out);
int bit_system_destroy(bit_system *sys);
int bit_system_numtree(const bit_system *sys);
int bit_system_maxdepth(const bit_system *sys);
/* Host copy of the tree: boxes[13*numtree] = x0,x1,y0,y1,z0,z1,xmid,ymid,zmid,
 * aspectxy,aspectxz,aspectyz,sqradius (TREE, tree3d.h:51-74); children[8*numtree];
 * begin/end[numtree] index perm[n]; node i's member list (TREE::members, ascending
 * original indices) is the sorted perm[begin[i] .. end[i]-1].  Slots inside a
 * leaf follow a Morton curve over the leaf box (the order targets are batched
 * in).  Any pointer may be NULL. */
int bit_system_export_tree(const bit_system *sys, double *boxes, int *children, int *begin,
                           int *end, int *perm);

/* Host-only tree build, no device needed: the same tree bit_system_create
 * builds (treemake, tree3d.cpp:1345-1551).  Export layout as
 * bit_system_export_tree. */
int bit_tree_build(int n, const double *x, const double *y, const double *z, int maxmemb,
                   bit_tree **out);
int bit_tree_destroy(bit_tree *t);
int bit_tree_numtree(const bit_tree *t);
int bit_tree_maxdepth(const bit_tree *t);
int bit_tree_export(const bit_tree *t, double *boxes, int *children, int *begin, int *end,
                    int *perm);

/* Cluster moments about each box midpoint, FP64, degrees < nterms.  Replaces
 * treemoments (tree3d.cpp:2631-2711; tree3d.h:126). */
int bit_moments(bit_system *sys, int nterms);
/* Host copy of the moments as the reference's cube: out[numtree][nterms]^3,
 * M[i][j][k] = sum q dx^i dy^j dz^k (entries with i+j+k >= nterms are 0). */
int bit_moments_download(const bit_system *sys, double *out);

/* Potential at every particle, self excluded.  Replaces the per-target loops
 * over shielded_treepot (tree3d.cpp:3448-3627) and treepotential
 * (tree3d.cpp:4679-4778), batched as in shielded_freespace_pot_ds
 * (tree3d.cpp:2218-2236).  out[n] receives +phi in the caller's ORIGINAL
 * particle order; out_on_device selects a device (1) or host (0) pointer.
 * Computes moments first if bit_moments was not called for params->nterms. */
int bit_eval_potential(bit_system *sys, const bit_params *params, double *out,
                       int out_on_device);
/* Moments and interaction lists for params, without evaluating (so a later
 * evaluation, alone or mixed, only runs the far and near kernels). */
int bit_system_prepare(bit_system *sys, const bit_params *params);

/* No counterpart in the reference: bit_eval_potential for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  It allocates nothing: the near-field
 * results and the evaluation descriptor go to a static device array of
 * 32 MiB + 1 KiB per GPU (held until the process ends), so the system may hold
 * at most 4,194,304 particles (BIT_E_INVALID beyond).  The system must have
 * been prepared for exactly these params (bit_system_prepare; BIT_E_STATE
 * otherwise).  d_out is a DEVICE array of n doubles that receives +phi in the
 * caller's original order, bit-identical to bit_eval_potential's.  Returns once
 * the work is queued on the context's stream (read d_out after bit_ctx_sync or
 * later work on that stream).  Every call on a GPU shares that GPU's array:
 * calls that overlap on one GPU (other streams, other contexts) race. */
int bit_eval_potential_noalloc(bit_system *sys, const bit_params *params, double *d_out);

/* Potentials of `count` independent systems (any sizes, per-system params) in
 * one persistent launch per precision on ctx's stream: as many work-groups as
 * the device holds at once take (system, 64-target batch) items from a device
 * counter -- systems of larger nterms, then larger size, first -- and run each
 * batch's near field and then its far field (the sweep chosen per item from
 * the system's nterms and kernel).  outs[i] receives system i's +phi
 * (original order); every target's sums are formed exactly as in
 * bit_eval_potential, so results are bit-identical alone and in any mix.
 * Systems may belong to other contexts on the same device.  kernel_ms (may be
 * NULL) gets the GPU time of the launches.  No reference counterpart: batches
 * the reference's per-system drivers (tree3d.cpp:2368-2436) across requests. */
int bit_eval_potential_multi(bit_ctx *ctx, int count, bit_system *const *systems,
                             const bit_params *params, double *const *outs, int out_on_device,
                             double *kernel_ms);

/* Field E_i = sum_{j != i} q_j (x_i - x_j)/R^3 for the regularized Coulomb
 * kernel (kappa must be 0); replaces treeforce (tree3d.cpp:851-971) as used by
 * freespace_particles_tree (tree3d.cpp:2268-2335).  out[3n] = Ex,Ey,Ez
 * interleaved per particle, original order. */
int bit_eval_field(bit_system *sys, const bit_params *params, double *out,
                   int out_on_device);
int bit_system_stats(const bit_system *sys, bit_stats *out);

/* Potential at m arbitrary points (host arrays, any order; out[m] in that
 * order): every source contributes, pairs with R^2 <= 1e-14 are skipped.
 * Replaces the reference's point evaluations treepotential(0, -1, x, y, z, ...)
 * (tree3d.cpp:4744-4757) and shielded_treepot(0, -1, ...) (3538-3551).
 * m = 0 returns BIT_OK and touches nothing. */
int bit_eval_potential_points(bit_system *sys, const bit_params *params, int m, const double *x,
                              const double *y, const double *z, double *out);
/* Field at m arbitrary points (kappa 0), out[3m]; treeforce(0, -1, ...)
 * (tree3d.cpp:937-950): every source contributes (the regularized kernel's
 * gradient vanishes at a coincident source).  m = 0 returns BIT_OK. */
int bit_eval_field_points(bit_system *sys, const bit_params *params, int m, const double *x,
                          const double *y, const double *z, double *out);

/* The batch interaction lists the evaluation uses for params' MAC (acc,
 * delta2, mac_sub), i.e. the reference traversal's per-target decisions
 * (tree3d.cpp:3473 accept -> far, childless -> near) grouped by batches of 64
 * consecutive sorted targets.  First call with NULL arrays to get the sizes
 * (nbatch, nfar, nnear), then with arrays of nbatch+1 / nfar / nnear entries.
 * Item t of batch b covers sorted slot 64*b + lane for every set lane bit.
 * Near items are in traversal order; a batch's far items are in the FP32 far
 * kernel's pairing order (sorted by target count, the largest paired with the
 * smallest that fits), so a target's far clusters are the traversal's as a
 * set. */
int bit_system_export_lists(bit_system *sys, const bit_params *params, int *nbatch, int *nfar,
                            int *nnear, int *far_off, uint64_t *far_mask, int *far_node,
                            int *near_off, uint64_t *near_mask, int *near_node);

/* Replaces the charges of a system (same positions, same tree and lists):
 * q[n] in the caller's original order, host (on_device 0) or device (1)
 * array.  Moments are recomputed by the next evaluation.  Used by iterative
 * solvers whose every matrix-vector product is a treecode evaluation. */
int bit_system_set_charges(bit_system *sys, const double *q, int on_device);

/* Second-kind integral equation on the particles of a system, solved by
 * restarted GMRES with FP64 Arnoldi (modified Gram-Schmidt), Givens rotations
 * and residual norms, every matrix-vector product a treecode evaluation:
 *     diag * s_i + sum_{j != i} w_j s_j G(x_i - x_j) + self * w_i s_i = f_i
 * (G as bit_eval_potential with params; w = quadrature weights).  The
 * reference's solver::gmres (solver.h:157-260) is dense and unrestarted; this
 * is its matrix-free counterpart, with the Krylov basis resident in HBM and
 * classical Gram-Schmidt applied twice (bit_krylov_* stepped to completion).
 * sigma holds the initial guess on entry.
 * iters/relres (may be NULL) receive the Arnoldi steps and ||r||/||f||. */
int bit_solve_second_kind(bit_system *sys, const bit_params *params, double diag, double self,
                          const double *w, const double *f, double *sigma, int restart,
                          double tol, int maxit, int *iters, double *relres);

/* The same solve as resumable state resident in HBM (basis, residual,
 * iterate, weights, right-hand side), for a request server that steps many
 * solves together: bit_serve_round advances it by one matrix-vector product
 * and the Arnoldi step after it.  sigma0 (may be NULL = zero) is the initial
 * guess.  No reference counterpart: the reference's solver::gmres
 * (solver.h:157-260) runs one dense system to completion. */
typedef struct bit_krylov bit_krylov;
int bit_krylov_create(bit_system *sys, const bit_params *params, double diag, double self,
                      const double *w, const double *f, const double *sigma0, int restart,
                      double tol, int maxit, bit_krylov **out);
int bit_krylov_destroy(bit_krylov *k);
/* done: 0 running, 1 converged (||b - A x|| <= tol ||b||), 2 stopped at maxit */
int bit_krylov_status(const bit_krylov *k, int *done, int *iters, double *relres);
int bit_krylov_solution(const bit_krylov *k, double *sigma);

/* One serving round on ctx's stream: the potentials of n_eval systems (host
 * outputs, as bit_eval_potential_multi) and one step of each running solve,
 * all products in the same persistent launch (one per precision) as
 * bit_eval_potential_multi's; each Gram-Schmidt pass, the
 * norms and the vector updates of all solves are one launch each, with one
 * read-back per pass.  Every result is bit-identical to what the request gets
 * alone (bit_eval_potential, bit_solve_second_kind).  A system may appear once
 * per round.  kernel_ms (may be NULL) gets the GPU time of the round. */
int bit_serve_round(bit_ctx *ctx, int n_eval, bit_system *const *systems,
                    const bit_params *params, double *const *outs, int n_solve,
                    bit_krylov *const *solves, double *kernel_ms);

/* Host-only restarted GMRES on a dense row-major n x n matrix (same FP64
 * core as bit_solve_second_kind; no device needed). */
int bit_gmres_dense(int n, const double *A, const double *b, double *x, int restart, double tol,
                    int maxit, int *iters, double *relres);

/* The same FP64 GMRES(restart) with the caller's operator: matvec(user, in,
 * out) computes out = A in (n doubles, host).  For solvers whose products are
 * scheduled elsewhere (e.g. batched with other requests). */
typedef void (*bit_matvec_fn)(void *user, const double *in, double *out);
int bit_gmres(int n, bit_matvec_fn matvec, void *user, const double *b, double *x, int restart,
              double tol, int maxit, int *iters, double *relres);

/* FP64 direct sum on the GPU at ntargets particle targets (original indices
 * tidx, host array): phi_t = sum_{j != tidx[t]} q_j exp(-kappa R)/R.
 * Replaces shielded_pot_particle_ds (tree3d.cpp:1966-2016) with +phi sign.
 * out is a host array of ntargets doubles. */
/* Diagnostic, with no counterpart in the reference: the FMA issue ceiling of
 * this library's instruction mix on the context's device (no packed FP32),
 * measured with eight independent FMA chains per thread at eight waves per
 * SIMD in the fastest FMA form (v_fmac: accumulator tied to the destination,
 * VGPR multiplicands); lane TFLOP/s (an FMA is two flops) for FP32 and FP64.
 * The kernels' measured VALU throughput is quoted against it beside the
 * published peaks. */
int bit_valu_ceiling(bit_ctx *ctx, double *fp32_tflops, double *fp64_tflops);

int bit_direct_potential(bit_system *sys, double kappa, double delta2, int ntargets,
                         const int *tidx, double *out);

/* FP64 direct-sum field at ntargets particle targets (original indices tidx,
 * host array), own particle skipped:
 *     E_t = sum_{j != tidx[t]} q_j (x_t - x_j) G(rs),  rs = |x_t - x_j|^2 + delta2,
 *     G = rs^-3/2 (+ the regularized family up to kernel_order, as bit_params).
 * out[3*ntargets] = Ex,Ey,Ez per target (host).  With delta2 = 0 and
 * kernel_order 0 it replaces force_particle_ds_noreg (tree3d.cpp:182-222),
 * with delta2 = DEL and kernel_order = KERNEL_ORDER force_particle_ds_reg
 * (tree3d.cpp:298-330); freespace_particles_ds_noreg/_reg (359-426) are all
 * particles. */
int bit_direct_field(bit_system *sys, double delta2, int kernel_order, int ntargets,
                     const int *tidx, double *out);
/* The same at m arbitrary points (host arrays), every source included as in
 * force_point_ds_noreg/_reg (tree3d.cpp:125-158, 241-274): a point on a
 * source with delta2 = 0 gives the reference's 0/0.  m = 0 returns BIT_OK. */
int bit_direct_field_points(bit_system *sys, double delta2, int kernel_order, int m,
                            const double *x, const double *y, const double *z, double *out);

#ifdef __cplusplus
}
#endif
#endif
