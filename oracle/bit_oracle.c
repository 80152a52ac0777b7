/*
 * bit_oracle.c -- FP64 CPU restatement of ongbw/bit's treecode (test infrastructure
 * only: the checker for the HIP path, never linked into the product).
 *
 * Compile with -ffp-contract=off (oracle/Makefile): the tree boxes, the
 * multipole acceptance test and the moments must round exactly as the
 * reference's scalar x86-64 build does.
 */
#include "bit_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

struct ora_tree {
  int numtree;
  int cap;
  int quirk_hits; /* z-split checks whose slot differs from the y-split's upper half */
  ora_node *nodes;
};

/* treeinit, tree3d.cpp:1045-1073: box, midpoint, aspect ratios, squared
 * half-diagonal; children cleared. */
static void node_init(ora_node *t, double x0, double x1, double y0, double y1, double z0,
                      double z1) {
  double xm = 0.5 * (x1 + x0), ym = 0.5 * (y1 + y0), zm = 0.5 * (z1 + z0);
  t->x0 = x0; t->x1 = x1; t->y0 = y0; t->y1 = y1; t->z0 = z0; t->z1 = z1;
  t->xmid = xm; t->ymid = ym; t->zmid = zm;
  t->aspectxy = fabs((x1 - x0) / (y1 - y0));
  t->aspectxz = fabs((x1 - x0) / (z1 - z0));
  t->aspectyz = fabs((y1 - y0) / (z1 - z0));
  double hx = x1 - xm, hy = y1 - ym, hz = z1 - zm;
  t->sqradius = hx * hx + hy * hy + hz * hz;
  for (int c = 0; c < 8; c++) t->children[c] = -1;
}

/* treeshrink, tree3d.cpp:1099-1127: tight bounding box of the members. */
static void node_shrink(ora_node *t, const double *x, const double *y, const double *z) {
  int j = t->members[0];
  double mnx = x[j], mxx = x[j], mny = y[j], mxy = y[j], mnz = z[j], mxz = z[j];
  for (int i = 0; i < t->nummemb; i++) {
    j = t->members[i];
    if (x[j] > mxx) mxx = x[j];
    if (x[j] < mnx) mnx = x[j];
    if (y[j] > mxy) mxy = y[j];
    if (y[j] < mny) mny = y[j];
    if (z[j] > mxz) mxz = z[j];
    if (z[j] < mnz) mnz = z[j];
  }
  node_init(t, mnx, mxx, mny, mxy, mnz, mxz);
}

/* treedivide, tree3d.cpp:1163-1304: stable split of `split`'s members at its
 * midpoint along dir (coordinate <= mid goes first).  The first non-empty side
 * goes to c1, the second to c2; c1 may alias split.  Returns #created. */
static int node_divide(ora_tree *T, int dir, int split, int c1, int c2, const double *x,
                       const double *y, const double *z) {
  ora_node *s = &T->nodes[split];
  const double *coord = dir == 0 ? x : (dir == 1 ? y : z);
  double mid = dir == 0 ? s->xmid : (dir == 1 ? s->ymid : s->zmid);
  int n = s->nummemb;
  int *side[2];
  int cnt[2] = {0, 0};
  side[0] = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  side[1] = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
  for (int j = 0; j < n; j++) {
    int idx = s->members[j];
    if (coord[idx] <= mid) side[0][cnt[0]++] = idx;
    else side[1][cnt[1]++] = idx;
  }
  int target = c1, made = 0;
  for (int sd = 0; sd < 2; sd++) {
    if (cnt[sd] > 0) {
      ora_node *p = &T->nodes[target];
      p->nummemb = cnt[sd];
      free(p->members);
      p->members = (int *)malloc(sizeof(int) * (size_t)cnt[sd]);
      memcpy(p->members, side[sd], sizeof(int) * (size_t)cnt[sd]);
      node_shrink(p, x, y, z);
      made++;
      target = c2;
    }
  }
  free(side[0]);
  free(side[1]);
  return made;
}

static int ensure_cap(ora_tree *T, int need) {
  if (need <= T->cap) return 1;
  int nc = T->cap * 2;
  while (nc < need) nc *= 2;
  ora_node *nn = (ora_node *)realloc(T->nodes, sizeof(ora_node) * (size_t)nc);
  if (!nn) return 0;
  memset(nn + T->cap, 0, sizeof(ora_node) * (size_t)(nc - T->cap));
  T->nodes = nn;
  T->cap = nc;
  return 1;
}

/* Split predicates of treemake, tree3d.cpp:1418-1541 (aspectlimit = 1/sqrt(2)). */
#define LIM (1.0 / sqrt(2.0))
static int want_y(const ora_node *c, int maxmemb) {
  return c->aspectxy < LIM && c->aspectyz > LIM && c->nummemb > maxmemb;
}
static int want_z(const ora_node *c, int maxmemb) {
  return c->aspectxz < LIM && c->aspectyz < LIM && c->nummemb > maxmemb;
}

void ora_tree_free(ora_tree *T) {
  if (!T) return;
  for (int i = 0; i < T->cap; i++) free(T->nodes[i].members);
  free(T->nodes);
  free(T);
}

/* treemake, tree3d.cpp:1345-1551.  Nodes are processed in index order; node i's
 * children occupy a block starting at the current numtree.  The z-split check
 * after a y-split addresses node numtree+j+k*(nchild2+1) (tree3d.cpp:1458-1477),
 * which is reproduced literally; block slots not yet created are empty
 * (nummemb 0) and never split.  A node cap of 4n+64 guards inputs that cannot
 * be separated (coincident particles), where the reference overruns its array. */
ora_tree *ora_tree_build(int n, const double *x, const double *y, const double *z,
                         int maxmemb) {
  if (n <= 0) return NULL;
  ora_tree *T = (ora_tree *)calloc(1, sizeof(ora_tree));
  T->cap = 64;
  T->nodes = (ora_node *)calloc((size_t)T->cap, sizeof(ora_node));
  const int node_cap = 4 * n + 64;
  ora_node *r = &T->nodes[0];
  r->nummemb = n;
  r->members = (int *)malloc(sizeof(int) * (size_t)n);
  double mnx = x[0], mxx = x[0], mny = y[0], mxy = y[0], mnz = z[0], mxz = z[0];
  for (int i = 0; i < n; i++) {
    r->members[i] = i;
    if (mnx > x[i]) mnx = x[i];
    if (mxx < x[i]) mxx = x[i];
    if (mny > y[i]) mny = y[i];
    if (mxy < y[i]) mxy = y[i];
    if (mnz > z[i]) mnz = z[i];
    if (mxz < z[i]) mxz = z[i];
  }
  node_init(r, mnx, mxx, mny, mxy, mnz, mxz);
  int numtree = 1;
  for (int i = 0; i < numtree; i++) {
    if (T->nodes[i].nummemb <= maxmemb) continue;
    /* a node's block holds at most 8 children plus the quirk's reach (+4) */
    if (numtree + 16 > node_cap || !ensure_cap(T, numtree + 16)) goto fail;
    int nchild = 0, nchild2 = 0;
    int kids[16];
    ora_node *nd = &T->nodes[i];
    if (nd->aspectxy > LIM && nd->aspectxz > LIM) {
      int xc = node_divide(T, 0, i, numtree, numtree + 1, x, y, z);
      int ind2 = numtree + xc;
      for (int j = 0; j < xc; j++) {
        kids[nchild++] = numtree + j;
        if (want_y(&T->nodes[numtree + j], maxmemb)) {
          int yc = node_divide(T, 1, numtree + j, numtree + j, ind2, x, y, z);
          int yup = ind2;
          if (yc == 2) { kids[nchild++] = ind2; ind2++; }
          for (int k = 0; k < yc; k++) {
            int t = numtree + j + k * (nchild2 + 1);
            if (k == 1 && t != yup) T->quirk_hits++;
            if (want_z(&T->nodes[t], maxmemb)) {
              int zc = node_divide(T, 2, t, t, ind2, x, y, z);
              if (zc == 2) { kids[nchild++] = ind2; nchild2++; ind2++; }
            }
          }
        } else if (want_z(&T->nodes[numtree + j], maxmemb)) {
          int zc = node_divide(T, 2, numtree + j, numtree + j, ind2, x, y, z);
          if (zc == 2) { kids[nchild++] = ind2; ind2++; }
        }
        if (nchild > 8) goto fail;
      }
    } else if (nd->aspectxy < LIM && nd->aspectyz > LIM) {
      int yc = node_divide(T, 1, i, numtree, numtree + 1, x, y, z);
      int ind2 = numtree + yc;
      for (int j = 0; j < yc; j++) {
        kids[nchild++] = numtree + j;
        if (want_z(&T->nodes[numtree + j], maxmemb)) {
          int zc = node_divide(T, 2, numtree + j, numtree + j, ind2, x, y, z);
          if (zc == 2) { kids[nchild++] = ind2; ind2++; }
        }
      }
    } else {
      int zc = node_divide(T, 2, i, numtree, numtree + 1, x, y, z);
      for (int j = 0; j < zc; j++) kids[nchild++] = numtree + j;
    }
    if (nchild > 8) goto fail;
    for (int c = 0; c < nchild; c++) T->nodes[i].children[c] = kids[c];
    numtree += nchild;
  }
  T->numtree = numtree;
  return T;
fail:
  ora_tree_free(T);
  return NULL;
}

int ora_tree_numtree(const ora_tree *t) { return t ? t->numtree : 0; }
int ora_tree_quirk_hits(const ora_tree *t) { return t ? t->quirk_hits : 0; }
const ora_node *ora_tree_node(const ora_tree *t, int i) { return &t->nodes[i]; }

int64_t ora_tree_total_members(const ora_tree *t) {
  int64_t s = 0;
  for (int i = 0; i < t->numtree; i++) s += t->nodes[i].nummemb;
  return s;
}

void ora_tree_export(const ora_tree *t, double *boxes, int *children, int *nummemb,
                     int *members) {
  int64_t off = 0;
  for (int i = 0; i < t->numtree; i++) {
    const ora_node *d = &t->nodes[i];
    double b[13] = {d->x0, d->x1, d->y0, d->y1, d->z0, d->z1, d->xmid, d->ymid,
                    d->zmid, d->aspectxy, d->aspectxz, d->aspectyz, d->sqradius};
    if (boxes) memcpy(boxes + 13 * (int64_t)i, b, sizeof(b));
    if (children) memcpy(children + 8 * (int64_t)i, d->children, sizeof(int) * 8);
    if (nummemb) nummemb[i] = d->nummemb;
    if (members) {
      memcpy(members + off, d->members, sizeof(int) * (size_t)d->nummemb);
    }
    off += d->nummemb;
  }
}

/* treemoments, tree3d.cpp:2631-2711: M[i][j][k] = sum q dx^i dy^j dz^k about the
 * box midpoint, i+j+k < nterms, with the reference's product order
 * (q * dz^k within dy^j within dx^i). */
void ora_moments(const ora_tree *t, const double *x, const double *y, const double *z,
                 const double *q, int nterms, double *out) {
  const int64_t nc = (int64_t)nterms * nterms * nterms;
#pragma omp parallel for schedule(dynamic, 1)
  for (int ind = 0; ind < t->numtree; ind++) {
    const ora_node *d = &t->nodes[ind];
    double *M = out + nc * ind;
    for (int64_t c = 0; c < nc; c++) M[c] = 0.0;
    for (int m = 0; m < d->nummemb; m++) {
      int p = d->members[m];
      double dx = x[p] - d->xmid, dy = y[p] - d->ymid, dz = z[p] - d->zmid;
      double mult = q[p], xmul = mult, ymul = mult;
      for (int i = 0; i < nterms; i++) {
        for (int j = 0; j < nterms - i; j++) {
          for (int k = 0; k < nterms - i - j; k++) {
            M[((int64_t)i * nterms + j) * nterms + k] += mult;
            mult *= dz;
          }
          ymul *= dy;
          mult = ymul;
        }
        xmul *= dx;
        ymul = xmul;
        mult = xmul;
      }
    }
  }
}

/* Taylor coefficients of G(x-y) = exp(-kappa R)/R in the source offset y,
 * shielded_treetaylorpot, tree3d.cpp:3004-3227.  The reference computes each
 * index family with its own loop; every family applies the same two rules along
 * one direction d:
 *   b_k = (kappa/k_d) (x_d a_{k-e_d} - a_{k-2e_d})
 *   R^2 a_k = sum_e (2 x_e a_{k-e} - a_{k-2e})
 *             - (1/k_d)(x_d a_{k-e_d} - a_{k-2e_d}) + (kappa/k_d)(x_d b_{k-e_d} - b_{k-2e_d})
 * with b the coefficients of exp(-kappa R).  The direction is the first index
 * equal to 1 when there is one (families a[1][i][0], a[i][1][1], ...), else the
 * first nonzero index (a[i][0][0], a[i][j][0], a[0][i][j], a[i][j][k]).  Here
 * the families are walked as one loop over total degree. */
void ora_taylor_coeffs(int nterms, double dx, double dy, double dz, double R2,
                       double kappa, double *a) {
  const int S = nterms + 1;
  double *b = (double *)calloc((size_t)S * S * S, sizeof(double));
  const double xv[3] = {dx, dy, dz};
  const double fac = 1.0 / R2;
  const double R = sqrt(R2);
#define IX(i, j, k) (((i) * S + (j)) * S + (k))
  for (int c = 0; c < S * S * S; c++) a[c] = 0.0;
  b[IX(0, 0, 0)] = exp(-kappa * R);
  a[IX(0, 0, 0)] = b[IX(0, 0, 0)] / R;
  for (int n = 1; n < nterms; n++) {
    for (int i = n; i >= 0; i--) {
      for (int j = n - i; j >= 0; j--) {
        int k = n - i - j;
        int kv[3] = {i, j, k};
        int d = -1;
        for (int e = 0; e < 3; e++)
          if (kv[e] == 1) { d = e; break; }
        if (d < 0)
          for (int e = 0; e < 3; e++)
            if (kv[e] > 0) { d = e; break; }
        double sum = 0.0;
        for (int e = 0; e < 3; e++) {
          if (kv[e] >= 1) {
            int m[3] = {i, j, k};
            m[e] -= 1;
            sum += 2.0 * xv[e] * a[IX(m[0], m[1], m[2])];
          }
          if (kv[e] >= 2) {
            int m[3] = {i, j, k};
            m[e] -= 2;
            sum -= a[IX(m[0], m[1], m[2])];
          }
        }
        int m1[3] = {i, j, k}, m2[3] = {i, j, k};
        m1[d] -= 1;
        m2[d] -= 2;
        double a1 = a[IX(m1[0], m1[1], m1[2])];
        double b1 = b[IX(m1[0], m1[1], m1[2])];
        double a2 = kv[d] >= 2 ? a[IX(m2[0], m2[1], m2[2])] : 0.0;
        double b2 = kv[d] >= 2 ? b[IX(m2[0], m2[1], m2[2])] : 0.0;
        double inv = 1.0 / kv[d];
        double v = inv * (xv[d] * a1 - a2);
        b[IX(i, j, k)] = kappa * v;
        a[IX(i, j, k)] = fac * (sum - v + kappa * inv * (xv[d] * b1 - b2));
      }
    }
  }
#undef IX
  free(b);
}

/* MAC of the traversals, tree3d.cpp:3469-3473 / 4687-4691:
 * rs = dx^2+dy^2+dz^2 (+ delta2); accept = sqradius < acc*(rs - mac_sub). */
static int mac_accept(const ora_node *d, double px, double py, double pz, ora_params prm,
                      double *dx, double *dy, double *dz, double *rs) {
  *dx = px - d->xmid;
  *dy = py - d->ymid;
  *dz = pz - d->zmid;
  *rs = *dx * *dx + *dy * *dy + *dz * *dz + prm.delta2;
  return d->sqradius < prm.acc * (*rs - prm.mac_sub);
}

typedef struct {
  double pot;
  int64_t nnear, nfar;
} pot_acc;

/* shielded_treepot / treepotential traversal, tree3d.cpp:3448-3627 and
 * 4679-4778: accept -> far field; reject -> recurse over children[0..7];
 * childless -> direct sum over members excluding the target itself. */
static void pot_rec(const ora_tree *T, int ind, int p, double px, double py, double pz,
                    const double *x, const double *y, const double *z, const double *q,
                    const double *mom, ora_params prm, double *coef, pot_acc *acc) {
  const ora_node *d = &T->nodes[ind];
  double dx, dy, dz, rs;
  if (mac_accept(d, px, py, pz, prm, &dx, &dy, &dz, &rs)) {
    const int nt = prm.nterms, S = nt + 1;
    ora_taylor_coeffs(nt, dx, dy, dz, rs, prm.kappa, coef);
    const double *M = mom + (int64_t)nt * nt * nt * ind;
    double s = 0.0;
    for (int i = 0; i < nt; i++)
      for (int j = 0; j < nt - i; j++)
        for (int k = 0; k < nt - i - j; k++)
          s += M[(i * nt + j) * nt + k] * coef[(i * S + j) * S + k];
    acc->pot += s;
    acc->nfar++;
    return;
  }
  int go = 1;
  for (int c = 0; c < 8; c++) {
    if (d->children[c] != -1) {
      go = 0;
      pot_rec(T, d->children[c], p, px, py, pz, x, y, z, q, mom, prm, coef, acc);
    }
  }
  if (go) {
    for (int m = 0; m < d->nummemb; m++) {
      int j = d->members[m];
      if (j == p) continue;
      double ex = px - x[j], ey = py - y[j], ez = pz - z[j];
      double R2 = ex * ex + ey * ey + ez * ez + prm.delta2;
      /* point targets (p == -1): coincident sources are skipped as in
       * tree3d.cpp:3547 / 4753 (rs > eps) */
      if (p < 0 && !(R2 > 1e-14)) continue;
      double R = sqrt(R2);
      acc->pot += q[j] * exp(-prm.kappa * R) / R;
      acc->nnear++;
    }
  }
}

static int thread_count(int nthreads) {
#ifdef _OPENMP
  return nthreads > 0 ? nthreads : omp_get_max_threads();
#else
  (void)nthreads;
  return 1;
#endif
}

void ora_potential(const ora_tree *T, const double *x, const double *y, const double *z,
                   const double *q, const double *moments, ora_params prm, int nt,
                   const int *tidx, double *pot, int64_t *n_near, int64_t *n_far,
                   int nthreads) {
  const int S = prm.nterms + 1;
#pragma omp parallel num_threads(thread_count(nthreads))
  {
    double *coef = (double *)malloc(sizeof(double) * (size_t)S * S * S);
#pragma omp for schedule(dynamic, 16)
    for (int t = 0; t < nt; t++) {
      pot_acc acc = {0.0, 0, 0};
      int p = tidx[t];
      pot_rec(T, prm.root, p, x[p], y[p], z[p], x, y, z, q, moments, prm, coef, &acc);
      pot[t] = acc.pot;
      if (n_near) n_near[t] = acc.nnear;
      if (n_far) n_far[t] = acc.nfar;
    }
    free(coef);
  }
}

void ora_potential_points(const ora_tree *T, const double *x, const double *y,
                          const double *z, const double *q, const double *moments,
                          ora_params prm, int m, const double *px, const double *py,
                          const double *pz, double *pot, int nthreads) {
  const int S = prm.nterms + 1;
#pragma omp parallel num_threads(thread_count(nthreads))
  {
    double *coef = (double *)malloc(sizeof(double) * (size_t)S * S * S);
#pragma omp for schedule(dynamic, 16)
    for (int t = 0; t < m; t++) {
      pot_acc acc = {0.0, 0, 0};
      pot_rec(T, prm.root, -1, px[t], py[t], pz[t], x, y, z, q, moments, prm, coef, &acc);
      pot[t] = acc.pot;
    }
    free(coef);
  }
}

static void lists_rec(const ora_tree *T, int ind, int p, const double *x, const double *y,
                      const double *z, ora_params prm, int *far, int *nfar, int *near,
                      int *nnear, int cap) {
  const ora_node *d = &T->nodes[ind];
  double dx, dy, dz, rs;
  if (mac_accept(d, x[p], y[p], z[p], prm, &dx, &dy, &dz, &rs)) {
    if (*nfar < cap) far[*nfar] = ind;
    (*nfar)++;
    return;
  }
  int go = 1;
  for (int c = 0; c < 8; c++)
    if (d->children[c] != -1) {
      go = 0;
      lists_rec(T, d->children[c], p, x, y, z, prm, far, nfar, near, nnear, cap);
    }
  if (go) {
    if (*nnear < cap) near[*nnear] = ind;
    (*nnear)++;
  }
}

void ora_lists(const ora_tree *T, const double *x, const double *y, const double *z,
               ora_params prm, int ti, int *far, int *nfar, int *near, int *nnear, int cap) {
  *nfar = 0;
  *nnear = 0;
  lists_rec(T, prm.root, ti, x, y, z, prm, far, nfar, near, nnear, cap);
}

/* shielded_pot_particle_ds, tree3d.cpp:1966-2016 (sign: +phi), generalised to
 * R = sqrt(r^2 + delta2). */
void ora_direct(int n, const double *x, const double *y, const double *z, const double *q,
                double kappa, double delta2, int nt, const int *tidx, double *pot,
                int nthreads) {
#pragma omp parallel for schedule(static) num_threads(thread_count(nthreads))
  for (int t = 0; t < nt; t++) {
    int p = tidx[t];
    double s = 0.0;
    for (int j = 0; j < n; j++) {
      if (j == p) continue;
      double ex = x[p] - x[j], ey = y[p] - y[j], ez = z[p] - z[j];
      double R = sqrt(ex * ex + ey * ey + ez * ez + delta2);
      s += q[j] * exp(-kappa * R) / R;
    }
    pot[t] = s;
  }
}

/* The regularized kernel family of treetaylorforce / eval_grad_kernel_reg
 * (tree3d.cpp:614-791, 73-97; order L = KERNEL_ORDER of problem.h):
 *   G_L(R) = sum_{l=0..L} w_l delta2^l R^-(2l+1),  w_l = prod_{j<=l} (j - 1/2)/j
 * (w_l = choose(-1/2, l) (-1)^l, the reference's binomial/factorial).  Its
 * field is  q (x_t - x_s) sum_l f_l  with f_0 = 1/(R^2 R) and
 * f_l = f_{l-1} (l + 1/2)/l delta2/R^2, as eval_grad_kernel_reg sums it.
 *
 * Coefficients of g = R^-(2l+1) = u^-s, u = R^2, s = l + 1/2: from
 * u(y) dg/dy_e = 2 s (x_e - y_e) g, matching powers of y gives
 *   |k| R^2 a_k = (2|k| + 2l - 1) sum_e x_e a_{k-e} - (|k| + 2l - 1) sum_e a_{k-2e},
 * the relation the reference applies family by family (its signs differ by
 * (-1)^|k|: it expands in the target offset). */
void ora_powinv_coeffs(int nterms, int l, double dx, double dy, double dz, double R2,
                       double *a) {
  const int S = nterms + 1;
  const double xv[3] = {dx, dy, dz};
  const double fac = 1.0 / R2;
#define IX(i, j, k) (((i) * S + (j)) * S + (k))
  for (int c = 0; c < S * S * S; c++) a[c] = 0.0;
  a[IX(0, 0, 0)] = pow(R2, -(2.0 * l + 1.0) / 2.0);
  for (int n = 1; n < nterms; n++)
    for (int i = n; i >= 0; i--)
      for (int j = n - i; j >= 0; j--) {
        int k = n - i - j, kv[3] = {i, j, k};
        double s1 = 0.0, s2 = 0.0;
        for (int e = 0; e < 3; e++) {
          int m[3] = {i, j, k};
          if (kv[e] >= 1) {
            m[e] -= 1;
            s1 += xv[e] * a[IX(m[0], m[1], m[2])];
          }
          if (kv[e] >= 2) {
            m[e] -= 1;
            s2 += a[IX(m[0], m[1], m[2])];
          }
        }
        a[IX(i, j, k)] = fac / n * ((2.0 * n + 2.0 * l - 1.0) * s1 - (n + 2.0 * l - 1.0) * s2);
      }
#undef IX
}

static double family_weight(int l) {
  double w = 1.0;
  for (int j = 1; j <= l; j++) w *= (j - 0.5) / j;
  return w;
}

/* Near-field weight of one source: q/(R^2 R) for L = 0 (the expression used
 * since the first field restatement), else eval_grad_kernel_reg's sum
 * (tree3d.cpp:80-91) times q, in its operation order. */
static double family_near(double q, double r2, double delta2, int L) {
  if (L == 0) return q / (r2 * sqrt(r2));
  double r = sqrt(r2);
  double temp = 1.0 / (r * r2), factor = temp;
  for (int l = 1; l <= L; l++) {
    factor = factor * (l + 0.5) / l * delta2 / r2;
    temp += factor;
  }
  return q * temp;
}

/* treeforce, tree3d.cpp:851-971, with the far field of treetaylorforce
 * (tree3d.cpp:456-810, KERNEL_ORDER 0): for potential coefficients a_k of 1/R
 * in the source offset, the field is E_x = sum_k (k_x+1) M_k a_{k+e_x}, which is
 * the reference's signed sum over derivs[i+1][k][m]*(i+1).  The near field is
 * eval_grad_kernel_reg (tree3d.cpp:73-97): q (x_t - x_s)/R^3. */
static void force_rec(const ora_tree *T, int ind, int p, double px, double py, double pz,
                      const double *x, const double *y, const double *z, const double *q,
                      const double *mom, ora_params prm, double *coef, double *E) {
  const ora_node *d = &T->nodes[ind];
  double dx, dy, dz, rs;
  if (mac_accept(d, px, py, pz, prm, &dx, &dy, &dz, &rs)) {
    const int nt = prm.nterms, S = nt + 2;
    const double *M = mom + (int64_t)nt * nt * nt * ind;
    for (int l = 0; l <= prm.kernel_order; l++) {
      double w = 1.0;
      if (l == 0) {
        ora_taylor_coeffs(nt + 1, dx, dy, dz, rs, 0.0, coef);
      } else {
        ora_powinv_coeffs(nt + 1, l, dx, dy, dz, rs, coef);
        w = family_weight(l) * pow(prm.delta2, l);
      }
      for (int i = 0; i < nt; i++)
        for (int j = 0; j < nt - i; j++)
          for (int k = 0; k < nt - i - j; k++) {
            double m = w * M[(i * nt + j) * nt + k];
            E[0] += m * (i + 1) * coef[((i + 1) * S + j) * S + k];
            E[1] += m * (j + 1) * coef[(i * S + j + 1) * S + k];
            E[2] += m * (k + 1) * coef[(i * S + j) * S + k + 1];
          }
    }
    return;
  }
  int go = 1;
  for (int c = 0; c < 8; c++)
    if (d->children[c] != -1) {
      go = 0;
      force_rec(T, d->children[c], p, px, py, pz, x, y, z, q, mom, prm, coef, E);
    }
  if (go) {
    for (int m = 0; m < d->nummemb; m++) {
      int j = d->members[m];
      if (j == p) continue;
      double ex = px - x[j], ey = py - y[j], ez = pz - z[j];
      double r2 = ex * ex + ey * ey + ez * ez + prm.delta2;
      double w = family_near(q[j], r2, prm.delta2, prm.kernel_order);
      E[0] += ex * w;
      E[1] += ey * w;
      E[2] += ez * w;
    }
  }
}

void ora_treeforce(const ora_tree *T, const double *x, const double *y, const double *z,
                   const double *q, const double *moments, ora_params prm, int nt,
                   const int *tidx, double *ex, double *ey, double *ez, int nthreads) {
  const int S = prm.nterms + 2;
#pragma omp parallel num_threads(thread_count(nthreads))
  {
    double *coef = (double *)malloc(sizeof(double) * (size_t)S * S * S);
#pragma omp for schedule(dynamic, 16)
    for (int t = 0; t < nt; t++) {
      double E[3] = {0, 0, 0};
      int p = tidx[t];
      force_rec(T, prm.root, p, x[p], y[p], z[p], x, y, z, q, moments, prm, coef, E);
      ex[t] = E[0];
      ey[t] = E[1];
      ez[t] = E[2];
    }
    free(coef);
  }
}

void ora_treeforce_points(const ora_tree *T, const double *x, const double *y,
                          const double *z, const double *q, const double *moments,
                          ora_params prm, int m, const double *px, const double *py,
                          const double *pz, double *ex, double *ey, double *ez, int nthreads) {
  const int S = prm.nterms + 2;
#pragma omp parallel num_threads(thread_count(nthreads))
  {
    double *coef = (double *)malloc(sizeof(double) * (size_t)S * S * S);
#pragma omp for schedule(dynamic, 16)
    for (int t = 0; t < m; t++) {
      double E[3] = {0, 0, 0};
      force_rec(T, prm.root, -1, px[t], py[t], pz[t], x, y, z, q, moments, prm, coef, E);
      ex[t] = E[0];
      ey[t] = E[1];
      ez[t] = E[2];
    }
    free(coef);
  }
}

/* force_particle_ds_reg, tree3d.cpp:298-330 (field q (x_t - x_s)/R^3). */
void ora_direct_force(int n, const double *x, const double *y, const double *z,
                      const double *q, double delta2, int kernel_order, int nt, const int *tidx,
                      double *ex, double *ey, double *ez, int nthreads) {
#pragma omp parallel for schedule(static) num_threads(thread_count(nthreads))
  for (int t = 0; t < nt; t++) {
    int p = tidx[t];
    double E[3] = {0, 0, 0};
    for (int j = 0; j < n; j++) {
      if (j == p) continue;
      double dx = x[p] - x[j], dy = y[p] - y[j], dz = z[p] - z[j];
      double r2 = dx * dx + dy * dy + dz * dz + delta2;
      double w = family_near(q[j], r2, delta2, kernel_order);
      E[0] += dx * w;
      E[1] += dy * w;
      E[2] += dz * w;
    }
    ex[t] = E[0];
    ey[t] = E[1];
    ez[t] = E[2];
  }
}
