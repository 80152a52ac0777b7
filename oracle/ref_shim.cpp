// ref_shim.cpp -- C entry points over the reference library, for fixture
// generation and cross-checks only (test infrastructure).
//
// Compiled together with the reference's own src/tree3d.cpp, where it lies
// under the reference checkout, by oracle/Makefile (target `ref`); the result
// goes to oracle/_ref/libbitref.so.  This file is original code: it only calls
// the reference API declared in src/tree3d.h and exports flat arrays.
//
// The reference's DEL (problem.h:12) is a compile-time constant 0.01, so every
// quantity produced here uses delta^2 = 0.01 wherever the reference applies
// DEL (treeforce/treepotential) and the MAC shift 0.01 everywhere.
#include <cstdint>
#include <cstring>
#include <vector>

#include "tree3d.h"

namespace {

struct RefSystem {
  int n = 0;
  int numtree = 0;
  int maxtrees = 0;
  PARTICLE *part = nullptr;
  TREE *tree = nullptr;
};

// Particles with charge q, weight 1 (tot_charge = q): tree3d.h:33-47.
void fill_particles(RefSystem &s, int n, const double *x, const double *y, const double *z,
                    const double *q) {
  s.n = n;
  s.part = new PARTICLE[n];
  for (int i = 0; i < n; i++) {
    std::memset(&s.part[i], 0, sizeof(PARTICLE));
    s.part[i].charge = q[i];
    s.part[i].x = x[i];
    s.part[i].y = y[i];
    s.part[i].z = z[i];
    s.part[i].weight = 1.0;
    s.part[i].tot_charge = q[i];
  }
}

// Tree array as the reference's tests allocate it (tests/treecode.cpp:24-38),
// sized generously; moments optionally pre-allocated (nterms+1)^3 and zeroed so
// that treetaylorpotential's degree-nterms reads stay in bounds.
void build(RefSystem &s, int maxmemb, int nterms, bool padded_moments) {
  s.maxtrees = 4 * s.n + 64;
  s.tree = new TREE[s.maxtrees];
  for (int i = 0; i < s.maxtrees; i++) {
    s.tree[i].members = nullptr;
    s.tree[i].moments = nullptr;
  }
  s.numtree = treemake(s.n, nterms, maxmemb, s.part, s.tree);
  if (padded_moments) {
    int m = nterms + 1;
    for (int t = 0; t < s.numtree; t++) {
      s.tree[t].moments = new double **[m];
      for (int i = 0; i < m; i++) {
        s.tree[t].moments[i] = new double *[m];
        for (int j = 0; j < m; j++) {
          s.tree[t].moments[i][j] = new double[m];
          for (int k = 0; k < m; k++) s.tree[t].moments[i][j][k] = 0.0;
        }
      }
    }
  }
  treemoments(s.numtree, nterms, s.tree, s.part);
}

void destroy(RefSystem &s, int nterms_alloc) {
  for (int t = 0; t < s.maxtrees; t++) {
    delete[] s.tree[t].members;
    if (s.tree[t].moments) {
      for (int i = 0; i < nterms_alloc; i++) {
        for (int j = 0; j < nterms_alloc; j++) delete[] s.tree[t].moments[i][j];
        delete[] s.tree[t].moments[i];
      }
      delete[] s.tree[t].moments;
    }
  }
  delete[] s.tree;
  delete[] s.part;
}

}  // namespace

extern "C" {

// treemake + treemoments; exports the tree (boxes[13], children[8], nummemb,
// members) and the moments cube [numtree][nterms]^3.  Arrays must hold
// cap_nodes nodes and cap_members members; returns numtree (or -1 if too small).
int ref_tree(int n, const double *x, const double *y, const double *z, const double *q,
             int maxmemb, int nterms, int cap_nodes, int64_t cap_members, double *boxes,
             int *children, int *nummemb, int *members, double *moments) {
  RefSystem s;
  fill_particles(s, n, x, y, z, q);
  build(s, maxmemb, nterms, false);
  int nt = s.numtree;
  int64_t total = 0;
  for (int t = 0; t < nt; t++) total += s.tree[t].nummemb;
  if (nt > cap_nodes || total > cap_members) {
    destroy(s, nterms);
    return -1;
  }
  int64_t off = 0;
  for (int t = 0; t < nt; t++) {
    const TREE &d = s.tree[t];
    double b[13] = {d.x0, d.x1, d.y0, d.y1, d.z0, d.z1, d.xmid, d.ymid, d.zmid,
                    d.aspectxy, d.aspectxz, d.aspectyz, d.sqradius};
    std::memcpy(boxes + 13 * (int64_t)t, b, sizeof(b));
    std::memcpy(children + 8 * (int64_t)t, d.children, sizeof(int) * 8);
    nummemb[t] = d.nummemb;
    std::memcpy(members + off, d.members, sizeof(int) * d.nummemb);
    off += d.nummemb;
    if (moments) {
      double *M = moments + (int64_t)nterms * nterms * nterms * t;
      for (int i = 0; i < nterms; i++)
        for (int j = 0; j < nterms; j++)
          for (int k = 0; k < nterms; k++)
            M[(i * nterms + j) * nterms + k] = d.moments[i][j][k];
    }
  }
  destroy(s, nterms);
  return nt;
}

// shielded_treepot(0, i, ...) for particle targets tidx (tree3d.cpp:3448): the
// raw reference value (-phi plus its self term q_i*intG/w_i).
void ref_shielded_treepot(int n, const double *x, const double *y, const double *z,
                          const double *q, int maxmemb, int nterms, double acc, double kappa,
                          int nt, const int *tidx, double *out) {
  RefSystem s;
  fill_particles(s, n, x, y, z, q);
  build(s, maxmemb, nterms, false);
#pragma omp parallel for schedule(dynamic, 16)
  for (int t = 0; t < nt; t++) {
    int i = tidx[t];
    double pot = 0.0;
    shielded_treepot(0, i, nterms, s.part[i].x, s.part[i].y, s.part[i].z, kappa, acc, &pot,
                     s.tree, s.part, 100, 100, 100, 1.0, 1.0, 1.0);
    out[t] = pot;
  }
  destroy(s, nterms);
}

// treepotential(0, -1, ...) evaluated at the positions of particles tidx
// (tree3d.cpp:4679): regularized near and far field (rs = r^2 + DEL), the
// target's own charge included as q_i/sqrt(DEL).  Moments padded (F4).
void ref_treepotential_point(int n, const double *x, const double *y, const double *z,
                             const double *q, int maxmemb, int nterms, double acc, int nt,
                             const int *tidx, double *out) {
  RefSystem s;
  fill_particles(s, n, x, y, z, q);
  build(s, maxmemb, nterms, true);
#pragma omp parallel for schedule(dynamic, 16)
  for (int t = 0; t < nt; t++) {
    int i = tidx[t];
    double pot = 0.0;
    treepotential(0, -1, nterms, s.part[i].x, s.part[i].y, s.part[i].z, acc, &pot, s.tree,
                  s.part);
    out[t] = pot;
  }
  destroy(s, nterms + 1);
}

// treeforce(0, i, ...) fields for particle targets (tree3d.cpp:851).
void ref_treeforce(int n, const double *x, const double *y, const double *z, const double *q,
                   int maxmemb, int nterms, double acc, int nt, const int *tidx, double *ex,
                   double *ey, double *ez) {
  RefSystem s;
  fill_particles(s, n, x, y, z, q);
  build(s, maxmemb, nterms, false);
#pragma omp parallel for schedule(dynamic, 16)
  for (int t = 0; t < nt; t++) {
    int i = tidx[t];
    double fx = 0, fy = 0, fz = 0;
    treeforce(0, i, nterms, s.part[i].x, s.part[i].y, s.part[i].z, acc, &fx, &fy, &fz,
              s.tree, s.part);
    ex[t] = fx;
    ey[t] = fy;
    ez[t] = fz;
  }
  destroy(s, nterms);
}

// Calls that start the traversal at node inds[t] (the reference's `ind`) for
// particle tidx[t]: shielded_treepot(ind, i, ...) raw (tree3d.cpp:3448),
// treepotential(ind, -1, x_i, ...) (tree3d.cpp:4679, moments padded as in
// ref_treepotential_point) and treeforce(ind, i, ...) (tree3d.cpp:851).
void ref_subtree_calls(int n, const double *x, const double *y, const double *z,
                       const double *q, int maxmemb, int nterms, double acc, double kappa,
                       int npairs, const int *inds, const int *tidx, double *shielded,
                       double *pot_point, double *ex, double *ey, double *ez) {
  RefSystem s, sp;
  fill_particles(s, n, x, y, z, q);
  build(s, maxmemb, nterms, false);
  fill_particles(sp, n, x, y, z, q);
  build(sp, maxmemb, nterms, true);
#pragma omp parallel for schedule(dynamic, 16)
  for (int t = 0; t < npairs; t++) {
    const int ind = inds[t], i = tidx[t];
    double pot = 0.0;
    shielded_treepot(ind, i, nterms, s.part[i].x, s.part[i].y, s.part[i].z, kappa, acc, &pot,
                     s.tree, s.part, 100, 100, 100, 1.0, 1.0, 1.0);
    shielded[t] = pot;
    pot = 0.0;
    treepotential(ind, -1, nterms, sp.part[i].x, sp.part[i].y, sp.part[i].z, acc, &pot,
                  sp.tree, sp.part);
    pot_point[t] = pot;
    double fx = 0, fy = 0, fz = 0;
    treeforce(ind, i, nterms, s.part[i].x, s.part[i].y, s.part[i].z, acc, &fx, &fy, &fz,
              s.tree, s.part);
    ex[t] = fx;
    ey[t] = fy;
    ez[t] = fz;
  }
  destroy(s, nterms);
  destroy(sp, nterms + 1);
}

// shielded_pot_particle_ds (tree3d.cpp:1966): -sum_{j != i} q_j exp(-k r)/r.
void ref_shielded_pot_ds(int n, const double *x, const double *y, const double *z,
                         const double *q, double kappa, int nt, const int *tidx,
                         double *out) {
  RefSystem s;
  fill_particles(s, n, x, y, z, q);
#pragma omp parallel for schedule(static)
  for (int t = 0; t < nt; t++) {
    double pot = 0.0;
    shielded_pot_particle_ds(tidx[t], n, s.part, &pot, kappa);
    out[t] = pot;
  }
  delete[] s.part;
}

// force_particle_ds_reg (tree3d.cpp:298): regularized field, DEL.
void ref_force_ds_reg(int n, const double *x, const double *y, const double *z,
                      const double *q, int nt, const int *tidx, double *ex, double *ey,
                      double *ez) {
  RefSystem s;
  fill_particles(s, n, x, y, z, q);
#pragma omp parallel for schedule(static)
  for (int t = 0; t < nt; t++) {
    double fx = 0, fy = 0, fz = 0;
    force_particle_ds_reg(tidx[t], n, s.part, &fx, &fy, &fz);
    ex[t] = fx;
    ey[t] = fy;
    ez[t] = fz;
  }
  delete[] s.part;
}

// boxparticles (tree3d.cpp:1698): positions of the reference's test set.
void ref_boxparticles(int n, double *x, double *y, double *z, double *q) {
  PARTICLE *p = new PARTICLE[n];
  boxparticles(n, 1.0, 1.0, 1.0, p, 1.0);
  for (int i = 0; i < n; i++) {
    x[i] = p[i].x;
    y[i] = p[i].y;
    z[i] = p[i].z;
    q[i] = p[i].tot_charge;
  }
  delete[] p;
}

// The value the reference adds for the target itself in shielded_treepot
// (tree3d.cpp:3575-3596), per unit q/weight.
double ref_self_intG(void) { return -0.0000483811229725; }

}  // extern "C"
