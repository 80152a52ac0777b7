// tree3d.h -- the reference library interface (tree3d.h of ongbw/bit) for the
// treecode path, served by the MI355X library through include/bit_amd.h.
//
// PARTICLE and TREE keep the reference's field order and types (tree3d.h:33-74)
// so caller-allocated arrays are interchangeable.  Functions keep the
// reference's names, signatures, accumulate-into-output conventions and
// "no error codes" behaviour: failures print a message on std::cout (as the
// reference does for its warnings) and leave outputs unchanged.
//
// Where the work goes:
//   treemake            host, bit-exact bisection tree (bit_tree_build)
//   treemoments         GPU FP64 moments, copied into TREE::moments
//   treeforce, treepotential, shielded_treepot  (particle targets p >= 0)
//                       one batched GPU evaluation of all particles the first
//                       time a (tree, start node ind, parameters) triple is
//                       asked for; later calls, from any thread, read that
//                       result.  ind != 0 starts the traversal at that node
//                       (the subtree's contribution); p == -1 is one GPU
//                       evaluation at the point.
//   freespace_particles_tree, freespace_pot_tree, shielded_freespace_pot_tree
//                       batched GPU evaluation
//   direct sums (force_*_ds_*, freespace_particles_ds_*, shielded_pot_particle_ds,
//   shielded_freespace_pot_ds)
//                       GPU FP64 direct sum over all particles; per-particle
//                       calls are served from one evaluation of all particles
//                       while the array holds the same positions and charges
//   eval_grad_kernel_*, treetaylorpotential, shielded_treetaylorpot, treetaylorforce
//                       one pair / one cluster on the host (primitives)
#ifndef tree3D_h
#define tree3D_h 1

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <string>

#include "constants.h"

using namespace std;

struct PARTICLE {
  double charge;
  double x;
  double y;
  double z;
  double u;
  double v;
  double w;
  double xforce;
  double yforce;
  double zforce;
  double weight;
  double tot_charge;  // charge x weight: the source strength q
};

struct TREE {
  int nummemb;
  double x0;
  double x1;
  double y0;
  double y1;
  double z0;
  double z1;
  double xmid;
  double ymid;
  double zmid;
  double aspectxy;
  double aspectxz;
  double aspectyz;
  double sqradius;
  int children[NUMCHILD];
  int* members;
  double*** moments;

  TREE() : nummemb(0), members(0x0), moments(0x0) {}
};

// ---- tree ----
int treemake(int num, int nterms, int maxmemb, PARTICLE* part, TREE* tree);
void treeempty(int numtree, TREE* tree);
void treeinit(TREE* tree1, double x0, double x1, double y0, double y1, double z0, double z1);
int treedivide(int dir, TREE* split, TREE* c1, TREE* c2, PARTICLE* part);
void treeshrink(TREE* tree1, PARTICLE* part);
void treemoments(int numtree, int nterms, TREE* tree, PARTICLE* part);

// ---- kernel primitives (one pair; the outputs are set) ----
void eval_grad_kernel_noreg(double dx, double dy, double dz, double rs, double* xfor,
                            double* yfor, double* zfor);
void eval_grad_kernel_reg(double dx, double dy, double dz, double rs, double* xfor,
                          double* yfor, double* zfor);

// ---- direct sums (outputs set, as the reference) ----
void force_point_ds_noreg(double x, double y, double z, int npart, PARTICLE* part,
                          double* xfor, double* yfor, double* zfor);
void force_particle_ds_noreg(int ind, int npart, PARTICLE* part, double* xfor, double* yfor,
                             double* zfor);
void force_point_ds_reg(double x, double y, double z, int npart, PARTICLE* part, double* xfor,
                        double* yfor, double* zfor);
void force_particle_ds_reg(int ind, int npart, PARTICLE* part, double* xfor, double* yfor,
                           double* zfor);
void freespace_particles_ds_noreg(int npart, PARTICLE* part);
void freespace_particles_ds_reg(int npart, PARTICLE* part);
void shielded_pot_particle_ds(int ind, int npart, PARTICLE* part, double* pot, double alpha);

// ---- one cluster's Taylor sum at offset (dx,dy,dz) = target - centre ----
// (degrees below nterms; pot -= for the screened kernel, += for 1/R)
void shielded_treetaylorpot(int nterms, double dx, double dy, double dz, double rs,
                            double* pot, double*** moment, double kappa);
void treetaylorpotential(int nterms, double dx, double dy, double dz, double rs, double* pot,
                         double*** moment);
// the field of one cluster (rs = |d|^2 + DEL; KERNEL_ORDER family; += into the outputs)
void treetaylorforce(int nterms, double dx, double dy, double dz, double rs, double* xfor,
                     double* yfor, double* zfor, double*** moment);

// ---- per-target evaluation (accumulate into the outputs) ----
double treeforce(int ind, int p, int nterms, double x, double y, double z, double acc,
                 double* xfor, double* yfor, double* zfor, TREE* tree, PARTICLE* part);
double treepotential(int ind, int p, int nterms, double x, double y, double z, double acc,
                     double* pot, TREE* tree, PARTICLE* part);
double shielded_treepot(int ind, int p, int nterms, double x, double y, double z,
                        double alpha, double acc, double* pot, TREE* tree, PARTICLE* part,
                        int Nx, int Ny, int Nz, double dlength, double dheight, double ddepth);

// ---- whole-system drivers ----
void freespace_particles_tree(int npart, PARTICLE* part, int nterms, double acc, TREE* tree,
                              int numtree, int maxmemb);
// Batch potentials (no reference counterpart: the reference evaluates them
// target by target); part[i].xforce = pot_i * charge_constant * tot_charge_i as
// in shielded_freespace_pot_ds.  pot_i is what treepotential(0, i, ...) resp.
// shielded_treepot(0, i, ...) would return.
void freespace_pot_tree(int npart, PARTICLE* part, int nterms, double acc, TREE* tree,
                        int numtree, int maxmemb);
void shielded_freespace_pot_tree(int npart, PARTICLE* part, double alpha, int nterms,
                                 double acc, TREE* tree, int numtree, int maxmemb);
void shielded_freespace_pot_ds(int npart, PARTICLE* part, double alpha);

// ---- particle set-up ----
double RandL(void);
void boxparticles(int npart, double dlength, double dheight, double ddepth, PARTICLE* part,
                  double density);
int bounding_box(int npart, PARTICLE* part, double& Mx, double& mx, double& My, double& my,
                 double& Mz, double& mz);
double sq(double x);

#endif
