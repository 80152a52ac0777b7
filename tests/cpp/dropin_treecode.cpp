// dropin_treecode.cpp -- caller written against the reference interface
// (tree3d.h, problem.h, constants.h), linked with the MI355X drop-in
// libtree3d.so.  It follows the reference's own drivers tests/treecode_mp.cpp
// (serial vs OpenMP treeforce over all particles) and tests/treecode.cpp
// (convergence in nterms), plus the screened potential path.
//
//   dropin_treecode field  OUT        boxparticles(16000), maxmemb 100, nterms 5
//   dropin_treecode conv   OUT        nterms 1..4, fields of each
//   dropin_treecode screened IN OUT   particles from IN (n, x, y, z, q), params from IN
//   dropin_treecode points IN OUT     p = -1 calls (treepotential, treeforce,
//                                     shielded_treepot) at the points given in IN
//   dropin_treecode directsum OUT     tests/directsum_mp.cpp's recipe: serial and
//                                     OpenMP force_particle_ds_noreg agree
//   dropin_treecode treecode OUT      tests/treecode.cpp's recipe: treeforce
//                                     nterms 1..4 against force_particle_ds_reg,
//                                     max-abs error not growing
//   dropin_treecode recharge IN OUT   treemake once, shielded_treepot for two
//                                     charge sets (treemoments after each)
//   dropin_treecode points_ds IN OUT  force_point_ds_noreg/_reg, shielded_pot_particle_ds,
//                                     freespace_particles_ds_noreg/_reg
//   dropin_treecode primitives IN OUT one-cluster Taylor sums and one-pair gradients
//                                     (host only)
//   dropin_treecode subtree IN OUT    calls that start at node ind != 0:
//                                     shielded_treepot(ind, i), treepotential(ind, -1,
//                                     x_i), treeforce(ind, i) for the (ind, i) in IN
// OUT receives raw doubles; "PASSED"/"FAILED" on stdout as the reference tests.
#include <omp.h>

#include <algorithm>
#include <cstdio>
#include <vector>

#include "constants.h"
#include "problem.h"
#include "tree3d.h"

static void write_all(const char* path, const std::vector<double>& v) {
  FILE* f = fopen(path, "wb");
  fwrite(v.data(), sizeof(double), v.size(), f);
  fclose(f);
}

static int field_mode(const char* out) {
  int npart = 16000, maxmemb = 100, nterms = 5;
  PARTICLE* part = new PARTICLE[npart];
  int maxtrees = 3 * npart / maxmemb + 10;
  TREE* tree = new TREE[maxtrees];
  boxparticles(npart, 1.0, 1.0, 1.0, part, 1.0);
  int numtree = treemake(npart, nterms, maxmemb, part, tree);
  treemoments(numtree, nterms, tree, part);
  std::vector<double> a(3 * npart, 0.0), b(3 * npart, 0.0);
  for (int i = 0; i < npart; i++)
    treeforce(0, i, nterms, part[i].x, part[i].y, part[i].z, ACC, &a[3 * i], &a[3 * i + 1],
              &a[3 * i + 2], tree, part);
#pragma omp parallel for
  for (int i = 0; i < npart; i++)
    treeforce(0, i, nterms, part[i].x, part[i].y, part[i].z, ACC, &b[3 * i], &b[3 * i + 1],
              &b[3 * i + 2], tree, part);
  bool same = a == b;
  write_all(out, a);
  // moments were copied back into the TREE array
  bool mom = tree[0].moments != nullptr && tree[numtree - 1].moments != nullptr;
  treeempty(numtree, tree);
  delete[] tree;
  delete[] part;
  printf("numtree %d serial==parallel %d moments %d\n", numtree, (int)same, (int)mom);
  printf(same && mom ? "Test PASSED\n" : "Test FAILED\n");
  return same && mom ? 0 : 1;
}

static int conv_mode(const char* out) {
  int npart = 16000, maxmemb = 100;
  PARTICLE* part = new PARTICLE[npart];
  int maxtrees = 3 * npart / maxmemb + 10;
  boxparticles(npart, 1.0, 1.0, 1.0, part, 1.0);
  std::vector<double> all;
  for (int nterms = 1; nterms < 5; nterms++) {
    TREE* tree = new TREE[maxtrees];
    int numtree = treemake(npart, nterms, maxmemb, part, tree);
    treemoments(numtree, nterms, tree, part);
    std::vector<double> f(3 * npart, 0.0);
#pragma omp parallel for
    for (int i = 0; i < npart; i++)
      treeforce(0, i, nterms, part[i].x, part[i].y, part[i].z, ACC, &f[3 * i], &f[3 * i + 1],
                &f[3 * i + 2], tree, part);
    all.insert(all.end(), f.begin(), f.end());
    treeempty(numtree, tree);
    delete[] tree;
  }
  write_all(out, all);
  delete[] part;
  printf("Test PASSED\n");
  return 0;
}

static int screened_mode(const char* in, const char* out) {
  FILE* f = fopen(in, "rb");
  double hdr[5];
  if (fread(hdr, sizeof(double), 5, f) != 5) return 2;
  int n = (int)hdr[0], maxmemb = (int)hdr[1], nterms = (int)hdr[2];
  double acc = hdr[3], alpha = hdr[4];
  std::vector<double> buf(4 * (size_t)n);
  if (fread(buf.data(), sizeof(double), buf.size(), f) != buf.size()) return 2;
  fclose(f);
  PARTICLE* part = new PARTICLE[n];
  for (int i = 0; i < n; i++) {
    part[i] = PARTICLE();
    part[i].x = buf[i];
    part[i].y = buf[n + i];
    part[i].z = buf[2 * n + i];
    part[i].charge = buf[3 * n + i];
    part[i].weight = 1.0;
    part[i].tot_charge = buf[3 * n + i];
  }
  TREE* tree = new TREE[4 * n + 64];
  int numtree = treemake(n, nterms, maxmemb, part, tree);
  treemoments(numtree, nterms, tree, part);
  std::vector<double> pot(n, 0.0);
#pragma omp parallel for
  for (int i = 0; i < n; i++)
    shielded_treepot(0, i, nterms, part[i].x, part[i].y, part[i].z, alpha, acc, &pot[i], tree,
                     part, 100, 100, 100, 1.0, 1.0, 1.0);
  treeempty(numtree, tree);
  // the batch driver gives the same values times charge_constant * tot_charge
  shielded_freespace_pot_tree(n, part, alpha, nterms, acc, tree, numtree, maxmemb);
  bool ok = true;
  for (int i = 0; i < n; i++) {
    double want = pot[i] * charge_constant * part[i].tot_charge;
    if (part[i].xforce != want) ok = false;
  }
  write_all(out, pot);
  delete[] tree;
  delete[] part;
  printf(ok ? "Test PASSED\n" : "Test FAILED\n");
  return ok ? 0 : 1;
}

static int points_mode(const char* in, const char* out) {
  FILE* f = fopen(in, "rb");
  double hdr[6];
  if (fread(hdr, sizeof(double), 6, f) != 6) return 2;
  int n = (int)hdr[0], maxmemb = (int)hdr[1], nterms = (int)hdr[2], m = (int)hdr[5];
  double acc = hdr[3], alpha = hdr[4];
  std::vector<double> buf(4 * (size_t)n + 3 * (size_t)m);
  if (fread(buf.data(), sizeof(double), buf.size(), f) != buf.size()) return 2;
  fclose(f);
  PARTICLE* part = new PARTICLE[n];
  for (int i = 0; i < n; i++) {
    part[i] = PARTICLE();
    part[i].x = buf[i];
    part[i].y = buf[n + i];
    part[i].z = buf[2 * n + i];
    part[i].charge = buf[3 * n + i];
    part[i].weight = 1.0;
    part[i].tot_charge = buf[3 * n + i];
  }
  const double* px = &buf[4 * (size_t)n];
  const double* py = px + m;
  const double* pz = py + m;
  TREE* tree = new TREE[4 * n + 64];
  int numtree = treemake(n, nterms, maxmemb, part, tree);
  treemoments(numtree, nterms, tree, part);
  std::vector<double> res(5 * (size_t)m, 0.0);  // pot, Ex, Ey, Ez, shielded
#pragma omp parallel for
  for (int i = 0; i < m; i++) {
    treepotential(0, -1, nterms, px[i], py[i], pz[i], acc, &res[i], tree, part);
    treeforce(0, -1, nterms, px[i], py[i], pz[i], acc, &res[m + 3 * (size_t)i],
              &res[m + 3 * (size_t)i + 1], &res[m + 3 * (size_t)i + 2], tree, part);
    shielded_treepot(0, -1, nterms, px[i], py[i], pz[i], alpha, acc, &res[4 * (size_t)m + i],
                     tree, part, 100, 100, 100, 1.0, 1.0, 1.0);
  }
  treeempty(numtree, tree);
  write_all(out, res);
  delete[] tree;
  delete[] part;
  printf("Test PASSED\n");
  return 0;
}

// Subtree calls.  IN: n, maxmemb, nterms, acc, alpha, m, then x, y, z, q (n each)
// and the pairs' nodes and particles (m each, as doubles).  OUT: shielded raw
// [m], treepotential at the particle [m], field [m][3].  A call with a node
// outside the tree adds nothing (it prints a message).
static int subtree_mode(const char* in, const char* out) {
  FILE* f = fopen(in, "rb");
  double hdr[6];
  if (fread(hdr, sizeof(double), 6, f) != 6) return 2;
  int n = (int)hdr[0], maxmemb = (int)hdr[1], nterms = (int)hdr[2], m = (int)hdr[5];
  double acc = hdr[3], alpha = hdr[4];
  std::vector<double> buf(4 * (size_t)n + 2 * (size_t)m);
  if (fread(buf.data(), sizeof(double), buf.size(), f) != buf.size()) return 2;
  fclose(f);
  PARTICLE* part = new PARTICLE[n];
  for (int i = 0; i < n; i++) {
    part[i] = PARTICLE();
    part[i].x = buf[i];
    part[i].y = buf[n + i];
    part[i].z = buf[2 * n + i];
    part[i].charge = buf[3 * n + i];
    part[i].weight = 1.0;
    part[i].tot_charge = buf[3 * n + i];
  }
  const double* ind = &buf[4 * (size_t)n];
  const double* tid = ind + m;
  TREE* tree = new TREE[4 * n + 64];
  int numtree = treemake(n, nterms, maxmemb, part, tree);
  treemoments(numtree, nterms, tree, part);
  std::vector<double> res(5 * (size_t)m, 0.0);  // shielded, pot, Ex, Ey, Ez
#pragma omp parallel for schedule(dynamic, 8)
  for (int t = 0; t < m; t++) {
    const int nd = (int)ind[t], i = (int)tid[t];
    shielded_treepot(nd, i, nterms, part[i].x, part[i].y, part[i].z, alpha, acc, &res[t], tree,
                     part, 100, 100, 100, 1.0, 1.0, 1.0);
    treepotential(nd, -1, nterms, part[i].x, part[i].y, part[i].z, acc, &res[m + t], tree,
                  part);
    treeforce(nd, i, nterms, part[i].x, part[i].y, part[i].z, acc, &res[2 * (size_t)m + 3 * t],
              &res[2 * (size_t)m + 3 * t + 1], &res[2 * (size_t)m + 3 * t + 2], tree, part);
  }
  double bad = 0.0;
  treepotential(numtree, 0, nterms, part[0].x, part[0].y, part[0].z, acc, &bad, tree, part);
  treeempty(numtree, tree);
  write_all(out, res);
  delete[] tree;
  delete[] part;
  printf(bad == 0.0 ? "Test PASSED\n" : "Test FAILED\n");
  return bad == 0.0 ? 0 : 1;
}

// tests/directsum_mp.cpp: the same direct sums serially and from OpenMP threads.
static int directsum_mode(const char* out) {
  int npart = 16000;
  PARTICLE* part = new PARTICLE[npart];
  boxparticles(npart, 1.0, 1.0, 1.0, part, 1.0);
  std::vector<double> a(3 * npart), b(3 * npart);
  double xfor, yfor, zfor;
  for (int i = 0; i < npart; i++) {
    xfor = yfor = zfor = 0.0;
    force_particle_ds_noreg(i, npart, part, &xfor, &yfor, &zfor);
    a[3 * i] = xfor * charge_constant * part[i].tot_charge;
    a[3 * i + 1] = yfor * charge_constant * part[i].tot_charge;
    a[3 * i + 2] = zfor * charge_constant * part[i].tot_charge;
  }
#pragma omp parallel for private(xfor, yfor, zfor)
  for (int i = 0; i < npart; i++) {
    xfor = yfor = zfor = 0.0;
    force_particle_ds_noreg(i, npart, part, &xfor, &yfor, &zfor);
    b[3 * i] = xfor * charge_constant * part[i].tot_charge;
    b[3 * i + 1] = yfor * charge_constant * part[i].tot_charge;
    b[3 * i + 2] = zfor * charge_constant * part[i].tot_charge;
  }
  double err[3] = {0, 0, 0};
  for (int i = 0; i < npart; i++)
    for (int c = 0; c < 3; c++) err[c] = std::max(err[c], std::abs(b[3 * i + c] - a[3 * i + c]));
  const bool ok = std::abs(err[0]) + std::abs(err[1]) + std::abs(err[2]) < 1.0e-15;
  std::vector<double> raw(3 * npart);  // unscaled fields for the comparison with the oracle
  for (int i = 0; i < npart; i++)
    force_particle_ds_noreg(i, npart, part, &raw[3 * i], &raw[3 * i + 1], &raw[3 * i + 2]);
  write_all(out, raw);
  delete[] part;
  printf(ok ? "Test PASSED\n" : "Test FAILED\n");
  return ok ? 0 : 1;
}

// tests/treecode.cpp: treeforce for nterms 1..4 against the regularized direct
// sum; the max-abs error must not grow with nterms.
static int treecode_mode(const char* out) {
  int npart = 16000, maxmemb = 100;
  PARTICLE* part = new PARTICLE[npart];
  int maxtrees = 3 * npart / maxmemb + 10;
  boxparticles(npart, 1.0, 1.0, 1.0, part, 1.0);
  std::vector<double> fx(npart), fy(npart), fz(npart), fx2(npart), fy2(npart), fz2(npart);
  double xfor, yfor, zfor;
#pragma omp parallel for private(xfor, yfor, zfor)
  for (int i = 0; i < npart; i++) {
    xfor = yfor = zfor = 0.0;
    force_particle_ds_reg(i, npart, part, &xfor, &yfor, &zfor);
    fx[i] = xfor * charge_constant * part[i].tot_charge;
    fy[i] = yfor * charge_constant * part[i].tot_charge;
    fz[i] = zfor * charge_constant * part[i].tot_charge;
  }
  std::vector<double> rec;
  for (int i = 0; i < npart; i++) {
    rec.push_back(fx[i]);
    rec.push_back(fy[i]);
    rec.push_back(fz[i]);
  }
  int failflag = 0;
  double errstore = 0.0;
  for (int nterms = 1; nterms < 5; nterms++) {
    TREE* tree = new TREE[maxtrees];
    int numtree = treemake(npart, nterms, maxmemb, part, tree);
    treemoments(numtree, nterms, tree, part);
#pragma omp parallel for private(xfor, yfor, zfor)
    for (int i = 0; i < npart; i++) {
      xfor = yfor = zfor = 0.0;
      treeforce(0, i, nterms, part[i].x, part[i].y, part[i].z, ACC, &xfor, &yfor, &zfor, tree,
                part);
      fx2[i] = xfor * charge_constant * part[i].tot_charge;
      fy2[i] = yfor * charge_constant * part[i].tot_charge;
      fz2[i] = zfor * charge_constant * part[i].tot_charge;
    }
    double errx = 0.0, erry = 0.0, errz = 0.0;
    for (int i = 0; i < npart; i++) {
      errx = std::max(errx, std::abs(fx2[i] - fx[i]));
      erry = std::max(erry, std::abs(fy2[i] - fy[i]));
      errz = std::max(errz, std::abs(fz2[i] - fz[i]));
    }
    printf("nterms = %d, errx = %g, erry = %g, errz = %g\n", nterms, errx, erry, errz);
    rec.push_back(errx);
    rec.push_back(erry);
    rec.push_back(errz);
    const double errnew = errx + erry + errz;
    if (nterms == 1)
      errstore = errnew;
    else if (errnew > errstore)
      failflag = 1;
    else
      errstore = errnew;
    treeempty(numtree, tree);
    delete[] tree;
  }
  write_all(out, rec);
  delete[] part;
  printf(failflag ? "Test FAILED\n" : "Test PASSED\n");
  return failflag;
}

static PARTICLE* read_particles(FILE* f, int n) {
  std::vector<double> buf(4 * (size_t)n);
  if (fread(buf.data(), sizeof(double), buf.size(), f) != buf.size()) return nullptr;
  PARTICLE* part = new PARTICLE[n];
  for (int i = 0; i < n; i++) {
    part[i] = PARTICLE();
    part[i].x = buf[i];
    part[i].y = buf[n + i];
    part[i].z = buf[2 * n + i];
    part[i].charge = buf[3 * n + i];
    part[i].weight = 1.0;
    part[i].tot_charge = buf[3 * n + i];
  }
  return part;
}

// One tree, two charge sets: the second treemoments re-reads part[].tot_charge.
// IN: n, maxmemb, nterms, acc, alpha, then x, y, z, q1, q2.  OUT: the raw
// shielded_treepot values for q1, then for q2.
static int recharge_mode(const char* in, const char* out) {
  FILE* f = fopen(in, "rb");
  double hdr[5];
  if (fread(hdr, sizeof(double), 5, f) != 5) return 2;
  int n = (int)hdr[0], maxmemb = (int)hdr[1], nterms = (int)hdr[2];
  double acc = hdr[3], alpha = hdr[4];
  PARTICLE* part = read_particles(f, n);
  std::vector<double> q2(n);
  if (!part || fread(q2.data(), sizeof(double), n, f) != (size_t)n) return 2;
  fclose(f);
  TREE* tree = new TREE[4 * n + 64];
  int numtree = treemake(n, nterms, maxmemb, part, tree);
  std::vector<double> pot(2 * (size_t)n, 0.0);
  for (int set = 0; set < 2; set++) {
    if (set == 1)
      for (int i = 0; i < n; i++) part[i].charge = part[i].tot_charge = q2[i];
    treemoments(numtree, nterms, tree, part);
#pragma omp parallel for
    for (int i = 0; i < n; i++)
      shielded_treepot(0, i, nterms, part[i].x, part[i].y, part[i].z, alpha, acc,
                       &pot[(size_t)set * n + i], tree, part, 100, 100, 100, 1.0, 1.0, 1.0);
  }
  // the moments copied into the TREE array follow the second charge set
  double m0 = 0.0;
  for (int i = 0; i < n; i++) m0 += q2[i];
  const bool mom = std::abs(tree[0].moments[0][0][0] - m0) <= 1e-9 * (1.0 + std::abs(m0));
  treeempty(numtree, tree);
  write_all(out, pot);
  delete[] tree;
  delete[] part;
  printf(mom ? "Test PASSED\n" : "Test FAILED\n");
  return mom ? 0 : 1;
}

// Direct sums of the reference interface.  IN: n, m, alpha, then x, y, z, q,
// then the m points px, py, pz.  OUT: force_point_ds_noreg [3m],
// force_point_ds_reg [3m], shielded_pot_particle_ds [n] (raw), and
// freespace_particles_ds_noreg / _reg part[].xforce.. [3n each].
static int points_ds_mode(const char* in, const char* out) {
  FILE* f = fopen(in, "rb");
  double hdr[3];
  if (fread(hdr, sizeof(double), 3, f) != 3) return 2;
  int n = (int)hdr[0], m = (int)hdr[1];
  double alpha = hdr[2];
  PARTICLE* part = read_particles(f, n);
  std::vector<double> pts(3 * (size_t)m);
  if (!part || fread(pts.data(), sizeof(double), pts.size(), f) != pts.size()) return 2;
  fclose(f);
  std::vector<double> res;
  std::vector<double> e(6 * (size_t)m);
#pragma omp parallel for
  for (int i = 0; i < m; i++) {
    force_point_ds_noreg(pts[i], pts[m + i], pts[2 * m + i], n, part, &e[3 * i], &e[3 * i + 1],
                         &e[3 * i + 2]);
    force_point_ds_reg(pts[i], pts[m + i], pts[2 * m + i], n, part, &e[3 * (m + i)],
                       &e[3 * (m + i) + 1], &e[3 * (m + i) + 2]);
  }
  res.insert(res.end(), e.begin(), e.end());
  std::vector<double> pot(n);
#pragma omp parallel for
  for (int i = 0; i < n; i++) shielded_pot_particle_ds(i, n, part, &pot[i], alpha);
  res.insert(res.end(), pot.begin(), pot.end());
  for (int reg = 0; reg < 2; reg++) {
    if (reg) freespace_particles_ds_reg(n, part);
    else freespace_particles_ds_noreg(n, part);
    for (int i = 0; i < n; i++) {
      res.push_back(part[i].xforce);
      res.push_back(part[i].yforce);
      res.push_back(part[i].zforce);
    }
  }
  write_all(out, res);
  delete[] part;
  printf("Test PASSED\n");
  return 0;
}

// The one-cluster and one-pair primitives (host functions of the drop-in; no
// device).  IN: nterms, kappa, ncase, then per case dx, dy, dz, rs and the
// moment cube [nterms]^3.  OUT per case: shielded_treetaylorpot's pot (from 0),
// treetaylorpotential's pot (from 0), eval_grad_kernel_noreg (3) and
// eval_grad_kernel_reg (3) at (dx, dy, dz, rs), treetaylorforce's field (3,
// from 0).
static int primitives_mode(const char* in, const char* out) {
  FILE* f = fopen(in, "rb");
  double hdr[3];
  if (fread(hdr, sizeof(double), 3, f) != 3) return 2;
  const int nt = (int)hdr[0], ncase = (int)hdr[2];
  const double kappa = hdr[1];
  std::vector<double> res;
  std::vector<double> cube((size_t)nt * nt * nt);
  std::vector<double*> rows((size_t)nt * nt);
  std::vector<double**> planes(nt);
  for (int i = 0; i < nt; i++) {
    planes[i] = &rows[(size_t)i * nt];
    for (int j = 0; j < nt; j++) rows[(size_t)i * nt + j] = &cube[((size_t)i * nt + j) * nt];
  }
  for (int c = 0; c < ncase; c++) {
    double d[4];
    if (fread(d, sizeof(double), 4, f) != 4) return 2;
    if (fread(cube.data(), sizeof(double), cube.size(), f) != cube.size()) return 2;
    double p1 = 0.0, p2 = 0.0, g[6];
    shielded_treetaylorpot(nt, d[0], d[1], d[2], d[3], &p1, planes.data(), kappa);
    treetaylorpotential(nt, d[0], d[1], d[2], d[3], &p2, planes.data());
    eval_grad_kernel_noreg(d[0], d[1], d[2], d[3], &g[0], &g[1], &g[2]);
    eval_grad_kernel_reg(d[0], d[1], d[2], d[3], &g[3], &g[4], &g[5]);
    double e[3] = {0.0, 0.0, 0.0};
    treetaylorforce(nt, d[0], d[1], d[2], d[3], &e[0], &e[1], &e[2], planes.data());
    res.push_back(p1);
    res.push_back(p2);
    res.insert(res.end(), g, g + 6);
    res.insert(res.end(), e, e + 3);
  }
  fclose(f);
  write_all(out, res);
  printf("Test PASSED\n");
  return 0;
}

int main(int argc, char** argv) {
  if (argc >= 3 && std::string(argv[1]) == "field") return field_mode(argv[2]);
  if (argc >= 3 && std::string(argv[1]) == "conv") return conv_mode(argv[2]);
  if (argc >= 4 && std::string(argv[1]) == "screened") return screened_mode(argv[2], argv[3]);
  if (argc >= 4 && std::string(argv[1]) == "points") return points_mode(argv[2], argv[3]);
  if (argc >= 3 && std::string(argv[1]) == "directsum") return directsum_mode(argv[2]);
  if (argc >= 3 && std::string(argv[1]) == "treecode") return treecode_mode(argv[2]);
  if (argc >= 4 && std::string(argv[1]) == "recharge") return recharge_mode(argv[2], argv[3]);
  if (argc >= 4 && std::string(argv[1]) == "points_ds") return points_ds_mode(argv[2], argv[3]);
  if (argc >= 4 && std::string(argv[1]) == "primitives") return primitives_mode(argv[2], argv[3]);
  if (argc >= 4 && std::string(argv[1]) == "subtree") return subtree_mode(argv[2], argv[3]);
  fprintf(stderr,
          "usage: dropin_treecode field|conv|directsum|treecode OUT | "
          "screened|points|recharge|points_ds IN OUT\n");
  return 2;
}
