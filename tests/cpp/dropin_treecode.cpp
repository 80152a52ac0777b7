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
// OUT receives raw doubles; "PASSED"/"FAILED" on stdout as the reference tests.
#include <omp.h>

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

int main(int argc, char** argv) {
  if (argc >= 3 && std::string(argv[1]) == "field") return field_mode(argv[2]);
  if (argc >= 3 && std::string(argv[1]) == "conv") return conv_mode(argv[2]);
  if (argc >= 4 && std::string(argv[1]) == "screened") return screened_mode(argv[2], argv[3]);
  if (argc >= 4 && std::string(argv[1]) == "points") return points_mode(argv[2], argv[3]);
  fprintf(stderr, "usage: dropin_treecode field|conv OUT | screened|points IN OUT\n");
  return 2;
}
