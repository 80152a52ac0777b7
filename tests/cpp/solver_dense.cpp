// solver_dense.cpp -- caller of the reference GMRES interface (solver.h):
// reads n, A (row-major), b from IN; writes x from solver::gmres(a, b, x, n)
// (reference signature and stopping rule) followed by x from the matrix-free
// overload (tol 1e-12, restart 20) to OUT.
#include <cstdio>
#include <vector>

#include "solver.h"

int main(int argc, char** argv) {
  if (argc < 3) return 2;
  FILE* f = fopen(argv[1], "rb");
  double nd;
  if (fread(&nd, sizeof(double), 1, f) != 1) return 2;
  int n = (int)nd;
  std::vector<double> A((size_t)n * n), b(n);
  if (fread(A.data(), sizeof(double), A.size(), f) != A.size()) return 2;
  if (fread(b.data(), sizeof(double), b.size(), f) != b.size()) return 2;
  fclose(f);
  std::vector<double*> rows(n);
  for (int i = 0; i < n; i++) rows[i] = &A[(size_t)i * n];
  std::vector<double> x1(n, 0.0), x2(n, 0.0);
  solver s;
  s.gmres(rows.data(), b.data(), x1.data(), n);
  auto mv = [&](const double* in, double* out) {
    for (int i = 0; i < n; i++) {
      double acc = 0;
      for (int j = 0; j < n; j++) acc += A[(size_t)i * n + j] * in[j];
      out[i] = acc;
    }
  };
  int it = s.gmres(mv, b.data(), x2.data(), n, 20, 1e-12, 10 * n);
  FILE* o = fopen(argv[2], "wb");
  fwrite(x1.data(), sizeof(double), n, o);
  fwrite(x2.data(), sizeof(double), n, o);
  fclose(o);
  printf("iters %d relres %.3e\n", it, s.relative_residual());
  return it < 0 ? 1 : 0;
}
