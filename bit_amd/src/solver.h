// solver.h -- the reference's GMRES class interface (solver.h of ongbw/bit),
// header-only, on the FP64 cores of bit_amd/csrc/gmres.h.
//
//   solver s;
//   s.gmres(a, b, x, num);                         // reference signature: dense a[num][num],
//                                                  // reference stopping rule, no restart
//   s.gmres(matvec, b, x, num, restart, tol, maxit) // matrix-free GMRES(restart), residual
//                                                  // tolerance; matvec(const double*, double*)
//
// makeFram / clear_all / minimize keep the reference's names.  Work arrays are
// owned by the cores (std::vector), so makeFram only records the size.
#ifndef solveh
#define solveh 1

#include <cmath>
#include <iostream>

#include "../csrc/gmres.h"

using namespace std;

typedef double* pdoub;
typedef pdoub* ppdoub;

class solver {
  int MAXPAN;
  double bnorm;
  bit::GmresResult last_;

 public:
  solver() : MAXPAN(0), bnorm(1.0) {}
  ~solver() { clear_all(); }

  void makeFram(int pan) {
    if (pan > MAXPAN) MAXPAN = pan;
  }
  void clear_all(void) { MAXPAN = 0; }

  // Dense system a x = b (a[i][j], row i); the reference algorithm
  // (solver.h:157-260): MGS Arnoldi without restart, stop when the largest
  // relative change of x between checks is below 1e-5.  x holds the previous
  // iterate used by the first check (the reference's callers pass zeros).
  int gmres(double** a, double* b, double* x, int num) {
    if (num > MAXPAN - 1) makeFram(num + 1);
    auto matvec = [&](const double* in, double* out) {
      for (int i = 0; i < num; i++) {
        double s = 0.0;
        for (int j = 0; j < num; j++) s += a[i][j] * in[j];
        out[i] = s;
      }
    };
    last_ = bit::reference_gmres(matvec, b, x, num);
    bnorm = 0.0;
    for (int i = 0; i < num; i++) bnorm += b[i] * b[i];
    bnorm = std::sqrt(bnorm);
    return 0;
  }

  // Matrix-free restarted GMRES: FP64 Arnoldi (modified Gram-Schmidt), Givens
  // rotations, stop at ||b - A x|| <= tol ||b||.  Returns the Arnoldi steps
  // taken, or -1 if maxit was reached first.
  template <class MatVec>
  int gmres(MatVec&& matvec, const double* b, double* x, int num, int restart, double tol,
            int maxit) {
    last_ = bit::restarted_gmres(matvec, b, x, num, restart, tol, maxit);
    return last_.converged ? last_.iters : -1;
  }

  // Kept for the interface: the least-squares step is inside the cores.
  int minimize(int n) {
    (void)n;
    return 0;
  }

  double relative_residual() const { return last_.relres; }
  int iterations() const { return last_.iters; }
};

#endif
