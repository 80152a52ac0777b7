// gmres.h -- FP64 GMRES cores behind solver.h and bit_solve_second_kind.
//
// restarted_gmres: matrix-free GMRES(m), modified Gram-Schmidt Arnoldi, Givens
// rotations on the Hessenberg column, residual-norm stopping ||r|| <= tol ||b||.
// reference_gmres: the algorithm of the reference's solver::gmres
// (solver.h:157-260 of ongbw/bit): no restart, MGS Arnoldi, stop when the
// largest relative change of x between two checks (every other iteration)
// falls below 1e-5, breakdown below 1e-14.  Its least-squares solve uses the
// same Givens recurrence (the reference re-factorises the Hessenberg by
// Gram-Schmidt every step; both give the minimiser of ||beta e1 - H y||).
// Everything is FP64.
#pragma once

#include <cmath>
#include <cstdio>
#include <vector>

namespace bit {

struct GmresResult {
  int iters = 0;          // matrix-vector products inside Arnoldi steps
  int restarts = 0;
  double relres = 0.0;    // ||b - A x|| / ||b|| estimated by the Givens recurrence
  bool converged = false;
  bool breakdown = false;
};

namespace detail {
inline double dot(const double* a, const double* b, int n) {
  double s = 0.0;
  for (int i = 0; i < n; i++) s += a[i] * b[i];
  return s;
}

// Upper-triangular solve of the first k Givens-reduced rows: H y = g.
inline void back_solve(const std::vector<double>& H, int ldh, const std::vector<double>& g,
                       int k, std::vector<double>& y) {
  y.assign(k, 0.0);
  for (int i = k - 1; i >= 0; i--) {
    double s = g[i];
    for (int j = i + 1; j < k; j++) s -= H[(size_t)i * ldh + j] * y[j];
    y[i] = s / H[(size_t)i * ldh + i];
  }
}

// Applies the stored rotations to column j of H and creates rotation j.
inline void givens_column(std::vector<double>& H, int ldh, int j, std::vector<double>& cs,
                          std::vector<double>& sn, std::vector<double>& g) {
  for (int i = 0; i < j; i++) {
    double a = H[(size_t)i * ldh + j], b = H[(size_t)(i + 1) * ldh + j];
    H[(size_t)i * ldh + j] = cs[i] * a + sn[i] * b;
    H[(size_t)(i + 1) * ldh + j] = -sn[i] * a + cs[i] * b;
  }
  double a = H[(size_t)j * ldh + j], b = H[(size_t)(j + 1) * ldh + j];
  double r = std::hypot(a, b);
  cs[j] = r > 0 ? a / r : 1.0;
  sn[j] = r > 0 ? b / r : 0.0;
  H[(size_t)j * ldh + j] = r;
  H[(size_t)(j + 1) * ldh + j] = 0.0;
  g[j + 1] = -sn[j] * g[j];
  g[j] = cs[j] * g[j];
}
}  // namespace detail

// x holds the initial guess on entry and the solution on return.
template <class MatVec>
GmresResult restarted_gmres(MatVec&& A, const double* b, double* x, int n, int restart,
                            double tol, int maxit) {
  GmresResult res;
  if (restart < 1) restart = 1;
  const double bnorm = std::sqrt(detail::dot(b, b, n));
  if (bnorm == 0.0) {
    for (int i = 0; i < n; i++) x[i] = 0.0;
    res.converged = true;
    return res;
  }
  const int ldh = restart;
  std::vector<double> V((size_t)(restart + 1) * n), H((size_t)(restart + 1) * ldh), cs(restart),
      sn(restart), g(restart + 1), w(n), y;
  while (true) {
    A(x, w.data());
    for (int i = 0; i < n; i++) w[i] = b[i] - w[i];
    double beta = std::sqrt(detail::dot(w.data(), w.data(), n));
    res.relres = beta / bnorm;
    if (res.relres <= tol) {
      res.converged = true;
      return res;
    }
    if (res.iters >= maxit) return res;
    for (int i = 0; i < n; i++) V[i] = w[i] / beta;
    std::fill(H.begin(), H.end(), 0.0);
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int k = 0;
    for (int j = 0; j < restart && res.iters < maxit; j++) {
      const double* vj = &V[(size_t)j * n];
      A(vj, w.data());
      res.iters++;
      for (int i = 0; i <= j; i++) {  // modified Gram-Schmidt
        const double* vi = &V[(size_t)i * n];
        double h = detail::dot(w.data(), vi, n);
        H[(size_t)i * ldh + j] = h;
        for (int t = 0; t < n; t++) w[t] -= h * vi[t];
      }
      double hn = std::sqrt(detail::dot(w.data(), w.data(), n));
      H[(size_t)(j + 1) * ldh + j] = hn;
      if (hn > 0)
        for (int t = 0; t < n; t++) V[(size_t)(j + 1) * n + t] = w[t] / hn;
      detail::givens_column(H, ldh, j, cs, sn, g);
      k = j + 1;
      res.relres = std::fabs(g[j + 1]) / bnorm;
      if (hn == 0.0) res.breakdown = true;  // the Krylov space is invariant: exact
      if (res.relres <= tol || hn == 0.0) break;
    }
    detail::back_solve(H, ldh, g, k, y);
    for (int j = 0; j < k; j++)
      for (int t = 0; t < n; t++) x[t] += y[j] * V[(size_t)j * n + t];
    res.restarts++;
  }
}

// The reference's stopping rule and breakdown test on a caller's operator.
template <class MatVec>
GmresResult reference_gmres(MatVec&& A, const double* b, double* x, int n) {
  GmresResult res;
  const double errlim = 1e-5;
  const int m = n;  // no restart
  const int ldh = m;
  std::vector<double> Q((size_t)(m + 1) * n), H((size_t)(m + 1) * ldh, 0.0), cs(m), sn(m),
      g(m + 1, 0.0), v(n), y;
  const double bnorm = std::sqrt(detail::dot(b, b, n));
  if (bnorm == 0.0) {
    for (int i = 0; i < n; i++) x[i] = 0.0;
    res.converged = true;
    return res;
  }
  for (int i = 0; i < n; i++) Q[i] = b[i] / bnorm;
  g[0] = bnorm;
  bool go = true;
  for (int k = 0; go; k++) {
    const int k1 = k + 1;
    A(&Q[(size_t)k * n], v.data());
    res.iters++;
    for (int j = 0; j < k1; j++) {
      double h = detail::dot(&Q[(size_t)j * n], v.data(), n);
      H[(size_t)j * ldh + k] = h;
      for (int i = 0; i < n; i++) v[i] -= h * Q[(size_t)j * n + i];
    }
    double hn = std::sqrt(detail::dot(v.data(), v.data(), n));
    H[(size_t)k1 * ldh + k] = hn;
    if (k1 < n && hn < 1e-14) {
      std::printf("ERROR: Breakdown!  Quitting...%d\n", k);
      res.breakdown = true;
      go = false;
    }
    for (int i = 0; i < n && hn > 0; i++) Q[(size_t)k1 * n + i] = v[i] / hn;
    detail::givens_column(H, ldh, k, cs, sn, g);
    res.relres = std::fabs(g[k1]) / bnorm;
    if (k1 == n) go = false;
    if (k1 % 2 || !go) {
      detail::back_solve(H, ldh, g, k1, y);
      double maxerr = 0.0;
      for (int i = 0; i < n; i++) {
        double prev = x[i], xi = 0.0;
        for (int j = 0; j < k1; j++) xi += Q[(size_t)j * n + i] * y[j];
        x[i] = xi;
        double e = std::fabs((prev - xi) / prev);
        if (e > maxerr) maxerr = e;
      }
      if (maxerr < errlim && k > 1) go = false;
    }
  }
  res.converged = !res.breakdown;
  return res;
}

}  // namespace bit
