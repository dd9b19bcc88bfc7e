// ratematrix.cpp -- codon rate matrices and their eigensystems (host, FP64).
//
// The reference diagonalises Q with GSL's nonsymmetric eigensolver and a
// complex LU inverse (lib/CamlPaml/Q.ml:124-141).  Every PhyloCSF ECM is
// reversible (q_ij = s_ij pi_j with symmetric s), so A = Pi^1/2 Q Pi^-1/2 is
// symmetric and is diagonalised here with cyclic Jacobi rotations
// (Jacobi 1846; Golub & Van Loan, Matrix Computations, sec. 8.5), which gives
// orthogonal eigenvectors to working precision; V and V^-1 follow without
// an inverse.
#include "ratematrix.hpp"

#include <cmath>
#include <cstdio>

namespace pcsf {

bool build_rate_matrix(int k, const double* S, const double* pi, double* Q, std::string* msg) {
  for (int i = 0; i < k; ++i) {
    double off = 0.0;
    for (int j = 0; j < k; ++j) {
      Q[i * k + j] = (i == j) ? 0.0 : S[i * k + j] * pi[j];
      if (j != i) off = Q[i * k + j] + off;
    }
    Q[i * k + i] = 0.0 - off;
  }
  double scale = 0.0;
  for (int i = 0; i < k; ++i)
    if (Q[i * k + i] != 0.0) scale = scale - pi[i] * Q[i * k + i];
  if (!(scale > 0.0)) {
    if (msg) *msg = "CamlPaml.P14n.instantiate_q: Q scale evaluated to a non-positive value";
    return false;
  }
  for (int i = 0; i < k * k; ++i) Q[i] = Q[i] / scale;
  return true;
}

bool eigen_reversible(int k, const double* Q, const double* pi, Eigen* out, std::string* msg) {
  for (int i = 0; i < k; ++i)
    if (!(pi[i] > 0.0)) {
      if (msg) *msg = "rate matrix: stationary distribution must be strictly positive";
      return false;
    }
  std::vector<double> A(k * k), U(k * k, 0.0), sq(k);
  for (int i = 0; i < k; ++i) sq[i] = std::sqrt(pi[i]);
  for (int i = 0; i < k; ++i)
    for (int j = 0; j < k; ++j)
      A[i * k + j] = 0.5 * (sq[i] * Q[i * k + j] / sq[j] + sq[j] * Q[j * k + i] / sq[i]);
  for (int i = 0; i < k; ++i) U[i * k + i] = 1.0;
  auto off_norm = [&]() {
    double s = 0.0;
    for (int p = 0; p < k; ++p)
      for (int q = 0; q < k; ++q)
        if (p != q) s += A[p * k + q] * A[p * k + q];
    return s;
  };
  double total = 0.0;
  for (double a : A) total += a * a;
  for (int sweep = 0; sweep < 64 && off_norm() > 1e-32 * total; ++sweep) {
    for (int p = 0; p + 1 < k; ++p) {
      for (int q = p + 1; q < k; ++q) {
        const double apq = A[p * k + q];
        if (std::fabs(apq) < 1e-300) continue;
        // rotation angle that zeroes A[p][q]
        const double tau = (A[q * k + q] - A[p * k + p]) / (2.0 * apq);
        const double t = std::copysign(1.0, tau) / (std::fabs(tau) + std::hypot(1.0, tau));
        const double c = 1.0 / std::hypot(1.0, t);
        const double s = t * c;
        for (int r = 0; r < k; ++r) {
          const double x = A[r * k + p], y = A[r * k + q];
          A[r * k + p] = c * x - s * y;
          A[r * k + q] = s * x + c * y;
        }
        for (int r = 0; r < k; ++r) {
          const double x = A[p * k + r], y = A[q * k + r];
          A[p * k + r] = c * x - s * y;
          A[q * k + r] = s * x + c * y;
        }
        for (int r = 0; r < k; ++r) {
          const double x = U[r * k + p], y = U[r * k + q];
          U[r * k + p] = c * x - s * y;
          U[r * k + q] = s * x + c * y;
        }
      }
    }
  }
  out->k = k;
  out->lam.resize(k);
  out->V.resize(k * k);
  out->Vinv.resize(k * k);
  for (int m = 0; m < k; ++m) out->lam[m] = A[m * k + m];
  for (int i = 0; i < k; ++i)
    for (int m = 0; m < k; ++m) {
      out->V[i * k + m] = U[i * k + m] / sq[i];
      out->Vinv[m * k + i] = U[i * k + m] * sq[i];
    }
  return true;
}

bool transition_host(const Eigen& e, double t, double tol, double* P, std::string* msg) {
  const int k = e.k;
  if (t < 0.0) {
    if (msg) *msg = "CamlPaml.Q.to_Pt";
    return false;
  }
  std::vector<double> ex(k);
  for (int m = 0; m < k; ++m) ex[m] = std::exp(t * e.lam[m]);
  for (int i = 0; i < k; ++i)
    for (int j = 0; j < k; ++j) {
      double s = 0.0;
      for (int m = 0; m < k; ++m) s += e.V[i * k + m] * (ex[m] * e.Vinv[m * k + j]);
      P[i * k + j] = s;
    }
  char buf[160];
  for (int i = 0; i < k; ++i) {
    double tot = 0.0, diag = 1.0;
    for (int j = 0; j < k; ++j) {
      double& x = P[i * k + j];
      tot += x;
      if (x < 0.0) {
        if (std::fabs(x) > tol) {
          std::snprintf(buf, sizeof buf,
                        "CamlPaml.Q.substition_matrix: expm(%.2e*Q)[%d,%d] = %e < 0", t, i, j, x);
          if (msg) *msg = buf;
          return false;
        }
        x = 0.0;
      }
      if (j != i) diag -= x;
    }
    if (std::fabs(tot - 1.0) > tol) {
      std::snprintf(buf, sizeof buf,
                    "CamlPaml.Q.substitution matrix: sum(expm(%.2e*Q)[%d,] = %e > 1", t, i, tot);
      if (msg) *msg = buf;
      return false;
    }
    if (!(diag <= 1.0 && diag > 0.0)) {
      if (msg) *msg = "Assert_failure (Q.ml:245)";
      return false;
    }
    P[i * k + i] = diag;
  }
  return true;
}

}  // namespace pcsf
