// ratematrix.hpp -- codon rate matrices and their eigensystems (host, FP64).
#pragma once
#include <string>
#include <vector>

namespace pcsf {

// Eigensystem of a reversible rate matrix Q with stationary pi:
//   Q = V diag(lam) Vinv,  V = Pi^-1/2 U,  Vinv = U^T Pi^1/2,  U orthogonal.
struct Eigen {
  int k = 0;
  std::vector<double> lam;   // k
  std::vector<double> V;     // k x k row-major (right eigenvectors in columns)
  std::vector<double> Vinv;  // k x k row-major (left eigenvectors in rows)
};

// Q from an ECM (src/PhyloCSFModel.ml:11-30 + lib/CamlPaml/PhyloModel.ml:76-106):
// q_ij = s_ij pi_j (i != j), q_ii = -sum_j q_ij, Q = q / sum_i pi_i (-q_ii).
// Returns false (with msg) when the scale is not positive.
bool build_rate_matrix(int k, const double* S, const double* pi, double* Q, std::string* msg);

// Symmetric similarity transform + cyclic Jacobi rotations.  Returns false
// if pi has a non-positive entry (the reversible path needs pi > 0).
bool eigen_reversible(int k, const double* Q, const double* pi, Eigen* out, std::string* msg);

// Q.Diag.real_to_Pt (lib/CamlPaml/Q.ml:211-249) on the host: used by the
// simulator.  Returns false with the reference's message on failure.
bool transition_host(const Eigen& e, double t, double tol, double* P, std::string* msg);

}  // namespace pcsf
