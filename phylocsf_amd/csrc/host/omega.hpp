// omega.hpp -- the codon model of the omega (dN/dS) test, host side.
//
//   pi_exprs, q_p14n, q_scale ...... src/OmegaModel.ml:21-80
//   update_f3x4 .................... src/OmegaModel.ml:102-134
//   priors of kr_map ............... src/OmegaModel.ml:146-157
#pragma once
#include <cstdint>

namespace pcsf {

// Settings of the model, as the reference numbers them: 0 kappa
// (transition/transversion), 1 omega (dN/dS), 2 sigma (stop-codon factor),
// then per codon position p = 0, 1, 2 the frequencies of A, C, G relative
// to T at 3 + 3p .. 5 + 3p.
constexpr int kOmegaSettings = 12;

// Codon frequencies of the settings (not normalised: every codon is divided
// by 1 - (1 - sigma) (f(TAA) + f(TAG) + f(TGA)), stops included).
void omega_pi(const double* qs, double* pi);
// Exchangeabilities: q_ij = pi_j s_ij, s_ij = kappa (transition) x omega
// (nonsynonymous between sense codons) for codons one nucleotide apart, 0
// otherwise.  Q = rows filled to zero sum, divided by sum_i -pi_i q_ii.
void omega_exchangeabilities(const double* qs, double* S);
// F3x4 settings 3..11 from leaf codes (n_cols x nl, 255 = `Marginalize):
// counts start at 1; A, C, G over T per codon position.
void omega_f3x4(const uint8_t* codes, int64_t ncols, int nl, double* qs9);
// log prior of the tree scale: half-Cauchy, mode 1, scale 0.5
double omega_lpr_rho(double rho);
// log prior of kappa: log gamma density (shape 7, scale 0.25) at kappa - 1
double omega_lpr_kappa(double kappa);

}  // namespace pcsf
