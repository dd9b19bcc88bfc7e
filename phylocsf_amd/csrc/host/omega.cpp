// omega.cpp -- the codon model of the omega (dN/dS) test, host side: codon
// frequencies and exchangeabilities from the twelve settings, the F3x4
// frequency settings of a region, and the two priors of the MAP fit.
// Paths are relative to the mlin/PhyloCSF tree.
#include "omega.hpp"

#include <cmath>

namespace pcsf {

namespace {
// Code.Codon64: index 16 n1 + 4 n2 + n3 with A, C, G, T = 0..3; the amino
// acid of each codon (lib/CamlPaml/Code.ml:153-173) and the stops TAA, TAG,
// TGA (:146-150)
constexpr char kAmino[65] = "KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVV*Y*YSSSS*CWCLFLF";
constexpr int kTAA = 48, kTAG = 50, kTGA = 56;
inline bool stop_codon(int c) { return c == kTAA || c == kTAG || c == kTGA; }
inline bool transition(int x, int y) {  // A <-> G, C <-> T
  return (x ^ y) == 2;
}
}  // namespace

// src/OmegaModel.ml:24-40
void omega_pi(const double* qs, double* pi) {
  double f[3][4];
  for (int p = 0; p < 3; ++p) {
    const double* r = qs + 3 + 3 * p;
    const double tot = r[0] + (r[1] + (r[2] + 1.0));
    for (int n = 0; n < 4; ++n) f[p][n] = (n == 3 ? 1.0 : r[n]) / tot;
  }
  double sc[64];
  for (int c = 0; c < 64; ++c) sc[c] = f[0][c >> 4] * (f[1][(c >> 2) & 3] * f[2][c & 3]);
  const double denom = 1.0 - (1.0 - qs[2]) * (sc[kTAA] + (sc[kTAG] + sc[kTGA]));
  for (int c = 0; c < 64; ++c) pi[c] = sc[c] / denom;
}

// src/OmegaModel.ml:42-74
void omega_exchangeabilities(const double* qs, double* S) {
  for (int i = 0; i < 64; ++i)
    for (int j = 0; j < 64; ++j) {
      const int d = i ^ j;
      // one nucleotide apart: the differing bits lie in one 2-bit field
      int field = -1;
      for (int p = 0; p < 3; ++p)
        if (d == (d & (3 << (2 * p)))) field = p;
      double s = 0.0;
      if (d != 0 && field >= 0) {
        const int x = (i >> (2 * field)) & 3, y = (j >> (2 * field)) & 3;
        const double kp = transition(x, y) ? qs[0] : 1.0;
        const double op =
            (!stop_codon(i) && !stop_codon(j) && kAmino[i] != kAmino[j]) ? qs[1] : 1.0;
        s = kp * op;
      }
      S[i * 64 + j] = s;
    }
}

// src/OmegaModel.ml:102-134
void omega_f3x4(const uint8_t* codes, int64_t ncols, int nl, double* qs9) {
  int64_t cnt[3][4];
  for (auto& row : cnt)
    for (auto& x : row) x = 1;
  for (int64_t k = 0; k < ncols * nl; ++k) {
    const int v = codes[k];
    if (v > 63) continue;  // `Marginalize
    ++cnt[0][v >> 4];
    ++cnt[1][(v >> 2) & 3];
    ++cnt[2][v & 3];
  }
  for (int p = 0; p < 3; ++p)
    for (int n = 0; n < 3; ++n) qs9[3 * p + n] = (double)cnt[p][n] / (double)cnt[p][3];
}

// half_cauchy_lpdf ~mode:1.0 ~scale:0.5 (src/OmegaModel.ml:147-156)
double omega_lpr_rho(double rho) {
  const double pi = std::acos(-1.0), mode = 1.0, scale = 0.5;
  const double numer = 1.0 / (pi * scale * (1.0 + std::pow((rho - mode) / scale, 2.0)));
  const double denom = 1.0 - (std::atan((0.0 - mode) / scale) / pi + 0.5);
  return std::log(numer) - std::log(denom);
}

// log (gsl_ran_gamma_pdf (kappa - 1 + DBL_EPSILON, 7, 0.25)) (src/OmegaModel.ml:157):
// GSL's density exp((a - 1) ln(x / b) - x / b - ln Gamma(a)) / b
double omega_lpr_kappa(double kappa) {
  const double a = 7.0, b = 0.25, x = kappa - 1.0 + 2.220446049250313e-16;
  if (!(x > 0.0)) return -INFINITY;
  return std::log(std::exp((a - 1.0) * std::log(x / b) - x / b - std::lgamma(a)) / b);
}

}  // namespace pcsf
