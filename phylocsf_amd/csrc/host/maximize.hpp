// maximize.hpp -- the reference's 1-D maximiser of lpr over the tree scale,
// as resumable state machines so many regions advance in lock step.
//
//   maximize_lpr ............ src/PhyloCSFModel.ml:84-99
//   Fit.find_init ........... lib/CamlPaml/Fit.ml:27-48 (Random.init 0, log space)
//   Fit.maximizer ........... lib/CamlPaml/Fit.ml:3-24 (GSL Brent on -f)
#pragma once
#include <cstdint>
#include <vector>

namespace pcsf {

// OCaml 4.x Stdlib.Random after Random.init seed: Random.float draws.
class OcamlRandom {
 public:
  explicit OcamlRandom(int seed);
  double uniform(double bound);  // Random.float bound

 private:
  int64_t bits();
  int64_t st_[55];
  int idx_ = 0;
};

void md5_digest(const uint8_t* msg, uint64_t len, uint8_t out[16]);

// The abscissae Fit.find_init tries after its first three evaluations:
// x_k = exp(log lo + Random.float (log hi - log lo)), k = 1..n, Random.init 0.
std::vector<double> find_init_retries(double lo, double hi, int n);

// GSL's Brent minimiser (gsl_min_fminimizer_brent), driven one function value
// at a time.  Minimises f = -lpr.
struct BrentMin {
  double x_min, f_min, x_lo, f_lo, x_hi, f_hi;
  double v, w, f_v, f_w, d, e;
  double u;  // pending abscissa
  // smallest step relative to |x_min|: GSL's sqrt(DBL_EPSILON); callers whose
  // function values carry more rounding noise than a step of that size can
  // resolve (FP32 / FP16 likelihoods) raise it
  double rel_tol = 1.4901161193847656e-08;

  // gsl_min_fminimizer_set_with_values + brent_init with f(v) already known
  void init(double xmin, double fmin, double lo, double flo, double hi, double fhi,
            double fv);
  double golden_point() const;  // v of brent_init for [lo, hi]
  void propose();               // first half of an iteration: sets u
  void accept(double fu);       // second half: bracket update with f(u)
};

constexpr double kGolden = 0.3819660;  // GSL's (3 - sqrt 5) / 2

}  // namespace pcsf
