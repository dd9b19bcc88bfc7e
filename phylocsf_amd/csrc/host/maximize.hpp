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

// PhyloCSFModel.maximize_lpr ~init ~lo ~hi ~accuracy f g
// (src/PhyloCSFModel.ml:84-99) for one function, one evaluation at a time:
// Fit.find_init (f at lo, hi, init, then Random.init 0 retries in log space,
// at most 250) and, when its point is interior, GSL Brent until the bracket
// is within `accuracy` of the estimate; the result is (x, f x).  Values at
// abscissae already evaluated are reused (the reference evaluates them again
// and gets the same numbers), so pending() is always a new abscissa.
class MaximizeLpr {
 public:
  void start(double init, double lo, double hi, double accuracy, double rel_tol);
  bool done() const { return phase_ == kDone; }
  double pending() const { return pending_; }  // the abscissa to evaluate next
  void feed(double f);                         // f(pending())
  double x = 0, fx = 0;  // result once done
  bool failed = false;   // a non-finite value where Brent needs a finite one
  int evals = 0;         // evaluations fed

 private:
  enum Phase { kLo, kHi, kInit, kRetry, kBrent, kFinal, kDone };
  void after_find_init();
  void brent_next();
  bool recall(double xx, double* f) const;
  void take(double f);  // one step of the state machine with f(pending())
  Phase phase_ = kDone;
  double lo_ = 0, hi_ = 0, acc_ = 0, rel_tol_ = 0;
  double flo_ = 0, fhi_ = 0, xc_ = 0, fc_ = 0;
  int tries_ = 0;
  BrentMin br_;
  bool brent_started_ = false;
  double pending_ = 0;
  std::vector<double> seen_x_, seen_f_;
};

}  // namespace pcsf
