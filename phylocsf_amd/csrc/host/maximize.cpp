// maximize.cpp -- the reference's tree-scale maximiser, resumable.
//
// Third-party algorithms restated (not present in the reference tree):
//  * GSL gsl_min_fminimizer_brent (min/brent.c, GNU Scientific Library;
//    version unpinned by the reference's CI): Brent 1973 with v = w = lower +
//    0.3819660 (upper - lower), tolerance sqrt(DBL_EPSILON) |x|, parabolic
//    step when acceptable, golden section otherwise, and GSL's bracket and
//    (v, w) update rules.
//  * OCaml 4.x Stdlib.Random: 55-word additive lagged Fibonacci generator
//    with 30-bit outputs; Random.init s seeds it through chained MD5
//    (Digest.string) of "x" ^ string_of_int s; Random.float b =
//    ((r1 / 2^30 + r2) / 2^30) * b.
//  * MD5, RFC 1321.
#include "maximize.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>

namespace pcsf {

namespace {
inline uint32_t rol(uint32_t x, int s) { return (x << s) | (x >> (32 - s)); }
}  // namespace

void md5_digest(const uint8_t* msg, uint64_t len, uint8_t out[16]) {
  static const uint8_t kShift[4][4] = {{7, 12, 17, 22}, {5, 9, 14, 20}, {4, 11, 16, 23},
                                       {6, 10, 15, 21}};
  uint32_t T[64];
  for (int i = 0; i < 64; ++i) T[i] = (uint32_t)(std::fabs(std::sin(i + 1.0)) * 4294967296.0);
  std::vector<uint8_t> buf(msg, msg + len);
  buf.push_back(0x80);
  while (buf.size() % 64 != 56) buf.push_back(0);
  for (int i = 0; i < 8; ++i) buf.push_back((uint8_t)((len * 8) >> (8 * i)));
  uint32_t h[4] = {0x67452301u, 0xefcdab89u, 0x98badcfeu, 0x10325476u};
  for (size_t off = 0; off < buf.size(); off += 64) {
    uint32_t X[16];
    for (int i = 0; i < 16; ++i)
      X[i] = buf[off + 4 * i] | (buf[off + 4 * i + 1] << 8) | (buf[off + 4 * i + 2] << 16) |
             ((uint32_t)buf[off + 4 * i + 3] << 24);
    uint32_t a = h[0], b = h[1], c = h[2], d = h[3];
    for (int i = 0; i < 64; ++i) {
      const int round = i / 16;
      uint32_t f;
      int k;
      switch (round) {
        case 0: f = (b & c) | (~b & d); k = i; break;
        case 1: f = (b & d) | (c & ~d); k = (5 * i + 1) & 15; break;
        case 2: f = b ^ c ^ d; k = (3 * i + 5) & 15; break;
        default: f = c ^ (b | ~d); k = (7 * i) & 15; break;
      }
      const uint32_t nb = b + rol(a + f + T[i] + X[k], kShift[round][i & 3]);
      a = d;
      d = c;
      c = b;
      b = nb;
    }
    h[0] += a;
    h[1] += b;
    h[2] += c;
    h[3] += d;
  }
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) out[4 * i + j] = (uint8_t)(h[i] >> (8 * j));
}

OcamlRandom::OcamlRandom(int seed) {
  for (int i = 0; i < 55; ++i) st_[i] = i;
  const std::string num = std::to_string(seed);
  std::string accu = "x";
  for (int i = 0; i < 55 + 55; ++i) {  // 0 .. 54 + max 55 (length [|seed|])
    std::string m = accu + num;
    uint8_t d[16];
    md5_digest(reinterpret_cast<const uint8_t*>(m.data()), m.size(), d);
    accu.assign(reinterpret_cast<const char*>(d), 16);
    const int64_t ex = d[0] | (d[1] << 8) | (d[2] << 16) | ((int64_t)d[3] << 24);
    st_[i % 55] = (st_[i % 55] ^ ex) & 0x3FFFFFFF;
  }
  idx_ = 0;
}

int64_t OcamlRandom::bits() {
  idx_ = (idx_ + 1) % 55;
  const int64_t cur = st_[idx_];
  const int64_t nv = (st_[(idx_ + 24) % 55] + (cur ^ ((cur >> 25) & 0x1F))) & 0x3FFFFFFF;
  st_[idx_] = nv;
  return nv;
}

double OcamlRandom::uniform(double bound) {
  const double scale = 1073741824.0;
  const double r1 = (double)bits();
  const double r2 = (double)bits();
  return ((r1 / scale + r2) / scale) * bound;
}

std::vector<double> find_init_retries(double lo, double hi, int n) {
  OcamlRandom rng(0);
  const double width = std::log(hi) - std::log(lo);
  std::vector<double> xs(n);
  for (int i = 0; i < n; ++i) xs[i] = std::exp(std::log(lo) + rng.uniform(width));
  return xs;
}

void BrentMin::init(double xmin, double fmin, double lo, double flo, double hi, double fhi,
                    double fv) {
  x_min = xmin;
  f_min = fmin;
  x_lo = lo;
  f_lo = flo;
  x_hi = hi;
  f_hi = fhi;
  v = w = golden_point();
  f_v = f_w = fv;
  d = e = 0.0;
  u = xmin;
}

double BrentMin::golden_point() const { return x_lo + kGolden * (x_hi - x_lo); }

void BrentMin::propose() {
  const double tol = rel_tol * std::fabs(x_min);  // GSL: GSL_SQRT_DBL_EPSILON * |x|
  const double mid = 0.5 * (x_lo + x_hi);
  const double z = x_min;
  // GSL keeps the last two steps in (d, e) and swaps them on entry
  double step = e, prev = d;
  double p = 0.0, q = 0.0, r = 0.0;
  if (std::fabs(prev) > tol) {
    r = (z - w) * (f_min - f_v);
    q = (z - v) * (f_min - f_w);
    p = (z - v) * q - (z - w) * r;
    q = 2.0 * (q - r);
    if (q > 0.0) p = -p;
    else q = -q;
    r = prev;
    prev = step;
  }
  if (std::fabs(p) < std::fabs(0.5 * q * r) && p < q * (z - x_lo) && p < q * (x_hi - z)) {
    step = p / q;
    const double trial = z + step;
    if ((trial - x_lo) < 2.0 * tol || (x_hi - trial) < 2.0 * tol)
      step = (z < mid) ? tol : -tol;
  } else {
    prev = (z < mid) ? x_hi - z : -(z - x_lo);
    step = kGolden * prev;
  }
  u = (std::fabs(step) >= tol) ? z + step : z + ((step > 0.0) ? tol : -tol);
  e = prev;
  d = step;
}

void BrentMin::accept(double fu) {
  const double z = x_min, fz = f_min;
  if (fu <= fz) {
    if (u < z) {
      x_hi = z;
      f_hi = fz;
    } else {
      x_lo = z;
      f_lo = fz;
    }
    v = w;
    f_v = f_w;
    w = z;
    f_w = fz;
    x_min = u;
    f_min = fu;
    return;
  }
  if (u < z) {
    x_lo = u;
    f_lo = fu;
  } else {
    x_hi = u;
    f_hi = fu;
  }
  if (fu <= f_w || w == z) {
    v = w;
    f_v = f_w;
    w = u;
    f_w = fu;
  } else if (fu <= f_v || v == z || v == w) {
    v = u;
    f_v = fu;
  }
}

namespace {
const std::vector<double>& retries_of(double lo, double hi) {
  // the retry abscissae depend only on (lo, hi): Random.init 0 each time
  static thread_local double c_lo = -1, c_hi = -1;
  static thread_local std::vector<double> c;
  if (lo != c_lo || hi != c_hi) {
    c = find_init_retries(lo, hi, 250);
    c_lo = lo;
    c_hi = hi;
  }
  return c;
}
}  // namespace

void MaximizeLpr::start(double init, double lo, double hi, double accuracy, double rel_tol) {
  lo_ = lo;
  hi_ = hi;
  acc_ = accuracy;
  rel_tol_ = rel_tol;
  xc_ = init;
  tries_ = 0;
  failed = false;
  evals = 0;
  brent_started_ = false;
  seen_x_.clear();
  seen_f_.clear();
  phase_ = kLo;
  pending_ = lo;
}

bool MaximizeLpr::recall(double xx, double* f) const {
  for (size_t k = seen_x_.size(); k-- > 0;)
    if (seen_x_[k] == xx) {
      *f = seen_f_[k];
      return true;
    }
  return false;
}

// Fit.find_init's loop condition and exit (lib/CamlPaml/Fit.ml:36-48), then
// the start of the maximiser when the point is interior.
void MaximizeLpr::after_find_init() {
  if (tries_ < 250 && (fc_ <= flo_ || fc_ <= fhi_)) {
    xc_ = retries_of(lo_, hi_)[tries_];
    ++tries_;
    phase_ = kRetry;
    pending_ = xc_;
    return;
  }
  // Fit.ml:43-47: after the 250th try the better end is taken, whatever that
  // try gave
  if (tries_ == 250) xc_ = (flo_ > fhi_) ? lo_ : hi_;
  if (lo_ < xc_ && xc_ < hi_) {
    // gsl_min_fminimizer_set: lower, upper and the minimum must be finite
    // and enclose the minimum of -f
    const double fmin = -fc_, flo = -flo_, fhi = -fhi_;
    if (!std::isfinite(fmin) || !std::isfinite(flo) || !std::isfinite(fhi) ||
        !(fmin < flo && fmin < fhi)) {
      failed = true;
      phase_ = kDone;
      return;
    }
    // brent_init's f at the golden point
    br_.x_lo = lo_;
    br_.x_hi = hi_;
    phase_ = kBrent;
    brent_started_ = false;
    pending_ = br_.golden_point();
    return;
  }
  phase_ = kFinal;
  pending_ = xc_;
}

void MaximizeLpr::brent_next() {
  br_.propose();
  pending_ = br_.u;
}

void MaximizeLpr::feed(double f) {
  ++evals;
  seen_x_.push_back(pending_);
  seen_f_.push_back(f);
  take(f);
  double known;
  while (phase_ != kDone && recall(pending_, &known)) take(known);
}

void MaximizeLpr::take(double f) {
  switch (phase_) {
    case kLo:
      flo_ = f;
      phase_ = kHi;
      pending_ = hi_;
      break;
    case kHi:
      fhi_ = f;
      phase_ = kInit;
      pending_ = xc_;
      break;
    case kInit:
    case kRetry:
      fc_ = f;
      after_find_init();
      break;
    case kBrent:
      if (!brent_started_) {
        br_.init(xc_, -fc_, lo_, -flo_, hi_, -fhi_, -f);
        br_.rel_tol = rel_tol_;
        brent_started_ = true;
        brent_next();
        break;
      }
      if (!std::isfinite(f)) {
        failed = true;
        phase_ = kDone;
        break;
      }
      br_.accept(-f);
      if ((br_.x_hi - br_.x_lo) / br_.x_min > acc_) {
        brent_next();
      } else {
        phase_ = kFinal;
        pending_ = br_.x_min;
      }
      break;
    case kFinal:
      x = pending_;
      fx = f;
      phase_ = kDone;
      return;
    default:
      return;
  }
}

}  // namespace pcsf
