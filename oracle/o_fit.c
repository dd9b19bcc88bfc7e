/*
 * o_fit.c -- oracle restatement of the reference's 1-D maximisation.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 *   Fit.maximizer / make_maximizer ... lib/CamlPaml/Fit.ml:3-24
 *   Fit.find_init .................... lib/CamlPaml/Fit.ml:27-48
 *   maximize_lpr stop rule ........... src/PhyloCSFModel.ml:84-99
 *
 * Third-party algorithms restated here (absent from the reference tree):
 *   - GSL's gsl_min_fminimizer_brent (GNU Scientific Library 1.x/2.x,
 *     min/brent.c; version unpinned by the reference's CI, .travis-ci.sh:1,5):
 *     Brent's method (R. P. Brent, Algorithms for Minimization without
 *     Derivatives, 1973, ch. 5) with GSL's initialisation v = w = lower +
 *     golden*(upper-lower), tolerance sqrt(DBL_EPSILON)*|x|, and its
 *     bracket-update rules.  gsl_min_fminimizer_set evaluates f at lower,
 *     upper and the initial minimum and requires f(min) < f(lower), f(upper).
 *   - OCaml 4.x Stdlib.Random (stdlib/random.ml): 55-word additive lagged
 *     Fibonacci generator with 30-bit outputs, seeded by full_init through
 *     MD5 (Digest.string) of "x" ^ string_of_int seed, chained.
 *   - MD5, RFC 1321.
 */
#include "oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

static int fail(char *err, int errlen, int code, const char *fmt, ...) {
  if (err && errlen > 0) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(err, (size_t)errlen, fmt, ap);
    va_end(ap);
  }
  return code;
}

/* ---- MD5 (RFC 1321) ----------------------------------------------------- */

static uint32_t rotl(uint32_t x, int c) { return (x << c) | (x >> (32 - c)); }

void o_md5(const uint8_t *msg, uint64_t len, uint8_t out[16]) {
  static const int r[64] = {7, 12, 17, 22, 7, 12, 17, 22, 7, 12, 17, 22, 7, 12, 17, 22,
                            5, 9,  14, 20, 5, 9,  14, 20, 5, 9,  14, 20, 5, 9,  14, 20,
                            4, 11, 16, 23, 4, 11, 16, 23, 4, 11, 16, 23, 4, 11, 16, 23,
                            6, 10, 15, 21, 6, 10, 15, 21, 6, 10, 15, 21, 6, 10, 15, 21};
  uint32_t K[64];
  for (int i = 0; i < 64; i++) K[i] = (uint32_t)(fabs(sin((double)(i + 1))) * 4294967296.0);
  uint32_t h0 = 0x67452301u, h1 = 0xefcdab89u, h2 = 0x98badcfeu, h3 = 0x10325476u;
  uint64_t nblk = (len + 8) / 64 + 1;
  for (uint64_t b = 0; b < nblk; b++) {
    uint8_t blk[64];
    for (int i = 0; i < 64; i++) {
      uint64_t pos = b * 64 + (uint64_t)i;
      uint8_t v;
      if (pos < len) v = msg[pos];
      else if (pos == len) v = 0x80;
      else if (pos >= nblk * 64 - 8) v = (uint8_t)(((len * 8) >> (8 * (pos - (nblk * 64 - 8)))) & 0xff);
      else v = 0;
      blk[i] = v;
    }
    uint32_t w[16];
    for (int i = 0; i < 16; i++)
      w[i] = (uint32_t)blk[4 * i] | ((uint32_t)blk[4 * i + 1] << 8) |
             ((uint32_t)blk[4 * i + 2] << 16) | ((uint32_t)blk[4 * i + 3] << 24);
    uint32_t a = h0, bb = h1, c = h2, d = h3;
    for (int i = 0; i < 64; i++) {
      uint32_t f;
      int g;
      if (i < 16) { f = (bb & c) | (~bb & d); g = i; }
      else if (i < 32) { f = (d & bb) | (~d & c); g = (5 * i + 1) % 16; }
      else if (i < 48) { f = bb ^ c ^ d; g = (3 * i + 5) % 16; }
      else { f = c ^ (bb | ~d); g = (7 * i) % 16; }
      uint32_t tmp = d;
      d = c;
      c = bb;
      bb = bb + rotl(a + f + K[i] + w[g], r[i]);
      a = tmp;
    }
    h0 += a;
    h1 += bb;
    h2 += c;
    h3 += d;
  }
  uint32_t hs[4] = {h0, h1, h2, h3};
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) out[4 * i + j] = (uint8_t)((hs[i] >> (8 * j)) & 0xff);
}

/* ---- OCaml 4.x Random --------------------------------------------------- */

typedef struct {
  int64_t st[55];
  int idx;
} ocaml_rng;

static void ocaml_full_init(ocaml_rng *s, int seed) {
  for (int i = 0; i < 55; i++) s->st[i] = i;
  uint8_t accu[64];
  size_t alen = 1;
  accu[0] = 'x';
  char num[32];
  int nl = snprintf(num, sizeof num, "%d", seed);
  for (int i = 0; i <= 54 + 55; i++) { /* 54 + max 55 (length seed = 1) */
    int j = i % 55;
    uint8_t msg[64];
    memcpy(msg, accu, alen);
    memcpy(msg + alen, num, (size_t)nl);
    uint8_t d[16];
    o_md5(msg, alen + (size_t)nl, d);
    memcpy(accu, d, 16);
    alen = 16;
    int64_t ex = (int64_t)d[0] + ((int64_t)d[1] << 8) + ((int64_t)d[2] << 16) + ((int64_t)d[3] << 24);
    s->st[j] = (s->st[j] ^ ex) & 0x3FFFFFFF;
  }
  s->idx = 0;
}

static int64_t ocaml_bits(ocaml_rng *s) {
  s->idx = (s->idx + 1) % 55;
  int64_t curval = s->st[s->idx];
  int64_t newval = s->st[(s->idx + 24) % 55] + (curval ^ ((curval >> 25) & 0x1F));
  int64_t newval30 = newval & 0x3FFFFFFF;
  s->st[s->idx] = newval30;
  return newval30;
}

static double ocaml_float(ocaml_rng *s, double bound) {
  const double scale = 1073741824.0;
  double r1 = (double)ocaml_bits(s);
  double r2 = (double)ocaml_bits(s);
  return ((r1 / scale + r2) / scale) * bound;
}

void o_ocaml_random_floats(int seed, double bound, int n, double *out) {
  ocaml_rng s;
  ocaml_full_init(&s, seed);
  for (int i = 0; i < n; i++) out[i] = ocaml_float(&s, bound);
}

/* ---- Fit.find_init (Fit.ml:27-48) ---------------------------------------- */

double o_find_init(o_fn f, void *ctx, int maxtries, int logspace, double init, double lo,
                   double hi, int *evals) {
  double width = logspace ? log(hi) - log(lo) : hi - lo;
  double flo = f(lo, ctx);
  double fhi = f(hi, ctx);
  double x = init;
  double fx = f(init, ctx);
  int e = 3;
  int i = 0;
  ocaml_rng rng;
  ocaml_full_init(&rng, 0); /* Random.init 0 */
  while (i < maxtries && (fx <= flo || fx <= fhi)) {
    if (logspace) x = exp(log(lo) + ocaml_float(&rng, width));
    else x = lo + ocaml_float(&rng, width);
    fx = f(x, ctx);
    e++;
    i++;
  }
  if (i == maxtries) x = (flo > fhi) ? lo : hi;
  if (evals) *evals = e;
  return x;
}

/* ---- GSL Brent on -f (Fit.ml:3-24), stopped per PhyloCSFModel.ml:89-94 --- */

typedef struct {
  double d, e, v, w, f_v, f_w;
} brent_state;

int o_brent_max(o_fn f, void *ctx, double lo, double hi, double accuracy, double *x_out,
                int *evals_out, char *err, int errlen) {
  /* the caller guarantees lo < init < hi with f(init) above both ends; the
   * initial minimum is passed through *x_out */
  double x_minimum = *x_out, x_lower = lo, x_upper = hi;
  int evals = 0;
#define FN(x) (evals++, 0. - f((x), ctx))
  double f_lower = FN(x_lower);
  double f_upper = FN(x_upper);
  double f_minimum = FN(x_minimum);
  if (!isfinite(f_lower) || !isfinite(f_upper) || !isfinite(f_minimum))
    return fail(err, errlen, O_EFAIL, "GSL: function value is not finite");
  if (x_lower > x_upper) return fail(err, errlen, O_EINVAL, "GSL: invalid interval");
  if (x_minimum >= x_upper || x_minimum <= x_lower)
    return fail(err, errlen, O_EINVAL, "GSL: x_minimum must lie inside interval");
  if (f_minimum >= f_lower || f_minimum >= f_upper)
    return fail(err, errlen, O_EINVAL, "GSL: endpoints do not enclose a minimum");
  const double golden = 0.3819660;
  const double sqrt_eps = 1.4901161193847656e-08;
  brent_state st;
  st.v = x_lower + golden * (x_upper - x_lower);
  st.w = st.v;
  st.d = 0.;
  st.e = 0.;
  {
    double fv = FN(st.v);
    if (!isfinite(fv)) return fail(err, errlen, O_EFAIL, "GSL: function value is not finite");
    st.f_v = fv;
    st.f_w = fv;
  }
  for (int it = 0; it < 10000; it++) {
    /* ---- one gsl brent iteration ---- */
    const double x_left = x_lower, x_right = x_upper, z = x_minimum;
    double d = st.e, e = st.d, u, f_u;
    const double v = st.v, w = st.w, f_v = st.f_v, f_w = st.f_w, f_z = f_minimum;
    const double w_lower = z - x_left, w_upper = x_right - z;
    const double tolerance = sqrt_eps * fabs(z);
    double p = 0, q = 0, r = 0;
    const double midpoint = 0.5 * (x_left + x_right);
    if (fabs(e) > tolerance) {
      r = (z - w) * (f_z - f_v);
      q = (z - v) * (f_z - f_w);
      p = (z - v) * q - (z - w) * r;
      q = 2 * (q - r);
      if (q > 0) p = -p;
      else q = -q;
      r = e;
      e = d;
    }
    if (fabs(p) < fabs(0.5 * q * r) && p < q * w_lower && p < q * w_upper) {
      double t2 = 2 * tolerance;
      d = p / q;
      u = z + d;
      if ((u - x_left) < t2 || (x_right - u) < t2) d = (z < midpoint) ? tolerance : -tolerance;
    } else {
      e = (z < midpoint) ? x_right - z : -(z - x_left);
      d = golden * e;
    }
    if (fabs(d) >= tolerance) u = z + d;
    else u = z + ((d > 0) ? tolerance : -tolerance);
    st.e = e;
    st.d = d;
    f_u = FN(u);
    if (!isfinite(f_u)) return fail(err, errlen, O_EFAIL, "GSL: function value is not finite");
    if (f_u <= f_z) {
      if (u < z) {
        x_upper = z;
        f_upper = f_z;
      } else {
        x_lower = z;
        f_lower = f_z;
      }
      st.v = w;
      st.f_v = f_w;
      st.w = z;
      st.f_w = f_z;
      x_minimum = u;
      f_minimum = f_u;
    } else {
      if (u < z) {
        x_lower = u;
        f_lower = f_u;
      } else {
        x_upper = u;
        f_upper = f_u;
      }
      if (f_u <= f_w || w == z) {
        st.v = w;
        st.f_v = f_w;
        st.w = u;
        st.f_w = f_u;
      } else if (f_u <= f_v || v == z || v == w) {
        st.v = u;
        st.f_v = f_u;
      }
    }
    /* ---- maximize_lpr's stop rule ---- */
    if (!((x_upper - x_lower) / x_minimum > accuracy)) break;
  }
#undef FN
  (void)f_lower;
  (void)f_upper;
  *x_out = x_minimum;
  if (evals_out) *evals_out = evals;
  return 0;
}
