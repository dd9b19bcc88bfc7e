/*
 * o_model.c -- oracle restatement of the reference's likelihood arithmetic,
 * in FP64 with the reference's loop structure.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 *   Q build ............... src/PhyloCSFModel.ml:11-30, lib/CamlPaml/PhyloModel.ml:76-106
 *   P(t) .................. lib/CamlPaml/Q.ml:211-249 (clamp + diagonal fix-up)
 *   pruning ............... lib/CamlPaml/PhyloLik.ml:9-138 (ensure_alpha / ensure_beta)
 *   branch posteriors ..... lib/CamlPaml/PhyloLik.ml:140-174
 *   d/dt, d/drho .......... lib/CamlPaml/PhyloEM.ml:140-182, Q.ml:260-275
 *   lpr_leaves ............ src/PhyloCSFModel.ml:67-82
 *   llr_FixedLik .......... src/PhyloCSFModel.ml:113,122-128
 *
 * The eigendecomposition is a cyclic Jacobi method (Golub & Van Loan,
 * Matrix Computations, 4th ed., sec. 8.5) applied to the symmetric matrix
 * Pi^1/2 Q Pi^-1/2 of a reversible Q; the reference uses GSL nonsymmv on Q
 * itself, and P(t) is the same matrix up to rounding.
 *
 * Build with -ffp-contract=off: GSL's ddot (a plain sequential loop) does
 * separate multiplies and adds on x86-64.
 */
#include "oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
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

/* GSL cblas ddot: r += x[i]*y[i] in index order (PhyloLik.ml:90,92). */
static inline double ddot(int k, const double *x, int incx, const double *y, int incy) {
  double r = 0.;
  for (int i = 0; i < k; i++) r += x[i * incx] * y[i * incy];
  return r;
}

/* ------------------------------------------------------------------------ */

int o_build_q(int k, const double *S, const double *pi, double *Q, char *err, int errlen) {
  /* q[i,j] = s[i,j] * pi[j]   (PhyloCSFModel.ml:11-21) */
  for (int i = 0; i < k; i++)
    for (int j = 0; j < k; j++) Q[i * k + j] = (i == j) ? 0. : S[i * k + j] * pi[j];
  /* diagonal: 0 - (q[i,n-1] + (... + (q[i,0] + 0)))   (PhyloModel.ml:76-84) */
  for (int i = 0; i < k; i++) {
    double tot = 0.;
    for (int j = 0; j < k; j++)
      if (j != i) tot = Q[i * k + j] + tot;
    Q[i * k + i] = 0. - tot;
  }
  /* scale = 0 - pi_0 q_00 - pi_1 q_11 - ...   (PhyloCSFModel.ml:24-30) */
  double scale = 0.;
  for (int i = 0; i < k; i++)
    if (Q[i * k + i] != 0.) scale = scale - pi[i] * Q[i * k + i];
  if (!(scale > 0.))
    return fail(err, errlen, O_EFAIL,
                "CamlPaml.P14n.instantiate_q: Q scale evaluated to a non-positive value");
  /* Q = q / scale   (PhyloModel.ml:102) */
  for (int i = 0; i < k * k; i++) Q[i] = Q[i] / scale;
  return 0;
}

/* cyclic Jacobi on symmetric A (k x k, destroyed); eigenvalues in w,
 * eigenvectors in the columns of U */
static void jacobi_eig(int k, double *A, double *w, double *U) {
  for (int i = 0; i < k; i++)
    for (int j = 0; j < k; j++) U[i * k + j] = (i == j) ? 1. : 0.;
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0., dia = 0.;
    for (int p = 0; p < k; p++) {
      dia += A[p * k + p] * A[p * k + p];
      for (int q = p + 1; q < k; q++) off += A[p * k + q] * A[p * k + q];
    }
    if (off <= 1e-34 * dia || off == 0.) break;
    for (int p = 0; p < k - 1; p++)
      for (int q = p + 1; q < k; q++) {
        double apq = A[p * k + q];
        if (apq == 0.) continue;
        double app = A[p * k + p], aqq = A[q * k + q];
        double theta = (aqq - app) / (2. * apq);
        double t = (theta >= 0. ? 1. : -1.) / (fabs(theta) + sqrt(theta * theta + 1.));
        double c = 1. / sqrt(t * t + 1.), s = t * c;
        for (int r = 0; r < k; r++) { /* A <- A J (columns p,q) */
          double arp = A[r * k + p], arq = A[r * k + q];
          A[r * k + p] = c * arp - s * arq;
          A[r * k + q] = s * arp + c * arq;
        }
        for (int r = 0; r < k; r++) { /* A <- J^T A (rows p,q) */
          double apr = A[p * k + r], aqr = A[q * k + r];
          A[p * k + r] = c * apr - s * aqr;
          A[q * k + r] = s * apr + c * aqr;
        }
        A[p * k + q] = A[q * k + p] = 0.;
        for (int r = 0; r < k; r++) { /* U <- U J */
          double urp = U[r * k + p], urq = U[r * k + q];
          U[r * k + p] = c * urp - s * urq;
          U[r * k + q] = s * urp + c * urq;
        }
      }
  }
  for (int i = 0; i < k; i++) w[i] = A[i * k + i];
}

typedef struct {
  int k;
  double *w;  /* eigenvalues */
  double *L;  /* L[i][m] = U[i][m] / sqrt(pi_i)   (right eigenvectors) */
  double *R;  /* R[m][j] = U[j][m] * sqrt(pi_j)   (left eigenvectors)  */
} o_eig;

static int eig_of_q(int k, const double *Q, const double *pi, o_eig *E, char *err, int errlen) {
  for (int i = 0; i < k; i++)
    if (!(pi[i] > 0.))
      return fail(err, errlen, O_EINVAL, "oracle: stationary distribution must be positive");
  double *A = (double *)malloc(sizeof(double) * (size_t)(k * k));
  double *U = (double *)malloc(sizeof(double) * (size_t)(k * k));
  for (int i = 0; i < k; i++)
    for (int j = 0; j < k; j++) {
      double a = sqrt(pi[i]) * Q[i * k + j] / sqrt(pi[j]);
      double b = sqrt(pi[j]) * Q[j * k + i] / sqrt(pi[i]);
      A[i * k + j] = 0.5 * (a + b);
    }
  E->k = k;
  E->w = (double *)malloc(sizeof(double) * (size_t)k);
  E->L = (double *)malloc(sizeof(double) * (size_t)(k * k));
  E->R = (double *)malloc(sizeof(double) * (size_t)(k * k));
  jacobi_eig(k, A, E->w, U);
  for (int i = 0; i < k; i++)
    for (int m = 0; m < k; m++) {
      E->L[i * k + m] = U[i * k + m] / sqrt(pi[i]);
      E->R[m * k + i] = U[i * k + m] * sqrt(pi[i]);
    }
  free(A);
  free(U);
  return 0;
}

static void eig_free(o_eig *E) {
  free(E->w);
  free(E->L);
  free(E->R);
}

/* Q.Diag.real_to_Pt (Q.ml:211-249) */
static int eig_pt(const o_eig *E, double t, double tol, double *P, char *err, int errlen) {
  int k = E->k;
  if (t < 0.) return fail(err, errlen, O_EINVAL, "CamlPaml.Q.to_Pt");
  double ex[256];
  for (int m = 0; m < k; m++) ex[m] = exp(t * E->w[m]);
  for (int i = 0; i < k; i++)
    for (int j = 0; j < k; j++) {
      double s = 0.;
      for (int m = 0; m < k; m++) s += E->L[i * k + m] * (ex[m] * E->R[m * k + j]);
      P[i * k + j] = s;
    }
  for (int i = 0; i < k; i++) {
    double tot = 0., smii = 1.;
    for (int j = 0; j < k; j++) {
      tot += P[i * k + j];
      if (P[i * k + j] < 0.) {
        if (fabs(P[i * k + j]) > tol)
          return fail(err, errlen, O_EFAIL,
                      "CamlPaml.Q.substition_matrix: expm(%.2e*Q)[%d,%d] = %e < 0", t, i, j,
                      P[i * k + j]);
        P[i * k + j] = 0.;
      }
      if (i != j) smii -= P[i * k + j];
    }
    if (fabs(tot - 1.) > tol)
      return fail(err, errlen, O_EFAIL,
                  "CamlPaml.Q.substitution matrix: sum(expm(%.2e*Q)[%d,] = %e > 1", t, i, tot);
    if (!(smii <= 1. && smii > 0.)) return fail(err, errlen, O_EFAIL, "Assert_failure Q.ml:245");
    P[i * k + i] = smii;
  }
  return 0;
}

int o_pt(int k, const double *Q, const double *pi, double t, double tol, double *P, char *err,
         int errlen) {
  if (k > 256) return fail(err, errlen, O_EINVAL, "k too large");
  o_eig E;
  int rc = eig_of_q(k, Q, pi, &E, err, errlen);
  if (rc) return rc;
  rc = eig_pt(&E, t, tol, P, err, errlen);
  eig_free(&E);
  return rc;
}

int o_dpt(int k, const double *Q, const double *P, double *dP) {
  for (int i = 0; i < k; i++)
    for (int j = 0; j < k; j++) {
      double s = 0.;
      for (int m = 0; m < k; m++) s += Q[i * k + m] * P[m * k + j];
      dP[i * k + j] = s;
    }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* PhyloLik (PhyloLik.ml:9-174)                                             */

typedef struct {
  int k, n, nl;
  const int *lc, *rc;
  int *parent, *sib;
  double *alpha; /* (n) x k : leaf vectors in rows 0..nl-1, internal alphas after */
  double *beta;  /* n x k */
  double *inter;
} o_ws;

static void ws_init(o_ws *W, int k, int n, const int *lc, const int *rc) {
  W->k = k;
  W->n = n;
  W->nl = (n + 1) / 2;
  W->lc = lc;
  W->rc = rc;
  W->parent = (int *)malloc(sizeof(int) * (size_t)n);
  W->sib = (int *)malloc(sizeof(int) * (size_t)n);
  for (int i = 0; i < n; i++) W->parent[i] = W->sib[i] = -1;
  for (int i = W->nl; i < n; i++) {
    W->parent[lc[i]] = i;
    W->parent[rc[i]] = i;
    W->sib[lc[i]] = rc[i];
    W->sib[rc[i]] = lc[i];
  }
  W->alpha = (double *)calloc((size_t)(n * k), sizeof(double));
  W->beta = (double *)calloc((size_t)(n * k), sizeof(double));
  W->inter = (double *)calloc((size_t)k, sizeof(double));
}

static void ws_free(o_ws *W) {
  free(W->parent);
  free(W->sib);
  free(W->alpha);
  free(W->beta);
  free(W->inter);
}

/* get_leaf (PhyloLik.ml:11-19): `Certain i -> e_i, `Marginalize -> ones */
static void set_leaves(o_ws *W, const int *leaves) {
  int k = W->k;
  for (int j = 0; j < W->nl; j++) {
    double *a = W->alpha + (size_t)j * k;
    int c = leaves[j];
    for (int x = 0; x < k; x++) a[x] = (c < 0) ? 1. : (x == c ? 1. : 0.);
  }
}

/* ensure_alpha (PhyloLik.ml:73-93); returns z */
static double inside(o_ws *W, const double *P, const double *prior) {
  int k = W->k, n = W->n;
  for (int i = W->nl; i < n; i++) {
    int l = W->lc[i], r = W->rc[i];
    const double *ls = P + (size_t)l * k * k, *rs = P + (size_t)r * k * k;
    const double *al = W->alpha + (size_t)l * k, *ar = W->alpha + (size_t)r * k;
    double *ai = W->alpha + (size_t)i * k;
    for (int a = 0; a < k; a++)
      ai[a] = ddot(k, ls + (size_t)a * k, 1, al, 1) * ddot(k, rs + (size_t)a * k, 1, ar, 1);
  }
  return ddot(k, W->alpha + (size_t)(n - 1) * k, 1, prior, 1);
}

/* ensure_beta (PhyloLik.ml:96-120) */
static void outside(o_ws *W, const double *P, const double *prior) {
  int k = W->k, n = W->n;
  double *broot = W->beta + (size_t)(n - 1) * k;
  for (int a = 0; a < k; a++) broot[a] = prior[a];
  for (int i = n - 2; i >= 0; i--) {
    int p = W->parent[i], s = W->sib[i];
    const double *ps = P + (size_t)i * k * k, *ss = P + (size_t)s * k * k;
    const double *bp = W->beta + (size_t)p * k, *xas = W->alpha + (size_t)s * k;
    for (int a = 0; a < k; a++) W->inter[a] = bp[a] * ddot(k, ss + (size_t)a * k, 1, xas, 1);
    double *bi = W->beta + (size_t)i * k;
    for (int b = 0; b < k; b++) bi[b] = ddot(k, W->inter, 1, ps + b, k);
  }
}

/* d ln z / d t_br = sum_ab xi_br[a,b] * dP_br[a,b] / P_br[a,b] with xi the
 * branch posterior (PhyloLik.ml:140-174 feeding PhyloEM.ml:149-161) */
static double dlogz_dt_branch(const o_ws *W, const double *P, const double *dP, double z, int br) {
  int k = W->k;
  if (!(z > 0.)) return 0.;
  int p = W->parent[br], sib = W->sib[br];
  const double *sm = P + (size_t)br * k * k, *sms = P + (size_t)sib * k * k;
  const double *dsm = dP + (size_t)br * k * k;
  const double *bp = W->beta + (size_t)p * k, *ab = W->alpha + (size_t)br * k;
  const double *xas = W->alpha + (size_t)sib * k;
  double tot = 0.;
  for (int a = 0; a < k; a++) {
    double bpa = bp[a];
    if (!(bpa > 0.)) continue;
    double bpa_sibtot = bpa * ddot(k, sms + (size_t)a * k, 1, xas, 1);
    for (int b = 0; b < k; b++) {
      double pr = bpa_sibtot * sm[(size_t)a * k + b] * ab[b] / z;
      if (pr > 0.) tot += pr * dsm[(size_t)a * k + b] / sm[(size_t)a * k + b];
    }
  }
  return tot;
}

int o_prune(int k, int n, const int *lc, const int *rc, const double *P, const double *prior,
            const int *leaves, int64_t ncols, double *z, double *post, const double *dP,
            double *dlogz_dt) {
  o_ws W;
  ws_init(&W, k, n, lc, rc);
  int nl = W.nl;
  for (int64_t c = 0; c < ncols; c++) {
    set_leaves(&W, leaves + c * nl);
    double zc = inside(&W, P, prior);
    z[c] = zc;
    int need_beta = (post != NULL) || (dlogz_dt != NULL);
    if (need_beta) outside(&W, P, prior);
    if (post) {
      /* node_posterior (PhyloLik.ml:127-138) */
      double *pc = post + (size_t)c * n * k;
      for (int i = 0; i < n; i++) {
        double *pi_ = pc + (size_t)i * k;
        if (zc == 0.) {
          for (int x = 0; x < k; x++) pi_[x] = 0.;
        } else if (i < nl) {
          for (int x = 0; x < k; x++) pi_[x] = W.alpha[(size_t)i * k + x];
        } else {
          for (int x = 0; x < k; x++)
            pi_[x] = W.alpha[(size_t)i * k + x] * W.beta[(size_t)i * k + x] / zc;
        }
      }
    }
    if (dlogz_dt && dP) {
      for (int br = 0; br < n - 1; br++)
        dlogz_dt[c * (n - 1) + br] = dlogz_dt_branch(&W, P, dP, zc, br);
    }
  }
  ws_free(&W);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* PhyloCSFModel (src/PhyloCSFModel.ml)                                     */

struct o_model {
  int n, nl, nm;
  int *lc, *rc;
  double *branch;
  double *Q;  /* nm x 64 x 64 */
  double *pi; /* nm x 64 */
  o_eig *eig;
};

o_model *o_model_new(int n, const int *lc, const int *rc, const double *branch, int nm,
                     const double *S, const double *pi, char *err, int errlen) {
  o_model *m = (o_model *)calloc(1, sizeof(o_model));
  m->n = n;
  m->nl = (n + 1) / 2;
  m->nm = nm;
  m->lc = (int *)malloc(sizeof(int) * (size_t)n);
  m->rc = (int *)malloc(sizeof(int) * (size_t)n);
  m->branch = (double *)malloc(sizeof(double) * (size_t)n);
  memcpy(m->lc, lc, sizeof(int) * (size_t)n);
  memcpy(m->rc, rc, sizeof(int) * (size_t)n);
  memcpy(m->branch, branch, sizeof(double) * (size_t)n);
  m->Q = (double *)malloc(sizeof(double) * 64 * 64 * (size_t)nm);
  m->pi = (double *)malloc(sizeof(double) * 64 * (size_t)nm);
  m->eig = (o_eig *)calloc((size_t)nm, sizeof(o_eig));
  memcpy(m->pi, pi, sizeof(double) * 64 * (size_t)nm);
  for (int q = 0; q < nm; q++) {
    if (o_build_q(64, S + (size_t)q * 4096, pi + (size_t)q * 64, m->Q + (size_t)q * 4096, err,
                  errlen) ||
        eig_of_q(64, m->Q + (size_t)q * 4096, pi + (size_t)q * 64, &m->eig[q], err, errlen)) {
      o_model_free(m);
      return NULL;
    }
  }
  return m;
}

void o_model_free(o_model *m) {
  if (!m) return;
  for (int q = 0; q < m->nm; q++)
    if (m->eig[q].w) eig_free(&m->eig[q]);
  free(m->eig);
  free(m->lc);
  free(m->rc);
  free(m->branch);
  free(m->Q);
  free(m->pi);
  free(m);
}

/* PhyloModel.make via P14n.update (PhyloModel.ml:12-25,86-92,132-149):
 * branch br gets t = rho * b_br; the root branch is unused. */
static int model_pms(o_model *m, int q, double rho, double *P, char *err, int errlen) {
  if (!(rho > 0.))
    return fail(err, errlen, O_EINVAL,
                "CamlPaml.P14n.instantiate_tree: domain violation on variable 0");
  for (int br = 0; br < m->n - 1; br++) {
    double t = rho * m->branch[br];
    if (t < 0.) return fail(err, errlen, O_EINVAL, "CamlPaml.PhyloModel.make");
    int rc = eig_pt(&m->eig[q], t, 1e-6, P + (size_t)br * 4096, err, errlen);
    if (rc) return rc;
  }
  return 0;
}

static int lpr_impl(o_model *m, int q, double rho, const uint8_t *codes, int64_t ncols,
                    double *lpr_out, double *elpr_out, double *dlpr_drho, double *dlpr_dt,
                    char *err, int errlen) {
  if (q < 0 || q >= m->nm) return fail(err, errlen, O_EINVAL, "model index");
  int n = m->n, nl = m->nl;
  double *P = (double *)malloc(sizeof(double) * 4096 * (size_t)(n - 1));
  double *dP = NULL;
  int rc = model_pms(m, q, rho, P, err, errlen);
  if (rc) {
    free(P);
    return rc;
  }
  const double *prior = m->pi + (size_t)q * 64;
  if (dlpr_drho || dlpr_dt) {
    dP = (double *)malloc(sizeof(double) * 4096 * (size_t)(n - 1));
    for (int br = 0; br < n - 1; br++)
      o_dpt(64, m->Q + (size_t)q * 4096, P + (size_t)br * 4096, dP + (size_t)br * 4096);
  }
  double anc_lprior[64];
  for (int x = 0; x < 64; x++) anc_lprior[x] = log(prior[x]);
  o_ws W;
  ws_init(&W, 64, n, m->lc, m->rc);
  int leaves[1024];
  double lpr = 0., elpr = 0.;
  double *gt = NULL;
  if (dP) gt = (double *)calloc((size_t)(n - 1), sizeof(double));
  for (int64_t c = 0; c < ncols; c++) {
    for (int j = 0; j < nl; j++) {
      uint8_t v = codes[c * nl + j];
      leaves[j] = (v == 255) ? -1 : (int)v;
    }
    set_leaves(&W, leaves);
    double z = inside(&W, P, prior);
    lpr = lpr + log(z);
    /* node_posterior root, then dot pr_root anc_lprior (PhyloCSFModel.ml:80-81) */
    double tot = 0.;
    if (z != 0.) {
      const double *ar = W.alpha + (size_t)(n - 1) * 64;
      for (int x = 0; x < 64; x++) tot = tot + (ar[x] * prior[x] / z) * anc_lprior[x];
    }
    elpr = elpr + tot;
    if (dP) {
      outside(&W, P, prior);
      for (int br = 0; br < n - 1; br++) gt[br] += dlogz_dt_branch(&W, P, dP, z, br);
    }
  }
  *lpr_out = lpr;
  if (elpr_out) *elpr_out = elpr;
  if (dP) {
    double d = 0.;
    for (int br = 0; br < n - 1; br++) {
      d += gt[br] * m->branch[br]; /* dt/drho = b_br (PhyloEM.ml:162-182) */
      if (dlpr_dt) dlpr_dt[br] = gt[br];
    }
    if (dlpr_drho) *dlpr_drho = d;
  }
  free(gt);
  ws_free(&W);
  free(P);
  free(dP);
  return 0;
}

int o_lpr(o_model *m, int model, double rho, const uint8_t *codes, int64_t ncols, double *lpr,
          double *elpr, char *err, int errlen) {
  return lpr_impl(m, model, rho, codes, ncols, lpr, elpr, NULL, NULL, err, errlen);
}

int o_lpr_grad(o_model *m, int model, double rho, const uint8_t *codes, int64_t ncols,
               double *lpr, double *dlpr_drho, double *dlpr_dt, char *err, int errlen) {
  return lpr_impl(m, model, rho, codes, ncols, lpr, NULL, dlpr_drho, dlpr_dt, err, errlen);
}

static double db(double x) { return 10. * x / log(10.); } /* PhyloCSFModel.ml:113 */

int o_score_fixed(o_model *m, const uint8_t *codes, int64_t ncols, double *out, char *err,
                  int errlen) {
  double lc, ec, ln, en;
  int rc = o_lpr(m, 0, 1.0, codes, ncols, &lc, &ec, err, errlen);
  if (rc) return rc;
  rc = o_lpr(m, 1, 1.0, codes, ncols, &ln, &en, err, errlen);
  if (rc) return rc;
  out[0] = db(lc - ln);
  out[1] = db(ec - en);
  out[2] = lc;
  out[3] = ln;
  return 0;
}

/* ---- mle (uses o_fit.c) ------------------------------------------------ */

typedef struct {
  o_model *m;
  int q;
  const uint8_t *codes;
  int64_t ncols;
  int evals;
  int failed;
  char msg[256];
} lpr_ctx;

static double lpr_cb(double x, void *vctx) {
  lpr_ctx *c = (lpr_ctx *)vctx;
  double l, e;
  c->evals++;
  int rc = o_lpr(c->m, c->q, x, c->codes, c->ncols, &l, &e, c->msg, sizeof c->msg);
  if (rc) {
    c->failed = 1;
    return NAN;
  }
  return l;
}

/* maximize_lpr (PhyloCSFModel.ml:84-99) for one model */
static int maximize_lpr(o_model *m, int q, const uint8_t *codes, int64_t ncols, double *rho_out,
                        double *lpr_out, double *elpr_out, int *evals, char *err, int errlen) {
  const double lo = 1e-2, hi = 10., init = 1., accuracy = 0.01;
  lpr_ctx c = {m, q, codes, ncols, 0, 0, {0}};
  int fe = 0;
  double good = o_find_init(lpr_cb, &c, 250, 1, init, lo, hi, &fe);
  if (c.failed) return fail(err, errlen, O_EFAIL, "%s", c.msg);
  double x = good;
  if (lo < good && good < hi) {
    int be = 0;
    int rc = o_brent_max(lpr_cb, &c, lo, hi, accuracy, &x, &be, err, errlen);
    if (c.failed) return fail(err, errlen, O_EFAIL, "%s", c.msg);
    if (rc) return rc;
  }
  int rc = o_lpr(m, q, x, codes, ncols, lpr_out, elpr_out, err, errlen);
  c.evals++;
  *rho_out = x;
  *evals += c.evals;
  return rc;
}

int o_score_mle(o_model *m, const uint8_t *codes, int64_t ncols, double *out, char *err,
                int errlen) {
  double rc_, rn, lc, ln, ec, en;
  int evals = 0;
  int rc = maximize_lpr(m, 0, codes, ncols, &rc_, &lc, &ec, &evals, err, errlen);
  if (rc) return rc;
  rc = maximize_lpr(m, 1, codes, ncols, &rn, &ln, &en, &evals, err, errlen);
  if (rc) return rc;
  out[0] = db(lc - ln);
  out[1] = db(ec - en);
  out[2] = lc;
  out[3] = ln;
  out[4] = rc_;
  out[5] = rn;
  out[6] = (double)evals;
  return 0;
}
