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

/* PhyloLik.prepare with any leaf kind, likelihood, node_posterior of every
 * node and add_branch_posteriors of every branch summed over the columns with
 * weights -- PhyloEM.collect_sufficient_statistics (lib/CamlPaml/PhyloEM.ml:
 * 14-20).  leaves[c*nl + j]: state index, -1 `Marginalize, -2 `Distribution
 * (leaf_dist[(c*nl + j)*k + x]).  weights (NULL: 1).  z[ncols]; post
 * (optional) [ncols][n][k]; ecounts (optional) [(n-1)][k][k], added to. */
int o_lik_posteriors(int k, int n, const int *lc, const int *rc, const double *P,
                     const double *prior, const int *leaves, const double *leaf_dist,
                     int64_t ncols, const double *weights, double *z, double *post,
                     double *ecounts) {
  o_ws W;
  ws_init(&W, k, n, lc, rc);
  const int nl = W.nl;
  for (int64_t c = 0; c < ncols; c++) {
    /* get_leaf (PhyloLik.ml:16-19) */
    for (int j = 0; j < nl; j++) {
      double *a = W.alpha + (size_t)j * k;
      const int v = leaves[c * nl + j];
      for (int x = 0; x < k; x++)
        a[x] = (v == -2) ? leaf_dist[((size_t)c * nl + j) * k + x]
               : (v < 0) ? 1. : (x == v ? 1. : 0.);
    }
    const double zc = inside(&W, P, prior);
    z[c] = zc;
    if (!post && !ecounts) continue;
    outside(&W, P, prior);
    if (post) {
      double *pc = post + (size_t)c * n * k;
      for (int i = 0; i < n; i++)
        for (int x = 0; x < k; x++)
          pc[(size_t)i * k + x] = (zc == 0.) ? 0.
                                  : (i < nl) ? W.alpha[(size_t)i * k + x]
                                             : W.alpha[(size_t)i * k + x] *
                                                   W.beta[(size_t)i * k + x] / zc;
    }
    if (ecounts && zc > 0.) {
      /* add_branch_posteriors ?weight (PhyloLik.ml:140-174) for every branch */
      const double w = weights ? weights[c] : 1.0;
      for (int br = 0; br < n - 1; br++) {
        const int p = W.parent[br], sib = W.sib[br];
        const double *sm = P + (size_t)br * k * k, *sms = P + (size_t)sib * k * k;
        const double *bp = W.beta + (size_t)p * k, *ab = W.alpha + (size_t)br * k;
        const double *xas = W.alpha + (size_t)sib * k;
        double *ec = ecounts + (size_t)br * k * k;
        for (int a = 0; a < k; a++) {
          const double bpa = bp[a];
          if (!(bpa > 0.)) continue;
          const double bpa_sibtot = bpa * ddot(k, sms + (size_t)a * k, 1, xas, 1);
          for (int b = 0; b < k; b++) {
            const double pr = bpa_sibtot * sm[(size_t)a * k + b] * ab[b] / zc;
            ec[(size_t)a * k + b] = ec[(size_t)a * k + b] + w * pr;
          }
        }
      }
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

/* ------------------------------------------------------------------------ */
/* OmegaModel (src/OmegaModel.ml): the omega (dN/dS) test                   */

/* Code.Codon64.translation_table (lib/CamlPaml/Code.ml:153-173), codon index
 * 16 n1 + 4 n2 + n3 with A, C, G, T = 0..3; stops TAA, TAG, TGA. */
static const char kAA[65] = "KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVV*Y*YSSSS*CWCLFLF";
static int is_stop_index(int c) { return c == 48 || c == 50 || c == 56; }

/* pi_exprs (OmegaModel.ml:24-40): settings qs[0..11] = kappa, omega, sigma,
 * then per codon position the frequencies of A, C, G relative to T.  The
 * values are not normalised (the stop-codon term only divides them all). */
void o_omega_pi(const double *qs, double *pi) {
  double sc[64];
  const double d1 = qs[3] + (qs[4] + (qs[5] + 1.0));
  const double d2 = qs[6] + (qs[7] + (qs[8] + 1.0));
  const double d3 = qs[9] + (qs[10] + (qs[11] + 1.0));
  for (int c = 0; c < 64; c++) {
    const int i1 = c / 16, i2 = (c / 4) % 4, i3 = c % 4;
    const double f1 = (i1 == 3 ? 1.0 : qs[3 + i1]) / d1;
    const double f2 = (i2 == 3 ? 1.0 : qs[6 + i2]) / d2;
    const double f3 = (i3 == 3 ? 1.0 : qs[9 + i3]) / d3;
    sc[c] = f1 * (f2 * f3);
  }
  const double denom = 1.0 - (1.0 - qs[2]) * (sc[48] + (sc[50] + sc[56]));
  for (int c = 0; c < 64; c++) pi[c] = sc[c] / denom;
}

/* q_p14n (OmegaModel.ml:42-74) as exchangeabilities: q_ij = pi_j s_ij with
 * s_ij = kappa_part * omega_part for codons one nucleotide apart, else 0. */
void o_omega_s(const double *qs, double *S) {
  for (int i = 0; i < 64; i++)
    for (int j = 0; j < 64; j++) {
      const int a[3] = {i / 16, (i / 4) % 4, i % 4}, b[3] = {j / 16, (j / 4) % 4, j % 4};
      int nd = 0, p = -1;
      for (int k = 0; k < 3; k++)
        if (a[k] != b[k]) {
          nd++;
          p = k;
        }
      double s = 0.0;
      if (nd == 1) {
        const int x = a[p], y = b[p];
        /* A<->G, C<->T (A, C, G, T = 0..3) */
        const int transition = (x == 0 && y == 2) || (x == 2 && y == 0) || (x == 1 && y == 3) ||
                               (x == 3 && y == 1);
        const double kp = transition ? qs[0] : 1.0;
        const double op =
            (!is_stop_index(i) && !is_stop_index(j) && kAA[i] != kAA[j]) ? qs[1] : 1.0;
        s = kp * op;
      }
      S[i * 64 + j] = s;
    }
}

/* update_f3x4 (OmegaModel.ml:102-134): counts from 1 per codon position and
 * nucleotide over every `Certain leaf; qs9[3 p + n] = count[p][n] / count[p][T]. */
void o_omega_f3x4(const uint8_t *codes, int64_t ncols, int nl, double *qs9) {
  long cnt[3][4];
  for (int p = 0; p < 3; p++)
    for (int n = 0; n < 4; n++) cnt[p][n] = 1;
  for (int64_t c = 0; c < ncols; c++)
    for (int j = 0; j < nl; j++) {
      const int v = codes[c * nl + j];
      if (v == 255) continue;
      cnt[0][v / 16]++;
      cnt[1][(v / 4) % 4]++;
      cnt[2][v % 4]++;
    }
  for (int p = 0; p < 3; p++)
    for (int n = 0; n < 3; n++) qs9[3 * p + n] = (double)cnt[p][n] / (double)cnt[p][3];
}

/* half_cauchy_lpdf ~mode ~scale (OmegaModel.ml:147-154) */
double o_half_cauchy_lpdf(double mode, double scale, double x) {
  const double pi = acos(-1.0);
  const double numer = 1.0 / (pi * scale * (1.0 + pow((x - mode) / scale, 2.0)));
  const double denom = 1.0 - (atan((0.0 - mode) / scale) / pi + 0.5);
  return log(numer) - log(denom);
}

/* lpr_kappa k = log (Gsl.Randist.gamma_pdf ~a:7 ~b:0.25 (k - 1 + epsilon_float))
 * (OmegaModel.ml:157), with GSL's gamma density gsl_ran_gamma_pdf:
 * exp((a - 1) log(x / b) - x / b - lngamma(a)) / b for x > 0, a != 1. */
double o_lpr_kappa(double k) {
  const double a = 7.0, b = 0.25, x = k - 1.0 + 2.220446049250313e-16;
  if (x < 0) return -INFINITY;
  if (x == 0) return -INFINITY;
  const double p = exp((a - 1.0) * log(x / b) - x / b - lgamma(a)) / b;
  return log(p);
}

/* One omega-model instance over a fixed region: Q and its eigensystem for
 * the current settings (rebuilt when they change) and the tree. */
typedef struct {
  int n, nl;
  const int *lc, *rc;
  const double *branch;
  const uint8_t *codes;
  int64_t ncols;
  double qs[12];
  int have_eig;
  double Q[4096], pi[64], prior[64];
  o_eig E;
  int evals;
  int failed;
  char msg[256];
} omega_inst;

static int omega_set_q(omega_inst *I, const double *qs) {
  if (I->have_eig && memcmp(qs, I->qs, sizeof I->qs) == 0) return 0;
  double S[4096];
  memcpy(I->qs, qs, sizeof I->qs);
  o_omega_s(qs, S);
  o_omega_pi(qs, I->pi);
  if (I->have_eig) eig_free(&I->E);
  I->have_eig = 0;
  int rc = o_build_q(64, S, I->pi, I->Q, I->msg, sizeof I->msg);
  if (rc) return rc;
  rc = eig_of_q(64, I->Q, I->pi, &I->E, I->msg, sizeof I->msg);
  if (rc) return rc;
  I->have_eig = 1;
  /* the root prior: Q.Diag.equilibrium, the stationary distribution, which
   * for this reversible Q is pi normalised */
  double tot = 0.0;
  for (int c = 0; c < 64; c++) tot += I->pi[c];
  for (int c = 0; c < 64; c++) I->prior[c] = I->pi[c] / tot;
  return 0;
}

/* OmegaModel.lpr_leaves (OmegaModel.ml:137-144) at tree scale rho */
static int omega_lpr(omega_inst *I, double rho, double *lpr_out) {
  const int n = I->n, nl = I->nl;
  double *P = (double *)malloc(sizeof(double) * 4096 * (size_t)(n - 1));
  for (int br = 0; br < n - 1; br++) {
    int rc = eig_pt(&I->E, rho * I->branch[br], 1e-6, P + (size_t)br * 4096, I->msg,
                    sizeof I->msg);
    if (rc) {
      free(P);
      return rc;
    }
  }
  o_ws W;
  ws_init(&W, 64, n, I->lc, I->rc);
  int leaves[1024];
  double lpr = 0.;
  for (int64_t c = 0; c < I->ncols; c++) {
    for (int j = 0; j < nl; j++) {
      const uint8_t v = I->codes[c * nl + j];
      leaves[j] = (v == 255) ? -1 : (int)v;
    }
    set_leaves(&W, leaves);
    lpr = lpr + log(inside(&W, P, I->prior));
  }
  ws_free(&W);
  free(P);
  *lpr_out = lpr;
  return 0;
}

int o_omega_lpr(int n, const int *lc, const int *rc, const double *branch, const double *qs,
                double rho, const uint8_t *codes, int64_t ncols, double *lpr, char *err,
                int errlen) {
  omega_inst I;
  memset(&I, 0, sizeof I);
  I.n = n;
  I.nl = (n + 1) / 2;
  I.lc = lc;
  I.rc = rc;
  I.branch = branch;
  I.codes = codes;
  I.ncols = ncols;
  int r = omega_set_q(&I, qs);
  if (!r) r = omega_lpr(&I, rho, lpr);
  if (r) fail(err, errlen, r, "%s", I.msg);
  if (I.have_eig) eig_free(&I.E);
  return r;
}

/* f_rho / f_kappa of kr_map (OmegaModel.ml:160-170) */
typedef struct {
  omega_inst *I;
  int which; /* 0: rho varies (kappa fixed), 1: kappa varies (rho fixed) */
  double rho, kappa;
} omega_fn;

static double omega_cb(double x, void *vctx) {
  omega_fn *F = (omega_fn *)vctx;
  omega_inst *I = F->I;
  I->evals++;
  double qs[12];
  memcpy(qs, I->qs, sizeof qs);
  double rho = F->rho, prior_term;
  if (F->which == 0) {
    rho = x;
    prior_term = o_half_cauchy_lpdf(1.0, 0.5, x);
  } else {
    qs[0] = x;
    prior_term = o_lpr_kappa(x);
  }
  double l = NAN;
  if (omega_set_q(I, qs) || omega_lpr(I, rho, &l)) {
    I->failed = 1;
    return NAN;
  }
  return prior_term + l;
}

/* PhyloCSFModel.maximize_lpr ~init ~lo ~hi ~accuracy:0.01 (PhyloCSFModel.ml:84-99) */
static int omega_maximize(omega_fn *F, double init, double lo, double hi, double *x_out,
                          double *f_out, char *err, int errlen) {
  int fe = 0;
  const double good = o_find_init(omega_cb, F, 250, 1, init, lo, hi, &fe);
  if (F->I->failed) return fail(err, errlen, O_EFAIL, "%s", F->I->msg);
  double x = good;
  if (lo < good && good < hi) {
    int be = 0;
    const int rc = o_brent_max(omega_cb, F, lo, hi, 0.01, &x, &be, err, errlen);
    if (F->I->failed) return fail(err, errlen, O_EFAIL, "%s", F->I->msg);
    if (rc) return rc;
  }
  *f_out = omega_cb(x, F);
  if (F->I->failed) return fail(err, errlen, O_EFAIL, "%s", F->I->msg);
  *x_out = x;
  return 0;
}

/* kr_map (OmegaModel.ml:160-190): three rounds of coordinate ascent, rho in
 * [0.001, 10] then kappa in [1, 10]; the lpr of the last kappa step
 * (lpr_kappa + lpr_leaves). */
static int omega_kr_map(omega_inst *I, double *rho, double *kappa, double *lpr, char *err,
                        int errlen) {
  for (int round = 0; round < 3; round++) {
    omega_fn F = {I, 0, *rho, *kappa};
    double x, f;
    int rc = omega_maximize(&F, *rho, 0.001, 10.0, &x, &f, err, errlen);
    if (rc) return rc;
    *rho = x;
    omega_fn G = {I, 1, *rho, *kappa};
    rc = omega_maximize(&G, *kappa, 1.0, 10.0, &x, &f, err, errlen);
    if (rc) return rc;
    *kappa = x;
    *lpr = f;
    double qs[12];
    memcpy(qs, I->qs, sizeof qs);
    qs[0] = *kappa;
    if (omega_set_q(I, qs)) return fail(err, errlen, O_EFAIL, "%s", I->msg);
  }
  return 0;
}

/* OmegaModel.score (OmegaModel.ml:195-219).  out[8] = score (decibans),
 * lpr_H0, lpr_H1, rho_H0, kappa_H0, rho_H1, kappa_H1, evaluations. */
int o_omega_score(int n, const int *lc, const int *rc, const double *branch,
                  const uint8_t *codes, int64_t ncols, double omega_H1, double sigma_H1,
                  double *out, char *err, int errlen) {
  omega_inst I;
  memset(&I, 0, sizeof I);
  I.n = n;
  I.nl = (n + 1) / 2;
  I.lc = lc;
  I.rc = rc;
  I.branch = branch;
  I.codes = codes;
  I.ncols = ncols;
  double qs[12] = {2.5, 1.0, 1.0};
  o_omega_f3x4(codes, ncols, I.nl, qs + 3);
  int r = omega_set_q(&I, qs);
  if (r) {
    fail(err, errlen, r, "%s", I.msg);
    goto done;
  }
  double rho0 = 1.0, kappa0 = 2.5, l0 = 0;
  r = omega_kr_map(&I, &rho0, &kappa0, &l0, err, errlen);
  if (r) goto done;
  memcpy(qs, I.qs, sizeof qs);
  qs[1] = omega_H1;
  qs[2] = sigma_H1;
  r = omega_set_q(&I, qs);
  if (r) {
    fail(err, errlen, r, "%s", I.msg);
    goto done;
  }
  double rho1 = rho0, kappa1 = kappa0, l1 = 0;
  r = omega_kr_map(&I, &rho1, &kappa1, &l1, err, errlen);
  if (r) goto done;
  out[0] = 10. * (l1 - l0) / log(10.);
  out[1] = l0;
  out[2] = l1;
  out[3] = rho0;
  out[4] = kappa0;
  out[5] = rho1;
  out[6] = kappa1;
  out[7] = (double)I.evals;
done:
  if (I.have_eig) eig_free(&I.E);
  return r;
}
