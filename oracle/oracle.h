/*
 * oracle.h -- FP64 CPU restatement of the PhyloCSF scoring path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: the tests, the
 * smoke entry and the bench's cpu_baseline leg load it to compare against
 * the HIP product path.  No product code links, imports or calls it.
 *
 * Every function cites the reference file:line it restates
 * (paths relative to the mlin/PhyloCSF source tree).
 *
 * Conventions: all functions return 0 on success, negative on error with a
 * message in err[errlen] (the reference raises Invalid_argument / Failure at
 * the same places).  Trees use the reference's T numbering: leaves 0..nl-1
 * left to right, internal nodes in post-order, root n-1
 * (lib/CamlPaml/T.ml:57-112).
 */
#ifndef PHYLOCSF_ORACLE_H
#define PHYLOCSF_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define O_EINVAL (-1) /* Invalid_argument in the reference */
#define O_EFAIL (-2)  /* Failure in the reference */

/* ---- parsing (o_parse.c) ---------------------------------------------- */

/* ECM.import_parameters (src/ECM.ml:18-72): S[64*64] symmetric, pi[64]. */
int o_ecm_read(const char *path, double *S, double *pi, char *err, int errlen);

/* NewickParser/NewickLexer + optional Newick.subtree (keep_csv != NULL,
 * comma separated species) + T.of_newick.  Arrays sized max_n.
 * labels_buf receives the n labels separated by '\n'. */
int o_newick_tree(const char *text, const char *keep_csv, int max_n, int *n_out,
                  int *parent, int *lc, int *rc, double *branch, char *labels_buf,
                  int labels_len, char *err, int errlen);

/* Newick.total_length (lib/CamlPaml/Newick.ml:52-61) of the (optionally
 * subtree-pruned by keep_csv) tree; count_root=0 like the CLI's bls. */
int o_newick_total_length(const char *text, const char *keep_csv, double *out,
                          char *err, int errlen);

/* ---- models and likelihood (o_model.c) -------------------------------- */

/* PhyloCSFModel.make_q_p14n + qscale + P14n.instantiate_q
 * (src/PhyloCSFModel.ml:11-30, lib/CamlPaml/PhyloModel.ml:94-106). */
int o_build_q(int k, const double *S, const double *pi, double *Q, char *err, int errlen);

/* Q.Diag.real_to_Pt (lib/CamlPaml/Q.ml:211-249) for a reversible Q with
 * stationary pi, eigendecomposed through the symmetric similarity
 * transform.  tol as Q.Diag's (1e-6). */
int o_pt(int k, const double *Q, const double *pi, double t, double tol, double *P,
         char *err, int errlen);

/* dP/dt = Q P(t) (lib/CamlPaml/Q.ml:260-275 states it for the complex path). */
int o_dpt(int k, const double *Q, const double *P, double *dP);

/* PhyloLik.prepare + ensure_alpha + ensure_beta + node_posterior
 * (lib/CamlPaml/PhyloLik.ml:46-138) for ncols columns.
 * leaves[c*nl + j] = state index or -1 (`Marginalize).
 * z[ncols]; post (optional) [ncols][n][k] node posteriors.
 * dlogz_dt (optional, with dP) [ncols][n-1]: d ln z / d t_br through the
 * branch posteriors (PhyloLik.ml:140-174, PhyloEM.ml:140-161). */
int o_prune(int k, int n, const int *lc, const int *rc, const double *P,
            const double *prior, const int *leaves, int64_t ncols, double *z,
            double *post, const double *dP, double *dlogz_dt);

/* PhyloLik with `Distribution leaves (-2, rows of leaf_dist), node
 * posteriors of every node and add_branch_posteriors summed over columns with
 * weights (PhyloLik.ml:9-174, PhyloEM.ml:14-20).  post [ncols][n][k] and
 * ecounts [(n-1)][k][k] (added to) are optional. */
int o_lik_posteriors(int k, int n, const int *lc, const int *rc, const double *P,
                     const double *prior, const int *leaves, const double *leaf_dist,
                     int64_t ncols, const double *weights, double *z, double *post,
                     double *ecounts);

typedef struct o_model o_model;

/* PhyloCSFModel.make (src/PhyloCSFModel.ml:107-110) generalised to nm
 * ECMs sharing one tree: S[nm][64*64], pi[nm][64]. */
o_model *o_model_new(int n, const int *lc, const int *rc, const double *branch, int nm,
                     const double *S, const double *pi, char *err, int errlen);
void o_model_free(o_model *m);

/* PhyloCSFModel.lpr_leaves (src/PhyloCSFModel.ml:67-82).
 * codes[c*nl + j]: codon 0..63 or 255 for `Marginalize. */
int o_lpr(o_model *m, int model, double rho, const uint8_t *codes, int64_t ncols,
          double *lpr, double *elpr, char *err, int errlen);

/* Same plus d lpr / d rho and d lpr / d t_br (branch lengths rho*b). */
int o_lpr_grad(o_model *m, int model, double rho, const uint8_t *codes, int64_t ncols,
               double *lpr, double *dlpr_drho, double *dlpr_dt, char *err, int errlen);

/* llr_FixedLik 1.0 (src/PhyloCSFModel.ml:122-128): models 0 = coding,
 * 1 = noncoding.  out[4] = score(dB), anc_comp(dB), lpr_c, lpr_n. */
int o_score_fixed(o_model *m, const uint8_t *codes, int64_t ncols, double *out,
                  char *err, int errlen);

/* llr_MaxLik ~init:1 (src/PhyloCSFModel.ml:84-99,130-136) with the GSL
 * Brent minimiser and Fit.find_init (lib/CamlPaml/Fit.ml:3-48).
 * out[7] = score, anc_comp, lpr_c, lpr_n, rho_c, rho_n, n_evals. */
int o_score_mle(o_model *m, const uint8_t *codes, int64_t ncols, double *out,
                char *err, int errlen);

/* ---- OmegaModel (src/OmegaModel.ml), in o_model.c ------------------------ */

/* pi_exprs (OmegaModel.ml:24-40) for settings qs[12] (kappa, omega, sigma,
 * then A, C, G relative to T at codon positions 1, 2, 3): pi[64], not
 * normalised. */
void o_omega_pi(const double *qs, double *pi);
/* exchangeabilities of q_p14n (OmegaModel.ml:42-74): S[64*64] */
void o_omega_s(const double *qs, double *S);
/* update_f3x4 (OmegaModel.ml:102-134): qs9 = settings 3..11 */
void o_omega_f3x4(const uint8_t *codes, int64_t ncols, int nl, double *qs9);
/* the priors of kr_map (OmegaModel.ml:147-157) */
double o_half_cauchy_lpdf(double mode, double scale, double x);
double o_lpr_kappa(double k);
/* OmegaModel.lpr_leaves (OmegaModel.ml:137-144) */
int o_omega_lpr(int n, const int *lc, const int *rc, const double *branch, const double *qs,
                double rho, const uint8_t *codes, int64_t ncols, double *lpr, char *err,
                int errlen);
/* OmegaModel.score (OmegaModel.ml:195-219): out[8] = score, lpr_H0, lpr_H1,
 * rho_H0, kappa_H0, rho_H1, kappa_H1, evaluations */
int o_omega_score(int n, const int *lc, const int *rc, const double *branch,
                  const uint8_t *codes, int64_t ncols, double omega_H1, double sigma_H1,
                  double *out, char *err, int errlen);

/* ---- fit helpers (o_fit.c) --------------------------------------------- */

/* OCaml 4.x Random: init(seed) then n draws of Random.float bound. */
void o_ocaml_random_floats(int seed, double bound, int n, double *out);

/* GSL Brent minimiser on a callback (for unit tests of the restatement). */
typedef double (*o_fn)(double x, void *ctx);
int o_brent_max(o_fn f, void *ctx, double lo, double hi, double accuracy, double *x_out,
                int *evals_out, char *err, int errlen);
/* Fit.find_init with logspace (lib/CamlPaml/Fit.ml:27-48). */
double o_find_init(o_fn f, void *ctx, int maxtries, int logspace, double init, double lo,
                   double hi, int *evals);
/* MD5 digest (RFC 1321) used by the OCaml Random seeding. */
void o_md5(const uint8_t *msg, uint64_t len, uint8_t out[16]);

#ifdef __cplusplus
}
#endif
#endif
