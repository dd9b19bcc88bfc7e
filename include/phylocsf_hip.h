/*
 * phylocsf_hip.h -- C ABI of the MI355X-native PhyloCSF scoring path.
 *
 * The boundary sits where the reference's per-column loop sits: one batched
 * call replaces the column loop of PhyloCSFModel.lpr_leaves
 * (src/PhyloCSFModel.ml:67-82) and the PhyloLik calls it makes
 * (lib/CamlPaml/PhyloLik.ml:46-138); the P(t) builder replaces the to_Pt
 * loop of PhyloModel.make (lib/CamlPaml/PhyloModel.ml:12-25,
 * lib/CamlPaml/Q.ml:211-256).  Paths are relative to the mlin/PhyloCSF tree.
 *
 * Conventions
 *   - every call returns int: 0 = ok, or one of the PCSF_E* codes below,
 *     which mirror the reference's exceptions (Invalid_argument -> EINVAL,
 *     Failure -> EFAIL); pcsf_last_error(ctx) gives the message.
 *   - no C++ exceptions cross this boundary; no torch types appear in it.
 *   - one context per host thread; a context owns one HIP stream (default
 *     priority) and the device memory of its models and batches.
 *   - trees use the reference's T numbering (lib/CamlPaml/T.mli:1-40):
 *     leaves 0..nl-1 in left-to-right order, internal nodes in post-order,
 *     root n-1; children[(i-nl)*2 + {0,1}] = (left, right) of internal i.
 *   - leaf codes: codon index 16*n1+4*n2+n3 (lib/CamlPaml/Code.ml:140-141)
 *     or PCSF_MARGINALIZE (255) for `Marginalize.
 */
#ifndef PHYLOCSF_HIP_H
#define PHYLOCSF_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define PCSF_OK 0
#define PCSF_EINVAL (-1)  /* Invalid_argument in the reference */
#define PCSF_EFAIL (-2)   /* Failure in the reference (numerics, parse) */
#define PCSF_EDEVICE (-3) /* HIP runtime error */
#define PCSF_ENOMEM (-4)  /* device or host allocation failed */

#define PCSF_MARGINALIZE 255

/* arithmetic of the pruning contraction.  Every MFMA precision keeps live
 * partials in FP32 registers with per-column power-of-two rescaling, the
 * diagonal of P in FP32 VALU arithmetic, and FP64 log sums; they differ in
 * how P and the partial enter the matrix cores. */
#define PCSF_FP64 0  /* plain FP64 kernel: one column per workgroup (GEMV) */
#define PCSF_FP32 1  /* fused whole-tree kernel, FP32 MFMA (16x16x4) */
#define PCSF_FP16S 2 /* split FP16: P and partial as hi + lo FP16 pairs (scaled
                        by 2^14), hi*hi + hi*lo + lo*hi on FP16 MFMA (16x16x32),
                        FP32 accumulation */
#define PCSF_FP16 3  /* FP16 P and partial (scaled by 2^14), FP32 accumulation */
#define PCSF_BF16 4  /* BF16 P and partial, FP32 accumulation */
#define PCSF_FP64M 5 /* FP64 on matrix cores (v_mfma_f64_16x16x4f64): 64
                        columns per workgroup, messages in FP64 as PCSF_FP64
                        with another summation order; no tree-size limit.
                        Not an MFMA precision in the sense above (no
                        rescaling, no FP32 rounding): the maximisers treat it
                        as PCSF_FP64, and take their final evaluations with it */

typedef struct pcsf_ctx pcsf_ctx;
typedef struct pcsf_model pcsf_model;
typedef struct pcsf_batch pcsf_batch;

/* ---- context --------------------------------------------------------- */

int pcsf_ctx_create(int device, pcsf_ctx **out);
void pcsf_ctx_destroy(pcsf_ctx *ctx);
const char *pcsf_last_error(const pcsf_ctx *ctx);
/* 1: record HIP events around every kernel of pcsf_forward (for timing) */
int pcsf_ctx_set_timing(pcsf_ctx *ctx, int enable);
/* the HIP stream (hipStream_t) every call of this context is ordered on */
void *pcsf_ctx_stream(pcsf_ctx *ctx);
int pcsf_ctx_synchronize(pcsf_ctx *ctx);

/* ---- host-side parsing helpers (no device work) ----------------------- */

/* ECM.import_parameters (src/ECM.ml:18-72): S[64*64] row-major symmetric,
 * pi[64].  err may be NULL. */
int pcsf_ecm_read(const char *path, double *S, double *pi, char *err, int errlen);

/* NewickParser.parse (lib/CamlPaml/NewickParser.mly:7-24) + optional
 * Newick.subtree (lib/CamlPaml/Newick.ml:34-41, keep_csv = "sp1,sp2,...")
 * + T.of_newick (lib/CamlPaml/T.ml:57-112).  Arrays sized max_nodes;
 * children is (n-nl)*2; branch[n] (NaN where unspecified); labels are
 * written '\n'-separated into labels_buf.  *n_out = node count. */
int pcsf_newick_tree(const char *text, const char *keep_csv, int max_nodes, int *n_out,
                     int32_t *children, double *branch, char *labels_buf, int labels_len,
                     char *err, int errlen);

/* The op program the fused kernel runs for a tree (host only): a post-order
 * walk that keeps the fewest partial tiles live, four int32 words per op
 * (opcode | depth << 8 | prefetched << 16, leaf/branch args, prefetch leaf or
 * table index, 0).  with_tables = 1 gives the fixed-scale program in which
 * messages above cherries come from lookup tables.  No counterpart in the
 * reference (it is this implementation's schedule of PhyloLik.ml:78-90);
 * exposed so the schedule can be checked without a GPU. */
int pcsf_program(int n_nodes, const int32_t *children, int with_tables, int32_t *ops, int cap,
                 int *n_ops, int *slots);

/* ---- model: tree + rate models --------------------------------------- */

/* PhyloCSFModel.make generalised to n_models ECMs on one tree
 * (src/PhyloCSFModel.ml:11-56,107-110; lib/CamlPaml/PhyloModel.ml:94-106):
 * Q_m = (S_m o pi_m)/scale_m, root prior pi_m, branch lengths rho * b.
 * Eigendecomposes each Q_m (FP64, host) and uploads it.
 * S is n_models*64*64, pi is n_models*64, branch has n_nodes entries (the
 * root's is ignored).  Every S must be symmetric (a reversible model, as
 * ECM.import_parameters builds it); PCSF_EINVAL otherwise.
 * n_models = 0 (S, pi may be NULL) makes a tree-only model for
 * pcsf_score_omega, whose sets carry rate models of their own.
 * ctx may be NULL: a host-only model, for pcsf_simulate,
 * pcsf_simulate_nucleotides and pcsf_model_info (no device memory; the
 * device entry points refuse it with PCSF_EINVAL). */
int pcsf_model_create(pcsf_ctx *ctx, int n_nodes, const int32_t *children,
                      const double *branch, int n_models, const double *S, const double *pi,
                      pcsf_model **out);
void pcsf_model_destroy(pcsf_model *m);
int pcsf_model_info(const pcsf_model *m, int *n_nodes, int *n_leaves, int *n_models,
                    int *slots_needed);

/* P(t) for every branch of model `model` at tree scale rho, as the reference
 * builds it (Q.ml:211-249: eigen form, clamp |neg|<=1e-6 to 0, diagonal
 * 1 - sum(offdiag)); built on the GPU in FP64.  P_out is (n_nodes-1)*64*64
 * row-major FP64 (host). */
int pcsf_model_transition(pcsf_model *m, int model, double rho, double *P_out);

/* ---- batch: packed columns of many regions ---------------------------- */

/* codes: n_cols x n_leaves (row per codon column, as PhyloCSF.pleaves
 * yields them, src/PhyloCSF.ml:219-246), host memory; uploaded once and
 * kept leaf-major in HBM.  region_offsets[n_regions+1]: columns of region r
 * are [off[r], off[r+1]).  PCSF_EINVAL if a code is neither 0..63 nor
 * PCSF_MARGINALIZE (checked on the device while the codes are transposed). */
int pcsf_batch_create(pcsf_model *m, const uint8_t *codes, int64_t n_cols,
                      const int64_t *region_offsets, int32_t n_regions, pcsf_batch **out);
/* A batch of every reading frame of packed nucleotide regions: PhyloCSF's
 * AsIs candidate regions with --frames 1, 3 or 6 (src/PhyloCSF.ml:198-206)
 * read by pleaves (:219-246), built on the GPU.  nuc: n_leaves rows of
 * n_nuc characters in tree-leaf order (row l at nuc + l * n_nuc; a species
 * absent from the alignment is a row of '-'); region r spans characters
 * [region_offsets[r], region_offsets[r+1]) of every row.  Batch region
 * r * frames + f is frame f of region r: forward strand offsets 0, 1, 2,
 * then reverse-complement offsets 0, 1, 2, each floor((len - f%3) / 3)
 * codon columns.  Codons with a character outside ACGTacgt are
 * `Marginalize (the reference's rejection of other characters, in
 * Code.DNA.revcomp, is the caller's check). */
int pcsf_batch_create_frames(pcsf_model *m, const uint8_t *nuc, int64_t n_nuc,
                             const int64_t *region_offsets, int32_t n_regions, int frames,
                             pcsf_batch **out);
void pcsf_batch_destroy(pcsf_batch *b);
int64_t pcsf_batch_columns(const pcsf_batch *b);
/* The batch's leaf codes as the kernels read them, back on the host:
 * codes (may be NULL) gets n_cols x n_leaves bytes, a row per codon column as
 * PhyloCSF.pleaves yields them (src/PhyloCSF.ml:219-246); region_offsets
 * (may be NULL) gets the n_regions + 1 column offsets (for a frame batch,
 * region r * frames + f).  Lets a caller check GPU-built codes byte for byte. */
int pcsf_batch_codes(pcsf_batch *b, uint8_t *codes, int64_t *region_offsets);

/* Forward: for every region r and model m,
 *   loglik[r*n_models+m] = sum_col ln z_col          (lpr_leaves)
 *   elpr[r*n_models+m]   = sum_col post_root . ln pi (elpr_anc)
 * rho: n_models values (every region shares them; the fixed strategy passes
 * 1.0) or, with rho_per_region = 1, n_regions*n_models values.
 * loglik / elpr may be NULL (results stay on the device).
 * save_for_backward: keep what pcsf_backward needs (FP64 path only). */
int pcsf_forward(pcsf_batch *b, const double *rho, int rho_per_region, int precision,
                 double *loglik, double *elpr, int save_for_backward);

/* pcsf_forward (shared tree scales) over codes already in device memory,
 * allocating nothing.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The caller passes
 *   d_codes           n_cols x n_leaves bytes on the device, the layout of
 *                     pcsf_batch_create's codes (0..63 or PCSF_MARGINALIZE);
 *   d_region_offsets  n_regions + 1 int64 on the device;
 *   rho               n_models tree scales (host);
 *   d_loglik, d_elpr  n_regions * n_models doubles on the device (outputs).
 * Transition matrices are the model's own (rebuilt in place when rho or the
 * precision changes).  Per-column work goes to one fixed device array of
 * PCSF_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends: a call
 * needs about n_cols * (n_leaves + 16 * n_models) bytes of it (columns
 * rounded up to 128) and is refused with PCSF_EINVAL beyond that (about
 * 370,000 columns for 58 leaves).  Every call on one GPU shares that array,
 * so calls that overlap on one GPU (other threads or streams) race.  Bad
 * codes or offsets give PCSF_EINVAL.  Returns after the results are written
 * (synchronous on the context's stream). */
#define PCSF_NOALLOC_SCRATCH_BYTES (32u << 20)
int pcsf_forward_noalloc(pcsf_model *m, const uint8_t *d_codes, int64_t n_cols,
                         const int64_t *d_region_offsets, int32_t n_regions, const double *rho,
                         int precision, double *d_loglik, double *d_elpr);

/* Backward of the last forward on this batch (same tree scales): given
 * dL/dloglik (n_regions*n_models), returns dL/drho (n_regions*n_models) and,
 * if grad_branch is non-NULL, dL/dt_br per region and model
 * (n_regions*n_models*(n_nodes-1); t_br = rho * b_br is the branch length,
 * so dL/drho = sum_br b_br dL/dt_br).  The gradient the reference's PhyloEM
 * defines (lib/CamlPaml/PhyloEM.ml:140-182: d ELL/dt through the branch
 * posteriors of PhyloLik.ml:140-174, equal to d ln z/dt), computed by the
 * outside (preorder) pass of PhyloLik.ensure_beta (PhyloLik.ml:96-120) in
 * FP64; the inside pass is recomputed, so forward saves no partials.
 * Per-branch sums use FP64 atomics (not bitwise reproducible run to run).
 * After a forward at an MFMA precision (PCSF_FP32 .. PCSF_BF16) and with
 * grad_branch NULL, dL/drho comes from a forward-mode (tangent) pass of the
 * fused kernel in FP32 instead: d a_v/d rho propagated beside a_v through
 * D = b Q P(rho b) on every branch. */
int pcsf_backward(pcsf_batch *b, const double *grad_loglik, double *grad_rho,
                  double *grad_branch);

/* Per-region scores, PhyloCSFModel.score (src/PhyloCSFModel.ml:138-146) for
 * a batch of a two-model (coding, noncoding) model:
 *   PCSF_STRATEGY_FIXED  llr_FixedLik 1.0 (src/PhyloCSFModel.ml:122-128)
 *   PCSF_STRATEGY_MLE    llr_MaxLik ~init:1 (src/PhyloCSFModel.ml:84-99,
 *                        130-136): Fit.find_init then GSL Brent per region and
 *                        model, all regions advancing in lock step on the GPU
 * At an MFMA precision (PCSF_FP32 .. PCSF_BF16) the maximisers search on
 * that precision's likelihoods and take lpr and elpr at their estimates from
 * one FP64 forward (PCSF_FP64M) (the MFMA rounding, up to ~6e-9 of lpr on regions of
 * thousands of codons at per-region scales, would otherwise reach the score).
 * score / anc_comp in decibans (n_regions); rho, lpr, elpr per region and
 * model (n_regions*2, coding first); status per region (PCSF_OK, or
 * PCSF_EFAIL where the reference raises inside the maximiser);
 * n_evals = likelihood evaluations per region.  Any output may be NULL. */
#define PCSF_STRATEGY_FIXED 0
#define PCSF_STRATEGY_MLE 1
/* mle with the gradient: root of h = d lpr / d ln rho on [0.01, 10].  lpr
 * of every pair first at seven tree scales shared by all pairs (ln rho =
 * -2 .. 2, the fixed-scale path); the search starts at the maximum of the
 * cubic through the best of them and three neighbours, with that cubic's
 * curvature, inside the bracket of the neighbours.  Then Newton steps inside
 * the bracket the signs of h leave -- to the maximum of the cubic through
 * the last two points' lpr and h while they are more than 0.01 apart, the
 * secant's Newton point when closer, the column scores' centred outer
 * product as the curvature where neither is known --
 * bisecting when a step leaves the bracket or two steps fail to halve |h|;
 * a pair stops when its next step would leave less than 1e-12 |lpr| (FP64,
 * FP64M) or 2e-5 nats (the MFMA precisions), at 1e-6 (1e-4) in ln rho, or at
 * a bound where h keeps its sign.  Finds the maximum the reference's
 * 1%-tolerance Brent approximates.  The gradient is the FP64 outside pass at
 * PCSF_FP64 and PCSF_FP64M, the FP32 tangent pass otherwise. */
#define PCSF_STRATEGY_MLE_GRAD 2
int pcsf_score(pcsf_batch *b, int strategy, int precision, double *score, double *anc_comp,
               double *rho, double *lpr, double *elpr, int32_t *status, int32_t *n_evals);

/* The omega (dN/dS) test, OmegaModel.score (src/OmegaModel.ml:195-219) for
 * every region of a batch: H0 (omega = sigma = 1) and H1 (omega = omega_h1,
 * sigma = sigma_h1; the reference's defaults are 0.2 and 0.01, a parameter
 * set's `_omega` file may name others, src/PhyloCSF.ml:452-463), each the
 * MAP fit kr_map (:160-190): three rounds of maximize_lpr over the tree scale
 * in [0.001, 10] (half-Cauchy prior, mode 1, scale 0.5) then over kappa in
 * [1, 10] (gamma(7, 0.25) prior on kappa - 1), with F3x4 codon frequencies
 * counted from the region (:102-134).  The codon model (:21-80) is rebuilt
 * and diagonalised per (region, hypothesis, kappa) on the host; every
 * likelihood is a batched per-region forward on the GPU.
 * The batch's model is a tree-only model (pcsf_model_create with n_models =
 * 0) or any two-model model (its rate models are not used).
 * score in decibans (n_regions; anc_comp is not defined for this test);
 * lpr (log prior + lpr at the estimates), rho, kappa per region and
 * hypothesis (n_regions * 2, H0 first); status per region (PCSF_EFAIL where
 * the reference raises); n_evals = likelihood evaluations per region.
 * Any output may be NULL. */
int pcsf_score_omega(pcsf_batch *b, double omega_h1, double sigma_h1, int precision,
                     double *score, double *lpr, double *rho, double *kappa, int32_t *status,
                     int32_t *n_evals);

/* Autotuner of the fused kernel at one MFMA precision (PCSF_FP32 ..
 * PCSF_BF16): times every launch shape (waves per workgroup sharing one LDS
 * P tile x register cap x gather prefetch x 16-column tiles per wave) on
 * this batch at tree scale 1,
 * `reps` launches each, keeps the fastest on the batch's model for later
 * forwards at that precision and returns the per-shape milliseconds in
 * ms_out[PCSF_N_VARIANTS].  Until then a model uses shape 2 if its tree has
 * generated programs (see pcsf_model_specialized), else shape 1.  No
 * counterpart in the reference (a tuning step of this implementation). */
#define PCSF_N_VARIANTS 7
int pcsf_autotune(pcsf_batch *b, int precision, int reps, double *ms_out, int *chosen);
/* Select a launch shape by index (0..PCSF_N_VARIANTS-1) for one MFMA
 * precision without timing. */
int pcsf_set_variant(pcsf_model *m, int precision, int variant);
/* Tree-specialised kernels: the library carries straight-line op programs
 * generated at build time for the shipped trees; a model whose tree matches
 * one uses it (FP32 and split-FP16 GEMMs, launch shapes 1, 2, 4 and 6) instead
 * of the op-program interpreter.  enable = 1 / 0 allows / forbids them
 * (-1 leaves the setting); *available (may be NULL) gets 1 if the plain
 * program is specialised, | 2 if the cherry-table program is.  No
 * counterpart in the reference. */
int pcsf_model_specialized(pcsf_model *m, int enable, int *available);

/* Device pointers of the last forward's per-region results (FP64,
 * n_regions*n_models each) and the kernel times when timing is enabled.
 * pcsf_backward does not change them (its passes reduce into buffers of
 * their own); pcsf_score with an mle strategy runs forwards of its own, so
 * afterwards they hold its last evaluation's values. */
int pcsf_batch_device_results(pcsf_batch *b, const double **d_loglik, const double **d_elpr);
int pcsf_batch_last_timing(const pcsf_batch *b, double *prune_ms, double *reduce_ms,
                           double *pbuild_ms);

/* ---- PhyloLik compatibility entry ------------------------------------ */

/* PhyloLik.prepare + likelihood + node_posterior (lib/CamlPaml/PhyloLik.ml:
 * 46-138) over n_cols columns with caller-supplied P matrices, any alphabet
 * size k <= 64 and any tree size:
 *   P[(n_nodes-1)*k*k] row-major, P[br] for the branch leading to node br;
 *   prior[k]; leaves[n_cols*nl] state index or -1 (`Marginalize).
 * Outputs z[n_cols] and, if non-NULL, root_posterior[n_cols*k].  FP64. */
int pcsf_lik_columns(pcsf_ctx *ctx, int k, int n_nodes, const int32_t *children,
                     const double *P, const double *prior, const int32_t *leaves,
                     int64_t n_cols, double *z, double *root_posterior);

/* The rest of PhyloLik (lib/CamlPaml/PhyloLik.mli:6-35) over n_cols columns,
 * FP64, any k <= 64 and tree size:
 *   leaves[n_cols*nl]: state index (`Certain; >= k gives a leaf vector of
 *     zeros, as the reference's), -1 (`Marginalize) or -2 (`Distribution,
 *     its k probabilities at leaf_dist[(col*nl + leaf)*k ..]);
 *   z[n_cols]: likelihood;
 *   node_post (may be NULL) [n_cols][n_nodes][k]: node_posterior of every
 *     node (a leaf's is its leaf vector; all zero where z = 0);
 *   ecounts (may be NULL) [(n_nodes-1)][k][k]: add_branch_posteriors of every
 *     branch (:140-174) with weight weights[col] (NULL: 1), summed over the
 *     columns and ADDED to what the array holds -- the E-step of
 *     PhyloEM.collect_sufficient_statistics (lib/CamlPaml/PhyloEM.ml:14-20):
 *     ecounts[br][a][b] += w P(parent = a, br = b | leaves), computed as one
 *     rank-n_cols update per branch. */
int pcsf_lik_posteriors(pcsf_ctx *ctx, int k, int n_nodes, const int32_t *children,
                        const double *P, const double *prior, const int32_t *leaves,
                        const double *leaf_dist, int64_t n_cols, const double *weights,
                        double *z, double *node_post, double *ecounts);

/* ---- synthetic alignments (bench / tests) ----------------------------- */

/* Simulates codes for packed regions: region r is drawn from model
 * region_model[r] at tree scale region_rho[r]: root ~ pi, each child ~ row
 * of P(rho*b) (lib/CamlPaml/PhyloModel.ml:38-54).  Leaves 1..nl-1 are then
 * masked to PCSF_MARGINALIZE: gap_mode 0 = i.i.d. with probability gap_p,
 * gap_mode 1 = runs of mean length gap_run (codons) covering gap_p; and
 * with probability n_p a codon holds an N (also `Marginalize).
 * Counter-based Philox4x32-10 streams keyed by (seed, region, column): the
 * result does not depend on the thread count.  codes_out is n_cols*nl. */
int pcsf_simulate(pcsf_model *m, int32_t n_regions, const int64_t *region_offsets,
                  const int32_t *region_model, const double *region_rho, int gap_mode,
                  double gap_p, double gap_run, double n_p, uint64_t seed, int n_threads,
                  uint8_t *codes_out);
/* The same alignments as nucleotide characters: nuc_out holds n_leaves rows
 * of 3 * n_cols characters (row l at nuc_out + l * 3 * n_cols; codon column
 * c at characters 3c .. 3c+2); a gap is "---", an N codon carries 'N' at
 * one of its three positions.  pcsf_batch_create_frames of this output
 * with frames = 1 gives the codes of pcsf_simulate. */
int pcsf_simulate_nucleotides(pcsf_model *m, int32_t n_regions, const int64_t *region_offsets,
                              const int32_t *region_model, const double *region_rho,
                              int gap_mode, double gap_p, double gap_run, double n_p,
                              uint64_t seed, int n_threads, uint8_t *nuc_out);

#ifdef __cplusplus
}
#endif
#endif
