// internal.hpp -- shared declarations of the MI355X PhyloCSF library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/phylocsf_hip.h"
#include "host/program.hpp"
#include "host/ratematrix.hpp"

constexpr int kTileCols = 16;   // columns per wave
constexpr int kGroupCols = 128; // columns per workgroup (8 waves)
constexpr int kTileFloats = 4096 + 64;  // P tile staged in LDS: fragments + diagonal
constexpr int kLeafRows = 65;  // leaf messages per branch: codons 0..63, 64 = `Marginalize

// Transition-matrix sets on the device: set s is model s % nm at tree scale
// rho[s] (the fixed strategy uses nm sets; per-region scales use one set per
// (region, model) of a chunk of regions).
// Formats of the P tile (PSets::d_Pfrag, kTileFloats floats per branch):
//   kFragF32   off-diagonal FP32 in v_mfma_f32_16x16x4f32 A order + diagonal
//   kFragF16   x = offdiag(P) x 2^13 as FP16 hi = RN(x), then the FP16 residual
//              lo = RN((x - hi) x 2^13), in v_mfma_f32_16x16x32_f16 A order
//              (4096 halves each) + diagonal
//   kFragBF16  the same in BF16, unscaled, residual unscaled
// The 64 FP32 values P[a][a] - 1 follow at float 4096 in every format.
constexpr int kFragF32 = 0, kFragF16 = 1, kFragBF16 = 2;
// 2^kHalfExp: the scale of the FP16 P tiles and of their residuals.  2^13
// keeps |x| < 2^13 and |lo| <= 2^14 inside FP16's range and their entries
// above 2^-27 normal numbers (the matrix cores read FP16 subnormals as zero).
constexpr int kHalfExp = 13;
int frag_format(int precision);  // -1 for a precision without MFMA tiles

struct PSets {
  int cap = 0;                 // sets allocated
  int fmt = -1;                // format of d_Pfrag as last built
  bool fp64 = false;           // P64T and P64F allocated
  double* d_P64T = nullptr;    // cap * nbr * 4096 FP64 P^T (P64T[b][a] = P[a][b])
  double* d_P64F = nullptr;    // cap * nbr * kTile64 FP64 P for the matrix-core forward:
                               //   an internal branch's P in v_mfma_f64_16x16x4f64
                               //   A-fragment order, a leaf branch's 65 leaf messages
  double* d_P64FT = nullptr;   // cap * nbr * 4096 FP64 P^T in that A-fragment order (the
                               //   outside pass's products; internal branches)
  float* d_PT32 = nullptr;     // cap * nbr * 65 * 64 FP32 leaf messages: row c < 64 is
                               //   P[a][c] / P[c][c], row 64 all ones (`Marginalize)
  float* d_Pfrag = nullptr;    // cap * nbr * kTileFloats: off-diagonal of P in MFMA
                               //   A-fragment order, then the 64 values P[a][a] - 1
  double* d_lnPcc = nullptr;   // cap * nbr * 65 FP64 ln P[c][c] (0 for row 64)
  // tangent data (allocated on request): D = dP/drho = b Q P(rho b)
  float* d_DT32 = nullptr;     // cap * nbr * 65 * 64: row c < 64 D[a][c] / P[c][c], row 64 0
  float* d_Dfrag = nullptr;    // cap * nbr * 4096: D in the FP32 A-operand order
  double* d_rho = nullptr;     // cap tree scales
  uint8_t* d_active = nullptr; // cap flags: 0 = set not needed by this launch
  // cherry tables (fixed scales only): message P_c (P_a[:,x] o P_b[:,y]) of
  // cherry c above its two leaves a, b, normalised to max 1, and its ln max
  float* d_ctab = nullptr;     // cap * nctab * 65 * 65 * 64
  double* d_clnt = nullptr;    // cap * nctab * 65 * 65
  // rate models per set (allocated on request: a model whose Q differs from
  // set to set, as the omega test's per region and hypothesis); nullptr: set
  // s uses model s % nm's eigensystem and root prior
  double* d_eVT = nullptr;     // cap * 4096, V transposed
  double* d_eVinv = nullptr;   // cap * 4096
  double* d_elam = nullptr;    // cap * 64
  double* d_epi = nullptr;     // cap * 64 root prior
  double* d_elnpi = nullptr;   // cap * 64 its logarithm
};

// Device and pinned buffers of the E-step (capi.hip estep_mfma), kept by the
// context between calls and grown as needed: the u, v hand-over of a chunk
// (gigabytes), the passes' node tiles, and two leaf-state slots that a second
// stream fills (from pinned staging) while the previous chunk computes.
struct EstepCache {
  double *U = nullptr, *V = nullptr, *scr = nullptr;
  size_t cap_uv = 0, cap_scr = 0;  // doubles of U (= of V) and of scr
  int32_t* leaves[2] = {nullptr, nullptr};
  int32_t* staged[2] = {nullptr, nullptr};  // pinned host
  size_t cap_leaves = 0;                    // int32 per slot
  hipStream_t copy = nullptr;               // default priority, like the context's
  hipEvent_t uploaded[2] = {nullptr, nullptr}, consumed[2] = {nullptr, nullptr};
};

struct pcsf_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  bool timing = false;
  int num_cus = 256;
  EstepCache es;
};

struct pcsf_model {
  pcsf_ctx* ctx = nullptr;
  int n = 0, nl = 0, nm = 0, nbr = 0;
  // a tree without rate models (made with n_models = 0, for the omega test,
  // whose sets carry rate models of their own): nm = 2 hypotheses
  bool tree_only = false;
  std::vector<int> lc, rc, parent;
  std::vector<double> branch;
  std::vector<pcsf::Eigen> eig;
  std::vector<double> pi;  // nm * 64
  // programs for the fused kernel: `prog` walks every node; `prog_t` reads
  // the messages above cherries from tables (fixed scales)
  std::vector<int32_t> prog, prog_t;
  int32_t* d_prog_t = nullptr;
  int first_gemm_br_t = kNoBranch;
  std::vector<int32_t> cherries;  // cherry nodes with a GEMM above them (table order)
  int32_t* d_cherries = nullptr;  // node, left leaf, right leaf per table
  int slots = 0;
  int n_gemm = 0;
  int mfma_variant[5] = {1, 1, 1, 1, 1};  // launch shape per precision (pcsf_autotune;
                                          // 2 for trees with generated programs)
  // tree-specialised programs matching prog / prog_t (nullptr: none), and
  // whether launches may use them
  const void* tree_p = nullptr;
  const void* tree_t = nullptr;
  bool specialized = true;
  int first_gemm_br = kNoBranch;
  // device: model constants
  double* d_VT = nullptr;    // nm * 4096, V transposed (VT[k][i] = V[i][k])
  double* d_Vinv = nullptr;  // nm * 4096
  double* d_lam = nullptr;   // nm * 64
  double* d_pi = nullptr;    // nm * 64
  double* d_lnpi = nullptr;  // nm * 64, ln pi
  double* d_QT = nullptr;    // nm * 4096, Q transposed (QT[k][a] = Q[a][k])
  double* d_branch = nullptr;  // nbr
  int32_t* d_lc = nullptr;     // n
  int32_t* d_rc = nullptr;
  int32_t* d_prog = nullptr;
  // leaves gathered by prog / prog_t (the others are read through tables)
  int16_t* d_lsl = nullptr;
  int16_t* d_lsl_t = nullptr;
  int n_lsl = 0, n_lsl_t = 0;
  // device: transition matrices at the cached tree scales (set q = model q)
  std::vector<double> cached_rho;
  int cached_fmt = -1;
  PSets fixed;
  // failure record of the P builder: the minimum packed (set, branch, row,
  // column, check) key over failing rows, all ones when none failed
  unsigned long long* d_perr = nullptr;
};

// Region-aligned copy of a batch's codes: every region starts on a multiple
// of kAlignCols, so a workgroup's columns share one region (and one P set).
constexpr int kAlignCols = 64;
struct Aligned {
  int64_t ld = 0;                   // aligned columns (multiple of kGroupCols)
  uint8_t* d_codes = nullptr;       // nl * ld
  int32_t* d_block_region = nullptr;// per kAlignCols block: region (-1 = none)
  int64_t* d_astart = nullptr;      // per region: first aligned column
  std::vector<int64_t> astart;      // host copy
  std::vector<int64_t> block0;      // per region: first block
};

struct pcsf_batch {
  pcsf_model* m = nullptr;
  int64_t ncols = 0, ld = 0;  // ld = padded column count (multiple of kGroupCols)
  int32_t nreg = 0;
  uint8_t* d_codes = nullptr;      // nl * ld, leaf-major
  int64_t* d_offsets = nullptr;    // nreg + 1
  double* d_logz = nullptr;        // nm * ld
  double* d_colelpr = nullptr;     // nm * ld
  double* d_loglik = nullptr;      // nreg * nm
  double* d_elpr = nullptr;        // nreg * nm
  hipEvent_t ev[6] = {};
  float t_pbuild = 0, t_prune = 0, t_reduce = 0;
  std::vector<int64_t> offsets;     // host copy
  Aligned al;                       // built on the first per-region forward
  std::vector<double> last_rho;     // tree scales of the last forward (nreg * nm)
  int last_precision = 0;           // PCSF_FP64 .. PCSF_BF16 of the last forward
  PSets rs;                         // per-region transition sets (one chunk)
  double* d_alogz = nullptr;        // nm * al.ld, per aligned column
  double* d_acolelpr = nullptr;
  // gradient passes (pcsf_backward, mle_grad) reduce into these, so that the
  // last forward's d_loglik / d_elpr stay as they were
  double* d_gloglik = nullptr;      // nreg * nm: lpr at the differentiated scales
  double* d_grad = nullptr;         // nreg * nm: d lpr / d rho
  double* d_gsq = nullptr;          // nreg * nm: sum over columns of (d ln z / d rho)^2
  double* d_adj_scratch = nullptr;  // FP64 outside pass: node tiles per resident wave
  size_t adj_scratch = 0;           // its size in doubles
  double* d_setqs = nullptr;        // per-set model settings of a chunk (omega test)
  size_t setqs_cap = 0;             // its size in doubles
  uint8_t* d_eigmode = nullptr;     // per set: kEigCold / kEigWarm / kEigKeep
  double* d_eigref = nullptr;       // per set: the reference basis of its warm starts
  // the settings each set of `rs` was last built from and those of its
  // reference basis (host), valid while a pass fits in one chunk (set k is
  // pair k) and `rs` is not reallocated
  std::vector<double> eig_qs, eig_ref_qs;
  std::vector<uint8_t> eig_ok, eig_ref_ok;
};

// Per-region-scale launch description (one chunk of regions).
struct RegionChunk {
  int32_t r0 = 0, nreg = 0;         // regions [r0, r0 + nreg)
  int64_t block0 = 0, nblocks = 0;  // aligned blocks covering them
  const PSets* sets = nullptr;      // set (r - r0) * nm + q; set q when `sets` is the
                                    // model's own (fixed) sets, shared by every region
};

// ---- kernel launchers (kernels/*.hip) --------------------------------------
// bad (may be null) gets a nonzero value if a code is neither a codon index
// (0..63) nor PCSF_MARGINALIZE
hipError_t launch_transpose_codes(const uint8_t* src, int64_t ncols, int nl, uint8_t* dst,
                                  int64_t ld, int* bad, hipStream_t s);
// What a transition-set build writes (the sets hold room for it): FP64 P^T
// (P64T, if the sets have it), the MFMA kernels' leaf rows and tiles in
// format fmt (PT32, Pfrag, lnPcc), the tangent's D (DT32, Dfrag, if the sets
// have them).
constexpr unsigned kBuildP64 = 1, kBuildMfma = 2, kBuildTangent = 4, kBuildP64M = 8,
                   kBuildP64MT = 16;
// Doubles per branch of PSets::d_P64F.  Internal branch: P[a][b] at
// ((a / 16) * 16 + b / 4) * 64 + a % 16 + 16 (b % 4) (the A-fragment order of
// fp64_tiles.hpp).  Leaf branch: 65 messages of 64 entries, message c < 64
// the column P[.][c] and message 64 all ones (`Marginalize), each stored so
// that lane group g's 16 rows are contiguous: row 16 mo + 4 v + g at
// c * 64 + 16 g + 4 mo + v.
constexpr int kTile64 = 65 * 64;
hipError_t launch_pbuild(const pcsf_model* m, const PSets& sets, int nsets, int fmt,
                         unsigned outputs, hipStream_t s);
hipError_t launch_cherry_tables(const pcsf_model* m, const PSets& sets, int nsets,
                                hipStream_t s);
hipError_t launch_prune_fp64(const pcsf_model* m, const pcsf_batch* b, hipStream_t s);
hipError_t launch_prune_fp64_regions(const pcsf_model* m, const pcsf_batch* b,
                                     const RegionChunk& ch, hipStream_t s);
hipError_t launch_prune_mfma_regions(const pcsf_model* m, const pcsf_batch* b, int precision,
                                     int variant, const RegionChunk& ch, hipStream_t s);
hipError_t launch_align_codes(const pcsf_batch* b, hipStream_t s);
hipError_t launch_frames(const uint8_t* nuc, int64_t nlen, const int64_t* d_roff,
                         const int64_t* d_foff, int32_t nreg, int frames, int nl, uint8_t* dst,
                         int64_t ld, hipStream_t s);
// The outside pass on FP64 matrix cores (kernels/adjoint_mfma.hip): region-
// aligned blocks of 64 columns per workgroup; scratch holds one 64 x 16 FP64
// tile per internal node for each of the grid's waves.
struct AdjointArgs {
  const uint8_t* codes;
  int64_t ld, ald;           // leading dimensions of codes and of the per-column outputs
  int64_t c0, nblocks;       // first column and block count of the chunk
  int n, nl, nm, nbr;
  const int32_t *lc, *rc;
  const double* P64F;        // sets: P per branch (PSets::d_P64F layout)
  const double* P64FT;       // sets: P^T per branch in A-fragment order
  const double* QT;          // per model: Q^T
  const double* pi;          // per model, or per set with prior_per_set
  const double* lnpi;        // its logarithm
  int prior_per_set;
  const double* branch;      // nbr
  const int32_t* block_region;
  const int64_t* astart;
  const int64_t* off;
  const uint8_t* active;
  int r0;
  int shared_sets;           // 1: set q for every region (the model's fixed sets)
  double* logz;              // nm x ald
  double* coldrho;           // nm x ald: d ln z / d rho per column
  double* grad_branch;       // (region - r0) x nm x nbr, added to (may be null)
  double* scratch;
};
int adjoint_mfma_grid(const pcsf_model* m);
size_t adjoint_mfma_scratch_doubles(const pcsf_model* m, int grid);
hipError_t launch_adjoint_mfma(const pcsf_model* m, const pcsf_batch* b, const RegionChunk& ch,
                               double* grad_branch, double* scratch, int grid, hipStream_t s);
// The FP64 forward on matrix cores (precision PCSF_FP64M, prune_fp64m.hip):
// ln z and the root's expected ln pi per column of the chunk's region blocks
// into the batch's aligned per-column buffers, from the sets' P64F.
hipError_t launch_forward_fp64m(const pcsf_model* m, const pcsf_batch* b, const RegionChunk& ch,
                                hipStream_t s);
// the omega test's rate model of every active set of a chunk from its twelve
// settings (d_qs: nsets * 12): eigensystem and root prior into the sets'
// own-model arrays
// Per-set modes of launch_omega_eigen: build from scratch (and store the
// eigenvectors as the set's reference basis in d_ref, 4096 doubles a set),
// start from that basis (same pi), or keep the outputs.
constexpr uint8_t kEigCold = 0, kEigWarm = 1, kEigKeep = 2;
hipError_t launch_omega_eigen(const double* d_qs, const uint8_t* d_mode, double* d_ref,
                              int nsets, const PSets& sets, hipStream_t s);
hipError_t launch_prune_tangent_regions(const pcsf_model* m, const pcsf_batch* b,
                                        const RegionChunk& ch, hipStream_t s);
// sums over each region of the chunk into loglik / elpr (nreg * nm)
// (sq, optional: the sums of the squared per-column second values -- with
// d ln z / d rho there, the outer-product estimate of lpr's curvature)
hipError_t launch_region_reduce_aligned(const pcsf_batch* b, const RegionChunk& ch,
                                        double* loglik, double* elpr, hipStream_t s,
                                        double* sq = nullptr);
// Launch shapes of the fused kernel that the autotuner chooses among.
struct MfmaVariant {
  int wpg;        // waves per workgroup (16 columns each) sharing one LDS P tile
  int minw;       // minimum waves per SIMD the compiler must allow (register cap)
  bool prefetch;  // GEMMs issue the next gather op's leaf or table row ahead (+18
                  // VGPRs; the program names them only at GEMMs above its deepest)
  int g;          // 16-column tiles per wave, walked side by side
  const char* what;
};
constexpr int kNumMfmaVariants = 7;
extern const MfmaVariant kMfmaVariants[kNumMfmaVariants];
hipError_t launch_prune_mfma(const pcsf_model* m, const pcsf_batch* b, int precision,
                             int variant, hipStream_t s);
// The generated straight-line program whose op words equal `ops` (built
// with or without cherry tables), or nullptr.
const void* find_tree_program(const std::vector<int32_t>& ops, bool tables);
hipError_t launch_region_reduce(const pcsf_batch* b, hipStream_t s);
// The PhyloLik compatibility path (kernels/lik.hip), device pointers.
struct LikArgs {
  int k = 0, n = 0;
  const int32_t *lc = nullptr, *rc = nullptr, *parent = nullptr, *sib = nullptr;
  const double* P = nullptr;      // (n-1) x k x k, row-major per branch
  const double* prior = nullptr;  // k
  const int32_t* leaves = nullptr;  // ncols x nl: state, -1 `Marginalize, -2 `Distribution
  const double* leaf_dist = nullptr;  // ncols x nl x k (rows used where leaves = -2)
  int64_t ncols = 0;
  const double* weights = nullptr;  // ncols (null: 1)
  double* ws = nullptr;             // grid x 2n x k workspace
  double* z = nullptr;              // ncols
  int post_mode = 0;                // 0 none, 1 root (ncols x k), 2 every node (ncols x n x k)
  double* post = nullptr;
  double* U = nullptr;              // (n-1) x grid x k, with V: the E-step factors
  double* V = nullptr;
  double* ecounts = nullptr;        // (n-1) x k x k, added to
};
// grid workgroups; with U set, ncols must not exceed grid
hipError_t launch_lik_posteriors(const LikArgs& a, int grid, hipStream_t s);
// The E-step on FP64 matrix cores for k = 64 without `Distribution leaves
// (kernels/estep.hip): u, v per (branch, column), z per column, then
// ecounts += P o sum_col u v^T per branch.
struct EstepArgs {
  int n = 0, nl = 0;
  const int32_t *lc = nullptr, *rc = nullptr;
  const double* P = nullptr;        // (n-1) x 64 x 64, row-major per branch
  const double* Pf = nullptr;       // (n-1) x 4096: P in A-fragment order (fp64_tiles.hpp)
  const double* PTf = nullptr;      // (n-1) x 4096: P^T in A-fragment order
  const int32_t* seq = nullptr;     // the passes' products in order: branch b for P_b,
                                    //   -1 - b for P_b^T (estep_sequence)
  int nseq = 0;
  const double* rowsum = nullptr;   // (n-1) x 64: P's row sums (`Marginalize leaf messages)
  const double* prior = nullptr;    // 64
  const int32_t* leaves = nullptr;  // ncols x nl: state, -1 `Marginalize, >= 64 zeros
  const double* weights = nullptr;  // ncols (null: 1)
  int64_t ncols = 0;
  double* z = nullptr;              // ncols
  double* U = nullptr;              // (n-1) x ldu x 64
  double* V = nullptr;
  int64_t ldu = 0;                  // columns per branch in U, V (>= ncols rounded up to 64)
  double* scratch = nullptr;        // estep_scratch_doubles(n, grid)
};
int estep_grid(const pcsf_ctx* c);
size_t estep_scratch_doubles(int n, int grid);
// the products of estep_passes_kernel in the order it runs them
std::vector<int32_t> estep_sequence(int n, const std::vector<int32_t>& lc,
                                    const std::vector<int32_t>& rc);
// Pf and PTf of EstepArgs from P (nbr row-major 64 x 64 matrices)
hipError_t launch_estep_frags(const double* P, int nbr, double* Pf, double* PTf, hipStream_t s);
hipError_t launch_estep(const EstepArgs& a, const double* P, double* ecounts, int grid,
                        hipStream_t s);
