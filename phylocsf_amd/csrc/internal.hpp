// internal.hpp -- shared declarations of the MI355X PhyloCSF library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/phylocsf_hip.h"
#include "host/ratematrix.hpp"

// Op program of the fused whole-tree kernel: a post-order walk of the tree
// that keeps at most `slots` partial-likelihood tiles live in registers.
// word0 = opcode | depth << 8, word1 = arg0 | arg1 << 16.
enum PcsfOp : int {
  OP_CHERRY = 1,   // slot[d] = gather(leaf a) * gather(leaf b); rescale
  OP_GEMM = 2,     // slot[d] = P_a slot[d]  (MFMA, P tile staged in LDS); arg1 = next GEMM branch
  OP_LEAFMUL = 3,  // slot[d] *= gather(leaf a); rescale
  OP_MULPOP = 4,   // slot[d] *= slot[d+1]; rescale
  OP_ROOT = 5,     // write ln z and elpr for slot[d]
  OP_END = 6,
};
constexpr int kNoBranch = 0xFFFF;
constexpr int kMaxSlots = 6;
constexpr int kTileCols = 16;   // columns per wave
constexpr int kGroupCols = 128; // columns per workgroup (8 waves)
constexpr int kTileFloats = 4096 + 64;  // P tile staged in LDS: fragments + diagonal

struct pcsf_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  bool timing = false;
  int num_cus = 256;
};

struct pcsf_model {
  pcsf_ctx* ctx = nullptr;
  int n = 0, nl = 0, nm = 0, nbr = 0;
  std::vector<int> lc, rc, parent;
  std::vector<double> branch;
  std::vector<pcsf::Eigen> eig;
  std::vector<double> pi;  // nm * 64
  // program for the fused kernel
  std::vector<int32_t> prog;  // pairs
  int slots = 0;
  int n_gemm = 0;
  int mfma_variant = 1;  // launch shape of the fused kernel (pcsf_autotune)
  int first_gemm_br = kNoBranch;
  // device: model constants
  double* d_V = nullptr;     // nm * 4096
  double* d_Vinv = nullptr;  // nm * 4096
  double* d_lam = nullptr;   // nm * 64
  double* d_pi = nullptr;    // nm * 64
  double* d_lnpi = nullptr;  // nm * 64, ln pi
  double* d_branch = nullptr;  // nbr
  int32_t* d_lc = nullptr;     // n
  int32_t* d_rc = nullptr;
  int32_t* d_prog = nullptr;
  // device: transition matrices at the cached tree scales
  std::vector<double> cached_rho;
  double* d_P64T = nullptr;  // nm * nbr * 4096  FP64, P^T (col-major P)
  float* d_PT32 = nullptr;   // nm * nbr * 4096  FP32, P^T[c][a] = P[a][c] / P[c][c]
  float* d_Pfrag = nullptr;  // nm * nbr * 4160  FP32, off-diagonal of P in MFMA A-fragment
                             //   order, then the 64 diagonal entries P[a][a] - 1
  double* d_lnPcc = nullptr; // nm * nbr * 64    FP64, ln P[c][c]
  double* d_rho = nullptr;   // nm
  int* d_perr = nullptr;     // error record of the P builder
  double* d_perrval = nullptr;
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
};

// ---- kernel launchers (kernels/*.hip) --------------------------------------
hipError_t launch_transpose_codes(const uint8_t* src, int64_t ncols, int nl, uint8_t* dst,
                                  int64_t ld, hipStream_t s);
hipError_t launch_pbuild(const pcsf_model* m, const double* d_rho, hipStream_t s);
hipError_t launch_prune_fp64(const pcsf_model* m, const pcsf_batch* b, hipStream_t s);
// Launch shapes of the fused kernel that the autotuner chooses among.
struct MfmaVariant {
  int wpg;   // waves per workgroup (16 columns each) sharing one LDS P tile
  int minw;  // minimum waves per SIMD the compiler must allow (register cap)
  const char* what;
};
constexpr int kNumMfmaVariants = 5;
extern const MfmaVariant kMfmaVariants[kNumMfmaVariants];
hipError_t launch_prune_mfma(const pcsf_model* m, const pcsf_batch* b, int variant,
                             hipStream_t s);
hipError_t launch_region_reduce(const pcsf_batch* b, hipStream_t s);
hipError_t launch_lik_columns(int k, int n, const int32_t* d_lc, const int32_t* d_rc,
                              const double* d_P, const double* d_prior, const int32_t* d_leaves,
                              int64_t ncols, double* d_z, double* d_post, hipStream_t s);
