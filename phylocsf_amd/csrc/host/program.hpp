// program.hpp -- the op program of the fused pruning kernel (host side).
//
// The tree is compiled into a post-order walk that keeps at most `slots`
// partial-likelihood tiles live in registers.  Four words per op:
// word0 = opcode | depth << 8 | use_prefetch << 16, word1 = arg0 | arg1 << 16,
// word2 = leaf whose gather a GEMM issues ahead (kNoBranch if none) or the
// table index of OP_CHERRYT, word3 = 0 or, for a GEMM, the table row it
// issues ahead: kPfTable | table << 24 | leaf1 << 12 | leaf0.  use_prefetch
// marks a gather op whose first leaf (arg0), or whose table row, was issued
// by the preceding GEMM.  This is this
// implementation's schedule of PhyloLik.ensure_alpha
// (lib/CamlPaml/PhyloLik.ml:78-90); the reference has no counterpart.
#pragma once
#include <cstdint>
#include <vector>

enum PcsfOp : int {
  OP_CHERRY = 1,   // slot[d] = gather(leaf a0) * gather(leaf a1); rescale
  OP_GEMM = 2,     // slot[d] = P_a0 slot[d] (MFMA, P tile in LDS); a1 = next GEMM's branch
  OP_LEAFMUL = 3,  // slot[d] *= gather(leaf a0); rescale
  OP_MULPOP = 4,   // slot[d] *= slot[d+1]; rescale
  OP_ROOT = 5,     // write ln z and elpr for slot[d]
  OP_END = 6,
  OP_CHERRYT = 7,  // slot[d] = message of cherry node w2 above it, from its table at the
                   // codes of leaves a0, a1 (replaces CHERRY + GEMM; fixed scales only)
};
constexpr int kTabCodes = 65;  // table index per leaf: codon 0..63, 64 = `Marginalize
constexpr int kNoBranch = 0xFFFF;
constexpr int kOpWords = 4;
constexpr int kMaxSlots = 6;
constexpr uint32_t kPfTable = 0x80000000u;  // word3 of a GEMM: a table row is issued ahead
constexpr int kPfLeafMax = 0xfff, kPfTableMax = 0x7f;

namespace pcsf {

struct Program {
  std::vector<int32_t> ops;       // kOpWords words per op, ending with OP_END
  int slots = 0;                  // live tiles needed
  int n_gemm = 0;                 // GEMM ops
  int first_gemm_br = kNoBranch;  // branch of the first GEMM
  std::vector<int32_t> cherries;  // tables: (node, left leaf, right leaf) per table index
};

// lc / rc: children of every node (-1 for leaves), leaves 0..nl-1, root
// n-1.  with_tables: messages above cherries (other than the root) come
// from lookup tables.
Program make_program(const std::vector<int>& lc, const std::vector<int>& rc, int nl,
                     bool with_tables);

}  // namespace pcsf
