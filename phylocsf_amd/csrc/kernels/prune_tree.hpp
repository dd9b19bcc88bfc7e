// prune_tree.hpp -- tree-specialised instances of the fused MFMA kernel.
//
// A straight-line op program (generated at build time from a shipped tree,
// host/gen_tree_programs.cpp) replaces the interpreter in
// prune_mfma_kernel; everything else -- ops, task loop, layouts -- is the
// same code.  Specialised kernels exist for the FP32 and split-FP16 GEMMs
// and the launch shapes the autotuner most often picks; other requests fall
// back to the interpreter.
#pragma once
#include "prune_fp64m.hpp"
#include "prune_mfma.hpp"
#include "prune_tangent.hpp"

namespace pcsf_mfma {

typedef hipError_t (*TreeLaunch)(const MfmaParams&, int precision, int variant, int cus,
                                 hipStream_t);

// One generated program: the tree's name, whether it reads cherry tables,
// its op words (to match a model's program), its launcher and (programs
// without tables) the launchers of its tangent pass and FP64 matrix-core
// forward.
struct TreeProgram {
  const char* name;
  int tables;
  const int32_t* ops;  // nullptr: the tree has no specialised program
  int n_words;
  TreeLaunch launch;
  TreeLaunch tangent;
  TreeLaunch fp64m;
};

// Every generated program (host function of the generated registry).
const TreeProgram* const* tree_programs(int* n);

template <int PREC, class PROG>
hipError_t launch_tree_prec(const MfmaParams& prm, int variant, int cus, hipStream_t s) {
  switch (variant) {
    case 1: return launch_ns<PREC, PROG::NS, 4, 2, true, PROG>(prm, 0, cus, s);
    case 2: return launch_ns<PREC, PROG::NS, 4, 3, false, PROG>(prm, 0, cus, s);
    case 4: return launch_ns<PREC, PROG::NS, 4, 2, false, PROG>(prm, 0, cus, s);
    case 5: return launch_ns<PREC, PROG::NS, 4, 3, true, PROG>(prm, 0, cus, s);
    case 6: return launch_ns<PREC, PROG::NS, 2, 1, false, PROG, 2>(prm, 0, cus, s);
    default: return hipErrorNotSupported;
  }
}

template <class PROG>
hipError_t launch_tree(const MfmaParams& prm, int precision, int variant, int cus,
                       hipStream_t s) {
  switch (precision) {
    case kPrecF32: return launch_tree_prec<kPrecF32, PROG>(prm, variant, cus, s);
    case kPrecF16S: return launch_tree_prec<kPrecF16S, PROG>(prm, variant, cus, s);
    default: return hipErrorNotSupported;
  }
}

}  // namespace pcsf_mfma
