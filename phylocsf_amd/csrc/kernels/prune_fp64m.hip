// prune_fp64m.hip -- launcher of the FP64 forward on matrix cores
// (prune_fp64m.hpp): the tree's generated program when it has one, the
// op-program interpreter otherwise.
#include "prune_tree.hpp"

using pcsf_mfma::Interp;
using pcsf_mfma::MfmaParams;

hipError_t launch_forward_fp64m(const pcsf_model* m, const pcsf_batch* b, const RegionChunk& ch,
                                hipStream_t s) {
  if (!ch.sets->d_P64F) return hipErrorInvalidValue;
  MfmaParams prm{};
  prm.codes = b->al.d_codes;
  prm.ld = b->al.ld;
  prm.g0 = ch.block0;  // one group of kGC = kAlignCols columns per block
  prm.ngroups = ch.nblocks;
  prm.nl = m->nl;
  prm.nm = m->nm;
  prm.nbr = m->nbr;
  prm.first_gemm_br = m->first_gemm_br;
  prm.prog = m->d_prog;
  prm.pi = ch.sets->d_epi ? ch.sets->d_epi : m->d_pi;
  prm.lnpi = ch.sets->d_epi ? ch.sets->d_elnpi : m->d_lnpi;
  prm.prior_per_set = ch.sets->d_epi ? 1 : 0;
  prm.logz = b->d_alogz;
  prm.colelpr = b->d_acolelpr;
  prm.block_region = b->al.d_block_region;
  prm.active = ch.sets->d_active;
  prm.r0 = ch.r0;
  prm.shared_sets = ch.sets == &m->fixed ? 1 : 0;
  prm.P64F = ch.sets->d_P64F;
  if (prm.ngroups == 0) return hipSuccess;
  const int cus = m->ctx->num_cus;
  if (m->specialized && m->tree_p) {
    const auto* t = static_cast<const pcsf_mfma::TreeProgram*>(m->tree_p);
    if (t->fp64m) return t->fp64m(prm, 0, 0, cus, s);
  }
  using pcsf_mfma::f64m::launch_ns;
  switch (m->slots) {
    case 1:
    case 2: return launch_ns<2, Interp<2>>(prm, cus, s);
    case 3: return launch_ns<3, Interp<3>>(prm, cus, s);
    case 4: return launch_ns<4, Interp<4>, 1>(prm, cus, s);
    case 5: return launch_ns<5, Interp<5>, 1>(prm, cus, s);
    case 6: return launch_ns<6, Interp<6>, 1>(prm, cus, s);
    default: return hipErrorInvalidValue;
  }
}
