// prune_tangent.hip -- launcher of the tangent (forward-mode) pass
// (prune_tangent.hpp): the tree-specialised kernel when the model's program
// has one, the interpreter otherwise.
#include "prune_tree.hpp"

using namespace pcsf_mfma;

hipError_t launch_prune_tangent_regions(const pcsf_model* m, const pcsf_batch* b,
                                        const RegionChunk& ch, hipStream_t s) {
  const PSets& sets = *ch.sets;
  if (!sets.d_Dfrag || !sets.d_DT32 || sets.fmt != kFragF32) return hipErrorInvalidValue;
  constexpr int gc = 16 * kTWpg;
  static_assert(gc <= kAlignCols, "a group must not span regions");
  MfmaParams prm{};
  prm.nl = m->nl;
  prm.nm = m->nm;
  prm.nbr = m->nbr;
  prm.first_gemm_br = m->first_gemm_br;
  prm.prog = m->d_prog;
  prm.Pfrag = sets.d_Pfrag;
  prm.PT32 = sets.d_PT32;
  prm.lnPcc = sets.d_lnPcc;
  prm.Dfrag = sets.d_Dfrag;
  prm.DT32 = sets.d_DT32;
  prm.pi = m->d_pi;
  prm.lnpi = m->d_lnpi;
  prm.codes = b->al.d_codes;
  prm.ld = b->al.ld;
  prm.g0 = ch.block0 * (kAlignCols / gc);
  prm.ngroups = ch.nblocks * (kAlignCols / gc);
  prm.logz = b->d_alogz;
  prm.colelpr = b->d_acolelpr;
  prm.block_region = b->al.d_block_region;
  prm.active = sets.d_active;
  prm.r0 = ch.r0;
  if (prm.ngroups == 0) return hipSuccess;
  if (m->specialized && m->tree_p) {
    const TreeProgram* t = static_cast<const TreeProgram*>(m->tree_p);
    if (t->tangent) return t->tangent(prm, PCSF_FP32, 0, m->ctx->num_cus, s);
  }
  switch (m->slots) {
    case 1:
    case 2: return launch_tangent_ns<2>(prm, m->ctx->num_cus, s);
    case 3: return launch_tangent_ns<3>(prm, m->ctx->num_cus, s);
    case 4: return launch_tangent_ns<4>(prm, m->ctx->num_cus, s);
    case 5: return launch_tangent_ns<5>(prm, m->ctx->num_cus, s);
    case 6: return launch_tangent_ns<6>(prm, m->ctx->num_cus, s);
    default: return hipErrorInvalidValue;
  }
}
