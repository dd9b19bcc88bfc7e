// prune_mfma.hip -- launchers of the fused MFMA pruning kernel
// (prune_mfma.hpp) and its FP32 instantiation.
#include <algorithm>

#include "prune_tree.hpp"

namespace pcsf_mfma {
template hipError_t launch_variant<kPrecF32>(const MfmaParams&, int, int, int, hipStream_t);
}  // namespace pcsf_mfma

using namespace pcsf_mfma;

const MfmaVariant kMfmaVariants[kNumMfmaVariants] = {
    {8, 1, true, 1, "8 waves/WG, gather prefetch"},
    {4, 2, true, 1, "4 waves/WG, >=2 waves/SIMD, gather prefetch"},
    {4, 3, false, 1, "4 waves/WG, >=3 waves/SIMD"},
    {2, 2, true, 1, "2 waves/WG, >=2 waves/SIMD, gather prefetch"},
    {4, 2, false, 1, "4 waves/WG, >=2 waves/SIMD"},
    {4, 3, true, 1, "4 waves/WG, >=3 waves/SIMD, gather prefetch above the deepest GEMMs"},
    {2, 1, false, 2, "2 waves/WG, 2 column tiles per wave"},
};

namespace {
hipError_t dispatch(const MfmaParams& prm, int precision, int variant, int slots, int cus,
                    hipStream_t s) {
  switch (precision) {
    case kPrecF32: return launch_variant<kPrecF32>(prm, variant, slots, cus, s);
    case kPrecF16S: return launch_variant<kPrecF16S>(prm, variant, slots, cus, s);
    case kPrecF16: return launch_variant<kPrecF16>(prm, variant, slots, cus, s);
    case kPrecBF16: return launch_variant<kPrecBF16>(prm, variant, slots, cus, s);
    default: return hipErrorInvalidValue;
  }
}

// The specialised launch when the model's program has one and it covers
// this precision and shape; hipErrorNotSupported otherwise.
hipError_t launch_special(const pcsf_model* m, const MfmaParams& prm, int precision,
                          int variant) {
  if (!m->specialized) return hipErrorNotSupported;
  const void* t = (prm.prog == m->d_prog_t) ? m->tree_t : m->tree_p;
  if (!t) return hipErrorNotSupported;
  return static_cast<const TreeProgram*>(t)->launch(prm, precision, variant, m->ctx->num_cus,
                                                     m->ctx->stream);
}

MfmaParams base_params(const pcsf_model* m, const PSets& sets) {
  MfmaParams prm{};
  prm.nl = m->nl;
  prm.nm = m->nm;
  prm.nbr = m->nbr;
  const bool tables = sets.d_ctab != nullptr && !m->cherries.empty();
  prm.first_gemm_br = tables ? m->first_gemm_br_t : m->first_gemm_br;
  prm.prog = tables ? m->d_prog_t : m->d_prog;
  prm.ctab = tables ? sets.d_ctab : nullptr;
  prm.clnt = tables ? sets.d_clnt : nullptr;
  prm.nctab = (int)m->cherries.size() / 3;
  prm.Pfrag = sets.d_Pfrag;
  prm.PT32 = sets.d_PT32;
  prm.lnPcc = sets.d_lnPcc;
  prm.pi = sets.d_epi ? sets.d_epi : m->d_pi;
  prm.lnpi = sets.d_epi ? sets.d_elnpi : m->d_lnpi;
  prm.prior_per_set = sets.d_epi ? 1 : 0;
  prm.ls_leaves = tables ? m->d_lsl_t : m->d_lsl;
  prm.n_ls = tables ? m->n_lsl_t : m->n_lsl;
  return prm;
}
}  // namespace

const void* find_tree_program(const std::vector<int32_t>& ops, bool tables) {
  int n = 0;
  const TreeProgram* const* all = tree_programs(&n);
  for (int i = 0; i < n; ++i) {
    const TreeProgram* t = all[i];
    if (!t->ops || t->tables != (tables ? 1 : 0) || t->n_words != (int)ops.size()) continue;
    if (std::equal(ops.begin(), ops.end(), t->ops)) return t;
  }
  return nullptr;
}

int frag_format(int precision) {
  switch (precision) {
    case PCSF_FP32: return kFragF32;
    case PCSF_FP16S:
    case PCSF_FP16: return kFragF16;
    case PCSF_BF16: return kFragBF16;
    default: return -1;
  }
}

hipError_t launch_prune_mfma(const pcsf_model* m, const pcsf_batch* b, int precision,
                             int variant, hipStream_t s) {
  if (variant < 0 || variant >= kNumMfmaVariants) return hipErrorInvalidValue;
  if (m->fixed.fmt != frag_format(precision)) return hipErrorInvalidValue;
  const int gc = 16 * kMfmaVariants[variant].wpg * kMfmaVariants[variant].g;
  MfmaParams prm = base_params(m, m->fixed);
  prm.codes = b->d_codes;
  prm.ld = b->ld;
  prm.ngroups = b->ld / gc;
  prm.logz = b->d_logz;
  prm.colelpr = b->d_colelpr;
  if (prm.ngroups == 0) return hipSuccess;
  if (s == m->ctx->stream) {
    const hipError_t e = launch_special(m, prm, precision, variant);
    if (e != hipErrorNotSupported) return e;
  }
  return dispatch(prm, precision, variant, m->slots, m->ctx->num_cus, s);
}

hipError_t launch_prune_mfma_regions(const pcsf_model* m, const pcsf_batch* b, int precision,
                                     int variant, const RegionChunk& ch, hipStream_t s) {
  if (variant < 0 || variant >= kNumMfmaVariants) return hipErrorInvalidValue;
  if (ch.sets->fmt != frag_format(precision)) return hipErrorInvalidValue;
  const int gc = 16 * kMfmaVariants[variant].wpg * kMfmaVariants[variant].g;
  if (gc > kAlignCols) return hipErrorInvalidValue;  // a group must not span regions
  MfmaParams prm = base_params(m, *ch.sets);
  prm.codes = b->al.d_codes;
  prm.ld = b->al.ld;
  prm.g0 = ch.block0 * (kAlignCols / gc);
  prm.ngroups = ch.nblocks * (kAlignCols / gc);
  prm.logz = b->d_alogz;
  prm.colelpr = b->d_acolelpr;
  prm.block_region = b->al.d_block_region;
  prm.active = ch.sets->d_active;
  prm.r0 = ch.r0;
  if (prm.ngroups == 0) return hipSuccess;
  if (s == m->ctx->stream) {
    const hipError_t e = launch_special(m, prm, precision, variant);
    if (e != hipErrorNotSupported) return e;
  }
  return dispatch(prm, precision, variant, m->slots, m->ctx->num_cus, s);
}
