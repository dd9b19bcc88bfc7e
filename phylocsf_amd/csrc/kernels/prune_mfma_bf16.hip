// prune_mfma_bf16.hip -- instantiation of the fused MFMA pruning kernel
// (prune_mfma.hpp) for the kPrecBF16 GEMM precision.
#include "prune_mfma.hpp"

namespace pcsf_mfma {
template hipError_t launch_variant<kPrecBF16>(const MfmaParams&, int, int, int, hipStream_t);
}  // namespace pcsf_mfma
