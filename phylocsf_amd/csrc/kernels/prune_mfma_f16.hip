// prune_mfma_f16.hip -- instantiation of the fused MFMA pruning kernel
// (prune_mfma.hpp) for the kPrecF16 GEMM precision.
#include "prune_mfma.hpp"

namespace pcsf_mfma {
template hipError_t launch_variant<kPrecF16>(const MfmaParams&, int, int, int, hipStream_t);
}  // namespace pcsf_mfma
