// prune_fp64.hip -- plain FP64 Felsenstein pruning, one column per workgroup.
//
// The straightforward GPU form of PhyloLik.ensure_alpha
// (lib/CamlPaml/PhyloLik.ml:73-93) as PhyloCSFModel.lpr_leaves drives it
// (src/PhyloCSFModel.ml:67-82): a 64-lane workgroup takes one
// (column, model), lane a owns codon row a, and internal nodes are visited
// in the reference's index (post) order with every alpha kept in LDS like
// the reference's workspace (PhyloLik.ml:21-24,41-43).  Each internal-child
// message is a 64-long FP64 dot per lane against P^T read coalesced from HBM;
// leaf children gather a column of P (one-hot leaf) or are 1 (`Marginalize).
// No rescaling, as in the reference.  z = alpha_root . pi and the root
// posterior's elpr are summed sequentially by lane 0 (PhyloLik.ml:92,
// PhyloCSFModel.ml:80-81).
#include "../internal.hpp"

namespace {

__global__ __launch_bounds__(64) void prune_fp64_kernel(
    const uint8_t* __restrict__ codes, int64_t ncols, int64_t ld, int n, int nl, int nm,
    int nbr, const int32_t* __restrict__ lc, const int32_t* __restrict__ rc,
    const double* __restrict__ P64T, const double* __restrict__ pi,
    const double* __restrict__ lnpi, double* __restrict__ logz,
    double* __restrict__ colelpr) {
  extern __shared__ double alpha[];  // (n - nl) x 64
  const int a = threadIdx.x;
  const int64_t ntask = ncols * nm;
  for (int64_t task = blockIdx.x; task < ntask; task += gridDim.x) {
    const int64_t col = task / nm;
    const int q = (int)(task - col * nm);
    const double* Pq = P64T + (size_t)q * nbr * 4096;
    for (int i = nl; i < n; ++i) {
      double msg[2];
      const int kid[2] = {lc[i], rc[i]};
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const int c = kid[s];
        const double* PT = Pq + (size_t)c * 4096;  // PT[b*64 + a] = P[a][b]
        if (c < nl) {
          const int code = codes[(size_t)c * ld + col];
          msg[s] = (code == PCSF_MARGINALIZE) ? 1.0 : PT[code * 64 + a];
        } else {
          const double* al = alpha + (size_t)(c - nl) * 64;
          double acc = 0.0;
          for (int b = 0; b < 64; ++b) acc = fma(PT[b * 64 + a], al[b], acc);
          msg[s] = acc;
        }
      }
      alpha[(size_t)(i - nl) * 64 + a] = msg[0] * msg[1];
      __syncthreads();
    }
    if (a == 0) {
      const double* ar = alpha + (size_t)(n - 1 - nl) * 64;
      const double* p = pi + q * 64;
      const double* pl = lnpi + q * 64;
      double z = 0.0;
      for (int x = 0; x < 64; ++x) z += ar[x] * p[x];
      double e = 0.0;
      if (z != 0.0)
        for (int x = 0; x < 64; ++x) e += (ar[x] * p[x] / z) * pl[x];
      logz[(size_t)q * ld + col] = log(z);
      colelpr[(size_t)q * ld + col] = e;
    }
    __syncthreads();
  }
}

// Per-region tree scales: aligned columns [c0, c0 + ncols) of a chunk whose
// regions start at r0; set (r - r0) * nm + q holds P at that region's scale.
__global__ __launch_bounds__(64) void prune_fp64_regions_kernel(
    const uint8_t* __restrict__ codes, int64_t c0, int64_t ncols, int64_t ld, int n, int nl,
    int nm, int nbr, const int32_t* __restrict__ lc, const int32_t* __restrict__ rc,
    const double* __restrict__ P64T, const double* __restrict__ pi,
    const double* __restrict__ lnpi, const int32_t* __restrict__ block_region,
    const int64_t* __restrict__ astart, const int64_t* __restrict__ off,
    const uint8_t* __restrict__ active, int r0, int prior_per_set,
    double* __restrict__ logz, double* __restrict__ colelpr) {
  extern __shared__ double alpha[];  // (n - nl) x 64
  const int a = threadIdx.x;
  const int64_t ntask = ncols * nm;
  for (int64_t task = blockIdx.x; task < ntask; task += gridDim.x) {
    const int64_t col = c0 + task / nm;
    const int q = (int)(task % nm);
    const int r = block_region[col / kAlignCols];
    if (r < 0 || col - astart[r] >= off[r + 1] - off[r]) continue;
    const int set = (r - r0) * nm + q;
    if (!active[set]) continue;
    const double* Pq = P64T + (size_t)set * nbr * 4096;
    for (int i = nl; i < n; ++i) {
      double msg[2];
      const int kid[2] = {lc[i], rc[i]};
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const int c = kid[s];
        const double* PT = Pq + (size_t)c * 4096;
        if (c < nl) {
          const int code = codes[(size_t)c * ld + col];
          msg[s] = (code == PCSF_MARGINALIZE) ? 1.0 : PT[code * 64 + a];
        } else {
          const double* al = alpha + (size_t)(c - nl) * 64;
          double acc = 0.0;
          for (int b = 0; b < 64; ++b) acc = fma(PT[b * 64 + a], al[b], acc);
          msg[s] = acc;
        }
      }
      alpha[(size_t)(i - nl) * 64 + a] = msg[0] * msg[1];
      __syncthreads();
    }
    if (a == 0) {
      const double* ar = alpha + (size_t)(n - 1 - nl) * 64;
      const int pq = prior_per_set ? set : q;  // one prior per set, or per model
      const double* p = pi + pq * 64;
      const double* pl = lnpi + pq * 64;
      double z = 0.0;
      for (int x = 0; x < 64; ++x) z += ar[x] * p[x];
      double e = 0.0;
      if (z != 0.0)
        for (int x = 0; x < 64; ++x) e += (ar[x] * p[x] / z) * pl[x];
      logz[(size_t)q * ld + col] = log(z);
      colelpr[(size_t)q * ld + col] = e;
    }
    __syncthreads();
  }
}

}  // namespace

hipError_t launch_prune_fp64(const pcsf_model* m, const pcsf_batch* b, hipStream_t s) {
  const size_t lds = (size_t)(m->n - m->nl) * 64 * sizeof(double);
  if (lds > 65536) return hipErrorInvalidValue;
  const int64_t ntask = b->ncols * m->nm;
  const int grid = (int)std::min<int64_t>(ntask, (int64_t)m->ctx->num_cus * 32);
  if (grid <= 0) return hipSuccess;
  hipLaunchKernelGGL(prune_fp64_kernel, dim3(grid), dim3(64), lds, s, b->d_codes, b->ncols,
                     b->ld, m->n, m->nl, m->nm, m->nbr, m->d_lc, m->d_rc, m->fixed.d_P64T, m->d_pi,
                     m->d_lnpi, b->d_logz, b->d_colelpr);
  return hipGetLastError();
}

hipError_t launch_prune_fp64_regions(const pcsf_model* m, const pcsf_batch* b,
                                     const RegionChunk& ch, hipStream_t s) {
  const size_t lds = (size_t)(m->n - m->nl) * 64 * sizeof(double);
  if (lds > 65536) return hipErrorInvalidValue;
  const int64_t c0 = ch.block0 * kAlignCols, ncols = ch.nblocks * kAlignCols;
  const int64_t ntask = ncols * m->nm;
  const int grid = (int)std::min<int64_t>(ntask, (int64_t)m->ctx->num_cus * 32);
  if (grid <= 0) return hipSuccess;
  hipLaunchKernelGGL(prune_fp64_regions_kernel, dim3(grid), dim3(64), lds, s, b->al.d_codes, c0,
                     ncols, b->al.ld, m->n, m->nl, m->nm, m->nbr, m->d_lc, m->d_rc,
                     ch.sets->d_P64T, ch.sets->d_epi ? ch.sets->d_epi : m->d_pi,
                     ch.sets->d_epi ? ch.sets->d_elnpi : m->d_lnpi, b->al.d_block_region,
                     b->al.d_astart, b->d_offsets, ch.sets->d_active, ch.r0,
                     ch.sets->d_epi ? 1 : 0, b->d_alogz, b->d_acolelpr);
  return hipGetLastError();
}
