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
    const uint8_t* __restrict__ active, int r0, double* __restrict__ logz,
    double* __restrict__ colelpr) {
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

// Generic alphabet (k <= 64) PhyloLik.prepare/likelihood/node_posterior root
// for caller-supplied row-major P (the CamlPaml compatibility entry).
__global__ __launch_bounds__(64) void lik_columns_kernel(
    int k, int n, const int32_t* __restrict__ lc, const int32_t* __restrict__ rc,
    const double* __restrict__ P, const double* __restrict__ prior,
    const int32_t* __restrict__ leaves, int64_t ncols, double* __restrict__ zout,
    double* __restrict__ post) {
  extern __shared__ double alpha[];  // n x k (leaf vectors too)
  const int a = threadIdx.x;
  const int nl = (n + 1) / 2;
  for (int64_t col = blockIdx.x; col < ncols; col += gridDim.x) {
    if (a < k)
      for (int j = 0; j < nl; ++j) {
        const int c = leaves[col * nl + j];
        alpha[j * k + a] = (c < 0) ? 1.0 : (c == a ? 1.0 : 0.0);
      }
    __syncthreads();
    for (int i = nl; i < n; ++i) {
      if (a < k) {
        const double* Pl = P + (size_t)lc[i] * k * k + (size_t)a * k;
        const double* Pr = P + (size_t)rc[i] * k * k + (size_t)a * k;
        const double* al = alpha + (size_t)lc[i] * k;
        const double* ar = alpha + (size_t)rc[i] * k;
        double x = 0.0, y = 0.0;
        for (int b = 0; b < k; ++b) x += Pl[b] * al[b];
        for (int b = 0; b < k; ++b) y += Pr[b] * ar[b];
        alpha[(size_t)i * k + a] = x * y;
      }
      __syncthreads();
    }
    const double* aroot = alpha + (size_t)(n - 1) * k;
    double z = 0.0;
    for (int x = 0; x < k; ++x) z += aroot[x] * prior[x];
    if (a == 0) zout[col] = z;
    if (post && a < k) post[col * k + a] = (z == 0.0) ? 0.0 : aroot[a] * prior[a] / z;
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
                     ch.sets->d_P64T, m->d_pi, m->d_lnpi, b->al.d_block_region, b->al.d_astart,
                     b->d_offsets, ch.sets->d_active, ch.r0, b->d_alogz, b->d_acolelpr);
  return hipGetLastError();
}

hipError_t launch_lik_columns(int k, int n, const int32_t* d_lc, const int32_t* d_rc,
                              const double* d_P, const double* d_prior, const int32_t* d_leaves,
                              int64_t ncols, double* d_z, double* d_post, hipStream_t s) {
  const size_t lds = (size_t)n * k * sizeof(double);
  if (lds > 65536) return hipErrorInvalidValue;
  const int grid = (int)std::min<int64_t>(ncols, 8192);
  if (grid <= 0) return hipSuccess;
  hipLaunchKernelGGL(lik_columns_kernel, dim3(grid), dim3(64), lds, s, k, n, d_lc, d_rc, d_P,
                     d_prior, d_leaves, ncols, d_z, d_post);
  return hipGetLastError();
}
