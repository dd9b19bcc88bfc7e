// adjoint.hip -- outside (preorder) pass: d ln z / d t_br and d ln z / d rho.
//
// The gradient the reference's PhyloEM defines (lib/CamlPaml/PhyloEM.ml:
// 140-182): d ELL / d t_br = sum_ab xi_br[a,b] (dP/dt)_ab / P_ab with the
// branch posteriors xi of PhyloLik.add_branch_posteriors (PhyloLik.ml:
// 140-174) and dP/dt = Q P(t) (Q.ml:260-275); by the Fisher identity it is
// d ln z / d t_br, and with every branch scaled by rho, d ln z / d rho =
// sum_br b_br d ln z / d t_br (PhyloEM.ml:162-182).  Written without the
// posteriors: for the branch into child c of node p with sibling s,
//   w_c = beta_p o (P_s alpha_s),   d ln z / d t_c = w_c . (Q P_c alpha_c) / z,
//   beta_c = P_c^T w_c            (PhyloLik.ensure_beta, PhyloLik.ml:96-120),
// beta_root = pi.  One 64-lane workgroup per (column, model), FP64, no
// rescaling (as the reference).  The inside pass is recomputed in LDS; in
// reverse index order each internal node's alpha is dead once its parent is
// processed, and its slot then holds beta, so one (n - nl) x 64 LDS array
// serves both passes.
#include "../internal.hpp"

namespace {

__device__ __forceinline__ double message(int c, int nl, int a, const double* __restrict__ PT,
                                          const uint8_t* __restrict__ codes, int64_t ld,
                                          int64_t col, const double* slot) {
  // (P_c v_c)[a]: v_c = leaf vector or alpha_c (PT[b*64 + a] = P[a][b])
  if (c < nl) {
    const int code = codes[(size_t)c * ld + col];
    return (code == PCSF_MARGINALIZE) ? 1.0 : PT[code * 64 + a];
  }
  double acc = 0.0;
  for (int b = 0; b < 64; ++b) acc = fma(PT[b * 64 + a], slot[b], acc);
  return acc;
}

__global__ __launch_bounds__(64) void adjoint_fp64_kernel(
    const uint8_t* __restrict__ codes, int64_t c0, int64_t ncols, int64_t ld, int n, int nl,
    int nm, int nbr, const int32_t* __restrict__ lc, const int32_t* __restrict__ rc,
    const double* __restrict__ P64T, const double* __restrict__ QT,
    const double* __restrict__ pi, const double* __restrict__ branch,
    const int32_t* __restrict__ block_region, const int64_t* __restrict__ astart,
    const int64_t* __restrict__ off, const uint8_t* __restrict__ active, int r0,
    double* __restrict__ logz, double* __restrict__ coldrho,
    double* __restrict__ grad_branch /* nreg*nm*nbr, region-relative to r0 */) {
  extern __shared__ double lds[];
  double* slot = lds;                           // (n - nl) x 64: alpha, then beta
  double* red = lds + (size_t)(n - nl) * 64;    // 64: reduction scratch
  double* vec = red + 64;                       // 2 x 64: messages of the children
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
    const double* Qq = QT + (size_t)q * 4096;  // Qq[k*64 + a] = Q[a][k]
    // inside pass (PhyloLik.ml:78-90)
    for (int i = nl; i < n; ++i) {
      const double ml = message(lc[i], nl, a, Pq + (size_t)lc[i] * 4096, codes, ld, col,
                                slot + (size_t)(lc[i] - nl) * 64);
      const double mr = message(rc[i], nl, a, Pq + (size_t)rc[i] * 4096, codes, ld, col,
                                slot + (size_t)(rc[i] - nl) * 64);
      __syncthreads();
      slot[(size_t)(i - nl) * 64 + a] = ml * mr;
      __syncthreads();
    }
    double* sroot = slot + (size_t)(n - 1 - nl) * 64;
    red[a] = sroot[a] * pi[q * 64 + a];
    __syncthreads();
    double z = 0.0;
    for (int x = 0; x < 64; ++x) z += red[x];
    __syncthreads();
    sroot[a] = pi[q * 64 + a];  // beta_root = prior (PhyloLik.ml:62)
    __syncthreads();
    double drho = 0.0;
    // outside pass, parents before children
    for (int p = n - 1; p >= nl; --p) {
      const int kid[2] = {lc[p], rc[p]};
      const double bp = slot[(size_t)(p - nl) * 64 + a];
      double m[2];
#pragma unroll
      for (int s = 0; s < 2; ++s)
        m[s] = message(kid[s], nl, a, Pq + (size_t)kid[s] * 4096, codes, ld, col,
                       slot + (size_t)(kid[s] - nl) * 64);
      vec[a] = m[0];
      vec[64 + a] = m[1];
      __syncthreads();
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const int c = kid[s];
        const double w = bp * m[1 - s];  // w_c[a]
        // (Q m_c)[a]
        double qm = 0.0;
        for (int k = 0; k < 64; ++k) qm = fma(Qq[k * 64 + a], vec[s * 64 + k], qm);
        red[a] = w * qm;
        __syncthreads();
        double g = 0.0;
        for (int x = 0; x < 64; ++x) g += red[x];
        g = (z != 0.0) ? g / z : 0.0;  // d ln z / d t_c
        drho = fma(branch[c], g, drho);
        if (a == 0 && grad_branch)
          atomicAdd(grad_branch + ((size_t)(r - r0) * nm + q) * nbr + c, g);
        __syncthreads();
        if (c >= nl) {
          // beta_c[b] = sum_a w_c[a] P_c[a][b]  (lane a plays b here)
          red[a] = w;
          __syncthreads();
          const double* PT = Pq + (size_t)c * 4096;
          double bc = 0.0;
          for (int x = 0; x < 64; ++x) bc = fma(red[x], PT[a * 64 + x], bc);
          __syncthreads();
          slot[(size_t)(c - nl) * 64 + a] = bc;
        }
        __syncthreads();
      }
    }
    if (a == 0) {
      logz[(size_t)q * ld + col] = log(z);
      coldrho[(size_t)q * ld + col] = drho;
    }
    __syncthreads();
  }
}

}  // namespace

hipError_t launch_adjoint_fp64(const pcsf_model* m, const pcsf_batch* b, const RegionChunk& ch,
                               double* grad_branch, hipStream_t s) {
  const size_t lds = ((size_t)(m->n - m->nl) * 64 + 64 + 128) * sizeof(double);
  if (lds > 65536) return hipErrorInvalidValue;
  const int64_t c0 = ch.block0 * kAlignCols, ncols = ch.nblocks * kAlignCols;
  const int64_t ntask = ncols * m->nm;
  const int grid = (int)std::min<int64_t>(ntask, (int64_t)m->ctx->num_cus * 32);
  if (grid <= 0) return hipSuccess;
  hipLaunchKernelGGL(adjoint_fp64_kernel, dim3(grid), dim3(64), lds, s, b->al.d_codes, c0, ncols,
                     b->al.ld, m->n, m->nl, m->nm, m->nbr, m->d_lc, m->d_rc, ch.sets->d_P64T,
                     m->d_QT, m->d_pi, m->d_branch, b->al.d_block_region, b->al.d_astart,
                     b->d_offsets, ch.sets->d_active, ch.r0, b->d_alogz, b->d_acolelpr,
                     grad_branch);
  return hipGetLastError();
}
