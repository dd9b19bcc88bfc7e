// lik.hip -- the CamlPaml.PhyloLik compatibility path on the GPU, FP64, for
// any alphabet size k <= 64 and caller-supplied P matrices:
//
//   lik_columns_kernel   PhyloLik.prepare (with `Certain, `Distribution and
//                        `Marginalize leaves), ensure_alpha / likelihood,
//                        ensure_beta and node_posterior of every node
//                        (lib/CamlPaml/PhyloLik.ml:9-138), one column per
//                        64-lane workgroup; and, for the E-step, the two
//                        vectors of every branch's posterior
//                        xi_br[a][b] = u[a] P_br[a][b] v[b] with
//                        u = w beta_parent o (P_sibling alpha_sibling) / z and
//                        v = alpha_br (PhyloLik.ml:140-174).
//   branch_counts_kernel the sum over a chunk of columns of u v^T per branch,
//                        then ecounts_br += P_br o that sum: the expected
//                        substitution counts of PhyloEM.collect_sufficient_
//                        statistics (lib/CamlPaml/PhyloEM.ml:14-20) as one
//                        rank-C update per branch instead of k^2 updates per
//                        column.
//
// Lane a owns state a.  alpha (rows 0..nl-1: the leaf vectors) and beta live
// in a global workspace per workgroup, so the tree size is not bounded by LDS.
// With the E-step factors requested (U non-null) every workgroup takes one
// column per launch: U and V hold one slot per workgroup and branch.
// Paths are relative to the mlin/PhyloCSF tree.
#include "../internal.hpp"

namespace {

__global__ __launch_bounds__(64) void lik_columns_kernel(
    int k, int n, const int32_t* __restrict__ lc, const int32_t* __restrict__ rc,
    const int32_t* __restrict__ parent, const int32_t* __restrict__ sib,
    const double* __restrict__ P, const double* __restrict__ prior,
    const int32_t* __restrict__ leaves, const double* __restrict__ leaf_dist, int64_t ncols,
    const double* __restrict__ weights, double* __restrict__ ws, double* __restrict__ zout,
    int post_mode, double* __restrict__ post, double* __restrict__ U, double* __restrict__ V) {
  __shared__ double inter[64];
  const int a = threadIdx.x;
  const int nl = (n + 1) / 2;
  const bool on = a < k;
  double* alpha = ws + (size_t)blockIdx.x * 2 * n * k;  // n x k
  double* beta = alpha + (size_t)n * k;                 // n x k
  for (int64_t col = blockIdx.x; col < ncols; col += gridDim.x) {
    // get_leaf (PhyloLik.ml:16-19): e_i, the given distribution, or ones
    if (on)
      for (int j = 0; j < nl; ++j) {
        const int c = leaves[col * nl + j];
        alpha[(size_t)j * k + a] =
            (c == -2) ? leaf_dist[((size_t)col * nl + j) * k + a] : (c < 0 || c == a) ? 1.0 : 0.0;
      }
    __syncthreads();
    // ensure_alpha (PhyloLik.ml:73-93)
    for (int i = nl; i < n; ++i) {
      if (on) {
        const double* Pl = P + (size_t)lc[i] * k * k + (size_t)a * k;
        const double* Pr = P + (size_t)rc[i] * k * k + (size_t)a * k;
        const double* al = alpha + (size_t)lc[i] * k;
        const double* ar = alpha + (size_t)rc[i] * k;
        double x = 0.0, y = 0.0;
        for (int b = 0; b < k; ++b) x = fma(Pl[b], al[b], x);
        for (int b = 0; b < k; ++b) y = fma(Pr[b], ar[b], y);
        alpha[(size_t)i * k + a] = x * y;
      }
      __syncthreads();
    }
    double z = 0.0;
    {
      const double* aroot = alpha + (size_t)(n - 1) * k;
      for (int x = 0; x < k; ++x) z = fma(aroot[x], prior[x], z);
    }
    if (a == 0) zout[col] = z;
    if (post_mode == 1 && on)  // node_posterior at the root: beta_root = prior
      post[(size_t)col * k + a] = (z == 0.0) ? 0.0 : alpha[(size_t)(n - 1) * k + a] * prior[a] / z;
    if (post_mode == 2 || U) {
      // ensure_beta (PhyloLik.ml:96-120)
      if (on) beta[(size_t)(n - 1) * k + a] = prior[a];
      __syncthreads();
      const double w = weights ? weights[col] : 1.0;
      for (int i = n - 2; i >= 0; --i) {
        const int p = parent[i], s = sib[i];
        if (on) {
          const double* Ps = P + (size_t)s * k * k + (size_t)a * k;
          const double* as = alpha + (size_t)s * k;
          double d = 0.0;
          for (int b = 0; b < k; ++b) d = fma(Ps[b], as[b], d);
          inter[a] = beta[(size_t)p * k + a] * d;
        }
        __syncthreads();
        if (on) {
          const double* Pi = P + (size_t)i * k * k;
          double bt = 0.0;
          for (int x = 0; x < k; ++x) bt = fma(inter[x], Pi[(size_t)x * k + a], bt);
          beta[(size_t)i * k + a] = bt;
          if (U) {
            // this column's factors of xi_i (add_branch_posteriors): zero
            // where beta_parent[a] <= 0 or z <= 0, as the reference skips them
            const size_t o = ((size_t)i * gridDim.x + blockIdx.x) * k + a;
            const bool live = z > 0.0 && beta[(size_t)p * k + a] > 0.0;
            U[o] = live ? w * inter[a] / z : 0.0;
            V[o] = alpha[(size_t)i * k + a];
          }
        }
        __syncthreads();
      }
      if (post_mode == 2 && on)  // node_posterior (PhyloLik.ml:127-138)
        for (int i = 0; i < n; ++i) {
          double v = 0.0;
          if (z != 0.0)
            v = (i < nl) ? alpha[(size_t)i * k + a]
                         : alpha[(size_t)i * k + a] * beta[(size_t)i * k + a] / z;
          post[((size_t)col * n + i) * k + a] = v;
        }
    }
    __syncthreads();
  }
}

// ecounts_br[a][b] += P_br[a][b] sum_j U[br][j][a] V[br][j][b] over the
// nj column slots of one pass; one 256-thread workgroup per branch, each
// thread 16 outputs (row a = t / 4, columns 16 (t % 4) ..), the slots staged
// 16 at a time in LDS.
__global__ __launch_bounds__(256) void branch_counts_kernel(int k, int nj,
                                                            const double* __restrict__ U,
                                                            const double* __restrict__ V,
                                                            const double* __restrict__ P,
                                                            double* __restrict__ ecounts) {
  __shared__ double su[16][64], sv[16][64];
  const int br = blockIdx.x, t = threadIdx.x;
  const int a = t >> 2, b0 = 16 * (t & 3);
  double acc[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) acc[j] = 0.0;
  const double* Ub = U + (size_t)br * nj * k;
  const double* Vb = V + (size_t)br * nj * k;
  for (int j0 = 0; j0 < nj; j0 += 16) {
    for (int e = t; e < 16 * 64; e += 256) {
      const int jj = e >> 6, x = e & 63;
      const bool in = j0 + jj < nj && x < k;
      su[jj][x] = in ? Ub[(size_t)(j0 + jj) * k + x] : 0.0;
      sv[jj][x] = in ? Vb[(size_t)(j0 + jj) * k + x] : 0.0;
    }
    __syncthreads();
#pragma unroll 4
    for (int jj = 0; jj < 16; ++jj) {
      const double u = su[jj][a];
#pragma unroll
      for (int j = 0; j < 16; ++j) acc[j] = fma(u, sv[jj][b0 + j], acc[j]);
    }
    __syncthreads();
  }
  if (a < k)
#pragma unroll
    for (int j = 0; j < 16; ++j)
      if (b0 + j < k) {
        const size_t o = ((size_t)br * k + a) * k + b0 + j;
        ecounts[o] += P[o] * acc[j];
      }
}

}  // namespace

hipError_t launch_lik_posteriors(const LikArgs& x, int grid, hipStream_t s) {
  if (x.ncols <= 0) return hipSuccess;
  hipLaunchKernelGGL(lik_columns_kernel, dim3(grid), dim3(64), 0, s, x.k, x.n, x.lc, x.rc,
                     x.parent, x.sib, x.P, x.prior, x.leaves, x.leaf_dist, x.ncols, x.weights,
                     x.ws, x.z, x.post_mode, x.post, x.U, x.V);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || !x.U) return e;
  hipLaunchKernelGGL(branch_counts_kernel, dim3(x.n - 1), dim3(256), 0, s, x.k, grid, x.U, x.V,
                     x.P, x.ecounts);
  return hipGetLastError();
}
