// estep.hip -- PhyloEM's expected substitution counts on FP64 matrix cores
// (collect_sufficient_statistics, lib/CamlPaml/PhyloEM.ml:14-20, over
// PhyloLik.add_branch_posteriors, lib/CamlPaml/PhyloLik.ml:140-174) for the
// 64-codon alphabet with `Certain and `Marginalize leaves (and states >= 64,
// the reference's vector of zeros).  Paths are relative to the mlin/PhyloCSF
// tree.
//
// estep_passes_kernel: per 64-column block (four waves of 16 columns, the
//   tiles of fp64_tiles.hpp), the inside pass (ensure_alpha, PhyloLik.ml:
//   73-93) with every internal node's message to its parent in a per-wave
//   device tile, then the outside pass (ensure_beta, :96-120) from the root:
//   for the branch into child c of p with sibling s,
//     u_c = w beta_p o m_s / z   (zero where beta_p <= 0 or z <= 0, the
//                                 entries the reference skips)
//     v_c = alpha_c              (internal c; a leaf's v is its leaf
//                                 vector, rebuilt from the leaf's state by
//                                 the counts kernel)
//     beta_c = P_c^T (beta_p o m_s),
//   u and v written per (branch, column) for the next kernel, z per column.
// estep_counts_kernel: per branch, S = sum_col u v^T as a 64 x 64 product
//   over the chunk's columns on v_mfma_f64_16x16x4f64 (K = columns), then
//   ecounts_br += P_br o S -- the branch posteriors xi = diag(u) P diag(v)
//   summed over columns (:140-174) without k^2 updates per column.
#include "fp64_tiles.hpp"

namespace {

using namespace fp64t;

// m = P_c x for the leaf vector x of `code`: column code of P_c (P row-major),
// the row sums for `Marginalize (x = ones; code -1), zeros for a state >= 64
__device__ __forceinline__ Tile leaf_message(const double* __restrict__ Pc,
                                             const double* __restrict__ rowsum, int code,
                                             int grp) {
  Tile t;
#pragma unroll
  for (int mo = 0; mo < 4; ++mo)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int row = 16 * mo + 4 * v + grp;
      t.r[mo][v] = (code >= 0 && code < 64) ? Pc[row * 64 + code]
                   : (code < 0)             ? rowsum[row]
                                            : 0.0;
    }
  return t;
}

// this lane's part of a column tile into U or V at (branch, column)
__device__ __forceinline__ void put_column(double* __restrict__ dst, const Tile& t, int grp) {
#pragma unroll
  for (int mo = 0; mo < 4; ++mo)
#pragma unroll
    for (int v = 0; v < 4; ++v) dst[16 * mo + 4 * v + grp] = t.r[mo][v];
}

__global__ __launch_bounds__(64 * kWaves, 2) void estep_passes_kernel(EstepArgs p) {
  // P_c or P_c^T in A-fragment order, double-buffered: product k of a block
  // (p.seq[k]) reads buffer k & 1 while product k + 1's tile is copied into
  // the other
  __shared__ __attribute__((aligned(16))) double Af[2][4096];
  const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63,
            grp = lane >> 4;
  const int j = lane & 15;
  const int n = p.n, nl = p.nl;
  double* slots = p.scratch + ((size_t)blockIdx.x * kWaves + wave) * (n - nl) * kTileDoubles;
  const int64_t nblocks = (p.ncols + kCols - 1) / kCols;
  auto issue = [&](int k) {  // product k's tile into buffer k & 1
    if (k >= p.nseq) return;
    const int code = p.seq[k];
    const double* src = (code >= 0) ? p.Pf + (size_t)code * 4096 : p.PTf + (size_t)(-1 - code) * 4096;
    copy_tile((L3*)&Af[k & 1][0], src, wave, tid);
  };
  for (int64_t blk = blockIdx.x; blk < nblocks; blk += gridDim.x) {
    __syncthreads();  // the previous block is done with both buffers
    issue(0);
    int kp = 0;  // products done in this block
    const int64_t col = blk * kCols + wave * 16 + j;
    const bool valid = col < p.ncols;
    const int64_t ccol = valid ? col : p.ncols - 1;  // a readable column for padding lanes
    const int32_t* codes = p.leaves + (size_t)ccol * nl;
    auto msg = [&](int c) -> Tile {
      return (c < nl) ? leaf_message(p.P + (size_t)c * 4096, p.rowsum + (size_t)c * 64, codes[c],
                                     grp)
                      : load(slots + (size_t)(c - nl) * kTileDoubles, lane);
    };
    // ---- inside pass (PhyloLik.ml:73-93)
    double z = 0.0;
    for (int i = nl; i < n; ++i) {
      Tile a = msg(p.lc[i]);
      const Tile mr = msg(p.rc[i]);
#pragma unroll
      for (int mo = 0; mo < 4; ++mo) a.r[mo] *= mr.r[mo];  // alpha_i
      if (i == n - 1) {
        Tile pr;
#pragma unroll
        for (int mo = 0; mo < 4; ++mo)
#pragma unroll
          for (int v = 0; v < 4; ++v) pr.r[mo][v] = p.prior[16 * mo + 4 * v + grp];
        z = column_dot(a, pr);
        break;
      }
      __syncthreads();  // P_i landed; the other buffer is no longer read
      issue(kp + 1);
      store(slots + (size_t)(i - nl) * kTileDoubles, gemm(Af[kp & 1], a, lane), lane);
      ++kp;
    }
    if (valid && grp == 0) p.z[col] = z;
    const double w = valid ? (p.weights ? p.weights[col] : 1.0) : 0.0;
    const double scale = (valid && z > 0.0) ? w / z : 0.0;
    // ---- outside pass (PhyloLik.ml:96-120), parents before children
    for (int pn = n - 1; pn >= nl; --pn) {
      const int kid[2] = {p.lc[pn], p.rc[pn]};
      Tile m[2];
#pragma unroll
      for (int s = 0; s < 2; ++s) m[s] = msg(kid[s]);
      Tile bp;
      if (pn == n - 1) {
#pragma unroll
        for (int mo = 0; mo < 4; ++mo)
#pragma unroll
          for (int v = 0; v < 4; ++v) bp.r[mo][v] = p.prior[16 * mo + 4 * v + grp];
      } else {
        bp = load(slots + (size_t)(pn - nl) * kTileDoubles, lane);
        Tile al;  // alpha_pn: v of the branch above pn
#pragma unroll
        for (int mo = 0; mo < 4; ++mo)
          al.r[mo] = valid ? m[0].r[mo] * m[1].r[mo] : (f64x4){0.0, 0.0, 0.0, 0.0};
        put_column(p.V + ((size_t)pn * p.ldu + blk * kCols + wave * 16 + j) * 64, al, grp);
      }
      Tile wt[2];  // beta_p o m_sibling
#pragma unroll
      for (int mo = 0; mo < 4; ++mo) {
        wt[0].r[mo] = bp.r[mo] * m[1].r[mo];
        wt[1].r[mo] = bp.r[mo] * m[0].r[mo];
      }
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const int c = kid[s];
        const size_t o = ((size_t)c * p.ldu + blk * kCols + wave * 16 + j) * 64;
        Tile u;
#pragma unroll
        for (int mo = 0; mo < 4; ++mo)
#pragma unroll
          for (int v = 0; v < 4; ++v)
            u.r[mo][v] = (bp.r[mo][v] > 0.0) ? wt[s].r[mo][v] * scale : 0.0;
        put_column(p.U + o, u, grp);
        if (c >= nl) {  // beta_c = P_c^T (beta_p o m_s) into c's slot
          // (a leaf branch's v is its leaf vector, which the counts kernel
          // rebuilds from the codes)
          __syncthreads();  // P_c^T landed; the other buffer is no longer read
          issue(kp + 1);
          store(slots + (size_t)(c - nl) * kTileDoubles, gemm(Af[kp & 1], wt[s], lane), lane);
          ++kp;
        }
      }
    }
  }
}

// ecounts_br += P_br o sum_col u v^T over columns [0, ncols) (ncols a
// multiple of 4; U and V hold ldu columns per branch; for a leaf branch
// br < nl, v is the leaf vector of the column's state in leaves: one-hot,
// ones for `Marginalize, zeros for a state >= 64 or a column past nvalid).
// Workgroup
// (branch, split) takes the split's share of the columns -- wave w the rows
// 16 w .. 16 w + 15 and all 64 columns (four 16 x 16 accumulators) -- and
// adds P o its partial sum with FP64 atomics (P o S is linear in S).
__global__ __launch_bounds__(256) void estep_counts_kernel(const double* __restrict__ U,
                                                           const double* __restrict__ V,
                                                           int64_t ldu, int64_t ncols,
                                                           const int32_t* __restrict__ leaves,
                                                           int nl, int64_t nvalid,
                                                           const double* __restrict__ P,
                                                           double* __restrict__ ecounts) {
  const int br = blockIdx.x, tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
  const int64_t steps = ncols / 4, per = (steps + gridDim.y - 1) / gridDim.y;
  const int64_t jb = 4 * per * blockIdx.y, je = std::min<int64_t>(ncols, jb + 4 * per);
  const int i = lane & 15, kk = lane >> 4;
  const double* Ub = U + (size_t)br * ldu * 64;
  const double* Vb = V + (size_t)br * ldu * 64;
  f64x4 acc[4];
#pragma unroll
  for (int no = 0; no < 4; ++no) acc[no] = (f64x4){0.0, 0.0, 0.0, 0.0};
  if (br < nl) {  // (uniform over the workgroup)
    for (int64_t j0 = jb; j0 < je; j0 += 4) {
      const double a = Ub[(j0 + kk) * 64 + 16 * w + i];  // A[16 w + i][j0 + kk]
      const int code = (j0 + kk < nvalid) ? leaves[(size_t)(j0 + kk) * nl + br] : 64;
#pragma unroll
      for (int no = 0; no < 4; ++no) {
        const double v = (code < 0 || code == 16 * no + i) ? 1.0 : 0.0;
        acc[no] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, v, acc[no], 0, 0, 0);
      }
    }
  } else {
    for (int64_t j0 = jb; j0 < je; j0 += 4) {
      const double a = Ub[(j0 + kk) * 64 + 16 * w + i];  // A[16 w + i][j0 + kk]
#pragma unroll
      for (int no = 0; no < 4; ++no)
        acc[no] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, Vb[(j0 + kk) * 64 + 16 * no + i],
                                                       acc[no], 0, 0, 0);
    }
  }
  // acc[no][v] of lane i + 16 g: S[16 w + 4 v + g][16 no + i]
  const double* Pb = P + (size_t)br * 4096;
  double* E = ecounts + (size_t)br * 4096;
#pragma unroll
  for (int no = 0; no < 4; ++no)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int a = 16 * w + 4 * v + kk, b = 16 * no + i;
      atomicAdd(E + a * 64 + b, Pb[a * 64 + b] * acc[no][v]);
    }
}

// P and P^T of every branch into the A-fragment order of fp64_tiles.hpp:
// M[row][col] at ((row / 16) * 16 + col / 4) * 64 + row % 16 + 16 (col % 4)
__global__ __launch_bounds__(256) void estep_frags_kernel(const double* __restrict__ P,
                                                          double* __restrict__ Pf,
                                                          double* __restrict__ PTf) {
  const int br = blockIdx.x;
  const double* src = P + (size_t)br * 4096;
  for (int e = threadIdx.x; e < 4096; e += 256) {
    const int row = e >> 6, col = e & 63;
    const int f = ((row >> 4) * 16 + (col >> 2)) * 64 + (row & 15) + 16 * (col & 3);
    Pf[(size_t)br * 4096 + f] = src[row * 64 + col];   // M = P
    PTf[(size_t)br * 4096 + f] = src[col * 64 + row];  // M = P^T
  }
}

}  // namespace

hipError_t launch_estep_frags(const double* P, int nbr, double* Pf, double* PTf, hipStream_t s) {
  if (nbr <= 0) return hipSuccess;
  hipLaunchKernelGGL(estep_frags_kernel, dim3(nbr), dim3(256), 0, s, P, Pf, PTf);
  return hipGetLastError();
}

std::vector<int32_t> estep_sequence(int n, const std::vector<int32_t>& lc,
                                    const std::vector<int32_t>& rc) {
  const int nl = (n + 1) / 2;
  std::vector<int32_t> seq;
  for (int i = nl; i < n - 1; ++i) seq.push_back(i);  // inside: P_i, nodes in order
  for (int pn = n - 1; pn >= nl; --pn)                // outside: P_c^T, internal children
    for (int c : {lc[pn], rc[pn]})
      if (c >= nl) seq.push_back(-1 - c);
  return seq;
}

int estep_grid(const pcsf_ctx* c) { return c->num_cus * 2; }

size_t estep_scratch_doubles(int n, int grid) {
  const int nl = (n + 1) / 2;
  return (size_t)grid * kWaves * (n - nl) * kTileDoubles;
}

hipError_t launch_estep(const EstepArgs& a, const double* P, double* ecounts, int grid,
                        hipStream_t s) {
  if (a.ncols <= 0) return hipSuccess;
  const int64_t nblocks = (a.ncols + kCols - 1) / kCols;
  if (nblocks * kCols > a.ldu) return hipErrorInvalidValue;
  const int g = (int)std::min<int64_t>(nblocks, grid);
  hipLaunchKernelGGL(estep_passes_kernel, dim3(g), dim3(64 * kWaves), 0, s, a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // enough workgroups to fill the device: splits of the columns per branch
  const int splits = (int)std::max<int64_t>(
      1, std::min<int64_t>({16, nblocks, (4 * grid + a.n - 2) / (a.n - 1)}));
  hipLaunchKernelGGL(estep_counts_kernel, dim3(a.n - 1, splits), dim3(256), 0, s, a.U, a.V,
                     a.ldu, nblocks * kCols, a.leaves, a.nl, a.ncols, P, ecounts);
  return hipGetLastError();
}
