// adjoint_mfma.hip -- the outside (preorder) pass on FP64 matrix cores:
// d ln z / d t_br per region and branch, and d ln z / d rho per column.
//
// The gradient the reference's PhyloEM defines (lib/CamlPaml/PhyloEM.ml:
// 140-182): d ELL / d t_br = sum_ab xi_br[a,b] (dP/dt)_ab / P_ab with the
// branch posteriors xi of PhyloLik.add_branch_posteriors (PhyloLik.ml:
// 140-174) and dP/dt = Q P(t) (Q.ml:260-275); by the Fisher identity it is
// d ln z / d t_br, and with every branch scaled by rho, d ln z / d rho =
// sum_br b_br d ln z / d t_br (PhyloEM.ml:162-182).  For the branch into
// child c of node p with sibling s, with messages m_c = P_c alpha_c:
//   w_c = beta_p o m_s,   d ln z / d t_c = w_c . (Q m_c) / z,
//   beta_c = P_c^T w_c    (PhyloLik.ensure_beta, PhyloLik.ml:96-120),
// beta_root = pi (PhyloLik.ml:62), alpha_p = m_l o m_r (:78-90).
//
// Layout for CDNA4: a wave owns 16 columns as 64 x 16 FP64 tiles (lane
// j + 16 g holds column j, rows 16 mo + 4 v + g in element v of r[mo]: the
// accumulator layout of v_mfma_f64_16x16x4f64, which is also its B operand
// for the K-steps taken in natural order), and the four waves of a workgroup
// (64 columns: one region-aligned block) walk the tree in lock step sharing
// one 64 x 64 FP64 matrix in LDS, stored in the MFMA's A-fragment order
// (element (mo, s, lane) at ((16 mo + s) 64 + lane): conflict-free 8-byte
// reads), with Q resident beside it.  The inside pass stores each internal
// node's message to its parent in a per-wave global slot; the outside pass
// reads the two messages of each node, overwrites them with the children's
// beta, and so needs no more memory.  Per-branch sums are reduced over the
// wave's 16 columns, then over the workgroup in LDS, and added to the region's
// total with one FP64 atomic per (workgroup, branch).  No rescaling, as the
// reference (FP64 range).  (The FP64 forward on matrix cores, which needs
// no outside pass, keeps its tiles in registers: prune_fp64m.hip.)
#include "fp64_tiles.hpp"

namespace {

using namespace fp64t;

// the message of leaf branch `br` for `code`: column `code` of P, or ones
// for `Marginalize (message 64), from the branch's leaf messages in P64F
// (row 16 mo + 4 v + g at code * 64 + 16 g + 4 mo + v)
__device__ __forceinline__ Tile leaf_message(const double* __restrict__ PF, int code, int grp) {
  const int c = (code == PCSF_MARGINALIZE) ? 64 : code;
  const double2* src = reinterpret_cast<const double2*>(PF + c * 64 + 16 * grp);
  Tile t;
#pragma unroll
  for (int mo = 0; mo < 4; ++mo)
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const double2 d = src[2 * mo + h];
      t.r[mo][2 * h] = d.x;
      t.r[mo][2 * h + 1] = d.y;
    }
  return t;
}

__global__ __launch_bounds__(64 * kWaves, 2) void adjoint_mfma_kernel(AdjointArgs p) {
  // the inside pass alternates its P tiles between the two buffers; the
  // outside pass holds Q in Qf and each P^T in Af
  __shared__ __attribute__((aligned(16))) double Qf[4096];
  __shared__ __attribute__((aligned(16))) double Af[4096];
  extern __shared__ double gsum[];        // nbr: per-branch sums of the task's workgroup
  const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63,
            grp = lane >> 4;
  const int j = lane & 15;
  const int n = p.n, nl = p.nl, nm = p.nm, nbr = p.nbr;
  // this wave's message / beta slots, one tile per internal node
  double* slots = p.scratch + ((size_t)blockIdx.x * kWaves + wave) * (n - nl) * kTileDoubles;
  const int64_t ntask = p.nblocks * nm;
  for (int64_t task = blockIdx.x; task < ntask; task += gridDim.x) {
    const int64_t blk = task / nm;
    const int q = (int)(task - blk * nm);
    const int64_t b0 = p.c0 + blk * kCols;
    const int r = p.block_region[b0 / kAlignCols];
    if (r < 0) continue;  // padding block (uniform over the workgroup)
    const int set = p.shared_sets ? q : (r - p.r0) * nm + q;
    if (!p.active[set]) continue;
    const int64_t col = b0 + wave * 16 + j;
    const bool valid = col - p.astart[r] < p.off[r + 1] - p.off[r];
    const double* PF = p.P64F + (size_t)set * nbr * kTile64;    // PF[br * kTile64 + ...]
    const double* PFT = p.P64FT + (size_t)set * nbr * 4096;     // PFT[br * 4096 + ...]
    auto msg = [&](int c) -> Tile {
      return (c < nl) ? leaf_message(PF + (size_t)c * kTile64, p.codes[(size_t)c * p.ld + col], grp)
                      : load(slots + (size_t)(c - nl) * kTileDoubles, lane);
    };
    __syncthreads();  // the previous task is done with Qf, Af and gsum
    for (int br = tid; br < nbr; br += 64 * kWaves) gsum[br] = 0.0;
    // ---- inside pass (PhyloLik.ml:78-92), messages to the slots; nodes
    // nl .. n - 2 take a product with their P, each tile copied into the
    // idle buffer one node ahead
    if (nl <= n - 2) copy_tile((L3*)Af, PF + (size_t)nl * kTile64, wave, tid);
    double z = 0.0;
    for (int i = nl; i < n; ++i) {
      Tile a = msg(p.lc[i]);
      const Tile mr = msg(p.rc[i]);
#pragma unroll
      for (int mo = 0; mo < 4; ++mo) a.r[mo] *= mr.r[mo];  // alpha_i
      if (i == n - 1) {
        const int pq = p.prior_per_set ? set : q;  // one prior per model, or per set
        Tile pr;
#pragma unroll
        for (int mo = 0; mo < 4; ++mo)
#pragma unroll
          for (int v = 0; v < 4; ++v) pr.r[mo][v] = p.pi[pq * 64 + 16 * mo + 4 * v + grp];
        z = column_dot(a, pr);
        break;
      }
      const bool odd = (i - nl) & 1;  // tile i in Qf, else in Af
      __syncthreads();  // tile i landed; the other buffer was last read a node ago
      if (i + 1 <= n - 2) copy_tile((L3*)(odd ? Af : Qf), PF + (size_t)(i + 1) * kTile64, wave, tid);
      store(slots + (size_t)(i - nl) * kTileDoubles, gemm(odd ? Qf : Af, a, lane), lane);
    }
    __syncthreads();  // the inside pass is done with both buffers (and gsum is zero)
    stage(Qf, p.QT + (size_t)q * 4096, false, tid);  // QT[k * 64 + a] = Q[a][k]
    __syncthreads();
    const double zinv = (z != 0.0) ? 1.0 / z : 0.0;
    // ---- outside pass, parents before children; each P^T is copied into
    // Af while the child's Q product runs
    double drho = 0.0;
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
          for (int v = 0; v < 4; ++v) bp.r[mo][v] = p.pi[q * 64 + 16 * mo + 4 * v + grp];
      } else {
        bp = load(slots + (size_t)(pn - nl) * kTileDoubles, lane);
      }
      // w_c = beta_p o m_sibling
      Tile w[2];
#pragma unroll
      for (int mo = 0; mo < 4; ++mo) {
        w[0].r[mo] = bp.r[mo] * m[1].r[mo];
        w[1].r[mo] = bp.r[mo] * m[0].r[mo];
      }
#pragma unroll
      for (int s = 0; s < 2; ++s) {
        const int c = kid[s];
        if (c >= nl) {
          __syncthreads();  // Af no longer read
          copy_tile((L3*)Af, PFT + (size_t)c * 4096, wave, tid);
        }
        // d ln z / d t_c = w_c . (Q m_c) / z for this column.  The product
        // runs on every lane: an MFMA takes its A rows from all 64 lanes, so
        // a lane of a padding column that skipped the fragment loads would
        // hand stale rows to the valid columns.
        const double wqm = column_dot(w[s], gemm(Qf, m[s], lane));
        const double g = valid ? wqm * zinv : 0.0;
        drho = fma(p.branch[c], g, drho);
        if (p.grad_branch) {
          double t = g;  // over the wave's 16 columns
          t += __shfl_xor(t, 1);
          t += __shfl_xor(t, 2);
          t += __shfl_xor(t, 4);
          t += __shfl_xor(t, 8);
          if (lane == 0) atomicAdd(&gsum[c], t);
        }
        if (c >= nl) {  // beta_c = P_c^T w_c into c's slot
          __syncthreads();  // P_c^T landed
          store(slots + (size_t)(c - nl) * kTileDoubles, gemm(Af, w[s], lane), lane);
        }
      }
    }
    if (grp == 0) {
      p.logz[(size_t)q * p.ald + col] = log(z);
      p.coldrho[(size_t)q * p.ald + col] = drho;
    }
    if (p.grad_branch) {
      __syncthreads();
      double* gb = p.grad_branch + ((size_t)(r - p.r0) * nm + q) * nbr;
      for (int br = tid; br < nbr; br += 64 * kWaves)
        if (gsum[br] != 0.0) atomicAdd(gb + br, gsum[br]);
    }
  }
}

}  // namespace

size_t adjoint_mfma_scratch_doubles(const pcsf_model* m, int grid) {
  return (size_t)grid * kWaves * (m->n - m->nl) * kTileDoubles;
}

// resident workgroups per CU: two (243 registers, 64 KiB of LDS)
int adjoint_mfma_grid(const pcsf_model* m) { return m->ctx->num_cus * 2; }

namespace {

AdjointArgs make_args(const pcsf_model* m, const pcsf_batch* b, const RegionChunk& ch,
                      double* scratch) {
  AdjointArgs a;
  a.codes = b->al.d_codes;
  a.ld = b->al.ld;
  a.ald = b->al.ld;
  a.c0 = ch.block0 * kAlignCols;
  a.nblocks = ch.nblocks;
  a.n = m->n;
  a.nl = m->nl;
  a.nm = m->nm;
  a.nbr = m->nbr;
  a.lc = m->d_lc;
  a.rc = m->d_rc;
  a.P64F = ch.sets->d_P64F;
  a.P64FT = ch.sets->d_P64FT;
  a.QT = m->d_QT;
  a.prior_per_set = ch.sets->d_epi ? 1 : 0;
  a.pi = a.prior_per_set ? ch.sets->d_epi : m->d_pi;
  a.lnpi = a.prior_per_set ? ch.sets->d_elnpi : m->d_lnpi;
  a.branch = m->d_branch;
  a.block_region = b->al.d_block_region;
  a.astart = b->al.d_astart;
  a.off = b->d_offsets;
  a.active = ch.sets->d_active;
  a.r0 = ch.r0;
  a.logz = b->d_alogz;
  a.coldrho = b->d_acolelpr;
  a.grad_branch = nullptr;
  a.scratch = scratch;
  a.shared_sets = ch.sets == &m->fixed ? 1 : 0;
  return a;
}

}  // namespace

hipError_t launch_adjoint_mfma(const pcsf_model* m, const pcsf_batch* b, const RegionChunk& ch,
                               double* grad_branch, double* scratch, int grid, hipStream_t s) {
  AdjointArgs a = make_args(m, b, ch, scratch);
  if (a.prior_per_set) return hipErrorInvalidValue;  // Q is per model
  a.grad_branch = grad_branch;
  const int64_t ntask = ch.nblocks * m->nm;
  const int g = (int)std::min<int64_t>(ntask, grid);
  if (g <= 0) return hipSuccess;
  hipLaunchKernelGGL(adjoint_mfma_kernel, dim3(g), dim3(64 * kWaves),
                     m->nbr * sizeof(double), s, a);
  return hipGetLastError();
}
