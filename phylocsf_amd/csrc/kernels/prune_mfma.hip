// prune_mfma.hip -- fused whole-tree Felsenstein pruning on MFMA (gfx950).
//
// Same result as PhyloLik.ensure_alpha + the root posterior
// (lib/CamlPaml/PhyloLik.ml:73-93,127-138) summed by
// PhyloCSFModel.lpr_leaves (src/PhyloCSFModel.ml:67-82), for every packed
// codon column and every model, organised for CDNA4:
//
//  * A wave owns a tile of 16 columns; the WPG waves of a workgroup walk the
//    same tree in lock step, so one P matrix staged in LDS feeds 16*WPG
//    columns.  Partial likelihoods never leave registers: the host compiles
//    the tree into an op program (post-order, subtree needing more live
//    tiles first) that keeps at most NS 64x16 tiles (16 VGPRs each) live.
//  * An internal-child message m = P alpha is a 64x64 by 64x16 GEMM on
//    v_mfma_f32_16x16x4f32.  Its accumulator layout (lane l holds column
//    l%16, rows 16*mb + 4*(l/16) + v in register 4*mb + v) is used directly
//    as the B operand of the next GEMM by permuting the K index, so partials
//    move between GEMMs without any shuffle; the P tile is stored in the
//    matching permuted order (pbuild.hip) and read with ds_read_b128.
//  * FP32 MFMA accumulation rounds with a bias toward zero (measured on
//    MI355X: about -2e-6 per column and model when the accumulator carries
//    alpha).  So the MFMA only accumulates the off-diagonal mass of P from a
//    zero accumulator, and m = alpha + fma(P_aa - 1, alpha, offdiag) is
//    finished in round-to-nearest VALU arithmetic.
//  * Leaf children are gathers of a column of P normalised by its diagonal
//    entry (so the dominant entry is exactly 1; ln P[c][c] is summed per
//    column in FP64), or 1 for `Marginalize; cherries need no GEMM.  A lane mask per live tile marks columns whose
//    partial is a constant vector (a fully marginalised subtree); P c1 = c1
//    exactly, so those columns skip the GEMM update instead of picking up
//    the rounding of the FP32 row sums of P - I.  After every product each
//    column is rescaled by the power of two of its maximum (exact) and the
//    exponents are summed per column; the root dot and the logarithms are
//    FP64.
//  * P tiles are double-buffered in LDS: while a GEMM runs, every thread
//    prefetches its share of the next GEMM's tile from L2.
//  * Workgroups are mapped to models by XCD (blockIdx % 8), so each XCD's
//    L2 holds one model's matrices.
#include "../internal.hpp"

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));

struct MfmaParams {
  const uint8_t* codes;
  int64_t ld;
  int64_t ngroups;
  int nl, nm, nbr, first_gemm_br;
  const int32_t* prog;
  const float* Pfrag;
  const float* PT32;
  const double* pi;
  const double* lnpi;
  const double* lnPcc;
  double* logz;
  double* colelpr;
};

struct Task {
  int q;
  int64_t g;
};

// Workgroup -> (model, group of columns).  When the grid is a multiple of 8
// and nm divides 8, model q is served by the workgroups on XCDs
// [q*8/nm, (q+1)*8/nm).
__device__ __forceinline__ bool task_at(const MfmaParams& p, int64_t j, Task* t) {
  const int nb = gridDim.x;
  if (p.nm <= 8 && (8 % p.nm) == 0 && (nb % 8) == 0) {
    const int xcd = blockIdx.x & 7, per = 8 / p.nm;
    const int64_t rank = (int64_t)(blockIdx.x >> 3) * per + (xcd % per);
    const int64_t count = nb / p.nm;
    t->q = xcd / per;
    t->g = rank + j * count;
    return t->g < p.ngroups;
  }
  const int64_t idx = blockIdx.x + j * (int64_t)nb;
  t->q = (int)(idx % p.nm);
  t->g = idx / p.nm;
  return t->g < p.ngroups;
}

// A 64x16 tile: t[mb][v] = row 16*mb + 4*grp + v of this lane's column.
struct Tile {
  f32x4 r[4];
};

__device__ __forceinline__ bool gather(Tile& g, const MfmaParams& p, int q, int leaf,
                                       int64_t col, int grp, double& ls) {
  const int code = p.codes[(size_t)leaf * p.ld + col];
  const bool marg = code == PCSF_MARGINALIZE;
  const int c = marg ? 0 : code;
  const size_t bc = ((size_t)q * p.nbr + leaf) * 64 + c;
  const float* src = p.PT32 + bc * 64 + 4 * grp;
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) {
    const f32x4 v = *reinterpret_cast<const f32x4*>(src + 16 * mb);
    g.r[mb] = marg ? (f32x4)(1.0f) : v;
  }
  const double l = p.lnPcc[bc];
  ls += marg ? 0.0 : l;
  return marg;
}

__device__ __forceinline__ void mul_rescale(Tile& s, const Tile& o, int& E) {
  float mx = 0.0f;
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) {
    s.r[mb] *= o.r[mb];
#pragma unroll
    for (int v = 0; v < 4; ++v) mx = fmaxf(mx, s.r[mb][v]);
  }
  mx = fmaxf(mx, __shfl_xor(mx, 16));
  mx = fmaxf(mx, __shfl_xor(mx, 32));
  const int e = (mx > 0.0f) ? __builtin_amdgcn_frexp_expf(mx) : 0;
#pragma unroll
  for (int mb = 0; mb < 4; ++mb)
#pragma unroll
    for (int v = 0; v < 4; ++v) s.r[mb][v] = __builtin_amdgcn_ldexpf(s.r[mb][v], -e);
  E += e;
}

template <int NS, int WPG, int MINW>
__global__ __launch_bounds__(64 * WPG, MINW) void prune_mfma_kernel(MfmaParams p) {
  constexpr int NT = 64 * WPG;          // threads
  constexpr int PER = 4096 / NT / 4;    // float4 of the fragments per thread
  constexpr int GC = 16 * WPG;          // columns per workgroup
  static_assert(NT >= 16, "the diagonal (16 float4) is copied by the first 16 threads");
  __shared__ __attribute__((aligned(16))) float Pbuf[2][kTileFloats];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int grp = lane >> 4, lcol = lane & 15;

  Tile S[NS];
  uint64_t K[NS];  // lanes whose tile column is a constant vector
#pragma unroll
  for (int d = 0; d < NS; ++d) {
    K[d] = 0;
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) S[d].r[mb] = (f32x4)(0.0f);
  }

  Task cur;
  if (!task_at(p, 0, &cur)) return;
  int kg = 0;
  if (p.first_gemm_br != kNoBranch) {
    const float* src = p.Pfrag + ((size_t)cur.q * p.nbr + p.first_gemm_br) * kTileFloats;
#pragma unroll
    for (int r = 0; r < PER; ++r)
      *reinterpret_cast<f32x4*>(&Pbuf[0][(tid + NT * r) * 4]) =
          *reinterpret_cast<const f32x4*>(src + (tid + NT * r) * 4);
    if (tid < 16)
      *reinterpret_cast<f32x4*>(&Pbuf[0][4096 + tid * 4]) =
          *reinterpret_cast<const f32x4*>(src + 4096 + tid * 4);
  }

  for (int64_t j = 0;; ++j) {
    Task nxt;
    const bool has_next = task_at(p, j + 1, &nxt);
    const int q = cur.q;
    const int64_t col = cur.g * GC + wave * 16 + lcol;
    int E = 0;
    double LS = 0.0;  // sum of ln P[c][c] of the gathered leaves
    for (int pc = 0;; ++pc) {
      const int w0 = __builtin_amdgcn_readfirstlane(p.prog[2 * pc]);
      const int w1 = __builtin_amdgcn_readfirstlane(p.prog[2 * pc + 1]);
      const int op = w0 & 0xff, d = w0 >> 8;
      const int a0 = w1 & 0xffff, a1 = (w1 >> 16) & 0xffff;
      if (op == OP_END) break;

      auto do_cherry = [&](Tile& s, uint64_t& k) {
        Tile g;
        const bool m0 = gather(s, p, q, a0, col, grp, LS);
        const bool m1 = gather(g, p, q, a1, col, grp, LS);
        k = __ballot(m0 && m1);
        mul_rescale(s, g, E);
      };
      auto do_leafmul = [&](Tile& s, uint64_t& k) {
        Tile g;
        const bool m0 = gather(g, p, q, a0, col, grp, LS);
        k &= __ballot(m0);
        mul_rescale(s, g, E);
      };
      auto do_gemm = [&](Tile& s, uint64_t k) {
        __syncthreads();  // Pbuf[kg&1] complete; Pbuf[(kg&1)^1] no longer read
        const float* nsrc = nullptr;
        if (a1 != kNoBranch) nsrc = p.Pfrag + ((size_t)q * p.nbr + a1) * kTileFloats;
        else if (has_next)
          nsrc = p.Pfrag + ((size_t)nxt.q * p.nbr + p.first_gemm_br) * kTileFloats;
        f32x4 pf[PER + 1];
        if (nsrc) {
#pragma unroll
          for (int r = 0; r < PER; ++r)
            pf[r] = *reinterpret_cast<const f32x4*>(nsrc + (tid + NT * r) * 4);
          if (tid < 16) pf[PER] = *reinterpret_cast<const f32x4*>(nsrc + 4096 + tid * 4);
        }
        const float* Ab = Pbuf[kg & 1];
        f32x4 acc[4];
#pragma unroll
        for (int mo = 0; mo < 4; ++mo) acc[mo] = (f32x4)(0.0f);
#pragma unroll
        for (int mb = 0; mb < 4; ++mb) {
          f32x4 A[4];
#pragma unroll
          for (int mo = 0; mo < 4; ++mo)
            A[mo] = *reinterpret_cast<const f32x4*>(Ab + ((mo * 4 + mb) * 64 + lane) * 4);
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const float bv = s.r[mb][v];
#pragma unroll
            for (int mo = 0; mo < 4; ++mo)
              acc[mo] = __builtin_amdgcn_mfma_f32_16x16x4f32(A[mo][v], bv, acc[mo], 0, 0, 0);
          }
        }
        // m = alpha + (P_aa - 1) alpha + offdiag(P) alpha, in round-to-nearest VALU
#pragma unroll
        for (int mo = 0; mo < 4; ++mo) {
          const f32x4 dd = *reinterpret_cast<const f32x4*>(Ab + 4096 + 16 * mo + 4 * grp);
#pragma unroll
          for (int v = 0; v < 4; ++v)
            acc[mo][v] = s.r[mo][v] + fmaf(dd[v], s.r[mo][v], acc[mo][v]);
        }
        const bool cst = (k >> lane) & 1;
#pragma unroll
        for (int mo = 0; mo < 4; ++mo) s.r[mo] = cst ? s.r[mo] : acc[mo];
        if (nsrc) {
          float* dst = Pbuf[(kg & 1) ^ 1];
#pragma unroll
          for (int r = 0; r < PER; ++r)
            *reinterpret_cast<f32x4*>(&dst[(tid + NT * r) * 4]) = pf[r];
          if (tid < 16) *reinterpret_cast<f32x4*>(&dst[4096 + tid * 4]) = pf[PER];
        }
        ++kg;
      };
      auto do_root = [&](const Tile& s) {
        const double* piq = p.pi + q * 64;
        const double* lnq = p.lnpi + q * 64;
        double z = 0.0, el = 0.0;
#pragma unroll
        for (int mb = 0; mb < 4; ++mb)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            const int row = 16 * mb + 4 * grp + v;
            const double x = (double)s.r[mb][v] * piq[row];
            z += x;
            el = fma(x, lnq[row], el);
          }
        z += __shfl_xor(z, 16);
        el += __shfl_xor(el, 16);
        z += __shfl_xor(z, 32);
        el += __shfl_xor(el, 32);
        if (grp == 0) {
          p.logz[(size_t)q * p.ld + col] = log(z) + (double)E * 0.69314718055994530942 + LS;
          p.colelpr[(size_t)q * p.ld + col] = (z == 0.0) ? 0.0 : el / z;
        }
      };

#define PCSF_CASES(D)                                  \
  case OP_CHERRY * 8 + D:                              \
    if constexpr (D < NS) do_cherry(S[D], K[D]);       \
    break;                                             \
  case OP_GEMM * 8 + D:                                \
    if constexpr (D < NS) do_gemm(S[D], K[D]);         \
    break;                                             \
  case OP_LEAFMUL * 8 + D:                             \
    if constexpr (D < NS) do_leafmul(S[D], K[D]);      \
    break;                                             \
  case OP_MULPOP * 8 + D:                              \
    if constexpr (D + 1 < NS) {                        \
      K[D] &= K[D + 1];                                \
      mul_rescale(S[D], S[D + 1], E);                  \
    }                                                  \
    break;

      switch (op * 8 + d) {
        PCSF_CASES(0)
        PCSF_CASES(1)
        PCSF_CASES(2)
        PCSF_CASES(3)
        PCSF_CASES(4)
        PCSF_CASES(5)
        case OP_ROOT * 8 + 0:
          do_root(S[0]);
          break;
        default:
          break;
      }
#undef PCSF_CASES
    }
    if (!has_next) break;
    cur = nxt;
  }
}

template <int NS, int WPG, int MINW>
hipError_t launch_ns(const MfmaParams& prm, int grid_per_cu, int num_cus, hipStream_t s) {
  auto kern = prune_mfma_kernel<NS, WPG, MINW>;
  if (grid_per_cu <= 0) {
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * WPG, 0);
    if (e != hipSuccess) return e;
    grid_per_cu = per_cu > 0 ? per_cu : 1;
  }
  int grid = (grid_per_cu * num_cus + 7) / 8 * 8;
  hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * WPG), 0, s, prm);
  return hipGetLastError();
}

template <int WPG, int MINW>
hipError_t launch_v(const MfmaParams& prm, int slots, int gpc, int cus, hipStream_t s) {
  switch (slots) {
    case 1:
    case 2: return launch_ns<2, WPG, MINW>(prm, gpc, cus, s);
    case 3: return launch_ns<3, WPG, MINW>(prm, gpc, cus, s);
    case 4: return launch_ns<4, WPG, MINW>(prm, gpc, cus, s);
    case 5: return launch_ns<5, WPG, MINW>(prm, gpc, cus, s);
    case 6: return launch_ns<6, WPG, MINW>(prm, gpc, cus, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

const MfmaVariant kMfmaVariants[kNumMfmaVariants] = {
    {8, 1, "8 waves/WG, >=1 wave/SIMD"}, {4, 2, "4 waves/WG, >=2 waves/SIMD"},
    {4, 3, "4 waves/WG, >=3 waves/SIMD"}, {8, 3, "8 waves/WG, >=3 waves/SIMD"},
    {2, 2, "2 waves/WG, >=2 waves/SIMD"},
};

hipError_t launch_prune_mfma(const pcsf_model* m, const pcsf_batch* b, int variant,
                             hipStream_t s) {
  if (variant < 0 || variant >= kNumMfmaVariants) return hipErrorInvalidValue;
  const int wpg = kMfmaVariants[variant].wpg;
  MfmaParams prm;
  prm.codes = b->d_codes;
  prm.ld = b->ld;
  prm.ngroups = b->ld / (16 * wpg);
  prm.nl = m->nl;
  prm.nm = m->nm;
  prm.nbr = m->nbr;
  prm.first_gemm_br = m->first_gemm_br;
  prm.prog = m->d_prog;
  prm.Pfrag = m->d_Pfrag;
  prm.PT32 = m->d_PT32;
  prm.pi = m->d_pi;
  prm.lnpi = m->d_lnpi;
  prm.lnPcc = m->d_lnPcc;
  prm.logz = b->d_logz;
  prm.colelpr = b->d_colelpr;
  if (prm.ngroups == 0) return hipSuccess;
  const int cus = m->ctx->num_cus;
  switch (variant) {
    case 0: return launch_v<8, 1>(prm, m->slots, 0, cus, s);
    case 1: return launch_v<4, 2>(prm, m->slots, 0, cus, s);
    case 2: return launch_v<4, 3>(prm, m->slots, 0, cus, s);
    case 3: return launch_v<8, 3>(prm, m->slots, 0, cus, s);
    case 4: return launch_v<2, 2>(prm, m->slots, 0, cus, s);
    default: return hipErrorInvalidValue;
  }
}
