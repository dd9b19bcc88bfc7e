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
//    prefetches its share of the next GEMM's tile from L2.  The workgroup's
//    leaf codes are staged in LDS once per task, and each GEMM also issues
//    the loads of the next leaf gather (named by the host's op program), so
//    gathers do not wait on a dependent chain of global loads.
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
  const float* ctab;   // cherry tables of the sets (nullptr: the program has none)
  const double* clnt;
  int nctab;
  double* logz;
  double* colelpr;
  // per-region tree scales (block_region == nullptr: one set per model)
  const int32_t* block_region;  // per kAlignCols block of columns: region
  const uint8_t* active;        // per set: 0 = skip
  int r0;                       // first region of the chunk
  int64_t g0;                   // first group of the chunk
};

struct Task {
  int q;       // model
  int set;     // transition-matrix set
  int64_t g;   // group within the launch
};

// Workgroup -> (model, group of columns).  When the grid is a multiple of 8
// and nm divides 8, model q is served by the workgroups on XCDs
// [q*8/nm, (q+1)*8/nm).  Returns false past the last task.
__device__ __forceinline__ bool task_at(const MfmaParams& p, int64_t j, Task* t) {
  const int nb = gridDim.x;
  if (p.nm <= 8 && (8 % p.nm) == 0 && (nb % 8) == 0) {
    const int xcd = blockIdx.x & 7, per = 8 / p.nm;
    const int64_t rank = (int64_t)(blockIdx.x >> 3) * per + (xcd % per);
    const int64_t count = nb / p.nm;
    t->q = xcd / per;
    t->g = rank + j * count;
  } else {
    const int64_t idx = blockIdx.x + j * (int64_t)nb;
    t->q = (int)(idx % p.nm);
    t->g = idx / p.nm;
  }
  return t->g < p.ngroups;
}

// The next task from index j on that has work (region mode skips padding
// blocks and inactive sets); j is left at the task returned.
template <int GC>
__device__ __forceinline__ bool next_task(const MfmaParams& p, int64_t& j, Task* t) {
  for (;; ++j) {
    if (!task_at(p, j, t)) return false;
    if (!p.block_region) {
      t->set = t->q;
      return true;
    }
    const int r = p.block_region[((p.g0 + t->g) * GC) / kAlignCols];
    if (r < 0) continue;
    t->set = (r - p.r0) * p.nm + t->q;
    if (p.active[t->set]) return true;
  }
}

// A 64x16 tile: t[mb][v] = row 16*mb + 4*grp + v of this lane's column.
struct Tile {
  f32x4 r[4];
};

struct Gathered {
  Tile t;
  double ln;   // ln P[c][c] (0 for `Marginalize)
  bool marg;
};

// Loads the normalised column of P of leaf `leaf` for this lane's codon.
__device__ __forceinline__ void gather(Gathered& g, const MfmaParams& p, const uint8_t* codes_s,
                                       int gc, int set, int leaf, int ccol, int grp) {
  const int code = codes_s[leaf * gc + ccol];
  g.marg = code == PCSF_MARGINALIZE;
  const int c = g.marg ? 0 : code;
  const size_t bc = ((size_t)set * p.nbr + leaf) * 64 + c;
  const float* src = p.PT32 + bc * 64 + 4 * grp;
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) g.t.r[mb] = *reinterpret_cast<const f32x4*>(src + 16 * mb);
  g.ln = p.lnPcc[bc];
}

// Applies a gathered leaf: `Marginalize is the all-ones message.
__device__ __forceinline__ const Tile& leaf_tile(Gathered& g, double& ls) {
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) g.t.r[mb] = g.marg ? (f32x4)(1.0f) : g.t.r[mb];
  ls += g.marg ? 0.0 : g.ln;
  return g.t;
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

template <int NS, int WPG, int MINW, bool PREF>
__global__ __launch_bounds__(64 * WPG, MINW) void prune_mfma_kernel(MfmaParams p) {
  constexpr int NT = 64 * WPG;          // threads
  constexpr int PER = 4096 / NT / 4;    // float4 of the fragments per thread
  constexpr int GC = 16 * WPG;          // columns per workgroup
  static_assert(NT >= 16, "the diagonal (16 float4) is copied by the first 16 threads");
  __shared__ __attribute__((aligned(16))) float Pbuf[2][kTileFloats];
  extern __shared__ __attribute__((aligned(16))) uint8_t codes_s[];  // nl x GC
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
  int64_t jc = 0;
  if (!next_task<GC>(p, jc, &cur)) return;
  int kg = 0;
  if (p.first_gemm_br != kNoBranch) {
    const float* src = p.Pfrag + ((size_t)cur.set * p.nbr + p.first_gemm_br) * kTileFloats;
#pragma unroll
    for (int r = 0; r < PER; ++r)
      *reinterpret_cast<f32x4*>(&Pbuf[0][(tid + NT * r) * 4]) =
          *reinterpret_cast<const f32x4*>(src + (tid + NT * r) * 4);
    if (tid < 16)
      *reinterpret_cast<f32x4*>(&Pbuf[0][4096 + tid * 4]) =
          *reinterpret_cast<const f32x4*>(src + 4096 + tid * 4);
  }

  Gathered PF;  // leaf gather issued ahead by the previous GEMM
  PF.marg = true;
  PF.ln = 0.0;
  const int ccol = wave * 16 + lcol;  // column within the workgroup's group
  for (;;) {
    Task nxt;
    int64_t jn = jc + 1;
    const bool has_next = next_task<GC>(p, jn, &nxt);
    const int q = cur.q, set = cur.set;
    const int64_t gabs = p.g0 + cur.g;
    const int64_t col = gabs * GC + ccol;
    // stage this group's leaf codes (nl x GC bytes) in LDS
    __syncthreads();  // the previous task's gathers are done with codes_s
    {
      const uint32_t* src = reinterpret_cast<const uint32_t*>(p.codes);
      uint32_t* dst = reinterpret_cast<uint32_t*>(codes_s);
      const int words = GC / 4;
      for (int e = tid; e < p.nl * words; e += NT) {
        const int l = e / words, w = e - l * words;
        dst[e] = src[((size_t)l * p.ld + (size_t)gabs * GC) / 4 + w];
      }
    }
    __syncthreads();
    int E = 0;
    double LS = 0.0;  // sum of ln P[c][c] of the gathered leaves
    for (int pc = 0;; ++pc) {
      const int w0 = __builtin_amdgcn_readfirstlane(p.prog[4 * pc]);
      const int w1 = __builtin_amdgcn_readfirstlane(p.prog[4 * pc + 1]);
      const int w2 = __builtin_amdgcn_readfirstlane(p.prog[4 * pc + 2]);
      const int op = w0 & 0xff, d = (w0 >> 8) & 0xff, use_pf = PREF ? (w0 >> 16) & 1 : 0;
      const int a0 = w1 & 0xffff, a1 = (w1 >> 16) & 0xffff;
      if (op == OP_END) break;

      auto first_leaf = [&](int leaf) -> Gathered& {
        if (!use_pf) gather(PF, p, codes_s, GC, set, leaf, ccol, grp);
        return PF;
      };
      auto do_cherry = [&](Tile& s, uint64_t& k) {
        Gathered g1;
        gather(g1, p, codes_s, GC, set, a1, ccol, grp);
        Gathered& g0 = first_leaf(a0);
        s = leaf_tile(g0, LS);
        k = __ballot(g0.marg && g1.marg);
        mul_rescale(s, leaf_tile(g1, LS), E);
      };
      auto do_cherryt = [&](Tile& s, uint64_t& k) {
        const int ca = codes_s[a0 * GC + ccol], cb = codes_s[a1 * GC + ccol];
        const bool ma = ca == PCSF_MARGINALIZE, mb = cb == PCSF_MARGINALIZE;
        const size_t e = ((size_t)set * p.nctab + w2) * (kTabCodes * kTabCodes) +
                         (ma ? 64 : ca) * kTabCodes + (mb ? 64 : cb);
        const float* src = p.ctab + e * 64 + 4 * grp;
#pragma unroll
        for (int mbk = 0; mbk < 4; ++mbk) s.r[mbk] = *reinterpret_cast<const f32x4*>(src + 16 * mbk);
        LS += p.clnt[e];
        k = __ballot(ma && mb);
      };
      auto do_leafmul = [&](Tile& s, uint64_t& k) {
        Gathered& g0 = first_leaf(a0);
        k &= __ballot(g0.marg);
        mul_rescale(s, leaf_tile(g0, LS), E);
      };
      auto do_gemm = [&](Tile& s, uint64_t k) {
        __syncthreads();  // Pbuf[kg&1] complete; Pbuf[(kg&1)^1] no longer read
        if (PREF && (w2 & 0xffff) != kNoBranch)
          gather(PF, p, codes_s, GC, set, w2 & 0xffff, ccol, grp);
        const float* nsrc = nullptr;
        if (a1 != kNoBranch) nsrc = p.Pfrag + ((size_t)set * p.nbr + a1) * kTileFloats;
        else if (has_next)
          nsrc = p.Pfrag + ((size_t)nxt.set * p.nbr + p.first_gemm_br) * kTileFloats;
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
  case OP_CHERRYT * 8 + D:                             \
    if constexpr (D < NS) do_cherryt(S[D], K[D]);      \
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
    jc = jn;
  }
}

template <int NS, int WPG, int MINW, bool PREF>
hipError_t launch_ns(const MfmaParams& prm, int grid_per_cu, int num_cus, hipStream_t s) {
  auto kern = prune_mfma_kernel<NS, WPG, MINW, PREF>;
  const size_t dyn = ((size_t)prm.nl * 16 * WPG + 15) / 16 * 16;  // staged leaf codes
  if (grid_per_cu <= 0) {
    int per_cu = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * WPG, dyn);
    if (e != hipSuccess) return e;
    grid_per_cu = per_cu > 0 ? per_cu : 1;
  }
  int grid = (grid_per_cu * num_cus + 7) / 8 * 8;
  hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * WPG), dyn, s, prm);
  return hipGetLastError();
}

template <int WPG, int MINW, bool PREF>
hipError_t launch_v(const MfmaParams& prm, int slots, int gpc, int cus, hipStream_t s) {
  switch (slots) {
    case 1:
    case 2: return launch_ns<2, WPG, MINW, PREF>(prm, gpc, cus, s);
    case 3: return launch_ns<3, WPG, MINW, PREF>(prm, gpc, cus, s);
    case 4: return launch_ns<4, WPG, MINW, PREF>(prm, gpc, cus, s);
    case 5: return launch_ns<5, WPG, MINW, PREF>(prm, gpc, cus, s);
    case 6: return launch_ns<6, WPG, MINW, PREF>(prm, gpc, cus, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace

const MfmaVariant kMfmaVariants[kNumMfmaVariants] = {
    {8, 1, true, "8 waves/WG, gather prefetch"},
    {4, 2, true, "4 waves/WG, >=2 waves/SIMD, gather prefetch"},
    {4, 3, false, "4 waves/WG, >=3 waves/SIMD"},
    {2, 2, true, "2 waves/WG, >=2 waves/SIMD, gather prefetch"},
    {4, 2, false, "4 waves/WG, >=2 waves/SIMD"},
    {2, 3, false, "2 waves/WG, >=3 waves/SIMD"},
};

namespace {
hipError_t launch_variant(const MfmaParams& prm, int variant, int slots, int cus,
                          hipStream_t s) {
  switch (variant) {
    case 0: return launch_v<8, 1, true>(prm, slots, 0, cus, s);
    case 1: return launch_v<4, 2, true>(prm, slots, 0, cus, s);
    case 2: return launch_v<4, 3, false>(prm, slots, 0, cus, s);
    case 3: return launch_v<2, 2, true>(prm, slots, 0, cus, s);
    case 4: return launch_v<4, 2, false>(prm, slots, 0, cus, s);
    case 5: return launch_v<2, 3, false>(prm, slots, 0, cus, s);
    default: return hipErrorInvalidValue;
  }
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
  prm.pi = m->d_pi;
  prm.lnpi = m->d_lnpi;
  return prm;
}
}  // namespace

hipError_t launch_prune_mfma(const pcsf_model* m, const pcsf_batch* b, int variant,
                             hipStream_t s) {
  if (variant < 0 || variant >= kNumMfmaVariants) return hipErrorInvalidValue;
  const int wpg = kMfmaVariants[variant].wpg;
  MfmaParams prm = base_params(m, m->fixed);
  prm.codes = b->d_codes;
  prm.ld = b->ld;
  prm.ngroups = b->ld / (16 * wpg);
  prm.logz = b->d_logz;
  prm.colelpr = b->d_colelpr;
  if (prm.ngroups == 0) return hipSuccess;
  return launch_variant(prm, variant, m->slots, m->ctx->num_cus, s);
}

hipError_t launch_prune_mfma_regions(const pcsf_model* m, const pcsf_batch* b, int variant,
                                     const RegionChunk& ch, hipStream_t s) {
  if (variant < 0 || variant >= kNumMfmaVariants) return hipErrorInvalidValue;
  const int gc = 16 * kMfmaVariants[variant].wpg;
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
  return launch_variant(prm, variant, m->slots, m->ctx->num_cus, s);
}
