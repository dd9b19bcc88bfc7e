// prune_tangent.hip -- lpr and d lpr / d rho in one fused pass (FP32 MFMA).
//
// The forward-mode (tangent) companion of the outside-pass adjoint
// (adjoint.hip) for the one derivative the mle strategy needs: every
// branch length is t_c = rho * b_c, so with D_c = dP_c/drho = b_c Q P_c
// (Q.ml:260-275) the inside recursion of PhyloLik.ensure_alpha
// (lib/CamlPaml/PhyloLik.ml:73-93) differentiates node by node:
//   message   m = P a,            m' = P a' + D a
//   node      a = m_l o m_r,      a' = m_l' o m_r + m_l o m_r'
//   root      z = pi . a,         d ln z / d rho = (pi . a') / z
// (PhyloEM.ml:162-182 defines the same derivative through branch
// posteriors; both equal d ln z / d rho.)  Each tile carries its tangent
// tile, scaled by the same power of two, so the ratio at the root is exact
// of the scaling.  Per internal-child branch: three 64x64x16 FP32 MFMA
// products (P a, P a', D a), P and D staged in LDS.  The column layout, the
// op program, leaf messages (65 rows; the tangent rows D[a][c] / P[c][c],
// zero for `Marginalize) and the rescaling are those of prune_mfma.hpp.
// Writes ln z (as the forward) and d ln z / d rho into the elpr slot.
#include "prune_mfma.hpp"

namespace pcsf_mfma {
namespace {

struct TGathered {
  Tile m, d;
  double ln;
  bool marg;
};

__device__ __forceinline__ void tgather(TGathered& g, const MfmaParams& p,
                                        const uint8_t* codes_s, int gc, int set, int leaf,
                                        int ccol, int grp) {
  const int code = codes_s[leaf * gc + ccol];
  g.marg = code == PCSF_MARGINALIZE;
  const int c = g.marg ? 64 : code;
  const size_t bc = ((size_t)set * p.nbr + leaf) * kLeafRows + c;
  const float* sm = p.PT32 + bc * 64 + 4 * grp;
  const float* sd = p.DT32 + bc * 64 + 4 * grp;
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) {
    g.m.r[mb] = *reinterpret_cast<const f32x4*>(sm + 16 * mb);
    g.d.r[mb] = *reinterpret_cast<const f32x4*>(sd + 16 * mb);
  }
  g.ln = p.lnPcc[bc];
}

// both tiles scaled by the power of two that brings s's column maximum to [0.5, 1)
__device__ __forceinline__ void rescale2(Tile& s, Tile& t, int& E) {
  float mx = 0.0f;
#pragma unroll
  for (int mb = 0; mb < 4; ++mb)
#pragma unroll
    for (int v = 0; v < 4; ++v) mx = fmaxf(mx, s.r[mb][v]);
  mx = fmaxf(mx, __shfl_xor(mx, 16));
  mx = fmaxf(mx, __shfl_xor(mx, 32));
  const int e = (mx > 0.0f) ? __builtin_amdgcn_frexp_expf(mx) : 0;
#pragma unroll
  for (int mb = 0; mb < 4; ++mb)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      s.r[mb][v] = __builtin_amdgcn_ldexpf(s.r[mb][v], -e);
      t.r[mb][v] = __builtin_amdgcn_ldexpf(t.r[mb][v], -e);
    }
  E += e;
}

constexpr int kTWpg = 4;  // waves per workgroup (64 columns: one aligned block)

template <int NS>
__global__ __launch_bounds__(64 * kTWpg, 1) void prune_tangent_kernel(MfmaParams p) {
  constexpr int NT = 64 * kTWpg, GC = 16 * kTWpg;
  constexpr int PER = 4096 / NT / 4;  // float4 per thread of a 4096-float tile
  __shared__ __attribute__((aligned(16))) float Pbuf[2][kTileFloats];
  __shared__ __attribute__((aligned(16))) float Dbuf[2][4096];
  extern __shared__ __attribute__((aligned(16))) uint8_t codes_s[];  // nl x GC
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int grp = lane >> 4, lcol = lane & 15;
  const int ccol = wave * 16 + lcol;

  Tile S[NS], T[NS];
  uint64_t K[NS];
#pragma unroll
  for (int d = 0; d < NS; ++d) {
    K[d] = 0;
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) S[d].r[mb] = T[d].r[mb] = (f32x4)(0.0f);
  }
  auto load_tiles = [&](int set, int br, f32x4 (&pp)[PER + 1], f32x4 (&dd)[PER]) {
    const float* ps = p.Pfrag + ((size_t)set * p.nbr + br) * kTileFloats;
    const float* ds = p.Dfrag + ((size_t)set * p.nbr + br) * 4096;
#pragma unroll
    for (int r = 0; r < PER; ++r) {
      pp[r] = *reinterpret_cast<const f32x4*>(ps + (tid + NT * r) * 4);
      dd[r] = *reinterpret_cast<const f32x4*>(ds + (tid + NT * r) * 4);
    }
    if (tid < 16) pp[PER] = *reinterpret_cast<const f32x4*>(ps + 4096 + tid * 4);
  };
  auto store_tiles = [&](int buf, const f32x4 (&pp)[PER + 1], const f32x4 (&dd)[PER]) {
#pragma unroll
    for (int r = 0; r < PER; ++r) {
      *reinterpret_cast<f32x4*>(&Pbuf[buf][(tid + NT * r) * 4]) = pp[r];
      *reinterpret_cast<f32x4*>(&Dbuf[buf][(tid + NT * r) * 4]) = dd[r];
    }
    if (tid < 16) *reinterpret_cast<f32x4*>(&Pbuf[buf][4096 + tid * 4]) = pp[PER];
  };

  Task cur;
  int64_t jc = 0;
  if (!next_task<GC>(p, jc, &cur)) return;
  if (p.first_gemm_br != kNoBranch) {
    f32x4 pp[PER + 1], dd[PER];
    load_tiles(cur.set, p.first_gemm_br, pp, dd);
    store_tiles(0, pp, dd);
  }
  int kg = 0;
  for (;;) {
    Task nxt;
    int64_t jn = jc + 1;
    const bool has_next = next_task<GC>(p, jn, &nxt);
    const int q = cur.q, set = cur.set;
    const int64_t gabs = p.g0 + cur.g;
    const int64_t col = gabs * GC + ccol;
    __syncthreads();
    {
      const uint32_t* src = reinterpret_cast<const uint32_t*>(p.codes);
      uint32_t* dst = reinterpret_cast<uint32_t*>(codes_s);
      const int words = GC / 4;
      for (int e = tid; e < p.nl * words; e += NT) {
        const int l = e / words, wd = e - l * words;
        dst[e] = src[((size_t)l * p.ld + (size_t)gabs * GC) / 4 + wd];
      }
    }
    __syncthreads();
    int E = 0;
    double LS = 0.0;

    auto cherry = [&](Tile& s, Tile& t, uint64_t& k, int a0, int a1) {
      TGathered g0, g1;
      tgather(g0, p, codes_s, GC, set, a0, ccol, grp);
      tgather(g1, p, codes_s, GC, set, a1, ccol, grp);
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) {
        s.r[mb] = g0.m.r[mb] * g1.m.r[mb];
        t.r[mb] = g0.d.r[mb] * g1.m.r[mb] + g0.m.r[mb] * g1.d.r[mb];
      }
      LS += g0.ln + g1.ln;
      k = __ballot(g0.marg && g1.marg);
      rescale2(s, t, E);
    };
    auto leafmul = [&](Tile& s, Tile& t, uint64_t& k, int a0) {
      TGathered g;
      tgather(g, p, codes_s, GC, set, a0, ccol, grp);
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) {
        t.r[mb] = t.r[mb] * g.m.r[mb] + s.r[mb] * g.d.r[mb];
        s.r[mb] = s.r[mb] * g.m.r[mb];
      }
      LS += g.ln;
      k &= __ballot(g.marg);
      rescale2(s, t, E);
    };
    auto mulpop = [&](Tile& s, Tile& t, uint64_t& k, const Tile& s2, const Tile& t2,
                      uint64_t k2) {
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) {
        t.r[mb] = t.r[mb] * s2.r[mb] + s.r[mb] * t2.r[mb];
        s.r[mb] = s.r[mb] * s2.r[mb];
      }
      k &= k2;
      rescale2(s, t, E);
    };
    auto gemm = [&](Tile& s, Tile& t, uint64_t k, int next_br) {
      __syncthreads();  // tiles of buffer kg&1 complete; the other buffer is free
      int nset = -1, nbr = 0;
      if (next_br != kNoBranch) {
        nset = set;
        nbr = next_br;
      } else if (has_next) {
        nset = nxt.set;
        nbr = p.first_gemm_br;
      }
      f32x4 pp[PER + 1], dd[PER];
      if (nset >= 0) load_tiles(nset, nbr, pp, dd);
      const float* Ab = Pbuf[kg & 1];
      const float* Db = Dbuf[kg & 1];
      f32x4 as[4], at[4], ad[4];
#pragma unroll
      for (int mo = 0; mo < 4; ++mo) as[mo] = at[mo] = ad[mo] = (f32x4)(0.0f);
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) {
        f32x4 A[4];
#pragma unroll
        for (int mo = 0; mo < 4; ++mo)
          A[mo] = *reinterpret_cast<const f32x4*>(Ab + ((mo * 4 + mb) * 64 + lane) * 4);
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
          for (int mo = 0; mo < 4; ++mo) {
            as[mo] = __builtin_amdgcn_mfma_f32_16x16x4f32(A[mo][v], s.r[mb][v], as[mo], 0, 0, 0);
            at[mo] = __builtin_amdgcn_mfma_f32_16x16x4f32(A[mo][v], t.r[mb][v], at[mo], 0, 0, 0);
          }
#pragma unroll
        for (int mo = 0; mo < 4; ++mo)
          A[mo] = *reinterpret_cast<const f32x4*>(Db + ((mo * 4 + mb) * 64 + lane) * 4);
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
          for (int mo = 0; mo < 4; ++mo)
            ad[mo] = __builtin_amdgcn_mfma_f32_16x16x4f32(A[mo][v], s.r[mb][v], ad[mo], 0, 0, 0);
      }
      const bool cst = (k >> lane) & 1;
#pragma unroll
      for (int mo = 0; mo < 4; ++mo) {
        const f32x4 dg = *reinterpret_cast<const f32x4*>(Ab + 4096 + 16 * mo + 4 * grp);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const float sn = s.r[mo][v] + fmaf(dg[v], s.r[mo][v], as[mo][v]);
          const float tn = t.r[mo][v] + fmaf(dg[v], t.r[mo][v], at[mo][v]) + ad[mo][v];
          s.r[mo][v] = cst ? s.r[mo][v] : sn;
          t.r[mo][v] = cst ? t.r[mo][v] : tn;
        }
      }
      if (nset >= 0) store_tiles((kg & 1) ^ 1, pp, dd);
      ++kg;
    };
    auto root = [&](const Tile& s, const Tile& t) {
      const double* piq = p.pi + q * 64;
      double z = 0.0, zt = 0.0;
#pragma unroll
      for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = 16 * mb + 4 * grp + v;
          z = fma((double)s.r[mb][v], piq[row], z);
          zt = fma((double)t.r[mb][v], piq[row], zt);
        }
      z += __shfl_xor(z, 16);
      zt += __shfl_xor(zt, 16);
      z += __shfl_xor(z, 32);
      zt += __shfl_xor(zt, 32);
      if (grp == 0) {
        p.logz[(size_t)q * p.ld + col] = log(z) + (double)E * 0.69314718055994530942 + LS;
        p.colelpr[(size_t)q * p.ld + col] = (z == 0.0) ? 0.0 : zt / z;
      }
    };

    for (int pc = 0;; ++pc) {
      const int w0 = __builtin_amdgcn_readfirstlane(p.prog[4 * pc]);
      const int w1 = __builtin_amdgcn_readfirstlane(p.prog[4 * pc + 1]);
      const int op = w0 & 0xff, d = (w0 >> 8) & 0xff;
      const int a0 = w1 & 0xffff, a1 = (w1 >> 16) & 0xffff;
      if (op == OP_END) break;
#define PCSF_TCASES(D)                                                       \
  case OP_CHERRY * 8 + D:                                                    \
    if constexpr (D < NS) cherry(S[D], T[D], K[D], a0, a1);                  \
    break;                                                                   \
  case OP_GEMM * 8 + D:                                                      \
    if constexpr (D < NS) gemm(S[D], T[D], K[D], a1);                        \
    break;                                                                   \
  case OP_LEAFMUL * 8 + D:                                                   \
    if constexpr (D < NS) leafmul(S[D], T[D], K[D], a0);                     \
    break;                                                                   \
  case OP_MULPOP * 8 + D:                                                    \
    if constexpr (D + 1 < NS) mulpop(S[D], T[D], K[D], S[D + 1], T[D + 1], K[D + 1]); \
    break;
      switch (op * 8 + d) {
        PCSF_TCASES(0)
        PCSF_TCASES(1)
        PCSF_TCASES(2)
        PCSF_TCASES(3)
        PCSF_TCASES(4)
        PCSF_TCASES(5)
        case OP_ROOT * 8 + 0:
          root(S[0], T[0]);
          break;
        default:
          break;
      }
#undef PCSF_TCASES
    }
    if (!has_next) break;
    cur = nxt;
    jc = jn;
  }
}

template <int NS>
hipError_t launch_tangent_ns(const MfmaParams& prm, int num_cus, hipStream_t s) {
  auto kern = prune_tangent_kernel<NS>;
  const size_t dyn = ((size_t)prm.nl * 16 * kTWpg + 15) / 16 * 16;
  int per_cu = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * kTWpg, dyn);
  if (e != hipSuccess) return e;
  const int grid = ((per_cu > 0 ? per_cu : 1) * num_cus + 7) / 8 * 8;
  hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * kTWpg), dyn, s, prm);
  return hipGetLastError();
}

}  // namespace
}  // namespace pcsf_mfma

using namespace pcsf_mfma;

hipError_t launch_prune_tangent_regions(const pcsf_model* m, const pcsf_batch* b,
                                        const RegionChunk& ch, hipStream_t s) {
  const PSets& sets = *ch.sets;
  if (!sets.d_Dfrag || !sets.d_DT32 || sets.fmt != kFragF32) return hipErrorInvalidValue;
  constexpr int gc = 16 * kTWpg;
  static_assert(gc <= kAlignCols, "a group must not span regions");
  MfmaParams prm{};
  prm.nl = m->nl;
  prm.nm = m->nm;
  prm.nbr = m->nbr;
  prm.first_gemm_br = m->first_gemm_br;
  prm.prog = m->d_prog;
  prm.Pfrag = sets.d_Pfrag;
  prm.PT32 = sets.d_PT32;
  prm.lnPcc = sets.d_lnPcc;
  prm.Dfrag = sets.d_Dfrag;
  prm.DT32 = sets.d_DT32;
  prm.pi = m->d_pi;
  prm.lnpi = m->d_lnpi;
  prm.codes = b->al.d_codes;
  prm.ld = b->al.ld;
  prm.g0 = ch.block0 * (kAlignCols / gc);
  prm.ngroups = ch.nblocks * (kAlignCols / gc);
  prm.logz = b->d_alogz;
  prm.colelpr = b->d_acolelpr;
  prm.block_region = b->al.d_block_region;
  prm.active = sets.d_active;
  prm.r0 = ch.r0;
  if (prm.ngroups == 0) return hipSuccess;
  switch (m->slots) {
    case 1:
    case 2: return launch_tangent_ns<2>(prm, m->ctx->num_cus, s);
    case 3: return launch_tangent_ns<3>(prm, m->ctx->num_cus, s);
    case 4: return launch_tangent_ns<4>(prm, m->ctx->num_cus, s);
    case 5: return launch_tangent_ns<5>(prm, m->ctx->num_cus, s);
    case 6: return launch_tangent_ns<6>(prm, m->ctx->num_cus, s);
    default: return hipErrorInvalidValue;
  }
}
