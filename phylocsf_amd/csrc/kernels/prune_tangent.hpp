// prune_tangent.hpp -- lpr and d lpr / d rho in one fused pass (FP32 MFMA).
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
// op programs (without cherry tables: a table message has no tangent row),
// leaf messages (65 rows; the tangent rows D[a][c] / P[c][c], zero for
// `Marginalize) and the rescaling are those of prune_mfma.hpp; the walk
// exposes the same ops, so the interpreter and the tree-specialised
// straight-line programs drive it as they drive the forward.  Writes ln z
// (as the forward) and d ln z / d rho into the elpr slot.
#pragma once
#include "prune_mfma.hpp"

namespace pcsf_mfma {

constexpr int kTWpg = 4;  // waves per workgroup (64 columns: one aligned block)

struct TTile {  // a partial and its tangent
  Tile s, t;
};

struct TGathered {
  Tile m, d;
  bool marg;
};

// both tiles scaled by the power of two that brings s's column maximum to
// [0.5, 1) (column maximum as in mul_rescale)
__device__ __forceinline__ void rescale2(Tile& s, Tile& t, int& E) {
  int mx = 0;
#pragma unroll
  for (int mb = 0; mb < 4; ++mb)
#pragma unroll
    for (int v = 0; v < 4; ++v) mx = max(mx, __float_as_int(s.r[mb][v]));
  {
    const auto a = __builtin_amdgcn_permlane16_swap(mx, mx, false, false);
    mx = max((int)a[0], (int)a[1]);
    const auto b = __builtin_amdgcn_permlane32_swap(mx, mx, false, false);
    mx = max((int)b[0], (int)b[1]);
  }
  const int e = (mx > 0) ? __builtin_amdgcn_frexp_expf(__int_as_float(mx)) : 0;
#pragma unroll
  for (int mb = 0; mb < 4; ++mb)
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      s.r[mb][v] = __builtin_amdgcn_ldexpf(s.r[mb][v], -e);
      t.r[mb][v] = __builtin_amdgcn_ldexpf(t.r[mb][v], -e);
    }
  E += e;
  asm("" : "+v"(E));  // a running sum (see mul_rescale)
}

struct TWalk {
  static constexpr int NT = 64 * kTWpg, GC = 16 * kTWpg;
  static constexpr int PER = 4096 / NT / 4;  // float4 per thread of a 4096-float tile
  typedef __attribute__((address_space(3))) float L3;
  const MfmaParams& p;
  const float (*Pbuf)[kTileFloats];  // LDS: double-buffered P and D tiles
  const float (*Dbuf)[4096];
  L3 (*Pbuf3)[kTileFloats];          // the same, as LDS addresses
  L3 (*Dbuf3)[4096];
  const uint8_t* codes_s;
  int tid, lane, grp, ccol;
  int q = 0, set = 0;
  int64_t col = 0;
  bool has_next = false;
  int next_set = 0;
  int kg = 0;
  int E = 0;
  double LS = 0.0;

  __device__ __forceinline__ int oset() const {
    int s = set;
    asm volatile("" : "+s"(s));
    return s;
  }
  __device__ __forceinline__ void tgather(TGathered& g, int leaf) {
    const int code = codes_s[leaf * GC + ccol];
    g.marg = code == PCSF_MARGINALIZE;
    const uint32_t c = g.marg ? 64 : code;
    constexpr uint32_t kRowsBytes = kLeafRows * 64 * 4;
    const size_t s0 = (size_t)oset() * p.nbr;
    const Rsrc rm = make_rsrc(p.PT32 + s0 * (kLeafRows * 64), p.nbr * kRowsBytes);
    const Rsrc rd = make_rsrc(p.DT32 + s0 * (kLeafRows * 64), p.nbr * kRowsBytes);
    const uint32_t off = c * 256u + (uint32_t)grp * 16u;
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      g.m.r[mb] = bload4(rm, off + 64u * mb, leaf * kRowsBytes);
      g.d.r[mb] = bload4(rd, off + 64u * mb, leaf * kRowsBytes);
    }
    const Rsrc rl = make_rsrc(p.lnPcc + s0 * kLeafRows, p.nbr * kLeafRows * 8u);
    LS += bload_f64(rl, c * 8u, leaf * kLeafRows * 8u);
    asm("" : "+v"(LS));
  }
  __device__ __forceinline__ void cherry(TTile& x, uint64_t& k, int a0, int a1, bool) {
    TGathered g0, g1;
    tgather(g0, a0);
    tgather(g1, a1);
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      x.s.r[mb] = g0.m.r[mb] * g1.m.r[mb];
      x.t.r[mb] = g0.d.r[mb] * g1.m.r[mb] + g0.m.r[mb] * g1.d.r[mb];
    }
    k = __ballot(g0.marg && g1.marg);
    rescale2(x.s, x.t, E);
  }
  // programs with cherry tables never reach this kernel (a table message has
  // no tangent row); the op exists so the program drivers compile
  __device__ __forceinline__ void cherryt(TTile&, uint64_t&, int, int, int, bool) {}
  __device__ __forceinline__ void leafmul(TTile& x, uint64_t& k, int a0, bool) {
    TGathered g;
    tgather(g, a0);
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      x.t.r[mb] = x.t.r[mb] * g.m.r[mb] + x.s.r[mb] * g.d.r[mb];
      x.s.r[mb] = x.s.r[mb] * g.m.r[mb];
    }
    k &= __ballot(g.marg);
    rescale2(x.s, x.t, E);
  }
  __device__ __forceinline__ void mulpop(TTile& x, uint64_t& k, const TTile& o, uint64_t ko) {
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) {
      x.t.r[mb] = x.t.r[mb] * o.s.r[mb] + x.s.r[mb] * o.t.r[mb];
      x.s.r[mb] = x.s.r[mb] * o.s.r[mb];
    }
    k &= ko;
    rescale2(x.s, x.t, E);
  }
  // m = P s, m' = P t + D s for the tiles staged in LDS; the next P and D
  // tiles go straight from memory into the idle LDS buffers (as in Walk::gemm)
  __device__ __forceinline__ void gemm(TTile& x, uint64_t k, int next_br, int, uint32_t) {
    __syncthreads();  // tiles of buffer kg&1 complete; the other buffers are free
    {
      typedef __attribute__((address_space(3))) void* LdsPtr;
      L3* dp = Pbuf3[(kg & 1) ^ 1];
      L3* dd = Dbuf3[(kg & 1) ^ 1];
      const int nset = (next_br != kNoBranch) ? oset() : next_set;
      const int nbr = (next_br != kNoBranch) ? next_br : p.first_gemm_br;
      if (next_br != kNoBranch || has_next) {
        constexpr uint32_t kTileBytes = kTileFloats * 4;
        const Rsrc rp = make_rsrc(p.Pfrag + (size_t)nset * p.nbr * kTileFloats, p.nbr * kTileBytes);
        const Rsrc rd = make_rsrc(p.Dfrag + (size_t)nset * p.nbr * 4096, p.nbr * 16384u);
#pragma unroll
        for (int q4 = 0; q4 < PER; ++q4) {
          const uint32_t w0 = (uint32_t)((tid & ~63) + NT * q4) * 4;
          __builtin_amdgcn_raw_ptr_buffer_load_lds(rp, (LdsPtr)(dp + w0), 16, (uint32_t)tid * 16u,
                                                   nbr * kTileBytes + 16u * NT * q4, 0, 0);
          __builtin_amdgcn_raw_ptr_buffer_load_lds(rd, (LdsPtr)(dd + w0), 16, (uint32_t)tid * 16u,
                                                   nbr * 16384u + 16u * NT * q4, 0, 0);
        }
        if (tid < 16)
          __builtin_amdgcn_raw_ptr_buffer_load_lds(rp, (LdsPtr)(dp + 4096), 16,
                                                   (uint32_t)tid * 16u, nbr * kTileBytes + 16384u,
                                                   0, 0);
      }
    }
    const float* Ab = Pbuf[kg & 1];
    const float* Db = Dbuf[kg & 1];
    // P a' and D a stay in separate accumulators: summed in one (the MFMA's
    // FP32 accumulation rounds toward zero) the gradients measured 2-4x less
    // accurate against the FP64 adjoint
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
          as[mo] = __builtin_amdgcn_mfma_f32_16x16x4f32(A[mo][v], x.s.r[mb][v], as[mo], 0, 0, 0);
          at[mo] = __builtin_amdgcn_mfma_f32_16x16x4f32(A[mo][v], x.t.r[mb][v], at[mo], 0, 0, 0);
        }
#pragma unroll
      for (int mo = 0; mo < 4; ++mo)
        A[mo] = *reinterpret_cast<const f32x4*>(Db + ((mo * 4 + mb) * 64 + lane) * 4);
#pragma unroll
      for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int mo = 0; mo < 4; ++mo)
          ad[mo] = __builtin_amdgcn_mfma_f32_16x16x4f32(A[mo][v], x.s.r[mb][v], ad[mo], 0, 0, 0);
    }
    const bool cst = (k >> lane) & 1;
#pragma unroll
    for (int mo = 0; mo < 4; ++mo) {
      const f32x4 dg = *reinterpret_cast<const f32x4*>(Ab + 4096 + 16 * mo + 4 * grp);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const float sn = x.s.r[mo][v] + fmaf(dg[v], x.s.r[mo][v], as[mo][v]);
        const float tn = x.t.r[mo][v] + fmaf(dg[v], x.t.r[mo][v], at[mo][v]) + ad[mo][v];
        x.s.r[mo][v] = cst ? x.s.r[mo][v] : sn;
        x.t.r[mo][v] = cst ? x.t.r[mo][v] : tn;
      }
    }
    ++kg;
  }
  __device__ __forceinline__ void root(const TTile& x) {
    const double* piq = p.pi + q * 64;
    double z = 0.0, zt = 0.0;
#pragma unroll
    for (int mb = 0; mb < 4; ++mb)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = 16 * mb + 4 * grp + v;
        z = fma((double)x.s.r[mb][v], piq[row], z);
        zt = fma((double)x.t.r[mb][v], piq[row], zt);
      }
    z += __shfl_xor(z, 16);
    zt += __shfl_xor(zt, 16);
    z += __shfl_xor(z, 32);
    zt += __shfl_xor(zt, 32);
    if (grp == 0) {
      p.logz[(size_t)q * p.ld + col] = log(z) + (double)E * 0.69314718055994530942 + LS;
      p.colelpr[(size_t)q * p.ld + col] = (z == 0.0) ? 0.0 : zt / z;
    }
  }
};

template <int NS, class PROG>
__global__ __launch_bounds__(64 * kTWpg, 1) void prune_tangent_kernel(MfmaParams p) {
  constexpr int NT = TWalk::NT, GC = TWalk::GC, PER = TWalk::PER;
  __shared__ __attribute__((aligned(16))) float Pbuf[2][kTileFloats];
  __shared__ __attribute__((aligned(16))) float Dbuf[2][4096];
  extern __shared__ __attribute__((aligned(16))) uint8_t codes_s[];  // nl x GC
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;

  TTile S[NS];
  uint64_t K[NS];
#pragma unroll
  for (int d = 0; d < NS; ++d) {
    K[d] = 0;
#pragma unroll
    for (int mb = 0; mb < 4; ++mb) S[d].s.r[mb] = S[d].t.r[mb] = (f32x4)(0.0f);
  }
  Task cur;
  int64_t jc = 0;
  if (!next_task<GC>(p, jc, &cur)) return;
  if (p.first_gemm_br != kNoBranch) {
    const float* ps = p.Pfrag + ((size_t)cur.set * p.nbr + p.first_gemm_br) * kTileFloats;
    const float* ds = p.Dfrag + ((size_t)cur.set * p.nbr + p.first_gemm_br) * 4096;
#pragma unroll
    for (int r = 0; r < PER; ++r) {
      *reinterpret_cast<f32x4*>(&Pbuf[0][(tid + NT * r) * 4]) =
          *reinterpret_cast<const f32x4*>(ps + (tid + NT * r) * 4);
      *reinterpret_cast<f32x4*>(&Dbuf[0][(tid + NT * r) * 4]) =
          *reinterpret_cast<const f32x4*>(ds + (tid + NT * r) * 4);
    }
    if (tid < 16)
      *reinterpret_cast<f32x4*>(&Pbuf[0][4096 + tid * 4]) =
          *reinterpret_cast<const f32x4*>(ps + 4096 + tid * 4);
  }
  TWalk w{p, Pbuf, Dbuf, (TWalk::L3 (*)[kTileFloats])Pbuf, (TWalk::L3 (*)[4096])Dbuf,
          codes_s, tid, lane, lane >> 4, wave * 16 + (lane & 15)};
  for (;;) {
    Task nxt;
    int64_t jn = jc + 1;
    const bool has_next = next_task<GC>(p, jn, &nxt);
    const int64_t gabs = p.g0 + cur.g;
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
    w.q = cur.q;
    w.set = cur.set;
    w.col = gabs * GC + w.ccol;
    w.has_next = has_next;
    w.next_set = has_next ? nxt.set : 0;
    w.E = 0;
    w.LS = 0.0;
    PROG::run(w, S, K);
    if (!has_next) break;
    cur = nxt;
    jc = jn;
  }
}

template <int NS, class PROG = Interp<NS>>
hipError_t launch_tangent_ns(const MfmaParams& prm, int num_cus, hipStream_t s) {
  auto kern = prune_tangent_kernel<NS, PROG>;
  const size_t dyn = ((size_t)prm.nl * 16 * kTWpg + 15) / 16 * 16;
  int per_cu = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, 64 * kTWpg, dyn);
  if (e != hipSuccess) return e;
  const int grid = ((per_cu > 0 ? per_cu : 1) * num_cus + 7) / 8 * 8;
  hipLaunchKernelGGL(kern, dim3(grid), dim3(64 * kTWpg), dyn, s, prm);
  return hipGetLastError();
}

// The tree-specialised tangent launch (precision and variant are those of
// the forward's launcher signature; the tangent pass has one shape).
template <class PROG>
hipError_t launch_tangent_tree(const MfmaParams& prm, int, int, int cus, hipStream_t s) {
  return launch_tangent_ns<PROG::NS, PROG>(prm, cus, s);
}

}  // namespace pcsf_mfma
