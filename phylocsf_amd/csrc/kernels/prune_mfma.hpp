// prune_mfma.hpp -- fused whole-tree Felsenstein pruning on MFMA (gfx950).
// The kernel template is instantiated per GEMM precision in its own
// translation unit (prune_mfma_{f32,f16s,f16,bf16}.hip).
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
//  * GEMM precision (template PREC): FP32 on v_mfma_f32_16x16x4f32; or P and
//    the partial enter v_mfma_f32_16x16x32_{f16,bf16} in half precision with
//    FP32 accumulation.  FP16 operands are scaled by 2^14 (P's off-diagonal
//    entries are below 1; the live tiles of the FP16 kernels are kept with
//    their column maximum in [2^13, 2^14) by the rescaling itself, so nothing
//    overflows and values down to 2^-28 of the maximum stay normal) and P's
//    2^14 is taken off the product exactly.  The split-FP16 precision carries each
//    operand as hi + lo, lo = the FP16 rounding of the residual scaled by 2^11
//    (so that it stays a normal number: the matrix cores treat FP16
//    subnormals as zero, which measurably biased an unscaled lo), and forms
//    hi*hi in one accumulator and hi*lo + lo*hi in another: about 22
//    significant bits per product at three FP16 MFMAs, against sixteen FP32
//    MFMAs for the same 64x64x16 product.
//    Live partials stay FP32 in registers in every precision.  The FP32 path
//    keeps the diagonal of P in an FP32 VALU term; the half-precision tiles
//    hold P - I, so their product is the whole (small) update of alpha.
#include <type_traits>

#include "../internal.hpp"

namespace pcsf_mfma {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// precisions of the GEMM (values of PCSF_FP32 .. PCSF_BF16)
constexpr int kPrecF32 = PCSF_FP32, kPrecF16S = PCSF_FP16S, kPrecF16 = PCSF_FP16,
              kPrecBF16 = PCSF_BF16;
constexpr int kF16Exp = 14;                 // FP16 operands are scaled by 2^14
constexpr float kF16Unscale = 1.0f / 16384.0f;  // P's scale, taken off the product
constexpr float kLoScale = 2048.0f;             // lo parts of split FP16 operands: x 2^11
constexpr float kLoUnscale = 1.0f / 2048.0f;

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

// Loads the leaf message of leaf `leaf` for this lane's codon: the
// normalised column of P, or row 64 (all ones) for `Marginalize.
__device__ __forceinline__ void gather(Gathered& g, const MfmaParams& p, const uint8_t* codes_s,
                                       int gc, int set, int leaf, int ccol, int grp) {
  const int code = codes_s[leaf * gc + ccol];
  g.marg = code == PCSF_MARGINALIZE;
  const int c = g.marg ? 64 : code;
  const size_t bc = ((size_t)set * p.nbr + leaf) * kLeafRows + c;
  const float* src = p.PT32 + bc * 64 + 4 * grp;
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) g.t.r[mb] = *reinterpret_cast<const f32x4*>(src + 16 * mb);
  g.ln = p.lnPcc[bc];
}

__device__ __forceinline__ const Tile& leaf_tile(Gathered& g, double& ls) {
  ls += g.ln;
  return g.t;
}

// s *= o, then every column is scaled by the power of two that brings its
// maximum to [2^(KE-1), 2^KE); E accumulates the exponents, so the true
// product of the column is s * 2^E.
template <int KE>
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
    for (int v = 0; v < 4; ++v) s.r[mb][v] = __builtin_amdgcn_ldexpf(s.r[mb][v], KE - e);
  E += e - KE;
}

// 64x64 (off-diagonal P, LDS tile Ab) times the 64x16 tile s in half
// precision; acc = the product, unscaled.  Tile layout of the halves
// (pbuild.hip): hi fragments, then lo fragments, 512 x 8 values each, in the
// A-operand order of v_mfma_f32_16x16x32_{f16,bf16} with K permuted so that
// K-step st of a lane in group g takes rows 16*(2*st + j/4) + 4*g + j%4 of
// its column: exactly what the accumulator layout leaves in registers
// s.r[2*st + j/4][j%4].
template <int PREC>
__device__ __forceinline__ void gemm_half(f32x4 (&acc)[4], f32x4 (&accx)[4], const Tile& s,
                                          const float* Ab, int lane) {
  constexpr bool kBF = PREC == kPrecBF16;
  using V8 = typename std::conditional<kBF, bf16x8, f16x8>::type;
  using E = typename std::conditional<kBF, __bf16, _Float16>::type;
  V8 bh[2], bl[2];
#pragma unroll
  for (int st = 0; st < 2; ++st)
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const float x = s.r[2 * st + (j >> 2)][j & 3];
      if constexpr (PREC == kPrecF16S) {
        // hi = x cut to 11 significant bits (exact in FP16), lo = the exact
        // remainder, scaled and rounded
        const float h = __uint_as_float(__float_as_uint(x) & 0xFFFFE000u);
        bh[st][j] = (E)h;
        bl[st][j] = (E)((x - h) * kLoScale);
      } else {
        bh[st][j] = (E)x;
      }
    }
  const V8* Ah = reinterpret_cast<const V8*>(Ab);
  const V8* Al = Ah + 512;
#pragma unroll
  for (int mo = 0; mo < 4; ++mo) acc[mo] = (f32x4)(0.0f);
  auto mma = [](const V8& a, const V8& b, const f32x4& c) -> f32x4 {
    if constexpr (kBF) return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
    else return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
  };
  if constexpr (PREC == kPrecF16S) {  // cross terms, carrying the lo scale
#pragma unroll
    for (int mo = 0; mo < 4; ++mo) accx[mo] = (f32x4)(0.0f);
#pragma unroll
    for (int st = 0; st < 2; ++st)
#pragma unroll
      for (int mo = 0; mo < 4; ++mo) {
        const int f = (mo * 2 + st) * 64 + lane;
        accx[mo] = mma(Al[f], bh[st], accx[mo]);
        accx[mo] = mma(Ah[f], bl[st], accx[mo]);
      }
  }
#pragma unroll
  for (int st = 0; st < 2; ++st)
#pragma unroll
    for (int mo = 0; mo < 4; ++mo) acc[mo] = mma(Ah[(mo * 2 + st) * 64 + lane], bh[st], acc[mo]);
}

// One wave's walk of the op program over its 16 columns, and the ops.  The
// interpreter (Interp) and the tree-specialised straight-line programs
// (generated per shipped tree at build time) both drive these members, so
// the arithmetic is the same code in both.
template <int PREC, int WPG, bool PREF>
struct Walk {
  // live tiles of the FP16 GEMMs are kept scaled by 2^14 (their maximum in
  // [2^13, 2^14)), the scale their GEMM operand needs
  static constexpr int KE = (PREC == kPrecF16S || PREC == kPrecF16) ? kF16Exp : 0;
  static constexpr int NT = 64 * WPG;        // threads
  static constexpr int PER = 4096 / NT / 4;  // float4 of the fragments per thread
  static constexpr int GC = 16 * WPG;        // columns per workgroup
  const MfmaParams& p;
  float (*Pbuf)[kTileFloats];  // LDS: double-buffered P tile
  const uint8_t* codes_s;      // LDS: the group's leaf codes, nl x GC
  int tid, lane, grp, ccol;
  int q = 0, set = 0;          // model and transition set of the current task
  int64_t col = 0;             // this lane's column
  bool has_next = false;       // a next task exists (its first tile is prefetched)
  int next_set = 0;
  int kg = 0;                  // GEMMs done (selects the LDS buffer)
  int E = 0;                   // sum of the power-of-two rescaling exponents
  double LS = 0.0;             // sum of ln P[c][c] of the gathered leaves
  Gathered PF;                 // leaf gather issued ahead by the previous GEMM

  __device__ __forceinline__ Walk(const MfmaParams& prm, float (*pb)[kTileFloats],
                                  const uint8_t* cs, int t, int l, int g, int cc)
      : p(prm), Pbuf(pb), codes_s(cs), tid(t), lane(l), grp(g), ccol(cc) {
    PF.marg = true;
    PF.ln = 0.0;
  }

  // The transition set as an opaque scalar: each op recomputes its table
  // addresses from it instead of the compiler keeping every op's address
  // live across a straight-line program (which overflows the SGPRs).
  __device__ __forceinline__ int oset() const {
    int s = set;
    asm volatile("" : "+s"(s));
    return s;
  }
  __device__ __forceinline__ Gathered& first_leaf(int leaf, bool use_pf) {
    if (!(PREF && use_pf)) gather(PF, p, codes_s, GC, oset(), leaf, ccol, grp);
    return PF;
  }
  // s = leaf a0 o leaf a1 (rescaled)
  __device__ __forceinline__ void cherry(Tile& s, uint64_t& k, int a0, int a1, bool use_pf) {
    Gathered g1;
    gather(g1, p, codes_s, GC, oset(), a1, ccol, grp);
    Gathered& g0 = first_leaf(a0, use_pf);
    s = leaf_tile(g0, LS);
    k = __ballot(g0.marg && g1.marg);
    mul_rescale<KE>(s, leaf_tile(g1, LS), E);
  }
  // s = message above the cherry (a0, a1), from table `tab`
  __device__ __forceinline__ void cherryt(Tile& s, uint64_t& k, int a0, int a1, int tab) {
    const int ca = codes_s[a0 * GC + ccol], cb = codes_s[a1 * GC + ccol];
    const bool ma = ca == PCSF_MARGINALIZE, mb = cb == PCSF_MARGINALIZE;
    const size_t e = ((size_t)oset() * p.nctab + tab) * (kTabCodes * kTabCodes) +
                     (ma ? 64 : ca) * kTabCodes + (mb ? 64 : cb);
    const float* src = p.ctab + e * 64 + 4 * grp;
#pragma unroll
    for (int mbk = 0; mbk < 4; ++mbk) s.r[mbk] = *reinterpret_cast<const f32x4*>(src + 16 * mbk);
    // (a table message always meets a sibling in LEAFMUL / MULPOP, whose
    // rescaling sets the tile scale, before it reaches a GEMM)
    LS += p.clnt[e];
    k = __ballot(ma && mb);
  }
  // s = s o leaf a0 (rescaled)
  __device__ __forceinline__ void leafmul(Tile& s, uint64_t& k, int a0, bool use_pf) {
    Gathered& g0 = first_leaf(a0, use_pf);
    k &= __ballot(g0.marg);
    mul_rescale<KE>(s, leaf_tile(g0, LS), E);
  }
  // s = s o o (rescaled)
  __device__ __forceinline__ void mulpop(Tile& s, uint64_t& k, const Tile& o, uint64_t ko) {
    k &= ko;
    mul_rescale<KE>(s, o, E);
  }
  // s = P s for the P tile staged in LDS; prefetches the tile of branch
  // next_br (or the next task's first tile) and, if pf_leaf names one, the
  // next leaf gather
  __device__ __forceinline__ void gemm(Tile& s, uint64_t k, int next_br, int pf_leaf) {
    __syncthreads();  // Pbuf[kg&1] complete; Pbuf[(kg&1)^1] no longer read
    const int st = oset();
    if (PREF && pf_leaf != kNoBranch) gather(PF, p, codes_s, GC, st, pf_leaf, ccol, grp);
    const float* nsrc = nullptr;
    if (next_br != kNoBranch) nsrc = p.Pfrag + ((size_t)st * p.nbr + next_br) * kTileFloats;
    else if (has_next)
      nsrc = p.Pfrag + ((size_t)next_set * p.nbr + p.first_gemm_br) * kTileFloats;
    f32x4 pf[PER + 1];
    if (nsrc) {
#pragma unroll
      for (int r = 0; r < PER; ++r)
        pf[r] = *reinterpret_cast<const f32x4*>(nsrc + (tid + NT * r) * 4);
      if (tid < 16) pf[PER] = *reinterpret_cast<const f32x4*>(nsrc + 4096 + tid * 4);
    }
    const float* Ab = Pbuf[kg & 1];
    f32x4 acc[4], accx[4];
    if constexpr (PREC == kPrecF32) {
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
    } else {
      gemm_half<PREC>(acc, accx, s, Ab, lane);
    }
    // FP32: m = alpha + (P_aa - 1) alpha + offdiag(P) alpha, the diagonal
    // term in round-to-nearest VALU arithmetic; half precisions: the product
    // is (P - I) alpha (carrying P's 2^14 for FP16, taken off in the FMA)
#pragma unroll
    for (int mo = 0; mo < 4; ++mo) {
      if constexpr (PREC == kPrecF32) {
        const f32x4 dd = *reinterpret_cast<const f32x4*>(Ab + 4096 + 16 * mo + 4 * grp);
#pragma unroll
        for (int v = 0; v < 4; ++v)
          acc[mo][v] = s.r[mo][v] + fmaf(dd[v], s.r[mo][v], acc[mo][v]);
      } else {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          if constexpr (PREC == kPrecF16S)
            acc[mo][v] =
                fmaf(fmaf(accx[mo][v], kLoUnscale, acc[mo][v]), kF16Unscale, s.r[mo][v]);
          else if constexpr (KE != 0)
            acc[mo][v] = fmaf(acc[mo][v], kF16Unscale, s.r[mo][v]);
          else
            acc[mo][v] = s.r[mo][v] + acc[mo][v];
        }
      }
    }
    if (k == 0) {  // no lane of this wave holds a constant column (k is wave-uniform)
#pragma unroll
      for (int mo = 0; mo < 4; ++mo) s.r[mo] = acc[mo];
    } else {
      const bool cst = (k >> lane) & 1;
#pragma unroll
      for (int mo = 0; mo < 4; ++mo) s.r[mo] = cst ? s.r[mo] : acc[mo];
    }
    if (nsrc) {
      float* dst = Pbuf[(kg & 1) ^ 1];
#pragma unroll
      for (int r = 0; r < PER; ++r) *reinterpret_cast<f32x4*>(&dst[(tid + NT * r) * 4]) = pf[r];
      if (tid < 16) *reinterpret_cast<f32x4*>(&dst[4096 + tid * 4]) = pf[PER];
    }
    ++kg;
  }
  // ln z and elpr of this lane's column from the root tile
  __device__ __forceinline__ void root(const Tile& s) {
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
  }
};

// The op-program interpreter: reads the host's program (four words per op)
// and dispatches on (opcode, slot).
template <int NS>
struct Interp {
  template <class W>
  __device__ __forceinline__ static void run(W& w, Tile (&S)[NS], uint64_t (&K)[NS]) {
    const int32_t* prog = w.p.prog;
    for (int pc = 0;; ++pc) {
      const int w0 = __builtin_amdgcn_readfirstlane(prog[4 * pc]);
      const int w1 = __builtin_amdgcn_readfirstlane(prog[4 * pc + 1]);
      const int w2 = __builtin_amdgcn_readfirstlane(prog[4 * pc + 2]);
      const int op = w0 & 0xff, d = (w0 >> 8) & 0xff;
      const bool use_pf = (w0 >> 16) & 1;
      const int a0 = w1 & 0xffff, a1 = (w1 >> 16) & 0xffff;
      if (op == OP_END) break;
#define PCSF_CASES(D)                                          \
  case OP_CHERRY * 8 + D:                                      \
    if constexpr (D < NS) w.cherry(S[D], K[D], a0, a1, use_pf); \
    break;                                                     \
  case OP_GEMM * 8 + D:                                        \
    if constexpr (D < NS) w.gemm(S[D], K[D], a1, w2 & 0xffff); \
    break;                                                     \
  case OP_LEAFMUL * 8 + D:                                     \
    if constexpr (D < NS) w.leafmul(S[D], K[D], a0, use_pf);    \
    break;                                                     \
  case OP_CHERRYT * 8 + D:                                     \
    if constexpr (D < NS) w.cherryt(S[D], K[D], a0, a1, w2);   \
    break;                                                     \
  case OP_MULPOP * 8 + D:                                      \
    if constexpr (D + 1 < NS) w.mulpop(S[D], K[D], S[D + 1], K[D + 1]); \
    break;
      switch (op * 8 + d) {
        PCSF_CASES(0)
        PCSF_CASES(1)
        PCSF_CASES(2)
        PCSF_CASES(3)
        PCSF_CASES(4)
        PCSF_CASES(5)
        case OP_ROOT * 8 + 0:
          w.root(S[0]);
          break;
        default:
          break;
      }
#undef PCSF_CASES
    }
  }
};

template <int PREC, int NS, int WPG, int MINW, bool PREF, class PROG>
__global__ __launch_bounds__(64 * WPG, MINW) void prune_mfma_kernel(MfmaParams p) {
  using W = Walk<PREC, WPG, PREF>;
  constexpr int NT = W::NT, PER = W::PER, GC = W::GC;
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

  const int ccol = wave * 16 + lcol;  // column within the workgroup's group
  W w(p, Pbuf, codes_s, tid, lane, grp, ccol);
  for (;;) {
    Task nxt;
    int64_t jn = jc + 1;
    const bool has_next = next_task<GC>(p, jn, &nxt);
    const int64_t gabs = p.g0 + cur.g;
    // stage this group's leaf codes (nl x GC bytes) in LDS
    __syncthreads();  // the previous task's gathers are done with codes_s
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
    w.col = gabs * GC + ccol;
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

template <int PREC, int NS, int WPG, int MINW, bool PREF, class PROG = Interp<NS>>
hipError_t launch_ns(const MfmaParams& prm, int grid_per_cu, int num_cus, hipStream_t s) {
  auto kern = prune_mfma_kernel<PREC, NS, WPG, MINW, PREF, PROG>;
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

template <int PREC, int WPG, int MINW, bool PREF>
hipError_t launch_v(const MfmaParams& prm, int slots, int gpc, int cus, hipStream_t s) {
  switch (slots) {
    case 1:
    case 2: return launch_ns<PREC, 2, WPG, MINW, PREF>(prm, gpc, cus, s);
    case 3: return launch_ns<PREC, 3, WPG, MINW, PREF>(prm, gpc, cus, s);
    case 4: return launch_ns<PREC, 4, WPG, MINW, PREF>(prm, gpc, cus, s);
    case 5: return launch_ns<PREC, 5, WPG, MINW, PREF>(prm, gpc, cus, s);
    case 6: return launch_ns<PREC, 6, WPG, MINW, PREF>(prm, gpc, cus, s);
    default: return hipErrorInvalidValue;
  }
}

// Launch shape `variant` (kMfmaVariants) of the precision-PREC kernel; each
// precision is instantiated in its own translation unit.
template <int PREC>
hipError_t launch_variant(const MfmaParams& prm, int variant, int slots, int cus,
                          hipStream_t s) {
  switch (variant) {
    case 0: return launch_v<PREC, 8, 1, true>(prm, slots, 0, cus, s);
    case 1: return launch_v<PREC, 4, 2, true>(prm, slots, 0, cus, s);
    case 2: return launch_v<PREC, 4, 3, false>(prm, slots, 0, cus, s);
    case 3: return launch_v<PREC, 2, 2, true>(prm, slots, 0, cus, s);
    case 4: return launch_v<PREC, 4, 2, false>(prm, slots, 0, cus, s);
    case 5: return launch_v<PREC, 2, 3, false>(prm, slots, 0, cus, s);
    default: return hipErrorInvalidValue;
  }
}

extern template hipError_t launch_variant<kPrecF32>(const MfmaParams&, int, int, int, hipStream_t);
extern template hipError_t launch_variant<kPrecF16S>(const MfmaParams&, int, int, int, hipStream_t);
extern template hipError_t launch_variant<kPrecF16>(const MfmaParams&, int, int, int, hipStream_t);
extern template hipError_t launch_variant<kPrecBF16>(const MfmaParams&, int, int, int, hipStream_t);

}  // namespace pcsf_mfma

