// prune_mfma.hpp -- fused whole-tree Felsenstein pruning on MFMA (gfx950).
// The kernel template is instantiated per GEMM precision in its own
// translation unit (prune_mfma_{f32,f16s,f16,bf16}.hip).
//
// Same result as PhyloLik.ensure_alpha + the root posterior
// (lib/CamlPaml/PhyloLik.ml:73-93,127-138) summed by
// PhyloCSFModel.lpr_leaves (src/PhyloCSFModel.ml:67-82), for every packed
// codon column and every model, organised for CDNA4:
//
//  * A wave owns G tiles of 16 columns (G = 1 but for one launch shape);
//    the WPG waves of a workgroup walk the same tree in lock step, so one P
//    matrix staged in LDS feeds 16*G*WPG columns.  Partial likelihoods never leave registers: the host compiles
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
//    alpha).  So in every precision the MFMA only accumulates the
//    off-diagonal mass of P from a zero accumulator (nonnegative terms, no
//    cancellation), and m = alpha + fma(P_aa - 1, alpha, offdiag) is
//    finished in round-to-nearest VALU arithmetic.
//  * Leaf children are gathers of a column of P normalised by its diagonal
//    entry (so the dominant entry is exactly 1; ln P[c][c] is summed per
//    column in FP64), or 1 for `Marginalize; cherries need no GEMM.  A lane
//    mask per live tile marks columns whose partial is a constant vector (a
//    fully marginalised subtree); P c1 = c1 exactly, so those columns skip
//    the GEMM update instead of picking up the rounding of the row sums of
//    P.  After every product each column is rescaled by the power of two of
//    its maximum (exact) and the exponents are summed per column; the root
//    dot and the logarithms are FP64.
//  * P tiles are double-buffered in LDS: while a GEMM runs, the next GEMM's
//    tile is loaded from memory straight into the idle buffer.  The
//    workgroup's leaf codes are staged in LDS once per task, and (launch
//    shapes with gather prefetch) each GEMM also issues the loads of the next
//    leaf gather named by the host's op program.
//  * Workgroups are mapped to models by XCD (blockIdx % 8), so each XCD's
//    L2 holds one model's matrices.
//  * GEMM precision (template PREC): FP32 on v_mfma_f32_16x16x4f32; or P and
//    the partial enter v_mfma_f32_16x16x32_{f16,bf16} in half precision with
//    FP32 accumulation.  FP16 tiles hold offdiag(P) 2^13 (entries below
//    2^13); the live tiles of the FP16 kernels are kept with their column
//    maximum in [2^25, 2^26) by the rescaling itself and enter the
//    GEMM as RN(s 2^-13), maximum in [2^12, 2^13): nothing overflows, values
//    down to 2^-26 of the maximum stay normal, and the product comes out at
//    the tile's own scale.  The split-FP16 precision carries each operand as
//    hi + lo, hi rounded to nearest and lo = the FP16 rounding of the
//    residual scaled by 2^13 (so that it stays a normal number: the matrix
//    cores treat FP16 subnormals as zero, which measurably biased an
//    unscaled lo), and forms hi*hi in one accumulator and hi*lo + lo*hi in
//    another: about 23 significant bits per product at three FP16 MFMAs,
//    against sixteen FP32 MFMAs for the same 64x64x16 product.
//    Live partials stay FP32 in registers in every precision.
#pragma once
#include <type_traits>

#include "../internal.hpp"

namespace pcsf_mfma {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// precisions of the GEMM (values of PCSF_FP32 .. PCSF_BF16)
constexpr int kPrecF32 = PCSF_FP32, kPrecF16S = PCSF_FP16S, kPrecF16 = PCSF_FP16,
              kPrecBF16 = PCSF_BF16;
constexpr float kHalfUnscale = 1.0f / (float)(1 << kHalfExp);  // off the cross terms

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
  // tangent kernel only: D = dP/drho tiles and leaf rows (prune_tangent.hip)
  const float* Dfrag;
  const float* DT32;
  // leaves whose message is gathered (not read through a cherry table): their
  // ln P[c][c] are summed once per task, before the walk
  const int16_t* ls_leaves;
  int n_ls;
  // pi / lnpi hold one root prior per set instead of per model (sets with
  // rate models of their own)
  int prior_per_set;
  // region mode over the model's own sets (set q in every region)
  int shared_sets;
  // the FP64 matrix-core forward only: P of the sets (PSets::d_P64F)
  const double* P64F;
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
    t->set = p.shared_sets ? t->q : (r - p.r0) * p.nm + t->q;
    if (p.active[t->set]) return true;
  }
}

// A 64x16 tile: t[mb][v] = row 16*mb + 4*grp + v of this lane's column.
struct Tile {
  f32x4 r[4];
};

// Raw buffer loads: a wave-uniform base in a buffer resource (scalar
// registers) plus a 32-bit lane offset and a scalar offset, so the address
// arithmetic of every table read is scalar but for one VGPR; reads past
// `bytes` return zero instead of touching memory.
typedef __amdgpu_buffer_rsrc_t Rsrc;
__device__ __forceinline__ Rsrc make_rsrc(const void* base, uint32_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, (int)bytes, 0x00020000);
}
__device__ __forceinline__ f32x4 bload4(Rsrc r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(f32x4, __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, 0));
}
__device__ __forceinline__ double bload_f64(Rsrc r, uint32_t voff, uint32_t soff) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, voff, soff, 0));
}

struct Gathered {
  Tile t;
  bool marg;
};

// Loads the leaf message of leaf `leaf` for this lane's codon: the
// normalised column of P, or row 64 (all ones) for `Marginalize (its
// ln P[c][c] is summed before the walk: leaf_log_sum).
__device__ __forceinline__ void gather(Gathered& g, const MfmaParams& p, const uint8_t* codes_s,
                                       int gc, int set, int leaf, int ccol, int grp) {
  const int code = codes_s[leaf * gc + ccol];
  g.marg = code == PCSF_MARGINALIZE;
  const uint32_t c = g.marg ? 64 : code;
  constexpr uint32_t kRowsBytes = kLeafRows * 64 * 4;  // one branch's leaf rows
  const Rsrc r = make_rsrc(p.PT32 + (size_t)set * p.nbr * (kLeafRows * 64), p.nbr * kRowsBytes);
  const uint32_t off = c * 256u + (uint32_t)grp * 16u;
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) g.t.r[mb] = bload4(r, off + 64u * mb, leaf * kRowsBytes);
}

// Sum of ln P[c][c] over the gathered leaves of this lane's column.
__device__ __forceinline__ double leaf_log_sum(const MfmaParams& p, const uint8_t* codes_s,
                                              int gc, int set, int ccol) {
  double ls = 0.0;
#pragma unroll 8
  for (int i = 0; i < p.n_ls; ++i) {
    const int leaf = p.ls_leaves[i];
    const int code = codes_s[leaf * gc + ccol];
    const int c = code == PCSF_MARGINALIZE ? 64 : code;
    ls += p.lnPcc[((size_t)set * p.nbr + leaf) * kLeafRows + c];
  }
  return ls;
}

__device__ __forceinline__ const Tile& leaf_tile(Gathered& g) { return g.t; }

// s *= o, then every column is scaled by the power of two that brings its
// maximum to [2^(KE-1), 2^KE); E accumulates the exponents, so the true
// product of the column is s * 2^E.
template <int KE>
__device__ __forceinline__ void mul_rescale(Tile& s, const Tile& o, int& E) {
  // column maximum on the bit patterns: positive floats order as signed
  // integers, and negative values (a rounding residue below zero) and -0
  // compare below +0 as they should
  int mx = 0;
#pragma unroll
  for (int mb = 0; mb < 4; ++mb) {
    s.r[mb] *= o.r[mb];
#pragma unroll
    for (int v = 0; v < 4; ++v) mx = max(mx, __float_as_int(s.r[mb][v]));
  }
  // lanes l, l^16, l^32, l^48 hold the column
  // (two cross-row swaps, each returning the value and its partner's: VALU
  // work, no LDS round trip)
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
    for (int v = 0; v < 4; ++v) s.r[mb][v] = __builtin_amdgcn_ldexpf(s.r[mb][v], KE - e);
  E += e - KE;
  // keep the running sum here: left to itself the scheduler of a
  // straight-line program defers every exponent to the root, holding one
  // register per product live across the whole tree
  asm("" : "+v"(E));
}

// G tiles of 16 columns each, walked by one wave side by side: the ops of
// the groups are independent, so the compiler interleaves their dependency
// chains, and one A fragment read from LDS feeds the MFMAs of every group.
template <int G>
struct TileG {
  Tile g[G];
};
template <int G>
struct MaskG {  // per group: lanes whose column is a constant vector
  uint64_t m[G];
};

// hi = RN(x 2^-13) and lo = RN(x - 2^13 hi) of eight FP32 values (a, then
// b), packed in pairs: v_fma_mixlo_f16 / v_fma_mixhi_f16 round one FMA
// straight into the low / high half of a register -- x 2^-13 + 0 and
// hi (-2^13) + x, both exact in FP32 before their one rounding, so the
// halves are those of a multiply and v_cvt_pk_f16_f32 (which the compiler
// emits instead: three instructions per pair for hi and lo each, against
// two).  The hazard recognizer does not see an assembly block as a VALU
// write, so the block ends with the wait states such a write needs before a
// matrix instruction reads its register.
__device__ __forceinline__ void split_f16x8(const f32x4& a, const f32x4& b, f16x8& hi,
                                            f16x8& lo) {
  typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
  u32x4 h, l;
  asm("v_fma_mixlo_f16 %0, %8, %16, 0 op_sel_hi:[0,0,0]\n\t"
      "v_fma_mixhi_f16 %0, %9, %16, 0 op_sel_hi:[0,0,0]\n\t"
      "v_fma_mixlo_f16 %1, %10, %16, 0 op_sel_hi:[0,0,0]\n\t"
      "v_fma_mixhi_f16 %1, %11, %16, 0 op_sel_hi:[0,0,0]\n\t"
      "v_fma_mixlo_f16 %2, %12, %16, 0 op_sel_hi:[0,0,0]\n\t"
      "v_fma_mixhi_f16 %2, %13, %16, 0 op_sel_hi:[0,0,0]\n\t"
      "v_fma_mixlo_f16 %3, %14, %16, 0 op_sel_hi:[0,0,0]\n\t"
      "v_fma_mixhi_f16 %3, %15, %16, 0 op_sel_hi:[0,0,0]\n\t"
      "v_fma_mixlo_f16 %4, %0, %17, %8 op_sel_hi:[1,0,0]\n\t"
      "v_fma_mixhi_f16 %4, %0, %17, %9 op_sel:[1,0,0] op_sel_hi:[1,0,0]\n\t"
      "v_fma_mixlo_f16 %5, %1, %17, %10 op_sel_hi:[1,0,0]\n\t"
      "v_fma_mixhi_f16 %5, %1, %17, %11 op_sel:[1,0,0] op_sel_hi:[1,0,0]\n\t"
      "v_fma_mixlo_f16 %6, %2, %17, %12 op_sel_hi:[1,0,0]\n\t"
      "v_fma_mixhi_f16 %6, %2, %17, %13 op_sel:[1,0,0] op_sel_hi:[1,0,0]\n\t"
      "v_fma_mixlo_f16 %7, %3, %17, %14 op_sel_hi:[1,0,0]\n\t"
      "v_fma_mixhi_f16 %7, %3, %17, %15 op_sel:[1,0,0] op_sel_hi:[1,0,0]\n\t"
      "s_nop 4"
      : "=&v"(h[0]), "=&v"(h[1]), "=&v"(h[2]), "=&v"(h[3]), "=&v"(l[0]), "=&v"(l[1]),
        "=&v"(l[2]), "=&v"(l[3])
      : "v"(a[0]), "v"(a[1]), "v"(a[2]), "v"(a[3]), "v"(b[0]), "v"(b[1]), "v"(b[2]), "v"(b[3]),
        "s"(kHalfUnscale), "s"(-(float)(1 << kHalfExp)));
  hi = __builtin_bit_cast(f16x8, h);
  lo = __builtin_bit_cast(f16x8, l);
}

// s = P s in half precision for the tile staged in LDS (Ab) and each group's
// 64x16 tile.  The tile is kept at 2^kHalfExp above its GEMM operand (FP16:
// operand maximum in [2^12, 2^13)), so the product comes out at the tile's
// scale, and hi = RN(s 2^-13) and the residual RN(s - 2^13 hi) are one
// mixed-precision FMA each.  Per 16-row block mo: hi*hi into acc and (F16S)
// the cross terms hi*lo + lo*hi, which carry 2^kHalfExp, into accx, both from
// zero (the matrix cores' FP32 accumulation rounds toward zero, which would
// bias s if it were the initial value); then the block's message
// m = s + fma(P_aa - 1, s, acc + 2^-13 accx) in round-to-nearest VALU
// arithmetic, which runs while the next block's MFMAs do.  Columns whose
// partial is a constant vector (mask k) keep it.  Tile layout of the halves
// (pbuild.hip): hi fragments, then lo fragments, 512 x 8 values each, in the
// A-operand order of v_mfma_f32_16x16x32_{f16,bf16} with K permuted so that
// K-step st of a lane in group g takes rows 16*(2*st + j/4) + 4*g + j%4 of
// its column: exactly what the accumulator layout leaves in registers
// s.r[2*st + j/4][j%4].  PRIO raises the wave's issue priority over its MFMA
// burst: with three waves per SIMD the arbiter then keeps the matrix pipe fed
// ahead of the other waves' rescaling and combining (measured: 6% on the
// 58mammals forward at that shape, a few percent slower at two waves per
// SIMD, hence per shape).
template <int PREC, int G, bool PRIO>
__device__ __forceinline__ void gemm_half(TileG<G>& s, const MaskG<G>& k, const float* Ab,
                                          int lane, int grp) {
  constexpr bool kBF = PREC == kPrecBF16;
  using V8 = typename std::conditional<kBF, bf16x8, f16x8>::type;
  using E = typename std::conditional<kBF, __bf16, _Float16>::type;
  constexpr float kDown = kBF ? 1.0f : 1.0f / (float)(1 << kHalfExp);
  V8 bh[G][2], bl[G][2];
#pragma unroll
  for (int c = 0; c < G; ++c) {
#pragma unroll
    for (int st = 0; st < 2; ++st) {
      if constexpr (PREC == kPrecF16S) {
        split_f16x8(s.g[c].r[2 * st], s.g[c].r[2 * st + 1], bh[c][st], bl[c][st]);
      } else {
#pragma unroll
        for (int j = 0; j < 8; ++j)
          bh[c][st][j] = (E)(s.g[c].r[2 * st + (j >> 2)][j & 3] * kDown);
      }
    }
  }
  const V8* Ah = reinterpret_cast<const V8*>(Ab);
  const V8* Al = Ah + 512;
  auto mma = [](const V8& a, const V8& b, const f32x4& cc) -> f32x4 {
    if constexpr (kBF) return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, cc, 0, 0, 0);
    else return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, cc, 0, 0, 0);
  };
  TileG<G> m;
  if constexpr (PRIO) __builtin_amdgcn_s_setprio(1);
#pragma unroll
  for (int mo = 0; mo < 4; ++mo) {
    V8 ah[2], al[2];
#pragma unroll
    for (int st = 0; st < 2; ++st) {
      const int f = (mo * 2 + st) * 64 + lane;
      ah[st] = Ah[f];
      if constexpr (PREC == kPrecF16S) al[st] = Al[f];
    }
    const f32x4 dd = *reinterpret_cast<const f32x4*>(Ab + 4096 + 16 * mo + 4 * grp);
#pragma unroll
    for (int c = 0; c < G; ++c) {
      f32x4 acc = (f32x4)(0.0f), accx = (f32x4)(0.0f);
#pragma unroll
      for (int st = 0; st < 2; ++st) {
        if constexpr (PREC == kPrecF16S) {  // cross terms, carrying 2^kHalfExp
          accx = mma(al[st], bh[c][st], accx);
          accx = mma(ah[st], bl[c][st], accx);
        }
        acc = mma(ah[st], bh[c][st], acc);
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        float u = acc[v];
        if constexpr (PREC == kPrecF16S) u = fmaf(accx[v], kHalfUnscale, u);
        m.g[c].r[mo][v] = s.g[c].r[mo][v] + fmaf(dd[v], s.g[c].r[mo][v], u);
      }
    }
  }
  if constexpr (PRIO) __builtin_amdgcn_s_setprio(0);
#pragma unroll
  for (int c = 0; c < G; ++c) {
    if (k.m[c] == 0) {  // no lane holds a constant column (wave-uniform)
      s.g[c] = m.g[c];
    } else {
      const bool cst = (k.m[c] >> lane) & 1;
#pragma unroll
      for (int mo = 0; mo < 4; ++mo) s.g[c].r[mo] = cst ? s.g[c].r[mo] : m.g[c].r[mo];
    }
  }
}

// One wave's walk of the op program over its G x 16 columns, and the ops.
// The interpreter (Interp) and the tree-specialised straight-line programs
// (generated per shipped tree at build time) both drive these members, so
// the arithmetic is the same code in both.
template <int PREC, int WPG, bool PREF, int G, bool PRIO>
struct Walk {
  // live tiles of the FP16 GEMMs are kept with their maximum in
  // [2^25, 2^26): 2^13 above the operand range [2^12, 2^13) gemm_half needs
  static constexpr int KE = (PREC == kPrecF16S || PREC == kPrecF16) ? 2 * kHalfExp : 0;
  static constexpr int NT = 64 * WPG;        // threads
  static constexpr int PER = 4096 / NT / 4;  // float4 of the fragments per thread
  static constexpr int GC = 16 * WPG * G;    // columns per workgroup
  using T = TileG<G>;
  using M = MaskG<G>;
  const MfmaParams& p;
  float (*Pbuf)[kTileFloats];  // LDS: double-buffered P tile
  typedef __attribute__((address_space(3))) float L3;
  L3 (*Pbuf3)[kTileFloats] = nullptr;  // the same, as LDS addresses
  const uint8_t* codes_s;      // LDS: the group's leaf codes, nl x GC
  int tid, lane, grp;
  int wave_s = 0;              // the wave's index in the workgroup (scalar)
  int ccol[G];                 // this lane's columns within the workgroup's GC
  int q = 0, set = 0;          // model and transition set of the current task
  int64_t col[G];              // this lane's columns
  bool has_next = false;       // a next task exists (its first tile is prefetched)
  int next_set = 0;
  int kg = 0;                  // GEMMs done (selects the LDS buffer)
  int E[G];                    // sums of the power-of-two rescaling exponents
  double LS[G];                // sums of ln P[c][c] of the gathered leaves
  Gathered PF[G];              // leaf gathers or table rows issued ahead by the previous GEMM
  double LSpf[G];              // log factor of a table row issued ahead

  __device__ __forceinline__ Walk(const MfmaParams& prm, float (*pb)[kTileFloats],
                                  const uint8_t* cs, int t, int l, int g, int cc0)
      : p(prm), Pbuf(pb), codes_s(cs), tid(t), lane(l), grp(g) {
#pragma unroll
    for (int c = 0; c < G; ++c) {
      ccol[c] = cc0 + 16 * c;
      col[c] = 0;
      E[c] = 0;
      LS[c] = 0.0;
      LSpf[c] = 0.0;
      PF[c].marg = true;
    }
  }

  // The transition set as an opaque scalar: each op recomputes its table
  // addresses from it instead of the compiler keeping every op's address
  // live across a straight-line program (which overflows the SGPRs).
  __device__ __forceinline__ int oset() const {
    int s = set;
    asm volatile("" : "+s"(s));
    return s;
  }
  __device__ __forceinline__ Gathered& first_leaf(int c, int leaf, bool use_pf) {
    if (!(PREF && use_pf)) gather(PF[c], p, codes_s, GC, oset(), leaf, ccol[c], grp);
    return PF[c];
  }
  // s = leaf a0 o leaf a1 (rescaled)
  __device__ __forceinline__ void cherry(T& s, M& k, int a0, int a1, bool use_pf) {
#pragma unroll
    for (int c = 0; c < G; ++c) {
      Gathered g1;
      gather(g1, p, codes_s, GC, oset(), a1, ccol[c], grp);
      Gathered& g0 = first_leaf(c, a0, use_pf);
      s.g[c] = leaf_tile(g0);
      k.m[c] = __ballot(g0.marg && g1.marg);
      mul_rescale<KE>(s.g[c], leaf_tile(g1), E[c]);
    }
  }
  // Loads row (a0, a1) of this set's cherry table `tab` for every column
  // tile: the message into g[c].t, its log factor into ls[c]; m[c] = both
  // leaves marginalised.
  __device__ __forceinline__ void table_row(int st, int a0, int a1, int tab, Gathered (&g)[G],
                                            double (&ls)[G]) {
    constexpr uint32_t kPairs = kTabCodes * kTabCodes;
    const size_t t0 = (size_t)st * p.nctab * kPairs;  // this set's first table entry
    const Rsrc rt = make_rsrc(p.ctab + t0 * 64, p.nctab * kPairs * 256u);
    const Rsrc rl = make_rsrc(p.clnt + t0, p.nctab * kPairs * 8u);
#pragma unroll
    for (int c = 0; c < G; ++c) {
      const int ca = codes_s[a0 * GC + ccol[c]], cb = codes_s[a1 * GC + ccol[c]];
      const bool ma = ca == PCSF_MARGINALIZE, mb = cb == PCSF_MARGINALIZE;
      const uint32_t e = (uint32_t)(ma ? 64 : ca) * kTabCodes + (mb ? 64 : cb);
#pragma unroll
      for (int mbk = 0; mbk < 4; ++mbk)
        g[c].t.r[mbk] =
            bload4(rt, e * 256u + (uint32_t)grp * 16u + 64u * mbk, tab * kPairs * 256u);
      ls[c] = bload_f64(rl, e * 8u, tab * kPairs * 8u);
      g[c].marg = ma && mb;
    }
  }
  // s = message above the cherry (a0, a1), from table `tab`: its row issued
  // ahead by the preceding GEMM (use_pf), or loaded now
  __device__ __forceinline__ void cherryt(T& s, M& k, int a0, int a1, int tab, bool use_pf) {
    if (!(PREF && use_pf)) table_row(oset(), a0, a1, tab, PF, LSpf);
#pragma unroll
    for (int c = 0; c < G; ++c) {
      s.g[c] = PF[c].t;
      // (a table message always meets a sibling in LEAFMUL / MULPOP, whose
      // rescaling sets the tile scale, before it reaches a GEMM)
      LS[c] += LSpf[c];
      asm("" : "+v"(LS[c]));  // as E in mul_rescale
      k.m[c] = __ballot(PF[c].marg);
    }
  }
  // s = s o leaf a0 (rescaled)
  __device__ __forceinline__ void leafmul(T& s, M& k, int a0, bool use_pf) {
#pragma unroll
    for (int c = 0; c < G; ++c) {
      Gathered& g0 = first_leaf(c, a0, use_pf);
      k.m[c] &= __ballot(g0.marg);
      mul_rescale<KE>(s.g[c], leaf_tile(g0), E[c]);
    }
  }
  // s = s o o (rescaled)
  __device__ __forceinline__ void mulpop(T& s, M& k, const T& o, const M& ko) {
#pragma unroll
    for (int c = 0; c < G; ++c) {
      k.m[c] &= ko.m[c];
      mul_rescale<KE>(s.g[c], o.g[c], E[c]);
    }
  }
  // s = P s for the P tile staged in LDS; prefetches the tile of branch
  // next_br (or the next task's first tile) and, if the program names one,
  // the next gather op's first leaf (pf_leaf) or table row (pf_tab:
  // kPfTable | table << 24 | leaf1 << 12 | leaf0)
  __device__ __forceinline__ void gemm(T& s, const M& k, int next_br, int pf_leaf,
                                       uint32_t pf_tab) {
    __syncthreads();  // Pbuf[kg&1] complete; Pbuf[(kg&1)^1] no longer read
    const int st = oset();
    if (PREF && pf_leaf != kNoBranch) {
#pragma unroll
      for (int c = 0; c < G; ++c) gather(PF[c], p, codes_s, GC, st, pf_leaf, ccol[c], grp);
    } else if (PREF && (pf_tab & kPfTable)) {
      table_row(st, pf_tab & kPfLeafMax, (pf_tab >> 12) & kPfLeafMax, (pf_tab >> 24) & kPfTableMax,
                PF, LSpf);
    }
    // the next tile, of this set (buffer loads from the set's tiles) or the
    // next task's first (plain loads, once per task), goes straight from
    // memory into the idle LDS buffer: no registers, no LDS stores, and the
    // barrier that opens the next GEMM waits for it.  Each instruction
    // writes 16 bytes per lane at the wave's LDS base + 16 * lane, the
    // layout the tile has in memory.
    {
      L3* dst = Pbuf3[(kg & 1) ^ 1];
      typedef __attribute__((address_space(3))) void* LdsPtr;
      if (next_br != kNoBranch) {
        constexpr uint32_t kTileBytes = kTileFloats * 4;
        const Rsrc r = make_rsrc(p.Pfrag + (size_t)st * p.nbr * kTileFloats, p.nbr * kTileBytes);
        const uint32_t so = next_br * kTileBytes;
#pragma unroll
        for (int q4 = 0; q4 < PER; ++q4)
          __builtin_amdgcn_raw_ptr_buffer_load_lds(
              r, (LdsPtr)(dst + (64 * wave_s + NT * q4) * 4), 16, (uint32_t)tid * 16u,
              so + 16u * NT * q4, 0, 0);
        if (tid < 16)
          __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (LdsPtr)(dst + 4096), 16,
                                                   (uint32_t)tid * 16u, so + 16384u, 0, 0);
      } else if (has_next) {
        typedef const __attribute__((address_space(1))) void* GPtr;
        const float* nsrc = p.Pfrag + ((size_t)next_set * p.nbr + p.first_gemm_br) * kTileFloats;
#pragma unroll
        for (int q4 = 0; q4 < PER; ++q4)
          __builtin_amdgcn_global_load_lds((GPtr)(nsrc + (tid + NT * q4) * 4),
                                           (LdsPtr)(dst + (64 * wave_s + NT * q4) * 4), 16, 0, 0);
        if (tid < 16)
          __builtin_amdgcn_global_load_lds((GPtr)(nsrc + 4096 + tid * 4), (LdsPtr)(dst + 4096),
                                           16, 0, 0);
      }
    }
    const float* Ab = Pbuf[kg & 1];
    if constexpr (PREC == kPrecF32) {
      f32x4 acc[G][4];
#pragma unroll
      for (int c = 0; c < G; ++c)
#pragma unroll
        for (int mo = 0; mo < 4; ++mo) acc[c][mo] = (f32x4)(0.0f);
      if constexpr (PRIO) __builtin_amdgcn_s_setprio(1);
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) {
        f32x4 A[4];
#pragma unroll
        for (int mo = 0; mo < 4; ++mo)
          A[mo] = *reinterpret_cast<const f32x4*>(Ab + ((mo * 4 + mb) * 64 + lane) * 4);
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
          for (int c = 0; c < G; ++c) {
            const float bv = s.g[c].r[mb][v];
#pragma unroll
            for (int mo = 0; mo < 4; ++mo)
              acc[c][mo] = __builtin_amdgcn_mfma_f32_16x16x4f32(A[mo][v], bv, acc[c][mo], 0, 0, 0);
          }
      }
      if constexpr (PRIO) __builtin_amdgcn_s_setprio(0);
      // m = alpha + (P_aa - 1) alpha + offdiag(P) alpha, the diagonal term in
      // round-to-nearest VALU arithmetic
#pragma unroll
      for (int mo = 0; mo < 4; ++mo) {
        const f32x4 dd = *reinterpret_cast<const f32x4*>(Ab + 4096 + 16 * mo + 4 * grp);
#pragma unroll
        for (int c = 0; c < G; ++c)
#pragma unroll
          for (int v = 0; v < 4; ++v)
            acc[c][mo][v] = s.g[c].r[mo][v] + fmaf(dd[v], s.g[c].r[mo][v], acc[c][mo][v]);
      }
#pragma unroll
      for (int c = 0; c < G; ++c) {
        if (k.m[c] == 0) {  // no lane holds a constant column (wave-uniform)
#pragma unroll
          for (int mo = 0; mo < 4; ++mo) s.g[c].r[mo] = acc[c][mo];
        } else {
          const bool cst = (k.m[c] >> lane) & 1;
#pragma unroll
          for (int mo = 0; mo < 4; ++mo) s.g[c].r[mo] = cst ? s.g[c].r[mo] : acc[c][mo];
        }
      }
    } else {
      gemm_half<PREC, G, PRIO>(s, k, Ab, lane, grp);
    }
    ++kg;
  }
  // ln z and elpr of this lane's columns from the root tiles
  __device__ __forceinline__ void root(const T& s) {
    const int pq = p.prior_per_set ? set : q;
    const double* piq = p.pi + pq * 64;
    const double* lnq = p.lnpi + pq * 64;
#pragma unroll
    for (int c = 0; c < G; ++c) {
      double z = 0.0, el = 0.0;
#pragma unroll
      for (int mb = 0; mb < 4; ++mb)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = 16 * mb + 4 * grp + v;
          const double x = (double)s.g[c].r[mb][v] * piq[row];
          z += x;
          el = fma(x, lnq[row], el);
        }
      z += __shfl_xor(z, 16);
      el += __shfl_xor(el, 16);
      z += __shfl_xor(z, 32);
      el += __shfl_xor(el, 32);
      if (grp == 0) {
        p.logz[(size_t)q * p.ld + col[c]] = log(z) + (double)E[c] * 0.69314718055994530942 + LS[c];
        p.colelpr[(size_t)q * p.ld + col[c]] = (z == 0.0) ? 0.0 : el / z;
      }
    }
  }
};

// The op-program interpreter: reads the host's program (four words per op)
// and dispatches on (opcode, slot).
template <int NS>
struct Interp {
  template <class W, class T, class M>
  __device__ __forceinline__ static void run(W& w, T (&S)[NS], M (&K)[NS]) {
    const int32_t* prog = w.p.prog;
    for (int pc = 0;; ++pc) {
      const int w0 = __builtin_amdgcn_readfirstlane(prog[4 * pc]);
      const int w1 = __builtin_amdgcn_readfirstlane(prog[4 * pc + 1]);
      const int w2 = __builtin_amdgcn_readfirstlane(prog[4 * pc + 2]);
      const uint32_t w3 = __builtin_amdgcn_readfirstlane(prog[4 * pc + 3]);
      const int op = w0 & 0xff, d = (w0 >> 8) & 0xff;
      const bool use_pf = (w0 >> 16) & 1;
      const int a0 = w1 & 0xffff, a1 = (w1 >> 16) & 0xffff;
      if (op == OP_END) break;
#define PCSF_CASES(D)                                          \
  case OP_CHERRY * 8 + D:                                      \
    if constexpr (D < NS) w.cherry(S[D], K[D], a0, a1, use_pf); \
    break;                                                     \
  case OP_GEMM * 8 + D:                                        \
    if constexpr (D < NS) w.gemm(S[D], K[D], a1, w2 & 0xffff, w3); \
    break;                                                     \
  case OP_LEAFMUL * 8 + D:                                     \
    if constexpr (D < NS) w.leafmul(S[D], K[D], a0, use_pf);    \
    break;                                                     \
  case OP_CHERRYT * 8 + D:                                     \
    if constexpr (D < NS) w.cherryt(S[D], K[D], a0, a1, w2, use_pf); \
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

template <int PREC, int NS, int WPG, int MINW, bool PREF, class PROG, int G>
__global__ __launch_bounds__(64 * WPG, MINW) void prune_mfma_kernel(MfmaParams p) {
  using W = Walk<PREC, WPG, PREF, G, (MINW >= 3)>;
  constexpr int NT = W::NT, PER = W::PER, GC = W::GC;
  static_assert(NT >= 16, "the diagonal (16 float4) is copied by the first 16 threads");
  __shared__ __attribute__((aligned(16))) float Pbuf[2][kTileFloats];
  extern __shared__ __attribute__((aligned(16))) uint8_t codes_s[];  // nl x GC
  const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const int grp = lane >> 4, lcol = lane & 15;

  TileG<G> S[NS];
  MaskG<G> K[NS];
#pragma unroll
  for (int d = 0; d < NS; ++d)
#pragma unroll
    for (int c = 0; c < G; ++c) {
      K[d].m[c] = 0;
#pragma unroll
      for (int mb = 0; mb < 4; ++mb) S[d].g[c].r[mb] = (f32x4)(0.0f);
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

  // wave w holds columns [16 G w, 16 G (w + 1)) of the workgroup's group
  W w(p, Pbuf, codes_s, tid, lane, grp, wave * 16 * G + lcol);
  w.Pbuf3 = (typename W::L3 (*)[kTileFloats])Pbuf;
  w.wave_s = wave;
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
    w.has_next = has_next;
    w.next_set = has_next ? nxt.set : 0;
#pragma unroll
    for (int c = 0; c < G; ++c) {
      w.col[c] = gabs * GC + w.ccol[c];
      w.E[c] = 0;
      w.LS[c] = leaf_log_sum(p, codes_s, GC, cur.set, w.ccol[c]);
    }
    PROG::run(w, S, K);
    if (!has_next) break;
    cur = nxt;
    jc = jn;
  }
}

template <int PREC, int NS, int WPG, int MINW, bool PREF, class PROG = Interp<NS>, int G = 1>
hipError_t launch_ns(const MfmaParams& prm, int grid_per_cu, int num_cus, hipStream_t s) {
  auto kern = prune_mfma_kernel<PREC, NS, WPG, MINW, PREF, PROG, G>;
  const size_t dyn = ((size_t)prm.nl * 16 * WPG * G + 15) / 16 * 16;  // staged leaf codes
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

template <int PREC, int WPG, int MINW, bool PREF, int G = 1>
hipError_t launch_v(const MfmaParams& prm, int slots, int gpc, int cus, hipStream_t s) {
  switch (slots) {
    case 1:
    case 2: return launch_ns<PREC, 2, WPG, MINW, PREF, Interp<2>, G>(prm, gpc, cus, s);
    case 3: return launch_ns<PREC, 3, WPG, MINW, PREF, Interp<3>, G>(prm, gpc, cus, s);
    case 4: return launch_ns<PREC, 4, WPG, MINW, PREF, Interp<4>, G>(prm, gpc, cus, s);
    case 5: return launch_ns<PREC, 5, WPG, MINW, PREF, Interp<5>, G>(prm, gpc, cus, s);
    case 6: return launch_ns<PREC, 6, WPG, MINW, PREF, Interp<6>, G>(prm, gpc, cus, s);
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
    case 5: return launch_v<PREC, 4, 3, true>(prm, slots, 0, cus, s);
    case 6: return launch_v<PREC, 2, 1, false, 2>(prm, slots, 0, cus, s);
    default: return hipErrorInvalidValue;
  }
}

extern template hipError_t launch_variant<kPrecF32>(const MfmaParams&, int, int, int, hipStream_t);
extern template hipError_t launch_variant<kPrecF16S>(const MfmaParams&, int, int, int, hipStream_t);
extern template hipError_t launch_variant<kPrecF16>(const MfmaParams&, int, int, int, hipStream_t);
extern template hipError_t launch_variant<kPrecBF16>(const MfmaParams&, int, int, int, hipStream_t);

}  // namespace pcsf_mfma
