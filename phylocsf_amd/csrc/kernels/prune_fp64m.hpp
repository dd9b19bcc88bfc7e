// prune_fp64m.hpp -- the FP64 forward on matrix cores (precision
// PCSF_FP64M), driven by an op program: the interpreter (prune_fp64m.hip) or
// a tree's straight-line program (generated per shipped tree).
//
// Same result as PhyloLik.ensure_alpha and the root posterior
// (lib/CamlPaml/PhyloLik.ml:73-93,127-138), per column ln z and the root's
// expected ln pi (the elpr_anc term of src/PhyloCSF.ml), in FP64 throughout
// with no rescaling, as the reference.  Organised like the fused MFMA kernel
// (prune_mfma.hpp): a wave owns 16 columns; the four waves of a workgroup
// walk the host's op program (post-order, fewest live tiles) in lock step;
// partials never leave registers -- a 64 x 16 FP64 tile is 32 VGPRs in the
// accumulator layout of v_mfma_f64_16x16x4f64 (fp64_tiles.hpp), which is
// also the next product's B operand -- and each internal-child message
// m = P alpha is a 64 x 64 by 64 x 16 product on those MFMAs with P staged in
// LDS.  The P tiles are double-buffered: every product issues the next
// tile's loads straight from memory into the idle buffer (the sets' P64F
// holds P in the MFMA's A-fragment order, so the copy is a plain one), and
// the barrier opening the next product is the only wait.  Leaf messages are
// the columns of P (the leaf-branch layout of P64F keeps each lane group's
// 16 rows of a column contiguous), all ones for `Marginalize.  Paths are
// relative to the mlin/PhyloCSF tree.
#pragma once
#include "fp64_tiles.hpp"
#include "prune_mfma.hpp"

namespace pcsf_mfma {

namespace f64m {

using Tile64 = fp64t::Tile;
using fp64t::f64x4;

constexpr int kWaves = fp64t::kWaves;
constexpr int kNT = 64 * kWaves;     // threads per workgroup
constexpr int kGC = fp64t::kCols;    // columns per workgroup
constexpr int kTileBytes = 4096 * 8; // one P tile

struct NoMask {};  // FP64 needs no constant-column masks (no rescaling)

struct Walk64 {
  typedef __attribute__((address_space(3))) double L3;
  typedef __attribute__((address_space(3))) void* LdsPtr;
  const MfmaParams& p;
  const double (*Pbuf)[4096];  // LDS: double-buffered P tile
  L3 (*Pbuf3)[4096];           // the same, as LDS addresses
  const uint8_t* codes_s;      // LDS: the group's leaf codes, nl x kGC
  int tid, lane, grp, wave_s, ccol;
  int q = 0, set = 0;
  int64_t col = 0;
  bool has_next = false;
  int next_set = 0;
  int kg = 0;  // products done (selects the LDS buffer)

  __device__ __forceinline__ int oset() const {
    int s = set;
    asm volatile("" : "+s"(s));
    return s;
  }
  // the message of leaf `leaf` for this lane's codon (message 64: ones)
  __device__ __forceinline__ Tile64 leaf(int leaf) const {
    const int code = codes_s[leaf * kGC + ccol];
    const uint32_t c = code == PCSF_MARGINALIZE ? 64u : (uint32_t)code;
    constexpr uint32_t kBranchBytes = kTile64 * 8;
    const Rsrc r = make_rsrc(p.P64F + (size_t)oset() * p.nbr * kTile64,
                                        p.nbr * kBranchBytes);
    const uint32_t off = c * 512u + (uint32_t)grp * 128u;
    Tile64 t;
#pragma unroll
    for (int mo = 0; mo < 4; ++mo)
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const auto x = __builtin_amdgcn_raw_buffer_load_b128(r, off + 32u * mo + 16u * h,
                                                             leaf * kBranchBytes, 0);
        const double2 d = __builtin_bit_cast(double2, x);
        t.r[mo][2 * h] = d.x;
        t.r[mo][2 * h + 1] = d.y;
      }
    return t;
  }
  __device__ __forceinline__ static void mul(Tile64& s, const Tile64& o) {
#pragma unroll
    for (int mo = 0; mo < 4; ++mo) s.r[mo] *= o.r[mo];
  }
  // s = leaf a0 o leaf a1
  __device__ __forceinline__ void cherry(Tile64& s, NoMask&, int a0, int a1, bool) {
    s = leaf(a0);
    mul(s, leaf(a1));
  }
  // the no-table programs have no table reads
  __device__ __forceinline__ void cherryt(Tile64&, NoMask&, int, int, int, bool) {}
  // s = s o leaf a0
  __device__ __forceinline__ void leafmul(Tile64& s, NoMask&, int a0, bool) { mul(s, leaf(a0)); }
  // s = s o o
  __device__ __forceinline__ void mulpop(Tile64& s, NoMask&, const Tile64& o, const NoMask&) {
    mul(s, o);
  }
  // s = P s for the tile staged in LDS; the next product's tile (branch
  // next_br of this set, or the next task's first) is loaded meanwhile into
  // the idle buffer, 16 bytes per lane at the wave's base + 16 lane: the
  // layout the tile has in memory
  __device__ __forceinline__ void gemm(Tile64& s, const NoMask&, int next_br, int, uint32_t) {
    __syncthreads();  // Pbuf[kg&1] complete; Pbuf[(kg&1)^1] no longer read
    L3* dst = Pbuf3[(kg & 1) ^ 1];
    constexpr int kPer = kTileBytes / 16 / kNT;  // 16-byte pieces per thread
    if (next_br != kNoBranch) {
      const Rsrc r = make_rsrc(p.P64F + (size_t)oset() * p.nbr * kTile64,
                                          p.nbr * kTile64 * 8u);
      const uint32_t so = next_br * (kTile64 * 8u);
#pragma unroll
      for (int k = 0; k < kPer; ++k)
        __builtin_amdgcn_raw_ptr_buffer_load_lds(r, (LdsPtr)(dst + (64 * wave_s + kNT * k) * 2),
                                                 16, (uint32_t)tid * 16u, so + 16u * kNT * k, 0,
                                                 0);
    } else if (has_next) {
      typedef const __attribute__((address_space(1))) void* GPtr;
      const double* nsrc = p.P64F + ((size_t)next_set * p.nbr + p.first_gemm_br) * kTile64;
#pragma unroll
      for (int k = 0; k < kPer; ++k)
        __builtin_amdgcn_global_load_lds((GPtr)(nsrc + (tid + kNT * k) * 2),
                                         (LdsPtr)(dst + (64 * wave_s + kNT * k) * 2), 16, 0, 0);
    }
    s = fp64t::gemm(Pbuf[kg & 1], s, lane);
    ++kg;
  }
  // ln z and the root posterior's expected ln pi of this lane's column
  __device__ __forceinline__ void root(const Tile64& s) {
    const int pq = p.prior_per_set ? set : q;
    Tile64 pr, w;
#pragma unroll
    for (int mo = 0; mo < 4; ++mo)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = 16 * mo + 4 * v + grp;
        pr.r[mo][v] = p.pi[pq * 64 + row];
        w.r[mo][v] = s.r[mo][v] * pr.r[mo][v];
        pr.r[mo][v] = p.lnpi[pq * 64 + row];
      }
    Tile64 one;
#pragma unroll
    for (int mo = 0; mo < 4; ++mo) one.r[mo] = (f64x4){1.0, 1.0, 1.0, 1.0};
    const double z = fp64t::column_dot(w, one);
    const double e = fp64t::column_dot(w, pr);  // sum_x alpha_x pi_x ln pi_x
    if (grp == 0) {
      p.logz[(size_t)q * p.ld + col] = log(z);
      p.colelpr[(size_t)q * p.ld + col] = (z != 0.0) ? e / z : 0.0;
    }
  }
};

// MINB workgroups per CU: two (a generated program keeps only its live
// tiles in registers), one for the interpreter above three live tiles (it
// keeps every slot live through its dispatch; at two per CU it spills)
template <int NS, class PROG, int MINB>
__global__ __launch_bounds__(kNT, MINB) void prune_fp64m_kernel(MfmaParams p) {
  __shared__ __attribute__((aligned(16))) double Pbuf[2][4096];
  extern __shared__ __attribute__((aligned(16))) uint8_t codes_s[];  // nl x kGC
  const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), lane = tid & 63;
  const int grp = lane >> 4, lcol = lane & 15;
  Tile64 S[NS];
  NoMask K[NS];
#pragma unroll
  for (int d = 0; d < NS; ++d)
#pragma unroll
    for (int mo = 0; mo < 4; ++mo) S[d].r[mo] = (f64x4){0.0, 0.0, 0.0, 0.0};

  Task cur;
  int64_t jc = 0;
  if (!next_task<kGC>(p, jc, &cur)) return;
  if (p.first_gemm_br != kNoBranch) {
    const double* src = p.P64F + ((size_t)cur.set * p.nbr + p.first_gemm_br) * kTile64;
#pragma unroll
    for (int k = 0; k < 4096 / 2 / kNT; ++k)
      reinterpret_cast<double2*>(Pbuf[0])[tid + kNT * k] =
          reinterpret_cast<const double2*>(src)[tid + kNT * k];
  }
  Walk64 w{p, Pbuf, (Walk64::L3 (*)[4096])Pbuf, codes_s, tid, lane, grp, wave,
           wave * 16 + lcol};
  for (;;) {
    Task nxt;
    int64_t jn = jc + 1;
    const bool has_next = next_task<kGC>(p, jn, &nxt);
    const int64_t gabs = p.g0 + cur.g;
    __syncthreads();  // the previous task's gathers are done with codes_s
    {
      const uint32_t* src = reinterpret_cast<const uint32_t*>(p.codes);
      uint32_t* dst = reinterpret_cast<uint32_t*>(codes_s);
      constexpr int words = kGC / 4;
      for (int e = tid; e < p.nl * words; e += kNT) {
        const int l = e / words, wd = e - l * words;
        dst[e] = src[((size_t)l * p.ld + (size_t)gabs * kGC) / 4 + wd];
      }
    }
    __syncthreads();
    w.q = cur.q;
    w.set = cur.set;
    w.has_next = has_next;
    w.next_set = has_next ? nxt.set : 0;
    w.col = gabs * kGC + w.ccol;
    PROG::run(w, S, K);
    if (!has_next) break;
    cur = nxt;
    jc = jn;
  }
}

template <int NS, class PROG, int MINB = 2>
hipError_t launch_ns(const MfmaParams& prm, int cus, hipStream_t s) {
  auto kern = prune_fp64m_kernel<NS, PROG, MINB>;
  const size_t dyn = ((size_t)prm.nl * kGC + 15) / 16 * 16;
  int per_cu = 0;
  hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kNT, dyn);
  if (e != hipSuccess) return e;
  const int grid = (std::max(per_cu, 1) * cus + 7) / 8 * 8;
  hipLaunchKernelGGL(kern, dim3(grid), dim3(kNT), dyn, s, prm);
  return hipGetLastError();
}

}  // namespace f64m

// the forward of a tree's generated program PROG (TreeLaunch signature;
// precision and launch shape do not apply)
template <class PROG>
hipError_t launch_fp64m_tree(const MfmaParams& prm, int, int, int cus, hipStream_t s) {
  return f64m::launch_ns<PROG::NS, PROG>(prm, cus, s);
}

}  // namespace pcsf_mfma
