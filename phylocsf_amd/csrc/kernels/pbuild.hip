// pbuild.hip -- FP64 transition matrices P(t) = V diag(exp(lam t)) V^-1 on the GPU.
//
// Set s is model s % nm at tree scale rho[s] and is skipped when
// active[s] == 0; a workgroup makes the matrices of a few consecutive
// branches of one set.  Reproduces Q.Diag.real_to_Pt
// (lib/CamlPaml/Q.ml:211-249): entries below zero are an error when
// |x| > 1e-6 and are set to 0 otherwise; the row sum must be within 1e-6 of
// 1; the diagonal is replaced by 1 - sum(off-diagonal).  The failure the
// reference raises first -- lowest set, then branch, then row, then column,
// with a row's negative entries before its sum and diagonal checks -- is the
// minimum of a packed key over every failing (set, branch, row), taken with
// one 64-bit atomicMin, so the record does not depend on which workgroup ran
// first (fail_key below; all ones when nothing failed).
//
// Outputs, all [set][branch][...] (which of them: kBuild* of launch_pbuild);
// a leaf branch (br < nl) is read only
// through its leaf messages and an internal one only through its tile, so
// PT32 and lnPcc are written for leaf branches and Pfrag for internal ones:
//   P64T  FP64 P transposed (P64T[b][a] = P[a][b]): coalesced GEMV operand of
//         the FP64 kernel and its leaf gathers.
//   PT32  FP32 leaf messages of the MFMA kernel, 65 rows of 64: row c < 64 is
//         P[a][c] / P[c][c] with lnPcc[c] = ln P[c][c] in FP64 (the dominant
//         entry of a leaf message is exactly 1 and its value travels in FP64);
//         row 64 is the all-ones message of `Marginalize (lnPcc 0).
//   Pfrag tile of kTileFloats floats in format fmt, in the A-operand order of
//         the MFMA the kernel uses: FP32 off-diagonal of P (diagonal entries
//         zero) for v_mfma_f32_16x16x4f32, or hi and lo halves of the same
//         off-diagonal (FP16 scaled by 2^kHalfExp, or BF16) for
//         v_mfma_f32_16x16x32_{f16,bf16};
//         followed by the 64 FP32 values P[a][a] - 1 (the diagonal term, which
//         every precision adds in VALU arithmetic).
//   P64F  FP64 P for the matrix-core forward (kTile64 doubles per branch):
//         an internal branch's P in v_mfma_f64_16x16x4f64 A-fragment order,
//         a leaf branch's 65 messages (columns of P, then all ones)
//   P64FT FP64 P^T of internal branches in that A-fragment order (the outside
//         pass's products)
//   DT32, Dfrag  (tangent sets only) D = dP(rho b)/drho = b Q P in the layouts
//         of PT32 (divided by P[c][c]; row 64 zero) and of the FP32 tile
//         (dense), for leaf and internal branches respectively.
#include "../internal.hpp"

namespace {

constexpr double kTol = 1e-6;

typedef double f64x4 __attribute__((ext_vector_type(4)));

// Sum of v over the lanes xor 16 (row swap) and xor 32 (half swap): each swap
// hands a lane its own value and its partner's, and the two are added in
// either order to the same result, so all four lanes end with the same sum.
__device__ __forceinline__ double lanes_sum(double v) {
  {
    const auto l = __builtin_amdgcn_permlane16_swap(__double2loint(v), __double2loint(v), false,
                                                    false);
    const auto h = __builtin_amdgcn_permlane16_swap(__double2hiint(v), __double2hiint(v), false,
                                                    false);
    v = __hiloint2double(h[0], l[0]) + __hiloint2double(h[1], l[1]);
  }
  {
    const auto l = __builtin_amdgcn_permlane32_swap(__double2loint(v), __double2loint(v), false,
                                                    false);
    const auto h = __builtin_amdgcn_permlane32_swap(__double2hiint(v), __double2hiint(v), false,
                                                    false);
    v = __hiloint2double(h[0], l[0]) + __hiloint2double(h[1], l[1]);
  }
  return v;
}

// P^T = (V^-T diag(e)) V^T with e = exp(lam t), on v_mfma_f64_16x16x4f64.
// Wave w makes rows 16w..16w+15 of P^T (columns of P) as four 16x16 tiles
// ct; at step s (k = 4s..4s+3)
//   A[i][k] = V^-1[k][16w + i] e[k]   in lane i + 16 (k % 4)
//   B[k][j] = V[16 ct + j][k]          in lane j + 16 (k % 4)
//   D[i][j]                            in lane j + 16 (i % 4), element i / 4
// so element r of tile ct in lane (lr, lk) = (lane % 16, lane / 16) is
//   P[a][b] with a = 16 ct + lr, b = 16 w + 4 r + lk.
// In this layout every output below is written straight from the
// accumulators: P^T rows are 16 consecutive entries per lane group and the
// FP32 tile order is 64 consecutive floats per store.  With WITH_D the same
// B also gives D^T = (V^-T diag(b lam e)) V^T, D = dP(rho b)/drho = b Q P.
//
// A workgroup makes the matrices of `nbw` consecutive branches of one set:
// A and B do not depend on t, so each wave reads its 16 A values (V^-1
// unscaled) and B = V^T sits in LDS for all of them; only e changes from
// branch to branch.  B is stored with row k's 16-column halves swapped when
// k is odd, so the two rows a half-wave reads per step fall in disjoint banks.
// Failure key: ((set * nbr + br) * 64 + row) * 65 + col, col = 64 for the
// row-sum and diagonal checks (made after the row's entries), shifted left by
// 2 over the check: 0 t < 0 (raised before any row), 1 negative entry, 2 row
// sum, 3 diagonal.  The host decodes it and recomputes the message.
__device__ __forceinline__ unsigned long long fail_key(int set, int nbr, int br, int row, int col,
                                                      int check) {
  const unsigned long long pos =
      (((unsigned long long)set * nbr + br) * 64 + row) * 65 + col;
  return (pos << 2) | (unsigned long long)check;
}

template <bool WITH_D>
__global__ __launch_bounds__(256, WITH_D ? 2 : 3) void pbuild_kernel(
    const double* __restrict__ VT, const double* __restrict__ Vinv,
    const double* __restrict__ lam, const double* __restrict__ branch,
    const double* __restrict__ rho, const uint8_t* __restrict__ active, int nm, int nbr,
    int nbw, int nl, double* __restrict__ P64T, float* __restrict__ PT32,
    float* __restrict__ Pfrag, int fmt, double* __restrict__ lnPcc,
    unsigned long long* __restrict__ perr, float* __restrict__ DT32,
    float* __restrict__ Dfrag, double* __restrict__ P64F, double* __restrict__ P64FT) {
  __shared__ double Bs[64 * 64];  // V^T, swizzled
  __shared__ double ptot[4][64], poff[4][64];  // per wave, per row of P
  __shared__ int pjbad[4][64];
  __shared__ double pdiag[64], dinv[64];  // fixed-up diagonal and its inverse
  const int set = blockIdx.y, q = set % nm, tid = threadIdx.x;
  if (active && !active[set]) return;
  const int br0 = blockIdx.x * nbw, br1 = min(nbr, br0 + nbw);
  for (int e = tid; e < 4096; e += 256)
    Bs[(e & ~63) | ((e & 63) ^ (((e >> 6) & 1) << 4))] = VT[(size_t)q * 4096 + e];
  const double lm = lam[q * 64 + (tid & 63)];
  __syncthreads();
  for (int br = br0; br < br1; ++br) {
    // the lane indices are made opaque at every branch, so the compiler
    // cannot hoist the many lane-dependent addresses and masks below out of
    // the loop, where they would hold registers for its whole length
    int tl = tid;
    asm volatile("" : "+v"(tl));
    const int w = tid >> 6, lane = tl & 63, lr = lane & 15, lk = lane >> 4;
    const double* A = Vinv + (size_t)q * 4096 + 16 * w + lr;
    const double bl = branch[br];
    const double t = rho[set] * bl;
    if (t < 0.0 && tid == 0) atomicMin(perr, fail_key(set, nbr, br, 0, 0, 0));
    f64x4 acc[4] = {};
    f64x4 dacc[WITH_D ? 4 : 1] = {};
    {
      // every wave makes e[lane] itself and takes e[k] from lane k
      const double el = exp(t * lm);
      const double dl = WITH_D ? bl * lm * el : 0.0;
      const int sw = (lk & 1) << 4;
      __builtin_amdgcn_s_setprio(1);
#pragma unroll 4
      for (int s = 0; s < 16; ++s) {
        const int k = 4 * s + lk;
        const double v = A[k * 64];
        const double a = v * __shfl(el, k);
        double ad = 0.0;
        if constexpr (WITH_D) ad = v * __shfl(dl, k);
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
          // column 16 ct + lr of B as stored in row k
          const double bv = Bs[k * 64 + ((16 * ct + lr) ^ sw)];
          acc[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, bv, acc[ct], 0, 0, 0);
          if constexpr (WITH_D)
            dacc[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(ad, bv, dacc[ct], 0, 0, 0);
        }
      }
      __builtin_amdgcn_s_setprio(0);
    }
    // Row fix-up (rows a of P).  An entry below -kTol fails the build: only
    // then is each row's first such column searched for (the four lanes
    // sharing lr, then the four waves in column order).  Otherwise every
    // lane clamps its entries at 0 and sums them, raw and clamped, with the
    // clamped diagonal taken out of the latter; the lanes sharing lr and then
    // the waves join the sums.
    {
      double mn = acc[0][0];
#pragma unroll
      for (int ct = 0; ct < 4; ++ct)
#pragma unroll
        for (int r = 0; r < 4; ++r) mn = fmin(mn, acc[ct][r]);
      if (__any(mn < -kTol)) {
        int jb[4];
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
          jb[ct] = 64;
#pragma unroll
          for (int r = 0; r < 4; ++r)
            if (acc[ct][r] < -kTol && jb[ct] == 64) jb[ct] = 16 * w + 4 * r + lk;
#pragma unroll
          for (int m = 16; m <= 32; m <<= 1) jb[ct] = min(jb[ct], __shfl_xor(jb[ct], m));
        }
        // lane (lr, lk) stores row a = 16 lk + lr = lane, i.e. tile ct = lk
        pjbad[w][lane] = lk == 0 ? jb[0] : lk == 1 ? jb[1] : lk == 2 ? jb[2] : jb[3];
      } else {
        pjbad[w][lane] = 64;
      }
      double tot[4], off[4];
#pragma unroll
      for (int ct = 0; ct < 4; ++ct) {
        tot[ct] = 0.0;
        off[ct] = 0.0;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const double x = acc[ct][r];
          tot[ct] += x;
          const double xc = fmax(x, 0.0);
          acc[ct][r] = xc;
          off[ct] += xc;
        }
        // the diagonal P[a][a] sits in tile ct = w, element r, of lane lr = 4 r + lk
        if (ct == w) {
#pragma unroll
          for (int r = 0; r < 4; ++r)
            if (lr == 4 * r + lk) off[ct] -= acc[ct][r];
        }
        tot[ct] = lanes_sum(tot[ct]);
        off[ct] = lanes_sum(off[ct]);
      }
      ptot[w][lane] = lk == 0 ? tot[0] : lk == 1 ? tot[1] : lk == 2 ? tot[2] : tot[3];
      poff[w][lane] = lk == 0 ? off[0] : lk == 1 ? off[1] : lk == 2 ? off[2] : off[3];
    }
    __syncthreads();
    if (tid < 64) {
      const int i = tid;
      double tot = 0.0, off = 0.0;
      int code = 0, jbad = 64;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        tot += ptot[v][i];
        off += poff[v][i];
        if (code == 0 && pjbad[v][i] < 64) { code = 1; jbad = pjbad[v][i]; }
      }
      const double diag = 1.0 - off;
      if (code == 0 && fabs(tot - 1.0) > kTol) code = 2;
      if (code == 0 && !(diag <= 1.0 && diag > 0.0)) code = 3;
      if (code != 0) atomicMin(perr, fail_key(set, nbr, br, i, jbad, code));
      pdiag[i] = diag;
      dinv[i] = 1.0 / diag;
    }
    __syncthreads();
    // the diagonal P[a][a] sits in tile ct = w, element r, of lane lr = 4 r + lk
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const double d = pdiag[16 * w + 4 * r + lk];
#pragma unroll
      for (int ct = 0; ct < 4; ++ct)
        if (ct == w && lr == 4 * r + lk) acc[ct][r] = d;
    }
    const size_t sb = (size_t)set * nbr + br;
    if (P64T) {
#pragma unroll
      for (int ct = 0; ct < 4; ++ct)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          P64T[sb * 4096 + (16 * w + 4 * r + lk) * 64 + 16 * ct + lr] = acc[ct][r];
    }
    // A leaf branch is read only through its leaf messages, an internal branch
    // only through its tile: each writes just its own kind.
    const bool leaf = br < nl;
    if (P64F) {
      // FP64 P for the matrix-core forward (layout at kTile64): element r of
      // tile ct is P[a][b], a = 16 ct + lr = 16 mo + 4 v + g (mo = ct,
      // v = lr / 4, g = lr % 4), b = 16 w + 4 r + lk
      double* out = P64F + sb * kTile64;
      if (leaf) {
#pragma unroll
        for (int ct = 0; ct < 4; ++ct)
#pragma unroll
          for (int r = 0; r < 4; ++r)
            out[(16 * w + 4 * r + lk) * 64 + 16 * (lr & 3) + 4 * ct + (lr >> 2)] = acc[ct][r];
        if (tid < 64) out[4096 + tid] = 1.0;
      } else {
#pragma unroll
        for (int ct = 0; ct < 4; ++ct)
#pragma unroll
          for (int r = 0; r < 4; ++r) out[(ct * 16 + 4 * w + r) * 64 + lane] = acc[ct][r];
      }
    }
    if (P64FT && !leaf) {
      // P^T in the same order: P^T[b][a] = P[a][b] at
      // ((b / 16) * 16 + a / 4) * 64 + b % 16 + 16 (a % 4)
      double* out = P64FT + sb * 4096;
#pragma unroll
      for (int ct = 0; ct < 4; ++ct)
#pragma unroll
        for (int r = 0; r < 4; ++r)
          out[(16 * w + 4 * ct + (lr >> 2)) * 64 + 4 * r + lk + 16 * (lr & 3)] = acc[ct][r];
    }
    if (leaf && PT32) {
      float* out = PT32 + sb * (kLeafRows * 64);
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int b = 16 * w + 4 * r + lk;
        const double di = dinv[b];
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) out[b * 64 + 16 * ct + lr] = (float)(acc[ct][r] * di);
      }
      if (tid < 64) {
        out[4096 + tid] = 1.0f;
        lnPcc[sb * kLeafRows + tid] = log(pdiag[tid]);
      }
      if (tid == 64) lnPcc[sb * kLeafRows + 64] = 0.0;
    } else if (!leaf && Pfrag) {
      // The tile: the off-diagonal of P (diagonal entries 0), so the products
      // sum nonnegative terms, and the diagonal P[a][a] - 1 joins in
      // round-to-nearest VALU arithmetic.  FP32: A-operand order of
      // v_mfma_f32_16x16x4f32 with the K index permuted to match the
      // accumulator layout (prune_mfma.hpp),
      //   e = ((mo*4 + mb)*64 + L)*4 + v:  row 16 mo + L%16, col 16 mb + 4 (L/16) + v
      // i.e. mo = ct, mb = w, L = lr + 16 r, v = lk.  Half formats: the
      // v_mfma_f32_16x16x32_{f16,bf16} A order, K permuted the same way,
      //   e = ((mo*2 + st)*64 + L)*8 + j:  row 16 mo + L%16,
      //                                    col 16 (2 st + j/4) + 4 (L/16) + j%4
      // i.e. st = w / 2, j = 4 (w % 2) + lk; hi and lo halves in two planes.
      // From x = the FP32 value of the entry: hi = x rounded to nearest, lo =
      // the residual x - hi (exact in FP32) rounded to nearest, so hi + lo is x
      // to about 2^-22 relative.  FP16: x = offdiag(P) 2^13 and lo carries a
      // further 2^13 (a normal number wherever hi is: the matrix cores read
      // FP16 subnormals as zero).
      float* out = Pfrag + sb * kTileFloats;
      if (fmt == kFragF32) {
#pragma unroll
        for (int ct = 0; ct < 4; ++ct)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const bool dg = (ct == w && lr == 4 * r + lk);
            out[((ct * 4 + w) * 64 + lr + 16 * r) * 4 + lk] = dg ? 0.0f : (float)acc[ct][r];
          }
      } else {
        const float sc = (fmt == kFragF16) ? (float)(1 << kHalfExp) : 1.0f;
#pragma unroll
        for (int ct = 0; ct < 4; ++ct)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const bool dg = (ct == w && lr == 4 * r + lk);
            const float x = dg ? 0.0f : (float)acc[ct][r] * sc;
            const int e = ((ct * 2 + (w >> 1)) * 64 + lr + 16 * r) * 8 + 4 * (w & 1) + lk;
            if (fmt == kFragF16) {
              _Float16* h = reinterpret_cast<_Float16*>(out);
              const _Float16 hi = (_Float16)x;
              h[e] = hi;
              h[4096 + e] = (_Float16)((x - (float)hi) * sc);
            } else {
              __bf16* h = reinterpret_cast<__bf16*>(out);
              const __bf16 hi = (__bf16)x;
              h[e] = hi;
              h[4096 + e] = (__bf16)(x - (float)hi);
            }
          }
      }
      if (tid < 64) out[4096 + tid] = (float)(pdiag[tid] - 1.0);
    }
    if constexpr (WITH_D) {
      // D = dP(rho b)/drho = b Q P(t) for the tangent kernel, in the layout
      // of P: leaf rows D[a][c] / P[c][c] (row 64 zero: `Marginalize does not
      // depend on rho) for a leaf branch, the dense FP32 tile in the order of
      // the FP32 P tile for an internal one
      if (leaf) {
        float* out = DT32 + sb * (kLeafRows * 64);
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int b = 16 * w + 4 * r + lk;
          const double di = dinv[b];
#pragma unroll
          for (int ct = 0; ct < 4; ++ct) out[b * 64 + 16 * ct + lr] = (float)(dacc[ct][r] * di);
        }
        if (tid < 64) out[4096 + tid] = 0.0f;
      } else {
        float* out = Dfrag + sb * 4096;
#pragma unroll
        for (int ct = 0; ct < 4; ++ct)
#pragma unroll
          for (int r = 0; r < 4; ++r)
            out[((ct * 4 + w) * 64 + lr + 16 * r) * 4 + lk] = (float)dacc[ct][r];
      }
    }
  }
}

// Cherry tables: for cherry c with leaves a, b (branches into them from c),
// T[x][y] = P_c (P_a[:, x] o P_b[:, y]) for codons x, y or `Marginalize (64,
// the all-ones leaf vector), in FP64, stored normalised to max 1 (FP32) with
// its ln max (FP64).  One 64-thread workgroup per (x, y) and (set, table).
__global__ __launch_bounds__(64) void cherry_table_kernel(const double* __restrict__ P64T,
                                                         const int32_t* __restrict__ cher,
                                                         int nctab, int nbr,
                                                         float* __restrict__ ctab,
                                                         double* __restrict__ clnt) {
  __shared__ double al[64];
  __shared__ double mx[64];
  const int xy = blockIdx.x, x = xy / kTabCodes, y = xy % kTabCodes;
  const int st = blockIdx.y, set = st / nctab, t = st % nctab, k = threadIdx.x;
  const int c = cher[3 * t], a = cher[3 * t + 1], b = cher[3 * t + 2];
  const double* Pa = P64T + ((size_t)set * nbr + a) * 4096;  // Pa[col * 64 + row]
  const double* Pb = P64T + ((size_t)set * nbr + b) * 4096;
  const double* Pc = P64T + ((size_t)set * nbr + c) * 4096;
  al[k] = (x == 64 ? 1.0 : Pa[x * 64 + k]) * (y == 64 ? 1.0 : Pb[y * 64 + k]);
  __syncthreads();
  double m = 0.0;
  for (int j = 0; j < 64; ++j) m = fma(Pc[j * 64 + k], al[j], m);  // row k of P_c al
  mx[k] = m;
  __syncthreads();
  for (int w = 32; w > 0; w >>= 1) {
    if (k < w) mx[k] = fmax(mx[k], mx[k + w]);
    __syncthreads();
  }
  const double top = mx[0];
  const size_t base = ((size_t)st * kTabCodes * kTabCodes + xy);
  const bool ones = (x == 64 && y == 64);  // P 1 = 1 exactly
  ctab[base * 64 + k] = ones ? 1.0f : (float)(m / top);
  if (k == 0) clnt[base] = ones ? 0.0 : log(top);
}

}  // namespace

hipError_t launch_cherry_tables(const pcsf_model* m, const PSets& sets, int nsets,
                                hipStream_t s) {
  const int nct = (int)m->cherries.size() / 3;
  if (nct == 0 || nsets == 0) return hipSuccess;
  dim3 grid(kTabCodes * kTabCodes, nsets * nct);
  hipLaunchKernelGGL(cherry_table_kernel, grid, dim3(64), 0, s, sets.d_P64T, m->d_cherries, nct,
                     m->nbr, sets.d_ctab, sets.d_clnt);
  return hipGetLastError();
}

hipError_t launch_pbuild(const pcsf_model* m, const PSets& sets, int nsets, int fmt,
                         unsigned outputs, hipStream_t s) {
  if (nsets <= 0) return hipSuccess;
  const bool p64 = (outputs & kBuildP64) && sets.fp64;
  const bool p64m = (outputs & kBuildP64M) && sets.fp64;
  const bool p64mt = (outputs & kBuildP64MT) && sets.fp64;
  const bool mfma = (outputs & kBuildMfma) != 0;
  const bool tangent = (outputs & kBuildTangent) && sets.d_Dfrag;
  // several branches per workgroup once there are enough matrices to fill
  // the device; one each for the few sets of a fixed-strategy model
  const int nbw = ((size_t)nsets * m->nbr >= 16384) ? 8 : 1;
  dim3 grid((m->nbr + nbw - 1) / nbw, nsets);
  auto* k = tangent ? pbuild_kernel<true> : pbuild_kernel<false>;
  // the eigensystem of set s is model s % nm's, or the set's own when the
  // sets carry rate models (then s % cap = s)
  const bool own = sets.d_eVT != nullptr;
  hipLaunchKernelGGL(k, grid, dim3(256), 0, s, own ? sets.d_eVT : m->d_VT,
                     own ? sets.d_eVinv : m->d_Vinv, own ? sets.d_elam : m->d_lam,
                     m->d_branch, sets.d_rho, sets.d_active, own ? sets.cap : m->nm, m->nbr,
                     nbw, m->nl,
                     p64 ? sets.d_P64T : nullptr, mfma ? sets.d_PT32 : nullptr,
                     mfma ? sets.d_Pfrag : nullptr, fmt, sets.d_lnPcc, m->d_perr, sets.d_DT32,
                     sets.d_Dfrag, p64m ? sets.d_P64F : nullptr,
                     p64mt ? sets.d_P64FT : nullptr);
  return hipGetLastError();
}
