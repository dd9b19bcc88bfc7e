// pbuild.hip -- FP64 transition matrices P(t) = V diag(exp(lam t)) V^-1 on the GPU.
//
// One workgroup per (branch, set); set s is model s % nm at tree scale
// rho[s] and is skipped when active[s] == 0.  Reproduces Q.Diag.real_to_Pt
// (lib/CamlPaml/Q.ml:211-249): entries below zero are an error when
// |x| > 1e-6 and are set to 0 otherwise; the row sum must be within 1e-6 of
// 1; the diagonal is replaced by 1 - sum(off-diagonal).  The first failure
// is recorded in perr = {code, branch, model, row, col} for the host.
//
// Outputs, all [set][branch][...]; a leaf branch (br < nl) is read only
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
//         followed by the 64 FP32 values P[a][a] - 1 (read by the FP32 path).
#include "../internal.hpp"

namespace {

constexpr double kTol = 1e-6;

// LDS row stride in doubles: with 65, row i starts 2i banks along, so a wave
// reading one column (64 rows) or a 4x16 block is free of bank conflicts
constexpr int kS = 65;

typedef double f64x4 __attribute__((ext_vector_type(4)));

// P = (V diag(e)) V^-1 with e = exp(lam t), on v_mfma_f64_16x16x4f64.  Wave w
// makes rows 16w..16w+15 of P as four 16x16 tiles; at step s (k = 4s..4s+3)
//   A[i][k] = V[16w + i][k] e[k]   in lane i + 16 (k % 4)
//   B[k][j] = V^-1[k][16 ct + j]   in lane j + 16 (k % 4)
//   D[i][j]                        in lane j + 16 (i % 4), element i / 4
// VT is V transposed, so both operands load 16 consecutive doubles per k.
template <bool WITH_D>
__global__ __launch_bounds__(256) void pbuild_kernel(
    const double* __restrict__ VT, const double* __restrict__ Vinv,
    const double* __restrict__ lam, const double* __restrict__ branch,
    const double* __restrict__ rho, const uint8_t* __restrict__ active, int nm, int nbr,
    int nl, double* __restrict__ P64T, float* __restrict__ PT32, float* __restrict__ Pfrag, int fmt,
    double* __restrict__ lnPcc, int* __restrict__ perr, double* __restrict__ perrval,
    const double* __restrict__ QT, float* __restrict__ DT32, float* __restrict__ Dfrag) {
  __shared__ double Ps[64 * kS];                // P, then fixed up
  __shared__ double Ds[WITH_D ? 64 * kS : 1];   // b Q P
  __shared__ double ex[64];
  __shared__ double dinv[64];                   // 1 / P[c][c] after the fix-up
  __shared__ double ptot[4][64], poff[4][64], pbad[4][64];
  __shared__ int pjbad[4][64];
  const int br = blockIdx.x, set = blockIdx.y, q = set % nm, tid = threadIdx.x;
  if (active && !active[set]) return;
  const double t = rho[set] * branch[br];
  if (tid < 64) ex[tid] = exp(t * lam[q * 64 + tid]);
  if (t < 0.0) {
    if (tid == 0 && atomicCAS(perr, 0, 4) == 0) {
      perr[1] = br; perr[2] = set; perr[3] = 0; perr[4] = 0; *perrval = t;
    }
  }
  __syncthreads();
  const int w = tid >> 6, lane = tid & 63, lr = lane & 15, lk = lane >> 4;
  {
    const double* A = VT + (size_t)q * 4096 + 16 * w + lr;
    const double* B = Vinv + (size_t)q * 4096 + lr;
    f64x4 acc[4] = {};
#pragma unroll 4
    for (int s = 0; s < 16; ++s) {
      const int k = 4 * s + lk;
      const double a = A[k * 64] * ex[k];
#pragma unroll
      for (int ct = 0; ct < 4; ++ct)
        acc[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, B[k * 64 + 16 * ct], acc[ct], 0, 0, 0);
    }
#pragma unroll
    for (int ct = 0; ct < 4; ++ct)
#pragma unroll
      for (int r = 0; r < 4; ++r) Ps[(16 * w + 4 * r + lk) * kS + 16 * ct + lr] = acc[ct][r];
  }
  __syncthreads();
  // Row fix-up.  Thread (row i = lane, quarter w) scans P[i][16w..16w+15]:
  // entries below zero are an error beyond kTol and become 0; it keeps the
  // row sum, the sum of the (clamped) off-diagonal entries and the first bad
  // column; one thread per row then joins the quarters in column order.
  {
    const int i = lane;
    double tot = 0.0, off = 0.0, bad = 0.0;
    int jbad = -1;
#pragma unroll
    for (int jj = 0; jj < 16; ++jj) {
      const int j = 16 * w + jj;
      double x = Ps[i * kS + j];
      tot += x;
      if (x < 0.0) {
        if (fabs(x) > kTol && jbad < 0) { jbad = j; bad = x; }
        x = 0.0;
        Ps[i * kS + j] = 0.0;
      }
      if (j != i) off += x;
    }
    ptot[w][i] = tot;
    poff[w][i] = off;
    pbad[w][i] = bad;
    pjbad[w][i] = jbad;
  }
  __syncthreads();
  if (tid < 64) {
    const int i = tid;
    double tot = 0.0, off = 0.0, bad = 0.0;
    int code = 0, jbad = 0;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      tot += ptot[v][i];
      off += poff[v][i];
      if (code == 0 && pjbad[v][i] >= 0) { code = 1; jbad = pjbad[v][i]; bad = pbad[v][i]; }
    }
    const double diag = 1.0 - off;
    if (code == 0 && fabs(tot - 1.0) > kTol) { code = 2; bad = tot; }
    if (code == 0 && !(diag <= 1.0 && diag > 0.0)) { code = 3; bad = diag; }
    if (code != 0 && atomicCAS(perr, 0, code) == 0) {
      perr[1] = br; perr[2] = set; perr[3] = i; perr[4] = jbad; *perrval = bad;
    }
    Ps[i * kS + i] = diag;
    dinv[i] = 1.0 / diag;
  }
  __syncthreads();
  const size_t sb = (size_t)set * nbr + br;
  if (P64T) {
    for (int e = tid; e < 4096; e += 256) P64T[sb * 4096 + e] = Ps[(e & 63) * kS + (e >> 6)];
  }
  // A leaf branch is read only through its leaf messages, an internal branch
  // only through its tile: each writes just its own kind.
  const bool leaf = br < nl;
  if (leaf) {
    for (int e = tid; e < 4096; e += 256) {
      const int b = e >> 6, a = e & 63;  // transposed: [b][a] = P[a][b]
      PT32[sb * (kLeafRows * 64) + e] = (float)(Ps[a * kS + b] * dinv[b]);
    }
    if (tid < 64) {
      PT32[sb * (kLeafRows * 64) + 4096 + tid] = 1.0f;
      lnPcc[sb * kLeafRows + tid] = log(Ps[tid * kS + tid]);
    }
    if (tid == 64) lnPcc[sb * kLeafRows + 64] = 0.0;
  } else if (fmt == kFragF32) {
    // fragment order of v_mfma_f32_16x16x4f32's A operand with the K index
    // permuted to match the accumulator layout (prune_mfma.hpp):
    //   e = ((mo*4 + mb)*64 + lane)*4 + v
    //   row = 16*mo + lane%16, col = 16*mb + 4*(lane/16) + v
    for (int e = tid; e < 4096; e += 256) {
      const int v = e & 3, lane = (e >> 2) & 63, mb = (e >> 8) & 3, mo = e >> 10;
      const int row = 16 * mo + (lane & 15);
      const int col = 16 * mb + 4 * (lane >> 4) + v;
      Pfrag[sb * kTileFloats + e] = (row == col) ? 0.0f : (float)Ps[row * kS + col];
    }
  } else {
    // v_mfma_f32_16x16x32_{f16,bf16} A order, K permuted the same way:
    //   e = ((mo*2 + st)*64 + lane)*8 + j
    //   row = 16*mo + lane%16, col = 16*(2*st + j/4) + 4*(lane/16) + j%4
    // From x = the FP32 value of the entry: hi = x rounded to nearest, lo =
    // the residual x - hi (exact in FP32) rounded to nearest, so hi + lo is x
    // to about 2^-22 relative.  FP16: x = offdiag(P) 2^13 and lo carries a
    // further 2^13 (a normal number wherever hi is: the matrix cores read FP16
    // subnormals as zero).
    const float sc = (fmt == kFragF16) ? (float)(1 << kHalfExp) : 1.0f;
    for (int e = tid; e < 4096; e += 256) {
      const int j = e & 7, lane = (e >> 3) & 63, st = (e >> 9) & 1, mo = e >> 10;
      const int row = 16 * mo + (lane & 15);
      const int col = 16 * (2 * st + (j >> 2)) + 4 * (lane >> 4) + (j & 3);
      // the off-diagonal of P, as in the FP32 tile: the products then sum
      // nonnegative terms, and the diagonal P[a][a] - 1 (float 4096 on)
      // joins in round-to-nearest VALU arithmetic
      const float x = (row == col) ? 0.0f : (float)Ps[row * kS + col] * sc;
      if (fmt == kFragF16) {
        _Float16* h = reinterpret_cast<_Float16*>(Pfrag + sb * kTileFloats);
        const _Float16 hi = (_Float16)x;
        h[e] = hi;
        h[4096 + e] = (_Float16)((x - (float)hi) * sc);
      } else {
        __bf16* h = reinterpret_cast<__bf16*>(Pfrag + sb * kTileFloats);
        const __bf16 hi = (__bf16)x;
        h[e] = hi;
        h[4096 + e] = (__bf16)(x - (float)hi);
      }
    }
  }
  if (!leaf && tid < 64) Pfrag[sb * kTileFloats + 4096 + tid] = (float)(Ps[tid * kS + tid] - 1.0);
  if constexpr (WITH_D) {
    // D = dP(rho b)/drho = b Q P(t) for the tangent kernel, on the same MFMA
    // with A[i][k] = Q[16w + i][k] (QT[k][i]) and B[k][j] = P[k][16 ct + j]:
    // leaf rows D[a][c] / P[c][c] (row 64 zero: `Marginalize does not depend
    // on rho) for a leaf branch, the dense FP32 tile in the same A-operand
    // order as Pfrag for an internal one
    const double* A = QT + (size_t)q * 4096 + 16 * w + lr;
    f64x4 acc[4] = {};
#pragma unroll 4
    for (int s = 0; s < 16; ++s) {
      const int k = 4 * s + lk;
      const double a = A[k * 64];
#pragma unroll
      for (int ct = 0; ct < 4; ++ct)
        acc[ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, Ps[k * kS + 16 * ct + lr], acc[ct], 0,
                                                       0, 0);
    }
    const double bl = branch[br];
#pragma unroll
    for (int ct = 0; ct < 4; ++ct)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        Ds[(16 * w + 4 * r + lk) * kS + 16 * ct + lr] = bl * acc[ct][r];
    __syncthreads();
    if (leaf) {
      for (int e = tid; e < 4096; e += 256) {
        const int c = e >> 6, a = e & 63;
        DT32[sb * (kLeafRows * 64) + e] = (float)(Ds[a * kS + c] * dinv[c]);
      }
      if (tid < 64) DT32[sb * (kLeafRows * 64) + 4096 + tid] = 0.0f;
    } else {
      for (int e = tid; e < 4096; e += 256) {
        const int v = e & 3, lane = (e >> 2) & 63, mb = (e >> 8) & 3, mo = e >> 10;
        Dfrag[sb * 4096 + e] =
            (float)Ds[(16 * mo + (lane & 15)) * kS + 16 * mb + 4 * (lane >> 4) + v];
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
                         hipStream_t s) {
  if (nsets <= 0) return hipSuccess;
  dim3 grid(m->nbr, nsets);
  auto* k = sets.d_Dfrag ? pbuild_kernel<true> : pbuild_kernel<false>;
  hipLaunchKernelGGL(k, grid, dim3(256), 0, s, m->d_VT, m->d_Vinv, m->d_lam,
                     m->d_branch, sets.d_rho, sets.d_active, m->nm, m->nbr, m->nl,
                     sets.fp64 ? sets.d_P64T : nullptr, sets.d_PT32, sets.d_Pfrag, fmt,
                     sets.d_lnPcc, m->d_perr, m->d_perrval, m->d_QT, sets.d_DT32,
                     sets.d_Dfrag);
  return hipGetLastError();
}
