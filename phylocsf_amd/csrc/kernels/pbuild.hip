// pbuild.hip -- FP64 transition matrices P(t) = V diag(exp(lam t)) V^-1 on the GPU.
//
// One workgroup per (branch, set); set s is model s % nm at tree scale
// rho[s] and is skipped when active[s] == 0.  Reproduces Q.Diag.real_to_Pt
// (lib/CamlPaml/Q.ml:211-249): entries below zero are an error when
// |x| > 1e-6 and are set to 0 otherwise; the row sum must be within 1e-6 of
// 1; the diagonal is replaced by 1 - sum(off-diagonal).  The first failure
// is recorded in perr = {code, branch, model, row, col} for the host.
//
// Outputs (all [model][branch][4096]):
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

__global__ __launch_bounds__(256) void pbuild_kernel(
    const double* __restrict__ V, const double* __restrict__ Vinv,
    const double* __restrict__ lam, const double* __restrict__ branch,
    const double* __restrict__ rho, const uint8_t* __restrict__ active, int nm, int nbr,
    double* __restrict__ P64T, float* __restrict__ PT32, float* __restrict__ Pfrag, int fmt,
    double* __restrict__ lnPcc, int* __restrict__ perr, double* __restrict__ perrval,
    const double* __restrict__ QT, float* __restrict__ DT32, float* __restrict__ Dfrag) {
  __shared__ double W[64 * 64];  // diag(exp(lam t)) V^-1, then P
  __shared__ double Vs[64 * 64];
  const int br = blockIdx.x, set = blockIdx.y, q = set % nm, tid = threadIdx.x;
  if (active && !active[set]) return;
  const double t = rho[set] * branch[br];
  const double* Vq = V + (size_t)q * 4096;
  const double* Viq = Vinv + (size_t)q * 4096;
  for (int e = tid; e < 4096; e += 256) {
    const int m = e >> 6;
    W[e] = exp(t * lam[q * 64 + m]) * Viq[e];
    Vs[e] = Vq[e];
  }
  __syncthreads();
  // 4x4 register block per thread
  const int r0 = (tid >> 4) * 4, c0 = (tid & 15) * 4;
  double acc[4][4] = {};
  for (int m = 0; m < 64; ++m) {
    double a[4], w[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) a[i] = Vs[(r0 + i) * 64 + m];
#pragma unroll
    for (int j = 0; j < 4; ++j) w[j] = W[m * 64 + c0 + j];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = fma(a[i], w[j], acc[i][j]);
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) Vs[(r0 + i) * 64 + c0 + j] = acc[i][j];
  __syncthreads();
  if (t < 0.0) {
    if (tid == 0 && atomicCAS(perr, 0, 4) == 0) {
      perr[1] = br; perr[2] = set; perr[3] = 0; perr[4] = 0; *perrval = t;
    }
  }
  if (tid < 64) {  // row fix-up, sequential over j like the reference
    const int i = tid;
    double tot = 0.0, diag = 1.0;
    int code = 0, jbad = 0;
    double bad = 0.0;
    for (int j = 0; j < 64; ++j) {
      double x = Vs[i * 64 + j];
      tot += x;
      if (x < 0.0) {
        if (fabs(x) > kTol && code == 0) { code = 1; jbad = j; bad = x; }
        x = 0.0;
        Vs[i * 64 + j] = 0.0;
      }
      if (j != i) diag -= x;
    }
    if (code == 0 && fabs(tot - 1.0) > kTol) { code = 2; bad = tot; }
    if (code == 0 && !(diag <= 1.0 && diag > 0.0)) { code = 3; bad = diag; }
    if (code != 0 && atomicCAS(perr, 0, code) == 0) {
      perr[1] = br; perr[2] = set; perr[3] = i; perr[4] = jbad; *perrval = bad;
    }
    Vs[i * 64 + i] = diag;
  }
  __syncthreads();
  const size_t base = ((size_t)set * nbr + br) * 4096;
  const size_t lbase = ((size_t)set * nbr + br) * (kLeafRows * 64);
  for (int e = tid; e < 4096; e += 256) {
    const int b = e >> 6, a = e & 63;  // transposed: [b][a] = P[a][b]
    const double p = Vs[a * 64 + b];
    if (P64T) P64T[base + e] = p;
    PT32[lbase + e] = (float)(p / Vs[b * 64 + b]);
  }
  if (tid < 64) {
    PT32[lbase + 4096 + tid] = 1.0f;
    lnPcc[((size_t)set * nbr + br) * kLeafRows + tid] = log(Vs[tid * 64 + tid]);
  }
  if (tid == 64) lnPcc[((size_t)set * nbr + br) * kLeafRows + 64] = 0.0;
  const size_t fbase = ((size_t)set * nbr + br) * kTileFloats;
  if (fmt == kFragF32) {
    // fragment order of v_mfma_f32_16x16x4f32's A operand with the K index
    // permuted to match the accumulator layout (prune_mfma.hpp):
    //   e = ((mo*4 + mb)*64 + lane)*4 + v
    //   row = 16*mo + lane%16, col = 16*mb + 4*(lane/16) + v
    for (int e = tid; e < 4096; e += 256) {
      const int v = e & 3, lane = (e >> 2) & 63, mb = (e >> 8) & 3, mo = e >> 10;
      const int row = 16 * mo + (lane & 15);
      const int col = 16 * mb + 4 * (lane >> 4) + v;
      Pfrag[fbase + e] = (row == col) ? 0.0f : (float)Vs[row * 64 + col];
    }
  } else {
    // v_mfma_f32_16x16x32_{f16,bf16} A order, K permuted the same way:
    //   e = ((mo*2 + st)*64 + lane)*8 + j
    //   row = 16*mo + lane%16, col = 16*(2*st + j/4) + 4*(lane/16) + j%4
    // hi = x rounded to nearest, lo = the residual x - hi rounded to nearest;
    // FP16: x = offdiag(P) 2^13 and lo carries a further 2^13 (a normal number
    // wherever hi is: the matrix cores read FP16 subnormals as zero)
    const double sc = (fmt == kFragF16) ? (double)(1 << kHalfExp) : 1.0;
    for (int e = tid; e < 4096; e += 256) {
      const int j = e & 7, lane = (e >> 3) & 63, st = (e >> 9) & 1, mo = e >> 10;
      const int row = 16 * mo + (lane & 15);
      const int col = 16 * (2 * st + (j >> 2)) + 4 * (lane >> 4) + (j & 3);
      // the off-diagonal of P, as in the FP32 tile: the products then sum
      // nonnegative terms, and the diagonal P[a][a] - 1 (float 4096 on)
      // joins in round-to-nearest VALU arithmetic
      const double x = (row == col) ? 0.0 : Vs[row * 64 + col] * sc;
      if (fmt == kFragF16) {
        _Float16* h = reinterpret_cast<_Float16*>(Pfrag + fbase);
        const _Float16 hi = (_Float16)x;
        h[e] = hi;
        h[4096 + e] = (_Float16)((x - (double)hi) * sc);
      } else {
        __bf16* h = reinterpret_cast<__bf16*>(Pfrag + fbase);
        const __bf16 hi = (__bf16)x;
        h[e] = hi;
        h[4096 + e] = (__bf16)(x - (double)hi);
      }
    }
  }
  if (tid < 64) Pfrag[fbase + 4096 + tid] = (float)(Vs[tid * 64 + tid] - 1.0);
  if (!Dfrag) return;
  // D = dP(rho b)/drho = b Q P(t) for the tangent kernel: leaf rows
  // D[a][c] / P[c][c] (row 64 zero: `Marginalize does not depend on rho) and
  // the dense FP32 tile in the same A-operand order as Pfrag
  __syncthreads();
  const double* Qq = QT + (size_t)q * 4096;  // QT[k*64 + a] = Q[a][k]
  for (int e = tid; e < 4096; e += 256) W[(e & 63) * 64 + (e >> 6)] = Qq[e];
  __syncthreads();
  double dacc[4][4] = {};
  for (int k = 0; k < 64; ++k) {
    double a[4], p[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) a[i] = W[(r0 + i) * 64 + k];
#pragma unroll
    for (int j = 0; j < 4; ++j) p[j] = Vs[k * 64 + c0 + j];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) dacc[i][j] = fma(a[i], p[j], dacc[i][j]);
  }
  __syncthreads();  // every thread is done reading Q from W
  const double bl = branch[br];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) W[(r0 + i) * 64 + c0 + j] = bl * dacc[i][j];
  __syncthreads();
  for (int e = tid; e < 4096; e += 256) {
    const int c = e >> 6, a = e & 63;
    DT32[lbase + e] = (float)(W[a * 64 + c] / Vs[c * 64 + c]);
    const int v = e & 3, lane = (e >> 2) & 63, mb = (e >> 8) & 3, mo = e >> 10;
    Dfrag[((size_t)set * nbr + br) * 4096 + e] =
        (float)W[(16 * mo + (lane & 15)) * 64 + 16 * mb + 4 * (lane >> 4) + v];
  }
  if (tid < 64) DT32[lbase + 4096 + tid] = 0.0f;
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
  hipLaunchKernelGGL(pbuild_kernel, grid, dim3(256), 0, s, m->d_V, m->d_Vinv, m->d_lam,
                     m->d_branch, sets.d_rho, sets.d_active, m->nm, m->nbr,
                     sets.fp64 ? sets.d_P64T : nullptr, sets.d_PT32, sets.d_Pfrag, fmt,
                     sets.d_lnPcc, m->d_perr, m->d_perrval, m->d_QT, sets.d_DT32,
                     sets.d_Dfrag);
  return hipGetLastError();
}
