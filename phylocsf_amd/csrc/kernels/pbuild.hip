// pbuild.hip -- FP64 transition matrices P(t) = V diag(exp(lam t)) V^-1 on the GPU.
//
// One workgroup per (branch, model).  Reproduces Q.Diag.real_to_Pt
// (lib/CamlPaml/Q.ml:211-249): entries below zero are an error when
// |x| > 1e-6 and are set to 0 otherwise; the row sum must be within 1e-6 of
// 1; the diagonal is replaced by 1 - sum(off-diagonal).  The first failure
// is recorded in perr = {code, branch, model, row, col} for the host.
//
// Outputs (all [model][branch][4096]):
//   P64T  FP64 P transposed (P64T[b][a] = P[a][b]): coalesced GEMV operand of
//         the FP64 kernel and its leaf gathers.
//   PT32  FP32 leaf messages of the MFMA kernel: PT32[c][a] = P[a][c] / P[c][c],
//         with lnPcc[c] = ln P[c][c] in FP64 (the dominant entry of a leaf
//         message is exactly 1 and its value travels in FP64).
//   Pfrag FP32 tile of kTileFloats: the off-diagonal of P in the A-operand
//         order of v_mfma_f32_16x16x4f32 as the MFMA kernel consumes it
//         (diagonal entries zero), followed by the 64 values P[a][a] - 1.
#include "../internal.hpp"

namespace {

constexpr double kTol = 1e-6;

__global__ __launch_bounds__(256) void pbuild_kernel(
    const double* __restrict__ V, const double* __restrict__ Vinv,
    const double* __restrict__ lam, const double* __restrict__ branch,
    const double* __restrict__ rho, int nbr, double* __restrict__ P64T,
    float* __restrict__ PT32, float* __restrict__ Pfrag, double* __restrict__ lnPcc,
    int* __restrict__ perr, double* __restrict__ perrval) {
  __shared__ double W[64 * 64];  // diag(exp(lam t)) V^-1, then P
  __shared__ double Vs[64 * 64];
  const int br = blockIdx.x, q = blockIdx.y, tid = threadIdx.x;
  const double t = rho[q] * branch[br];
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
      perr[1] = br; perr[2] = q; perr[3] = 0; perr[4] = 0; *perrval = t;
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
      perr[1] = br; perr[2] = q; perr[3] = i; perr[4] = jbad; *perrval = bad;
    }
    Vs[i * 64 + i] = diag;
  }
  __syncthreads();
  const size_t base = ((size_t)q * nbr + br) * 4096;
  for (int e = tid; e < 4096; e += 256) {
    const int b = e >> 6, a = e & 63;  // transposed: [b][a] = P[a][b]
    const double p = Vs[a * 64 + b];
    P64T[base + e] = p;
    PT32[base + e] = (float)(p / Vs[b * 64 + b]);
  }
  if (tid < 64) lnPcc[((size_t)q * nbr + br) * 64 + tid] = log(Vs[tid * 64 + tid]);
  const size_t fbase = ((size_t)q * nbr + br) * kTileFloats;
  // fragment order of v_mfma_f32_16x16x4f32's A operand with the K index
  // permuted to match the accumulator layout (prune_mfma.hip):
  //   e = ((mo*4 + mb)*64 + lane)*4 + v
  //   row = 16*mo + lane%16, col = 16*mb + 4*(lane/16) + v
  for (int e = tid; e < 4096; e += 256) {
    const int v = e & 3, lane = (e >> 2) & 63, mb = (e >> 8) & 3, mo = e >> 10;
    const int row = 16 * mo + (lane & 15);
    const int col = 16 * mb + 4 * (lane >> 4) + v;
    Pfrag[fbase + e] = (row == col) ? 0.0f : (float)Vs[row * 64 + col];
  }
  if (tid < 64) Pfrag[fbase + 4096 + tid] = (float)(Vs[tid * 64 + tid] - 1.0);
}

}  // namespace

hipError_t launch_pbuild(const pcsf_model* m, const double* d_rho, hipStream_t s) {
  dim3 grid(m->nbr, m->nm);
  hipLaunchKernelGGL(pbuild_kernel, grid, dim3(256), 0, s, m->d_V, m->d_Vinv, m->d_lam,
                     m->d_branch, d_rho, m->nbr, m->d_P64T, m->d_PT32, m->d_Pfrag, m->d_lnPcc,
                     m->d_perr, m->d_perrval);
  return hipGetLastError();
}
