// omega.hip -- the omega test's rate models on the GPU, one per set (region,
// hypothesis) of a chunk, rebuilt at every lock-step evaluation from the
// set's twelve settings (kappa, omega, sigma, nine F3x4 frequency ratios):
//
//   pi   codon frequencies, the stops' share scaled by sigma
//        (src/OmegaModel.ml:24-40)
//   S    exchangeabilities: one nucleotide apart, kappa on transitions, omega
//        on nonsynonymous changes between sense codons (:42-74)
//   Q    q_ij = s_ij pi_j / scale, scale = -sum_i pi_i q_ii (the mean-rate
//        normalisation of P14n.instantiate_q, lib/CamlPaml/P14n.ml)
//   eigensystem of Q through the symmetric A = Pi^1/2 Q Pi^-1/2 by cyclic
//        Jacobi rotations in parallel order (Golub & Van Loan, Matrix
//        Computations, sec. 8.5: the round-robin ordering applies 32 disjoint
//        rotations at once, 63 rounds per sweep), until the off-diagonal mass
//        is below 1e-30 of the total or stops falling (it converges
//        quadratically to a rounding floor near 1e-31: eight sweeps for the
//        omega models), at most 64 sweeps; V = Pi^-1/2 U and V^-1 = U^T Pi^1/2.
//        A cold build also stores U as the set's reference basis U0; a later
//        build of the same set with the same pi (kappa or omega changed)
//        starts from it: A' = U0^T A U0 is near-diagonal, the rotations go on
//        from U = U0, and the stopping rule is unchanged (a few sweeps instead
//        of eight; Golub & Van Loan on Jacobi for nearby matrices).  Starting
//        from the fixed U0, not from the previous build, keeps the result a
//        function of the settings and U0's (the fit's starting point), not of
//        the path between them.  A set whose settings did not change keeps its
//        eigensystem.
//
// One 256-thread workgroup per set, A and U in LDS (rows padded to 65
// doubles).  Outputs in the sets' own-model layout: V transposed, V^-1, the
// eigenvalues, the normalised root prior and its log.  Paths are relative to
// the mlin/PhyloCSF tree.
#include "../internal.hpp"

namespace {

constexpr int kLd = 65;  // LDS row stride of A and U (doubles)

// Code.Codon64 amino acids (lib/CamlPaml/Code.ml:153-173); stops TAA, TAG, TGA
__constant__ char kAmino[65] = "KNKNTTTTRSRSIIMIQHQHPPPPRRRRLLLLEDEDAAAAGGGGVVVV*Y*YSSSS*CWCLFLF";
constexpr int kTAA = 48, kTAG = 50, kTGA = 56;

__device__ __forceinline__ bool stop_codon(int c) { return c == kTAA || c == kTAG || c == kTGA; }

// s_ij of the omega model (0 unless i, j are one nucleotide apart)
__device__ __forceinline__ double exchangeability(const double* qs, int i, int j) {
  const int d = i ^ j;
  int field = -1;
#pragma unroll
  for (int p = 0; p < 3; ++p)
    if (d == (d & (3 << (2 * p)))) field = p;
  if (d == 0 || field < 0) return 0.0;
  const int x = (i >> (2 * field)) & 3, y = (j >> (2 * field)) & 3;
  const double kp = ((x ^ y) == 2) ? qs[0] : 1.0;  // A <-> G, C <-> T
  const double op = (!stop_codon(i) && !stop_codon(j) && kAmino[i] != kAmino[j]) ? qs[1] : 1.0;
  return kp * op;
}

// mode per set: kEigCold builds from scratch and stores U in ref (64 x 64
// per set), kEigWarm starts from ref, kEigKeep leaves the outputs (unchanged
// settings, or an inactive set)
__global__ __launch_bounds__(256) void omega_eigen_kernel(
    const double* __restrict__ qs_all, const uint8_t* __restrict__ mode, double* __restrict__ ref,
    double* __restrict__ VT, double* __restrict__ Vinv, double* __restrict__ lam,
    double* __restrict__ pi_out, double* __restrict__ lnpi_out) {
  __shared__ double A[64 * kLd], U[64 * kLd];
  __shared__ double pi[64], sq[64], diag[64];
  __shared__ double cs[32], sn[32];
  __shared__ int pp[32], qq[32];
  __shared__ double red[256];
  __shared__ double total_s, scale_s;
  const int set = blockIdx.x, tid = threadIdx.x;
  const int md = mode[set];
  if (md == kEigKeep) return;  // uniform over the workgroup
  const double* qs = qs_all + (size_t)set * 12;
  // codon frequencies (OmegaModel.ml:24-40)
  if (tid < 64) {
    double f[3];
#pragma unroll
    for (int p = 0; p < 3; ++p) {
      const double* r = qs + 3 + 3 * p;
      const double tot = r[0] + (r[1] + (r[2] + 1.0));
      const int n = (tid >> (2 * (2 - p))) & 3;
      f[p] = (n == 3 ? 1.0 : r[n]) / tot;
    }
    pi[tid] = f[0] * (f[1] * f[2]);  // sc[c], divided by the stop factor below
  }
  __syncthreads();
  if (tid == 0) {
    const double denom = 1.0 - (1.0 - qs[2]) * (pi[kTAA] + (pi[kTAG] + pi[kTGA]));
    red[0] = denom;
  }
  __syncthreads();
  if (tid < 64) {
    pi[tid] = pi[tid] / red[0];
    sq[tid] = sqrt(pi[tid]);
  }
  __syncthreads();
  // Q's off-diagonal into A (row i: s_ij pi_j), its diagonal -sum_j q_ij in
  // the host's order, then the mean-rate scale
  for (int e = tid; e < 4096; e += 256) {
    const int i = e >> 6, j = e & 63;
    A[i * kLd + j] = (i == j) ? 0.0 : exchangeability(qs, i, j) * pi[j];
  }
  __syncthreads();
  if (tid < 64) {
    double off = 0.0;
    for (int j = 0; j < 64; ++j)
      if (j != tid) off = A[tid * kLd + j] + off;
    diag[tid] = 0.0 - off;
  }
  __syncthreads();
  if (tid == 0) {
    double scale = 0.0;
    for (int i = 0; i < 64; ++i)
      if (diag[i] != 0.0) scale = scale - pi[i] * diag[i];
    scale_s = scale;
  }
  __syncthreads();
  const double scale = scale_s;
  // A = Pi^1/2 Q Pi^-1/2 symmetrised (the host's formula), U = I; the total
  // squared mass for the stopping rule
  double part = 0.0;
  for (int e = tid; e < 4096; e += 256) {
    const int i = e >> 6, j = e & 63;
    const double qij = ((i == j) ? diag[i] : A[i * kLd + j]) / scale;
    const double qji = ((i == j) ? diag[i] : A[j * kLd + i]) / scale;
    U[i * kLd + j] = 0.5 * (sq[i] * qij / sq[j] + sq[j] * qji / sq[i]);
  }
  __syncthreads();
  double* vt = VT + (size_t)set * 4096;
  double* vi = Vinv + (size_t)set * 4096;
  double* u0 = ref + (size_t)set * 4096;
  for (int e = tid; e < 4096; e += 256) {
    const int i = e >> 6, j = e & 63;
    const double x = U[i * kLd + j];
    A[i * kLd + j] = x;
    U[i * kLd + j] = (md == kEigWarm) ? u0[e] : (i == j) ? 1.0 : 0.0;
    part = fma(x, x, part);
  }
  if (md == kEigWarm) {
    // A <- U0^T (A U0), each product through registers (16 entries per
    // thread) so that no third LDS matrix is needed
    double t[16];
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int e = tid + 256 * k, i = e >> 6, j = e & 63;
      double x = 0.0;
      for (int l = 0; l < 64; ++l) x = fma(A[i * kLd + l], U[l * kLd + j], x);
      t[k] = x;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int e = tid + 256 * k;
      A[(e >> 6) * kLd + (e & 63)] = t[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int e = tid + 256 * k, i = e >> 6, j = e & 63;
      double x = 0.0;
      for (int l = 0; l < 64; ++l) x = fma(U[l * kLd + i], A[l * kLd + j], x);
      t[k] = x;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 16; ++k) {
      const int e = tid + 256 * k, i = e >> 6, j = e & 63;
      A[i * kLd + j] = t[k];
    }
    __syncthreads();
    if (tid < 64) {  // symmetric up to rounding: the mean of the two halves
      for (int j = tid + 1; j < 64; ++j) {
        const double m = 0.5 * (A[tid * kLd + j] + A[j * kLd + tid]);
        A[tid * kLd + j] = m;
        A[j * kLd + tid] = m;
      }
    }
  }
  red[tid] = part;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) red[tid] += red[tid + w];
    __syncthreads();
  }
  if (tid == 0) total_s = red[0];
  __syncthreads();
  const double total = total_s;
  double prev = INFINITY;
  for (int sweep = 0; sweep < 64; ++sweep) {
    // off-diagonal mass
    double o = 0.0;
    for (int e = tid; e < 4096; e += 256) {
      const int i = e >> 6, j = e & 63;
      if (i != j) o = fma(A[i * kLd + j], A[i * kLd + j], o);
    }
    red[tid] = o;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
      if (tid < w) red[tid] += red[tid + w];
      __syncthreads();
    }
    const double off = red[0];
    const bool converged = !(off > 1e-30 * total) || !(off < prev);
    prev = off;
    __syncthreads();
    if (converged) break;
    for (int round = 0; round < 63; ++round) {
      if (tid < 32) {
        // round-robin pairing: 63 fixed, the others on a circle
        int a, b;
        if (tid == 0) {
          a = 63;
          b = round;
        } else {
          a = (round + tid) % 63;
          b = (round - tid + 63) % 63;
        }
        const int p = min(a, b), q = max(a, b);
        const double apq = A[p * kLd + q];
        double c = 1.0, s = 0.0;
        if (fabs(apq) >= 1e-300) {
          const double tau = (A[q * kLd + q] - A[p * kLd + p]) / (2.0 * apq);
          const double t = copysign(1.0, tau) / (fabs(tau) + hypot(1.0, tau));
          c = 1.0 / hypot(1.0, t);
          s = t * c;
        }
        pp[tid] = p;
        qq[tid] = q;
        cs[tid] = c;
        sn[tid] = s;
      }
      __syncthreads();
      // A <- A J and U <- U J (columns p, q of every pair)
      for (int e = tid; e < 64 * 32; e += 256) {
        const int r = e & 63, k = e >> 6;
        const int p = pp[k], q = qq[k];
        const double c = cs[k], s = sn[k];
        const double x = A[r * kLd + p], y = A[r * kLd + q];
        A[r * kLd + p] = c * x - s * y;
        A[r * kLd + q] = s * x + c * y;
        const double u = U[r * kLd + p], v = U[r * kLd + q];
        U[r * kLd + p] = c * u - s * v;
        U[r * kLd + q] = s * u + c * v;
      }
      __syncthreads();
      // A <- J^T A (rows p, q of every pair)
      for (int e = tid; e < 64 * 32; e += 256) {
        const int r = e & 63, k = e >> 6;
        const int p = pp[k], q = qq[k];
        const double c = cs[k], s = sn[k];
        const double x = A[p * kLd + r], y = A[q * kLd + r];
        A[p * kLd + r] = c * x - s * y;
        A[q * kLd + r] = s * x + c * y;
      }
      __syncthreads();
    }
  }
  // outputs: V^T[m][i] = U[i][m] / sq[i], V^-1[m][i] = U[i][m] sq[i]; a
  // cold build's U is the set's reference basis
  __syncthreads();
  for (int e = tid; e < 4096; e += 256) {
    const int m = e >> 6, i = e & 63;
    const double u = U[i * kLd + m];
    vt[e] = u / sq[i];
    vi[e] = u * sq[i];
    if (md == kEigCold) u0[e] = U[(e >> 6) * kLd + (e & 63)];
  }
  if (tid < 64) lam[(size_t)set * 64 + tid] = A[tid * kLd + tid];
  if (tid == 0) {
    double tot = 0.0;
    for (int i = 0; i < 64; ++i) tot += pi[i];
    red[0] = tot;
  }
  __syncthreads();
  if (tid < 64) {
    const double pn = pi[tid] / red[0];
    pi_out[(size_t)set * 64 + tid] = pn;
    lnpi_out[(size_t)set * 64 + tid] = log(pn);
  }
}

}  // namespace

hipError_t launch_omega_eigen(const double* d_qs, const uint8_t* d_mode, double* d_ref,
                              int nsets, const PSets& sets, hipStream_t s) {
  if (nsets <= 0) return hipSuccess;
  hipLaunchKernelGGL(omega_eigen_kernel, dim3(nsets), dim3(256), 0, s, d_qs, d_mode, d_ref,
                     sets.d_eVT, sets.d_eVinv, sets.d_elam, sets.d_epi, sets.d_elnpi);
  return hipGetLastError();
}
