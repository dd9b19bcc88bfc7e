// misc.hip -- batch packing and per-region reduction kernels.
//
//   transpose_codes: the column-major leaf arrays PhyloCSF.pleaves yields
//     (src/PhyloCSF.ml:219-246; here n_cols x nl bytes) become leaf-major
//     rows of ld bytes, so a wave reads 32 consecutive columns of one leaf
//     in one transaction; padding columns are `Marginalize.
//   region_reduce: lpr = sum ln z and elpr_anc = sum post_root . ln pi over
//     each region's columns (src/PhyloCSFModel.ml:76-81), FP64, one
//     workgroup per (region, model), fixed order so results are bitwise
//     reproducible run to run.
#include "../internal.hpp"

namespace {

__global__ void transpose_codes_kernel(const uint8_t* __restrict__ src, int64_t ncols, int nl,
                                       uint8_t* __restrict__ dst, int64_t ld,
                                       int* __restrict__ bad) {
  // 64 columns x nl leaves per workgroup, staged through LDS
  __shared__ uint8_t tile[64 * 257];
  const int64_t c0 = (int64_t)blockIdx.x * 64;
  bool invalid = false;
  for (int l0 = 0; l0 < nl; l0 += 256) {
    const int lw = min(256, nl - l0);
    for (int e = threadIdx.x; e < 64 * lw; e += blockDim.x) {
      const int c = e / lw, l = e - c * lw;
      const int64_t col = c0 + c;
      const uint8_t x = (col < ncols) ? src[col * nl + l0 + l] : (uint8_t)PCSF_MARGINALIZE;
      // a code that is neither a codon nor `Marginalize would index past the
      // 65-row leaf tables: it is stored as `Marginalize and reported
      const bool ok = x < 64 || x == PCSF_MARGINALIZE;
      invalid = invalid || !ok;
      tile[c * 257 + l] = ok ? x : (uint8_t)PCSF_MARGINALIZE;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 64 * lw; e += blockDim.x) {
      const int l = e / 64, c = e - l * 64;
      const int64_t col = c0 + c;
      if (col < ld) dst[(size_t)(l0 + l) * ld + col] = tile[c * 257 + l];
    }
    __syncthreads();
  }
  if (bad && __any(invalid) && (threadIdx.x & 63) == 0) atomicOr(bad, 1);
}

__global__ __launch_bounds__(256) void region_reduce_kernel(
    const double* __restrict__ logz, const double* __restrict__ colelpr,
    const int64_t* __restrict__ off, int nm, int64_t ld, double* __restrict__ loglik,
    double* __restrict__ elpr) {
  __shared__ double s0[256], s1[256];
  const int r = blockIdx.x, q = blockIdx.y, tid = threadIdx.x;
  const int64_t lo = off[r], hi = off[r + 1];
  double a = 0.0, b = 0.0;
  for (int64_t c = lo + tid; c < hi; c += 256) {
    a += logz[(size_t)q * ld + c];
    b += colelpr[(size_t)q * ld + c];
  }
  s0[tid] = a;
  s1[tid] = b;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) {
      s0[tid] += s0[tid + w];
      s1[tid] += s1[tid + w];
    }
    __syncthreads();
  }
  if (tid == 0) {
    loglik[(size_t)r * nm + q] = s0[0];
    elpr[(size_t)r * nm + q] = s1[0];
  }
}

// Region-aligned copy: aligned column a of region r (astart[r] <= a <
// astart[r] + len) holds packed column off[r] + (a - astart[r]); the rest of
// the region's last block is `Marginalize.
__global__ void align_codes_kernel(const uint8_t* __restrict__ src, int64_t ld, int nl,
                                   const int32_t* __restrict__ block_region,
                                   const int64_t* __restrict__ astart,
                                   const int64_t* __restrict__ off, uint8_t* __restrict__ dst,
                                   int64_t ald) {
  const int64_t a = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (a >= ald) return;
  const int r = block_region[a / kAlignCols];
  int64_t pc = -1;
  if (r >= 0) {
    const int64_t k = a - astart[r];
    if (k < off[r + 1] - off[r]) pc = off[r] + k;
  }
  for (int l = 0; l < nl; ++l)
    dst[(size_t)l * ald + a] = (pc >= 0) ? src[(size_t)l * ld + pc] : (uint8_t)PCSF_MARGINALIZE;
}

__global__ __launch_bounds__(256) void region_reduce_aligned_kernel(
    const double* __restrict__ logz, const double* __restrict__ colelpr,
    const int64_t* __restrict__ off, const int64_t* __restrict__ astart,
    const uint8_t* __restrict__ active, int r0, int nm, int64_t ald,
    double* __restrict__ loglik, double* __restrict__ elpr, double* __restrict__ sq) {
  __shared__ double s0[256], s1[256], s2[256];
  const int r = r0 + blockIdx.x, q = blockIdx.y, tid = threadIdx.x;
  if (active && !active[(size_t)blockIdx.x * nm + q]) return;
  const int64_t lo = astart[r], hi = lo + (off[r + 1] - off[r]);
  double a = 0.0, b = 0.0, d = 0.0;
  for (int64_t c = lo + tid; c < hi; c += 256) {
    a += logz[(size_t)q * ald + c];
    const double e = colelpr[(size_t)q * ald + c];
    b += e;
    d = fma(e, e, d);
  }
  s0[tid] = a;
  s1[tid] = b;
  s2[tid] = d;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (tid < w) {
      s0[tid] += s0[tid + w];
      s1[tid] += s1[tid + w];
      s2[tid] += s2[tid + w];
    }
    __syncthreads();
  }
  if (tid == 0) {
    loglik[(size_t)r * nm + q] = s0[0];
    elpr[(size_t)r * nm + q] = s1[0];
    if (sq) sq[(size_t)r * nm + q] = s2[0];
  }
}

}  // namespace

hipError_t launch_align_codes(const pcsf_batch* b, hipStream_t s) {
  const int64_t blocks = (b->al.ld + 255) / 256;
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(align_codes_kernel, dim3((unsigned)blocks), dim3(256), 0, s, b->d_codes,
                     b->ld, b->m->nl, b->al.d_block_region, b->al.d_astart, b->d_offsets,
                     b->al.d_codes, b->al.ld);
  return hipGetLastError();
}

hipError_t launch_region_reduce_aligned(const pcsf_batch* b, const RegionChunk& ch,
                                        double* loglik, double* elpr, hipStream_t s,
                                        double* sq) {
  if (ch.nreg == 0) return hipSuccess;
  dim3 grid(ch.nreg, b->m->nm);
  // the model's own sets (one per model, shared by every region) are all live
  const uint8_t* active = (ch.sets == &b->m->fixed) ? nullptr : ch.sets->d_active;
  hipLaunchKernelGGL(region_reduce_aligned_kernel, grid, dim3(256), 0, s, b->d_alogz,
                     b->d_acolelpr, b->d_offsets, b->al.d_astart, active, ch.r0,
                     b->m->nm, b->al.ld, loglik, elpr, sq);
  return hipGetLastError();
}

hipError_t launch_transpose_codes(const uint8_t* src, int64_t ncols, int nl, uint8_t* dst,
                                  int64_t ld, int* bad, hipStream_t s) {
  const int64_t blocks = (ld + 63) / 64;
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(transpose_codes_kernel, dim3((unsigned)blocks), dim3(256), 0, s, src,
                     ncols, nl, dst, ld, bad);
  return hipGetLastError();
}

namespace {

__device__ __forceinline__ int nuc_index(uint8_t ch) {  // Code.DNA.index; -1: not ACGTacgt
  switch (ch) {
    case 'A': case 'a': return 0;
    case 'C': case 'c': return 1;
    case 'G': case 'g': return 2;
    case 'T': case 't': return 3;
    default: return -1;
  }
}

// Codon codes of every reading frame of packed nucleotide regions: the
// candidate regions of PhyloCSF's AsIs mode with --frames 3 or 6
// (src/PhyloCSF.ml:198-206) read by pleaves (:219-246).  One workgroup per
// (region, frame); frame f reads codons from offset f % 3 of the forward
// strand (f < 3) or of the reverse complement (Code.DNA.revcomp: the
// complement of A,C,G,T is T,G,C,A, i.e. index 3 - i).  A codon with any
// character outside ACGTacgt is `Marginalize.  nuc holds nl rows of nlen
// characters; dst is the batch's leaf-major code array.
__global__ __launch_bounds__(256) void frames_kernel(const uint8_t* __restrict__ nuc,
                                                     int64_t nlen,
                                                     const int64_t* __restrict__ roff,
                                                     const int64_t* __restrict__ foff, int F,
                                                     int nl, uint8_t* __restrict__ dst,
                                                     int64_t ld) {
  const int64_t rf = blockIdx.x;
  const int64_t r = rf / F;
  const int f = (int)(rf - r * F), o = f % 3;
  const bool rc = f >= 3;
  const int64_t base = roff[r], L = roff[r + 1] - base;
  const int64_t col0 = foff[rf], count = foff[rf + 1] - col0;
  for (int64_t e = threadIdx.x; e < count * nl; e += blockDim.x) {
    const int leaf = (int)(e / count);
    const int64_t j = e - (int64_t)leaf * count;
    const uint8_t* row = nuc + (size_t)leaf * nlen + base;
    int a, b, c;
    if (!rc) {
      const int64_t p = o + 3 * j;
      a = nuc_index(row[p]);
      b = nuc_index(row[p + 1]);
      c = nuc_index(row[p + 2]);
    } else {
      const int64_t q = L - 1 - (o + 3 * j);
      a = nuc_index(row[q]);
      b = nuc_index(row[q - 1]);
      c = nuc_index(row[q - 2]);
      a = a < 0 ? -1 : 3 - a;
      b = b < 0 ? -1 : 3 - b;
      c = c < 0 ? -1 : 3 - c;
    }
    dst[(size_t)leaf * ld + col0 + j] =
        (a < 0 || b < 0 || c < 0) ? (uint8_t)PCSF_MARGINALIZE : (uint8_t)(16 * a + 4 * b + c);
  }
}

}  // namespace

hipError_t launch_frames(const uint8_t* nuc, int64_t nlen, const int64_t* d_roff,
                         const int64_t* d_foff, int32_t nreg, int frames, int nl, uint8_t* dst,
                         int64_t ld, hipStream_t s) {
  const int64_t blocks = (int64_t)nreg * frames;
  if (blocks == 0) return hipSuccess;
  hipLaunchKernelGGL(frames_kernel, dim3((unsigned)blocks), dim3(256), 0, s, nuc, nlen, d_roff,
                     d_foff, frames, nl, dst, ld);
  return hipGetLastError();
}

hipError_t launch_region_reduce(const pcsf_batch* b, hipStream_t s) {
  if (b->nreg == 0) return hipSuccess;
  dim3 grid(b->nreg, b->m->nm);
  hipLaunchKernelGGL(region_reduce_kernel, grid, dim3(256), 0, s, b->d_logz, b->d_colelpr,
                     b->d_offsets, b->m->nm, b->ld, b->d_loglik, b->d_elpr);
  return hipGetLastError();
}
