// fp64_tiles.hpp -- FP64 64 x 16 column tiles on v_mfma_f64_16x16x4f64,
// shared by the FP64 matrix-core kernels (the outside pass and forward in
// adjoint_mfma.hip, the E-step in estep.hip).
//
// A wave owns 16 columns as a 64 x 16 tile: lane j + 16 g holds column j,
// rows 16 mo + 4 v + g in element v of r[mo] -- the accumulator layout of
// v_mfma_f64_16x16x4f64, which is also its B operand for the K-steps taken
// in natural order.  The four waves of a workgroup share 64 x 64 matrices in
// LDS in the MFMA's A-fragment order (element (mo, s, lane) at
// ((16 mo + s) 64 + lane): conflict-free 8-byte reads).  Every product runs
// in wave-uniform control flow: the A fragments come from LDS through all 64
// lanes.
#pragma once
#include "../internal.hpp"

namespace fp64t {

typedef double f64x4 __attribute__((ext_vector_type(4)));

struct Tile {
  f64x4 r[4];
};

constexpr int kWaves = 4;                    // waves per workgroup
constexpr int kCols = 16 * kWaves;           // columns per workgroup (= kAlignCols)
static_assert(kCols == kAlignCols, "a workgroup covers one region-aligned block");
constexpr int kTileDoubles = 64 * 16;        // one 64 x 16 tile

// A 64 x 64 matrix M into LDS in A-fragment order: M[row][col] goes to
// ((row / 16) * 16 + col / 4) * 64 + (row % 16) + 16 (col % 4).  `src`
// gives M either as M[row][col] = src[row * 64 + col] (by_rows) or as
// M[row][col] = src[col * 64 + row]; the global reads run in src order.
// fetch() takes a thread's share (elements tid + 256 k) into registers, so
// that the next tile's reads are in flight during a matrix product; put()
// writes the share to LDS.
constexpr int kShare = 4096 / (64 * kWaves);
struct Share {
  double v[kShare];
};
__device__ __forceinline__ void fetch(Share& r, const double* __restrict__ src, int tid) {
#pragma unroll
  for (int k = 0; k < kShare; ++k) r.v[k] = src[tid + 64 * kWaves * k];
}
__device__ __forceinline__ void put(double* __restrict__ frag, const Share& r, bool by_rows,
                                    int tid) {
#pragma unroll
  for (int k = 0; k < kShare; ++k) {
    const int e = tid + 64 * kWaves * k, hi = e >> 6, lo = e & 63;
    const int row = by_rows ? hi : lo, col = by_rows ? lo : hi;
    frag[((row >> 4) * 16 + (col >> 2)) * 64 + (row & 15) + 16 * (col & 3)] = r.v[k];
  }
}
// both at once, a few elements at a time (fewer registers)
__device__ __forceinline__ void stage(double* __restrict__ frag, const double* __restrict__ src,
                                      bool by_rows, int tid) {
#pragma unroll 4
  for (int e = tid; e < 4096; e += 64 * kWaves) {
    const int hi = e >> 6, lo = e & 63;
    const int row = by_rows ? hi : lo, col = by_rows ? lo : hi;
    frag[((row >> 4) * 16 + (col >> 2)) * 64 + (row & 15) + 16 * (col & 3)] = src[e];
  }
}

typedef __attribute__((address_space(3))) double L3;
typedef __attribute__((address_space(3))) void* LdsPtr;

// A 4096-double tile already in A-fragment order in memory, copied straight
// into LDS (16 bytes per lane at the wave's base + 16 lane, the order the tile
// has in memory): no registers, no LDS stores.  A barrier before its first
// read waits for it (its release fence waits for the copy).
__device__ __forceinline__ void copy_tile(L3* dst, const double* src, int wave, int tid) {
  typedef const __attribute__((address_space(1))) void* GPtr;
#pragma unroll
  for (int k = 0; k < 4096 / 2 / (64 * kWaves); ++k)
    __builtin_amdgcn_global_load_lds((GPtr)(src + (tid + 64 * kWaves * k) * 2),
                                     (LdsPtr)(dst + (64 * wave + 64 * kWaves * k) * 2), 16, 0, 0);
}

// acc = M B for the staged M and this wave's tile B
__device__ __forceinline__ Tile gemm(const double* __restrict__ frag, const Tile& b, int lane) {
  Tile acc;
#pragma unroll
  for (int mo = 0; mo < 4; ++mo) {
    f64x4 c = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int s = 0; s < 16; ++s)
      c = __builtin_amdgcn_mfma_f64_16x16x4f64(frag[(mo * 16 + s) * 64 + lane], b.r[s >> 2][s & 3],
                                               c, 0, 0, 0);
    acc.r[mo] = c;
  }
  return acc;
}

__device__ __forceinline__ void store(double* __restrict__ slot, const Tile& t, int lane) {
#pragma unroll
  for (int mo = 0; mo < 4; ++mo)
#pragma unroll
    for (int v = 0; v < 4; ++v) slot[(mo * 4 + v) * 64 + lane] = t.r[mo][v];
}

__device__ __forceinline__ Tile load(const double* __restrict__ slot, int lane) {
  Tile t;
#pragma unroll
  for (int mo = 0; mo < 4; ++mo)
#pragma unroll
    for (int v = 0; v < 4; ++v) t.r[mo][v] = slot[(mo * 4 + v) * 64 + lane];
  return t;
}

// sum over the 64 rows of this lane's column of x o y (all four lane groups
// of the column end with it)
__device__ __forceinline__ double column_dot(const Tile& x, const Tile& y) {
  double s = 0.0;
#pragma unroll
  for (int mo = 0; mo < 4; ++mo)
#pragma unroll
    for (int v = 0; v < 4; ++v) s = fma(x.r[mo][v], y.r[mo][v], s);
  s += __shfl_xor(s, 16);
  s += __shfl_xor(s, 32);
  return s;
}

}  // namespace fp64t
