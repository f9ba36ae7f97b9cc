// bsr.h -- the block-CSR row kernel shared by the SpMV, the CG and the
// multigrid smoothers: one lane per b x b block, LPR lanes per block row.
#pragma once

#include "lpmg_internal.h"
#include "reduce.h"

namespace lpmg {

constexpr int kBS = 256;         // vector kernels
constexpr int kSpBS = 1024;      // SpMV kernels: 16 wavefronts per workgroup
constexpr int kMaxGrid = 1024;   // workgroups of the grid-strided vector kernels
constexpr int kMaxParts = 4096;  // SpMV partials summed directly by the consumer

// y = A x for the rows of wavefronts of workgroup `blk` (64 / LPR rows per
// wavefront, LPR lanes each).  With DOT, returns this lane's share of x.y
// (nonzero only in the row's first b lanes).
template <int D, typename T, int LPR, bool DOT>
__device__ __forceinline__ double bsr_rows(int blk, const int* __restrict__ cptr, const int* __restrict__ conns,
                                           const T* __restrict__ val, const T* __restrict__ x,
                                           T* __restrict__ y, int N) {
  constexpr int DD = D * D;
  constexpr int RPW = 64 / LPR;
  const int lane = threadIdx.x & 63, sub = lane / LPR, sl = lane - sub * LPR;
  const int wave = blk * (kSpBS / 64) + (threadIdx.x >> 6);
  const int row = wave * RPW + sub;
  const bool live = row < N;
  const int rw = live ? row : N - 1;  // dead rows shadow the last row: all lanes stay in the shuffles
  const int lo = cptr[rw], hi = cptr[rw + 1];
  T acc[D];
#pragma unroll
  for (int q = 0; q < D; ++q) acc[q] = T(0);
  for (int e0 = lo; e0 < hi; e0 += LPR) {
    const int e = e0 + sl;
    const bool ok = e < hi;
    const int ec = ok ? e : hi - 1;  // every row holds its self block, so hi - 1 >= lo
    const int col = conns[ec];
    const T* vb = val + (size_t)ec * DD;
    T v[DD], xv[D];
#pragma unroll
    for (int k = 0; k < DD; ++k) v[k] = vb[k];
#pragma unroll
    for (int k = 0; k < D; ++k) xv[k] = x[(size_t)col * D + k];
#pragma unroll
    for (int q = 0; q < D; ++q) {
      T t = v[q * D] * xv[0];
#pragma unroll
      for (int k = 1; k < D; ++k) t += v[q * D + k] * xv[k];
      acc[q] += ok ? t : T(0);
    }
  }
#pragma unroll
  for (int o = LPR / 2; o > 0; o >>= 1)
#pragma unroll
    for (int q = 0; q < D; ++q) acc[q] += __shfl_xor(acc[q], o, 64);
  double dot = 0;
  if (live && sl < D) {
    T yv = acc[0];
#pragma unroll
    for (int q = 1; q < D; ++q)
      if (sl == q) yv = acc[q];
    y[(size_t)row * D + sl] = yv;
    if (DOT) dot = (double)x[(size_t)row * D + sl] * (double)yv;
  }
  return dot;
}

}  // namespace lpmg
