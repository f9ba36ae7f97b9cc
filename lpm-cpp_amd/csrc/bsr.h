// bsr.h -- the block-CSR row kernel shared by the SpMV, the CG and the
// multigrid smoothers: one lane per b x b block, LPR lanes per block row.
#pragma once

#include "lpmg_internal.h"
#include "reduce.h"

namespace lpmg {

constexpr int kBS = 256;         // vector kernels
// SpMV kernels: 8 wavefronts per workgroup.  The CG SpMV ends in a workgroup
// reduction of its dot product, whose barrier holds every wavefront until the
// slowest row is done; 8 wavefronts wait less than 16 (SC 40^3: 51.9 vs
// 58.2 us per CG SpMV).
constexpr int kSpBS = 512;
constexpr int kMaxGrid = 1024;   // workgroups of the grid-strided vector kernels
constexpr int kMaxParts = 8192;  // SpMV partials summed directly by the consumer

// One chunk of LPR consecutive blocks of a row: this lane's block e0 + sl
// (clamped to the row's last block when past the end; `ok` says whether it counts).
template <int D, typename T, int LPR>
struct BsrChunk {
  int col;
  bool ok;
  T v[D * D];
  __device__ __forceinline__ void load(int e0, int sl, int hi, const int* __restrict__ conns,
                                       const T* __restrict__ val) {
    const int e = e0 + sl;
    ok = e < hi;
    const int ec = ok ? e : hi - 1;  // every row holds its self block, so hi - 1 >= lo
    col = conns[ec];
    const T* vb = val + (size_t)ec * D * D;
#pragma unroll
    for (int k = 0; k < D * D; ++k) v[k] = vb[k];
  }
  __device__ __forceinline__ void apply(const T* __restrict__ x, T* acc) const {
    T xv[D];
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
};

// y = A x for the rows of wavefronts of workgroup `blk` (64 / LPR rows per
// wavefront, LPR lanes each).  With DOT, returns this lane's share of x.y
// (nonzero only in the row's first b lanes).
//
// A row's pass is latency-bound: row start -> column index -> x gather are
// three dependent loads.  With LPR < 64 lanes per row a pass covers 64 / LPR
// chunks (64 blocks) and issues every chunk's loads before the first chunk's x
// gather waits, so 64 / LPR rows per wavefront keep that many times the bytes
// in flight; a chunk that no row of the wavefront reaches is skipped.
template <int D, typename T, int LPR, bool DOT>
__device__ __forceinline__ double bsr_rows(int blk, const int* __restrict__ cptr, const int* __restrict__ conns,
                                           const T* __restrict__ val, const T* __restrict__ x,
                                           T* __restrict__ y, int N) {
  constexpr int RPW = 64 / LPR;
  constexpr int U = 64 / LPR;
  const int lane = threadIdx.x & 63, sub = lane / LPR, sl = lane - sub * LPR;
  const int wave = blk * (kSpBS / 64) + (threadIdx.x >> 6);
  const int row = wave * RPW + sub;
  const bool live = row < N;
  const int rw = live ? row : N - 1;  // dead rows shadow the last row: all lanes stay in the shuffles
  const int lo = cptr[rw], hi = cptr[rw + 1];
  T acc[D];
#pragma unroll
  for (int q = 0; q < D; ++q) acc[q] = T(0);
  for (int e0 = lo; e0 < hi; e0 += U * LPR) {
    BsrChunk<D, T, LPR> ch[U];
    ch[0].load(e0, sl, hi, conns, val);
    int nc = 1;
#pragma unroll
    for (int c = 1; c < U; ++c)
      if (__any(e0 + c * LPR < hi)) {  // wavefront-uniform
        ch[c].load(e0 + c * LPR, sl, hi, conns, val);
        nc = c + 1;
      }
#pragma unroll
    for (int c = 0; c < U; ++c)
      if (c < nc) ch[c].apply(x, acc);
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
