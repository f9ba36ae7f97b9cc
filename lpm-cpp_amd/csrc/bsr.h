// bsr.h -- the block-CSR row kernel shared by the SpMV, the CG and the
// multigrid smoothers: one lane per b x b block, LPR lanes per block row.
#pragma once

#include "lpmg_internal.h"
#include "reduce.h"

namespace lpmg {

constexpr int kBS = 256;         // vector kernels
// SpMV kernels: 8 wavefronts (32 block rows) per workgroup.  The CG SpMV ends
// in a workgroup reduction of its dot product, whose barrier holds every
// wavefront until the slowest row is done.  Measured per CG SpMV with 16-lane
// rows: 1024 threads 52.5 us on SC 40^3; 128 and 256 threads within 2% of 512
// on SC 40^3 but 4% slower on the 250k CT plate (45.0 vs 43.2 us back to back).
constexpr int kSpBS = 512;
constexpr int kMaxGrid = 1024;   // workgroups of the grid-strided vector kernels
constexpr int kMaxParts = 8192;  // SpMV partials summed directly by the consumer

// Lanes per block row of the SpMV and CG kernels: four rows per wavefront.
constexpr int kLPR = 16;

// Chunks of kLPR blocks per pass over a row: enough for the system's longest
// row (2 for the hexagonal lattice's 31 blocks, 3 for BCC's 41, 4 for SC's and
// FCC's 61); longer rows take several passes.  Chunks past every row of a
// wavefront cost issue slots, so the pass is sized to the longest row.
inline int chunks_per_pass(const Context* c) {
  return std::min(4, std::max(2, (c->max_nconn + kLPR - 1) / kLPR));
}

// acc = this lane's row's share of A x for blocks lo..hi-1 (LPR lanes per row,
// U chunks of LPR blocks per pass), summed across the row's lanes: every lane
// of the row ends with the row's D sums.
template <int D, typename T, int LPR, int U>
__device__ __forceinline__ void bsr_row_acc(int lo, int hi, int sl, const int* __restrict__ conns,
                                            const T* __restrict__ val, const T* __restrict__ x, T acc[D]) {
#pragma unroll
  for (int q = 0; q < D; ++q) acc[q] = T(0);
  for (int e0 = lo; e0 < hi; e0 += U * LPR) {
    // Block columns and values of every chunk (unconditional: past the row's
    // end they repeat its last block and `ok` drops the product), then the x
    // gathers the columns address, all chunks' gathers in flight together.
    // With three or four chunks the values are issued before the gathers (they
    // stream in while the gathers wait for their columns); with two, after
    // them, skipping a chunk that no row of the wavefront reaches.  The
    // scheduling barriers keep each group together (measured on the SC 40^3
    // CG SpMV: 44.1 us against 46.5 with the values last; FP32 CT plate, two
    // chunks: 25.1 us with the values last against 27.3).
    int col[U];
    bool ok[U];
    int ec[U];
    T v[U][D * D];
    T xv[U][D];
#pragma unroll
    for (int c = 0; c < U; ++c) {
      const int e = e0 + c * LPR + sl;
      ok[c] = e < hi;
      ec[c] = ok[c] ? e : hi - 1;
      col[c] = conns[ec[c]];
    }
    if constexpr (U >= 3) {
#pragma unroll
      for (int c = 0; c < U; ++c)
#pragma unroll
        for (int k = 0; k < D * D; ++k) v[c][k] = val[(size_t)ec[c] * D * D + k];
      __builtin_amdgcn_sched_barrier(0);
#pragma unroll
      for (int c = 0; c < U; ++c)
#pragma unroll
        for (int k = 0; k < D; ++k) xv[c][k] = x[(size_t)col[c] * D + k];
      __builtin_amdgcn_sched_barrier(0);
    } else {
#pragma unroll
      for (int c = 0; c < U; ++c)
#pragma unroll
        for (int k = 0; k < D; ++k) xv[c][k] = x[(size_t)col[c] * D + k];
#pragma unroll
      for (int c = 0; c < U; ++c) {
        if (c == 0 || __any(e0 + c * LPR < hi)) {  // wavefront-uniform
#pragma unroll
          for (int k = 0; k < D * D; ++k) v[c][k] = val[(size_t)ec[c] * D * D + k];
        } else {
#pragma unroll
          for (int k = 0; k < D * D; ++k) v[c][k] = T(0);
        }
      }
    }
#pragma unroll
    for (int c = 0; c < U; ++c)
#pragma unroll
      for (int q = 0; q < D; ++q) {
        T t = v[c][q * D] * xv[c][0];
#pragma unroll
        for (int k = 1; k < D; ++k) t += v[c][q * D + k] * xv[c][k];
        acc[q] += ok[c] ? t : T(0);
      }
  }
#pragma unroll
  for (int o = LPR / 2; o > 0; o >>= 1)
#pragma unroll
    for (int q = 0; q < D; ++q) acc[q] += __shfl_xor(acc[q], o, 64);
}

// y = A x for the rows of wavefronts of workgroup `blk` (64 / LPR rows per
// wavefront, LPR lanes each, U chunks of LPR blocks per pass over a row).
// With DOT, returns this lane's share of x.y (nonzero only in the row's first
// b lanes).
//
// A row's pass is latency-bound: row start -> block columns -> x gathers are
// three dependent loads.  A pass issues the block columns and values of all U
// chunks and all their x gathers before it waits for the products: one latency
// for the gathers of the whole pass.  With LPR < 64, 64 / LPR rows per
// wavefront keep that many times the bytes in flight.
//
// Rows go in launch order: consecutive workgroups (dispatched to the XCDs in
// turn) sweep the matrix as one sequential stream.  Giving each XCD's
// workgroups one contiguous range of rows instead (so each L2 gathers only its
// rows' part of x) measured 4-7% faster on the 250k CT plate and FP32 SC 40^3
// but 2-8% slower on the 1.0M polycrystal and 1.6 us slower in the SC 40^3
// FP64 CG SpMV, so it is not used.  Each lane reads its block's b*b values as
// one contiguous run; interleaving the values of every 16 consecutive blocks
// (value k of the 16 blocks contiguous, so a row's lanes read 128-byte runs)
// was measured 6-14% slower in FP64 and 4-22% in FP32 on SC 40^3, the 250k CT
// plate, the BCC bar and the polycrystal (510 -> 580 us FP64 there).  Reading
// the blocks below the diagonal as transposes of their mirrors (storing the
// upper triangle only, half the value bytes, the same bits) measured slower
// too except on the BCC bar: SC 40^3 42.6 -> 53.1 us, polycrystal 512 -> 558
// us FP64 -- a row's transposed blocks sit in 16 unrelated places.
template <int D, typename T, int LPR, int U, bool DOT>
__device__ __forceinline__ double bsr_rows(int blk, const int* __restrict__ cptr, const int* __restrict__ conns,
                                           const T* __restrict__ val, const T* __restrict__ x,
                                           T* __restrict__ y, int N) {
  constexpr int RPW = 64 / LPR;
  const int lane = threadIdx.x & 63, sub = lane / LPR, sl = lane - sub * LPR;
  const int wave = blk * (kSpBS / 64) + (threadIdx.x >> 6);
  const int row = wave * RPW + sub;
  const bool live = row < N;
  const int rw = live ? row : N - 1;  // dead rows shadow the last row: all lanes stay in the shuffles
  const int lo = cptr[rw], hi = cptr[rw + 1];
  // With DOT: this lane's own entry of x for the dot product, loaded with the
  // row's extent.
  const T xo = DOT ? x[(size_t)rw * D + min(sl, D - 1)] : T(0);
  T acc[D];
  bsr_row_acc<D, T, LPR, U>(lo, hi, sl, conns, val, x, acc);
  double dot = 0;
  if (live && sl < D) {
    T yv = acc[0];
#pragma unroll
    for (int q = 1; q < D; ++q)
      if (sl == q) yv = acc[q];
    y[(size_t)row * D + sl] = yv;
    if (DOT) dot = (double)xo * (double)yv;
  }
  return dot;
}

}  // namespace lpmg
