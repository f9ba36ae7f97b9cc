// reduce.h -- deterministic block and grid reductions for 64-wide wavefronts.
//
// A grid reduction is split across a kernel boundary: each workgroup of the
// producing kernel writes its partial sum (plain stores); a later kernel sums
// the partials in a fixed order.  No kernel waits on other workgroups of the
// same launch, so no agent-scope fence is issued inside a kernel (on gfx950 such
// a fence writes back and invalidates the XCD's L2, which costs more than the
// reduction itself), and the result does not depend on scheduling.
#pragma once

#include <hip/hip_runtime.h>

namespace lpmg {

template <int NV>
struct Vec {
  double v[NV];
};

// Workgroups are dispatched to the 8 XCDs in turn (workgroup b to XCD b % 8),
// and each XCD has its own L2.  This maps the launch's workgroup b to a virtual
// index such that each XCD takes one contiguous range of about n / 8 virtual
// workgroups, in order: kernels whose neighbouring rows read the same data
// (particle neighbourhoods) then share it in one L2.  A bijection on [0, n).
__device__ __forceinline__ int xcd_contiguous(int b, int n) {
  constexpr int kXcd = 8;
  const int q = n / kXcd, rem = n % kXcd;
  const int x = b % kXcd, l = b / kXcd;
  return x * q + min(x, rem) + l;
}

template <int NV>
__device__ inline Vec<NV> wave_sum(Vec<NV> a) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int k = 0; k < NV; ++k) a.v[k] += __shfl_xor(a.v[k], o, 64);
  return a;
}

// Inclusive prefix sum across one 64-wide wavefront.
template <typename T>
__device__ inline T wave_incl_scan(T v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    T t = __shfl_up(v, o, 64);
    if (lane >= o) v += t;
  }
  return v;
}

// Sum over the workgroup; the result is valid in every thread.
template <int BS, int NV>
__device__ inline Vec<NV> block_sum(Vec<NV> a) {
  __shared__ double red[BS / 64][NV];
  a = wave_sum(a);
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) red[w][k] = a.v[k];
  __syncthreads();
  Vec<NV> t;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    t.v[k] = 0;
    for (int q = 0; q < BS / 64; ++q) t.v[k] += red[q][k];
  }
  __syncthreads();
  return t;
}

// Workgroup `blk`'s partial goes to part[NV*blk ...].  Every thread of the
// workgroup must call it.
template <int BS, int NV>
__device__ inline void write_partial_at(double* __restrict__ part, int blk, Vec<NV> a) {
  a = block_sum<BS, NV>(a);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) part[(size_t)NV * blk + k] = a.v[k];
}

template <int BS, int NV>
__device__ inline void write_partial(double* __restrict__ part, Vec<NV> a) {
  write_partial_at<BS, NV>(part, blockIdx.x, a);
}

// This thread's share of n partials (entries threadIdx.x, +BS, +2BS, ...,
// NV values each) added into a in that order.  Eight entries are loaded before
// any is added, so their loads are in flight together (a plain loop waits for
// each load before issuing the next); the loads are unconditional -- entries
// past the end read the last one and are not added -- since a load inside a
// branch is waited for at the branch's end.
template <int BS, int NV>
__device__ inline void add_partials(const double* __restrict__ part, int n, Vec<NV>& a) {
  constexpr int U = 8;
  for (int b0 = threadIdx.x; b0 < n; b0 += U * BS) {
    double v[U][NV];
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const int bj = min(b0 + j * BS, n - 1);
#pragma unroll
      for (int k = 0; k < NV; ++k) v[j][k] = part[(size_t)NV * bj + k];
    }
#pragma unroll
    for (int j = 0; j < U; ++j)
      if (b0 + j * BS < n)
#pragma unroll
        for (int k = 0; k < NV; ++k) a.v[k] += v[j][k];
  }
}

// Sum of n partials written by an earlier kernel, in a fixed order; the same
// bits in every thread of every workgroup.
template <int BS, int NV>
__device__ inline Vec<NV> sum_partials(const double* __restrict__ part, int n) {
  Vec<NV> a;
#pragma unroll
  for (int k = 0; k < NV; ++k) a.v[k] = 0;
  add_partials<BS, NV>(part, n, a);
  return block_sum<BS, NV>(a);
}

}  // namespace lpmg
