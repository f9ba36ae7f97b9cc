// reduce.h -- deterministic block and grid reductions for 64-wide wavefronts.
//
// Grid reductions use the last-block-done pattern: each workgroup writes its
// partial sums, publishes them with an agent-scope release fence and takes a
// ticket; the workgroup that draws the last ticket acquires and sums all the
// partials in a fixed order, so the result does not depend on scheduling.  The
// ticket counter is reset by that last workgroup (no memset is needed, which
// keeps these kernels safe to capture in graphs).
#pragma once

#include <hip/hip_runtime.h>

namespace lpmg {

template <int NV>
struct Vec {
  double v[NV];
};

template <int NV>
__device__ inline Vec<NV> wave_sum(Vec<NV> a) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int k = 0; k < NV; ++k) a.v[k] += __shfl_xor(a.v[k], o, 64);
  return a;
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

// Publish this block's partial; returns true (in every thread of the block) for
// the last block, which then holds the grid total in *out.  Must be reached by
// every thread of every block.
template <int BS, int NV>
__device__ inline bool last_block_sum(Vec<NV> blocktot, double* partials, unsigned int* counter,
                                      Vec<NV>* out) {
  __shared__ int is_last;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) partials[(size_t)NV * blockIdx.x + k] = blocktot.v[k];
    __threadfence();
    const unsigned int ticket = atomicAdd(counter, 1u);
    is_last = (ticket == gridDim.x - 1);
  }
  __syncthreads();
  if (!is_last) return false;
  __threadfence();
  Vec<NV> a;
#pragma unroll
  for (int k = 0; k < NV; ++k) a.v[k] = 0;
  for (int b = threadIdx.x; b < (int)gridDim.x; b += BS)
#pragma unroll
    for (int k = 0; k < NV; ++k)
      a.v[k] += __hip_atomic_load(&partials[(size_t)NV * b + k], __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
  *out = block_sum<BS, NV>(a);
  if (threadIdx.x == 0) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

template <int BS>
__device__ inline double2 block_sum2(double2 v) {
  Vec<2> a;
  a.v[0] = v.x;
  a.v[1] = v.y;
  a = block_sum<BS, 2>(a);
  return make_double2(a.v[0], a.v[1]);
}

template <int BS = 256>
__device__ inline bool last_block_sum2(double2 blocktot, double* partials, unsigned int* counter,
                                       double2* out) {
  Vec<2> a, o;
  a.v[0] = blocktot.x;
  a.v[1] = blocktot.y;
  const bool last = last_block_sum<BS, 2>(a, partials, counter, &o);
  if (last) *out = make_double2(o.v[0], o.v[1]);
  return last;
}

}  // namespace lpmg
