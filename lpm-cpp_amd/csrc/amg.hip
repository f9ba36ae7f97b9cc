// amg.hip -- aggregation multigrid preconditioners for the linear solve of the
// Newton step (opts precond = 2: tentative, unsmoothed aggregation; 4:
// smoothed aggregation, described before its kernels below), a replacement for
// the Jacobi-PCG that keeps the solve's stopping test (||r||^2 <= rtol ||r0||^2
// + atol, solver.h:312-313).
//
// Hierarchy (built once per handle from the particle positions and the
// block-CSR pattern, on the host):
//   level 0      the handle's BSR stiffness (b = dim)
//   level l+1    aggregates of level-l nodes: the nonempty cells of a uniform
//                grid (cell edge 3 lattice spacings at level 1, x3.5 per level;
//                x2.5 took 3% longer on the 205k- and 109k-particle CT
//                requests, x4 2% longer, 4 spacings at level 1 up to 14%)
//                over the nodes' positions (particles, then aggregate centroids)
//   prolongation piecewise constant per aggregate and component (the
//                translations, unsmoothed), so the Galerkin coarse operator is
//                a plain block sum: A_{l+1}[a,b] = sum over level-l blocks (i,j)
//                with agg(i) = a, agg(j) = b.  Each coarse block owns the list of
//                its level-l blocks, summed in a fixed order (deterministic).
//   coarsest     <= 128 unknowns, inverted in LDS by one workgroup
//                (Gauss-Jordan, no pivoting: SPD, or semidefinite when a crack
//                has cut a piece loose -- vanishing pivots are then dropped and
//                the inverse is taken on the range)
//   values       summed in FP64 from the FP64 stiffness, rounded afterwards
//                for the FP32 inner solve
// The values are rebuilt on the device whenever the stiffness was assembled
// again (every modified-Newton pass).
//
// Preconditioner: one V-cycle with one damped-Jacobi sweep (omega = 2/3) before
// and after the coarse correction at every level.  Per level: a fused
// residual kernel r = b - A (omega D^-1 b), the restriction (sum over each
// aggregate's members), and the post-smoothing sweep, which forms the
// prolonged correction plus the pre-smoothed value where it reads them; the
// finest sweep also writes the partials of
// (r, z) and zeroes z on the DOFs the Dirichlet elimination decoupled (the
// coarse correction would otherwise move them: the preconditioner then acts
// on the free DOFs only, where it stays symmetric positive definite).
//
// Per FP32 iteration on the 205,645-particle CT specimen this takes 153 us
// (the fine-level SpMV alone 21.5 us): the levels below about 8,000 unknowns
// are a few microseconds of work each but three launches.  Running them all
// in one launch of one workgroup (one thread per row, barriers between the
// phases) was measured at 211 us: one CU cannot hide the row loops'
// dependent loads that a launch over the whole GPU covers in one round.
//
// PCG (standard form, two reductions per iteration): w = A p with partials of
// (p,w); x += alpha p, r -= alpha w with partials of (r,r); z = M r (V-cycle)
// with partials of (r,z); stop test, p = z + beta p.  As in the Jacobi CG, no
// kernel waits on other workgroups of its own launch; the iteration index a
// kernel reads is written by an earlier launch, and scalars live across
// iterations in slots indexed by its parity.  The host replays a captured graph
// of 16 iterations between reads of the convergence flag.
#include <algorithm>
#include <array>
#include <cmath>
#include <numeric>
#include <thread>
#include <type_traits>
#include <vector>

#include "bsr.h"

namespace lpmg {

namespace {

constexpr int kCoarsestMax = 128;  // unknowns of the dense coarsest level (fits LDS)
constexpr int kTailChunk = 4;      // PCG iterations of the short chunk that ends a solve
constexpr double kOmega = 2.0 / 3.0;
// Scale of the coarse-grid correction.  Unsmoothed aggregation's piecewise-
// constant prolongation under-corrects the smooth error; scaling the correction
// (below 2, so the symmetric V-cycle stays a positive definite preconditioner)
// takes 9% fewer CG iterations on the 205,645-particle CT specimen at 1.5 than
// at 1 (1.8: 1% more, 2.0: 10% more).
constexpr double kCoarseScale = 1.5;

// ---- device kernels ------------------------------------------------------

// Row gather of A x for this lane's row, reduced over the row's lanes.  XM 0:
// x_j = x[j]; XM 1: x_j = om * dinv[j] * b[j] (one Jacobi sweep from zero);
// XM 2: x_j = om * dinv[j] * b[j] + kCoarseScale * xc[agg[j]] (that sweep plus the coarse
// correction, prolonged piecewise constant).
// A pass covers U chunks of LPR blocks, issued as bsr_rows issues them (block
// columns and values, then the gathers, each group together).
// `wave` is the wavefront's index over the level's rows (RPW = 64 / LPR rows
// each); a row's sums do not depend on which wavefront forms them.
// Block e's D*D values as T, from values stored as T or -- the smoothers of
// the finest level of an FP32 V-cycle -- as bfloat16 bit patterns (round to
// nearest even of the FP32 stiffness): a 2x2 block in one 8-byte word, a 3x3
// block in 20 bytes (nine values and a zero, read as five 4-byte words).  The
// preconditioner only has to be symmetric and close to K: the PCG's own
// operator stays FP32, and the bf16 smoothers take the same CG iterations on
// CT specimens (CPU experiment: 23 / 56 / 80 PCG iterations to 1e-4 / 1e-8 /
// 1e-12 with FP64, FP32 or bf16 smoother values) while reading 12 instead of
// 20 bytes per block: 153 -> 148 us per FP32 multigrid iteration on the
// 205,645-particle specimen (the smoothers' dependent gathers, not their
// bytes, bound them).
constexpr int bf16_stride(int D) { return D == 2 ? 4 : 10; }  // bfloat16 values per stored block

template <int D, typename T, typename TV>
__device__ __forceinline__ void load_block(const TV* __restrict__ val, int e, T v[D * D]) {
  if constexpr (std::is_same<TV, uint16_t>::value) {
    static_assert(std::is_same<T, float>::value, "bfloat16 blocks expand to FP32");
    if constexpr (D == 2) {
      const uint2 w = reinterpret_cast<const uint2*>(val)[e];
      v[0] = __uint_as_float(w.x << 16);
      v[1] = __uint_as_float(w.x & 0xffff0000u);
      v[2] = __uint_as_float(w.y << 16);
      v[3] = __uint_as_float(w.y & 0xffff0000u);
    } else {
      const uint32_t* w = reinterpret_cast<const uint32_t*>(val) + (size_t)e * 5;
      uint32_t x[5];
#pragma unroll
      for (int k = 0; k < 5; ++k) x[k] = w[k];
#pragma unroll
      for (int k = 0; k < D * D; ++k) v[k] = __uint_as_float((k & 1) ? (x[k >> 1] & 0xffff0000u) : (x[k >> 1] << 16));
    }
  } else {
#pragma unroll
    for (int k = 0; k < D * D; ++k) v[k] = val[(size_t)e * D * D + k];
  }
}

// b = bfloat16 of the FP32 blocks a (nb blocks of D*D values), bf16_stride(D)
// values per stored block (the padding zero)
template <int D>
__global__ void bf16_round_kernel(const float* __restrict__ a, uint16_t* __restrict__ b, int64_t nb) {
  constexpr int S = bf16_stride(D);
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nb * S) return;
  const int64_t e = t / S;
  const int k = (int)(t - e * S);
  if (k >= D * D) {
    b[t] = 0;
    return;
  }
  uint32_t u = __float_as_uint(a[e * D * D + k]);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest, ties to even (finite values)
  b[t] = (uint16_t)(u >> 16);
}

template <int D, typename T, int LPR, int U, int XM, typename TV = T>
__device__ __forceinline__ void row_gather(int wave, const int* __restrict__ cptr, const int* __restrict__ conns,
                                           const TV* __restrict__ val, const T* __restrict__ x,
                                           const T* __restrict__ dinv, const T* __restrict__ b, T om, int N,
                                           int* row_out, bool* live_out, int* sl_out, T acc[D],
                                           const int* __restrict__ agg = nullptr, const T* __restrict__ xc = nullptr) {
  constexpr int DD = D * D;
  constexpr int RPW = 64 / LPR;
  const int lane = threadIdx.x & 63, sub = lane / LPR, sl = lane - sub * LPR;
  const int row = wave * RPW + sub;
  const bool live = row < N;
  const int rw = live ? row : N - 1;
  const int lo = cptr[rw], hi = cptr[rw + 1];
#pragma unroll
  for (int q = 0; q < D; ++q) acc[q] = T(0);
  for (int e0 = lo; e0 < hi; e0 += U * LPR) {
    int col[U], ec[U];
    bool ok[U];
    T v[U][DD], xv[U][D];
#pragma unroll
    for (int c = 0; c < U; ++c) {
      const int e = e0 + c * LPR + sl;
      ok[c] = e < hi;
      ec[c] = ok[c] ? e : hi - 1;
      col[c] = conns[ec[c]];
    }
#pragma unroll
    for (int c = 0; c < U; ++c) load_block<D, T, TV>(val, ec[c], v[c]);
    if constexpr (U > 1) __builtin_amdgcn_sched_barrier(0);
    int ag[U];
    if constexpr (XM == 2)
#pragma unroll
      for (int c = 0; c < U; ++c) ag[c] = agg[col[c]];
#pragma unroll
    for (int c = 0; c < U; ++c)
#pragma unroll
      for (int k = 0; k < D; ++k) {
        const size_t q = (size_t)col[c] * D + k;
        if constexpr (XM == 0) xv[c][k] = x[q];
        else if constexpr (XM == 1) xv[c][k] = om * dinv[q] * b[q];
        else xv[c][k] = om * dinv[q] * b[q] + (T)kCoarseScale * xc[(size_t)ag[c] * D + k];
      }
    if constexpr (U > 1) __builtin_amdgcn_sched_barrier(0);
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
  *row_out = row;
  *live_out = live;
  *sl_out = sl;
}

template <int D, typename T>
__device__ __forceinline__ T pick(const T acc[D], int sl) {
  T v = acc[0];
#pragma unroll
  for (int q = 1; q < D; ++q)
    if (sl == q) v = acc[q];
  return v;
}

// r = b - A (om D^-1 b): the residual after one Jacobi sweep from zero
template <int D, typename T, int LPR, int U, typename TV = T>
__global__ void __launch_bounds__(kSpBS) amg_presmooth_resid_kernel(const int* __restrict__ cptr,
                                                                    const int* __restrict__ conns,
                                                                    const TV* __restrict__ val, int N,
                                                                    const T* __restrict__ b,
                                                                    const T* __restrict__ dinv,
                                                                    const double* __restrict__ omp,
                                                                    T* __restrict__ r, const Scalars* sc) {
  if (sc->done) return;
  const T om = (T)omp[0];
  int row, sl;
  bool live;
  T acc[D];
  row_gather<D, T, LPR, U, 1, TV>(blockIdx.x * (kSpBS / 64) + (threadIdx.x >> 6), cptr, conns, val, nullptr, dinv, b,
                                  om, N, &row, &live, &sl, acc);
  if (live && sl < D) {
    const size_t q = (size_t)row * D + sl;
    r[q] = b[q] - pick<D, T>(acc, sl);
  }
}

// r = b - A x1 with x1 = om D^-1 b written by the kernel that produced b (the
// smoothed hierarchy's pre-smoothing: one gathered vector per block, as an
// SpMV, where forming x1 at every gathered column reads two)
template <int D, typename T, int LPR, int U, typename TV = T>
__global__ void __launch_bounds__(kSpBS) amg_presmooth_x1_kernel(const int* __restrict__ cptr,
                                                                 const int* __restrict__ conns,
                                                                 const TV* __restrict__ val, int N,
                                                                 const T* __restrict__ b, const T* __restrict__ x1,
                                                                 T* __restrict__ r, const Scalars* sc) {
  if (sc->done) return;
  int row, sl;
  bool live;
  T acc[D];
  row_gather<D, T, LPR, U, 0, TV>(blockIdx.x * (kSpBS / 64) + (threadIdx.x >> 6), cptr, conns, val, x1, nullptr,
                                  nullptr, T(0), N, &row, &live, &sl, acc);
  if (live && sl < D) {
    const size_t q = (size_t)row * D + sl;
    r[q] = b[q] - pick<D, T>(acc, sl);
  }
}

// z = x + om D^-1 (b - A x) with x = om D^-1 b + P xc, the pre-smoothed
// iterate plus the prolonged coarse correction, formed where it is read (for
// the row itself and for every gathered column: no prolongation launch and
// no vector for it); with `mask` (finest level), z = 0 on the eliminated
// DOFs, so the CG never moves them however inexact the solve; with `part`,
// the workgroup's partial of (b, z)
template <int D, typename T, int LPR, int U, typename TV = T>
__global__ void __launch_bounds__(kSpBS) amg_postsmooth_kernel(const int* __restrict__ cptr,
                                                               const int* __restrict__ conns,
                                                               const TV* __restrict__ val, int N,
                                                               const T* __restrict__ b, const int* __restrict__ agg,
                                                               const T* __restrict__ xc,
                                                               const T* __restrict__ dinv,
                                                               const double* __restrict__ omp,
                                                               const T* __restrict__ mask,
                                                               T* __restrict__ z, double* __restrict__ part,
                                                               const Scalars* sc) {
  if (sc->done) return;
  const T om = (T)omp[0];
  int row, sl;
  bool live;
  T acc[D];
  row_gather<D, T, LPR, U, 2, TV>(blockIdx.x * (kSpBS / 64) + (threadIdx.x >> 6), cptr, conns, val, nullptr, dinv, b,
                                  om, N, &row, &live, &sl, acc, agg, xc);
  double dot = 0;
  if (live && sl < D) {
    const size_t q = (size_t)row * D + sl;
    const T xq = om * dinv[q] * b[q] + (T)kCoarseScale * xc[(size_t)agg[row] * D + sl];
    T zq = xq + om * dinv[q] * (b[q] - pick<D, T>(acc, sl));
    if (mask) zq *= mask[q];
    z[q] = zq;
    dot = (double)b[q] * (double)zq;
  }
  if (part) {
    Vec<1> a;
    a.v[0] = dot;
    write_partial_at<kSpBS, 1>(part, blockIdx.x, a);
  }
}

// b_c[a] = sum of r over the aggregate's members: one wavefront per
// aggregate, lanes over members, then a fixed butterfly (deterministic)
template <int D, typename T>
__global__ void __launch_bounds__(256) amg_restrict_kernel(const int* __restrict__ mptr, const int* __restrict__ midx,
                                                           int Nc, const T* __restrict__ r, T* __restrict__ bc,
                                                           const Scalars* sc) {
  if (sc->done) return;
  const int a = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (a >= Nc) return;  // whole wavefronts leave together
  double s[D];
#pragma unroll
  for (int k = 0; k < D; ++k) s[k] = 0;
  for (int q = mptr[a] + lane; q < mptr[a + 1]; q += 64) {
    const size_t i = (size_t)midx[q] * D;
#pragma unroll
    for (int k = 0; k < D; ++k) s[k] += (double)r[i + k];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int k = 0; k < D; ++k) s[k] += __shfl_xor(s[k], o, 64);
  if (lane < D) {
    double v = s[0];
#pragma unroll
    for (int k = 1; k < D; ++k)
      if (lane == k) v = s[k];
    bc[(size_t)a * D + lane] = (T)v;
  }
}

// x = Ainv b on the coarsest level (n <= kCoarsestMax)
template <typename T>
__global__ void amg_coarsest_kernel(const T* __restrict__ inv, int n, const T* __restrict__ b, T* __restrict__ x,
                                    double* __restrict__ part, const Scalars* sc) {
  if (sc->done) return;
  const int r = threadIdx.x;
  double s = 0;
  if (r < n) {
    for (int j = 0; j < n; ++j) s += (double)inv[(size_t)r * n + j] * (double)b[j];
    x[r] = (T)s;
  }
  if (part) {  // the whole system is the coarsest level: partial of (b, x)
    Vec<1> a;
    a.v[0] = r < n ? (double)b[r] * (double)(T)s : 0.0;
    write_partial_at<kBS, 1>(part, 0, a);
  }
}

// coarse values: block c = sum of its fine blocks (fixed order), in FP64
template <typename Tf, typename Tc, int DD>
__global__ void amg_gather_kernel(const int64_t* __restrict__ gptr, const int* __restrict__ gidx, int64_t nb,
                                  const Tf* __restrict__ fine, Tc* __restrict__ coarse) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nb * DD) return;
  const int64_t c = t / DD;
  const int k = (int)(t - c * DD);
  double s = 0;
  for (int64_t q = gptr[c]; q < gptr[c + 1]; ++q) s += (double)fine[(size_t)gidx[q] * DD + k];
  coarse[t] = (Tc)s;
}

template <typename T, int D>
__global__ void amg_dinv_kernel(const int* __restrict__ dpos, const T* __restrict__ val, int N, T* __restrict__ dinv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
#pragma unroll
  for (int r = 0; r < D; ++r) dinv[(size_t)D * i + r] = T(1) / val[(size_t)dpos[i] * D * D + r * D + r];
}

// Inverse of the coarsest matrix (n <= kCoarsestMax, SPD) by in-place
// Gauss-Jordan in LDS, one workgroup; the BSR level is expanded to dense first.
template <typename Tv, typename Ti, int D>
__global__ void __launch_bounds__(1024) amg_invert_kernel(const int* __restrict__ cptr,
                                                          const int* __restrict__ conns,
                                                          const Tv* __restrict__ val, int N,
                                                          Ti* __restrict__ inv) {
  __shared__ double A[kCoarsestMax * kCoarsestMax];
  __shared__ double fcol[kCoarsestMax];
  __shared__ int skip[kCoarsestMax];
  __shared__ double dmax;
  const int n = N * D;
  for (int t = threadIdx.x; t < n * n; t += blockDim.x) A[t] = 0;
  __syncthreads();
  for (int t = threadIdx.x; t < N * D; t += blockDim.x) {  // thread per (row node, component)
    const int i = t / D, ri = t - i * D;
    for (int e = cptr[i]; e < cptr[i + 1]; ++e) {
      const int j = conns[e];
      for (int cj = 0; cj < D; ++cj) A[(i * D + ri) * n + j * D + cj] = (double)val[(size_t)e * D * D + ri * D + cj];
    }
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = 0;
    for (int k = 0; k < n; ++k) m = fmax(m, fabs(A[k * n + k]));
    dmax = m;
  }
  for (int k = threadIdx.x; k < n; k += blockDim.x) skip[k] = 0;
  __syncthreads();
  for (int k = 0; k < n; ++k) {
    const double p = A[k * n + k];
    __syncthreads();
    // a vanishing pivot of the (semidefinite) Schur complement: a rigid motion
    // the constraints leave free (a specimen split by a crack); its row and
    // column are zero, and the inverse is taken on the range only
    if (!(p > 1e-10 * dmax)) {
      if (threadIdx.x == 0) skip[k] = 1;
      continue;
    }
    for (int j = threadIdx.x; j < n; j += blockDim.x) A[k * n + j] = (j == k ? 1.0 : A[k * n + j]) / p;
    for (int i = threadIdx.x; i < n; i += blockDim.x) fcol[i] = i == k ? 0.0 : A[i * n + k];
    __syncthreads();
    for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
      const int i = t / n, j = t - i * n;
      if (i == k) continue;
      A[t] = (j == k ? 0.0 : A[t]) - fcol[i] * A[k * n + j];
    }
    __syncthreads();
  }
  __syncthreads();
  for (int t = threadIdx.x; t < n * n; t += blockDim.x) {
    const int i = t / n, j = t - i * n;
    inv[t] = (skip[i] || skip[j]) ? Ti(0) : (Ti)A[t];
  }
}

// ---- smoothed aggregation (preconditioner option 4) ----
//
// The prolongation of a level is the tentative one (piecewise constant per
// aggregate and component) smoothed by one damped Jacobi step,
//   P = (I - om D^-1 A) T,  P(i,a) = [a = agg(i)] I - om D_i^-1 sum_{j in a} A(i,j),
// with om = 2/3 (4 / (3 rho) for rho(D^-1 A) taken as 2; the CT lattice's is
// 1.86), and the coarse operator is the Galerkin product P^T A P, formed as
// Y = A P, then P^T Y.  The patterns and the pair lists of both products are
// built on the host once per handle (sa_symbolic); every assembly recomputes
// the values on the device, each block a fixed-order FP64 sum (thread per
// block: deterministic).  In the V-cycle the restriction is P^T r (wave per
// aggregate over its P blocks) and the prolongation an explicit x = om D^-1 b +
// P xc before the post-smoothing sweep.  On CT stiffnesses this takes 2.2-2.7x
// fewer PCG iterations than the tentative prolongation (CPU experiment: 20k
// particles 24 -> 9, 80k 33 -> 15 to ||r||^2 <= 1e-4 ||r0||^2), at a level-1
// operator about 5x denser.  Each aggregate's P blocks are also stored in
// column order with their rows, so the restriction's loads are two deep, and
// the prolongation issues the loads of up to 8 P blocks of a row at once:
// measured on the GPU, an FP32 iteration at 80k particles costs 143.5 us
// against 102.9 for the tentative hierarchy, the iterations drop 33 -> 17, and
// the CT stream's 205,645- and 109,174-particle requests (30 load steps, FP32
// inner solves) take 15.0 and 6.1 s against 20.6 and 8.6 s.

// ---- smoother damping from the level's largest eigenvalue (smoothed aggregation) ----
//
// One damped-Jacobi sweep damps the error components of D^-1 A's eigenvalues
// lambda by |1 - om lambda|: it converges (and the V-cycle stays a positive
// definite preconditioner) only for om lambda_max < 2, and a larger om damps
// the oscillatory components more.  lambda_max is 1.86 on the intact
// hexagonal lattice, 2.0-2.4 on the square, SC, BCC and FCC lattices and grows
// as bonds break (2.31 on a cracked 6k-particle CT specimen); the coarse levels
// of the smoothed hierarchy measured 1.47-1.67.  After every value rebuild,
// kLzSteps Lanczos steps on D^-1 A (a fixed start vector, deterministic
// reductions) estimate lambda_max from below, within 2.7% on those matrices,
// and the level's damping becomes min(kOmegaMax, kOmegaTarget / lambda_est),
// so om lambda_max stays below 1.7.  CPU experiment (CT specimens, PCG
// iterations to ||r||^2 <= 1e-4 ||r0||^2): 19 -> 14 at 80k particles, 22 -> 16
// at 150k, against the fixed 2/3.
constexpr int kLzSteps = 12;
constexpr double kOmegaMax = 0.9, kOmegaTarget = 1.65;
// The prolongator's smoothing P = (I - om_p D^-1 A) T takes om_p = kPTarget /
// lambda_est (at most 1): 1.5 / 1.86 = 0.81 on the intact CT lattice instead of
// 2/3 (CPU experiment, PCG iterations to ||r||^2 <= 1e-4 ||r0||^2 with the
// adaptive smoother: 9 -> 7 at 20k particles, 15 -> 12 at 80k, 10 -> 10 on a
// cracked 6k specimen; 4/3 took 8, 13 and 10).  Any om_p gives a valid
// prolongator: the V-cycle stays a positive definite preconditioner.
constexpr double kPTarget = 1.5;

__device__ __forceinline__ double lz_start(int64_t k) {  // start vector entries in [-0.5, 0.5)
  const uint32_t h = (uint32_t)((uint64_t)k * 2654435761ull) >> 16;
  return (double)(h & 1023u) / 1024.0 - 0.5;
}

template <int D>
__device__ __forceinline__ double lz_diag(const int* __restrict__ dpos, const double* __restrict__ val, int64_t i) {
  const int64_t row = i / D;
  const int c = (int)(i - row * D);
  return val[(size_t)dpos[row] * D * D + c * (D + 1)];
}

// The iterates are kept as u = D^-1/2 v (v the Lanczos vectors of the
// symmetric D^-1/2 A D^-1/2), so a step needs a plain A u: t = D^-1/2 v0
// (unnormalized; partials of ||v0||^2)
template <int D>
__global__ void lz_init_kernel(const int* __restrict__ dpos, const double* __restrict__ val, int64_t n,
                               double* __restrict__ t, double* __restrict__ part) {
  double s = 0;
  for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS) {
    const double h = lz_start(i);
    t[i] = h / sqrt(lz_diag<D>(dpos, val, i));
    s += h * h;
  }
  Vec<1> a;
  a.v[0] = s;
  write_partial<kBS, 1>(part, a);
}

// y = A t with partials of (t, y)
template <int D, int LPR, int U>
__global__ void __launch_bounds__(kSpBS) lz_spmv_kernel(const int* __restrict__ cptr, const int* __restrict__ conns,
                                                        const double* __restrict__ val, const double* __restrict__ t,
                                                        double* __restrict__ y, int N, double* __restrict__ part) {
  Vec<1> a;
  a.v[0] = bsr_rows<D, double, LPR, U, true>(blockIdx.x, cptr, conns, val, t, y, N);
  write_partial_at<kSpBS, 1>(part, blockIdx.x, a);
}

// Step k: beta_k = ||t||_D (the D-norm of the unnormalized iterate, from the
// partials pb), alpha_k = (t, A t) / beta_k^2 (partials pa); u_k = t / beta_k;
// t <- D^-1 A u_k - alpha_k u_k - beta_k u_{k-1} with partials of ||t||_D^2
// into pb_next.  Workgroup 0 records alpha_k and beta_k.  A vanishing or
// non-finite beta (the Krylov space exhausted) ends the recurrence: u and t
// become zero.
template <int D>
__global__ void lz_update_kernel(const int* __restrict__ dpos, const double* __restrict__ val, int64_t n, int k,
                                 const double* __restrict__ pb, int nb, const double* __restrict__ pa, int na,
                                 const double* __restrict__ y, double* __restrict__ t,
                                 const double* __restrict__ uprev, double* __restrict__ ucur,
                                 double* __restrict__ pb_next, double* __restrict__ coef) {
  const double b2 = sum_partials<kBS, 1>(pb, nb).v[0];
  const double ty = sum_partials<kBS, 1>(pa, na).v[0];
  const bool ok = b2 > 0 && isfinite(b2) && isfinite(ty);
  const double beta = ok ? sqrt(b2) : 0.0, ib = ok ? 1.0 / beta : 0.0;
  const double alpha = ok ? ty / b2 : 0.0;
  double s = 0;
  for (int64_t i = (int64_t)blockIdx.x * kBS + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBS) {
    const double d = lz_diag<D>(dpos, val, i);
    const double u = t[i] * ib;
    const double up = k > 0 ? uprev[i] : 0.0;
    const double tn = ok ? y[i] * ib / d - alpha * u - beta * up : 0.0;
    ucur[i] = u;
    t[i] = tn;
    s += d * tn * tn;
  }
  Vec<1> a;
  a.v[0] = s;
  write_partial<kBS, 1>(pb_next, a);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    coef[k] = alpha;
    coef[kLzSteps + k] = beta;
  }
}

// The largest eigenvalue of the Lanczos tridiagonal (alpha_0..alpha_{m-1},
// off-diagonals beta_1..beta_{m-1}; m stops at the first vanishing beta) by
// bisection on its Sturm count, and the damping it gives (2/3 if the estimate
// is not usable).  One thread.
__global__ void lz_omega_kernel(const double* __restrict__ coef, double* __restrict__ omega,
                                double* __restrict__ omega_p) {
  if (threadIdx.x != 0) return;
  const double* al = coef;
  const double* be = coef + kLzSteps;
  int m = 1;
  double scale = fabs(al[0]);
  while (m < kLzSteps && be[m] > 1e-12 * (scale + 1e-300) && isfinite(be[m])) {
    scale = fmax(scale, fabs(al[m]));
    ++m;
  }
  double lo = al[0], hi = al[0];
  for (int i = 0; i < m; ++i) {
    const double r = (i > 0 ? fabs(be[i]) : 0.0) + (i + 1 < m ? fabs(be[i + 1]) : 0.0);
    lo = fmin(lo, al[i] - r);
    hi = fmax(hi, al[i] + r);
  }
  for (int it = 0; it < 200 && hi - lo > 1e-12 * fmax(fabs(hi), 1e-300); ++it) {
    const double x = 0.5 * (lo + hi);
    int below = 0;  // eigenvalues below x
    double q = 1.0;
    for (int i = 0; i < m; ++i) {
      q = al[i] - x - (i > 0 ? be[i] * be[i] / q : 0.0);
      if (q == 0.0) q = -1e-300;
      below += q < 0.0;
    }
    if (below == m) hi = x;
    else lo = x;
  }
  const double lam = hi;
  const bool ok = lam > 0 && isfinite(lam);
  *omega = ok ? fmin(kOmegaMax, kOmegaTarget / lam) : kOmega;
  *omega_p = ok ? fmin(1.0, kPTarget / lam) : kOmega;
}

// P's blocks: the sum of the level's blocks each one covers (pair list atidx,
// contiguous per P block), one wavefront per G consecutive P blocks.  The
// wavefront walks their lists 64 entries at a time -- lane t loads entry
// base+t's block into LDS (coalesced index loads, 64 gathers in flight) --
// and each P block's lane adds its own entries in list order; the block then
// becomes [aggregate of the row] I - om D_i^-1 (sum), written in row order (Pv)
// and column order (PvT).
template <int D>
__global__ void __launch_bounds__(64) sa_p_kernel(const int* __restrict__ atptr, const int* __restrict__ atidx,
                                                  const double* __restrict__ A, const int* __restrict__ prow,
                                                  const int* __restrict__ pcol, const int* __restrict__ agg,
                                                  const int* __restrict__ dpos, const int* __restrict__ qpos,
                                                  int nP, int G, const double* __restrict__ omp,
                                                  double* __restrict__ Pv,
                                                  double* __restrict__ PvT) {
  constexpr int DD = D * D;
  __shared__ double term[DD * 64];
  const int lane = threadIdx.x;
  const int o0 = blockIdx.x * G;
  const int o = o0 + lane;
  const bool mine = lane < G && o < nP;
  const int k0 = atptr[o0], k1 = atptr[min(o0 + G, nP)];
  const int lo = mine ? atptr[o] : 0, hi = mine ? atptr[o + 1] : 0;
  double s[DD];
#pragma unroll
  for (int t = 0; t < DD; ++t) s[t] = 0;
  for (int base = k0; base < k1; base += 64) {
    const int k = base + lane;
    if (k < k1) {
      const double* a = A + (size_t)atidx[k] * DD;
#pragma unroll
      for (int t = 0; t < DD; ++t) term[t * 64 + lane] = a[t];
    }
    __syncthreads();
    const int a0 = max(lo, base), a1 = min(hi, base + 64);
    for (int j = a0; j < a1; ++j)
#pragma unroll
      for (int t = 0; t < DD; ++t) s[t] += term[t * 64 + (j - base)];
    __syncthreads();
  }
  if (!mine) return;
  const double om = omp[0];
  const int i = prow[o];
  const size_t dg = (size_t)dpos[i] * DD;
  const bool own = pcol[o] == agg[i];
  const size_t pt = (size_t)qpos[o] * DD;
#pragma unroll
  for (int r = 0; r < D; ++r) {
    const double w = om / A[dg + r * D + r];
#pragma unroll
    for (int q = 0; q < D; ++q) {
      const double v = ((own && r == q) ? 1.0 : 0.0) - w * s[r * D + q];
      Pv[(size_t)o * DD + r * D + q] = v;
      PvT[pt + r * D + q] = v;
    }
  }
}

// P blocks per wavefront of sa_p_kernel: about 10 passes of 64 list entries
// on average, at most 64 blocks.
inline int sa_outputs_per_wave(int64_t pairs, int64_t outputs) {
  const double avg = outputs > 0 ? (double)pairs / (double)outputs : 1.0;
  return (int)std::max(1.0, std::min(64.0, std::floor(640.0 / std::max(avg, 1.0))));
}

// Row products by scattering into LDS: LPR lanes per output row, whose
// blocks accumulate in LDS (`width` blocks per row).  The row's sources are
// taken in order; each source contributes to a contiguous run of blocks (a P
// row for Y = A P, a Y row for P^T Y), one per lane, each lane adding its term
// to the slot the host listed for it (a source's terms go to distinct slots,
// and a wavefront's LDS operations complete in order).  Every output block is
// therefore the sum of its terms in source order -- the sequence a thread
// summing its own pair list adds -- and each source block is read once per
// row, its run contiguously.  Loads are batched so that no source waits for
// the previous one's: the lanes load LPR sources' indices, slot offsets and
// run bounds at once (broadcast by shuffles), and a batch of SB sources issues
// all its block and slot loads before its LDS updates.  (Taking the sources
// one at a time, each behind its three dependent loads, measured 462 + 340 us
// for Y and P^T Y on the finest level of the 205,645-particle CT specimen.)
//   MODE 1  Y(i,:) = sum over A's blocks e of row i of A(e) P(conns[e], :)
//   MODE 2  C(a,:) = sum over P's blocks k of column a of P(k)^T Y(qrow[k], :)
// ONE: every run fits the row's lanes (the longest run <= LPR).
template <int D, int MODE, int LPR, bool ONE>
__global__ void sa_rows_kernel(int nrows, const int* __restrict__ sptr, const int* __restrict__ sidx,
                               const double* __restrict__ S, const int* __restrict__ rptr,
                               const double* __restrict__ R, const int64_t* __restrict__ slot0,
                               const int* __restrict__ soff, const uint16_t* __restrict__ slot,
                               const int* __restrict__ optr, int width, double* __restrict__ out) {
  constexpr int DD = D * D;
  constexpr int SB = D == 2 ? 4 : 2;  // sources per batch
  extern __shared__ double acc[];
  const int g = threadIdx.x / LPR, t = threadIdx.x - g * LPR;
  const int row = blockIdx.x * (blockDim.x / LPR) + g;
  if (row >= nrows) return;  // whole groups leave together; no barriers below
  double* my = acc + (size_t)g * width * DD;
  const int o0 = optr[row], no = optr[row + 1] - o0;
  for (int u = t; u < no * DD; u += LPR) my[u] = 0;
  const uint16_t* rs = slot + slot0[row];
  const int s0 = sptr[row], s1 = sptr[row + 1];
  for (int kb = s0; kb < s1; kb += LPR) {
    // lane t: source kb + t's run start, length and first slot
    const int nk = min(LPR, s1 - kb);
    int r0l = 0, ml = 0, ol = 0;
    if (t < nk) {
      const int j = sidx[kb + t];
      ol = soff[kb + t];
      r0l = rptr[j];
      ml = rptr[j + 1] - r0l;
    }
    for (int b0 = 0; b0 < nk; b0 += SB) {
      double sv[SB][DD], rv[SB][DD];
      int sl[SB], m[SB];
#pragma unroll
      for (int b = 0; b < SB; ++b) {
        // unconditional loads from addresses that are always valid (the
        // batch's first source, the run's first block): a lane without a term
        // never uses what it loaded, and no load waits behind a branch
        const int src = b0 + b;
        const bool ok = src < nk;
        const int r0 = __shfl(r0l, src & (LPR - 1), LPR);
        const int off = __shfl(ol, src & (LPR - 1), LPR);
        m[b] = ok ? __shfl(ml, src & (LPR - 1), LPR) : 0;
        const bool act = t < m[b];
        const double* sp = S + (size_t)(kb + (ok ? src : b0)) * DD;
        const double* rp = R + (size_t)(act ? r0 + t : 0) * DD;
#pragma unroll
        for (int q = 0; q < DD; ++q) sv[b][q] = sp[q];
        sl[b] = rs[act ? off + t : 0];
#pragma unroll
        for (int q = 0; q < DD; ++q) rv[b][q] = rp[q];
      }
#pragma unroll
      for (int b = 0; b < SB; ++b) {
        // the head of source b's run (preloaded), then, when it is longer
        // than the row's lanes, its tail
        const int r0 = ONE ? 0 : __shfl(r0l, (b0 + b) & (LPR - 1), LPR);
        const int off = ONE ? 0 : __shfl(ol, (b0 + b) & (LPR - 1), LPR);
        for (int c = t; c < m[b]; c += LPR) {
          int sc;
          double rc[DD];
          if (ONE || c == t) {
            sc = sl[b];
#pragma unroll
            for (int q = 0; q < DD; ++q) rc[q] = rv[b][q];
          } else {
            sc = rs[off + c];
#pragma unroll
            for (int q = 0; q < DD; ++q) rc[q] = R[(size_t)(r0 + c) * DD + q];
          }
#pragma unroll
          for (int r = 0; r < D; ++r)
#pragma unroll
            for (int q = 0; q < D; ++q) {
              // MODE 1: sum_m s(r,m) R(m,q); MODE 2: sum_m s(m,r) R(m,q)
              double v = (MODE == 1 ? sv[b][r * D] : sv[b][r]) * rc[q];
#pragma unroll
              for (int mm = 1; mm < D; ++mm)
                v += (MODE == 1 ? sv[b][r * D + mm] : sv[b][mm * D + r]) * rc[mm * D + q];
              my[sc * DD + r * D + q] += v;
            }
          if (ONE) break;
        }
      }
    }
  }
  for (int u = t; u < no * DD; u += LPR) out[(size_t)o0 * DD + u] = my[u];
}

// ---- smoothed aggregation, symbolic lists on the device ----
//
// The slot lists of Y = A P and of P^T Y (and their row patterns) as the host
// builds them -- each output row's distinct columns ascending, and for every
// term, in source then run order, its column's position -- one wavefront (a
// 64-thread workgroup) per output row: the row's sources' run offsets by a
// wavefront scan, its columns deduplicated in an LDS hash table (bounded
// probing), sorted by a bitonic sort in LDS, each term's slot by a binary
// search.  MODE 1: rows of Y (sources: A's blocks of row i, runs: the P rows
// of their columns, pcol); MODE 2: coarse rows (sources: P's blocks of column
// a in column order, runs: the Y rows of their rows, ycol).  A row with more
// sources or distinct columns than the LDS arrays hold sets *overflow, and the
// level's lists are then built on the host.  On the 1.0M-particle BCC
// polycrystal the host took 0.39 s for these lists.
constexpr int kSymMaxSrc = 1024;            // sources per output row
constexpr int kSymHash = 2048;              // hash slots per output row
constexpr int kSymMaxU = kSymHash / 2;      // distinct columns per output row

__device__ __forceinline__ int sym_source(const int* off, int ns, int t) {  // k with off[k] <= t < off[k+1]
  int lo = 0, hi = ns - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= t) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// FILL 0: ucnt[row] = distinct columns, ecnt[row] = terms, soff[s0 + k] =
// source k's first term within the row.  FILL 1: ocol[optr[row] ...] = the
// distinct columns ascending, oslot[eptr[row] + t] = term t's position, and
// (MODE 2) odiag[row] = the position of column `row` (-1 if absent).
template <int FILL>
__global__ void __launch_bounds__(64) sa_sym_kernel(int nrows, const int* __restrict__ sptr,
                                                    const int* __restrict__ sidx, const int* __restrict__ rptr,
                                                    const int* __restrict__ rcol, int* __restrict__ ucnt,
                                                    int64_t* __restrict__ ecnt, int* __restrict__ soff,
                                                    const int* __restrict__ optr, const int64_t* __restrict__ eptr,
                                                    int* __restrict__ ocol, uint16_t* __restrict__ oslot,
                                                    int* __restrict__ odiag, int* __restrict__ overflow) {
  __shared__ int off[kSymMaxSrc + 1];
  __shared__ int tab[kSymHash];
  __shared__ int U[kSymMaxU];
  __shared__ int nu, bad;
  const int row = blockIdx.x, lane = threadIdx.x;
  if (row >= nrows) return;  // the whole workgroup (one wavefront)
  const int s0 = sptr[row], ns = sptr[row + 1] - s0;
  if (ns > kSymMaxSrc) {
    if (lane == 0) atomicOr(overflow, 1);
    return;
  }
  // runs' exclusive offsets: off[k] = terms of sources before k
  int carry = 0;
  for (int b = 0; b < ns; b += 64) {
    const int k = b + lane;
    int m = 0;
    if (k < ns) {
      const int j = sidx[s0 + k];
      m = rptr[j + 1] - rptr[j];
    }
    const int incl = wave_incl_scan(m) + carry;
    if (k < ns) off[k + 1] = incl;
    carry = __shfl(incl, 63, 64);
  }
  if (lane == 0) {
    off[0] = 0;
    nu = 0;
    bad = 0;
  }
  for (int h = lane; h < kSymHash; h += 64) tab[h] = -1;
  __syncthreads();
  const int E = ns > 0 ? off[ns] : 0;
  if (FILL == 0)
    for (int k = lane; k < ns; k += 64) soff[s0 + k] = off[k];
  // distinct columns: insert every term's column (at most kSymHash probes)
  for (int t = lane; t < E; t += 64) {
    const int k = sym_source(off, ns, t);
    const int col = rcol[rptr[sidx[s0 + k]] + (t - off[k])];
    int h = (int)(((uint32_t)col * 2654435761u) >> 21) & (kSymHash - 1);
    bool placed = false;
    for (int probe = 0; probe < kSymHash; ++probe) {
      const int prev = atomicCAS(&tab[h], -1, col);
      if (prev == -1) {
        atomicAdd(&nu, 1);
        placed = true;
        break;
      }
      if (prev == col) {
        placed = true;
        break;
      }
      h = (h + 1) & (kSymHash - 1);
    }
    if (!placed) atomicOr(&bad, 1);
  }
  __syncthreads();
  const int u = nu;
  if (bad || u > kSymMaxU) {
    if (lane == 0) atomicOr(overflow, 1);
    return;
  }
  if (FILL == 0) {
    if (lane == 0) {
      ucnt[row] = u;
      ecnt[row] = E;
    }
    return;
  }
  // the distinct columns, sorted (bitonic, padded to a power of two)
  int P = 1;
  while (P < u) P <<= 1;
  __syncthreads();
  if (lane == 0) nu = 0;
  __syncthreads();
  for (int h = lane; h < kSymHash; h += 64) {
    const int v = tab[h];
    if (v >= 0) U[atomicAdd(&nu, 1)] = v;
  }
  __syncthreads();
  for (int i = u + lane; i < P; i += 64) U[i] = 0x7fffffff;
  __syncthreads();
  for (int kk = 2; kk <= P; kk <<= 1)
    for (int jj = kk >> 1; jj > 0; jj >>= 1) {
      for (int i = lane; i < P; i += 64) {
        const int l = i ^ jj;
        if (l > i) {
          const int x = U[i], y = U[l];
          const bool up = (i & kk) == 0;
          if ((x > y) == up) {
            U[i] = y;
            U[l] = x;
          }
        }
      }
      __syncthreads();
    }
  int* oc = ocol + optr[row];
  for (int k = lane; k < u; k += 64) oc[k] = U[k];
  if (odiag && lane == 0) {
    int d = -1, lo = 0, hi = u - 1;
    while (lo <= hi) {
      const int mid = (lo + hi) >> 1;
      if (U[mid] == row) {
        d = mid;
        break;
      }
      if (U[mid] < row) lo = mid + 1;
      else hi = mid - 1;
    }
    odiag[row] = d;
  }
  uint16_t* os = oslot + eptr[row];
  for (int t = lane; t < E; t += 64) {
    const int k = sym_source(off, ns, t);
    const int col = rcol[rptr[sidx[s0 + k]] + (t - off[k])];
    int lo = 0, hi = u - 1;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (U[mid] < col) lo = mid + 1;
      else hi = mid;
    }
    os[t] = (uint16_t)lo;
  }
}

// Launch sa_rows_kernel: LPR from the longest run (16, 32 or 64 lanes), rows
// per workgroup so that their accumulators take at most 32 KB of LDS.
template <int D, int MODE>
void launch_sa_rows(hipStream_t st, int nrows, int longest_run, int width, const int* sptr, const int* sidx,
                    const double* S, const int* rptr, const double* R, const int64_t* slot0, const int* soff,
                    const uint16_t* slot, const int* optr, double* out) {
  const size_t row_bytes = (size_t)std::max(width, 1) * D * D * sizeof(double);
  auto go = [&](auto LP) {
    constexpr int LPR = decltype(LP)::value;
    const int rpb = (int)std::max<size_t>(1, std::min<size_t>(256 / LPR, 32768 / row_bytes));
    const unsigned grid = (unsigned)((nrows + rpb - 1) / rpb);
    if (longest_run <= LPR)
      sa_rows_kernel<D, MODE, LPR, true><<<grid, rpb * LPR, rpb * row_bytes, st>>>(nrows, sptr, sidx, S, rptr, R,
                                                                                  slot0, soff, slot, optr, width, out);
    else
      sa_rows_kernel<D, MODE, LPR, false><<<grid, rpb * LPR, rpb * row_bytes, st>>>(nrows, sptr, sidx, S, rptr, R,
                                                                                   slot0, soff, slot, optr, width, out);
  };
  if (longest_run <= 16) go(std::integral_constant<int, 16>());
  else if (longest_run <= 32) go(std::integral_constant<int, 32>());
  else go(std::integral_constant<int, 64>());
}

// bc(a) = sum over the aggregate's P blocks of P(i,a)^T r(i): one wavefront per
// aggregate, lanes over the blocks, a fixed butterfly (deterministic)
// The aggregate's P blocks and their rows are stored contiguously in column
// order (PvT, qrow), so a lane's loads are two deep: its row and block, then
// r of that row.
template <int D, typename T>
__global__ void __launch_bounds__(256) sa_restrict_kernel(const int* __restrict__ qptr, const int* __restrict__ qrow,
                                                          const T* __restrict__ PvT, int Nc, const T* __restrict__ r,
                                                          T* __restrict__ bc, const T* __restrict__ dinvc,
                                                          const double* __restrict__ omc, T* __restrict__ x1c,
                                                          const Scalars* sc) {
  if (sc->done) return;
  constexpr int DD = D * D;
  const int a = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (a >= Nc) return;  // whole wavefronts leave together
  double s[D];
#pragma unroll
  for (int k = 0; k < D; ++k) s[k] = 0;
  for (int k = qptr[a] + lane; k < qptr[a + 1]; k += 64) {
    const size_t i = (size_t)qrow[k] * D;
#pragma unroll
    for (int q = 0; q < D; ++q)
#pragma unroll
      for (int m = 0; m < D; ++m) s[q] += (double)PvT[(size_t)k * DD + m * D + q] * (double)r[i + m];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int k = 0; k < D; ++k) s[k] += __shfl_xor(s[k], o, 64);
  if (lane < D) {
    double v = s[0];
#pragma unroll
    for (int k = 1; k < D; ++k)
      if (lane == k) v = s[k];
    const size_t q = (size_t)a * D + lane;
    const T bq = (T)v;
    bc[q] = bq;
    // the coarse level's pre-smoothed iterate om D^-1 bc (when it has a smoother)
    if (x1c) x1c[q] = (T)omc[0] * dinvc[q] * bq;
  }
}

// x = x1 + P xc, x1 = om D^-1 b the pre-smoothed iterate (thread per row; the
// row's P blocks in order)
template <int D, typename T>
__global__ void __launch_bounds__(256) sa_prolong_kernel(const int* __restrict__ pptr, const int* __restrict__ pcol,
                                                         const T* __restrict__ Pv, int N, const T* __restrict__ x1,
                                                         const T* __restrict__ xc, T* __restrict__ x,
                                                         const Scalars* sc) {
  if (sc->done) return;
  constexpr int DD = D * D;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  T acc[D];
#pragma unroll
  for (int q = 0; q < D; ++q) acc[q] = T(0);
  // the row's blocks eight at a time: columns and blocks, then the gathers,
  // each group's loads in flight together (indices past the row's end repeat
  // its last block and are not added)
  constexpr int G = 8;
  const int p0 = pptr[i], p1 = pptr[i + 1];
  for (int pb = p0; pb < p1; pb += G) {
    int pc[G];
    T pv[G][DD], xv[G][D];
#pragma unroll
    for (int t = 0; t < G; ++t) {
      const int p = min(pb + t, p1 - 1);
      pc[t] = pcol[p];
#pragma unroll
      for (int k = 0; k < DD; ++k) pv[t][k] = Pv[(size_t)p * DD + k];
    }
#pragma unroll
    for (int t = 0; t < G; ++t)
#pragma unroll
      for (int m = 0; m < D; ++m) xv[t][m] = xc[(size_t)pc[t] * D + m];
#pragma unroll
    for (int t = 0; t < G; ++t)
#pragma unroll
      for (int q = 0; q < D; ++q) {
        T v = pv[t][q * D] * xv[t][0];
#pragma unroll
        for (int m = 1; m < D; ++m) v += pv[t][q * D + m] * xv[t][m];
        acc[q] += pb + t < p1 ? v : T(0);
      }
  }
#pragma unroll
  for (int q = 0; q < D; ++q) {
    const size_t iq = (size_t)i * D + q;
    x[iq] = x1[iq] + acc[q];
  }
}

// z = x + om D^-1 (b - A x) with x read from memory (the prolongation's
// output); with `mask` (finest level) z = 0 on the eliminated DOFs; with
// `part`, the workgroup's partial of (b, z)
template <int D, typename T, int LPR, int U, typename TV = T>
__global__ void __launch_bounds__(kSpBS) amg_postsmooth_x_kernel(const int* __restrict__ cptr,
                                                                 const int* __restrict__ conns,
                                                                 const TV* __restrict__ val, int N,
                                                                 const T* __restrict__ b, const T* __restrict__ x,
                                                                 const T* __restrict__ dinv,
                                                                 const double* __restrict__ omp,
                                                                 const T* __restrict__ mask, T* __restrict__ z,
                                                                 double* __restrict__ part, const Scalars* sc) {
  if (sc->done) return;
  const T om = (T)omp[0];
  int row, sl;
  bool live;
  T acc[D];
  row_gather<D, T, LPR, U, 0, TV>(blockIdx.x * (kSpBS / 64) + (threadIdx.x >> 6), cptr, conns, val, x, dinv, b,
                                  om, N, &row, &live, &sl, acc);
  double dot = 0;
  if (live && sl < D) {
    const size_t q = (size_t)row * D + sl;
    T zq = x[q] + om * dinv[q] * (b[q] - pick<D, T>(acc, sl));
    if (mask) zq *= mask[q];
    z[q] = zq;
    dot = (double)b[q] * (double)zq;
  }
  if (part) {
    Vec<1> a;
    a.v[0] = dot;
    write_partial_at<kSpBS, 1>(part, blockIdx.x, a);
  }
}

// ---- PCG kernels (one system, its own stream) ----

template <int D, typename T, int LPR, int U>
__global__ void __launch_bounds__(kSpBS) pcg_spmv_kernel(const int* __restrict__ cptr, const int* __restrict__ conns,
                                                         const T* __restrict__ val, const T* __restrict__ p,
                                                         T* __restrict__ w, int N, double* __restrict__ part,
                                                         Scalars* sc) {
  if (sc->done) return;
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->icur = sc->inext;  // read by the next kernels
  Vec<1> a;
  a.v[0] = bsr_rows<D, T, LPR, U, true>(blockIdx.x, cptr, conns, val, p, w, N);
  write_partial<kSpBS, 1>(part, a);
}

__global__ void __launch_bounds__(kBS) pcg_reduce_kernel(const double* __restrict__ in, int n,
                                                         double* __restrict__ out, const Scalars* __restrict__ sc) {
  if (sc->done) return;
  const int per = (n + gridDim.x - 1) / gridDim.x;
  const int lo = blockIdx.x * per, hi = min(n, lo + per);
  Vec<1> a;
  a.v[0] = 0;
  for (int b = lo + threadIdx.x; b < hi; b += kBS) a.v[0] += in[b];
  write_partial<kBS, 1>(out, a);
}

// x += alpha p; r -= alpha w; partials of (r,r).  alpha = gamma_i / (p,w).
// With x1 (the smoothed hierarchy): x1 = om D^-1 r, the finest level's
// pre-smoothed iterate.
template <typename T>
__global__ void __launch_bounds__(kBS) pcg_vec1_kernel(T* __restrict__ x, T* __restrict__ r,
                                                       const T* __restrict__ p, const T* __restrict__ w, int64_t n,
                                                       const double* __restrict__ pw_part, int npw,
                                                       double* __restrict__ rr_part, const T* __restrict__ dinv,
                                                       const double* __restrict__ omp, T* __restrict__ x1,
                                                       Scalars* sc) {
  if (sc->done) return;
  const int i = sc->icur;
  const double pw = sum_partials<kBS, 1>(pw_part, npw).v[0];
  // (p, A p) <= 0: breakdown (a singular stiffness); no update, and the
  // vector step that follows stops the solve
  const double alpha = pw > 0 ? sc->gam[i & 1] / pw : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->alp[i & 1] = alpha;
    sc->pq = pw;
    sc->alpha = alpha;
  }
  const T a = (T)alpha;
  const T om = x1 ? (T)omp[0] : T(0);
  double rr = 0;
  for (int64_t k = (int64_t)blockIdx.x * kBS + threadIdx.x; k < n; k += (int64_t)gridDim.x * kBS) {
    x[k] += a * p[k];
    const T rk = r[k] - a * w[k];
    r[k] = rk;
    if (x1) x1[k] = om * dinv[k] * rk;
    rr += (double)rk * (double)rk;
  }
  Vec<1> v;
  v.v[0] = rr;
  write_partial<kBS, 1>(rr_part, v);
}

// stop test on (r,r); beta = (r,z)_{i+1} / (r,z)_i; p = z + beta p.  Workgroup 0
// sets `done`, so each workgroup reads it once through LDS.
template <typename T>
__global__ void __launch_bounds__(kBS) pcg_vec2_kernel(T* __restrict__ p, const T* __restrict__ z, int64_t n,
                                                       const double* __restrict__ rz_part, int nrz,
                                                       const double* __restrict__ rr_part, int nrr, Scalars* sc) {
  __shared__ int done_s;
  if (threadIdx.x == 0) done_s = sc->done;
  __syncthreads();
  if (done_s) return;
  const int i = sc->icur;
  const double rz = sum_partials<kBS, 1>(rz_part, nrz).v[0];
  const double rr = sum_partials<kBS, 1>(rr_part, nrr).v[0];
  const bool stop = rr <= sc->thresh || i + 1 >= sc->maxit || !(sc->pq > 0) || !(rz > 0);
  const double beta = rz / sc->gam[i & 1];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->rr = rr;
    sc->rz = rz;
    sc->iter = i + 1;
    sc->inext = i + 1;
    sc->gam[(i + 1) & 1] = rz;
    sc->beta = beta;
    if (stop) sc->done = 1;
  }
  if (stop) return;
  const T b = (T)beta;
  for (int64_t k = (int64_t)blockIdx.x * kBS + threadIdx.x; k < n; k += (int64_t)gridDim.x * kBS) p[k] = z[k] + b * p[k];
}

// start: x = 0, r = b (rhs); partials of (r,r); with x1, x1 = om D^-1 r
template <typename T>
__global__ void __launch_bounds__(kBS) pcg_init_kernel(const double* __restrict__ rhs, T* __restrict__ x,
                                                       T* __restrict__ r, int64_t n, double* __restrict__ rr_part,
                                                       const T* __restrict__ dinv, const double* __restrict__ omp,
                                                       T* __restrict__ x1) {
  const T om = x1 ? (T)omp[0] : T(0);
  double rr = 0;
  for (int64_t k = (int64_t)blockIdx.x * kBS + threadIdx.x; k < n; k += (int64_t)gridDim.x * kBS) {
    const T rk = (T)rhs[k];
    x[k] = 0;
    r[k] = rk;
    if (x1) x1[k] = om * dinv[k] * rk;
    rr += (double)rk * (double)rk;
  }
  Vec<1> v;
  v.v[0] = rr;
  write_partial<kBS, 1>(rr_part, v);
}

// after z = M r: gamma_0 = (r,z), the stop threshold, p = z
template <typename T>
__global__ void __launch_bounds__(kBS) pcg_init2_kernel(T* __restrict__ p, const T* __restrict__ z, int64_t n,
                                                        const double* __restrict__ rz_part, int nrz,
                                                        const double* __restrict__ rr_part, int nrr, Scalars* sc) {
  const double rz = sum_partials<kBS, 1>(rz_part, nrz).v[0];
  const double rr = sum_partials<kBS, 1>(rr_part, nrr).v[0];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->gam[0] = rz;
    sc->rz = rz;
    sc->rr = rr;
    sc->rr0 = rr;
    sc->thresh = sc->rtol * rr + sc->atol;
    sc->iter = 0;
    sc->icur = sc->inext = 0;
    if (rr <= sc->thresh || sc->maxit <= 0) sc->done = 1;
  }
  for (int64_t k = (int64_t)blockIdx.x * kBS + threadIdx.x; k < n; k += (int64_t)gridDim.x * kBS) p[k] = z[k];
}

__global__ void amg_setup_scalars_kernel(Scalars* sc, double rtol, double atol, int64_t maxit) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    sc->rtol = rtol;
    sc->atol = atol;
    sc->maxit = (int)maxit;
    sc->iter = 0;
    sc->icur = sc->inext = 0;
    sc->done = 0;
  }
}

// finest level: point-Jacobi diagonal and the free-DOF mask (1 free, 0 eliminated)
template <int D, typename T>
__global__ void amg_dinv0_kernel(const int* __restrict__ dpos, const T* __restrict__ val, const int* __restrict__ cons,
                                 const double* __restrict__ dvis, const int* __restrict__ floating, int N,
                                 T* __restrict__ dinv, T* __restrict__ mask) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
#pragma unroll
  for (int r = 0; r < D; ++r) {
    dinv[(size_t)D * i + r] = T(1) / val[(size_t)dpos[i] * D * D + r * D + r];
    mask[(size_t)D * i + r] = dof_fixed(cons, dvis, floating, i, r) ? T(0) : T(1);
  }
}

int grid_n(int64_t work, int per) {
  int64_t g = (work + per - 1) / per;
  g = std::max<int64_t>(8, std::min<int64_t>(g, kMaxGrid));
  return (int)g;
}

// Row shape of the finest level: 16 lanes a row and chunks_per_pass(c) chunks
// per pass (as the Jacobi CG's SpMV) when every row fits four chunks, else 64
// lanes.  Coarser levels take 32 or 64 lanes, one chunk per pass.
void fine_shape(const Context* c, int* lpr, int* u) {
  if (c->max_nconn <= 4 * kLPR) {
    *lpr = kLPR;
    *u = chunks_per_pass(c);
  } else {
    *lpr = 64;
    *u = 1;
  }
}

// f(lanes per row, chunks per pass) with both as compile-time constants.
template <class F>
void with_shape(int lpr, int u, F&& f) {
  using std::integral_constant;
  if (lpr == 16) {
    if (u == 2) f(integral_constant<int, 16>{}, integral_constant<int, 2>{});
    else if (u == 3) f(integral_constant<int, 16>{}, integral_constant<int, 3>{});
    else f(integral_constant<int, 16>{}, integral_constant<int, 4>{});
  } else if (lpr == 32) {
    f(integral_constant<int, 32>{}, integral_constant<int, 1>{});
  } else {
    f(integral_constant<int, 64>{}, integral_constant<int, 1>{});
  }
}

int rows_grid(int N, int lpr) {
  const int rows_per_block = (kSpBS / 64) * (64 / lpr);
  return (N + rows_per_block - 1) / rows_per_block;
}

// flag = 1 when any bond slot is broken (|bdamage - 1| < kEPS)
__global__ void any_broken_kernel(const double* __restrict__ bdamage, int64_t nb, int* __restrict__ flag) {
  for (int64_t b = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; b < nb; b += (int64_t)gridDim.x * blockDim.x)
    if (fabs(bdamage[b] - 1.0) < kEPS) atomicOr(flag, 1);
}

}  // namespace

// ---- hierarchy ------------------------------------------------------------

struct AmgLevel {  // level l >= 1; its nodes are the aggregates of level l-1
  int N = 0, lpr = 32;
  int64_t nnzb = 0;
  int *cptr = nullptr, *conns = nullptr, *dpos = nullptr;
  int64_t* gptr = nullptr;   // nnzb + 1: this level's blocks -> ranges of gidx
  int* gidx = nullptr;       // level l-1 block indices, grouped by coarse block
  int *mptr = nullptr, *midx = nullptr;  // members (level l-1 nodes) of each aggregate
  int* agg = nullptr;        // aggregate of each level l-1 node
  void* val[2] = {nullptr, nullptr};   // FP64 / FP32 values
  void* dinv[2] = {nullptr, nullptr};
  void* vec[2][5] = {{nullptr}};       // b, r, x, y, x1 (smoothed hierarchy) per precision
  // smoothed aggregation (level l-1 -> l): P's blocks by row (pptr over level
  // l-1 nodes; column, row, and the level l-1 blocks each one sums), by
  // column (qptr over the aggregates), and the pair lists of Y = A P and of
  // this level's blocks P^T Y
  int nP = 0;
  int64_t nY = 0;
  int sa_gp = 1;                 // P blocks per wavefront of sa_p_kernel
  int pmax = 0, ymax = 0, cmax = 0;  // longest P, Y and coarse row
  int *pptr = nullptr, *pcol = nullptr, *prow = nullptr, *atptr = nullptr, *atidx = nullptr;
  int *qptr = nullptr, *qrow = nullptr, *qpos = nullptr;
  int* yptr = nullptr;                       // Y's rows
  int64_t *ysp = nullptr, *csp = nullptr;     // each row's start in the slot lists
  int *yso = nullptr, *cso = nullptr;         // each source's first slot within its row's list
  uint16_t *ys = nullptr, *cs = nullptr;      // slot lists of the Y and coarse-block sums
  void* pval[2] = {nullptr, nullptr};   // P per precision, by row
  void* pvalT[2] = {nullptr, nullptr};  // the same blocks by column (qpos gives each one's place)
  double* yval = nullptr;
};

struct Amg {
  bool ok = false;             // false: the hierarchy could not be built (falls back to Jacobi)
  bool smoothed = false;       // smoothed aggregation (option 4) instead of the tentative prolongation
  std::vector<AmgLevel> lv;    // levels 1..L
  int nc = 0;                  // unknowns of the coarsest level
  void* inv[2] = {nullptr, nullptr};
  void* mask[2] = {nullptr, nullptr};  // finest-level free-DOF mask per precision
  int64_t version[2] = {-1, -1};       // the assembly the values were built from
  int64_t values_step[2] = {-1, -1};   // the load step they were built in (-1: none)
  int64_t version16 = -1;              // the assembly the bfloat16 copy was rounded from
  uint16_t* fine16 = nullptr;  // FP32: the finest level's values as bfloat16, for the smoothers
  hipGraphExec_t graph[2] = {nullptr, nullptr};
  std::vector<uintptr_t> graph_sig[2];  // pointers and shapes captured into graph[P]
  hipGraphExec_t graph_tail[2] = {nullptr, nullptr};  // the short chunk that ends a solve
  std::vector<uintptr_t> graph_tail_sig[2];
  int last_iters[2] = {0, 0};  // the previous solve's iterations per precision
  // smoother damping per level with a smoother (device memory, read by the
  // V-cycle's kernels): 2/3 for the tentative hierarchy; for the smoothed one
  // set after every value rebuild from a Lanczos estimate of the level's
  // largest eigenvalue (lz_coef: its coefficients, lz_part: its partial sums)
  double* omega = nullptr;
  double* omega_p = nullptr;  // the prolongator's smoothing factor per level (smoothed hierarchy)
  double* lz_coef = nullptr;
  double* lz_part = nullptr;
  int lz_nparts = 0;
  std::vector<hipGraphExec_t> lz_graphs;  // each level's estimate, captured once
  std::vector<std::vector<uintptr_t>> lz_sigs;
  // the hierarchy's device arrays (freed with it) and their bytes
  std::vector<void*> allocs;
  size_t bytes = 0;
  // smoothed hierarchy: the finest level's grid cell and aggregate of each
  // node at build time, host copies of the bond lists, the context's
  // break_version the aggregates were last checked against, and whether a
  // load step has begun since (the check runs once per step)
  std::vector<int> cell0, agg0, bonds_h;
  int64_t break_version = -1;
  bool recheck = false;
  int* flag = nullptr;  // device word for any_broken
};

namespace {

// Stream-ordered allocation owned by the hierarchy: amg_free returns it (a
// handle that switches between the two kinds of hierarchy keeps one).
template <typename T>
int amg_alloc(Context* c, T** p, size_t n) {
  if (n == 0) n = 1;
  void* q = nullptr;
  hipError_t e = hipMallocAsync(&q, n * sizeof(T), c->stream);
  if (e != hipSuccess) return hip_fail(e, "hipMallocAsync");
  c->amg->allocs.push_back(q);
  c->amg->bytes += n * sizeof(T);
  c->bytes += n * sizeof(T);
  *p = static_cast<T*>(q);
  return 0;
}

// Smoothed aggregation, symbolic (host, once per handle): for a level of N
// nodes with block pattern (cptr, conns) and aggregates agg (Nc of them), the
// patterns of P, of Y = A P and of the coarse operator P^T Y, with the lists
// each of their blocks sums (in a fixed order: by column, then by the blocks'
// positions), and the coarse level's pattern.
struct SaSym {
  std::vector<int> pptr, pcol, prow, atptr, atidx, qptr, qidx, qrow, qpos;
  std::vector<int> yptr, ycol;      // pattern of Y = A P
  std::vector<uint16_t> ys, cs;     // slot lists of the Y and coarse-block sums
  std::vector<int64_t> ysp, csp;    // each row's start in them
  std::vector<int> yso, cso;        // each source's first slot within its row's list
  std::vector<int> ccptr, cconns, dpos;
  int pmax = 0, ymax = 0, cmax = 0;  // longest P, Y and coarse row
  int64_t nY = 0;                    // blocks of Y
  bool device_lists = false;         // the Y and coarse lists were built on the device
};

// Runs f(t, r0, r1) over the row ranges of nt host threads (thread 0 is the
// caller).
template <class F>
void host_rows(int nrows, F&& f) {
  const int nt = std::max(1, std::min<int>(16, (int)std::thread::hardware_concurrency()));
  const int per = (nrows + nt - 1) / nt;
  std::vector<std::thread> th;
  for (int t = 1; t < nt; ++t) {
    const int r0 = std::min(nrows, t * per), r1 = std::min(nrows, r0 + per);
    if (r0 < r1) th.emplace_back([&f, t, r0, r1] { f(t, r0, r1); });
  }
  f(0, 0, std::min(nrows, per));
  for (auto& x : th) x.join();
}

// Rows [0, nrows) each emit columns through gen(row, emit).  Out, in row
// order: each row's distinct columns ascending (rptr, cols), and for every
// emission, in emission order, the position of its column among them (slot,
// rows starting at sptr).  Two passes over host threads -- counts, then the
// rows written in place at their offsets -- so nothing is copied or grown;
// the result does not depend on the split.  Returns the longest row, or -1
// past 65,535 columns.
template <class Gen>
int slot_rows(int nrows, int ncols, Gen gen, std::vector<int>* rptr, std::vector<int>* cols,
              std::vector<uint16_t>* slot, std::vector<int64_t>* sptr) {
  std::vector<int> ucnt(nrows);
  std::vector<int64_t> ecnt(nrows);
  std::vector<int> longest(16, 0);
  host_rows(nrows, [&](int t, int r0, int r1) {
    std::vector<int> stamp(ncols, -1);
    for (int i = r0; i < r1; ++i) {
      int u = 0;
      int64_t e = 0;
      gen(i, [&](int col) {
        if (stamp[col] != i) {
          stamp[col] = i;
          ++u;
        }
        ++e;
      });
      ucnt[i] = u;
      ecnt[i] = e;
      longest[t] = std::max(longest[t], u);
    }
  });
  rptr->assign(nrows + 1, 0);
  sptr->assign(nrows + 1, 0);
  for (int i = 0; i < nrows; ++i) {
    (*rptr)[i + 1] = (*rptr)[i] + ucnt[i];
    (*sptr)[i + 1] = (*sptr)[i] + ecnt[i];
  }
  const int lmax = *std::max_element(longest.begin(), longest.end());
  if (lmax > 65535) return -1;
  cols->resize((*rptr)[nrows]);
  slot->resize((*sptr)[nrows]);
  host_rows(nrows, [&](int, int r0, int r1) {
    std::vector<int> stamp(ncols, -1), pos(ncols, 0), em;
    for (int i = r0; i < r1; ++i) {
      int* rc = cols->data() + (*rptr)[i];
      int u = 0;
      em.clear();
      gen(i, [&](int col) {
        em.push_back(col);
        if (stamp[col] != i) {
          stamp[col] = i;
          rc[u++] = col;
        }
      });
      std::sort(rc, rc + u);
      for (int k = 0; k < u; ++k) pos[rc[k]] = k;
      uint16_t* rs = slot->data() + (*sptr)[i];
      for (size_t e = 0; e < em.size(); ++e) rs[e] = (uint16_t)pos[em[e]];
    }
  });
  return lmax;
}

int sa_symbolic_p(int N, int Nc, const std::vector<int>& cptr, const std::vector<int>& conns,
                  const std::vector<int>& agg, SaSym* s) {
  // P: a row's blocks are the distinct aggregates of its columns (ascending);
  // each lists the row's level blocks of that aggregate in order (atidx, the
  // row's entries at the row's own offset: every level block belongs to one
  // P block).  Two passes over host threads: counts, then the rows in place.
  std::vector<int> ucnt(N);
  host_rows(N, [&](int, int r0, int r1) {
    std::vector<int> stamp(Nc, -1);
    for (int i = r0; i < r1; ++i) {
      int u = 0;
      for (int e = cptr[i]; e < cptr[i + 1]; ++e) {
        const int a = agg[conns[e]];
        if (stamp[a] != i) {
          stamp[a] = i;
          ++u;
        }
      }
      ucnt[i] = u;
    }
  });
  s->pptr.assign(N + 1, 0);
  for (int i = 0; i < N; ++i) s->pptr[i + 1] = s->pptr[i] + ucnt[i];
  const int nP = s->pptr[N];
  s->pcol.resize(nP);
  s->prow.resize(nP);
  s->atptr.assign(nP + 1, 0);
  s->atidx.resize(cptr[N]);
  host_rows(N, [&](int, int r0, int r1) {
    std::vector<int> stamp(Nc, -1), pos(Nc, 0), fill;
    for (int i = r0; i < r1; ++i) {
      int* pc = s->pcol.data() + s->pptr[i];
      int u = 0;
      for (int e = cptr[i]; e < cptr[i + 1]; ++e) {
        const int a = agg[conns[e]];
        if (stamp[a] != i) {
          stamp[a] = i;
          pc[u++] = a;
        }
      }
      std::sort(pc, pc + u);
      fill.assign(u + 1, 0);
      for (int k = 0; k < u; ++k) pos[pc[k]] = k;
      for (int e = cptr[i]; e < cptr[i + 1]; ++e) ++fill[pos[agg[conns[e]]] + 1];
      for (int k = 0; k < u; ++k) fill[k + 1] += fill[k];
      for (int k = 0; k < u; ++k) {
        s->prow[s->pptr[i] + k] = i;
        s->atptr[s->pptr[i] + k + 1] = cptr[i] + fill[k + 1];
      }
      for (int e = cptr[i]; e < cptr[i + 1]; ++e) s->atidx[cptr[i] + fill[pos[agg[conns[e]]]]++] = e;
    }
  });
  // P's blocks by aggregate, in row order (a stable counting sort: per-thread
  // counts over row ranges, then each thread's blocks after the earlier
  // threads' in every column)
  const int nt = std::max(1, std::min<int>(16, (int)std::thread::hardware_concurrency()));
  std::vector<std::vector<int>> tc(nt, std::vector<int>(Nc, 0));
  host_rows(nP, [&](int t, int k0, int k1) {
    for (int p = k0; p < k1; ++p) ++tc[t][s->pcol[p]];
  });
  s->qptr.assign(Nc + 1, 0);
  {
    int run = 0;
    for (int a = 0; a < Nc; ++a) {
      s->qptr[a] = run;
      for (int t = 0; t < nt; ++t) {
        const int n = tc[t][a];
        tc[t][a] = run;
        run += n;
      }
    }
    s->qptr[Nc] = run;
  }
  s->qidx.assign(nP, 0);
  host_rows(nP, [&](int t, int k0, int k1) {
    for (int p = k0; p < k1; ++p) s->qidx[tc[t][s->pcol[p]]++] = p;
  });
  s->qrow.assign(nP, 0);
  s->qpos.assign(nP, 0);
  host_rows(nP, [&](int, int k0, int k1) {
    for (int k = k0; k < k1; ++k) {
      s->qrow[k] = s->prow[s->qidx[k]];
      s->qpos[s->qidx[k]] = k;
    }
  });
  for (int i = 0; i < N; ++i) s->pmax = std::max(s->pmax, s->pptr[i + 1] - s->pptr[i]);
  return 0;
}

// The Y and coarse lists on the host (when the device cannot hold a row).
int sa_symbolic_yc(int N, int Nc, const std::vector<int>& cptr, const std::vector<int>& conns, SaSym* s) {
  const int nP = (int)s->pcol.size();
  // Y = A P: row i's blocks by column b; a slot per (A block e, P block q) in
  // (e, q) order
  s->ymax = slot_rows(N, Nc, [&](int i, auto emit) {
    for (int e = cptr[i]; e < cptr[i + 1]; ++e) {
      const int j = conns[e];
      for (int q = s->pptr[j]; q < s->pptr[j + 1]; ++q) emit(s->pcol[q]);
    }
  }, &s->yptr, &s->ycol, &s->ys, &s->ysp);
  // coarse blocks (a,b) = P^T Y: a slot per (P block k of column a, Y block y
  // of row i(k)) in (k, y) order
  s->cmax = slot_rows(Nc, Nc, [&](int a, auto emit) {
    for (int k = s->qptr[a]; k < s->qptr[a + 1]; ++k) {
      const int i = s->qrow[k];
      for (int y = s->yptr[i]; y < s->yptr[i + 1]; ++y) emit(s->ycol[y]);
    }
  }, &s->ccptr, &s->cconns, &s->cs, &s->csp);
  if (s->ymax < 0 || s->cmax < 0) {
    set_error("multigrid: a smoothed-aggregation row exceeds 65,535 blocks");
    return -1;
  }
  s->yso.assign(cptr[N], 0);
  host_rows(N, [&](int, int r0, int r1) {
    for (int i = r0; i < r1; ++i)
      for (int e = cptr[i], off = 0; e < cptr[i + 1]; ++e) {
        s->yso[e] = off;
        off += s->pptr[conns[e] + 1] - s->pptr[conns[e]];
      }
  });
  s->cso.assign(nP, 0);
  host_rows(Nc, [&](int, int r0, int r1) {
    for (int a = r0; a < r1; ++a)
      for (int k = s->qptr[a], off = 0; k < s->qptr[a + 1]; ++k) {
        s->cso[k] = off;
        off += s->yptr[s->qrow[k] + 1] - s->yptr[s->qrow[k]];
      }
  });
  s->dpos.assign(Nc, -1);
  for (int a = 0; a < Nc; ++a) {
    for (int k = s->ccptr[a]; k < s->ccptr[a + 1]; ++k)
      if (s->cconns[k] == a) s->dpos[a] = k;
    if (s->dpos[a] < 0) {
      set_error("multigrid: aggregate without a diagonal block");
      return -1;
    }
  }
  s->nY = (int64_t)s->ycol.size();
  return 0;
}

template <typename T>
int sa_upload(Context* c, T** d, const std::vector<T>& h) {
  if (amg_alloc(c, d, h.size())) return -3;
  if (!h.empty()) LPMG_HIP(copy_sync(c, *d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
  return 0;
}

// The Y and coarse lists on the device (sa_sym_kernel) for a level of N rows
// with block pattern (dcptr, dconns) on the device and P's lists already in
// L (pptr, pcol, qptr, qrow).  Fills L's yptr, ysp, ys, yso, csp, cs, cso and,
// on the host, s's coarse pattern (ccptr, cconns, dpos), ymax, cmax and nY.
// *done stays false when a row does not fit the kernel's LDS arrays (the
// caller then builds the lists on the host).
int sa_symbolic_yc_device(Context* c, int N, int Nc, int64_t nnzb, const int* dcptr, const int* dconns,
                          AmgLevel* L, SaSym* s, bool* done) {
  *done = false;
  const int nP = (int)s->pcol.size();
  const int nr = std::max(N, Nc);
  int* flag = nullptr;
  int* ucnt = nullptr;
  int64_t* ecnt = nullptr;
  int* ycol = nullptr;
  int* dccptr = nullptr;
  int* dcconns = nullptr;
  int* ddiag = nullptr;
  struct Tmp {
    Context* c;
    std::vector<void*> p;
    ~Tmp() {
      for (void* q : p)
        if (q) (void)hipFreeAsync(q, c->stream);
    }
  } tmp{c, {}};
  auto talloc = [&](auto** q, size_t bytes) -> int {
    void* v = nullptr;
    hipError_t e = hipMallocAsync(&v, std::max<size_t>(bytes, 8), c->stream);
    if (e != hipSuccess) return hip_fail(e, "hipMallocAsync");
    tmp.p.push_back(v);
    *q = static_cast<std::remove_pointer_t<std::remove_reference_t<decltype(q)>>>(v);
    return 0;
  };
  if (talloc(&flag, sizeof(int)) || talloc(&ucnt, sizeof(int) * nr) || talloc(&ecnt, sizeof(int64_t) * nr))
    return -3;
  LPMG_HIP(hipMemsetAsync(flag, 0, sizeof(int), c->stream));
  int hflag = 0;
  std::vector<int> hu(nr);
  std::vector<int64_t> he(nr);
  auto counts = [&](int rows) -> int {
    LPMG_HIP(copy_sync(c, &hflag, flag, sizeof(int), hipMemcpyDeviceToHost));
    if (hflag) return 0;
    LPMG_HIP(copy_sync(c, hu.data(), ucnt, sizeof(int) * rows, hipMemcpyDeviceToHost));
    LPMG_HIP(copy_sync(c, he.data(), ecnt, sizeof(int64_t) * rows, hipMemcpyDeviceToHost));
    return 0;
  };
  // Y = A P: counts and the sources' offsets
  if (amg_alloc(c, &L->yso, (size_t)nnzb)) return -3;
  sa_sym_kernel<0><<<N, 64, 0, c->stream>>>(N, dcptr, dconns, L->pptr, L->pcol, ucnt, ecnt, L->yso, nullptr,
                                            nullptr, nullptr, nullptr, nullptr, flag);
  LPMG_HIP(hipGetLastError());
  if (int rc = counts(N)) return rc;
  if (hflag) return 0;
  std::vector<int> yptr(N + 1, 0);
  std::vector<int64_t> ysp(N + 1, 0);
  int ymax = 0;
  for (int i = 0; i < N; ++i) {
    yptr[i + 1] = yptr[i] + hu[i];
    ysp[i + 1] = ysp[i] + he[i];
    ymax = std::max(ymax, hu[i]);
    if ((int64_t)yptr[i] + hu[i] >= ((int64_t)1 << 31)) {
      set_error("multigrid: smoothed-aggregation rows too long for the device sums");
      return -1;
    }
  }
  const int64_t nY = yptr[N];
  if (sa_upload(c, &L->yptr, yptr) || sa_upload(c, &L->ysp, ysp)) return -3;
  if (amg_alloc(c, &L->ys, (size_t)ysp[N]) || talloc(&ycol, sizeof(int) * (size_t)nY)) return -3;
  sa_sym_kernel<1><<<N, 64, 0, c->stream>>>(N, dcptr, dconns, L->pptr, L->pcol, nullptr, nullptr, nullptr, L->yptr,
                                            L->ysp, ycol, L->ys, nullptr, flag);
  LPMG_HIP(hipGetLastError());
  // coarse rows P^T Y: counts and the sources' offsets
  if (amg_alloc(c, &L->cso, (size_t)std::max(nP, 1))) return -3;
  sa_sym_kernel<0><<<Nc, 64, 0, c->stream>>>(Nc, L->qptr, L->qrow, L->yptr, ycol, ucnt, ecnt, L->cso, nullptr,
                                             nullptr, nullptr, nullptr, nullptr, flag);
  LPMG_HIP(hipGetLastError());
  if (int rc = counts(Nc)) return rc;
  if (hflag) return 0;
  s->ccptr.assign(Nc + 1, 0);
  std::vector<int64_t> csp(Nc + 1, 0);
  int cmax = 0;
  for (int a = 0; a < Nc; ++a) {
    s->ccptr[a + 1] = s->ccptr[a] + hu[a];
    csp[a + 1] = csp[a] + he[a];
    cmax = std::max(cmax, hu[a]);
  }
  const int nC = s->ccptr[Nc];
  if (sa_upload(c, &L->csp, csp)) return -3;
  if (amg_alloc(c, &L->cs, (size_t)csp[Nc]) || talloc(&dccptr, sizeof(int) * (Nc + 1)) ||
      talloc(&dcconns, sizeof(int) * (size_t)nC) || talloc(&ddiag, sizeof(int) * Nc))
    return -3;
  LPMG_HIP(copy_sync(c, dccptr, s->ccptr.data(), sizeof(int) * (Nc + 1), hipMemcpyHostToDevice));
  sa_sym_kernel<1><<<Nc, 64, 0, c->stream>>>(Nc, L->qptr, L->qrow, L->yptr, ycol, nullptr, nullptr, nullptr, dccptr,
                                             L->csp, dcconns, L->cs, ddiag, flag);
  LPMG_HIP(hipGetLastError());
  LPMG_HIP(copy_sync(c, &hflag, flag, sizeof(int), hipMemcpyDeviceToHost));
  if (hflag) return 0;
  s->cconns.resize(nC);
  s->dpos.resize(Nc);
  LPMG_HIP(copy_sync(c, s->cconns.data(), dcconns, sizeof(int) * (size_t)nC, hipMemcpyDeviceToHost));
  LPMG_HIP(copy_sync(c, s->dpos.data(), ddiag, sizeof(int) * Nc, hipMemcpyDeviceToHost));
  for (int a = 0; a < Nc; ++a) {
    if (s->dpos[a] < 0) {
      set_error("multigrid: aggregate without a diagonal block");
      return -1;
    }
    s->dpos[a] += s->ccptr[a];
  }
  s->ymax = ymax;
  s->cmax = cmax;
  s->nY = nY;
  s->device_lists = true;
  *done = true;
  return 0;
}

// The intact bonds as node pairs (i, j), i < j: a bond joins its particles
// while either of its two slots is intact (|bdamage - 1| >= kEPS, the fragment
// labelling's test).  The bond lists are copied to the host once per hierarchy.
int intact_edges(Context* c, std::vector<int>& bl, std::vector<std::pair<int, int>>* e) {
  const int N = c->N;
  const int64_t NB = c->NB;
  if (bl.empty()) {  // bptr (N + 1), then p2 and rev (NB each)
    bl.resize(N + 1 + 2 * NB);
    LPMG_HIP(copy_sync(c, bl.data(), c->bptr, sizeof(int) * (N + 1), hipMemcpyDeviceToHost));
    if (NB) {
      LPMG_HIP(copy_sync(c, bl.data() + N + 1, c->p2, sizeof(int) * NB, hipMemcpyDeviceToHost));
      LPMG_HIP(copy_sync(c, bl.data() + N + 1 + NB, c->rev, sizeof(int) * NB, hipMemcpyDeviceToHost));
    }
  }
  const int* bptr = bl.data();
  const int* p2 = bptr + N + 1;
  const int* rev = p2 + NB;
  std::vector<double> bd(std::max<int64_t>(NB, 1));
  if (NB) LPMG_HIP(copy_sync(c, bd.data(), c->bdamage, sizeof(double) * NB, hipMemcpyDeviceToHost));
  e->clear();
  e->reserve(NB / 2 + 1);
  for (int i = 0; i < N; ++i)
    for (int b = bptr[i]; b < bptr[i + 1]; ++b)
      if (i < p2[b] && (fabs(bd[b] - 1.0) >= kEPS || fabs(bd[rev[b]] - 1.0) >= kEPS)) e->emplace_back(i, p2[b]);
  return 0;
}

// Whether the aggregates should follow the bonds: in 2D always (a notch cut
// into the lattice splits cells before any bond breaks; 1.3M bonds cost ~10 ms
// on the 205k CT request), in 3D once some bond is broken (the 1.0M-particle
// polycrystal's 7M bonds cost 0.17 s per build and split no cell).
int split_by_bonds(Context* c, bool* split) {
  *split = true;
  if (c->dim == 2 || c->NB == 0) return 0;
  Amg* A = c->amg;
  if (!A->flag && amg_alloc(c, &A->flag, 1)) return -3;
  LPMG_HIP(hipMemsetAsync(A->flag, 0, sizeof(int), c->stream));
  any_broken_kernel<<<kMaxGrid, kBS, 0, c->stream>>>(c->bdamage, c->NB, A->flag);
  LPMG_HIP(hipGetLastError());
  int any = 0;
  LPMG_HIP(copy_sync(c, &any, A->flag, sizeof(int), hipMemcpyDeviceToHost));
  *split = any != 0;
  return 0;
}

// Union-find over the edges whose two nodes share a group; each node's root
// is the smallest node of its component (union by the smaller index).
std::vector<int> group_components(int N, const std::vector<std::pair<int, int>>& e, const std::vector<int>& group) {
  std::vector<int> par(N);
  std::iota(par.begin(), par.end(), 0);
  auto find = [&](int x) {
    while (par[x] != x) {
      par[x] = par[par[x]];
      x = par[x];
    }
    return x;
  };
  for (const auto& uv : e) {
    if (group[uv.first] != group[uv.second]) continue;
    const int a = find(uv.first), b = find(uv.second);
    if (a != b) par[std::max(a, b)] = std::min(a, b);
  }
  for (int i = 0; i < N; ++i) par[i] = find(i);
  return par;
}

// Split every aggregate into the connected components of the intact bonds
// (at coarser levels: the bonds between aggregates) inside it, so that no
// aggregate spans a crack.  The new aggregates keep the order of their cells;
// a cell's components follow their smallest node.  Members stay in node order.
// A split that would leave more than 80% of the nodes as aggregates (a level
// falling apart into fragments) is not made.
void split_aggregates(int N, const std::vector<std::pair<int, int>>& e, std::vector<int>& agg,
                      std::vector<int>& mptr, std::vector<int>& midx, int& Nc) {
  const std::vector<int> root = group_components(N, e, agg);
  std::vector<int> rid(N, -1);
  int n2 = 0;
  for (int q = 0; q < N; ++q) {  // members run in cell order, node order inside a cell
    const int r = root[midx[q]];
    if (rid[r] < 0) rid[r] = n2++;
  }
  if (n2 == Nc || n2 > (int)(0.8 * N)) return;
  std::vector<int> nptr(n2 + 1, 0), nmidx(N);
  for (int i = 0; i < N; ++i) {
    agg[i] = rid[root[i]];
    ++nptr[agg[i] + 1];
  }
  for (int a = 0; a < n2; ++a) nptr[a + 1] += nptr[a];
  std::vector<int> fill(nptr.begin(), nptr.end() - 1);
  for (int q = 0; q < N; ++q) nmidx[fill[agg[midx[q]]]++] = midx[q];
  mptr.swap(nptr);
  midx.swap(nmidx);
  Nc = n2;
}

// The next level's edges: each pair of distinct aggregates joined by an edge,
// once (bucketed by the smaller aggregate, repeats dropped with a marker).
void coarse_edges(const std::vector<int>& agg, int Nc, std::vector<std::pair<int, int>>& e) {
  std::vector<int> start(Nc + 1, 0);
  for (const auto& uv : e) {
    const int a = agg[uv.first], b = agg[uv.second];
    if (a != b) ++start[std::min(a, b) + 1];
  }
  for (int a = 0; a < Nc; ++a) start[a + 1] += start[a];
  std::vector<int> fill(start.begin(), start.end() - 1), other(start[Nc]);
  for (const auto& uv : e) {
    const int a = agg[uv.first], b = agg[uv.second];
    if (a != b) other[fill[std::min(a, b)]++] = std::max(a, b);
  }
  std::vector<int> mark(Nc, -1);
  e.clear();
  for (int a = 0; a < Nc; ++a)
    for (int k = start[a]; k < start[a + 1]; ++k)
      if (mark[other[k]] != a) {
        mark[other[k]] = a;
        e.emplace_back(a, other[k]);
      }
}

// Whether the finest level's cells, split by the bonds intact now, give other
// aggregates than the hierarchy's (bonds broke across a cell, or healed).
int amg_split_stale(Context* c, bool* stale) {
  Amg* A = c->amg;
  *stale = false;
  if (A->cell0.empty()) return 0;  // a single-level hierarchy: nothing aggregated
  const int N = c->N;
  bool split = false;
  if (int rc = split_by_bonds(c, &split)) return rc;
  std::vector<int> root = A->cell0;
  if (split) {
    std::vector<std::pair<int, int>> e;
    if (int rc = intact_edges(c, A->bonds_h, &e)) return rc;
    root = group_components(N, e, A->cell0);
    int n2 = 0;
    for (int i = 0; i < N; ++i) n2 += root[i] == i;
    if (n2 > (int)(0.8 * N)) root = A->cell0;  // such a split is not made
  }
  // the two partitions agree when each aggregate maps to one root and back
  std::vector<int> a2r(N, -1), r2a(N, -1);
  for (int i = 0; i < N; ++i) {
    const int a = A->agg0[i], r = root[i];
    if (a2r[a] < 0) a2r[a] = r;
    if (r2a[r] < 0) r2a[r] = a;
    if (a2r[a] != r || r2a[r] != a) {
      *stale = true;
      return 0;
    }
  }
  return 0;
}

// Build the aggregation hierarchy on the host from the initial positions and
// the level-0 block pattern.
int amg_build(Context* c) {
  Amg* A = c->amg;
  const int D = c->dim;
  int N = c->N;
  std::vector<int> cptr(N + 1), conns(c->nnzb);
  std::vector<double4> x4(N);
  LPMG_HIP(copy_sync(c, cptr.data(), c->cptr, sizeof(int) * (N + 1), hipMemcpyDeviceToHost));
  LPMG_HIP(copy_sync(c, conns.data(), c->conns, sizeof(int) * c->nnzb, hipMemcpyDeviceToHost));
  LPMG_HIP(copy_sync(c, x4.data(), c->xyz0, sizeof(double4) * N, hipMemcpyDeviceToHost));
  std::vector<double> pos(3 * (size_t)N);
  for (int i = 0; i < N; ++i) {
    pos[3 * i] = x4[i].x;
    pos[3 * i + 1] = x4[i].y;
    pos[3 * i + 2] = x4[i].z;
  }
  double h = 3.0 * 2.0 * c->cell.radius;
  // the smoothed hierarchy's aggregates follow the intact bonds
  std::vector<std::pair<int, int>> edges;
  bool split = false;
  if (A->smoothed) {
    if (int rc = split_by_bonds(c, &split)) return rc;
    if (split)
      if (int rc = intact_edges(c, A->bonds_h, &edges)) return rc;
    A->break_version = c->break_version;
  }
  while ((int64_t)N * D > kCoarsestMax) {
    // aggregates: nonempty cells of a grid of edge h, numbered in cell order
    double lo[3] = {1e300, 1e300, 1e300};
    for (int i = 0; i < N; ++i)
      for (int k = 0; k < 3; ++k) lo[k] = std::min(lo[k], pos[3 * i + k]);
    std::vector<int64_t> key(N);
    int64_t dims[3] = {1, 1, 1};
    std::vector<int64_t> cc(3 * (size_t)N);
    for (int i = 0; i < N; ++i)
      for (int k = 0; k < 3; ++k) {
        cc[3 * i + k] = (int64_t)std::floor((pos[3 * i + k] - lo[k]) / h);
        dims[k] = std::max(dims[k], cc[3 * i + k] + 1);
      }
    for (int i = 0; i < N; ++i) key[i] = (cc[3 * i + 2] * dims[1] + cc[3 * i + 1]) * dims[0] + cc[3 * i];
    // nodes in (cell, node) order: a counting sort over the cells when the
    // grid has at most a few cells per node (the 1.0M-particle polycrystal's
    // finest cells with their keys: 71 -> 16 ms), else a stable sort
    std::vector<int> order(N);
    const int64_t ncell = dims[0] * dims[1] * dims[2];
    if (ncell <= 4 * (int64_t)N + 1024) {
      std::vector<int> at(ncell + 1, 0);
      for (int i = 0; i < N; ++i) ++at[key[i] + 1];
      for (int64_t k = 0; k < ncell; ++k) at[k + 1] += at[k];
      for (int i = 0; i < N; ++i) order[at[key[i]]++] = i;
    } else {
      std::iota(order.begin(), order.end(), 0);
      std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return key[a] < key[b]; });
    }
    std::vector<int> agg(N), mptr(1, 0), midx(N);
    int Nc = 0;
    for (int t = 0; t < N; ++t) {
      if (t > 0 && key[order[t]] != key[order[t - 1]]) {
        mptr.push_back(t);
        ++Nc;
      }
      agg[order[t]] = Nc;
      midx[t] = order[t];
    }
    ++Nc;
    mptr.push_back(N);
    if (Nc > (int)(0.8 * N)) {  // too little coarsening at this cell size
      h *= 2.0;
      if (h > 1e6 * c->cell.radius) break;
      continue;
    }
    if (A->smoothed) {
      if (A->lv.empty()) A->cell0 = agg;
      if (split) split_aggregates(N, edges, agg, mptr, midx, Nc);
      if (A->lv.empty()) A->agg0 = agg;
    }
    // coarse pattern and gather lists: per aggregate row, the (column, fine
    // block) pairs sorted by column then fine block; each column run is one
    // coarse block and its gather list
    std::vector<int> ccptr(Nc + 1, 0), cconns, dpos(Nc, -1), gidx;
    std::vector<int64_t> gptr(1, 0);
    std::vector<std::pair<int, int>> pairs;
    SaSym sa;
    AmgLevel L;
    if (A->smoothed) {
      if (int rc = sa_symbolic_p(N, Nc, cptr, conns, agg, &sa)) return rc;
      // P's lists on the device, then the Y and coarse lists: built on the
      // device, or on the host when a row does not fit the device kernel
      const std::pair<int**, const std::vector<int>*> ups[] = {
          {&L.pptr, &sa.pptr}, {&L.pcol, &sa.pcol}, {&L.prow, &sa.prow}, {&L.atptr, &sa.atptr},
          {&L.atidx, &sa.atidx}, {&L.qptr, &sa.qptr}, {&L.qrow, &sa.qrow}, {&L.qpos, &sa.qpos}};
      for (const auto& up : ups)
        if (int r2 = sa_upload(c, up.first, *up.second)) return r2;
      const int* dcptr = A->lv.empty() ? c->cptr : A->lv.back().cptr;
      const int* dconns = A->lv.empty() ? c->conns : A->lv.back().conns;
      // LPMG_SA_HOST_SYMBOLIC set: the host builder only (the two give the
      // same lists; tests compare them through their results)
      bool done = false;
      if (!std::getenv("LPMG_SA_HOST_SYMBOLIC"))
        if (int rc = sa_symbolic_yc_device(c, N, Nc, (int64_t)cptr[N], dcptr, dconns, &L, &sa, &done)) return rc;
      if (!done)
        if (int rc = sa_symbolic_yc(N, Nc, cptr, conns, &sa)) return rc;
      // a row's accumulators live in LDS (sa_rows_kernel)
      if ((int64_t)std::max(sa.ymax, sa.cmax) * D * D * (int64_t)sizeof(double) > 65536 ||
          sa.nY >= ((int64_t)1 << 31)) {
        set_error("multigrid: smoothed-aggregation rows too long for the device sums");
        return -1;
      }
      ccptr = sa.ccptr;
      cconns = sa.cconns;
      dpos = sa.dpos;
    }
    for (int a = 0; a < Nc && !A->smoothed; ++a) {
      pairs.clear();
      for (int q = mptr[a]; q < mptr[a + 1]; ++q) {
        const int i = midx[q];
        for (int e = cptr[i]; e < cptr[i + 1]; ++e) pairs.emplace_back(agg[conns[e]], e);
      }
      std::sort(pairs.begin(), pairs.end());
      const size_t np = pairs.size();
      for (size_t t = 0; t < np; ++t) {
        if (t == 0 || pairs[t].first != pairs[t - 1].first) {
          if (pairs[t].first == a) dpos[a] = (int)cconns.size();
          cconns.push_back(pairs[t].first);
        }
        gidx.push_back(pairs[t].second);
        if (t + 1 == np || pairs[t + 1].first != pairs[t].first) gptr.push_back((int64_t)gidx.size());
      }
      ccptr[a + 1] = (int)cconns.size();
      if (dpos[a] < 0) {
        set_error("multigrid: aggregate without a diagonal block");
        return -1;
      }
    }
    L.N = Nc;
    L.nnzb = (int64_t)cconns.size();
    int maxrow = 0;
    for (int a = 0; a < Nc; ++a) maxrow = std::max(maxrow, ccptr[a + 1] - ccptr[a]);
    L.lpr = maxrow <= 32 ? 32 : 64;
    int rc = 0;
    rc |= amg_alloc(c, &L.cptr, Nc + 1);
    rc |= amg_alloc(c, &L.conns, cconns.size());
    rc |= amg_alloc(c, &L.dpos, Nc);
    rc |= amg_alloc(c, &L.mptr, Nc + 1);
    rc |= amg_alloc(c, &L.midx, N);
    rc |= amg_alloc(c, &L.agg, N);
    if (!A->smoothed) {
      rc |= amg_alloc(c, &L.gptr, gptr.size());
      rc |= amg_alloc(c, &L.gidx, gidx.size());
    }
    if (rc) return -3;
    LPMG_HIP(copy_sync(c, L.cptr, ccptr.data(), sizeof(int) * (Nc + 1), hipMemcpyHostToDevice));
    LPMG_HIP(copy_sync(c, L.conns, cconns.data(), sizeof(int) * cconns.size(), hipMemcpyHostToDevice));
    LPMG_HIP(copy_sync(c, L.dpos, dpos.data(), sizeof(int) * Nc, hipMemcpyHostToDevice));
    LPMG_HIP(copy_sync(c, L.mptr, mptr.data(), sizeof(int) * (Nc + 1), hipMemcpyHostToDevice));
    LPMG_HIP(copy_sync(c, L.midx, midx.data(), sizeof(int) * N, hipMemcpyHostToDevice));
    LPMG_HIP(copy_sync(c, L.agg, agg.data(), sizeof(int) * N, hipMemcpyHostToDevice));
    if (!A->smoothed) {
      LPMG_HIP(copy_sync(c, L.gptr, gptr.data(), sizeof(int64_t) * gptr.size(), hipMemcpyHostToDevice));
      LPMG_HIP(copy_sync(c, L.gidx, gidx.data(), sizeof(int) * gidx.size(), hipMemcpyHostToDevice));
    } else {
      L.nP = (int)sa.pcol.size();
      L.nY = sa.nY;
      L.sa_gp = sa_outputs_per_wave((int64_t)sa.atidx.size(), L.nP);
      L.pmax = sa.pmax;
      L.ymax = sa.ymax;
      L.cmax = sa.cmax;
      if (!sa.device_lists) {
        const std::pair<int**, const std::vector<int>*> ups[] = {
            {&L.yptr, &sa.yptr}, {&L.yso, &sa.yso}, {&L.cso, &sa.cso}};
        for (const auto& u : ups)
          if (int r2 = sa_upload(c, u.first, *u.second)) return r2;
        if (int r2 = sa_upload(c, &L.ysp, sa.ysp)) return r2;
        if (int r2 = sa_upload(c, &L.csp, sa.csp)) return r2;
        if (int r2 = sa_upload(c, &L.ys, sa.ys)) return r2;
        if (int r2 = sa_upload(c, &L.cs, sa.cs)) return r2;
      }
    }
    A->lv.push_back(L);
    // next level: the aggregates' centroids and pattern
    std::vector<double> cen(3 * (size_t)Nc, 0.0);
    for (int a = 0; a < Nc; ++a) {
      const int m = mptr[a + 1] - mptr[a];
      for (int q = mptr[a]; q < mptr[a + 1]; ++q)
        for (int k = 0; k < 3; ++k) cen[3 * a + k] += pos[3 * midx[q] + k];
      for (int k = 0; k < 3; ++k) cen[3 * a + k] /= m;
    }
    pos.swap(cen);
    cptr.swap(ccptr);
    conns.swap(cconns);
    if (split) coarse_edges(agg, Nc, edges);
    N = Nc;
    h *= 3.5;
  }
  if ((int64_t)N * D > kCoarsestMax) {
    set_error("multigrid: the coarsest level is too large");
    return -1;
  }
  A->nc = N * D;
  // smoother damping per level (2/3 until a smoothed hierarchy's estimate
  // replaces it) and the estimate's buffers
  const int L = (int)A->lv.size();
  if (L > 0) {
    if (amg_alloc(c, &A->omega, L) || amg_alloc(c, &A->omega_p, L)) return -3;
    const std::vector<double> om(L, kOmega);
    LPMG_HIP(copy_sync(c, A->omega, om.data(), sizeof(double) * L, hipMemcpyHostToDevice));
    LPMG_HIP(copy_sync(c, A->omega_p, om.data(), sizeof(double) * L, hipMemcpyHostToDevice));
    if (A->smoothed) {
      // partials of the largest level's estimate (the levels' estimates run in turn)
      int lpr = 64, u = 1, gmax = 0;
      fine_shape(c, &lpr, &u);
      gmax = rows_grid(c->N, lpr);
      for (int l = 1; l < L; ++l) gmax = std::max(gmax, rows_grid(A->lv[l - 1].N, A->lv[l - 1].lpr));
      A->lz_nparts = gmax + 2 * kMaxGrid;
      if (amg_alloc(c, &A->lz_part, A->lz_nparts) || amg_alloc(c, &A->lz_coef, 2 * kLzSteps * L)) return -3;
    }
  }
  A->ok = true;
  return 0;
}

template <typename T>
constexpr int prec_of() { return sizeof(T) == 4 ? 1 : 0; }

template <typename T>
int amg_alloc_prec(Context* c) {
  Amg* A = c->amg;
  const int P = prec_of<T>();
  const int D = c->dim;
  if (P == 1)  // the FP32 levels are rounded from FP64 ones
    if (int rc = amg_alloc_prec<double>(c)) return rc;
  for (auto& L : A->lv) {
    if (A->smoothed && !L.pval[P]) {
      T *pv = nullptr, *pvt = nullptr;
      if (amg_alloc(c, &pv, (size_t)L.nP * D * D) || amg_alloc(c, &pvt, (size_t)L.nP * D * D)) return -3;
      L.pval[P] = pv;
      L.pvalT[P] = pvt;
      if (P == 0 && amg_alloc(c, &L.yval, (size_t)L.nY * D * D)) return -3;
    }
    if (L.val[P]) continue;
    T *v = nullptr, *di = nullptr, *w[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    int rc = amg_alloc(c, &v, (size_t)L.nnzb * D * D) | amg_alloc(c, &di, (size_t)L.N * D);
    for (auto& q : w) rc |= amg_alloc(c, &q, (size_t)L.N * D);
    if (rc) return -3;
    L.val[P] = v;
    L.dinv[P] = di;
    for (int k = 0; k < 5; ++k) L.vec[P][k] = w[k];
  }
  if (P == 1 && !A->fine16)
    if (amg_alloc(c, &A->fine16, (size_t)c->nnzb * bf16_stride(D))) return -3;
  if (!A->inv[P]) {
    T *iv = nullptr, *mk = nullptr;
    if (amg_alloc(c, &iv, (size_t)A->nc * A->nc) || amg_alloc(c, &mk, (size_t)c->N * D)) return -3;
    A->inv[P] = iv;
    A->mask[P] = mk;
  }
  return 0;
}

// Level values from the current FP64 stiffness.  The coarse operators are
// block sums whose entries cancel more at every level (their rows add up to
// nearly nothing for the rigid motions), so they are always summed from the
// FP64 values and only then rounded for the FP32 solve.
template <typename T>
__global__ void amg_round_kernel(const double* __restrict__ a, T* __restrict__ b, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = (T)a[t];
}

// Level l's smoother damping and prolongator smoothing factor (smoothed
// hierarchy, levels 0..L-1) from kLzSteps Lanczos steps on its FP64 values,
// enqueued before its prolongator is formed; the vectors used are free while
// the values are rebuilt (level 0: the FP64 CG's, coarser levels: their own
// FP64 work vectors).  A level's launches (two per step) are captured into a
// graph on first use and replayed after every later value rebuild (eager, the
// launches of all levels took 1.3 ms per rebuild on the 205,645-particle CT
// specimen).
template <int D>
int amg_damping_level(Context* c, int l) {
  Amg* A = c->amg;
  const AmgLevel* B = l == 0 ? nullptr : &A->lv[l - 1];
  const int N = l == 0 ? c->N : B->N;
  const int* cptr = l == 0 ? c->cptr : B->cptr;
  const int* conns = l == 0 ? c->conns : B->conns;
  const int* dpos = l == 0 ? c->dpos : B->dpos;
  const double* val = l == 0 ? c->Kval : (const double*)B->val[0];
  double* t = l == 0 ? c->cg_p : (double*)B->vec[0][0];
  double* y = l == 0 ? c->cg_w : (double*)B->vec[0][1];
  double* u[2] = {l == 0 ? c->cg_z : (double*)B->vec[0][2], l == 0 ? c->cg_r : (double*)B->vec[0][3]};
  int lpr = 64, uc = 1;
  if (l == 0) fine_shape(c, &lpr, &uc);
  else lpr = B->lpr;
  const int g = rows_grid(N, lpr);
  const int64_t n = (int64_t)D * N;
  const int gv = grid_n(n, kBS * 4);
  if (g + 2 * kMaxGrid > A->lz_nparts) {
    set_error("multigrid: damping estimate partials buffer too small");
    return -1;
  }
  double* pa = A->lz_part;
  double* pb[2] = {pa + g, pa + g + kMaxGrid};
  double* coef = A->lz_coef + 2 * kLzSteps * l;
  std::vector<uintptr_t> sig;
  for (const void* q : {(const void*)cptr, (const void*)conns, (const void*)dpos, (const void*)val, (const void*)t,
                        (const void*)y, (const void*)u[0], (const void*)u[1], (const void*)A->lz_part,
                        (const void*)A->lz_coef, (const void*)A->omega, (const void*)A->omega_p})
    graph_sig_push(&sig, q);
  for (int64_t x : {(int64_t)N, (int64_t)lpr, (int64_t)uc}) graph_sig_push(&sig, x);
  if ((int)A->lz_graphs.size() <= l) {
    A->lz_graphs.resize(l + 1, nullptr);
    A->lz_sigs.resize(l + 1);
  }
  if (A->lz_graphs[l]) {
    if (int rc = check_graph_signature(A->lz_sigs[l], sig, "multigrid damping estimate")) return rc;
  } else {
    hipGraph_t graph = nullptr;
    LPMG_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    lz_init_kernel<D><<<gv, kBS, 0, c->stream>>>(dpos, val, n, t, pb[0]);
    for (int k = 0; k < kLzSteps; ++k) {
      with_shape(lpr, uc, [&](auto LP, auto UC) {
        lz_spmv_kernel<D, decltype(LP)::value, decltype(UC)::value>
            <<<g, kSpBS, 0, c->stream>>>(cptr, conns, val, t, y, N, pa);
      });
      lz_update_kernel<D><<<gv, kBS, 0, c->stream>>>(dpos, val, n, k, pb[k & 1], gv, pa, g, y, t, u[(k + 1) & 1],
                                                     u[k & 1], pb[(k + 1) & 1], coef);
    }
    lz_omega_kernel<<<1, 64, 0, c->stream>>>(coef, A->omega + l, A->omega_p + l);
    const hipError_t e1 = hipGetLastError();
    const hipError_t e2 = hipStreamEndCapture(c->stream, &graph);
    LPMG_HIP(e1);
    LPMG_HIP(e2);
    LPMG_HIP(hipGraphInstantiate(&A->lz_graphs[l], graph, nullptr, nullptr, 0));
    LPMG_HIP(hipGraphDestroy(graph));
    A->lz_sigs[l] = sig;
  }
  LPMG_HIP(hipGraphLaunch(A->lz_graphs[l], c->stream));
  return 0;
}

template <int D, typename T>
int amg_values(Context* c) {
  Amg* A = c->amg;
  const int P = prec_of<T>();
  const double* fine = c->Kval;
  const int* fdpos = c->dpos;
  const int *fcptr = c->cptr, *fconns = c->conns;
  int fN = c->N;
  for (int l = 0; l < (int)A->lv.size(); ++l) {
    AmgLevel& L = A->lv[l];
    const int64_t nt = L.nnzb * D * D;
    const unsigned gb = (unsigned)((nt + 255) / 256);
    if (A->smoothed) {
      if (int rc = amg_damping_level<D>(c, l)) return rc;
      sa_p_kernel<D><<<(unsigned)((L.nP + L.sa_gp - 1) / L.sa_gp), 64, 0, c->stream>>>(
          L.atptr, L.atidx, fine, L.prow, L.pcol, L.agg, fdpos, L.qpos, L.nP, L.sa_gp, A->omega_p + l,
          (double*)L.pval[0],
          (double*)L.pvalT[0]);
      // Y = A P over the level's rows (sources: A's blocks, runs: P's rows)
      launch_sa_rows<D, 1>(c->stream, fN, L.pmax, L.ymax, fcptr, fconns, fine, L.pptr, (const double*)L.pval[0],
                           L.ysp, L.yso, L.ys, L.yptr, L.yval);
      // coarse blocks P^T Y over the aggregates (sources: P's columns, runs: Y's rows)
      launch_sa_rows<D, 2>(c->stream, L.N, L.ymax, L.cmax, L.qptr, L.qrow, (const double*)L.pvalT[0], L.yptr,
                           L.yval, L.csp, L.cso, L.cs, L.cptr, (double*)L.val[0]);
      if (P) {
        const int64_t np = (int64_t)L.nP * D * D;
        amg_round_kernel<T><<<(unsigned)((np + 255) / 256), 256, 0, c->stream>>>((const double*)L.pval[0],
                                                                                (T*)L.pval[P], np);
        amg_round_kernel<T><<<(unsigned)((np + 255) / 256), 256, 0, c->stream>>>((const double*)L.pvalT[0],
                                                                                (T*)L.pvalT[P], np);
      }
    } else {
      amg_gather_kernel<double, double, D * D><<<gb, 256, 0, c->stream>>>(L.gptr, L.gidx, L.nnzb, fine,
                                                                         (double*)L.val[0]);
    }
    fdpos = L.dpos;
    fcptr = L.cptr;
    fconns = L.conns;
    fN = L.N;
    if (P) amg_round_kernel<T><<<gb, 256, 0, c->stream>>>((const double*)L.val[0], (T*)L.val[P], nt);
    amg_dinv_kernel<T, D><<<(L.N + 255) / 256, 256, 0, c->stream>>>(L.dpos, (const T*)L.val[P], L.N, (T*)L.dinv[P]);
    fine = (const double*)L.val[0];
  }
  if (A->lv.empty())
    amg_invert_kernel<double, T, D><<<1, 1024, 0, c->stream>>>(c->cptr, c->conns, c->Kval, c->N, (T*)A->inv[P]);
  else {
    const AmgLevel& L = A->lv.back();
    amg_invert_kernel<double, T, D><<<1, 1024, 0, c->stream>>>(L.cptr, L.conns, (const double*)L.val[0], L.N,
                                                              (T*)A->inv[P]);
  }
  LPMG_HIP(hipGetLastError());
  return 0;
}

// Enqueue z = M b at level l (l = 0: the handle's operator, b = the CG
// residual, z written with partials of (b, z) into rz_part).
template <int D, typename T>
// x1: the smoothed hierarchy's om D^-1 b at level 0 (written with b by the
// caller; coarser levels take theirs from the restriction), else unused.
void vcycle(Context* c, hipStream_t st, int l, const T* b, T* out, T* scratch_r, double* rz_part, const T* x1) {
  Amg* A = c->amg;
  const int P = prec_of<T>();
  const int L = (int)A->lv.size();
  const double* om = A->omega + l;  // this level's smoother damping (device memory)
  if (l == L) {  // coarsest (the eliminated DOFs of a single-level system
                 // decouple in the exact inverse: z = 0 there already)
    amg_coarsest_kernel<T><<<1, kBS, 0, st>>>((const T*)A->inv[P], A->nc, b, out, l == 0 ? rz_part : nullptr, c->sc);
    return;
  }
  const int N = l == 0 ? c->N : A->lv[l - 1].N;
  const int* cptr = l == 0 ? c->cptr : A->lv[l - 1].cptr;
  const int* conns = l == 0 ? c->conns : A->lv[l - 1].conns;
  const T* val = l == 0 ? (const T*)(P ? (const void*)c->Kval32 : (const void*)c->Kval) : (const T*)A->lv[l - 1].val[P];
  const T* dinv = l == 0 ? (const T*)(P ? (const void*)c->f_dinv : (const void*)c->dinv) : (const T*)A->lv[l - 1].dinv[P];
  int lpr = 64, u = 1;
  if (l == 0) fine_shape(c, &lpr, &u);
  else lpr = A->lv[l - 1].lpr;
  T* r = l == 0 ? scratch_r : (T*)A->lv[l - 1].vec[P][1];
  const AmgLevel& C = A->lv[l];  // the next level
  T* bc = (T*)C.vec[P][0];
  T* xc = (l + 1 == L) ? (T*)C.vec[P][2] : (T*)C.vec[P][3];
  const int g = rows_grid(N, lpr);
  // FP32: the finest level's smoothers read the bfloat16 copy of the values
  const uint16_t* v16 = nullptr;
  if constexpr (std::is_same<T, float>::value) v16 = l == 0 ? A->fine16 : nullptr;
  const T* mask = l == 0 ? (const T*)A->mask[P] : nullptr;
  if (A->smoothed) {
    // the smoothed hierarchy: x1 = om D^-1 b came with b (from the CG's
    // vector kernel at the finest level, from the restriction below it)
    if (l > 0) x1 = (const T*)A->lv[l - 1].vec[P][4];
    with_shape(lpr, u, [&](auto LP, auto UC) {
      if constexpr (std::is_same<T, float>::value) {
        if (v16) {
          amg_presmooth_x1_kernel<D, T, decltype(LP)::value, decltype(UC)::value, uint16_t>
              <<<g, kSpBS, 0, st>>>(cptr, conns, v16, N, b, x1, r, c->sc);
          return;
        }
      }
      amg_presmooth_x1_kernel<D, T, decltype(LP)::value, decltype(UC)::value>
          <<<g, kSpBS, 0, st>>>(cptr, conns, val, N, b, x1, r, c->sc);
    });
    // restriction P^T r (with the coarse level's x1 when it has a smoother),
    // the coarse cycle, prolongation x = x1 + P xc (into r, which is dead by
    // then), post-smoothing from x
    const bool coarse_smoothed = l + 1 < L;
    sa_restrict_kernel<D, T><<<(C.N + 3) / 4, 256, 0, st>>>(
        C.qptr, C.qrow, (const T*)C.pvalT[P], C.N, r, bc, (const T*)C.dinv[P], A->omega + l + 1,
        coarse_smoothed ? (T*)C.vec[P][4] : nullptr, c->sc);
    vcycle<D, T>(c, st, l + 1, bc, xc, nullptr, nullptr, nullptr);
    sa_prolong_kernel<D, T><<<(N + 255) / 256, 256, 0, st>>>(C.pptr, C.pcol, (const T*)C.pval[P], N, x1, xc, r,
                                                             c->sc);
    with_shape(lpr, u, [&](auto LP, auto UC) {
      if constexpr (std::is_same<T, float>::value) {
        if (v16) {
          amg_postsmooth_x_kernel<D, T, decltype(LP)::value, decltype(UC)::value, uint16_t>
              <<<g, kSpBS, 0, st>>>(cptr, conns, v16, N, b, r, dinv, om, mask, out, l == 0 ? rz_part : nullptr,
                                    c->sc);
          return;
        }
      }
      amg_postsmooth_x_kernel<D, T, decltype(LP)::value, decltype(UC)::value>
          <<<g, kSpBS, 0, st>>>(cptr, conns, val, N, b, r, dinv, om, mask, out, l == 0 ? rz_part : nullptr, c->sc);
    });
    return;
  }
  // the tentative hierarchy: pre-smoothing forms om D^-1 b where it reads it
  with_shape(lpr, u, [&](auto LP, auto UC) {
    if constexpr (std::is_same<T, float>::value) {
      if (v16) {
        amg_presmooth_resid_kernel<D, T, decltype(LP)::value, decltype(UC)::value, uint16_t>
            <<<g, kSpBS, 0, st>>>(cptr, conns, v16, N, b, dinv, om, r, c->sc);
        return;
      }
    }
    amg_presmooth_resid_kernel<D, T, decltype(LP)::value, decltype(UC)::value>
        <<<g, kSpBS, 0, st>>>(cptr, conns, val, N, b, dinv, om, r, c->sc);
  });
  amg_restrict_kernel<D, T><<<(C.N + 3) / 4, 256, 0, st>>>(C.mptr, C.midx, C.N, r, bc, c->sc);
  vcycle<D, T>(c, st, l + 1, bc, xc, nullptr, nullptr, nullptr);
  with_shape(lpr, u, [&](auto LP, auto UC) {
    if constexpr (std::is_same<T, float>::value) {
      if (v16) {
        amg_postsmooth_kernel<D, T, decltype(LP)::value, decltype(UC)::value, uint16_t>
            <<<g, kSpBS, 0, st>>>(cptr, conns, v16, N, b, C.agg, xc, dinv, om, mask, out,
                                  l == 0 ? rz_part : nullptr, c->sc);
        return;
      }
    }
    amg_postsmooth_kernel<D, T, decltype(LP)::value, decltype(UC)::value>
        <<<g, kSpBS, 0, st>>>(cptr, conns, val, N, b, C.agg, xc, dinv, om, mask, out, l == 0 ? rz_part : nullptr,
                              c->sc);
  });
}

}  // namespace

void amg_step_begins(Context* c) {
  if (c->amg && c->amg->smoothed) c->amg->recheck = true;
}

void amg_free(Context* c) {
  if (!c->amg) return;
  for (auto& g : c->amg->graph)
    if (g) (void)hipGraphExecDestroy(g);
  for (auto& g : c->amg->graph_tail)
    if (g) (void)hipGraphExecDestroy(g);
  for (auto& g : c->amg->lz_graphs)
    if (g) (void)hipGraphExecDestroy(g);
  for (void* q : c->amg->allocs) (void)hipFreeAsync(q, c->stream);  // after the work queued on them
  c->bytes -= c->amg->bytes;
  delete c->amg;
  c->amg = nullptr;
}

// PCG with the multigrid preconditioner on the context's stream; x = du (FP64)
// or f_x (FP32), rhs = residual.
template <int D, typename T>
int run_pcg_amg(Context* c, const SolveOpts& o, int* iters, double* rr_final) {
  const bool smoothed = o.precond == 4;
  if (c->amg && c->amg->smoothed != smoothed) amg_free(c);  // the other kind of hierarchy
  // at a load step's first solve, when bonds broke or healed since the last
  // check: rebuild a smoothed hierarchy whose finest aggregates no longer
  // follow the intact bonds
  if (c->amg && c->amg->smoothed && c->amg->recheck) {
    c->amg->recheck = false;
    if (c->amg->break_version != c->break_version) {
      bool stale = false;
      if (int rc = amg_split_stale(c, &stale)) return rc;
      c->amg->break_version = c->break_version;
      if (stale) amg_free(c);
    }
  }
  if (!c->amg) {
    c->amg = new Amg;
    c->amg->smoothed = smoothed;
    if (int rc = amg_build(c)) {
      amg_free(c);
      return rc;
    }
  }
  Amg* A = c->amg;
  const int P = prec_of<T>();
  if (int rc = amg_alloc_prec<T>(c)) return rc;
  const T* val0 = (const T*)(P ? (const void*)c->Kval32 : (const void*)c->Kval);
  T* dinv0 = (T*)(P ? (void*)c->f_dinv : (void*)c->dinv);
  T *x, *r, *z, *p, *w;
  if (P) {
    x = (T*)c->f_x; r = (T*)c->f_r; z = (T*)c->f_z; p = (T*)c->f_p; w = (T*)c->f_w;
  } else {
    x = (T*)c->du; r = (T*)c->cg_r; z = (T*)c->cg_z; p = (T*)c->cg_p; w = (T*)c->cg_w;
  }
  const int N = c->N;
  const int64_t n = (int64_t)D * N;
  amg_dinv0_kernel<D, T><<<(N + 255) / 256, 256, 0, c->stream>>>(c->dpos, val0, c->cons, c->damage_visual,
                                                                 c->floating, N, dinv0, (T*)A->mask[P]);
  if (A->version[P] != c->asm_version) {
    // inside a load step the hierarchy's values from the step's first solve
    // serve its later damage passes: a preconditioner from a slightly older
    // stiffness (the CG's operator and the finest smoother stay current)
    if (!(c->in_step && A->values_step[P] == c->step_serial)) {
      if (int rc = amg_values<D, T>(c)) return rc;
      A->version[P] = c->asm_version;
      A->values_step[P] = c->in_step ? c->step_serial : -1;
    }
    if (P == 1 && A->version16 != c->asm_version) {
      const int64_t nv = (int64_t)c->nnzb * bf16_stride(D);
      bf16_round_kernel<D><<<(unsigned)((nv + 255) / 256), 256, 0, c->stream>>>(c->Kval32, A->fine16, c->nnzb);
      LPMG_HIP(hipGetLastError());
      A->version16 = c->asm_version;
    }
  }
  int lpr = 64, u = 1;
  fine_shape(c, &lpr, &u);
  const int gs = rows_grid(N, lpr);
  const int gv = grid_n(n, kBS * 4);
  const bool two = gs > kMaxParts;
  const bool coarse0 = A->lv.empty();  // the whole system is the coarsest level
  // partial buffers: (r,r) [gv], (p,w) [gs (+kBS)], (r,z) [gs (+kBS)]
  double* rr_part = c->partials;
  double* pw_raw = c->partials + 2 * kMaxGrid;
  double* pw2 = pw_raw + gs;
  double* rz_raw = pw2 + kBS;
  double* rz2 = rz_raw + gs;
  if ((int64_t)(2 * kMaxGrid + 2 * (gs + kBS)) > 2 * (int64_t)c->n_partials) {
    set_error("partials buffer too small for the multigrid CG");
    return -1;
  }
  const int nrz = coarse0 ? 1 : (two ? kBS : gs);
  const int npw = two ? kBS : gs;
  const int64_t maxit = o.maxit > 0 ? o.maxit : n;
  // the smoothed hierarchy's finest pre-smoothed iterate x1 = om D^-1 r,
  // written with r (the multigrid PCG does not use the q vector)
  T* x1 = (A->smoothed && !coarse0) ? (T*)(P ? (void*)c->f_q : (void*)c->cg_q) : nullptr;
  amg_setup_scalars_kernel<<<1, 64, 0, c->stream>>>(c->sc, o.rtol, o.atol, maxit);
  pcg_init_kernel<T><<<gv, kBS, 0, c->stream>>>(c->residual, x, r, n, rr_part, dinv0, A->omega, x1);
  vcycle<D, T>(c, c->stream, 0, r, z, w, rz_raw, x1);
  if (two && !coarse0) pcg_reduce_kernel<<<kBS, kBS, 0, c->stream>>>(rz_raw, gs, rz2, c->sc);
  pcg_init2_kernel<T><<<gv, kBS, 0, c->stream>>>(p, z, n, (two && !coarse0) ? rz2 : rz_raw, nrz, rr_part, gv, c->sc);
  LPMG_HIP(hipGetLastError());
  auto enqueue_chunk = [&](int iters) -> int {
    for (int k = 0; k < iters; ++k) {
      with_shape(lpr, u, [&](auto LP, auto UC) {
        pcg_spmv_kernel<D, T, decltype(LP)::value, decltype(UC)::value>
            <<<gs, kSpBS, 0, c->stream>>>(c->cptr, c->conns, val0, p, w, N, pw_raw, c->sc);
      });
      if (two) pcg_reduce_kernel<<<kBS, kBS, 0, c->stream>>>(pw_raw, gs, pw2, c->sc);
      pcg_vec1_kernel<T><<<gv, kBS, 0, c->stream>>>(x, r, p, w, n, two ? pw2 : pw_raw, npw, rr_part, dinv0, A->omega,
                                                    x1, c->sc);
      vcycle<D, T>(c, c->stream, 0, r, z, w, rz_raw, x1);
      if (two && !coarse0) pcg_reduce_kernel<<<kBS, kBS, 0, c->stream>>>(rz_raw, gs, rz2, c->sc);
      pcg_vec2_kernel<T><<<gv, kBS, 0, c->stream>>>(p, z, n, (two && !coarse0) ? rz2 : rz_raw, nrz, rr_part, gv,
                                                   c->sc);
    }
    LPMG_HIP(hipGetLastError());
    return 0;
  };
  std::vector<uintptr_t> sig;
  for (const void* q : {(const void*)c->cptr, (const void*)c->conns, (const void*)val0, (const void*)dinv0,
                        (const void*)x, (const void*)r, (const void*)z, (const void*)p, (const void*)w,
                        (const void*)c->partials, (const void*)c->sc, (const void*)A->inv[P],
                        (const void*)A->mask[P], (const void*)A->fine16, (const void*)A->omega,
                        (const void*)x1})
    graph_sig_push(&sig, q);
  for (const auto& L : A->lv)
    for (const void* q : {(const void*)L.cptr, (const void*)L.conns, (const void*)L.mptr, (const void*)L.midx,
                          (const void*)L.agg, (const void*)L.val[P], (const void*)L.dinv[P],
                          (const void*)L.vec[P][0], (const void*)L.vec[P][1], (const void*)L.vec[P][2],
                          (const void*)L.vec[P][3], (const void*)L.vec[P][4], (const void*)L.pptr,
                          (const void*)L.pcol, (const void*)L.qrow,
                          (const void*)L.qptr, (const void*)L.pvalT[P], (const void*)L.pval[P]})
      graph_sig_push(&sig, q);
  graph_sig_push(&sig, (int64_t)A->smoothed);
  for (int64_t v : {(int64_t)N, (int64_t)gs, (int64_t)gv, (int64_t)lpr, (int64_t)u, (int64_t)two, (int64_t)A->nc})
    graph_sig_push(&sig, v);
  // two captured chunks: kCgChunk iterations, and kTailChunk for the end of
  // a solve -- once the iterations launched pass the previous solve's count
  // on this handle, short chunks follow, so that fewer iterations past
  // convergence are launched (each costs its launches though every kernel
  // returns at once); chunks of either length run the same iterations
  auto graph_for = [&](int which) -> hipGraphExec_t* {
    hipGraphExec_t* g = which ? &A->graph_tail[P] : &A->graph[P];
    std::vector<uintptr_t>& gs_sig = which ? A->graph_tail_sig[P] : A->graph_sig[P];
    if (*g) {
      if (check_graph_signature(gs_sig, sig, "multigrid PCG chunk")) return nullptr;
      return g;
    }
    hipGraph_t graph = nullptr;
    if (hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return nullptr;
    const int rc = enqueue_chunk(which ? kTailChunk : kCgChunk);
    const hipError_t e2 = hipStreamEndCapture(c->stream, &graph);
    if (rc || e2 != hipSuccess) {
      if (!rc) hip_fail(e2, "hipStreamEndCapture");
      return nullptr;
    }
    const hipError_t e3 = hipGraphInstantiate(g, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e3 != hipSuccess) {
      hip_fail(e3, "hipGraphInstantiate");
      *g = nullptr;
      return nullptr;
    }
    gs_sig = sig;
    return g;
  };
  const int pred = A->last_iters[P];
  for (int launched = 0;;) {
    const bool tail = pred > 0 && launched + kCgChunk > pred + kTailChunk && launched < pred + 2 * kCgChunk;
    hipGraphExec_t* g = graph_for(tail ? 1 : 0);
    if (!g) return -1;
    LPMG_HIP(hipGraphLaunch(*g, c->stream));
    launched += tail ? kTailChunk : kCgChunk;
    LPMG_HIP(hipMemcpyAsync(c->sc_host, c->sc, sizeof(Scalars), hipMemcpyDeviceToHost, c->stream));
    LPMG_HIP(hipStreamSynchronize(c->stream));
    if (c->sc_host->done) break;
  }
  A->last_iters[P] = c->sc_host->iter;
  *iters = c->sc_host->iter;
  *rr_final = c->sc_host->rr;
  return 0;
}

template int run_pcg_amg<2, double>(Context*, const SolveOpts&, int*, double*);
template int run_pcg_amg<3, double>(Context*, const SolveOpts&, int*, double*);
template int run_pcg_amg<2, float>(Context*, const SolveOpts&, int*, double*);
template int run_pcg_amg<3, float>(Context*, const SolveOpts&, int*, double*);

int amg_levels(Context* c, int* out, int cap) {
  if (!c->amg || !c->amg->ok) return 0;
  int k = 0;
  if (k < cap) out[k++] = c->N * c->dim;
  for (auto& L : c->amg->lv)
    if (k < cap) out[k++] = L.N * c->dim;
  return k;
}

}  // namespace lpmg
