// solver.hip -- block-CSR SpMV and the (Jacobi-preconditioned) CG that replaces
// MKL's RCI CG (Solver::LPM_CG, solver.h:276-354) and PARDISO
// (Solver::LPM_PARDISO, solver.h:190-274) as the linear solve of the Newton step.
//
// SpMV: one lane per b x b block.  A row of up to 64 blocks (every lattice here)
// is owned by 16 lanes, four rows per wavefront, and a pass issues up to four
// 16-block chunks of the row: each lane loads its block's b*b values (a
// contiguous 72/32-byte run, the row's loads together one contiguous segment),
// its block column and the b entries of x, multiplies, and the row's b sums are
// reduced across the row's lanes.  All loads are unconditional straight-line
// code (indices past the row's end are clamped and their products dropped), so
// a row's loads are in flight together.  Measured on the SC 40^3 pattern this
// streams at 6.3 TB/s back to back (FP64), 79% of the 8 TB/s HBM peak.  The
// operator is sparse, so MFMA is not used.
//
// CG is the Chronopoulos-Gear single-reduction form of preconditioned CG: per
// iteration one SpMV kernel (w = A u, with the partials of (w,u)) and one fused
// vector kernel (p = u + beta p; s = w + beta s; x += alpha p; r -= alpha s;
// u = M r; partials of (r,u) and (r,r)).  No kernel waits on other workgroups of
// its own launch: each kernel writes one partial per workgroup and a later
// kernel sums them in a fixed order (deterministic) -- the SpMV's partials in
// every workgroup of the vector step, the vector step's in the next SpMV's
// workgroup 0, which leaves the two totals for the next vector step -- so no
// intra-kernel fence is ever issued (on gfx950 an agent-scope fence writes back
// and invalidates the XCD's L2).  Scalars that live across iterations are written
// by workgroup 0 into slots indexed by the iteration parity.  The stopping test
// is the one the reference configures MKL with: ||r||^2 <= rtol*||r0||^2 + atol
// (dpar[0], dpar[1], solver.h:312-313), at most maxit (ipar[4] = n) iterations.
// The host launches `chunk` iterations between reads of the convergence flag;
// kernels of iterations after convergence return immediately.
//
// The same iteration bodies also run as fused launches over many systems (the
// serving runtime): one launch per kernel kind covers every system of a batch
// class (dimension, precision, chunks per pass); workgroup g finds its system by
// a binary search over the systems' first workgroups and then computes exactly
// what that system's own launch would compute in its workgroup g - first, so a
// result does not depend on the batch it ran in.
#include <algorithm>
#include <vector>

#include <hip/hip_ext.h>

#include "bsr.h"

namespace lpmg {

namespace {

// CG vector step: one element per thread, 4 wavefronts per workgroup.  Every
// workgroup also sums the SpMV's partials, but small systems need the
// workgroups to fill the CUs: per Jacobi-PCG iteration, 1024 / 512 / 256 /
// 128 threads took 53.1 / 52.7 / 52.1 / 52.3 us on SC 40^3 (FP64) and
// 12.4 / 10.3 / 9.9 / 10.0 us on a 20k-particle CT plate (FP32), when each
// workgroup still summed the vector step's own partials too.  With those
// summed once by the SpMV's workgroup 0 and u no longer read: 51.1 us (FP64)
// and 30.0 us (FP32) per iteration on SC 40^3, from 52.2 and 31.2 in the same
// run (fixed 1,600-iteration solves); the vector step is 7.0 us of it, about
// 4 us of which an empty launch of the same shape also takes.
constexpr int kVBS = 256;

template <int D, typename T, int U>
__global__ void __launch_bounds__(kSpBS) spmv_kernel(const int* __restrict__ cptr, const int* __restrict__ conns,
                                                     const T* __restrict__ val, const T* __restrict__ x,
                                                     T* __restrict__ y, int N) {
  bsr_rows<D, T, kLPR, U, false>(blockIdx.x, cptr, conns, val, x, y, N);
}

// ---- one CG iteration, written for workgroup `blk` of `nblk` of one system.
// The same bodies run in the per-system launches and in the fused launches
// over several systems, so a system's arithmetic does not depend on which
// other systems share its launches.

// SpMV: w = A u and the per-workgroup partials of (w, u).  `done` is written
// only by the vector step, never during this launch.  Workgroup 0 also sums
// the last vector step's nV partials of (r,u), (r,r) into sc->vsum, so the
// next vector step reads two totals instead of every partial (one memory
// round trip less in each of its workgroups); nothing else of this launch
// reads vsum.
template <int D, typename T, int U>
__device__ __forceinline__ void cg_spmv_body(int blk, const int* __restrict__ cptr, const int* __restrict__ conns,
                                             const T* __restrict__ val, const T* __restrict__ u,
                                             T* __restrict__ w, int N, double* __restrict__ partS,
                                             const double* __restrict__ partV, int nV, Scalars* sc) {
  if (sc->done) return;
  if (blk == 0) {
    if (threadIdx.x == 0) sc->icur = sc->inext;  // iteration index of the next vector step
    const Vec<2> v = sum_partials<kSpBS, 2>(partV, nV);
    if (threadIdx.x == 0) {
      sc->vsum[0] = v.v[0];
      sc->vsum[1] = v.v[1];
    }
  }
  Vec<1> a;
  a.v[0] = bsr_rows<D, T, kLPR, U, true>(blk, cptr, conns, val, u, w, N);
  write_partial_at<kSpBS, 1>(partS, blk, a);
}

// Second level for large grids: n partials -> nblk partials (contiguous slices)
__device__ __forceinline__ void reduce_partials_body(int blk, int nblk, const double* __restrict__ in, int n,
                                                     double* __restrict__ out, const Scalars* __restrict__ sc) {
  if (sc->done) return;
  const int per = (n + nblk - 1) / nblk;
  const int lo = blk * per, hi = min(n, lo + per);
  Vec<1> a;
  a.v[0] = 0;
  for (int b = lo + threadIdx.x; b < hi; b += kBS) a.v[0] += in[b];
  write_partial_at<kBS, 1>(out, blk, a);
}

// The scalars of iteration i from the partials: gamma = (r,u), delta = (w,u), rr = (r,r)
struct CGStep {
  double gamma, delta, rr, alpha, beta;
  bool stop;
};

// gamma = (r,u), rr = (r,r), delta = (w,u) summed over the workgroup give the
// step's coefficients and stop decision.
__device__ __forceinline__ CGStep cg_step(const Vec<3>& a, const Scalars* sc, int i) {
  CGStep st;
  st.gamma = a.v[0];
  st.rr = a.v[1];
  st.delta = a.v[2];
  if (i == 0) {
    st.beta = 0;
    st.alpha = st.gamma / st.delta;
    st.stop = st.rr <= sc->rtol * st.rr + sc->atol || sc->maxit <= 0;
  } else {
    const double gamma_old = sc->gam[(i - 1) & 1], alpha_old = sc->alp[(i - 1) & 1];
    st.beta = st.gamma / gamma_old;
    st.alpha = st.gamma / (st.delta - st.beta * st.gamma / alpha_old);
    st.stop = st.rr <= sc->thresh || i >= sc->maxit;
  }
  return st;
}

// Vector step of iteration i: p = u + beta p; s = w + beta s; x += alpha p;
// r -= alpha s; u = M r; partials of (r,u), (r,r).  Workgroup 0 sets `done`
// during this launch, so each workgroup reads the flag once, through LDS: all
// wavefronts of a workgroup take the same branch into the barriers below.  A
// workgroup that still sees 0 reaches the same stop decision from the same
// partials and returns without writing.
//
// Every load the step needs is issued before its first barrier (whose fence
// waits for them): this thread's first element, its share of the SpMV's
// partials, the two totals of (r,u), (r,r) that the SpMV's workgroup 0 formed
// and, in thread 0, the flag -- one memory latency instead of several in a
// row.  u = M r is not read: the last step stored u = dinv * r from the same r
// and dinv, so the product gives its bits again.
template <typename T>
__device__ __forceinline__ void cg_vec_body(int blk, int nblk, T* __restrict__ x, T* __restrict__ r,
                                            T* __restrict__ u, T* __restrict__ p, T* __restrict__ s,
                                            const T* __restrict__ w, const T* __restrict__ dinv, int64_t n,
                                            const double* __restrict__ partS, int nS,
                                            double* __restrict__ partV_out, Scalars* sc) {
  __shared__ int done_s;
  __shared__ double red[kVBS / 64];
  const int64_t k0 = (int64_t)blk * kVBS + threadIdx.x;
  const bool has0 = k0 < n;
  T p0 = 0, w0 = 0, s0 = 0, x0 = 0, r0 = 0, d0 = 0;
  if (has0) {
    p0 = p[k0]; w0 = w[k0]; s0 = s[k0]; x0 = x[k0]; r0 = r[k0]; d0 = dinv[k0];
  }
  const double gam = sc->vsum[0], rrt = sc->vsum[1];
  Vec<1> a;
  a.v[0] = 0;
  if (nS <= 8 * kVBS) {  // one share of the set: all loads in flight at once
    // unconditional loads (indices past the end read the last entry and are
    // not added): a load inside a branch is waited for at the branch's end
    const int b = threadIdx.x;
    double sv[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) sv[j] = partS[min(b + j * kVBS, max(nS - 1, 0))];
#pragma unroll
    for (int j = 0; j < 8; ++j)
      if (b + j * kVBS < nS) a.v[0] += sv[j];
  } else {
    add_partials<kVBS, 1>(partS, nS, a);
  }
  const int done0 = threadIdx.x == 0 ? sc->done : 0;
  a = wave_sum(a);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = a.v[0];
  if (threadIdx.x == 0) done_s = done0;
  __syncthreads();
  if (done_s) return;
  Vec<3> tot;
  tot.v[0] = gam;
  tot.v[1] = rrt;
  tot.v[2] = 0;
  for (int q = 0; q < kVBS / 64; ++q) tot.v[2] += red[q];
  const int i = sc->icur;
  const CGStep st = cg_step(tot, sc, i);
  if (blk == 0 && threadIdx.x == 0) {
    if (i == 0) {
      sc->rr0 = st.rr;
      sc->thresh = sc->rtol * st.rr + sc->atol;
    }
    sc->gam[i & 1] = st.gamma;
    sc->alp[i & 1] = st.alpha;
    sc->rz = st.gamma;
    sc->rr = st.rr;
    sc->pq = st.delta;
    sc->alpha = st.alpha;
    sc->beta = st.beta;
    sc->iter = i;
    sc->inext = i + 1;
    if (st.stop) sc->done = 1;
  }
  if (st.stop) return;  // every workgroup reaches the same decision
  const T alpha = (T)st.alpha, beta = (T)st.beta;
  double ru = 0, rr = 0;
  if (has0) {
    const T u0 = d0 * r0;
    const T pk = u0 + beta * p0;
    const T sk = w0 + beta * s0;
    p[k0] = pk;
    s[k0] = sk;
    x[k0] = x0 + alpha * pk;
    const T rk = r0 - alpha * sk;
    r[k0] = rk;
    const T uk = d0 * rk;
    u[k0] = uk;
    ru += (double)rk * (double)uk;
    rr += (double)rk * (double)rk;
  }
  for (int64_t k = k0 + (int64_t)nblk * kVBS; k < n; k += (int64_t)nblk * kVBS) {
    const T dk = dinv[k], ro = r[k];
    const T pk = dk * ro + beta * p[k];
    const T sk = w[k] + beta * s[k];
    p[k] = pk;
    s[k] = sk;
    x[k] += alpha * pk;
    const T rk = ro - alpha * sk;
    r[k] = rk;
    const T uk = dk * rk;
    u[k] = uk;
    ru += (double)rk * (double)uk;
    rr += (double)rk * (double)rk;
  }
  Vec<2> part;
  part.v[0] = ru;
  part.v[1] = rr;
  write_partial_at<kVBS, 2>(partV_out, blk, part);
}

// ---- per-system launches
template <int D, typename T, int U>
__global__ void __launch_bounds__(kSpBS) cg_spmv_kernel(const int* __restrict__ cptr, const int* __restrict__ conns,
                                                        const T* __restrict__ val, const T* __restrict__ u,
                                                        T* __restrict__ w, int N, double* __restrict__ partS,
                                                        const double* __restrict__ partV, int nV, Scalars* sc) {
  cg_spmv_body<D, T, U>(blockIdx.x, cptr, conns, val, u, w, N, partS, partV, nV, sc);
}

__global__ void __launch_bounds__(kBS) reduce_partials_kernel(const double* __restrict__ in, int n,
                                                              double* __restrict__ out,
                                                              const Scalars* __restrict__ sc) {
  reduce_partials_body(blockIdx.x, gridDim.x, in, n, out, sc);
}

template <typename T>
__global__ void __launch_bounds__(kVBS) cg_vec_kernel(T* __restrict__ x, T* __restrict__ r, T* __restrict__ u,
                                                     T* __restrict__ p, T* __restrict__ s,
                                                     const T* __restrict__ w, const T* __restrict__ dinv,
                                                     int64_t n, const double* __restrict__ partS, int nS,
                                                     double* __restrict__ partV_out, Scalars* sc) {
  cg_vec_body<T>(blockIdx.x, gridDim.x, x, r, u, p, s, w, dinv, n, partS, nS, partV_out, sc);
}

// ---- fused launches over the systems of a CgSys table.  Workgroup g belongs
// to the last system whose first workgroup (of this kernel's kind) is <= g.
template <int KIND>
__device__ __forceinline__ int cg_sys_of(const CgSys* __restrict__ tab, int nsys, int g) {
  int lo = 0, hi = nsys - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    const int b = KIND == 0 ? tab[mid].sp_begin : KIND == 1 ? tab[mid].red_begin : tab[mid].vec_begin;
    if (b <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// The launches are sized for a bucket >= the live workgroup counts in `hdr`;
// workgroups past them return before any barrier.
template <int D, typename T, int U>
__global__ void __launch_bounds__(kSpBS) fused_cg_spmv_kernel(const CgBatchHdr* __restrict__ hdr,
                                                              const CgSys* __restrict__ tab, int slot) {
  if ((int)blockIdx.x >= hdr->sp_blocks) return;
  const CgSys& q = tab[cg_sys_of<0>(tab, hdr->nsys, blockIdx.x)];
  cg_spmv_body<D, T, U>(blockIdx.x - q.sp_begin, q.cptr, q.conns, (const T*)q.val, (const T*)q.u, (T*)q.w, q.N,
                          q.partS_raw, slot ? q.partV1 : q.partV0, q.gv, q.sc);
}

__global__ void __launch_bounds__(kBS) fused_reduce_partials_kernel(const CgBatchHdr* __restrict__ hdr,
                                                                    const CgSys* __restrict__ tab) {
  if ((int)blockIdx.x >= hdr->red_blocks) return;
  const CgSys& q = tab[cg_sys_of<1>(tab, hdr->nsys, blockIdx.x)];
  if (!q.two_level) return;  // systems without a second level own no workgroups here
  reduce_partials_body(blockIdx.x - q.red_begin, kBS, q.partS_raw, q.gs, q.partS, q.sc);
}

template <typename T>
__global__ void __launch_bounds__(kVBS) fused_cg_vec_kernel(const CgBatchHdr* __restrict__ hdr,
                                                           const CgSys* __restrict__ tab, int slot) {
  if ((int)blockIdx.x >= hdr->vec_blocks) return;
  const CgSys& q = tab[cg_sys_of<2>(tab, hdr->nsys, blockIdx.x)];
  cg_vec_body<T>(blockIdx.x - q.vec_begin, q.gv, (T*)q.x, (T*)q.r, (T*)q.u, (T*)q.p, (T*)q.s, (const T*)q.w,
                 (const T*)q.dinv, q.n, q.partS, q.nS, slot ? q.partV0 : q.partV1, q.sc);
}

// done flag and iteration count of every system of the table
__global__ void cg_flags_kernel(const CgSys* __restrict__ tab, int nsys, int* __restrict__ out) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < nsys; k += gridDim.x * blockDim.x) {
    out[2 * k] = tab[k].sc->done;
    out[2 * k + 1] = tab[k].sc->iter;
  }
}

// Start: x = 0, r = b, u = M b, p = s = 0; partials of (r,u), (r,r)
template <typename T>
__global__ void __launch_bounds__(kVBS) cg_init_kernel(const double* __restrict__ b, T* __restrict__ x,
                                                      T* __restrict__ r, T* __restrict__ u,
                                                      T* __restrict__ p, T* __restrict__ s,
                                                      const T* __restrict__ dinv, int64_t n,
                                                      double* __restrict__ partV) {
  double ru = 0, rr = 0;
  for (int64_t k = (int64_t)blockIdx.x * kVBS + threadIdx.x; k < n; k += (int64_t)gridDim.x * kVBS) {
    const T rk = (T)b[k];
    x[k] = 0;
    p[k] = 0;
    s[k] = 0;
    r[k] = rk;
    const T uk = dinv[k] * rk;
    u[k] = uk;
    ru += (double)rk * (double)uk;
    rr += (double)rk * (double)rk;
  }
  Vec<2> a;
  a.v[0] = ru;
  a.v[1] = rr;
  write_partial<kVBS, 2>(partV, a);
}

__global__ void cg_setup_kernel(Scalars* sc, double rtol, double atol, int64_t maxit) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    sc->rtol = rtol;
    sc->atol = atol;
    sc->maxit = (int)maxit;
    sc->iter = 0;
    sc->icur = 0;
    sc->inext = 0;
    sc->done = 0;
  }
}

// Jacobi preconditioner from the BSR diagonal blocks (1 for the plain CG).
template <int D, typename T>
__global__ void dinv_kernel(const int* __restrict__ dpos, const T* __restrict__ val, int N, int precond,
                            T* __restrict__ dinv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
#pragma unroll
  for (int r = 0; r < D; ++r)
    dinv[(size_t)D * i + r] = precond ? T(1) / val[(size_t)dpos[i] * D * D + r * D + r] : T(1);
}

__global__ void f32_to_f64_kernel(const float* __restrict__ a, double* __restrict__ b, int64_t n) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
    b[t] = (double)a[t];
}

int grid_for(int64_t work, int per) {
  int64_t g = (work + per - 1) / per;
  g = std::max<int64_t>(8, std::min<int64_t>(g, kMaxGrid));
  return (int)g;
}

// A row takes 16 lanes, four rows per wavefront, in both precisions, and a pass
// covers chunks_per_pass(c) chunks of 16 blocks: a row's dependent loads (row
// start, block columns, x) bound one row per wavefront, and 16 lanes keep four
// rows' bytes in flight (back to back on SC 40^3: FP64 42.2 us with 16 lanes,
// 47.2 us with 64, 57.0 us with 32; the BCC J2 bar: 58.4 vs 68.7 us with 64;
// the 250k CT plate: 43.9 vs 50.4 us with 32).
int spmv_grid(const Context* c) {
  const int rows_per_block = (kSpBS / 64) * (64 / kLPR);
  return (c->N + rows_per_block - 1) / rows_per_block;
}

template <int D, typename T>
void launch_bsr(Context* c, const T* val, const T* x, T* y) {
  const int g = spmv_grid(c);
  switch (chunks_per_pass(c)) {
    case 2: spmv_kernel<D, T, 2><<<g, kSpBS, 0, c->stream>>>(c->cptr, c->conns, val, x, y, c->N); break;
    case 3: spmv_kernel<D, T, 3><<<g, kSpBS, 0, c->stream>>>(c->cptr, c->conns, val, x, y, c->N); break;
    default: spmv_kernel<D, T, 4><<<g, kSpBS, 0, c->stream>>>(c->cptr, c->conns, val, x, y, c->N); break;
  }
}

// With events, the kernel's own start and end are recorded with its dispatch
// (hipExtLaunchKernelGGL): the interval is the kernel, without the gaps that
// separately enqueued event records add.
// Iteration k of a chunk reads the vector-step partials of slot k & 1.
template <int D, typename T, int U>
void launch_cg_spmv_u(hipStream_t st, const CgSys& q, int k, hipEvent_t e0, hipEvent_t e1) {
  const double* partV = (k & 1) ? q.partV1 : q.partV0;
  if (e0)  // instrumented (eager) launches only; captured chunks use plain launches
    hipExtLaunchKernelGGL(cg_spmv_kernel<D, T, U>, dim3(q.gs), dim3(kSpBS), 0, st, e0, e1, 0, q.cptr, q.conns,
                          (const T*)q.val, (const T*)q.u, (T*)q.w, q.N, q.partS_raw, partV, q.gv, q.sc);
  else
    cg_spmv_kernel<D, T, U><<<q.gs, kSpBS, 0, st>>>(q.cptr, q.conns, (const T*)q.val, (const T*)q.u, (T*)q.w, q.N,
                                                    q.partS_raw, partV, q.gv, q.sc);
}

template <int D, typename T>
void launch_cg_spmv(hipStream_t st, const CgSys& q, int chunks, int k, hipEvent_t e0 = nullptr,
                    hipEvent_t e1 = nullptr) {
  switch (chunks) {
    case 2: launch_cg_spmv_u<D, T, 2>(st, q, k, e0, e1); break;
    case 3: launch_cg_spmv_u<D, T, 3>(st, q, k, e0, e1); break;
    default: launch_cg_spmv_u<D, T, 4>(st, q, k, e0, e1); break;
  }
}

// Vectors of the CG in the given precision (FP32: the mixed-precision copies)
template <typename T>
void cg_vectors(Context* c, CgSys* q);
template <>
void cg_vectors<double>(Context* c, CgSys* q) {
  q->val = c->Kval; q->dinv = c->dinv;
  q->x = c->du; q->r = c->cg_r; q->u = c->cg_z; q->p = c->cg_p; q->s = c->cg_q; q->w = c->cg_w;
}
template <>
void cg_vectors<float>(Context* c, CgSys* q) {
  q->val = c->Kval32; q->dinv = c->f_dinv;
  q->x = c->f_x; q->r = c->f_r; q->u = c->f_z; q->p = c->f_p; q->s = c->f_q; q->w = c->f_w;
}

// Fill the system's descriptor and enqueue the start of its solve (Jacobi
// diagonal, scalars, x = 0, r = b and the first partials) on its stream.
template <int D, typename T>
int cg_begin(Context* c, const SolveOpts& o, CgSys* q) {
  *q = CgSys{};
  cg_vectors<T>(c, q);
  q->cptr = c->cptr;
  q->conns = c->conns;
  q->N = c->N;
  q->n = (int64_t)D * c->N;
  q->gv = grid_for(q->n, kVBS);
  q->gs = spmv_grid(c);
  q->two_level = q->gs > kMaxParts;
  q->nS = q->two_level ? kBS : q->gs;
  q->partV0 = c->partials;
  q->partV1 = c->partials + 2 * kMaxGrid;
  q->partS_raw = c->partials + 4 * kMaxGrid;
  q->partS = q->two_level ? q->partS_raw + q->gs : q->partS_raw;
  q->sc = c->sc;
  q->cls = cg_class(D, sizeof(T) == 4, chunks_per_pass(c));
  dinv_kernel<D, T><<<(c->N + 255) / 256, 256, 0, c->stream>>>(c->dpos, (const T*)q->val, c->N, o.precond,
                                                              (T*)q->dinv);
  const int64_t maxit = o.maxit > 0 ? o.maxit : q->n;
  cg_setup_kernel<<<1, 64, 0, c->stream>>>(c->sc, o.rtol, o.atol, maxit);
  cg_init_kernel<T><<<q->gv, kVBS, 0, c->stream>>>(c->residual, (T*)q->x, (T*)q->r, (T*)q->u, (T*)q->p, (T*)q->s,
                                                  (const T*)q->dinv, q->n, q->partV0);
  LPMG_HIP(hipGetLastError());
  return 0;
}

// One system, its own stream: chunks of kCgChunk iterations between reads of
// the convergence flag; a chunk always starts at an even iteration, so the
// parity of the vector-step partial buffers is fixed per launch slot.
template <int D, typename T>
int run_cg(Context* c, const CgSys& q, int* iters, double* rr_final) {
  const int chunks = chunks_per_pass(c);
  c->sc_host_iter_launched = 0;
  auto enqueue_chunk = [&](std::vector<hipEvent_t>* evs) -> int {
    for (int k = 0; k < kCgChunk; ++k) {
      if (evs)
        launch_cg_spmv<D, T>(c->stream, q, chunks, k, (*evs)[2 * k], (*evs)[2 * k + 1]);
      else
        launch_cg_spmv<D, T>(c->stream, q, chunks, k);
      if (q.two_level) reduce_partials_kernel<<<kBS, kBS, 0, c->stream>>>(q.partS_raw, q.gs, q.partS, q.sc);
      cg_vec_kernel<T><<<q.gv, kVBS, 0, c->stream>>>((T*)q.x, (T*)q.r, (T*)q.u, (T*)q.p, (T*)q.s, (const T*)q.w,
                                                    (const T*)q.dinv, q.n, q.partS, q.nS,
                                                    (k & 1) ? q.partV0 : q.partV1, q.sc);
    }
    LPMG_HIP(hipGetLastError());
    return 0;
  };
  std::vector<hipEvent_t> evs;
  const int P = sizeof(T) == 4 ? 1 : 0;
  hipGraphExec_t* gexec = &c->cg_graph[P];
  const std::vector<uintptr_t> sig = cg_graph_signature(q, chunks);
  if (c->prof) {
    evs.resize(2 * kCgChunk);
    for (auto& e : evs) LPMG_HIP(hipEventCreate(&e));
  } else if (*gexec) {
    if (int rc = check_graph_signature(c->cg_graph_sig[P], sig, "CG chunk")) return rc;
  } else {
    // capture one chunk once per handle and precision; every pointer, size and
    // launch shape in it is fixed for the handle's lifetime
    hipGraph_t graph = nullptr;
    LPMG_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_chunk(nullptr);
    hipError_t e2 = hipStreamEndCapture(c->stream, &graph);
    if (rc) return rc;
    LPMG_HIP(e2);
    LPMG_HIP(hipGraphInstantiate(gexec, graph, nullptr, nullptr, 0));
    LPMG_HIP(hipGraphDestroy(graph));
    c->cg_graph_sig[P] = sig;
  }
  for (;;) {
    const int i0 = c->sc_host_iter_launched;
    if (c->prof) {
      if (int rc = enqueue_chunk(&evs)) return rc;
    } else {
      LPMG_HIP(hipGraphLaunch(*gexec, c->stream));
    }
    c->sc_host_iter_launched += kCgChunk;
    // the flag is read after the chunk; iterations past convergence were no-ops
    LPMG_HIP(hipMemcpyAsync(c->sc_host, c->sc, sizeof(Scalars), hipMemcpyDeviceToHost, c->stream));
    LPMG_HIP(hipStreamSynchronize(c->stream));
    if (c->prof) {
      // SpMV launches S_i0 .. S_stop did work; the ones after the stop returned at once
      const int worked = c->sc_host->done ? std::min(kCgChunk, c->sc_host->iter - i0 + 1) : kCgChunk;
      for (int k = 0; k < worked; ++k) {
        float ms = 0;
        LPMG_HIP(hipEventElapsedTime(&ms, evs[2 * k], evs[2 * k + 1]));
        c->prof_spmv_ms += ms;
        c->prof_spmv_n += 1;
      }
    }
    if (c->sc_host->done) break;
  }
  if (c->prof)
    for (auto& e : evs) (void)hipEventDestroy(e);
  *iters = c->sc_host->iter;
  *rr_final = c->sc_host->rr;
  return 0;
}

template <int D, typename T, int U>
void launch_fused_class(hipStream_t st, const CgBatchClass& b, int k) {
  const CgSys* tab = reinterpret_cast<const CgSys*>(reinterpret_cast<const char*>(b.hdr) + kCgHdrBytes);
  fused_cg_spmv_kernel<D, T, U><<<b.sp_grid, kSpBS, 0, st>>>(b.hdr, tab, k & 1);
  if (b.red_grid) fused_reduce_partials_kernel<<<b.red_grid, kBS, 0, st>>>(b.hdr, tab);
  fused_cg_vec_kernel<T><<<b.vec_grid, kVBS, 0, st>>>(b.hdr, tab, k & 1);
}

}  // namespace

std::vector<uintptr_t> cg_graph_signature(const CgSys& q, int chunks) {
  std::vector<uintptr_t> s;
  for (const void* p : {(const void*)q.cptr, (const void*)q.conns, q.val, q.dinv, (const void*)q.x, (const void*)q.r,
                        (const void*)q.u, (const void*)q.p, (const void*)q.s, (const void*)q.w, (const void*)q.partV0,
                        (const void*)q.partV1, (const void*)q.partS_raw, (const void*)q.partS, (const void*)q.sc})
    graph_sig_push(&s, p);
  for (int64_t v : {(int64_t)q.n, (int64_t)q.N, (int64_t)q.gs, (int64_t)q.gv, (int64_t)q.nS, (int64_t)q.two_level,
                    (int64_t)chunks})
    graph_sig_push(&s, v);
  return s;
}

int check_graph_signature(const std::vector<uintptr_t>& captured, const std::vector<uintptr_t>& now,
                          const char* which) {
  if (captured == now) return 0;
  set_error(std::string("captured ") + which + " graph no longer matches its buffers or launch shape");
  return -1;
}

int cg_class(int dim, int fp32, int chunks) {
  return (dim == 3 ? 6 : 0) + (fp32 ? 3 : 0) + (chunks - 2);
}

// Workgroup offsets of each system in the fused grids; the live totals.
void cg_batch_layout(CgSys* tab, int nsys, CgBatchHdr* h) {
  int sp = 0, red = 0, vec = 0;
  for (int k = 0; k < nsys; ++k) {
    tab[k].sp_begin = sp;
    tab[k].red_begin = red;
    tab[k].vec_begin = vec;
    sp += tab[k].gs;
    red += tab[k].two_level ? kBS : 0;
    vec += tab[k].gv;
  }
  h->nsys = nsys;
  h->sp_blocks = sp;
  h->red_blocks = red;
  h->vec_blocks = vec;
}

int cg_batch_enqueue_chunk(hipStream_t st, const CgBatchClass* classes, int nclass) {
  for (int k = 0; k < kCgChunk; ++k)
    for (int c = 0; c < nclass; ++c) {
      const CgBatchClass& b = classes[c];
      if (b.sp_grid == 0) continue;
      switch (b.cls) {
        case 0: launch_fused_class<2, double, 2>(st, b, k); break;
        case 1: launch_fused_class<2, double, 3>(st, b, k); break;
        case 2: launch_fused_class<2, double, 4>(st, b, k); break;
        case 3: launch_fused_class<2, float, 2>(st, b, k); break;
        case 4: launch_fused_class<2, float, 3>(st, b, k); break;
        case 5: launch_fused_class<2, float, 4>(st, b, k); break;
        case 6: launch_fused_class<3, double, 2>(st, b, k); break;
        case 7: launch_fused_class<3, double, 3>(st, b, k); break;
        case 8: launch_fused_class<3, double, 4>(st, b, k); break;
        case 9: launch_fused_class<3, float, 2>(st, b, k); break;
        case 10: launch_fused_class<3, float, 3>(st, b, k); break;
        case 11: launch_fused_class<3, float, 4>(st, b, k); break;
        default: set_error("bad CG batch class"); return -1;
      }
    }
  LPMG_HIP(hipGetLastError());
  return 0;
}

int cg_batch_flags(hipStream_t st, const CgSys* tab, int nsys, int* out) {
  if (nsys == 0) return 0;
  cg_flags_kernel<<<(nsys + 63) / 64, 64, 0, st>>>(tab, nsys, out);
  LPMG_HIP(hipGetLastError());
  return 0;
}

double spmv_bytes(const Context* c, int precision) {
  const double w = precision ? 4.0 : 8.0;
  const double dd = (double)c->dim * c->dim;
  return (double)c->nnzb * (dd * w + 4.0) + (double)(c->N + 1) * 4.0 + 2.0 * c->dim * (double)c->N * w;
}

int launch_spmv(Context* c, int precision, const void* x, void* y) {
  if (precision == 0) {
    if (c->dim == 3) launch_bsr<3, double>(c, c->Kval, (const double*)x, (double*)y);
    else launch_bsr<2, double>(c, c->Kval, (const double*)x, (double*)y);
  } else {
    if (!c->Kval32) {
      set_error("FP32 stiffness not assembled");
      return -1;
    }
    if (c->dim == 3) launch_bsr<3, float>(c, c->Kval32, (const float*)x, (float*)y);
    else launch_bsr<2, float>(c, c->Kval32, (const float*)x, (float*)y);
  }
  LPMG_HIP(hipGetLastError());
  return 0;
}

// The linear solve K du = residual of one Newton iteration.  With a CG
// executor installed on the context (the serving runtime's fused launches),
// the iterations run there; otherwise on the context's own stream.
int solve_linear(Context* c, const SolveOpts& o, int* iters, double* rr_final) {
  const int64_t n = (int64_t)c->dim * c->N;
  if (o.precision == 0 ? !c->assembled64 : !c->assembled32) {
    set_error(o.precision == 0 ? "stiffness not assembled" : "FP32 stiffness not assembled");
    return -1;
  }
  int rc;
  if (o.precond >= 2 && o.precond <= 4) {  // multigrid PCG or BiCGSTAB, on the context's own stream
    if (o.precond == 3) {
      if (o.precision == 0)
        rc = c->dim == 3 ? run_bicgstab<3, double>(c, o, iters, rr_final)
                         : run_bicgstab<2, double>(c, o, iters, rr_final);
      else
        rc = c->dim == 3 ? run_bicgstab<3, float>(c, o, iters, rr_final)
                         : run_bicgstab<2, float>(c, o, iters, rr_final);
    } else if (o.precision == 0) {
      rc = c->dim == 3 ? run_pcg_amg<3, double>(c, o, iters, rr_final) : run_pcg_amg<2, double>(c, o, iters, rr_final);
    } else {
      rc = c->dim == 3 ? run_pcg_amg<3, float>(c, o, iters, rr_final) : run_pcg_amg<2, float>(c, o, iters, rr_final);
    }
    if (rc) return rc;
    if (o.precision != 0) {
      f32_to_f64_kernel<<<grid_for(n, 1024), 256, 0, c->stream>>>(c->f_x, c->du, n);
      LPMG_HIP(hipGetLastError());
    }
    return 0;
  }
  CgSys q;
  if (o.precision == 0) rc = c->dim == 3 ? cg_begin<3, double>(c, o, &q) : cg_begin<2, double>(c, o, &q);
  else rc = c->dim == 3 ? cg_begin<3, float>(c, o, &q) : cg_begin<2, float>(c, o, &q);
  if (rc) return rc;
  if (c->cg_exec && !c->prof) {
    if ((rc = c->cg_exec(c->cg_exec_arg, c, q))) return rc;
    // the executor has ordered this stream after the system's last fused launch
    LPMG_HIP(hipMemcpyAsync(c->sc_host, c->sc, sizeof(Scalars), hipMemcpyDeviceToHost, c->stream));
    LPMG_HIP(hipStreamSynchronize(c->stream));
    *iters = c->sc_host->iter;
    *rr_final = c->sc_host->rr;
  } else if (o.precision == 0) {
    rc = c->dim == 3 ? run_cg<3, double>(c, q, iters, rr_final) : run_cg<2, double>(c, q, iters, rr_final);
  } else {
    rc = c->dim == 3 ? run_cg<3, float>(c, q, iters, rr_final) : run_cg<2, float>(c, q, iters, rr_final);
  }
  if (rc) return rc;
  if (o.precision != 0) {
    f32_to_f64_kernel<<<grid_for(n, 1024), 256, 0, c->stream>>>(c->f_x, c->du, n);
    LPMG_HIP(hipGetLastError());
  }
  return 0;
}

}  // namespace lpmg
