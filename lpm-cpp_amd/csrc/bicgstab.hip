// bicgstab.hip -- Jacobi-preconditioned BiCGSTAB on the block-CSR stiffness
// (solver option opts[0] = 3).  No reference counterpart: LPM-CPP solves the
// Newton step with MKL's RCI CG or PARDISO (Solver::LPM_CG / LPM_PARDISO,
// solver.h:190-354).  The stiffness is symmetric, so CG is the default; this
// solver is for callers that ask for BiCGSTAB (it does not need symmetry).
//
// Method: preconditioned BiCGSTAB as in Barrett et al., "Templates for the
// Solution of Linear Systems" (SIAM 1994), Fig. 2.10 (van der Vorst 1992),
// with M = the diagonal of K (Jacobi) and r~ = r0:
//   rho_i = (r~, r_i);  beta = (rho_i / rho_{i-1}) (alpha_{i-1} / omega_{i-1})
//   p = r + beta (p - omega v);  y = M^-1 p;  v = K y;  alpha = rho_i / (r~, v)
//   s = r - alpha v;  z = M^-1 s;  t = K z;  omega = (t, s) / (t, t)
//   x += alpha y + omega z;  r = s - omega t
// Stopping test as the CG's: ||r||^2 <= rtol ||r0||^2 + atol, at most maxit
// iterations.  rho = 0, (r~, v) = 0 or omega = 0 before convergence is a
// breakdown: the solve stops and reports it.
//
// Per iteration five kernels, each grid reduction split across a kernel
// boundary as in the CG (per-workgroup partials, summed in a fixed order by the
// consumer; no intra-kernel fence): SpMV v = K y with the partials of (r~, v);
// vector step B (alpha, s, x += alpha y, z); SpMV t = K z with the partials of
// (t, s), (t, t); vector step C (omega, x += omega z, r, partials of (r~, r),
// (r, r)); vector step A (rho, the stop test, beta, p, y).  Scalars that
// outlive a launch sit in parity slots (rho, alpha, omega of iteration i in
// slot i & 1) written by workgroup 0; the iteration index is copied by the
// first SpMV of each iteration, a launch that does not read it.  A chunk of 16
// iterations is captured once per handle and precision and replayed, with the
// done flag read by the host after each chunk; launches after convergence
// return at once.  Deterministic: a repeated solve gives the same bits.
#include <algorithm>
#include <cmath>
#include <vector>

#include "bsr.h"

namespace lpmg {

namespace {

constexpr int kBVS = 256;  // vector kernels: threads per workgroup

// Done flag of the system, read once per workgroup (thread 0) and shared
// through LDS, so every wavefront takes the same branch.
__device__ __forceinline__ bool bicg_stopped(const Scalars* sc) {
  __shared__ int d;
  if (threadIdx.x == 0) d = sc->done;
  __syncthreads();
  return d != 0;
}

// y = K x for the rows of this workgroup and the partials of (a, y) and, with
// NDOT = 2, (y, y).  With SET_ICUR, workgroup 0 copies the iteration index for
// this iteration's vector steps (no other launch of the iteration writes it).
template <int D, typename T, int U, int NDOT, bool SET_ICUR>
__global__ void __launch_bounds__(kSpBS) bicg_spmv_kernel(const int* __restrict__ cptr,
                                                          const int* __restrict__ conns,
                                                          const T* __restrict__ val, const T* __restrict__ x,
                                                          T* __restrict__ y, int N, const T* __restrict__ a,
                                                          double* __restrict__ part, Scalars* sc) {
  if (sc->done) return;
  if (SET_ICUR && blockIdx.x == 0 && threadIdx.x == 0) sc->icur = sc->inext;
  constexpr int RPW = 64 / kLPR;
  const int lane = threadIdx.x & 63, sub = lane / kLPR, sl = lane - sub * kLPR;
  const int wave = blockIdx.x * (kSpBS / 64) + (threadIdx.x >> 6);
  const int row = wave * RPW + sub;
  const bool live = row < N;
  const int rw = live ? row : N - 1;  // dead rows shadow the last row: all lanes stay in the shuffles
  const int lo = cptr[rw], hi = cptr[rw + 1];
  const T ao = a[(size_t)rw * D + min(sl, D - 1)];
  T acc[D];
  bsr_row_acc<D, T, kLPR, U>(lo, hi, sl, conns, val, x, acc);
  Vec<NDOT> d;
#pragma unroll
  for (int k = 0; k < NDOT; ++k) d.v[k] = 0;
  if (live && sl < D) {
    T yv = acc[0];
#pragma unroll
    for (int q = 1; q < D; ++q)
      if (sl == q) yv = acc[q];
    y[(size_t)row * D + sl] = yv;
    d.v[0] = (double)ao * (double)yv;
    if (NDOT == 2) d.v[NDOT - 1] = (double)yv * (double)yv;
  }
  write_partial<kSpBS, NDOT>(part, d);
}

// Second level for large SpMV grids: n partials -> gridDim.x partials
// (contiguous slices, NV values each).
template <int NV>
__global__ void __launch_bounds__(kBS) bicg_reduce_kernel(const double* __restrict__ in, int n,
                                                          double* __restrict__ out, const Scalars* __restrict__ sc) {
  if (sc->done) return;
  const int per = (n + gridDim.x - 1) / gridDim.x;
  const int lo = blockIdx.x * per, hi = min(n, lo + per);
  Vec<NV> a;
#pragma unroll
  for (int k = 0; k < NV; ++k) a.v[k] = 0;
  for (int b = lo + threadIdx.x; b < hi; b += kBS)
#pragma unroll
    for (int k = 0; k < NV; ++k) a.v[k] += in[(size_t)NV * b + k];
  write_partial<kBS, NV>(out, a);
}

template <int D, typename T>
__global__ void bicg_dinv_kernel(const int* __restrict__ dpos, const T* __restrict__ val, int N,
                                 T* __restrict__ dinv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
#pragma unroll
  for (int r = 0; r < D; ++r) dinv[(size_t)D * i + r] = T(1) / val[(size_t)dpos[i] * D * D + r * D + r];
}

// x = 0, r = r~ = p = b, y = M^-1 b; partials of (r~, r) and (r, r)
template <typename T>
__global__ void __launch_bounds__(kBVS) bicg_init_kernel(const double* __restrict__ b, T* __restrict__ x,
                                                        T* __restrict__ r, T* __restrict__ rh, T* __restrict__ p,
                                                        T* __restrict__ y, const T* __restrict__ dinv, int64_t n,
                                                        double* __restrict__ partV) {
  Vec<2> a;
  a.v[0] = a.v[1] = 0;
  for (int64_t k = (int64_t)blockIdx.x * kBVS + threadIdx.x; k < n; k += (int64_t)gridDim.x * kBVS) {
    const T bk = (T)b[k];
    x[k] = 0;
    r[k] = bk;
    rh[k] = bk;
    p[k] = bk;
    y[k] = dinv[k] * bk;
    a.v[0] += (double)bk * (double)bk;
    a.v[1] += (double)bk * (double)bk;
  }
  write_partial<kBVS, 2>(partV, a);
}

// One workgroup: rho_0, ||r0||^2, the threshold and the stop test before the
// first iteration.
__global__ void __launch_bounds__(kBVS) bicg_start_kernel(const double* __restrict__ partV, int nV, Scalars* sc,
                                                         double rtol, double atol, int64_t maxit) {
  const Vec<2> a = sum_partials<kBVS, 2>(partV, nV);
  if (threadIdx.x == 0) {
    const double rho = a.v[0], rr = a.v[1];
    sc->rtol = rtol;
    sc->atol = atol;
    sc->maxit = (int)maxit;
    sc->rr0 = rr;
    sc->rr = rr;
    sc->thresh = rtol * rr + atol;
    sc->rho[0] = rho;
    sc->iter = 0;
    sc->icur = 0;
    sc->inext = 0;
    const bool stop = rr <= sc->thresh || maxit <= 0;
    sc->done = stop ? 1 : (rho == 0.0 || !std::isfinite(rho)) ? 2 : 0;
  }
}

// Vector step B of iteration i: alpha = rho_i / (r~, v); s = r - alpha v;
// x += alpha y; z = M^-1 s.
template <typename T>
__global__ void __launch_bounds__(kBVS) bicg_b_kernel(const T* __restrict__ r, const T* __restrict__ v,
                                                     const T* __restrict__ y, T* __restrict__ x,
                                                     T* __restrict__ s, T* __restrict__ z,
                                                     const T* __restrict__ dinv, int64_t n,
                                                     const double* __restrict__ part1, int n1, Scalars* sc) {
  if (bicg_stopped(sc)) return;
  const Vec<1> rv = sum_partials<kBVS, 1>(part1, n1);
  const int i = sc->icur;
  const double alpha = sc->rho[i & 1] / rv.v[0];
  const bool broke = rv.v[0] == 0.0 || !std::isfinite(alpha);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->alp[i & 1] = alpha;
    if (broke) sc->done = 2;
  }
  if (broke) return;  // every workgroup reaches the same decision
  const T a = (T)alpha;
  for (int64_t k = (int64_t)blockIdx.x * kBVS + threadIdx.x; k < n; k += (int64_t)gridDim.x * kBVS) {
    const T sk = r[k] - a * v[k];
    s[k] = sk;
    x[k] += a * y[k];
    z[k] = dinv[k] * sk;
  }
}

// Vector step C: omega = (t, s) / (t, t); x += omega z; r = s - omega t;
// partials of (r~, r), (r, r).  (t, t) = 0 means s = 0: omega = 0 and r = s.
template <typename T>
__global__ void __launch_bounds__(kBVS) bicg_c_kernel(T* __restrict__ x, const T* __restrict__ z,
                                                     const T* __restrict__ s, const T* __restrict__ t,
                                                     T* __restrict__ r, const T* __restrict__ rh, int64_t n,
                                                     const double* __restrict__ part2, int n2,
                                                     double* __restrict__ partV, Scalars* sc) {
  if (bicg_stopped(sc)) return;
  const Vec<2> ts = sum_partials<kBVS, 2>(part2, n2);
  const int i = sc->icur;
  const double omega = ts.v[1] > 0.0 ? ts.v[0] / ts.v[1] : 0.0;
  if (blockIdx.x == 0 && threadIdx.x == 0) sc->om[i & 1] = omega;
  const T w = (T)omega;
  Vec<2> a;
  a.v[0] = a.v[1] = 0;
  for (int64_t k = (int64_t)blockIdx.x * kBVS + threadIdx.x; k < n; k += (int64_t)gridDim.x * kBVS) {
    x[k] += w * z[k];
    const T rk = s[k] - w * t[k];
    r[k] = rk;
    a.v[0] += (double)rh[k] * (double)rk;
    a.v[1] += (double)rk * (double)rk;
  }
  write_partial<kBVS, 2>(partV, a);
}

// Vector step A, the end of iteration i: rho_{i+1} = (r~, r), ||r||^2, the stop
// test; beta = (rho_{i+1} / rho_i) (alpha_i / omega_i); p = r + beta (p -
// omega_i v); y = M^-1 p.  Workgroup 0 sets `done` during this launch; a
// workgroup that still reads 0 reaches the same decision from the same
// partials and returns without writing.
template <typename T>
__global__ void __launch_bounds__(kBVS) bicg_a_kernel(const T* __restrict__ r, T* __restrict__ p,
                                                     const T* __restrict__ v, T* __restrict__ y,
                                                     const T* __restrict__ dinv, int64_t n,
                                                     const double* __restrict__ partV, int nV, Scalars* sc) {
  if (bicg_stopped(sc)) return;
  const Vec<2> a = sum_partials<kBVS, 2>(partV, nV);
  const int i = sc->icur;
  const double rho = a.v[0], rr = a.v[1];
  const double omega = sc->om[i & 1];
  const double beta = (rho / sc->rho[i & 1]) * (sc->alp[i & 1] / omega);
  const bool stop = rr <= sc->thresh || i + 1 >= sc->maxit;
  const bool broke = !stop && (rho == 0.0 || omega == 0.0 || !std::isfinite(beta));
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    sc->rho[(i + 1) & 1] = rho;
    sc->rr = rr;
    sc->iter = i + 1;
    sc->inext = i + 1;
    if (stop) sc->done = 1;
    else if (broke) sc->done = 2;
  }
  if (stop || broke) return;
  const T b = (T)beta, w = (T)omega;
  for (int64_t k = (int64_t)blockIdx.x * kBVS + threadIdx.x; k < n; k += (int64_t)gridDim.x * kBVS) {
    const T pk = r[k] + b * (p[k] - w * v[k]);
    p[k] = pk;
    y[k] = dinv[k] * pk;
  }
}

int bicg_grid(int64_t work) {
  int64_t g = (work + kBVS - 1) / kBVS;
  return (int)std::max<int64_t>(8, std::min<int64_t>(g, kMaxGrid));
}

// The system's vectors in precision T: x, r, p, v, y, s share the CG's
// buffers; r~, z, t are BiCGSTAB's own (allocated on first use).
template <typename T>
struct BicgVecs {
  const T* val;
  T *dinv, *x, *r, *rh, *p, *v, *y, *s, *z, *t;
};

template <typename T>
int bicg_vectors(Context* c, BicgVecs<T>* b) {
  const int P = sizeof(T) == 4 ? 1 : 0;
  const size_t n = (size_t)c->dim * c->N;
  if (!c->bicg_vec[P]) {
    T* q = nullptr;
    if (dalloc(c, &q, 3 * n)) return -3;
    c->bicg_vec[P] = q;
  }
  T* e = (T*)c->bicg_vec[P];
  if (P) {
    b->val = (const T*)(const void*)c->Kval32;
    b->dinv = (T*)(void*)c->f_dinv;
    b->x = (T*)(void*)c->f_x; b->r = (T*)(void*)c->f_r; b->p = (T*)(void*)c->f_p;
    b->v = (T*)(void*)c->f_q; b->y = (T*)(void*)c->f_z; b->s = (T*)(void*)c->f_w;
  } else {
    b->val = (const T*)(const void*)c->Kval;
    b->dinv = (T*)(void*)c->dinv;
    b->x = (T*)(void*)c->du; b->r = (T*)(void*)c->cg_r; b->p = (T*)(void*)c->cg_p;
    b->v = (T*)(void*)c->cg_q; b->y = (T*)(void*)c->cg_z; b->s = (T*)(void*)c->cg_w;
  }
  b->rh = e;
  b->z = e + n;
  b->t = e + 2 * n;
  return 0;
}

template <int D, typename T, int U>
void bicg_spmv1(Context* c, const BicgVecs<T>& b, int gs, double* part) {
  bicg_spmv_kernel<D, T, U, 1, true><<<gs, kSpBS, 0, c->stream>>>(c->cptr, c->conns, b.val, b.y, b.v, c->N, b.rh,
                                                                  part, c->sc);
}

template <int D, typename T, int U>
void bicg_spmv2(Context* c, const BicgVecs<T>& b, int gs, double* part) {
  bicg_spmv_kernel<D, T, U, 2, false><<<gs, kSpBS, 0, c->stream>>>(c->cptr, c->conns, b.val, b.z, b.t, c->N, b.s,
                                                                   part, c->sc);
}

}  // namespace

template <int D, typename T>
int run_bicgstab(Context* c, const SolveOpts& o, int* iters, double* rr_final) {
  BicgVecs<T> b;
  if (int rc = bicg_vectors<T>(c, &b)) return rc;
  const int N = c->N;
  const int64_t n = (int64_t)D * N;
  const int gv = bicg_grid(n);
  const int rows_per_block = (kSpBS / 64) * (64 / kLPR);
  const int gs = (N + rows_per_block - 1) / rows_per_block;
  const bool two = gs > kMaxParts;
  // partials: (r~,r),(r,r) [2 gv]; SpMV 1 [gs (+ kBS)]; SpMV 2 [2 gs (+ 2 kBS)]
  double* partV = c->partials;
  double* p1 = c->partials + 2 * kMaxGrid;
  double* p1b = p1 + gs;
  double* p2 = p1b + kBS;
  double* p2b = p2 + 2 * (size_t)gs;
  if ((int64_t)(2 * kMaxGrid + 3 * (int64_t)gs + 3 * kBS) > 2 * (int64_t)c->n_partials) {
    set_error("partials buffer too small for BiCGSTAB");
    return -1;
  }
  const int n1 = two ? kBS : gs, n2 = two ? kBS : gs;
  const double* q1 = two ? p1b : p1;
  const double* q2 = two ? p2b : p2;
  const int chunks = chunks_per_pass(c);
  bicg_dinv_kernel<D, T><<<(N + 255) / 256, 256, 0, c->stream>>>(c->dpos, b.val, N, b.dinv);
  bicg_init_kernel<T><<<gv, kBVS, 0, c->stream>>>(c->residual, b.x, b.r, b.rh, b.p, b.y, b.dinv, n, partV);
  const int64_t maxit = o.maxit > 0 ? o.maxit : n;
  bicg_start_kernel<<<1, kBVS, 0, c->stream>>>(partV, gv, c->sc, o.rtol, o.atol, maxit);
  LPMG_HIP(hipGetLastError());
  auto enqueue_chunk = [&]() -> int {
    for (int k = 0; k < kCgChunk; ++k) {
      switch (chunks) {
        case 2: bicg_spmv1<D, T, 2>(c, b, gs, p1); break;
        case 3: bicg_spmv1<D, T, 3>(c, b, gs, p1); break;
        default: bicg_spmv1<D, T, 4>(c, b, gs, p1); break;
      }
      if (two) bicg_reduce_kernel<1><<<kBS, kBS, 0, c->stream>>>(p1, gs, p1b, c->sc);
      bicg_b_kernel<T><<<gv, kBVS, 0, c->stream>>>(b.r, b.v, b.y, b.x, b.s, b.z, b.dinv, n, q1, n1, c->sc);
      switch (chunks) {
        case 2: bicg_spmv2<D, T, 2>(c, b, gs, p2); break;
        case 3: bicg_spmv2<D, T, 3>(c, b, gs, p2); break;
        default: bicg_spmv2<D, T, 4>(c, b, gs, p2); break;
      }
      if (two) bicg_reduce_kernel<2><<<kBS, kBS, 0, c->stream>>>(p2, gs, p2b, c->sc);
      bicg_c_kernel<T><<<gv, kBVS, 0, c->stream>>>(b.x, b.z, b.s, b.t, b.r, b.rh, n, q2, n2, partV, c->sc);
      bicg_a_kernel<T><<<gv, kBVS, 0, c->stream>>>(b.r, b.p, b.v, b.y, b.dinv, n, partV, gv, c->sc);
    }
    LPMG_HIP(hipGetLastError());
    return 0;
  };
  const int P = sizeof(T) == 4 ? 1 : 0;
  std::vector<uintptr_t> sig;
  for (const void* q : {(const void*)c->cptr, (const void*)c->conns, (const void*)b.val, (const void*)b.dinv,
                        (const void*)b.x, (const void*)b.r, (const void*)b.rh, (const void*)b.p, (const void*)b.v,
                        (const void*)b.y, (const void*)b.s, (const void*)b.z, (const void*)b.t,
                        (const void*)c->partials, (const void*)c->sc})
    graph_sig_push(&sig, q);
  for (int64_t v : {(int64_t)N, (int64_t)gs, (int64_t)gv, (int64_t)two, (int64_t)chunks})
    graph_sig_push(&sig, v);
  hipGraphExec_t* gexec = &c->bicg_graph[P];
  if (*gexec) {
    if (int rc = check_graph_signature(c->bicg_graph_sig[P], sig, "BiCGSTAB chunk")) return rc;
  } else {
    hipGraph_t graph = nullptr;
    LPMG_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const int rc = enqueue_chunk();
    hipError_t e2 = hipStreamEndCapture(c->stream, &graph);
    if (rc) return rc;
    LPMG_HIP(e2);
    LPMG_HIP(hipGraphInstantiate(gexec, graph, nullptr, nullptr, 0));
    LPMG_HIP(hipGraphDestroy(graph));
    c->bicg_graph_sig[P] = sig;
  }
  for (;;) {
    LPMG_HIP(hipMemcpyAsync(c->sc_host, c->sc, sizeof(Scalars), hipMemcpyDeviceToHost, c->stream));
    LPMG_HIP(hipStreamSynchronize(c->stream));
    if (c->sc_host->done) break;
    LPMG_HIP(hipGraphLaunch(*gexec, c->stream));
  }
  *iters = c->sc_host->iter;
  *rr_final = c->sc_host->rr;
  if (c->sc_host->done == 2) {
    set_error("BiCGSTAB breakdown (rho, (r~, v) or omega vanished before convergence)");
    return -4;
  }
  return 0;
}

template int run_bicgstab<2, double>(Context*, const SolveOpts&, int*, double*);
template int run_bicgstab<3, double>(Context*, const SolveOpts&, int*, double*);
template int run_bicgstab<2, float>(Context*, const SolveOpts&, int*, double*);
template int run_bicgstab<3, float>(Context*, const SolveOpts&, int*, double*);

}  // namespace lpmg
