// solver.hip -- block-CSR SpMV and the (Jacobi-preconditioned) CG that replaces
// MKL's RCI CG (Solver::LPM_CG, solver.h:276-354) and PARDISO
// (Solver::LPM_PARDISO, solver.h:190-274) as the linear solve of the Newton step.
//
// SpMV streams the BSR values once: a wavefront owns a contiguous run of block
// rows; per row, lane l < (64/(b*b))*(b*b) always takes value l of each group of
// 64/(b*b) blocks, so its (block, r, s) position is fixed and the value loads of
// the 63 (b=3) or 64 (b=2) active lanes are one contiguous, coalesced segment.
// Workgroups are remapped so that each of the 8 XCDs owns one contiguous eighth
// of the rows (its L2 then holds the x entries its rows gather).  The operator is
// sparse, so MFMA is not used.  FP32 values stream half the bytes of FP64.
//
// CG keeps every scalar on the device.  Dot products are reduced per workgroup
// and finished by the last workgroup in a fixed order (deterministic), which
// also computes alpha / beta and the stopping test.  The stopping test is the
// one the reference configures MKL with: ||r||^2 <= rtol*||r0||^2 + atol
// (dpar[0], dpar[1], solver.h:312-313), at most maxit (ipar[4] = n) iterations.
// The host launches `chunk` iterations between reads of the convergence flag;
// kernels of iterations after convergence return immediately.
#include <algorithm>
#include <vector>

#include "lpmg_internal.h"
#include "reduce.h"

namespace lpmg {

namespace {

constexpr int kBS = 256;
constexpr int kMaxGrid = 2048;  // workgroups of the grid-strided kernels (multiple of 8)

__device__ __forceinline__ int xcd_remap(int b, int nb) {
  const int per = nb >> 3, rem = nb & 7;
  const int x = b & 7, idx = b >> 3;
  return x < rem ? x * (per + 1) + idx : rem * (per + 1) + (x - rem) * per + idx;
}

// y = A x for the wavefront's rows [row0, row1); returns the wave's partial of
// dot(x, y) in every lane when DOT.
template <int D, typename T, bool DOT>
__device__ __forceinline__ double spmv_rows(const int* __restrict__ cptr, const int* __restrict__ conns,
                                            const T* __restrict__ val, const T* __restrict__ x,
                                            T* __restrict__ y, int row0, int row1) {
  constexpr int DD = D * D;
  constexpr int BPI = 64 / DD;        // blocks per iteration
  constexpr int LANES = BPI * DD;     // active lanes
  const int lane = threadIdx.x & 63;
  const int boff = lane / DD, rs = lane - boff * DD, r = rs / D, s = rs - r * D;
  const bool on = lane < LANES;
  double dot = 0;
  for (int row = row0; row < row1; ++row) {
    const int lo = cptr[row], hi = cptr[row + 1];
    T acc = 0;
    int blk = lo + boff;
    const T* vp = val + (size_t)lo * DD + lane;
#pragma unroll 2
    for (int b0 = lo; b0 < hi; b0 += BPI, blk += BPI, vp += BPI * DD) {
      if (on && blk < hi) {
        const int col = conns[blk];
        acc += *vp * x[(size_t)col * D + s];
      }
    }
    // reduce lanes by output component r
    T out[D];
#pragma unroll
    for (int q = 0; q < D; ++q) {
      T v = (on && r == q) ? acc : T(0);
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
      out[q] = v;
    }
    if (lane < D) {
      T yv = out[0];
#pragma unroll
      for (int q = 1; q < D; ++q)
        if (lane == q) yv = out[q];
      y[(size_t)row * D + lane] = yv;
      if (DOT) dot += (double)x[(size_t)row * D + lane] * (double)yv;
    }
  }
  return dot;
}

__device__ __forceinline__ void wave_rows(int N, int* r0, int* r1) {
  const int nwaves = gridDim.x * (kBS / 64);
  const int wid = xcd_remap(blockIdx.x, gridDim.x) * (kBS / 64) + (threadIdx.x >> 6);
  const int per = (N + nwaves - 1) / nwaves;
  *r0 = min(N, wid * per);
  *r1 = min(N, *r0 + per);
}

template <int D, typename T>
__global__ void __launch_bounds__(kBS) spmv_kernel(const int* __restrict__ cptr,
                                                   const int* __restrict__ conns,
                                                   const T* __restrict__ val, const T* __restrict__ x,
                                                   T* __restrict__ y, int N) {
  int r0, r1;
  wave_rows(N, &r0, &r1);
  spmv_rows<D, T, false>(cptr, conns, val, x, y, r0, r1);
}

// CG step A: q = A p, pq = p.q, alpha = rz / pq
template <int D, typename T>
__global__ void __launch_bounds__(kBS) cg_spmv_kernel(const int* __restrict__ cptr,
                                                      const int* __restrict__ conns,
                                                      const T* __restrict__ val, const T* __restrict__ p,
                                                      T* __restrict__ q, int N, double* partials,
                                                      Scalars* sc) {
  if (sc->done) return;
  int r0, r1;
  wave_rows(N, &r0, &r1);
  Vec<1> a;
  a.v[0] = spmv_rows<D, T, true>(cptr, conns, val, p, q, r0, r1);
  // only lanes < D carry a dot partial; block_sum adds all lanes
  a = block_sum<kBS, 1>(a);
  Vec<1> tot;
  if (last_block_sum<kBS, 1>(a, partials, &sc->counter[2], &tot))
    if (threadIdx.x == 0) {
      sc->pq = tot.v[0];
      sc->alpha = sc->rz / tot.v[0];
    }
}

// CG step B: x += alpha p; r -= alpha q; z = dinv r; rz, rr; beta; stopping test
template <typename T>
__global__ void __launch_bounds__(kBS) cg_update_kernel(T* __restrict__ x, T* __restrict__ r,
                                                        T* __restrict__ z, const T* __restrict__ p,
                                                        const T* __restrict__ q,
                                                        const T* __restrict__ dinv, int64_t n,
                                                        double* partials, Scalars* sc) {
  if (sc->done) return;
  const T alpha = (T)sc->alpha;
  double rz = 0, rr = 0;
  for (int64_t t = (int64_t)blockIdx.x * kBS + threadIdx.x; t < n; t += (int64_t)gridDim.x * kBS) {
    x[t] += alpha * p[t];
    const T rv = r[t] - alpha * q[t];
    r[t] = rv;
    const T zv = dinv[t] * rv;
    z[t] = zv;
    rz += (double)rv * (double)zv;
    rr += (double)rv * (double)rv;
  }
  Vec<2> a;
  a.v[0] = rz;
  a.v[1] = rr;
  a = block_sum<kBS, 2>(a);
  Vec<2> tot;
  if (last_block_sum<kBS, 2>(a, partials, &sc->counter[3], &tot))
    if (threadIdx.x == 0) {
      sc->beta = tot.v[0] / sc->rz;
      sc->rz = tot.v[0];
      sc->rr = tot.v[1];
      sc->iter += 1;
      if (tot.v[1] <= sc->thresh || sc->iter >= sc->maxit) sc->done = 1;
    }
}

// CG step C: p = z + beta p
template <typename T>
__global__ void __launch_bounds__(kBS) cg_dir_kernel(T* __restrict__ p, const T* __restrict__ z,
                                                     int64_t n, const Scalars* sc) {
  if (sc->done) return;
  const T beta = (T)sc->beta;
  for (int64_t t = (int64_t)blockIdx.x * kBS + threadIdx.x; t < n; t += (int64_t)gridDim.x * kBS)
    p[t] = z[t] + beta * p[t];
}

// Start: x = 0, r = b, z = dinv b, p = z, rz, rr0, threshold.
template <typename T>
__global__ void __launch_bounds__(kBS) cg_init_kernel(const double* __restrict__ b, T* __restrict__ x,
                                                      T* __restrict__ r, T* __restrict__ z,
                                                      T* __restrict__ p, const T* __restrict__ dinv,
                                                      int64_t n, double rtol, double atol,
                                                      int64_t maxit, double* partials, Scalars* sc) {
  double rz = 0, rr = 0;
  for (int64_t t = (int64_t)blockIdx.x * kBS + threadIdx.x; t < n; t += (int64_t)gridDim.x * kBS) {
    const T rv = (T)b[t];
    x[t] = 0;
    r[t] = rv;
    const T zv = dinv[t] * rv;
    z[t] = zv;
    p[t] = zv;
    rz += (double)rv * (double)zv;
    rr += (double)rv * (double)rv;
  }
  Vec<2> a;
  a.v[0] = rz;
  a.v[1] = rr;
  a = block_sum<kBS, 2>(a);
  Vec<2> tot;
  if (last_block_sum<kBS, 2>(a, partials, &sc->counter[3], &tot))
    if (threadIdx.x == 0) {
      sc->rz = tot.v[0];
      sc->rr = tot.v[1];
      sc->rr0 = tot.v[1];
      sc->thresh = rtol * tot.v[1] + atol;
      sc->iter = 0;
      sc->maxit = (int)maxit;
      sc->alpha = 0;
      sc->beta = 0;
      sc->done = (tot.v[1] <= sc->thresh || maxit <= 0) ? 1 : 0;
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

template <int D, typename T>
int run_cg(Context* c, const SolveOpts& o, T* x, T* r, T* z, T* p, T* q, T* dinv, const T* val,
           int* iters, double* rr_final) {
  const int N = c->N;
  const int64_t n = (int64_t)D * N;
  const int gv = grid_for(n, kBS * 4);
  const int gs = grid_for(N, 4 * 4);  // ~4 rows per wavefront
  dinv_kernel<D, T><<<(N + 255) / 256, 256, 0, c->stream>>>(c->dpos, val, N, o.precond, dinv);
  const int64_t maxit = o.maxit > 0 ? o.maxit : n;
  cg_init_kernel<T><<<gv, kBS, 0, c->stream>>>(c->residual, x, r, z, p, dinv, n, o.rtol, o.atol, maxit,
                                               c->partials, c->sc);
  LPMG_HIP(hipGetLastError());
  const int chunk = std::max(1, o.chunk);
  std::vector<hipEvent_t> evs;
  if (c->prof) {
    evs.resize(2 * chunk);
    for (auto& e : evs) LPMG_HIP(hipEventCreate(&e));
  }
  int rc = 0;
  int iter_before = 0;
  for (;;) {
    for (int k = 0; k < chunk; ++k) {
      if (c->prof) LPMG_HIP(hipEventRecord(evs[2 * k], c->stream));
      cg_spmv_kernel<D, T><<<gs, kBS, 0, c->stream>>>(c->cptr, c->conns, val, p, q, N, c->partials, c->sc);
      if (c->prof) LPMG_HIP(hipEventRecord(evs[2 * k + 1], c->stream));
      cg_update_kernel<T><<<gv, kBS, 0, c->stream>>>(x, r, z, p, q, dinv, n, c->partials, c->sc);
      cg_dir_kernel<T><<<gv, kBS, 0, c->stream>>>(p, z, n, c->sc);
    }
    LPMG_HIP(hipGetLastError());
    // the flag is read after the chunk; iterations past convergence were no-ops
    LPMG_HIP(hipMemcpyAsync(c->sc_host, c->sc, sizeof(Scalars), hipMemcpyDeviceToHost, c->stream));
    LPMG_HIP(hipStreamSynchronize(c->stream));
    if (c->prof) {
      // only the first (iter - iter_before) SpMV launches of the chunk did work;
      // the ones after convergence returned at once
      const int worked = std::min(chunk, c->sc_host->iter - iter_before);
      for (int k = 0; k < worked; ++k) {
        float ms = 0;
        LPMG_HIP(hipEventElapsedTime(&ms, evs[2 * k], evs[2 * k + 1]));
        c->prof_spmv_ms += ms;
        c->prof_spmv_n += 1;
      }
    }
    iter_before = c->sc_host->iter;
    if (c->sc_host->done) break;
  }
  if (c->prof)
    for (auto& e : evs) (void)hipEventDestroy(e);
  *iters = c->sc_host->iter;
  *rr_final = c->sc_host->rr;
  return rc;
}

}  // namespace

double spmv_bytes(const Context* c, int precision) {
  const double w = precision ? 4.0 : 8.0;
  const double dd = (double)c->dim * c->dim;
  return (double)c->nnzb * (dd * w + 4.0) + (double)(c->N + 1) * 4.0 + 2.0 * c->dim * (double)c->N * w;
}

int launch_spmv(Context* c, int precision, const void* x, void* y) {
  const int gs = grid_for(c->N, 16);
  if (precision == 0) {
    if (c->dim == 3)
      spmv_kernel<3, double><<<gs, kBS, 0, c->stream>>>(c->cptr, c->conns, c->Kval, (const double*)x,
                                                        (double*)y, c->N);
    else
      spmv_kernel<2, double><<<gs, kBS, 0, c->stream>>>(c->cptr, c->conns, c->Kval, (const double*)x,
                                                        (double*)y, c->N);
  } else {
    if (!c->Kval32) {
      set_error("FP32 stiffness not assembled");
      return -1;
    }
    if (c->dim == 3)
      spmv_kernel<3, float><<<gs, kBS, 0, c->stream>>>(c->cptr, c->conns, c->Kval32, (const float*)x,
                                                       (float*)y, c->N);
    else
      spmv_kernel<2, float><<<gs, kBS, 0, c->stream>>>(c->cptr, c->conns, c->Kval32, (const float*)x,
                                                       (float*)y, c->N);
  }
  LPMG_HIP(hipGetLastError());
  return 0;
}

int solve_linear(Context* c, const SolveOpts& o, int* iters, double* rr_final) {
  const int64_t n = (int64_t)c->dim * c->N;
  int rc;
  if (o.precision == 0) {
    if (!c->assembled64) {
      set_error("stiffness not assembled");
      return -1;
    }
    rc = c->dim == 3 ? run_cg<3, double>(c, o, c->du, c->cg_r, c->cg_z, c->cg_p, c->cg_q, c->dinv,
                                         c->Kval, iters, rr_final)
                     : run_cg<2, double>(c, o, c->du, c->cg_r, c->cg_z, c->cg_p, c->cg_q, c->dinv,
                                         c->Kval, iters, rr_final);
    if (rc) return rc;
  } else {
    if (!c->assembled32) {
      set_error("FP32 stiffness not assembled");
      return -1;
    }
    rc = c->dim == 3 ? run_cg<3, float>(c, o, c->f_x, c->f_r, c->f_z, c->f_p, c->f_q, c->f_dinv,
                                        c->Kval32, iters, rr_final)
                     : run_cg<2, float>(c, o, c->f_x, c->f_r, c->f_z, c->f_p, c->f_q, c->f_dinv,
                                        c->Kval32, iters, rr_final);
    if (rc) return rc;
    f32_to_f64_kernel<<<grid_for(n, 1024), 256, 0, c->stream>>>(c->f_x, c->du, n);
    LPMG_HIP(hipGetLastError());
  }
  return 0;
}

}  // namespace lpmg
