// api.hip -- the extern "C" boundary of liblpmg.so (declared in include/lpm_gpu.h)
// and the host-side drivers of the Newton step (NewtonIteration,
// solveProblemStep) that sequence the device kernels.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/lpm_gpu.h"
#include "lpmg_internal.h"

namespace lpmg {

static thread_local std::string g_err;

void set_error(const std::string& s) { g_err = s; }

int hip_fail(hipError_t e, const char* what) {
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  return e == hipErrorOutOfMemory ? -3 : -2;
}

// UnitCell::set* (include/unit_cell.h:44-122)
bool make_cell(int lattice, double r, Cell* c) {
  c->lattice = lattice;
  c->radius = r;
  switch (lattice) {
    case kSquare2D: {
      const double e[9] = {1. / 2., -1. / 2., 0., 0., 0., 1. / 2., 0., 1. / 12., -1. / 12.};
      std::memcpy(c->el, e, sizeof e);
      c->dim = 2; c->nneighbors = 8;
      c->cut[0] = 2.0 * r; c->cut[1] = 2.0 * std::sqrt(2.0) * r;
      c->volume = 4 * (r * r);
      return true;
    }
    case kHexagon2D: {
      const double e[9] = {std::sqrt(3.0) / 12.0, -std::sqrt(3.0) / 12.0, 0., -std::sqrt(3.0) / 144.0,
                           std::sqrt(3.0) / 48.0, 0.0, 0.0, 0.0, 1.0};
      std::memcpy(c->el, e, sizeof e);
      c->dim = 2; c->nneighbors = 12;
      c->cut[0] = 2.0 * r; c->cut[1] = 2.0 * std::sqrt(3.0) * r;
      c->volume = 2 * std::sqrt(3.0) * (r * r);
      return true;
    }
    case kSC3D: {
      const double e[9] = {1, -1, -1.0, 0.0, 0.0, 1., 0.0, 1 / 18.0, -1 / 18.0};
      std::memcpy(c->el, e, sizeof e);
      c->dim = 3; c->nneighbors = 18;
      c->cut[0] = 2.0 * r; c->cut[1] = 2.0 * std::sqrt(2.0) * r;
      c->volume = std::pow(2 * r, 3.0);
      return true;
    }
    case kFCC3D: {
      const double e[9] = {0.0, 0.0, std::sqrt(2.0), std::sqrt(2.0) / 4, -std::sqrt(2.0) / 4,
                           -std::sqrt(2.0) / 4, 0.0, std::sqrt(2.0) / 24, -std::sqrt(2.0) / 24};
      std::memcpy(c->el, e, sizeof e);
      c->dim = 3; c->nneighbors = 18;
      c->cut[0] = 2.0 * r; c->cut[1] = 2.0 * std::sqrt(2.0) * r;
      c->volume = 4.0 * std::sqrt(2.0) * std::pow(r, 3.0);
      return true;
    }
    case kBCC3D: {
      const double e[9] = {0.0, 0.0, std::sqrt(3.0), 1 / std::sqrt(3.0), -1 / std::sqrt(3.0), 0.,
                           0.0, std::sqrt(3.0) / 14, std::sqrt(3.0) / 14};
      std::memcpy(c->el, e, sizeof e);
      c->dim = 3; c->nneighbors = 14;
      c->cut[0] = 2.0 * r; c->cut[1] = 4.0 / std::sqrt(3.0) * r;
      c->volume = 32.0 * std::sqrt(3.0) / 9.0 * std::pow(r, 3.0);
      return true;
    }
  }
  return false;
}

}  // namespace lpmg

using namespace lpmg;

struct lpmg_context {
  Context c;
  // host mirrors of the material arrays (small; uploaded after each change)
  std::vector<double> Kn, Tv, crit, nlL, mE, mMu, mSy, mXi, mH, mA, k0, k1, dthr, ctr, fA, fB, fD;
};

namespace lpmg {
Context* handle_context(lpmg_context* h) { return &h->c; }
}  // namespace lpmg

namespace {

#define API_HIP(call)                                  \
  do {                                                 \
    hipError_t _e = (call);                            \
    if (_e != hipSuccess) return hip_fail(_e, #call);  \
  } while (0)

#define CHECK_HANDLE(h)                  \
  do {                                   \
    if (!(h)) {                          \
      set_error("null handle");          \
      return -1;                         \
    }                                    \
  } while (0)

__global__ void pack_xyz_kernel(const double* __restrict__ xyz, double4* __restrict__ out, int N) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) out[i] = make_double4(xyz[3 * (size_t)i], xyz[3 * (size_t)i + 1], xyz[3 * (size_t)i + 2], 0.0);
}
__global__ void unpack_xyz_kernel(const double4* __restrict__ in, double* __restrict__ xyz, int N) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) {
    const double4 p = in[i];
    xyz[3 * (size_t)i] = p.x; xyz[3 * (size_t)i + 1] = p.y; xyz[3 * (size_t)i + 2] = p.z;
  }
}
__global__ void d2f_kernel(const double* __restrict__ a, float* __restrict__ b, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = (float)a[t];
}
__global__ void f2d_kernel(const float* __restrict__ a, double* __restrict__ b, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = (double)a[t];
}
// Plain streaming copy of known size (read n doubles, write n doubles): the
// reference kernel that calibrates HBM byte counters against algorithmic bytes.
__global__ void __launch_bounds__(256) stream_copy_kernel(const double* __restrict__ a,
                                                          double* __restrict__ b, int64_t n) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (int64_t)gridDim.x * blockDim.x)
    b[t] = a[t];
}
// Plain streaming read of known size (n2 16-byte words summed, one partial
// per workgroup): the attainable read rate that a read-dominated kernel such as
// the SpMV is measured against.  Each workgroup reads one contiguous chunk of
// kReadChunk words, four loads in flight per thread -- the fastest of a sweep
// of read kernels on a 3 GiB buffer (grid-stride loops with 2,048-32,768
// workgroups and 4 or 8 loads in flight: 5.5-6.2 TB/s; chunks of 4,096-65,536
// words: 6.26-6.37 TB/s, 79.6% of the 8 TB/s peak).
constexpr int kReadChunk = 4096;
__global__ void __launch_bounds__(256) stream_read_kernel(const double2* __restrict__ a, int64_t n2,
                                                          double* __restrict__ part) {
  const int64_t lo = (int64_t)blockIdx.x * kReadChunk;
  const int64_t hi = lo + kReadChunk < n2 ? lo + kReadChunk : n2;
  double s0 = 0, s1 = 0, s2 = 0, s3 = 0;
  int64_t t = lo + threadIdx.x;
  for (; t + 3 * 256 < hi; t += 4 * 256) {
    const double2 v0 = a[t], v1 = a[t + 256], v2 = a[t + 512], v3 = a[t + 768];
    s0 += v0.x + v0.y;
    s1 += v1.x + v1.y;
    s2 += v2.x + v2.y;
    s3 += v3.x + v3.y;
  }
  for (; t < hi; t += 256) s0 += a[t].x + a[t].y;
  __shared__ double red[256];
  red[threadIdx.x] = (s0 + s1) + (s2 + s3);
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
__global__ void i2d_kernel(const int* __restrict__ a, double* __restrict__ b, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = (double)a[t];
}
__global__ void d2i_kernel(const double* __restrict__ a, int* __restrict__ b, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) b[t] = (int)a[t];
}

int upload_material(lpmg_context* h) {
  Context* c = &h->c;
  const size_t N = c->N;
  struct {
    double* d;
    std::vector<double>* v;
  } m[] = {{c->Kn, &h->Kn}, {c->Tv, &h->Tv}, {c->crit, &h->crit}, {c->nlL, &h->nlL}, {c->mE, &h->mE},
           {c->mMu, &h->mMu}, {c->mSy, &h->mSy}, {c->mXi, &h->mXi}, {c->mH, &h->mH}, {c->mA, &h->mA},
           {c->k0, &h->k0}, {c->k1, &h->k1}, {c->dthr, &h->dthr}, {c->ctr, &h->ctr},
           {c->fA, &h->fA}, {c->fB, &h->fB}, {c->fD, &h->fD}};
  for (auto& e : m)
    API_HIP(hipMemcpyAsync(e.d, e.v->data(), sizeof(double) * e.v->size(), hipMemcpyHostToDevice, c->stream));
  API_HIP(hipStreamSynchronize(c->stream));
  (void)N;
  return 0;
}

int alloc_context(lpmg_context* h, const double* xyz) {
  Context* c = &h->c;
  const size_t N = c->N, d = c->dim;
  API_HIP(hipGetDevice(&c->device));
  API_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
  int rc = 0;
  rc |= dalloc(c, &c->xyz, N);
  rc |= dalloc(c, &c->xyz0, N);
  rc |= dalloc(c, &c->xyz_last, N);
  for (double** p : {&c->Pex, &c->Pin})
    rc |= dalloc(c, p, 3 * N);
  rc |= dalloc(c, &c->stress, 6 * N);
  for (double** p : {&c->damage, &c->damage_last, &c->damage_visual, &c->Ddot_local, &c->Ddot_nonlocal,
                     &c->Tv, &c->crit, &c->nlL, &c->mE, &c->mMu, &c->mSy, &c->mXi, &c->mH, &c->mA,
                     &c->k0, &c->k1, &c->dthr, &c->ctr, &c->fA, &c->fB, &c->fD, &c->diag_sq})
    rc |= dalloc(c, p, N);
  rc |= dalloc(c, &c->sv, 7 * N);
  rc |= dalloc(c, &c->sv_last, 7 * N);
  rc |= dalloc(c, &c->dLe_tot, 2 * N);
  rc |= dalloc(c, &c->TdLe_tot, 2 * N);
  rc |= dalloc(c, &c->cs_sum, 6 * N);
  rc |= dalloc(c, &c->cs_wsum, 6 * N);
  rc |= dalloc(c, &c->Kn, 2 * N);
  rc |= dalloc(c, &c->type, N);
  rc |= dalloc(c, &c->floating, N);
  rc |= dalloc(c, &c->cons, 3 * N);
  for (double** p : {&c->residual, &c->du, &c->cg_r, &c->cg_p, &c->cg_q, &c->cg_z, &c->cg_w, &c->dinv})
    rc |= dalloc(c, p, d * N);
  for (float** p : {&c->f_r, &c->f_p, &c->f_q, &c->f_z, &c->f_w, &c->f_x, &c->f_dinv})
    rc |= dalloc(c, p, d * N);
  // reduction partials: two vector-kernel buffers (2 x 2048), SpMV partials
  // (one per SpMV workgroup of >= 16 rows, plus a second-level buffer), and the
  // per-workgroup partials of the N-sized reductions (2 per 256 particles)
  c->n_partials = std::max<int>((int)((N + 255) / 256), 4096) + (int)((N + 7) / 8) + 4096 + 64;
  rc |= dalloc(c, &c->partials, 2 * (size_t)c->n_partials);
  rc |= dalloc(c, &c->sc, 1);
  rc |= dalloc(c, &c->err_flag, 1);
  if (rc) return -3;
  API_HIP(hipHostMalloc((void**)&c->sc_host, sizeof(Scalars), hipHostMallocDefault));
  API_HIP(hipMemsetAsync(c->sc, 0, sizeof(Scalars), c->stream));
  API_HIP(hipMemsetAsync(c->err_flag, 0, sizeof(int), c->stream));
  for (double* p : {c->Pex, c->Pin})
    API_HIP(hipMemsetAsync(p, 0, sizeof(double) * 3 * N, c->stream));
  API_HIP(hipMemsetAsync(c->stress, 0, sizeof(double) * 6 * N, c->stream));
  for (double* p : {c->damage, c->damage_last, c->damage_visual, c->Ddot_local, c->Ddot_nonlocal})
    API_HIP(hipMemsetAsync(p, 0, sizeof(double) * N, c->stream));
  if (int rc = launch_init_state_var(c)) return rc;
  API_HIP(hipMemsetAsync(c->type, 0, sizeof(int) * N, c->stream));
  API_HIP(hipMemsetAsync(c->floating, 0, sizeof(int) * N, c->stream));
  API_HIP(hipMemsetAsync(c->cons, 0, sizeof(int) * 3 * N, c->stream));
  for (double* p : {c->residual, c->du, c->cg_r, c->cg_p, c->cg_q, c->cg_z, c->cg_w, c->dinv})
    API_HIP(hipMemsetAsync(p, 0, sizeof(double) * d * N, c->stream));
  for (float* p : {c->f_r, c->f_p, c->f_q, c->f_z, c->f_w, c->f_x, c->f_dinv})
    API_HIP(hipMemsetAsync(p, 0, sizeof(float) * d * N, c->stream));
  for (auto* v : {&h->Tv, &h->crit, &h->nlL, &h->mE, &h->mMu, &h->mSy, &h->mXi, &h->mH, &h->mA, &h->k0,
                  &h->k1, &h->dthr, &h->ctr, &h->fA, &h->fB, &h->fD})
    v->assign(N, 0.0);
  h->Kn.assign(2 * N, 0.0);
  // positions
  double* tmp = nullptr;
  API_HIP(hipMallocAsync((void**)&tmp, sizeof(double) * 3 * N, c->stream));
  API_HIP(hipMemcpyAsync(tmp, xyz, sizeof(double) * 3 * N, hipMemcpyHostToDevice, c->stream));
  pack_xyz_kernel<<<(c->N + 255) / 256, 256, 0, c->stream>>>(tmp, c->xyz, c->N);
  API_HIP(hipGetLastError());
  API_HIP(hipMemcpyAsync(c->xyz0, c->xyz, sizeof(double4) * N, hipMemcpyDeviceToDevice, c->stream));
  API_HIP(hipMemcpyAsync(c->xyz_last, c->xyz, sizeof(double4) * N, hipMemcpyDeviceToDevice, c->stream));
  API_HIP(hipFreeAsync(tmp, c->stream));
  if (int r2 = upload_material(h)) return r2;
  return 0;
}

int finish_create(lpmg_context* h) {
  Context* c = &h->c;
  if (int rc = build_topology(c)) return rc;
  if ((int64_t)c->nnzb * c->dim * c->dim >= (int64_t)1 << 31) {
    set_error("stiffness exceeds 2^31 values per handle");
    return -1;
  }
  if (dalloc(c, &c->Kval, (size_t)c->nnzb * c->dim * c->dim)) return -3;
  API_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int free_context(lpmg_context* h) {
  Context* c = &h->c;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (hipGraphExec_t& g : c->cg_graph)
    if (g) (void)hipGraphExecDestroy(g);
  for (hipGraphExec_t& g : c->bicg_graph)
    if (g) (void)hipGraphExecDestroy(g);
  amg_free(c);
  for (void* p : c->allocs) (void)hipFreeAsync(p, c->stream);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->sc_host) (void)hipHostFree(c->sc_host);
  if (c->cg_ready) (void)hipEventDestroy(c->cg_ready);
  if (c->cg_released) (void)hipEventDestroy(c->cg_released);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  return 0;
}

int validate_create(int lattice, double radius, int ptype, int n, const double* xyz, lpmg_handle* out) {
  Cell cell;
  if (!out) { set_error("null output handle"); return -1; }
  *out = nullptr;
  if (!make_cell(lattice, radius, &cell)) { set_error("unknown lattice"); return -1; }
  if (!(radius > 0) || n <= 0 || !xyz || ptype < 0 || ptype > 3) {
    set_error("bad radius, particle count, coordinates or particle type");
    return -1;
  }
  return 0;
}

lpmg_context* new_context(int lattice, double radius, int ptype, int n) {
  auto* h = new lpmg_context;
  make_cell(lattice, radius, &h->c.cell);
  h->c.dim = h->c.cell.dim;
  h->c.ptype = ptype;
  h->c.N = n;
  return h;
}

void parse_opts(const double* o, SolveOpts* so) {
  if (!o) return;
  so->precond = (int)o[0];
  so->rtol = o[1];
  so->atol = o[2];
  so->maxit = (int64_t)o[3];
  so->precision = (int)o[4];
  so->chunk = o[5] > 0 ? (int)o[5] : 16;
  so->predict = o[6] > 0 ? o[6] : 0;
}

int do_update_rr(lpmg_context* h, double* nr, double* nreac) {
  double reac = 0;
  if (int rc = launch_update_rr(&h->c, nr, &reac)) return rc;
  h->c.reac_sumsq += reac;
  *nreac = std::sqrt(h->c.reac_sumsq);
  return 0;
}

int do_solve_linear(lpmg_context* h, const SolveOpts& so, double* stats) {
  Context* c = &h->c;
  int it = 0;
  double rr = 0;
  if (int rc = solve_linear(c, so, &it, &rr)) return rc;
  if (int rc = launch_move_by(c, c->du)) return rc;
  if (stats) { stats[0] = it; stats[1] = rr; }
  const int64_t maxit = so.maxit > 0 ? so.maxit : (int64_t)c->dim * c->N;
  if (it >= maxit && !(rr <= c->sc_host->thresh)) {
    set_error("CG reached the iteration limit");
    return -4;
  }
  return 0;
}

// Solver::NewtonIteration (solver.h:54-90)
int do_newton(lpmg_context* h, int max_iter, double tol, const SolveOpts& so, double* stats) {
  Context* c = &h->c;
  if (int rc = launch_forward(c)) return rc;
  double nr = 0, nreac = 0;
  if (int rc = do_update_rr(h, &nr, &nreac)) return rc;
  const double mult = std::max(nr, nreac);
  int ni = 0;
  double cgit = 0;
  while (nr > tol * mult) {
    if (++ni > max_iter) { ni = max_iter; break; }
    if (!std::isfinite(nr)) break;  // reported below
    double st[2] = {0, 0};
    int rc = do_solve_linear(h, so, st);
    if (rc && rc != -4) return rc;  // a CG that stops at maxit keeps its partial update, as MKL's
    cgit += st[0];
    if ((rc = launch_forward(c))) return rc;
    double dummy;
    if ((rc = do_update_rr(h, &nr, &dummy))) return rc;
  }
  if (stats) { stats[0] = cgit; stats[1] = nr; stats[2] = mult; stats[3] = ni; }
  // A NaN residual would pass the loop test above as converged (the
  // reference's test, solver.h:74, has the same blind spot): report it.
  if (!std::isfinite(nr) || !std::isfinite(mult)) {
    set_error("Newton residual is not finite");
    return -4;
  }
  return ni;
}

size_t field_len(const Context* c, int f) {
  const size_t N = c->N, NB = c->NB;
  switch (f) {
    case LPMG_F_XYZ: case LPMG_F_XYZ_INITIAL: case LPMG_F_XYZ_LAST: case LPMG_F_PIN: case LPMG_F_PEX:
    case LPMG_F_CONSTRAINT: return 3 * N;
    case LPMG_F_STRESS: case LPMG_F_CS_SUM: return 6 * N;
    case LPMG_F_DAMAGE: case LPMG_F_DAMAGE_VISUAL: case LPMG_F_DDOT_LOCAL: case LPMG_F_DDOT_NONLOCAL:
    case LPMG_F_DAMAGE_LAST: case LPMG_F_TYPE: case LPMG_F_FLOATING: return N;
    case LPMG_F_STATE_VAR: return 7 * N;
    case LPMG_F_RESIDUAL: return (size_t)c->dim * N;
    case LPMG_F_BDAMAGE: case LPMG_F_DLP: case LPMG_F_BFORCE: case LPMG_F_DIS: case LPMG_F_BSTRAIN:
    case LPMG_F_DLE: return NB;
    case LPMG_F_CS: return 3 * NB;
    case LPMG_F_DLE_TOTAL: case LPMG_F_TDLE_TOTAL: return 2 * N;
  }
  return 0;
}

double* field_ptr(Context* c, int f) {
  switch (f) {
    case LPMG_F_PIN: return c->Pin;
    case LPMG_F_PEX: return c->Pex;
    case LPMG_F_STRESS: return c->stress;
    case LPMG_F_DAMAGE: return c->damage;
    case LPMG_F_DAMAGE_VISUAL: return c->damage_visual;
    case LPMG_F_STATE_VAR: return c->sv;
    case LPMG_F_RESIDUAL: return c->residual;
    case LPMG_F_BDAMAGE: return c->bdamage;
    case LPMG_F_DLP: return c->dLp;
    case LPMG_F_BFORCE: return c->bforce;
    case LPMG_F_DIS: return c->dis;
    case LPMG_F_CS: return c->cs;
    case LPMG_F_DDOT_LOCAL: return c->Ddot_local;
    case LPMG_F_DDOT_NONLOCAL: return c->Ddot_nonlocal;
    case LPMG_F_DLE_TOTAL: return c->dLe_tot;
    case LPMG_F_TDLE_TOTAL: return c->TdLe_tot;
    case LPMG_F_CS_SUM: return c->cs_sum;
    case LPMG_F_BSTRAIN: return c->bstrain;
    case LPMG_F_DLE: return c->dLe;
    case LPMG_F_DAMAGE_LAST: return c->damage_last;
  }
  return nullptr;
}

double4* xyz_ptr(Context* c, int f) {
  if (f == LPMG_F_XYZ) return c->xyz;
  if (f == LPMG_F_XYZ_INITIAL) return c->xyz0;
  if (f == LPMG_F_XYZ_LAST) return c->xyz_last;
  return nullptr;
}

double assembly_bytes(const Context* c) {
  const double DD = (double)c->dim * c->dim;
  return (double)c->nnzb * (DD * 8.0 + 4.0) + (double)c->NB * (4.0 + 24.0 + 8.0) +
         (double)c->N * (48.0 + 16.0 + 8.0 + 8.0 + 8.0 + 4.0 + 12.0 + 8.0);
}

double forward_bytes(const Context* c) {
  return (double)c->NB * (28.0 + 56.0 + 60.0) + (double)c->N * (32.0 + 40.0 + 80.0 + 24.0 + 48.0 + 8.0 + 16.0);
}

}  // namespace

// Allocation-free SpMV: the same kernels as lpmg_spmv, with the staged x and y
// (and their FP32 copies) in one file-scope device array instead of
// stream-ordered allocations.  Each GPU holds its own copy of the array from
// module load until the process ends; its address on a device comes from
// hipGetSymbolAddress on the first call there (the array's name is a host-side
// stand-in that cannot be passed to a kernel).  Every call on one GPU shares
// the array, so calls that overlap on a GPU race; nothing in it needs clearing
// (x is copied in whole and every row of y is written).
namespace {
constexpr int64_t kScratchMaxDofs = int64_t(1) << 20;
constexpr size_t kScratchBytes = (size_t)kScratchMaxDofs * (2 * sizeof(double) + 2 * sizeof(float));
constexpr int kScratchDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kScratchDevices] = {};

int scratch_on_device(int dev, unsigned char** out) {
  if (dev < 0 || dev >= kScratchDevices) {
    set_error("allocation-free entry points: device ordinal out of range");
    return -1;
  }
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int cur = -1;
    API_HIP(hipGetDevice(&cur));
    if (cur != dev) API_HIP(hipSetDevice(dev));
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (cur != dev) API_HIP(hipSetDevice(cur));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    g_scratch_ptr[dev] = (unsigned char*)p;
  }
  *out = g_scratch_ptr[dev];
  return 0;
}
}  // namespace

extern "C" {

int lpmg_version(void) { return 1; }

const char* lpmg_last_error(void) { return g_err.c_str(); }

int lpmg_unit_cell(int lattice, double radius, double* out) {
  Cell c;
  if (!out || !make_cell(lattice, radius, &c)) { set_error("unknown lattice"); return -1; }
  out[0] = c.dim; out[1] = c.nneighbors; out[2] = c.volume; out[3] = c.cut[0]; out[4] = c.cut[1];
  for (int k = 0; k < 9; ++k) out[5 + k] = c.el[k];
  return 0;
}

int lpmg_create(int lattice, double radius, int ptype, int n, const double* xyz, lpmg_handle* out) {
  if (int rc = validate_create(lattice, radius, ptype, n, xyz, out)) return rc;
  lpmg_context* h = new_context(lattice, radius, ptype, n);
  int rc = alloc_context(h, xyz);
  if (!rc) rc = build_bonds_from_search(&h->c, xyz);
  if (!rc) rc = finish_create(h);
  if (rc) {
    free_context(h);
    delete h;
    return rc;
  }
  *out = h;
  return 0;
}

int lpmg_create_with_bonds(int lattice, double radius, int ptype, int n, const double* xyz, int64_t nbond,
                           const int32_t* p1, const int32_t* p2, const int32_t* layer, lpmg_handle* out) {
  if (int rc = validate_create(lattice, radius, ptype, n, xyz, out)) return rc;
  if (nbond < 0 || (nbond > 0 && (!p1 || !p2 || !layer))) { set_error("bad bond list"); return -1; }
  lpmg_context* h = new_context(lattice, radius, ptype, n);
  int rc = alloc_context(h, xyz);
  if (!rc) rc = build_bonds_from_list(&h->c, nbond, p1, p2, layer);
  if (!rc) rc = finish_create(h);
  if (rc) {
    free_context(h);
    delete h;
    return rc;
  }
  *out = h;
  return 0;
}

void lpmg_destroy(lpmg_handle h) {
  if (!h) return;
  free_context(h);
  delete h;
}

int lpmg_sizes(lpmg_handle h, int64_t* out) {
  CHECK_HANDLE(h);
  const Context& c = h->c;
  out[0] = c.N; out[1] = c.dim; out[2] = c.NB; out[3] = c.nnz_upper; out[4] = c.nnzb;
  out[5] = c.max_nb; out[6] = c.max_nconn; out[7] = c.ptype;
  return 0;
}

int lpmg_export_bonds(lpmg_handle h, int32_t* bptr, int32_t* nb0, int32_t* p2) {
  CHECK_HANDLE(h);
  Context* c = &h->c;
  if (bptr) API_HIP(copy_sync(c, bptr, c->bptr, sizeof(int) * (c->N + 1), hipMemcpyDeviceToHost));
  if (nb0) API_HIP(copy_sync(c, nb0, c->nb0, sizeof(int) * c->N, hipMemcpyDeviceToHost));
  if (p2) API_HIP(copy_sync(c, p2, c->p2, sizeof(int) * c->NB, hipMemcpyDeviceToHost));
  return 0;
}

int lpmg_export_conns(lpmg_handle h, int32_t* cptr, int32_t* conns, int32_t* ncl) {
  CHECK_HANDLE(h);
  Context* c = &h->c;
  if (cptr) API_HIP(copy_sync(c, cptr, c->cptr, sizeof(int) * (c->N + 1), hipMemcpyDeviceToHost));
  if (conns) API_HIP(copy_sync(c, conns, c->conns, sizeof(int) * c->nnzb, hipMemcpyDeviceToHost));
  if (ncl) API_HIP(copy_sync(c, ncl, c->ncl, sizeof(int) * c->N, hipMemcpyDeviceToHost));
  return 0;
}

int lpmg_export_csr(lpmg_handle h, int64_t* Kp, int64_t* IK, int64_t* JK, double* K) {
  CHECK_HANDLE(h);
  Context* c = &h->c;
  API_HIP(hipStreamSynchronize(c->stream));
  if (Kp) API_HIP(copy_sync(c, Kp, c->Kp, sizeof(int64_t) * (c->N + 1), hipMemcpyDeviceToHost));
  if (!IK && !JK && !K) return 0;
  int64_t *dIK = nullptr, *dJK = nullptr;
  double* dK = nullptr;
  const size_t nIK = (size_t)c->dim * c->N + 1, nnz = (size_t)c->nnz_upper;
  API_HIP(hipMallocAsync((void**)&dIK, sizeof(int64_t) * nIK, c->stream));
  API_HIP(hipMallocAsync((void**)&dJK, sizeof(int64_t) * nnz, c->stream));
  API_HIP(hipMallocAsync((void**)&dK, sizeof(double) * nnz, c->stream));
  int rc = launch_export_csr(c, nullptr, dIK, dJK, dK);
  if (!rc) {
    API_HIP(hipStreamSynchronize(c->stream));
    if (IK) API_HIP(copy_sync(c, IK, dIK, sizeof(int64_t) * nIK, hipMemcpyDeviceToHost));
    if (JK) API_HIP(copy_sync(c, JK, dJK, sizeof(int64_t) * nnz, hipMemcpyDeviceToHost));
    if (K) API_HIP(copy_sync(c, K, dK, sizeof(double) * nnz, hipMemcpyDeviceToHost));
  }
  (void)hipFreeAsync(dIK, c->stream);
  (void)hipFreeAsync(dJK, c->stream);
  (void)hipFreeAsync(dK, c->stream);
  return rc;
}

int lpmg_set_types(lpmg_handle h, const int32_t* type) {
  CHECK_HANDLE(h);
  API_HIP(copy_sync(&h->c, h->c.type, type, sizeof(int) * h->c.N, hipMemcpyHostToDevice));
  return 0;
}

// Particle*::setParticleProperty: Ce from (E, mu) or (C11, C12, C44), KnTv =
// alpha * el_mapping * Ce (alpha = 1 in 2D, radius in 3D; particle_elastic.h:87-90),
// hexagon bonds share Kn = KnTv[0], Tv = KnTv[1], other lattices Kn = KnTv[layer],
// Tv = KnTv[2] (particle_elastic.h:92-107).
int lpmg_set_material(lpmg_handle h, int n_idx, const int32_t* idx, int model, const double* p) {
  CHECK_HANDLE(h);
  Context* c = &h->c;
  if (!p || model < 0 || model > 5) { set_error("unknown material model"); return -1; }
  if ((model == LPMG_MAT_FATIGUE) != (c->ptype == kFatigue)) {
    set_error("the fatigue material needs fatigue particles and vice versa");
    return -1;
  }
  double Ce[3];
  if (model == LPMG_MAT_ELASTIC || model == LPMG_MAT_DAMAGE || model == LPMG_MAT_J2 ||
      model == LPMG_MAT_FATIGUE) {
    const double E = p[2], mu = p[3];
    if (!(int)p[1]) {
      Ce[0] = E * (1.0 - mu) / (1.0 + mu) / (1.0 - 2.0 * mu);
      Ce[1] = E * mu / (1.0 + mu) / (1.0 - 2.0 * mu);
      Ce[2] = E / 2.0 / (1.0 + mu);
    } else {
      Ce[0] = E / (1.0 - mu) / (1.0 + mu);
      Ce[1] = E * mu / (1.0 + mu) / (1.0 - mu);
      Ce[2] = E / 2.0 / (1.0 + mu);
    }
  } else {
    Ce[0] = p[1]; Ce[1] = p[2]; Ce[2] = p[3];
  }
  const double alpha = c->cell.dim == 2 ? 1.0 : c->cell.radius;
  double KnTv[3];
  for (int r = 0; r < 3; ++r) {
    double s = 0;
    for (int l = 0; l < 3; ++l) s += (alpha * Ce[l]) * c->cell.el[3 * r + l];
    KnTv[r] = s;
  }
  const int n = idx ? n_idx : c->N;
  for (int t = 0; t < n; ++t) {
    const int i = idx ? idx[t] : t;
    if (i < 0 || i >= c->N) { set_error("particle index out of range"); return -1; }
    if (c->cell.lattice == kHexagon2D) {
      h->Kn[2 * i] = KnTv[0]; h->Kn[2 * i + 1] = KnTv[0]; h->Tv[i] = KnTv[1];
    } else {
      h->Kn[2 * i] = KnTv[0]; h->Kn[2 * i + 1] = KnTv[1]; h->Tv[i] = KnTv[2];
    }
    h->nlL[i] = p[0];
    if (model == LPMG_MAT_ELASTIC || model == LPMG_MAT_ELASTIC_CUBIC) h->crit[i] = p[4];
    if (model == LPMG_MAT_DAMAGE || model == LPMG_MAT_DAMAGE_CUBIC) {
      h->k0[i] = p[4]; h->k1[i] = p[5]; h->ctr[i] = p[6]; h->dthr[i] = p[7];
    }
    if (model == LPMG_MAT_J2) {
      h->mE[i] = p[2]; h->mMu[i] = p[3]; h->mSy[i] = p[4]; h->mXi[i] = p[5];
      h->mH[i] = p[6]; h->mA[i] = p[7]; h->crit[i] = p[8];
    }
    if (model == LPMG_MAT_FATIGUE) {  // particle_fatigue_hcf.h:163-172
      h->mE[i] = p[2]; h->mMu[i] = p[3]; h->fA[i] = p[4]; h->fB[i] = p[5];
      h->fD[i] = p[6]; h->dthr[i] = p[7];
    }
  }
  return upload_material(h);
}

int lpmg_forward(lpmg_handle h) {
  CHECK_HANDLE(h);
  return launch_forward(&h->c);
}

int lpmg_update_rr(lpmg_handle h, double* norms) {
  CHECK_HANDLE(h);
  double nr, nreac;
  if (int rc = do_update_rr(h, &nr, &nreac)) return rc;
  if (norms) { norms[0] = nr; norms[1] = nreac; }
  return 0;
}

int lpmg_assemble(lpmg_handle h, int precision) {
  CHECK_HANDLE(h);
  return launch_assemble(&h->c, precision ? 1 : 0);
}

int lpmg_spmv(lpmg_handle h, int precision, const double* x, double* y) {
  CHECK_HANDLE(h);
  Context* c = &h->c;
  const int64_t n = (int64_t)c->dim * c->N;
  if (precision == 0 && !c->assembled64) { set_error("stiffness not assembled"); return -1; }
  if (precision == 1 && !c->assembled32) { set_error("FP32 stiffness not assembled"); return -1; }
  double *dx = nullptr, *dy = nullptr;
  float *fx = nullptr, *fy = nullptr;
  API_HIP(hipMallocAsync((void**)&dx, sizeof(double) * n, c->stream));
  API_HIP(hipMallocAsync((void**)&dy, sizeof(double) * n, c->stream));
  API_HIP(hipMemcpyAsync(dx, x, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  int rc;
  const int nb = (int)((n + 255) / 256);
  if (precision == 0) {
    rc = launch_spmv(c, 0, dx, dy);
  } else {
    API_HIP(hipMallocAsync((void**)&fx, sizeof(float) * n, c->stream));
    API_HIP(hipMallocAsync((void**)&fy, sizeof(float) * n, c->stream));
    d2f_kernel<<<nb, 256, 0, c->stream>>>(dx, fx, n);
    rc = launch_spmv(c, 1, fx, fy);
    f2d_kernel<<<nb, 256, 0, c->stream>>>(fy, dy, n);
  }
  if (!rc) {
    API_HIP(hipMemcpyAsync(y, dy, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    API_HIP(hipStreamSynchronize(c->stream));
  }
  (void)hipFreeAsync(dx, c->stream); (void)hipFreeAsync(dy, c->stream);
  if (fx) (void)hipFreeAsync(fx, c->stream);
  if (fy) (void)hipFreeAsync(fy, c->stream);
  return rc;
}

int lpmg_spmv_noalloc(lpmg_handle h, int precision, const double* x, double* y) {
  CHECK_HANDLE(h);
  Context* c = &h->c;
  const int64_t n = (int64_t)c->dim * c->N;
  if (n > kScratchMaxDofs) {
    set_error("lpmg_spmv_noalloc: " + std::to_string(n) + " unknowns exceed the " +
              std::to_string(kScratchMaxDofs) + " its static scratch holds (use lpmg_spmv)");
    return -1;
  }
  if (precision == 0 && !c->assembled64) { set_error("stiffness not assembled"); return -1; }
  if (precision == 1 && !c->assembled32) { set_error("FP32 stiffness not assembled"); return -1; }
  unsigned char* base = nullptr;
  if (int rc = scratch_on_device(c->device, &base)) return rc;
  double* dx = (double*)base;
  double* dy = dx + kScratchMaxDofs;
  float* fx = (float*)(dy + kScratchMaxDofs);
  float* fy = fx + kScratchMaxDofs;
  API_HIP(hipMemcpyAsync(dx, x, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  int rc;
  const int nb = (int)((n + 255) / 256);
  if (precision == 0) {
    rc = launch_spmv(c, 0, dx, dy);
  } else {
    d2f_kernel<<<nb, 256, 0, c->stream>>>(dx, fx, n);
    rc = launch_spmv(c, 1, fx, fy);
    f2d_kernel<<<nb, 256, 0, c->stream>>>(fy, dy, n);
  }
  if (rc) return rc;
  API_HIP(hipMemcpyAsync(y, dy, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  API_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int lpmg_solve_linear(lpmg_handle h, const double* opts, double* stats) {
  CHECK_HANDLE(h);
  SolveOpts so;
  parse_opts(opts, &so);
  return do_solve_linear(h, so, stats);
}

int lpmg_newton(lpmg_handle h, int max_iter, double tol, const double* opts, double* stats) {
  CHECK_HANDLE(h);
  SolveOpts so;
  parse_opts(opts, &so);
  return do_newton(h, max_iter, tol, so, stats);
}

int lpmg_update_state_var(lpmg_handle h) {
  CHECK_HANDLE(h);
  return launch_update_state_var(&h->c);
}

int lpmg_update_static_damage(lpmg_handle h, int undamaged) {
  CHECK_HANDLE(h);
  int any = 0;
  if (int rc = launch_static_damage(&h->c, undamaged, &any)) return rc;
  return any ? 1 : 0;
}

int lpmg_update_fatigue_damage(lpmg_handle h, int undamaged, double dNdt) {
  CHECK_HANDLE(h);
  int any = 0;
  if (int rc = launch_fatigue_damage(&h->c, undamaged, dNdt, &any)) return rc;
  return any ? 1 : 0;
}

int lpmg_store_state(lpmg_handle h) {
  CHECK_HANDLE(h);
  return store_state(&h->c);
}

int lpmg_reset_state(lpmg_handle h, int reset_xyz) {
  CHECK_HANDLE(h);
  return reset_state(&h->c, reset_xyz);
}

int lpmg_search_nonlocal(lpmg_handle h, double ratio) {
  CHECK_HANDLE(h);
  return search_nonlocal(&h->c, ratio);
}

int lpmg_export_nonlocal(lpmg_handle h, int32_t* nlptr, int32_t* nlidx) {
  CHECK_HANDLE(h);
  Context* c = &h->c;
  if (!c->nlptr || !c->nlidx) { set_error("no nonlocal neighbours: lpmg_search_nonlocal first"); return -1; }
  if (nlptr) API_HIP(copy_sync(c, nlptr, c->nlptr, sizeof(int) * (c->N + 1), hipMemcpyDeviceToHost));
  if (nlidx && c->n_nl > 0) API_HIP(copy_sync(c, nlidx, c->nlidx, sizeof(int) * c->n_nl, hipMemcpyDeviceToHost));
  return 0;
}

int lpmg_define_group(lpmg_handle h, int n, const int32_t* idx) {
  CHECK_HANDLE(h);
  Context* c = &h->c;
  if (n < 0 || (n > 0 && !idx)) { set_error("bad group"); return -1; }
  for (int t = 0; t < n; ++t)
    if (idx[t] < 0 || idx[t] >= c->N) { set_error("group index out of range"); return -1; }
  Group g;
  g.n = n;
  if (dalloc(c, &g.idx, (size_t)std::max(n, 1))) return -3;
  if (n) API_HIP(copy_sync(c, g.idx, idx, sizeof(int) * n, hipMemcpyHostToDevice));
  c->groups.push_back(g);
  return (int)c->groups.size() - 1;
}

int lpmg_disp_bc(lpmg_handle h, int gid, int mode, int flag, double step) {
  CHECK_HANDLE(h);
  Context* c = &h->c;
  if (gid < 0 || gid >= (int)c->groups.size()) { set_error("bad group id"); return -1; }
  const int k = flag - 'x';
  if (k < 0 || k > 2 || (mode != 0 && mode != 1)) { set_error("bad flag or load mode"); return -1; }
  return launch_disp_bc(c, c->groups[gid], mode, k, step);
}

int lpmg_force_bc(lpmg_handle h, int gid, int mode, double fx, double fy, double fz) {
  CHECK_HANDLE(h);
  Context* c = &h->c;
  if (gid < 0 || gid >= (int)c->groups.size()) { set_error("bad group id"); return -1; }
  if (mode != 0 && mode != 1) { set_error("bad load mode"); return -1; }
  return launch_force_bc(c, c->groups[gid], mode, fx, fy, fz);
}

// SolverStatic::solveProblemStep (solver_static.h:97-152)
int lpmg_solve_step(lpmg_handle h, int undamaged, int max_iter, double tol, const double* opts,
                    double* stats) {
  CHECK_HANDLE(h);
  Context* c = &h->c;
  SolveOpts so;
  parse_opts(opts, &so);
  int n_newton = 0, passes = 0;
  double cgit = 0;
  bool new_damaged = false;
  if (so.predict > 0 && c->inc_valid)
    if (int rc = launch_predict(c, so.predict)) return rc;
  struct StepScope {
    Context* c;
    explicit StepScope(Context* x) : c(x) {
      c->in_step = true;
      ++c->step_serial;
    }
    ~StepScope() { c->in_step = false; }
  } scope(c);
  amg_step_begins(c);
  do {
    if (int rc = launch_assemble(c, so.precision)) return rc;
    double st[4] = {0, 0, 0, 0};
    n_newton = do_newton(h, max_iter, tol, so, st);
    if (n_newton < 0) return n_newton;
    cgit += st[0];
    ++passes;
    if (n_newton >= max_iter) break;
    if (int rc = launch_update_state_var(c)) return rc;
    int any = 0;
    if (int rc = launch_static_damage(c, undamaged, &any)) return rc;
    new_damaged = any != 0;
    if (int rc = store_state(c)) return rc;
    if (new_damaged)
      if (int rc = launch_forward(c)) return rc;
  } while (new_damaged);
  if (stats) { stats[0] = cgit; stats[1] = passes; stats[2] = n_newton; }
  if (so.predict > 0 && n_newton < max_iter)
    if (int rc = record_step_end(c)) return rc;
  API_HIP(hipStreamSynchronize(c->stream));
  return n_newton < max_iter ? 1 : 0;
}

int64_t lpmg_field_len(lpmg_handle h, int f) {
  if (!h) return -1;
  return (int64_t)field_len(&h->c, f);
}

int lpmg_get_field(lpmg_handle h, int f, double* out) {
  CHECK_HANDLE(h);
  Context* c = &h->c;
  const size_t n = field_len(c, f);
  if (!n) { set_error("unknown field"); return -1; }
  if (double4* p = xyz_ptr(c, f)) {
    double* tmp = nullptr;
    API_HIP(hipMallocAsync((void**)&tmp, sizeof(double) * n, c->stream));
    unpack_xyz_kernel<<<(c->N + 255) / 256, 256, 0, c->stream>>>(p, tmp, c->N);
    API_HIP(hipMemcpyAsync(out, tmp, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    API_HIP(hipFreeAsync(tmp, c->stream));
  } else if (f == LPMG_F_CONSTRAINT || f == LPMG_F_TYPE || f == LPMG_F_FLOATING) {
    double* tmp = nullptr;
    API_HIP(hipMallocAsync((void**)&tmp, sizeof(double) * n, c->stream));
    const int* src = f == LPMG_F_TYPE ? c->type : f == LPMG_F_FLOATING ? c->floating : c->cons;
    i2d_kernel<<<(int)((n + 255) / 256), 256, 0, c->stream>>>(src, tmp, n);
    API_HIP(hipMemcpyAsync(out, tmp, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    API_HIP(hipFreeAsync(tmp, c->stream));
  } else {
    API_HIP(hipMemcpyAsync(out, field_ptr(c, f), sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
  }
  API_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int lpmg_set_field(lpmg_handle h, int f, const double* in) {
  CHECK_HANDLE(h);
  h->c.floating_dirty = true;
  ++h->c.break_version;
  Context* c = &h->c;
  const size_t n = field_len(c, f);
  const bool writable = f == LPMG_F_XYZ || f == LPMG_F_XYZ_LAST || f == LPMG_F_PEX || f == LPMG_F_DAMAGE ||
                        f == LPMG_F_STATE_VAR || f == LPMG_F_BDAMAGE || f == LPMG_F_DLP ||
                        f == LPMG_F_CONSTRAINT || f == LPMG_F_DDOT_LOCAL;
  if (!n || !writable) { set_error("field is read-only or unknown"); return -1; }
  double* tmp = nullptr;
  API_HIP(hipMallocAsync((void**)&tmp, sizeof(double) * n, c->stream));
  API_HIP(hipMemcpyAsync(tmp, in, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
  if (double4* p = xyz_ptr(c, f)) {
    c->step_valid = c->inc_valid = false;  // positions no longer follow the load steps
    pack_xyz_kernel<<<(c->N + 255) / 256, 256, 0, c->stream>>>(tmp, p, c->N);
  } else if (f == LPMG_F_CONSTRAINT) {
    d2i_kernel<<<(int)((n + 255) / 256), 256, 0, c->stream>>>(tmp, c->cons, n);
    for (Group& g : c->groups) g.cons_set = 0;
  } else {
    API_HIP(hipMemcpyAsync(field_ptr(c, f), tmp, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
  }
  API_HIP(hipGetLastError());
  API_HIP(hipFreeAsync(tmp, c->stream));
  API_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

int lpmg_restart(lpmg_handle h) {
  CHECK_HANDLE(h);
  Context* c = &h->c;
  const size_t N = c->N, NB = c->NB;
  API_HIP(hipMemcpyAsync(c->xyz, c->xyz0, sizeof(double4) * N, hipMemcpyDeviceToDevice, c->stream));
  API_HIP(hipMemcpyAsync(c->xyz_last, c->xyz0, sizeof(double4) * N, hipMemcpyDeviceToDevice, c->stream));
  for (double* p : {c->Pex, c->Pin})
    API_HIP(hipMemsetAsync(p, 0, sizeof(double) * 3 * N, c->stream));
  API_HIP(hipMemsetAsync(c->cons, 0, sizeof(int) * 3 * N, c->stream));
  for (Group& g : c->groups) g.cons_set = 0;
  for (double* p : {c->damage, c->damage_last, c->Ddot_local, c->Ddot_nonlocal})
    API_HIP(hipMemsetAsync(p, 0, sizeof(double) * N, c->stream));
  if (int rc = launch_init_state_var(c)) return rc;
  for (double* p : {c->dLp, c->dLp_last, c->bdamage, c->bdamage_last, c->bforce, c->bforce_last})
    API_HIP(hipMemsetAsync(p, 0, sizeof(double) * NB, c->stream));
  API_HIP(hipMemcpyAsync(c->dis, c->dis0, sizeof(double) * NB, hipMemcpyDeviceToDevice, c->stream));
  API_HIP(hipMemcpyAsync(c->dis_last, c->dis0, sizeof(double) * NB, hipMemcpyDeviceToDevice, c->stream));
  API_HIP(hipMemsetAsync(c->floating, 0, sizeof(int) * N, c->stream));
  c->floating_dirty = true;
  ++c->break_version;
  c->step_valid = c->inc_valid = false;
  c->reac_sumsq = 0;
  c->assembled64 = c->assembled32 = false;
  API_HIP(hipStreamSynchronize(c->stream));
  return launch_forward(c);
}

int lpmg_profile(lpmg_handle h, int enable) {
  CHECK_HANDLE(h);
  h->c.prof = enable != 0;
  h->c.prof_spmv_ms = 0;
  h->c.prof_spmv_n = 0;
  return 0;
}

int lpmg_profile_read(lpmg_handle h, double* out) {
  CHECK_HANDLE(h);
  out[0] = h->c.prof_spmv_ms;
  out[1] = (double)h->c.prof_spmv_n;
  out[2] = spmv_bytes(&h->c, 0);
  out[3] = spmv_bytes(&h->c, 1);
  return 0;
}

int lpmg_time_kernel(lpmg_handle h, int which, int reps, double* out) {
  CHECK_HANDLE(h);
  Context* c = &h->c;
  if (reps <= 0 || which < 0 || which > 11) { set_error("bad kernel id or repetitions"); return -1; }
  if ((which == 0 || which == 2 || which == 5 || which == 7 || which == 9 || which == 10) && !c->assembled64) {
    set_error("stiffness not assembled");
    return -1;
  }
  if ((which == 1 || which == 6 || which == 8 || which == 11) && !c->assembled32) {
    set_error("FP32 stiffness not assembled");
    return -1;
  }
  if ((which >= 5 && which <= 8) || which >= 10) {  // solver iterations: a fixed-length solve between events
    SolveOpts so;
    so.precond = which <= 6 ? 1 : which <= 8 ? 2 : 4;
    so.precision = (which == 6 || which == 8 || which == 11) ? 1 : 0;
    so.rtol = 0;
    so.atol = 0;
    so.maxit = reps;
    hipEvent_t s0, s1;
    API_HIP(hipEventCreate(&s0));
    API_HIP(hipEventCreate(&s1));
    int it = 0;
    double rr = 0;
    API_HIP(hipEventRecord(s0, c->stream));
    int rc = solve_linear(c, so, &it, &rr);
    API_HIP(hipEventRecord(s1, c->stream));
    API_HIP(hipEventSynchronize(s1));
    float ms = 0;
    API_HIP(hipEventElapsedTime(&ms, s0, s1));
    (void)hipEventDestroy(s0);
    (void)hipEventDestroy(s1);
    if (rc) return rc;
    if (it <= 0) {
      set_error("the solve stopped at once: the residual is zero (load the system and call lpmg_update_rr)");
      return -1;
    }
    const double w = so.precision ? 4.0 : 8.0;
    out[0] = ms / it;
    out[1] = so.precond == 1 ? spmv_bytes(c, so.precision) + 11.0 * w * c->dim * (double)c->N : 0.0;
    return 0;
  }
  const int64_t ncopy = (int64_t)c->nnzb * c->dim * c->dim;
  double* copy_dst = nullptr;
  if (which == 4) API_HIP(hipMallocAsync((void**)&copy_dst, sizeof(double) * ncopy, c->stream));
  double* read_part = nullptr;  // one partial per workgroup of the streaming read
  if (which == 9)
    API_HIP(hipMallocAsync((void**)&read_part, sizeof(double) * (size_t)((ncopy / 2 + kReadChunk - 1) / kReadChunk + 1),
                           c->stream));
  hipEvent_t e0, e1;
  API_HIP(hipEventCreate(&e0));
  API_HIP(hipEventCreate(&e1));
  int rc = 0;
  API_HIP(hipEventRecord(e0, c->stream));
  for (int r = 0; r < reps && !rc; ++r) {
    if (which == 0) rc = launch_spmv(c, 0, c->cg_p, c->cg_q);
    else if (which == 1) rc = launch_spmv(c, 1, c->f_p, c->f_q);
    else if (which == 2) rc = launch_assemble(c, c->assembled32 ? 1 : 0);
    else if (which == 3) rc = launch_forward(c);
    else if (which == 9)
      stream_read_kernel<<<(unsigned)((ncopy / 2 + kReadChunk - 1) / kReadChunk), 256, 0, c->stream>>>(
          (const double2*)c->Kval, ncopy / 2, read_part);
    else stream_copy_kernel<<<2048, 256, 0, c->stream>>>(c->Kval, copy_dst, ncopy);
  }
  API_HIP(hipEventRecord(e1, c->stream));
  API_HIP(hipEventSynchronize(e1));
  float ms = 0;
  API_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (copy_dst) (void)hipFreeAsync(copy_dst, c->stream);
  if (read_part) (void)hipFreeAsync(read_part, c->stream);
  out[0] = ms / reps;
  out[1] = which == 0 ? spmv_bytes(c, 0) : which == 1 ? spmv_bytes(c, 1)
           : which == 2 ? assembly_bytes(c) : which == 3 ? forward_bytes(c)
           : which == 9 ? 8.0 * (double)(ncopy / 2 * 2) : 16.0 * (double)ncopy;
  return rc;
}

int lpmg_multigrid_levels(lpmg_handle h, int32_t* out, int cap) {
  CHECK_HANDLE(h);
  if (cap < 0 || (cap > 0 && !out)) { set_error("bad output"); return -1; }
  return amg_levels(&h->c, out, cap);
}

int lpmg_synchronize(lpmg_handle h) {
  CHECK_HANDLE(h);
  API_HIP(hipStreamSynchronize(h->c.stream));
  return 0;
}

}  // extern "C"
