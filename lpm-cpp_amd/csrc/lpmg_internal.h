// lpmg_internal.h -- device-resident context and kernel entry points of liblpmg.
//
// Layout in HBM (one context = one LPM particle system, ids 0..N-1 local to it):
//   particles   xyz/xyz0/xyz_last as double4 {x,y,z,0}; per-particle SoA arrays
//   bonds       per particle: layer-0 bonds then layer-1 bonds, each ascending by
//               partner id (the order of Assembly::createBonds); SoA per bond
//   conns       sorted 1-/2-hop connection lists; they are also the block-column
//               lists of the full block-CSR (BSR, b = dim) stiffness used by SpMV
//   stiffness   BSR values, b*b row-major per block, FP64 and (mixed) FP32 copies
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace lpmg {

constexpr double kEPS = 1e-6;  // include/lpm.h:41

enum Lattice { kSquare2D = 0, kHexagon2D = 1, kSC3D = 2, kFCC3D = 3, kBCC3D = 4 };
enum PType { kElastic = 0, kElasticDamage = 1, kJ2 = 2, kFatigue = 3 };  // lpm.h:91-97

struct Cell {  // include/unit_cell.h:7-27
  int lattice = 0, dim = 3, nneighbors = 0;
  double radius = 0, volume = 0;
  double cut[2] = {0, 0};
  double el[9] = {0};
};
bool make_cell(int lattice, double r, Cell* c);

// Device scalars of one CG solve and of the reductions around it.
struct Scalars {
  double rz, rr, rr0, alpha, beta, pq, thresh;  // CG (last iteration)
  double rtol, atol;                            // stopping test of the current solve
  double gam[2], alp[2];                        // (r,u) and alpha of iteration i in slot i & 1
  double vsum[2];                               // (r,u), (r,r) of the last vector step, summed
  double rho[2], om[2];                         // BiCGSTAB: rho and omega of iteration i in slot i & 1
  double rnorm_sq, reac_sq;                     // residual / reaction (updateRR)
  double norm_diag;                             // updateStiffnessDispBC
  int iter, done, maxit, any_flag;
  int icur, inext;                              // iteration index of the CG vector step
};

template <typename T>
struct DArr {  // device array owned by the context
  T* p = nullptr;
  size_t n = 0;
};

struct Group {
  int* idx = nullptr;  // device
  int n = 0;
  // Bit k: a displacement condition on component k has set cons[3*i+k] for
  // every member since the constraint array was last cleared or overwritten,
  // so another one cannot change the constrained set (or the fragment labels).
  unsigned cons_set = 0;
};

// One system of a CG launch: its operator, vectors, reduction buffers, scalars
// and launch shape.  Device-visible (the fused launches read a table of them).
struct CgSys {
  const int* cptr;
  const int* conns;
  const void* val;
  const void* dinv;
  void *x, *r, *u, *p, *s, *w;
  double *partV0, *partV1, *partS_raw, *partS;
  Scalars* sc;
  int64_t n;                      // dim * N unknowns
  int N, gs, gv, nS, two_level;   // block rows, SpMV / vector workgroups, SpMV partials
  int sp_begin, red_begin, vec_begin;  // first workgroup in each fused grid
  int cls;                        // batch class: cg_class(dim, fp32, chunks per pass)
};

// Live size of one batch class, read by the fused kernels from device memory
// so that a captured launch (grid sized for a bucket) serves any membership
// that fits it: workgroups past the live counts return at once.
struct CgBatchHdr {
  int nsys, sp_blocks, red_blocks, vec_blocks;
};
constexpr size_t kCgHdrBytes = 256;  // the table follows the header

// The systems of one batch class in a fused launch: header + table in device
// memory (`hdr`, table at hdr + kCgHdrBytes), live counts, and the grids the
// launches use (>= the live counts)
struct CgBatchClass {
  int cls = 0;
  const CgBatchHdr* hdr = nullptr;
  CgBatchHdr live{0, 0, 0, 0};
  int sp_grid = 0, red_grid = 0, vec_grid = 0;
};

constexpr int kCgChunk = 16;     // CG iterations per launch batch (even)
constexpr int kCgClasses = 12;

struct Context;
struct Amg;  // multigrid hierarchy (amg.hip)
// Runs the iterations of a CG whose start cg_begin has enqueued on c->stream;
// returns after ordering c->stream behind the system's last iteration.
typedef int (*CgExec)(void* arg, Context* c, const CgSys& q);

struct Context {
  int device = 0;
  hipStream_t stream = nullptr;
  Cell cell;
  int ptype = 0, N = 0, dim = 3;
  int64_t NB = 0, nnzb = 0, nnz_upper = 0;
  int max_nb = 0, max_nconn = 0;
  int max_cand = 0;  // largest sum over a particle's bonds of (1 + the bond end's same-layer bonds)
  bool bonds_unique = true;  // no particle has two same-layer bonds to one particle, or a bond to itself

  // particles
  double4 *xyz = nullptr, *xyz0 = nullptr, *xyz_last = nullptr;
  // load-step predictor: positions at the end of the last converged load step
  // and that step's displacement increment (allocated on first use)
  double4 *xyz_step = nullptr, *xyz_inc = nullptr;
  bool step_valid = false, inc_valid = false;
  double *Pex = nullptr, *Pin = nullptr, *stress = nullptr;
  double *damage = nullptr, *damage_last = nullptr, *damage_visual = nullptr;
  double *Ddot_local = nullptr, *Ddot_nonlocal = nullptr;
  double *sv = nullptr, *sv_last = nullptr;                 // 7 per particle
  double *dLe_tot = nullptr, *TdLe_tot = nullptr;           // [2i+L]
  double *cs_sum = nullptr;                                 // [(2i+L)*3+c]
  double *cs_wsum = nullptr;                                // [(2i+L)*3+c], sum of (1 - bdamage) * cs
  int *type = nullptr, *cons = nullptr;                     // cons[3i+k]
  // material
  double *Kn = nullptr, *Tv = nullptr;                      // Kn[2i+L], Tv[i]
  double *crit = nullptr, *nlL = nullptr, *mE = nullptr, *mMu = nullptr, *mSy = nullptr,
         *mXi = nullptr, *mH = nullptr, *mA = nullptr, *k0 = nullptr, *k1 = nullptr,
         *dthr = nullptr, *ctr = nullptr;
  double *fA = nullptr, *fB = nullptr, *fD = nullptr;       // fatigue law A, B, d
  // bonds
  int *bptr = nullptr, *nb0 = nullptr, *p2 = nullptr, *rev = nullptr;
  double *dis0 = nullptr, *dis = nullptr, *dLe = nullptr, *bstrain = nullptr, *dLp = nullptr,
         *dLp_last = nullptr, *cs = nullptr, *bforce = nullptr, *bforce_last = nullptr,
         *bdamage = nullptr, *bdamage_last = nullptr, *dis_last = nullptr;
  // conns / BSR pattern
  int *cptr = nullptr, *conns = nullptr, *ncl = nullptr, *tpos = nullptr, *brow = nullptr,
      *dpos = nullptr;
  int* fixm = nullptr;      // per particle: bit k set when DOF k is eliminated (assembly scratch)
  int64_t* Kp = nullptr;
  double* Kval = nullptr;   // nnzb*b*b
  double* Kloc = nullptr;   // nnzb*b*b, unsymmetrised tangent blocks (two-pass assembly only; lazy)
  // one-pass assembly's path plan (topology only; built on first use):
  // per row, path entries at stride plan_E, per connection lane words at stride 64
  long long* plan_e = nullptr;
  int* plan_r = nullptr;
  int* plan_l = nullptr;
  int* plan_n = nullptr;    // N + 1: the last holds the largest slot count of a row
  int plan_E = 0, plan_S = 0;
  float* Kval32 = nullptr;  // nnzb*b*b (mixed precision)
  double* diag_sq = nullptr;
  // nonlocal neighbours
  int *nlptr = nullptr, *nlidx = nullptr;
  int64_t n_nl = 0;
  // solver vectors
  double *residual = nullptr, *du = nullptr, *cg_r = nullptr, *cg_p = nullptr, *cg_q = nullptr,
         *cg_z = nullptr, *cg_w = nullptr, *dinv = nullptr;
  float *f_r = nullptr, *f_p = nullptr, *f_q = nullptr, *f_z = nullptr, *f_w = nullptr, *f_x = nullptr,
        *f_dinv = nullptr;
  double* partials = nullptr;  // per-block reduction partials
  int n_partials = 0;
  Scalars* sc = nullptr;       // device
  Scalars* sc_host = nullptr;  // pinned host mirror
  int* err_flag = nullptr;     // device error flag (setup kernels)
  double reac_sumsq = 0;       // running sum of squares of every reaction pushed
  bool assembled64 = false, assembled32 = false;
  std::vector<Group> groups;
  std::vector<void*> allocs;
  // per-kernel event timing (bench instrumentation)
  bool prof = false;
  hipEvent_t ev[2] = {nullptr, nullptr};
  double prof_spmv_ms = 0;
  int64_t prof_spmv_n = 0;
  int prof_spmv_prec = -1;
  // captured CG chunk per precision (FP64, FP32), with the device pointers
  // and launch shapes captured into it (checked at every launch)
  hipGraphExec_t cg_graph[2] = {nullptr, nullptr};
  std::vector<uintptr_t> cg_graph_sig[2];
  // BiCGSTAB (bicgstab.hip): captured chunk and its own vectors per precision
  hipGraphExec_t bicg_graph[2] = {nullptr, nullptr};
  std::vector<uintptr_t> bicg_graph_sig[2];
  void* bicg_vec[2] = {nullptr, nullptr};  // r~, z, t (3 * dim * N of the precision)
  int sc_host_iter_launched = 0;
  // CG executor (serving runtime); null: the context's own stream
  CgExec cg_exec = nullptr;
  void* cg_exec_arg = nullptr;
  hipEvent_t cg_ready = nullptr, cg_released = nullptr;  // CG start / end of its last fused launch
  size_t bytes = 0;  // device memory held by the context
  int64_t asm_version = 0;  // number of assemblies so far (multigrid values follow it)
  // fragments: component label per particle (intact bonds), 1 for particles of
  // components without a displacement constraint; recomputed at the next
  // assembly after bonds broke or constraints changed
  int *comp = nullptr, *compfix = nullptr, *floating = nullptr, *changed = nullptr;
  bool floating_dirty = true;
  int n_floating = 0;
  // bumped whenever bonds may have broken or healed (damage passes, state
  // resets, field writes): the smoothed hierarchy's aggregates follow the
  // intact bonds
  int64_t break_version = 0;
  // lpmg_solve_step in progress, and the count of load steps begun
  bool in_step = false;
  int64_t step_serial = 0;
  Amg* amg = nullptr;
};

// ---- helpers (api.hip)
void set_error(const std::string& s);
int hip_fail(hipError_t e, const char* what);
#define LPMG_HIP(call)                                   \
  do {                                                   \
    hipError_t _e = (call);                              \
    if (_e != hipSuccess) return hip_fail(_e, #call);    \
  } while (0)

// Stream-ordered allocation on the context's stream: freeing it later (also
// stream-ordered) does not synchronize the device, so handles come and go
// while other handles' work runs.
template <typename T>
int dalloc(Context* c, T** p, size_t n) {
  if (n == 0) n = 1;
  void* q = nullptr;
  hipError_t e = hipMallocAsync(&q, n * sizeof(T), c->stream);
  if (e != hipSuccess) return hip_fail(e, "hipMallocAsync");
  c->allocs.push_back(q);
  c->bytes += n * sizeof(T);
  *p = static_cast<T*>(q);
  return 0;
}

// A DOF eliminated from the linear solve: a displacement constraint or a
// fully damaged particle, as Stiffness::updateStiffnessDispBC
// (stiffness.h:368-410), or a particle of a fragment that intact bonds no
// longer tie to any constrained particle (`floating`, no counterpart in the
// reference, whose stiffness is singular then).
__device__ __forceinline__ bool dof_fixed(const int* cons, const double* dvis, const int* floating, int i, int k) {
  return cons[3 * i + k] == 1 || fabs(dvis[i] - 1.0) < kEPS || floating[i] != 0;
}

// Copy on the context's stream and wait for it.  Never the null stream: that
// would also wait for every other handle's work on the device.
inline hipError_t copy_sync(Context* c, void* dst, const void* src, size_t bytes, hipMemcpyKind kind) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, c->stream);
  if (e != hipSuccess) return e;
  return hipStreamSynchronize(c->stream);
}

// A captured graph bakes device pointers and launch shapes into its nodes;
// these record them at capture and compare them at every launch, so a graph
// whose buffers were replaced (a reallocation, a freed temporary) is never
// replayed: the launch fails with -1 instead.
inline void graph_sig_push(std::vector<uintptr_t>* sig, const void* p) { sig->push_back((uintptr_t)p); }
inline void graph_sig_push(std::vector<uintptr_t>* sig, int64_t v) { sig->push_back((uintptr_t)v); }
std::vector<uintptr_t> cg_graph_signature(const CgSys& q, int chunks);
int check_graph_signature(const std::vector<uintptr_t>& captured, const std::vector<uintptr_t>& now,
                          const char* which);

// the context behind a C-ABI handle (api.hip)
Context* handle_context(struct lpmg_context* h);

// ---- setup.hip
int build_bonds_from_search(Context* c, const double* host_xyz);
int build_bonds_from_list(Context* c, int64_t nbond, const int32_t* p1, const int32_t* p2,
                          const int32_t* layer);
int build_topology(Context* c);  // rev, conns, Kp, BSR helpers
int search_nonlocal(Context* c, double ratio);
int exclusive_scan_i32(Context* c, const int* in, int* out, int n, int* total);
int exclusive_scan_i64(Context* c, const int64_t* in, int64_t* out, int n);

// ---- forward.hip
int launch_forward(Context* c);
int launch_update_rr(Context* c, double* norm_res, double* norm_reac_step_sq);
int launch_disp_bc(Context* c, Group& g, int mode, int k, double step);
int launch_force_bc(Context* c, const Group& g, int mode, double fx, double fy, double fz);
int launch_move_by(Context* c, const double* du);

// ---- assemble.hip
int launch_assemble(Context* c, int want32);
int launch_export_csr(Context* c, int64_t* dKp, int64_t* dIK, int64_t* dJK, double* dK);

// ---- solver.hip
struct SolveOpts {
  int precond = 1;      // 0 plain CG (the reference's MKL call), 1 Jacobi PCG, 2 multigrid PCG,
                        // 3 Jacobi BiCGSTAB, 4 smoothed-aggregation multigrid PCG
  int precision = 0;    // 0 FP64, 1 FP32 inner solve
  double rtol = 1e-12, atol = 1e-12;  // ||r||^2 <= rtol*||r0||^2 + atol
  int64_t maxit = 0;    // 0 -> problem size
  int chunk = 16;       // unused: the chunk length is fixed (graph capture)
  double predict = 0;   // > 0: lpmg_solve_step starts from this multiple of the last step's increment
};
int solve_linear(Context* c, const SolveOpts& o, int* iters, double* rr_final);
// ---- bicgstab.hip
template <int D, typename T>
int run_bicgstab(Context* c, const SolveOpts& o, int* iters, double* rr_final);
// ---- amg.hip
template <int D, typename T>
int run_pcg_amg(Context* c, const SolveOpts& o, int* iters, double* rr_final);
void amg_free(Context* c);
void amg_step_begins(Context* c);  // a load step starts: the aggregates may be rechecked
int amg_levels(Context* c, int* out, int cap);
int cg_class(int dim, int fp32, int chunks);
void cg_batch_layout(CgSys* tab, int nsys, CgBatchHdr* h);
int cg_batch_enqueue_chunk(hipStream_t st, const CgBatchClass* classes, int nclass);
int cg_batch_flags(hipStream_t st, const CgSys* tab, int nsys, int* out);
int launch_spmv(Context* c, int precision, const void* x, void* y);
double spmv_bytes(const Context* c, int precision);

// ---- state.hip
int launch_update_state_var(Context* c);
int launch_static_damage(Context* c, int undamaged, int* any);
int launch_fatigue_damage(Context* c, int undamaged, double dNdt, int* any);
int launch_init_state_var(Context* c);
int launch_predict(Context* c, double scale);
int record_step_end(Context* c);
int update_floating(Context* c);
int store_state(Context* c);
int reset_state(Context* c, int reset_xyz);

}  // namespace lpmg
