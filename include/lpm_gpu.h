/* lpm_gpu.h -- C ABI of liblpmg.so, the MI355X (gfx950) implementation of
 * LPM-CPP's quasi-static Newton step: bond geometry and forces, the analytic
 * tangent stiffness and its assembly, and the linear solve.
 *
 * Every entry point replaces one reference interface; the citation beside each
 * declaration names it (paths relative to the LPM-CPP repository root).  All
 * buffers in these signatures are HOST buffers; the particle system itself
 * (positions, bonds, state, stiffness) lives in device memory for the lifetime
 * of the handle.  Particle ids are local to a handle (0..n-1), so any number of
 * systems can coexist in one process (the reference's process-global
 * Particle::_ID, particle.h:17, allows one).
 *
 * Return values: 0 (or a non-negative count / flag) on success, negative on
 * error: -1 invalid argument or state, -2 HIP error, -3 out of device memory,
 * -4 not converged.  lpmg_last_error() describes the last error of the calling
 * thread.  No entry point calls exit().
 *
 * Threading: one host thread at a time per handle; handles are independent and
 * each owns its own HIP stream.
 */
#ifndef LPM_GPU_H
#define LPM_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lpmg_context* lpmg_handle;

/* LatticeType (include/lpm.h:99-108); the BCC slip variants share one cell. */
enum { LPMG_SQUARE2D = 0, LPMG_HEXAGON2D = 1, LPMG_SC3D = 2, LPMG_FCC3D = 3, LPMG_BCC3D = 4 };
/* ParticleType (include/lpm.h:91-97) */
enum { LPMG_ELASTIC = 0, LPMG_ELASTIC_DAMAGE = 1, LPMG_J2 = 2, LPMG_FATIGUE_HCF = 3 };
/* LoadMode (include/lpm.h:67-71) */
enum { LPMG_RELATIVE = 0, LPMG_ABSOLUTE = 1 };
/* material models of lpmg_set_material */
enum {
  LPMG_MAT_ELASTIC = 0,        /* nonlocal_L, plane_stress, E, mu, critical_bstrain */
  LPMG_MAT_ELASTIC_CUBIC = 1,  /* nonlocal_L, C11, C12, C44, critical_bstrain */
  LPMG_MAT_DAMAGE = 2,         /* nonlocal_L, plane_stress, E, mu, k0, k1, ct_ratio, d_thres */
  LPMG_MAT_DAMAGE_CUBIC = 3,   /* nonlocal_L, C11, C12, C44, k0, k1, ct_ratio, d_thres */
  LPMG_MAT_J2 = 4,             /* nonlocal_L, plane_stress, E, mu, sigmay, xi, H, A, crit */
  LPMG_MAT_FATIGUE = 5         /* nonlocal_L, plane_stress, E, mu, A, B, d, d_thres, f_lmt_ratio
                                  (fatigue particles only; particle_fatigue_hcf.h:163) */
};
/* fields of lpmg_get_field / lpmg_set_field (lengths in doubles) */
enum {
  LPMG_F_XYZ = 0,          /* n*3   Particle::xyz          (set: yes) */
  LPMG_F_XYZ_INITIAL = 1,  /* n*3   Particle::xyz_initial  */
  LPMG_F_XYZ_LAST = 2,     /* n*3   Particle::xyz_last     (set: yes) */
  LPMG_F_PIN = 3,          /* n*3   Particle::Pin          */
  LPMG_F_PEX = 4,          /* n*3   Particle::Pex          (set: yes) */
  LPMG_F_STRESS = 5,       /* n*6   Particle::stress       */
  LPMG_F_DAMAGE = 6,       /* n     Particle::damage       (set: yes) */
  LPMG_F_DAMAGE_VISUAL = 7,
  LPMG_F_STATE_VAR = 8,    /* n*7   Particle::state_var    (set: yes; fatigue particles
                                    use [0] history max, [1] cycle min, [2] cycle max
                                    of the strain energy density) */
  LPMG_F_RESIDUAL = 9,     /* n*dim Stiffness::residual    */
  LPMG_F_BDAMAGE = 10,     /* nbond Bond::bdamage          (set: yes) */
  LPMG_F_DLP = 11,         /* nbond Bond::dLp              (set: yes) */
  LPMG_F_BFORCE = 12,
  LPMG_F_DIS = 13,
  LPMG_F_CS = 14,          /* nbond*3 Bond::csx/csy/csz    */
  LPMG_F_CONSTRAINT = 15,  /* n*3   Particle::disp_constraint (set: yes) */
  LPMG_F_DDOT_LOCAL = 16,  /* n     Particle::Ddot_local   (set: yes) */
  LPMG_F_DDOT_NONLOCAL = 17,
  LPMG_F_DLE_TOTAL = 18,   /* n*2   Particle::dLe_total    */
  LPMG_F_CS_SUM = 19,      /* n*6   Particle::cs_sum{x,y,z}[layer] */
  LPMG_F_BSTRAIN = 20,
  LPMG_F_DLE = 21,
  LPMG_F_DAMAGE_LAST = 22,
  LPMG_F_TYPE = 23,        /* n     Particle::type         */
  LPMG_F_TDLE_TOTAL = 24,  /* n*2   Particle::TdLe_total   */
  LPMG_F_FLOATING = 25     /* n     1: particle of a fragment no intact bond ties to a
                                    constrained particle (held in place; no
                                    reference counterpart), read-only */
};

/* Solver options, opts[8]:
 *   [0] preconditioner: 0 none (the reference's MKL CG, solver.h:311), 1 Jacobi,
 *       2 aggregation multigrid (one V-cycle; built on first use per handle);
 *       3 Jacobi-preconditioned BiCGSTAB instead of CG (no reference
 *       counterpart; same stopping test, -4 with a message on a breakdown);
 *       4 smoothed-aggregation multigrid (as 2, with the prolongation
 *       smoothed by one damped-Jacobi step and Galerkin coarse operators;
 *       aggregates split along broken bonds, in 3D once a bond is broken,
 *       rebuilt at a lpmg_solve_step's first solve when that split changed;
 *       fewer iterations, more setup and memory; no reference counterpart)
 *   [1] rtol, [2] atol: stop when ||r||^2 <= rtol*||r0||^2 + atol
 *       (the reference's dpar[0], dpar[1], solver.h:312-313)
 *   [3] maximum CG iterations (<= 0: the problem size, as ipar[4], solver.h:308)
 *   [4] precision: 0 FP64 stiffness and CG; 1 FP32 stiffness and CG inside the
 *       FP64 Newton (residual, refinement update and state stay FP64)
 *   [5] unused (16 CG iterations run between convergence checks)
 *   [6] load-step predictor (lpmg_solve_step only; no reference counterpart):
 *       s > 0 records where each converged load step ends, and a step after
 *       one so recorded starts with every DOF that the stiffness does not
 *       eliminate moved by s times its displacement over that step (pass
 *       s = this step's load increment / the last converged step's, 1 for
 *       uniform steps).  The equilibrium reached is the same; Newton starts
 *       nearer to it.  0: the reference's start (only the BCs move particles).
 * Solver statistics, stats[4]: [0] CG iterations, [1] final ||r||^2 (CG) or
 * ||r|| (Newton), [2] Newton tolerance multiplier, [3] Newton iterations. */

int lpmg_version(void);
const char* lpmg_last_error(void);

/* UnitCell (include/unit_cell.h:29-122): out[16] = dim, nneighbors, volume,
 * cutoff[0], cutoff[1], el_mapping[9] */
int lpmg_unit_cell(int lattice, double radius, double* out);

/* Assembly(p_xyz, box, cell, ptype) (include/assembly.h:97-106): createParticles,
 * createBonds (assembly.h:212-234), updateConnections (assembly.h:236-266), plus
 * Stiffness::initialize (stiffness.h:54-71).  xyz: n*3. */
int lpmg_create(int lattice, double radius, int ptype, int n, const double* xyz, lpmg_handle* out);

/* Assembly(dumpFile, bondFile, cell, ptype) (include/assembly.h:83-95) with the
 * bond list of Assembly::readBond (assembly.h:390-424): nbond directed bonds
 * (p1, p2, layer) in file order. */
int lpmg_create_with_bonds(int lattice, double radius, int ptype, int n, const double* xyz,
                           int64_t nbond, const int32_t* p1, const int32_t* p2,
                           const int32_t* layer, lpmg_handle* out);

/* Releases all device memory of the handle (the reference never frees its particles). */
void lpmg_destroy(lpmg_handle h);

/* out[8] = n, dim, nbond, nnz of the reference upper CSR, BSR blocks, max bonds
 * per particle, max conns per particle, particle type */
int lpmg_sizes(lpmg_handle h, int64_t* out);

/* Bond lists: bptr[n+1], nb0[n] (layer-0 count), p2[nbond]; per particle the
 * layer-0 bonds precede the layer-1 bonds (Particle::bond_layers, particle.h:33). */
int lpmg_export_bonds(lpmg_handle h, int32_t* bptr, int32_t* nb0, int32_t* p2);

/* Particle::conns / nconn_largeq (particle.h:22,35): cptr[n+1], conns[nnzb], ncl[n] */
int lpmg_export_conns(lpmg_handle h, int32_t* cptr, int32_t* conns, int32_t* ncl);

/* Stiffness::K_pointer, IK, JK, K_global (stiffness.h:24-26): the reference's
 * upper-triangular 1-based CSR of the last assembled stiffness.  Any pointer may be NULL. */
int lpmg_export_csr(lpmg_handle h, int64_t* K_pointer, int64_t* IK, int64_t* JK, double* K);

/* Particle::type for every particle (used by undamaged_pt_type, solver_static.h:23) */
int lpmg_set_types(lpmg_handle h, const int32_t* type);

/* Particle*::setParticleProperty (particle_elastic.h:70-140,
 * particle_elastic_damage.h:175-250, particle_j2plasticity.h:154-198,
 * particle_fatigue_hcf.h:162-206) for the
 * n_idx particles in idx, or for all particles when idx is NULL. */
int lpmg_set_material(lpmg_handle h, int n_idx, const int32_t* idx, int model, const double* params);

/* Assembly::updateGeometry + Assembly::updateForceState (assembly.h:132-138, 178-191) */
int lpmg_forward(lpmg_handle h);

/* Solver::updateRR (solver.h:111-126) and the norms of NewtonIteration
 * (solver.h:63-64): norms[0] = ||residual||, norms[1] = ||all reactions so far||. */
int lpmg_update_rr(lpmg_handle h, double* norms);

/* Stiffness::reset + calcStiffness2D/3D + updateStiffnessDispBC
 * (stiffness.h:73-89, 197-438), from the geometry of the last lpmg_forward.
 * precision 1 also keeps an FP32 copy for the mixed-precision solve. */
int lpmg_assemble(lpmg_handle h, int precision);

/* y = K x with the assembled stiffness (the reference's mkl_sparse_d_mv on the
 * symmetric upper CSR, solver.h:334).  x, y: n*dim host buffers. */
int lpmg_spmv(lpmg_handle h, int precision, const double* x, double* y);

/* lpmg_spmv without any device allocation: no counterpart in the reference.
 * For callers that cannot allocate (tight serving loops, harnesses that forbid
 * hipMalloc).  Same kernels and bit-identical result; x and y are staged in one
 * 24 MiB static device array per GPU, so n*dim must be at most 1,048,576
 * (larger systems return -1 with a message: use lpmg_spmv).  Calls that overlap
 * on one GPU share that array and race. */
int lpmg_spmv_noalloc(lpmg_handle h, int precision, const double* x, double* y);

/* Solver::solveLinearSystem (solver.h:92-109): solve K du = residual by
 * (P)CG (replaces LPM_CG solver.h:276-354 and LPM_PARDISO solver.h:190-274),
 * then move every particle by du.  Returns -4 if maxit was reached. */
int lpmg_solve_linear(lpmg_handle h, const double* opts, double* stats);

/* Solver::NewtonIteration (solver.h:54-90); returns the iteration count
 * (max_iter on the abnormal return, like the reference). */
int lpmg_newton(lpmg_handle h, int max_iter, double tol, const double* opts, double* stats);

/* Assembly::updateStateVar (assembly.h:161-166) */
int lpmg_update_state_var(lpmg_handle h);
/* SolverStatic::updateStaticDamage (solver_static.h:31-62); returns 1 if any new damage */
int lpmg_update_static_damage(lpmg_handle h, int undamaged_type);
/* SolverFatigue::updateFatigueDamage (solver_fatigue.h:75-126) with
 * ParticleFatigueHCF::updateParticleFatigueDamage (particle_fatigue_hcf.h:147-160):
 * nonlocal averaging of the damage rate, damage += rate * dNdt (1 at the
 * threshold) and the reset of the cycle's energy extremes.  Fatigue particles
 * only; returns 1 if any damage changed by more than 1e-3, else 0. */
int lpmg_update_fatigue_damage(lpmg_handle h, int undamaged_type, double dNdt);
/* Assembly::storeStateVar / resetStateVar (assembly.h:140-159) */
int lpmg_store_state(lpmg_handle h);
int lpmg_reset_state(lpmg_handle h, int reset_xyz);
/* Assembly::searchNonlocalNeighbors (assembly.h:117-130) */
int lpmg_search_nonlocal(lpmg_handle h, double cutoff_ratio);
/* Particle::neighbors_nonlocal (particle.h:36) as the last lpmg_search_nonlocal
 * left them: nlptr[n+1] (nlptr[n] = total entries) and nlidx[total], each
 * particle's list in id order with the particle itself included, as the
 * reference's loop over pt_sys pushes them (assembly.h:120-127).  Either pointer
 * may be NULL (fetch nlptr first to size nlidx); -1 before any search. */
int lpmg_export_nonlocal(lpmg_handle h, int32_t* nlptr, int32_t* nlidx);

/* A particle group for boundary conditions (DispBC/ForceBC::group, boundary.h:12,31);
 * returns the group id. */
int lpmg_define_group(lpmg_handle h, int n, const int32_t* idx);
/* Solver::updateDisplacementBC for one DispBC (solver.h:128-164); flag 'x','y','z' */
int lpmg_disp_bc(lpmg_handle h, int group, int load_mode, int flag, double step);
/* Solver::updateForceBC for one ForceBC (solver.h:166-188) */
int lpmg_force_bc(lpmg_handle h, int group, int load_mode, double fx, double fy, double fz);

/* SolverStatic::solveProblemStep (solver_static.h:97-152) after the step's BCs
 * were applied with lpmg_force_bc / lpmg_disp_bc.  Returns 1 converged, 0 not
 * converged (the caller halves the step, as SolverStatic::solveProblem,
 * solver_static.h:64-95).  stats: [0] CG iterations, [1] damage passes,
 * [2] Newton iterations of the last pass. */
int lpmg_solve_step(lpmg_handle h, int undamaged_type, int max_iter, double tol,
                    const double* opts, double* stats);

/* Restart from the initial configuration without rebuilding the topology (no
 * reference counterpart; the reference builds a new Assembly per run): positions
 * back to xyz_initial, every load, constraint, damage, plastic and reaction
 * history cleared, then updateGeometry + updateForceState as the examples do
 * before solving (e.g. examples/ex1_isoelasticity_3d_sc.cpp:104-105).  Groups
 * and materials are kept. */
int lpmg_restart(lpmg_handle h);

int64_t lpmg_field_len(lpmg_handle h, int field);
int lpmg_get_field(lpmg_handle h, int field, double* out);
int lpmg_set_field(lpmg_handle h, int field, const double* in);

/* Bench instrumentation (no reference counterpart): while enabled, every CG
 * SpMV dispatch carries start/stop HIP events (hipExtLaunchKernelGGL).
 * lpmg_profile(h, enable) resets the counters; lpmg_profile_read returns
 * out[0] = total SpMV ms, out[1] = number of SpMV launches, out[2] = algorithmic
 * bytes of one FP64 SpMV, out[3] = of one FP32 SpMV. */
int lpmg_profile(lpmg_handle h, int enable);
int lpmg_profile_read(lpmg_handle h, double* out);

/* Time `reps` back-to-back launches of one kernel on the handle's stream with
 * HIP events (no reference counterpart): which 0 FP64 SpMV, 1 FP32 SpMV,
 * 2 assembly (FP64), 3 forward pass, 4 a plain copy of the FP64 stiffness
 * values (byte-counter calibration), 9 a plain streaming read of the FP64
 * stiffness values (the attainable read rate).  out[0] = average ms per
 * launch, out[1] = algorithmic bytes per launch.
 * which 5-8, 10 and 11 time one linear-solver iteration: a solve of the
 * current residual (lpmg_update_rr) stopped after exactly `reps` iterations,
 * its captured chunks replayed as in any solve: 5 FP64 Jacobi CG, 6 FP32
 * Jacobi CG, 7 FP64 multigrid CG, 8 FP32 multigrid CG, 10 FP64 and 11 FP32
 * smoothed-aggregation multigrid CG (the stiffness of that precision
 * assembled).  out[0] = ms per iteration, out[1] = algorithmic
 * bytes of a Jacobi-CG iteration (the SpMV's and 11 vectors of the vector
 * step), 0 for multigrid.  The solver's vectors are overwritten; positions
 * are not moved. */
int lpmg_time_kernel(lpmg_handle h, int which, int reps, double* out);

/* Unknowns per level of the handle's multigrid hierarchy (finest first), built
 * on its first multigrid solve (opts[0] = 2 or 4); returns the number of levels,
 * 0 before.  No reference counterpart (the reference solves with MKL CG or
 * PARDISO). */
int lpmg_multigrid_levels(lpmg_handle h, int32_t* out, int cap);

int lpmg_synchronize(lpmg_handle h);

/* ---- Serving runtime ---------------------------------------------------
 * No counterpart in the reference: the reference runs one Assembly per
 * process (Particle::_ID is a process-wide counter, particle.h:17,79, and
 * Stiffness indexes by it, stiffness.h:57-65).  A server runs many
 * load-stepped quasi-static solves (SolverStatic::solveProblem,
 * solver_static.h:64-95) at once on the current GPU:
 *   - requests wait in a FIFO and are admitted while a worker slot is free
 *     and the request's device-memory estimate fits the budget next to the
 *     admitted ones; a request larger than the whole budget fails at submit
 *     with -3 instead of waiting for ever;
 *   - each admitted request owns a handle (its particle, bond, stiffness and
 *     solver arrays stay in HBM for the whole solve) and a stream; its
 *     forward pass, assembly and state updates run there, overlapping the
 *     other requests' work;
 *   - the CG iterations of every (small enough) request that is in its
 *     linear solve run as fused launches: one SpMV, one vector and (large
 *     systems) one reduction launch per iteration cover all of them, grouped
 *     by dimension, precision and row width.  A request's arithmetic is the
 *     same alone and in any mix: the result is bit-identical.
 * Requests join or leave the fused launches every 16 iterations; a chunk of
 * 16 iterations is one captured hipGraph per bucket of launch sizes. */
typedef struct lpmg_server_s* lpmg_server;

/* One boundary condition of a load step: DispBC (boundary.h:9-25) or
 * ForceBC (boundary.h:27-44) on one of the request's groups. */
typedef struct lpmg_bc {
  int kind;         /* 0 displacement, 1 force */
  int group;        /* index into the request's groups */
  int mode;         /* LPMG_RELATIVE or LPMG_ABSOLUTE */
  int flag;         /* displacement: 'x', 'y' or 'z', as lpmg_disp_bc */
  double value[3];  /* displacement: value[0] per step; force: fx, fy, fz */
} lpmg_bc;

/* A field of the final state copied into caller memory (lpmg_get_field) */
typedef struct lpmg_out_field {
  int field;         /* LPMG_F_* */
  int64_t capacity;  /* doubles available at data; too small fails the request */
  double* data;
} lpmg_out_field;

/* One quasi-static solve.  Every input array is copied by
 * lpmg_server_submit; the output arrays must stay valid until
 * lpmg_server_wait reports the request finished. */
typedef struct lpmg_request {
  int lattice;                 /* LPMG_SQUARE2D ... */
  double radius;
  int ptype;                   /* LPMG_ELASTIC ... */
  int n;                       /* particles */
  const double* xyz;           /* n*3 */
  int64_t nbond;               /* > 0: bonds given (bond-file order); 0: bond search */
  const int32_t* p1;
  const int32_t* p2;
  const int32_t* layer;
  const int32_t* type;         /* n, or NULL for all 0 */
  int model;                   /* material of every particle, as lpmg_set_material */
  int n_params;
  const double* params;
  double nonlocal_ratio;       /* > 0: searchNonlocalNeighbors(ratio) */
  int prepare_state;           /* 1: updateStateVar after the first updateForceState */
  int n_groups;
  const int32_t* group_size;   /* n_groups sizes */
  const int32_t* group_idx;    /* the groups' particle indices, concatenated */
  int n_bc;                    /* the boundary conditions of one load step ... */
  const lpmg_bc* bc;
  int n_steps;                 /* ... applied in each of n_steps steps */
  int undamaged;               /* undamaged particle type */
  int max_iter;                /* Newton iterations per pass */
  double tol;                  /* Newton tolerance */
  double opts[8];              /* solver options as lpmg_solve_linear; opts[0] < 0:
                                  the server picks fused Jacobi for requests it
                                  fuses, smoothed-aggregation multigrid (4) for
                                  larger 2D and BCC requests and Jacobi for the
                                  larger ones on other 3D lattices */
  int max_halvings;            /* step halvings allowed before the request fails */
  int n_out;
  const lpmg_out_field* out;
  double* step_stats;          /* per converged step: Newton iterations, CG
                                  iterations, damage passes; or NULL */
  int step_stats_cap;          /* steps of room at step_stats */
} lpmg_request;

/* Statistics of a finished request, filled by lpmg_server_wait */
enum {
  LPMG_RS_STATUS = 0,    /* 0 or the error code */
  LPMG_RS_STEPS = 1,     /* converged load steps, halved ones included */
  LPMG_RS_HALVINGS = 2,
  LPMG_RS_NEWTON = 3,    /* Newton iterations of all steps */
  LPMG_RS_CG = 4,        /* CG iterations of all steps */
  LPMG_RS_PASSES = 5,    /* damage passes of all steps */
  LPMG_RS_QUEUE_MS = 6,  /* submit -> admission */
  LPMG_RS_RUN_MS = 7,    /* admission -> finished */
  LPMG_RS_TOTAL_MS = 8,  /* submit -> finished (latency) */
  LPMG_RS_BYTES_EST = 9, /* the admission estimate */
  LPMG_RS_BYTES = 10,    /* device memory the request's handle held */
  LPMG_RS_N = 11
};

/* Device-memory estimate used for admission (an upper bound of what the
 * request's handle allocates); < 0 on a malformed request. */
int64_t lpmg_request_bytes(const lpmg_request* r);

/* max_active: worker slots (requests in flight); mem_budget: bytes (<= 0:
 * 90% of the free device memory at creation); fuse_max_rows: requests with at
 * most this many particles (block rows) run their CG iterations in the fused
 * launches, larger ones on their own stream, where one system already streams
 * HBM at full rate and its phases overlap the fused batch; < 0 fuses every
 * request, 0 none. */
int lpmg_server_create(int max_active, int64_t mem_budget, int64_t fuse_max_rows, lpmg_server* out);
/* Queue a request; *ticket identifies it.  -1 malformed, -3 larger than the
 * budget (never admissible). */
int lpmg_server_submit(lpmg_server s, const lpmg_request* r, int64_t* ticket);
/* Wait up to timeout_s seconds (< 0: no limit) for a request: 0 finished and
 * succeeded, 1 still running (timeout), < 0 finished with that error code
 * (message in lpmg_last_error) or unknown ticket.  stats[LPMG_RS_N] is filled
 * when finished; a finished ticket is released. */
int lpmg_server_wait(lpmg_server s, int64_t ticket, double timeout_s, double* stats);
/* out[0] queued, [1] running, [2] reserved bytes, [3] budget, [4] peak
 * running, [5] peak reserved bytes, [6] finished, [7] fused CG chunks,
 * [8] system-chunks in them, [9] largest batch (systems in one chunk),
 * [10] chunk graphs captured. */
int lpmg_server_info(lpmg_server s, double* out);
/* Finish every queued and running request, then free the server. */
void lpmg_server_destroy(lpmg_server s);

#ifdef __cplusplus
}
#endif

#endif /* LPM_GPU_H */
