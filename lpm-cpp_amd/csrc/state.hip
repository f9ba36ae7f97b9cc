// state.hip -- converged-state updates of the quasi-static step.
//
// Reference: Assembly::updateStateVar (assembly.h:161-166) with
// ParticleJ2Plasticity::updateParticleStateVariables (particle_j2plasticity.h:78-140)
// and ParticleElasticDamage::updateParticleStateVariables /
// calcStrainEnergyDensity (particle_elastic_damage.h:89-104, 150-161);
// SolverStatic::updateStaticDamage (solver_static.h:31-62) with the three
// updateParticleStaticDamage variants (particle_elastic.h:52-68,
// particle_elastic_damage.h:163-173, particle_j2plasticity.h:142-152) and
// func_phi (lpm.h:49-53); Assembly::storeStateVar / resetStateVar
// (assembly.h:140-159).  All arithmetic is FP64: these are the threshold tests
// (yield, bond stretch, damage) whose flags must match the reference exactly.
// The fatigue particle (particle_fatigue_hcf.h) adds its cycle energy extremes
// and damage rate here, and SolverFatigue::updateFatigueDamage
// (solver_fatigue.h:75-126) its once-per-cycle damage update.
#include "lpmg_internal.h"

namespace lpmg {

namespace {

__global__ void __launch_bounds__(256)
j2_update_kernel(const int* __restrict__ bptr, const double* __restrict__ cs,
                 const double* __restrict__ dis, const double* __restrict__ stress,
                 const double* __restrict__ mE, const double* __restrict__ mMu,
                 const double* __restrict__ mSy, const double* __restrict__ mXi,
                 const double* __restrict__ mH, const double* __restrict__ mA, int N,
                 double* __restrict__ sv, double* __restrict__ dLp, double* __restrict__ Ddot_local) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double* s = sv + 7 * (size_t)i;
  double st[6];
  for (int k = 0; k < 6; ++k) st[k] = stress[6 * (size_t)i + k];
  double dpl[6] = {0, 0, 0, 0, 0, 0};
  double sigma_eq = 0, dlambda = 0, triax = 0;
  const double sigma_m = 1.0 / 3.0 * (st[0] + st[1] + st[2]);
  for (int k = 0; k < 3; ++k) st[k] -= sigma_m;
  for (int k = 0; k < 6; ++k) st[k] -= s[k + 1];
  for (int k = 0; k < 6; ++k) sigma_eq += (k < 3) ? st[k] * st[k] : 2.0 * st[k] * st[k];
  sigma_eq = sqrt(3.0 / 2.0 * sigma_eq);
  const double H = mH[i], xi = mXi[i];
  const double yield = sigma_eq - (mSy[i] + (1.0 - xi) * H * s[0]);
  const double G = mE[i] / 2 / (1 + mMu[i]);
  if (yield > 0.0) dlambda = yield / (3 * G + H);
  s[0] += dlambda;
  for (int k = 0; k < 6; ++k)
    if (fabs(sigma_eq) > kEPS) {
      dpl[k] = dlambda * 1.5 * st[k] / sigma_eq;
      s[k + 1] += 2. / 3. * xi * H * dpl[k];
    }
  for (int b = bptr[i]; b < bptr[i + 1]; ++b) {
    const double cx = cs[3 * b], cy = cs[3 * b + 1], cz = cs[3 * b + 2];
    const double ddLp = dis[b] * (dpl[0] * cx * cx + dpl[1] * cy * cy + dpl[2] * cz * cz +
                                  2 * dpl[3] * cy * cz + 2 * dpl[4] * cx * cz + 2 * dpl[5] * cx * cy);
    dLp[b] += ddLp;
  }
  if (sigma_eq > kEPS) triax = sigma_m / sigma_eq;
  Ddot_local[i] = dlambda * (1.0 + mA[i] * triax);
}

// ParticleElasticDamage::calcStrainEnergyDensity (particle_elastic_damage.h:89-104);
// ParticleFatigueHCF::calcEqEnergy returns the same sum in the same order
// (energy_total, particle_fatigue_hcf.h:96-106, 123).
__device__ __forceinline__ double particle_energy(const int* __restrict__ bptr, const int* __restrict__ nb0,
                                                  const double* __restrict__ dLe, const double* __restrict__ Kn,
                                                  const double* __restrict__ dLe_tot,
                                                  const double* __restrict__ TdLe_tot, double volume,
                                                  double nneighbors, int i) {
  const int lo = bptr[i], mid = lo + nb0[i], hi = bptr[i + 1];
  const double V_m = volume * (hi - lo) / nneighbors;
  double e = 0;
  for (int L = 0; L < 2; ++L) {
    const double kn = Kn[2 * i + L];
    for (int b = (L == 0 ? lo : mid); b < (L == 0 ? mid : hi); ++b) e += 0.5 * (kn * dLe[b] * dLe[b]) / V_m;
    e += 0.5 * TdLe_tot[2 * i + L] * dLe_tot[2 * i + L] / V_m;
  }
  return e;
}

__global__ void __launch_bounds__(256)
damage_energy_kernel(const int* __restrict__ bptr, const int* __restrict__ nb0,
                     const double* __restrict__ dLe, const double* __restrict__ Kn,
                     const double* __restrict__ dLe_tot, const double* __restrict__ TdLe_tot,
                     const double* __restrict__ damage, const double* __restrict__ k0,
                     const double* __restrict__ k1, double volume, double nneighbors, int N,
                     double* __restrict__ sv, const double* __restrict__ sv_last,
                     double* __restrict__ Ddot_local) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double e = particle_energy(bptr, nb0, dLe, Kn, dLe_tot, TdLe_tot, volume, nneighbors, i);
  sv[7 * (size_t)i] = e;
  const double delta = e - sv_last[7 * (size_t)i];
  const double f = e - k1[i] * damage[i] - k0[i];
  Ddot_local[i] = (f > 0 && delta > 0) ? (1 / k1[i]) * delta : 0;
}

// ParticleFatigueHCF::updateParticleStateVariables (particle_fatigue_hcf.h:127-145):
// sv[1] / sv[2] are the least / greatest energy of the current cycle, sv[0] the
// greatest over the whole history, and the local damage rate is
// A * sv[0]^B * (sv[2] - sv[1])^d while the cycle has a range.  The min/max
// keep std::min/std::max's operand order, so a NaN energy propagates the same way.
__global__ void __launch_bounds__(256)
fatigue_state_kernel(const int* __restrict__ bptr, const int* __restrict__ nb0,
                     const double* __restrict__ dLe, const double* __restrict__ Kn,
                     const double* __restrict__ dLe_tot, const double* __restrict__ TdLe_tot,
                     const double* __restrict__ fA, const double* __restrict__ fB,
                     const double* __restrict__ fD, double volume, double nneighbors, int N,
                     double* __restrict__ sv, double* __restrict__ Ddot_local) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double e = particle_energy(bptr, nb0, dLe, Kn, dLe_tot, TdLe_tot, volume, nneighbors, i);
  double* s = sv + 7 * (size_t)i;
  const double lo = (s[1] < e) ? s[1] : e;
  const double hi = (e < s[2]) ? s[2] : e;
  const double his = (hi < s[0]) ? s[0] : hi;
  s[0] = his;
  s[1] = lo;
  s[2] = hi;
  Ddot_local[i] = (hi - lo > 0) ? fA[i] * pow(his, fB[i]) * pow(hi - lo, fD[i]) : 0;
}

__device__ __forceinline__ double func_phi(double x, double L) {
  const double t = 1 / (1 + pow(x / L, 8.0));
  return t * t;
}

__global__ void __launch_bounds__(256)
nonlocal_average_kernel(const double4* __restrict__ xyz, const int* __restrict__ type,
                        const int* __restrict__ bptr, const int* __restrict__ nlptr,
                        const int* __restrict__ nlidx, const double* __restrict__ nlL,
                        const double* __restrict__ Ddot_local, double volume, double nneighbors,
                        int undamaged, int N, double* __restrict__ Ddot_nonlocal) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N || type[i] == undamaged) return;
  double A = 0, Dn = 0;
  const double4 pi = xyz[i];
  const double L = nlL[i];
  if (nlptr) {
    for (int t = nlptr[i]; t < nlptr[i + 1]; ++t) {
      const int j = nlidx[t];
      if (type[j] == undamaged) continue;
      const double4 pj = xyz[j];
      const double dx = pi.x - pj.x, dy = pi.y - pj.y, dz = pi.z - pj.z;
      const double d = sqrt(dx * dx + dy * dy + dz * dz);
      const double V_m = volume * (bptr[j + 1] - bptr[j]) / nneighbors;
      Dn += Ddot_local[j] * func_phi(d, L) * V_m;
      A += func_phi(d, L) * V_m;
    }
  }
  Ddot_nonlocal[i] = Dn / A;  // 0/0 = NaN without nonlocal neighbours, as the reference
}

__global__ void __launch_bounds__(256)
particle_damage_kernel(const int* __restrict__ type, const int* __restrict__ bptr,
                       const double* __restrict__ bstrain, const double* __restrict__ crit,
                       const double* __restrict__ Ddot_nonlocal, const double* __restrict__ dthr,
                       const double* __restrict__ damage_last, int ptype, int undamaged, int N,
                       double* __restrict__ bdamage, double* __restrict__ damage, int* __restrict__ any) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N || type[i] == undamaged) return;
  int flags = 0;  // 1: noticeable new damage (the return value), 2: a bond may have broken
  if (ptype == kElastic) {
    for (int b = bptr[i]; b < bptr[i + 1]; ++b)
      if (bstrain[b] >= crit[i] && fabs(bdamage[b] - 1.0) > kEPS) {
        flags = 3;
        bdamage[b] = 1;
      }
  } else if (ptype != kFatigue) {  // the fatigue particle keeps Particle's no-op (particle.h:63)
    const double d0 = damage[i];
    double d = d0 + Ddot_nonlocal[i];
    if (d >= dthr[i]) d = 1;
    damage[i] = d;
    if (fabs(d - damage_last[i]) > 1e-3) flags |= 1;
    if (fabs(d - 1.0) < kEPS && !(fabs(d0 - 1.0) < kEPS)) flags |= 2;
  }
  if (flags) atomicOr(any, flags);
}

// ParticleFatigueHCF::updateParticleFatigueDamage (particle_fatigue_hcf.h:147-160)
// for every particle that is not of the undamaged type (solver_fatigue.h:118-123).
__global__ void __launch_bounds__(256)
fatigue_damage_kernel(const int* __restrict__ type, const double* __restrict__ Ddot_nonlocal,
                      const double* __restrict__ dthr, const double* __restrict__ damage_last,
                      double dNdt, int undamaged, int N, double* __restrict__ damage,
                      double* __restrict__ sv, int* __restrict__ any) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N || type[i] == undamaged) return;
  const double d0 = damage[i];
  double d = d0 + Ddot_nonlocal[i] * dNdt;
  if (d >= dthr[i]) d = 1;
  damage[i] = d;
  sv[7 * (size_t)i + 1] = 2147483647.0;  // INT_MAX
  sv[7 * (size_t)i + 2] = 0;
  int flags = 0;
  if (fabs(d - damage_last[i]) > 1e-3) flags |= 1;
  if (fabs(d - 1.0) < kEPS && !(fabs(d0 - 1.0) < kEPS)) flags |= 2;
  if (flags) atomicOr(any, flags);
}

// Particle state variables at construction: zeros, except the fatigue
// particle's cycle minimum, which starts at INT_MAX (particle_fatigue_hcf.h:25-35).
__global__ void init_state_var_kernel(double* __restrict__ sv, double* __restrict__ sv_last, int fatigue,
                                      int N) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 7 * N) return;
  const double v = (fatigue && t % 7 == 1) ? 2147483647.0 : 0.0;
  sv[t] = v;
  sv_last[t] = v;
}

__global__ void copy_xyz_kernel(double4* __restrict__ dst, const double4* __restrict__ src, int N) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) dst[i] = src[i];
}

// Load-step predictor: every DOF the stiffness does not eliminate (no
// displacement BC, not fully damaged, not in a loose fragment) moves by
// `scale` times its displacement over the last converged load step.
__global__ void predict_kernel(double4* __restrict__ xyz, const double4* __restrict__ inc,
                               const int* __restrict__ cons, const double* __restrict__ dvis,
                               const int* __restrict__ floating, double scale, int N) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double4 p = xyz[i];
  const double4 d = inc[i];
  if (!dof_fixed(cons, dvis, floating, i, 0)) p.x += scale * d.x;
  if (!dof_fixed(cons, dvis, floating, i, 1)) p.y += scale * d.y;
  if (!dof_fixed(cons, dvis, floating, i, 2)) p.z += scale * d.z;
  xyz[i] = p;
}

// End of a converged load step: increment since the last one, and the new end
__global__ void step_end_kernel(const double4* __restrict__ xyz, double4* __restrict__ step,
                                double4* __restrict__ inc, int valid, int N) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double4 p = xyz[i];
  if (valid) {
    const double4 q = step[i];
    inc[i] = make_double4(p.x - q.x, p.y - q.y, p.z - q.z, 0.0);
  }
  step[i] = p;
}

}  // namespace

int launch_update_state_var(Context* c) {
  const int N = c->N, tb = 256, nb = (N + tb - 1) / tb;
  if (c->ptype == kJ2)
    j2_update_kernel<<<nb, tb, 0, c->stream>>>(c->bptr, c->cs, c->dis, c->stress, c->mE, c->mMu, c->mSy,
                                               c->mXi, c->mH, c->mA, N, c->sv, c->dLp, c->Ddot_local);
  else if (c->ptype == kElasticDamage)
    damage_energy_kernel<<<nb, tb, 0, c->stream>>>(c->bptr, c->nb0, c->dLe, c->Kn, c->dLe_tot,
                                                   c->TdLe_tot, c->damage, c->k0, c->k1, c->cell.volume,
                                                   (double)c->cell.nneighbors, N, c->sv, c->sv_last,
                                                   c->Ddot_local);
  else if (c->ptype == kFatigue)
    fatigue_state_kernel<<<nb, tb, 0, c->stream>>>(c->bptr, c->nb0, c->dLe, c->Kn, c->dLe_tot, c->TdLe_tot,
                                                   c->fA, c->fB, c->fD, c->cell.volume,
                                                   (double)c->cell.nneighbors, N, c->sv, c->Ddot_local);
  LPMG_HIP(hipGetLastError());
  return 0;
}

// The damage kernels OR bit 1 (noticeable new damage) and bit 2 (a particle
// reached full damage or a bond broke, so the fragment labels are stale) into err_flag.
static int read_damage_flags(Context* c, int* any) {
  int h = 0;
  LPMG_HIP(hipMemcpyAsync(&h, c->err_flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  LPMG_HIP(hipStreamSynchronize(c->stream));
  LPMG_HIP(hipMemsetAsync(c->err_flag, 0, sizeof(int), c->stream));
  *any = h & 1;
  if (h & 2) {
    c->floating_dirty = true;
    ++c->break_version;
  }
  return 0;
}

int launch_static_damage(Context* c, int undamaged, int* any) {
  const int N = c->N, tb = 256, nb = (N + tb - 1) / tb;
  LPMG_HIP(hipMemsetAsync(c->err_flag, 0, sizeof(int), c->stream));
  nonlocal_average_kernel<<<nb, tb, 0, c->stream>>>(c->xyz, c->type, c->bptr, c->nlptr, c->nlidx,
                                                    c->nlL, c->Ddot_local, c->cell.volume,
                                                    (double)c->cell.nneighbors, undamaged, N,
                                                    c->Ddot_nonlocal);
  particle_damage_kernel<<<nb, tb, 0, c->stream>>>(c->type, c->bptr, c->bstrain, c->crit,
                                                   c->Ddot_nonlocal, c->dthr, c->damage_last, c->ptype,
                                                   undamaged, N, c->bdamage, c->damage, c->err_flag);
  LPMG_HIP(hipGetLastError());
  return read_damage_flags(c, any);
}

int launch_fatigue_damage(Context* c, int undamaged, double dNdt, int* any) {
  const int N = c->N, tb = 256, nb = (N + tb - 1) / tb;
  LPMG_HIP(hipMemsetAsync(c->err_flag, 0, sizeof(int), c->stream));
  nonlocal_average_kernel<<<nb, tb, 0, c->stream>>>(c->xyz, c->type, c->bptr, c->nlptr, c->nlidx,
                                                    c->nlL, c->Ddot_local, c->cell.volume,
                                                    (double)c->cell.nneighbors, undamaged, N,
                                                    c->Ddot_nonlocal);
  if (c->ptype == kFatigue)  // other particles keep Particle's no-op (particle.h:64)
    fatigue_damage_kernel<<<nb, tb, 0, c->stream>>>(c->type, c->Ddot_nonlocal, c->dthr, c->damage_last,
                                                    dNdt, undamaged, N, c->damage, c->sv, c->err_flag);
  LPMG_HIP(hipGetLastError());
  return read_damage_flags(c, any);
}

int launch_init_state_var(Context* c) {
  const int n = 7 * c->N, tb = 256;
  init_state_var_kernel<<<(n + tb - 1) / tb, tb, 0, c->stream>>>(c->sv, c->sv_last, c->ptype == kFatigue,
                                                                  c->N);
  LPMG_HIP(hipGetLastError());
  return 0;
}

namespace {

// ---- fragments: connected components of the intact-bond graph
// Labels start as the particle index; each round hooks the larger of two
// labels met across an intact bond onto the smaller (atomicMin on the label's
// own entry) and then shortens every particle's label chain to its root.
// Labels only decrease, so the rounds end; at the end every component carries
// its smallest particle index, whatever the order of the atomics.
__global__ void comp_init_kernel(int* __restrict__ comp, int* __restrict__ compfix, int N) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) {
    comp[i] = i;
    compfix[i] = 0;
  }
}

__global__ void comp_hook_kernel(const int* __restrict__ bptr, const int* __restrict__ p2,
                                 const double* __restrict__ bdamage, int N, int* comp, int* __restrict__ changed) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  for (int b = bptr[i]; b < bptr[i + 1]; ++b) {
    if (fabs(bdamage[b] - 1.0) < kEPS) continue;  // broken bond
    const int li = __atomic_load_n(&comp[i], __ATOMIC_RELAXED);
    const int lj = __atomic_load_n(&comp[p2[b]], __ATOMIC_RELAXED);
    if (li == lj) continue;
    const int lo = li < lj ? li : lj, hi = li < lj ? lj : li;
    atomicMin(&comp[hi], lo);
    *changed = 1;
  }
}

__global__ void comp_jump_kernel(int* comp, int N) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int l = __atomic_load_n(&comp[i], __ATOMIC_RELAXED);
  for (;;) {
    const int m = __atomic_load_n(&comp[l], __ATOMIC_RELAXED);
    if (m == l) break;
    l = m;
  }
  atomicMin(&comp[i], l);
}

__global__ void comp_fixed_kernel(const int* __restrict__ comp, const int* __restrict__ cons, int N, int d,
                                  int* __restrict__ compfix, int* __restrict__ anyfix) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  bool c = false;
  for (int k = 0; k < d; ++k) c = c || cons[3 * i + k] == 1;
  if (c) {
    compfix[comp[i]] = 1;
    *anyfix = 1;
  }
}

// With no displacement constraint anywhere (a pure traction problem) nothing
// is held: the stiffness is singular as the user posed it.
__global__ void comp_floating_kernel(const int* __restrict__ comp, const int* __restrict__ compfix,
                                     const int* __restrict__ anyfix, int N, int* __restrict__ floating,
                                     int* __restrict__ count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int f = (*anyfix && !compfix[comp[i]]) ? 1 : 0;
  floating[i] = f;
  if (f) atomicAdd(count, 1);
}

}  // namespace

int update_floating(Context* c) {
  const int N = c->N, tb = 256, nb = (N + tb - 1) / tb;
  if (!c->comp) {
    if (dalloc(c, &c->comp, N) || dalloc(c, &c->compfix, N) || dalloc(c, &c->changed, 3)) return -3;
  }
  comp_init_kernel<<<nb, tb, 0, c->stream>>>(c->comp, c->compfix, N);
  int h[2] = {0, 0};
  int rounds = 0;
  for (;; ++rounds) {
    if (rounds > 4096) {
      set_error("fragment labelling did not settle");
      return -1;
    }
    LPMG_HIP(hipMemsetAsync(c->changed, 0, sizeof(int), c->stream));
    comp_hook_kernel<<<nb, tb, 0, c->stream>>>(c->bptr, c->p2, c->bdamage, N, c->comp, c->changed);
    comp_jump_kernel<<<nb, tb, 0, c->stream>>>(c->comp, N);
    LPMG_HIP(hipGetLastError());
    LPMG_HIP(copy_sync(c, h, c->changed, sizeof(int), hipMemcpyDeviceToHost));
    if (!h[0]) break;
  }
  LPMG_HIP(hipMemsetAsync(c->changed + 1, 0, 2 * sizeof(int), c->stream));
  comp_fixed_kernel<<<nb, tb, 0, c->stream>>>(c->comp, c->cons, N, c->dim, c->compfix, c->changed + 2);
  comp_floating_kernel<<<nb, tb, 0, c->stream>>>(c->comp, c->compfix, c->changed + 2, N, c->floating,
                                                 c->changed + 1);
  LPMG_HIP(hipGetLastError());
  LPMG_HIP(copy_sync(c, h + 1, c->changed + 1, sizeof(int), hipMemcpyDeviceToHost));
  c->n_floating = h[1];
  c->floating_dirty = false;
  return 0;
}

int store_state(Context* c) {
  const size_t N = c->N, NB = c->NB;
  auto cp = [&](void* d, const void* s, size_t bytes) {
    return hipMemcpyAsync(d, s, bytes, hipMemcpyDeviceToDevice, c->stream);
  };
  copy_xyz_kernel<<<(c->N + 255) / 256, 256, 0, c->stream>>>(c->xyz_last, c->xyz, c->N);
  LPMG_HIP(cp(c->damage_last, c->damage, N * sizeof(double)));
  LPMG_HIP(cp(c->sv_last, c->sv, 7 * N * sizeof(double)));
  LPMG_HIP(cp(c->dis_last, c->dis, NB * sizeof(double)));
  LPMG_HIP(cp(c->bdamage_last, c->bdamage, NB * sizeof(double)));
  LPMG_HIP(cp(c->bforce_last, c->bforce, NB * sizeof(double)));
  LPMG_HIP(cp(c->dLp_last, c->dLp, NB * sizeof(double)));
  LPMG_HIP(hipGetLastError());
  return 0;
}

int launch_predict(Context* c, double scale) {
  if (c->floating_dirty)
    if (int rc = update_floating(c)) return rc;
  predict_kernel<<<(c->N + 255) / 256, 256, 0, c->stream>>>(c->xyz, c->xyz_inc, c->cons, c->damage_visual,
                                                             c->floating, scale, c->N);
  LPMG_HIP(hipGetLastError());
  return 0;
}

int record_step_end(Context* c) {
  if (!c->xyz_step && (dalloc(c, &c->xyz_step, c->N) || dalloc(c, &c->xyz_inc, c->N))) return -3;
  step_end_kernel<<<(c->N + 255) / 256, 256, 0, c->stream>>>(c->xyz, c->xyz_step, c->xyz_inc, c->step_valid ? 1 : 0,
                                                              c->N);
  LPMG_HIP(hipGetLastError());
  c->inc_valid = c->step_valid;
  c->step_valid = true;
  return 0;
}

int reset_state(Context* c, int reset_xyz) {
  const size_t N = c->N, NB = c->NB;
  c->floating_dirty = true;
  ++c->break_version;
  auto cp = [&](void* d, const void* s, size_t bytes) {
    return hipMemcpyAsync(d, s, bytes, hipMemcpyDeviceToDevice, c->stream);
  };
  if (reset_xyz) copy_xyz_kernel<<<(c->N + 255) / 256, 256, 0, c->stream>>>(c->xyz, c->xyz_last, c->N);
  LPMG_HIP(cp(c->damage, c->damage_last, N * sizeof(double)));
  LPMG_HIP(cp(c->sv, c->sv_last, 7 * N * sizeof(double)));
  LPMG_HIP(cp(c->dis, c->dis_last, NB * sizeof(double)));
  LPMG_HIP(cp(c->bdamage, c->bdamage_last, NB * sizeof(double)));
  LPMG_HIP(cp(c->bforce, c->bforce_last, NB * sizeof(double)));
  LPMG_HIP(cp(c->dLp, c->dLp_last, NB * sizeof(double)));
  LPMG_HIP(hipGetLastError());
  return 0;
}

}  // namespace lpmg
