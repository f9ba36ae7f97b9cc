// forward.hip -- the forward pass f_int(x) and the residual r = f_ext - f_int.
//
// Reference: Assembly::updateGeometry / updateForceState (assembly.h:132-138,
// 178-191), Bond::updatebGeometry (bond.h:30-39), Particle::updateBondsGeometry
// (particle.h:163-181), updateBondsForce of the three particle types
// (particle_elastic.h:38-50, particle_elastic_damage.h:74-87,
// particle_j2plasticity.h:63-76), Particle::updateParticleForce / Stress /
// DamageVisual (particle.h:150-161, 197-240), Solver::updateRR (solver.h:111-126),
// Solver::updateDisplacementBC / updateForceBC (solver.h:128-188),
// Solver::solveLinearSystem's moveBy (solver.h:100-108).
//
// Rows of up to 32 bonds take 32 lanes per particle (geometry_force_lanes_kernel,
// particle_force_lanes_kernel); longer rows one thread per particle.  Either
// way the sums of a particle are formed in the reference's bond order.
//
// One thread owns one particle and walks its bonds in the reference's order, so
// the per-layer dilatation sums, the forces and the stresses are summed in the
// same order as the reference's serial loops (bit-identical FP64, the build uses
// -ffp-contract=off).  The opposite bond's force comes from the precomputed
// reverse index instead of std::find_if.  Bytes per bond per pass: the geometry
// kernel reads p2, dis0, dLp, bdamage (28 B) and writes dis, bstrain, dLe, cs,
// bforce (56 B); the force kernel reads rev, bforce x2, cs, dis, bdamage (60 B).
#include "lpmg_internal.h"
#include "reduce.h"

namespace lpmg {

namespace {

__device__ __forceinline__ double std_max(double a, double b) { return (a < b) ? b : a; }

__global__ void __launch_bounds__(256)
geometry_force_kernel(const double4* __restrict__ xyz, const int* __restrict__ bptr,
                      const int* __restrict__ nb0, const int* __restrict__ p2,
                      const double* __restrict__ dis0, const double* __restrict__ dLp,
                      double* __restrict__ bdamage, const double* __restrict__ Kn,
                      const double* __restrict__ Tv, const double* __restrict__ damage, int ptype,
                      int N, double* __restrict__ dis, double* __restrict__ bstrain,
                      double* __restrict__ dLe, double* __restrict__ cs, double* __restrict__ bforce,
                      double* __restrict__ dLe_tot, double* __restrict__ TdLe_tot,
                      double* __restrict__ cs_sum) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double4 xi = xyz[i];
  const double tv = Tv[i];
  const double dmg_i = damage[i];
  const int b_lo = bptr[i], b_mid = b_lo + nb0[i], b_hi = bptr[i + 1];
  for (int L = 0; L < 2; ++L) {
    const int lo = L == 0 ? b_lo : b_mid, hi = L == 0 ? b_mid : b_hi;
    double dLe_total = 0, TdLe_total = 0, sx = 0, sy = 0, sz = 0;
    for (int b = lo; b < hi; ++b) {
      const double4 xj = xyz[p2[b]];
      const double dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
      const double d = sqrt(dx * dx + dy * dy + dz * dz);
      const bool broken = fabs(bdamage[b] - 1.0) < kEPS;
      const double d0 = dis0[b];
      const double dL = broken ? 0.0 : (d - d0);
      const double le = broken ? 0.0 : (d - d0 - dLp[b]);
      const double cx = broken ? 0.0 : dx / d, cy = broken ? 0.0 : dy / d, cz = broken ? 0.0 : dz / d;
      dis[b] = d;
      bstrain[b] = broken ? 0.0 : (dL / d);
      dLe[b] = le;
      cs[3 * b] = cx;
      cs[3 * b + 1] = cy;
      cs[3 * b + 2] = cz;
      dLe_total += le;
      TdLe_total += tv * le;
      sx += cx;
      sy += cy;
      sz += cz;
    }
    dLe_tot[2 * i + L] = dLe_total;
    TdLe_tot[2 * i + L] = TdLe_total;
    cs_sum[(2 * i + L) * 3] = sx;
    cs_sum[(2 * i + L) * 3 + 1] = sy;
    cs_sum[(2 * i + L) * 3 + 2] = sz;
    // updateBondsForce needs only this particle's layer sum, so it runs here.
    const double kn = Kn[2 * i + L];
    for (int b = lo; b < hi; ++b) {
      double f = 2. * kn * dLe[b] + 2. * tv * dLe_total;
      double bd = bdamage[b];
      if (ptype != kElastic) {
        bd = std_max(bd, std_max(dmg_i, damage[p2[b]]));
        bdamage[b] = bd;
      }
      f *= (1.0 - bd);
      bforce[b] = f;
    }
  }
}

__global__ void __launch_bounds__(256)
particle_force_kernel(const int* __restrict__ bptr, const int* __restrict__ rev,
                      const double* __restrict__ bforce, const double* __restrict__ cs,
                      const double* __restrict__ dis, const double* __restrict__ bdamage,
                      double volume, double nneighbors, int N, double* __restrict__ Pin,
                      double* __restrict__ stress, double* __restrict__ damage_visual) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int lo = bptr[i], hi = bptr[i + 1];
  const double V_m = volume * (hi - lo) / nneighbors;
  double P0 = 0., P1 = 0., P2 = 0.;
  double s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0, dv = 0;
  for (int b = lo; b < hi; ++b) {
    const double fsum = bforce[b] + bforce[rev[b]];
    const double cx = cs[3 * b], cy = cs[3 * b + 1], cz = cs[3 * b + 2];
    P0 += cx * 0.5 * fsum;
    P1 += cy * 0.5 * fsum;
    P2 += cz * 0.5 * fsum;
    const double w = 0.5 / V_m * dis[b] * 0.5 * fsum;
    s0 += w * cx * cx;
    s1 += w * cy * cy;
    s2 += w * cz * cz;
    s3 += w * cy * cz;
    s4 += w * cx * cz;
    s5 += w * cx * cy;
    dv += bdamage[b];
  }
  Pin[3 * i] = P0;
  Pin[3 * i + 1] = P1;
  Pin[3 * i + 2] = P2;
  double* s = stress + 6 * (size_t)i;
  s[0] = s0; s[1] = s1; s[2] = s2; s[3] = s3; s[4] = s4; s[5] = s5;
  damage_visual[i] = dv / (double)(hi - lo);
}

// The same two kernels for particles of at most 32 bonds (every lattice here):
// 32 lanes per particle, one per bond, so the bond arrays are read and written
// in contiguous runs instead of one strided run per thread.  Each bond lane
// forms its terms; the per-particle sums are then formed by one lane per sum,
// reading the terms back from LDS in bond order -- the reference's summation
// order, so the bits equal the thread-per-particle kernels'.  LDS is written
// and read by the lanes of one wavefront only.
constexpr int kFP = 32;  // lanes per particle

__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ void __launch_bounds__(256)
geometry_force_lanes_kernel(const double4* __restrict__ xyz, const int* __restrict__ bptr,
                            const int* __restrict__ nb0, const int* __restrict__ p2,
                            const double* __restrict__ dis0, const double* __restrict__ dLp,
                            double* __restrict__ bdamage, const double* __restrict__ Kn,
                            const double* __restrict__ Tv, const double* __restrict__ damage, int ptype,
                            int N, double* __restrict__ dis, double* __restrict__ bstrain,
                            double* __restrict__ dLe, double* __restrict__ cs, double* __restrict__ bforce,
                            double* __restrict__ dLe_tot, double* __restrict__ TdLe_tot,
                            double* __restrict__ cs_sum) {
  __shared__ double terms[256 / 64][64][5];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, base = lane & ~(kFP - 1), sl = lane & (kFP - 1);
  const int i = (blockIdx.x * blockDim.x + threadIdx.x) / kFP;
  const bool live = i < N;
  const int ii = live ? i : N - 1;
  const double4 xi = xyz[ii];
  const double tv = Tv[ii];
  const int b_lo = bptr[ii], n0 = nb0[ii], nb = bptr[ii + 1] - b_lo;
  const bool has = live && sl < nb;
  const int b = b_lo + (has ? sl : 0);
  // Bond::updatebGeometry (bond.h:30-39)
  const int j = p2[b];
  const double4 xj = xyz[j];
  const double bd0 = bdamage[b], d0 = dis0[b], dlp = dLp[b];
  const double dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
  const double d = sqrt(dx * dx + dy * dy + dz * dz);
  const bool broken = fabs(bd0 - 1.0) < kEPS;
  const double dL = broken ? 0.0 : (d - d0);
  const double le = broken ? 0.0 : (d - d0 - dlp);
  const double cx = broken ? 0.0 : dx / d, cy = broken ? 0.0 : dy / d, cz = broken ? 0.0 : dz / d;
  if (has) {
    dis[b] = d;
    bstrain[b] = broken ? 0.0 : (dL / d);
    dLe[b] = le;
    cs[3 * b] = cx;
    cs[3 * b + 1] = cy;
    cs[3 * b + 2] = cz;
  }
  terms[w][lane][0] = le;
  terms[w][lane][1] = tv * le;
  terms[w][lane][2] = cx;
  terms[w][lane][3] = cy;
  terms[w][lane][4] = cz;
  wave_lds_sync();
  // Particle::updateBondsGeometry (particle.h:163-181): lane q < 10 of the
  // particle sums term q % 5 over the bonds of layer q / 5, in bond order
  double tot = 0;
  if (live && sl < 10) {
    const int L = sl / 5, q = sl - 5 * L;
    const int tlo = L == 0 ? 0 : n0, thi = L == 0 ? n0 : nb;
    for (int t = tlo; t < thi; ++t) tot += terms[w][base + t][q];
    if (q == 0) dLe_tot[2 * i + L] = tot;
    else if (q == 1) TdLe_tot[2 * i + L] = tot;
    else cs_sum[(2 * i + L) * 3 + (q - 2)] = tot;
  }
  // updateBondsForce (particle_elastic.h:38-50 and the damage models'):
  // dLe_total of the bond's layer from the lane that summed it
  const int L = sl < n0 ? 0 : 1;
  const double dlt = __shfl(tot, base + 5 * L, 64);
  if (has) {
    const double kn = Kn[2 * i + L];
    double f = 2. * kn * le + 2. * tv * dlt;
    double bd = bd0;
    if (ptype != kElastic) {
      bd = std_max(bd, std_max(damage[i], damage[j]));
      bdamage[b] = bd;
    }
    f *= (1.0 - bd);
    bforce[b] = f;
  }
}

__global__ void __launch_bounds__(256)
particle_force_lanes_kernel(const int* __restrict__ bptr, const int* __restrict__ rev,
                            const double* __restrict__ bforce, const double* __restrict__ cs,
                            const double* __restrict__ dis, const double* __restrict__ bdamage,
                            double volume, double nneighbors, int N, double* __restrict__ Pin,
                            double* __restrict__ stress, double* __restrict__ damage_visual) {
  __shared__ double terms[256 / 64][64][10];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, base = lane & ~(kFP - 1), sl = lane & (kFP - 1);
  const int i = (blockIdx.x * blockDim.x + threadIdx.x) / kFP;
  const bool live = i < N;
  const int ii = live ? i : N - 1;
  const int lo = bptr[ii], nb = bptr[ii + 1] - lo;
  const bool has = live && sl < nb;
  const int b = lo + (has ? sl : 0);
  // Particle::updateParticleForce / Stress / DamageVisual (particle.h:150-161, 197-240)
  const double V_m = volume * nb / nneighbors;
  const double fsum = bforce[b] + bforce[rev[b]];
  const double cx = cs[3 * b], cy = cs[3 * b + 1], cz = cs[3 * b + 2];
  const double wv = 0.5 / V_m * dis[b] * 0.5 * fsum;
  double* const tm = terms[w][lane];
  tm[0] = cx * 0.5 * fsum;
  tm[1] = cy * 0.5 * fsum;
  tm[2] = cz * 0.5 * fsum;
  tm[3] = wv * cx * cx;
  tm[4] = wv * cy * cy;
  tm[5] = wv * cz * cz;
  tm[6] = wv * cy * cz;
  tm[7] = wv * cx * cz;
  tm[8] = wv * cx * cy;
  tm[9] = bdamage[b];
  wave_lds_sync();
  if (live && sl < 10) {
    double tot = 0;
    for (int t = 0; t < nb; ++t) tot += terms[w][base + t][sl];
    if (sl < 3) Pin[3 * i + sl] = tot;
    else if (sl < 9) stress[6 * (size_t)i + (sl - 3)] = tot;
    else damage_visual[i] = tot / (double)nb;
  }
}

// Solver::updateRR: residual of free DOFs, and the sum of squares of the
// reactions at constrained DOFs (the reference appends them to a vector that is
// only ever read through its norm).
__global__ void __launch_bounds__(256)
update_rr_kernel(const double* __restrict__ Pex, const double* __restrict__ Pin,
                 const int* __restrict__ cons, const int* __restrict__ floating, int N, int d,
                 double* __restrict__ residual, double* __restrict__ partials, Scalars* sc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double rr = 0, reac = 0;
  if (i < N) {
    const int fl = floating[i];
    for (int k = 0; k < d; ++k) {
      const int cn = cons[3 * i + k];
      const double pin = Pin[3 * i + k];
      // a floating fragment's particles are held where they are (no residual)
      const double r = fl ? 0.0 : (1 - cn) * (Pex[3 * i + k] - pin);
      residual[(size_t)d * i + k] = r;
      rr += r * r;
      if (cn == 1) reac += pin * pin;
    }
  }
  Vec<2> a;
  a.v[0] = rr;
  a.v[1] = reac;
  write_partial<256, 2>(partials, a);
}

__global__ void __launch_bounds__(256) update_rr_finish_kernel(const double* __restrict__ partials, int n,
                                                               Scalars* sc) {
  const Vec<2> a = sum_partials<256, 2>(partials, n);
  if (threadIdx.x == 0) {
    sc->rnorm_sq = a.v[0];
    sc->reac_sq = a.v[1];
  }
}

__global__ void disp_bc_kernel(double4* __restrict__ xyz, int* __restrict__ cons,
                               const int* __restrict__ idx, int n, int mode, int k, double step) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = idx[t];
  double4 p = xyz[i];
  const double d0 = k == 0 ? step : 0.0, d1 = k == 1 ? step : 0.0, d2 = k == 2 ? step : 0.0;
  if (mode == 0) {
    p.x = p.x + d0; p.y = p.y + d1; p.z = p.z + d2;
  } else {
    p.x = d0; p.y = d1; p.z = d2;
  }
  xyz[i] = p;
  cons[3 * i + k] = 1;
}

__global__ void force_bc_kernel(double* __restrict__ Pex, const int* __restrict__ idx, int n,
                                int mode, double fx, double fy, double fz) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int i = idx[t];
  if (mode == 0) {
    Pex[3 * i] += fx / n;
    Pex[3 * i + 1] += fy / n;
    Pex[3 * i + 2] += fz / n;
  } else {
    Pex[3 * i] = fx / n;
    Pex[3 * i + 1] = fy / n;
    Pex[3 * i + 2] = fz / n;
  }
}

__global__ void move_by_kernel(double4* __restrict__ xyz, const double* __restrict__ du, int N, int d) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  double4 p = xyz[i];
  p.x = p.x + du[(size_t)d * i];
  p.y = p.y + du[(size_t)d * i + 1];
  if (d == 3) p.z = p.z + du[(size_t)d * i + 2];
  xyz[i] = p;
}

}  // namespace

int launch_forward(Context* c) {
  const int N = c->N, tb = 256, nb = (N + tb - 1) / tb;
  if (c->max_nb <= kFP) {
    const int nbl = (int)(((int64_t)N * kFP + tb - 1) / tb);
    geometry_force_lanes_kernel<<<nbl, tb, 0, c->stream>>>(
        c->xyz, c->bptr, c->nb0, c->p2, c->dis0, c->dLp, c->bdamage, c->Kn, c->Tv, c->damage,
        c->ptype, N, c->dis, c->bstrain, c->dLe, c->cs, c->bforce, c->dLe_tot, c->TdLe_tot, c->cs_sum);
    particle_force_lanes_kernel<<<nbl, tb, 0, c->stream>>>(c->bptr, c->rev, c->bforce, c->cs, c->dis,
                                                          c->bdamage, c->cell.volume,
                                                          (double)c->cell.nneighbors, N, c->Pin, c->stress,
                                                          c->damage_visual);
    LPMG_HIP(hipGetLastError());
    return 0;
  }
  geometry_force_kernel<<<nb, tb, 0, c->stream>>>(
      c->xyz, c->bptr, c->nb0, c->p2, c->dis0, c->dLp, c->bdamage, c->Kn, c->Tv, c->damage,
      c->ptype, N, c->dis, c->bstrain, c->dLe, c->cs, c->bforce, c->dLe_tot, c->TdLe_tot, c->cs_sum);
  particle_force_kernel<<<nb, tb, 0, c->stream>>>(c->bptr, c->rev, c->bforce, c->cs, c->dis,
                                                  c->bdamage, c->cell.volume,
                                                  (double)c->cell.nneighbors, N, c->Pin, c->stress,
                                                  c->damage_visual);
  LPMG_HIP(hipGetLastError());
  return 0;
}

int launch_update_rr(Context* c, double* norm_res, double* reac_sq) {
  const int N = c->N, tb = 256, nb = (N + tb - 1) / tb;
  if (nb > c->n_partials) {
    set_error("partials buffer too small");
    return -1;
  }
  update_rr_kernel<<<nb, tb, 0, c->stream>>>(c->Pex, c->Pin, c->cons, c->floating, N, c->dim, c->residual,
                                             c->partials, c->sc);
  update_rr_finish_kernel<<<1, tb, 0, c->stream>>>(c->partials, nb, c->sc);
  LPMG_HIP(hipGetLastError());
  LPMG_HIP(hipMemcpyAsync(c->sc_host, c->sc, sizeof(Scalars), hipMemcpyDeviceToHost, c->stream));
  LPMG_HIP(hipStreamSynchronize(c->stream));
  *norm_res = sqrt(c->sc_host->rnorm_sq);
  *reac_sq = c->sc_host->reac_sq;
  return 0;
}

int launch_disp_bc(Context* c, Group& g, int mode, int k, double step) {
  if (g.n == 0) return 0;
  if (!(g.cons_set >> k & 1u)) {  // the constrained set may grow
    c->floating_dirty = true;
    g.cons_set |= 1u << k;
  }
  disp_bc_kernel<<<(g.n + 255) / 256, 256, 0, c->stream>>>(c->xyz, c->cons, g.idx, g.n, mode, k, step);
  LPMG_HIP(hipGetLastError());
  return 0;
}

int launch_force_bc(Context* c, const Group& g, int mode, double fx, double fy, double fz) {
  if (g.n == 0) return 0;
  force_bc_kernel<<<(g.n + 255) / 256, 256, 0, c->stream>>>(c->Pex, g.idx, g.n, mode, fx, fy, fz);
  LPMG_HIP(hipGetLastError());
  return 0;
}

int launch_move_by(Context* c, const double* du) {
  move_by_kernel<<<(c->N + 255) / 256, 256, 0, c->stream>>>(c->xyz, du, c->N, c->dim);
  LPMG_HIP(hipGetLastError());
  return 0;
}

}  // namespace lpmg
