// assemble.hip -- analytic tangent stiffness and its block-CSR assembly.
//
// Reference: Stiffness::calcStiffness3D/2D (stiffness.h:197-361),
// Stiffness::localStiffnessANA (stiffness.h:108-142), Particle::hasAFEMneighbor
// (particle.h:242-256), fdu2dxyz / fdu2dxyz1 / fdu2dxyz2 (derivative.h:7-239),
// Stiffness::updateStiffnessDispBC (stiffness.h:363-438).
//
// Kernel 1 (tangent): one wavefront per particle row i, one lane per connection
// j.  Every lane walks the same bond paths i->k->m of the row and keeps the
// terms whose end is its own j, in the reference's loop order, so each lane's
// K_local(i,j) is the reference's value bit for bit.  Rows of up to 64 bonds
// whose two-hop path data fits in LDS (every lattice here) gather it in
// batched rounds and rank each path under the lane it ends at
// (tangent_lds_kernel); otherwise rows of up to 64 bonds keep
// them in lanes (tangent_lanes_kernel) and longer rows walk them with
// wave-uniform loads (tangent_kernel).
// Kernel 2 (symmetrise + Dirichlet): owner-computes symmetrisation with no
// atomics: the lane of block (i,j), i<j, reads K_local(i,j) and K_local(j,i)
// through the precomputed transpose position, writes
// 0.5*K_local(i,j) + 0.5*K_local(j,i)^T to (i,j) and its transpose to (j,i) --
// the two-term sum the reference's scatter forms (stiffness.h:241-249, 267-275)
// -- and applies the row/column elimination of constrained or fully damaged
// DOFs with the diagonal set to ||diag(K)||_2.
#include "lpmg_internal.h"
#include "reduce.h"

namespace lpmg {

namespace {

constexpr int kWaves = 4;
// Path entries per wavefront in tangent_lds_kernel's LDS slice (36 bytes each,
// plus 1.5 KB of per-bond and per-connection tables and 72 bytes per bond of
// direct terms; a workgroup's kWaves slices stay within 64 KB)
constexpr int kTangentLdsMax = 384;

template <int D>
__global__ void __launch_bounds__(64 * kWaves)
tangent_kernel(const int* __restrict__ bptr, const int* __restrict__ nb0,
               const int* __restrict__ p2, const double* __restrict__ cs,
               const double* __restrict__ bdamage, const double* __restrict__ cs_sum,
               const double* __restrict__ Kn, const double* __restrict__ Tv,
               const int* __restrict__ cptr, const int* __restrict__ conns, int N,
               double* __restrict__ H, double* __restrict__ diag_sq) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = blockIdx.x * kWaves + w;
  if (i >= N) return;
  const int c_lo = cptr[i], c_hi = cptr[i + 1];
  const int b_lo = bptr[i], b_mid = b_lo + nb0[i], b_hi = bptr[i + 1];
  const double tv_i = Tv[i];
  for (int e0 = c_lo; e0 < c_hi; e0 += 64) {
    const int e = e0 + lane;
    const bool active = e < c_hi;
    const int j = active ? conns[e] : -1;
    double K[D][D];
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
      for (int s = 0; s < D; ++s) K[r][s] = 0.0;
    if (j == i) {
      // fdu2dxyz (derivative.h:7-45): upper triangle of the diagonal block
      for (int L = 0; L < 2; ++L) {
        const int lo = L == 0 ? b_lo : b_mid, hi = L == 0 ? b_mid : b_hi;
        const double kn = Kn[2 * i + L];
        const double* csum = cs_sum + (2 * i + L) * 3;
        for (int b = lo; b < hi; ++b) {
          const double om = 1 - bdamage[b];
          const double c[3] = {cs[3 * b], cs[3 * b + 1], cs[3 * b + 2]};
#pragma unroll
          for (int r = 0; r < D; ++r)
#pragma unroll
            for (int s = r; s < D; ++s) {
              const double geo = (r == s) ? (kn + tv_i) * (c[r] * c[r]) : (kn + tv_i) * c[s] * c[r];
              K[r][s] += om * ((kn * c[s] + tv_i * csum[s]) * c[r] + geo);
            }
        }
      }
      double dsq = 0;
#pragma unroll
      for (int r = 0; r < D; ++r) dsq += K[r][r] * K[r][r];
      diag_sq[i] = dsq;
    } else if (active) {
      for (int L = 0; L < 2; ++L) {
        const int lo = L == 0 ? b_lo : b_mid, hi = L == 0 ? b_mid : b_hi;
        int b1 = -1;
        for (int b = lo; b < hi; ++b)
          if (p2[b] == j && b1 < 0) b1 = b;
        bool found = b1 >= 0;
        double T[D][D];
#pragma unroll
        for (int r = 0; r < D; ++r)
#pragma unroll
          for (int s = 0; s < D; ++s) T[r][s] = 0.0;
        double c1[3] = {0, 0, 0}, om1 = 0;
        if (b1 >= 0) {
          c1[0] = cs[3 * b1]; c1[1] = cs[3 * b1 + 1]; c1[2] = cs[3 * b1 + 2];
          om1 = 1 - bdamage[b1];
        }
        const double kn = Kn[2 * i + L];
        for (int b2 = lo; b2 < hi; ++b2) {
          const int k = p2[b2];
          const double c2[3] = {cs[3 * b2], cs[3 * b2 + 1], cs[3 * b2 + 2]};
          const double om2 = 1 - bdamage[b2];
          if (b1 >= 0 && k == j) {
            // direct bond term (derivative.h:59-88 / 156-185)
            const double* csum = cs_sum + (2 * k + L) * 3;
#pragma unroll
            for (int r = 0; r < D; ++r)
#pragma unroll
              for (int s = 0; s < D; ++s) {
                const int hi_rs = r > s ? r : s, lo_rs = r > s ? s : r;
                const double geo = (r == s) ? (kn + tv_i) * (c2[r] * c2[r])
                                            : (kn + tv_i) * c2[hi_rs] * c2[lo_rs];
                T[r][s] += om2 * ((-kn * c2[s] + tv_i * csum[s]) * c2[r] - geo);
              }
            continue;
          }
          if (b1 >= 0) {
            // common-neighbour term through i's dilatation (derivative.h:91-99 / 188-196)
            const double coef = 0.5 * (om1 * tv_i + om2 * tv_i);
#pragma unroll
            for (int r = 0; r < D; ++r)
#pragma unroll
              for (int s = 0; s < D; ++s) T[r][s] -= coef * c2[r] * c1[s];
          }
          // two-hop paths i->k->j through k's dilatation
          // (derivative.h:101-115 / 122-139 and 198-212 / 219-236)
          const double tv_k = Tv[k];
          const int klo = L == 0 ? bptr[k] : bptr[k] + nb0[k];
          const int khi = L == 0 ? bptr[k] + nb0[k] : bptr[k + 1];
          for (int b3 = klo; b3 < khi; ++b3) {
            if (p2[b3] != j) continue;
            found = true;
            const double c3[3] = {cs[3 * b3], cs[3 * b3 + 1], cs[3 * b3 + 2]};
            const double coef3 = 0.5 * ((1 - bdamage[b3]) * tv_k + om2 * tv_i);
#pragma unroll
            for (int r = 0; r < D; ++r)
#pragma unroll
              for (int s = 0; s < D; ++s) T[r][s] -= coef3 * c2[r] * c3[s];
          }
        }
        if (found)
#pragma unroll
          for (int r = 0; r < D; ++r)
#pragma unroll
            for (int s = 0; s < D; ++s) K[r][s] = K[r][s] + T[r][s];
      }
    }
    if (active) {
      double* out = H + (size_t)e * D * D;
#pragma unroll
      for (int r = 0; r < D; ++r)
#pragma unroll
        for (int s = 0; s < D; ++s) out[r * D + s] = K[r][s];
    }
  }
}

// The same tangent with the row's bond paths staged in lanes instead of walked
// through wave-uniform loads (rows of at most 64 bonds).  Lane t holds bond
// b_lo + t of particle i (its end k, direction, 1 - bdamage, k's layer ranges
// and Tv[k]); for each bond b2 of a layer the lanes load k's bonds of that
// layer in one parallel round, and every lane scans them through cross-lane
// broadcasts for its own j.  The walk -- b2 ascending, the direct or
// common-neighbour term, then k's bonds ascending -- and every expression are
// those of tangent_kernel, so each K_local(i,j) has the same bits; what changes
// is that one path step costs one parallel load round instead of a chain of
// dependent scalar loads.
template <int D>
__global__ void __launch_bounds__(64 * kWaves)
tangent_lanes_kernel(const int* __restrict__ bptr, const int* __restrict__ nb0,
                     const int* __restrict__ p2, const double* __restrict__ cs,
                     const double* __restrict__ bdamage, const double* __restrict__ cs_sum,
                     const double* __restrict__ Kn, const double* __restrict__ Tv,
                     const int* __restrict__ cptr, const int* __restrict__ conns, int N,
                     double* __restrict__ H, double* __restrict__ diag_sq) {
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = blockIdx.x * kWaves + w;
  if (i >= N) return;  // wave-uniform
  const int c_lo = cptr[i], c_hi = cptr[i + 1];
  const int b_lo = bptr[i], n0 = nb0[i], nbi = bptr[i + 1] - b_lo;
  const double tv_i = Tv[i];
  // lane t: bond b_lo + t of i
  int bk = 0, kr[3] = {0, 0, 0};
  double bc[3] = {0, 0, 0}, bom = 0, btv = 0;
  if (lane < nbi) {
    const int b = b_lo + lane;
    bk = p2[b];
    bc[0] = cs[3 * b]; bc[1] = cs[3 * b + 1]; bc[2] = cs[3 * b + 2];
    bom = 1 - bdamage[b];
    kr[0] = bptr[bk]; kr[1] = kr[0] + nb0[bk]; kr[2] = bptr[bk + 1];
    btv = Tv[bk];
  }
  for (int e0 = c_lo; e0 < c_hi; e0 += 64) {
    const int e = e0 + lane;
    const bool active = e < c_hi;
    const int j = active ? conns[e] : -1;
    double K[D][D];
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
      for (int s = 0; s < D; ++s) K[r][s] = 0.0;
    for (int L = 0; L < 2; ++L) {
      const int tlo = L == 0 ? 0 : n0, thi = L == 0 ? n0 : nbi;
      int t1 = -1;
      for (int t = tlo; t < thi; ++t)
        if (__shfl(bk, t, 64) == j && t1 < 0) t1 = t;
      bool found = t1 >= 0;
      // Every lane takes part in the broadcasts (a lane reads another lane's
      // registers only while both run the same instruction); lanes without a
      // direct bond discard what they read.
      const int src1 = t1 >= 0 ? t1 : 0;
      double c1[3];
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const double v = __shfl(bc[q], src1, 64);
        c1[q] = t1 >= 0 ? v : 0.0;
      }
      const double om1v = __shfl(bom, src1, 64);
      const double om1 = t1 >= 0 ? om1v : 0.0;
      double T[D][D];
#pragma unroll
      for (int r = 0; r < D; ++r)
#pragma unroll
        for (int s = 0; s < D; ++s) T[r][s] = 0.0;
      const double kn = Kn[2 * i + L];
      for (int t = tlo; t < thi; ++t) {
        const int k = __shfl(bk, t, 64);
        double c2[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) c2[q] = __shfl(bc[q], t, 64);
        const double om2 = __shfl(bom, t, 64);
        const int klo = __shfl(kr[L], t, 64), khi = __shfl(kr[L + 1], t, 64);
        const double tv_k = __shfl(btv, t, 64);
        const bool direct = t1 >= 0 && k == j;
        if (direct) {
          // direct bond term (derivative.h:59-88 / 156-185)
          const double* csum = cs_sum + (2 * k + L) * 3;
#pragma unroll
          for (int r = 0; r < D; ++r)
#pragma unroll
            for (int s = 0; s < D; ++s) {
              const int hi_rs = r > s ? r : s, lo_rs = r > s ? s : r;
              const double geo = (r == s) ? (kn + tv_i) * (c2[r] * c2[r])
                                          : (kn + tv_i) * c2[hi_rs] * c2[lo_rs];
              T[r][s] += om2 * ((-kn * c2[s] + tv_i * csum[s]) * c2[r] - geo);
            }
        } else if (t1 >= 0) {
          // common-neighbour term through i's dilatation (derivative.h:91-99 / 188-196)
          const double coef = 0.5 * (om1 * tv_i + om2 * tv_i);
#pragma unroll
          for (int r = 0; r < D; ++r)
#pragma unroll
            for (int s = 0; s < D; ++s) T[r][s] -= coef * c2[r] * c1[s];
        }
        // two-hop paths i->k->j through k's dilatation (derivative.h:101-115 /
        // 122-139 and 198-212 / 219-236): k's bonds of layer L, one per lane
        const int nk = khi - klo;
        int q3 = -1;
        double c3l[3] = {0, 0, 0}, om3l = 0;
        if (lane < nk) {
          const int b3 = klo + lane;
          q3 = p2[b3];
          c3l[0] = cs[3 * b3]; c3l[1] = cs[3 * b3 + 1]; c3l[2] = cs[3 * b3 + 2];
          om3l = 1 - bdamage[b3];
        }
        for (int u = 0; u < nk; ++u) {
          const int m = __shfl(q3, u, 64);
          double c3[3];
#pragma unroll
          for (int q = 0; q < 3; ++q) c3[q] = __shfl(c3l[q], u, 64);
          const double om3 = __shfl(om3l, u, 64);
          if (direct || m != j) continue;
          found = true;
          const double coef3 = 0.5 * (om3 * tv_k + om2 * tv_i);
#pragma unroll
          for (int r = 0; r < D; ++r)
#pragma unroll
            for (int s = 0; s < D; ++s) T[r][s] -= coef3 * c2[r] * c3[s];
        }
      }
      if (found)
#pragma unroll
        for (int r = 0; r < D; ++r)
#pragma unroll
          for (int s = 0; s < D; ++s) K[r][s] = K[r][s] + T[r][s];
    }
    if (j == i) {
      // fdu2dxyz (derivative.h:7-45): upper triangle of the diagonal block
#pragma unroll
      for (int r = 0; r < D; ++r)
#pragma unroll
        for (int s = 0; s < D; ++s) K[r][s] = 0.0;
      for (int L = 0; L < 2; ++L) {
        const int lo = L == 0 ? b_lo : b_lo + n0, hi = L == 0 ? b_lo + n0 : b_lo + nbi;
        const double kn = Kn[2 * i + L];
        const double* csum = cs_sum + (2 * i + L) * 3;
        for (int b = lo; b < hi; ++b) {
          const double om = 1 - bdamage[b];
          const double c[3] = {cs[3 * b], cs[3 * b + 1], cs[3 * b + 2]};
#pragma unroll
          for (int r = 0; r < D; ++r)
#pragma unroll
            for (int s = r; s < D; ++s) {
              const double geo = (r == s) ? (kn + tv_i) * (c[r] * c[r]) : (kn + tv_i) * c[s] * c[r];
              K[r][s] += om * ((kn * c[s] + tv_i * csum[s]) * c[r] + geo);
            }
        }
      }
      double dsq = 0;
#pragma unroll
      for (int r = 0; r < D; ++r) dsq += K[r][r] * K[r][r];
      diag_sq[i] = dsq;
    }
    if (active) {
      double* out = H + (size_t)e * D * D;
#pragma unroll
      for (int r = 0; r < D; ++r)
#pragma unroll
        for (int s = 0; s < D; ++s) out[r * D + s] = K[r][s];
    }
  }
}

// The tangent with every load of a row issued in batched rounds and no search
// in the path walk.  The row's bonds sit one per lane (as in
// tangent_lanes_kernel) and its connections one per lane (j = conns[lane]).
// All two-hop path data -- for each bond i->k, k's bonds of the same layer
// (end m, direction, 1 - bdamage) -- is gathered into this wavefront's slice of
// LDS, and each path entry finds its owner, the lane whose j is m, by a binary
// search over the row's sorted connections.  Owners are recorded as one bit per
// bond t in the owner's mask, and the entries are ranked into per-owner lists in
// bond order (the rank of entry (t, owner) is the owner's list start plus the
// bits of its mask below t).  k's bonds end at distinct particles, so a lane
// has at most one path per bond; the walk then takes the lane's next list entry
// at each bond whose bit is set, all lanes together, instead of scanning every
// path end for its own j.  A bond list in which two of k's bonds share an end
// takes the scan (detected while ranking).  Each bond's direct term is formed
// by its bond lane into LDS and added by the lane whose j is the bond's end at
// that bond's step, instead of one lane computing it while the wavefront waits.
// Walk order and expressions are those of tangent_kernel, so the bits are the
// same.  Rows of up to 64 connections.
template <int D>
__global__ void __launch_bounds__(64 * kWaves)
tangent_lds_kernel(const int* __restrict__ bptr, const int* __restrict__ nb0,
                   const int* __restrict__ p2, const double* __restrict__ cs,
                   const double* __restrict__ bdamage, const double* __restrict__ cs_sum,
                   const double* __restrict__ Kn, const double* __restrict__ Tv,
                   const int* __restrict__ cptr, const int* __restrict__ conns, int N, int E, int NBM,
                   double* __restrict__ H, double* __restrict__ diag_sq) {
  extern __shared__ double lds_tangent[];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = blockIdx.x * kWaves + w;
  const bool row = i < N;  // wave-uniform; every wavefront still reaches the barriers
  double* const Lcs = lds_tangent + (size_t)w * (4 * E + 64);  // [E][3]
  double* const Lom = Lcs + 3 * E;                              // [E]
  unsigned long long* const Lmask = reinterpret_cast<unsigned long long*>(Lom + E);  // [64] per owner
  int* const Loff = (int*)(lds_tangent + (size_t)kWaves * (4 * E + 64)) + (size_t)w * 256;  // [64] first entry of bond t
  int* const Lklo = Loff + 64;    // [64] k's first bond of the layer
  int* const Lconn = Lklo + 64;   // [64] the row's connections, padded with INT_MAX
  int* const Lbase = Lconn + 64;  // [64] first list entry of each owner
  short* const Lot = (short*)(lds_tangent + (size_t)kWaves * (4 * E + 64) + (size_t)kWaves * 128) +
                     (size_t)w * 2 * E;  // [E] owner << 8 | bond, -1 without owner
  short* const Llist = Lot + E;          // [E] entries by owner, in bond order
  double* const Ld = (double*)(lds_tangent + (size_t)kWaves * (4 * E + 64) + (size_t)kWaves * 128 +
                               (size_t)kWaves * E / 2) + (size_t)w * NBM * D * D;  // [NBM][D][D] direct terms
  int c_lo = 0, nc = 0, b_lo = 0, n0 = 0, nbi = 0;
  double tv_i = 0;
  if (row) {
    c_lo = cptr[i]; nc = cptr[i + 1] - c_lo;
    b_lo = bptr[i]; n0 = nb0[i]; nbi = bptr[i + 1] - b_lo;
    tv_i = Tv[i];
  }
  const bool active = lane < nc;
  const int j = active ? conns[c_lo + lane] : -1;
  // lane t: bond b_lo + t of i and the layer range of its end's bonds
  int bk = 0, klo = 0, nk = 0;
  double bc[3] = {0, 0, 0}, bom = 0, btv = 0;
  if (lane < nbi) {
    const int b = b_lo + lane;
    bk = p2[b];
    bc[0] = cs[3 * b]; bc[1] = cs[3 * b + 1]; bc[2] = cs[3 * b + 2];
    bom = 1 - bdamage[b];
    const int k0 = bptr[bk], k1 = k0 + nb0[bk], k2 = bptr[bk + 1];
    klo = lane < n0 ? k0 : k1;
    nk = lane < n0 ? k1 - k0 : k2 - k1;
    btv = Tv[bk];
    // bond t's direct term (derivative.h:59-88 / 156-185), added at step t of
    // the walk by the lane whose j is the bond's end
    const int L = lane < n0 ? 0 : 1;
    const double kn = Kn[2 * i + L];
    const double* csum = cs_sum + (2 * bk + L) * 3;
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
      for (int s = 0; s < D; ++s) {
        const int hi_rs = r > s ? r : s, lo_rs = r > s ? s : r;
        const double geo = (r == s) ? (kn + tv_i) * (bc[r] * bc[r]) : (kn + tv_i) * bc[hi_rs] * bc[lo_rs];
        Ld[(lane * D + r) * D + s] = bom * ((-kn * bc[s] + tv_i * csum[s]) * bc[r] - geo);
      }
  }
  const int incl = wave_incl_scan(nk);
  const int off = incl - nk;
  const int T = __shfl(incl, 63, 64);
  if (lane < nbi) {
    Loff[lane] = off;
    Lklo[lane] = klo;
  }
  Lconn[lane] = active ? j : INT_MAX;
  Lmask[lane] = 0ull;
  __syncthreads();
  // gather: entry e = off_t + u is bond klo_t + u; four entries per lane in
  // flight.  Each entry's owner sets bit t of its mask.
  bool dup = false;
  for (int e0 = lane; e0 < T; e0 += 4 * 64) {
    int b3[4], t3[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = min(e0 + 64 * q, T - 1);
      int lo = 0, hi = nbi - 1;  // last t with Loff[t] <= e
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (Loff[mid] <= e) lo = mid;
        else hi = mid - 1;
      }
      t3[q] = lo;
      b3[q] = Lklo[lo] + (e - Loff[lo]);
    }
    int m[4];
    double c3[4][3], om3[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      m[q] = p2[b3[q]];
      c3[q][0] = cs[3 * b3[q]]; c3[q][1] = cs[3 * b3[q] + 1]; c3[q][2] = cs[3 * b3[q] + 2];
      om3[q] = 1 - bdamage[b3[q]];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = e0 + 64 * q;
      if (e < T) {
        Lcs[3 * e] = c3[q][0]; Lcs[3 * e + 1] = c3[q][1]; Lcs[3 * e + 2] = c3[q][2];
        Lom[e] = om3[q];
        int pos = 0;  // last connection <= m
#pragma unroll
        for (int st = 32; st > 0; st >>= 1)
          if (Lconn[pos + st] <= m[q]) pos += st;
        const bool own = Lconn[pos] == m[q];
        Lot[e] = (short)(own ? (pos << 8 | t3[q]) : -1);
        if (own) {
          const unsigned long long bit = 1ull << t3[q];
          dup = dup || (atomicOr(&Lmask[pos], bit) & bit) != 0;
        }
      }
    }
  }
  __syncthreads();
  const bool scan = __any(dup);  // wave-uniform
  const unsigned long long mine = active ? Lmask[lane] : 0ull;
  const int cnt = __popcll(mine);
  const int base = wave_incl_scan(cnt) - cnt;
  Lbase[lane] = base;
  __syncthreads();
  if (!scan)
    for (int e = lane; e < T; e += 64) {
      const int v = Lot[e];
      if (v >= 0) {
        const int o = v >> 8, t = v & 255;
        Llist[Lbase[o] + __popcll(Lmask[o] & ((1ull << t) - 1ull))] = (short)e;
      }
    }
  __syncthreads();
  if (!row) return;
  double K[D][D];
#pragma unroll
  for (int r = 0; r < D; ++r)
#pragma unroll
    for (int s = 0; s < D; ++s) K[r][s] = 0.0;
  int cursor = base;
  for (int L = 0; L < 2; ++L) {
    const int tlo = L == 0 ? 0 : n0, thi = L == 0 ? n0 : nbi;
    int t1 = -1;
    for (int t = tlo; t < thi; ++t)
      if (__shfl(bk, t, 64) == j && t1 < 0) t1 = t;
    const unsigned long long below_n0 = n0 >= 64 ? ~0ull : (1ull << n0) - 1ull;
    const unsigned long long layer = L == 0 ? below_n0 : ~below_n0;
    bool found = t1 >= 0 || (!scan && (mine & layer) != 0ull);
    const int src1 = t1 >= 0 ? t1 : 0;
    double c1[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double v = __shfl(bc[q], src1, 64);
      c1[q] = t1 >= 0 ? v : 0.0;
    }
    const double om1v = __shfl(bom, src1, 64);
    const double om1 = t1 >= 0 ? om1v : 0.0;
    double Tm[D][D];
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
      for (int s = 0; s < D; ++s) Tm[r][s] = 0.0;
    for (int t = tlo; t < thi; ++t) {
      const int k = __shfl(bk, t, 64);
      double c2[3];
#pragma unroll
      for (int q = 0; q < 3; ++q) c2[q] = __shfl(bc[q], t, 64);
      const double om2 = __shfl(bom, t, 64);
      const double tv_k = __shfl(btv, t, 64);
      int toff = 0, tnk = 0;
      if (scan) {  // wave-uniform: every lane runs the broadcasts
        toff = __shfl(off, t, 64);
        tnk = __shfl(nk, t, 64);
      }
      const bool direct = t1 >= 0 && k == j;
      const bool path = !scan && ((mine >> t) & 1ull);
      const int qp = path ? Llist[cursor] : 0;
      cursor += path ? 1 : 0;
      if (direct) {
#pragma unroll
        for (int r = 0; r < D; ++r)
#pragma unroll
          for (int s = 0; s < D; ++s) Tm[r][s] += Ld[(t * D + r) * D + s];
      } else {
        if (t1 >= 0) {
          // common-neighbour term through i's dilatation (derivative.h:91-99 / 188-196)
          const double coef = 0.5 * (om1 * tv_i + om2 * tv_i);
#pragma unroll
          for (int r = 0; r < D; ++r)
#pragma unroll
            for (int s = 0; s < D; ++s) Tm[r][s] -= coef * c2[r] * c1[s];
        }
        // two-hop paths i->k->j through k's dilatation (derivative.h:101-115 /
        // 122-139 and 198-212 / 219-236)
        if (path) {
          const double coef3 = 0.5 * (Lom[qp] * tv_k + om2 * tv_i);
          const double c3[3] = {Lcs[3 * qp], Lcs[3 * qp + 1], Lcs[3 * qp + 2]};
#pragma unroll
          for (int r = 0; r < D; ++r)
#pragma unroll
            for (int s = 0; s < D; ++s) Tm[r][s] -= coef3 * c2[r] * c3[s];
        }
        if (scan) {  // every path end of bond t, matches in ascending order
          for (int u = 0; u < tnk; ++u) {
            const int qq = toff + u;
            const int v = Lot[qq];  // the owner is the lane whose j is the path's end
            if (v < 0 || (v >> 8) != lane) continue;
            found = true;
            const double coef3 = 0.5 * (Lom[qq] * tv_k + om2 * tv_i);
            const double c3[3] = {Lcs[3 * qq], Lcs[3 * qq + 1], Lcs[3 * qq + 2]};
#pragma unroll
            for (int r = 0; r < D; ++r)
#pragma unroll
              for (int s = 0; s < D; ++s) Tm[r][s] -= coef3 * c2[r] * c3[s];
          }
        }
      }
    }
    if (found)
#pragma unroll
      for (int r = 0; r < D; ++r)
#pragma unroll
        for (int s = 0; s < D; ++s) K[r][s] = K[r][s] + Tm[r][s];
  }
  {
    // fdu2dxyz (derivative.h:7-45): upper triangle of the diagonal block.
    // Lane t forms bond t's term from the bond it holds; the sums over the
    // bonds, in bond order, are formed in every lane and kept by lane j == i.
    double G[D][D];
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
      for (int s = 0; s < D; ++s) G[r][s] = 0.0;
    if (lane < nbi) {
      const int L = lane < n0 ? 0 : 1;
      const double kn = Kn[2 * i + L];
      const double* csum = cs_sum + (2 * i + L) * 3;
#pragma unroll
      for (int r = 0; r < D; ++r)
#pragma unroll
        for (int s = r; s < D; ++s) {
          const double geo = (r == s) ? (kn + tv_i) * (bc[r] * bc[r]) : (kn + tv_i) * bc[s] * bc[r];
          G[r][s] = bom * ((kn * bc[s] + tv_i * csum[s]) * bc[r] + geo);
        }
    }
    double Kd[D][D];
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
      for (int s = 0; s < D; ++s) Kd[r][s] = 0.0;
    for (int t = 0; t < nbi; ++t)
#pragma unroll
      for (int r = 0; r < D; ++r)
#pragma unroll
        for (int s = r; s < D; ++s) Kd[r][s] += __shfl(G[r][s], t, 64);
    if (j == i) {
#pragma unroll
      for (int r = 0; r < D; ++r)
#pragma unroll
        for (int s = 0; s < D; ++s) K[r][s] = Kd[r][s];
      double dsq = 0;
#pragma unroll
      for (int r = 0; r < D; ++r) dsq += K[r][r] * K[r][r];
      diag_sq[i] = dsq;
    }
  }
  if (active) {
    double* out = H + (size_t)(c_lo + lane) * D * D;
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
      for (int s = 0; s < D; ++s) out[r * D + s] = K[r][s];
  }
}

__global__ void __launch_bounds__(256)
norm_diag_kernel(const double* __restrict__ diag_sq, int N, double* __restrict__ partials,
                 Scalars* sc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  Vec<1> a;
  a.v[0] = i < N ? diag_sq[i] : 0.0;
  write_partial<256, 1>(partials, a);
}

__global__ void __launch_bounds__(256) norm_diag_finish_kernel(const double* __restrict__ partials, int n,
                                                               Scalars* sc) {
  const Vec<1> tot = sum_partials<256, 1>(partials, n);
  if (threadIdx.x == 0) sc->norm_diag = sqrt(tot.v[0]);
}

// Bit k of fixm[i]: DOF k of particle i is eliminated (dof_fixed).
template <int D>
__global__ void __launch_bounds__(256)
fixed_dofs_kernel(const int* __restrict__ cons, const double* __restrict__ dvis, const int* __restrict__ floating,
                  int N, int* __restrict__ fixm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int m = 0;
#pragma unroll
  for (int k = 0; k < D; ++k) m |= dof_fixed(cons, dvis, floating, i, k) ? 1 << k : 0;
  fixm[i] = m;
}

// One thread per block e = (i,j), in storage order: K(i,j) = 0.5 L(i,j) +
// 0.5 L(j,i)^T from the tangent's unsymmetrised blocks L (the upper block's
// term first, so a block and its transpose carry the same bits), then the
// elimination of fixed DOFs (rows and columns zeroed, diagonal ||diag(K)||_2)
// and the FP32 copy.  Every write is coalesced; only the L(j,i) reads scatter.
template <int D>
__global__ void __launch_bounds__(256)
symmetrize_bc_kernel(int nnzb, const int* __restrict__ brow, const int* __restrict__ conns,
                     const int* __restrict__ tpos, const int* __restrict__ fixm, const Scalars* __restrict__ sc,
                     const double* __restrict__ L, double* __restrict__ Kval, float* __restrict__ Kval32) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nnzb) return;
  const int i = brow[e], j = conns[e];
  const int mi = fixm[i], mj = fixm[j];
  const double* Le = L + (size_t)e * D * D;
  double v[D][D];
  if (i == j) {
    const double nd = sc->norm_diag;
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
      for (int s = 0; s < D; ++s) {
        v[r][s] = (s >= r) ? Le[r * D + s] : Le[s * D + r];
        if (((mi >> r) & 1) || ((mi >> s) & 1)) v[r][s] = (r == s) ? nd : 0.0;
      }
  } else {
    const double* Lt = L + (size_t)tpos[e] * D * D;
    double a[D][D], b[D][D];
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
      for (int s = 0; s < D; ++s) {
        a[r][s] = Le[r * D + s];
        b[r][s] = Lt[r * D + s];
      }
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
      for (int s = 0; s < D; ++s) {
        v[r][s] = i < j ? 0.5 * a[r][s] + 0.5 * b[s][r] : 0.5 * b[s][r] + 0.5 * a[r][s];
        if (((mi >> r) & 1) || ((mj >> s) & 1)) v[r][s] = 0.0;
      }
  }
  double* Ke = Kval + (size_t)e * D * D;
#pragma unroll
  for (int r = 0; r < D; ++r)
#pragma unroll
    for (int s = 0; s < D; ++s) Ke[r * D + s] = v[r][s];
  if (Kval32) {
    float* K32 = Kval32 + (size_t)e * D * D;
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
      for (int s = 0; s < D; ++s) K32[r * D + s] = (float)v[r][s];
  }
}

// Reference upper-triangular 1-based CSR (IK/JK/K_global) from the BSR.
template <int D>
__global__ void export_csr_kernel(const int64_t* __restrict__ Kp, const int* __restrict__ cptr,
                                  const int* __restrict__ conns, const int* __restrict__ dpos,
                                  const int* __restrict__ ncl, const double* __restrict__ Kval,
                                  int N, int64_t* __restrict__ IK, int64_t* __restrict__ JK,
                                  double* __restrict__ K) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int64_t kp = Kp[i];
  const int n = ncl[i];
  const int dp = dpos[i];
#pragma unroll
  for (int r = 0; r < D; ++r) {
    const int64_t start = kp + (int64_t)r * D * n - r * (r + 1) / 2 + r;
    IK[(size_t)D * i + r] = start + 1;
    int64_t pos = start;
    for (int s = r; s < D; ++s, ++pos) {
      JK[pos] = (int64_t)D * i + s + 1;
      K[pos] = Kval[(size_t)dp * D * D + r * D + s];
    }
    for (int e = dp + 1; e < cptr[i + 1]; ++e) {
      const int j = conns[e];
      for (int s = 0; s < D; ++s, ++pos) {
        JK[pos] = (int64_t)D * j + s + 1;
        K[pos] = Kval[(size_t)e * D * D + r * D + s];
      }
    }
  }
  if (i == N - 1) IK[(size_t)D * N] = Kp[N] + 1;
}

// ------------------------------------------------------------ one-pass assembly
//
// The assembled block is K(i,j) = 0.5 L(i,j) + 0.5 L(j,i)^T (stiffness.h:241-249,
// 267-275), L the unsymmetrised local tangent.  Every term of L(j,i) runs
// through the same particles as a term of L(i,j): the direct bond j->i is the
// reverse of i->j, and a two-hop path j->k->i uses the reverses of the bonds of
// a path i->k->j (bond lists are symmetric within each layer).  So the lane of
// block (i,j) forms both halves and writes the final block -- symmetrised,
// eliminated, and its FP32 copy -- once, with no unsymmetrised blocks in HBM:
//   direct + common-neighbour term of L(i,j) in layer L through bond b = i->j
//   (derivative.h:59-99 / 156-196), with the common-neighbour sum over the
//   layer's other bonds in closed form from per-particle sums
//     Q_i = sum_t c_t (cs_sum),  S_i = sum_t (1 - bdamage_t) c_t,
//     om(b) ((-Kn_i c_b[s] + Tv_i Q_j[s]) c_b[r] - (Kn_i + Tv_i) c_b[r] c_b[s]) - v_i[r] c_b[s],
//     v_i = 0.5 Tv_i (om(b) (Q_i - c_b) + (S_i - om(b) c_b)),
//   and the same with i and j exchanged through the reverse bond for L(j,i);
//   path i->k->j (bonds t = i->k, u = k->j; derivative.h:101-139 / 198-236)
//   with the path j->k->i (bonds rev u, rev t) of L(j,i):
//     -0.5 (0.5 (om(u) Tv_k + om(t) Tv_i) c_t[r] c_u[s]
//           + 0.5 (om(rev t) Tv_k + om(rev u) Tv_j) c_rev t[r] c_rev u[s]).
// Each term is formed so that the lanes of (i,j) and (j,i) compute it from the
// same values with operands exchanged, and the terms are added in the same order
// (direct term, then paths by ascending k, which is bond order for bond lists
// sorted by particle id): the two blocks are exact transposes of each other.
// The rounding differs from the two-pass assembly by a few ulps of each term.
// Used when no particle has two same-layer bonds to one particle, rows have at
// most 64 bonds and 64 connections, and the path data fits LDS.

// Per particle: eliminated DOFs (fixm), S_i per layer, the diagonal block
// (fdu2dxyz, derivative.h:7-45, in the reference's bond order) written into the
// stiffness before elimination, and the workgroup's share of ||diag(K)||^2.
template <int D>
__global__ void __launch_bounds__(256)
assembly_prep_kernel(const int* __restrict__ bptr, const int* __restrict__ nb0, const double* __restrict__ cs,
                     const double* __restrict__ bdamage, const double* __restrict__ cs_sum,
                     const double* __restrict__ Kn, const double* __restrict__ Tv, const int* __restrict__ dpos,
                     const int* __restrict__ cons, const double* __restrict__ dvis,
                     const int* __restrict__ floating, int N, int* __restrict__ fixm,
                     double* __restrict__ cs_wsum, double* __restrict__ diag_sq, double* __restrict__ Kval,
                     double* __restrict__ partials) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  Vec<1> a;
  a.v[0] = 0.0;
  if (i < N) {
    int m = 0;
#pragma unroll
    for (int k = 0; k < D; ++k) m |= dof_fixed(cons, dvis, floating, i, k) ? 1 << k : 0;
    fixm[i] = m;
    const int b_lo = bptr[i], b_mid = b_lo + nb0[i], b_hi = bptr[i + 1];
    const double tv_i = Tv[i];
    double K[D][D];
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
      for (int s = 0; s < D; ++s) K[r][s] = 0.0;
    for (int L = 0; L < 2; ++L) {
      const int lo = L == 0 ? b_lo : b_mid, hi = L == 0 ? b_mid : b_hi;
      const double kn = Kn[2 * i + L];
      const double* csum = cs_sum + (2 * i + L) * 3;
      double w[3] = {0.0, 0.0, 0.0};
      for (int b = lo; b < hi; ++b) {
        const double om = 1 - bdamage[b];
        const double c[3] = {cs[3 * b], cs[3 * b + 1], cs[3 * b + 2]};
#pragma unroll
        for (int q = 0; q < 3; ++q) w[q] += om * c[q];
#pragma unroll
        for (int r = 0; r < D; ++r)
#pragma unroll
          for (int s = r; s < D; ++s) {
            const double geo = (r == s) ? (kn + tv_i) * (c[r] * c[r]) : (kn + tv_i) * c[s] * c[r];
            K[r][s] += om * ((kn * c[s] + tv_i * csum[s]) * c[r] + geo);
          }
      }
#pragma unroll
      for (int q = 0; q < 3; ++q) cs_wsum[(2 * i + L) * 3 + q] = w[q];
    }
    double dsq = 0;
#pragma unroll
    for (int r = 0; r < D; ++r) dsq += K[r][r] * K[r][r];
    diag_sq[i] = dsq;
    a.v[0] = dsq;
    double* out = Kval + (size_t)dpos[i] * D * D;
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
      for (int s = 0; s < D; ++s) out[r * D + s] = s >= r ? K[r][s] : K[s][r];
  }
  write_partial<256, 1>(partials, a);
}

// Direct and common-neighbour terms of L(a,b) in one layer through bond a->b:
// direction c, om = 1 - bdamage, a's Kn, Tv, Q, S and b's Q.
template <int D>
__device__ __forceinline__ void direct_common_half(double kn, double tv, const double* Qa, const double* Sa,
                                                   const double* Qb, const double* c, double om,
                                                   double H[D][D]) {
  double v[D];
#pragma unroll
  for (int r = 0; r < D; ++r) v[r] = 0.5 * tv * (om * (Qa[r] - c[r]) + (Sa[r] - om * c[r]));
#pragma unroll
  for (int r = 0; r < D; ++r)
#pragma unroll
    for (int s = 0; s < D; ++s) {
      const double cc = r >= s ? c[r] * c[s] : c[s] * c[r];
      H[r][s] = om * ((-kn * c[s] + tv * Qb[s]) * c[r] - (kn + tv) * cc) - v[r] * c[s];
    }
}

// Doubles per bond t = i->k in the bond table: c_t[3], om(t), c_rev t[3],
// om(rev t), Tv(k), and k's Q[3], S[3] and Kn of t's layer (for the direct term
// of the lane whose connection is k).
constexpr int kSymTable = 16;
// Path entries per lane whose loads are issued before the first barrier (rows
// of more entries take further rounds after it).
constexpr int kSymGather = 3;
// Flag of a path entry's slot word: a bond of the path is broken on one side
// only (the two halves then have different directions).
constexpr int kOneSided = 1 << 8;

// The path plan depends on the bond topology only, so it is built once per
// handle.  For row i (a wavefront), path entry e (bond u = klo_t + v of the end
// k of i's bond t; entries ordered by t, then v) is owned by the lane whose
// connection is u's end m; entries with m == i belong to the diagonal block and
// are left out.  Each owner's entries are ranked in bond order into slots:
//   plan_e[i*E + e] = u | (t | owner << 6 | slot << 12) << 32, or -1,
//   plan_r[i*E + e] = the reverse of bond u,
//   plan_l[i*64 + lane] = first slot | count << 10 | (direct bond of layer 0
//       + 1) << 17 | (direct bond of layer 1 + 1) << 24,
//   plan_n[i] = number of entries;
// plan_smax[0] is raised to the largest number of slots of a row.
__host__ __device__ constexpr size_t plan_slice_bytes(int E) {
  return ((size_t)64 * 8 + 6 * 64 * 4 + (size_t)2 * E + 15) / 16 * 16;
}

__global__ void __launch_bounds__(64 * kWaves)
tangent_plan_kernel(const int* __restrict__ bptr, const int* __restrict__ nb0, const int* __restrict__ p2,
                    const int* __restrict__ rev, const int* __restrict__ cptr, const int* __restrict__ conns,
                    int N, int E, long long* __restrict__ plan_e, int* __restrict__ plan_r,
                    int* __restrict__ plan_l, int* __restrict__ plan_n, int* __restrict__ plan_smax) {
  extern __shared__ double lds_plan[];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = __builtin_amdgcn_readfirstlane(blockIdx.x * kWaves + w);
  const bool row = i < N;  // wave-uniform; every wavefront still reaches the barriers
  char* const slice = reinterpret_cast<char*>(lds_plan) + (size_t)w * plan_slice_bytes(E);
  unsigned long long* const Lmask = reinterpret_cast<unsigned long long*>(slice);  // [64] per owner
  int* const Lconn = reinterpret_cast<int*>(Lmask + 64);  // [64] connections, padded with INT_MAX
  int* const Loff = Lconn + 64;                           // [64] first entry of bond t
  int* const Lklo = Loff + 64;                            // [64] k's first bond of the layer
  int* const Lbase = Lklo + 64;                           // [64] first slot of each owner
  int* const Ldir = Lbase + 64;                           // [2][64] direct bond of each owner per layer
  short* const Lot = reinterpret_cast<short*>(Ldir + 128);  // [E] owner << 8 | bond, -1 without owner
  int c_lo = 0, nc = 0, b_lo = 0, n0 = 0, nbi = 0;
  if (row) {
    c_lo = cptr[i]; nc = cptr[i + 1] - c_lo;
    b_lo = bptr[i]; n0 = nb0[i]; nbi = bptr[i + 1] - b_lo;
  }
  const bool active = lane < nc;
  const int j = active ? conns[c_lo + lane] : -1;
  Lconn[lane] = active ? j : INT_MAX;
  Lmask[lane] = 0ull;
  Ldir[lane] = -1;
  Ldir[64 + lane] = -1;
  int bk = 0, bL = 0, klo = 0, nk = 0;
  if (lane < nbi) {
    bk = p2[b_lo + lane];
    bL = lane < n0 ? 0 : 1;
    const int k0 = bptr[bk], k1 = k0 + nb0[bk], k2 = bptr[bk + 1];
    klo = bL == 0 ? k0 : k1;
    nk = bL == 0 ? k1 - k0 : k2 - k1;
  }
  const int incl = wave_incl_scan(nk);
  const int off = incl - nk;
  const int T = __builtin_amdgcn_readfirstlane(__shfl(incl, 63, 64));
  if (lane < nbi) {
    Loff[lane] = off;
    Lklo[lane] = klo;
  }
  __syncthreads();
  if (lane < nbi) {  // the owner of bond t's end takes it as its direct bond
    int pos = 0;
#pragma unroll
    for (int st = 32; st > 0; st >>= 1)
      if (Lconn[pos + st] <= bk) pos += st;
    if (Lconn[pos] == bk) Ldir[bL * 64 + pos] = lane;
  }
  for (int e = lane; e < T; e += 64) {
    int lo = 0, hi = nbi - 1;  // last t with Loff[t] <= e
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (Loff[mid] <= e) lo = mid;
      else hi = mid - 1;
    }
    const int m = p2[Lklo[lo] + (e - Loff[lo])];
    int pos = 0;
#pragma unroll
    for (int st = 32; st > 0; st >>= 1)
      if (Lconn[pos + st] <= m) pos += st;
    const bool own = Lconn[pos] == m && m != i;
    Lot[e] = (short)(own ? (pos << 8 | lo) : -1);
    if (own) atomicOr(&Lmask[pos], 1ull << lo);
  }
  __syncthreads();
  const unsigned long long mine = active ? Lmask[lane] : 0ull;
  const int cnt = __popcll(mine);
  const int incl_c = wave_incl_scan(cnt);
  const int base = incl_c - cnt;
  Lbase[lane] = base;
  __syncthreads();
  if (!row) return;  // no barrier below
  if (lane == 63) atomicMax(plan_smax, incl_c);
  for (int e = lane; e < T; e += 64) {
    const int v = Lot[e];
    long long rec = -1;
    int ru = 0;
    if (v >= 0) {
      const int o = v >> 8, t = v & 255;
      const int slot = Lbase[o] + __popcll(Lmask[o] & ((1ull << t) - 1ull));
      const int u = Lklo[t] + (e - Loff[t]);
      rec = (long long)(unsigned)u | ((long long)(t | o << 6 | slot << 12) << 32);
      ru = rev[u];
    }
    plan_e[(size_t)i * E + e] = rec;
    plan_r[(size_t)i * E + e] = ru;
  }
  if (active)
    plan_l[(size_t)i * 64 + lane] = base | cnt << 10 | (Ldir[lane] + 1) << 17 | (Ldir[64 + lane] + 1) << 24;
  if (lane == 0) plan_n[i] = T;
}

// Per wavefront: the bond table, a region holding first the path entries'
// coefficients and second directions (S slots) and then the row's finished
// blocks on their way to HBM, Tv of each connection, the diagonal block, and per
// slot the bond index and the bond t.
__host__ __device__ constexpr int sym_region(int S, int DD) { return 4 * S > 64 * DD ? 4 * S : 64 * DD; }
__host__ __device__ constexpr size_t sym_slice_bytes(int S, int NBM, int DD) {
  return ((size_t)kSymTable * NBM * 8 + (size_t)8 * sym_region(S, DD) + 64 * 8 + 16 * 8 + (size_t)4 * S +
          (size_t)2 * S + 15) / 16 * 16;
}

template <int D>
__global__ void __launch_bounds__(64 * kWaves)
tangent_sym_kernel(const int* __restrict__ bptr, const int* __restrict__ nb0, const int* __restrict__ p2,
                   const int* __restrict__ rev, const double* __restrict__ cs, const double* __restrict__ bdamage,
                   const double* __restrict__ cs_sum, const double* __restrict__ cs_wsum,
                   const double* __restrict__ Kn, const double* __restrict__ Tv, const int* __restrict__ cptr,
                   const int* __restrict__ conns, const int* __restrict__ fixm, const Scalars* __restrict__ sc,
                   const long long* __restrict__ plan_e, const int* __restrict__ plan_r,
                   const int* __restrict__ plan_l, const int* __restrict__ plan_n, int N, int E, int S,
                   int NBM, double* __restrict__ Kval,
                   float* __restrict__ Kval32) {
  constexpr int DD = D * D;
  extern __shared__ double lds_sym[];
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = __builtin_amdgcn_readfirstlane(xcd_contiguous(blockIdx.x, gridDim.x) * kWaves + w);
  const bool row = i < N;  // wave-uniform; every wavefront still reaches the barriers
  char* const slice = reinterpret_cast<char*>(lds_sym) + (size_t)w * sym_slice_bytes(S, NBM, DD);
  double* const Lt = reinterpret_cast<double*>(slice);  // [kSymTable][NBM]
  double* const Lreg = Lt + kSymTable * NBM;            // [sym_region]
  double* const Lcc = Lreg;                             // [S] combined path coefficient, by slot
  double* const Lc3 = Lreg + S;                         // [S][3] direction of the path's second bond
  double* const Ltvj = Lreg + sym_region(S, DD);        // [64] Tv of each connection
  double* const Ldg = Ltvj + 64;                        // [16] the diagonal block
  int* const Lu = reinterpret_cast<int*>(Ldg + 16);     // [S] the path's second bond, by slot
  short* const Lst = reinterpret_cast<short*>(Lu + S);  // [S] bond t (| kOneSided), by slot
  int c_lo = 0, nc = 0, b_lo = 0, n0 = 0, nbi = 0, T = 0, mi = 0;
  double tv_i = 0, kni[2] = {0, 0}, Qi[2][3] = {}, Si[2][3] = {};
  if (row) {
    c_lo = cptr[i]; nc = cptr[i + 1] - c_lo;
    b_lo = bptr[i]; n0 = nb0[i]; nbi = bptr[i + 1] - b_lo;
    tv_i = Tv[i];
    T = plan_n[i];
    mi = fixm[i];
#pragma unroll
    for (int L = 0; L < 2; ++L) {
      kni[L] = Kn[2 * i + L];
#pragma unroll
      for (int u = 0; u < 3; ++u) {
        Qi[L][u] = cs_sum[(2 * i + L) * 3 + u];
        Si[L][u] = cs_wsum[(2 * i + L) * 3 + u];
      }
    }
  }
  const bool active = row && lane < nc;
  const int pw = active ? plan_l[(size_t)i * 64 + lane] : 0;
  const int j = active ? conns[c_lo + lane] : -1;
  // the first round of path entries, loaded before the barrier
  long long rec[kSymGather];
  int u3[kSymGather], r3[kSymGather];
#pragma unroll
  for (int q = 0; q < kSymGather; ++q) {
    const int e = lane + 64 * q;
    rec[q] = e < T ? plan_e[(size_t)i * E + e] : -1ll;
    r3[q] = e < T ? plan_r[(size_t)i * E + e] : 0;
  }
  double c3[kSymGather][3], cr3[kSymGather][3], om3[kSymGather], omr3[kSymGather];
#pragma unroll
  for (int q = 0; q < kSymGather; ++q) {
    u3[q] = rec[q] >= 0 ? (int)(unsigned)(rec[q] & 0xffffffffll) : 0;
#pragma unroll
    for (int u = 0; u < 3; ++u) {
      c3[q][u] = cs[3 * u3[q] + u];
      cr3[q][u] = cs[3 * r3[q] + u];
    }
    om3[q] = 1 - bdamage[u3[q]];
    omr3[q] = 1 - bdamage[r3[q]];
  }
  // lane t: bond t of i, its reverse, and its end's direct-term inputs
  if (lane < nbi) {
    const int b = b_lo + lane, rb = rev[b], k = p2[b];
    const int L = lane < n0 ? 0 : 1;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      Lt[q * NBM + lane] = cs[3 * b + q];
      Lt[(4 + q) * NBM + lane] = cs[3 * rb + q];
      Lt[(9 + q) * NBM + lane] = cs_sum[(2 * k + L) * 3 + q];
      Lt[(12 + q) * NBM + lane] = cs_wsum[(2 * k + L) * 3 + q];
    }
    Lt[3 * NBM + lane] = 1 - bdamage[b];
    Lt[7 * NBM + lane] = 1 - bdamage[rb];
    Lt[8 * NBM + lane] = Tv[k];
    Lt[15 * NBM + lane] = Kn[2 * k + L];
  }
  Ltvj[lane] = active ? Tv[j] : 0.0;
  const int mj = active ? fixm[j] : 0;
  if (active && j == i) {  // the diagonal block from the per-particle pass
    const double* in = Kval + (size_t)(c_lo + lane) * DD;
#pragma unroll
    for (int q = 0; q < DD; ++q) Ldg[q] = in[q];
  }
  __syncthreads();
  // Each entry's coefficient 0.5 (cA + cB) and second direction go to its slot.
  auto place = [&](long long r, const double* c3e, const double* cr3e, double om3e, double omr3e) {
    const int word = (int)(r >> 32);
    const int t = word & 63, o = (word >> 6) & 63, slot = word >> 12;
    const double tvk = Lt[8 * NBM + t];
    const double cA = 0.5 * (om3e * tvk + Lt[3 * NBM + t] * tv_i);
    const double cB = 0.5 * (Lt[7 * NBM + t] * tvk + omr3e * Ltvj[o]);
    bool anti = true;
#pragma unroll
    for (int u = 0; u < D; ++u)
      anti = anti && Lt[(4 + u) * NBM + t] == -Lt[u * NBM + t] && cr3e[u] == -c3e[u];
    Lcc[slot] = 0.5 * (cA + cB);
#pragma unroll
    for (int u = 0; u < 3; ++u) Lc3[3 * slot + u] = c3e[u];
    Lu[slot] = (int)(unsigned)(r & 0xffffffffll);
    Lst[slot] = (short)(t | (anti ? 0 : kOneSided));
  };
#pragma unroll
  for (int q = 0; q < kSymGather; ++q)
    if (rec[q] >= 0) place(rec[q], c3[q], cr3[q], om3[q], omr3[q]);
  for (int e = lane + 64 * kSymGather; e < T; e += 64) {  // rows with more entries
    const long long r = plan_e[(size_t)i * E + e];
    if (r < 0) continue;
    const int u = (int)(unsigned)(r & 0xffffffffll), ru = plan_r[(size_t)i * E + e];
    double a[3], b[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      a[k] = cs[3 * u + k];
      b[k] = cs[3 * ru + k];
    }
    place(r, a, b, 1 - bdamage[u], 1 - bdamage[ru]);
  }
  __syncthreads();
  if (!row) return;  // no barrier below
  double K[D][D];
#pragma unroll
  for (int r = 0; r < D; ++r)
#pragma unroll
    for (int s = 0; s < D; ++s) K[r][s] = 0.0;
  if (active && j == i) {  // the diagonal block from the per-particle pass, eliminated
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
      for (int s = 0; s < D; ++s) K[r][s] = Ldg[r * D + s];
    if (mi) {
      const double nd = sc->norm_diag;
#pragma unroll
      for (int r = 0; r < D; ++r)
#pragma unroll
        for (int s = 0; s < D; ++s)
          if (((mi >> r) & 1) || ((mi >> s) & 1)) K[r][s] = (r == s) ? nd : 0.0;
    }
  } else if (active) {
    const double tv_j = Ltvj[lane];
    int q = pw & 1023;
    const int qend = q + ((pw >> 10) & 127);
    for (int L = 0; L < 2; ++L) {
      double acc[D][D];
#pragma unroll
      for (int r = 0; r < D; ++r)
#pragma unroll
        for (int s = 0; s < D; ++s) acc[r][s] = 0.0;
      const int t1 = ((pw >> (L == 0 ? 17 : 24)) & 127) - 1;
      bool found = t1 >= 0;
      if (found) {  // bond t1 = i -> j; the table holds j's inputs for t1's layer
        double Qj[3], Sj[3], c1[3], cr1[3];
#pragma unroll
        for (int u = 0; u < 3; ++u) {
          Qj[u] = Lt[(9 + u) * NBM + t1];
          Sj[u] = Lt[(12 + u) * NBM + t1];
          c1[u] = Lt[u * NBM + t1];
          cr1[u] = Lt[(4 + u) * NBM + t1];
        }
        double A[D][D], B[D][D];
        direct_common_half<D>(kni[L], tv_i, Qi[L], Si[L], Qj, c1, Lt[3 * NBM + t1], A);
        direct_common_half<D>(Lt[15 * NBM + t1], tv_j, Qj, Sj, Qi[L], cr1, Lt[7 * NBM + t1], B);
#pragma unroll
        for (int r = 0; r < D; ++r)
#pragma unroll
          for (int s = 0; s < D; ++s) acc[r][s] = 0.5 * A[r][s] + 0.5 * B[s][r];
      }
      const int thi = L == 0 ? n0 : nbi;
      while (q < qend) {
        const int st = Lst[q];
        const int t = st & 63;
        if (t >= thi) break;
        found = true;
        double ct[3], cu[3];
#pragma unroll
        for (int u = 0; u < 3; ++u) {
          ct[u] = Lt[u * NBM + t];
          cu[u] = Lc3[3 * q + u];
        }
        if (!(st & kOneSided)) {  // c_rev t = -c_t, c_rev u = -c_u: the two products are equal
          const double cc = Lcc[q];
#pragma unroll
          for (int r = 0; r < D; ++r)
#pragma unroll
            for (int s = 0; s < D; ++s) acc[r][s] -= cc * (ct[r] * cu[s]);
        } else {  // a bond broken on one side only: the halves separately
          const int u = Lu[q], ru = rev[u];
          double crt[3], cru[3];
#pragma unroll
          for (int k = 0; k < 3; ++k) {
            crt[k] = Lt[(4 + k) * NBM + t];
            cru[k] = cs[3 * ru + k];
          }
          const double tvk = Lt[8 * NBM + t];
          const double cA = 0.5 * ((1 - bdamage[u]) * tvk + Lt[3 * NBM + t] * tv_i);
          const double cB = 0.5 * (Lt[7 * NBM + t] * tvk + (1 - bdamage[ru]) * tv_j);
#pragma unroll
          for (int r = 0; r < D; ++r)
#pragma unroll
            for (int s = 0; s < D; ++s) acc[r][s] -= 0.5 * (cA * (ct[r] * cu[s]) + cB * (crt[r] * cru[s]));
        }
        ++q;
      }
      if (found)
#pragma unroll
        for (int r = 0; r < D; ++r)
#pragma unroll
          for (int s = 0; s < D; ++s) K[r][s] = K[r][s] + acc[r][s];
    }
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
      for (int s = 0; s < D; ++s)
        if (((mi >> r) & 1) || ((mj >> s) & 1)) K[r][s] = 0.0;
  }
  // The row's blocks are contiguous in HBM: stage them in this wavefront's
  // slice (its path data is no longer read: the wavefront's LDS operations
  // complete in order) and store them with consecutive lanes on consecutive
  // values.
  double* const Ks = Lreg;
  if (active)
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
      for (int s = 0; s < D; ++s) Ks[lane * DD + r * D + s] = K[r][s];
  const int nv = nc * DD;
  double* const out = Kval + (size_t)c_lo * DD;
  for (int x = lane; x < nv; x += 64) out[x] = Ks[x];
  if (Kval32) {
    float* const o32 = Kval32 + (size_t)c_lo * DD;
    for (int x = lane; x < nv; x += 64) o32[x] = (float)Ks[x];
  }
}

}  // namespace

int launch_assemble(Context* c, int want32) {
  const int N = c->N;
  const int nbw = (N + kWaves - 1) / kWaves;
  const int tb = 256;
  const int nbn = (N + tb - 1) / tb;
  const int nbk = (int)((c->nnzb + tb - 1) / tb);
  if (nbn > c->n_partials) {
    set_error("partials buffer too small");
    return -1;
  }
  if (want32 && !c->Kval32 && dalloc(c, &c->Kval32, (size_t)c->nnzb * c->dim * c->dim)) return -3;
  if (c->floating_dirty)
    if (int rc = update_floating(c)) return rc;
  float* k32 = want32 ? c->Kval32 : nullptr;
  // one pass (tangent_sym_kernel) when every row fits a wavefront and LDS
  {
    const int E = std::max(32, (c->max_cand + 7) / 8 * 8);
    const int NBM = std::max(1, c->max_nb);
    const int DD = c->dim * c->dim;
    // slots per row never exceed E, so this bounds the LDS of any plan
    if (c->bonds_unique && c->max_nb <= 64 && c->max_nconn <= 64 && E <= 1023 &&
        (size_t)kWaves * sym_slice_bytes(E, NBM, DD) <= 65536) {
      if (c->plan_E != E) {  // the path plan, once per topology
        if (dalloc(c, &c->plan_e, (size_t)N * E) || dalloc(c, &c->plan_r, (size_t)N * E) ||
            dalloc(c, &c->plan_l, (size_t)N * 64) || dalloc(c, &c->plan_n, (size_t)N + 1))
          return -3;
        LPMG_HIP(hipMemsetAsync(c->plan_n + N, 0, sizeof(int), c->stream));
        tangent_plan_kernel<<<nbw, 64 * kWaves, (size_t)kWaves * plan_slice_bytes(E), c->stream>>>(
            c->bptr, c->nb0, c->p2, c->rev, c->cptr, c->conns, N, E, c->plan_e, c->plan_r, c->plan_l, c->plan_n,
            c->plan_n + N);
        LPMG_HIP(hipGetLastError());
        int smax = 0;
        LPMG_HIP(copy_sync(c, &smax, c->plan_n + N, sizeof(int), hipMemcpyDeviceToHost));
        c->plan_E = E;
        c->plan_S = std::max(1, smax);
      }
      const int S = c->plan_S;
      const size_t lds = (size_t)kWaves * sym_slice_bytes(S, NBM, DD);
      if (lds > 65536) {
        set_error("one-pass assembly: path data exceeds LDS");
        return -1;
      }
      if (c->dim == 3) {
        assembly_prep_kernel<3><<<nbn, tb, 0, c->stream>>>(c->bptr, c->nb0, c->cs, c->bdamage, c->cs_sum, c->Kn,
                                                          c->Tv, c->dpos, c->cons, c->damage_visual, c->floating,
                                                          N, c->fixm, c->cs_wsum, c->diag_sq, c->Kval,
                                                          c->partials);
        norm_diag_finish_kernel<<<1, tb, 0, c->stream>>>(c->partials, nbn, c->sc);
        tangent_sym_kernel<3><<<nbw, 64 * kWaves, lds, c->stream>>>(
            c->bptr, c->nb0, c->p2, c->rev, c->cs, c->bdamage, c->cs_sum, c->cs_wsum, c->Kn, c->Tv, c->cptr,
            c->conns, c->fixm, c->sc, c->plan_e, c->plan_r, c->plan_l, c->plan_n, N, E, S, NBM, c->Kval, k32);
      } else {
        assembly_prep_kernel<2><<<nbn, tb, 0, c->stream>>>(c->bptr, c->nb0, c->cs, c->bdamage, c->cs_sum, c->Kn,
                                                          c->Tv, c->dpos, c->cons, c->damage_visual, c->floating,
                                                          N, c->fixm, c->cs_wsum, c->diag_sq, c->Kval,
                                                          c->partials);
        norm_diag_finish_kernel<<<1, tb, 0, c->stream>>>(c->partials, nbn, c->sc);
        tangent_sym_kernel<2><<<nbw, 64 * kWaves, lds, c->stream>>>(
            c->bptr, c->nb0, c->p2, c->rev, c->cs, c->bdamage, c->cs_sum, c->cs_wsum, c->Kn, c->Tv, c->cptr,
            c->conns, c->fixm, c->sc, c->plan_e, c->plan_r, c->plan_l, c->plan_n, N, E, S, NBM, c->Kval, k32);
      }
      LPMG_HIP(hipGetLastError());
      c->assembled64 = true;
      c->assembled32 = want32 != 0;
      c->asm_version += 1;
      return 0;
    }
  }
  // two passes: unsymmetrised blocks, then symmetrisation and elimination
  if (!c->Kloc && dalloc(c, &c->Kloc, (size_t)c->nnzb * c->dim * c->dim)) return -3;
  // LDS-staged tangent when a row's two-hop path data and direct terms fit
  // 64 KB per workgroup and its connections one wavefront
  const int nbm = c->max_nb;
  int lds_E = (c->max_nb <= 64 && c->max_nconn <= 64 && c->max_cand <= kTangentLdsMax)
                  ? std::max(32, (c->max_cand + 7) / 8 * 8) : 0;
  const size_t lds_bytes = (size_t)kWaves * ((size_t)(4 * lds_E + 64) * sizeof(double) + 256 * sizeof(int) +
                                             (size_t)2 * lds_E * sizeof(short) +
                                             (size_t)nbm * c->dim * c->dim * sizeof(double));
  if (lds_bytes > 65536) lds_E = 0;
  if (c->dim == 3) {
    if (lds_E > 0)
      tangent_lds_kernel<3><<<nbw, 64 * kWaves, lds_bytes, c->stream>>>(c->bptr, c->nb0, c->p2, c->cs, c->bdamage,
                                                                       c->cs_sum, c->Kn, c->Tv, c->cptr, c->conns,
                                                                       N, lds_E, nbm, c->Kloc, c->diag_sq);
    else if (c->max_nb <= 64)
      tangent_lanes_kernel<3><<<nbw, 64 * kWaves, 0, c->stream>>>(c->bptr, c->nb0, c->p2, c->cs, c->bdamage,
                                                                 c->cs_sum, c->Kn, c->Tv, c->cptr, c->conns,
                                                                 N, c->Kloc, c->diag_sq);
    else
      tangent_kernel<3><<<nbw, 64 * kWaves, 0, c->stream>>>(c->bptr, c->nb0, c->p2, c->cs, c->bdamage,
                                                           c->cs_sum, c->Kn, c->Tv, c->cptr, c->conns,
                                                           N, c->Kloc, c->diag_sq);
    norm_diag_kernel<<<nbn, tb, 0, c->stream>>>(c->diag_sq, N, c->partials, c->sc);
    norm_diag_finish_kernel<<<1, tb, 0, c->stream>>>(c->partials, nbn, c->sc);
    fixed_dofs_kernel<3><<<nbn, tb, 0, c->stream>>>(c->cons, c->damage_visual, c->floating, N, c->fixm);
    symmetrize_bc_kernel<3><<<nbk, tb, 0, c->stream>>>((int)c->nnzb, c->brow, c->conns, c->tpos, c->fixm, c->sc,
                                                      c->Kloc, c->Kval, k32);
  } else {
    if (lds_E > 0)
      tangent_lds_kernel<2><<<nbw, 64 * kWaves, lds_bytes, c->stream>>>(c->bptr, c->nb0, c->p2, c->cs, c->bdamage,
                                                                       c->cs_sum, c->Kn, c->Tv, c->cptr, c->conns,
                                                                       N, lds_E, nbm, c->Kloc, c->diag_sq);
    else if (c->max_nb <= 64)
      tangent_lanes_kernel<2><<<nbw, 64 * kWaves, 0, c->stream>>>(c->bptr, c->nb0, c->p2, c->cs, c->bdamage,
                                                                 c->cs_sum, c->Kn, c->Tv, c->cptr, c->conns,
                                                                 N, c->Kloc, c->diag_sq);
    else
      tangent_kernel<2><<<nbw, 64 * kWaves, 0, c->stream>>>(c->bptr, c->nb0, c->p2, c->cs, c->bdamage,
                                                           c->cs_sum, c->Kn, c->Tv, c->cptr, c->conns,
                                                           N, c->Kloc, c->diag_sq);
    norm_diag_kernel<<<nbn, tb, 0, c->stream>>>(c->diag_sq, N, c->partials, c->sc);
    norm_diag_finish_kernel<<<1, tb, 0, c->stream>>>(c->partials, nbn, c->sc);
    fixed_dofs_kernel<2><<<nbn, tb, 0, c->stream>>>(c->cons, c->damage_visual, c->floating, N, c->fixm);
    symmetrize_bc_kernel<2><<<nbk, tb, 0, c->stream>>>((int)c->nnzb, c->brow, c->conns, c->tpos, c->fixm, c->sc,
                                                      c->Kloc, c->Kval, k32);
  }
  LPMG_HIP(hipGetLastError());
  c->assembled64 = true;
  c->assembled32 = want32 != 0;
  c->asm_version += 1;
  return 0;
}

int launch_export_csr(Context* c, int64_t* dKp, int64_t* dIK, int64_t* dJK, double* dK) {
  (void)dKp;
  const int N = c->N, tb = 256, nb = (N + tb - 1) / tb;
  if (c->dim == 3)
    export_csr_kernel<3><<<nb, tb, 0, c->stream>>>(c->Kp, c->cptr, c->conns, c->dpos, c->ncl, c->Kval,
                                                  N, dIK, dJK, dK);
  else
    export_csr_kernel<2><<<nb, tb, 0, c->stream>>>(c->Kp, c->cptr, c->conns, c->dpos, c->ncl, c->Kval,
                                                  N, dIK, dJK, dK);
  LPMG_HIP(hipGetLastError());
  return 0;
}

}  // namespace lpmg
