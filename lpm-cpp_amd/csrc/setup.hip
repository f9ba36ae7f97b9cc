// setup.hip -- particle-system topology on the GPU: bonds, opposite bonds,
// connections (conns) and the static CSR/BSR pattern derived from them.
//
// Reference: Assembly::createBonds (assembly.h:212-234), Assembly::readBond
// (assembly.h:390-424), Assembly::updateConnections (assembly.h:236-266),
// Stiffness::initialize (stiffness.h:54-71), Assembly::searchNonlocalNeighbors
// (assembly.h:117-130).  The reference searches all N^2 pairs; here a uniform
// cell grid (cell edge >= the search radius) visits a superset of the candidate
// pairs, the pair test is the reference's own expression, and every per-particle
// list is ordered by partner id -- the order the reference's id-ordered loop
// produces -- so bonds, layers and conns come out identical.
#include <algorithm>
#include <cmath>

#include "lpmg_internal.h"
#include "reduce.h"

namespace lpmg {

namespace {

constexpr int kScanBlock = 256;
constexpr int kScanItems = 4;  // elements per thread
constexpr int kScanTile = kScanBlock * kScanItems;

// Exclusive scan of one tile of kScanTile elements; the block total goes to sums[blockIdx.x].
template <typename T>
__global__ void __launch_bounds__(kScanBlock) scan_tile_kernel(const T* __restrict__ in,
                                                               T* __restrict__ out,
                                                               T* __restrict__ sums, int n) {
  __shared__ T wsum[kScanBlock / 64];
  const int base = blockIdx.x * kScanTile + threadIdx.x * kScanItems;
  T v[kScanItems];
  T s = 0;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = (base + k < n) ? in[base + k] : T(0);
    s += v[k];
  }
  T incl = wave_incl_scan(s);
  const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
  if (lane == 63) wsum[wid] = incl;
  __syncthreads();
  T wpre = 0;
  for (int w = 0; w < wid; ++w) wpre += wsum[w];
  T run = wpre + incl - s;
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) out[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == kScanBlock - 1) {
    T tot = 0;
    for (int w = 0; w < kScanBlock / 64; ++w) tot += wsum[w];
    sums[blockIdx.x] = tot;
  }
}

// Exclusive scan of the tile sums by one workgroup, looping with a carry.
template <typename T>
__global__ void __launch_bounds__(kScanBlock) scan_sums_kernel(T* sums, int nt, T* total) {
  __shared__ T wsum[kScanBlock / 64];
  __shared__ T carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int b = 0; b < nt; b += kScanBlock) {
    const int t = b + threadIdx.x;
    T s = (t < nt) ? sums[t] : T(0);
    T incl = wave_incl_scan(s);
    const int wid = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (lane == 63) wsum[wid] = incl;
    __syncthreads();
    T wpre = 0;
    for (int w = 0; w < wid; ++w) wpre += wsum[w];
    if (t < nt) sums[t] = carry + wpre + incl - s;
    __syncthreads();
    if (threadIdx.x == 0) {
      T tot = 0;
      for (int w = 0; w < kScanBlock / 64; ++w) tot += wsum[w];
      carry += tot;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

template <typename T>
__global__ void scan_add_kernel(T* out, const T* sums, int n) {
  const int i = blockIdx.x * kScanTile + threadIdx.x;
  const T add = sums[blockIdx.x];
#pragma unroll
  for (int k = 0; k < kScanItems; ++k) {
    const int q = i + k * kScanBlock;
    if (q < n) out[q] += add;
  }
}

template <typename T>
int exclusive_scan(Context* c, const T* in, T* out, int n) {
  // out has n+1 entries; out[n] = total
  const int nt = (n + kScanTile - 1) / kScanTile;
  T* sums = nullptr;
  LPMG_HIP(hipMallocAsync((void**)&sums, sizeof(T) * (nt + 1), c->stream));
  if (n > 0) {
    scan_tile_kernel<T><<<nt, kScanBlock, 0, c->stream>>>(in, out, sums, n);
  }
  scan_sums_kernel<T><<<1, kScanBlock, 0, c->stream>>>(sums, nt, out + n);
  if (n > 0) scan_add_kernel<T><<<nt, kScanBlock, 0, c->stream>>>(out, sums, n);
  LPMG_HIP(hipGetLastError());
  LPMG_HIP(hipFreeAsync(sums, c->stream));
  return 0;
}

struct Grid {
  double lo[3];
  double h;
  int nc[3];
};

// Particle::distanceTo (particle.h:276-280); built with -ffp-contract=off so the
// sum of squares rounds exactly as the reference's.
__device__ __forceinline__ double pdist(double4 a, double4 b) {
  const double dx = a.x - b.x, dy = a.y - b.y, dz = a.z - b.z;
  return sqrt(dx * dx + dy * dy + dz * dz);
}

__device__ __forceinline__ int cell_coord(double x, double lo, double h, int n) {
  int q = (int)floor((x - lo) / h);
  return q < 0 ? 0 : (q >= n ? n - 1 : q);
}

__global__ void cell_id_kernel(const double4* __restrict__ xyz, int N, Grid g, int* __restrict__ cid,
                               int* __restrict__ ccount) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double4 p = xyz[i];
  const int cx = cell_coord(p.x, g.lo[0], g.h, g.nc[0]);
  const int cy = cell_coord(p.y, g.lo[1], g.h, g.nc[1]);
  const int cz = cell_coord(p.z, g.lo[2], g.h, g.nc[2]);
  const int q = (cz * g.nc[1] + cy) * g.nc[0] + cx;
  cid[i] = q;
  atomicAdd(&ccount[q], 1);
}

__global__ void cell_fill_kernel(const int* __restrict__ cid, int N, const int* __restrict__ cstart,
                                 int* __restrict__ cfill, int* __restrict__ items) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const int q = cid[i];
  items[cstart[q] + atomicAdd(&cfill[q], 1)] = i;
}

// mode 0: count layer-0 / layer-1 partners; mode 1: write them (unsorted)
__global__ void bond_search_kernel(const double4* __restrict__ xyz, int N, Grid g,
                                   const int* __restrict__ cid, const int* __restrict__ cstart,
                                   const int* __restrict__ items, double rc0, double rc1, int mode,
                                   int* __restrict__ n0, int* __restrict__ n1,
                                   const int* __restrict__ bptr, int* __restrict__ p2,
                                   double* __restrict__ dis0) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double4 pi = xyz[i];
  const int q = cid[i];
  const int cx = q % g.nc[0], cy = (q / g.nc[0]) % g.nc[1], cz = q / (g.nc[0] * g.nc[1]);
  int c0 = 0, c1 = 0;
  int w0 = 0, w1 = 0;
  if (mode == 1) { w0 = bptr[i]; w1 = bptr[i] + n0[i]; }
  for (int z = max(0, cz - 1); z <= min(g.nc[2] - 1, cz + 1); ++z)
    for (int y = max(0, cy - 1); y <= min(g.nc[1] - 1, cy + 1); ++y)
      for (int x = max(0, cx - 1); x <= min(g.nc[0] - 1, cx + 1); ++x) {
        const int cc = (z * g.nc[1] + y) * g.nc[0] + x;
        for (int t = cstart[cc]; t < cstart[cc + 1]; ++t) {
          const int j = items[t];
          const double d = pdist(pi, xyz[j]);
          if (d < rc1 && j != i) {
            if (d < rc0) {
              if (mode == 1) { p2[w0] = j; dis0[w0] = d; ++w0; }
              ++c0;
            } else {
              if (mode == 1) { p2[w1] = j; dis0[w1] = d; ++w1; }
              ++c1;
            }
          }
        }
      }
  if (mode == 0) { n0[i] = c0; n1[i] = c1; }
}

__global__ void nb_total_kernel(const int* n0, const int* n1, int* nb, int N) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) nb[i] = n0[i] + n1[i];
}

// Insertion sort of each layer range by partner id, carrying dis0.
__global__ void bond_sort_kernel(const int* __restrict__ bptr, const int* __restrict__ nb0, int N,
                                 int* __restrict__ p2, double* __restrict__ dis0) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  for (int L = 0; L < 2; ++L) {
    const int lo = L == 0 ? bptr[i] : bptr[i] + nb0[i];
    const int hi = L == 0 ? bptr[i] + nb0[i] : bptr[i + 1];
    for (int a = lo + 1; a < hi; ++a) {
      const int key = p2[a];
      const double dv = dis0[a];
      int b = a - 1;
      while (b >= lo && p2[b] > key) {
        p2[b + 1] = p2[b];
        dis0[b + 1] = dis0[b];
        --b;
      }
      p2[b + 1] = key;
      dis0[b + 1] = dv;
    }
  }
}

// Bond(p1,p2,layer,dis) initial direction cosines (bond.h:59-72) and the
// opposite-bond index that replaces std::find_if (particle.h:207-209).
__global__ void bond_init_kernel(const double4* __restrict__ xyz, const int* __restrict__ bptr,
                                 const int* __restrict__ nb0, const int* __restrict__ p2,
                                 const double* __restrict__ dis0, int N, double* __restrict__ cs,
                                 int* __restrict__ rev, int* __restrict__ err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double4 pi = xyz[i];
  for (int b = bptr[i]; b < bptr[i + 1]; ++b) {
    const int j = p2[b];
    const int L = (b < bptr[i] + nb0[i]) ? 0 : 1;
    const double4 pj = xyz[j];
    const double d = dis0[b];
    cs[3 * b] = (pi.x - pj.x) / d;
    cs[3 * b + 1] = (pi.y - pj.y) / d;
    cs[3 * b + 2] = (pi.z - pj.z) / d;
    const int lo = L == 0 ? bptr[j] : bptr[j] + nb0[j];
    const int hi = L == 0 ? bptr[j] + nb0[j] : bptr[j + 1];
    int r = -1;
    for (int q = lo; q < hi; ++q)
      if (p2[q] == i) { r = q; break; }
    rev[b] = r;
    if (r < 0) atomicOr(err, 1);
  }
}

// Candidates of a particle are its bonds' ends and their same-layer bonds' ends
// (a regular lattice: about 200-350).  The LDS per wavefront is sized to the
// largest candidate count of the system (cand_count_kernel), so regular
// lattices keep many wavefronts per CU and bond lists with hub particles still
// fit: 4 wavefronts per workgroup up to kConnCap4 candidates, 1 up to
// kConnCap1 (5 bytes per candidate, at most 64 KB per workgroup).
constexpr int kConnCap4 = 3072;
constexpr int kConnCap1 = 12288;

// cmax[0]: the largest candidate count; cmax[1] set to 1 when some particle has
// two bonds of one layer to the same particle, or a bond to itself.
__global__ void cand_count_kernel(const int* __restrict__ bptr, const int* __restrict__ nb0,
                                  const int* __restrict__ p2, int N, int* __restrict__ cmax) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  int C = 0;
  bool dup = false;
  for (int L = 0; L < 2; ++L) {
    const int lo = L == 0 ? bptr[i] : bptr[i] + nb0[i];
    const int hi = L == 0 ? bptr[i] + nb0[i] : bptr[i + 1];
    for (int b = lo; b < hi; ++b) {
      const int k = p2[b];
      C += 1 + (L == 0 ? nb0[k] : bptr[k + 1] - bptr[k] - nb0[k]);
      dup = dup || k == i;
      for (int q = lo; q < b; ++q) dup = dup || p2[q] == k;
    }
  }
  atomicMax(cmax, C);
  if (dup) atomicOr(cmax + 1, 1);
}

// Assembly::updateConnections: one wavefront per particle gathers the 1- and
// 2-hop same-layer candidates into LDS, then each lane ranks its candidates
// (first occurrence, number of distinct smaller ids) -- a sort+unique without
// a sort.  mode 0 counts (nconn, nconn_largeq), mode 1 writes conns and dpos.
__global__ void __launch_bounds__(256)
conns_kernel(const int* __restrict__ bptr, const int* __restrict__ nb0, const int* __restrict__ p2,
             int N, int mode, int cap, int* __restrict__ nconn, int* __restrict__ ncl,
             const int* __restrict__ cptr, int* __restrict__ conns, int* __restrict__ dpos,
             int* __restrict__ err) {
  extern __shared__ int smem[];  // cand[waves][cap] then first[waves][cap]
  const int waves = blockDim.x >> 6;
  const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
  int* const candw = smem + (size_t)w * cap;
  unsigned char* const firstw = (unsigned char*)(smem + (size_t)waves * cap) + (size_t)w * cap;
  const int i = blockIdx.x * waves + w;
  const bool active = i < N;
  int C = 0;
  if (active) {
    for (int L = 0; L < 2; ++L) {
      const int lo = L == 0 ? bptr[i] : bptr[i] + nb0[i];
      const int hi = L == 0 ? bptr[i] + nb0[i] : bptr[i + 1];
      for (int b0 = lo; b0 < hi; b0 += 64) {
        const int b = b0 + lane;
        int k = -1, klo = 0, khi = 0, cnt = 0;
        if (b < hi) {
          k = p2[b];
          klo = L == 0 ? bptr[k] : bptr[k] + nb0[k];
          khi = L == 0 ? bptr[k] + nb0[k] : bptr[k + 1];
          cnt = 1 + (khi - klo);
        }
        const int incl = wave_incl_scan(cnt);
        const int off = C + incl - cnt;
        if (b < hi) {
          if (off + cnt <= cap) {
            candw[off] = k;
            for (int q = klo; q < khi; ++q) candw[off + 1 + (q - klo)] = p2[q];
          }
        }
        C += __shfl(incl, 63, 64);
      }
    }
    if (C > cap) {
      if (lane == 0) atomicOr(err, 2);
      C = 0;
    }
  }
  __syncthreads();
  for (int t = lane; t < C; t += 64) {
    const int v = candw[t];
    unsigned char f = 1;
    for (int u = 0; u < t; ++u)
      if (candw[u] == v) { f = 0; break; }
    firstw[t] = f;
  }
  __syncthreads();
  int my_distinct = 0, my_ge = 0;
  for (int t = lane; t < C; t += 64) {
    if (!firstw[t]) continue;
    const int v = candw[t];
    ++my_distinct;
    if (v >= i) ++my_ge;
    if (mode == 1) {
      int rank = 0;
      for (int u = 0; u < C; ++u) rank += (firstw[u] && candw[u] < v) ? 1 : 0;
      conns[cptr[i] + rank] = v;
      if (v == i) dpos[i] = cptr[i] + rank;
    }
  }
  if (mode == 0) {
    for (int o = 32; o > 0; o >>= 1) {
      my_distinct += __shfl_xor(my_distinct, o, 64);
      my_ge += __shfl_xor(my_ge, o, 64);
    }
    if (active && lane == 0) {
      nconn[i] = my_distinct;
      ncl[i] = my_ge;
      if (my_ge == 0) atomicOr(err, 4);
    }
  }
}

// K_pointer increments d*d*nconn_largeq - (1|3) (stiffness.h:61-64)
__global__ void kp_incr_kernel(const int* __restrict__ ncl, int N, int d, int64_t* __restrict__ inc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < N) inc[i] = (int64_t)d * d * ncl[i] - (d == 2 ? 1 : 3);
}

// Block-row index and transpose position of every BSR block.
__global__ void bsr_aux_kernel(const int* __restrict__ cptr, const int* __restrict__ conns, int N,
                               int* __restrict__ brow, int* __restrict__ tpos, int* __restrict__ err) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  for (int e = cptr[i]; e < cptr[i + 1]; ++e) {
    const int j = conns[e];
    brow[e] = i;
    int lo = cptr[j], hi = cptr[j + 1] - 1, pos = -1;
    while (lo <= hi) {
      const int mid = (lo + hi) >> 1;
      const int v = conns[mid];
      if (v == i) { pos = mid; break; }
      if (v < i) lo = mid + 1; else hi = mid - 1;
    }
    tpos[e] = pos;
    if (pos < 0) atomicOr(err, 8);
  }
}

int grid_for(const std::vector<double>& hx, int N, double radius_search, Grid* g) {
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (int i = 0; i < N; ++i)
    for (int k = 0; k < 3; ++k) {
      lo[k] = std::min(lo[k], hx[3 * (size_t)i + k]);
      hi[k] = std::max(hi[k], hx[3 * (size_t)i + k]);
    }
  g->h = radius_search * 1.0000001;
  int64_t total = 1;
  for (int k = 0; k < 3; ++k) {
    g->lo[k] = lo[k];
    const int64_t n = std::max<int64_t>(1, (int64_t)std::floor((hi[k] - lo[k]) / g->h) + 1);
    g->nc[k] = (int)std::min<int64_t>(n, 1 << 20);
    total *= g->nc[k];
  }
  // keep the cell table modest: coarsen the grid when particles are sparse
  while (total > 8 * (int64_t)N + 1024) {
    g->h *= 1.5;
    total = 1;
    for (int k = 0; k < 3; ++k) {
      g->nc[k] = std::max(1, (int)std::floor((hi[k] - lo[k]) / g->h) + 1);
      total *= g->nc[k];
    }
  }
  return 0;
}

// Cell lists of the particles (counting sort by cell).
int bin_particles(Context* c, const Grid& g, int** cid, int** cstart, int** items) {
  const int N = c->N;
  const int ncell = g.nc[0] * g.nc[1] * g.nc[2];
  int *ccount = nullptr, *cfill = nullptr;
  LPMG_HIP(hipMallocAsync((void**)cid, sizeof(int) * N, c->stream));
  LPMG_HIP(hipMallocAsync((void**)&ccount, sizeof(int) * (ncell + 1), c->stream));
  LPMG_HIP(hipMallocAsync((void**)cstart, sizeof(int) * (ncell + 1), c->stream));
  LPMG_HIP(hipMallocAsync((void**)&cfill, sizeof(int) * (ncell + 1), c->stream));
  LPMG_HIP(hipMallocAsync((void**)items, sizeof(int) * N, c->stream));
  LPMG_HIP(hipMemsetAsync(ccount, 0, sizeof(int) * (ncell + 1), c->stream));
  LPMG_HIP(hipMemsetAsync(cfill, 0, sizeof(int) * (ncell + 1), c->stream));
  const int tb = 256, nb = (N + tb - 1) / tb;
  cell_id_kernel<<<nb, tb, 0, c->stream>>>(c->xyz, N, g, *cid, ccount);
  if (int rc = exclusive_scan<int>(c, ccount, *cstart, ncell)) return rc;
  cell_fill_kernel<<<nb, tb, 0, c->stream>>>(*cid, N, *cstart, cfill, *items);
  LPMG_HIP(hipGetLastError());
  LPMG_HIP(hipFreeAsync(ccount, c->stream));
  LPMG_HIP(hipFreeAsync(cfill, c->stream));
  return 0;
}

int alloc_bond_arrays(Context* c) {
  const size_t NB = (size_t)c->NB;
  int rc = 0;
  rc |= dalloc(c, &c->rev, NB);
  for (double** p : {&c->dis, &c->dLe, &c->bstrain, &c->dLp, &c->dLp_last, &c->bforce,
                     &c->bforce_last, &c->bdamage, &c->bdamage_last, &c->dis_last})
    rc |= dalloc(c, p, NB);
  rc |= dalloc(c, &c->cs, 3 * NB);
  if (rc) return -3;
  for (double* p : {c->dLe, c->bstrain, c->dLp, c->dLp_last, c->bforce, c->bforce_last, c->bdamage,
                    c->bdamage_last})
    LPMG_HIP(hipMemsetAsync(p, 0, sizeof(double) * NB, c->stream));
  LPMG_HIP(hipMemcpyAsync(c->dis, c->dis0, sizeof(double) * NB, hipMemcpyDeviceToDevice, c->stream));
  LPMG_HIP(hipMemcpyAsync(c->dis_last, c->dis0, sizeof(double) * NB, hipMemcpyDeviceToDevice, c->stream));
  return 0;
}

int check_err(Context* c, const char* what) {
  int h = 0;
  LPMG_HIP(hipMemcpyAsync(&h, c->err_flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  LPMG_HIP(hipStreamSynchronize(c->stream));
  if (h) {
    set_error(std::string(what) + " failed (device flag " + std::to_string(h) + ")");
    LPMG_HIP(hipMemsetAsync(c->err_flag, 0, sizeof(int), c->stream));
    return -1;
  }
  return 0;
}

}  // namespace

int exclusive_scan_i32(Context* c, const int* in, int* out, int n, int* total) {
  if (int rc = exclusive_scan<int>(c, in, out, n)) return rc;
  if (total) {
    LPMG_HIP(hipMemcpyAsync(total, out + n, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    LPMG_HIP(hipStreamSynchronize(c->stream));
  }
  return 0;
}

int exclusive_scan_i64(Context* c, const int64_t* in, int64_t* out, int n) {
  return exclusive_scan<int64_t>(c, in, out, n);
}

int build_bonds_from_search(Context* c, const double* host_xyz) {
  const int N = c->N;
  std::vector<double> hx(host_xyz, host_xyz + 3 * (size_t)N);
  const double rc1 = 1.01 * c->cell.cut[1], rc0 = 1.01 * c->cell.cut[0];
  Grid g;
  grid_for(hx, N, rc1, &g);
  int *cid = nullptr, *cstart = nullptr, *items = nullptr, *n0 = nullptr, *n1 = nullptr, *nbv = nullptr;
  if (int rc = bin_particles(c, g, &cid, &cstart, &items)) return rc;
  LPMG_HIP(hipMallocAsync((void**)&n0, sizeof(int) * N, c->stream));
  LPMG_HIP(hipMallocAsync((void**)&n1, sizeof(int) * N, c->stream));
  LPMG_HIP(hipMallocAsync((void**)&nbv, sizeof(int) * N, c->stream));
  const int tb = 256, nb = (N + tb - 1) / tb;
  bond_search_kernel<<<nb, tb, 0, c->stream>>>(c->xyz, N, g, cid, cstart, items, rc0, rc1, 0, n0, n1,
                                               nullptr, nullptr, nullptr);
  nb_total_kernel<<<nb, tb, 0, c->stream>>>(n0, n1, nbv, N);
  LPMG_HIP(hipGetLastError());
  if (dalloc(c, &c->bptr, N + 1) || dalloc(c, &c->nb0, N)) return -3;
  int total = 0;
  if (int rc = exclusive_scan_i32(c, nbv, c->bptr, N, &total)) return rc;
  c->NB = total;
  LPMG_HIP(hipMemcpyAsync(c->nb0, n0, sizeof(int) * N, hipMemcpyDeviceToDevice, c->stream));
  if (dalloc(c, &c->p2, c->NB) || dalloc(c, &c->dis0, c->NB)) return -3;
  bond_search_kernel<<<nb, tb, 0, c->stream>>>(c->xyz, N, g, cid, cstart, items, rc0, rc1, 1, n0, n1,
                                               c->bptr, c->p2, c->dis0);
  bond_sort_kernel<<<nb, tb, 0, c->stream>>>(c->bptr, c->nb0, N, c->p2, c->dis0);
  LPMG_HIP(hipGetLastError());
  for (void* p : {(void*)cid, (void*)cstart, (void*)items, (void*)n0, (void*)n1, (void*)nbv})
    LPMG_HIP(hipFreeAsync(p, c->stream));
  return alloc_bond_arrays(c);
}

__global__ void bond_dist_kernel(const double4* __restrict__ xyz, const int* __restrict__ bptr,
                                 const int* __restrict__ p2, int N, double* __restrict__ dis0) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  for (int b = bptr[i]; b < bptr[i + 1]; ++b) dis0[b] = pdist(xyz[i], xyz[p2[b]]);
}

// Assembly::readBond: bonds keep the list's order within each (particle, layer).
int build_bonds_from_list(Context* c, int64_t nbond, const int32_t* p1, const int32_t* p2h,
                          const int32_t* layer) {
  const int N = c->N;
  std::vector<int> cnt0(N, 0), cnt1(N, 0);
  for (int64_t t = 0; t < nbond; ++t) {
    if (p1[t] < 0 || p1[t] >= N || p2h[t] < 0 || p2h[t] >= N || (layer[t] != 0 && layer[t] != 1)) {
      set_error("bond list holds an id or layer out of range");
      return -1;
    }
    (layer[t] == 0 ? cnt0 : cnt1)[p1[t]]++;
  }
  std::vector<int> bptr(N + 1, 0), fill0(N), fill1(N), hp2(nbond);
  for (int i = 0; i < N; ++i) bptr[i + 1] = bptr[i] + cnt0[i] + cnt1[i];
  for (int i = 0; i < N; ++i) { fill0[i] = bptr[i]; fill1[i] = bptr[i] + cnt0[i]; }
  for (int64_t t = 0; t < nbond; ++t) hp2[(layer[t] == 0 ? fill0 : fill1)[p1[t]]++] = p2h[t];
  c->NB = nbond;
  if (dalloc(c, &c->bptr, N + 1) || dalloc(c, &c->nb0, N) || dalloc(c, &c->p2, c->NB) ||
      dalloc(c, &c->dis0, c->NB))
    return -3;
  LPMG_HIP(hipMemcpyAsync(c->bptr, bptr.data(), sizeof(int) * (N + 1), hipMemcpyHostToDevice, c->stream));
  LPMG_HIP(hipMemcpyAsync(c->nb0, cnt0.data(), sizeof(int) * N, hipMemcpyHostToDevice, c->stream));
  LPMG_HIP(hipMemcpyAsync(c->p2, hp2.data(), sizeof(int) * nbond, hipMemcpyHostToDevice, c->stream));
  bond_dist_kernel<<<(N + 255) / 256, 256, 0, c->stream>>>(c->xyz, c->bptr, c->p2, N, c->dis0);
  LPMG_HIP(hipGetLastError());
  LPMG_HIP(hipStreamSynchronize(c->stream));  // host vectors go out of scope
  return alloc_bond_arrays(c);
}

int build_topology(Context* c) {
  const int N = c->N;
  const int tb = 256, nb = (N + tb - 1) / tb;
  bond_init_kernel<<<nb, tb, 0, c->stream>>>(c->xyz, c->bptr, c->nb0, c->p2, c->dis0, N, c->cs,
                                             c->rev, c->err_flag);
  LPMG_HIP(hipGetLastError());
  if (check_err(c, "opposite-bond lookup (bond list not symmetric)")) return -1;
  int* nconn = nullptr;
  LPMG_HIP(hipMallocAsync((void**)&nconn, sizeof(int) * N, c->stream));
  if (dalloc(c, &c->ncl, N) || dalloc(c, &c->cptr, N + 1) || dalloc(c, &c->dpos, N)) return -3;
  int* cmax = nullptr;
  int hcm[2] = {0, 0};
  LPMG_HIP(hipMallocAsync((void**)&cmax, 2 * sizeof(int), c->stream));
  LPMG_HIP(hipMemsetAsync(cmax, 0, 2 * sizeof(int), c->stream));
  cand_count_kernel<<<nb, tb, 0, c->stream>>>(c->bptr, c->nb0, c->p2, N, cmax);
  LPMG_HIP(hipGetLastError());
  LPMG_HIP(copy_sync(c, hcm, cmax, 2 * sizeof(int), hipMemcpyDeviceToHost));
  LPMG_HIP(hipFreeAsync(cmax, c->stream));
  const int hmax = hcm[0];
  c->bonds_unique = hcm[1] == 0;
  if (hmax > kConnCap1) {
    set_error("a particle has " + std::to_string(hmax) + " connection candidates, more than " +
              std::to_string(kConnCap1));
    return -1;
  }
  c->max_cand = hmax;
  const int waves = hmax <= kConnCap4 ? 4 : 1;
  const int cap = std::max(64, (hmax + 63) / 64 * 64);
  const size_t shm = (size_t)waves * cap * (sizeof(int) + 1);
  const int cb = (N + waves - 1) / waves;
  conns_kernel<<<cb, 64 * waves, shm, c->stream>>>(c->bptr, c->nb0, c->p2, N, 0, cap, nconn, c->ncl, nullptr,
                                                   nullptr, nullptr, c->err_flag);
  LPMG_HIP(hipGetLastError());
  if (check_err(c, "connection search (capacity, or a particle without bonds)")) return -1;
  int total = 0;
  if (int rc = exclusive_scan_i32(c, nconn, c->cptr, N, &total)) return rc;
  c->nnzb = total;
  if (dalloc(c, &c->conns, c->nnzb) || dalloc(c, &c->tpos, c->nnzb) || dalloc(c, &c->brow, c->nnzb))
    return -3;
  conns_kernel<<<cb, 64 * waves, shm, c->stream>>>(c->bptr, c->nb0, c->p2, N, 1, cap, nconn, c->ncl, c->cptr,
                                                   c->conns, c->dpos, c->err_flag);
  bsr_aux_kernel<<<nb, tb, 0, c->stream>>>(c->cptr, c->conns, N, c->brow, c->tpos, c->err_flag);
  LPMG_HIP(hipGetLastError());
  if (check_err(c, "BSR transpose lookup")) return -1;
  if (dalloc(c, &c->fixm, N)) return -3;
  int64_t* inc = nullptr;
  LPMG_HIP(hipMallocAsync((void**)&inc, sizeof(int64_t) * N, c->stream));
  if (dalloc(c, &c->Kp, N + 1)) return -3;
  kp_incr_kernel<<<nb, tb, 0, c->stream>>>(c->ncl, N, c->dim, inc);
  if (int rc = exclusive_scan_i64(c, inc, c->Kp, N)) return rc;
  LPMG_HIP(hipMemcpyAsync(&c->nnz_upper, c->Kp + N, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
  LPMG_HIP(hipFreeAsync(inc, c->stream));
  LPMG_HIP(hipFreeAsync(nconn, c->stream));
  LPMG_HIP(hipStreamSynchronize(c->stream));
  // maxima (host copies are small)
  std::vector<int> hb(N + 1), hc(N + 1);
  LPMG_HIP(copy_sync(c, hb.data(), c->bptr, sizeof(int) * (N + 1), hipMemcpyDeviceToHost));
  LPMG_HIP(copy_sync(c, hc.data(), c->cptr, sizeof(int) * (N + 1), hipMemcpyDeviceToHost));
  c->max_nb = c->max_nconn = 0;
  for (int i = 0; i < N; ++i) {
    c->max_nb = std::max(c->max_nb, hb[i + 1] - hb[i]);
    c->max_nconn = std::max(c->max_nconn, hc[i + 1] - hc[i]);
  }
  return 0;
}

// ---------------------------------------------------------------- nonlocal

__global__ void nonlocal_kernel(const double4* __restrict__ xyz, int N, Grid g,
                                const int* __restrict__ cid, const int* __restrict__ cstart,
                                const int* __restrict__ items, const double* __restrict__ nlL,
                                double ratio, int mode, int* __restrict__ cnt,
                                const int* __restrict__ nlptr, int* __restrict__ nlidx) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= N) return;
  const double4 pi = xyz[i];
  const double lim = ratio * nlL[i];
  const int q = cid[i];
  const int cx = q % g.nc[0], cy = (q / g.nc[0]) % g.nc[1], cz = q / (g.nc[0] * g.nc[1]);
  const int span = 1 + (int)ceil(lim / g.h);
  int n = 0, w = (mode == 1) ? nlptr[i] : 0;
  for (int z = max(0, cz - span); z <= min(g.nc[2] - 1, cz + span); ++z)
    for (int y = max(0, cy - span); y <= min(g.nc[1] - 1, cy + span); ++y)
      for (int x = max(0, cx - span); x <= min(g.nc[0] - 1, cx + span); ++x) {
        const int cc = (z * g.nc[1] + y) * g.nc[0] + x;
        for (int t = cstart[cc]; t < cstart[cc + 1]; ++t) {
          const int j = items[t];
          if (pdist(pi, xyz[j]) < lim) {
            if (mode == 1) nlidx[w++] = j;
            ++n;
          }
        }
      }
  if (mode == 0) cnt[i] = n;
  if (mode == 1) {  // id order, as the reference's loop over pt_sys
    const int lo = nlptr[i], hi = nlptr[i + 1];
    for (int a = lo + 1; a < hi; ++a) {
      const int key = nlidx[a];
      int b = a - 1;
      while (b >= lo && nlidx[b] > key) { nlidx[b + 1] = nlidx[b]; --b; }
      nlidx[b + 1] = key;
    }
  }
}

// Assembly::searchNonlocalNeighbors (assembly.h:117-130)
int search_nonlocal(Context* c, double ratio) {
  const int N = c->N;
  std::vector<double> hx(3 * (size_t)N), hl(N);
  {
    std::vector<double4> h4(N);
    LPMG_HIP(copy_sync(c, h4.data(), c->xyz, sizeof(double4) * N, hipMemcpyDeviceToHost));
    for (int i = 0; i < N; ++i) { hx[3 * i] = h4[i].x; hx[3 * i + 1] = h4[i].y; hx[3 * i + 2] = h4[i].z; }
    LPMG_HIP(copy_sync(c, hl.data(), c->nlL, sizeof(double) * N, hipMemcpyDeviceToHost));
  }
  double maxL = 0;
  for (double v : hl) maxL = std::max(maxL, v);
  const double rad = std::max(ratio * maxL, 1.01 * c->cell.cut[1]);
  Grid g;
  grid_for(hx, N, rad, &g);
  int *cid = nullptr, *cstart = nullptr, *items = nullptr, *cnt = nullptr;
  if (int rc = bin_particles(c, g, &cid, &cstart, &items)) return rc;
  LPMG_HIP(hipMallocAsync((void**)&cnt, sizeof(int) * N, c->stream));
  const int tb = 256, nb = (N + tb - 1) / tb;
  nonlocal_kernel<<<nb, tb, 0, c->stream>>>(c->xyz, N, g, cid, cstart, items, c->nlL, ratio, 0, cnt,
                                            nullptr, nullptr);
  LPMG_HIP(hipGetLastError());
  if (!c->nlptr && dalloc(c, &c->nlptr, N + 1)) return -3;
  int total = 0;
  if (int rc = exclusive_scan_i32(c, cnt, c->nlptr, N, &total)) return rc;
  if (c->nlidx) {
    LPMG_HIP(hipFreeAsync(c->nlidx, c->stream));
    c->allocs.erase(std::find(c->allocs.begin(), c->allocs.end(), (void*)c->nlidx));
    c->nlidx = nullptr;
  }
  if (dalloc(c, &c->nlidx, (size_t)std::max(total, 1))) return -3;
  c->n_nl = total;
  nonlocal_kernel<<<nb, tb, 0, c->stream>>>(c->xyz, N, g, cid, cstart, items, c->nlL, ratio, 1, cnt,
                                            c->nlptr, c->nlidx);
  LPMG_HIP(hipGetLastError());
  for (void* p : {(void*)cid, (void*)cstart, (void*)items, (void*)cnt}) LPMG_HIP(hipFreeAsync(p, c->stream));
  LPMG_HIP(hipStreamSynchronize(c->stream));
  return 0;
}

}  // namespace lpmg
