// MI355X (gfx950) implementation of the ndspmhd hot path behind the C-ABI in
// include/ndsp.h.  One context holds one simulation resident in HBM.
//
// Pipeline of one derivative evaluation (src/derivs.f90:27):
//   k_boundary      periodic wrap / wall reflection (src/boundaryND.f90:66-160)
//   k_binstats      max h and bounding box (src/ghostND_mhd.f90:77, linkND.f90:78-90)
//   k_cellid, scan, k_scatter, k_cellsort_wave
//                   counting sort of particles by Morton cell key, stable
//                   by slot: replaces the head/next link lists of
//                   set_linklist (src/linkND.f90:45-161)
//   k_pack_density  sorted neighbour view: FP64 (x,y,z,1/h) + payload
//   k_nr_pass       h-rho Newton-Raphson passes, persistent once a pass is partial
//                   (src/iterate_density.f90:43-360 with density_sums.f90),
//                   fused conservative->primitive + EOS
//                   (src/conservative2primitive.f90:42-470, src/eos.f90:40-106)
//   k_pack_force    sorted neighbour view for the force pass
//   k_force         fused pair sums of get_rates (src/ratesND_mhd.f90:304-467,
//                   1175-1690, 1700-1894, 2154-2206, 2377-2720, 2813-2878)
//   k_finalize      per-particle finalisation (src/ratesND_mhd.f90:482-943)
//   k_dt            timestep (src/stepND_leapfrog_mhd.f90:252)
// and of one step: k_predict, derivs with the corrector in the finalize launch
// (k_finalize_correct) (stepND_leapfrog_mhd.f90:31-275).
//
// Periodic and reflecting boundaries use image cells instead of ghost copies;
// the image coordinates reproduce the ghost arithmetic of
// src/ghostND_mhd.f90:180-200 bit for bit, so neighbour sets are identical.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstddef>
#include <cstdlib>
#include <initializer_list>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>
#include <chrono>

#include "ndsp.h"
#include "ndsp_kernels.h"
#include "ndsp_mlp.h"

static thread_local std::string g_err;

static int set_err(int code, const std::string& m) {
  g_err = m;
  return code;
}

#define HIPCHK(call)                                                              \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess)                                                         \
      return set_err(NDSP_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

static const int BS = 256;

// ===========================================================================
// small kernels: boundary, statistics, binning, scan
// ===========================================================================

// boundary, src/boundaryND.f90:66-160
__device__ __forceinline__ void b_boundary(KArgs a, Phys p, Grid g, int in_step, VL vl) {
#pragma clang fp contract(off)
  int i = vl.b * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  double* xs[3] = {a.f[F_X], a.f[F_Y], a.f[F_Z]};
  double* vs[3] = {a.f[F_VX], a.f[F_VY], a.f[F_VZ]};
  double* xins[3] = {a.f[F_XIN], a.f[F_YIN], a.f[F_ZIN]};
  double* vins[3] = {a.f[F_VXIN], a.f[F_VYIN], a.f[F_VZIN]};
  double* fs[3] = {a.f[F_FX], a.f[F_FY], a.f[F_FZ]};
  for (int d = 0; d < 3; d++) if (d < p.ndim) {
    if (g.bound[d] != 3) continue;
    double x = xs[d][i];
    if (x > g.xmax[d]) x = g.xmin[d] + x - g.xmax[d];
    else if (x < g.xmin[d]) x = g.xmax[d] - (g.xmin[d] - x);
    xs[d][i] = x;
  }
  for (int d = 0; d < 3; d++) if (d < p.ndim) {
    if (g.bound[d] <= 1) continue;
    double x = xs[d][i];
    bool hit = false;
    if (x > g.xmax[d]) { x = g.xmax[d] - (x - g.xmax[d]); hit = true; }
    else if (x < g.xmin[d]) { x = g.xmin[d] + (g.xmin[d] - x); hit = true; }
    if (hit) {
      xs[d][i] = x;
      vs[d][i] = -vs[d][i];
      if (in_step) { xins[d][i] = x; vins[d][i] = vs[d][i]; }
      fs[d][i] = 0.0;
    }
  }
}
__global__ void k_boundary(KArgs a, Phys p, Grid g, int in_step) {
  b_boundary(a, p, g, in_step, VL{(int)blockIdx.x, (int)gridDim.x});
}

// per-block max h and bounding box -> part[blk*8 + 0..6]
__device__ __forceinline__ void b_binstats(KArgs a, int ndim, double* part, VL vl) {
  __shared__ double sm[7][BS];
  int t = threadIdx.x;
  double hm = 0.0, lo[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, hi[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
  const double* xs[3] = {a.f[F_X], a.f[F_Y], a.f[F_Z]};
  for (int i = vl.b * blockDim.x + t; i < a.n; i += vl.nb * blockDim.x) {
    hm = fmax(hm, a.f[F_H][i]);
    for (int d = 0; d < ndim; d++) {
      lo[d] = fmin(lo[d], xs[d][i]);
      hi[d] = fmax(hi[d], xs[d][i]);
    }
  }
  sm[0][t] = hm;
  for (int d = 0; d < 3; d++) { sm[1 + d][t] = lo[d]; sm[4 + d][t] = hi[d]; }
  __syncthreads();
  for (int s = BS / 2; s > 0; s >>= 1) {
    if (t < s) {
      sm[0][t] = fmax(sm[0][t], sm[0][t + s]);
      for (int d = 0; d < 3; d++) {
        sm[1 + d][t] = fmin(sm[1 + d][t], sm[1 + d][t + s]);
        sm[4 + d][t] = fmax(sm[4 + d][t], sm[4 + d][t + s]);
      }
    }
    __syncthreads();
  }
  if (t < 7) part[vl.b * 8 + t] = sm[t][0];
}
__global__ void k_binstats(KArgs a, int ndim, double* part) {
  b_binstats(a, ndim, part, VL{(int)blockIdx.x, (int)gridDim.x});
}

// one block of BS threads folds the per-block partials of k_binstats
__device__ __forceinline__ void b_binstats_final(DScal* sc, const double* part, int nb, VL vl) {
  __shared__ double sm[7][BS];
  double v[7] = {0.0, DBL_MAX, DBL_MAX, DBL_MAX, -DBL_MAX, -DBL_MAX, -DBL_MAX};
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    v[0] = fmax(v[0], part[b * 8]);
    for (int d = 0; d < 3; d++) {
      v[1 + d] = fmin(v[1 + d], part[b * 8 + 1 + d]);
      v[4 + d] = fmax(v[4 + d], part[b * 8 + 4 + d]);
    }
  }
  for (int q = 0; q < 7; q++) sm[q][threadIdx.x] = v[q];
  __syncthreads();
  for (int st = BS / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      const int t = threadIdx.x;
      sm[0][t] = fmax(sm[0][t], sm[0][t + st]);
      for (int d = 0; d < 3; d++) {
        sm[1 + d][t] = fmin(sm[1 + d][t], sm[1 + d][t + st]);
        sm[4 + d][t] = fmax(sm[4 + d][t], sm[4 + d][t + st]);
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    sc->hmax_bin = sm[0][0];
    for (int d = 0; d < 3; d++) { sc->bbox_lo[d] = sm[1 + d][0]; sc->bbox_hi[d] = sm[4 + d][0]; }
  }
}
__global__ __launch_bounds__(BS) void k_binstats_final(DScal* sc, const double* part, int nb) {
  b_binstats_final(sc, part, nb, VL{(int)blockIdx.x, (int)gridDim.x});
}

// Cell grid of this binning, on the device (no host round trip): cells of
// side >= radkern*hmax/cell_sub (src/linkND.f90:72 uses radkern*hhmax);
// walls tile their dimension exactly; the dimensions with the most key bits
// give up bits (their cells widen) while the Morton key space would exceed
// 2^cap_bits.  Also decides whether neighbour lists can be used
// (radkern*hmax*(1+skin) below every periodic length, so every image within
// a list radius lies in the image cells one period away).
__device__ __forceinline__ void b_grid_setup(DScal* sc, Grid* gd, Grid box, Phys p, int cap_bits, VL vl) {
  if (threadIdx.x | vl.b) return;
  Grid g = box;
  const double hmax = sc->hmax_bin;
  const double rmax = p.radkern * hmax;
  sc->grid_clamped = 0;
  if (!(rmax > 0.0) || !(rmax < DBL_MAX)) sc->err = NDSP_ERR_CELLS;
  double cs = rmax / p.cell_sub;
  double ncd[3] = {1.0, 1.0, 1.0};
  for (int d = 0; d < p.ndim; d++) {
    if (g.bound[d] >= 2) {
      const double L = g.xmax[d] - g.xmin[d];
      if (!(L >= rmax)) sc->err = NDSP_ERR_CELLS;
      ncd[d] = floor(L / cs);
      if (!(ncd[d] >= 1.0)) ncd[d] = 1.0;
    } else {
      ncd[d] = floor((sc->bbox_hi[d] - sc->bbox_lo[d]) / cs) + 1.0;
      if (!(ncd[d] >= 1.0)) ncd[d] = 1.0;
    }
    ncd[d] = fmin(ncd[d], (double)(1 << 28));
  }
  int kb[3] = {0, 0, 0}, tot = 0;
  for (int d = 0; d < p.ndim; d++) {
    while ((double)(1 << kb[d]) < ncd[d]) kb[d]++;
    tot += kb[d];
  }
  bool clamped[3] = {false, false, false};
  while (tot > cap_bits) {
    int dm = 0;
    for (int d = 1; d < p.ndim; d++)
      if (kb[d] > kb[dm]) dm = d;
    kb[dm]--;
    tot--;
    clamped[dm] = true;
    sc->grid_clamped = 1;
  }
  for (int d = 0; d < 3; d++) {
    g.kb[d] = kb[d];
    if (d >= p.ndim) {
      g.nc[d] = 1; g.lo[d] = 0.0; g.w[d] = 1.0; g.inv_w[d] = 1.0;
      continue;
    }
    if (clamped[d]) ncd[d] = fmin(ncd[d], (double)(1 << kb[d]));
    g.nc[d] = (int)ncd[d];
    if (g.bound[d] >= 2) {
      g.lo[d] = g.xmin[d];
      g.w[d] = (g.xmax[d] - g.xmin[d]) / g.nc[d];
    } else {
      const double ext = sc->bbox_hi[d] - sc->bbox_lo[d];
      g.lo[d] = sc->bbox_lo[d];
      g.w[d] = !clamped[d] ? cs
               : g.nc[d] > 1 ? fmax(cs, ext / (g.nc[d] - 1.0) * (1.0 + 1.0e-12))
                             : fmax(cs, ext * (1.0 + 1.0e-12));
    }
    g.inv_w[d] = 1.0 / g.w[d];
  }
  g.ncell = 1LL << tot;
  {
    // key layout: the bits every dimension has, then the next ones of the
    // two longest (u < v), then the rest of the longest
    int m1 = kb[0];
    for (int d = 1; d < p.ndim; d++) m1 = min(m1, kb[d]);
    int u = -1, v = -1;
    for (int d = 0; d < p.ndim; d++)
      if (kb[d] > m1) {
        if (u < 0) u = d;
        else v = d;
      }
    if (u < 0) u = 0;
    if (v < 0) v = u;
    g.km[0] = m1;
    g.km[1] = v != u ? min(kb[u], kb[v]) - m1 : 0;
    g.km[2] = u;
    g.km[3] = v;
  }
  int on = p.use_list;
  for (int d = 0; d < p.ndim; d++)
    if (g.bound[d] == 3 && !(rmax * (1.0 + p.skin) * 1.001 < g.xmax[d] - g.xmin[d])) on = 0;
  sc->list_on = on;
  sc->list_overflow = 0;
  *gd = g;
}
__global__ void k_grid_setup(DScal* sc, Grid* gd, Grid box, Phys p, int cap_bits) {
  b_grid_setup(sc, gd, box, p, cap_bits, VL{(int)blockIdx.x, (int)gridDim.x});
}

__device__ __forceinline__ void b_fill_cells(int* __restrict__ a, const Grid* __restrict__ gd, VL vl) {
  const long long m = gd->ncell + 1;
  for (long long i = (long long)vl.b * blockDim.x + threadIdx.x; i < m; i += (long long)vl.nb * blockDim.x)
    a[i] = 0;
}
__global__ void k_fill_cells(int* __restrict__ a, const Grid* __restrict__ gd) {
  b_fill_cells(a, gd, VL{(int)blockIdx.x, (int)gridDim.x});
}

__device__ __forceinline__ void b_fill_i(int* __restrict__ a, int n, int v, VL vl) {
  for (int i = vl.b * blockDim.x + threadIdx.x; i < n; i += vl.nb * blockDim.x) a[i] = v;
}
__global__ void k_fill_i(int* __restrict__ a, int n, int v) {
  b_fill_i(a, n, v, VL{(int)blockIdx.x, (int)gridDim.x});
}

__device__ __forceinline__ int cell_coord(const Grid& g, int d, double x) {
  int c = (int)floor((x - g.lo[d]) * g.inv_w[d]);
  if (c < 0) c = 0;
  if (c > g.nc[d] - 1) c = g.nc[d] - 1;
  return c;
}

// Consecutive lanes of a wave holding the same key (-1: none) form a run; the
// run's first lane adds the run's length to ctr[key] once, and each lane gets
// the old value plus its place in the run.  Slots are in cell order after a
// reorder, so a wave meets few cells: few atomics, and the particles of a
// cell land in slot order within each run.
__device__ __forceinline__ int run_atomic_add(int* __restrict__ ctr, int key) {
  const int lane = threadIdx.x & 63;
  const int prev = __shfl_up(key, 1);
  const bool head = lane == 0 || prev != key;
  const unsigned long long heads = __ballot(head);
  const unsigned long long upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
  const int myhead = 63 - __clzll(heads & upto);
  const unsigned long long after = heads & ~upto;
  const int next = after ? __ffsll((long long)after) - 1 : 64;
  int base = 0;
  if (head && key >= 0) base = atomicAdd(&ctr[key], next - lane);
  return __shfl(base, myhead) + (lane - myhead);
}

__device__ __forceinline__ void b_cellid(KArgs a, const Grid* __restrict__ gp, VL vl) {
  const int i = vl.b * blockDim.x + threadIdx.x;
  int c = -1;
  if (i < a.n) {
    const Grid g = *gp;
    int cx = cell_coord(g, 0, a.f[F_X][i]);
    int cy = g.ndim >= 2 ? cell_coord(g, 1, a.f[F_Y][i]) : 0;
    int cz = g.ndim >= 3 ? cell_coord(g, 2, a.f[F_Z][i]) : 0;
    c = cell_key(g, cx, cy, cz);
    a.cellid[i] = c;
  }
  run_atomic_add(a.cfill, c);
}
__global__ void k_cellid(KArgs a, const Grid* __restrict__ gp) {
  b_cellid(a, gp, VL{(int)blockIdx.x, (int)gridDim.x});
}

// exclusive scan, 3 phases (1024 elements per block)
__device__ __forceinline__ void b_scan_local(const int* in, int* out, int* bsum, const Grid* __restrict__ gp, VL vl) {
  __shared__ int sm[BS];
  const long long n = gp->ncell + 1;
  if ((long long)vl.b * BS * 4 >= n) return;
  const int ITEMS = 4;
  long long base = (long long)vl.b * BS * ITEMS + (long long)threadIdx.x * ITEMS;
  int v[ITEMS];
  int tot = 0;
  for (int k = 0; k < ITEMS; k++) {
    long long idx = base + k;
    v[k] = idx < n ? in[idx] : 0;
    tot += v[k];
  }
  sm[threadIdx.x] = tot;
  __syncthreads();
  for (int s = 1; s < BS; s <<= 1) {
    int x = threadIdx.x >= s ? sm[threadIdx.x - s] : 0;
    __syncthreads();
    sm[threadIdx.x] += x;
    __syncthreads();
  }
  int run = sm[threadIdx.x] - tot;
  for (int k = 0; k < ITEMS; k++) {
    long long idx = base + k;
    if (idx < n) out[idx] = run;
    run += v[k];
  }
  if (threadIdx.x == BS - 1) bsum[vl.b] = sm[BS - 1];
}
__global__ void k_scan_local(const int* in, int* out, int* bsum, const Grid* __restrict__ gp) {
  b_scan_local(in, out, bsum, gp, VL{(int)blockIdx.x, (int)gridDim.x});
}

__device__ __forceinline__ void b_scan_bsum(int* bsum, const Grid* __restrict__ gp, int* total_out, VL vl) {
  __shared__ int sm[BS];
  const int nb = (int)((gp->ncell + 1 + BS * 4 - 1) / (BS * 4));
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nb; base += BS) {
    int idx = base + threadIdx.x;
    int v = idx < nb ? bsum[idx] : 0;
    sm[threadIdx.x] = v;
    __syncthreads();
    for (int s = 1; s < BS; s <<= 1) {
      int x = threadIdx.x >= s ? sm[threadIdx.x - s] : 0;
      __syncthreads();
      sm[threadIdx.x] += x;
      __syncthreads();
    }
    if (idx < nb) bsum[idx] = carry + sm[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == BS - 1) carry += sm[BS - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0 && total_out) *total_out = carry;
}
__global__ void k_scan_bsum(int* bsum, const Grid* __restrict__ gp, int* total_out) {
  b_scan_bsum(bsum, gp, total_out, VL{(int)blockIdx.x, (int)gridDim.x});
}

__device__ __forceinline__ void b_scan_add(int* out, const int* bsum, const Grid* __restrict__ gp, VL vl) {
  const long long n = gp->ncell + 1;
  long long idx = (long long)vl.b * blockDim.x + threadIdx.x;
  if (idx < n) out[idx] += bsum[idx / (BS * 4)];
}
__global__ void k_scan_add(int* out, const int* bsum, const Grid* __restrict__ gp) {
  b_scan_add(out, bsum, gp, VL{(int)blockIdx.x, (int)gridDim.x});
}

__device__ __forceinline__ void b_scatter(KArgs a, VL vl) {
  const int i = vl.b * blockDim.x + threadIdx.x;
  const int c = i < a.n ? a.cellid[i] : -1;
  const int r = run_atomic_add(a.cfill, c);
  if (c >= 0) a.sidx[a.cstart[c] + r] = i;
}
__global__ void k_scatter(KArgs a) {
  b_scatter(a, VL{(int)blockIdx.x, (int)gridDim.x});
}

// ascending order of N values held in registers (bitonic network; every
// index is fixed at compile time, so v stays in registers)
template <int N, typename T>
__device__ __forceinline__ void sort_regs(T (&v)[N]) {
#pragma unroll
  for (int k = 2; k <= N; k <<= 1)
#pragma unroll
    for (int j = k >> 1; j > 0; j >>= 1)
#pragma unroll
      for (int i = 0; i < N; i++) {
        const int l = i ^ j;
        if (l > i) {
          const T x = v[i], y = v[l];
          const bool up = (i & k) == 0;
          v[i] = up ? min(x, y) : max(x, y);
          v[l] = up ? max(x, y) : min(x, y);
        }
      }
}
// cell [b, b+m) of sorted positions into ascending slot order by one lane,
// m <= N: every entry is read before any is written back
template <int N>
__device__ __forceinline__ void sort_cell_lane(int* __restrict__ sidx, int b, int m) {
  int v[N];
#pragma unroll
  for (int q = 0; q < N; q++) v[q] = q < m ? sidx[b + q] : INT_MAX;
  sort_regs<N>(v);
#pragma unroll
  for (int q = 0; q < N; q++)
    if (q < m) sidx[b + q] = v[q];
}

// make the order inside each cell ascending in slot index (deterministic,
// and a stable sort by cell key overall).  A lane sorts its own cell in
// registers when the cell holds at most CS_LANE particles, or at most
// CS_LANE2 while at least CS_LANE2_MIN cells of its wave are of that size
// (the lanes then work side by side); the other cells are taken by the whole
// wave in turn: a cell of at most 256 particles is ranked in registers (each
// slot index counts the smaller ones of its cell), a larger one is
// insertion-sorted by the first lane.
#define CS_LANE 16
#define CS_LANE2 64
#define CS_LANE2_MIN 4
__device__ __forceinline__ void b_cellsort_wave(KArgs a, const Grid* __restrict__ gp, VL vl) {
  const long long c = (long long)vl.b * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  int b = 0, m = 0;
  if (c < gp->ncell) {
    b = a.cstart[c];
    m = a.cstart[c + 1] - b;
  }
  if (m > 1 && m <= CS_LANE) sort_cell_lane<CS_LANE>(a.sidx, b, m);
  const bool mid = m > CS_LANE && m <= CS_LANE2;
  const bool lanes_mid = __popcll(__ballot(mid)) >= CS_LANE2_MIN;
  if (lanes_mid && mid) sort_cell_lane<CS_LANE2>(a.sidx, b, m);
  unsigned long long todo = __ballot(m > CS_LANE && !(lanes_mid && mid));
  while (todo) {
    const int src = __ffsll((long long)todo) - 1;
    todo &= todo - 1ull;
    const int cb = __shfl(b, src), cm = __shfl(m, src);
    if (cm <= 256) {
      // up to four entries per lane (positions lane + 64 q)
      const int v0 = lane < cm ? a.sidx[cb + lane] : INT_MAX;
      const int v1 = lane + 64 < cm ? a.sidx[cb + lane + 64] : INT_MAX;
      const int v2 = lane + 128 < cm ? a.sidx[cb + lane + 128] : INT_MAX;
      const int v3 = lane + 192 < cm ? a.sidx[cb + lane + 192] : INT_MAX;
      int r0 = 0, r1 = 0, r2 = 0, r3 = 0;
      for (int q = 0; q * 64 < cm; q++) {
        const int vq = q == 0 ? v0 : q == 1 ? v1 : q == 2 ? v2 : v3;
        const int nj = min(64, cm - q * 64);
        for (int j = 0; j < nj; j++) {
          const int u = __shfl(vq, j);
          r0 += u < v0;
          r1 += u < v1;
          r2 += u < v2;
          r3 += u < v3;
        }
      }
      if (lane < cm) a.sidx[cb + r0] = v0;
      if (lane + 64 < cm) a.sidx[cb + r1] = v1;
      if (lane + 128 < cm) a.sidx[cb + r2] = v2;
      if (lane + 192 < cm) a.sidx[cb + r3] = v3;
    } else if (lane == 0) {
      for (int k = cb + 1; k < cb + cm; k++) {
        int v = a.sidx[k];
        int j = k - 1;
        while (j >= cb && a.sidx[j] > v) {
          a.sidx[j + 1] = a.sidx[j];
          j--;
        }
        a.sidx[j + 1] = v;
      }
    }
  }
}
__global__ __launch_bounds__(BS) void k_cellsort_wave(KArgs a, const Grid* __restrict__ gp) {
  b_cellsort_wave(a, gp, VL{(int)blockIdx.x, (int)gridDim.x});
}

// ===========================================================================
// neighbour views
// ===========================================================================

// bytes of a density record: the position and 1/h (double4), then one quad
// of reals, two with dust (48 bytes in FP32: records straddle cache lines,
// and the pass runs faster on fewer bytes than on 64-byte aligned records)
template <typename PT>
__host__ __device__ constexpr int dr_bytes(bool dust) {
  return (32 + (dust ? 2 : 1) * 4 * (int)sizeof(PT) + 15) / 16 * 16;
}

// The pack kernels form each particle's record in LDS, then the workgroup
// writes its records (one contiguous run) with consecutive 16-byte stores:
// records written in place by their own lanes would be strided stores.
__device__ __forceinline__ void flush_records(char* __restrict__ dst, const float4* __restrict__ stage, int cnt,
                                              int rb) {
  const int nq = cnt * rb / 16;
  float4* d = (float4*)dst;
  for (int q = threadIdx.x; q < nq; q += blockDim.x) d[q] = stage[q];
}

template <typename PT>
__device__ __forceinline__ void b_pack_density(KArgs a, Phys p, VL vl) {
  alignas(32) __shared__ float4 stage[BS * dr_bytes<PT>(true) / 16];
  const int rb = dr_bytes<PT>(p.idust != 0);
  const int k0 = vl.b * blockDim.x, k = k0 + threadIdx.x;
  if (k < a.n) {
    int i = a.sidx[k];
    double h = a.f[F_H][i];
    const double4 pk = make_double4(a.f[F_X][i], a.f[F_Y][i], a.f[F_Z][i], 1.0 / h);
    a.P4[k] = pk;
    a.dxb[k] = p.radkern * h;
    typedef typename V4<PT>::type T4;
    char* rec = (char*)stage + threadIdx.x * rb;
    *(double4*)rec = pk;
    T4* R = (T4*)(rec + 32);
    R[0] = T4{(PT)a.f[F_VX][i], (PT)a.f[F_VY][i], (PT)a.f[F_VZ][i], (PT)a.f[F_M][i]};
    // the density sums read the dust fraction of the last conservative2primitive
    // (dustfrac, src/density_sums.f90:276-278), not the predicted dustevol
    if (p.idust) R[1] = T4{(PT)a.f[F_DUSTFRAC][i], PT(0), PT(0), PT(0)};
  }
  __syncthreads();
  flush_records((char*)a.DR + (size_t)k0 * rb, stage, min((int)blockDim.x, a.n - k0), rb);
}
template <typename PT>
__global__ void k_pack_density(KArgs a, Phys p) {
  b_pack_density<PT>(a, p, VL{(int)blockIdx.x, (int)gridDim.x});
}

// quads of a force record: the physics variant of the force pass (MHD,
// else dust, else hydro) fixes the layout
__host__ __device__ __forceinline__ int fr_phys(const Phys& p) { return p.imhd ? 1 : p.idust ? 2 : 0; }
__host__ __device__ constexpr int fr_slots(int phys) { return phys == 1 ? 4 : phys == 2 ? 5 : 3; }
// bytes of a record: the position and 1/h (double4), then the quads, padded
// to 32 bytes
template <typename PT>
__host__ __device__ constexpr size_t fr_bytes(int phys) {
  return (32 + (size_t)fr_slots(phys) * 4 * sizeof(PT) + 31) / 32 * 32;
}

// the force record of sorted particle k into rec, and the copies the force
// pass leaves behind
template <typename PT>
__device__ __forceinline__ void pack_force_one(KArgs& a, const Phys& p, int k, char* rec) {
  int i = a.sidx[k];
  double h = a.f[F_H][i];
  double4 pk = a.P4[k];
  pk.w = 1.0 / h;
  a.P4[k].w = pk.w;
  // a fixed particle's h was replaced by its partner's after the iteration
  if (p.fixed_on && p.use_list && a.itype[i] == 1 && h > a.hcap[k]) atomicOr(&a.sc->list_invalid, 1);
  typedef typename V4<PT>::type T4;
  const int phys = fr_phys(p);
  *(double4*)rec = pk;
  T4* R = (T4*)(rec + 32);
  R[0] = T4{(PT)a.f[F_VX][i], (PT)a.f[F_VY][i], (PT)a.f[F_VZ][i], (PT)a.f[F_M][i]};
  R[1] = T4{(PT)a.f[F_RHO][i], (PT)a.f[F_PR][i], (PT)a.f[F_CS][i], (PT)a.f[F_GRADH][i]};
  R[2] = T4{(PT)a.f[F_UU][i], (PT)a.f[F_A0][i], (PT)a.f[F_A1][i], (PT)a.f[F_A2][i]};
  if (phys == 1) R[3] = T4{(PT)a.f[F_BX][i], (PT)a.f[F_BY][i], (PT)a.f[F_BZ][i], (PT)a.f[F_PSI][i]};
  if (p.imhd) {
    // Bfield of conservative2primitive (src/conservative2primitive.f90:152),
    // kept for the outputs after the corrector has moved B on
    for (int d = 0; d < 3; d++) a.f[F_BFX + d][i] = a.f[F_BX + d][i];
  }
  if (phys == 2) {
    R[3] = T4{(PT)h, (PT)a.f[F_DENS][i], (PT)a.f[F_DUSTFRAC][i], (PT)a.f[F_RHOGAS][i]};
    R[4] = T4{(PT)a.f[F_RHODUST][i], PT(0), PT(0), PT(0)};
  }
  if (p.fixed_on) {
    const bool fx = a.itype[i] == 1;
    ((T4*)a.GF)[k] = T4{(PT)a.f[fx ? F_MIMG : F_M][i], (PT)a.f[fx ? F_A0IMG : F_A0][i],
                        (PT)a.f[fx ? F_A1IMG : F_A1][i], (PT)a.f[fx ? F_A2IMG : F_A2][i]};
  }
}
template <typename PT>
__device__ __forceinline__ void b_pack_force(KArgs a, Phys p, VL vl) {
  alignas(32) __shared__ float4 stage[BS * fr_bytes<PT>(2) / 16];
  const int phys = fr_phys(p);
  const int rb = (int)fr_bytes<PT>(phys);
  const int k0 = vl.b * blockDim.x, k = k0 + threadIdx.x;
  if (k < a.n) pack_force_one<PT>(a, p, k, (char*)stage + threadIdx.x * rb);
  __syncthreads();
  flush_records((char*)a.FR + (size_t)k0 * rb, stage, min((int)blockDim.x, a.n - k0), rb);
}

template <typename PT>
__global__ void k_pack_force(KArgs a, Phys p) {
  b_pack_force<PT>(a, p, VL{(int)blockIdx.x, (int)gridDim.x});
}

// ===========================================================================
// neighbour iteration over image cells
// ===========================================================================

// Calls body(kk, xj_img[3], flip[3], code) for every particle of every cell
// image that can hold a point within rad of xi; xj_img holds the image
// coordinates (ghost arithmetic of src/ghostND_mhd.f90:180-200), flip[d] = 1
// where the image's velocity component is reversed (image_flip), code = the image
// kinds kx + 5 ky + 25 kz (0 for a particle seen in place).
template <int NDIM, typename F>
__device__ __forceinline__ void for_each_candidate(const KArgs& a, const Grid& g, const double* xi,
                                                   double rad, F&& body) {
  // rows (cy, cz) whose cells lie wholly outside the sphere are skipped and
  // each row is cut to the chord of the sphere; cell faces are widened by a
  // relative margin so that a particle binned by rounding into a cell is
  // never culled (the FP64 pair test downstream is what decides membership)
  const double rad2 = rad * rad * (1.0 + 1.0e-12);
  int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  for (int d = 1; d < NDIM; d++) cell_range(g, d, xi[d], rad, lo[d], hi[d]);
  for (int cz = lo[2]; cz <= hi[2]; cz++) {
    int kz = 0, sz = 0;
    double gz2 = 0.0;
    if (NDIM >= 3) {
      sz = cell_image(g, 2, cz, kz);
      if (sz < 0) continue;
      const double gz = cell_gap(g, 2, cz, xi[2]);
      gz2 = gz * gz;
      if (gz2 > rad2) continue;
    }
    for (int cy = lo[1]; cy <= hi[1]; cy++) {
      int ky = 0, sy = 0;
      double r2 = rad2 - gz2;
      if (NDIM >= 2) {
        sy = cell_image(g, 1, cy, ky);
        if (sy < 0) continue;
        const double gy = cell_gap(g, 1, cy, xi[1]);
        r2 -= gy * gy;
        if (r2 < 0.0) continue;
      }
      int lx, hx;
      cell_range(g, 0, xi[0], NDIM >= 2 ? sqrt(r2) : rad, lx, hx);
      for (int cx = lx; cx <= hx; cx++) {
        int kx = 0;
        const int sx = cell_image(g, 0, cx, kx);
        if (sx < 0) continue;
        const int kind[3] = {kx, ky, kz};
        const int code = kx + 5 * ky + 25 * kz;
        int fl[3];
        image_flip(kind, fl);
        const int key = cell_key(g, sx, sy, sz);
        const int b = a.cstart[key], e = a.cstart[key + 1];
        for (int kk = b; kk < e; kk++) {
          const double4 pj = a.P4[kk];
          double xj[3] = {pj.x, pj.y, pj.z};
          bool ok = true;
          if (code) {
            for (int d = 0; d < NDIM; d++)
              if (kind[d]) ok = ok && image_coord(g, d, kind[d], xj[d], a.dxb[kk], xj[d]);
          }
          if (ok) body(kk, xj, fl, code);
        }
      }
    }
  }
}

// Image code of a list entry: 2 bits per dimension, kind 0 (in place),
// 1 (periodic image left of xmin) or 2 (right of xmax); lists exist only
// without reflecting walls, so no other kind occurs.
__device__ __forceinline__ unsigned list_code(int kx, int ky, int kz) {
  return (unsigned)(kx | (ky << 2) | (kz << 4));
}
// the code of k seen from j, given the code of j seen from k (1 <-> 2)
__device__ __forceinline__ unsigned inverse_code(unsigned code) {
  return ((code & 0x15u) << 1) | ((code & 0x2Au) >> 1);
}

// Iterates the first cnt entries of the neighbour list of sorted particle k,
// calling body(kk, r, xj_img, flip, code) with r = load(kk) (a struct holding
// the FP64 position and 1/h in r.p and whatever the pair sum reads) and the
// image coordinates of the ghost arithmetic (so the exact FP64 tests see the
// values of the cell scan).  Entries come four per load; the record of entry
// m + PF (PF 1 or 2) is fetched while entry m is processed, so a wavefront's
// gathers overlap its arithmetic (the pair passes use one ahead: a second
// record in flight costs them a wave per SIMD).  The loop is unrolled over
// the PF + 1 record buffers, which take turns as the current one, so no
// record is copied.
template <int NDIM, int PF = 1, typename LOAD, typename F>
__device__ __forceinline__ void for_each_listed_pf(const KArgs& a, const Grid& g, int k, int cnt, LOAD&& load,
                                                   F&& body) {
  static_assert(PF == 1 || PF == 2, "prefetch distance 1 or 2");
  if (cnt <= 0) return;
  const uint4* L4 = (const uint4*)(a.nbl + (size_t)k * a.maxnb);
  const int ng = (cnt + 3) >> 2;
  // entries are read four at a time into an eight-entry window (c0..c3 the
  // current ones, n0..n3 the next group) that shifts by one per entry, so
  // that no entry is picked by a run-time index (an indexed uint4 would live
  // in scratch memory)
  const uint4 g0 = L4[0];
  const uint4 g1 = ng > 1 ? L4[1] : make_uint4(0u, 0u, 0u, 0u);
  unsigned c0 = g0.x, c1 = g0.y, c2 = g0.z, c3 = g0.w, n0 = g1.x, n1 = g1.y, n2 = g1.z, n3 = g1.w;
  int m = 0;
  // entry m, its record in cur; the record of entry m + PF goes into nxt
  auto visit = [&](const auto& cur, auto& nxt) {
    const unsigned e = c0;
    const unsigned en = m + PF < cnt ? (PF == 2 ? c2 : c1) : e;
    nxt = load((int)(en >> 7));
    const int kk = (int)(e >> 7);
    const int code = (int)(e & 63u);
    double xj[3] = {cur.p.x, cur.p.y, cur.p.z};
    const int flip[3] = {0, 0, 0};
    bool ok = true;
    if (code) {
      // ghost arithmetic of src/ghostND_mhd.f90:180-200; a periodic image
      // of a particle exactly on the wall does not exist (:172)
      for (int d = 0; d < NDIM; d++) {
        const int kd = (code >> (2 * d)) & 3;
        const double x0 = xj[d];
        if (kd == 1) {
          ok = ok && g.xmax[d] - x0 > 0.0;
          xj[d] = g.xmin[d] + (x0 - g.xmax[d]);
        } else if (kd == 2) {
          ok = ok && x0 - g.xmin[d] > 0.0;
          xj[d] = g.xmax[d] + (x0 - g.xmin[d]);
        }
      }
    }
    if (ok) body(kk, cur, xj, flip, code);
    c0 = c1; c1 = c2; c2 = c3; c3 = n0; n0 = n1; n1 = n2; n2 = n3;
    if ((m & 3) == 3) {
      const int gi2 = (m >> 2) + 2;
      const uint4 t = gi2 < ng ? L4[gi2] : make_uint4(0u, 0u, 0u, 0u);
      n0 = t.x; n1 = t.y; n2 = t.z; n3 = t.w;
    }
    m++;
  };
  using R = decltype(load(0));
  R rA = load((int)(c0 >> 7));
  if constexpr (PF == 1) {
    R rB;
    for (;;) {
      if (m >= cnt) break;
      visit(rA, rB);
      if (m >= cnt) break;
      visit(rB, rA);
    }
  } else {
    R rB = load((int)((cnt > 1 ? c1 : c0) >> 7));
    R rC;
    for (;;) {
      if (m >= cnt) break;
      visit(rA, rC);
      if (m >= cnt) break;
      visit(rB, rA);
      if (m >= cnt) break;
      visit(rC, rB);
    }
  }
}

// cell-scan counterpart of for_each_listed_pf (record loaded in place)
template <int NDIM, typename LOAD, typename F>
__device__ __forceinline__ void for_each_candidate_rec(const KArgs& a, const Grid& g, const double* xi,
                                                       double rad, LOAD&& load, F&& body) {
  for_each_candidate<NDIM>(a, g, xi, rad, [&](int kk, const double* xj, const int* flip, int code) {
    const auto r = load(kk);
    body(kk, r, xj, flip, code);
  });
}

// Neighbour lists (periodic or open domains).  Entry m of sorted particle k
// sits at nbl[k*maxnb + m] = (sorted index << 7) | image code.  Entries
// [0, nbc) are the gather list G(k): every image of a particle j with
// r_kj < radkern*h_k*(1+skin), in the candidate order of the cell scan; then
// nbx transposed extras: every image of j whose opposite image of k is in
// G(j) while it is not in G(k) itself.  Together they hold every pair (and
// pair of periodic images) with r < radkern*max(h_k, h_j) while no h has
// grown past hcap = h*(1+skin).  Membership uses the separation
// fl(x_k - x_j) + s L per dimension, the shift s = +1, -1 or 0 given by the
// image code, summed without contraction: it is exactly antisymmetric in
// FP64 (fl(a - b) = -fl(b - a), and the shifts of two opposite images are
// exact negatives), so "k in G(j)" is decided identically from k's and from
// j's side and no pair is lost or doubled.  Lists need
// radkern*hmax*(1+skin) < L along periodic dimensions (checked on the device
// at every binning), so every image within a list radius lies in the image
// cells one period away; where 2*radkern*hmax*(1+skin) >= L a list can hold
// two images of one neighbour, or an image of the particle itself, as the
// reference's ghost copies interact (src/ghostND_mhd.f90:180-200).
template <int NDIM>
__device__ __forceinline__ double image_r2(const Grid& g, const double* xa, const double* xb, unsigned code) {
#pragma clang fp contract(off)
  double r2 = 0.0;
  for (int d = 0; d < NDIM; d++) {
    double dx = xa[d] - xb[d];
    const unsigned kd = (code >> (2 * d)) & 3u;
    // kind 1: the image left of xmin (at x_j - L); kind 2: right of xmax
    // (selected without a branch; adding 0 leaves dx as it is)
    const double L = g.xmax[d] - g.xmin[d];
    dx = dx + (kd == 1u ? L : kd == 2u ? -L : 0.0);
    r2 = r2 + dx * dx;
  }
  return r2;
}

// streaming stores (each output written once per kernel): non-temporal, so
// that they do not evict the neighbour records the gathers reuse from L2
__device__ __forceinline__ void st_stream(double* p, double v) { __builtin_nontemporal_store(v, p); }

// a value every lane of the wave holds, moved to scalar registers (frees
// vector registers in the register-bound list builder)
__device__ __forceinline__ double wave_uniform(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readfirstlane((int)(unsigned)(b & 0xffffffffll));
  const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ int wave_min_i(int v) {
  for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ int wave_max_i(int v) {
  for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o));
  return v;
}

// One wavefront per 64 consecutive sorted particles (a compact group: the
// order is by Morton cell key).  The wave enumerates the cells of the union
// of its particles' gather boxes (rows (cy, cz) cut to the union of the
// particles' chords), in the nested order of for_each_candidate, so each
// list keeps the scan's candidate order.  For a batch of up to NBL_CELLS
// cells it loads all cell ranges at once, forms their prefix sums in LDS,
// and then streams the batch's candidates 64 at a time through LDS (the next
// 64 are fetched while the current ones are tested): FP64 raw positions for
// the exact test and FP32 image positions relative to the wave's first
// particle for a screening test.  A decision of the FP32 test taken outside
// a narrow band around the radius (its width bounded by the FP32 rounding,
// see below) is the FP64 test's decision; a candidate inside the band takes
// the exact FP64 image test.  A pair k-j with j in G(k) and k not in G(j) is
// appended to the extras of j, which fill j's row from its end
// (k_nblist_sort moves them into place).
#define NBL_CELLS 128
template <int NDIM>
__device__ __forceinline__ void b_nblist(KArgs a, Phys p, const Grid* __restrict__ gp, VL vl) {
  __shared__ double4 sraw[128];
  __shared__ float4 srel[128];
  __shared__ unsigned se[128];
  __shared__ int ckey[NBL_CELLS];
  __shared__ int ccode[NBL_CELLS];
  __shared__ int cbeg[NBL_CELLS];
  __shared__ int cpre[NBL_CELLS + 1];
  const int lane = threadIdx.x;
  const int k = xcd_tile(vl) * 64 + lane;
  const bool act = k < a.n;
  if (!a.sc->list_on) {
    if (act) a.nbc[k] = -1;
    return;
  }
  const Grid g = *gp;
  const double s1 = 1.0 + p.skin;
  const double lim2 = p.radkern2 * s1 * s1 * (1.0 + 1.0e-12);
  const double4 pk = a.P4[act ? k : 0];
  const double xi[3] = {pk.x, pk.y, pk.z};
  const double hi21 = pk.w * pk.w;
  const double rad = act ? p.radkern * s1 * (1.0 + 1.0e-9) / pk.w : 0.0;
  const double rad2 = rad * rad * (1.0 + 1.0e-12);
  const double org[3] = {wave_uniform(__shfl(xi[0], 0)), wave_uniform(__shfl(xi[1], 0)),
                          wave_uniform(__shfl(xi[2], 0))};
  const float xr0 = (float)(xi[0] - org[0]);
  const float xr1 = NDIM >= 2 ? (float)(xi[1] - org[1]) : 0.f;
  const float xr2 = NDIM >= 3 ? (float)(xi[2] - org[2]) : 0.f;
  unsigned* L = a.nbl + (size_t)(act ? k : 0) * a.maxnb;
  int cnt = 0;
  bool over = false;
  // box of the group's particles and the largest search radius
  // (candidates are kept within that radius of the box itself, which drops
  // the corners of the widened box: about a fifth of them)
  double gbox_lo[3], gbox_hi[3], rw2;
  {
    double rw = rad;
    for (int o = 32; o > 0; o >>= 1) rw = fmax(rw, __shfl_xor(rw, o));
    rw *= 1.0 + 1.0e-9;
    rw2 = wave_uniform(rw * rw);
    for (int d = 0; d < 3; d++) {
      double lo_ = act ? xi[d] : DBL_MAX, hi_ = act ? xi[d] : -DBL_MAX;
      for (int o = 32; o > 0; o >>= 1) {
        lo_ = fmin(lo_, __shfl_xor(lo_, o));
        hi_ = fmax(hi_, __shfl_xor(hi_, o));
      }
      gbox_lo[d] = wave_uniform(lo_);
      gbox_hi[d] = wave_uniform(hi_);
    }
  }
  int fill = 0;  // staged candidates awaiting the test (wave-uniform)

  // FP32 screening.  Positions are rounded to FP32 relative to the origin
  // org; every staged candidate (and particle of the group) lies within Mg
  // of it (the group box's farthest corner plus the largest radius).  With
  // u = 2^-24, the FP32 squared separation misses r^2 by at most
  // 5u r^2 + 7u Mg r, i.e. near a radius R by a relative 5u + 7u Mg/R; the
  // band eps(R) = 4e-6 (4 + Mg/R) is more than twice that, thresholds'
  // rounding included, so a decision outside the band is the FP64 one.
  float Mg = 0.0f;
  {
    double m2 = 0.0;
    for (int d = 0; d < NDIM; d++) {
      const double md = fmax(org[d] - gbox_lo[d], gbox_hi[d] - org[d]);
      m2 += md * md;
    }
    Mg = (float)((sqrt(m2) + sqrt(rw2)) * (1.0 + 1.0e-6));
  }
  const double rl2 = act ? lim2 / hi21 : 0.0;  // squared list radius of this particle
  const float eps_k = act ? 4.0e-6f * (4.0f + Mg / (float)sqrt(rl2)) : 0.0f;
  const float rl2_out = (float)rl2 * (1.0f + eps_k), rl2_in = (float)rl2 * (1.0f - eps_k);
  const float lim2_f = (float)lim2;
  // lim2 * eps(R) for a candidate's radius R = sqrt(lim2) / w: band0 + band1 w
  const float band0 = lim2_f * 1.6e-5f, band1 = lim2_f * 4.0e-6f * Mg / (float)sqrt(lim2);
  // test the staged candidates [0, m) against this lane's particle (every
  // lane of the wave takes part; an idle lane has no candidates)
  auto test = [&](int m) {
    unsigned long long mask = 0ull;
    if (act) {
#pragma unroll 1
      for (int t0 = 0; t0 < m; t0 += 8) {
        // 8 candidates with a constant bit each (entries past m are masked)
        unsigned m8 = 0u;
#pragma unroll
        for (int u = 0; u < 8; u++) {
          const float4 c = srel[t0 + u];
          const float dx = xr0 - c.x, dy = xr1 - c.y, dz = xr2 - c.z;
          const float r2f = dx * dx + dy * dy + dz * dz;
          m8 |= r2f < rl2_out ? (1u << u) : 0u;
        }
        mask |= (unsigned long long)m8 << t0;
      }
      if (m < 64) mask &= (1ull << m) - 1ull;
    }
    // an extra's slot is reserved when it is found and written when the next
    // one is found (or after the loop), so the atomic's latency overlaps the
    // tests in between
    int pend_row = -1, pend_slot = 0;
    unsigned pend_ent = 0u;
    while (mask) {
      const int t = __builtin_ctzll(mask);
      mask &= mask - 1ull;
      const unsigned e = se[t];
      const float4 cr = srel[t];
      const float dx = xr0 - cr.x, dy = xr1 - cr.y, dz = xr2 - cr.z;
      const float r2f = dx * dx + dy * dy + dz * dz;
      double rr = -1.0;
      if (!(r2f < rl2_in)) {
        const double4 c = sraw[t];
        const double xr[3] = {c.x, c.y, c.z};
        rr = image_r2<NDIM>(g, xi, xr, e & 63u);
        if (!(rr * hi21 < lim2)) continue;
      }
      if (cnt < a.maxnb) L[cnt] = e;
      else over = true;
      cnt++;
      const int kk = (int)(e >> 7);
      if (kk != k) {
        // q - lim2 against the band lim2 * eps(R_kk), R_kk = rlim / w_kk
        const float dq = r2f * (cr.w * cr.w) - lim2_f;
        const float band = band0 + band1 * cr.w;
        int kin = dq < -band ? 1 : dq > band ? 0 : -1;
        if (kin < 0) {
          const double4 c = sraw[t];
          if (rr < 0.0) {
            const double xr[3] = {c.x, c.y, c.z};
            rr = image_r2<NDIM>(g, xi, xr, e & 63u);
          }
          kin = rr * (c.w * c.w) < lim2 ? 1 : 0;
        }
        if (!kin) {
          if (pend_row >= 0 && pend_slot < a.maxnb)
            a.nbl[(size_t)pend_row * a.maxnb + (a.maxnb - 1 - pend_slot)] = pend_ent;
          pend_slot = atomicAdd(&a.nbx[kk], 1);
          pend_row = kk;
          pend_ent = ((unsigned)k << 7) | inverse_code(e & 63u);
        }
      }
    }
    if (pend_row >= 0 && pend_slot < a.maxnb)
      a.nbl[(size_t)pend_row * a.maxnb + (a.maxnb - 1 - pend_slot)] = pend_ent;
  };

  int ncl = 0;  // cells in the current batch (wave-uniform)
  auto process_batch = [&]() {
    __syncthreads();
    for (int c = lane; c < ncl; c += 64) {
      const int key = ckey[c];
      const int b = key >= 0 ? a.cstart[key] : 0;
      const int e = key >= 0 ? a.cstart[key + 1] : 0;
      cbeg[c] = b;
      cpre[c] = e - b;
    }
    __syncthreads();
    int run = 0;
    for (int c0 = 0; c0 < ncl; c0 += 64) {
      const int v = c0 + lane < ncl ? cpre[c0 + lane] : 0;
      int x = v;
      for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o);
        if (lane >= o) x += y;
      }
      if (c0 + lane < ncl) cpre[c0 + lane] = run + x - v;
      run += __shfl(x, 63);
    }
    if (lane == 0) cpre[ncl] = run;
    __syncthreads();
    const int T = run;
    // candidate q of the batch -> (raw record, list entry)
    auto fetch = [&](int q, double4& pj, unsigned& ent) {
      int lo = 0, hi = ncl - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (cpre[mid] <= q) lo = mid;
        else hi = mid - 1;
      }
      const int kk = cbeg[lo] + (q - cpre[lo]);
      ent = ((unsigned)kk << 7) | (unsigned)ccode[lo];
      pj = a.P4[kk];
    };
    double4 pj = make_double4(0.0, 0.0, 0.0, 0.0);
    unsigned ent = 0u;
    if (lane < T) fetch(lane, pj, ent);
    for (int q0 = 0; q0 < T; q0 += 64) {
      // keep the candidates inside the group's box widened by the largest
      // radius, compacted in order behind the ones already staged
      const int code = (int)(ent & 63u);
      double xj[3] = {pj.x, pj.y, pj.z};
      for (int d = 0; d < NDIM; d++) {
        const int kd = (code >> (2 * d)) & 3;
        if (kd) image_coord(g, d, kd, xj[d], 0.0, xj[d]);
      }
      bool keep = lane < T - q0;
      double gd2 = 0.0;
      for (int d = 0; d < NDIM; d++) {
        const double e = fmax(fmax(gbox_lo[d] - xj[d], xj[d] - gbox_hi[d]), 0.0);
        gd2 += e * e;
      }
      keep = keep && gd2 <= rw2;
      const unsigned long long bal = __ballot(keep);
      if (keep) {
        const int pos = fill + __popcll(bal & ((1ull << lane) - 1ull));
        sraw[pos] = pj;
        srel[pos] = make_float4((float)(xj[0] - org[0]), NDIM >= 2 ? (float)(xj[1] - org[1]) : 0.f,
                                NDIM >= 3 ? (float)(xj[2] - org[2]) : 0.f, (float)pj.w);
        se[pos] = ent;
      }
      fill += __popcll(bal);
      if (q0 + 64 + lane < T) fetch(q0 + 64 + lane, pj, ent);  // next 64, in flight during the test
      if (fill >= 64) {
        __syncthreads();
        test(64);
        __syncthreads();
        if (lane < fill - 64) {
          sraw[lane] = sraw[64 + lane];
          srel[lane] = srel[64 + lane];
          se[lane] = se[64 + lane];
        }
        fill -= 64;
        __syncthreads();
      }
    }
    ncl = 0;
  };

  int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  for (int d = 1; d < NDIM; d++) {
    int l = INT_MAX, h = INT_MIN;
    if (act) cell_range(g, d, xi[d], rad, l, h);
    lo[d] = wave_min_i(l);
    hi[d] = wave_max_i(h);
  }
  for (int cz = lo[2]; cz <= hi[2]; cz++) {
    int kz = 0, sz = 0;
    double gz2 = 0.0;
    if (NDIM >= 3) {
      sz = cell_image(g, 2, cz, kz);
      if (sz < 0) continue;
      const double gz = cell_gap(g, 2, cz, xi[2]);
      gz2 = gz * gz;
    }
    for (int cy = lo[1]; cy <= hi[1]; cy++) {
      int ky = 0, sy = 0;
      double r2 = rad2 - gz2;
      if (NDIM >= 2) {
        sy = cell_image(g, 1, cy, ky);
        if (sy < 0) continue;
        const double gy = cell_gap(g, 1, cy, xi[1]);
        r2 -= gy * gy;
      }
      int lx = INT_MAX, hx = INT_MIN;
      if (act && r2 >= 0.0) cell_range(g, 0, xi[0], NDIM >= 2 ? sqrt(r2) : rad, lx, hx);
      lx = wave_min_i(lx);
      hx = wave_max_i(hx);
      // the row's key bits (y and z parts of cell_key)
      const unsigned kyz = NDIM == 1 ? 0u : key_part(g, 1, sy) | (NDIM == 3 ? key_part(g, 2, sz) : 0u);
      for (int cx0 = lx; cx0 <= hx; cx0 += 64) {
        const int nadd = min(64, hx - cx0 + 1);
        if (ncl + nadd > NBL_CELLS) process_batch();
        if (lane < nadd) {
          int kx = 0;
          const int sx = cell_image(g, 0, cx0 + lane, kx);
          ckey[ncl + lane] = sx >= 0 ? (int)(key_part(g, 0, sx) | kyz) : -1;
          ccode[ncl + lane] = (int)list_code(kx, ky, kz);
        }
        ncl += nadd;
      }
    }
  }
  if (ncl) process_batch();
  if (fill) {
    __syncthreads();
    test(fill);
  }
  if (act) {
    a.nbc[k] = over ? -1 : cnt;
    a.hcap[k] = s1 / pk.w;
  }
}
template <int NDIM>
// four waves per SIMD (a few registers spill; the LDS of a wave fits 4.5):
// the builder waits on its candidate fetches, and 3 -> 4 waves took a
// 128^3 build from 2.31 to 2.04 ms
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void k_nblist(KArgs a, Phys p,
                                                                                        const Grid* __restrict__ gp) {
  b_nblist<NDIM>(a, p, gp, VL{(int)blockIdx.x, (int)gridDim.x});
}

// nx <= N entries E[0..nx) sorted ascending into D[0..nx) (D may overlap E:
// every entry is read before any is written)
template <int N>
__device__ __forceinline__ void sort_extras(const unsigned* E, unsigned* D, int nx) {
  unsigned v[N];
#pragma unroll
  for (int m = 0; m < N; m++) v[m] = m < nx ? E[m] : 0xffffffffu;
  sort_regs<N>(v);
#pragma unroll
  for (int m = 0; m < N; m++)
    if (m < nx) D[m] = v[m];
}

// extras (written from the end of the row) into [nbc, nbc + nbx) in
// ascending entry order, so sums are deterministic; a row that did not hold
// the gather list and the extras falls back to the cell scan
__device__ __forceinline__ void b_nblist_sort(KArgs a, VL vl) {
  const int k = vl.b * blockDim.x + threadIdx.x;
  if (k >= a.n || !a.sc->list_on) return;
  const int c0 = a.nbc[k], nx = a.nbx[k];
  if (c0 < 0 || c0 + nx > a.maxnb) {
    if (c0 >= 0) a.nbc[k] = -1;
    atomicAdd(&a.sc->list_overflow, 1);
    return;
  }
  // extras occupy E = L[maxnb-nx .. maxnb-1]; they end at L[c0 ..] ascending
  unsigned* E = a.nbl + (size_t)k * a.maxnb + (a.maxnb - nx);
  unsigned* D = a.nbl + (size_t)k * a.maxnb + c0;
  if (nx <= 32) {
    // up to 32 of them sorted in registers (a bitonic network of 8, 16 or
    // 32, the smallest that holds every such row of the wave; every index a
    // constant), then stored in place
    const bool w16 = __ballot(nx > 16) != 0, w8 = __ballot(nx > 8) != 0;
    if (w16) sort_extras<32>(E, D, nx);
    else if (w8) sort_extras<16>(E, D, nx);
    else sort_extras<8>(E, D, nx);
    return;
  }
  // more: insertion sort in place, then move down (ascending copy,
  // destination not above source)
  for (int m = 1; m < nx; m++) {
    const unsigned v = E[m];
    int q = m - 1;
    while (q >= 0 && E[q] > v) {
      E[q + 1] = E[q];
      q--;
    }
    E[q + 1] = v;
  }
  if (D != E)
    for (int m = 0; m < nx; m++) D[m] = E[m];
}
__global__ __launch_bounds__(BS) void k_nblist_sort(KArgs a) {
  b_nblist_sort(a, VL{(int)blockIdx.x, (int)gridDim.x});
}

// ===========================================================================
// global reductions without contended atomics: every block reduces its
// values (wave shuffles, then LDS across waves) and writes one partial per
// quantity; k_red_final folds the partials into the scalar record.
// ===========================================================================

__device__ __forceinline__ double red_combine(int kind, double a, double b) {
  return kind == RED_SUM ? a + b : kind == RED_MAX ? fmax(a, b) : fmin(a, b);
}

// v[q] of every thread of the block -> red[q * rcap + (virtual block)]
// (kinds are compile-time per call site; every thread must call this)
template <int NQ>
__device__ __forceinline__ void block_partials(double (&v)[NQ], const int (&kind)[NQ], double* red,
                                               int rcap, VL vl) {
  __shared__ double sred[NQ][BS / 64];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    double x = v[q];
    for (int o = 32; o > 0; o >>= 1) x = red_combine(kind[q], x, __shfl_xor(x, o));
    if (lane == 0) sred[q][wv] = x;
  }
  __syncthreads();
  if (threadIdx.x < NQ) {
    const int q = threadIdx.x;
    double x = sred[q][0];
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) x = red_combine(kind[q], x, sred[q][w]);
    red[(size_t)q * rcap + vl.b] = x;
  }
}

__device__ __forceinline__ void b_red_final(const double* __restrict__ red, int rcap, int nblocks,
                                                  RedSpec spec, DScal* sc, const int* skip, VL vl) {
  if (skip && *skip) return;
  // every quantity at once: each thread folds its blocks for all of them
  // (eight blocks a step, so that up to 8 x nq loads are in flight), then a
  // butterfly per wave and the waves in order (every fold here is a min, a
  // max or a sum of integers: exact in any order)
  __shared__ double sw[BS / 64][RED_MAXQ];
  const int nq = spec.nq;
  int kind[RED_MAXQ];
  double acc[RED_MAXQ];
#pragma unroll
  for (int q = 0; q < RED_MAXQ; q++) {
    kind[q] = q < nq ? (spec.op[q] & 15) : RED_SUM;
    acc[q] = kind[q] == RED_SUM ? 0.0 : kind[q] == RED_MAX ? -DBL_MAX : DBL_MAX;
  }
  int b = threadIdx.x;
  for (; b + 7 * BS < nblocks; b += 8 * BS) {
#pragma unroll
    for (int q = 0; q < RED_MAXQ; q++) {
      if (q >= nq) continue;
      const double* __restrict__ rq = red + (size_t)q * rcap + b;
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; u++) v[u] = rq[u * BS];
#pragma unroll
      for (int u = 0; u < 8; u++) acc[q] = red_combine(kind[q], acc[q], v[u]);
    }
  }
  for (; b < nblocks; b += BS) {
#pragma unroll
    for (int q = 0; q < RED_MAXQ; q++)
      if (q < nq) acc[q] = red_combine(kind[q], acc[q], red[(size_t)q * rcap + b]);
  }
#pragma unroll
  for (int q = 0; q < RED_MAXQ; q++) {
    double x = acc[q];
    for (int o = 32; o > 0; o >>= 1) x = red_combine(kind[q], x, __shfl_xor(x, o, 64));
    acc[q] = x;
  }
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int q = 0; q < RED_MAXQ; q++) sw[threadIdx.x >> 6][q] = acc[q];
  }
  __syncthreads();
  if ((int)threadIdx.x < nq) {
    const int q = threadIdx.x;
    const int kq = spec.op[q] & 15, type = spec.op[q] >> 4;
    double r = sw[0][q];
    for (int w = 1; w < BS / 64; w++) r = red_combine(kq, r, sw[w][q]);
    char* base = (char*)sc + spec.off[q];
    if (type == RED_F64) {
      double* d = (double*)base;
      *d = red_combine(kq, *d, r);
    } else if (type == RED_I32) {
      int* d = (int*)base;
      *d = (int)red_combine(kq, (double)*d, r);
    } else {
      unsigned long long* d = (unsigned long long*)base;
      *d = kq == RED_SUM ? *d + (unsigned long long)r : (unsigned long long)red_combine(kq, (double)*d, r);
    }
  }
}
__global__ __launch_bounds__(BS) void k_red_final(const double* __restrict__ red, int rcap, int nblocks,
                                                  RedSpec spec, DScal* sc, const int* skip) {
  b_red_final(red, rcap, nblocks, spec, sc, skip, VL{(int)blockIdx.x, (int)gridDim.x});
}

static RedSpec red_spec(std::initializer_list<std::pair<int, int>> q) {
  RedSpec s{};
  for (auto& e : q) {
    s.op[s.nq] = e.first;
    s.off[s.nq] = e.second;
    s.nq++;
  }
  return s;
}
#define RED_Q(kind, type, field) std::make_pair((kind) | ((type) << 4), (int)offsetof(DScal, field))

// ===========================================================================
// density: h-rho Newton-Raphson, src/iterate_density.f90:43-360
//
// One launch runs pass `its0` for the listed particles (all when list is
// null).  sym = 1 is a pass of the symmetric routine density
// (src/density_sums.f90:38), which the reference uses as long as every
// particle is still iterating (ncalc = npart, iterate_density.f90:131-135):
// the neighbour count is over the symmetric set and needs each neighbour's
// current h (P4.w), and psi is left alone unless MHD is off.  sym = 0 is
// density_partial (src/density_sums.f90:396): gather count, psi zeroed when
// MHD is on (:630-635).  With persistent = 1 each particle keeps iterating
// (partial passes) until it converges; otherwise an unconverged particle is
// appended to next_list for the next launch.  Sums depend only on the
// particle's own h, so the per-particle sequence of h is the reference's.
// ===========================================================================


// Newton-Raphson step for h of one particle from its sums at the current h
// (src/iterate_density.f90:162-278; FP64, reference operation order)
struct NRStep {
  double hnew, gradh, ggh, dhdrho;
  bool converged, relink;
};

template <int NDIM>
__device__ __forceinline__ void nr_update(const Phys& p, double hh, double hin, double mi, double rho_v,
                                          double gradh_sum, double ggh_sum, int nn, int its, NRStep& r) {
#pragma clang fp contract(off)
  const double t = hh / p.hfact;
  const double rhoi = mi / (NDIM == 1 ? t : NDIM == 2 ? t * t : t * t * t);
  const double dhdrhoi = -hh / (NDIM * rho_v);
  const double dwdhsumi = gradh_sum;
  double omegai = 1.0 - dhdrhoi * gradh_sum;
  if (omegai < 1.e-5 && fabs(omegai) == 0.0) omegai = 1.0;
  r.gradh = 1.0 / omegai;
  const double func = rhoi - rho_v;
  const double dfdh = omegai / dhdrhoi;
  const double rn = rho_v * NDIM;
  const double d2hdrho2i = hh * (NDIM + 1) / (rn * rn);
  r.ggh = rho_v * (d2hdrho2i * dwdhsumi + dhdrhoi * dhdrhoi * ggh_sum);
  double hnew = hh - func / dfdh;
  if (hnew > 1.2 * hh) hnew = 1.2 * hh;
  else if (hnew < 0.8 * hh) hnew = 0.8 * hh;
  if (hnew <= 0.0 || r.gradh <= DBL_MIN) hnew = p.hfact * pow(mi / rho_v, p.dndim);
  else if (its > 100) hnew = p.hfact * pow(mi / rho_v, p.dndim);
  r.relink = false;
  if (nn <= 1) { hnew = hh + p.psep; r.relink = true; }
  r.converged = (fabs((hnew - hh) / hin) < p.tolh && omegai > 0.0) || p.maxits == 0;
  r.hnew = hnew;
  r.dhdrho = dhdrhoi;
}

template <int NDIM, typename PT, bool DUST, int TB>
__device__ __forceinline__ void b_nr_pass(KArgs a, Phys p, const Grid* __restrict__ gp, int* __restrict__ list_a,
                                                int* __restrict__ list_b, VL vl) {
  // pass parameters from the device-side control (k_pass_ctl): pass its0
  // reads the active list written by the previous pass (list_a after odd
  // passes, list_b after even ones) and appends its unconverged particles to
  // the other one
  if (a.sc->pc_done) return;
  const int nlist = a.sc->pc_nlist, its0 = a.sc->pc_its, sym = a.sc->pc_sym;
  const int persistent = a.sc->pc_persistent;
  const int* __restrict__ list = its0 == 1 ? nullptr : ((its0 & 1) ? list_b : list_a);
  int* __restrict__ next_list = (its0 & 1) ? list_a : list_b;
  // a workgroup with no particle of this pass skips loading the tables (a
  // later pass holds only the particles still iterating); it still writes
  // its neutral reduction partials, which the final fold reads
  const int tile = xcd_tile_n(vl, (nlist + (int)blockDim.x - 1) / (int)blockDim.x);
  const bool idle = tile * (int)blockDim.x >= nlist;
  __shared__ PT tw[NDSP_IKERN + 1];
  __shared__ PT tg[NDSP_IKERN + 1];
  __shared__ PT tgg[NDSP_IKERN + 1];
  if (!idle) {
    const PT* src = sizeof(PT) == 4 ? (const PT*)a.tabf : (const PT*)a.tabd;
    for (int t = threadIdx.x; t <= NDSP_IKERN; t += blockDim.x) {
      tw[t] = src[t];
      tg[t] = src[NDSP_IKERN + 1 + t];
      tgg[t] = src[2 * (NDSP_IKERN + 1) + t];
    }
  }
  __syncthreads();
  const Grid g = *gp;
  const int tid = tile * (int)blockDim.x + (int)threadIdx.x;
  const bool active = tid < nlist;
  bool unconverged = false;
  const int k = active ? (list ? list[tid] : tid) : 0;
  typedef typename V4<PT>::type T4;
  constexpr int RB = dr_bytes<PT>(DUST);
  const char* DR = (const char*)a.DR;
  // this thread's share of the global reductions
  double r_stress = 0.0, r_hmax = 0.0, r_its = 0.0, r_nmin = DBL_MAX, r_nmax = 0.0, r_pairs = 0.0, r_ncalc = 0.0;
  // fixed particles are not iterated (src/iterate_density.f90:166,253) and
  // take their partner's values after the iteration; the first symmetric pass
  // still counts their neighbours (src/density_sums.f90:166-169)
  const bool fixedp = active && p.fixed_on && a.itype[a.sidx[k]] == 1;
  if (active && (!fixedp || (its0 == 1 && sym))) {
    const int i = a.sidx[k];
    const double4 pk = a.P4[k];
    const double xi[3] = {pk.x, pk.y, pk.z};
    double h = a.f[F_H][i];
    double hin;
    if (its0 == 1) { hin = h; a.f[F_HNR0][i] = h; }
    else hin = a.f[F_HNR0][i];
    const double mi = a.f[F_M][i];
    const T4 dk = *(const T4*)(DR + (size_t)k * RB + 32);
    const PT vi[3] = {dk.x, dk.y, dk.z};
    const PT ddq2 = (PT)p.ddq2table, dq2 = (PT)p.dq2table;
    const double R2 = p.radkern2;
    const double hmaxc = a.sc->hmax_cur;
    constexpr bool dust = DUST;
    struct DRec {
      double4 p;
      T4 d;
      PT e;
    };
    auto loadj = [&](int kk) {
      DRec r;
      const char* rec = DR + (size_t)kk * RB;
      r.p = *(const double4*)rec;
      const T4* q = (const T4*)(rec + 32);
      r.d = q[0];
      r.e = dust ? q[1].x : PT(0);
      return r;
    };

    int its = its0 - 1, nn = 0;
    unsigned long long npairs = 0;
    double rho_v = 0, drhodt_v = 0, gradh_v = 0, ggh_v = 0, rhog_v = 0, rhod_v = 0, dhdt_v = 0;
    int err = 0;
    bool done = false;
    while (true) {
      its++;
      const bool symp = sym && its == its0;
      if (!symp && p.imhd != 0) a.f[F_PSI][i] = 0.0;
      if (symp && p.imhd == 0 && !fixedp) a.f[F_PSI][i] = 0.0;
      const double hi1 = 1.0 / h;
      const double hi21 = mul_exact(hi1, hi1);
      const PT hi1p = (PT)hi1;
      const PT hfacwabi = pown_t<NDIM>(hi1p);
      const PT hfacgrkerni = hfacwabi * hi1p;
      const PT hi1s = hi1p * hi1p;
      Acc<PT> rho, drhodt, gradh, ggh, rhog, rhod;
      rho.zero(); drhodt.zero(); gradh.zero(); ggh.zero(); rhog.zero(); rhod.zero();
      int nsym = 0, ngat = 0;
      const double rad = p.radkern * (symp ? fmax(h, hmaxc) : h);
      if (!(rad < p.boxmin)) { err = NDSP_ERR_CELLS; done = true; break; }
      // the list serves the gather while h stays within its margin, and the
      // symmetric count of the first pass (neighbours still at their build h)
      const bool uselist = p.use_list && a.nbc[k] >= 0 && h <= a.hcap[k] && (!symp || its == 1);
      const int nlisted = uselist ? (symp ? a.nbc[k] + a.nbx[k] : a.nbc[k]) : 0;
      auto body = [&](int kk, const DRec& rj, const double* xj, const int* flip, int code) {
        double dx[3] = {0.0, 0.0, 0.0};
        for (int d = 0; d < NDIM; d++) dx[d] = xi[d] - xj[d];
        const double rij2 = rij2_exact<NDIM>(dx);
        const double q2i = mul_exact(rij2, hi21);
        if (symp) {
          const double hj1 = rj.p.w;
          const double q2j = mul_exact(mul_exact(rij2, hj1), hj1);
          if (q2i < R2 || q2j < R2) nsym++;
        }
        if (!(q2i < R2)) return;
        ngat++;
        const PT rij = p_sqrt((PT)rij2);
        const PT q2 = (PT)q2i;
        const PT wabi = ktab(tw, q2, ddq2, dq2) * hfacwabi;
        const PT grkerni = ktab(tg, q2, ddq2, dq2) * hfacgrkerni;
        const PT grgrkerni = ktab(tgg, q2, ddq2, dq2) * hfacwabi * hi1s;
        const PT dwdhi = -(rij * grkerni * hi1p) - PT(NDIM) * wabi * hi1p;
        const PT dwdhdhi = PT(NDIM * (NDIM + 1)) * wabi * hi1s + PT(2 * (NDIM + 1)) * rij * hi1s * grkerni +
                           rij * rij * hi1s * grgrkerni;
        const T4 dj = rj.d;
        const PT mj = dj.w;
        rho.add(mj * wabi);
        gradh.add(mj * dwdhi);
        ggh.add(mj * dwdhdhi);
        if (dust) {
          const PT epsj = rj.e;
          rhod.add(mj * epsj * wabi);
          rhog.add(mj * (PT(1) - epsj) * wabi);
        }
        if (kk != k || code != 0) {
          const PT rinv = p_rcp(rij + (PT)DBL_EPSILON);
          const PT vj[3] = {flip[0] ? -dj.x : dj.x, flip[1] ? -dj.y : dj.y, flip[2] ? -dj.z : dj.z};
          PT dvdotr = PT(0);
          for (int d = 0; d < NDIM; d++) dvdotr += (vi[d] - vj[d]) * ((PT)dx[d] * rinv);
          drhodt.add(mj * dvdotr * grkerni);
        }
      };
      if (uselist) for_each_listed_pf<NDIM, 1>(a, g, k, nlisted, loadj, body);
      else for_each_candidate_rec<NDIM>(a, g, xi, rad, loadj, body);
      npairs += (unsigned long long)ngat;
      nn = symp ? nsym : ngat;
      if (fixedp) {
        a.numneigh[i] = nn;
        r_nmin = nn;
        r_nmax = nn;
        break;
      }
      // Newton-Raphson update, src/iterate_density.f90:162-278 (FP64)
      rho_v = rho.value();
      drhodt_v = drhodt.value();
      if (dust) { rhog_v = rhog.value(); rhod_v = rhod.value(); }
      if (rho_v <= 0.0) { err = NDSP_ERR_RHO_NONPOSITIVE; done = true; break; }
      NRStep r;
      nr_update<NDIM>(p, h, hin, mi, rho_v, gradh.value(), ggh.value(), nn, its, r);
      gradh_v = r.gradh;
      ggh_v = r.ggh;
      const double hnew = r.hnew;
      bool relink = r.relink;
      if (r.converged) {
        drhodt_v = mul_exact(drhodt_v, gradh_v);
        dhdt_v = mul_exact(r.dhdrho, drhodt_v);
        done = true;
        break;
      }
      if (its > p.maxits) { err = NDSP_ERR_NOT_CONVERGED; done = true; break; }
      if (hnew > a.sc->hhmax_ghost) relink = true;
      if (relink) atomicOr(&a.sc->relink, 1);
      h = hnew;
      if (!persistent) {
        a.f[F_H][i] = h;
        unconverged = true;
        break;
      }
    }
    if (done) {
      // ---- results, then conservative->primitive and EOS for this particle ----
      a.f[F_H][i] = h;
      st_stream(&a.f[F_RHO][i], rho_v);
      st_stream(&a.f[F_GRADH][i], gradh_v);
      st_stream(&a.f[F_GRADGRADH][i], ggh_v);
      st_stream(&a.f[F_DRHODT][i], drhodt_v);
      st_stream(&a.f[F_DHDT][i], dhdt_v);
      a.numneigh[i] = nn;
      a.nits[i] = its;
      double dens;
      if (dust) {
        double eps = a.f[F_EPS][i];
        if (eps > 1.0) { eps = 1.0; a.f[F_EPS][i] = 1.0; a.f[F_EPSIN][i] = 1.0; }
        a.f[F_DUSTFRAC][i] = eps;
        a.f[F_RHOGAS][i] = rhog_v;
        a.f[F_RHODUST][i] = rhod_v;
        dens = rho_v * (1.0 - eps);
      } else {
        dens = rho_v;
      }
      a.f[F_DENS][i] = dens;
      double uu = a.f[F_EN][i];
      double pr = 0.0, cs = 0.0;
      const double gamma1 = p.gamma - 1.0;
      if (p.iener == 0) {
        if (dens > 0.0) {
          pr = p.polyk * pow(dens, p.gamma);
          cs = sqrt(p.gamma * pr / dens);
          if (fabs(gamma1) > 1.e-3) uu = pr / (gamma1 * dens);
        }
      } else {
        if (uu < 0.0) err = err ? err : NDSP_ERR_EOS;
        if (dens > 0.0) {
          pr = gamma1 * uu * dens;
          cs = sqrt(p.gamma * pr / dens);
        }
      }
      a.f[F_UU][i] = uu;
      a.f[F_PR][i] = pr;
      a.f[F_CS][i] = cs;
      if (p.imhd != 0) {
        const double bx = a.f[F_BX][i], by = a.f[F_BY][i], bz = a.f[F_BZ][i];
        const double B2 = bx * bx + by * by + bz * bz;
        r_stress = fmax(0.5 * B2 - pr, 0.0);
      }
      r_hmax = h;
      if (p.use_list && h > a.hcap[k]) atomicOr(&a.sc->list_invalid, 1);
      r_its = its;
      r_nmin = nn;
      r_nmax = nn;
    }
    r_pairs = (double)npairs;
    r_ncalc = its - its0 + 1;
    if (err) atomicCAS(&a.sc->err, 0, err);
  }
  {
    double rv[7] = {r_stress, r_hmax, r_its, r_nmin, r_nmax, r_pairs, r_ncalc};
    const int rk[7] = {RED_MAX, RED_MAX, RED_MAX, RED_MIN, RED_MAX, RED_SUM, RED_SUM};
    block_partials<7>(rv, rk, a.red, a.rcap, vl);
  }
  // unconverged particles form the next pass's list: flagged here and
  // compacted in sorted order afterwards (k_cmp_*), so that the next pass's
  // workgroups keep the spatial order, or appended one atomic per wave
  if (!persistent && a.nrflag) {
    if (active) a.nrflag[k] = unconverged ? 1 : 0;
  } else if (!persistent) {
    const unsigned long long m = __ballot(unconverged);
    const int lane = threadIdx.x & 63;
    const int cnt = __popcll(m);
    int base = 0;
    if (lane == 0 && cnt) base = atomicAdd(&a.sc->ncalc_next, cnt);
    base = __shfl(base, 0);
    if (unconverged) next_list[base + __popcll(m & ((1ull << lane) - 1ull))] = k;
  }
}
// three waves per SIMD (an occupancy attribute): with one record in flight
// the FP32 pass fits in 168 registers, spilling two outside the pair loop
template <int NDIM, typename PT, bool DUST, int TB>
__global__ __launch_bounds__(TB) __attribute__((amdgpu_waves_per_eu(3))) void k_nr_pass(
    KArgs a, Phys p, const Grid* __restrict__ gp, int* __restrict__ list_a, int* __restrict__ list_b) {
  b_nr_pass<NDIM, PT, DUST, TB>(a, p, gp, list_a, list_b, VL{(int)blockIdx.x, (int)gridDim.x});
}

// P4.w = 1/h of every particle before a symmetric pass after the first, and
// the reflecting ghost extents after a relink (as decided by k_pass_ctl)
__device__ __forceinline__ void b_refresh_hinv(KArgs a, Phys p, VL vl) {
  if (a.sc->pc_done || !a.sc->pc_refresh) return;
  const int refresh_dxb = a.sc->pc_relink;
  int k = vl.b * blockDim.x + threadIdx.x;
  if (k >= a.n) return;
  const double h = a.f[F_H][a.sidx[k]];
  a.P4[k].w = 1.0 / h;
  ((double4*)((char*)a.DR + (size_t)k * a.dr_stride))->w = 1.0 / h;
  if (refresh_dxb) a.dxb[k] = p.radkern * h;
}
__global__ void k_refresh_hinv(KArgs a, Phys p) {
  b_refresh_hinv(a, p, VL{(int)blockIdx.x, (int)gridDim.x});
}

// max h over all particles -> hmax_cur (and the ghost extent on a relink)
__device__ __forceinline__ void b_hmax(KArgs a, DScal* sc, VL vl) {
  if (sc->pc_done) return;
  const int relink = sc->pc_relink;
  __shared__ double sm[BS];
  double hm = 0.0;
  for (int i = vl.b * blockDim.x + threadIdx.x; i < a.n; i += vl.nb * blockDim.x)
    hm = fmax(hm, a.f[F_H][i]);
  sm[threadIdx.x] = hm;
  __syncthreads();
  for (int s = BS / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    atomic_max_pos(&sc->hmax_cur, sm[0]);
    if (relink) atomic_max_pos(&sc->hhmax_ghost, sm[0]);
  }
}
__global__ void k_hmax(KArgs a, DScal* sc) {
  b_hmax(a, sc, VL{(int)blockIdx.x, (int)gridDim.x});
}

// After a pass (src/iterate_density.f90:280-300): finished when the pass was
// the persistent one, failed, or left no particle unconverged; otherwise the
// unconverged ones form the next pass, symmetric if they are all particles
// (:131-135), synchronous while walls reflect (a relink, :260-262, rebuilds
// the ghosts with every particle's current h).
__device__ __forceinline__ void b_pass_ctl(DScal* sc, int n, int reflect, int maxits, VL vl) {
  if (threadIdx.x | vl.b) return;
  if (sc->pc_done) return;
  if (sc->pc_persistent || sc->err || sc->ncalc_next == 0 || sc->pc_its > maxits) {
    sc->pc_done = 1;
    return;
  }
  const int relink = reflect && sc->relink;
  const int nl = sc->ncalc_next;
  sc->pc_nlist = nl;
  sc->pc_its = sc->pc_its + 1;
  sc->pc_sym = nl == n;
  sc->pc_persistent = !reflect && nl != n;
  sc->pc_relink = relink;
  sc->pc_refresh = nl == n || relink;
  sc->ncalc_next = 0;
  sc->relink = 0;
  sc->hmax_cur = 0.0;
  if (relink) sc->hhmax_ghost = 0.0;
}
__global__ void k_pass_ctl(DScal* sc, int n, int reflect, int maxits) {
  b_pass_ctl(sc, n, reflect, maxits, VL{(int)blockIdx.x, (int)gridDim.x});
}

// Ordered compaction of the flags a pass left (sorted positions of the
// unconverged particles) into the next pass's list, in three launches over
// blocks of CMP_ITEMS*BS flags: per-block counts, their exclusive scan (and
// the total, the next pass's size), and the scatter, which also clears the
// flags it reads.
#define CMP_ITEMS 4
__global__ __launch_bounds__(BS) void k_cmp_count(const int* __restrict__ flag, int n, int* __restrict__ bsum) {
  __shared__ int ws[BS / 64];
  const long long base = (long long)blockIdx.x * BS * CMP_ITEMS;
  int c = 0;
  for (int q = 0; q < CMP_ITEMS; q++) {
    const long long i = base + (long long)q * BS + threadIdx.x;
    if (i < n) c += flag[i];
  }
  for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o);
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    int t = 0;
    for (int w = 0; w < BS / 64; w++) t += ws[w];
    bsum[blockIdx.x] = t;
  }
}
__global__ __launch_bounds__(BS) void k_cmp_scan(int* __restrict__ bsum, int nb, DScal* sc) {
  __shared__ int sm[BS];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < nb; base += BS) {
    const int idx = base + threadIdx.x;
    const int v = idx < nb ? bsum[idx] : 0;
    sm[threadIdx.x] = v;
    __syncthreads();
    for (int s = 1; s < BS; s <<= 1) {
      const int x = (int)threadIdx.x >= s ? sm[threadIdx.x - s] : 0;
      __syncthreads();
      sm[threadIdx.x] += x;
      __syncthreads();
    }
    if (idx < nb) bsum[idx] = carry + sm[threadIdx.x] - v;
    __syncthreads();
    if (threadIdx.x == BS - 1) carry += sm[BS - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) sc->ncalc_next = carry;
}
__global__ __launch_bounds__(BS) void k_cmp_scatter(int* __restrict__ flag, int n, const int* __restrict__ bsum,
                                                    const DScal* __restrict__ sc, int* __restrict__ list_a,
                                                    int* __restrict__ list_b) {
  __shared__ int ws[BS / 64];
  // the list the pass that set the flags names as its next (b_nr_pass)
  int* __restrict__ out = (sc->pc_its & 1) ? list_a : list_b;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const long long base = (long long)blockIdx.x * BS * CMP_ITEMS;
  int run = bsum[blockIdx.x];
  for (int q = 0; q < CMP_ITEMS; q++) {
    const long long i = base + (long long)q * BS + threadIdx.x;
    const int f = i < n ? flag[i] : 0;
    const unsigned long long m = __ballot(f != 0);
    if (lane == 0) ws[w] = __popcll(m);
    __syncthreads();
    int off = run, tot = 0;
    for (int u = 0; u < BS / 64; u++) {
      if (u < w) off += ws[u];
      tot += ws[u];
    }
    if (f) {
      out[off + __popcll(m & ((1ull << lane) - 1ull))] = (int)i;
      flag[i] = 0;
    }
    run += tot;
    __syncthreads();
  }
}


// ===========================================================================
// force: fused pair sums of get_rates
// ===========================================================================

// PHYS: 0 hydro, 1 MHD (imhd 11), 2 one-fluid dust (idust 4); TB threads
// per workgroup (chosen by the autotuner)
template <int NDIM, typename PT, int PHYS, int TB>
__device__ __forceinline__ void b_force(KArgs a, Phys p, const Grid* __restrict__ gp, VL vl) {
  __shared__ PT tg[NDSP_IKERN + 1];
  {
    const PT* src = sizeof(PT) == 4 ? (const PT*)a.tabf : (const PT*)a.tabd;
    for (int t = threadIdx.x; t <= NDSP_IKERN; t += blockDim.x) tg[t] = src[NDSP_IKERN + 1 + t];
  }
  __syncthreads();
  const Grid g = *gp;
  const int k = xcd_tile(vl) * blockDim.x + threadIdx.x;
  const bool active = k < a.n;
  typedef typename V4<PT>::type T4;
  constexpr bool mhd = PHYS == 1;
  constexpr bool dust = PHYS == 2;
  constexpr size_t RB = fr_bytes<PT>(PHYS);
  const char* FR = (const char*)a.FR;
  const T4* GF = (const T4*)a.GF;
  const PT ddq2 = (PT)p.ddq2table, dq2 = (PT)p.dq2table;
  const double R2 = p.radkern2;
  const PT beta = (PT)p.beta;

  PT vsigmax_l = PT(0);
  double dtc_l = 1.0e6;
  // FP32 pair math: the largest vsig_dtc / min(h_i, h_j) = vsig_dtc max(1/h_i, 1/h_j)
  // over the pairs, one division at the end instead of two reciprocals a pair
  PT rdt_l = PT(0);
  unsigned long long npairs = 0;
  int err = 0;
  if (active) {
    const int i = a.sidx[k];
    const char* rec = FR + (size_t)k * RB;
    const double4 pk = *(const double4*)rec;
    const double xi[3] = {pk.x, pk.y, pk.z};
    const double hi1 = pk.w;
    const double hi21 = mul_exact(hi1, hi1);
    const T4* ri = (const T4*)(rec + 32);
    const T4 g0 = ri[0], g1 = ri[1], g2 = ri[2];
    const T4 g4 = dust ? ri[3] : T4{PT(0), PT(0), PT(0), PT(0)};
    const PT vi[3] = {g0.x, g0.y, g0.z};
    const PT mi = g0.w;
    const PT rhoi = g1.x, rho1i = p_rcp(rhoi), rho21i = rho1i * rho1i;
    const PT pri = fmax(g1.y, PT(0));
    const PT Prho2i = pri * rho21i;
    const PT csi = g1.z, gradhi = g1.w;
    const PT uui = g2.x, alphai = g2.y, alphaui = g2.z, alphaBi = g2.w;
    // h_i enters only the FP64 time-step bound
    const PT hi = dust ? g4.x : sizeof(PT) == 8 ? (PT)a.f[F_H][i] : PT(0);
    const PT hi1p = (PT)hi1;
    const PT hfacgrkerni = pown_t<NDIM>(hi1p) * hi1p;
    PT Bi[3] = {0, 0, 0}, psii = 0;
    if (mhd) { const T4 g3 = ri[3]; Bi[0] = g3.x; Bi[1] = g3.y; Bi[2] = g3.z; psii = g3.w; }
    const PT Brhoi[3] = {Bi[0] * rho1i, Bi[1] * rho1i, Bi[2] * rho1i};
    const PT B2i = Bi[0] * Bi[0] + Bi[1] * Bi[1] + Bi[2] * Bi[2];
    const PT Brho2i = B2i * rho21i;
    const PT valfven2i = B2i * rho1i;
    const PT epsi = dust ? g4.z : PT(0);
    const PT rhogasi = dust ? g4.w : rhoi;
    const PT rhodusti = dust ? ri[4].x : PT(0);
    const PT stressmax = (PT)a.sc->stressmax;
    const PT tstop = (PT)p.Kdrag;
    const double rad = p.radkern * a.sc->hmax_nr;
    if (!(rad < p.boxmin)) err = NDSP_ERR_CELLS;

    Acc<PT> fx, fy, fz, dudt, dbx, dby, dbz, divb, cbx, cby, cbz, gpx, gpy, gpz, del2u, deps, gdx, gdy, gdz;
    fx.zero(); fy.zero(); fz.zero(); dudt.zero(); dbx.zero(); dby.zero(); dbz.zero();
    divb.zero(); cbx.zero(); cby.zero(); cbz.zero(); gpx.zero(); gpy.zero(); gpz.zero();
    del2u.zero(); deps.zero(); gdx.zero(); gdy.zero(); gdz.zero();

    const bool uselist = p.use_list && a.nbc[k] >= 0 && !a.sc->list_invalid;
    struct FRec {
      double4 p;
      T4 g0, g1, g2, g3, g4;
      PT g5;
    };
    auto loadj = [&](int kk) {
      FRec r;
      const char* rec = FR + (size_t)kk * RB;
      r.p = *(const double4*)rec;
      const T4* rj = (const T4*)(rec + 32);
      r.g0 = rj[0];
      r.g1 = rj[1];
      r.g2 = rj[2];
      // (h, EOS density) of a gas without dust are 1/P4.w and rho: no load
      if (dust) r.g4 = rj[3];
      else r.g4 = T4{PT(0), PT(0), PT(0), PT(0)};
      if (mhd) r.g3 = rj[3];
      else r.g3 = T4{PT(0), PT(0), PT(0), PT(0)};
      r.g5 = dust ? rj[4].x : PT(0);
      return r;
    };
    auto body = [&](int kk, const FRec& rj, const double* xj, const int* flip, int code) {
      double dx[3] = {0.0, 0.0, 0.0};
      for (int d = 0; d < NDIM; d++) dx[d] = xi[d] - xj[d];
      const double rij2 = rij2_exact<NDIM>(dx);
      const double q2i = mul_exact(rij2, hi21);
      const double hj1 = rj.p.w;
      const double q2j = mul_exact(mul_exact(rij2, hj1), hj1);
      if (!(q2i < R2 || q2j < R2)) return;
      if (kk == k && code == 0) return;
      npairs++;
      if (rij2 <= 0.0) { err = NDSP_ERR_COINCIDENT; return; }
      const PT rij = p_sqrt((PT)rij2);
      const PT rinv = p_rcp(rij);
      PT dr[3] = {0, 0, 0};
      for (int d = 0; d < NDIM; d++) dr[d] = (PT)dx[d] * rinv;
      const T4 j1 = rj.g1, j4 = rj.g4;
      T4 j0 = rj.g0, j2 = rj.g2;
      if (p.fixed_on && code != 0) {
        // the ghosts of src/ghostND_mhd.f90:363-420 are copies made before
        // conservative2primitive hands a fixed particle its partner's mass
        // and alpha (src/conservative2primitive.f90:425-432), and keep them
        const T4 gf = GF[kk];
        j0.w = gf.x;
        j2.y = gf.y;
        j2.z = gf.z;
        j2.w = gf.w;
      }
      const PT hj1p = (PT)hj1;
      const PT hfacgrkernj = pown_t<NDIM>(hj1p) * hj1p;
      const PT grkerni = ktab(tg, (PT)q2i, ddq2, dq2) * hfacgrkerni * gradhi;
      const PT grkernj = ktab(tg, (PT)q2j, ddq2, dq2) * hfacgrkernj * j1.w;
      const PT grkern = PT(0.5) * (grkerni + grkernj);
      const PT vj[3] = {flip[0] ? -j0.x : j0.x, flip[1] ? -j0.y : j0.y, flip[2] ? -j0.z : j0.z};
      const PT mj = j0.w;
      const PT dvel[3] = {vi[0] - vj[0], vi[1] - vj[1], vi[2] - vj[2]};
      const PT dvdotr = dvel[0] * dr[0] + dvel[1] * dr[1] + dvel[2] * dr[2];
      const PT rhoj = j1.x, rho1j = p_rcp(rhoj), rho21j = rho1j * rho1j;
      const PT rhoav1 = PT(0.5) * (rho1i + rho1j);
      const PT prj = fmax(j1.y, PT(0));
      const PT Prho2j = prj * rho21j;
      const PT csj = j1.z;
      PT Bj[3] = {0, 0, 0}, psij = 0, dB[3] = {0, 0, 0};
      PT projBi = 0, projBj = 0, projdB = 0, projBrhoi = 0, projBrhoj = 0, Brho2j = 0, valfven2j = 0;
      PT Brhoj[3] = {0, 0, 0};
      if (mhd) {
        const T4 j3 = rj.g3;
        Bj[0] = j3.x; Bj[1] = j3.y; Bj[2] = j3.z; psij = j3.w;
        for (int c = 0; c < 3; c++) { Brhoj[c] = Bj[c] * rho1j; dB[c] = Bi[c] - Bj[c]; }
        projBi = Bi[0] * dr[0] + Bi[1] * dr[1] + Bi[2] * dr[2];
        projBj = Bj[0] * dr[0] + Bj[1] * dr[1] + Bj[2] * dr[2];
        projdB = dB[0] * dr[0] + dB[1] * dr[1] + dB[2] * dr[2];
        projBrhoi = Brhoi[0] * dr[0] + Brhoi[1] * dr[1] + Brhoi[2] * dr[2];
        projBrhoj = Brhoj[0] * dr[0] + Brhoj[1] * dr[1] + Brhoj[2] * dr[2];
        const PT B2j = Bj[0] * Bj[0] + Bj[1] * Bj[1] + Bj[2] * Bj[2];
        // B^2 / rho_j: the reciprocal formed above serves in FP32
        valfven2j = dust ? p_div(B2j, j4.y) : sizeof(PT) == 4 ? B2j * rho1j : p_div(B2j, rhoj);
        Brho2j = B2j * rho21j;
      }
      // signal velocities, src/ratesND_mhd.f90:1417-1482
      PT vsigi, vsigj, vsigB;
      if (mhd) {
        const PT vsig2i = csi * csi + valfven2i;
        const PT vsig2j = csj * csj + valfven2j;
        const PT ti = csi * projBi, tj = csj * projBj;
        const PT vsigproji = vsig2i * vsig2i - PT(4) * (ti * ti) * rho1i;
        const PT vsigprojj = vsig2j * vsig2j - PT(4) * (tj * tj) * rho1j;
        vsigi = p_sqrt(PT(0.5) * (vsig2i + p_sqrt(fmax(vsigproji, PT(0)))));
        vsigj = p_sqrt(PT(0.5) * (vsig2j + p_sqrt(fmax(vsigprojj, PT(0)))));
        if (vsigproji < PT(-1.e-4) * vsig2i * vsig2i || vsigprojj < PT(-1.e-4) * vsig2j * vsig2j)
          err = NDSP_ERR_VSIG;
        if (p.iavlim[2] != 2) vsigB = p_sqrt(dvel[0] * dvel[0] + dvel[1] * dvel[1] + dvel[2] * dvel[2]);
        else vsigB = PT(0.5) * (vsigi + vsigj) + fabs(dvdotr);
      } else {
        vsigi = csi;
        vsigj = csj;
        vsigB = PT(0);
      }
      const PT vsig = PT(0.5) * fmax(vsigi + vsigj - beta * dvdotr, PT(0));
      const PT vsigu = p_sqrt(fabs(pri - prj) * rhoav1);
      const PT vsigdtc = fmax(fmax(vsig, PT(0.5) * (vsigi + vsigj + beta * fabs(dvdotr))), vsigB);
      vsigmax_l = fmax(vsigmax_l, vsigdtc);
      if (vsigdtc > PT(1.e-10)) {
        if constexpr (sizeof(PT) == 4) {
          rdt_l = fmax(rdt_l, vsigdtc * fmax(hi1p, hj1p));
        } else {
          const PT hj = dust ? j4.x : p_rcp(hj1p);
          dtc_l = fmin(dtc_l, (double)p_div((PT)fmin(hi, hj), vsigdtc));
        }
      }
      PT f[3] = {0, 0, 0};
      PT du = PT(0);
      // dB/dt terms of the pair (dissipation and induction), summed in the
      // pair precision before the accumulation
      PT dbp[3] = {0, 0, 0};
      // artificial dissipation
      if (dust) {
        // src/ratesND_mhd.f90:2154-2206
        if (dvdotr < PT(0)) {
          const PT vsi = alphai * csi - beta * dvdotr;
          const PT vsj = j2.y * csj - beta * dvdotr;
          const PT qi = -rhogasi * vsi * dvdotr;
          const PT qj = -(dust ? j4.w : rhoj) * vsj * dvdotr;
          const PT visc = PT(0.5) * (qi * rho21i * grkerni + qj * rho21j * grkernj);
          for (int c = 0; c < 3; c++) f[c] -= mj * visc * dr[c];
          const PT duu = uui - j2.x;
          const PT cfaci = PT(0.5) * alphaui * rhoi * vsigu * duu;
          const PT cfacj = PT(0.5) * j2.z * rhoj * vsigu * duu;
          const PT diffu = cfaci * grkerni * rho21i + cfacj * grkernj * rho21j;
          du += PT(0.5) * qi * p_div(rho1i, rhogasi) * mj * dvdotr * grkerni + p_div(mj * diffu, PT(1) - epsi);
        }
      } else {
        // src/ratesND_mhd.f90:1700-1894
        const PT alphaav = PT(0.5) * (alphai + j2.y);
        const PT alphau = PT(0.5) * (alphaui + j2.z);
        const PT alphaB = PT(0.5) * (alphaBi + j2.w);
        const PT term = vsig * rhoav1 * grkern;
        const PT termu = vsigu * rhoav1 * grkern;
        const PT termB = vsigB * rhoav1 * grkern;
        if (dvdotr < PT(0)) {
          const PT visc = alphaav * term * (-dvdotr);
          for (int c = 0; c < 3; c++) f[c] -= mj * visc * dr[c];
        }
        if (mhd) {
          const PT fac = rhoi * mj * alphaB * termB * rhoav1;
          for (int c = 0; c < 3; c++) dbp[c] = fac * dB[c];
        }
        if (p.iener > 0) {
          PT vissv = PT(0);
          if (dvdotr < PT(0)) {
            const PT t = (vi[0] * dr[0] + vi[1] * dr[1] + vi[2] * dr[2]) - (vj[0] * dr[0] + vj[1] * dr[1] + vj[2] * dr[2]);
            vissv = -alphaav * PT(0.5) * (t * t);
          }
          const PT vissu = alphau * (uui - j2.x);
          const PT vissB = mhd ? -alphaB * PT(0.5) * (dB[0] * dB[0] + dB[1] * dB[1] + dB[2] * dB[2]) * rhoav1 : PT(0);
          du += mj * (term * vissv + termu * vissu + termB * vissB);
        }
      }
      // pressure, src/ratesND_mhd.f90:1538-1568
      {
        const PT prterm = Prho2i * grkerni + Prho2j * grkernj;
        for (int c = 0; c < 3; c++) f[c] -= mj * prterm * dr[c];
      }
      if (mhd) {
        // mhd_terms, src/ratesND_mhd.f90:2377-2720 (tensor force, imhd 11)
        const PT fiso = PT(0.5) * (Brho2i * grkerni + Brho2j * grkernj);
        for (int c = 0; c < 3; c++) {
          const PT faniso = (Brhoi[c] * projBrhoi - stressmax * dr[c] * rho21i) * grkerni +
                            (Brhoj[c] * projBrhoj - stressmax * dr[c] * rho21j) * grkernj;
          f[c] += mj * (faniso - fiso * dr[c]);
        }
        divb.add(-mj * projdB * grkern);
        const PT cb[3] = {dB[1] * dr[2] - dB[2] * dr[1], dB[2] * dr[0] - dB[0] * dr[2], dB[0] * dr[1] - dB[1] * dr[0]};
        cbx.add(mj * cb[0] * grkern);
        cby.add(mj * cb[1] * grkern);
        cbz.add(mj * cb[2] * grkern);
        const PT ind = mj * projBrhoi * grkerni;
        for (int c = 0; c < 3; c++) dbp[c] -= ind * dvel[c];
        dbx.add(dbp[0]);
        dby.add(dbp[1]);
        dbz.add(dbp[2]);
        if (p.idivbzero >= 2) {
          const PT gpt = mj * (psii * rho21i * grkerni + psij * rho21j * grkernj);
          gpx.add(-gpt * dr[0]);
          gpy.add(-gpt * dr[1]);
          gpz.add(-gpt * dr[2]);
        }
      }
      if (dust) {
        // dust_derivs_diffusion, idustevol 0, src/ratesND_mhd.f90:2813-2878
        const PT Di = rhodusti * rho1i * tstop;
        const PT Dj = rj.g5 * rho1j * tstop;
        const PT Dav = PT(0.5) * (Di + Dj);
        const PT grgrkern = PT(-2) * grkern * rinv;
        const PT diffterm = rho1i * rho1j * Dav * (pri - prj) * grgrkern;
        deps.add(mj * diffterm);
        if (p.iener > 0) {
          const PT duu = uui - j2.x;
          du += pri * p_div(rho1i, rhogasi) * mj * dvdotr * grkerni - p_div(PT(0.5) * mj * duu * diffterm, PT(1) - epsi);
        }
      }
      fx.add(f[0]);
      fy.add(f[1]);
      fz.add(f[2]);
      dudt.add(du);
      if (p.iavlim[1] > 0) del2u.add(mj * rho1j * ((uui - j2.x) * rinv) * grkerni);
      if (PHYS != 1) {
        // the "curl v" accumulation of the dissipation-switch block
        // (src/ratesND_mhd.f90:1653), written to dumps of non-MHD runs
        gdx.add(mj * (dvel[0] - dvdotr) * grkerni);
        gdy.add(mj * (dvel[1] - dvdotr) * grkerni);
        gdz.add(mj * (dvel[2] - dvdotr) * grkerni);
      }
    };
    if (!err) {
      // one record ahead: a second 96-byte record in flight would cost the
      // third wave per SIMD (162 registers with one, 180 with two)
      if (uselist) for_each_listed_pf<NDIM, 1>(a, g, k, a.nbc[k] + a.nbx[k], loadj, body);
      else for_each_candidate_rec<NDIM>(a, g, xi, rad, loadj, body);
    }
    st_stream(&a.f[F_FX][i], fx.value());
    st_stream(&a.f[F_FY][i], fy.value());
    st_stream(&a.f[F_FZ][i], fz.value());
    st_stream(&a.f[F_DUDT][i], dudt.value());
    st_stream(&a.f[F_DBX][i], dbx.value());
    st_stream(&a.f[F_DBY][i], dby.value());
    st_stream(&a.f[F_DBZ][i], dbz.value());
    st_stream(&a.f[F_DIVB][i], divb.value());
    st_stream(&a.f[F_CBX][i], cbx.value());
    st_stream(&a.f[F_CBY][i], cby.value());
    st_stream(&a.f[F_CBZ][i], cbz.value());
    st_stream(&a.f[F_GPX][i], gpx.value());
    st_stream(&a.f[F_GPY][i], gpy.value());
    st_stream(&a.f[F_GPZ][i], gpz.value());
    st_stream(&a.f[F_DEL2U][i], del2u.value());
    st_stream(&a.f[F_DEPS][i], deps.value());
    if (PHYS != 1) {
      st_stream(&a.f[F_GDVX][i], gdx.value());
      st_stream(&a.f[F_GDVY][i], gdy.value());
      st_stream(&a.f[F_GDVZ][i], gdz.value());
    }
  }
  if (err) atomicCAS(&a.sc->err, 0, err);
  {
    if (rdt_l > PT(0)) dtc_l = fmin(dtc_l, 1.0 / (double)rdt_l);
    double rv[3] = {(double)vsigmax_l, dtc_l, (double)npairs};
    const int rk[3] = {RED_MAX, RED_MIN, RED_SUM};
    block_partials<3>(rv, rk, a.red, a.rcap, vl);
  }
}
template <int NDIM, typename PT, int PHYS, int TB>
__global__ __launch_bounds__(TB) void k_force(KArgs a, Phys p, const Grid* __restrict__ gp) {
  b_force<NDIM, PT, PHYS, TB>(a, p, gp, VL{(int)blockIdx.x, (int)gridDim.x});
}

// per-particle finalisation, src/ratesND_mhd.f90:482-943 (FP64)
__device__ __forceinline__ void b_finalize(KArgs a, Phys p, VL vl) {
#pragma clang fp contract(off)
  int i = vl.b * blockDim.x + threadIdx.x;
  double fh = 0.0, dtd = DBL_MAX;
  if (i < a.n) {
    double** f = a.f;
    const double rhoi = f[F_RHO][i];
    const double rho1i = 1.0 / rhoi;
    const double h = f[F_H][i];
    const double cs = f[F_CS][i];
    if (p.idust == 4) {
      const double rhodusti = f[F_RHODUST][i], rhogasi = f[F_RHOGAS][i];
      const double dustfraci = rhodusti * rho1i;
      const double ts = p.Kdrag;
      f[F_DVX][i] = rhogasi > DBL_MIN ? -rhoi / rhogasi * f[F_FX][i] * ts : 0.0;
      f[F_DVY][i] = rhogasi > DBL_MIN ? -rhoi / rhogasi * f[F_FY][i] * ts : 0.0;
      f[F_DVZ][i] = rhogasi > DBL_MIN ? -rhoi / rhogasi * f[F_FZ][i] * ts : 0.0;
      dtd = fabs(0.5 * (h * h) / (dustfraci * ts * (cs * cs)));
    }
    const bool mhd = p.imhd != 0;
    if (mhd) {
      f[F_CBX][i] *= rho1i;
      f[F_CBY][i] *= rho1i;
      f[F_CBZ][i] *= rho1i;
      f[F_DIVB][i] *= rho1i;
    }
    const double fxv = f[F_FX][i], fyv = f[F_FY][i], fzv = f[F_FZ][i];
    fh = sqrt(fxv * fxv + fyv * fyv + fzv * fzv) / h;
    double valfven2i = 0.0;
    const double bx = f[F_BX][i], by = f[F_BY][i], bz = f[F_BZ][i];
    if (mhd) valfven2i = (bx * bx + by * by + bz * bz) / f[F_DENS][i];
    const double vsig = sqrt(cs * cs + valfven2i);
    if (p.imhd == 11) {
      const double drhodt = f[F_DRHODT][i];
      double dbx = f[F_DBX][i] + bx * rho1i * drhodt;
      double dby = f[F_DBY][i] + by * rho1i * drhodt;
      double dbz = f[F_DBZ][i] + bz * rho1i * drhodt;
      if (p.idivbzero >= 2) {
        const double gx = f[F_GPX][i] * rhoi, gy = f[F_GPY][i] * rhoi, gz = f[F_GPZ][i] * rhoi;
        f[F_GPX][i] = gx; f[F_GPY][i] = gy; f[F_GPZ][i] = gz;
        dbx = dbx + gx; dby = dby + gy; dbz = dbz + gz;
      }
      f[F_DBX][i] = dbx; f[F_DBY][i] = dby; f[F_DBZ][i] = dbz;
    }
    if (p.iener > 0 && p.idust == 0) {
      f[F_DUDT][i] = f[F_DUDT][i] + f[F_PR][i] * (rho1i * rho1i) * f[F_DRHODT][i];
    }
    f[F_DENDT][i] = f[F_DUDT][i];
    double da0 = 0.0, da1 = 0.0, da2 = 0.0;
    if (p.iavlim[0] || p.iavlim[1] || p.iavlim[2]) {
      const double tdecay1 = (p.avdecayconst * vsig) / h;
      const double al0 = f[F_A0][i], al1 = f[F_A1][i], al2 = f[F_A2][i];
      if (p.iavlim[0] == 1 || p.iavlim[0] == 2) {
        double source = fmax(f[F_DRHODT][i] * rho1i, 0.0);
        if (p.iavlim[0] == 2) source = source * (2.0 - al0);
        da0 = (p.alphamin - al0) * tdecay1 + p.avfact * source;
      }
      if (p.iener > 0 && p.iavlim[1] > 0) {
        const double uu = f[F_UU][i];
        const double sourceu = uu > DBL_EPSILON ? h * fabs(f[F_DEL2U][i]) / sqrt(uu) : 0.0;
        da1 = (p.alphaumin - al1) * tdecay1 + sourceu;
      }
      if (p.iavlim[2] != 0 && mhd) {
        const double c0 = f[F_CBX][i], c1 = f[F_CBY][i], c2 = f[F_CBZ][i];
        const double sourceJ = sqrt((c0 * c0 + c1 * c1 + c2 * c2) * rho1i);
        const double sourcedivB = 10. * fabs(f[F_DIVB][i]) * sqrt(rho1i);
        double sourceB = fmax(sourceJ, sourcedivB);
        if (p.iavlim[2] == 3) sourceB = sqrt(fmax(f[F_DRHODT][i] * rho1i, 0.0) * (2. - al2) * sourceB);
        da2 = (p.alphabmin - al2) * tdecay1 + sourceB;
      }
    }
    f[F_DA0][i] = da0;
    f[F_DA1][i] = da1;
    f[F_DA2][i] = da2;
    if (p.idivbzero >= 2 && p.idivbzero <= 7) {
      const double vsigmax = a.sc->vsigmax;
      f[F_DPSIDT][i] = -(vsigmax * vsigmax) * f[F_DIVB][i] - p.psidecayfact * f[F_PSI][i] * vsigmax / h;
    } else {
      f[F_DPSIDT][i] = 0.0;
    }
    if (p.fixed_on && a.itype[i] == 1) {
      // rates are zero on fixed particles, after their force entered fhmax
      // (src/ratesND_mhd.f90:949-966)
      static const int zf[] = {F_FX, F_FY, F_FZ, F_DRHODT, F_DHDT, F_DUDT, F_DENDT, F_DBX, F_DBY, F_DBZ,
                               F_DA0, F_DA1, F_DA2, F_DPSIDT, F_DIVB, F_CBX, F_CBY, F_CBZ, F_GPX, F_GPY,
                               F_GPZ};
      for (int q = 0; q < (int)(sizeof zf / sizeof zf[0]); q++) f[zf[q]][i] = 0.0;
    }
  }
  double rv[2] = {fh, dtd};
  const int rk[2] = {RED_MAX, RED_MIN};
  block_partials<2>(rv, rk, a.red, a.rcap, vl);
}
__global__ void k_finalize(KArgs a, Phys p) {
  b_finalize(a, p, VL{(int)blockIdx.x, (int)gridDim.x});
}

// timestep, src/ratesND_mhd.f90:938-943 and src/stepND_leapfrog_mhd.f90:252
__device__ __forceinline__ void b_dt(DScal* sc, Phys p, VL vl) {
  if (threadIdx.x | vl.b) return;
  double dtforce = DBL_MAX;
  if (sc->fhmax > 0.0) dtforce = fmin(dtforce, sqrt(1.0 / sc->fhmax));
  sc->dtforce = dtforce;
  sc->dt = fmin(fmin(fmin(p.C_force * dtforce, p.C_cour * sc->dtcourant), 0.9 * sc->dtdrag),
                p.C_force * DBL_MAX);
}
__global__ void k_dt(DScal* sc, Phys p) {
  b_dt(sc, p, VL{(int)blockIdx.x, (int)gridDim.x});
}

// end of a leapfrog step: advance the time by the step's dt and take the dt
// chosen by this step's rates for the next one (src/evolve.f90:106-110)
__device__ __forceinline__ void b_step_end(DScal* sc, VL vl) {
  if (threadIdx.x | vl.b) return;
  sc->time = sc->time + sc->dt_cur;
  sc->dt_cur = sc->dt;
}
__global__ void k_step_end(DScal* sc) {
  b_step_end(sc, VL{(int)blockIdx.x, (int)gridDim.x});
}

__device__ __forceinline__ void b_init_density_scalars(DScal* sc, double bconstmax, int n, VL vl) {
  if (threadIdx.x | vl.b) return;
  sc->pc_nlist = n;
  sc->pc_its = 1;
  sc->pc_sym = 1;
  sc->pc_persistent = 0;
  sc->pc_done = 0;
  sc->pc_relink = 0;
  sc->pc_refresh = 0;
  sc->hmax_nr = 0.0;
  sc->hmax_cur = sc->hmax_bin;
  sc->hhmax_ghost = sc->hmax_bin;
  sc->ncalc_next = 0;
  sc->relink = 0;
  sc->list_invalid = 0;
  sc->stressmax = fmax(0.0, bconstmax);
  sc->itsmax = 0;
  sc->nmin = INT_MAX;
  sc->nmax = 0;
  sc->pairs_density = 0;
  sc->ncalc_total = 0;
}
__global__ void k_init_density_scalars(DScal* sc, double bconstmax, int n) {
  b_init_density_scalars(sc, bconstmax, n, VL{(int)blockIdx.x, (int)gridDim.x});
}

__device__ __forceinline__ void b_init_force_scalars(DScal* sc, VL vl) {
  if (threadIdx.x | vl.b) return;
  sc->vsigmax = 0.0;
  sc->dtcourant = 1.0e6;
  sc->fhmax = 0.0;
  sc->dtdrag = DBL_MAX;
  sc->pairs_force = 0;
}
__global__ void k_init_force_scalars(DScal* sc) {
  b_init_force_scalars(sc, VL{(int)blockIdx.x, (int)gridDim.x});
}

// ===========================================================================
// leapfrog, src/stepND_leapfrog_mhd.f90:31-275
// ===========================================================================

__device__ __forceinline__ void b_predict(KArgs a, Phys p, VL vl) {
#pragma clang fp contract(off)
  int i = vl.b * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const double dt = a.sc->dt_cur;
  double** f = a.f;
  static const int xs[3] = {F_X, F_Y, F_Z}, vs[3] = {F_VX, F_VY, F_VZ}, bs[3] = {F_BX, F_BY, F_BZ};
  static const int as[3] = {F_A0, F_A1, F_A2}, fs[3] = {F_FX, F_FY, F_FZ}, dbs[3] = {F_DBX, F_DBY, F_DBZ};
  static const int das[3] = {F_DA0, F_DA1, F_DA2};
  // every input into registers first (the field pointers may alias as far as
  // the compiler knows, so reading back the copies would reload them)
  double x[3], v[3], b[3], al[3], fo[3], db[3], da[3];
  for (int c = 0; c < 3; c++) {
    x[c] = f[xs[c]][i];
    v[c] = f[vs[c]][i];
    b[c] = f[bs[c]][i];
    al[c] = f[as[c]][i];
    fo[c] = f[fs[c]][i];
    db[c] = f[dbs[c]][i];
    da[c] = f[das[c]][i];
  }
  const double h = f[F_H][i], en = f[F_EN][i], psi = f[F_PSI][i], eps = f[F_EPS][i];
  const double dhdt = f[F_DHDT][i], dendt = f[F_DENDT][i], dpsidt = f[F_DPSIDT][i], deps = f[F_DEPS][i];
  for (int c = 0; c < 3; c++) {
    f[F_XIN + c][i] = x[c];
    f[F_VXIN + c][i] = v[c];
    f[F_BXIN + c][i] = b[c];
    f[F_A0IN + c][i] = al[c];
    f[F_FXIN + c][i] = fo[c];
    f[F_DBXIN + c][i] = db[c];
    f[F_DA0IN + c][i] = da[c];
  }
  f[F_HIN][i] = h;
  f[F_ENIN][i] = en;
  f[F_PSIIN][i] = psi;
  f[F_EPSIN][i] = eps;
  f[F_DHDTIN][i] = dhdt;
  f[F_DENDTIN][i] = dendt;
  f[F_DPSIDTIN][i] = dpsidt;
  f[F_DEPSIN][i] = deps;
  for (int d = 0; d < 3; d++) if (d < p.ndim) f[xs[d]][i] = x[d] + dt * v[d] + 0.5 * dt * dt * fo[d];
  for (int c = 0; c < 3; c++) f[vs[c]][i] = (v[c] + dt * fo[c]) / 1.0;
  if (p.imhd != 0)
    for (int c = 0; c < 3; c++) f[bs[c]][i] = b[c] + dt * db[c];
  f[F_H][i] = h + dt * dhdt;
  if (p.iener != 0) f[F_EN][i] = en + dt * dendt;
  for (int c = 0; c < 3; c++)
    if (p.iavlim[c] != 0) f[as[c]][i] = fmin(al[c] + dt * da[c], 1.0);
  if (p.idivbzero >= 2) f[F_PSI][i] = psi + dt * dpsidt;
  if (p.idust) f[F_EPS][i] = eps + dt * deps;
}
__global__ void k_predict(KArgs a, Phys p) {
  b_predict(a, p, VL{(int)blockIdx.x, (int)gridDim.x});
}

// Predictor of the fixed particles, after b_predict has written every
// slot's step in-copies (src/stepND_leapfrog_mhd.f90:110-143): a fixed
// particle moves with the starting velocity and force of the particle it
// copies and holds its own velocity, B, h, energy, switches, psi and dust.
__device__ __forceinline__ void b_predict_fixed(KArgs a, Phys p, VL vl) {
#pragma clang fp contract(off)
  const int i = vl.b * blockDim.x + threadIdx.x;
  if (!p.fixed_on || i >= a.n || a.itype[i] != 1) return;
  const double dt = a.sc->dt_cur;
  double** f = a.f;
  const int j = a.slot_of[a.ireal[i]];
  static const int xs[3] = {F_X, F_Y, F_Z}, vs[3] = {F_VX, F_VY, F_VZ}, bs[3] = {F_BX, F_BY, F_BZ};
  static const int as[3] = {F_A0, F_A1, F_A2};
  for (int d = 0; d < 3; d++) if (d < p.ndim)
    f[xs[d]][i] = f[F_XIN + d][i] + dt * f[F_VXIN + d][j] + 0.5 * dt * dt * f[F_FXIN + d][j];
  for (int c = 0; c < 3; c++) {
    f[vs[c]][i] = f[F_VXIN + c][i];
    f[bs[c]][i] = f[F_BXIN + c][i];
    f[as[c]][i] = f[F_A0IN + c][i];
  }
  f[F_H][i] = f[F_HIN][i];
  f[F_EN][i] = f[F_ENIN][i];
  f[F_PSI][i] = f[F_PSIIN][i];
  f[F_EPS][i] = f[F_EPSIN][i];
  f[F_RHOIN][i] = f[F_RHO][i];
}
__global__ void k_predict_fixed(KArgs a, Phys p) {
  b_predict_fixed(a, p, VL{(int)blockIdx.x, (int)gridDim.x});
}

// After the h-rho iteration a fixed particle becomes a replica of the particle
// it copies (iterate_density.f90:310-326, then copy_particle from
// conservative2primitive.f90:425-432): mass, density, h, energies, B,
// switches, psi, the h terms, pressure and sound speed, and the dust
// fields; position and velocity stay its own.  Rates are formed afterwards.
__device__ __forceinline__ void b_fixed_copy(KArgs a, Phys p, VL vl) {
  const int i = vl.b * blockDim.x + threadIdx.x;
  if (!p.fixed_on || i >= a.n || a.itype[i] != 1) return;
  const int j = a.slot_of[a.ireal[i]];
  a.f[F_MIMG][i] = a.f[F_M][i];
  a.f[F_A0IMG][i] = a.f[F_A0][i];
  a.f[F_A1IMG][i] = a.f[F_A1][i];
  a.f[F_A2IMG][i] = a.f[F_A2][i];
  static const int cf[] = {F_M, F_RHO, F_H, F_UU, F_EN, F_BX, F_BY, F_BZ, F_A0, F_A1, F_A2, F_PSI,
                           F_GRADH, F_GRADGRADH, F_CS, F_PR, F_DENS, F_DRHODT, F_DHDT, F_DUSTFRAC,
                           F_EPS, F_RHOGAS, F_RHODUST, F_DVX, F_DVY, F_DVZ};
  for (int q = 0; q < (int)(sizeof cf / sizeof cf[0]); q++) a.f[cf[q]][i] = a.f[cf[q]][j];
}
__global__ void k_fixed_copy(KArgs a, Phys p) {
  b_fixed_copy(a, p, VL{(int)blockIdx.x, (int)gridDim.x});
}

// slot of every caller index (after uploads and reorders)
__device__ __forceinline__ void b_slot_of(KArgs a, VL vl) {
  const int s = vl.b * blockDim.x + threadIdx.x;
  if (s < a.n) a.slot_of[a.uid[s]] = s;
}
__global__ void k_slot_of(KArgs a) { b_slot_of(a, VL{(int)blockIdx.x, (int)gridDim.x}); }

__device__ __forceinline__ void b_correct(KArgs a, Phys p, VL vl) {
#pragma clang fp contract(off)
  int i = vl.b * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  const double dt = a.sc->dt_cur;
  double** f = a.f;
  const double hdt = 0.5 * dt;
  static const int vs[3] = {F_VX, F_VY, F_VZ}, bs[3] = {F_BX, F_BY, F_BZ}, as[3] = {F_A0, F_A1, F_A2};
  static const int fs[3] = {F_FX, F_FY, F_FZ}, dbs[3] = {F_DBX, F_DBY, F_DBZ}, das[3] = {F_DA0, F_DA1, F_DA2};
  f[F_HEVAL][i] = f[F_H][i];
  if (p.fixed_on && a.itype[i] == 1) {
    // fixed particles get back their own starting values
    // (src/stepND_leapfrog_mhd.f90:172-186)
    for (int c = 0; c < 3; c++) {
      f[vs[c]][i] = f[F_VXIN + c][i];
      f[bs[c]][i] = f[F_BXIN + c][i];
      f[as[c]][i] = f[F_A0IN + c][i];
    }
    f[F_H][i] = f[F_HIN][i];
    f[F_EN][i] = f[F_ENIN][i];
    f[F_PSI][i] = f[F_PSIIN][i];
    f[F_EPS][i] = f[F_EPSIN][i];
    f[F_RHO][i] = f[F_RHOIN][i];
    return;
  }
  for (int c = 0; c < 3; c++)
    f[vs[c]][i] = (f[F_VXIN + c][i] + hdt * (f[fs[c]][i] + f[F_FXIN + c][i])) / 1.0;
  if (p.imhd != 0)
    for (int c = 0; c < 3; c++) f[bs[c]][i] = f[F_BXIN + c][i] + hdt * (f[dbs[c]][i] + f[F_DBXIN + c][i]);
  const double hnew = f[F_HIN][i] + hdt * (f[F_DHDT][i] + f[F_DHDTIN][i]);
  f[F_H][i] = hnew;
  if (hnew <= 0.0) atomicCAS(&a.sc->err, 0, NDSP_ERR_H_NONPOSITIVE);
  if (p.iener != 0) f[F_EN][i] = f[F_ENIN][i] + hdt * (f[F_DENDT][i] + f[F_DENDTIN][i]);
  for (int c = 0; c < 3; c++)
    if (p.iavlim[c] != 0) f[as[c]][i] = fmin(f[F_A0IN + c][i] + hdt * (f[das[c]][i] + f[F_DA0IN + c][i]), 1.0);
  if (p.idivbzero >= 2) f[F_PSI][i] = f[F_PSIIN][i] + hdt * (f[F_DPSIDT][i] + f[F_DPSIDTIN][i]);
  if (p.idust) f[F_EPS][i] = f[F_EPSIN][i] + hdt * (f[F_DEPS][i] + f[F_DEPSIN][i]);
}
// finalize, then the corrector, particle by particle: neither reads another
// particle's fields, and the rates the corrector reads were just written by
// the same thread (no second pass over them from HBM)
__global__ void k_finalize_correct(KArgs a, Phys p) {
  const VL vl{(int)blockIdx.x, (int)gridDim.x};
  b_finalize(a, p, vl);
  b_correct(a, p, vl);
}

// ===========================================================================
// FP64 conserved-quantity reduction, src/evwrite_mhd.f90:124-248
// ===========================================================================

#define NCONS 17
__global__ void k_conserved(KArgs a, Phys p, double* part) {
  __shared__ double sm[NCONS][BS];
  int t = threadIdx.x;
  double v[NCONS];
  for (int q = 0; q < NCONS; q++) v[q] = 0.0;
  v[12] = DBL_MAX;
  v[13] = -DBL_MAX;
  double** f = a.f;
  for (int i = blockIdx.x * blockDim.x + t; i < a.n; i += gridDim.x * blockDim.x) {
    const double m = f[F_M][i];
    const double vx = f[F_VX][i], vy = f[F_VY][i], vz = f[F_VZ][i];
    const double x = f[F_X][i], y = p.ndim >= 2 ? f[F_Y][i] : 0.0, z = p.ndim >= 3 ? f[F_Z][i] : 0.0;
    v[4] += m * vx; v[5] += m * vy; v[6] += m * vz;
    if (p.ndim == 3) {
      v[8] += m * (y * vz - z * vy);
      v[9] += m * (z * vx - x * vz);
      v[10] += m * (x * vy - y * vx);
    } else if (p.ndim == 2) {
      v[10] += m * (x * vy - y * vx);
    }
    v[0] += 0.5 * m * (vx * vx + vy * vy + vz * vz);
    const double rho = f[F_RHO][i];
    if (p.idust) {
      const double eps = f[F_DUSTFRAC][i], dterm = 1.0 - eps;
      const double dx = f[F_DVX][i], dy = f[F_DVY][i], dz = f[F_DVZ][i];
      v[0] += 0.5 * m * eps * dterm * (dx * dx + dy * dy + dz * dz);
      v[1] += m * f[F_UU][i] * dterm;
      v[15] += m * dterm;
      v[16] += m * eps;
    } else {
      v[1] += m * f[F_UU][i];
    }
    if (p.imhd) {
      const double bx = f[F_BFX][i], by = f[F_BFY][i], bz = f[F_BFZ][i];
      v[2] += 0.5 * m * (bx * bx + by * by + bz * bz) / rho;
    }
    v[12] = fmin(v[12], rho);
    v[13] = fmax(v[13], rho);
    v[14] += rho;
  }
  for (int q = 0; q < NCONS; q++) sm[q][t] = v[q];
  __syncthreads();
  for (int s = BS / 2; s > 0; s >>= 1) {
    if (t < s) {
      for (int q = 0; q < NCONS; q++) {
        if (q == 12) sm[q][t] = fmin(sm[q][t], sm[q][t + s]);
        else if (q == 13) sm[q][t] = fmax(sm[q][t], sm[q][t + s]);
        else sm[q][t] += sm[q][t + s];
      }
    }
    __syncthreads();
  }
  if (t < NCONS) part[blockIdx.x * NCONS + t] = sm[t][0];
}

// folds the per-block partials part[nb][NQ] into v[NQ] (shared memory) in
// an order fixed by nb alone, so that the result does not depend on timing:
// thread t takes blocks t, t+256, ... in turn, then a butterfly over each
// wave and the four waves in order.  One block of 256 threads.
#define FOLD_T 256
template <int NQ, class Comb, class Ident>
__device__ __forceinline__ void fold_block_partials(const double* __restrict__ part, int nb, double* v, Comb comb,
                                                    Ident ident) {
  __shared__ double sw[FOLD_T / 64][NQ];
  double acc[NQ];
#pragma unroll
  for (int q = 0; q < NQ; q++) acc[q] = ident(q);
  for (int b = threadIdx.x; b < nb; b += FOLD_T) {
    const double* r = part + (size_t)b * NQ;
#pragma unroll
    for (int q = 0; q < NQ; q++) acc[q] = comb(q, acc[q], r[q]);
  }
#pragma unroll
  for (int q = 0; q < NQ; q++) {
    double x = acc[q];
    for (int o = 32; o > 0; o >>= 1) x = comb(q, x, __shfl_xor(x, o, 64));
    acc[q] = x;
  }
  if ((threadIdx.x & 63) == 0) {
#pragma unroll
    for (int q = 0; q < NQ; q++) sw[threadIdx.x >> 6][q] = acc[q];
  }
  __syncthreads();
  if (threadIdx.x < NQ) {
    const int q = threadIdx.x;
    double x = sw[0][q];
    for (int k = 1; k < FOLD_T / 64; k++) x = comb(q, x, sw[k][q]);
    v[q] = x;
  }
}

__global__ __launch_bounds__(FOLD_T) void k_conserved_final(const double* part, int nb, double* out, int n) {
  __shared__ double v[NCONS];
  fold_block_partials<NCONS>(
      part, nb, v,
      [](int q, double x, double y) { return q == 12 ? fmin(x, y) : q == 13 ? fmax(x, y) : x + y; },
      [](int q) { return q == 12 ? DBL_MAX : q == 13 ? -DBL_MAX : 0.0; });
  __syncthreads();
  if (threadIdx.x | blockIdx.x) return;
  v[3] = v[0] + v[2] + v[1];
  v[7] = sqrt(v[4] * v[4] + v[5] * v[5] + v[6] * v[6]);
  v[11] = sqrt(v[8] * v[8] + v[9] * v[9] + v[10] * v[10]);
  v[14] = v[14] / (double)n;
  for (int q = 0; q < NCONS; q++) out[q] = v[q];
}

// ===========================================================================
// the quantities of one .ev line, src/evwrite_mhd.f90:124-248, FP64
// ===========================================================================

// slots of the per-block partials: sums, then minima, then maxima
enum {
  EV_EKIN, EV_ETHERM, EV_EMAG, EV_EMAGP, EV_MOMX, EV_MOMY, EV_MOMZ, EV_DMOMX, EV_DMOMY, EV_DMOMZ,
  EV_ANGX, EV_ANGY, EV_ANGZ, EV_EKINY, EV_MGAS, EV_MDUST, EV_RHOSUM, EV_BETASUM, EV_DIVBSUM,
  EV_DIVBTOT, EV_OMEGASUM, EV_NDIVBOK, EV_CROSSHEL, EV_NSUM,
  EV_RHOMIN = EV_NSUM, EV_BETAMIN, EV_NMIN_END,
  EV_RHOMAX = EV_NMIN_END, EV_DIVBMAX, EV_OMEGAMAX, EV_NQ
};

__device__ __forceinline__ double ev_combine(int q, double x, double y) {
  return q < EV_NSUM ? x + y : q < EV_NMIN_END ? fmin(x, y) : fmax(x, y);
}
__device__ __forceinline__ double ev_identity(int q) {
  return q < EV_NSUM ? 0.0 : q < EV_NMIN_END ? DBL_MAX : -DBL_MAX;
}

// per-particle terms of evwrite's loop over the real particles
// (src/evwrite_mhd.f90:124-245) folded into per-block partials; the
// magnetic-force diagnostics (fdotB, force_err) are not summed: with the
// tensor-form Lorentz force fmag stays zero (src/ratesND_mhd.f90:202,
// 2621-2626), so both are zero for every particle
__global__ __launch_bounds__(BS) void k_ev_sums(KArgs a, Phys p, double* __restrict__ part) {
  __shared__ double sm[BS / 64][EV_NQ];
  double v[EV_NQ];
#pragma unroll
  for (int q = 0; q < EV_NQ; q++) v[q] = ev_identity(q);
  double** f = a.f;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += gridDim.x * blockDim.x) {
    const double m = f[F_M][i], rho = f[F_RHO][i];
    const double vx = f[F_VX][i], vy = f[F_VY][i], vz = f[F_VZ][i];
    v[EV_MOMX] += m * vx; v[EV_MOMY] += m * vy; v[EV_MOMZ] += m * vz;
    v[EV_DMOMX] += m * f[F_FX][i]; v[EV_DMOMY] += m * f[F_FY][i]; v[EV_DMOMZ] += m * f[F_FZ][i];
    const double x = f[F_X][i], y = p.ndim >= 2 ? f[F_Y][i] : 0.0, z = p.ndim >= 3 ? f[F_Z][i] : 0.0;
    if (p.ndim == 3) {
      v[EV_ANGX] += m * (y * vz - z * vy);
      v[EV_ANGY] += m * (z * vx - x * vz);
      v[EV_ANGZ] += m * (x * vy - y * vx);
    } else if (p.ndim == 2) {
      v[EV_ANGZ] += m * (x * vy - y * vx);
    }
    v[EV_EKIN] += 0.5 * m * (vx * vx + vy * vy + vz * vz);
    if (p.idust) {
      const double eps = f[F_DUSTFRAC][i], dterm = 1.0 - eps;
      const double dx = f[F_DVX][i], dy = f[F_DVY][i], dz = f[F_DVZ][i];
      const double ekdv = 0.5 * m * eps * dterm * (dx * dx + dy * dy + dz * dz);
      v[EV_EKIN] += ekdv;
      v[EV_EKINY] += ekdv;
      v[EV_ETHERM] += m * f[F_UU][i] * dterm;
      v[EV_MGAS] += m * dterm;
      v[EV_MDUST] += m * eps;
    } else {
      if (p.ndim >= 2) v[EV_EKINY] += 0.5 * m * vx * vx;
      v[EV_ETHERM] += m * f[F_UU][i];
    }
    if (p.imhd) {
      const double bx = f[F_BFX][i], by = f[F_BFY][i], bz = f[F_BFZ][i];
      const double b2 = bx * bx + by * by + bz * bz, bmag = sqrt(b2);
      const double divb = fabs(f[F_DIVB][i]);
      v[EV_EMAG] += 0.5 * m * b2 / rho;
      v[EV_EMAGP] += 0.5 * m * (bx * bx + by * by) / rho;
      const double beta = b2 < DBL_MIN ? 0.0 : f[F_PR][i] / (0.5 * b2);
      v[EV_BETASUM] += beta;
      v[EV_BETAMIN] = fmin(v[EV_BETAMIN], beta);
      v[EV_DIVBMAX] = fmax(v[EV_DIVBMAX], divb);
      v[EV_DIVBSUM] += divb;
      v[EV_DIVBTOT] += m * divb / rho;
      const double om = bmag < 1.0e-8 ? 0.0 : divb * f[F_H][i] / bmag;
      if (om < 1.0e-2) v[EV_NDIVBOK] += 1.0;
      v[EV_OMEGAMAX] = fmax(v[EV_OMEGAMAX], om);
      v[EV_OMEGASUM] += om;
      v[EV_CROSSHEL] += m * (vx * (bx / rho) + vy * (by / rho) + vz * (bz / rho));
    }
    v[EV_RHOSUM] += rho;
    v[EV_RHOMIN] = fmin(v[EV_RHOMIN], rho);
    v[EV_RHOMAX] = fmax(v[EV_RHOMAX], rho);
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < EV_NQ; q++) {
    double x = v[q];
    for (int o = 32; o > 0; o >>= 1) x = ev_combine(q, x, __shfl_xor(x, o, 64));
    if (lane == 0) sm[w][q] = x;
  }
  __syncthreads();
  if (threadIdx.x < EV_NQ) {
    const int q = threadIdx.x;
    double x = sm[0][q];
    for (int k = 1; k < BS / 64; k++) x = ev_combine(q, x, sm[k][q]);
    part[(size_t)blockIdx.x * EV_NQ + q] = x;
  }
}

// folds the block partials and forms the columns of the .ev line in the
// order evwrite writes them (src/evwrite_mhd.f90:247-308)
__global__ __launch_bounds__(FOLD_T) void k_ev_final(const double* __restrict__ part, int nb, int n, Phys p,
                                                     double t, double dt, double* __restrict__ out) {
  __shared__ double v[EV_NQ];
  fold_block_partials<EV_NQ>(
      part, nb, v, [](int q, double x, double y) { return ev_combine(q, x, y); },
      [](int q) { return ev_identity(q); });
  __syncthreads();
  if (threadIdx.x | blockIdx.x) return;
  const double epot = 0.0;  // no gravity or external potentials on this path
  const double ekin = v[EV_EKIN], etherm = v[EV_ETHERM], emag = v[EV_EMAG];
  // iprterm 0: the thermal energy counts (src/evwrite_mhd.f90:247-248)
  const double etot = ekin + emag + epot + etherm;
  const double momtot = sqrt(v[EV_MOMX] * v[EV_MOMX] + v[EV_MOMY] * v[EV_MOMY] + v[EV_MOMZ] * v[EV_MOMZ]);
  const double dmomtot = sqrt(v[EV_DMOMX] * v[EV_DMOMX] + v[EV_DMOMY] * v[EV_DMOMY] + v[EV_DMOMZ] * v[EV_DMOMZ]);
  const double angtot = sqrt(v[EV_ANGX] * v[EV_ANGX] + v[EV_ANGY] * v[EV_ANGY] + v[EV_ANGZ] * v[EV_ANGZ]);
  const double fn = (double)n;
  int k = 0;
  const double common[11] = {t, ekin, etherm, emag, epot, etot, momtot, angtot, v[EV_RHOMAX], v[EV_RHOSUM] / fn, dt};
  for (int q = 0; q < 11; q++) out[k++] = common[q];
  if (p.imhd) {
    out[k++] = v[EV_EMAGP];
    out[k++] = v[EV_CROSSHEL];
    out[k++] = v[EV_BETAMIN];
    out[k++] = v[EV_BETASUM] / fn;
    out[k++] = v[EV_DIVBSUM] / fn;
    out[k++] = v[EV_DIVBMAX];
    out[k++] = v[EV_DIVBTOT];
    out[k++] = 0.0;  // fdotBav
    out[k++] = 0.0;  // fdotBmax
    out[k++] = 0.0;  // force_err_av
    out[k++] = 0.0;  // force_err_max
    out[k++] = v[EV_OMEGASUM] / fn;
    out[k++] = v[EV_OMEGAMAX];
    out[k++] = 100.0 * v[EV_NDIVBOK] / fn;
  } else if (p.idust) {
    out[k++] = v[EV_EKINY];
    out[k++] = v[EV_MGAS];
    out[k++] = v[EV_MDUST];
  } else {
    out[k++] = v[EV_EKINY];
    out[k++] = dmomtot;
  }
}

// ===========================================================================
// reorder of the particle state into cell order
// ===========================================================================

__global__ void k_fill_f(float* __restrict__ a, int n, float v) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) a[k] = v;
}

__global__ void k_gather_d(const double* __restrict__ in, double* __restrict__ out, const int* sidx, int n) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = in[sidx[k]];
}
__global__ void k_fill_d(double* __restrict__ a, int n, double v) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) a[i] = v;
}
__global__ void k_copy_d(const double* __restrict__ in, double* __restrict__ out, int n) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = in[k];
}
// counter calibration for gather-shaped access: thread i reads the 64-byte
// line (i * mult) mod nl of src (a bijection for odd mult, nl a power of
// two: every line once, in a scattered order) and writes one float
__global__ void k_gather_probe(const double4* __restrict__ src, float* __restrict__ out, long long nl,
                               unsigned long long mult) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nl) return;
  const long long j = (long long)(((unsigned long long)i * mult) & (unsigned long long)(nl - 1));
  const double4 u = src[2 * j], v = src[2 * j + 1];
  out[i] = (float)(u.x + u.y + u.z + u.w + v.x + v.y + v.z + v.w);
}
__global__ void k_copy_i(const int* __restrict__ in, int* __restrict__ out, int n) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = in[k];
}
__global__ void k_gather_i(const int* __restrict__ in, int* __restrict__ out, const int* sidx, int n) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) out[k] = in[sidx[k]];
}

// up to REORDER_W fields of one kind per launch: gather into the scratch
// (slot k <- sidx[k]), then copy back, so the field pointers never change
#define REORDER_W 8
template <typename T>
struct FieldSet {
  T* f[REORDER_W];
  int nf;
};
template <typename T>
__global__ void k_gather_set(FieldSet<T> fs, T* __restrict__ scratch, const int* __restrict__ sidx, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int i = sidx[k];
#pragma unroll
  for (int q = 0; q < REORDER_W; q++)
    if (q < fs.nf) scratch[(size_t)q * n + k] = fs.f[q][i];
}
// src[q][sidx[k]] -> dst[q][k]: a reorder into the other field set
struct FieldPairs {
  const double* src[REORDER_W];
  double* dst[REORDER_W];
  int nf;
};
__global__ void k_gather_pairs(FieldPairs fp, const int* __restrict__ sidx, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  const int i = sidx[k];
#pragma unroll
  for (int q = 0; q < REORDER_W; q++)
    if (q < fp.nf) fp.dst[q][k] = fp.src[q][i];
}
template <typename T>
__global__ void k_copy_set(FieldSet<T> fs, const T* __restrict__ scratch, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
#pragma unroll
  for (int q = 0; q < REORDER_W; q++)
    if (q < fs.nf) fs.f[q][k] = scratch[(size_t)q * n + k];
}

// ===========================================================================
// neighbour-set enumeration (test support)
// ===========================================================================

template <int NDIM>
__global__ void k_nb(KArgs a, Phys p, const Grid* __restrict__ gp, int sym, const int* offs, int* out) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= a.n) return;
  const Grid g = *gp;
  const double4 pk = a.P4[k];
  const double xi[3] = {pk.x, pk.y, pk.z};
  const double hi21 = mul_exact(pk.w, pk.w);
  const double R2 = p.radkern2;
  const int i = a.sidx[k];
  const double rad = p.radkern * (sym ? a.sc->hmax_nr : a.f[F_H][i]);
  int cnt = 0;
  if (!(rad < p.boxmin)) { if (!offs) out[k] = 0; return; }
  int base = offs ? offs[k] : 0;
  for_each_candidate<NDIM>(a, g, xi, rad, [&](int kk, const double* xj, const int* flip, int icode) {
    double dx[3] = {0.0, 0.0, 0.0};
    for (int d = 0; d < NDIM; d++) dx[d] = xi[d] - xj[d];
    const double rij2 = rij2_exact<NDIM>(dx);
    const double q2i = mul_exact(rij2, hi21);
    const double hj1 = a.P4[kk].w;
    const double q2j = mul_exact(mul_exact(rij2, hj1), hj1);
    const bool in = sym ? (q2i < R2 || q2j < R2) : (q2i < R2);
    if (!in) return;
    if (offs) {
      int code = 0;
      if (icode) {
        const double4 pj = a.P4[kk];
        const double xo[3] = {pj.x, pj.y, pj.z};
        for (int d = NDIM - 1; d >= 0; d--) {
          code *= 3;
          if (xj[d] > xo[d]) code += 1;
          if (xj[d] < xo[d]) code += 2;
        }
      }
      out[3 * (base + cnt) + 0] = a.uid[i];
      out[3 * (base + cnt) + 1] = a.uid[a.sidx[kk]];
      out[3 * (base + cnt) + 2] = code;
    }
    cnt++;
  });
  if (!offs) out[k] = cnt;
}

// ===========================================================================
// h-predictor features and targets (MLP seed of the h-rho iteration)
// ===========================================================================

// after the h-rho iteration: target log(h_converged / h_predictor)
__global__ void k_mlp_target(KArgs a, float* __restrict__ X) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.n) return;
  float* x = X + (size_t)i * MLP_LDX;
  const double corr = log(a.f[F_H][i] / (double)x[MLP_COL_HPRED]);
  x[MLP_COL_TARGET] = (float)(MLP_TSCALE * corr);
  a.f[F_MLPCORR][i] = corr;
}

// ===========================================================================
// host side
// ===========================================================================

// kernel tables, src/kernelND.f90:127-161 (set-up), 211-269 (M6 quintic),
// 4143-4190 (M4 cubic), 4279-4289 (normalisation)
// ===========================================================================
// launches over a batch: one launch per phase covers every member of a
// group (same dimension, physics and precision).  Workgroup b belongs to the
// member m with off[m] <= b < off[m+1] and plays block b - off[m] of the
// member's own launch, so each member computes exactly what its own launch
// would.
// ===========================================================================

struct MArgs {
  KArgs a;
  Phys p;
  Grid* gdev;
  Grid box;
  double* part;   // binning partials
  int* bsum;      // scan block sums
  int* list_a;
  int* list_b;
  DScal* out;     // slot of this member in the batch's scalar array
  double bconstmax;
  int cap_bits;
  int reflect;
  int nbp;        // binning-statistics blocks
  int nb_nr, nb_force, nb_256;  // blocks of the h-rho, force and elementwise launches
  double* scratch_d;  // REORDER_W fields of n doubles (or the five int fields) for reorders
  int* scratch_i;
};

__device__ __forceinline__ int member_of(const int* __restrict__ off, int nm, int b) {
  int lo = 0, hi = nm - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (off[mid] <= b) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

#define NDSP_MEMBER                                        \
  const int m_ = member_of(off, nm, (int)blockIdx.x);      \
  const MArgs& M = ma[m_];                                 \
  const VL vl{(int)blockIdx.x - off[m_], off[m_ + 1] - off[m_]};
#define NDSP_MK(name) __global__ void name(const MArgs* __restrict__ ma, const int* __restrict__ off, int nm)

NDSP_MK(km_predict) { NDSP_MEMBER b_predict(M.a, M.p, vl); }
NDSP_MK(km_predict_fixed) { NDSP_MEMBER b_predict_fixed(M.a, M.p, vl); }
NDSP_MK(km_fixed_copy) { NDSP_MEMBER b_fixed_copy(M.a, M.p, vl); }
NDSP_MK(km_slot_of) { NDSP_MEMBER b_slot_of(M.a, vl); }
NDSP_MK(km_correct) { NDSP_MEMBER b_correct(M.a, M.p, vl); }
NDSP_MK(km_boundary_step) { NDSP_MEMBER b_boundary(M.a, M.p, M.box, 1, vl); }
NDSP_MK(km_err_reset) { NDSP_MEMBER if (threadIdx.x == 0) M.a.sc->err = 0; }
NDSP_MK(km_binstats) { NDSP_MEMBER b_binstats(M.a, M.p.ndim, M.part, vl); }
NDSP_MK(km_binstats_final) { NDSP_MEMBER b_binstats_final(M.a.sc, M.part, M.nbp, vl); }
NDSP_MK(km_grid_setup) { NDSP_MEMBER b_grid_setup(M.a.sc, M.gdev, M.box, M.p, M.cap_bits, vl); }
NDSP_MK(km_fill_cells) { NDSP_MEMBER b_fill_cells(M.a.cfill, M.gdev, vl); }
NDSP_MK(km_cellid) { NDSP_MEMBER b_cellid(M.a, M.gdev, vl); }
NDSP_MK(km_scan_local) { NDSP_MEMBER b_scan_local(M.a.cfill, M.a.cstart, M.bsum, M.gdev, vl); }
NDSP_MK(km_scan_bsum) { NDSP_MEMBER b_scan_bsum(M.bsum, M.gdev, nullptr, vl); }
NDSP_MK(km_scan_add) { NDSP_MEMBER b_scan_add(M.a.cstart, M.bsum, M.gdev, vl); }
NDSP_MK(km_scatter) { NDSP_MEMBER b_scatter(M.a, vl); }
NDSP_MK(km_cellsort) { NDSP_MEMBER b_cellsort_wave(M.a, M.gdev, vl); }
NDSP_MK(km_fill_nbx) { NDSP_MEMBER b_fill_i(M.a.nbx, M.a.n, 0, vl); }
NDSP_MK(km_nblist_sort) { NDSP_MEMBER b_nblist_sort(M.a, vl); }
NDSP_MK(km_init_density) { NDSP_MEMBER b_init_density_scalars(M.a.sc, M.bconstmax, M.a.n, vl); }
NDSP_MK(km_pass_ctl) { NDSP_MEMBER b_pass_ctl(M.a.sc, M.a.n, M.reflect, M.p.maxits, vl); }
NDSP_MK(km_hmax) { NDSP_MEMBER b_hmax(M.a, M.a.sc, vl); }
NDSP_MK(km_refresh_hinv) { NDSP_MEMBER b_refresh_hinv(M.a, M.p, vl); }
NDSP_MK(km_init_force) { NDSP_MEMBER b_init_force_scalars(M.a.sc, vl); }
NDSP_MK(km_finalize) { NDSP_MEMBER b_finalize(M.a, M.p, vl); }
NDSP_MK(km_dt) { NDSP_MEMBER b_dt(M.a.sc, M.p, vl); }
NDSP_MK(km_step_end) { NDSP_MEMBER b_step_end(M.a.sc, vl); }
// every member's scalars into the batch's contiguous array (one copy to the host)
NDSP_MK(km_gather_scalars) {
  NDSP_MEMBER
  const int* src = (const int*)M.a.sc;
  int* dst = (int*)M.out;
  for (int w = threadIdx.x; w < (int)(sizeof(DScal) / sizeof(int)); w += blockDim.x) dst[w] = src[w];
}
// a block of fields of the periodic reorder (slot k <- sorted position k) of
// the members flagged in due: phase 0 gathers fields f0 .. f0+REORDER_W-1
// (up to F_NPERSIST) into the member's scratch, phase 1 copies them back;
// f0 = F_NPERSIST takes the ids and counters (five int fields, held in the
// same scratch)
__global__ void km_reorder(const MArgs* __restrict__ ma, const int* __restrict__ off, int nm,
                           const int* __restrict__ due, int f0, int phase) {
  NDSP_MEMBER
  if (!due[m_]) return;
  const int k = vl.b * blockDim.x + threadIdx.x;
  const int n = M.a.n;
  if (k >= n) return;
  if (f0 < F_NPERSIST) {
    const int nf = min(REORDER_W, F_NPERSIST - f0);
    const int src = phase == 0 ? M.a.sidx[k] : 0;
    for (int q = 0; q < nf; q++) {
      if (phase == 0) M.scratch_d[(size_t)q * n + k] = M.a.f[f0 + q][src];
      else M.a.f[f0 + q][k] = M.scratch_d[(size_t)q * n + k];
    }
  } else {
    int* const iv[5] = {M.a.uid, M.a.numneigh, M.a.nits, M.a.itype, M.a.ireal};
    int* si = (int*)M.scratch_d;
    const int src = phase == 0 ? M.a.sidx[k] : 0;
    for (int q = 0; q < 5; q++) {
      if (phase == 0) si[(size_t)q * n + k] = iv[q][src];
      else iv[q][k] = si[(size_t)q * n + k];
    }
  }
}
// reductions: which[0] = h-rho pass (skipped once finished), 1 = force, 2 = finalize
__global__ void km_red_final(const MArgs* __restrict__ ma, const int* __restrict__ off, int nm, RedSpec spec,
                             int which) {
  NDSP_MEMBER
  const int nb = which == 0 ? M.nb_nr : which == 1 ? M.nb_force : M.nb_256;
  b_red_final(M.a.red, M.a.rcap, nb, spec, M.a.sc, which == 0 ? (const int*)&M.a.sc->pc_done : nullptr, vl);
}
template <typename PT>
NDSP_MK(km_pack_density) { NDSP_MEMBER b_pack_density<PT>(M.a, M.p, vl); }
template <typename PT>
NDSP_MK(km_pack_force) { NDSP_MEMBER b_pack_force<PT>(M.a, M.p, vl); }
template <int NDIM>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) void km_nblist(const MArgs* __restrict__ ma,
                                                                                         const int* __restrict__ off,
                                                                                         int nm) {
  NDSP_MEMBER b_nblist<NDIM>(M.a, M.p, M.gdev, vl);
}
template <int NDIM, typename PT, bool DUST, int TB>
__global__ __launch_bounds__(TB) void km_nr_pass(const MArgs* __restrict__ ma, const int* __restrict__ off, int nm) {
  NDSP_MEMBER b_nr_pass<NDIM, PT, DUST, TB>(M.a, M.p, M.gdev, M.list_a, M.list_b, vl);
}
template <int NDIM, typename PT, int PHYS, int TB>
__global__ __launch_bounds__(TB) void km_force(const MArgs* __restrict__ ma, const int* __restrict__ off, int nm) {
  NDSP_MEMBER b_force<NDIM, PT, PHYS, TB>(M.a, M.p, M.gdev, vl);
}

static int make_tables(int ikernel, int ndim, std::vector<double>& tab, double& radkern, double& dq2) {
  const int n = NDSP_IKERN + 1;
  tab.assign(3 * n, 0.0);
  double* w = tab.data();
  double* g = w + n;
  double* gg = g + n;
  const double pi = 3.141592653589;
  double cnormk;
  if (ikernel == 3) {
    radkern = 3.0;
    dq2 = radkern * radkern / (double)NDSP_IKERN;
    cnormk = ndim == 1 ? 1. / 120. : ndim == 2 ? 7. / (478 * pi) : 1. / (120. * pi);
    for (int i = 0; i < n; i++) {
      double q2 = i * dq2, q4 = q2 * q2, q = std::sqrt(q2);
      double term1 = -5. * std::pow(3. - q, 4.);
      if (q < 1.0) {
        w[i] = 66. - 60. * q2 + 30. * q4 - 10. * q4 * q;
        g[i] = term1 + 30 * std::pow(2. - q, 4) - 75. * std::pow(1. - q, 4);
        gg[i] = 20. * std::pow(3. - q, 3) - 120. * std::pow(2. - q, 3) + 300. * std::pow(1. - q, 3);
      } else if (q < 2.0) {
        w[i] = std::pow(3. - q, 5) - 6. * std::pow(2. - q, 5);
        g[i] = term1 + 30 * std::pow(2. - q, 4);
        gg[i] = 20. * std::pow(3. - q, 3) - 120. * std::pow(2. - q, 3);
      } else if (q <= 3.0) {
        w[i] = std::pow(3. - q, 5);
        g[i] = term1;
        gg[i] = 20. * std::pow(3. - q, 3);
      }
    }
  } else if (ikernel == 0) {
    radkern = 2.0;
    dq2 = radkern * radkern / (double)NDSP_IKERN;
    cnormk = ndim == 1 ? 0.66666666666 : ndim == 2 ? 10. / (7. * pi) : 1. / pi;
    for (int i = 0; i < n; i++) {
      double q2 = i * dq2, q = std::sqrt(q2);
      if (q < 1.0) {
        w[i] = 1. - 1.5 * q2 + 0.75 * q * q2;
        g[i] = -3. * q + 2.25 * q2;
        gg[i] = -3. + 4.5 * q;
      } else if (q <= 2.0) {
        double t = 2. - q;
        w[i] = 0.25 * (t * t * t);
        g[i] = -0.75 * (t * t);
        gg[i] = 1.5 * t;
      }
    }
  } else {
    return -1;
  }
  for (int i = 0; i < 3 * n; i++) tab[i] = cnormk * tab[i];
  return 0;
}

// expected neighbour count of the smoothing length rule h = hfact (m/rho)^(1/d)
static double nexp_of(const Phys& p) {
  const double r = p.radkern * p.hfact;
  if (p.ndim == 1) return 2.0 * r;
  if (p.ndim == 2) return M_PI * r * r;
  return 4.0 / 3.0 * M_PI * r * r * r;
}

struct ndsp_ctx {
  int device = 0;           // HIP device current at creation (holds every buffer)
  int tb_force = 256, tb_nr = 256;  // workgroup sizes of the pair kernels (autotuned)
  int nfixed = 0;           // fixed boundary particles uploaded
  bool ireal_set = true;    // their partners have been given
  bool reflect = false;     // some wall reflects
  Grid* gdev = nullptr;     // cell grid of the last binning (device)
  int cap_bits = 0;         // Morton key-space capacity of cstart/cfill
  int maxnb_limit = 0;      // largest list capacity the lists may grow to
  bool capturing = false;   // a step is being recorded into a graph
  bool use_graph = false;   // ndsp_step replays recorded graphs
  bool graph_valid = false;
  hipGraphExec_t graph_a = nullptr, graph_b = nullptr;
  // second set of the persistent fields: a reorder gathers into it and the
  // two sets trade places, each with its own recorded halves of a step
  // (members of a batch, whose argument tables hold the field addresses,
  // copy back instead)
  double* falt[F_NPERSIST] = {};
  bool falt_failed = false;
  bool in_batch = false;
  bool alt_graph_valid = false;
  hipGraphExec_t alt_graph_a = nullptr, alt_graph_b = nullptr;
  Phys p;
  int prec;
  int n;
  int ibound[3];
  double xmin[3], xmax[3];
  double bconstmax;
  hipStream_t st = nullptr;
  KArgs a;
  double* scratch_d = nullptr;  // REORDER_W fields of n doubles for reorders
  int* scratch_i = nullptr;
  int* bsum = nullptr;
  int* cmp_bsum = nullptr;      // block offsets of the compaction of unconverged flags
  int* list_a = nullptr;        // active lists of the h-rho passes
  int* list_b = nullptr;
  double* part = nullptr;       // reduction partials
  double* cons_out = nullptr;
  int nb_part = 0;
  Grid g;
  bool grid_valid = false;
  bool have_rates = false;
  int nsteps = 0;
  int reorder_every = 10;
  bool reorder_next = false;  // reorder at the next step (after a change of tile shapes)
  double dt = 0.0, time = 0.0;
  MlpState mlp;
  bool mlp_alloced = false;
  // timing
  bool timing = false;
  std::vector<std::string> tnames;
  std::vector<hipEvent_t> ev0, ev1;
  std::vector<double> tms;
  std::vector<int> tcount;
  std::vector<std::pair<int, int>> pending;  // (label, event pair index)
  int nev = 0;
};

// recorded steps hold field addresses and launch shapes: both sets' go stale
static void invalidate_graphs(ndsp_ctx* c) {
  c->graph_valid = false;
  c->alt_graph_valid = false;
}

// ---- timing helpers ------------------------------------------------------
static int label_id(ndsp_ctx* c, const char* name) {
  for (size_t k = 0; k < c->tnames.size(); k++)
    if (c->tnames[k] == name) return (int)k;
  c->tnames.push_back(name);
  c->tms.push_back(0.0);
  c->tcount.push_back(0);
  return (int)c->tnames.size() - 1;
}

struct TScope {
  ndsp_ctx* c;
  int lab, ev;
  TScope(ndsp_ctx* c_, const char* name) : c(c_), lab(-1), ev(-1) {
    if (!c->timing || c->capturing) return;
    lab = label_id(c, name);
    if (c->nev >= (int)c->ev0.size()) {
      hipEvent_t a0, a1;
      hipEventCreate(&a0);
      hipEventCreate(&a1);
      c->ev0.push_back(a0);
      c->ev1.push_back(a1);
    }
    ev = c->nev++;
    hipEventRecord(c->ev0[ev], c->st);
  }
  ~TScope() {
    if (ev < 0) return;
    hipEventRecord(c->ev1[ev], c->st);
    c->pending.push_back({lab, ev});
  }
};

static void timing_collect(ndsp_ctx* c) {
  if (!c->timing) return;
  for (auto& pe : c->pending) {
    float ms = 0.f;
    hipEventElapsedTime(&ms, c->ev0[pe.second], c->ev1[pe.second]);
    c->tms[pe.first] += ms;
    c->tcount[pe.first] += 1;
  }
  c->pending.clear();
  c->nev = 0;
}

static inline unsigned nblk(long long n, int bs = BS) { return (unsigned)((n + bs - 1) / bs); }

template <typename F>
static int dispatch_ndim(int ndim, F&& f) {
  if (ndim == 1) return f(std::integral_constant<int, 1>());
  if (ndim == 2) return f(std::integral_constant<int, 2>());
  return f(std::integral_constant<int, 3>());
}

static int sync_scalars(ndsp_ctx* c, DScal& h) {
  HIPCHK(hipMemcpyAsync(&h, c->a.sc, sizeof(DScal), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

static int check_err(ndsp_ctx* c, const DScal& h) {
  if (h.err) {
    static const char* names[] = {"ok", "h <= 0", "rho <= 0", "density not converged", "u < 0",
                                  "cells", "outside", "alloc", "coincident particles",
                                  "vsig determinant < 0"};
    int e = h.err;
    return set_err(e, std::string("device reported: ") + (e < 10 ? names[e] : "error"));
  }
  return 0;
}

// set_linklist: boundary, statistics, grid, counting sort, density view
// Morton key-space capacity (bits) for a grid of the current h: the key
// bits of every dimension (its cell count rounded up to a power of two) plus
// extra_bits each, at most 2^28 keys
static int size_key_space(ndsp_ctx* c, const DScal& hs, int extra_bits) {
  const double rmax = c->p.radkern * hs.hmax_bin;
  if (!(rmax > 0.0) || !std::isfinite(rmax)) return set_err(NDSP_ERR_CELLS, "link: max h <= 0");
  const double cs = rmax / c->p.cell_sub;
  int bits = 0;
  for (int d = 0; d < c->p.ndim; d++) {
    const double ext = c->ibound[d] >= 2 ? c->xmax[d] - c->xmin[d] : hs.bbox_hi[d] - hs.bbox_lo[d];
    const double ncd = std::min(std::floor(ext / cs) + 1.0, 1.0e8);
    int kb = 0;
    while ((double)(1 << kb) < ncd) kb++;
    bits += kb + extra_bits;
  }
  bits = std::min(28, bits);
  if (bits == c->cap_bits) return 0;
  const long long cap = (1LL << bits) + 1;
  if (c->a.cstart) hipFree(c->a.cstart);
  if (c->a.cfill) hipFree(c->a.cfill);
  if (c->bsum) hipFree(c->bsum);
  c->a.cstart = c->a.cfill = c->bsum = nullptr;
  HIPCHK(hipMalloc(&c->a.cstart, sizeof(int) * cap));
  HIPCHK(hipMalloc(&c->a.cfill, sizeof(int) * cap));
  HIPCHK(hipMalloc(&c->bsum, sizeof(int) * (cap / (BS * 4) + 2)));
  c->cap_bits = bits;
  invalidate_graphs(c);
  return 0;
}

// List capacity after a binning in which some rows overflowed (those
// particles scanned cells, the slowest part of their launch): 1.5 times the
// present one, up to four times the initial capacity; the old lists are
// dropped (the next binning rebuilds them) and recorded steps are recorded
// again.  A failed allocation keeps the present capacity for good.
static int grow_lists(ndsp_ctx* c) {
  if (!c->p.use_list || c->a.maxnb >= c->maxnb_limit) return 0;
  const int m = std::min(c->maxnb_limit, (c->a.maxnb * 3 / 2 + 31) / 32 * 32);
  unsigned* nbl = nullptr;
  if (hipMalloc(&nbl, sizeof(unsigned) * (size_t)c->n * (size_t)m) != hipSuccess) {
    (void)hipGetLastError();
    c->maxnb_limit = c->a.maxnb;
    return 0;
  }
  HIPCHK(hipStreamSynchronize(c->st));
  HIPCHK(hipFree(c->a.nbl));
  c->a.nbl = nbl;
  c->a.maxnb = m;
  invalidate_graphs(c);
  return 0;
}

// Binning (set_linklist, src/linkND.f90:45): wrap, statistics, device-side
// grid, counting sort by Morton cell key, packed views, neighbour lists.
// No host round trip except the first call, which sizes the key space.
static int do_linklist(ndsp_ctx* c, int in_step) {
  KArgs& a = c->a;
  const int n = c->n;
  if (!c->ireal_set) return set_err(NDSP_ERR_BAD_ARG, "fixed particles need their partners: call ndsp_set_ireal");
  hipLaunchKernelGGL(k_fill_i, dim3(1), dim3(64), 0, c->st, &a.sc->err, 1, 0);
  {
    TScope t(c, "boundary");
    hipLaunchKernelGGL(k_boundary, dim3(nblk(n)), dim3(BS), 0, c->st, a, c->p, c->g, in_step);
  }
  {
    TScope t(c, "binstats");
    hipLaunchKernelGGL(k_binstats, dim3(c->nb_part), dim3(BS), 0, c->st, a, c->p.ndim, c->part);
    hipLaunchKernelGGL(k_binstats_final, dim3(1), dim3(BS), 0, c->st, a.sc, c->part, c->nb_part);
  }
  if (c->cap_bits == 0) {
    if (c->capturing) return set_err(NDSP_ERR_BAD_ARG, "capture before the first binning");
    DScal hs;
    if (int e = sync_scalars(c, hs)) return e;
    if (int e = size_key_space(c, hs, 1)) return e;
  }
  const long long cap = (1LL << c->cap_bits) + 1;
  const unsigned cap_blocks = (unsigned)std::min<long long>(nblk(cap), 1 << 20);
  {
    TScope t(c, "cellsort");
    hipLaunchKernelGGL(k_grid_setup, dim3(1), dim3(64), 0, c->st, a.sc, c->gdev, c->g, c->p, c->cap_bits);
    hipLaunchKernelGGL(k_fill_cells, dim3(std::min(cap_blocks, 4096u)), dim3(BS), 0, c->st, a.cfill,
                       (const Grid*)c->gdev);
    hipLaunchKernelGGL(k_cellid, dim3(nblk(n)), dim3(BS), 0, c->st, a, (const Grid*)c->gdev);
    const unsigned nbs = (unsigned)((cap + BS * 4 - 1) / (BS * 4));
    hipLaunchKernelGGL(k_scan_local, dim3(nbs), dim3(BS), 0, c->st, (const int*)a.cfill, a.cstart, c->bsum,
                       (const Grid*)c->gdev);
    hipLaunchKernelGGL(k_scan_bsum, dim3(1), dim3(BS), 0, c->st, c->bsum, (const Grid*)c->gdev, (int*)nullptr);
    hipLaunchKernelGGL(k_scan_add, dim3(cap_blocks), dim3(BS), 0, c->st, a.cstart, (const int*)c->bsum,
                       (const Grid*)c->gdev);
    hipLaunchKernelGGL(k_fill_cells, dim3(std::min(cap_blocks, 4096u)), dim3(BS), 0, c->st, a.cfill,
                       (const Grid*)c->gdev);
    hipLaunchKernelGGL(k_scatter, dim3(nblk(n)), dim3(BS), 0, c->st, a);
    hipLaunchKernelGGL(k_cellsort_wave, dim3(cap_blocks), dim3(BS), 0, c->st, a, (const Grid*)c->gdev);
  }
  {
    TScope t(c, "pack_density");
    if (c->prec == NDSP_PREC_FP64)
      hipLaunchKernelGGL(k_pack_density<double>, dim3(nblk(n)), dim3(BS), 0, c->st, a, c->p);
    else
      hipLaunchKernelGGL(k_pack_density<float>, dim3(nblk(n)), dim3(BS), 0, c->st, a, c->p);
  }
  if (c->p.use_list) {
    hipLaunchKernelGGL(k_fill_i, dim3(std::min(nblk(n), 4096u)), dim3(BS), 0, c->st, a.nbx, n, 0);
    {
      TScope t(c, "nblist");
      dispatch_ndim(c->p.ndim, [&](auto nd) {
        constexpr int ND = decltype(nd)::value;
        hipLaunchKernelGGL((k_nblist<ND>), dim3(nblk(n, 64)), dim3(64), 0, c->st, a, c->p, (const Grid*)c->gdev);
        return 0;
      });
    }
    TScope t(c, "nblist_sort");
    hipLaunchKernelGGL(k_nblist_sort, dim3(nblk(n)), dim3(BS), 0, c->st, a);
  }
  HIPCHK(hipGetLastError());
  c->grid_valid = true;
  return 0;
}

// K pass launches of the h-rho iteration; each reads its parameters from the
// device-side control and does nothing once the iteration has finished
static void launch_passes(ndsp_ctx* c, int K) {
  KArgs& a = c->a;
  const int n = c->n;
  const bool reflect = c->reflect;
  for (int q = 0; q < K; q++) {
    dispatch_ndim(c->p.ndim, [&](auto nd) {
      constexpr int ND = decltype(nd)::value;
      const bool fp64 = c->prec == NDSP_PREC_FP64, dust = c->p.idust != 0;
      const bool small = c->tb_nr == 64;
      auto kern = fp64 ? (dust ? (small ? k_nr_pass<ND, double, true, 64> : k_nr_pass<ND, double, true, 256>)
                               : (small ? k_nr_pass<ND, double, false, 64> : k_nr_pass<ND, double, false, 256>))
                       : (dust ? (small ? k_nr_pass<ND, float, true, 64> : k_nr_pass<ND, float, true, 256>)
                               : (small ? k_nr_pass<ND, float, false, 64> : k_nr_pass<ND, float, false, 256>));
      hipLaunchKernelGGL(kern, dim3(nblk(n, c->tb_nr)), dim3(c->tb_nr), 0, c->st, a, c->p, (const Grid*)c->gdev,
                         c->list_a, c->list_b);
      return 0;
    });
    {
      const unsigned nb = (unsigned)((n + BS * CMP_ITEMS - 1) / (BS * CMP_ITEMS));
      hipLaunchKernelGGL(k_cmp_count, dim3(nb), dim3(BS), 0, c->st, (const int*)a.nrflag, n, c->cmp_bsum);
      hipLaunchKernelGGL(k_cmp_scan, dim3(1), dim3(BS), 0, c->st, c->cmp_bsum, (int)nb, a.sc);
      hipLaunchKernelGGL(k_cmp_scatter, dim3(nb), dim3(BS), 0, c->st, a.nrflag, n, (const int*)c->cmp_bsum,
                         (const DScal*)a.sc, c->list_a, c->list_b);
    }
    hipLaunchKernelGGL(k_red_final, dim3(1), dim3(BS), 0, c->st, (const double*)a.red, a.rcap, (int)nblk(n, c->tb_nr),
                       red_spec({RED_Q(RED_MAX, RED_F64, stressmax), RED_Q(RED_MAX, RED_F64, hmax_nr),
                                 RED_Q(RED_MAX, RED_I32, itsmax), RED_Q(RED_MIN, RED_I32, nmin),
                                 RED_Q(RED_MAX, RED_I32, nmax), RED_Q(RED_SUM, RED_U64, pairs_density),
                                 RED_Q(RED_SUM, RED_U64, ncalc_total)}),
                       a.sc, (const int*)&a.sc->pc_done);
    hipLaunchKernelGGL(k_pass_ctl, dim3(1), dim3(64), 0, c->st, a.sc, n, reflect ? 1 : 0, c->p.maxits);
    hipLaunchKernelGGL(k_hmax, dim3(std::min(1024u, nblk(n))), dim3(BS), 0, c->st, a, a.sc);
    hipLaunchKernelGGL(k_refresh_hinv, dim3(nblk(n)), dim3(BS), 0, c->st, a, c->p);
  }
}

// Passes launched without a host decision in between: synchronous symmetric
// passes while every particle iterates, then (no reflecting walls) one
// persistent pass for the rest.  A recorded step holds the first chunk only;
// density_finish completes the iteration after it when needed.
static int pass_chunk(const ndsp_ctx* c) { return c->reflect ? 4 : 3; }

static int do_density(ndsp_ctx* c) {
  KArgs& a = c->a;
  hipLaunchKernelGGL(k_init_density_scalars, dim3(1), dim3(1), 0, c->st, a.sc, c->bconstmax, c->n);
  TScope t(c, "density");
  launch_passes(c, pass_chunk(c));
  HIPCHK(hipGetLastError());
  return 0;
}

// until the device control reports the iteration finished: more passes
static int density_finish(ndsp_ctx* c) {
  for (int guard = 0; guard < 1000; guard++) {
    DScal hs;
    if (int e = sync_scalars(c, hs)) return e;
    if (hs.pc_done) return 0;
    TScope t(c, "density");
    launch_passes(c, pass_chunk(c));
    HIPCHK(hipGetLastError());
  }
  return set_err(NDSP_ERR_NOT_CONVERGED, "h-rho iteration did not finish");
}

// with_correct: the step's corrector runs in the finalize launch (each
// particle's final rates, then its corrected state)
static int do_rates(ndsp_ctx* c, bool with_correct = false) {
  KArgs& a = c->a;
  const int n = c->n;
  if (c->p.fixed_on) hipLaunchKernelGGL(k_fixed_copy, dim3(nblk(n)), dim3(BS), 0, c->st, a, c->p);
  {
    TScope t(c, "pack_force");
    if (c->prec == NDSP_PREC_FP64)
      hipLaunchKernelGGL(k_pack_force<double>, dim3(nblk(n)), dim3(BS), 0, c->st, a, c->p);
    else
      hipLaunchKernelGGL(k_pack_force<float>, dim3(nblk(n)), dim3(BS), 0, c->st, a, c->p);
  }
  hipLaunchKernelGGL(k_init_force_scalars, dim3(1), dim3(1), 0, c->st, a.sc);
  {
    TScope t(c, "force");
    dispatch_ndim(c->p.ndim, [&](auto nd) {
      constexpr int ND = decltype(nd)::value;
      const int phys = c->p.imhd ? 1 : c->p.idust ? 2 : 0;
      void (*kern)(KArgs, Phys, const Grid* __restrict__);
      const bool small = c->tb_force == 64;
      if (c->prec == NDSP_PREC_FP64) {
        if (small) kern = phys == 1 ? k_force<ND, double, 1, 64> : phys == 2 ? k_force<ND, double, 2, 64> : k_force<ND, double, 0, 64>;
        else kern = phys == 1 ? k_force<ND, double, 1, 256> : phys == 2 ? k_force<ND, double, 2, 256> : k_force<ND, double, 0, 256>;
      } else {
        if (small) kern = phys == 1 ? k_force<ND, float, 1, 64> : phys == 2 ? k_force<ND, float, 2, 64> : k_force<ND, float, 0, 64>;
        else kern = phys == 1 ? k_force<ND, float, 1, 256> : phys == 2 ? k_force<ND, float, 2, 256> : k_force<ND, float, 0, 256>;
      }
      hipLaunchKernelGGL(kern, dim3(nblk(n, c->tb_force)), dim3(c->tb_force), 0, c->st, a, c->p, (const Grid*)c->gdev);
      return 0;
    });
    hipLaunchKernelGGL(k_red_final, dim3(1), dim3(BS), 0, c->st, (const double*)a.red, a.rcap, (int)nblk(n, c->tb_force),
                       red_spec({RED_Q(RED_MAX, RED_F64, vsigmax), RED_Q(RED_MIN, RED_F64, dtcourant),
                                 RED_Q(RED_SUM, RED_U64, pairs_force)}),
                       a.sc, (const int*)nullptr);
  }
  {
    TScope t(c, "finalize");
    if (with_correct) hipLaunchKernelGGL(k_finalize_correct, dim3(nblk(n)), dim3(BS), 0, c->st, a, c->p);
    else hipLaunchKernelGGL(k_finalize, dim3(nblk(n)), dim3(BS), 0, c->st, a, c->p);
    hipLaunchKernelGGL(k_red_final, dim3(1), dim3(BS), 0, c->st, (const double*)a.red, a.rcap, (int)nblk(n),
                       red_spec({RED_Q(RED_MAX, RED_F64, fhmax), RED_Q(RED_MIN, RED_F64, dtdrag)}), a.sc,
                       (const int*)nullptr);
    hipLaunchKernelGGL(k_dt, dim3(1), dim3(1), 0, c->st, a.sc, c->p);
  }
  HIPCHK(hipGetLastError());
  c->have_rates = true;
  return 0;
}

static void fill_outputs(ndsp_ctx* c, const DScal& h, ndsp_stats* stats, ndsp_dt* dt) {
  if (stats) {
    stats->itsdensity = h.itsmax;
    stats->nmin = h.nmin;
    stats->nmax = h.nmax;
    stats->ncalc_mean = (double)h.ncalc_total / (double)c->n;
    stats->pairs_density = (int64_t)h.pairs_density;
    stats->pairs_force = (int64_t)(h.pairs_force / 2);
  }
  if (dt) {
    dt->dt = h.dt;
    dt->dtcourant = h.dtcourant;
    dt->dtforce = h.dtforce;
    dt->dtdrag = h.dtdrag;
    dt->vsig2max = (c->p.imhd && c->p.idivbzero >= 2) ? h.vsigmax * h.vsigmax : 0.0;
    dt->hhmax = h.hmax_nr;
    dt->stressmax = h.stressmax;
    dt->time = c->time;
  }
}

// Permutes the persistent per-particle state into the cell order of the last
// binning (slot k <- sorted position k), so that per-particle accesses of the
// pair kernels stay spatially coherent.  Each field is gathered into scratch
// and copied back: the field pointers never change, so launch parameters
// recorded earlier (a captured step graph) stay valid.
static int reorder(ndsp_ctx* c) {
  KArgs& a = c->a;
  const int n = c->n;
  TScope t(c, "reorder");
  if (!c->in_batch && !c->falt_failed && !c->falt[0]) {
    bool ok = true;
    for (int fi = 0; fi < F_NPERSIST && ok; fi++) ok = hipMalloc(&c->falt[fi], sizeof(double) * n) == hipSuccess;
    if (!ok) {
      (void)hipGetLastError();
      for (int fi = 0; fi < F_NPERSIST; fi++)
        if (c->falt[fi]) {
          (void)hipFree(c->falt[fi]);
          c->falt[fi] = nullptr;
        }
      c->falt_failed = true;
    }
  }
  const bool swap_sets = !c->in_batch && c->falt[0] != nullptr;
  for (int f0 = 0; swap_sets && f0 < F_NPERSIST; f0 += REORDER_W) {
    FieldPairs fp{};
    fp.nf = std::min(REORDER_W, (int)F_NPERSIST - f0);
    for (int q = 0; q < fp.nf; q++) {
      fp.src[q] = a.f[f0 + q];
      fp.dst[q] = c->falt[f0 + q];
    }
    hipLaunchKernelGGL(k_gather_pairs, dim3(nblk(n)), dim3(BS), 0, c->st, fp, (const int*)a.sidx, n);
  }
  if (swap_sets) {
    for (int fi = 0; fi < F_NPERSIST; fi++) std::swap(a.f[fi], c->falt[fi]);
    std::swap(c->graph_a, c->alt_graph_a);
    std::swap(c->graph_b, c->alt_graph_b);
    std::swap(c->graph_valid, c->alt_graph_valid);
  }
  for (int f0 = 0; !swap_sets && f0 < F_NPERSIST; f0 += REORDER_W) {
    FieldSet<double> fs{};
    fs.nf = std::min(REORDER_W, F_NPERSIST - f0);
    for (int q = 0; q < fs.nf; q++) fs.f[q] = a.f[f0 + q];
    hipLaunchKernelGGL(k_gather_set<double>, dim3(nblk(n)), dim3(BS), 0, c->st, fs, c->scratch_d,
                       (const int*)a.sidx, n);
    hipLaunchKernelGGL(k_copy_set<double>, dim3(nblk(n)), dim3(BS), 0, c->st, fs, (const double*)c->scratch_d, n);
  }
  FieldSet<int> is{};
  is.nf = 5;
  is.f[0] = a.uid;
  is.f[1] = a.numneigh;
  is.f[2] = a.nits;
  is.f[3] = a.itype;
  is.f[4] = a.ireal;
  int* si = (int*)c->scratch_d;
  hipLaunchKernelGGL(k_gather_set<int>, dim3(nblk(n)), dim3(BS), 0, c->st, is, si, (const int*)a.sidx, n);
  hipLaunchKernelGGL(k_copy_set<int>, dim3(nblk(n)), dim3(BS), 0, c->st, is, (const int*)si, n);
  hipLaunchKernelGGL(k_slot_of, dim3(nblk(n)), dim3(BS), 0, c->st, a);
  HIPCHK(hipGetLastError());
  c->grid_valid = false;
  return 0;
}

// ---------------------------------------------------------------------------
// C-ABI
// ---------------------------------------------------------------------------

extern "C" {

const char* ndsp_last_error(void) { return g_err.c_str(); }

// ---------------------------------------------------------------------------
// Fixed device staging of the allocation-free MLP inference
// (ndsp_mlp_forward_noalloc): NOALLOC_MAX_ROWS feature rows of MLP_NIN
// floats, then one output float per row.  A file-scope device array exists
// once per GPU (each holds its own copy until the process ends); its address
// on a device is taken once, with that device current, when the first
// context is created there, and only that address is used on the host.
// ---------------------------------------------------------------------------
#define NOALLOC_MAX_ROWS (1 << 16)
#define NOALLOC_MAX_DEVICES 64
static __device__ __align__(256) float scratch[(size_t)NOALLOC_MAX_ROWS * (MLP_NIN + 1)];
static float* g_scratch_ptr[NOALLOC_MAX_DEVICES];
static std::mutex g_scratch_mu;

static int noalloc_scratch_init(int* device) {
  int dev = 0;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return set_err(NDSP_ERR_NO_DEVICE, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return set_err(NDSP_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = static_cast<float*>(ptr);
  }
  *device = dev;
  return 0;
}

int ndsp_create(const int32_t* iopt, const double* dopt, int64_t npart, int32_t precision, ndsp_ctx** out) {
  if (!iopt || !dopt || !out || npart < 1 || npart > 1000000000LL)
    return set_err(NDSP_ERR_BAD_ARG, "ndsp_create: bad arguments");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return set_err(NDSP_ERR_NO_DEVICE, "no HIP device");
  Phys p{};
  p.ikernel = iopt[0]; p.iener = iopt[1]; p.imhd = iopt[2];
  const int imagforce = iopt[3];
  p.idivbzero = iopt[4];
  const int iav = iopt[5];
  for (int k = 0; k < 3; k++) p.iavlim[k] = iopt[6 + k];
  p.idust = iopt[9];
  const int idrag = iopt[10];
  int ib[3] = {iopt[11], iopt[12], iopt[13]};
  p.maxits = iopt[14];
  const int idustevol = iopt[15];
  p.ndim = iopt[16];
  p.hfact = dopt[0]; p.tolh = dopt[1]; p.gamma = dopt[2]; p.polyk = dopt[3]; p.Kdrag = dopt[4];
  p.alphamin = dopt[5]; p.alphaumin = dopt[6]; p.alphabmin = dopt[7]; p.beta = dopt[8];
  p.avdecayconst = dopt[9]; p.psidecayfact = dopt[10]; p.C_cour = dopt[11]; p.C_force = dopt[12];
  p.psep = dopt[19];
  std::string bad;
  if (p.ndim < 1 || p.ndim > 3) bad = "ndim must be 1, 2 or 3";
  else if (p.ikernel != 0 && p.ikernel != 3) bad = "only ikernel 0 (cubic) and 3 (quintic)";
  else if (p.iener != 0 && p.iener != 2) bad = "only iener 0 and 2";
  else if (p.imhd != 0 && p.imhd != 11) bad = "only imhd 0 and 11";
  else if (p.imhd && imagforce != 2) bad = "only imagforce 2";
  else if (p.idivbzero != 0 && p.idivbzero != 2) bad = "only idivbzero 0 and 2";
  else if (iav != 2) bad = "only iav 2";
  else if (p.iavlim[0] < 0 || p.iavlim[0] > 2) bad = "iavlim(1) must be 0, 1 or 2";
  else if (p.iavlim[1] < 0 || p.iavlim[1] > 1) bad = "iavlim(2) must be 0 or 1";
  else if (!(p.iavlim[2] == 0 || p.iavlim[2] == 1 || p.iavlim[2] == 3)) bad = "iavlim(3) must be 0, 1 or 3";
  else if (p.idust != 0 && p.idust != 4) bad = "only idust 0 and 4";
  else if (p.idust == 4 && (idrag != 2 || idustevol != 0)) bad = "dust: only idrag_nature 2, idustevol 0";
  else if (p.idust == 4 && p.imhd) bad = "dust with MHD not supported";
  else if (dopt[24] != 0.0) bad = "damp must be 0";
  else if (p.maxits < 0) bad = "maxdensits must be >= 0";
  else if (precision != NDSP_PREC_MIXED && precision != NDSP_PREC_FP64) bad = "unknown precision";
  int nrefl = 0;
  for (int d = 0; d < p.ndim && bad.empty(); d++) {
    if (!(ib[d] == 0 || ib[d] == 1 || ib[d] == 2 || ib[d] == 3)) bad = "ibound must be 0, 1, 2 or 3";
    if (ib[d] == 1) p.fixed_on = 1;
    if (ib[d] == 2) nrefl++;
    if (ib[d] >= 2 && !(dopt[16 + d] > dopt[13 + d])) bad = "xmax must exceed xmin";
  }
  if (nrefl > 1 && bad.empty()) bad = "reflecting walls in at most one dimension";
  // neighbour lists need ghost extents that do not change between the list
  // build and its use, which holds unless a wall reflects
  // (a list entry packs the sorted index in 25 bits and the image code in 7)
  p.use_list = (nrefl == 0 && npart < (int64_t(1) << 25)) ? 1 : 0;
  if (iopt[17] != 0) p.use_list = 0;
  p.skin = 0.04;
  p.cell_sub = 2;
  if (!bad.empty()) return set_err(NDSP_ERR_BAD_OPTION, bad);
  p.dndim = 1.0 / p.ndim;
  p.boxmin = DBL_MAX;
  for (int d = 0; d < p.ndim; d++)
    if (ib[d] >= 2) p.boxmin = std::min(p.boxmin, dopt[16 + d] - dopt[13 + d]);
  p.avfact = std::fabs(p.gamma - 1.) > 1.e-3 ? std::log(4.) / (std::log((p.gamma + 1.) / (p.gamma - 1.))) : 1.0;
  std::vector<double> tab;
  double dq2 = 0.0;
  make_tables(p.ikernel, p.ndim, tab, p.radkern, dq2);
  p.radkern2 = p.radkern * p.radkern;
  p.dq2table = dq2;
  p.ddq2table = 1.0 / dq2;

  int device = 0;
  if (int e = noalloc_scratch_init(&device)) return e;
  ndsp_ctx* c = new ndsp_ctx();
  c->device = device;
  c->p = p;
  c->prec = precision;
  c->n = (int)npart;
  c->bconstmax = std::max(std::max(dopt[20], dopt[21]), dopt[22]);
  for (int d = 0; d < 3; d++) {
    c->ibound[d] = d < p.ndim ? ib[d] : 0;
    c->xmin[d] = dopt[13 + d];
    c->xmax[d] = dopt[16 + d];
    // a wall of fixed particles (ibound 1) is open space to the neighbour search
    c->g.bound[d] = c->ibound[d] == 1 ? 0 : c->ibound[d];
    c->g.xmin[d] = c->xmin[d];
    c->g.xmax[d] = c->xmax[d];
    c->g.nc[d] = 1;
  }
  c->g.ndim = p.ndim;
  c->reflect = nrefl > 0;
  c->dt = dopt[23];
  c->time = 0.0;
  auto cleanup_fail = [&](int e) { ndsp_destroy(c); return e; };
  if (hipStreamCreate(&c->st) != hipSuccess) return cleanup_fail(set_err(NDSP_ERR_HIP, "stream"));
  KArgs& a = c->a;
  memset(&a, 0, sizeof(a));
  const size_t n = (size_t)npart;
  for (int fi = 0; fi < F_NFIELDS; fi++) {
    if (hipMalloc(&a.f[fi], sizeof(double) * n) != hipSuccess) return cleanup_fail(set_err(NDSP_ERR_ALLOC, "hipMalloc fields"));
    hipMemset(a.f[fi], 0, sizeof(double) * n);
  }
  bool ok = true;
  ok &= hipMalloc(&c->scratch_d, sizeof(double) * n * REORDER_W) == hipSuccess;
  ok &= hipMalloc(&c->scratch_i, sizeof(int) * n) == hipSuccess;
  ok &= hipMalloc(&a.itype, sizeof(int) * n) == hipSuccess;
  ok &= hipMalloc(&a.ireal, sizeof(int) * n) == hipSuccess;
  ok &= hipMalloc(&a.slot_of, sizeof(int) * n) == hipSuccess;
  ok &= hipMalloc(&a.numneigh, sizeof(int) * n) == hipSuccess;
  ok &= hipMalloc(&a.nits, sizeof(int) * n) == hipSuccess;
  ok &= hipMalloc(&a.uid, sizeof(int) * n) == hipSuccess;
  ok &= hipMalloc(&a.cellid, sizeof(int) * n) == hipSuccess;
  ok &= hipMalloc(&a.sidx, sizeof(int) * n) == hipSuccess;
  ok &= hipMalloc(&c->list_a, sizeof(int) * n) == hipSuccess;
  ok &= hipMalloc(&c->list_b, sizeof(int) * n) == hipSuccess;
  ok &= hipMalloc(&a.nrflag, sizeof(int) * n) == hipSuccess;
  ok &= hipMalloc(&c->cmp_bsum, sizeof(int) * ((size_t)n / (BS * CMP_ITEMS) + 2)) == hipSuccess;
  if (ok) ok &= hipMemset(a.nrflag, 0, sizeof(int) * n) == hipSuccess;
  ok &= hipMalloc(&a.P4, sizeof(double4) * n) == hipSuccess;
  ok &= hipMalloc(&a.dxb, sizeof(double) * n) == hipSuccess;
  const size_t v4 = sizeof(double4) * n;
  const bool fp64 = c->prec == NDSP_PREC_FP64;
  a.dr_stride = fp64 ? dr_bytes<double>(c->p.idust != 0) : dr_bytes<float>(c->p.idust != 0);
  ok &= hipMalloc(&a.DR, (size_t)a.dr_stride * n) == hipSuccess;
  const size_t fr_rec = fp64 ? fr_bytes<double>(fr_phys(c->p)) : fr_bytes<float>(fr_phys(c->p));
  ok &= hipMalloc(&a.FR, fr_rec * n) == hipSuccess;
  if (c->p.fixed_on) ok &= hipMalloc(&a.GF, v4) == hipSuccess;
  ok &= hipMalloc((void**)&a.sc, sizeof(DScal)) == hipSuccess;
  a.rcap = (int)nblk(n, 64);
  ok &= hipMalloc((void**)&a.red, sizeof(double) * RED_MAXQ * (size_t)a.rcap) == hipSuccess;
  {
    // list capacity: 1.8x the neighbour count of a uniform lattice within
    // radkern*hfact*(1+skin) (a particle with more falls back to the scan)
    const double r = p.radkern * p.hfact * (1.0 + p.skin);
    const double vol = p.ndim == 3 ? 4.18879020478639 * r * r * r : p.ndim == 2 ? M_PI * r * r : 2.0 * r;
    a.maxnb = ((int)(1.8 * vol + 8.0) + 31) / 32 * 32;
    c->maxnb_limit = 4 * a.maxnb;
  }
  if (p.use_list) {
    ok &= hipMalloc(&a.nbl, sizeof(unsigned) * n * (size_t)a.maxnb) == hipSuccess;
    ok &= hipMalloc(&a.nbc, sizeof(int) * n) == hipSuccess;
    ok &= hipMalloc(&a.nbx, sizeof(int) * n) == hipSuccess;
    ok &= hipMalloc(&a.hcap, sizeof(double) * n) == hipSuccess;
  }
  float* tabf = nullptr;
  double* tabd = nullptr;
  ok &= hipMalloc(&tabf, sizeof(float) * tab.size()) == hipSuccess;
  ok &= hipMalloc(&tabd, sizeof(double) * tab.size()) == hipSuccess;
  c->nb_part = 1024;
  ok &= hipMalloc(&c->part, sizeof(double) * 32 * c->nb_part) == hipSuccess;
  ok &= hipMalloc(&c->cons_out, sizeof(double) * 32) == hipSuccess;
  ok &= hipMalloc((void**)&c->gdev, sizeof(Grid)) == hipSuccess;
  if (!ok) return cleanup_fail(set_err(NDSP_ERR_ALLOC, "hipMalloc buffers"));
  hipMemcpy(c->gdev, &c->g, sizeof(Grid), hipMemcpyHostToDevice);
  a.tabf = tabf;
  a.tabd = tabd;
  a.n = c->n;
  std::vector<float> tf(tab.begin(), tab.end());
  hipMemcpy(tabf, tf.data(), sizeof(float) * tf.size(), hipMemcpyHostToDevice);
  hipMemcpy(tabd, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice);
  DScal hs{};
  hs.dt = c->dt;
  hs.dt_cur = c->dt;
  hs.dtdrag = DBL_MAX;
  hs.dtcourant = 1.0e6;
  hipMemcpy(a.sc, &hs, sizeof(DScal), hipMemcpyHostToDevice);
  hipMemset(a.numneigh, 0, sizeof(int) * n);
  hipMemset(a.nits, 0, sizeof(int) * n);
  if (hipDeviceSynchronize() != hipSuccess) return cleanup_fail(set_err(NDSP_ERR_HIP, "init sync"));
  *out = c;
  return 0;
}

int ndsp_destroy(ndsp_ctx* c) {
  if (!c) return 0;
  if (c->st) hipStreamSynchronize(c->st);
  KArgs& a = c->a;
  for (int fi = 0; fi < F_NFIELDS; fi++) if (a.f[fi]) hipFree(a.f[fi]);
  void* ptrs[] = {c->scratch_d, c->scratch_i, a.numneigh, a.nits, a.uid, a.itype, a.ireal, a.slot_of, a.cellid, a.sidx, a.P4, a.dxb,
                  a.DR, a.FR, a.GF, a.sc, (void*)a.tabf, (void*)a.tabd,
                  c->part, c->cons_out, a.cstart, a.cfill, c->bsum, c->list_a, c->list_b, a.nrflag, c->cmp_bsum,
                  a.nbl, a.nbc, a.nbx, a.hcap, a.red, c->gdev};
  for (void* q : ptrs) if (q) hipFree(q);
  if (c->graph_a) hipGraphExecDestroy(c->graph_a);
  if (c->graph_b) hipGraphExecDestroy(c->graph_b);
  if (c->alt_graph_a) hipGraphExecDestroy(c->alt_graph_a);
  if (c->alt_graph_b) hipGraphExecDestroy(c->alt_graph_b);
  for (int fi = 0; fi < F_NPERSIST; fi++) if (c->falt[fi]) hipFree(c->falt[fi]);
  if (c->mlp_alloced) mlp_free(c->mlp);
  for (auto e : c->ev0) hipEventDestroy(e);
  for (auto e : c->ev1) hipEventDestroy(e);
  if (c->st) hipStreamDestroy(c->st);
  delete c;
  return 0;
}

int ndsp_upload(ndsp_ctx* c, const double* x, const double* vel, const double* pmass, const double* hh,
                const double* en, const double* Bevol, const double* alpha, const double* psi,
                const double* dustevol, const int32_t* itype) {
  if (!c || !x || !vel || !pmass || !hh || !en || !alpha || !psi) return set_err(NDSP_ERR_BAD_ARG, "upload: null array");
  const int n = c->n, nd = c->p.ndim;
  std::vector<double> buf(n);
  auto put = [&](int fi) { return hipMemcpyAsync(c->a.f[fi], buf.data(), sizeof(double) * n, hipMemcpyHostToDevice, c->st); };
  for (int i = 0; i < n; i++) {
    if (itype && itype[i] != 0 && itype[i] != 1)
      return set_err(NDSP_ERR_BAD_OPTION, "only gas (itype 0) and fixed boundary (itype 1) particles are supported");
    if (itype && itype[i] == 1 && !c->p.fixed_on)
      return set_err(NDSP_ERR_BAD_OPTION, "fixed particles (itype 1) need a wall of fixed particles (ibound 1)");
    if (!(hh[i] > 0.0)) return set_err(NDSP_ERR_H_NONPOSITIVE, "upload: h <= 0");
  }
  for (int d = 0; d < 3; d++) {
    for (int i = 0; i < n; i++) buf[i] = d < nd ? x[(size_t)i * nd + d] : 0.0;
    HIPCHK(put(F_X + d));
    HIPCHK(hipStreamSynchronize(c->st));
  }
  const int vf[3] = {F_VX, F_VY, F_VZ}, bf[3] = {F_BX, F_BY, F_BZ}, af[3] = {F_A0, F_A1, F_A2};
  for (int d = 0; d < 3; d++) {
    for (int i = 0; i < n; i++) buf[i] = vel[(size_t)i * 3 + d];
    HIPCHK(put(vf[d])); HIPCHK(hipStreamSynchronize(c->st));
    for (int i = 0; i < n; i++) buf[i] = (c->p.imhd && Bevol) ? Bevol[(size_t)i * 3 + d] : 0.0;
    HIPCHK(put(bf[d])); HIPCHK(hipStreamSynchronize(c->st));
    for (int i = 0; i < n; i++) buf[i] = alpha[(size_t)i * 3 + d];
    HIPCHK(put(af[d])); HIPCHK(hipStreamSynchronize(c->st));
  }
  const double* one[4] = {pmass, hh, en, psi};
  const int onef[4] = {F_M, F_H, F_EN, F_PSI};
  for (int q = 0; q < 4; q++) {
    for (int i = 0; i < n; i++) buf[i] = one[q][i];
    HIPCHK(put(onef[q])); HIPCHK(hipStreamSynchronize(c->st));
  }
  for (int i = 0; i < n; i++) buf[i] = (c->p.idust && dustevol) ? dustevol[i] : 0.0;
  HIPCHK(put(F_EPS)); HIPCHK(hipStreamSynchronize(c->st));
  // dustfrac starts equal to the evolved dust fraction
  HIPCHK(hipMemcpy(c->a.f[F_DUSTFRAC], c->a.f[F_EPS], sizeof(double) * n, hipMemcpyDeviceToDevice));
  std::vector<int> ids(n);
  for (int i = 0; i < n; i++) ids[i] = i;
  HIPCHK(hipMemcpy(c->a.uid, ids.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->a.slot_of, ids.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  for (int i = 0; i < n; i++) ids[i] = itype ? itype[i] : 0;
  HIPCHK(hipMemcpy(c->a.itype, ids.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  for (int i = 0; i < n; i++) ids[i] = -1;
  HIPCHK(hipMemcpy(c->a.ireal, ids.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  c->nfixed = 0;
  for (int i = 0; i < n; i++) c->nfixed += itype && itype[i] == 1;
  c->ireal_set = c->nfixed == 0;
  // derived fields and rates start at zero
  for (int fi = F_RHO; fi < F_NFIELDS; fi++)
    if (fi != F_DUSTFRAC) HIPCHK(hipMemsetAsync(c->a.f[fi], 0, sizeof(double) * n, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  for (int d = 0; d < 3; d++)
    HIPCHK(hipMemcpy(c->a.f[F_BFX + d], c->a.f[F_BX + d], sizeof(double) * n, hipMemcpyDeviceToDevice));
  c->grid_valid = false;
  c->have_rates = false;
  c->nsteps = 0;
  return 0;
}

static const struct { const char* name; int f0; int w; } kFieldMap[] = {
    {"vel", F_VX, 3}, {"pmass", F_M, 1}, {"hh", F_H, 1}, {"en", F_EN, 1}, {"Bevol", F_BX, 3},
    {"alpha", F_A0, 3}, {"psi", F_PSI, 1}, {"dustevol", F_EPS, 1}, {"rho", F_RHO, 1},
    {"gradh", F_GRADH, 1}, {"gradgradh", F_GRADGRADH, 1}, {"drhodt", F_DRHODT, 1},
    {"dhdt", F_DHDT, 1}, {"dens", F_DENS, 1}, {"pr", F_PR, 1}, {"spsound", F_CS, 1},
    {"uu", F_UU, 1}, {"Bfield", F_BFX, 3}, {"force", F_FX, 3}, {"dudt", F_DUDT, 1},
    {"dendt", F_DENDT, 1}, {"dBevoldt", F_DBX, 3}, {"daldt", F_DA0, 3}, {"dpsidt", F_DPSIDT, 1},
    {"divB", F_DIVB, 1}, {"curlB", F_CBX, 3}, {"gradpsi", F_GPX, 3}, {"dustfrac", F_DUSTFRAC, 1},
    {"rhogas", F_RHOGAS, 1}, {"rhodust", F_RHODUST, 1}, {"ddustevoldt", F_DEPS, 1},
    {"deltav", F_DVX, 3}, {"del2u", F_DEL2U, 1}, {"graddivv", F_GDVX, 3}, {"hh_eval", F_HEVAL, 1}};

int ndsp_download(ndsp_ctx* c, const char* name, double* out, int32_t* width) {
  if (!c || !name || !out) return set_err(NDSP_ERR_BAD_ARG, "download: null argument");
  const int n = c->n;
  std::vector<int> uid(n);
  HIPCHK(hipStreamSynchronize(c->st));
  HIPCHK(hipMemcpy(uid.data(), c->a.uid, sizeof(int) * n, hipMemcpyDeviceToHost));
  std::vector<double> buf(n);
  if (!strcmp(name, "x")) {
    const int nd = c->p.ndim;
    for (int d = 0; d < nd; d++) {
      HIPCHK(hipMemcpy(buf.data(), c->a.f[F_X + d], sizeof(double) * n, hipMemcpyDeviceToHost));
      for (int s = 0; s < n; s++) out[(size_t)uid[s] * nd + d] = buf[s];
    }
    if (width) *width = nd;
    return 0;
  }
  if (!strcmp(name, "slot")) {
    // where each particle sits in device memory (the cell order of the last reorder)
    for (int s = 0; s < n; s++) out[uid[s]] = s;
    if (width) *width = 1;
    return 0;
  }
  if (!strcmp(name, "nblist_len")) {
    // entries of each particle's neighbour list at the last binning
    // (gather list + extras); -1 where it scans cells instead (an overflow,
    // or lists off)
    std::vector<int> nbc(n, -1), nbx(n, 0), sidx(n);
    if (c->p.use_list) {
      HIPCHK(hipMemcpy(nbc.data(), c->a.nbc, sizeof(int) * n, hipMemcpyDeviceToHost));
      HIPCHK(hipMemcpy(nbx.data(), c->a.nbx, sizeof(int) * n, hipMemcpyDeviceToHost));
    }
    HIPCHK(hipMemcpy(sidx.data(), c->a.sidx, sizeof(int) * n, hipMemcpyDeviceToHost));
    for (int k = 0; k < n; k++)
      out[uid[sidx[k]]] = nbc[k] < 0 || nbc[k] + nbx[k] > c->a.maxnb ? -1.0 : (double)(nbc[k] + nbx[k]);
    if (width) *width = 1;
    return 0;
  }
  if (!strcmp(name, "numneigh") || !strcmp(name, "nits")) {
    std::vector<int> ib(n);
    HIPCHK(hipMemcpy(ib.data(), !strcmp(name, "nits") ? c->a.nits : c->a.numneigh, sizeof(int) * n,
                     hipMemcpyDeviceToHost));
    for (int s = 0; s < n; s++) out[uid[s]] = ib[s];
    if (width) *width = 1;
    return 0;
  }
  for (const auto& e : kFieldMap) {
    if (strcmp(name, e.name)) continue;
    for (int q = 0; q < e.w; q++) {
      HIPCHK(hipMemcpy(buf.data(), c->a.f[e.f0 + q], sizeof(double) * n, hipMemcpyDeviceToHost));
      for (int s = 0; s < n; s++) out[(size_t)uid[s] * e.w + q] = buf[s];
    }
    if (width) *width = e.w;
    return 0;
  }
  return set_err(NDSP_ERR_BAD_ARG, std::string("download: unknown field ") + name);
}

int ndsp_set_linklist(ndsp_ctx* c) {
  if (!c) return set_err(NDSP_ERR_BAD_ARG, "null context");
  return do_linklist(c, 0);
}

int ndsp_iterate_density(ndsp_ctx* c, ndsp_stats* stats) {
  if (!c) return set_err(NDSP_ERR_BAD_ARG, "null context");
  if (!c->grid_valid) return set_err(NDSP_ERR_BAD_ARG, "iterate_density: call set_linklist first");
  if (int e = do_density(c)) return e;
  if (int e = density_finish(c)) return e;
  DScal hs;
  if (int e = sync_scalars(c, hs)) return e;
  timing_collect(c);
  fill_outputs(c, hs, stats, nullptr);
  return check_err(c, hs);
}

int ndsp_get_rates(ndsp_ctx* c, ndsp_dt* dt) {
  if (!c) return set_err(NDSP_ERR_BAD_ARG, "null context");
  if (!c->grid_valid) return set_err(NDSP_ERR_BAD_ARG, "get_rates: call set_linklist and iterate_density first");
  if (int e = do_rates(c)) return e;
  DScal hs;
  if (int e = sync_scalars(c, hs)) return e;
  timing_collect(c);
  fill_outputs(c, hs, nullptr, dt);
  return check_err(c, hs);
}

int ndsp_derivs(ndsp_ctx* c, ndsp_stats* stats, ndsp_dt* dt) {
  if (!c) return set_err(NDSP_ERR_BAD_ARG, "null context");
  if (int e = do_linklist(c, 0)) return e;
  if (int e = do_density(c)) return e;
  if (int e = density_finish(c)) return e;
  if (int e = do_rates(c)) return e;
  DScal hs;
  if (int e = sync_scalars(c, hs)) return e;
  timing_collect(c);
  fill_outputs(c, hs, stats, dt);
  return check_err(c, hs);
}

int ndsp_set_dt(ndsp_ctx* c, double dt, double time) {
  if (!c) return set_err(NDSP_ERR_BAD_ARG, "null context");
  c->dt = dt;
  c->time = time;
  HIPCHK(hipStreamSynchronize(c->st));
  HIPCHK(hipMemcpy(&c->a.sc->dt, &dt, sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(&c->a.sc->dt_cur, &dt, sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(&c->a.sc->time, &time, sizeof(double), hipMemcpyHostToDevice));
  return 0;
}

// First half of a leapfrog step (src/stepND_leapfrog_mhd.f90:107-167):
// predictor, binning and the first chunk of h-rho passes.
static int step_part_a(ndsp_ctx* c) {
  {
    TScope t(c, "predict");
    hipLaunchKernelGGL(k_predict, dim3(nblk(c->n)), dim3(BS), 0, c->st, c->a, c->p);
    if (c->p.fixed_on) hipLaunchKernelGGL(k_predict_fixed, dim3(nblk(c->n)), dim3(BS), 0, c->st, c->a, c->p);
  }
  MlpState& ml = c->mlp;
  if (ml.enabled) {
    // features of every particle (the training rows) and, once the network
    // has trained for seed_after steps, the seed itself, in one launch on
    // the matrix cores
    TScope t(c, "mlp_seed");
    MlpSeedFields f{};
    double** fd = c->a.f;
    f.hin = fd[F_HIN];
    f.rho = fd[F_RHO];
    f.m = fd[F_M];
    f.dhdtin = fd[F_DHDTIN];
    f.drhodt = fd[F_DRHODT];
    f.corr = fd[F_MLPCORR];
    f.numneigh = c->a.numneigh;
    f.nits = c->a.nits;
    f.itype = c->p.fixed_on ? c->a.itype : nullptr;
    f.h = fd[F_H];
    f.X = ml.X;
    f.dt = c->dt;
    f.nexp = nexp_of(c->p);
    f.hfact = c->p.hfact;
    f.dndim = c->p.dndim;
    f.n = c->n;
    f.apply = ml.have_params && ml.steps_seen >= ml.seed_after;
    mlp_launch_seed_mfma(ml, f, c->st);
  }
  if (int e = do_linklist(c, 1)) return e;
  if (int e = do_density(c)) return e;
  return 0;
}

// Second half (src/stepND_leapfrog_mhd.f90:167-252): rates, corrector,
// wrap, next dt; online training of the h predictor when enabled.
static int step_part_b(ndsp_ctx* c) {
  MlpState& ml = c->mlp;
  if (ml.enabled) {
    TScope t(c, "mlp_target");
    hipLaunchKernelGGL(k_mlp_target, dim3(nblk(c->n)), dim3(BS), 0, c->st, c->a, ml.X);
  }
  if (int e = do_rates(c, true)) return e;
  {
    TScope t(c, "correct");
    hipLaunchKernelGGL(k_boundary, dim3(nblk(c->n)), dim3(BS), 0, c->st, c->a, c->p, c->g, 1);
    hipLaunchKernelGGL(k_step_end, dim3(1), dim3(64), 0, c->st, c->a.sc);
  }
  if (ml.enabled && ml.train) {
    TScope t(c, "mlp_train");
    // online training on this step's (features, converged h): Adam steps on
    // pseudo-random subsets (stride coprime with n)
    const int B = std::min(ml.batch, c->n);
    for (int it = 0; it < ml.iters; it++) {
      long long stride = 1000003;
      while (std::__gcd(stride, (long long)c->n) != 1) stride += 2;
      const long long off = (long long)(((unsigned long long)(c->nsteps + 1) * 2654435761ULL + it * 40503ULL) % c->n);
      mlp_launch_grad(ml, ml.X, c->n, B, off, stride, c->st);
      mlp_launch_adam(ml, ml.lr, c->st);
    }
  }
  HIPCHK(hipGetLastError());
  return 0;
}

// records one half of a step into an executable graph
static int capture(ndsp_ctx* c, hipGraphExec_t& exec, int (*part)(ndsp_ctx*)) {
  if (exec) {
    hipGraphExecDestroy(exec);
    exec = nullptr;
  }
  hipGraph_t graph = nullptr;
  HIPCHK(hipStreamBeginCapture(c->st, hipStreamCaptureModeThreadLocal));
  c->capturing = true;
  const int e = part(c);
  c->capturing = false;
  const hipError_t ce = hipStreamEndCapture(c->st, &graph);
  if (e) {
    if (graph) hipGraphDestroy(graph);
    return e;
  }
  if (ce != hipSuccess) return set_err(NDSP_ERR_HIP, std::string("stream capture: ") + hipGetErrorString(ce));
  const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (ie != hipSuccess) return set_err(NDSP_ERR_HIP, std::string("graph instantiate: ") + hipGetErrorString(ie));
  return 0;
}

static bool graph_usable(const ndsp_ctx* c) { return c->use_graph && !c->mlp.enabled && !c->timing; }

int ndsp_step(ndsp_ctx* c, ndsp_stats* stats, ndsp_dt* dtout) {
  if (!c) return set_err(NDSP_ERR_BAD_ARG, "null context");
  if (!c->have_rates) {
    // the reference evaluates derivs once before the first step
    // (primitive2conservative, src/conservative2primitive.f90:799)
    if (int e = ndsp_derivs(c, nullptr, nullptr)) return e;
  }
  if (c->grid_valid && c->reorder_every > 0 && (c->nsteps % c->reorder_every == 0 || c->reorder_next)) {
    if (int e = reorder(c)) return e;
    c->reorder_next = false;
  }
  if (graph_usable(c) && c->cap_bits > 0) {
    // the step as two recorded graphs; the h-rho iteration is completed on
    // the host path in between when it needs more passes than recorded
    if (!c->graph_valid) {
      if (int e = capture(c, c->graph_a, step_part_a)) return e;
      if (int e = capture(c, c->graph_b, step_part_b)) return e;
      c->graph_valid = true;
    }
    HIPCHK(hipGraphLaunch(c->graph_a, c->st));
    if (int e = density_finish(c)) return e;
    HIPCHK(hipGraphLaunch(c->graph_b, c->st));
  } else {
    if (int e = step_part_a(c)) return e;
    if (int e = density_finish(c)) return e;
    if (int e = step_part_b(c)) return e;
  }
  DScal hs;
  if (int e = sync_scalars(c, hs)) return e;
  timing_collect(c);
  c->time = hs.time;
  c->dt = hs.dt_cur;
  c->nsteps++;
  if (hs.grid_clamped && c->cap_bits < 28)
    if (int e = size_key_space(c, hs, 2)) return e;
  if (hs.list_overflow)
    if (int e = grow_lists(c)) return e;
  if (c->mlp.enabled) c->mlp.steps_seen++;
  fill_outputs(c, hs, stats, dtout);
  return check_err(c, hs);
}

// ===========================================================================
// batches of independent simulations (a parameter sweep): one recorded graph
// per half step covering every member, members on their own streams
// ===========================================================================

static int ensure_key_space(ndsp_ctx* c);

// A group of batch members that share dimension, physics, neighbour-list use
// and walls: one launch per phase covers all of them.
struct BGroup {
  int ndim = 0, phys = 0, dust = 0, lists = 0, reflect = 0;
  int fixed = 0;  // some member holds fixed particles
  std::vector<int> mem;
  MArgs* args = nullptr;
  // workgroup offsets per launch shape (device; nm + 1 entries each)
  enum { S_256, S_64, S_ONE, S_BIN, S_FILL, S_SCAN, S_CAP, S_NR, S_FORCE, S_HMAX, NSHAPE };
  int* off[NSHAPE] = {};
  int tot[NSHAPE] = {};
  int* due = nullptr;  // members whose periodic reorder is due (device, nm entries)
};

struct ndsp_batch {
  std::vector<ndsp_ctx*> sims;
  hipStream_t st = nullptr;          // stream the graphs are launched on
  hipGraphExec_t ga = nullptr, gb = nullptr;
  bool valid = false;
  DScal* hs = nullptr;               // pinned host copies of every member's scalars
  DScal* dsc = nullptr;              // the members' scalars gathered on the device
  std::vector<int> caps;             // members' key-space sizes at recording
  std::vector<int> nbcaps;           // members' list capacities at recording
  std::vector<hipStream_t> gst;      // streams of the groups after the first (recording only)
  std::vector<hipEvent_t> gev;       // fork event, then one join event per further group
  std::vector<BGroup> groups;
  int prec = 0;
  int tb = 256;                      // workgroup size of the pair kernels in batch launches
};

static void batch_free_groups(ndsp_batch* b) {
  for (auto& g : b->groups) {
    if (g.args) hipFree(g.args);
    if (g.due) hipFree(g.due);
    for (int k = 0; k < BGroup::NSHAPE; k++)
      if (g.off[k]) hipFree(g.off[k]);
  }
  b->groups.clear();
}

// groups, per-member launch arguments and workgroup offsets of every shape
static int batch_build(ndsp_batch* b) {
  batch_free_groups(b);
  const int ns = (int)b->sims.size();
  for (int s = 0; s < ns; s++) {
    ndsp_ctx* c = b->sims[s];
    const int phys = c->p.imhd ? 1 : c->p.idust ? 2 : 0;
    BGroup* g = nullptr;
    for (auto& q : b->groups)
      if (q.ndim == c->p.ndim && q.phys == phys && q.lists == c->p.use_list && q.reflect == (int)c->reflect)
        g = &q;
    if (!g) {
      b->groups.emplace_back();
      g = &b->groups.back();
      g->ndim = c->p.ndim;
      g->phys = phys;
      g->dust = c->p.idust != 0;
      g->lists = c->p.use_list;
      g->reflect = c->reflect;
    }
    g->mem.push_back(s);
    g->fixed |= c->p.fixed_on;
  }
  while (b->gst.size() + 1 < b->groups.size()) {
    hipStream_t st = nullptr;
    HIPCHK(hipStreamCreate(&st));
    b->gst.push_back(st);
  }
  while (b->gev.size() < b->groups.size()) {
    hipEvent_t ev = nullptr;
    HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    b->gev.push_back(ev);
  }
  for (auto& g : b->groups) {
    const int nm = (int)g.mem.size();
    std::vector<MArgs> ma(nm);
    std::vector<int> off[BGroup::NSHAPE];
    for (auto& o : off) o.assign(1, 0);
    for (int q = 0; q < nm; q++) {
      ndsp_ctx* c = b->sims[g.mem[q]];
      const int n = c->n;
      const long long cap = (1LL << c->cap_bits) + 1;
      const int capb = (int)std::min<long long>(nblk(cap), 1 << 20);
      MArgs& M = ma[q];
      memset(&M, 0, sizeof(M));
      M.a = c->a;
      M.a.nrflag = nullptr;  // batched passes append their unconverged particles
      M.p = c->p;
      M.gdev = c->gdev;
      M.box = c->g;
      M.part = c->part;
      M.bsum = c->bsum;
      M.list_a = c->list_a;
      M.list_b = c->list_b;
      M.out = b->dsc + g.mem[q];
      M.bconstmax = c->bconstmax;
      M.cap_bits = c->cap_bits;
      M.reflect = c->reflect;
      M.nbp = (int)std::min<unsigned>(c->nb_part, nblk(n));
      M.nb_nr = (int)nblk(n, b->tb);
      M.nb_force = (int)nblk(n, b->tb);
      M.nb_256 = (int)nblk(n);
      M.scratch_d = c->scratch_d;
      M.scratch_i = c->scratch_i;
      const int cnt[BGroup::NSHAPE] = {(int)nblk(n), (int)nblk(n, 64), 1, M.nbp, std::min(capb, 4096),
                                       (int)((cap + BS * 4 - 1) / (BS * 4)), capb, M.nb_nr, M.nb_force,
                                       (int)std::min(1024u, nblk(n))};
      for (int k = 0; k < BGroup::NSHAPE; k++) off[k].push_back(off[k].back() + cnt[k]);
    }
    HIPCHK(hipMalloc((void**)&g.due, sizeof(int) * nm));
    HIPCHK(hipMalloc((void**)&g.args, sizeof(MArgs) * nm));
    HIPCHK(hipMemcpy(g.args, ma.data(), sizeof(MArgs) * nm, hipMemcpyHostToDevice));
    for (int k = 0; k < BGroup::NSHAPE; k++) {
      HIPCHK(hipMalloc((void**)&g.off[k], sizeof(int) * (nm + 1)));
      HIPCHK(hipMemcpy(g.off[k], off[k].data(), sizeof(int) * (nm + 1), hipMemcpyHostToDevice));
      g.tot[k] = off[k].back();
    }
  }
  return 0;
}

// launch of a batch kernel over shape k of group g, on stream bst
#define BLAUNCH(kern, g, k, threads, ...)                                                                   \
  hipLaunchKernelGGL(kern, dim3((g).tot[k]), dim3(threads), 0, bst, (const MArgs*)(g).args,                 \
                     (const int*)(g).off[k], (int)(g).mem.size(), ##__VA_ARGS__)

// In a recorded half step every group after the first runs on a stream of
// its own, forked from the batch stream and joined back before the scalars
// are copied out: the groups are independent, so the graph holds one branch
// per group and their phases overlap on the device.
static hipStream_t batch_group_stream(ndsp_batch* b, size_t gi) { return gi == 0 ? b->st : b->gst[gi - 1]; }
static int batch_fork(ndsp_batch* b) {
  if (b->groups.size() < 2) return 0;
  HIPCHK(hipEventRecord(b->gev[0], b->st));
  for (size_t gi = 1; gi < b->groups.size(); gi++) HIPCHK(hipStreamWaitEvent(b->gst[gi - 1], b->gev[0], 0));
  return 0;
}
static int batch_join(ndsp_batch* b) {
  for (size_t gi = 1; gi < b->groups.size(); gi++) {
    HIPCHK(hipEventRecord(b->gev[gi], b->gst[gi - 1]));
    HIPCHK(hipStreamWaitEvent(b->st, b->gev[gi], 0));
  }
  return 0;
}

// first half of a step for every member (as step_part_a), then the scalars
static int batch_part_a(ndsp_batch* b) {
  const bool fp64 = b->prec == NDSP_PREC_FP64;
  if (int e = batch_fork(b)) return e;
  for (size_t gi = 0; gi < b->groups.size(); gi++) {
    auto& g = b->groups[gi];
    const hipStream_t bst = batch_group_stream(b, gi);
    BLAUNCH(km_predict, g, BGroup::S_256, BS);
    if (g.fixed) BLAUNCH(km_predict_fixed, g, BGroup::S_256, BS);
    BLAUNCH(km_err_reset, g, BGroup::S_ONE, 64);
    BLAUNCH(km_boundary_step, g, BGroup::S_256, BS);
    BLAUNCH(km_binstats, g, BGroup::S_BIN, BS);
    BLAUNCH(km_binstats_final, g, BGroup::S_ONE, BS);
    BLAUNCH(km_grid_setup, g, BGroup::S_ONE, 64);
    BLAUNCH(km_fill_cells, g, BGroup::S_FILL, BS);
    BLAUNCH(km_cellid, g, BGroup::S_256, BS);
    BLAUNCH(km_scan_local, g, BGroup::S_SCAN, BS);
    BLAUNCH(km_scan_bsum, g, BGroup::S_ONE, BS);
    BLAUNCH(km_scan_add, g, BGroup::S_CAP, BS);
    BLAUNCH(km_fill_cells, g, BGroup::S_FILL, BS);
    BLAUNCH(km_scatter, g, BGroup::S_256, BS);
    BLAUNCH(km_cellsort, g, BGroup::S_CAP, BS);
    if (fp64) BLAUNCH(km_pack_density<double>, g, BGroup::S_256, BS);
    else BLAUNCH(km_pack_density<float>, g, BGroup::S_256, BS);
    if (g.lists) {
      BLAUNCH(km_fill_nbx, g, BGroup::S_256, BS);
      dispatch_ndim(g.ndim, [&](auto nd) {
        constexpr int ND = decltype(nd)::value;
        BLAUNCH(km_nblist<ND>, g, BGroup::S_64, 64);
        return 0;
      });
      BLAUNCH(km_nblist_sort, g, BGroup::S_256, BS);
    }
    BLAUNCH(km_init_density, g, BGroup::S_ONE, 64);
    const RedSpec spec = red_spec({RED_Q(RED_MAX, RED_F64, stressmax), RED_Q(RED_MAX, RED_F64, hmax_nr),
                                   RED_Q(RED_MAX, RED_I32, itsmax), RED_Q(RED_MIN, RED_I32, nmin),
                                   RED_Q(RED_MAX, RED_I32, nmax), RED_Q(RED_SUM, RED_U64, pairs_density),
                                   RED_Q(RED_SUM, RED_U64, ncalc_total)});
    const int K = g.reflect ? 4 : 3;
    for (int q = 0; q < K; q++) {
      dispatch_ndim(g.ndim, [&](auto nd) {
        constexpr int ND = decltype(nd)::value;
        if (b->tb == 64) {
          if (fp64) { if (g.dust) BLAUNCH((km_nr_pass<ND, double, true, 64>), g, BGroup::S_NR, 64); else BLAUNCH((km_nr_pass<ND, double, false, 64>), g, BGroup::S_NR, 64); }
          else { if (g.dust) BLAUNCH((km_nr_pass<ND, float, true, 64>), g, BGroup::S_NR, 64); else BLAUNCH((km_nr_pass<ND, float, false, 64>), g, BGroup::S_NR, 64); }
        } else {
          if (fp64) { if (g.dust) BLAUNCH((km_nr_pass<ND, double, true, 256>), g, BGroup::S_NR, 256); else BLAUNCH((km_nr_pass<ND, double, false, 256>), g, BGroup::S_NR, 256); }
          else { if (g.dust) BLAUNCH((km_nr_pass<ND, float, true, 256>), g, BGroup::S_NR, 256); else BLAUNCH((km_nr_pass<ND, float, false, 256>), g, BGroup::S_NR, 256); }
        }
        return 0;
      });
      BLAUNCH(km_red_final, g, BGroup::S_ONE, BS, spec, 0);
      BLAUNCH(km_pass_ctl, g, BGroup::S_ONE, 64);
      BLAUNCH(km_hmax, g, BGroup::S_HMAX, BS);
      BLAUNCH(km_refresh_hinv, g, BGroup::S_256, BS);
    }
    BLAUNCH(km_gather_scalars, g, BGroup::S_ONE, 64);
  }
  if (int e = batch_join(b)) return e;
  HIPCHK(hipMemcpyAsync(b->hs, b->dsc, sizeof(DScal) * b->sims.size(), hipMemcpyDeviceToHost, b->st));
  HIPCHK(hipGetLastError());
  return 0;
}

// second half of a step for every member (as step_part_b), then the scalars
static int batch_part_b(ndsp_batch* b) {
  const bool fp64 = b->prec == NDSP_PREC_FP64;
  if (int e = batch_fork(b)) return e;
  for (size_t gi = 0; gi < b->groups.size(); gi++) {
    auto& g = b->groups[gi];
    const hipStream_t bst = batch_group_stream(b, gi);
    if (g.fixed) BLAUNCH(km_fixed_copy, g, BGroup::S_256, BS);
    if (fp64) BLAUNCH(km_pack_force<double>, g, BGroup::S_256, BS);
    else BLAUNCH(km_pack_force<float>, g, BGroup::S_256, BS);
    BLAUNCH(km_init_force, g, BGroup::S_ONE, 64);
    dispatch_ndim(g.ndim, [&](auto nd) {
      constexpr int ND = decltype(nd)::value;
#define NDSP_FORCE_B(PTT, TBB)                                                                        \
  do {                                                                                                \
    if (g.phys == 1) BLAUNCH((km_force<ND, PTT, 1, TBB>), g, BGroup::S_FORCE, TBB);                   \
    else if (g.phys == 2) BLAUNCH((km_force<ND, PTT, 2, TBB>), g, BGroup::S_FORCE, TBB);              \
    else BLAUNCH((km_force<ND, PTT, 0, TBB>), g, BGroup::S_FORCE, TBB);                               \
  } while (0)
      if (b->tb == 64) { if (fp64) NDSP_FORCE_B(double, 64); else NDSP_FORCE_B(float, 64); }
      else { if (fp64) NDSP_FORCE_B(double, 256); else NDSP_FORCE_B(float, 256); }
#undef NDSP_FORCE_B
      return 0;
    });
    BLAUNCH(km_red_final, g, BGroup::S_ONE, BS,
            red_spec({RED_Q(RED_MAX, RED_F64, vsigmax), RED_Q(RED_MIN, RED_F64, dtcourant),
                      RED_Q(RED_SUM, RED_U64, pairs_force)}), 1);
    BLAUNCH(km_finalize, g, BGroup::S_256, BS);
    BLAUNCH(km_red_final, g, BGroup::S_ONE, BS,
            red_spec({RED_Q(RED_MAX, RED_F64, fhmax), RED_Q(RED_MIN, RED_F64, dtdrag)}), 2);
    BLAUNCH(km_dt, g, BGroup::S_ONE, 64);
    BLAUNCH(km_correct, g, BGroup::S_256, BS);
    BLAUNCH(km_boundary_step, g, BGroup::S_256, BS);
    BLAUNCH(km_step_end, g, BGroup::S_ONE, 64);
    BLAUNCH(km_gather_scalars, g, BGroup::S_ONE, 64);
  }
  if (int e = batch_join(b)) return e;
  HIPCHK(hipMemcpyAsync(b->hs, b->dsc, sizeof(DScal) * b->sims.size(), hipMemcpyDeviceToHost, b->st));
  HIPCHK(hipGetLastError());
  return 0;
}

// the periodic reorder of the flagged members (as reorder() does for one
// context), each field one launch over all members of a group
static int batch_reorder(ndsp_batch* b, const std::vector<int>& due) {
  const hipStream_t bst = b->st;
  for (auto& g : b->groups) {
    const int nm = (int)g.mem.size();
    std::vector<int> d(nm);
    bool any = false;
    for (int q = 0; q < nm; q++) any |= (d[q] = due[g.mem[q]]) != 0;
    if (!any) continue;
    HIPCHK(hipMemcpyAsync(g.due, d.data(), sizeof(int) * nm, hipMemcpyHostToDevice, b->st));
    // blocks of REORDER_W double fields, then the int fields (f0 = F_NPERSIST)
    for (int f0 = 0; f0 < F_NPERSIST + REORDER_W; f0 += REORDER_W) {
      const int b0 = std::min(f0, (int)F_NPERSIST);
      for (int phase = 0; phase < 2; phase++)
        BLAUNCH(km_reorder, g, BGroup::S_256, BS, (const int*)g.due, b0, phase);
      if (b0 == F_NPERSIST) break;
    }
    BLAUNCH(km_slot_of, g, BGroup::S_256, BS);
    // the host vector d must outlive the copy
    HIPCHK(hipStreamSynchronize(b->st));
  }
  HIPCHK(hipGetLastError());
  for (size_t s = 0; s < due.size(); s++)
    if (due[s]) b->sims[s]->grid_valid = false;
  return 0;
}

static int batch_capture(ndsp_batch* b, hipGraphExec_t& exec, int (*part)(ndsp_batch*)) {
  if (exec) {
    hipGraphExecDestroy(exec);
    exec = nullptr;
  }
  hipGraph_t graph = nullptr;
  HIPCHK(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
  const int err = part(b);
  const hipError_t ce = hipStreamEndCapture(b->st, &graph);
  if (err) {
    if (graph) hipGraphDestroy(graph);
    return err;
  }
  if (ce != hipSuccess) return set_err(NDSP_ERR_HIP, std::string("batch capture: ") + hipGetErrorString(ce));
  const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  if (ie != hipSuccess) return set_err(NDSP_ERR_HIP, std::string("batch instantiate: ") + hipGetErrorString(ie));
  return 0;
}

int ndsp_batch_create(ndsp_ctx* const* ctxs, int32_t nsim, ndsp_batch** out) {
  if (!ctxs || !out || nsim < 1) return set_err(NDSP_ERR_BAD_ARG, "batch_create: bad arguments");
  *out = nullptr;
  for (int s = 0; s < nsim; s++) {
    if (!ctxs[s]) return set_err(NDSP_ERR_BAD_ARG, "batch_create: null member");
    if (ctxs[s]->prec != ctxs[0]->prec) return set_err(NDSP_ERR_BAD_ARG, "batch_create: members differ in precision");
  }
  ndsp_batch* b = new ndsp_batch();
  b->sims.assign(ctxs, ctxs + nsim);
  for (int s = 0; s < nsim; s++) ctxs[s]->in_batch = true;
  b->prec = ctxs[0]->prec;
  bool ok = hipStreamCreate(&b->st) == hipSuccess;
  ok &= hipHostMalloc((void**)&b->hs, sizeof(DScal) * nsim, hipHostMallocDefault) == hipSuccess;
  ok &= hipMalloc((void**)&b->dsc, sizeof(DScal) * nsim) == hipSuccess;
  if (!ok) {
    ndsp_batch_destroy(b);
    return set_err(NDSP_ERR_ALLOC, "batch_create: stream or scalar buffers");
  }
  *out = b;
  return 0;
}

int ndsp_batch_destroy(ndsp_batch* b) {
  if (!b) return 0;
  if (b->st) hipStreamSynchronize(b->st);
  if (b->ga) hipGraphExecDestroy(b->ga);
  if (b->gb) hipGraphExecDestroy(b->gb);
  batch_free_groups(b);
  for (hipStream_t st : b->gst) {
    hipStreamSynchronize(st);
    hipStreamDestroy(st);
  }
  for (hipEvent_t ev : b->gev) hipEventDestroy(ev);
  if (b->hs) hipHostFree(b->hs);
  if (b->dsc) hipFree(b->dsc);
  if (b->st) hipStreamDestroy(b->st);
  delete b;
  return 0;
}

// One leapfrog step of every member (each as ndsp_step would take it): the
// two halves of all members are replayed as two graphs in which each phase
// is one launch over all members of a group; members whose h-rho iteration
// needs more passes than recorded complete it on the host path in between.
int ndsp_batch_step(ndsp_batch* b, ndsp_stats* stats, ndsp_dt* dt) {
  if (!b) return set_err(NDSP_ERR_BAD_ARG, "null batch");
  const int ns = (int)b->sims.size();
  std::vector<int> due(ns, 0);
  for (int s = 0; s < ns; s++) {
    ndsp_ctx* c = b->sims[s];
    if (c->mlp.enabled || c->timing)
      return set_err(NDSP_ERR_BAD_OPTION, "batch_step: members may not time kernels or use the h predictor");
    if (!c->have_rates)
      if (int e = ndsp_derivs(c, nullptr, nullptr)) return e;
    if (int e = ensure_key_space(c)) return e;
    due[s] = c->grid_valid && c->reorder_every > 0 && (c->nsteps % c->reorder_every == 0 || c->reorder_next);
    if (due[s]) c->reorder_next = false;
    if (b->caps.size() != (size_t)ns || b->caps[s] != c->cap_bits || b->nbcaps[s] != c->a.maxnb) b->valid = false;
  }
  for (int s = 0; s < ns; s++) HIPCHK(hipStreamSynchronize(b->sims[s]->st));
  if (!b->valid) {
    if (int e = batch_build(b)) return e;
    if (int e = batch_capture(b, b->ga, batch_part_a)) return e;
    if (int e = batch_capture(b, b->gb, batch_part_b)) return e;
    b->caps.resize(ns);
    b->nbcaps.resize(ns);
    for (int s = 0; s < ns; s++) {
      b->caps[s] = b->sims[s]->cap_bits;
      b->nbcaps[s] = b->sims[s]->a.maxnb;
    }
    b->valid = true;
  }
  if (std::find(due.begin(), due.end(), 1) != due.end())
    if (int e = batch_reorder(b, due)) return e;
  HIPCHK(hipGraphLaunch(b->ga, b->st));
  HIPCHK(hipStreamSynchronize(b->st));
  for (int s = 0; s < ns; s++) {
    // every member has been binned (as do_linklist leaves it): the next
    // periodic reorder is due by the member's own step count
    b->sims[s]->grid_valid = true;
    if (!b->hs[s].pc_done)
      if (int e = density_finish(b->sims[s])) return e;
  }
  HIPCHK(hipGraphLaunch(b->gb, b->st));
  HIPCHK(hipStreamSynchronize(b->st));
  int first_err = 0;
  for (int s = 0; s < ns; s++) {
    ndsp_ctx* c = b->sims[s];
    const DScal hs = b->hs[s];
    c->time = hs.time;
    c->dt = hs.dt_cur;
    c->nsteps++;
    if (hs.grid_clamped && c->cap_bits < 28)
      if (int e = size_key_space(c, hs, 2)) return e;
    if (hs.list_overflow)
      if (int e = grow_lists(c)) return e;
    fill_outputs(c, hs, stats ? stats + s : nullptr, dt ? dt + s : nullptr);
    const int e = check_err(c, hs);
    if (e && !first_err) first_err = e;
  }
  return first_err;
}


// Autotuner of the tile shapes: workgroup size of the force and h-rho
// kernels (64 or 256 threads: small simulations need more workgroups than
// large ones) and cells per search radius of the binning (1 or 2).  Each
// combination is timed over `reps` derivs evaluations (HIP events on the
// context's stream); the fastest is kept.  The evolved state is saved before
// and restored after the trials, so tuning leaves results unchanged.
// A copy of everything a step evolves (device fields in their current
// order, the particle ids and counters, the device and host scalars), so
// that trial evaluations can be undone bit for bit.
struct SavedState {
  double* buf = nullptr;
  DScal sc;
  double time = 0.0, dt = 0.0;
  int nsteps = 0;
  bool have_rates = false, grid_valid = false;
};

static int save_state(ndsp_ctx* c, SavedState& S) {
  const size_t n = (size_t)c->n;
  if (hipMalloc(&S.buf, sizeof(double) * n * (F_NPERSIST + 3)) != hipSuccess)
    return set_err(NDSP_ERR_ALLOC, "state copy");
  int* si = (int*)(S.buf + n * F_NPERSIST);  // uid, numneigh, nits, itype, ireal
  int* const iv[5] = {c->a.uid, c->a.numneigh, c->a.nits, c->a.itype, c->a.ireal};
  for (int fi = 0; fi < F_NPERSIST; fi++)
    HIPCHK(hipMemcpyAsync(S.buf + fi * n, c->a.f[fi], sizeof(double) * n, hipMemcpyDeviceToDevice, c->st));
  for (int q = 0; q < 5; q++)
    HIPCHK(hipMemcpyAsync(si + q * n, iv[q], sizeof(int) * n, hipMemcpyDeviceToDevice, c->st));
  if (int e = sync_scalars(c, S.sc)) return e;
  S.time = c->time;
  S.dt = c->dt;
  S.nsteps = c->nsteps;
  S.have_rates = c->have_rates;
  S.grid_valid = c->grid_valid;
  return 0;
}

static int restore_state(ndsp_ctx* c, const SavedState& S) {
  const size_t n = (size_t)c->n;
  const int* si = (const int*)(S.buf + n * F_NPERSIST);
  for (int fi = 0; fi < F_NPERSIST; fi++)
    HIPCHK(hipMemcpyAsync(c->a.f[fi], S.buf + fi * n, sizeof(double) * n, hipMemcpyDeviceToDevice, c->st));
  int* const iv[5] = {c->a.uid, c->a.numneigh, c->a.nits, c->a.itype, c->a.ireal};
  for (int q = 0; q < 5; q++)
    HIPCHK(hipMemcpyAsync(iv[q], si + q * n, sizeof(int) * n, hipMemcpyDeviceToDevice, c->st));
  hipLaunchKernelGGL(k_slot_of, dim3(nblk(c->n)), dim3(BS), 0, c->st, c->a);
  HIPCHK(hipMemcpyAsync(c->a.sc, &S.sc, sizeof(DScal), hipMemcpyHostToDevice, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  c->time = S.time;
  c->dt = S.dt;
  c->nsteps = S.nsteps;
  c->have_rates = S.have_rates;
  // the restored order may differ from the last binning's
  c->grid_valid = false;
  invalidate_graphs(c);
  return 0;
}

static void free_state(SavedState& S) {
  if (S.buf) hipFree(S.buf);
  S.buf = nullptr;
}

// Key space of a context whose cell size changed (cap_bits reset to 0): the
// statistics of the current positions, as the first binning would size it.
static int ensure_key_space(ndsp_ctx* c) {
  if (c->cap_bits > 0) return 0;
  hipLaunchKernelGGL(k_binstats, dim3(c->nb_part), dim3(BS), 0, c->st, c->a, c->p.ndim, c->part);
  hipLaunchKernelGGL(k_binstats_final, dim3(1), dim3(BS), 0, c->st, c->a.sc, c->part, c->nb_part);
  DScal hs;
  if (int e = sync_scalars(c, hs)) return e;
  return size_key_space(c, hs, 1);
}

int ndsp_set_tile_shapes(ndsp_ctx* c, int32_t force_wg, int32_t density_wg, int32_t cells_per_radius) {
  if (!c) return set_err(NDSP_ERR_BAD_ARG, "null context");
  if ((force_wg != 64 && force_wg != 256) || (density_wg != 64 && density_wg != 256) ||
      (cells_per_radius != 1 && cells_per_radius != 2))
    return set_err(NDSP_ERR_BAD_ARG, "tile shapes: workgroups 64 or 256, cells per radius 1 or 2");
  c->tb_force = force_wg;
  c->tb_nr = density_wg;
  c->p.cell_sub = cells_per_radius;
  c->cap_bits = 0;  // the key space follows the cell size at the next binning
  invalidate_graphs(c);  // recorded steps hold the old launch shapes
  c->reorder_next = true;  // the slot order follows the new cells
  return 0;
}

int ndsp_autotune(ndsp_ctx* c, int32_t reps, int32_t* choice) {
  if (!c) return set_err(NDSP_ERR_BAD_ARG, "null context");
  if (reps < 1) reps = 1;
  const bool timing = c->timing;
  c->timing = false;
  SavedState S;
  if (int e = save_state(c, S)) { free_state(S); c->timing = timing; return e; }
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  const int tbs[2] = {64, 256}, subs[2] = {1, 2};
  double best = 1e300;
  int bf = c->tb_force, bn = c->tb_nr, bc = c->p.cell_sub;
  int err = 0;
  for (int a = 0; a < 2 && !err; a++)
    for (int b = 0; b < 2 && !err; b++)
      for (int q = 0; q < 2 && !err; q++) {
        c->tb_force = tbs[a];
        c->tb_nr = tbs[b];
        c->p.cell_sub = subs[q];
        c->cap_bits = 0;  // the warm-up evaluation sizes the key space for this cell size
        err = ndsp_derivs(c, nullptr, nullptr);
        double tmin = 1e300;
        for (int r = 0; r < reps && !err; r++) {
          hipEventRecord(e0, c->st);
          err = ndsp_derivs(c, nullptr, nullptr);
          hipEventRecord(e1, c->st);
          hipEventSynchronize(e1);
          float ms = 0.f;
          hipEventElapsedTime(&ms, e0, e1);
          tmin = std::min(tmin, (double)ms);
        }
        if (!err && tmin < best) {
          best = tmin;
          bf = tbs[a];
          bn = tbs[b];
          bc = subs[q];
        }
      }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  c->tb_force = bf;
  c->tb_nr = bn;
  c->p.cell_sub = bc;
  c->cap_bits = 0;  // the key space follows the chosen cell size at the next binning
  invalidate_graphs(c);
  c->reorder_next = true;  // the restored slot order predates the chosen cells
  const int re = restore_state(c, S);
  free_state(S);
  c->timing = timing;
  if (err) return err;
  if (re) return re;
  if (choice) {
    choice[0] = bf;
    choice[1] = bn;
    choice[2] = bc;
  }
  return 0;
}

// Autotuner of a batch's tile shapes: the workgroup size of the merged pair
// kernels (64 or 256) and the cells per search radius of every member (1 or
// 2), each timed over `reps` batch steps after a recorded warm-up step.  All
// members' states are restored after every trial, so tuning leaves the
// results of later steps as they would be with the chosen shapes.
int ndsp_batch_autotune(ndsp_batch* b, int32_t reps, int32_t* choice) {
  if (!b) return set_err(NDSP_ERR_BAD_ARG, "null batch");
  if (reps < 1) reps = 1;
  const int ns = (int)b->sims.size();
  for (int s = 0; s < ns; s++)
    if (!b->sims[s]->have_rates)
      if (int e = ndsp_derivs(b->sims[s], nullptr, nullptr)) return e;
  std::vector<SavedState> S(ns);
  int err = 0;
  for (int s = 0; s < ns && !err; s++) err = save_state(b->sims[s], S[s]);
  const int tbs[2] = {64, 256}, subs[2] = {1, 2};
  int bt = b->tb, bc = b->sims[0]->p.cell_sub;
  double best = 1e300;
  auto restore_all = [&]() {
    int e = 0;
    for (int s = 0; s < ns; s++)
      if (int r = restore_state(b->sims[s], S[s])) e = e ? e : r;
    return e;
  };
  for (int a = 0; a < 2 && !err; a++)
    for (int q = 0; q < 2 && !err; q++) {
      b->tb = tbs[a];
      for (int s = 0; s < ns; s++) {
        b->sims[s]->p.cell_sub = subs[q];
        b->sims[s]->cap_bits = 0;
      }
      b->valid = false;
      err = ndsp_batch_step(b, nullptr, nullptr);  // sizes, records and warms up
      if (!err) err = restore_all();
      if (err) break;
      const auto t0 = std::chrono::steady_clock::now();
      for (int r = 0; r < reps && !err; r++) err = ndsp_batch_step(b, nullptr, nullptr);
      const double t = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
      if (!err) err = restore_all();
      if (!err && t < best) {
        best = t;
        bt = tbs[a];
        bc = subs[q];
      }
    }
  b->tb = bt;
  for (int s = 0; s < ns; s++) {
    b->sims[s]->p.cell_sub = bc;
    b->sims[s]->cap_bits = 0;  // sized again at the next step
    b->sims[s]->reorder_next = true;
  }
  b->valid = false;
  const int re = restore_all();
  for (auto& x : S) free_state(x);
  if (err) return err;
  if (re) return re;
  if (choice) {
    choice[0] = bt;
    choice[1] = bc;
  }
  return 0;
}

// Copy of nbytes between two fresh device buffers with the library's plain
// copy kernel: a reference of known traffic (nbytes read, nbytes written)
// for calibrating hardware memory counters.
int ndsp_copy_probe(ndsp_ctx* c, int64_t nbytes) {
  if (!c || nbytes < 8) return set_err(NDSP_ERR_BAD_ARG, "copy_probe: bad arguments");
  const long long nd = nbytes / 8;
  if (nd > INT_MAX) return set_err(NDSP_ERR_BAD_ARG, "copy_probe: at most 2^31 doubles");
  double *src = nullptr, *dst = nullptr;
  if (hipMalloc(&src, sizeof(double) * nd) != hipSuccess || hipMalloc(&dst, sizeof(double) * nd) != hipSuccess) {
    if (src) hipFree(src);
    return set_err(NDSP_ERR_ALLOC, "copy_probe: buffers");
  }
  hipLaunchKernelGGL(k_fill_d, dim3(4096), dim3(BS), 0, c->st, src, (int)nd, 1.0);
  HIPCHK(hipStreamSynchronize(c->st));
  hipLaunchKernelGGL(k_copy_d, dim3(nblk(nd)), dim3(BS), 0, c->st, (const double*)src, dst, (int)nd);
  HIPCHK(hipStreamSynchronize(c->st));
  // the same bytes read as scattered 64-byte lines (gather-shaped)
  long long nl = 1;
  while (nl * 2 * 64 <= (long long)nd * 8) nl *= 2;
  if (nl >= 64) {
    hipLaunchKernelGGL(k_fill_d, dim3(4096), dim3(BS), 0, c->st, dst, (int)(nl * 8), 2.0);
    HIPCHK(hipStreamSynchronize(c->st));
    hipLaunchKernelGGL(k_gather_probe, dim3((unsigned)((nl + BS - 1) / BS)), dim3(BS), 0, c->st,
                       (const double4*)dst, (float*)src, nl, 0x9E3779B97F4A7C15ull);
    HIPCHK(hipStreamSynchronize(c->st));
  }
  hipFree(src);
  hipFree(dst);
  return 0;
}

int ndsp_set_reorder(ndsp_ctx* c, int32_t every) {
  if (!c || every < 0) return set_err(NDSP_ERR_BAD_ARG, "set_reorder: bad arguments");
  c->reorder_every = every;
  return 0;
}

int ndsp_set_ireal(ndsp_ctx* c, const int32_t* ireal) {
  if (!c || !ireal) return set_err(NDSP_ERR_BAD_ARG, "set_ireal: null argument");
  const int n = c->n;
  std::vector<int> uid(n), it(n), ir(n);
  HIPCHK(hipStreamSynchronize(c->st));
  HIPCHK(hipMemcpy(uid.data(), c->a.uid, sizeof(int) * n, hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(it.data(), c->a.itype, sizeof(int) * n, hipMemcpyDeviceToHost));
  std::vector<int> type_of(n);
  for (int s = 0; s < n; s++) type_of[uid[s]] = it[s];
  for (int u = 0; u < n; u++) {
    const int j = ireal[u];
    if (j < -1 || j >= n) return set_err(NDSP_ERR_BAD_ARG, "set_ireal: index out of range");
    if (type_of[u] == 1 && (j < 0 || type_of[j] != 0))
      return set_err(NDSP_ERR_BAD_ARG, "set_ireal: a fixed particle must copy a gas particle");
  }
  for (int s = 0; s < n; s++) ir[s] = type_of[uid[s]] == 1 ? ireal[uid[s]] : -1;
  HIPCHK(hipMemcpy(c->a.ireal, ir.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  c->ireal_set = true;
  return 0;
}

int ndsp_set_graph(ndsp_ctx* c, int32_t on) {
  if (!c) return set_err(NDSP_ERR_BAD_ARG, "null context");
  c->use_graph = on != 0;
  invalidate_graphs(c);
  return 0;
}

int ndsp_conserved(ndsp_ctx* c, double* out) {
  if (!c || !out) return set_err(NDSP_ERR_BAD_ARG, "null argument");
  const int nb = std::min(c->nb_part, (int)nblk(c->n));
  hipLaunchKernelGGL(k_conserved, dim3(nb), dim3(BS), 0, c->st, c->a, c->p, c->part);
  hipLaunchKernelGGL(k_conserved_final, dim3(1), dim3(FOLD_T), 0, c->st, (const double*)c->part, nb, c->cons_out, c->n);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, c->cons_out, sizeof(double) * NCONS, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

int ndsp_ev_columns(ndsp_ctx* c, double t, double dt, double* cols, int32_t* ncols) {
  if (!c || !cols) return set_err(NDSP_ERR_BAD_ARG, "null argument");
  if (!c->have_rates) return set_err(NDSP_ERR_BAD_ARG, "ev_columns: call derivs first");
  const int nb = std::min(c->nb_part, (int)nblk(c->n));
  const int nc = c->p.imhd ? 25 : c->p.idust ? 14 : 13;
  hipLaunchKernelGGL(k_ev_sums, dim3(nb), dim3(BS), 0, c->st, c->a, c->p, c->part);
  hipLaunchKernelGGL(k_ev_final, dim3(1), dim3(FOLD_T), 0, c->st, (const double*)c->part, nb, c->n, c->p, t, dt,
                     c->cons_out);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(cols, c->cons_out, sizeof(double) * nc, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  if (ncols) *ncols = nc;
  return 0;
}

int ndsp_neighbours(ndsp_ctx* c, int32_t sym, int32_t* pairs, int64_t cap, int64_t* npairs) {
  if (!c || !npairs) return set_err(NDSP_ERR_BAD_ARG, "null argument");
  if (!c->grid_valid || !c->have_rates) return set_err(NDSP_ERR_BAD_ARG, "neighbours: call derivs first");
  const int n = c->n;
  int* counts = nullptr;
  int* offs = nullptr;
  int* dpairs = nullptr;
  HIPCHK(hipMalloc(&counts, sizeof(int) * n));
  dispatch_ndim(c->p.ndim, [&](auto nd) {
    constexpr int ND = decltype(nd)::value;
    hipLaunchKernelGGL((k_nb<ND>), dim3(nblk(n)), dim3(BS), 0, c->st, c->a, c->p, (const Grid*)c->gdev, (int)sym,
                       (const int*)nullptr, counts);
    return 0;
  });
  std::vector<int> hc(n), ho(n);
  HIPCHK(hipMemcpy(hc.data(), counts, sizeof(int) * n, hipMemcpyDeviceToHost));
  long long tot = 0;
  for (int k = 0; k < n; k++) { ho[k] = (int)tot; tot += hc[k]; }
  *npairs = tot;
  if (tot > cap || !pairs) {
    hipFree(counts);
    return set_err(NDSP_ERR_BAD_ARG, "neighbours: buffer too small");
  }
  HIPCHK(hipMalloc(&offs, sizeof(int) * n));
  HIPCHK(hipMalloc(&dpairs, sizeof(int) * 3 * std::max(1LL, tot)));
  HIPCHK(hipMemcpy(offs, ho.data(), sizeof(int) * n, hipMemcpyHostToDevice));
  dispatch_ndim(c->p.ndim, [&](auto nd) {
    constexpr int ND = decltype(nd)::value;
    hipLaunchKernelGGL((k_nb<ND>), dim3(nblk(n)), dim3(BS), 0, c->st, c->a, c->p, (const Grid*)c->gdev, (int)sym,
                       (const int*)offs, dpairs);
    return 0;
  });
  HIPCHK(hipMemcpy(pairs, dpairs, sizeof(int) * 3 * tot, hipMemcpyDeviceToHost));
  hipFree(counts);
  hipFree(offs);
  hipFree(dpairs);
  return 0;
}

int ndsp_set_timing(ndsp_ctx* c, int32_t on) {
  if (!c) return set_err(NDSP_ERR_BAD_ARG, "null context");
  c->timing = on != 0;
  for (auto& v : c->tms) v = 0.0;
  for (auto& v : c->tcount) v = 0;
  return 0;
}

int ndsp_kernel_times(ndsp_ctx* c, double* ms, int32_t cap, int32_t* nout, char* names, int32_t names_cap) {
  if (!c || !nout) return set_err(NDSP_ERR_BAD_ARG, "null argument");
  int n = (int)c->tnames.size();
  *nout = n;
  std::string all;
  for (int k = 0; k < n; k++) {
    if (ms && k < cap) ms[k] = c->tcount[k] ? c->tms[k] / c->tcount[k] : 0.0;
    if (k) all += ",";
    all += c->tnames[k];
  }
  if (names && names_cap > 0) {
    strncpy(names, all.c_str(), names_cap - 1);
    names[names_cap - 1] = 0;
  }
  return 0;
}

// ---- h-predictor MLP -------------------------------------------------------

static int mlp_ready(ndsp_ctx* c) {
  if (c->mlp_alloced) return 0;
  HIPCHK(mlp_alloc(c->mlp, c->n));
  c->mlp_alloced = true;
  return 0;
}

int ndsp_mlp_config(ndsp_ctx* c, int32_t enable, int32_t train, double lr, int32_t batch, int32_t iters,
                    int32_t seed_after) {
  if (!c || batch < 1 || iters < 0 || seed_after < 0) return set_err(NDSP_ERR_BAD_ARG, "mlp_config: bad arguments");
  if (int e = mlp_ready(c)) return e;
  MlpState& m = c->mlp;
  m.enabled = enable != 0;
  m.train = train != 0;
  m.lr = (float)lr;
  m.batch = batch;
  m.iters = iters;
  m.seed_after = seed_after;
  if (m.enabled && !m.have_params) {
    // Glorot-uniform weights, zero biases (same layout as ndsp_mlp_set_params)
    std::vector<float> w(MLP_NPARAM, 0.0f);
    unsigned long long st = 0x9E3779B97F4A7C15ULL;
    auto uni = [&]() {
      st = st * 6364136223846793005ULL + 1442695040888963407ULL;
      return (double)(st >> 11) * (1.0 / 9007199254740992.0);
    };
    int o = 0;
    const int shapes[3][2] = {{MLP_NH, MLP_NIN}, {MLP_NH, MLP_NH}, {1, MLP_NH}};
    for (auto& sh : shapes) {
      const double lim = std::sqrt(6.0 / (sh[0] + sh[1]));
      for (int k = 0; k < sh[0] * sh[1]; k++) w[o + k] = (float)((2.0 * uni() - 1.0) * lim);
      o += sh[0] * sh[1] + sh[0];
    }
    HIPCHK(mlp_set_params(m, w.data(), c->st));
    HIPCHK(hipStreamSynchronize(c->st));
  }
  return 0;
}

int ndsp_mlp_set_params(ndsp_ctx* c, const float* params, int32_t n) {
  if (!c || !params || n != MLP_NPARAM) return set_err(NDSP_ERR_BAD_ARG, "mlp_set_params: need 1313 floats");
  if (int e = mlp_ready(c)) return e;
  HIPCHK(mlp_set_params(c->mlp, params, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

int ndsp_mlp_get_params(ndsp_ctx* c, float* params, int32_t n, int32_t* adam_steps) {
  if (!c || !params || n != MLP_NPARAM) return set_err(NDSP_ERR_BAD_ARG, "mlp_get_params: need 1313 floats");
  if (int e = mlp_ready(c)) return e;
  HIPCHK(hipMemcpyAsync(params, c->mlp.w32, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  if (adam_steps) *adam_steps = c->mlp.adam_steps;
  return 0;
}

static int mlp_stage(ndsp_ctx* c, const float* X, const float* T, int32_t B) {
  MlpState& m = c->mlp;
  if (B > m.xext_rows) {
    if (m.Xext) (void)hipFree(m.Xext);
    m.Xext = nullptr;
    HIPCHK(hipMalloc(&m.Xext, sizeof(float) * MLP_LDX * (size_t)B));
    m.xext_rows = B;
  }
  std::vector<float> buf((size_t)B * MLP_LDX, 0.0f);
  for (int b = 0; b < B; b++) {
    for (int q = 0; q < MLP_NIN; q++) buf[(size_t)b * MLP_LDX + q] = X[(size_t)b * MLP_NIN + q];
    if (T) buf[(size_t)b * MLP_LDX + MLP_COL_TARGET] = T[b];
  }
  HIPCHK(hipMemcpy(m.Xext, buf.data(), sizeof(float) * buf.size(), hipMemcpyHostToDevice));
  return 0;
}

int ndsp_mlp_forward(ndsp_ctx* c, const float* X, int32_t B, float* y) {
  if (!c || !X || !y || B < 1) return set_err(NDSP_ERR_BAD_ARG, "mlp_forward: bad arguments");
  if (int e = mlp_ready(c)) return e;
  if (int e = mlp_stage(c, X, nullptr, B)) return e;
  float* dy = nullptr;
  HIPCHK(hipMalloc(&dy, sizeof(float) * B));
  mlp_launch_forward(c->mlp, c->mlp.Xext, MLP_LDX, B, dy, c->st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(y, dy, sizeof(float) * B, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  (void)hipFree(dy);
  return 0;
}

int ndsp_mlp_infer(ndsp_ctx* c, const float* X, int32_t B, float* y) {
  if (!c || !X || !y || B < 1) return set_err(NDSP_ERR_BAD_ARG, "mlp_infer: bad arguments");
  if (int e = mlp_ready(c)) return e;
  if (int e = mlp_stage(c, X, nullptr, B)) return e;
  float* dy = nullptr;
  HIPCHK(hipMalloc(&dy, sizeof(float) * B));
  mlp_launch_infer(c->mlp, c->mlp.Xext, MLP_LDX, B, dy, c->st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(y, dy, sizeof(float) * B, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  (void)hipFree(dy);
  return 0;
}

// ndsp_mlp_forward without any allocation: X and y go through the fixed
// device staging above (host or device memory, hipMemcpyDefault); the
// weights are the context's, which ndsp_mlp_set_params or ndsp_mlp_config
// must have set.  Calls that overlap on one GPU share that staging and race.
int ndsp_mlp_forward_noalloc(ndsp_ctx* c, const float* X, int32_t B, float* y) {
  if (!c || !X || !y || B < 1) return set_err(NDSP_ERR_BAD_ARG, "mlp_forward_noalloc: bad arguments");
  if (B > NOALLOC_MAX_ROWS)
    return set_err(NDSP_ERR_BAD_ARG, "mlp_forward_noalloc: at most 65536 rows per call");
  if (!c->mlp_alloced || !c->mlp.have_params)
    return set_err(NDSP_ERR_BAD_ARG, "mlp_forward_noalloc: set the weights first (ndsp_mlp_set_params)");
  if (c->device < 0 || c->device >= NOALLOC_MAX_DEVICES || !g_scratch_ptr[c->device])
    return set_err(NDSP_ERR_HIP, "mlp_forward_noalloc: no staging address for the context's device");
  float* dX = g_scratch_ptr[c->device];
  float* dy = dX + (size_t)NOALLOC_MAX_ROWS * MLP_NIN;
  HIPCHK(hipMemcpyAsync(dX, X, sizeof(float) * MLP_NIN * (size_t)B, hipMemcpyDefault, c->st));
  mlp_launch_forward(c->mlp, dX, MLP_NIN, B, dy, c->st);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(y, dy, sizeof(float) * (size_t)B, hipMemcpyDefault, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  return 0;
}

int ndsp_mlp_grad(ndsp_ctx* c, const float* X, const float* T, int32_t B, float* grad, double* loss) {
  if (!c || !X || !T || !grad || B < 1) return set_err(NDSP_ERR_BAD_ARG, "mlp_grad: bad arguments");
  if (int e = mlp_ready(c)) return e;
  if (int e = mlp_stage(c, X, T, B)) return e;
  MlpState& m = c->mlp;
  hipLaunchKernelGGL(k_fill_f, dim3(nblk(MLP_NPARAM)), dim3(BS), 0, c->st, m.grad, (int)MLP_NPARAM, 0.0f);
  mlp_launch_grad(m, m.Xext, B, B, 0, 1, c->st);
  HIPCHK(hipGetLastError());
  float l = 0.f;
  HIPCHK(hipMemcpyAsync(grad, m.grad, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipMemcpyAsync(&l, m.loss, sizeof(float), hipMemcpyDeviceToHost, c->st));
  HIPCHK(hipStreamSynchronize(c->st));
  hipLaunchKernelGGL(k_fill_f, dim3(nblk(MLP_NPARAM)), dim3(BS), 0, c->st, m.grad, (int)MLP_NPARAM, 0.0f);
  if (loss) *loss = l;
  return 0;
}

int ndsp_mlp_train(ndsp_ctx* c, const float* X, const float* T, int32_t B, int32_t iters, double lr,
                   double* loss) {
  if (!c || !X || !T || B < 1 || iters < 1) return set_err(NDSP_ERR_BAD_ARG, "mlp_train: bad arguments");
  if (int e = mlp_ready(c)) return e;
  if (int e = mlp_stage(c, X, T, B)) return e;
  MlpState& m = c->mlp;
  float l = 0.f;
  for (int it = 0; it < iters; it++) {
    mlp_launch_grad(m, m.Xext, B, B, 0, 1, c->st);
    HIPCHK(hipMemcpyAsync(&l, m.loss, sizeof(float), hipMemcpyDeviceToHost, c->st));
    mlp_launch_adam(m, (float)lr, c->st);
  }
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(c->st));
  if (loss) *loss = l;
  return 0;
}

}  // extern "C"

