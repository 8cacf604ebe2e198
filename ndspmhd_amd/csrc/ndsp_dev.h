// Device-side building blocks of the MI355X hot path: cell grid and periodic /
// reflecting image arithmetic, kernel-table interpolation, compensated sums.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define NDSP_IKERN 4000

// Cell grid.  Dimensions with walls (periodic or reflecting) are tiled exactly
// by nc cells of width w = (xmax - xmin)/nc >= radkern*hmax/cell_sub, so
// a periodic or mirror image of a cell is again a whole cell; open dimensions
// tile the particle bounding box.
struct Grid {
  int ndim;
  int nc[3];
  int kb[3];         // key bits per dimension (2^kb[d] >= nc[d])
  // layout of the key (cell_key): km[0] bits interleaved over every
  // dimension, then km[1] over dimensions km[2] < km[3], then the rest of
  // the longest
  int km[4];
  int bound[3];      // 0 open, 2 reflecting, 3 periodic
  double lo[3];      // grid origin
  double w[3];       // cell width
  double inv_w[3];
  double xmin[3], xmax[3];
  long long ncell;
};

// Morton (bit-interleaved) key of cell (cx, cy, cz): consecutive keys are
// mostly adjacent cells, so 64 consecutive particles of the sorted order form
// a compact group (x bit lowest).  Dimensions may hold different numbers of
// key bits kb[d] (an elongated box): the bits all dimensions have are
// interleaved first, then those the two longer ones have, then the rest of
// the longest (no branch: a dimension without bits at a level adds zeros),
// so the key space is 2^(kb[0]+kb[1]+kb[2]) and equal kb give the plain
// interleave.
__host__ __device__ __forceinline__ unsigned spread_bits3(unsigned x) {
  x &= 0x3ffu;
  x = (x | (x << 16)) & 0x030000FFu;
  x = (x | (x << 8)) & 0x0300F00Fu;
  x = (x | (x << 4)) & 0x030C30C3u;
  x = (x | (x << 2)) & 0x09249249u;
  return x;
}
__host__ __device__ __forceinline__ unsigned spread_bits2(unsigned x) {
  x &= 0xffffu;
  x = (x | (x << 8)) & 0x00FF00FFu;
  x = (x | (x << 4)) & 0x0F0F0F0Fu;
  x = (x | (x << 2)) & 0x33333333u;
  x = (x | (x << 1)) & 0x55555555u;
  return x;
}
// the key bits cell coordinate c of dimension d contributes (a key is the OR
// of its dimensions' parts)
__host__ __device__ __forceinline__ unsigned key_part(const Grid& g, int d, int c) {
  const unsigned x = (unsigned)c;
  if (g.ndim == 1) return x;
  const int m1 = g.km[0], m2 = g.km[1], nd = g.ndim;
  const unsigned M1 = (1u << m1) - 1u;
  unsigned k = (nd == 3 ? spread_bits3(x & M1) : spread_bits2(x & M1)) << d;
  // the two dimensions with the most bits (u = km[2] < v = km[3], or the
  // same one twice): their next m2 bits interleaved, then the bits only the
  // longest has (a dimension without bits at a level adds zeros)
  const int u = g.km[2], v = g.km[3];
  if (d == u || d == v) {
    const unsigned M2 = (1u << m2) - 1u;
    if (u != v) k |= spread_bits2((x >> m1) & M2) << (nd * m1 + (d == v ? 1 : 0));
    k |= (x >> (m1 + m2)) << (nd * m1 + 2 * m2);
  }
  return k;
}
__host__ __device__ __forceinline__ int cell_key(const Grid& g, int cx, int cy, int cz) {
  if (g.ndim == 1) return cx;
  unsigned k = key_part(g, 0, cx) | key_part(g, 1, cy);
  if (g.ndim == 3) k |= key_part(g, 2, cz);
  return (int)k;
}

// Image of a cell at unwrapped coordinate c along one dimension.
//   kind 0: inside, 1: periodic image left of xmin (source near xmax),
//   2: periodic image right of xmax (source near xmin),
//   3: mirror about xmin, 4: mirror about xmax.
// Returns the source cell, or -1 if the unwrapped cell is unreachable.
__device__ __forceinline__ int cell_image(const Grid& g, int d, int c, int& kind) {
  const int n = g.nc[d];
  if (c >= 0 && c < n) { kind = 0; return c; }
  if (g.bound[d] == 3) {
    if (c < 0 && c >= -n) { kind = 1; return c + n; }
    if (c >= n && c < 2 * n) { kind = 2; return c - n; }
    return -1;
  }
  if (g.bound[d] == 2) {
    if (c < 0 && c >= -n) { kind = 3; return -1 - c; }
    if (c >= n && c < 2 * n) { kind = 4; return 2 * n - 1 - c; }
    return -1;
  }
  return -1;
}

// Coordinate of an image along one dimension, with the reference's ghost
// arithmetic (src/ghostND_mhd.f90:180-200: xnew = xperbound + dxshift for
// periodic, xbound - dxshift for reflecting, dxshift = x - xbound), and the
// existence test of a ghost copy (0 < distance to the wall < dxbound,
// src/ghostND_mhd.f90:172).  dxb is the reflecting dxbound radkern*h_j at
// ghost-creation time; periodic images of true neighbours always satisfy the
// upper bound, only the strict positivity is tested.
__device__ __forceinline__ bool image_coord(const Grid& g, int d, int kind, double xj,
                                            double dxb, double& ximg) {
#pragma clang fp contract(off)
  switch (kind) {
    case 0: ximg = xj; return true;
    case 1: {  // source near xmax, ghost at xmin + (x - xmax)
      double dx = g.xmax[d] - xj;
      ximg = g.xmin[d] + (xj - g.xmax[d]);
      return dx > 0.0;
    }
    case 2: {  // source near xmin, ghost at xmax + (x - xmin)
      double dx = xj - g.xmin[d];
      ximg = g.xmax[d] + (xj - g.xmin[d]);
      return dx > 0.0;
    }
    case 3: {  // mirror about xmin: xmin - (x - xmin)
      double dx = xj - g.xmin[d];
      ximg = g.xmin[d] - (xj - g.xmin[d]);
      return dx > 0.0 && dx < dxb;
    }
    default: {  // mirror about xmax: xmax - (x - xmax)
      double dx = g.xmax[d] - xj;
      ximg = g.xmax[d] - (xj - g.xmax[d]);
      return dx > 0.0 && dx < dxb;
    }
  }
}

// Velocity of an image as the reference's ghost loop leaves it
// (src/ghostND_mhd.f90:153-330, at most one reflecting dimension): a ghost
// is made while processing its highest imaged dimension c, whose reflection
// flips v_c; an edge ghost whose lower imaged dimension reflects gets the
// unflipped velocity (:281-282); a 3D corner ghost whose lowest dimension
// reflects gets -v in every component (:305-306).  flip[d] = 1 negates v_d.
__host__ __device__ __forceinline__ void image_flip(const int* kind, int* flip) {
  flip[0] = flip[1] = flip[2] = 0;
  int nk = 0, c = -1, r = -1;
  for (int d = 0; d < 3; d++) {
    if (kind[d]) { nk++; c = d; }
    if (kind[d] >= 3) r = d;
  }
  if (r < 0) return;
  if (r == c) flip[r] = 1;
  else if (nk == 3 && r == 0) flip[0] = flip[1] = flip[2] = 1;
}

// unwrapped cell range covering [x - rad, x + rad] along d (a small margin on
// both sides keeps every true neighbour in range; the exact FP64 distance test
// decides membership)
__device__ __forceinline__ void cell_range(const Grid& g, int d, double x, double rad, int& lo,
                                           int& hi) {
  double a = (x - rad - g.lo[d]) * g.inv_w[d];
  double b = (x + rad - g.lo[d]) * g.inv_w[d];
  lo = (int)floor(a - 1.0e-9);
  hi = (int)floor(b + 1.0e-9);
  if (g.bound[d] == 0) {
    if (lo < 0) lo = 0;
    if (hi > g.nc[d] - 1) hi = g.nc[d] - 1;
  } else {
    if (lo < -g.nc[d]) lo = -g.nc[d];
    if (hi > 2 * g.nc[d] - 1) hi = 2 * g.nc[d] - 1;
  }
}

// distance along d from x to unwrapped cell c, less a margin of 1e-9 cell
// widths (0 when x lies within the widened cell)
__device__ __forceinline__ double cell_gap(const Grid& g, int d, int c, double x) {
  const double a = g.lo[d] + c * g.w[d], b = a + g.w[d];
  const double gap = fmax(fmax(a - x, x - b), 0.0);
  return fmax(gap - 1.0e-9 * g.w[d], 0.0);
}

// FP64 squared separation with the reference's operation order and no FMA
// contraction: rij2 = dot_product(dx,dx), left to right
// (src/density_sums.f90:181, src/ratesND_mhd.f90:405)
template <int NDIM>
__device__ __forceinline__ double rij2_exact(const double* dx) {
#pragma clang fp contract(off)
  double r = dx[0] * dx[0];
  if (NDIM >= 2) r = r + dx[1] * dx[1];
  if (NDIM >= 3) r = r + dx[2] * dx[2];
  return r;
}

__device__ __forceinline__ double mul_exact(double a, double b) {
#pragma clang fp contract(off)
  return a * b;
}

// Accumulator of the pair sums.  FP32 pair terms are summed in FP64: on this
// hardware (no packed FP32 here) an FP64 add costs what an FP32 add costs, so
// a convert and an add (2 operations) replace a TwoSum compensation (6) and
// the sum of the FP32 terms is exact to FP64 rounding.  FP64 terms: plain sum.
template <typename T>
struct Acc {
  double s;
  __device__ __forceinline__ void zero() { s = 0.0; }
  __device__ __forceinline__ void add(T x) { s += (double)x; }
  __device__ __forceinline__ double value() const { return s; }
};

// square root, reciprocal and quotient of the pair math: FP32 uses the
// hardware v_sqrt_f32 / v_rcp_f32 (about 1 ulp), FP64 the correctly rounded
// operations
__device__ __forceinline__ float p_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
__device__ __forceinline__ double p_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float p_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ __forceinline__ double p_rcp(double x) { return 1.0 / x; }
__device__ __forceinline__ float p_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double p_div(double a, double b) { return a / b; }

// linear interpolation in q^2 from a table held in LDS
// (interpolate_kernels_dens, src/kernelND.f90:4599-4638)
template <typename T>
__device__ __forceinline__ T ktab(const T* __restrict__ tab, T q2, T ddq2, T dq2) {
  T t = q2 * ddq2;
  int index = (t < T(NDSP_IKERN + 1)) ? (int)t : NDSP_IKERN + 1;
  int index1 = index + 1;
  if (index > NDSP_IKERN) index = NDSP_IKERN;
  if (index1 > NDSP_IKERN) index1 = NDSP_IKERN;
  T dxx = q2 - T(index) * dq2;
  T w0 = tab[index];
  return w0 + (tab[index1] - w0) * ddq2 * dxx;
}

template <int N, typename T>
__device__ __forceinline__ T pown_t(T x) {
  if (N == 1) return x;
  if (N == 2) return x * x;
  return x * x * x;
}

// Workgroups are dispatched to the 8 XCDs of the MI355X round robin
// (workgroup b to XCD b mod 8), and each XCD has its own L2.  The pair
// kernels map workgroup b to tile xcd_tile(b) so that every XCD works on one
// contiguous range of the (spatially sorted) particles: neighbour gathers of
// adjacent tiles then hit the same L2.  A bijection of [0, gridDim.x).
#define NDSP_NXCD 8
// virtual launch position: block vl.b of vl.nb (blockIdx/gridDim for a
// launch of one simulation; a block of one member in a launch over a batch)
struct VL {
  int b, nb;
};
__device__ __forceinline__ int xcd_tile(VL vl) {
  const int nb = vl.nb, b = vl.b;
  const int x = b % NDSP_NXCD, q = b / NDSP_NXCD;
  int start = 0;
  for (int y = 0; y < x; y++) start += (nb - y + NDSP_NXCD - 1) / NDSP_NXCD;
  return start + q;
}

// as xcd_tile, for the first nt tiles only (nt <= vl.nb, a pass over part of
// the particles): every XCD takes a contiguous share of them, where xcd_tile
// would hand them all to the first XCDs; other workgroups get tiles >= nt
__device__ __forceinline__ int xcd_tile_n(VL vl, int nt) {
  const int b = vl.b;
  const int x = b % NDSP_NXCD, q = b / NDSP_NXCD;
  if (q >= (nt - x + NDSP_NXCD - 1) / NDSP_NXCD) return nt + b;
  int start = 0;
  for (int y = 0; y < x; y++) start += (nt - y + NDSP_NXCD - 1) / NDSP_NXCD;
  return start + q;
}

// atomics on non-negative doubles through their IEEE bit patterns
__device__ __forceinline__ void atomic_max_pos(double* a, double v) {
  atomicMax((unsigned long long*)a, (unsigned long long)__double_as_longlong(v));
}
__device__ __forceinline__ void atomic_min_pos(double* a, double v) {
  atomicMin((unsigned long long*)a, (unsigned long long)__double_as_longlong(v));
}
