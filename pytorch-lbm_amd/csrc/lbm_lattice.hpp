// Lattice tables and device-side helpers shared by every kernel.
//
// D2Q9 / D3Q19 velocity sets, weights and opposites follow lbm/lattice.py:
// 36-96 of the reference (same direction numbering, since golden data and the
// public f/g layout depend on it).  Weights are the fp32-rounded values the
// reference keeps even for fp64 runs (lbm/lattice.py:21).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lbm {

template <int DIM> struct Lat;

template <> struct Lat<3> {
  static constexpr int Q = 19;
  static constexpr int D = 3;
};
template <> struct Lat<2> {
  static constexpr int Q = 9;
  static constexpr int D = 2;
};

// e[i][d] packed as constexpr functions so the compiler folds them into
// immediates after unrolling.
__host__ __device__ constexpr int ex3(int i) {
  constexpr int t[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, 0, 0, 0, 0};
  return t[i];
}
__host__ __device__ constexpr int ey3(int i) {
  constexpr int t[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, -1, 1};
  return t[i];
}
__host__ __device__ constexpr int ez3(int i) {
  constexpr int t[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
  return t[i];
}
__host__ __device__ constexpr int opp3(int i) {
  constexpr int t[19] = {0, 2, 1, 4, 3, 6, 5, 9, 10, 7, 8, 13, 14, 11, 12, 17, 18, 15, 16};
  return t[i];
}
__host__ __device__ constexpr int ex2(int i) {
  constexpr int t[9] = {0, 1, -1, 0, 0, 1, -1, -1, 1};
  return t[i];
}
__host__ __device__ constexpr int ey2(int i) {
  constexpr int t[9] = {0, 0, 0, 1, -1, 1, 1, -1, -1};
  return t[i];
}
__host__ __device__ constexpr int opp2(int i) {
  constexpr int t[9] = {0, 2, 1, 4, 3, 7, 8, 5, 6};
  return t[i];
}

template <int DIM> __host__ __device__ constexpr int E(int i, int d) {
  if constexpr (DIM == 3) {
    return d == 0 ? ex3(i) : (d == 1 ? ey3(i) : ez3(i));
  } else {
    return d == 0 ? ex2(i) : ey2(i);
  }
}
template <int DIM> __host__ __device__ constexpr int OPP(int i) {
  if constexpr (DIM == 3) return opp3(i); else return opp2(i);
}

// fp32 weights: 1/3, 1/18, 1/36 (D3Q19); 4/9, 1/9, 1/36 (D2Q9), each rounded
// to float once.
template <int DIM> __host__ __device__ constexpr float W32(int i) {
  if constexpr (DIM == 3) {
    return i == 0 ? (float)(1.0 / 3.0) : (i <= 6 ? (float)(1.0 / 18.0) : (float)(1.0 / 36.0));
  } else {
    return i == 0 ? (float)(4.0 / 9.0) : (i <= 4 ? (float)(1.0 / 9.0) : (float)(1.0 / 36.0));
  }
}

// Weight class of direction i: 0 rest, 1 axis, 2 diagonal (the lattice has
// three distinct weights).
template <int DIM> __host__ __device__ constexpr int wclass(int i) {
  if constexpr (DIM == 3) return i == 0 ? 0 : (i <= 6 ? 1 : 2);
  else return i == 0 ? 0 : (i <= 4 ? 1 : 2);
}

// Per-node geometry byte.
enum : uint8_t {
  M_SOLID = 1u,
  M_BOUNDARY = 2u,
  M_PARTIAL = 4u,
  M_TOP = 8u,          // boundary node whose wall normal is mostly vertical
};
__host__ __device__ inline int dist_class(uint8_t m) { return m >> 4; }  // 0..11, 15 = far

// Grid indexing.
//
// The public arrays (the reference's (Q, nx, ny[, nz]) tensors) run the last
// axis -- the wall-normal one -- fastest.  Inside the library nodes are
// ordered with that axis SLOWEST: 3D index = (z * nx + x) * ny + y, 2D index =
// y * nx + x.  A wave of 64 consecutive nodes then lies in one layer parallel
// to the substrate, so walls, solid and bulk fluid split the waves by layer
// instead of every wave holding a piece of a column from the floor up, and the
// wall-only work (bounce-back, wetting ghosts, gradient correction) runs in
// the few waves that hold wall nodes.  ext_index() maps a node to its position
// in the public arrays; every transfer across the API boundary goes through it.
struct Grid {
  int nx, ny, nz;        // nz == 1 for 2D
  int sx, sy, sz;        // internal strides of the x, y, z axes
  int wx, wy, wz;        // n_d * s_d: the jump of a periodic wrap along each axis
  long long N;
  // magic-number division by the fastest axis length (n1) and by the middle
  // one (n2; 3D only); exact for 0 <= v < 2^31
  uint32_t n1, m1, sh1, n2, m2, sh2;
};

__device__ __forceinline__ uint32_t fdiv(uint32_t v, uint32_t m, uint32_t sh, uint32_t d) {
  // d == 1 -> v; otherwise ((v * m) >> 32) >> sh  (Granlund-Montgomery)
  return d == 1 ? v : (__umulhi(v, m) >> sh);
}

template <int DIM> struct Coord {
  int x, y, z;
};

template <int DIM>
__device__ __forceinline__ Coord<DIM> decompose(const Grid& g, uint32_t idx) {
  Coord<DIM> c;
  if constexpr (DIM == 3) {
    const uint32_t q = fdiv(idx, g.m1, g.sh1, g.n1);     // idx / ny
    c.y = (int)(idx - q * g.n1);
    const uint32_t q2 = fdiv(q, g.m2, g.sh2, g.n2);      // / nx
    c.x = (int)(q - q2 * g.n2);
    c.z = (int)q2;
  } else {
    const uint32_t q = fdiv(idx, g.m1, g.sh1, g.n1);     // idx / nx
    c.x = (int)(idx - q * g.n1);
    c.y = (int)q;
    c.z = 0;
  }
  return c;
}

template <int DIM>
__device__ __forceinline__ uint32_t int_index(const Grid& g, int x, int y, int z) {
  return (uint32_t)(x * g.sx + y * g.sy + (DIM == 3 ? z * g.sz : 0));
}

// Position of a node in the public (last axis fastest) arrays.
template <int DIM>
__device__ __forceinline__ uint32_t ext_index(const Grid& g, const Coord<DIM>& c) {
  if constexpr (DIM == 3) return (uint32_t)((c.x * g.ny + c.y) * g.nz + c.z);
  else return (uint32_t)(c.x * g.ny + c.y);
}

// Periodic neighbour offsets of one node along each axis: p[d] to the +1
// neighbour, m[d] to the -1 neighbour (a wrap at the faces).  Computed once per
// node; the offset of any lattice direction is then a sum of at most DIM of
// them.
template <int DIM> struct AxisOff {
  int p[DIM], m[DIM];
};

template <int DIM>
__device__ __forceinline__ AxisOff<DIM> axis_offsets(const Grid& g, const Coord<DIM>& c) {
  AxisOff<DIM> o;
  const int cc[3] = {c.x, c.y, c.z};
  const int n[3] = {g.nx, g.ny, g.nz};
  const int st[3] = {g.sx, g.sy, g.sz};
  const int w[3] = {g.wx, g.wy, g.wz};
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    o.p[d] = cc[d] == n[d] - 1 ? st[d] - w[d] : st[d];
    o.m[d] = cc[d] == 0 ? w[d] - st[d] : -st[d];
  }
  return o;
}

// Offset of x + s * e_i (s = +1 or -1) from precomputed axis offsets.
template <int DIM>
__device__ __forceinline__ int dir_offset(const AxisOff<DIM>& o, int i, int s) {
  int off = 0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    const int e = E<DIM>(i, d) * s;
    if (e > 0) off += o.p[d];
    else if (e < 0) off += o.m[d];
  }
  return off;
}

// Linear offset of the periodic neighbour at c + (dx, dy, dz), |d*| <= 1.
template <int DIM>
__device__ __forceinline__ int nb_offset(const Grid& g, const Coord<DIM>& c, int dx, int dy,
                                         int dz) {
  const int cc[3] = {c.x, c.y, c.z};
  const int dd[3] = {dx, dy, dz};
  const int n[3] = {g.nx, g.ny, g.nz};
  const int st[3] = {g.sx, g.sy, g.sz};
  const int w[3] = {g.wx, g.wy, g.wz};
  int off = 0;
#pragma unroll
  for (int d = 0; d < DIM; ++d) {
    if (dd[d] > 0) off += cc[d] == n[d] - 1 ? st[d] - w[d] : st[d];
    else if (dd[d] < 0) off += cc[d] == 0 ? w[d] - st[d] : -st[d];
  }
  return off;
}

}  // namespace lbm
