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

// Per-node geometry byte.
enum : uint8_t {
  M_SOLID = 1u,
  M_BOUNDARY = 2u,
  M_PARTIAL = 4u,
  M_TOP = 8u,          // boundary node whose wall normal is mostly vertical
};
__host__ __device__ inline int dist_class(uint8_t m) { return m >> 4; }  // 0..11, 15 = far

// Grid index helpers.  Linear index = (ix*ny + iy)*nz + iz (3D) or ix*ny + iy
// (2D), last axis fastest, as the reference's (Q, nx, ny[, nz]) tensors.
struct Grid {
  int nx, ny, nz;        // nz == 1 for 2D
  int sx, sy;            // strides of x and y
  long long N;
  // magic-number division by nz and ny (unsigned, exact for 0 <= v < 2^31)
  uint32_t mz, sz_shift, my, sy_shift;
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
    uint32_t q = fdiv(idx, g.mz, g.sz_shift, (uint32_t)g.nz);
    c.z = (int)(idx - q * (uint32_t)g.nz);
    uint32_t q2 = fdiv(q, g.my, g.sy_shift, (uint32_t)g.ny);
    c.y = (int)(q - q2 * (uint32_t)g.ny);
    c.x = (int)q2;
  } else {
    uint32_t q = fdiv(idx, g.my, g.sy_shift, (uint32_t)g.ny);
    c.y = (int)(idx - q * (uint32_t)g.ny);
    c.x = (int)q;
    c.z = 0;
  }
  return c;
}

// Linear offset of the periodic neighbour at c + (dx, dy, dz).
template <int DIM>
__device__ __forceinline__ int nb_offset(const Grid& g, const Coord<DIM>& c, int dx, int dy,
                                         int dz) {
  int off = 0;
  if (dx) {
    int xx = c.x + dx;
    off += dx * g.sx;
    if (xx < 0) off += g.nx * g.sx;
    else if (xx >= g.nx) off -= g.nx * g.sx;
  }
  if (dy) {
    int yy = c.y + dy;
    off += dy * g.sy;
    if (yy < 0) off += g.ny * g.sy;
    else if (yy >= g.ny) off -= g.ny * g.sy;
  }
  if constexpr (DIM == 3) {
    if (dz) {
      int zz = c.z + dz;
      off += dz;
      if (zz < 0) off += g.nz;
      else if (zz >= g.nz) off -= g.nz;
    }
  }
  return off;
}

}  // namespace lbm
