// One-time geometry precompute on the device: the classification behind
// set_solid (lbm/fe_ac_solver.py:193-237, lbm/boundary.py:15-34) and the
// ConningtonLeeWetting constructor (lbm/fe_wetting.py:122-311).
//
// Every quantity is reproduced bit-exactly:
//  * boundary(x) = !solid(x) && solid(x + e_i) for some i >= 1 (periodic);
//  * wetting mask bit i = boundary(x) && solid(x + e_i);
//  * smooth normals: numpy.gradient of the fraction (central differences
//    inside, one-sided at the array ends, not periodic) in the fraction's own
//    precision, n = (-grad/|grad|) * boundary rounded to fp32;
//  * discrete normals (no fraction): minus the mean of the solid-neighbour
//    directions, in-bounds neighbours only, normalised in fp64, rounded to fp32;
//  * near-wall distance: 11 min-plus passes over the face neighbours
//    (periodic), integer valued, 100 at solid and unreached nodes.
#pragma once

#include "lbm_lattice.hpp"
#include "lbm_phys.hpp"

namespace lbm {

template <int DIM, int BS>
__global__ __launch_bounds__(BS) void k_boundary(Grid g, const uint8_t* __restrict__ solid,
                                                 uint8_t* __restrict__ bnd,
                                                 uint32_t* __restrict__ wmask) {
  constexpr int Q = Lat<DIM>::Q;
  const uint32_t x = blockIdx.x * BS + threadIdx.x;
  if (x >= g.N) return;
  const Coord<DIM> c = decompose<DIM>(g, x);
  uint32_t bits = 0;
  if (!solid[x]) {
#pragma unroll
    for (int i = 1; i < Q; ++i) {
      const int off = nb_offset<DIM>(g, c, E<DIM>(i, 0), E<DIM>(i, 1),
                                     DIM == 3 ? E<DIM>(i, 2) : 0);
      if (solid[(int)x + off]) bits |= 1u << i;
    }
  }
  bnd[x] = bits ? 1 : 0;
  wmask[x] = bits;
}

// numpy.gradient along one axis with unit spacing, edge_order=1.
template <typename FT>
__device__ __forceinline__ FT npgrad(const FT* __restrict__ f, uint32_t x, int i, int n,
                                     int stride) {
  if (n < 2) return FT(0);
  if (i == 0) return (f[x + stride] - f[x]) / FT(1);
  if (i == n - 1) return (f[x] - f[x - stride]) / FT(1);
  return (f[x + stride] - f[x - stride]) / FT(2);
}

template <typename FT, int DIM, int BS>
__global__ __launch_bounds__(BS) void k_normals_frac(Grid g, const FT* __restrict__ frac,
                                                     const uint8_t* __restrict__ bnd,
                                                     float* __restrict__ normal) {
  const uint32_t x = blockIdx.x * BS + threadIdx.x;
  if (x >= g.N) return;
  const Coord<DIM> c = decompose<DIM>(g, x);
  FT gr[DIM];
  gr[0] = npgrad<FT>(frac, x, c.x, g.nx, g.sx);
  gr[1] = npgrad<FT>(frac, x, c.y, g.ny, g.sy);
  if constexpr (DIM == 3) gr[2] = npgrad<FT>(frac, x, c.z, g.nz, g.sz);
  // Python's sum() starts from integer 0: ((0 + g0^2) + g1^2) + g2^2
  FT s = gr[0] * gr[0];
#pragma unroll
  for (int d = 1; d < DIM; ++d) s = s + gr[d] * gr[d];
  FT mag = (FT)sqrt(s);
  if (mag < FT(1e-10)) mag = FT(1e-10);
  const FT b = bnd[x] ? FT(1) : FT(0);
#pragma unroll
  for (int d = 0; d < DIM; ++d) normal[(long long)d * g.N + x] = (float)(((-gr[d]) / mag) * b);
}

template <int DIM, int BS>
__global__ __launch_bounds__(BS) void k_normals_discrete(Grid g,
                                                         const uint8_t* __restrict__ solid,
                                                         const uint8_t* __restrict__ bnd,
                                                         float* __restrict__ normal) {
  constexpr int Q = Lat<DIM>::Q;
  const uint32_t x = blockIdx.x * BS + threadIdx.x;
  if (x >= g.N) return;
  double v[DIM];
#pragma unroll
  for (int d = 0; d < DIM; ++d) v[d] = 0.0;
  if (bnd[x]) {
    const Coord<DIM> c = decompose<DIM>(g, x);
    int cnt = 0;
    for (int i = 1; i < Q; ++i) {
      const int px = c.x + E<DIM>(i, 0), py = c.y + E<DIM>(i, 1);
      const int pz = DIM == 3 ? c.z + E<DIM>(i, 2) : 0;
      if (px < 0 || px >= g.nx || py < 0 || py >= g.ny) continue;
      if (DIM == 3 && (pz < 0 || pz >= g.nz)) continue;
      const long long y = int_index<DIM>(g, px, py, pz);
      if (solid[y]) {
#pragma unroll
        for (int d = 0; d < DIM; ++d) v[d] = v[d] - (double)E<DIM>(i, d);
        ++cnt;
      }
    }
    if (cnt > 0) {
#pragma unroll
      for (int d = 0; d < DIM; ++d) v[d] = v[d] / (double)cnt;
      double s = v[0] * v[0];
#pragma unroll
      for (int d = 1; d < DIM; ++d) s = s + v[d] * v[d];
      const double nrm = sqrt(s);
      if (nrm > 1e-10) {
#pragma unroll
        for (int d = 0; d < DIM; ++d) v[d] = v[d] / nrm;
      }
    }
  }
#pragma unroll
  for (int d = 0; d < DIM; ++d) normal[(long long)d * g.N + x] = (float)v[d];
}

// Distance seed: 0 at boundary nodes, 100 elsewhere.
template <int BS>
__global__ __launch_bounds__(BS) void k_dist_seed(long long N, const uint8_t* __restrict__ bnd,
                                                  uint8_t* __restrict__ dist) {
  const long long x = (long long)blockIdx.x * BS + threadIdx.x;
  if (x >= N) return;
  dist[x] = bnd[x] ? 0 : 100;
}

template <int DIM, int BS>
__global__ __launch_bounds__(BS) void k_dist_pass(Grid g, const uint8_t* __restrict__ din,
                                                  uint8_t* __restrict__ dout) {
  const uint32_t x = blockIdx.x * BS + threadIdx.x;
  if (x >= g.N) return;
  const Coord<DIM> c = decompose<DIM>(g, x);
  int d = din[x];
#pragma unroll
  for (int ax = 0; ax < DIM; ++ax) {
#pragma unroll
    for (int s = -1; s <= 1; s += 2) {
      const int off = nb_offset<DIM>(g, c, ax == 0 ? s : 0, ax == 1 ? s : 0, ax == 2 ? s : 0);
      const int t = din[(int)x + off] + 1;
      d = t < d ? t : d;
    }
  }
  dout[x] = (uint8_t)(d > 100 ? 100 : d);
}

// Per-node geometry byte + partial-solid eps + S12 eligibility.
template <typename T, typename FT, int DIM, int BS>
__global__ __launch_bounds__(BS) void k_meta(Grid g, const uint8_t* __restrict__ solid,
                                             const uint8_t* __restrict__ bnd,
                                             const uint8_t* __restrict__ dist,
                                             const FT* __restrict__ frac,
                                             const float* __restrict__ normal,
                                             uint8_t* __restrict__ meta, T* __restrict__ eps,
                                             int wet) {
  const uint32_t x = blockIdx.x * BS + threadIdx.x;
  if (x >= g.N) return;
  const bool s = solid[x] != 0;
  uint8_t m = s ? M_SOLID : 0;
  // boundary classification is only used by the wetting path
  if (wet && bnd[x]) {
    m |= M_BOUNDARY;
    // |n_z| / max(|n|, 1e-10) > 0.5 in the run precision (fe_ac_solver.py:602-606)
    T n[DIM];
#pragma unroll
    for (int d = 0; d < DIM; ++d) n[d] = (T)normal[(long long)d * g.N + x];
    T nn = n[0] * n[0];
#pragma unroll
    for (int d = 1; d < DIM; ++d) nn = nn + n[d] * n[d];
    T mag = (T)sqrt(nn);
    if (mag < T(1e-10)) mag = T(1e-10);
    const T nz = n[DIM - 1] < T(0) ? -n[DIM - 1] : n[DIM - 1];
    if (nz / mag > T(0.5)) m |= M_TOP;
  }
  T e = T(0);
  if (frac) {
    const FT fr = frac[x];
    if (fr > FT(0.01) && fr < FT(0.99) && !s) {
      m |= M_PARTIAL;
      e = (T)fr;
    }
  }
  int d = dist[x];
  if (s) d = 100;
  const int cls = d <= 11 ? d : 15;
  m |= (uint8_t)(cls << 4);
  meta[x] = m;
  eps[x] = e;
}

// Direct wetting: for each boundary node the node one rounded wall-normal step
// into the fluid, coordinates clamped to the domain (not wrapped), or -1 when
// that node is solid (lbm/fe_wetting.py:313-359).  Internal node indices.
template <int DIM, int BS>
__global__ __launch_bounds__(BS) void k_nsw(Grid g, const uint8_t* __restrict__ bnd,
                                            const uint8_t* __restrict__ solid,
                                            const float* __restrict__ normal,
                                            int32_t* __restrict__ nsw) {
  const uint32_t x = blockIdx.x * BS + threadIdx.x;
  if (x >= g.N) return;
  int32_t r = -1;
  if (bnd[x]) {
    const Coord<DIM> c = decompose<DIM>(g, x);
    int cc[3] = {c.x, c.y, c.z};
    const int n[3] = {g.nx, g.ny, g.nz};
#pragma unroll
    for (int d = 0; d < DIM; ++d) {
      // torch.round: half to even, as rintf in the default rounding mode
      const int v = cc[d] + (int)rintf(normal[(long long)d * g.N + x]);
      cc[d] = v < 0 ? 0 : (v > n[d] - 1 ? n[d] - 1 : v);
    }
    const uint32_t q = int_index<DIM>(g, cc[0], cc[1], cc[2]);
    if (!solid[q]) r = (int32_t)q;
  }
  nsw[x] = r;
}

// Topmost solid index per column (last axis), -1 if the column has none.
// Columns are numbered as in the public arrays: x * ny + y (3D), x (2D).
template <int DIM, int BS>
__global__ __launch_bounds__(BS) void k_surface(Grid g, const uint8_t* __restrict__ solid,
                                                int16_t* __restrict__ surf) {
  const long long col = (long long)blockIdx.x * BS + threadIdx.x;
  if constexpr (DIM == 3) {
    if (col >= (long long)g.nx * g.ny) return;
    const int cx = (int)(col / g.ny), cy = (int)(col % g.ny);
    int h = -1;
    for (int z = g.nz - 1; z >= 0; --z)
      if (solid[int_index<3>(g, cx, cy, z)]) { h = z; break; }
    surf[col] = (int16_t)h;
  } else {
    if (col >= g.nx) return;
    int h = -1;
    for (int y = g.ny - 1; y >= 0; --y)
      if (solid[int_index<2>(g, (int)col, y, 0)]) { h = y; break; }
    surf[col] = (int16_t)h;
  }
}

}  // namespace lbm

namespace lbm {

// ---------------------------------------------------------------------------
// Device-side case generation (geometry/substrate.py:180-267 and
// lbm/fe_droplet.py:22-112 of the reference), used by the batch engine so a
// case is built without host round trips.  The substrate is written in the
// internal node order (it feeds the geometry precompute directly), the droplet
// in the public order (it goes through init_fields).  Same formulas and evaluation order
// as the numpy generators: fraction = clip(signed distance + 0.5, 0, 1) in
// fp64 (fp32 for flat), solid = fraction >= 0.5; droplet phi =
// clip(0.5 + 0.5 tanh(2 (R - r) / xi), 0, 1) evaluated in fp64 and rounded to
// fp32, u = u_impact * phi.
// ---------------------------------------------------------------------------
enum SubstrateKind : int { SUB_FLAT = 0, SUB_RIDGE = 1, SUB_CONVEX = 2, SUB_CONCAVE = 3 };

template <int DIM, int BS>
__global__ __launch_bounds__(BS) void k_gen_substrate(Grid g, int kind, double Rg, double* frac64,
                                                      float* frac32, uint8_t* solid) {
  const uint32_t x = blockIdx.x * BS + threadIdx.x;
  if (x >= g.N) return;
  const Coord<DIM> c = decompose<DIM>(g, x);
  const int k = DIM == 3 ? c.z : c.y;   // last axis: the wall-normal one
  double fr = k == 0 ? 1.0 : 0.0;
  const double cx = g.nx / 2.0, cy = g.ny / 2.0;
  if (kind != SUB_FLAT) {
    const double dxx = (double)c.x - cx;
    if (kind == SUB_RIDGE) {
      const double kk = (double)k;
      const double dist = sqrt(dxx * dxx + kk * kk);
      double part = (Rg - dist) + 0.5;
      part = part < 0.0 ? 0.0 : (part > 1.0 ? 1.0 : part);
      if (k > 0) fr = part > fr ? part : fr;
    } else {
      double s2 = dxx * dxx;
      if (DIM == 3) {
        const double dyy = (double)c.y - cy;
        s2 = s2 + dyy * dyy;
      }
      const double dkk = (double)k - Rg;
      const double dist = sqrt(s2 + dkk * dkk);
      if (kind == SUB_CONVEX) {
        double part = (Rg - dist) + 0.5;
        part = part < 0.0 ? 0.0 : (part > 1.0 ? 1.0 : part);
        fr = part > fr ? part : fr;
      } else {
        double part = (dist - Rg) + 0.5;
        part = part < 0.0 ? 0.0 : (part > 1.0 ? 1.0 : part);
        if ((double)k < Rg) fr = part > fr ? part : fr;
      }
    }
  }
  if (frac64) frac64[x] = fr;
  if (frac32) frac32[x] = (float)fr;
  solid[x] = fr >= 0.5 ? 1 : 0;
}

template <int DIM, int BS>
__global__ __launch_bounds__(BS) void k_gen_droplet(Grid g, double cx0, double cy0, double cz0,
                                                    double R, double xi, float ux, float uy,
                                                    float uz, float* phi, float* u) {
  const uint32_t x = blockIdx.x * BS + threadIdx.x;
  if (x >= g.N) return;
  const Coord<DIM> c = decompose<DIM>(g, x);
  // Python's sum() from integer 0: ((0 + a^2) + b^2) + c^2
  double r2 = 0.0;
  const double a = (double)c.x - cx0;
  r2 = r2 + a * a;
  const double b = (double)c.y - cy0;
  r2 = r2 + b * b;
  if (DIM == 3) {
    const double d = (double)c.z - cz0;
    r2 = r2 + d * d;
  }
  const double r = sqrt(r2);
  double C = 0.5 + 0.5 * tanh(2.0 * (R - r) / xi);
  C = C < 0.0 ? 0.0 : (C > 1.0 ? 1.0 : C);
  const float Cf = (float)C;
  // public order: these arrays go through init_fields like user input
  const uint32_t e = ext_index<DIM>(g, c);
  phi[e] = Cf;
  u[e] = ux * Cf;
  u[g.N + e] = uy * Cf;
  if (DIM == 3) u[2 * g.N + e] = uz * Cf;
}

}  // namespace lbm
